"""The file-driven job (psktm_main): the reference's on-disk formats end to end.

A small job is written the way scripts/ktmWriteTraceBlockConstOffset.m and
scripts/ktmWriteVelModel.m write it (IBM words or floats, float coordinates,
inp.txt), migrated by the driver through ktmMigrationGPU, and outputData.bin
is compared with the oracle on the same decoded traces.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import psktm
from segy_files import ibm_encode

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PSKTM_DRIVER: the host-sanitizer test run's instrumented build
DRIVER = os.environ.get("PSKTM_DRIVER") or os.path.join(ROOT, "seismic-migration_amd", "psktm",
                                                        "psktm_main")


def write_job(d, ntr=24, ns=200, grid=(8, 0.0, 20.0, 2, 0.0, 30.0, 150, 0.0, 0.004),
              float_mode=False, seed=3):
    rng = np.random.default_rng(seed)
    os.makedirs(d, exist_ok=True)
    sx, sy = rng.uniform(0, 150, ntr), rng.uniform(0, 40, ntr)
    rx, ry = rng.uniform(0, 150, ntr), rng.uniform(0, 40, ntr)
    tr = rng.standard_normal((ntr, ns)).astype(np.float32)
    if float_mode:
        tr.tofile(os.path.join(d, "inputData.bin"))
        dec = tr
    else:
        w = ibm_encode(tr)
        w.tofile(os.path.join(d, "inputData.bin"))
        dec = psktm.ibm32_to_ieee(w).reshape(ntr, ns)
    for name, v in (("inputSrcX", sx), ("inputSrcY", sy), ("inputRecX", rx), ("inputRecY", ry)):
        v.astype(np.float32).tofile(os.path.join(d, name + ".bin"))
    nx, x0, dx, ny, y0, dy, nt, t0, dtau = grid
    vel = np.linspace(1600, 2600, nt).astype(np.float32)
    vel.tofile(os.path.join(d, "velModel.bin"))
    a = [ns, 0.004, ntr, x0, dx, nx, y0, dy, ny, t0, dtau, nt]
    with open(os.path.join(d, "inp.txt"), "w") as f:
        f.write(" ".join(f"{v:f}" for v in a) + " ")
    f32 = lambda v: v.astype(np.float32)
    mx = (f32(rx) + f32(sx)) / np.float32(2.0)
    my = (f32(ry) + f32(sy)) / np.float32(2.0)
    hx = (f32(rx) - f32(sx)) / np.float32(2.0)
    hy = (f32(ry) - f32(sy)) / np.float32(2.0)
    return dec, (mx, my, hx, hy), vel, grid


def run(*args):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)


def test_driver_built():
    assert os.access(DRIVER, os.X_OK)


def test_missing_files_fail_cleanly(tmp_path):
    r = run("--data", str(tmp_path))
    assert r.returncode == 1 and "inp.txt" in r.stderr
    write_job(str(tmp_path))
    os.remove(tmp_path / "inputRecY.bin")
    r = run("--data", str(tmp_path))
    assert r.returncode == 1 and "inputRecY.bin" in r.stderr


def test_short_trace_file_rejected(tmp_path):
    write_job(str(tmp_path))
    with open(tmp_path / "inputData.bin", "r+b") as f:
        f.truncate(100)
    r = run("--data", str(tmp_path))
    assert r.returncode == 1 and "fewer samples" in r.stderr


def test_without_gpu_reports_device(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    write_job(str(tmp_path))
    r = run("--data", str(tmp_path))
    assert r.returncode == 1 and "device" in r.stderr.lower()


@pytest.mark.gpu
@pytest.mark.parametrize("float_mode,cpu_entry", [(False, False), (True, False), (False, True)])
def test_driver_end_to_end_bit_exact(tmp_path, float_mode, cpu_entry):
    """The job through ktmMigrationGPU, or through ktmMigrationCPU
    (--cpu-entry), gives the reference loop's image bit for bit."""
    dec, (mx, my, hx, hy), vel, grid = write_job(str(tmp_path), float_mode=float_mode)
    args = ["--data", str(tmp_path)] + (["--float"] if float_mode else []) + \
        (["--cpu-entry"] if cpu_entry else [])
    r = run(*args)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(tmp_path / "outputData.bin", dtype=np.float32)
    ref = oracle.pstm_f32(oracle.Grid(*grid), mx, my, hx, hy, dec, vel, np.float32(0.004))
    assert img.size == ref.size
    assert np.array_equal(img.view(np.uint32), ref.reshape(-1).view(np.uint32))
