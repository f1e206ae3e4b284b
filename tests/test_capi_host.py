"""C ABI of libpsktm.so on the CPU: exports, reference host API, error paths.

No compute kernel is launched here; the GPU parity tests live in test_gpu_*.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import psktm
from psktm import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "psktm.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.strip().startswith("#"))
    names = []
    for chunk in text.split(";"):
        chunk = chunk.strip()
        if "(" not in chunk or chunk.startswith("typedef"):
            continue
        m = re.search(r"(\w+)\s*\(", chunk)
        if m and m.group(1) not in ("if", "sizeof"):
            names.append(m.group(1))
    return names


def test_header_parsed():
    names = header_functions()
    assert "ktmMigrationGPU" in names and "psktm_migrate" in names and "initThreeAxes" in names
    assert "ktmMigrationCPU" in names
    assert len(names) >= 30


def test_library_exports_every_header_symbol():
    L = psktm.load()
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, f"declared in include/psktm.h but not exported: {missing}"


def test_binding_signatures_cover_header():
    names = set(header_functions())
    assert names <= set(_lib._SIGNATURES), names - set(_lib._SIGNATURES)


def test_version_and_device_count_never_fail():
    L = psktm.load()
    assert b"gfx950" in L.psktm_version()
    assert L.psktm_device_count() >= 0


def test_axes_match_oracle_bit_exact():
    X, Y, Z = psktm.ImageGrid(37, -12.5, 0.3), psktm.ImageGrid(5, 7.0, 12.5), psktm.ImageGrid(1500, 0.0, 0.002)
    xx, yy, zz = psktm.init_three_axes(X, Y, Z)
    ox, oy, oz = oracle.init_three_axes(37, -12.5, 0.3, 5, 7.0, 12.5, 1500, 0.0, 0.002)
    assert np.array_equal(xx, ox) and np.array_equal(yy, oy) and np.array_equal(zz, oz)
    y2, z2 = psktm.init_two_axes(Y, Z)
    assert np.array_equal(y2, oy) and np.array_equal(z2, oz)


def _ibm_words(rng, n=20000):
    w = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    # all exponents including the overflow/underflow ends, plus zero fractions
    special = [(s << 31) | (e << 24) | f for s in (0, 1) for e in (0, 1, 32, 64, 65, 90, 95, 96, 127)
               for f in (0, 1, 0x100000, 0xFFFFFF)]
    return np.concatenate([w, np.array(special, dtype=np.uint32)])


def test_ibm_reference_decode_matches_oracle_bit_exact():
    rng = np.random.default_rng(1)
    w = _ibm_words(rng)
    got = psktm.ibm32_to_ieee(w)
    ref = oracle.ibm32_to_ieee_ref(w)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_ibm_exact_decode_closed_form():
    rng = np.random.default_rng(2)
    w = _ibm_words(rng)
    got = psktm.ibm32_to_ieee_exact(w)
    frac = (w & 0xFFFFFF).astype(np.float64)
    e = ((w >> 24) & 0x7F).astype(np.int64) - 64
    val = np.ldexp(frac, 4 * e - 24)            # exact in float64 for this range
    val = np.where(w & 0x80000000, -val, val)
    with np.errstate(over="ignore"):
        expect = val.astype(np.float32)          # one rounding: binary64 -> binary32
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32))
    known = np.array([0x41100000, 0x42640000, 0xC276A000, 0], np.uint32)
    assert list(psktm.ibm32_to_ieee_exact(known)) == [1.0, 100.0, -118.625, 0.0]


def _libc():
    c = ctypes.CDLL(None)
    c.fopen.restype = ctypes.c_void_p
    c.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    c.fclose.argtypes = [ctypes.c_void_p]
    c.free.argtypes = [ctypes.c_void_p]
    return c


def test_binary_io_roundtrip(tmp_path):
    L, c = psktm.load(), _libc()
    vals = np.arange(10, dtype=np.float32) * 1.5 - 3
    p = str(tmp_path / "x.bin").encode()
    f = c.fopen(p, b"wb")
    L.fileWrite(f, vals.ctypes.data_as(ctypes.c_void_p), 10)
    c.fclose(f)
    with open(p, "ab") as fh:
        fh.write(b"\x01\x02")                   # trailing partial word is not counted
    f = c.fopen(p, b"rb")
    assert L.fileSizeFourBytes(f) == 10
    ptr = L.binReadFloats(f, 10)
    got = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_float)), (10,)).copy()
    c.free(ptr)
    assert np.array_equal(got, vals)
    ptr = L.binReadUints(f, 10)                  # the reader rewinds after each read
    gotu = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint)), (10,)).copy()
    c.free(ptr)
    assert np.array_equal(gotu, vals.view(np.uint32))
    c.fclose(f)


def test_load_input_data_flags(tmp_path):
    L = psktm.load()
    work = tmp_path / "run"
    data = tmp_path / "data"
    work.mkdir()
    data.mkdir()
    cwd = os.getcwd()
    try:
        os.chdir(work)
        fp = _lib.FileInputPtrs()
        L.loadInputData(ctypes.byref(fp))
        assert fp.flag == 1 and b"Can't open ../data/" in L.psktm_last_error()
        for n in ("inputData", "inputSrcX", "inputSrcY", "inputRecX", "inputRecY"):
            (data / f"{n}.bin").write_bytes(b"\0" * 8)
        fp = _lib.FileInputPtrs()
        L.loadInputData(ctypes.byref(fp))
        assert fp.flag == 0 and all([fp.inputData, fp.inputSrcX, fp.inputSrcY, fp.inputRecX, fp.inputRecY])
        c = _libc()
        for h in (fp.inputData, fp.inputSrcX, fp.inputSrcY, fp.inputRecX, fp.inputRecY):
            c.fclose(h)
    finally:
        os.chdir(cwd)


def test_load_ktm_config(tmp_path):
    p = tmp_path / "inp.txt"
    p.write_text("1000.000000 0.004000 2048.000000 0.000000 12.500000 64.000000 "
                 "0.000000 25.000000 1.000000 0.000000 0.002000 1000.000000 ")
    X, Y, Z, J = psktm.load_ktm_config(str(p))
    assert (J.tracePts, J.traceNo) == (1000, 2048) and J.traceSampleRate == np.float32(0.004)
    assert (X.imagePnts, X.imageStep, Y.imagePnts, Z.imagePnts, Z.imageStep) == \
        (64, 12.5, 1, 1000, np.float32(0.002))
    p.write_text("1 2 3")
    with pytest.raises(psktm.PsktmError):
        psktm.load_ktm_config(str(p))
    with pytest.raises(psktm.PsktmError):
        psktm.load_ktm_config(str(tmp_path / "missing.txt"))


def test_legacy_stubs_are_callable():
    L = psktm.load()
    L.loadRTM2dConfig()
    L.loadKTMConfig()
    L.loadKDMConfig()


def _small_job():
    X, Y, Z = psktm.ImageGrid(4, 0.0, 10.0), psktm.ImageGrid(1, 0.0, 10.0), psktm.ImageGrid(16, 0.0, 0.004)
    J = psktm.JobParams(3, 20, 0.004)
    z3 = np.zeros(3, np.float32)
    return X, Y, Z, J, z3, np.ones((3, 20), np.float32), np.full(16, 2000.0, np.float32)


@pytest.mark.parametrize("entry,name", [(psktm.ktm_migration_gpu, "ktmMigrationGPU"),
                                        (psktm.ktm_migration_cpu, "ktmMigrationCPU")])
def test_ktm_migration_invalid_inputs_leave_zero_image(entry, name):
    X, Y, Z, J, z3, tr, v = _small_job()
    img = np.full(64, 7.0, np.float32)
    J.tracePts = 0
    out = entry(X, Y, Z, J, z3, z3, z3, z3, tr, v, image=img)
    assert not np.any(out) and "invalid job" in psktm.last_error()
    assert psktm.last_error().startswith(name)


@pytest.mark.parametrize("entry", [psktm.ktm_migration_gpu, psktm.ktm_migration_cpu])
def test_ktm_migration_without_device_fails_cleanly(entry):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present; the no-device path is exercised on CPU hosts")
    X, Y, Z, J, z3, tr, v = _small_job()
    img = np.full(64, 7.0, np.float32)
    out = entry(X, Y, Z, J, z3, z3, z3, z3, tr, v, image=img)
    assert not np.any(out)
    assert "device" in psktm.last_error().lower()
