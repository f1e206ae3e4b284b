"""The oracle and the HIP kernels against the frozen golden fixtures.

tests/golden/make_golden.py computes the fixtures from the definitions alone
(FP64 numpy from explicit source/receiver positions; closed forms of the rho
filter on cosines; the IBM single format), independently of the oracle's
code, so these tests pin both the oracle and the GPU path to the same
numbers. Bars: oracle FP64 vs fixture 1e-10 relative; GPU FAST/FP32 vs
fixture 1e-4 relative L2 (as against the live FP64 oracle); rho filter 2e-6;
IBM decode bit-exact.
"""
import os

import numpy as np
import pytest

import oracle
import psktm

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def mig():
    z = load("migration_3d_small.npz")
    nx, ny, nt = (int(v) for v in z["grid"])
    x0, dx, y0, dy, tau0, dtau = (float(v) for v in z["axes"])
    return z, (nx, x0, dx, ny, y0, dy, nt, tau0, dtau)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- oracle (CPU)

def test_oracle_migration_vs_golden():
    z, g = mig()
    args = (oracle.Grid(*g), z["mid"][:, 0], z["mid"][:, 1], z["off"][:, 0], z["off"][:, 1])
    img = oracle.pstm_f64(*args, z["traces"], z["vel"], float(z["dt"]), interp=oracle.LINEAR)
    assert rel_l2(img, z["image_linear"]) <= 1e-10
    gn = (z["traces"].astype(np.float64)
          * np.sqrt(np.arange(z["traces"].shape[1]) * float(z["dt"]))).astype(np.float32)
    img_g = oracle.pstm_f64(*args, gn, z["vel"], float(z["dt"]), interp=oracle.LINEAR,
                            gained=True)
    # the fixture gains in FP64; the oracle reads the binary32 gained samples
    assert rel_l2(img_g, z["image_linear_gained"]) <= 1e-7


def test_oracle_rho_vs_golden():
    z = load("rho_cosines.npz")
    for opt in (1, 2, 3):
        got = oracle.apply_rho_filter(z["x"].astype(np.float64), float(z["dw_us"]), opt)
        ref = z[f"opt{opt}"]
        assert np.max(np.abs(got - ref)) <= 1e-6 * np.max(np.abs(ref)), opt


def test_oracle_and_host_ibm_vs_golden():
    z = load("ibm_words.npz")
    ref = oracle.ibm32_to_ieee_ref(z["words"])
    want = z["reference_formula"]
    for g, w in zip(ref, want):
        assert abs(float(g) - w) <= 2 * float(np.spacing(np.float32(abs(w)))), (g, w)
    exact = psktm.ibm32_to_ieee_exact(z["words"])
    assert np.array_equal(exact, z["exact"].astype(np.float32))


# --------------------------------------------------------------- HIP (MI355X)

@pytest.mark.gpu
def test_gpu_migration_vs_golden(cuda):
    import torch
    z, g = mig()
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(cuda)
    ns, dt = z["traces"].shape[1], float(z["dt"])
    s = psktm.Survey(t(z["mid"][:, 0]), t(z["mid"][:, 1]), t(z["off"][:, 0]), t(z["off"][:, 1]),
                     ns, dt)
    plan = psktm.Plan(psktm.Grid(*g), s)
    img = psktm.migrate(plan, t(z["traces"]), t(z["vel"])).cpu().numpy()
    e = rel_l2(img, z["image_linear"])
    gained = psktm.trace_gain(t(z["traces"]), dt)
    img_g = psktm.migrate(plan, gained, t(z["vel"]), gained=True).cpu().numpy()
    eg = rel_l2(img_g, z["image_linear_gained"])
    print(f"golden 3-D: per-update weight {e:.2e}, gained {eg:.2e}")
    assert e <= 1e-4 and eg <= 1e-4


@pytest.mark.gpu
def test_gpu_rho_vs_golden(cuda):
    import torch
    z = load("rho_cosines.npz")
    x = torch.as_tensor(z["x"]).to(cuda)
    for opt in (1, 2, 3):
        got = psktm.rho_filter(x, float(z["dw_us"]), opt).cpu().numpy()
        ref = z[f"opt{opt}"]
        assert rel_l2(got, ref) <= 2e-6, opt


@pytest.mark.gpu
def test_gpu_ibm_vs_golden(cuda):
    import torch
    z = load("ibm_words.npz")
    w = torch.as_tensor(z["words"].view(np.int32)).to(cuda)
    got = psktm.ibm32_to_ieee_device(w).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), z["exact"].astype(np.float32).view(np.uint32))
