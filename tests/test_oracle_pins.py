"""Pins of the CPU oracle against mathematics it must satisfy.

The reference ships no tests, fixtures or golden vectors for this path, and
its C sources cannot be built as shipped (undefined structs), so every oracle
function is pinned here by closed forms, invariants and brute force written
independently from source/receiver positions (not from the oracle's A/B
distance terms). Each check is chosen so that a plausible mistake - a wrong
sign on the half offset, a transposed axis, a dropped weight factor, an
off-by-one sample index - fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import Grid

# the reference's sample interval is a float (struct jobParams)
DT = float(np.float32(0.004))


def brute_dsr_f64(grid, src, rec, tau_idx, bin_ij, vel_l):
    """DSR time from explicit source/receiver positions, FP64."""
    xx, yy, zz = grid.axes()
    i, j = bin_ij
    x = np.array([xx[i], yy[j]], dtype=np.float64)
    tau = float(zz[tau_idx])
    v = float(vel_l)
    ds = np.linalg.norm(np.asarray(src, np.float64) - x)
    dr = np.linalg.norm(np.asarray(rec, np.float64) - x)
    return math.sqrt(tau * tau + ds * ds / (v * v)) + math.sqrt(tau * tau + dr * dr / (v * v))


def tiny_case(seed=0, ntr=6, nx=5, ny=3, nt=64, ns=80):
    rng = np.random.default_rng(seed)
    grid = Grid(nx, 100.0, 25.0, ny, 50.0, 25.0, nt, 0.0, 0.004)
    src = rng.uniform(0, 300, size=(ntr, 2)).astype(np.float32)
    rec = rng.uniform(0, 300, size=(ntr, 2)).astype(np.float32)
    mid = ((src + rec) / 2).astype(np.float32)
    off = ((rec - src) / 2).astype(np.float32)
    # positions the operator actually sees: s = mid - off, r = mid + off (FP64)
    src = mid.astype(np.float64) - off.astype(np.float64)
    rec = mid.astype(np.float64) + off.astype(np.float64)
    vel = np.linspace(1500, 2500, nt).astype(np.float32)
    traces = rng.standard_normal((ntr, ns)).astype(np.float32)
    return grid, src, rec, mid, off, vel, traces, DT


def test_axes_closed_form():
    xx, yy, zz = oracle.init_three_axes(7, -3.5, 0.1, 4, 1000.0, 12.5, 9, 0.0, 0.002)
    i7 = np.arange(7, dtype=np.float32)
    assert np.array_equal(xx, np.float32(-3.5) + i7 * np.float32(0.1))
    assert np.array_equal(yy, np.float32(1000.0) + np.arange(4, dtype=np.float32) * np.float32(12.5))
    assert np.array_equal(zz, np.arange(9, dtype=np.float32) * np.float32(0.002))


def test_single_spike_support_and_amplitude():
    """A spike at n0 lands exactly where floor(t/dt) == n0 with amplitude
    100 * t * sqrt(tau^2/t) (src/psktmCPU.c:80-85), t from positions."""
    grid = Grid(9, 0.0, 25.0, 2, 0.0, 30.0, 200, 0.0, 0.002)
    src = np.array([[40.0, 10.0]], np.float32)
    rec = np.array([[160.0, 35.0]], np.float32)
    mid = (src + rec) / 2
    off = (rec - src) / 2
    ns, dt, n0 = 300, DT, 137
    tr = np.zeros((1, ns), np.float32)
    tr[0, n0] = 100.0
    vel = np.full(grid.nt, 2000.0, np.float32)
    img = oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, dt)
    xx, yy, zz = grid.axes()
    checked = 0
    for i in range(grid.nx):
        for j in range(grid.ny):
            for l in range(grid.nt):
                t = brute_dsr_f64(grid, src[0], rec[0], l, (i, j), 2000.0)
                q = t / dt
                frac = q - math.floor(q)
                if min(frac, 1 - frac) < 1e-4:
                    continue      # too close to a sample boundary for FP32 vs FP64
                expect = 100.0 * float(zz[l]) * math.sqrt(t) if math.floor(q) == n0 else 0.0
                assert img[i, j, l] == pytest.approx(expect, rel=2e-5, abs=1e-9), (i, j, l)
                checked += expect != 0.0
    assert checked > 5


def test_zero_offset_is_twice_tau():
    """At the midpoint bin and zero offset, t = 2 tau: a spike at n0 fills
    exactly the tau samples with floor(2 tau / dt) == n0."""
    grid = Grid(1, 500.0, 25.0, 1, 0.0, 25.0, 400, 0.0, 0.001)
    ns, dt, n0 = 250, 0.004, 100
    tr = np.zeros((1, ns), np.float32)
    tr[0, n0] = 1.0
    z = np.zeros(1, np.float32)
    img = oracle.pstm_f32(grid, np.array([500.0], np.float32), z, z, z, tr,
                          np.full(400, 1800.0, np.float32), dt)[0, 0]
    _, _, zz = grid.axes()
    nz = np.nonzero(img)[0]
    expect = [l for l in range(400) if math.floor(2 * float(zz[l]) / dt + 1e-9) == n0]
    assert list(nz) == expect and len(expect) == 2


def test_source_receiver_symmetry_bit_exact():
    grid, src, rec, mid, off, vel, tr, dt = tiny_case(1)
    a = oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, dt)
    b = oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], -off[:, 0], -off[:, 1], tr, vel, dt)
    assert np.array_equal(a, b)
    # a wrong half-offset convention (full offset) must change the image
    c = oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], 2 * off[:, 0], 2 * off[:, 1], tr, vel, dt)
    assert not np.array_equal(a, c)


def test_linearity_and_zero():
    grid, src, rec, mid, off, vel, tr, dt = tiny_case(2)
    rng = np.random.default_rng(5)
    tr2 = rng.standard_normal(tr.shape).astype(np.float32)
    f = lambda d: oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], d, vel, dt).astype(np.float64)
    lhs = f((2.0 * tr - 3.0 * tr2).astype(np.float32))
    rhs = 2.0 * f(tr) - 3.0 * f(tr2)
    assert np.linalg.norm(lhs - rhs) <= 1e-5 * np.linalg.norm(rhs)
    assert not np.any(f(np.zeros_like(tr)))


def test_sample_zero_never_read_and_last_sample():
    grid, src, rec, mid, off, vel, tr, dt = tiny_case(3)
    base = tr.copy()
    base[:, 0] = 0.0
    hot = base.copy()
    hot[:, 0] = 1e30
    f = lambda d: oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], d, vel, dt)
    assert np.array_equal(f(base), f(hot))


def test_fp32_restatement_vs_brute_force_fp64():
    """Smooth data so that rare index flips cost little: FP32 restatement vs
    an FP64 evaluation from source/receiver positions."""
    grid, src, rec, mid, off, vel, _, dt = tiny_case(4, ntr=5, nx=4, ny=2, nt=48, ns=120)
    n = np.arange(120)
    tr = np.stack([np.sin(2 * np.pi * n / (37.0 + 3 * k)) for k in range(5)]).astype(np.float32)
    img = oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, dt)
    _, _, zz = grid.axes()
    ref = np.zeros(img.shape)
    for k in range(5):
        for i in range(grid.nx):
            for j in range(grid.ny):
                for l in range(grid.nt):
                    t = brute_dsr_f64(grid, src[k], rec[k], l, (i, j), vel[l])
                    m = math.floor(t / dt)
                    if 0 < m < 120:
                        ref[i, j, l] += float(tr[k, m]) * float(zz[l]) * math.sqrt(t)
    assert np.linalg.norm(img - ref) <= 1e-4 * np.linalg.norm(ref)


def test_threads_and_racy_single_thread_bit_identical():
    grid, src, rec, mid, off, vel, tr, dt = tiny_case(6, ntr=12, nx=6, ny=2, nt=50, ns=90)
    args = (grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, dt)
    one = oracle.pstm_f32(*args, threads=1)
    many = oracle.pstm_f32(*args, threads=4)
    racy1 = oracle.ktm_migration_racy(*args, threads=1)
    assert np.array_equal(one, many)
    assert np.array_equal(one, racy1)


def test_fp64_trunc_matches_brute_force():
    grid, src, rec, mid, off, vel, tr, dt = tiny_case(7, ntr=4, nx=3, ny=2, nt=30, ns=70)
    out = oracle.pstm_f64(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, dt,
                          interp=oracle.TRUNC)
    _, _, zz = grid.axes()
    ref = np.zeros(out.shape)
    for k in range(4):
        for i in range(3):
            for j in range(2):
                for l in range(30):
                    t = brute_dsr_f64(grid, src[k], rec[k], l, (i, j), vel[l])
                    m = math.floor(t / dt)
                    if 0 < m < 70:
                        ref[i, j, l] += float(tr[k, m]) * float(zz[l]) * math.sqrt(t)
    assert np.allclose(out, ref, rtol=1e-9, atol=1e-12)


def test_fp64_linear_exact_on_linear_trace():
    """Linear interpolation reproduces a trace linear in sample index."""
    grid = Grid(2, 0.0, 40.0, 1, 0.0, 10.0, 60, 0.0, 0.004)
    mid = np.array([20.0], np.float32)
    off = np.array([150.0], np.float32)
    z = np.zeros(1, np.float32)
    ns, dt = 200, DT
    tr = (3.0 + 0.5 * np.arange(ns)).astype(np.float32)[None]
    vel = np.full(60, 2100.0, np.float32)
    out = oracle.pstm_f64(grid, mid, z, off, z, tr, vel, dt, interp=oracle.LINEAR)
    _, _, zz = grid.axes()
    xx, _, _ = grid.axes()
    for i in range(2):
        for l in range(60):
            src = np.array([mid[0] - off[0], 0.0])
            rec = np.array([mid[0] + off[0], 0.0])
            t = brute_dsr_f64(grid, src, rec, l, (i, 0), 2100.0)
            q = t / dt
            expect = (3.0 + 0.5 * q) * float(zz[l]) * math.sqrt(t) if q < ns - 1 else 0.0
            assert out[i, 0, l] == pytest.approx(expect, rel=1e-11, abs=1e-12)


def test_fp64_gained_closed_form_on_linear_trace():
    """Gained traces (sample n carries sqrt(n dt)) are migrated with weight tau
    alone: on a trace that is linear in n the interpolation is exact, so the
    image is tau * g(t/dt), from source/receiver positions in FP64. A weight
    that kept sqrt(t), or lost tau, fails here."""
    grid = Grid(3, -20.0, 30.0, 2, 0.0, 15.0, 50, 0.002, 0.004)
    mid = np.array([10.0, 35.0], np.float32)
    midy = np.array([5.0, -12.0], np.float32)
    offx = np.array([120.0, -60.0], np.float32)
    offy = np.array([40.0, 90.0], np.float32)
    ns, dt = 180, DT
    g = np.stack([1.5 + 0.25 * np.arange(ns), -2.0 + 0.125 * np.arange(ns)]).astype(np.float32)
    vel = np.linspace(1800, 2600, 50).astype(np.float32)
    out = oracle.pstm_f64(grid, mid, midy, offx, offy, g, vel, dt, interp=oracle.LINEAR,
                          gained=True)
    _, _, zz = grid.axes()
    ref = np.zeros(out.shape)
    for k in range(2):
        src = np.array([mid[k] - offx[k], midy[k] - offy[k]], np.float64)
        rec = np.array([mid[k] + offx[k], midy[k] + offy[k]], np.float64)
        for i in range(3):
            for j in range(2):
                for l in range(50):
                    q = brute_dsr_f64(grid, src, rec, l, (i, j), vel[l]) / dt
                    if q < ns - 1:
                        ref[i, j, l] += float(zz[l]) * (float(g[k, 0]) + float(g[k, 1] - g[k, 0]) * q)
    assert np.any(ref != 0)
    assert np.allclose(out, ref, rtol=1e-11, atol=1e-12)


def test_fp64_gained_equals_weighted_where_the_gain_is_exact():
    """TRUNC reads one sample at n = floor(t/dt): the per-update weight
    tau*sqrt(t) and the gained form tau*d[n]*sqrt(n dt) differ only by
    sqrt(t) versus sqrt(n dt). With d[n] = c/sqrt(n dt) the gained trace is
    the constant c, so gained = sum tau c, while the weighted form gives
    sum tau c sqrt(t/(n dt)) >= the gained value, within the factor
    sqrt((n+1)/n) of it."""
    grid, src, rec, mid, off, vel, _, dt = tiny_case(9, ntr=3, nx=3, ny=2, nt=40, ns=300)
    n = np.arange(300, dtype=np.float64)
    d = np.zeros((3, 300), np.float32)
    d[:, 1:] = (2.0 / np.sqrt(n[1:] * dt)).astype(np.float32)
    gtr = (d.astype(np.float64) * np.sqrt(n * dt)).astype(np.float32)
    w = oracle.pstm_f64(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], d, vel, dt,
                        interp=oracle.TRUNC)
    gn = oracle.pstm_f64(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], gtr, vel, dt,
                         interp=oracle.TRUNC, gained=True)
    _, _, zz = grid.axes()
    assert np.all(gn >= 0) and np.any(gn > 0)
    nz = gn > 0
    ratio = w[nz] / gn[nz]
    assert np.all(ratio >= 1 - 1e-6)
    assert np.all(ratio <= math.sqrt(2.0) + 1e-6)


@pytest.mark.parametrize("interp", [oracle.TRUNC, oracle.LINEAR])
def test_fp64_adjoint_dot_product(interp):
    rng = np.random.default_rng(11)
    grid = Grid(6, 0.0, 20.0, 3, 0.0, 20.0, 40, 0.0, 0.004)
    ntr, ns, dt = 9, 90, 0.004
    mid = rng.uniform(-20, 120, (ntr, 2)).astype(np.float32)
    off = rng.uniform(-80, 80, (ntr, 2)).astype(np.float32)
    vel = rng.uniform(1500, 2500, (6, 3, 40)).astype(np.float32)
    d = rng.standard_normal((ntr, ns)).astype(np.float32)
    m = rng.standard_normal((6, 3, 40)).astype(np.float32)
    Fd = oracle.pstm_f64(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], d, vel, dt,
                         vel3d=True, aperture=90.0, interp=interp)
    Ftm = oracle.pstm_adj_f64(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], m, vel, dt, ns,
                              vel3d=True, aperture=90.0, interp=interp)
    lhs = float(np.sum(Fd * m.astype(np.float64)))
    rhs = float(np.sum(d.astype(np.float64) * Ftm))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    assert abs(lhs) > 1e-3


def test_aperture_mask_rule():
    """Traces beyond the aperture contribute nothing; inside, all of them."""
    grid = Grid(3, 0.0, 100.0, 1, 0.0, 10.0, 40, 0.0, 0.004)
    mid = np.array([0.0, 150.0], np.float32)
    z = np.zeros(2, np.float32)
    off = np.array([10.0, 10.0], np.float32)
    tr = np.ones((2, 100), np.float32)
    vel = np.full(40, 2000.0, np.float32)
    full = oracle.pstm_f32(grid, mid, z, off, z, tr, vel, 0.004)
    ap = oracle.pstm_f32(grid, mid, z, off, z, tr, vel, 0.004, aperture=120.0)
    only0 = oracle.pstm_f32(grid, mid[:1], z[:1], off[:1], z[:1], tr[:1], vel, 0.004)
    only1 = oracle.pstm_f32(grid, mid[1:], z[1:], off[1:], z[1:], tr[1:], vel, 0.004)
    # bin 0 (x=0): trace 0 (d=0) in, trace 1 (d=150) out
    assert np.array_equal(ap[0], only0[0])
    # bin 1 (x=100): both in (d=100, 50)
    assert np.array_equal(ap[1], full[1])
    # bin 2 (x=200): trace 1 in (d=50), trace 0 out (d=200)
    assert np.array_equal(ap[2], only1[2])


IBM_KNOWN = [
    (0x41100000, 1.0), (0x42640000, 100.0), (0xC276A000, -118.625), (0x00000000, 0.0),
    (0x40800000, 0.5), (0xC1100000, -1.0), (0x3F100000, 1.0 / 256.0),
]


def test_ibm_reference_formula_known_words():
    """The reference divides by 2^24-1 (src/ioFunctions.c:33): every known word
    decodes to (1 + 2^-24/(1-2^-24)) times the true value, rounded to float."""
    words = np.array([w for w, _ in IBM_KNOWN], np.uint32)
    got = oracle.ibm32_to_ieee_ref(words)
    for (w, true), g in zip(IBM_KNOWN, got):
        if true == 0.0:
            assert g == 0.0
            continue
        frac = w & 0xFFFFFF
        e = ((w >> 24) & 0x7F) - 64
        sign = -1.0 if w & 0x80000000 else 1.0
        q = np.float32(np.float32(frac) / np.float32(16777215.0))
        assert g == np.float32(np.float32(sign * q) * np.float32(16.0 ** e))
        assert abs(float(g) - true) <= 2 * np.spacing(np.float32(abs(true)))
    assert got[0] == np.float32(1.0) + np.spacing(np.float32(1.0))


# ---------------------------- rho filter ---------------------------------

@pytest.mark.parametrize("M,k0", [(64, 5), (256, 17), (1024, 100)])
def test_rho_closed_forms_cosine(M, k0):
    dw = 4000.0
    Dr = dw * 1e-6
    n = np.arange(M)
    th = 2 * np.pi * k0 * n / M
    x = np.cos(th)[None]
    w0 = 2 * np.pi * k0 / (M * Dr)        # physical angular frequency
    o1 = oracle.apply_rho_filter(x, dw, 1)[0]
    o2 = oracle.apply_rho_filter(x, dw, 2)[0]
    o3 = oracle.apply_rho_filter(x, dw, 3)[0]
    # the reference's omega omits 1/M (scripts/ktmApplyRhoFilter.m:33)
    np.testing.assert_allclose(o1, M * w0 * np.cos(th), atol=1e-9 * M * w0)
    np.testing.assert_allclose(o2, np.sqrt(M * w0) * np.cos(th + np.pi / 4), atol=1e-9 * np.sqrt(M * w0))
    np.testing.assert_allclose(o3, -M * w0 * np.sin(th), atol=1e-9 * M * w0)


def test_rho_dc_and_nyquist():
    M, dw = 128, 2000.0
    Dr = dw * 1e-6
    dc = np.ones((1, M))
    for opt in (1, 2, 3):
        assert np.max(np.abs(oracle.apply_rho_filter(dc, dw, opt))) < 1e-9
    ny = ((-1.0) ** np.arange(M))[None]
    wn = np.pi * M / Dr
    np.testing.assert_allclose(oracle.apply_rho_filter(ny, dw, 1)[0], wn * ny[0], rtol=1e-12)
    np.testing.assert_allclose(oracle.apply_rho_filter(ny, dw, 2)[0],
                               np.sqrt(wn) * np.cos(np.pi / 4) * ny[0], rtol=1e-12)
    assert np.max(np.abs(oracle.apply_rho_filter(ny, dw, 3))) < 1e-6 * wn


def test_rho_padding_and_truncation():
    """m not a power of two: pad to M, keep the first m (ktmApplyRhoFilter.m:26,44)."""
    rng = np.random.default_rng(3)
    m = 100
    x = rng.standard_normal((2, m))
    assert oracle.rho_fft_length(m) == 128
    out = oracle.apply_rho_filter(x, 4000.0, 2)
    assert out.shape == (2, m)
    xp = np.zeros((2, 128))
    xp[:, :m] = x
    np.testing.assert_allclose(oracle.apply_rho_filter(xp, 4000.0, 2)[:, :m], out, rtol=1e-12, atol=1e-9)


def test_rho_matlab_fftshift_rotates_odd_trace_counts():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3, 64))
    per = oracle.rho_filter_per_trace(x, 4000.0, 2)
    full = oracle.apply_rho_filter(x, 4000.0, 2)
    np.testing.assert_allclose(full, np.roll(per, -1, axis=0), rtol=1e-12, atol=1e-9)
    x4 = rng.standard_normal((4, 64))
    np.testing.assert_allclose(oracle.apply_rho_filter(x4, 4000.0, 2),
                               oracle.rho_filter_per_trace(x4, 4000.0, 2), rtol=1e-12, atol=1e-9)


# ---------------------------- trace preparation ---------------------------

def test_prep_restatements_closed_forms():
    # sort: magnitudes 5, 5, 1, 13 -> stable order 2, 0, 1, 3
    hx = np.array([3.0, 4.0, 1.0, 5.0]); hy = np.array([4.0, 3.0, 0.0, 12.0])
    assert list(oracle.sorted_offset_order(hx, hy)) == [2, 0, 1, 3]
    # selection: round(2.5) = 3 (halves away from zero), window open at both ends
    hx = np.array([2.5, 2.4, 7.0, 10.0]); hy = np.zeros(4)
    assert list(oracle.const_offset_selection(hx, hy, 3.0, 1.0)) == [0]
    assert list(oracle.const_offset_selection(hx, hy, 8.5, 2.0)) == [2, 3]
    # padding
    p = oracle.pad_block(np.ones((3, 5)), 4)
    assert p.shape == (4, 8) and p.sum() == 15 and p[:3, :5].all()
    # upsampling of a linear ramp is exact, length (m-1)*2^N + 1
    r = np.arange(6.0)[None] * 2 + 1
    u = oracle.upsample_time(r, 2)
    assert u.shape == (1, 21) and np.allclose(u[0], 1 + 2 * np.arange(21) / 4.0)


def _binary32_loop(grid, mid, off, traces, vel, dt, weight="ref", dist="ref", ratio="ref",
                   contract=False):
    """src/psktmCPU.c:42-87 evaluated with numpy binary32 scalars-as-arrays:
    every operation rounded to float32 once, in the association of the C
    expressions (C evaluates a * b * c as (a * b) * c), traces added in k order.
    Vectorised over (i, j, l), where each image sample's sum stays in k order.
    The keyword arguments select plausible re-associations, for the check that
    the comparison can tell them apart."""
    f = np.float32
    X = f(grid.x0) + np.arange(grid.nx, dtype=f) * f(grid.dx)
    Y = f(grid.y0) + np.arange(grid.ny, dtype=f) * f(grid.dy)
    Z = f(grid.tau0) + np.arange(grid.nt, dtype=f) * f(grid.dtau)
    tauS = Z * Z
    vv = vel.astype(f)
    sS = f(1.0) / (vv * vv)
    dtf = f(dt)
    ns = traces.shape[1]
    img = np.zeros((grid.nx, grid.ny, grid.nt), f)
    for k in range(traces.shape[0]):
        mx, my, hx, hy = f(mid[k, 0]), f(mid[k, 1]), f(off[k, 0]), f(off[k, 1])
        if dist == "ref":
            ax, bx = (mx - X) - hx, (mx - X) + hx                   # :65-66
            ay, by = (my - Y) - hy, (my - Y) + hy                   # :71-72
        else:
            ax, bx = mx - (X + hx), mx - (X - hx)
            ay, by = my - (Y + hy), my - (Y - hy)
        A = (ax * ax)[:, None] + (ay * ay)[None, :]
        B = (bx * bx)[:, None] + (by * by)[None, :]
        A, B = A[:, :, None], B[:, :, None]
        if contract:          # tauS + A*sS as one fused rounding
            ua = (A.astype(np.float64) * sS + tauS).astype(f)
            ub = (B.astype(np.float64) * sS + tauS).astype(f)
        else:
            ua, ub = tauS + A * sS, tauS + B * sS
        t = np.sqrt(ua)                                             # :78
        t = t + np.sqrt(ub)                                         # :79
        q = t / dtf if ratio == "ref" else t * (f(1.0) / dtf)       # :80
        with np.errstate(invalid="ignore"):
            n = np.where(np.isfinite(q) & (q < ns), q, 0).astype(np.int64)
        ok = (n > 0) & (n < ns)                                     # :82
        d = traces[k][np.clip(n, 0, ns - 1)]
        with np.errstate(invalid="ignore", divide="ignore"):
            if weight == "ref":
                w = (d * t) * np.sqrt(tauS / t)                     # :84-85
            elif weight == "right":
                w = d * (t * np.sqrt(tauS / t))
            else:             # the algebraic identity tau * sqrt(t)
                w = d * (Z * np.sqrt(t))
        img = img + np.where(ok, w, f(0.0)).astype(f)
    return img


def test_binary32_operation_order_bit_exact():
    """The oracle's binary32 loop is bit-identical to an independent numpy
    restatement of src/psktmCPU.c:42-87 in the C expressions' association,
    and each plausible re-association - the weight as d*(t*s), the distance as
    mx-(x+hx), a fused tauS + A*sS, t*(1/dt) for t/dt, tau*sqrt(t) for
    t*sqrt(tau^2/t) - changes the bits on the same inputs, so this comparison
    pins the order and not only the value."""
    rng = np.random.default_rng(77)

    def case(ntr, nt):
        grid = Grid(5, 100.0, 12.5, 3, 40.0, 25.0, nt, 0.0, 0.002)
        ns = nt * 5 // 4
        mid = rng.uniform(60, 180, (ntr, 2)).astype(np.float32)
        off = rng.uniform(-150, 150, (ntr, 2)).astype(np.float32)
        traces = rng.standard_normal((ntr, ns)).astype(np.float32)
        vel = np.linspace(1500, 2400, grid.nt).astype(np.float32)
        return grid, mid, off, traces, vel, np.float32(0.004)

    small, large = case(7, 96), case(200, 400)
    for grid, mid, off, traces, vel, dt in (small, large):
        got = oracle.pstm_f32(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], traces, vel, dt)
        ref = _binary32_loop(grid, mid, off, traces, vel, dt)
        assert np.count_nonzero(ref) > 0.5 * ref.size
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # the truncated index hides t*(1/dt) except next to sample boundaries:
    # the larger case has such samples
    for kw, c in (({"weight": "right"}, small), ({"weight": "identity"}, small),
                  ({"dist": "alt"}, small), ({"contract": True}, small),
                  ({"ratio": "recip"}, large)):
        ref = _binary32_loop(*c)
        alt = _binary32_loop(*c, **kw)
        changed = np.count_nonzero(alt.view(np.uint32) != ref.view(np.uint32))
        assert changed > 0, kw


def test_f64_bin_list_equals_rectangles():
    """oracle.pstm_f64_bins on scattered bins equals the rectangle form's
    columns bit for bit (the same per-column loop)."""
    rng = np.random.default_rng(5)
    grid = Grid(9, 0.0, 25.0, 7, 0.0, 25.0, 80, 0.0, 0.004)
    ntr, ns = 60, 100
    mid = rng.uniform(0, 200, (ntr, 2)).astype(np.float32)
    off = rng.uniform(-100, 100, (ntr, 2)).astype(np.float32)
    tr = rng.standard_normal((ntr, ns)).astype(np.float32)
    vel = rng.uniform(1800, 2200, (9, 7, 80)).astype(np.float32)
    bins = [(0, 0), (8, 6), (4, 3), (0, 6), (8, 0), (2, 5)]
    for gained in (False, True):
        full = oracle.pstm_f64(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, 0.004,
                               vel3d=True, aperture=120.0, gained=gained)
        cols = oracle.pstm_f64_bins(grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel,
                                    0.004, bins, vel3d=True, aperture=120.0, gained=gained)
        assert np.count_nonzero(cols) > 0
        for b, (i, j) in enumerate(bins):
            assert np.array_equal(cols[b], full[i, j])
