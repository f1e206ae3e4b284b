"""Fast HIP migration (plan + psktm_migrate) against the FP64 oracle.

Tolerances (relative L2 over the compared image, FP64 reference):
  LINEAR + FAST + FP32 traces   <= 1e-4
  LINEAR + FAST + FP16 traces   <= 1e-3   (the headline metric's bound)
  LINEAR + FAST + BF16 traces   <= 8e-3   (reported form, 8-bit mantissa)
  TRUNC  + FAST + FP32 traces   <= 2e-2   (index flips near sample
                                           boundaries; smooth data)
"""
import numpy as np
import pytest

import oracle
import psktm
from psktm import synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def small_3d(seed=3, ntr=3000, nx=13, ny=11, nt=300, ns=320, aperture=400.0, dxy=25.0):
    rng = np.random.default_rng(seed)
    ext = 8 * dxy
    mid = rng.uniform(-ext, (nx - 1) * dxy + ext, (ntr, 2)).astype(np.float32)
    mid[:, 1] = rng.uniform(-ext, (ny - 1) * dxy + ext, ntr).astype(np.float32)
    off = rng.uniform(-300, 300, (ntr, 2)).astype(np.float32)
    n = np.arange(ns)
    f = rng.uniform(0.01, 0.06, (ntr, 1))
    ph = rng.uniform(0, 2 * np.pi, (ntr, 1))
    tr = (np.sin(2 * np.pi * f * n[None] + ph) * rng.uniform(0.5, 2.0, (ntr, 1))).astype(np.float32)
    grid = (nx, 0.0, dxy, ny, 0.0, dxy, nt, 0.0, 0.002)
    xs = np.arange(nx, dtype=np.float32) * np.float32(dxy)
    ys = np.arange(ny, dtype=np.float32) * np.float32(dxy)
    t0 = 2 * np.arange(nt, dtype=np.float32) * np.float32(0.002)
    vel = (1500 + 600 * t0[None, None, :] + 200 * np.sin(xs[:, None, None] / 200.0) *
           np.cos(ys[None, :, None] / 300.0)).astype(np.float32)
    return grid, mid, off, tr, vel, 0.004, aperture


def gpu_plan(dev, grid, mid, off, ns, dt, aperture, interp=psktm.INTERP_LINEAR, **kw):
    import torch
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)
    s = psktm.Survey(t(mid[:, 0]), t(mid[:, 1]), t(off[:, 0]), t(off[:, 1]), ns, dt)
    return psktm.Plan(psktm.Grid(*grid), s, interp=interp, aperture=aperture, **kw), t


def oracle_f64(grid, mid, off, tr, vel, dt, aperture, interp, vel3d=True, bins=None):
    return oracle.pstm_f64(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr,
                           vel, dt, vel3d=vel3d, aperture=aperture, interp=interp, bins=bins,
                           threads=0)


def test_linear_fp32_3d_vs_fp64(cuda):
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d()
    print(f"case linear fp32 3d grid={grid} ntr={len(mid)}", flush=True)
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap)
    img = psktm.migrate(plan, t(tr), t(vel)).cpu().numpy()
    ref = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
    e = rel_l2(img, ref)
    print(f"rel-L2 = {e:.3e}")
    assert e <= 1e-4
    # exact update count: in-aperture (trace, bin) pairs x nt, binary32 rule
    xs = np.float32(0.0) + np.arange(grid[0], dtype=np.float32) * np.float32(grid[2])
    ys = np.float32(0.0) + np.arange(grid[3], dtype=np.float32) * np.float32(grid[5])
    ex = mid[:, 0][:, None, None] - xs[None, :, None]
    ey = mid[:, 1][:, None, None] - ys[None, None, :]
    pairs = np.count_nonzero((ex * ex + ey * ey) <= np.float32(ap) * np.float32(ap))
    assert plan.updates() == pairs * grid[6]
    # deterministic
    img2 = psktm.migrate(plan, t(tr), t(vel)).cpu().numpy()
    assert np.array_equal(img, img2)


def test_fp16_bf16_traces_vs_fp64(cuda):
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=4)
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap)
    ref = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
    errs = {}
    for dtype, tol in ((torch.float16, 1e-3), (torch.bfloat16, 8e-3)):
        q, sc = psktm.quantize(t(tr), dtype)
        img = psktm.migrate(plan, q, t(vel), scale=sc).cpu().numpy()
        errs[str(dtype)] = rel_l2(img, ref)
        assert errs[str(dtype)] <= tol, errs
    print("rel-L2", errs)


def test_trunc_fast_vs_reference_restatement(cuda):
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=5, ntr=1500)
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap, interp=psktm.INTERP_TRUNC)
    img = psktm.migrate(plan, t(tr), t(vel)).cpu().numpy()
    ref32 = oracle.pstm_f32(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr,
                            vel, dt, vel3d=True, aperture=ap)
    ref64 = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.TRUNC)
    e32, e64 = rel_l2(img, ref32), rel_l2(img, ref64)
    print(f"TRUNC+FAST rel-L2 vs binary32 reference {e32:.3e}, vs FP64 {e64:.3e}")
    assert e32 <= 2e-2 and e64 <= 2e-2


def test_line_2d_no_aperture_vs_fp64(cuda):
    d = synth.line_2d(n_cmp=32, n_off=16, ns=500, nt=500)
    mid = np.stack([d.mid_x, d.mid_y], 1)
    off = np.stack([d.off_x, d.off_y], 1)
    plan, t = gpu_plan(cuda, d.grid, mid, off, d.traces.shape[1], d.dt, 0.0)
    img = psktm.migrate(plan, t(d.traces), t(d.vel)).cpu().numpy()
    ref = oracle_f64(d.grid, mid, off, d.traces, d.vel, d.dt, 0.0, oracle.LINEAR, vel3d=False)
    assert rel_l2(img, ref) <= 1e-4
    assert plan.updates() == len(mid) * d.grid[0] * d.grid[6]


@pytest.mark.parametrize("tile,stage", [((4, 4), 1), ((8, 2), 3), ((2, 4), 8), ((4, 2), 4)])
def test_tile_shapes_and_stages_agree(cuda, tile, stage):
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=6, ntr=1200, nx=9, ny=7)
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap, tile_x=tile[0], tile_y=tile[1],
                       traces_per_stage=stage)
    img = psktm.migrate(plan, t(tr), t(vel)).cpu().numpy()
    ref = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
    assert rel_l2(img, ref) <= 1e-4


def test_accumulate_and_1d_velocity(cuda):
    import torch
    grid, mid, off, tr, _, dt, ap = small_3d(seed=7, ntr=800, nx=6, ny=5)
    v1 = np.linspace(1500, 3000, grid[6]).astype(np.float32)
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap)
    a = psktm.migrate(plan, t(tr), t(v1))
    b = a.clone()
    psktm.migrate(plan, t(tr), t(v1), image=b, accumulate=True)
    assert torch.allclose(b, 2 * a, rtol=1e-6, atol=0)
    ref = oracle_f64(grid, mid, off, tr, v1, dt, ap, oracle.LINEAR, vel3d=False)
    assert rel_l2(a.cpu().numpy(), ref) <= 1e-4


def test_traces_outside_aperture_and_empty(cuda):
    import torch
    grid = (5, 0.0, 20.0, 4, 0.0, 20.0, 64, 0.0, 0.004)
    mid = np.array([[5000.0, 5000.0], [-3000.0, 10.0]], np.float32)
    off = np.zeros((2, 2), np.float32)
    tr = np.ones((2, 80), np.float32)
    plan, t = gpu_plan(cuda, grid, mid, off, 80, 0.004, 200.0)
    img = psktm.migrate(plan, t(tr), torch.full((64,), 2000.0, device=cuda))
    assert plan.updates() == 0 and not torch.any(img)
    e = np.zeros((0, 2), np.float32)
    plan0, t = gpu_plan(cuda, grid, e, e, 80, 0.004, 200.0)
    img0 = psktm.migrate(plan0, torch.zeros((0, 80), device=cuda), torch.full((64,), 2000.0, device=cuda))
    assert not torch.any(img0)


def test_diffractor_survey_small(cuda):
    """The 3-D generator at reduced size: migrated energy focuses, FP64 parity."""
    s = synth.diffractor_3d(4000, nx=24, ny=20, nt=400, ns=420, n_diff=6, aperture=300.0,
                            device="cpu")
    mid = np.stack([s.mid_x.numpy(), s.mid_y.numpy()], 1)
    off = np.stack([s.off_x.numpy(), s.off_y.numpy()], 1)
    tr = s.traces.numpy()
    vel = s.vel.numpy()
    plan, t = gpu_plan(cuda, s.grid, mid, off, tr.shape[1], s.dt, s.aperture)
    img = psktm.migrate(plan, t(tr), t(vel)).cpu().numpy()
    ref = oracle_f64(s.grid, mid, off, tr, vel, s.dt, s.aperture, oracle.LINEAR,
                     bins=(8, 16, 6, 14))
    assert rel_l2(img[8:16, 6:14], ref) <= 1e-4


@pytest.mark.parametrize("ns,offset", [(7, 0), (301, 0), (333, 3), (1001, 5)])
def test_odd_sample_counts_and_unaligned_rows(cuda, ns, offset):
    """Trace rows at every 16-byte misalignment the LDS fill handles: odd
    sample counts, and trace arrays that start `offset` elements into a larger
    buffer. Each storage type is compared with the FP64 oracle run on exactly
    the values the kernel decodes (stored sample x scale in binary32)."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=11, ntr=900, nx=9, ny=8, nt=260, ns=ns)
    dt = max(dt, 1.2 / ns)          # traces long enough to reach the image
    print(f"case ns={ns} offset={offset} dt={dt}", flush=True)
    plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, ap)
    errs = {}
    for dtype in (torch.float32, torch.float16, torch.bfloat16):
        if dtype == torch.float32:
            data, sc, src = t(tr), None, tr
        else:
            data, sc = psktm.quantize(t(tr), dtype)
            src = (data.float() * sc[:, None]).cpu().numpy()
        big = torch.zeros(data.numel() + offset, dtype=dtype, device=cuda)
        view = big[offset:].view(data.shape)
        view.copy_(data)
        img = psktm.migrate(plan, view, t(vel), scale=sc).cpu().numpy()
        ref = oracle_f64(grid, mid, off, src, vel, dt, ap, oracle.LINEAR)
        assert np.linalg.norm(ref) > 0
        errs[str(dtype)] = rel_l2(img, ref)
        if dtype == torch.float16:
            # the same rows as gained FP16 traces: staged as half pairs
            img_g = psktm.migrate(plan, view, t(vel), scale=sc, gained=True).cpu().numpy()
            ref_g = oracle.pstm_f64(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0],
                                    off[:, 1], src, vel, dt, vel3d=True, aperture=ap,
                                    interp=oracle.LINEAR, gained=True)
            errs["float16 half pairs"] = rel_l2(img_g, ref_g)
    print("rel-L2", errs)
    assert max(errs.values()) <= 1e-4, errs


def test_split_tile_lists_agree(cuda):
    """Several workgroups per tile (partial images added in fixed order) give
    the one-workgroup image up to summation rounding, deterministically, also
    in accumulate mode and with more splits than a tile has traces."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=13, ntr=2500, nx=8, ny=6, nt=300)
    ref = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
    imgs = {}
    for splits in (1, 3, 7, 0, 1024):
        plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap, splits=splits)
        a = psktm.migrate(plan, t(tr), t(vel))
        b = psktm.migrate(plan, t(tr), t(vel))
        assert torch.equal(a, b), splits                      # deterministic
        c = a.clone()
        psktm.migrate(plan, t(tr), t(vel), image=c, accumulate=True)
        assert torch.equal(c, a + a), splits
        imgs[splits] = a.cpu().numpy()
        assert rel_l2(imgs[splits], ref) <= 1e-4, splits
    for k, v in imgs.items():
        assert rel_l2(v, imgs[1]) <= 2e-6, k
    with pytest.raises(psktm.PsktmError):
        gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap, splits=-1)


def _reach_bound(grid, mid, off, vel, dt, ap, vel3d):
    """Host restatement of the kernel's per-tile traveltime bound, over the
    whole image: 2 sqrt(tau'^2 + reach^2 s') maximised over tau and bins
    (tile maxima are at most this), in samples; plus the smallest tile value
    (over tau, for the bin of least slowness)."""
    nx, x0, dx, ny, y0, dy, nt, tau0, dtau = grid
    cx = np.array([x0, x0 + (nx - 1) * dx], np.float64)
    cy = np.array([y0, y0 + (ny - 1) * dy], np.float64)
    fx = np.max(np.abs(mid[:, 0:1].astype(np.float64) - cx[None]), 1)
    fy = np.max(np.abs(mid[:, 1:2].astype(np.float64) - cy[None]), 1)
    far = np.sqrt(fx * fx + fy * fy)
    h = np.hypot(off[:, 0].astype(np.float64), off[:, 1].astype(np.float64))
    reach = np.max((np.minimum(ap if ap > 0 else np.inf, far) + h)) * (1 + 1e-6)
    tau = (tau0 + np.arange(nt) * dtau) / dt
    v = vel.reshape(-1, nt) if vel3d else vel.reshape(1, nt)
    s = 1.0 / (v.astype(np.float64) * dt) ** 2
    xb = 2 * np.sqrt(tau[None] ** 2 + reach ** 2 * s)
    return float(xb.max()), float(xb.max(1).min())


@pytest.mark.parametrize("case", ["unclamped", "clamped"])
def test_row_padding_is_bit_identical(cuda, case):
    """Padded pair rows let a tile skip the index clamp when its traveltimes
    are bounded inside the rows; the image must not change. row_pad -1 always
    clamps. In the "unclamped" geometry every tile's bound lies inside the
    default (ns/4) and the 4000-entry rows; in the "clamped" one, one trace
    with a 10 km offset raises the plan's reach so that every tile's bound
    exceeds the default rows but lies inside the 4000-entry rows. Both code paths are compared for FP32 and FP16
    traces, LINEAR and TRUNC, 3-D and 1-D velocity, and the accumulate mode."""
    import torch
    if case == "unclamped":
        grid, mid, off, tr, vel, dt, ap = small_3d(seed=21, ntr=1500, nx=9, ny=7)
    else:
        grid, mid, off, tr, vel, dt, ap = small_3d(seed=22, ntr=1500, nx=9, ny=7)
        off[0] = (3000.0, 4000.0)
    ns = tr.shape[1]
    v1 = np.linspace(1500, 2600, grid[6]).astype(np.float32)
    hi, lo = _reach_bound(grid, mid, off, vel, dt, ap, True)
    lim_default = ((ns + ns // 4 + 17) & ~1) - 8
    print(f"case {case}: bound max {hi:.1f}, min {lo:.1f}, ns {ns}, default rows {lim_default}",
          flush=True)
    assert hi * 1.0001 + 2 < ((ns + 4000 + 17) & ~1) - 8
    if case == "unclamped":
        assert hi * 1.0001 + 2 < lim_default
    else:
        assert lo > lim_default
    q, sc = psktm.quantize(torch.as_tensor(tr).to(cuda), torch.float16)
    for interp in (psktm.INTERP_LINEAR, psktm.INTERP_TRUNC):
        outs = {}
        for pad in (-1, 0, 4000):
            plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, ap, interp=interp, row_pad=pad)
            a = psktm.migrate(plan, t(tr), t(vel))
            b = psktm.migrate(plan, q, t(v1), scale=sc)
            c = a.clone()
            psktm.migrate(plan, t(tr), t(vel), image=c, accumulate=True)
            outs[pad] = (a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy())
            plan.close()
        for pad in (0, 4000):
            for x, y in zip(outs[pad], outs[-1]):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (interp, pad)
        assert all(np.count_nonzero(x) > x.size // 2 for x in outs[-1])
        if interp == psktm.INTERP_LINEAR:
            ref = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
            assert rel_l2(outs[0][0], ref) <= 1e-4


def _gain_np(x, dt):
    """sample n times sqrtf(n dt) in binary32 (one rounding per operation)."""
    n = np.arange(x.shape[1], dtype=np.float32)
    g = np.sqrt(n * np.float32(dt)).astype(np.float32)
    return (x.astype(np.float32) * g[None]).astype(np.float32)


@pytest.mark.parametrize("m", [333, 1500, 4000])
def test_trace_gain_and_rho_gain_bit_exact(cuda, m):
    """psktm_trace_gain is in[k][n] * sqrtf(n dt) in binary32; the rho filter's
    fused gain equals the ungained filter output times the same factors, and its
    FP16 form equals quantising the FP32 gained output (radix-2 transform at
    m = 333, radix-16 at 1500 and 4000). 300 traces hold a few dozen values
    whose binary32 product y * (1/s) is an FP16 tie: rounding the product
    once from its exact value (a fused mixed-precision FMA) fails on them."""
    import torch
    rng = np.random.default_rng(41)
    x = rng.standard_normal((300, m)).astype(np.float32)
    dt = 0.004
    xt = torch.as_tensor(x).to(cuda)
    got = psktm.trace_gain(xt, dt).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), _gain_np(x, dt).view(np.uint32))
    plain = psktm.rho_filter(xt, dt * 1e6, 2).cpu().numpy()
    gained = psktm.rho_filter(xt, dt * 1e6, 2, gain=True)
    assert np.array_equal(gained.cpu().numpy().view(np.uint32),
                          _gain_np(plain, np.float32(dt * 1e6) * 1e-6).view(np.uint32)) or \
        np.array_equal(gained.cpu().numpy().view(np.uint32), _gain_np(plain, dt).view(np.uint32))
    for dtype in (torch.float16, torch.bfloat16):
        q1, s1 = psktm.rho_filter(xt, dt * 1e6, 2, out_dtype=dtype, gain=True)
        q2, s2 = psktm.quantize(gained, dtype)
        assert torch.equal(q1, q2) and torch.equal(s1, s2), dtype
        q3, s3 = psktm.trace_gain(torch.as_tensor(plain).to(cuda), dt, dtype)
        assert torch.equal(s3, s2) or float((s3 - s2).abs().max()) <= 1e-6 * float(s2.abs().max())


@pytest.mark.parametrize("interp", [psktm.INTERP_LINEAR, psktm.INTERP_TRUNC])
def test_gained_traces_vs_fp64(cuda, interp):
    """Gained traces migrated with tau alone: the GPU form against the FP64
    oracle's gained form on exactly the values the kernel reads (1e-4, as the
    per-update form), and - LINEAR - against the FP64 per-update weight on the
    raw traces, where the difference is the interpolation of the weight
    (bounded by the metric's 1e-3)."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=43, ntr=2000, nx=9, ny=7)
    ns = tr.shape[1]
    plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, ap, interp=interp)
    g = psktm.trace_gain(t(tr), dt)
    img = psktm.migrate(plan, g, t(vel), gained=True).cpu().numpy()
    oi = oracle.LINEAR if interp == psktm.INTERP_LINEAR else oracle.TRUNC
    ref_g = oracle.pstm_f64(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0], off[:, 1],
                            g.cpu().numpy(), vel, dt, vel3d=True, aperture=ap, interp=oi,
                            gained=True)
    e = rel_l2(img, ref_g)
    print(f"gained interp={interp}: rel-L2 vs FP64 gained form {e:.3e}")
    assert e <= (1e-4 if interp == psktm.INTERP_LINEAR else 2e-2)
    if interp == psktm.INTERP_LINEAR:
        ref_w = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
        ew = rel_l2(img, ref_w)
        print(f"gained vs per-update weight (FP64): {ew:.3e}")
        assert ew <= 1e-3
        # FP16 through the rho filter's fused gain: the headline form
        dw = dt * 1e6
        q, sc = psktm.rho_filter(t(tr), dw, 2, out_dtype=torch.float16, gain=True)
        img16 = psktm.migrate(plan, q, t(vel), scale=sc, gained=True).cpu().numpy()
        rho = psktm.rho_filter(t(tr), dw, 2).cpu().numpy()
        ref16 = oracle_f64(grid, mid, off, rho, vel, dt, ap, oracle.LINEAR)
        e16 = rel_l2(img16, ref16)
        print(f"FP16 gained (rho) vs FP64 per-update weight: {e16:.3e}")
        assert e16 <= 1e-3


def test_gained_flag_checked(cuda):
    """The trace form's gained flag must be 0 or 1 (precise mode takes both)."""
    import ctypes
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=44, ntr=500, nx=5, ny=5)
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap, precise=True)
    L = psktm._lib.load()
    d, v = t(tr), t(vel)
    img = psktm.migrate(plan, d, v, gained=True)
    trs = psktm._lib.PsktmTraces(psktm._lib.F32, d.data_ptr(), None, 2)
    vv = psktm._lib.PsktmVelocity(psktm._lib.VEL_3D, v.data_ptr())
    st = L.psktm_migrate(plan.handle, ctypes.byref(trs), ctypes.byref(vv), img.data_ptr(), 0, None)
    assert st == psktm._lib.E_INVAL


@pytest.mark.parametrize("axis", ["x", "y"])
def test_lines_take_1d_tiles_by_default(cuda, axis):
    """A line along x (ny = 1) or y (nx = 1) gets 8x1 or 1x8 tiles by default:
    the default plan's image is bit-identical to that explicit tile shape,
    matches the FP64 oracle, and passes the precise-mode dot-product test."""
    import torch
    nx, ny = (40, 1) if axis == "x" else (1, 40)
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=21, ntr=1500, nx=nx, ny=ny)
    vt = (8, 1) if axis == "x" else (1, 8)
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap)
    img = psktm.migrate(plan, t(tr), t(vel))
    plan_t, _ = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap, tile_x=vt[0], tile_y=vt[1])
    img_t = psktm.migrate(plan_t, t(tr), t(vel))
    assert torch.equal(img.view(torch.int32), img_t.view(torch.int32))
    ref = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
    assert rel_l2(img.cpu().numpy(), ref) <= 1e-4
    pp, _ = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap, precise=True)
    g = torch.Generator(device="cpu").manual_seed(2)
    d = torch.randn(tr.shape, generator=g).to(cuda)
    m = torch.randn(grid[0], grid[3], grid[6], generator=g).to(cuda)
    lhs = psktm.dot_f64(psktm.migrate(pp, d, t(vel)).reshape(-1), m.reshape(-1)).item()
    rhs = psktm.dot_f64(d.reshape(-1), psktm.demigrate(pp, m, t(vel)).reshape(-1)).item()
    print(f"line along {axis}: dot-product test rel={abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_longest_rows_and_rows_too_long_for_lds(cuda):
    """Traces near the LDS limit (one staged row of 16,000 samples per
    workgroup in the forward, one FP64 accumulator row in the adjoint) match
    the FP64 oracle with traveltimes reaching the end of the rows; rows longer
    than the LDS holds are refused with an error, not launched.

    Tolerance 5e-4 rather than 1e-4: an FP32 time of ~15,000 samples resolves
    only ~1e-3 of a sample (2^-24 x, summed over the two square roots), and
    the random sinusoids here reach 0.06 cycles per sample, so interpolated
    values at the end of the rows carry up to ~2 pi 0.06 1e-3 = 4e-4 relative
    error (measured 1.4e-4 over the image)."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=23, ntr=300, nx=3, ny=2, nt=200, ns=16000)
    grid = grid[:8] + (0.15,)               # tau to 30 s: times up to ~60 s of 64 s
    plan, t = gpu_plan(cuda, grid, mid, off, tr.shape[1], dt, ap)
    img = psktm.migrate(plan, t(tr), t(vel)).cpu().numpy()
    ref = oracle_f64(grid, mid, off, tr, vel, dt, ap, oracle.LINEAR)
    assert rel_l2(img, ref) <= 5e-4
    m = torch.as_tensor(ref.astype(np.float32)).to(cuda)
    back = psktm.demigrate(plan, m, t(vel))
    assert torch.isfinite(back).all() and back.abs().max() > 0
    lhs = psktm.dot_f64(psktm.migrate(plan, t(tr), t(vel)).reshape(-1), m.reshape(-1)).item()
    rhs = psktm.dot_f64(t(tr).reshape(-1), back.reshape(-1)).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    g2, mid2, off2, tr2, vel2, dt2, ap2 = small_3d(seed=24, ntr=20, nx=3, ny=2, nt=50, ns=30000)
    plan2, t2 = gpu_plan(cuda, g2, mid2, off2, tr2.shape[1], dt2, ap2)
    with pytest.raises(psktm.PsktmError):
        psktm.migrate(plan2, t2(tr2), t2(vel2))
    with pytest.raises(psktm.PsktmError):
        psktm.demigrate(plan2, torch.zeros(3, 2, 50, device=cuda), t2(vel2))


@pytest.mark.parametrize("interp", [psktm.INTERP_LINEAR, psktm.INTERP_TRUNC])
def test_half_pairs_vs_fp64_gained(cuda, interp):
    """Gained FP16 traces (staged as half pairs: the stored value and its FP16
    difference, the trace's scale applied in the accumulate) against the FP64
    gained operator on exactly the dequantised values the kernel reads. The
    only rounding beyond the FP32 form is the FP16 difference (2^-11 relative
    of one interpolation term), so LINEAR keeps the 1e-4 bar of the FP32
    form; TRUNC reads no difference and keeps its index-flip bar."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=47, ntr=2500, nx=9, ny=7)
    ns = tr.shape[1]
    plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, ap, interp=interp)
    q, sc = psktm.trace_gain(t(tr), dt, out_dtype=torch.float16)
    img = psktm.migrate(plan, q, t(vel), scale=sc, gained=True).cpu().numpy()
    deq = (q.float() * sc[:, None]).cpu().numpy()
    oi = oracle.LINEAR if interp == psktm.INTERP_LINEAR else oracle.TRUNC
    ref = oracle.pstm_f64(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], deq,
                          vel, dt, vel3d=True, aperture=ap, interp=oi, gained=True)
    e = rel_l2(img, ref)
    print(f"half pairs interp={interp}: rel-L2 vs FP64 gained on the stored values {e:.3e}")
    assert e <= (1e-4 if interp == psktm.INTERP_LINEAR else 2e-2)


def test_half_pairs_stage_sizes_and_padding_bit_identical(cuda):
    """Half pairs: every lane adds its traces in list order whatever the
    stage size, and padded rows read zeros past ns like the clamp, so 1, 3
    and 8 traces per stage, with or without padding, give the same bits
    (one workgroup per tile: the split-K count depends on the stage size)."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=48, ntr=1800, nx=9, ny=7)
    ns = tr.shape[1]
    q, sc = psktm.trace_gain(torch.as_tensor(tr).to(cuda), dt, out_dtype=torch.float16)
    imgs = []
    for stage, pad in [(0, 0), (1, 0), (3, 0), (8, 0), (8, -1), (3, 300)]:
        plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, ap, traces_per_stage=stage, row_pad=pad,
                           splits=1)
        imgs.append(psktm.migrate(plan, q, t(vel), scale=sc, gained=True))
    for im in imgs[1:]:
        assert torch.equal(im.view(torch.int32), imgs[0].view(torch.int32))
    assert imgs[0].abs().max() > 0


def test_half_pairs_packed_fill_equals_entrywise_fill(cuda):
    """The fill makes interior LINEAR chunks of 16-byte aligned rows with one
    packed FP16 FMA per entry; rows that start off a 16-byte boundary take the
    entry-by-entry path. The same FP16 values stored both ways (320-sample
    rows at a 16-byte aligned base, then the same rows one sample further
    into a buffer) give the same image bits, smooth traces and rows at the
    ends of the FP16 range alike."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=50, ntr=1500, nx=9, ny=7, ns=320)
    ntr, ns = tr.shape
    assert (ns * 2) % 16 == 0
    plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, ap, splits=1)
    q, sc = psktm.trace_gain(t(tr), dt, out_dtype=torch.float16)
    rng = np.random.default_rng(50)
    n = np.arange(ns)
    q[1::4] = torch.as_tensor(np.where(n % 2 == 0, 65504.0, -65504.0).astype(np.float16)).to(cuda)
    q[2::4] = torch.as_tensor(rng.uniform(-65504, 65504, (len(range(2, ntr, 4)), ns))
                              .astype(np.float16)).to(cuda)
    assert q.data_ptr() % 16 == 0
    buf = torch.zeros(ntr * ns + 8, dtype=torch.float16, device=cuda)
    shifted = buf[1:1 + ntr * ns].view(ntr, ns)
    shifted.copy_(q)
    assert shifted.data_ptr() % 16 == 2
    a = psktm.migrate(plan, q, t(vel), scale=sc, gained=True)
    b = psktm.migrate(plan, shifted, t(vel), scale=sc, gained=True)
    assert torch.isfinite(a).all() and a.abs().max() > 0
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))


def test_half_pairs_full_fp16_range(cuda):
    """Half pairs on samples at the ends of the FP16 range. The quantiser maps
    a trace alternating +1 / -0.9995 to +32768 / -32752, whose difference
    (65520) rounds to inf in FP16; caller rows alternating +-65504 and rows of
    full-range noise differ by up to 131008. Half pairs hold halved values and
    differences, so every image stays finite and matches the FP64 gained
    operator on the dequantised values (1e-3: the FP16 rounding of differences
    as large as the samples themselves), and the FP32-pair form on the same
    values (1e-4)."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=49, ntr=1200, nx=9, ny=7)
    ntr, ns = tr.shape
    rng = np.random.default_rng(49)
    n = np.arange(ns)
    tr2 = tr.copy()
    alt = np.where(n % 2 == 0, 1.0, -0.9995)
    tr2[::3] = (alt[None, :] * rng.uniform(0.5, 2.0, (len(tr2[::3]), 1))).astype(np.float32)
    plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, ap)
    q, sc = psktm.quantize(t(tr2), torch.float16)
    qa = q[::3].cpu().numpy().astype(np.float64)
    assert qa.max() == 32768.0 and qa.min() == -32752.0
    assert np.isinf(np.float16(qa[0, 1]) - np.float16(qa[0, 0]))
    qh = q.clone()
    qh[1::3] = torch.as_tensor(np.where(n % 2 == 0, 65504.0, -65504.0)
                               .astype(np.float16)).to(cuda)
    qh[2::3] = torch.as_tensor(rng.uniform(-65504, 65504, (len(range(2, ntr, 3)), ns))
                               .astype(np.float16)).to(cuda)
    for name, data in (("quantised", q), ("full range", qh)):
        print(f"case half pairs at the FP16 range ends: {name}", flush=True)
        img = psktm.migrate(plan, data, t(vel), scale=sc, gained=True).cpu().numpy()
        assert np.isfinite(img).all(), name
        deq = (data.float() * sc[:, None]).cpu().numpy()
        img32 = psktm.migrate(plan, torch.as_tensor(deq).to(cuda), t(vel),
                              gained=True).cpu().numpy()
        ref = oracle.pstm_f64(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0], off[:, 1],
                              deq, vel, dt, vel3d=True, aperture=ap, interp=oracle.LINEAR,
                              gained=True)
        e, e32 = rel_l2(img, ref), rel_l2(img32, ref)
        print(f"{name}: half pairs {e:.3e}, FP32 pairs {e32:.3e} vs FP64 gained")
        assert e <= 1e-3 and e32 <= 1e-4, (name, e, e32)


def test_nan_velocity_and_geometry_stay_local(cuda):
    """A NaN velocity sample and a trace with NaN coordinates: the tile
    traveltime bound fails on NaN, so those tiles take the clamped loop,
    where a NaN time reads the zero entry past the samples. The NaN velocity
    sample gets no contribution, the NaN trace contributes nothing, and every
    other image sample is bit-identical to the run without them (the clamped
    and unclamped loops give the same bits)."""
    import torch
    grid, mid, off, tr, vel, dt, ap = small_3d(seed=51, ntr=900, nx=9, ny=7)
    ns = tr.shape[1]
    for aperture in (ap, 0.0):
        print(f"case NaN inputs aperture={aperture}", flush=True)
        plan, t = gpu_plan(cuda, grid, mid, off, ns, dt, aperture)
        clean = psktm.migrate(plan, t(tr), t(vel), gained=True).cpu().numpy()
        vbad = vel.copy()
        vbad[4, 3, 150] = np.nan
        img = psktm.migrate(plan, t(tr), t(vbad), gained=True).cpu().numpy()
        assert np.isfinite(img).all()
        assert img[4, 3, 150] == 0.0
        mask = np.ones(img.shape, bool)
        mask[4, 3, 150] = False
        assert np.array_equal(img[mask].view(np.uint32), clean[mask].view(np.uint32))
        # one trace with a NaN midpoint: outside every aperture, or (no
        # aperture) in every tile, where it must add nothing
        # (one workgroup per tile: the split count follows the list lengths)
        mid2 = mid.copy()
        mid2[17, 0] = np.nan
        plan2, _ = gpu_plan(cuda, grid, mid2, off, ns, dt, aperture, splits=1)
        plan1, _ = gpu_plan(cuda, grid, mid, off, ns, dt, aperture, splits=1)
        tr2 = tr.copy()
        tr2[17] = 0.0
        img2 = psktm.migrate(plan2, t(tr), t(vel), gained=True).cpu().numpy()
        img2_ref = psktm.migrate(plan1, t(tr2), t(vel), gained=True).cpu().numpy()
        assert np.isfinite(img2).all()
        assert np.array_equal(img2.view(np.uint32), img2_ref.view(np.uint32))
