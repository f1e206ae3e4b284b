"""Parity at the workloads' stated sizes.

* 3-D PSTM, 1,000,000 traces x 1,500 samples (4 ms) -> 200 x 200 bins x
  1,500 tau, aperture 1500 m, v(x,y,t0): every storage form against the FP64
  oracle on 48 bins scattered over the central 40 x 40-bin sub-volume (bins
  80-119), the image edges and the four corners. The trace buffers exceed
  2^32 bytes (6 GB FP32), and the per-tile trace lists exceed 4 GiB, so the
  64-bit offsets of the fill and of the plan are exercised.
      FP32, per-update weight      vs its FP64 operator     rel-L2 <= 1e-4
      FP32 gained traces           vs its FP64 operator     rel-L2 <= 1e-4
      FP16 gained (the headline)   vs FP64 on its stored values <= 1e-4,
                                   vs the per-update FP64 oracle <= 1e-3
      FP16, per-update weight      vs the per-update FP64 oracle <= 1e-3
      BF16, per-update weight      vs the per-update FP64 oracle <= 8e-3
* 64 offset-class gathers at full size in one persistent launch replayed from
  a HIP graph: bit-identical to per-gather plans with one workgroup per tile,
  and two gathers against the FP64 oracle.
* Least squares on the 512-CMP line (32,768 traces x 2,000 samples): 20 CGLS
  iterations with a monotone residual; dot-product test and finite
  differences on a random 16-CMP slice.
"""
import time

import numpy as np
import pytest

import oracle
import psktm
from psktm import synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def log(msg):
    print(f"[fullsize] {msg}", flush=True)


def sample_bins(nx, ny, seed=20261205):
    """48 bins: 24 in the central 40 x 40 sub-volume, the 4 corners, the 4
    edge midpoints, 8 more on the edges and 8 anywhere."""
    rng = np.random.default_rng(seed)
    bins = [(int(i), int(j)) for i, j in rng.integers(80, 120, (24, 2))]
    bins += [(0, 0), (0, ny - 1), (nx - 1, 0), (nx - 1, ny - 1)]
    bins += [(0, ny // 2), (nx - 1, ny // 2), (nx // 2, 0), (nx // 2, ny - 1)]
    for k in range(8):
        t = int(rng.integers(1, nx - 1))
        bins.append([(0, t), (nx - 1, t), (t, 0), (t, ny - 1)][k % 4])
    bins += [(int(i), int(j)) for i, j in zip(rng.integers(0, nx, 8), rng.integers(0, ny, 8))]
    return bins


@pytest.mark.timeout(900)
def test_pstm3d_1m_traces_every_form_vs_fp64(cuda):
    import torch
    t0 = time.time()
    s = synth.diffractor_3d(1_000_000, device=cuda)
    grid = psktm.Grid(*s.grid)
    ns, dt = s.traces.shape[1], s.dt
    dw = dt * 1e6
    survey = psktm.Survey(s.mid_x, s.mid_y, s.off_x, s.off_y, ns, dt)
    plan = psktm.Plan(grid, survey, aperture=s.aperture)
    assert plan.nbytes() > 4 * 2**30            # per-tile lists past 4 GiB
    log(f"survey + plan in {time.time() - t0:.1f} s; {plan.updates():.4e} updates")
    bins = sample_bins(grid.nx, grid.ny)
    bi = torch.as_tensor([b[0] for b in bins], device=cuda)
    bj = torch.as_tensor([b[1] for b in bins], device=cuda)
    cols = lambda img: img[bi, bj].double().cpu().numpy()

    rho = psktm.rho_filter(s.traces, dw, 2)                     # FP32 filtered
    assert rho.numel() * 4 > 2**32
    imgs = {}
    imgs["f32"] = cols(psktm.migrate(plan, rho, s.vel))
    g32 = psktm.trace_gain(rho, dt)
    imgs["f32g"] = cols(psktm.migrate(plan, g32, s.vel, gained=True))
    q16, s16 = psktm.quantize(rho, torch.float16)
    imgs["f16"] = cols(psktm.migrate(plan, q16, s.vel, scale=s16))
    del q16, s16
    qb, sb = psktm.quantize(rho, torch.bfloat16)
    imgs["bf16"] = cols(psktm.migrate(plan, qb, s.vel, scale=sb))
    del qb, sb
    qg, sg = psktm.rho_filter(s.traces, dw, 2, out_dtype=torch.float16, gain=True)
    imgs["f16g"] = cols(psktm.migrate(plan, qg, s.vel, scale=sg, gained=True))
    torch.cuda.synchronize()
    log(f"five migrations in {time.time() - t0:.1f} s")

    host = lambda t: t.cpu().numpy()
    mx, my, hx, hy, vel = (host(a) for a in (s.mid_x, s.mid_y, s.off_x, s.off_y, s.vel))
    og = oracle.Grid(*s.grid)
    kw = dict(vel3d=True, aperture=s.aperture, interp=oracle.LINEAR)
    refs = {}
    refs["per_update"] = oracle.pstm_f64_bins(og, mx, my, hx, hy, host(rho), vel, dt, bins, **kw)
    log(f"FP64 per-update oracle on {len(bins)} bins at {time.time() - t0:.1f} s")
    del rho
    refs["f32g"] = oracle.pstm_f64_bins(og, mx, my, hx, hy, host(g32), vel, dt, bins,
                                        gained=True, **kw)
    del g32
    deq = (qg.float() * sg[:, None]).cpu().numpy()
    refs["f16g"] = oracle.pstm_f64_bins(og, mx, my, hx, hy, deq, vel, dt, bins, gained=True, **kw)
    del deq
    log(f"FP64 gained oracles at {time.time() - t0:.1f} s")

    for r in refs.values():
        assert np.count_nonzero(r) > 0.5 * r.size
    errs = {
        "f32 vs own": rel_l2(imgs["f32"], refs["per_update"]),
        "f32g vs own": rel_l2(imgs["f32g"], refs["f32g"]),
        "f16g vs stored": rel_l2(imgs["f16g"], refs["f16g"]),
        "f16g vs per-update": rel_l2(imgs["f16g"], refs["per_update"]),
        "f16 vs per-update": rel_l2(imgs["f16"], refs["per_update"]),
        "bf16 vs per-update": rel_l2(imgs["bf16"], refs["per_update"]),
    }
    mref = float(np.abs(refs["per_update"]).max())
    for k in ("f32", "f16", "bf16", "f16g"):
        log(f"{k}: max-abs error / max-ref {np.abs(imgs[k] - refs['per_update']).max() / mref:.3e}")
    # the edge and corner bins on their own (fewer traces, clipped apertures)
    edge = slice(24, 48)
    errs["f16g edges vs per-update"] = rel_l2(imgs["f16g"][edge], refs["per_update"][edge])
    errs["f32 edges vs own"] = rel_l2(imgs["f32"][edge], refs["per_update"][edge])
    log("rel-L2 " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    bars = {"f32 vs own": 1e-4, "f32g vs own": 1e-4, "f16g vs stored": 1e-4,
            "f16g vs per-update": 1e-3, "f16 vs per-update": 1e-3, "bf16 vs per-update": 8e-3,
            "f16g edges vs per-update": 1e-3, "f32 edges vs own": 1e-4}
    bad = {k: v for k, v in errs.items() if not v <= bars[k]}
    assert not bad, bad


@pytest.mark.timeout(900)
def test_batched_64_full_gathers(cuda, tmp_path):
    import torch
    from psktm.tune import CANDIDATES_3D, Autotuner, signature, time_gpu
    t0 = time.time()
    gd = synth.offset_class_gathers(64, device=cuda)
    ns = gd.traces.shape[1]
    assert len(gd.counts) == 64 and min(gd.counts) >= 2000 and max(gd.counts) <= 60000
    q, sc = psktm.rho_filter(gd.traces, gd.dt * 1e6, 2, out_dtype=torch.float16, gain=True)
    survey = psktm.Survey(gd.mid_x, gd.mid_y, gd.off_x, gd.off_y, ns, gd.dt)
    gathers = [(psktm.Grid(*g), f, c) for g, f, c in zip(gd.grids, gd.first, gd.counts)]
    log(f"64 gathers, {sum(gd.counts)} traces in {time.time() - t0:.1f} s")
    # tile shape from the autotuner, timed on every fourth gather
    tuner = Autotuner(str(tmp_path / "tune.json"))
    sub = gathers[::4]

    def trial(c):
        b = psktm.Batch(sub, survey, aperture=gd.aperture, tile_x=c[0], tile_y=c[1],
                        traces_per_stage=c[2])
        t = time_gpu(lambda: b.migrate(q, gd.vel, scale=sc, gained=True), repeats=1)
        b.close()
        return t

    choice = tuner.choose(signature("batch", ns=ns, gathers=64), CANDIDATES_3D, trial)
    log(f"autotuner choice {choice} at {time.time() - t0:.1f} s")
    batch = psktm.Batch(gathers, survey, aperture=gd.aperture, tile_x=choice[0],
                        tile_y=choice[1], traces_per_stage=choice[2])
    images = torch.empty(int(sum(batch.sizes)), dtype=torch.float32, device=cuda)
    batch.migrate(q, gd.vel, images, scale=sc, use_graph=True, gained=True)
    first = images.clone()
    batch.migrate(q, gd.vel, images, scale=sc, use_graph=True, gained=True)   # replay
    torch.cuda.synchronize()
    assert torch.equal(images, first)
    total = 0.0
    for g, (grid, f, c) in enumerate(gathers):
        sl = slice(f, f + c)
        s = psktm.Survey(gd.mid_x[sl].contiguous(), gd.mid_y[sl].contiguous(),
                         gd.off_x[sl].contiguous(), gd.off_y[sl].contiguous(), ns, gd.dt)
        plan = psktm.Plan(grid, s, aperture=gd.aperture, tile_x=choice[0], tile_y=choice[1],
                          traces_per_stage=choice[2], splits=1)
        total += plan.updates()
        one = psktm.migrate(plan, q[sl], gd.vel, scale=sc[sl], gained=True)
        assert torch.equal(batch.image(images, g), one), g
        plan.close()
    assert total == batch.updates()
    log(f"64 gathers bit-identical to per-gather plans at {time.time() - t0:.1f} s")
    # the smallest and the largest gather against the FP64 oracle
    rho = None
    for g in (int(np.argmin(gd.counts)), int(np.argmax(gd.counts))):
        grid, f, c = gathers[g]
        sl = slice(f, f + c)
        gg = gd.grids[g]
        bins = [(0, 0), (gg[0] - 1, gg[3] - 1), (gg[0] // 2, gg[3] // 2), (0, gg[3] - 1),
                (gg[0] - 1, 0), (gg[0] // 3, 2 * gg[3] // 3)]
        img = batch.image(images, g).cpu().numpy()
        got = np.stack([img[i, j] for i, j in bins]).astype(np.float64)
        host = lambda t: t[sl].cpu().numpy()
        og = oracle.Grid(*gg)
        deq = (q[sl].float() * sc[sl][:, None]).cpu().numpy()
        kw = dict(aperture=gd.aperture, interp=oracle.LINEAR)
        ref_s = oracle.pstm_f64_bins(og, host(gd.mid_x), host(gd.mid_y), host(gd.off_x),
                                     host(gd.off_y), deq, gd.vel.cpu().numpy(), gd.dt, bins,
                                     gained=True, **kw)
        rho = psktm.rho_filter(gd.traces[sl].contiguous(), gd.dt * 1e6, 2).cpu().numpy()
        ref_w = oracle.pstm_f64_bins(og, host(gd.mid_x), host(gd.mid_y), host(gd.off_x),
                                     host(gd.off_y), rho, gd.vel.cpu().numpy(), gd.dt, bins, **kw)
        e_s, e_w = rel_l2(got, ref_s), rel_l2(got, ref_w)
        log(f"gather {g} ({c} traces, {gg[0]}x{gg[3]} bins): rel-L2 {e_s:.2e} vs FP64 on the "
            f"stored values, {e_w:.2e} vs the per-update FP64 oracle")
        assert np.count_nonzero(ref_s) > 0.5 * ref_s.size
        assert e_s <= 1e-4 and e_w <= 1e-3, (g, e_s, e_w)


@pytest.mark.timeout(900)
def test_lsm_512_cmp_line(cuda):
    import torch
    from psktm import lsm
    t0 = time.time()
    d = synth.line_2d(seed=synth.SEED0 + 4, n_cmp=512, n_off=64, ns=2000, nt=2000, dt=0.002)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(cuda)
    assert d.traces.shape == (32768, 2000)
    grid = psktm.Grid(*d.grid)
    tau = np.arange(grid.nt, dtype=np.float32) * np.float32(d.grid[8])
    vel = t(1500.0 + 500.0 * 2.0 * tau.astype(np.float64))
    survey = psktm.Survey(t(d.mid_x), t(d.mid_y), t(d.off_x), t(d.off_y), 2000, 0.002)
    plan = psktm.Plan(grid, survey, precise=True)
    rng = np.random.default_rng(synth.SEED0 + 4)
    m = np.zeros(grid.shape, np.float32)
    for _ in range(8):                          # 8 random-dip interfaces
        l0, slope, a = rng.integers(100, grid.nt - 100), rng.uniform(-1.0, 1.0), rng.choice([-1, 1])
        for i in range(grid.nx):
            l = int(round(l0 + slope * (i - grid.nx / 2)))
            if 0 < l < grid.nt:
                m[i, 0, l] += a
    for _ in range(50):                         # 50 point scatterers
        m[rng.integers(0, grid.nx), 0, rng.integers(50, grid.nt - 50)] += rng.choice([-2.0, 2.0])
    m_true = t(m)
    data = psktm.demigrate(plan, m_true, vel)
    mcg, hist = lsm.cgls(plan, vel, data, iters=20)
    log(f"20 CGLS iterations at {time.time() - t0:.1f} s; ||r||^2 {hist[0]:.4e} -> {hist[-1]:.4e}")
    assert len(hist) == 21 and all(np.isfinite(hist))
    assert all(hist[k + 1] <= hist[k] * (1 + 1e-9) for k in range(20)), hist
    assert hist[-1] < 0.5 * hist[0]
    cos = float((mcg * m_true).sum() / (mcg.norm() * m_true.norm()))
    mig = psktm.migrate(plan, data, vel)
    cos_mig = float((mig * m_true).sum() / (mig.norm() * m_true.norm()))
    log(f"cosine to the model: CGLS {cos:.3f}, migration {cos_mig:.3f}")
    assert cos > cos_mig
    # dot-product test and finite differences on a random 16-CMP slice
    c0 = int(np.random.default_rng(5).integers(0, 512 - 16))
    sd = synth.line_2d(seed=synth.SEED0 + 4, n_cmp=16, n_off=64, ns=2000, nt=2000, dt=0.002)
    shift = np.float32(12.5 * c0)
    sgrid = psktm.Grid(16, float(shift), 12.5, 1, 0.0, 12.5, 2000, 0.0, float(d.grid[8]))
    ss = psktm.Survey(t(sd.mid_x + shift), t(sd.mid_y), t(sd.off_x), t(sd.off_y), 2000, 0.002)
    for precise in (True, False):
        sp = psktm.Plan(sgrid, ss, precise=precise)
        g = torch.Generator(device="cpu").manual_seed(7)
        x = torch.randn((ss.n_traces, 2000), generator=g).to(cuda)
        y = torch.randn(sgrid.shape, generator=g).to(cuda)
        ws = psktm.dot_workspace(cuda)
        lhs = psktm.dot_f64(psktm.migrate(sp, x, vel).reshape(-1), y.reshape(-1), workspace=ws).item()
        rhs = psktm.dot_f64(x.reshape(-1), psktm.demigrate(sp, y, vel).reshape(-1),
                            workspace=ws).item()
        rel = abs(lhs - rhs) / abs(lhs)
        log(f"16-CMP slice from CMP {c0}, precise={precise}: dot-product test {rel:.2e}")
        assert rel <= (1e-6 if precise else 5e-6), (precise, rel)
        if precise:
            mm = y.clone().requires_grad_(True)
            dd = torch.randn((ss.n_traces, 2000), generator=g).to(cuda)
            e = torch.randn(sgrid.shape, generator=g).to(cuda)
            f = lambda v: 0.5 * ((lsm.Demigrate.apply(v, sp, vel) - dd).double() ** 2).sum()
            L = f(mm)
            L.backward()
            gdot = float((mm.grad.double() * e.double()).sum())
            with torch.no_grad():
                fd = float(f(mm + 0.5 * e) - f(mm - 0.5 * e))
            log(f"finite differences {fd:.6e} vs gradient {gdot:.6e}")
            assert abs(fd - gdot) <= 1e-5 * abs(gdot)
