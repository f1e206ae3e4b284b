"""Batched offset-class gathers: one persistent launch, HIP-graph replay."""
import numpy as np
import pytest

import oracle
import psktm
from psktm import synth
from psktm.tune import CANDIDATES_3D, Autotuner, signature, time_gpu

from test_gpu_migrate import rel_l2

pytestmark = pytest.mark.gpu


def small_gathers(dev, n=12, scale=0.02, seed=31):
    return synth.offset_class_gathers(n, count_scale=scale, side_min=8, side_max=24,
                                      nt=400, ns=420, seed=seed, device=dev)


def make_batch(gd, dev, **kw):
    s = psktm.Survey(gd.mid_x, gd.mid_y, gd.off_x, gd.off_y, gd.traces.shape[1], gd.dt)
    gathers = [(psktm.Grid(*g), f, c) for g, f, c in zip(gd.grids, gd.first, gd.counts)]
    return psktm.Batch(gathers, s, aperture=gd.aperture, **kw), s


def test_batch_equals_per_gather_plans(cuda):
    import torch
    gd = small_gathers(cuda)
    q, sc = psktm.quantize(gd.traces, torch.float16)
    batch, _ = make_batch(gd, cuda)
    imgs = batch.migrate(q, gd.vel, scale=sc)
    torch.cuda.synchronize()
    assert batch.updates() > 0
    total = 0.0
    for g, (grid, f, c) in enumerate(zip(gd.grids, gd.first, gd.counts)):
        sl = slice(f, f + c)
        s = psktm.Survey(gd.mid_x[sl].contiguous(), gd.mid_y[sl].contiguous(),
                         gd.off_x[sl].contiguous(), gd.off_y[sl].contiguous(),
                         gd.traces.shape[1], gd.dt)
        # one workgroup per tile: the batch kernel's summation order
        plan = psktm.Plan(psktm.Grid(*grid), s, aperture=gd.aperture, splits=1)
        total += plan.updates()
        one = psktm.migrate(plan, q[sl].contiguous(), gd.vel, scale=sc[sl].contiguous())
        assert torch.equal(batch.image(imgs, g), one), g
    assert total == batch.updates()


def test_batch_row_padding_is_bit_identical(cuda):
    """The persistent kernel re-clears its pair table for every work item
    (gathers follow one another in a workgroup), so padded rows read zeros
    past each gather's samples: images equal the always-clamped form."""
    import torch
    gd = small_gathers(cuda, n=8, seed=33)
    q, sc = psktm.quantize(gd.traces, torch.float16)
    outs = []
    for pad in (-1, 0, 2000):
        batch, _ = make_batch(gd, cuda, row_pad=pad)
        outs.append(batch.migrate(q, gd.vel, scale=sc).clone())
        torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_batch_gained_equals_per_gather_plans(cuda):
    """Gained traces in the persistent launch (and its graph replay) equal
    one-workgroup-per-tile plans migrating the same gained traces."""
    import torch
    gd = small_gathers(cuda, n=6, seed=34)
    q, sc = psktm.rho_filter(gd.traces, gd.dt * 1e6, 2, out_dtype=torch.float16, gain=True)
    batch, _ = make_batch(gd, cuda)
    imgs = batch.migrate(q, gd.vel, scale=sc, gained=True).clone()
    rep = batch.migrate(q, gd.vel, scale=sc, gained=True, use_graph=True)
    rep = batch.migrate(q, gd.vel, scale=sc, gained=True, use_graph=True).clone()
    raw = batch.migrate(q, gd.vel, scale=sc).clone()
    torch.cuda.synchronize()
    assert torch.equal(imgs, rep)
    assert not torch.equal(imgs, raw)
    for g, (grid, f, c) in enumerate(zip(gd.grids, gd.first, gd.counts)):
        sl = slice(f, f + c)
        s = psktm.Survey(gd.mid_x[sl].contiguous(), gd.mid_y[sl].contiguous(),
                         gd.off_x[sl].contiguous(), gd.off_y[sl].contiguous(),
                         gd.traces.shape[1], gd.dt)
        plan = psktm.Plan(psktm.Grid(*grid), s, aperture=gd.aperture, splits=1)
        one = psktm.migrate(plan, q[sl].contiguous(), gd.vel, scale=sc[sl].contiguous(),
                            gained=True)
        assert torch.equal(batch.image(imgs, g), one), g


def test_batch_vs_fp64_oracle(cuda):
    gd = small_gathers(cuda, n=3, scale=0.05, seed=32)
    batch, _ = make_batch(gd, cuda)
    imgs = batch.migrate(gd.traces, gd.vel)
    for g, (grid, f, c) in enumerate(zip(gd.grids, gd.first, gd.counts)):
        sl = slice(f, f + c)
        h = lambda t: t[sl].cpu().numpy()
        ref = oracle.pstm_f64(oracle.Grid(*grid), h(gd.mid_x), h(gd.mid_y), h(gd.off_x),
                              h(gd.off_y), h(gd.traces), gd.vel.cpu().numpy(), gd.dt,
                              aperture=gd.aperture, interp=oracle.LINEAR)
        assert rel_l2(batch.image(imgs, g).cpu().numpy(), ref) <= 1e-4


def test_graph_replay_matches_direct(cuda):
    """Captured once, replayed several times with eager work in between; every
    replay equals the direct launch bit for bit (deterministic kernels)."""
    import torch
    gd = small_gathers(cuda, n=10, seed=33)
    q, sc = psktm.quantize(gd.traces, torch.float16)
    batch, _ = make_batch(gd, cuda)
    direct = batch.migrate(q, gd.vel, scale=sc).clone()
    out = torch.empty_like(direct)
    for rep in range(4):
        out.fill_(-1.0)
        batch.migrate(q, gd.vel, out, scale=sc, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(out, direct), rep
        _ = psktm.quantize(gd.traces, torch.float16)        # eager work between replays
        torch.cuda.synchronize()


def test_autotuner_on_batch(cuda, tmp_path):
    import torch
    gd = small_gathers(cuda, n=6, seed=34)
    q, sc = psktm.quantize(gd.traces, torch.float16)
    key = signature("batch", ns=gd.traces.shape[1], nt=gd.grids[0][6], dtype="f16",
                    aperture=gd.aperture, gathers=len(gd.counts))
    results = {}

    def run(c):
        b, _ = make_batch(gd, cuda, tile_x=c[0], tile_y=c[1], traces_per_stage=c[2])
        t = time_gpu(lambda: b.migrate(q, gd.vel, scale=sc), repeats=1)
        results[c] = b.migrate(q, gd.vel, scale=sc).clone()
        return t

    tuner = Autotuner(str(tmp_path / "tune.json"))
    choice = tuner.choose(key, CANDIDATES_3D[:4], run)
    assert choice in CANDIDATES_3D[:4]
    assert Autotuner(str(tmp_path / "tune.json")).lookup(key) == choice
    # every candidate computes the same images up to summation order
    ref = results[(4, 4, 4)]
    for c, img in results.items():
        assert torch.allclose(img, ref, rtol=1e-4, atol=1e-4 * float(ref.abs().max())), c
