"""Reference-exact HIP migration vs the oracle: bit-identical images.

The HIP form evaluates every update with the reference's binary32 operation
order (src/psktmCPU.c:78-85) and sums traces in k order per image sample, so
it must equal the single-thread restatement bit for bit.
"""
import numpy as np
import pytest

import oracle
import psktm
from psktm import synth

pytestmark = pytest.mark.gpu


def _to(dev, *arrs):
    import torch
    return [torch.as_tensor(np.ascontiguousarray(a)).to(dev) for a in arrs]


def _exact_gpu(dev, grid, mid_x, mid_y, off_x, off_y, traces, vel, dt):
    import torch
    g = psktm.Grid(*grid)
    mx, my, hx, hy, tr, v = _to(dev, mid_x, mid_y, off_x, off_y, traces, vel)
    s = psktm.Survey(mx, my, hx, hy, traces.shape[1], dt)
    img = psktm.migrate_exact(g, s, tr, v)
    torch.cuda.synchronize()
    return img.cpu().numpy()


def test_line_2d_bit_exact(cuda):
    d = synth.line_2d()
    print(f"case line_2d traces={d.traces.shape} grid={d.grid}", flush=True)
    gpu = _exact_gpu(cuda, d.grid, d.mid_x, d.mid_y, d.off_x, d.off_y, d.traces, d.vel, d.dt)
    ref = oracle.pstm_f32(oracle.Grid(*d.grid), d.mid_x, d.mid_y, d.off_x, d.off_y, d.traces,
                          d.vel, d.dt)
    assert np.count_nonzero(ref) > 0.5 * ref.size
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("entry", ["ktm_migration_gpu", "ktm_migration_cpu"])
def test_legacy_ktm_migration_bit_exact(cuda, entry):
    """Both reference entry points (src/psktmCUDA.cu:211-221,
    src/psktmCPU.c:13-23) through the C ABI on host arrays: the single-thread
    reference loop's image bit for bit."""
    d = synth.line_2d(n_cmp=24, n_off=12, ns=600, nt=500)
    nx, x0, dx, ny, y0, dy, nt, t0, dtau = d.grid
    X, Y, Z = psktm.ImageGrid(nx, x0, dx), psktm.ImageGrid(ny, y0, dy), psktm.ImageGrid(nt, t0, dtau)
    J = psktm.JobParams(d.traces.shape[0], d.traces.shape[1], d.dt)
    img = getattr(psktm, entry)(X, Y, Z, J, d.mid_x, d.mid_y, d.off_x, d.off_y, d.traces, d.vel)
    assert psktm.last_error() == ""
    ref = oracle.pstm_f32(oracle.Grid(*d.grid), d.mid_x, d.mid_y, d.off_x, d.off_y, d.traces,
                          d.vel, d.dt).reshape(-1)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("seed,ntr,nx,ny,nt,ns", [
    (0, 37, 5, 3, 300, 260),        # ragged: bins not a multiple of 8, nt not of 256
    (1, 1, 1, 1, 1, 2),             # single trace, single bin, single sample
    (2, 70, 9, 7, 513, 400),        # 3-D, nonzero offY, split form: 2 stages + a partial one
    (3, 100, 64, 64, 128, 300),     # direct form with many stages (512 workgroups)
    (4, 600, 30, 30, 1000, 1200),   # split form in three passes (288 + 288 + 24 traces)
])
def test_random_3d_bit_exact(cuda, seed, ntr, nx, ny, nt, ns):
    """Images under a few hundred workgroups of the direct form take the
    split form (contributions in parallel, then summed per sample in k
    order); both must equal the single-thread loop bit for bit."""
    rng = np.random.default_rng(seed)
    grid = (nx, -30.0, 17.5, ny, 12.0, 22.5, nt, 0.0, 0.003)
    mid = rng.uniform(-100, 300, (ntr, 2)).astype(np.float32)
    off = rng.uniform(-400, 400, (ntr, 2)).astype(np.float32)
    tr = rng.standard_normal((ntr, ns)).astype(np.float32)
    vel = rng.uniform(1400, 3500, nt).astype(np.float32)
    print(f"case random seed={seed} ntr={ntr} grid={grid}", flush=True)
    gpu = _exact_gpu(cuda, grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, 0.004)
    ref = oracle.pstm_f32(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, 0.004)
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


def test_zero_traces_gives_zero_image(cuda):
    import torch
    grid = psktm.Grid(3, 0, 10, 2, 0, 10, 40, 0, 0.004)
    e = torch.zeros(0, dtype=torch.float32, device=cuda)
    s = psktm.Survey(e, e, e, e, 50, 0.004)
    img = torch.full((3, 2, 40), 5.0, device=cuda)
    psktm.migrate_exact(grid, s, torch.zeros(0, device=cuda), torch.full((40,), 2000.0, device=cuda),
                        image=img)
    torch.cuda.synchronize()
    assert not torch.any(img)


@pytest.mark.parametrize("ntr", [8, 96])      # the direct form, then the split form
def test_degenerate_velocity_matches_reference(cuda, ntr):
    """v = 0 makes 1/v^2 infinite: inf and NaN traveltimes are skipped by both
    (the split form adds +0.0 for them)."""
    rng = np.random.default_rng(9)
    grid = (4, 0.0, 20.0, 2, 0.0, 20.0, 64, 0.0, 0.004)
    mid = rng.uniform(0, 60, (ntr, 2)).astype(np.float32)
    mid[0] = [20.0, 0.0]                      # a trace exactly on a bin: A = 0 -> 0*inf
    off = np.zeros((ntr, 2), np.float32)
    tr = rng.standard_normal((ntr, 100)).astype(np.float32)
    vel = rng.uniform(1500, 2500, 64).astype(np.float32)
    vel[10:14] = 0.0
    gpu = _exact_gpu(cuda, grid, mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, 0.004)
    ref = oracle.pstm_f32(oracle.Grid(*grid), mid[:, 0], mid[:, 1], off[:, 0], off[:, 1], tr, vel, 0.004)
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


def test_invalid_arguments_raise(cuda):
    import torch
    grid = psktm.Grid(0, 0, 10, 2, 0, 10, 40, 0, 0.004)
    e = torch.zeros(4, dtype=torch.float32, device=cuda)
    s = psktm.Survey(e, e, e, e, 50, 0.004)
    with pytest.raises(psktm.PsktmError):
        psktm.migrate_exact(grid, s, torch.zeros(200, device=cuda), torch.ones(40, device=cuda),
                            image=torch.zeros(0, device=cuda))
