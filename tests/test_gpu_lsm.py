"""Least-squares PSTM on the GPU: autograd wrappers, finite differences, CGLS.

The operator pair is A = demigration (image -> traces) and A' = migration on a
2-D line like the least-squares workload (full line, no aperture), reduced so
that each test runs in seconds.
"""
import numpy as np
import pytest

import psktm
from psktm import lsm, synth

pytestmark = pytest.mark.gpu


def line_problem(dev, n_cmp=48, n_off=32, ns=600, nt=600, seed=7):
    import torch
    d = synth.line_2d(seed=seed, n_cmp=n_cmp, n_off=n_off, ns=ns, nt=nt, dt=0.002)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)
    s = psktm.Survey(t(d.mid_x), t(d.mid_y), t(d.off_x), t(d.off_y), ns, d.dt)
    plan = psktm.Plan(psktm.Grid(*d.grid), s, precise=True)
    return plan, t(d.vel), d


def test_autograd_wrappers_and_finite_differences(cuda):
    """For the quadratic loss 0.5||A m - d||^2 the central difference along a
    direction e equals <grad, e> up to rounding (no truncation error)."""
    import torch
    plan, vel, _ = line_problem(cuda)
    g = torch.Generator(device="cpu").manual_seed(5)
    shape = plan.grid.shape
    m = torch.randn(shape, generator=g).to(cuda).requires_grad_(True)
    d = torch.randn((plan.survey.n_traces, plan.survey.n_samples), generator=g).to(cuda)
    e = torch.randn(shape, generator=g).to(cuda)

    def loss(mm):
        r = lsm.Demigrate.apply(mm, plan, vel) - d
        return 0.5 * (r.double() ** 2).sum()

    L = loss(m)
    L.backward()
    gdot = float((m.grad.double() * e.double()).sum())
    h = 0.5
    with torch.no_grad():
        fd = float((loss(m + h * e) - loss(m - h * e)) / (2 * h))
    print(f"FD {fd:.9e}  <grad,e> {gdot:.9e}  rel {abs(fd - gdot) / abs(gdot):.2e}")
    assert abs(fd - gdot) <= 1e-4 * abs(gdot)
    # the migration wrapper's gradient is demigration
    tr = torch.randn(d.shape, generator=g).to(cuda).requires_grad_(True)
    img = lsm.Migrate.apply(tr, plan, vel)
    w = torch.randn(shape, generator=g).to(cuda)
    (img * w).sum().backward()
    ref = psktm.demigrate(plan, w, vel)
    assert torch.allclose(tr.grad, ref, rtol=0, atol=0)


def test_cgls_residual_monotone_and_recovers_model(cuda):
    import torch
    plan, vel, _ = line_problem(cuda, n_cmp=64, n_off=32, ns=500, nt=500, seed=9)
    shape = plan.grid.shape
    m_true = torch.zeros(shape, device=cuda)
    rng = np.random.default_rng(3)
    for _ in range(6):                      # dipping interfaces
        l0, slope = rng.integers(80, 420), rng.uniform(-1.5, 1.5)
        for i in range(shape[0]):
            l = int(round(l0 + slope * i))
            if 0 < l < shape[2]:
                m_true[i, 0, l] = float(rng.choice([-1.0, 1.0]))
    data = psktm.demigrate(plan, m_true, vel)
    m, hist = lsm.cgls(plan, vel, data, iters=20)
    print("||r||^2:", " ".join(f"{h:.3e}" for h in hist))
    assert all(hist[k + 1] <= hist[k] * (1 + 1e-6) for k in range(len(hist) - 1))
    assert hist[-1] < 0.05 * hist[0]
    mig = psktm.migrate(plan, data, vel)
    # least squares sharpens the image: closer to the model than plain migration
    cos = lambda a, b: float((a * b).sum() / (a.norm() * b.norm()))
    assert cos(m, m_true) > cos(mig, m_true)


def test_cgls_gained_operator(cuda):
    """CGLS on the gained operator (trace gain + gained migration, and
    psktm_demigrate_gained): monotone residual, and a sharper image than
    plain migration of the same data."""
    import torch
    plan, vel, _ = line_problem(cuda, n_cmp=64, n_off=32, ns=500, nt=500, seed=9)
    shape = plan.grid.shape
    m_true = torch.zeros(shape, device=cuda)
    rng = np.random.default_rng(4)
    for _ in range(6):
        l0, slope = rng.integers(80, 420), rng.uniform(-1.5, 1.5)
        for i in range(shape[0]):
            l = int(round(l0 + slope * i))
            if 0 < l < shape[2]:
                m_true[i, 0, l] = float(rng.choice([-1.0, 1.0]))
    data = psktm.demigrate(plan, m_true, vel, gained=True)
    m, hist = lsm.cgls(plan, vel, data, iters=20, gained=True)
    print("gained ||r||^2:", " ".join(f"{h:.3e}" for h in hist))
    assert all(hist[k + 1] <= hist[k] * (1 + 1e-6) for k in range(len(hist) - 1))
    assert hist[-1] < 0.05 * hist[0]
    mig = psktm.migrate(plan, psktm.trace_gain(data, plan.survey.dt), vel, gained=True)
    cos = lambda a, b: float((a * b).sum() / (a.norm() * b.norm()))
    assert cos(m, m_true) > cos(mig, m_true)


def test_cgls_graph_replay_matches_eager(cuda):
    """Two CG iterations captured in one HIP graph, replayed several times with
    eager work in between; every replay must equal the eager iterations."""
    import torch
    plan, vel, _ = line_problem(cuda, n_cmp=32, n_off=16, ns=400, nt=400, seed=4)
    g = torch.Generator(device="cpu").manual_seed(2)
    data = torch.randn((plan.survey.n_traces, plan.survey.n_samples), generator=g).to(cuda)
    eager = lsm.CGLS(plan, vel, data)
    graphed = lsm.CGLS(plan, vel, data)
    # warm-up outside capture (allocates the plan's tables once)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        tmp = lsm.CGLS(plan, vel, data)
        tmp.step()
        tmp.step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        graphed.step()
        graphed.step()
    for rep in range(3):
        eager.step()
        eager.step()
        torch.cuda.synchronize()
        _ = psktm.migrate(plan, data, vel)            # unrelated eager work
        graph.replay()
        torch.cuda.synchronize()
        # atomics in the FP64 dot product make the last bits order-dependent
        assert torch.allclose(graphed.m, eager.m, rtol=1e-5, atol=1e-6 * float(eager.m.abs().max())), rep
        assert abs(float(graphed.rr) - float(eager.rr)) <= 1e-5 * float(eager.rr), rep
