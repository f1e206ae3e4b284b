"""Trace preparation kernels vs restatements of the MATLAB writers' steps."""
import numpy as np
import pytest

import oracle
import psktm

pytestmark = pytest.mark.gpu


def offsets(dev, n=5000, seed=1):
    import torch
    rng = np.random.default_rng(seed)
    # half offsets on a 0.5 m grid up to 400 m: magnitudes are either equal or
    # far apart compared with binary32 rounding, so orders are comparable
    hx = (rng.integers(-800, 800, n) * 0.5).astype(np.float32)
    hy = (rng.integers(-800, 800, n) * 0.5).astype(np.float32)
    hx[:50] = hx[50:100]            # exact duplicates: stability
    hy[:50] = hy[50:100]
    return hx, hy, torch.as_tensor(hx).to(dev), torch.as_tensor(hy).to(dev)


def test_offset_sort_stable(cuda):
    hx, hy, tx, ty = offsets(cuda)
    perm, key = psktm.offset_sort(tx, ty)
    ref = oracle.sorted_offset_order(hx, hy)
    assert np.array_equal(perm.cpu().numpy(), ref)
    assert np.allclose(key.cpu().numpy(), np.sqrt(hx.astype(float) ** 2 + hy.astype(float) ** 2)[ref],
                       rtol=1e-6)


def test_offset_classes(cuda):
    hx, hy, tx, ty = offsets(cuda, seed=2)
    width, nc = 62.5, 10
    perm, start = psktm.offset_classes(tx, ty, width, nc)
    full = 2 * np.sqrt(hx.astype(float) ** 2 + hy.astype(float) ** 2)
    cls = np.floor(full / width)
    st = start.cpu().numpy()
    p = perm.cpu().numpy()
    for c in range(nc):
        expect = np.nonzero(cls == c)[0]                 # trace order within a class
        assert np.array_equal(p[st[c]:st[c + 1]], expect), c
    assert st[-1] == np.count_nonzero(cls < nc) == len(p)


@pytest.mark.parametrize("h,tol", [(100.0, 10.0), (250.0, 0.5), (0.0, 3.0)])
def test_select_constant_offset(cuda, h, tol):
    hx, hy, tx, ty = offsets(cuda, seed=3)
    sel = psktm.select_offset(tx, ty, h, tol).cpu().numpy()
    assert np.array_equal(sel, oracle.const_offset_selection(hx, hy, h, tol))


def test_gather_pad(cuda):
    import torch
    rng = np.random.default_rng(4)
    x = rng.standard_normal((45, 70)).astype(np.float32)
    out = psktm.gather_pad(torch.as_tensor(x).to(cuda), 32).cpu().numpy()
    assert np.array_equal(out, oracle.pad_block(x, 32))
    rows = np.array([7, 3, 44, 0, 3], np.int32)
    out = psktm.gather_pad(torch.as_tensor(x).to(cuda), 16, torch.as_tensor(rows).to(cuda)).cpu().numpy()
    assert np.array_equal(out, oracle.pad_block(x[rows], 16))


@pytest.mark.parametrize("N", [0, 1, 3])
def test_upsample(cuda, N):
    import torch
    rng = np.random.default_rng(5)
    x = rng.standard_normal((9, 301)).astype(np.float32)
    out = psktm.upsample_traces(torch.as_tensor(x).to(cuda), N).cpu().numpy()
    ref = oracle.upsample_time(x, N)
    assert out.shape == ref.shape
    assert np.max(np.abs(out - ref)) <= 4e-7 * np.max(np.abs(x))
    assert np.array_equal(out[:, :: 2 ** N], x)          # original samples are kept exactly
