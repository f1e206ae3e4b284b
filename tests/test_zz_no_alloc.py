"""psktm_dot_f64_noalloc through the C ABI: the same bits as
psktm_dot_f64_ws on two calls in a row, the second with a smaller input, with
its partials in the library's static device array instead of a workspace."""
import ctypes

import numpy as np
import pytest

import psktm

pytestmark = pytest.mark.gpu

_p = ctypes.c_void_p


def _entry():
    L = psktm.load()
    f = L.psktm_dot_f64_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [_p, _p, ctypes.c_size_t, _p, _p]
    return f


def test_noalloc_dot_equals_dot_f64_ws(cuda):
    import torch
    f = _entry()
    st = torch.cuda.current_stream()
    sp = _p(st.cuda_stream)
    g = torch.Generator(device="cpu").manual_seed(31)
    for n in (3_000_017, 70_001):            # second call smaller than the first
        print(f"case psktm_dot_f64_noalloc n={n}", flush=True)
        a = torch.randn(n, generator=g).to(cuda)
        b = torch.randn(n, generator=g).to(cuda)
        out = torch.full((1,), float("nan"), dtype=torch.float64, device=cuda)
        assert f(_p(a.data_ptr()), _p(b.data_ptr()), n, _p(out.data_ptr()), sp) == 0, \
            psktm.last_error()
        ref = psktm.dot_f64(a, b, workspace=psktm.dot_workspace(cuda))
        torch.cuda.synchronize()
        assert out.item() == ref.item(), (n, out.item(), ref.item())
        exact = float((a.double() * b.double()).sum())
        assert abs(out.item() - exact) <= 1e-12 * n
    # empty input: result 0, no kernel reads beyond what it wrote
    out = torch.full((1,), float("nan"), dtype=torch.float64, device=cuda)
    assert f(None, None, 0, _p(out.data_ptr()), sp) == 0
    torch.cuda.synchronize()
    assert out.item() == 0.0


def test_noalloc_dot_rejects_null(cuda):
    f = _entry()
    assert f(None, None, 10, None, None) == 1          # PSKTM_E_INVAL
    assert "NULL" in psktm.last_error()
