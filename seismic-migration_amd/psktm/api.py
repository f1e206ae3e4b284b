"""Host-side API over libpsktm.so.

Two families, as in the C header:

* reference-named entry points on host numpy arrays: :func:`ktm_migration_gpu`
  (src/psktmCUDA.cu:211-221), :func:`init_three_axes` / :func:`init_two_axes`
  (src/psktmFunctions.c:1-39), :func:`ibm32_to_ieee` (src/ioFunctions.c:4-40);
* extended entry points on device tensors (HBM-resident, stream-ordered):
  :class:`Plan`, :func:`migrate`, :func:`demigrate`, :func:`migrate_exact`,
  :func:`rho_filter`, :func:`quantize`, :func:`dot_f64`.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (F32, INTERP_LINEAR, MATH_FAST, VEL_1D, VEL_3D, ImageGrid, JobParams,
                   PsktmGather, PsktmGrid, PsktmOptions, PsktmSurvey, PsktmTraces,
                   PsktmVelocity, check)

_p = ctypes.c_void_p


def _torch():
    import torch
    return torch


@dataclass(frozen=True)
class Grid:
    """Output image axes (x slowest, tau fastest; tau is one-way time)."""
    nx: int
    x0: float
    dx: float
    ny: int
    y0: float
    dy: float
    nt: int
    tau0: float
    dtau: float

    def c(self) -> PsktmGrid:
        return PsktmGrid(self.nx, self.ny, self.nt, self.x0, self.dx, self.y0, self.dy,
                         self.tau0, self.dtau)

    def image_grids(self):
        return (ImageGrid(self.nx, self.x0, self.dx), ImageGrid(self.ny, self.y0, self.dy),
                ImageGrid(self.nt, self.tau0, self.dtau))

    @property
    def shape(self):
        return (self.nx, self.ny, self.nt)


def _ptr(t) -> int:
    return t.data_ptr() if t is not None else 0


def _need(t, dtype, name, numel=None):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements, got {t.numel()}")


def _stream(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return _p(stream.cuda_stream)


class Survey:
    """Acquisition geometry on the device: midpoints and half offsets."""

    def __init__(self, mid_x, mid_y, off_x, off_y, n_samples: int, dt: float):
        torch = _torch()
        n = mid_x.numel()
        for name, t in (("mid_x", mid_x), ("mid_y", mid_y), ("off_x", off_x), ("off_y", off_y)):
            _need(t, torch.float32, name, n)
        self.mid_x, self.mid_y, self.off_x, self.off_y = mid_x, mid_y, off_x, off_y
        self.n_traces, self.n_samples, self.dt = int(n), int(n_samples), float(dt)

    def c(self) -> PsktmSurvey:
        return PsktmSurvey(self.n_traces, self.n_samples, self.dt, _ptr(self.mid_x),
                           _ptr(self.mid_y), _ptr(self.off_x), _ptr(self.off_y))


def _velocity(vel, grid: Grid):
    torch = _torch()
    _need(vel, torch.float32, "vel")
    if vel.numel() == grid.nt:
        return PsktmVelocity(VEL_1D, _ptr(vel))
    if vel.numel() == grid.nx * grid.ny * grid.nt:
        return PsktmVelocity(VEL_3D, _ptr(vel))
    raise ValueError(f"velocity must have nt={grid.nt} or nx*ny*nt elements, got {vel.numel()}")


def migrate_exact(grid: Grid, survey: Survey, traces, vel, image=None, stream=None):
    """Reference-exact migration (TRUNC + IEEE, 1-D velocity, no aperture).

    Bit-identical to the single-thread reference loop of src/psktmCPU.c.
    """
    torch = _torch()
    L = _lib.load()
    _need(traces, torch.float32, "traces", survey.n_traces * survey.n_samples)
    _need(vel, torch.float32, "vel", grid.nt)
    if image is None:
        image = torch.empty(grid.shape, dtype=torch.float32, device=traces.device)
    _need(image, torch.float32, "image", grid.nx * grid.ny * grid.nt)
    g, s = grid.c(), survey.c()
    check(L.psktm_migrate_exact(ctypes.byref(g), ctypes.byref(s), _p(_ptr(traces)),
                                _p(_ptr(vel)), _p(_ptr(image)), _stream(stream)))
    return image


class Plan:
    """psktm_plan: traces binned into output tiles with per-tile aperture lists."""

    def __init__(self, grid: Grid, survey: Survey, *, interp=INTERP_LINEAR, math=MATH_FAST,
                 aperture=0.0, tile_x=0, tile_y=0, traces_per_stage=0, precise=False,
                 splits=0, row_pad=0, stream=None):
        L = _lib.load()
        self.grid, self.survey = grid, survey
        self.opt = PsktmOptions(int(interp), int(math), float(aperture), int(tile_x),
                                int(tile_y), int(traces_per_stage), int(bool(precise)),
                                int(splits), int(row_pad))
        self._h = _p()
        g, s = grid.c(), survey.c()
        check(L.psktm_plan_create(ctypes.byref(self._h), ctypes.byref(g), ctypes.byref(s),
                                  ctypes.byref(self.opt), _stream(stream)))

    @property
    def handle(self):
        return self._h

    def updates(self) -> float:
        v = ctypes.c_double()
        check(_lib.load().psktm_plan_updates(self._h, ctypes.byref(v)))
        return v.value

    def nbytes(self) -> int:
        v = ctypes.c_size_t()
        check(_lib.load().psktm_plan_bytes(self._h, ctypes.byref(v)))
        return v.value

    def close(self):
        if self._h:
            _lib.load().psktm_plan_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """psktm_batch: independent gathers migrated by one persistent launch.

    ``gathers`` is a list of (Grid, first_trace, n_traces); traces of gather g
    are rows first_trace .. first_trace+n_traces-1 of the survey and trace
    arrays. Images are returned concatenated in gather order.
    """

    def __init__(self, gathers, survey: Survey, *, interp=INTERP_LINEAR, aperture=0.0,
                 tile_x=0, tile_y=0, traces_per_stage=0, precise=False, row_pad=0, stream=None):
        L = _lib.load()
        self.gathers = [(g, int(f), int(n)) for g, f, n in gathers]
        self.survey = survey
        arr = (PsktmGather * len(self.gathers))()
        for q, (g, f, n) in enumerate(self.gathers):
            arr[q] = PsktmGather(g.c(), f, n)
        self.opt = PsktmOptions(int(interp), int(MATH_FAST), float(aperture), int(tile_x),
                                int(tile_y), int(traces_per_stage), int(bool(precise)), 1,
                                int(row_pad))
        self._h = _p()
        s = survey.c()
        check(L.psktm_batch_create(ctypes.byref(self._h), arr, len(self.gathers),
                                   ctypes.byref(s), ctypes.byref(self.opt), _stream(stream)))
        self.sizes = [g.nx * g.ny * g.nt for g, _, _ in self.gathers]
        self.offsets = list(np.cumsum([0] + self.sizes[:-1]))
        self.nt = self.gathers[0][0].nt

    def updates(self) -> float:
        v = ctypes.c_double()
        check(_lib.load().psktm_batch_updates(self._h, ctypes.byref(v)))
        return v.value

    def migrate(self, traces, vel, images=None, *, scale=None, use_graph=False, gained=False,
                stream=None):
        torch = _torch()
        L = _lib.load()
        kinds = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.bfloat16: _lib.BF16}
        _need(traces, traces.dtype, "traces", self.survey.n_traces * self.survey.n_samples)
        kind = kinds[traces.dtype]
        if kind != _lib.F32:
            _need(scale, torch.float32, "scale", self.survey.n_traces)
        _need(vel, torch.float32, "vel", self.nt)
        total = int(sum(self.sizes))
        if images is None:
            images = torch.empty(total, dtype=torch.float32, device=traces.device)
        _need(images, torch.float32, "images", total)
        tr = PsktmTraces(kind, _ptr(traces), _ptr(scale) if kind != _lib.F32 else 0,
                         int(bool(gained)))
        v = PsktmVelocity(VEL_1D, _ptr(vel))
        check(L.psktm_batch_migrate(self._h, ctypes.byref(tr), ctypes.byref(v), _p(_ptr(images)),
                                    int(bool(use_graph)), _stream(stream)))
        return images

    def image(self, images, g):
        """View of gather g's patch image [nx, ny, nt] inside ``images``."""
        grid = self.gathers[g][0]
        o = int(self.offsets[g])
        return images[o:o + self.sizes[g]].view(grid.nx, grid.ny, grid.nt)

    def close(self):
        if self._h:
            _lib.load().psktm_batch_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def migrate(plan: Plan, traces, vel, image=None, *, scale=None, dtype=None, accumulate=False,
            gained=False, stream=None):
    """Migration traces -> image on the plan's grid (device tensors).

    ``traces`` is float32, float16 or bfloat16 [n_traces, n_samples]; for the
    16-bit forms ``scale`` [n_traces] float32 gives the per-trace scale.
    ``gained``: the samples already carry sqrt(n dt) (:func:`rho_filter` with
    ``gain=True`` or :func:`trace_gain`), so the kernel applies tau only.
    """
    torch = _torch()
    L = _lib.load()
    s, g = plan.survey, plan.grid
    kinds = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.bfloat16: _lib.BF16}
    if traces.dtype not in kinds:
        raise TypeError(f"unsupported trace dtype {traces.dtype}")
    _need(traces, traces.dtype, "traces", s.n_traces * s.n_samples)
    kind = kinds[traces.dtype]
    if kind != _lib.F32:
        _need(scale, torch.float32, "scale", s.n_traces)
    tr = PsktmTraces(kind, _ptr(traces), _ptr(scale) if kind != _lib.F32 else 0,
                     int(bool(gained)))
    v = _velocity(vel, g)
    if image is None:
        image = torch.empty(g.shape, dtype=torch.float32, device=traces.device)
    _need(image, torch.float32, "image", g.nx * g.ny * g.nt)
    check(L.psktm_migrate(plan.handle, ctypes.byref(tr), ctypes.byref(v), _p(_ptr(image)),
                          int(bool(accumulate)), _stream(stream)))
    return image


def demigrate(plan: Plan, image, vel, traces=None, *, accumulate=False, gained=False,
              stream=None):
    """Adjoint of :func:`migrate`: image -> float32 traces [n_traces, n_samples].

    ``gained``: the adjoint of the gained operator (:func:`trace_gain`, then
    :func:`migrate` with ``gained=True``): weight tau per update, output
    sample n times sqrtf(n dt).
    """
    torch = _torch()
    L = _lib.load()
    s, g = plan.survey, plan.grid
    _need(image, torch.float32, "image", g.nx * g.ny * g.nt)
    v = _velocity(vel, g)
    if traces is None:
        traces = torch.empty((s.n_traces, s.n_samples), dtype=torch.float32, device=image.device)
    _need(traces, torch.float32, "traces", s.n_traces * s.n_samples)
    fn = L.psktm_demigrate_gained if gained else L.psktm_demigrate
    check(fn(plan.handle, _p(_ptr(image)), ctypes.byref(v), _p(_ptr(traces)),
             int(bool(accumulate)), _stream(stream)))
    return traces


def rho_filter(traces, dw_us: float, opt: int = 2, out_dtype=None, *, matlab_trace_shift=False,
               gain=False, stream=None):
    """Rho filter (scripts/ktmApplyRhoFilter.m) on device traces [n, m].

    Returns float32 traces, or (quantised traces, per-trace scale) when
    ``out_dtype`` is torch.float16 / torch.bfloat16. ``gain=True`` multiplies
    sample n by sqrt(n dt) (dt = dw_us 1e-6 s) before quantising: migrate the
    result with ``gained=True``.
    """
    torch = _torch()
    L = _lib.load()
    _need(traces, torch.float32, "traces")
    n, m = traces.shape
    out_dtype = torch.float32 if out_dtype is None else out_dtype
    kinds = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.bfloat16: _lib.BF16}
    out = torch.empty((n, m), dtype=out_dtype, device=traces.device)
    scale = torch.empty((n,), dtype=torch.float32, device=traces.device) \
        if out_dtype != torch.float32 else None
    gain_dt = float(dw_us) * 1e-6 if gain else 0.0
    check(L.psktm_rho_filter_gain(_p(_ptr(traces)), int(n), int(m), float(dw_us), int(opt),
                                  kinds[out_dtype], _p(_ptr(out)), _p(_ptr(scale)),
                                  int(bool(matlab_trace_shift)), gain_dt, _stream(stream)))
    return out if scale is None else (out, scale)


def trace_gain(traces, dt: float, out_dtype=None, *, out=None, stream=None):
    """sample n times sqrtf(n dt) (the trace-domain part of the weight), as
    float32 (into ``out`` if given), or (quantised traces, per-trace scale)
    for float16 / bfloat16."""
    torch = _torch()
    L = _lib.load()
    _need(traces, torch.float32, "traces")
    n, m = traces.shape
    out_dtype = torch.float32 if out_dtype is None else out_dtype
    kinds = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.bfloat16: _lib.BF16}
    if out is None or out_dtype != torch.float32:
        out = torch.empty((n, m), dtype=out_dtype, device=traces.device)
    else:
        _need(out, torch.float32, "out", n * m)
    scale = torch.empty((n,), dtype=torch.float32, device=traces.device) \
        if out_dtype != torch.float32 else None
    check(L.psktm_trace_gain(_p(_ptr(traces)), int(n), int(m), float(dt), kinds[out_dtype],
                             _p(_ptr(out)), _p(_ptr(scale)), _stream(stream)))
    return out if scale is None else (out, scale)


def quantize(traces, out_dtype, *, stream=None):
    """FP32 -> FP16/BF16 with per-trace max-abs scales."""
    torch = _torch()
    L = _lib.load()
    _need(traces, torch.float32, "traces")
    n, m = traces.shape
    kinds = {torch.float16: _lib.F16, torch.bfloat16: _lib.BF16}
    out = torch.empty((n, m), dtype=out_dtype, device=traces.device)
    scale = torch.empty((n,), dtype=torch.float32, device=traces.device)
    check(L.psktm_quantize(_p(_ptr(traces)), int(n), int(m), kinds[out_dtype], _p(_ptr(out)),
                           _p(_ptr(scale)), _stream(stream)))
    return out, scale


DOT_WORKSPACE = 1024      # doubles in a psktm_dot_f64_ws workspace


def dot_workspace(device):
    """Workspace for :func:`dot_f64` in the fixed summation order."""
    torch = _torch()
    return torch.empty((DOT_WORKSPACE,), dtype=torch.float64, device=device)


def dot_f64(a, b, out=None, *, workspace=None, stream=None):
    """sum(a*b) of float32 device vectors accumulated in FP64 (device result).

    With ``workspace`` (:func:`dot_workspace`), psktm_dot_f64_ws: the partials
    are added in a fixed order, so equal inputs give equal bits on every call.
    Without it, psktm_dot_f64 (the workgroup sums meet in an FP64 atomic)."""
    torch = _torch()
    L = _lib.load()
    _need(a, torch.float32, "a")
    _need(b, torch.float32, "b", a.numel())
    if out is None:
        out = torch.empty((1,), dtype=torch.float64, device=a.device)
    if workspace is None:
        check(L.psktm_dot_f64(_p(_ptr(a)), _p(_ptr(b)), ctypes.c_size_t(a.numel()),
                              _p(_ptr(out)), _stream(stream)))
    else:
        _need(workspace, torch.float64, "workspace", DOT_WORKSPACE)
        check(L.psktm_dot_f64_ws(_p(_ptr(a)), _p(_ptr(b)), ctypes.c_size_t(a.numel()),
                                 _p(_ptr(out)), _p(_ptr(workspace)), _stream(stream)))
    return out


# --------------------------------------------------------------------------
# Trace preparation (the MATLAB writers' steps), device tensors.
# --------------------------------------------------------------------------

def offset_sort(off_x, off_y, *, stream=None):
    """Permutation sorting traces by half-offset magnitude (stable), and the
    sorted magnitudes (scripts/ktmGetSortedOffTraces.m:17-22)."""
    torch = _torch()
    n = off_x.numel()
    _need(off_x, torch.float32, "off_x")
    _need(off_y, torch.float32, "off_y", n)
    perm = torch.empty(n, dtype=torch.int32, device=off_x.device)
    key = torch.empty(n, dtype=torch.float32, device=off_x.device)
    check(_lib.load().psktm_offset_sort(_p(_ptr(off_x)), _p(_ptr(off_y)), n, _p(_ptr(perm)),
                                        _p(_ptr(key)), _stream(stream)))
    return perm, key


def offset_classes(off_x, off_y, width, n_classes, *, stream=None):
    """(perm of kept traces class by class, class_start[n_classes+1])."""
    torch = _torch()
    n = off_x.numel()
    _need(off_x, torch.float32, "off_x")
    _need(off_y, torch.float32, "off_y", n)
    perm = torch.empty(max(n, 1), dtype=torch.int32, device=off_x.device)
    start = torch.empty(n_classes + 1, dtype=torch.int32, device=off_x.device)
    kept = ctypes.c_int()
    check(_lib.load().psktm_offset_classes(_p(_ptr(off_x)), _p(_ptr(off_y)), n, float(width),
                                           int(n_classes), _p(_ptr(perm)), _p(_ptr(start)),
                                           ctypes.byref(kept), _stream(stream)))
    return perm[:kept.value], start


def select_offset(off_x, off_y, h, tol, *, stream=None):
    """Indices of traces with h - tol < round(|h|) < h + tol, in trace order
    (scripts/ktmGetConstOffTraces.m:18-28)."""
    torch = _torch()
    n = off_x.numel()
    _need(off_x, torch.float32, "off_x")
    _need(off_y, torch.float32, "off_y", n)
    sel = torch.empty(max(n, 1), dtype=torch.int32, device=off_x.device)
    kept = ctypes.c_int()
    check(_lib.load().psktm_select_offset(_p(_ptr(off_x)), _p(_ptr(off_y)), n, float(h),
                                          float(tol), _p(_ptr(sel)), ctypes.byref(kept),
                                          _stream(stream)))
    return sel[:kept.value]


def gather_pad(traces, sze, rows=None, *, stream=None):
    """Rows ``rows`` of ``traces`` (all if None), zero padded in both
    dimensions to multiples of ``sze`` (scripts/ktmWriteTraceBlockConstOffset.m:100-118)."""
    torch = _torch()
    _need(traces, torch.float32, "traces")
    n, m = traces.shape
    nr = n if rows is None else rows.numel()
    if rows is not None:
        _need(rows, torch.int32, "rows")
    R, C = -(-nr // sze) * sze, -(-m // sze) * sze
    out = torch.empty((R, C), dtype=torch.float32, device=traces.device)
    check(_lib.load().psktm_gather_pad(_p(_ptr(traces)), m, _p(_ptr(rows)), nr, int(sze),
                                       _p(_ptr(out)), _stream(stream)))
    return out


def upsample_traces(traces, N, *, stream=None):
    """Linear interpolation of every trace onto the sample grid
    0 : 2^-N : m-1 (scripts/ktmInterpTraceBlock.m:134-154)."""
    torch = _torch()
    _need(traces, torch.float32, "traces")
    n, m = traces.shape
    mo = (m - 1) * (1 << N) + 1
    out = torch.empty((n, mo), dtype=torch.float32, device=traces.device)
    check(_lib.load().psktm_upsample_traces(_p(_ptr(traces)), n, m, int(N), _p(_ptr(out)),
                                            _stream(stream)))
    return out


# --------------------------------------------------------------------------
# SEG-Y / SU ingest.
# --------------------------------------------------------------------------

SEGY_HEADER_COLUMNS = 101


def segy_text_header(path):
    """40 ASCII lines of the EBCDIC header (scripts/ktmGetInputMeta.m:30-56)."""
    buf = ctypes.create_string_buffer(3241)
    check(_lib.load().psktm_segy_text_header(os.fsencode(path), buf))
    return buf.raw[:3240].decode("latin-1").split("\n")[:40]


def segy_reel_header(path):
    """The 27 binary reel values (scripts/ktmGetInputMeta.m:58-76)."""
    import numpy as np
    v = (ctypes.c_longlong * 27)()
    check(_lib.load().psktm_segy_reel_header(os.fsencode(path), v))
    return np.array(v[:], dtype=np.int64)


def segy_open(path, big_endian=True):
    """Record layout and trace count (scripts/ktmGetTotalTraceNo.m)."""
    L = _lib.PsktmSegyLayout()
    check(_lib.load().psktm_segy_open(os.fsencode(path), int(bool(big_endian)), ctypes.byref(L)))
    return L


def segy_trace_headers(layout, path, first, count, *, saturate=False):
    """count x 101 header columns on the host (scripts/ktmGetInputMeta.m:78-153)."""
    import numpy as np
    out = np.zeros((int(count), SEGY_HEADER_COLUMNS), dtype=np.int32)
    check(_lib.load().psktm_segy_trace_headers(ctypes.byref(layout), os.fsencode(path),
                                               int(first), int(count), int(bool(saturate)),
                                               _p(out.ctypes.data) if count else None))
    return out


def segy_decode(layout, records, count, *, ibm_reference=True, headers=True, stream=None):
    """Decode raw records (a uint8 device tensor) into FP32 traces [count, ns]
    and, optionally, int32 headers [count, 101]."""
    torch = _torch()
    dev = records.device
    tr = torch.empty((int(count), layout.n_samples), dtype=torch.float32, device=dev)
    hd = torch.empty((int(count), SEGY_HEADER_COLUMNS), dtype=torch.int32, device=dev) \
        if headers else None
    check(_lib.load().psktm_segy_decode(ctypes.byref(layout), _p(_ptr(records)), int(count),
                                        int(bool(ibm_reference)), _p(_ptr(tr)),
                                        _p(_ptr(hd)) if hd is not None else None,
                                        _stream(stream)))
    return (tr, hd) if headers else tr


def segy_load(layout, path, first=0, count=None, *, ibm_reference=True, headers=True,
              device="cuda", stream=None):
    """Read records [first, first + count) from the file straight to device
    tensors: FP32 traces [count, ns] and optional int32 headers [count, 101]."""
    torch = _torch()
    count = layout.n_traces - int(first) if count is None else int(count)
    tr = torch.empty((count, layout.n_samples), dtype=torch.float32, device=device)
    hd = torch.empty((count, SEGY_HEADER_COLUMNS), dtype=torch.int32, device=device) \
        if headers else None
    check(_lib.load().psktm_segy_load(ctypes.byref(layout), os.fsencode(path), int(first), count,
                                      int(bool(ibm_reference)), _p(_ptr(tr)),
                                      _p(_ptr(hd)) if hd is not None else None,
                                      _stream(stream)))
    return (tr, hd) if headers else tr


# --------------------------------------------------------------------------
# Reference-named host entry points.
# --------------------------------------------------------------------------

def _host(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def _legacy_migration(entry, imageX, imageY, imageZ, config, arrays, image):
    L = _lib.load()
    arrs = [_host(a) for a in arrays]
    n = imageX.imagePnts * imageY.imagePnts * imageZ.imagePnts
    if image is None:
        image = np.empty(n, dtype=np.float32)
    getattr(L, entry)(ctypes.byref(imageX), ctypes.byref(imageY), ctypes.byref(imageZ),
                      ctypes.byref(config), *[a.ctypes.data_as(_p) for a in arrs],
                      image.ctypes.data_as(_p))
    return image


def ktm_migration_gpu(imageX: ImageGrid, imageY: ImageGrid, imageZ: ImageGrid,
                      config: JobParams, midX, midY, offX, offY, traces, slowness,
                      image=None):
    """ktmMigrationGPU (src/psktmCUDA.cu:211-221) on host arrays.

    Returns the image (zeroed when the call failed; see :func:`last_error`).
    """
    return _legacy_migration("ktmMigrationGPU", imageX, imageY, imageZ, config,
                             (midX, midY, offX, offY, traces, slowness), image)


def ktm_migration_cpu(imageX: ImageGrid, imageY: ImageGrid, imageZ: ImageGrid,
                      config: JobParams, midX, midY, offX, offY, traces, slowness,
                      image=None):
    """ktmMigrationCPU (src/psktmCPU.c:13-23) on host arrays: the same
    arguments and the image of the reference loop run on one thread,
    computed on the GPU by the reference-exact kernel.

    Returns the image (zeroed when the call failed; see :func:`last_error`).
    """
    return _legacy_migration("ktmMigrationCPU", imageX, imageY, imageZ, config,
                             (midX, midY, offX, offY, traces, slowness), image)


def init_three_axes(imageX: ImageGrid, imageY: ImageGrid, imageZ: ImageGrid):
    """initThreeAxes (src/psktmFunctions.c:19-39)."""
    L = _lib.load()
    xx = np.empty(imageX.imagePnts, np.float32)
    yy = np.empty(imageY.imagePnts, np.float32)
    zz = np.empty(imageZ.imagePnts, np.float32)
    L.initThreeAxes(ctypes.byref(imageX), ctypes.byref(imageY), ctypes.byref(imageZ),
                    xx.ctypes.data_as(_p), yy.ctypes.data_as(_p), zz.ctypes.data_as(_p))
    return xx, yy, zz


def init_two_axes(imageY: ImageGrid, imageZ: ImageGrid):
    """initTwoAxes (src/psktmFunctions.c:1-17)."""
    L = _lib.load()
    yy = np.empty(imageY.imagePnts, np.float32)
    zz = np.empty(imageZ.imagePnts, np.float32)
    L.initTwoAxes(ctypes.byref(imageY), ctypes.byref(imageZ), yy.ctypes.data_as(_p),
                  zz.ctypes.data_as(_p))
    return yy, zz


def ibm32_to_ieee(words):
    """ktmIBM32toIEEE (src/ioFunctions.c:4-40), reference arithmetic."""
    L = _lib.load()
    w = _host(words, np.uint32)
    out = np.empty(w.shape, np.float32)
    L.ktmIBM32toIEEE(w.ctypes.data_as(_p), out.ctypes.data_as(_p), int(w.size))
    return out


def ibm32_to_ieee_exact(words):
    """True IBM decode (fraction / 2^24) on the host."""
    L = _lib.load()
    w = _host(words, np.uint32)
    out = np.empty(w.shape, np.float32)
    L.psktm_ibm32_to_ieee_exact(w.ctypes.data_as(_p), out.ctypes.data_as(_p), int(w.size))
    return out


def ibm32_to_ieee_device(words, out=None, *, stream=None):
    """True IBM decode on the device: int32/uint32 words tensor -> float32."""
    torch = _torch()
    L = _lib.load()
    if words.dtype not in (torch.int32, torch.uint32):
        raise TypeError("words must be int32 or uint32")
    _need(words, words.dtype, "words")
    if out is None:
        out = torch.empty(words.shape, dtype=torch.float32, device=words.device)
    _need(out, torch.float32, "out", words.numel())
    check(L.psktm_ibm32_to_ieee_device(_p(_ptr(words)), _p(_ptr(out)),
                                       ctypes.c_size_t(words.numel()), _stream(stream)))
    return out


def load_ktm_config(path: str):
    """Read inp.txt into (imageX, imageY, imageZ, config)."""
    L = _lib.load()
    X, Y, Z, J = ImageGrid(), ImageGrid(), ImageGrid(), JobParams()
    check(L.psktm_load_ktm_config(path.encode(), ctypes.byref(X), ctypes.byref(Y),
                                  ctypes.byref(Z), ctypes.byref(J)))
    return X, Y, Z, J
