"""CPU oracle for the PSTM path.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, always as the checker, never as the measured or
shipped path. The product library never imports this package.

* ``liboracle.so`` (``psktm_oracle.c``): FP32 restatement of ktmMigrationCPU
  (src/psktmCPU.c:13-101), FP64 forward/adjoint, IBM decode
  (src/ioFunctions.c:4-40), axes (src/psktmFunctions.c:19-39).
* :func:`apply_rho_filter`: numpy restatement of scripts/ktmApplyRhoFilter.m.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f = ctypes.c_float
_i = ctypes.c_int
_p = ctypes.c_void_p


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        # PSKTM_ORACLE_LIBRARY: the host-sanitizer test run's instrumented build
        path = os.environ.get("PSKTM_ORACLE_LIBRARY") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_pstm_f32.argtypes = [_i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _i, _i, _f,
                                      _p, _p, _p, _p, _p, _p, _i, _f, _p, _i]
        L.oracle_pstm_f32.restype = None
        L.oracle_ktm_migration_racy.argtypes = [_i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _i, _f,
                                                _p, _p, _p, _p, _p, _p, _p, _i]
        L.oracle_ktm_migration_racy.restype = None
        L.oracle_pstm_f64.argtypes = [_i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _i, _f,
                                      _p, _p, _p, _p, _p, _p, _p, _i, _f, _i, _i,
                                      _i, _i, _i, _i, _p, _i]
        L.oracle_pstm_f64.restype = None
        L.oracle_pstm_f64_bins.argtypes = [_i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _i, _f,
                                           _p, _p, _p, _p, _p, _p, _p, _i, _f, _i, _i,
                                           _p, _p, _i, _p, _i]
        L.oracle_pstm_f64_bins.restype = None
        L.oracle_pstm_adj_f64.argtypes = [_i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _i, _f,
                                          _p, _p, _p, _p, _p, _p, _i, _f, _i, _p, _i]
        L.oracle_pstm_adj_f64.restype = None
        L.oracle_ibm32_to_ieee_ref.argtypes = [_p, _p, _i]
        L.oracle_ibm32_to_ieee_ref.restype = None
        L.oracle_init_three_axes.argtypes = [_p, _p, _p, _p, _p, _p]
        L.oracle_init_three_axes.restype = None
        _LIB = L
    return _LIB


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_p)


class Grid:
    """Image axes: x slowest, tau (one-way time) fastest."""

    def __init__(self, nx, x0, dx, ny, y0, dy, nt, tau0, dtau):
        self.nx, self.x0, self.dx = int(nx), float(x0), float(dx)
        self.ny, self.y0, self.dy = int(ny), float(y0), float(dy)
        self.nt, self.tau0, self.dtau = int(nt), float(tau0), float(dtau)

    def args(self):
        return (self.nx, self.x0, self.dx, self.ny, self.y0, self.dy,
                self.nt, self.tau0, self.dtau)

    def axes(self):
        """Binary32 axes, start + (float)i * step (src/psktmFunctions.c:28-34)."""
        def ax(n, s, d):
            return (np.float32(s) + np.arange(n, dtype=np.float32) * np.float32(d)).astype(np.float32)
        return ax(self.nx, self.x0, self.dx), ax(self.ny, self.y0, self.dy), ax(self.nt, self.tau0, self.dtau)


def pstm_f32(grid: Grid, mid_x, mid_y, off_x, off_y, traces, vel, dt, *, vel3d=False,
             aperture=0.0, k0=0, k1=None, threads=0):
    """Binary32 migration in the reference operation order (src/psktmCPU.c:42-87).

    Bit-identical to the single-thread reference loop; k0..k1 selects traces.
    """
    traces = np.ascontiguousarray(traces, dtype=np.float32)
    ntr, ns = traces.shape
    k1 = ntr if k1 is None else int(k1)
    out = np.zeros((grid.nx, grid.ny, grid.nt), dtype=np.float32)
    keep = []
    ptrs = []
    for a in (mid_x, mid_y, off_x, off_y, traces, vel):
        arr, p = _c(a, np.float32)
        keep.append(arr)
        ptrs.append(p)
    lib().oracle_pstm_f32(*grid.args(), int(k0), k1, ns, float(dt), *ptrs,
                          int(bool(vel3d)), float(aperture),
                          out.ctypes.data_as(_p), int(threads))
    return out


def ktm_migration_racy(grid: Grid, mid_x, mid_y, off_x, off_y, traces, vel, dt, threads=0):
    """The reference's trace-parallel loop as written (racy; timing only)."""
    traces = np.ascontiguousarray(traces, dtype=np.float32)
    ntr, ns = traces.shape
    out = np.zeros((grid.nx, grid.ny, grid.nt), dtype=np.float32)
    keep, ptrs = [], []
    for a in (mid_x, mid_y, off_x, off_y, traces, vel):
        arr, p = _c(a, np.float32)
        keep.append(arr)
        ptrs.append(p)
    lib().oracle_ktm_migration_racy(*grid.args(), ntr, ns, float(dt), *ptrs,
                                    out.ctypes.data_as(_p), int(threads))
    return out


TRUNC, LINEAR = 0, 1


def pstm_f64(grid: Grid, mid_x, mid_y, off_x, off_y, traces, vel, dt, *, vel3d=False,
             aperture=0.0, interp=LINEAR, scale=None, bins=None, gained=False, threads=0):
    """FP64 forward operator on bins ``(ix0, ix1, iy0, iy1)`` (default: all).

    ``gained``: the traces carry sqrt(n dt) per sample and the amplitude
    weight is tau alone (the GPU's gained-trace form)."""
    traces = np.ascontiguousarray(traces, dtype=np.float32)
    ntr, ns = traces.shape
    ix0, ix1, iy0, iy1 = bins if bins is not None else (0, grid.nx, 0, grid.ny)
    out = np.zeros((ix1 - ix0, iy1 - iy0, grid.nt), dtype=np.float64)
    keep, ptrs = [], []
    for a in (mid_x, mid_y, off_x, off_y, traces):
        arr, p = _c(a, np.float32)
        keep.append(arr)
        ptrs.append(p)
    if scale is not None:
        sc, sp = _c(scale, np.float32)
        keep.append(sc)
    else:
        sp = None
    va, vp = _c(vel, np.float32)
    lib().oracle_pstm_f64(*grid.args(), ntr, ns, float(dt), *ptrs, sp, vp,
                          int(bool(vel3d)), float(aperture), int(interp), int(bool(gained)),
                          int(ix0), int(ix1), int(iy0), int(iy1),
                          out.ctypes.data_as(_p), int(threads))
    return out


def pstm_f64_bins(grid: Grid, mid_x, mid_y, off_x, off_y, traces, vel, dt, bins, *,
                  vel3d=False, aperture=0.0, interp=LINEAR, scale=None, gained=False,
                  threads=0):
    """:func:`pstm_f64` on an explicit list of bins [(i, j), ...] -> [nb, nt]."""
    traces = np.ascontiguousarray(traces, dtype=np.float32)
    ntr, ns = traces.shape
    bi = np.ascontiguousarray([b[0] for b in bins], dtype=np.int32)
    bj = np.ascontiguousarray([b[1] for b in bins], dtype=np.int32)
    assert bi.min() >= 0 and bi.max() < grid.nx and bj.min() >= 0 and bj.max() < grid.ny
    out = np.zeros((len(bins), grid.nt), dtype=np.float64)
    keep, ptrs = [], []
    for a in (mid_x, mid_y, off_x, off_y, traces):
        arr, p = _c(a, np.float32)
        keep.append(arr)
        ptrs.append(p)
    if scale is not None:
        sc, sp = _c(scale, np.float32)
        keep.append(sc)
    else:
        sp = None
    va, vp = _c(vel, np.float32)
    lib().oracle_pstm_f64_bins(*grid.args(), ntr, ns, float(dt), *ptrs, sp, vp,
                               int(bool(vel3d)), float(aperture), int(interp), int(bool(gained)),
                               bi.ctypes.data_as(_p), bj.ctypes.data_as(_p), len(bins),
                               out.ctypes.data_as(_p), int(threads))
    return out


def pstm_adj_f64(grid: Grid, mid_x, mid_y, off_x, off_y, image, vel, dt, ns, *,
                 vel3d=False, aperture=0.0, interp=LINEAR, threads=0):
    """FP64 adjoint (demigration): image -> traces [ntr][ns]."""
    image = np.ascontiguousarray(image, dtype=np.float32).reshape(grid.nx, grid.ny, grid.nt)
    ntr = len(mid_x)
    out = np.zeros((ntr, int(ns)), dtype=np.float64)
    keep, ptrs = [], []
    for a in (mid_x, mid_y, off_x, off_y, image, vel):
        arr, p = _c(a, np.float32)
        keep.append(arr)
        ptrs.append(p)
    lib().oracle_pstm_adj_f64(*grid.args(), ntr, int(ns), float(dt), *ptrs,
                              int(bool(vel3d)), float(aperture), int(interp),
                              out.ctypes.data_as(_p), int(threads))
    return out


def ibm32_to_ieee_ref(words):
    """ktmIBM32toIEEE as written (src/ioFunctions.c:18-38)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.zeros(w.shape, dtype=np.float32)
    lib().oracle_ibm32_to_ieee_ref(w.ctypes.data_as(_p), out.ctypes.data_as(_p), int(w.size))
    return out


def init_three_axes(nx, x0, dx, ny, y0, dy, nt, t0, dt):
    """initThreeAxes through the C restatement (src/psktmFunctions.c:19-39)."""
    class IG(ctypes.Structure):
        _fields_ = [("imagePnts", ctypes.c_int), ("imageStart", ctypes.c_float),
                    ("imageStep", ctypes.c_float)]
    X, Y, Z = IG(nx, x0, dx), IG(ny, y0, dy), IG(nt, t0, dt)
    xx = np.zeros(nx, np.float32)
    yy = np.zeros(ny, np.float32)
    zz = np.zeros(nt, np.float32)
    lib().oracle_init_three_axes(ctypes.byref(X), ctypes.byref(Y), ctypes.byref(Z),
                                 xx.ctypes.data_as(_p), yy.ctypes.data_as(_p),
                                 zz.ctypes.data_as(_p))
    return xx, yy, zz


# --------------------------------------------------------------------------
# Rho filter: restatement of scripts/ktmApplyRhoFilter.m (Octave 3.2.0 fft).
# --------------------------------------------------------------------------

def rho_fft_length(m: int) -> int:
    """M = 2^ceil(log(m)/log(2)) (scripts/ktmApplyRhoFilter.m:26)."""
    return int(2 ** np.ceil(np.log(m) / np.log(2)))


def apply_rho_filter(trace, dw, opt):
    """rhoTrace = ktmApplyRhoFilter(trace, dw, opt), line by line.

    ``trace`` is [n traces][m samples]; ``dw`` is the sample interval in
    microseconds. MATLAB/Octave ``fftshift`` on a matrix shifts every
    dimension, exactly like ``numpy.fft.fftshift`` with ``axes=None``; this is
    why an odd trace count comes back rotated by one trace.
    """
    trace = np.asarray(trace, dtype=np.float64)
    if trace.ndim == 1:
        trace = trace[None, :]
    n, m = trace.shape                                                 # :22
    M = rho_fft_length(m)                                              # :26
    Dr = dw * 1e-6                                                     # :31
    omega = np.tile(2 * np.pi * np.arange(-M // 2, M // 2) / Dr, (n, 1)).T  # :33
    tT = trace.T                                                       # m x n
    X = np.zeros((M, n), dtype=np.complex128)
    X[:m, :] = tT
    fTrace = np.fft.fftshift(np.fft.fft(X, axis=0))                    # :37
    if opt == 1:
        G = np.abs(omega) * fTrace                                     # :43
    elif opt == 2:
        G = np.sqrt(np.abs(omega)) * np.exp((1j * np.pi / 4) * np.sign(omega)) * fTrace  # :50
    else:
        G = 1j * omega * fTrace                                        # :57
    rho = np.real(np.fft.ifft(np.fft.fftshift(G), axis=0)).T           # :43/:50/:57
    return rho[:, :m]                                                  # :44/:51/:58


def rho_filter_per_trace(trace, dw, opt):
    """The same filter applied trace by trace (no 2-D fftshift rotation)."""
    trace = np.asarray(trace, dtype=np.float64)
    return np.concatenate([apply_rho_filter(trace[c:c + 1], dw, opt) for c in range(trace.shape[0])], 0) \
        if trace.shape[0] else trace.copy()


# --------------------------------------------------------------------------
# Trace preparation: restatements of the MATLAB writers' steps.
# --------------------------------------------------------------------------

def sorted_offset_order(off_x, off_y):
    """[~, index] = sort(offset) (scripts/ktmGetSortedOffTraces.m:17-19): MATLAB's
    sort is stable; offset is the half-offset magnitude."""
    mag = np.sqrt(np.asarray(off_x, np.float64) ** 2 + np.asarray(off_y, np.float64) ** 2)
    return np.argsort(mag, kind="stable")


def const_offset_selection(off_x, off_y, h, tol):
    """find((offset > h - tol) & (offset < h + tol)) with offset =
    round(|h|) (scripts/ktmGetConstOffTraces.m:18-23; MATLAB round: halves
    away from zero)."""
    mag = np.sqrt(np.asarray(off_x, np.float64) ** 2 + np.asarray(off_y, np.float64) ** 2)
    r = np.floor(mag + 0.5)
    return np.nonzero((r > h - tol) & (r < h + tol))[0]


def pad_block(traces, sze):
    """Zero padding of records and samples to multiples of sze
    (scripts/ktmWriteTraceBlockConstOffset.m:100-118)."""
    traces = np.asarray(traces)
    n, m = traces.shape
    out = np.zeros((-(-n // sze) * sze, -(-m // sze) * sze), traces.dtype)
    out[:n, :m] = traces
    return out


def upsample_time(traces, N):
    """interp1(0:NN-1, trace, 0:1/2^N:NN-1) for every trace
    (scripts/ktmInterpTraceBlock.m:134-154)."""
    traces = np.asarray(traces, np.float64)
    n, m = traces.shape
    t_new = np.arange((m - 1) * 2 ** N + 1) / 2.0 ** N
    return np.stack([np.interp(t_new, np.arange(m), traces[k]) for k in range(n)]) \
        if n else np.zeros((0, len(t_new)))
