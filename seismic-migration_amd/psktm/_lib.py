"""ctypes binding of libpsktm.so (the C ABI declared in include/psktm.h).

The library is loaded from this directory only (or from PSKTM_LIBRARY, which
the host-sanitizer test run sets to its instrumented build); there is no
fallback. If it is missing, :func:`load` raises ImportError with the build
command.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PSKTM_LIBRARY") or os.path.join(HERE, "libpsktm.so")

_c_int, _c_float, _c_double, _p = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
_c_size = ctypes.c_size_t

# psktm_status
OK, E_INVAL, E_NODEV, E_HIP, E_ALLOC, E_LAUNCH, E_IO, E_UNSUPPORTED = range(8)
STATUS_NAMES = ["OK", "E_INVAL", "E_NODEV", "E_HIP", "E_ALLOC", "E_LAUNCH", "E_IO", "E_UNSUPPORTED"]
INTERP_TRUNC, INTERP_LINEAR = 0, 1
MATH_IEEE, MATH_FAST = 0, 1
F32, F16, BF16 = 0, 1, 2
VEL_1D, VEL_3D = 0, 1


class ImageGrid(ctypes.Structure):
    """struct imageGrid (src/psktmFunctions.c:8-12)."""
    _fields_ = [("imagePnts", _c_int), ("imageStart", _c_float), ("imageStep", _c_float)]


class JobParams(ctypes.Structure):
    """struct jobParams (src/psktmCPU.c:29,59,80,82)."""
    _fields_ = [("traceNo", _c_int), ("tracePts", _c_int), ("traceSampleRate", _c_float)]


class FileInputPtrs(ctypes.Structure):
    """struct fileInputPtrs (src/ioFunctions.c:94-153)."""
    _fields_ = [("inputData", _p), ("inputSrcX", _p), ("inputSrcY", _p),
                ("inputRecX", _p), ("inputRecY", _p), ("flag", _c_int)]


class PsktmGrid(ctypes.Structure):
    _fields_ = [("nx", _c_int), ("ny", _c_int), ("nt", _c_int),
                ("x0", _c_float), ("dx", _c_float), ("y0", _c_float), ("dy", _c_float),
                ("tau0", _c_float), ("dtau", _c_float)]


class PsktmSurvey(ctypes.Structure):
    _fields_ = [("n_traces", _c_int), ("n_samples", _c_int), ("dt", _c_float),
                ("mid_x", _p), ("mid_y", _p), ("off_x", _p), ("off_y", _p)]


class PsktmTraces(ctypes.Structure):
    _fields_ = [("dtype", _c_int), ("data", _p), ("scale", _p), ("gained", _c_int)]


class PsktmVelocity(ctypes.Structure):
    _fields_ = [("kind", _c_int), ("v", _p)]


class PsktmOptions(ctypes.Structure):
    _fields_ = [("interp", _c_int), ("math", _c_int), ("aperture", _c_float),
                ("tile_x", _c_int), ("tile_y", _c_int), ("traces_per_stage", _c_int),
                ("precise", _c_int), ("splits", _c_int), ("row_pad", _c_int)]


class PsktmSegyLayout(ctypes.Structure):
    _fields_ = [("data_offset", ctypes.c_longlong), ("n_traces", ctypes.c_longlong),
                ("record_bytes", ctypes.c_longlong), ("n_samples", ctypes.c_int),
                ("sample_interval_us", ctypes.c_int), ("format", ctypes.c_int),
                ("bytes_per_sample", ctypes.c_int), ("big_endian", ctypes.c_int),
                ("ragged", ctypes.c_int)]


class PsktmGather(ctypes.Structure):
    _fields_ = [("grid", PsktmGrid), ("first_trace", _c_int), ("n_traces", _c_int)]


_P = ctypes.POINTER
_SIGNATURES = {
    # reference API
    "initTwoAxes": (None, [_P(ImageGrid), _P(ImageGrid), _p, _p]),
    "initThreeAxes": (None, [_P(ImageGrid), _P(ImageGrid), _P(ImageGrid), _p, _p, _p]),
    "ktmIBM32toIEEE": (None, [_p, _p, _c_int]),
    "fileSizeFourBytes": (_c_int, [_p]),
    "binReadUints": (_p, [_p, _c_int]),
    "binReadFloats": (_p, [_p, _c_int]),
    "fileWrite": (None, [_p, _p, _c_int]),
    "errorMessage": (None, [ctypes.c_char_p]),
    "loadInputData": (None, [_P(FileInputPtrs)]),
    "loadRTM2dConfig": (None, []),
    "loadKTMConfig": (None, []),
    "loadKDMConfig": (None, []),
    "ktmMigrationGPU": (None, [_P(ImageGrid), _P(ImageGrid), _P(ImageGrid), _P(JobParams),
                               _p, _p, _p, _p, _p, _p, _p]),
    "ktmMigrationCPU": (None, [_P(ImageGrid), _P(ImageGrid), _P(ImageGrid), _P(JobParams),
                               _p, _p, _p, _p, _p, _p, _p]),
    # extended API
    "psktm_last_error": (ctypes.c_char_p, []),
    "psktm_version": (ctypes.c_char_p, []),
    "psktm_device_count": (_c_int, []),
    "psktm_plan_create": (_c_int, [_P(_p), _P(PsktmGrid), _P(PsktmSurvey), _P(PsktmOptions), _p]),
    "psktm_plan_destroy": (_c_int, [_p]),
    "psktm_plan_updates": (_c_int, [_p, _P(_c_double)]),
    "psktm_plan_bytes": (_c_int, [_p, _P(_c_size)]),
    "psktm_migrate": (_c_int, [_p, _P(PsktmTraces), _P(PsktmVelocity), _p, _c_int, _p]),
    "psktm_demigrate": (_c_int, [_p, _p, _P(PsktmVelocity), _p, _c_int, _p]),
    "psktm_demigrate_gained": (_c_int, [_p, _p, _P(PsktmVelocity), _p, _c_int, _p]),
    "psktm_batch_create": (_c_int, [_P(_p), _P(PsktmGather), _c_int, _P(PsktmSurvey),
                                    _P(PsktmOptions), _p]),
    "psktm_batch_destroy": (_c_int, [_p]),
    "psktm_batch_updates": (_c_int, [_p, _P(_c_double)]),
    "psktm_batch_migrate": (_c_int, [_p, _P(PsktmTraces), _P(PsktmVelocity), _p, _c_int, _p]),
    "psktm_migrate_exact": (_c_int, [_P(PsktmGrid), _P(PsktmSurvey), _p, _p, _p, _p]),
    "psktm_rho_filter": (_c_int, [_p, _c_int, _c_int, _c_float, _c_int, _c_int, _p, _p, _c_int, _p]),
    "psktm_rho_filter_gain": (_c_int, [_p, _c_int, _c_int, _c_float, _c_int, _c_int, _p, _p, _c_int,
                                       _c_float, _p]),
    "psktm_trace_gain": (_c_int, [_p, _c_int, _c_int, _c_float, _c_int, _p, _p, _p]),
    "psktm_quantize": (_c_int, [_p, _c_int, _c_int, _c_int, _p, _p, _p]),
    "psktm_ibm32_to_ieee_exact": (None, [_p, _p, _c_int]),
    "psktm_ibm32_to_ieee_device": (_c_int, [_p, _p, _c_size, _p]),
    "psktm_load_ktm_config": (_c_int, [ctypes.c_char_p, _P(ImageGrid), _P(ImageGrid),
                                       _P(ImageGrid), _P(JobParams)]),
    "psktm_offset_sort": (_c_int, [_p, _p, _c_int, _p, _p, _p]),
    "psktm_offset_classes": (_c_int, [_p, _p, _c_int, _c_float, _c_int, _p, _p, _P(_c_int), _p]),
    "psktm_select_offset": (_c_int, [_p, _p, _c_int, _c_float, _c_float, _p, _P(_c_int), _p]),
    "psktm_gather_pad": (_c_int, [_p, _c_int, _p, _c_int, _c_int, _p, _p]),
    "psktm_upsample_traces": (_c_int, [_p, _c_int, _c_int, _c_int, _p, _p]),
    "psktm_segy_text_header": (_c_int, [ctypes.c_char_p, ctypes.c_char_p]),
    "psktm_segy_reel_header": (_c_int, [ctypes.c_char_p, _P(ctypes.c_longlong)]),
    "psktm_segy_open": (_c_int, [ctypes.c_char_p, _c_int, _P(PsktmSegyLayout)]),
    "psktm_segy_trace_headers": (_c_int, [_P(PsktmSegyLayout), ctypes.c_char_p,
                                          ctypes.c_longlong, ctypes.c_longlong, _c_int, _p]),
    "psktm_segy_decode": (_c_int, [_P(PsktmSegyLayout), _p, ctypes.c_longlong, _c_int, _p, _p,
                                   _p]),
    "psktm_segy_load": (_c_int, [_P(PsktmSegyLayout), ctypes.c_char_p, ctypes.c_longlong,
                                 ctypes.c_longlong, _c_int, _p, _p, _p]),
    "psktm_dot_f64": (_c_int, [_p, _p, _c_size, _p, _p]),
    "psktm_dot_f64_ws": (_c_int, [_p, _p, _c_size, _p, _p, _p]),
    "psktm_dot_f64_noalloc": (_c_int, [_p, _p, _c_size, _p, _p]),
    "psktm_axpy_ratio": (_c_int, [_p, _p, _c_size, _p, _p, _c_float, _p]),
    "psktm_xpby_ratio": (_c_int, [_p, _p, _c_size, _p, _p, _p]),
}

_LIB = None


class PsktmError(RuntimeError):
    def __init__(self, status: int, message: str):
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"psktm {name}: {message}")
        self.status = status


def load() -> ctypes.CDLL:
    """Load libpsktm.so from this package directory (raises if not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run "
                          "`make -C seismic-migration_amd/csrc`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def exported_symbols():
    """Names from the signature table that the loaded library exports."""
    L = load()
    return sorted(n for n in _SIGNATURES if hasattr(L, n))


def last_error() -> str:
    return (load().psktm_last_error() or b"").decode()


def check(status: int) -> None:
    if status != OK:
        raise PsktmError(status, last_error())
