"""psktm: MI355X-native prestack Kirchhoff time migration (host side).

The Python layer mirrors the reference's C interface for the migration path
(src/psktmCPU.c:13-23, src/psktmCUDA.cu:211-221, src/psktmFunctions.c,
src/ioFunctions.c) over the C ABI of ``libpsktm.so``; every compute call runs
a HIP kernel on the current device. Torch tensors are used for device memory
and streams only.
"""
from ._lib import (BF16, F16, F32, INTERP_LINEAR, INTERP_TRUNC, MATH_FAST, MATH_IEEE,
                   VEL_1D, VEL_3D, ImageGrid, JobParams, PsktmError, last_error, load)
from .api import (Batch, Grid, Plan, Survey, demigrate, dot_f64, dot_workspace, ibm32_to_ieee,
                  ibm32_to_ieee_device, ibm32_to_ieee_exact, init_three_axes,
                  init_two_axes, ktm_migration_cpu, ktm_migration_gpu, load_ktm_config, migrate,
                  migrate_exact, quantize, rho_filter, trace_gain, offset_sort, offset_classes,
                  select_offset, gather_pad, upsample_traces, segy_text_header,
                  segy_reel_header, segy_open, segy_trace_headers, segy_decode, segy_load)

__all__ = [
    "BF16", "F16", "F32", "INTERP_LINEAR", "INTERP_TRUNC", "MATH_FAST", "MATH_IEEE",
    "VEL_1D", "VEL_3D", "ImageGrid", "JobParams", "PsktmError", "last_error", "load",
    "Batch", "Grid", "Plan", "Survey", "demigrate", "dot_f64", "dot_workspace", "ibm32_to_ieee",
    "ibm32_to_ieee_device", "ibm32_to_ieee_exact", "init_three_axes", "init_two_axes",
    "ktm_migration_cpu", "ktm_migration_gpu", "load_ktm_config", "migrate", "migrate_exact", "quantize",
    "rho_filter", "trace_gain", "offset_sort", "offset_classes", "select_offset", "gather_pad",
    "upsample_traces", "segy_text_header", "segy_reel_header", "segy_open",
    "segy_trace_headers", "segy_decode", "segy_load",
]
