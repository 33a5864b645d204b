"""ctypes binding of libsparc_hip.so (the C ABI in include/sparc_hip.h).

The library is required: importing the bindings on a machine without the
built library, or calling into it without a HIP device, raises -- there is no
CPU fallback for the time-step path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsparc_hip.so")

SPARC_OK = 0
PREC_FP64, PREC_FP32 = 0, 1
DEFER_FILTER = 1                         # SPARC_DEFER_FILTER
GUARD_OFF, GUARD_RECORD, GUARD_HALT = 0, 1, 2

DP = C.POINTER(C.c_double)


class SparcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("option", C.c_int), ("precision", C.c_int), ("periodic", C.c_int),
        ("use_pml", C.c_int), ("use_hpml", C.c_int),
        ("npml", C.c_int * 6),
        ("g_xy_filt", C.c_int), ("g_z_filt", C.c_int),
        ("s_xy", C.c_double), ("s_z", C.c_double),
        ("xyfreq", C.c_int), ("zfreq", C.c_int),
        ("pulse_dir", C.c_int), ("bg3d", C.c_int),
        ("deltat", C.c_double), ("betas", C.c_double * 5),
        ("stretchx", C.c_double), ("stretchy", C.c_double),
        ("stretch", DP),
        ("device", C.c_int),
        ("tile", C.c_int * 3),
    ]


class SparcBg(C.Structure):
    _fields_ = [
        ("rho0", DP), ("rhoinv", DP), ("c2rho0", DP), ("gradrho0_z", DP), ("gradp0_z", DP),
        ("spongexyz", DP), ("p0", DP), ("g", DP),
        ("box", DP), ("boy", DP), ("boz", DP), ("curlbox", DP), ("curlboy", DP), ("curlboz", DP),
        ("gradrho0_x", DP), ("gradrho0_y", DP), ("gradp0_x", DP), ("gradp0_y", DP),
        ("reduction", DP),
        ("pml_a", DP * 6), ("pml_b", DP * 6), ("pml_kappa", DP * 6),
        ("src_x", DP), ("src_y", DP), ("src_z", DP),
        ("dimc", C.c_double), ("dimrho", C.c_double),
    ]


EXPORTS = [
    "sparc_hip_create", "sparc_hip_destroy", "sparc_hip_set_background", "sparc_hip_set_damping",
    "sparc_hip_set_forcing",
    "sparc_hip_put_state",
    "sparc_hip_get_state", "sparc_hip_step", "sparc_hip_filter", "sparc_hip_advance",
    "sparc_hip_diag", "sparc_hip_ddx", "sparc_hip_ddy", "sparc_hip_ddz", "sparc_hip_filtx",
    "sparc_hip_filty", "sparc_hip_filtz", "sparc_hip_sync", "sparc_hip_timing",
    "sparc_hip_kernel_time", "sparc_hip_tile", "sparc_hip_strerror", "sparc_hip_last_error",
    "sparc_hip_version", "sparc_hip_batch_create", "sparc_hip_batch_advance",
    "sparc_hip_batch_info", "sparc_hip_batch_destroy", "sparc_hip_alfven_search",
    "sparc_hip_get_slab", "sparc_hip_autotune", "sparc_hip_set_tile", "sparc_hip_batch_time_stage",
    "sparc_hip_set_guard", "sparc_hip_guard", "sparc_hip_batch_filter", "sparc_hip_stream_bench",
    "sparc_hip_deriv_noalloc", "sparc_hip_device_memory",
]

_lib = None


class SparcError(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SparcError(
            f"{LIB_PATH} is missing: build it with `python -m sparc2_amd.build` "
            "(the time-step path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i = C.c_int
    L.sparc_hip_create.argtypes = [C.POINTER(SparcCfg), C.POINTER(vp)]
    L.sparc_hip_destroy.argtypes = [vp]
    L.sparc_hip_set_background.argtypes = [vp, C.POINTER(SparcBg)]
    L.sparc_hip_set_damping.argtypes = [vp, DP]
    L.sparc_hip_set_forcing.argtypes = [vp, DP, DP]
    L.sparc_hip_put_state.argtypes = [vp, DP, DP * 6]
    L.sparc_hip_get_state.argtypes = [vp, DP, DP * 6]
    L.sparc_hip_step.argtypes = [vp, C.c_double]
    L.sparc_hip_filter.argtypes = [vp, i]
    L.sparc_hip_advance.argtypes = [vp, i, i, DP]
    L.sparc_hip_diag.argtypes = [vp, DP]
    for n in ("sparc_hip_ddx", "sparc_hip_ddy", "sparc_hip_ddz"):
        getattr(L, n).argtypes = [vp, DP, DP, i]
    for n in ("sparc_hip_filtx", "sparc_hip_filty", "sparc_hip_filtz"):
        getattr(L, n).argtypes = [vp, DP, DP]
    L.sparc_hip_sync.argtypes = [vp]
    L.sparc_hip_timing.argtypes = [vp, i]
    L.sparc_hip_kernel_time.argtypes = [vp, i, C.POINTER(C.c_long), C.POINTER(C.c_double)]
    L.sparc_hip_tile.argtypes = [vp, C.c_int * 3, C.POINTER(C.c_long)]
    L.sparc_hip_batch_create.argtypes = [C.POINTER(vp), i, i, C.POINTER(vp)]
    L.sparc_hip_batch_advance.argtypes = [vp, i, C.POINTER(C.c_int), DP, i]
    L.sparc_hip_batch_filter.argtypes = [vp, C.POINTER(C.c_int)]
    L.sparc_hip_set_guard.argtypes = [vp, i]
    L.sparc_hip_stream_bench.argtypes = [i, C.c_long, i, i, DP]
    L.sparc_hip_deriv_noalloc.argtypes = [vp, i, DP, DP, i]
    L.sparc_hip_device_memory.argtypes = [i, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.sparc_hip_guard.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_long), C.POINTER(C.c_long), DP]
    L.sparc_hip_batch_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_long)]
    L.sparc_hip_batch_destroy.argtypes = [vp]
    L.sparc_hip_alfven_search.argtypes = [vp, DP, DP, DP, C.c_double, C.c_double, C.c_double,
                                          C.POINTER(C.c_double)]
    L.sparc_hip_get_slab.argtypes = [vp, i, i, i, DP]
    L.sparc_hip_strerror.restype = C.c_char_p
    L.sparc_hip_strerror.argtypes = [i]
    L.sparc_hip_last_error.restype = C.c_char_p
    L.sparc_hip_version.restype = i
    for n in EXPORTS:
        f = getattr(L, n)
        if f.restype is C.c_int or f.restype is None:
            f.restype = C.c_int
    L.sparc_hip_strerror.restype = C.c_char_p
    L.sparc_hip_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != SPARC_OK:
        L = lib()
        raise SparcError(f"{what}: {L.sparc_hip_strerror(rc).decode()} "
                         f"({L.sparc_hip_last_error().decode()})")


def dptr(a) -> DP:
    return a.ctypes.data_as(DP) if a is not None else DP()
