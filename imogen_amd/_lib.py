"""ctypes binding of libimogen_hip.so (the C ABI declared in include/imogen_hip.h).

The shared library is built in-tree by ``imogen_amd/build.py``.  There is no fallback:
if it is missing or a call fails, an exception is raised.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# IMH_LIB_PATH selects another build of the library (one made with build.py --out)
LIB_PATH = os.environ.get("IMH_LIB_PATH") or os.path.join(_HERE, "libimogen_hip.so")

IMH_OK = 0
STATUS = {-1: "IMH_EINVAL", -2: "IMH_EHIP", -3: "IMH_ENOMEM", -4: "IMH_ECFL",
          -5: "IMH_ESTATE", -6: "IMH_ENOTSUP"}
FP64, FP32_BULK = 0, 1

FIELD_IDS = {"mass": 0, "momx": 1, "momy": 2, "momz": 3, "ener": 4, "magx": 5, "magy": 6,
             "magz": 7, "cmagx": 8, "cmagy": 9, "cmagz": 10}


class ImogenError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS.get(status, str(status)), message))
        self.status = status


class CFLError(ImogenError):
    """TimeManager.update found no signal speed above 1e-2 (TimeManager.m:82-85)."""


class ImhConfig(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_int * 3),
        ("dgrid", ctypes.c_double * 3),
        ("gamma", ctypes.c_double),
        ("minmass", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("bc_infinity", ctypes.c_int),
        ("bc", (ctypes.c_int * 3) * 2),
        ("limiter", ctypes.c_int * 3),
        ("fluid_active", ctypes.c_int),
        ("magnet_active", ctypes.c_int),
        ("update_mode", ctypes.c_int),
        ("iter_max", ctypes.c_int),
        ("time_max", ctypes.c_double),
        ("perms_order", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


_lib = None

# Every entry point of include/imogen_hip.h with its argument types.
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int
SIGNATURES = {
    "imh_last_error": (ctypes.c_char_p, []),
    "imh_version": (ctypes.c_char_p, []),
    "imh_create": (_I, [ctypes.POINTER(ImhConfig), ctypes.POINTER(_P)]),
    "imh_destroy": (_I, [_P]),
    "imh_upload_state": (_I, [_P, _D, _D * 3, _D, _D * 3, ctypes.c_size_t]),
    "imh_upload_field": (_I, [_P, _I, _D, ctypes.c_size_t]),
    "imh_download_field": (_I, [_P, _I, _D, ctypes.c_size_t]),
    "imh_fluid_sweep": (_I, [_P, _I]),
    "imh_magnet_sweep": (_I, [_P, _I, _I * 2]),
    "imh_time_update": (_I, [_P, _I]),
    "imh_flux": (_I, [_P, _I]),
    "imh_iteration": (_I, [_P]),
    "imh_run": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "imh_set_dtime": (_I, [_P, ctypes.c_double]),
    "imh_get_time": (_I, [_P, _D, _D, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "imh_set_iteration": (_I, [_P, _I]),
    "imh_synchronize": (_I, [_P]),
    "imh_conservation_sums": (_I, [_P, _D]),
    "imh_conservation_sums_noalloc": (_I, [ctypes.POINTER(_P), _I, _I, ctypes.c_size_t, _P, _P]),
    "imh_memory_bytes": (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "imh_memory_estimate": (_I, [ctypes.POINTER(ImhConfig), ctypes.POINTER(ctypes.c_size_t)]),
    "imh_init_problem": (_I, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "imh_source": (_I, [_P]),
    "imh_set_gravity": (_I, [_P, _D, ctypes.c_size_t, ctypes.c_double]),
    "imh_set_spacing": (_I, [_P, _I, _D, ctypes.c_size_t]),
    "imh_set_radiation": (_I, [_P, _I, ctypes.c_double, ctypes.c_double]),
    "imh_get_radiation_strength": (_I, [_P, _D]),
    "imh_set_viscosity": (_I, [_P, _I, ctypes.c_double, ctypes.c_double]),
    "imh_batch_set_gravity": (_I, [_P, _I, _D, ctypes.c_size_t, ctypes.c_double]),
    "imh_batch_set_spacing": (_I, [_P, _I, _I, _D, ctypes.c_size_t]),
    "imh_batch_set_radiation": (_I, [_P, _I, _I, ctypes.c_double, ctypes.c_double]),
    "imh_batch_set_viscosity": (_I, [_P, _I, _I, ctypes.c_double, ctypes.c_double]),
    "imh_checkpoint_bytes": (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "imh_checkpoint_save": (_I, [_P, ctypes.c_void_p, ctypes.c_size_t]),
    "imh_checkpoint_load": (_I, [_P, ctypes.c_void_p, ctypes.c_size_t]),
    "imh_batch_create": (_I, [ctypes.POINTER(ImhConfig), _I, ctypes.POINTER(_P)]),
    "imh_batch_destroy": (_I, [_P]),
    "imh_batch_upload_state": (_I, [_P, _I, _D, _D * 3, _D, _D * 3, ctypes.c_size_t]),
    "imh_batch_download_field": (_I, [_P, _I, _I, _D, ctypes.c_size_t]),
    "imh_batch_create_view": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(_P)]),
    "imh_batch_step": (_I, [_P, _I]),
    "imh_batch_get_time": (_I, [_P, _I, _D, _D, ctypes.POINTER(_I)]),
    "imh_batch_synchronize": (_I, [_P]),
    "imh_batch_memory_bytes": (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "imh_serve_create": (_I, [_I, _I, ctypes.c_size_t, _I, ctypes.POINTER(_P)]),
    "imh_serve_destroy": (_I, [_P]),
    "imh_serve_submit": (_I, [_P, ctypes.POINTER(ImhConfig), _D, _D * 3, _D, _D * 3,
                              ctypes.c_size_t, ctypes.POINTER(ctypes.c_longlong)]),
    "imh_serve_submit_problem": (_I, [_P, ctypes.POINTER(ImhConfig), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_longlong)]),
    "imh_serve_step": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "imh_serve_status": (_I, [_P, ctypes.c_longlong, ctypes.POINTER(_I), _D,
                              ctypes.POINTER(_I), _D]),
    "imh_serve_download": (_I, [_P, ctypes.c_longlong, _I, _D, ctypes.c_size_t]),
    "imh_serve_release": (_I, [_P, ctypes.c_longlong]),
    "imh_serve_stats": (_I, [_P, ctypes.POINTER(ctypes.c_size_t),
                             ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_longlong)]),
}

# Tuning and measurement entry points (include/imogen_hip_ext.h; not part of the drop-in
# boundary).
NKIND = 8
KERNEL_KINDS = ["cfl", "fluid", "mag_adv", "mag_ct", "cellmag", "bookkeeping", "source",
                "mhd_1d"]
EXTRA = {
    "imh_set_fluid_tile": (_I, [_P, _I, _I, _I, _I]),
    "imh_get_fluid_tile": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                ctypes.POINTER(_I)]),
    "imh_set_graphs": (_I, [_P, _I]),
    "imh_set_fluid_kernel": (_I, [_P, _I]),
    "imh_get_fluid_kernel": (_I, [_P, ctypes.POINTER(_I)]),
    "imh_batch_set_fluid_tile": (_I, [_P, _I, _I, _I, _I, _I]),
    "imh_batch_set_fused_1d": (_I, [_P, _I]),
    "imh_profile_iterations": (_I, [_P, _I, _D, _D, ctypes.POINTER(_I)]),
    "imh_batch_profile_step": (_I, [_P, _I, _D, _D, ctypes.POINTER(_I)]),
    "imh_copy_probe": (_I, [ctypes.c_size_t, _I, _D]),
}


def run_profile(fn, handle, iterations):
    """Call one of the profiling entry points; returns {kind: (ms, bytes, launches)}."""
    ms = (ctypes.c_double * NKIND)()
    by = (ctypes.c_double * NKIND)()
    cnt = (ctypes.c_int * NKIND)()
    check(fn(handle, int(iterations), ms, by, cnt))
    return {KERNEL_KINDS[k]: (ms[k], by[k], cnt[k]) for k in range(NKIND)}


def lib():
    """Load the in-tree libimogen_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("imogen_amd: %s is missing; build it with "
                              "`python -m imogen_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in list(SIGNATURES.items()) + list(EXTRA.items()):
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != IMH_OK:
        msg = lib().imh_last_error().decode(errors="replace")
        if rc == -4:
            raise CFLError(rc, msg)
        raise ImogenError(rc, msg)
    return rc


def dptr(a):
    return a.ctypes.data_as(_D)


def as_field(a, ncell):
    """A host array in the reference layout (x fastest) as contiguous float64."""
    arr = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if arr.size != ncell:
        raise ValueError("array has %d elements, the grid has %d cells" % (arr.size, ncell))
    return arr
