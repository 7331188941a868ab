"""The allocation-free conservation sums (imh_conservation_sums_noalloc) against
imh_conservation_sums, through the C ABI.

The fields are downloaded from a context, copied into device buffers that the test owns
(allocated through the same HIP runtime the library links), and summed by the
allocation-free entry point into a device output; the result must equal the context's
sums bit for bit, on two calls in a row, the second over a smaller input (the scratch
array keeps the first call's partials), in FP64 and in FP32 bulk."""
import ctypes

import numpy as np
import pytest

import imogen_amd
from imogen_amd import _lib, problems
from imogen_amd._lib import FIELD_IDS

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
SUM_FIELDS = ["mass", "momx", "momy", "momz", "ener", "magx", "magy", "magz"]


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    for f in ("hipMalloc", "hipMemcpy", "hipFree", "hipDeviceSynchronize"):
        getattr(h, f).restype = ctypes.c_int
    return h


def hchk(rc, what):
    assert rc == 0, "%s: hip error %d" % (what, rc)


def context_sums(ini, st, precision, iterations):
    run = imogen_amd.initialize(ini, precision)
    imogen_amd.create_arrays(run, st)
    run.run_iterations(iterations)
    ref = run.conservation_sums()
    fields = [run.download(f) for f in SUM_FIELDS]
    run.close()
    return ref, fields


def noalloc_sums(fields, precision):
    h = hip()
    dt = np.float64 if precision == "fp64" else np.float32
    host = [np.ascontiguousarray(f.reshape(-1), dtype=dt) for f in fields]
    n = host[0].size
    bufs = []
    out = ctypes.c_void_p()
    try:
        for a in host:
            p = ctypes.c_void_p()
            hchk(h.hipMalloc(ctypes.byref(p), a.nbytes), "hipMalloc")
            bufs.append(p)
            hchk(h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        hchk(h.hipMalloc(ctypes.byref(out), 8 * 8), "hipMalloc out")
        ptrs = (ctypes.c_void_p * len(bufs))(*[b.value for b in bufs])
        prec = _lib.FP64 if precision == "fp64" else _lib.FP32_BULK
        _lib.check(_lib.lib().imh_conservation_sums_noalloc(ptrs, len(bufs), prec, n, out, None))
        hchk(h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        res = np.empty(len(bufs), dtype=np.float64)
        hchk(h.hipMemcpy(res.ctypes.data, out, res.nbytes, D2H), "hipMemcpy D2H")
    finally:
        for b in bufs:
            h.hipFree(b)
        if out.value:
            h.hipFree(out)
    return res


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_sums_equal_the_context_sums(precision):
    # first call: Orszag-Tang 96^2 (MHD, 8 fields); second: Sod 512 (smaller, B = 0)
    for name, (ini, st), iters in [("orszag-tang 96^2", problems.orszag_tang(96), 4),
                                   ("sod 512", problems.sod(512), 6)]:
        print("noalloc sums: %s %s" % (name, precision), flush=True)
        ref, fields = context_sums(ini, st, precision, iters)
        got = noalloc_sums(fields, precision)
        assert np.array_equal(got, ref), (name, got, ref)


def test_noalloc_refuses_too_many_fields():
    L = _lib.lib()
    ptrs = (ctypes.c_void_p * 9)(*([1] * 9))
    rc = L.imh_conservation_sums_noalloc(ptrs, 9, _lib.FP64, 16, ctypes.c_void_p(1), None)
    assert rc == -1  # IMH_EINVAL
    assert b"1..8 fields" in L.imh_last_error()
