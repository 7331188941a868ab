// noalloc.hip -- conservation totals over caller-owned device arrays, with no allocation.
//
// imh_conservation_sums_noalloc is the allocation-free form of imh_conservation_sums: the
// same kernels (launch_conservation, cfl.hip) and the same fixed reduction order, so both
// give the same bits, but the fields and the output are device memory the caller owns,
// the work is queued on the caller's stream without a wait, and the block partials live
// in `scratch`, a file-scope device array (one per GPU, held until the process ends)
// instead of a context's allocation.  For callers that cannot allocate device memory:
// code under stream capture, tight serving loops, harnesses that forbid hipMalloc.
//
// Host code never passes `scratch` itself to a kernel (in host code the name is a
// host-side stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device on that device's first call, and only that
// address is used afterwards.  Every call on one GPU shares its scratch: calls that
// overlap on one GPU (two streams, two threads) race and must be ordered by the caller.
// Nothing in scratch has to start at zero: every partial a call reads is written by the
// same call first.
#include <mutex>

#include "imh_host.h"
#include "imh_kernels.h"

namespace imh {
namespace {

constexpr int NA_BLOCKS = 1024;  // partial blocks per field, as imh_conservation_sums uses
constexpr int NA_MAX_FIELDS = IMH_NOALLOC_MAX_FIELDS;
constexpr size_t NA_SCRATCH_DOUBLES = size_t(NA_MAX_FIELDS) * 2 * NA_BLOCKS;  // 128 KiB

__device__ __align__(256) double scratch[NA_SCRATCH_DOUBLES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
double* na_ptr[NA_MAX_DEVICES] = {};

// Device address of this device's scratch (looked up on the device's first call).
hipError_t scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NA_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        na_ptr[dev] = static_cast<double*>(p);
    }
    *out = na_ptr[dev];
    return hipSuccess;
}

}  // namespace
}  // namespace imh

extern "C" int imh_conservation_sums_noalloc(const void* const* fields, int nfields,
                                             int precision, size_t ncell, double* out,
                                             void* stream) {
    using namespace imh;
    if (!fields || !out) return fail(IMH_EINVAL, "null argument");
    if (nfields < 1 || nfields > NA_MAX_FIELDS)
        return fail(IMH_EINVAL,
                    "nfields = %d: the allocation-free sums take 1..%d fields (their scratch "
                    "holds %d x %d block partials)",
                    nfields, NA_MAX_FIELDS, NA_MAX_FIELDS, NA_BLOCKS);
    if (ncell < 1 || ncell > (size_t(1) << 40)) return fail(IMH_EINVAL, "ncell out of range");
    if (precision != IMH_FP64 && precision != IMH_FP32_BULK)
        return fail(IMH_EINVAL, "precision must be IMH_FP64 or IMH_FP32_BULK");
    for (int i = 0; i < nfields; ++i)
        if (!fields[i]) return fail(IMH_EINVAL, "fields[%d] is null", i);
    double* part = nullptr;
    hipError_t e = scratch_ptr(&part);
    if (e != hipSuccess) return hip_fail(e, "scratch address (hipGetSymbolAddress)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == IMH_FP64)
        e = launch_conservation<double>(reinterpret_cast<const double* const*>(fields), nfields,
                                        (long long)ncell, part, NA_BLOCKS, out, st);
    else
        e = launch_conservation<float>(reinterpret_cast<const float* const*>(fields), nfields,
                                       (long long)ncell, part, NA_BLOCKS, out, st);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "allocation-free conservation sums");
}
