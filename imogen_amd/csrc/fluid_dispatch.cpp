// fluid_dispatch.cpp -- picks the compiled fluid sweep variant for a launch.
#include "imh_kernels.h"

namespace imh {

template <typename T>
hipError_t launch_fluid_batch(const FluidArgs<T>* args, const BatchTile* tiles, int ntiles,
                              int nblocks, int bs, int kmax, bool mhd, bool general,
                              hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    constexpr int esz = (int)sizeof(T);
#define IMH_FLUID_PICK(E, B, K)                                                          \
    if (esz == E && bs == B && kmax == K)                                                \
        return IMH_FLUID_LAUNCHER(E, B, K)(args, tiles, ntiles, nblocks, mhd, general, st);
    IMH_FLUID_VARIANTS(IMH_FLUID_PICK)
#undef IMH_FLUID_PICK
    return hipErrorInvalidValue;
}

template hipError_t launch_fluid_batch<double>(const FluidArgs<double>*, const BatchTile*, int,
                                               int, int, int, bool, bool, hipStream_t);
template hipError_t launch_fluid_batch<float>(const FluidArgs<float>*, const BatchTile*, int, int,
                                              int, int, bool, bool, hipStream_t);

bool fluid_variant_compiled(int esz, int bs, int kmax) {
#define IMH_FLUID_HAS(E, B, K) \
    if (esz == E && bs == B && kmax == K) return true;
    IMH_FLUID_VARIANTS(IMH_FLUID_HAS)
#undef IMH_FLUID_HAS
    return false;
}

}  // namespace imh
