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

template <typename T>
hipError_t launch_fluid_wave(const FluidArgs<T>* args, const BatchTile* tiles, int ntiles,
                             int nblocks, int bs, bool fast, int d, bool mhd, size_t lds,
                             hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    constexpr int esz = (int)sizeof(T);
    const int f = (esz == 8 && fast) ? 1 : 0;
#define IMH_WAVE_PICK(E, B, F)                                                           \
    if (esz == E && bs == B && f == F)                                                   \
        return IMH_WAVE_LAUNCHER(E, B, F)(args, tiles, ntiles, nblocks, d, mhd, lds, st);
    IMH_WAVE_VARIANTS(IMH_WAVE_PICK)
#undef IMH_WAVE_PICK
    return hipErrorInvalidValue;
}
template hipError_t launch_fluid_wave<double>(const FluidArgs<double>*, const BatchTile*, int,
                                              int, int, bool, int, bool, size_t, hipStream_t);
template hipError_t launch_fluid_wave<float>(const FluidArgs<float>*, const BatchTile*, int, int,
                                             int, bool, int, bool, size_t, hipStream_t);

// Block size of y / z tiles of the wave-spread sweep (more threads: more x-adjacent
// pencils per tile, longer contiguous row segments); a build-time knob.
#ifndef IMH_WAVE_YZ_BS
#define IMH_WAVE_YZ_BS 512
#endif

bool wave_tile(int N, int d, int esz, size_t max_lds, int* bs, int* px, int* tl, size_t* lds) {
    if (N % 4 != 0) return false;
    const int TL = N / 4;
    if (TL < 2 || TL > 512 || (TL & (TL - 1)) != 0) return false;
    const bool xm = d == 0;
    // y / z: the larger block where it gives at most 8 pencils per tile (64-byte FP64 row
    // segments; more pencils put most lanes at wave edges, whose halos go through LDS)
    const int b = (!xm && TL <= IMH_WAVE_YZ_BS && IMH_WAVE_YZ_BS / TL <= 8) ? IMH_WAVE_YZ_BS
                                                                             : (TL <= 256 ? 256 : 512);
    const int p = b / TL;
    if (p > 32 || (!xm && b / 64 < 2)) return false;
    const int dl = xm ? 1 : p;
    const bool multi = !xm || TL > 64;
    const size_t bytes = (size_t)wave_lds_layout(b / 64, p, dl, multi ? 1 : 0, esz).total;
    if (bytes > max_lds) return false;  // the device's LDS per workgroup
    *bs = b;
    *px = p;
    *tl = TL;
    *lds = bytes;
    return true;
}

bool fluid_variant_compiled(int esz, int bs, int kmax) {
#define IMH_FLUID_HAS(E, B, K) \
    if (esz == E && bs == B && kmax == K) return true;
    IMH_FLUID_VARIANTS(IMH_FLUID_HAS)
#undef IMH_FLUID_HAS
    return false;
}

}  // namespace imh
