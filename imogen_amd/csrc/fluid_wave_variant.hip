// One (precision, block size, arithmetic) variant of the wave-spread fluid sweep
// (fluid_wave_impl.h).  The build compiles this file once per entry of
// IMH_WAVE_VARIANTS with -DIMH_WAVE_ESZ=<4|8> -DIMH_WAVE_BS=<256|512> -DIMH_WAVE_FAST=<0|1>;
// exact FP64 units keep floating-point contraction off, the others contract.
#include "fluid_wave_impl.h"

namespace imh {

#if IMH_WAVE_ESZ == 8
using WaveT = double;
#else
using WaveT = float;
#endif

hipError_t IMH_WAVE_LAUNCHER(IMH_WAVE_ESZ, IMH_WAVE_BS, IMH_WAVE_FAST)(
    const void* args, const BatchTile* tiles, int ntiles, int nblocks, int d, bool mhd,
    size_t lds, hipStream_t st) {
    const FluidArgs<WaveT>* a = (const FluidArgs<WaveT>*)args;
    constexpr bool F = IMH_WAVE_FAST != 0;
#define IMH_WL(M, DD)                                                                       \
    if (mhd == M && d == DD) {                                                              \
        hipLaunchKernelGGL((fluid_wave_kernel<WaveT, IMH_WAVE_BS, M, F, DD>), dim3(nblocks),  \
                           dim3(IMH_WAVE_BS), lds, st, a, tiles, ntiles);                   \
        return hipGetLastError();                                                           \
    }
    IMH_WL(false, 0) IMH_WL(false, 1) IMH_WL(false, 2) IMH_WL(true, 0) IMH_WL(true, 1)
    IMH_WL(true, 2)
#undef IMH_WL
    return hipErrorInvalidValue;
}

}  // namespace imh
