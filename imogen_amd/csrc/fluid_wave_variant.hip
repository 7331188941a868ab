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
    const void* args, const BatchTile* tiles, int ntiles, int nblocks, bool mhd, size_t lds,
    hipStream_t st) {
    const FluidArgs<WaveT>* a = (const FluidArgs<WaveT>*)args;
    if (mhd)
        hipLaunchKernelGGL((fluid_wave_kernel<WaveT, IMH_WAVE_BS, true, IMH_WAVE_FAST != 0>),
                           dim3(nblocks), dim3(IMH_WAVE_BS), lds, st, a, tiles, ntiles);
    else
        hipLaunchKernelGGL((fluid_wave_kernel<WaveT, IMH_WAVE_BS, false, IMH_WAVE_FAST != 0>),
                           dim3(nblocks), dim3(IMH_WAVE_BS), lds, st, a, tiles, ntiles);
    return hipGetLastError();
}

}  // namespace imh
