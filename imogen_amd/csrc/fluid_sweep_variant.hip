// One (precision, block size, KMAX) variant of the fused fluid sweep; the build compiles
// this file once per entry of IMH_FLUID_VARIANTS with -DIMH_FLUID_ESZ=<4|8>
// -DIMH_FLUID_BS=<256|512> -DIMH_FLUID_KMAX=<k>.
//
// FP32 chunks of at most 4 cells fit in 128 registers without spilling: those variants
// ask the register allocator for 4 waves per SIMD (the allocator otherwise spends the
// 256 registers a 512-thread block allows, leaving 2).
#if !defined(IMH_FLUID_WAVES_PER_EU) && IMH_FLUID_ESZ == 4 && IMH_FLUID_KMAX <= 4
#define IMH_FLUID_WAVES_PER_EU 4
#endif
#include "fluid_sweep_impl.h"
