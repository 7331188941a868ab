// One (precision, block size, KMAX) variant of the fused fluid sweep; the build compiles
// this file once per entry of IMH_FLUID_VARIANTS with -DIMH_FLUID_ESZ=<4|8>
// -DIMH_FLUID_BS=<256|512> -DIMH_FLUID_KMAX=<k>.
#include "fluid_sweep_impl.h"
