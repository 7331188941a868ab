// FP32-bulk instantiations of the fused fluid sweep.
#define IMH_FLUID_T float
#define IMH_FLUID_VARIANTS IMH_CASE(256, 2) IMH_CASE(256, 4) IMH_CASE(512, 4) IMH_CASE(512, 8)
#include "fluid_sweep_impl.h"
