// FP64 instantiations of the fused fluid sweep (the plain path).
#define IMH_FLUID_T double
#define IMH_FLUID_VARIANTS IMH_CASE(256, 1) IMH_CASE(256, 2) IMH_CASE(512, 2) IMH_CASE(512, 4)
#include "fluid_sweep_impl.h"
