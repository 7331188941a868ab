// One (precision, KMAX) variant of the magnet advection and CT kernels; the build
// compiles this file once per entry of IMH_MAG_VARIANTS with -DIMH_MAG_ESZ=<4|8>
// -DIMH_MAG_KMAX=<k>.
#include "magnet_sweep_impl.h"
