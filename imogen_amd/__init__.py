"""imogen_amd -- MI355X (gfx950) implementation of Imogen's relaxing-TVD fluid/magnet
update behind the reference's FluidManager/MagnetManager call surface.

The hot path is hand-written HIP (imogen_amd/csrc) behind a C ABI
(include/imogen_hip.h); this package holds the ctypes binding and the host-side mirror
of the reference's functions.
"""
from .config import ENUM, FluxLimiterEnum, RunConfig, from_ini  # noqa: F401
from ._lib import CFLError, ImogenError  # noqa: F401
from .imogen import (FluidArray, ImogenManager, MagnetArray, create_arrays, flux,  # noqa: F401
                     imogen, initialize, iteration, magnetFlux, relaxingFluid, source)

__version__ = "0.1.0"
