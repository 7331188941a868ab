"""Run settings: the reference's ``ini`` struct parsed the way ``initialize.m`` does.

Only the settings that reach the relaxing-TVD path are parsed (``initialize.m:24-190``):
bcMode, bcInfinity, fluxLimiter, cfl, timeUpdateMode, minMass, iterMax, timeMax, mode,
dGrid, gamma and grid.  Errors are raised with the reference's message text.
"""
from dataclasses import dataclass, field
from typing import List, Tuple


class ENUM:
    """enums/ENUM.m (the subset used on this path)."""
    MASS = "mass"
    MOM = "mom"
    ENER = "ener"
    MAG = "mag"
    BCMODE_CIRCULAR = "circ"
    BCMODE_CONST = "const"
    BCMODE_FADE = "fade"
    BCMODE_FLIP = "flip"
    BCMODE_LINEAR = "linear"
    BCMODE_MIRROR = "mirror"
    BCMODE_TRANSPARENT = "trans"
    BCMODE_WALL = "wall"
    BCMODE_WORMHOLE = "wormhole"
    BCMODE_ZERO = "zero"
    TIMEUPDATE_PER_STEP = 0
    TIMEUPDATE_PER_ITERATION = 1
    PRESSURE_TOTAL_AND_SOUND = "totsnd"
    PRESSURE_SOUND_SPEED = "sound"


class FluxLimiterEnum:
    """enums/FluxLimiterEnum.m -- including the reference's 'minmond' spelling."""
    VAN_LEER = "vanleer"
    SUPERBEE = "superbee"
    MINMOD = "minmond"


# Mode codes of the C ABI (include/imogen_hip.h).
BC_CODES = {"circ": 0, "const": 1, "trans": 2, "fade": 3, "mirror": 4, "flip": 5,
            "wall": 6, "zero": 7, "wormhole": 8}
LIMITER_CODES = {FluxLimiterEnum.VAN_LEER: 0, FluxLimiterEnum.SUPERBEE: 1,
                 FluxLimiterEnum.MINMOD: 2}


def parse_flux_limiter(name):
    """FluidManager.parseFluxLimiterEnum (fluid/FluidManager.m:79-87)."""
    if name not in LIMITER_CODES:
        raise ValueError("Imogen:UnknownType: Unknown flux limiter type: %s" % name)
    return name


@dataclass
class RunConfig:
    grid: Tuple[int, int, int]
    dgrid: Tuple[float, float, float]
    gamma: float = 5.0 / 3.0
    minmass: float = 1e-5
    cfl: float = 0.7
    bc_infinity: int = 20
    bc: List[List[str]] = field(default_factory=lambda: [["circ"] * 3, ["circ"] * 3])
    limiter: List[str] = field(default_factory=lambda: [FluxLimiterEnum.VAN_LEER] * 3)
    fluid_active: bool = True
    magnet_active: bool = False
    update_mode: int = 1
    iter_max: int = 1000
    time_max: float = 5000.0
    perms_order: int = 0

    @property
    def ndims(self):
        return 3 if self.grid[2] > 1 else 2

    @property
    def ncell(self):
        return self.grid[0] * self.grid[1] * self.grid[2]

    @property
    def shape(self):
        """numpy shape (nz, ny, nx) of one field."""
        return (self.grid[2], self.grid[1], self.grid[0])


def parse_bc_modes(bc):
    """BCManager.populateBCModes (shifting/BCManager.m:97-133) for the global modes:
    a string applies to every side; a struct/dict gives x, y, z, each a mode or a cell
    {lower, ..., upper}."""
    if isinstance(bc, str):
        modes = {"x": bc, "y": bc, "z": bc}
    elif isinstance(bc, dict):
        modes = bc
    elif isinstance(bc, (list, tuple)) and len(bc) == 2:
        return [list(bc[0]), list(bc[1])]
    else:
        raise ValueError("BoundaryConditionError: Boundary condition field of type %s is "
                         "not recognized. bcMode recognizes string or structure input. "
                         "Run aborted." % type(bc).__name__)
    out = [[None] * 3, [None] * 3]
    for d, key in enumerate("xyz"):
        v = modes.get(key, "circ")
        if isinstance(v, (list, tuple)):
            out[0][d], out[1][d] = v[0], v[-1]
        else:
            out[0][d] = out[1][d] = v
    for s in range(2):
        for d in range(3):
            if out[s][d] not in BC_CODES:
                raise ValueError("Imogen:UnknownType: Unknown BC mode %s" % out[s][d])
    return out


def from_ini(ini, perms_order=0):
    """Parse an ini dict (keys as produced by ``Initializer.getRunSettings``)."""
    grid = tuple(int(g) for g in ini["grid"])
    if len(grid) != 3:
        raise ValueError("ini.grid must have 3 entries")
    dg = ini.get("dGrid")
    if dg is None:
        dgrid = (1.0, 1.0, 1.0)
    elif isinstance(dg, (int, float)):
        dgrid = (float(dg),) * 3
    else:
        dg = [float(v) for v in dg]
        if len(dg) == 2:
            dg = dg + [max(dg)]
        dgrid = tuple(dg)
    lim = ini.get("fluxLimiter", FluxLimiterEnum.VAN_LEER)
    if isinstance(lim, str):
        lims = [lim] * 3
    elif isinstance(lim, dict):
        lims = [lim.get(k, FluxLimiterEnum.VAN_LEER) for k in "xyz"]
    else:
        lims = list(lim)
    lims = [parse_flux_limiter(l) for l in lims]
    mode = ini.get("mode", {"fluid": True, "magnet": False})
    cfl = ini.get("cfl")
    if cfl is None:
        cfl = 0.35 if mode.get("magnet", False) else 0.7  # Initializer.m:143-150
    return RunConfig(
        grid=grid, dgrid=dgrid, gamma=float(ini.get("gamma", 5.0 / 3.0)),
        minmass=float(ini.get("minMass", 1e-5)), cfl=float(cfl),
        bc_infinity=int(ini.get("bcInfinity", 20)),
        bc=parse_bc_modes(ini.get("bcMode", "circ")), limiter=lims,
        fluid_active=bool(mode.get("fluid", True)), magnet_active=bool(mode.get("magnet", False)),
        update_mode=int(ini.get("timeUpdateMode", 1)), iter_max=int(ini.get("iterMax", 1000)),
        time_max=float(ini.get("timeMax", 5000.0)), perms_order=int(perms_order))
