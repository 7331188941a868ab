"""Run settings: the reference's ``ini`` struct parsed the way ``initialize.m`` does.

Only the settings that reach the relaxing-TVD path are parsed (``initialize.m:24-190``):
bcMode, bcInfinity, fluxLimiter, cfl, timeUpdateMode, minMass, iterMax, timeMax, mode,
dGrid, gamma and grid.  Errors are raised with the reference's message text.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


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
    # non-uniform DGRID{i}: a line of grid[i] spacings (dgrid[i] then holds its minimum,
    # MINDGRID), or None for the uniform dgrid[i]
    dgrid_lines: List[Optional[np.ndarray]] = field(default_factory=lambda: [None] * 3)

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


def pchip(x, y, xi):
    """MATLAB's interp1(x, y, xi, 'pchip'): piecewise cubic Hermite with the
    Fritsch-Carlson slopes (weighted harmonic mean of the neighbouring secants, zero at a
    sign change) and the three-point, shape-preserving end slopes; two points give the
    line; outside [x(1), x(end)] the end pieces extrapolate."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xi = np.asarray(xi, dtype=np.float64)
    n = x.size
    if n < 2 or np.any(np.diff(x) <= 0):
        raise ValueError("pchip needs at least two strictly increasing sample points")
    h = np.diff(x)
    delta = np.diff(y) / h
    d = np.zeros(n)
    if n == 2:
        d[:] = delta[0]
    else:
        for k in range(1, n - 1):
            if np.sign(delta[k - 1]) * np.sign(delta[k]) > 0:
                w1 = 2.0 * h[k] + h[k - 1]
                w2 = h[k] + 2.0 * h[k - 1]
                d[k] = (w1 + w2) / (w1 / delta[k - 1] + w2 / delta[k])

        def end(h1, h2, del1, del2):
            s = ((2.0 * h1 + h2) * del1 - h1 * del2) / (h1 + h2)
            if np.sign(s) != np.sign(del1):
                return 0.0
            if np.sign(del1) != np.sign(del2) and abs(s) > abs(3.0 * del1):
                return 3.0 * del1
            return s

        d[0] = end(h[0], h[1], delta[0], delta[1])
        d[-1] = end(h[-1], h[-2], delta[-1], delta[-2])
    k = np.clip(np.searchsorted(x, xi, side="right") - 1, 0, n - 2)
    t = xi - x[k]
    c2 = (3.0 * delta[k] - 2.0 * d[k] - d[k + 1]) / h[k]
    c3 = (d[k] - 2.0 * delta[k] + d[k + 1]) / (h[k] * h[k])
    return y[k] + t * (d[k] + t * (c2 + t * c3))


def grid_spacing(dg, grid):
    """ImogenManager.setGridSpacing (classes/ImogenManager.m:190-254).  A number or a
    3-vector is uniform.  A dict (the struct form) takes, per field x/y/z: a number (that
    spacing), or a dict with 'value' (else the top-level 'value', else 1) and 'points'
    (rows [percent, multiplier], default [[0, 1], [100, 1]]): the percentages map to grid
    points floor((N-1)*p/100)+1 and DGRID is interp1(points, value*multipliers, 1:N,
    'pchip'); a missing field takes the top-level 'value'.  Returns (dgrid, lines):
    dgrid[i] is MINDGRID(i) (minFinderND) and lines[i] the line of a non-uniform i."""
    if dg is None:
        return (1.0, 1.0, 1.0), [None] * 3
    if isinstance(dg, (int, float)):
        return (float(dg),) * 3, [None] * 3
    if not isinstance(dg, dict):
        v = [float(a) for a in dg]
        if len(v) == 2:
            v = v + [max(v)]
        return tuple(v), [None] * 3
    dgrid, lines = [], []
    for i, key in enumerate("xyz"):
        n = int(grid[i])
        if key in dg:
            f = dg[key]
            if isinstance(f, dict):
                value = float(f.get("value", dg.get("value", 1.0)))
                pts = np.asarray(f.get("points", [[0.0, 1.0], [100.0, 1.0]]), dtype=np.float64)
                px = np.floor((n - 1) * pts[:, 0] / 100.0) + 1.0
                line = pchip(px, value * pts[:, 1], np.arange(1, n + 1, dtype=np.float64))
            else:
                arr = np.asarray(f, dtype=np.float64)
                if arr.size == 1:
                    dgrid.append(float(arr.ravel()[0]))
                    lines.append(None)
                    continue
                line = arr.ravel()
                if line.size != n:
                    raise ValueError("dGrid.%s holds %d spacings, the grid %d cells"
                                     % (key, line.size, n))
        else:
            if "value" not in dg:
                raise ValueError("dGrid has no field %s and no value" % key)
            dgrid.append(float(dg["value"]))
            lines.append(None)
            continue
        if np.any(line <= 0):
            raise ValueError("dGrid.%s: spacings must be > 0" % key)
        dgrid.append(float(np.min(line)))
        lines.append(line)
    return tuple(dgrid), lines


def from_ini(ini, perms_order=0):
    """Parse an ini dict (keys as produced by ``Initializer.getRunSettings``)."""
    grid = tuple(int(g) for g in ini["grid"])
    if len(grid) != 3:
        raise ValueError("ini.grid must have 3 entries")
    dgrid, dlines = grid_spacing(ini.get("dGrid"), grid)
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
        time_max=float(ini.get("timeMax", 5000.0)), perms_order=int(perms_order),
        dgrid_lines=dlines)
