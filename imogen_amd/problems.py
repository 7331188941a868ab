"""Initial-condition generators for the benchmark problems.

Each generator returns ``(ini, state)``:

* ``ini`` -- the run settings, keyed like the reference's ``ini`` struct
  (``experiment/Initializer.m:73-107,143-184``): grid, dGrid, gamma, cfl, bcMode,
  bcInfinity, fluxLimiter, minMass, mode (fluid/magnet), iterMax, timeMax,
  timeUpdateMode.
* ``state`` -- ``mass``, ``mom`` (3 arrays), ``ener``, ``mag`` (3 face-centred arrays),
  numpy float64 arrays of shape ``(nz, ny, nx)`` (C order, so x is fastest in memory,
  the column-major ``[nx ny nz]`` layout of ``imogen.m:34-42``).

The formulas follow the reference initializers where they are well defined and the
problem statements of the benchmark configurations otherwise:

* Sod: ``experiment/SodShockTube/SodShockTubeInitializer.m:75-109`` -- right state from
  cell ``round(N/2)`` (1-based) inclusive; ``E = P/(gamma-1)``.
* Orszag-Tang: ``experiment/OrszagTangVortex/OrszagTangVortexInitializer.m:72-90`` -- B
  written straight into the face arrays.  The reference's ``(N-1)``-period grid is
  replaced by the periodic grid ``x_i = i*2pi/N``.
* Sedov-Taylor: energy 1 deposited uniformly, as internal energy, in the cells whose
  centres lie within 3 cells of the box centre (the reference's odd-N-only ``ndgrid``
  construction, ``SedovTaylorBlastWaveInitializer.m:77-84``, is not reproduced).
* Kelvin-Helmholtz: ``experiment/KelvinHelmholtz/KelvinHelmholtzInitializer.m:61-117`` --
  density ratio 8, equal pressure, counter-streaming x momentum, circular x and
  constant y boundaries; the shear speed from ``external/speedFromMach.m``.
* MHD Riemann problem: a seeded two-state shock tube on an ``N x 1 x 1`` grid with the
  same ``B_x`` on both sides (so div B = 0), constant x and circular y/z boundaries.
"""
import math

import numpy as np


def _base_ini(grid, dgrid, gamma, cfl, bc, limiter, magnet, iter_max=1000, time_max=5000.0):
    return {
        "grid": tuple(int(g) for g in grid),
        "dGrid": tuple(float(d) for d in dgrid),
        "gamma": float(gamma),
        "cfl": float(cfl),
        "bcMode": bc,
        "bcInfinity": 20,
        "fluxLimiter": limiter,
        "minMass": 1e-5,
        "mode": {"fluid": True, "magnet": bool(magnet)},
        "iterMax": int(iter_max),
        "timeMax": float(time_max),
        "timeUpdateMode": 1,
    }


def resize(ini, grid):
    """The configuration of the same problem (same domain) on another grid: the spacing
    scales with the cell count of each dimension."""
    out = dict(ini)
    old = ini["grid"]
    out["grid"] = tuple(int(g) for g in grid)
    out["dGrid"] = tuple(float(ini["dGrid"][d]) * old[d] / grid[d] for d in range(3))
    return out


def _zeros(shape):
    return np.zeros(shape, dtype=np.float64)


def sod(n=1024, ny=1, gamma=1.4, limiter="vanleer", bc="trans", iter_max=200,
        time_max=5000.0):
    """Sod shock tube along x on [0,1] (n x ny x 1 cells)."""
    shape = (1, ny, n)
    mass = np.ones(shape)
    ener = np.ones(shape) / (gamma - 1.0)
    half = int(math.floor(n / 2.0 + 0.5))  # MATLAB round(N/2)
    mass[..., half - 1:] *= 0.125
    ener[..., half - 1:] *= 0.1
    state = {"mass": mass, "mom": [_zeros(shape) for _ in range(3)], "ener": ener,
             "mag": [_zeros(shape) for _ in range(3)]}
    ini = _base_ini((n, ny, 1), (1.0 / n, 1.0 / ny, 1.0), gamma, 0.7, bc, limiter, False,
                    iter_max, time_max)
    return ini, state


def orszag_tang(n=512, gamma=5.0 / 3.0, limiter="superbee", iter_max=1000):
    """Orszag-Tang vortex, 2-D ideal MHD, n x n periodic on [0, 2pi]^2."""
    shape = (1, n, n)
    dx = 2.0 * math.pi / n
    x = np.arange(n) * dx
    X = np.broadcast_to(x[None, None, :], shape)
    Y = np.broadcast_to(x[None, :, None], shape)
    rho = 25.0 / (36.0 * math.pi)
    P = 5.0 / (12.0 * math.pi)
    mass = np.full(shape, rho)
    mom = [-mass * np.sin(Y), mass * np.sin(X), _zeros(shape)]
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    mag = [-b0 * np.sin(Y), b0 * np.sin(2.0 * X), _zeros(shape)]
    ener = (P / (gamma - 1.0) + 0.5 * (mom[0] ** 2 + mom[1] ** 2 + mom[2] ** 2) / mass
            + 0.5 * (mag[0] ** 2 + mag[1] ** 2 + mag[2] ** 2))
    state = {"mass": mass, "mom": [np.ascontiguousarray(m) for m in mom], "ener": ener,
             "mag": [np.ascontiguousarray(b) for b in mag]}
    ini = _base_ini((n, n, 1), (dx, dx, 1.0), gamma, 0.35, "circ", limiter, True, iter_max)
    return ini, state


def sedov(n=256, gamma=5.0 / 3.0, limiter="minmond", time_max=0.05, iter_max=100000,
          radius_cells=3.0, p_ambient=1e-5, energy=1.0):
    """3-D Sedov-Taylor blast (hydro), n^3 on the unit cube, transparent boundaries."""
    shape = (n, n, n)
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx - 0.5
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    inside = r2 <= (radius_cells * dx) ** 2
    n_in = int(inside.sum())
    mass = np.ones(shape)
    ener = np.full(shape, p_ambient / (gamma - 1.0))
    ener[inside] += energy / (n_in * dx ** 3)
    state = {"mass": mass, "mom": [_zeros(shape) for _ in range(3)], "ener": ener,
             "mag": [_zeros(shape) for _ in range(3)]}
    ini = _base_ini((n, n, n), (dx, dx, dx), gamma, 0.7, "trans", limiter, False, iter_max,
                    time_max)
    return ini, state


def speed_from_mach(mach, gamma, mass, ener):
    """external/speedFromMach.m with a zero magnetic field."""
    gm1 = gamma - 1.0
    return mach * math.sqrt((gamma * ener) / (1.0 / gm1 + gamma * mach * mach / 2.0)) \
        if mass else 0.0


def kelvin_helmholtz(n=256, mach=0.5, mass_ratio=8.0, gamma=5.0 / 3.0, perturb=0.01, seed=0,
                     limiter="vanleer", iter_max=1000):
    """Kelvin-Helmholtz shear layer, n x n, circular x / constant y boundaries."""
    rng = np.random.default_rng(seed)
    shape = (1, n, n)
    mass = np.ones(shape)
    half = int(math.ceil(n / 2.0))
    mass[:, half - 1:, :] /= mass_ratio
    speed = speed_from_mach(mach, gamma, 1.0, 1.0 / (gamma - 1.0))
    vx = np.full(shape, speed)
    vx[:, half - 1:, :] = -speed
    vx = vx * (1.0 + perturb * rng.uniform(-1.0, 1.0, shape))
    vy = perturb * speed * rng.uniform(-1.0, 1.0, shape)
    mom = [mass * vx, mass * vy, _zeros(shape)]
    P = 1.0  # (max rho)^gamma with max rho = 1: equal pressure across the layer
    ener = P / (gamma - 1.0) + 0.5 * (mom[0] ** 2 + mom[1] ** 2) / mass
    state = {"mass": mass, "mom": mom, "ener": ener, "mag": [_zeros(shape) for _ in range(3)]}
    ini = _base_ini((n, n, 1), (1.0 / n, 1.0 / n, 1.0), gamma, 0.7,
                    {"x": "circ", "y": "const", "z": "const"}, limiter, False, iter_max)
    return ini, state


def mhd_riemann(n=1024, seed=0, gamma=5.0 / 3.0, limiter="vanleer", iter_max=1000):
    """Seeded 1-D MHD Riemann problem on n x 1 x 1 cells, split at n/2."""
    rng = np.random.default_rng(seed)
    shape = (1, 1, n)
    split = n // 2

    def side():
        rho = math.exp(rng.uniform(math.log(0.1), math.log(10.0)))
        p = math.exp(rng.uniform(math.log(0.01), math.log(10.0)))
        v = rng.uniform(-1.0, 1.0, 3)
        b = rng.uniform(-1.0, 1.0, 3)
        return rho, p, v, b

    L, R = side(), side()
    bx = rng.uniform(-1.0, 1.0)
    mass = np.empty(shape)
    mom = [np.empty(shape) for _ in range(3)]
    mag = [np.empty(shape) for _ in range(3)]
    ener = np.empty(shape)
    for sl, (rho, p, v, b) in ((np.s_[..., :split], L), (np.s_[..., split:], R)):
        bb = np.array([bx, b[1], b[2]])
        mass[sl] = rho
        for k in range(3):
            mom[k][sl] = rho * v[k]
            mag[k][sl] = bb[k]
        ener[sl] = p / (gamma - 1.0) + 0.5 * rho * float(v @ v) + 0.5 * float(bb @ bb)
    state = {"mass": mass, "mom": mom, "ener": ener, "mag": mag}
    # The size-1 dims get the x spacing: TimeManager.update takes dt = CFL*MINDGRID(gi)/c
    # with gi the direction of the largest signal speed, and the transverse velocities
    # here can be the largest (a unit spacing there would give a 1/dx times too large dt).
    ini = _base_ini((n, 1, 1), (1.0 / n, 1.0 / n, 1.0 / n), gamma, 0.35,
                    {"x": "const", "y": "circ", "z": "circ"}, limiter, True, iter_max)
    return ini, state


# ---- the reference initializers themselves ---------------------------------------------
# Cell for cell what experiment/*Initializer.m compute (the benchmark generators above
# differ where the reference's construction is a quirk: an (N - 1)-period grid, an
# odd-N-only Sedov grid).  The device kinds ORSZAG_TANG_REF, SEDOV_REF and
# MAGNETIC_SHOCK_TUBE of imh_init_problem evaluate the same expressions.

def orszag_tang_reference(nx=256, ny=None, limiter="vanleer"):
    """OrszagTangVortexInitializer.m:42-58,72-90: dGrid = 1 ./ grid; x, y are the cell
    indices 0..N-1 and the sines have period N - 1 (so the first and last columns carry
    the same values); gamma 5/3, CFL 0.35, ITERMAX 1500, TIMEMAX 0.48, circular
    boundaries.  The initializer sets no flux limiter (the manager's default applies)."""
    ny = nx if ny is None else ny
    shape = (1, ny, nx)
    gamma = 5.0 / 3.0
    mass = 25.0 / (36.0 * math.pi) * np.ones(shape)
    x = np.broadcast_to(np.arange(nx, dtype=np.float64)[None, None, :], shape)
    y = np.broadcast_to(np.arange(ny, dtype=np.float64)[None, :, None], shape)
    mom = [-mass * np.sin(2 * math.pi * y / (ny - 1)),
           mass * np.sin(2 * math.pi * x / (nx - 1)), _zeros(shape)]
    mag0 = 1.0 / math.sqrt(4.0 * math.pi)
    mag = [-mag0 * np.sin(2 * math.pi * y / (ny - 1)),
           mag0 * np.sin(4 * math.pi * x / (nx - 1)), _zeros(shape)]
    S = (mom[0] * mom[0] + mom[1] * mom[1]) + mom[2] * mom[2]
    B = (mag[0] * mag[0] + mag[1] * mag[1]) + mag[2] * mag[2]
    ener = (5.0 / (12.0 * math.pi) / (gamma - 1.0) + 0.5 * S / mass) + 0.5 * B
    state = {"mass": mass, "mom": [np.ascontiguousarray(m) for m in mom], "ener": ener,
             "mag": [np.ascontiguousarray(b) for b in mag]}
    ini = _base_ini((nx, ny, 1), (1.0 / nx, 1.0 / ny, 1.0), gamma, 0.35, "circ", limiter,
                    True, 1500, 0.48)
    return ini, state


def sedov_reference(n=65, ny=None, nz=None, limiter="vanleer"):
    """SedovTaylorBlastWaveInitializer.m:30-50,71-86: dGrid = 1 ./ grid, integer offsets
    X, Y, Z from the centre cell by ndgrid(-half:half) (half = floor(grid / 2), so every
    dimension with more than one cell needs an odd count), radius max(5, round(0.06 *
    min(grid))) cells, pressure 3 (gamma - 1) / (4 pi (dGrid(1) radius)^3) where
    distance <= radius and 1e-5 elsewhere, E = P / (gamma - 1); gamma 1.4, CFL 0.7,
    ITERMAX 100, fade boundaries."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    grid = (n, ny, nz)
    for g in grid:
        if g > 1 and g % 2 == 0:
            raise ValueError("the reference Sedov-Taylor initializer needs odd cell counts "
                             "(ndgrid(-half:half) has 2*floor(n/2)+1 points), got %s" % (grid,))
    gamma = 1.4
    half = [g // 2 for g in grid]
    X = np.arange(-half[0], half[0] + 1)[None, None, :]
    Y = np.arange(-half[1], half[1] + 1)[None, :, None]
    Z = np.arange(-half[2], half[2] + 1)[:, None, None]
    d2 = X * X + Y * Y + Z * Z                     # distance^2, exact in integers
    radius = max(5, int(math.floor(0.06 * min(grid) + 0.5)))   # MATLAB round
    dx = 1.0 / n
    inside = 3 * (gamma - 1) / (4 * math.pi * (dx * radius) ** 3)
    pressure = np.where(d2 > radius * radius, 1e-5, inside).astype(np.float64)
    shape = (nz, ny, n)
    state = {"mass": np.ones(shape), "mom": [_zeros(shape) for _ in range(3)],
             "ener": pressure / (gamma - 1), "mag": [_zeros(shape) for _ in range(3)]}
    ini = _base_ini(grid, (1.0 / n, 1.0 / ny, 1.0 / nz), gamma, 0.7, "fade", limiter, False,
                    100)
    return ini, state


def magnetic_shock_tube(nx=1024, ny=3, nz=3, direction="X", fields=(True, True, False),
                        limiter="vanleer"):
    """MagneticShockTubeInitializer.m:42-66,79-117: unit spacing (the initializer leaves
    dGrid empty; initialize.m:166-168), rho 1, B = 5/sqrt(4 pi) in the components that are
    on, momentum 10 along the shock direction with the right side (from MATLAB round(N/2),
    1-based) set to -1.0 times the x momentum there (so a y or z shock gets -0.0 on the
    right: the .m text reads mom(1, ...)), E = kinetic + magnetic, plus 20/(gamma-1) on
    the left and 1/(gamma-1) on the right; gamma 1.4, CFL 0.35, ITERMAX 250, fade in x,
    circular in y and z."""
    grid = (nx, ny, nz)
    shape = (nz, ny, nx)
    gamma = 1.4
    dirn = "XYZ".index(direction.upper())
    half = int(math.floor(grid[dirn] / 2.0 + 0.5))
    mass = np.ones(shape)
    mom = [_zeros(shape) for _ in range(3)]
    b = 5 / math.sqrt(4 * math.pi)
    mag = [np.full(shape, b) if on else _zeros(shape) for on in fields]
    axis = 2 - dirn
    idx = [slice(None)] * 3
    idx[axis] = slice(half - 1, None)              # indeces{i} = half:grid(i)
    right = tuple(idx)
    mom[dirn][...] = 10.0
    mom[dirn][right] = -1.0 * mom[0][right]
    S = (mom[0] * mom[0] + mom[1] * mom[1]) + mom[2] * mom[2]
    B = (mag[0] * mag[0] + mag[1] * mag[1]) + mag[2] * mag[2]
    ener = 0.5 * S / mass + 0.5 * B
    ener[right] = ener[right] + 1 / (gamma - 1) * 1.0
    idx[axis] = slice(0, half - 1)                 # indeces{i} = 1:(half - 1)
    left = tuple(idx)
    ener[left] = ener[left] + 20 / (gamma - 1) * 1.0
    state = {"mass": mass, "mom": mom, "ener": ener, "mag": mag}
    ini = _base_ini(grid, (1.0, 1.0, 1.0), gamma, 0.35,
                    {"x": "fade", "y": "circ", "z": "circ"}, limiter, True, 250)
    return ini, state


def random_state(grid, seed=0, magnet=True, amplitude=0.2):
    """Smooth-ish random state for unit and parity tests (positive density/pressure)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid
    shape = (nz, ny, nx)
    mass = 1.0 + amplitude * rng.uniform(-1, 1, shape)
    vel = [amplitude * rng.uniform(-1, 1, shape) for _ in range(3)]
    mag = [(0.5 + amplitude * rng.uniform(-1, 1, shape)) if magnet else _zeros(shape)
           for _ in range(3)]
    P = 1.0 + amplitude * rng.uniform(-1, 1, shape)
    gamma = 5.0 / 3.0
    mom = [mass * v for v in vel]
    ener = P / (gamma - 1.0) + 0.5 * mass * sum(v * v for v in vel) + 0.5 * sum(b * b for b in mag)
    return {"mass": mass, "mom": mom, "ener": ener, "mag": mag}
