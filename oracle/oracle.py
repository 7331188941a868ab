"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(imogen_amd) never imports this module.

Arrays follow the reference layout (column-major [nx ny nz], x fastest), held in
numpy as C-ordered arrays of shape (nz, ny, nx).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

BC = {"circ": 0, "const": 1, "trans": 2, "fade": 3, "mirror": 4, "flip": 5, "wall": 6, "zero": 7,
      "wormhole": 8}
LIMITER = {"vanleer": 0, "superbee": 1, "minmod": 2}
FIELDS = ["mass", "momx", "momy", "momz", "ener", "magx", "magy", "magz",
          "cmagx", "cmagy", "cmagz"]


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_int * 3),
        ("dgrid", ctypes.c_double * 3),
        ("gamma", ctypes.c_double),
        ("minmass", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("bc_infinity", ctypes.c_int),
        ("bc", (ctypes.c_int * 3) * 2),
        ("limiter", ctypes.c_int * 3),
        ("fluid_active", ctypes.c_int),
        ("magnet_active", ctypes.c_int),
        ("update_mode", ctypes.c_int),
        ("iter_max", ctypes.c_int),
        ("time_max", ctypes.c_double),
        ("perms_order", ctypes.c_int),
    ]


_lib = None


def lib():
    """Load liboracle.so (built by `make -C oracle`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_state.argtypes = [ctypes.c_void_p, P, P * 3, P, P * 3]
        L.orc_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, P]
        L.orc_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, P]
        for name in ("orc_time_update", "orc_flux", "orc_relaxing_fluid"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_magnet_flux.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int * 2]
        L.orc_iteration.argtypes = [ctypes.c_void_p]
        L.orc_running.argtypes = [ctypes.c_void_p]
        for name in ("orc_time", "orc_dtime"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_double
        L.orc_iteration_count.argtypes = [ctypes.c_void_p]
        L.orc_last_grid_index.argtypes = [ctypes.c_void_p]
        L.orc_set_dtime.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_last_error.argtypes = [ctypes.c_void_p]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_sweep_order.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int * 3, (ctypes.c_int * 2) * 3]
        L.orc_pressure.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double] + [P] * 8 + \
            [ctypes.c_int, P, P]
        L.orc_limiter.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P]
        L.orc_pchip5.argtypes = [ctypes.c_double * 5, ctypes.c_double * 5, ctypes.c_double]
        L.orc_pchip5.restype = ctypes.c_double
        L.orc_fade_coeffs.argtypes = [ctypes.c_int, P]
        L.orc_shift_test.argtypes = [ctypes.c_int * 3, (ctypes.c_int * 3) * 2, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, P, P, P, P, ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_gravity.argtypes = [ctypes.c_void_p, P, ctypes.c_double]
        L.orc_set_spacing.argtypes = [ctypes.c_void_p, ctypes.c_int, P, ctypes.c_int]
        L.orc_set_radiation.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double]
        L.orc_radiation_strength.argtypes = [ctypes.c_void_p]
        L.orc_radiation_strength.restype = ctypes.c_double
        L.orc_source.argtypes = [ctypes.c_void_p]
        L.orc_set_viscosity.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double]
        L.orc_viscosity_term.argtypes = [ctypes.c_void_p, ctypes.c_int, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(grid, dgrid, gamma=5.0 / 3.0, minmass=1e-5, cfl=0.7, bc_infinity=20,
                bc="circ", limiter="vanleer", fluid=True, magnet=False, update_mode=1,
                iter_max=1000, time_max=5000.0, perms_order=0):
    """Build an OrcConfig.  bc: a mode name, a dict {x,y,z} of names or (lower, upper)
    pairs, or a [2][3] nested list of names.  limiter: a name or three names."""
    c = OrcConfig()
    for d in range(3):
        c.grid[d] = int(grid[d])
        c.dgrid[d] = float(dgrid[d])
    c.gamma, c.minmass, c.cfl = gamma, minmass, cfl
    c.bc_infinity = int(bc_infinity)
    modes = normalize_bc(bc)
    for s in range(2):
        for d in range(3):
            c.bc[s][d] = BC[modes[s][d]]
    lims = [limiter] * 3 if isinstance(limiter, str) else list(limiter)
    for d in range(3):
        c.limiter[d] = LIMITER[lims[d]]
    c.fluid_active, c.magnet_active = int(fluid), int(magnet)
    c.update_mode, c.iter_max, c.time_max = int(update_mode), int(iter_max), float(time_max)
    c.perms_order = int(perms_order)
    return c


def normalize_bc(bc):
    """Return [[lower x,y,z],[upper x,y,z]] mode names."""
    if isinstance(bc, str):
        return [[bc] * 3, [bc] * 3]
    if isinstance(bc, dict):
        out = [[None] * 3, [None] * 3]
        for d, key in enumerate("xyz"):
            v = bc.get(key, "circ")
            if isinstance(v, (list, tuple)):
                out[0][d], out[1][d] = v[0], v[-1]
            else:
                out[0][d] = out[1][d] = v
        return out
    return [list(bc[0]), list(bc[1])]


class Oracle:
    """One simulation on the CPU oracle."""

    def __init__(self, config, threads=1):
        self.L = lib()
        self.cfg = config
        err = ctypes.create_string_buffer(256)
        self.h = self.L.orc_create(ctypes.byref(config), err, 256)
        if not self.h:
            raise ValueError(err.value.decode())
        self.shape = (config.grid[2], config.grid[1], config.grid[0])
        self.L.orc_set_threads(int(threads))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, mass, mom, ener, mag):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
                for a in [mass, *mom, ener, *mag]]
        self._keep = arrs
        P = ctypes.POINTER(ctypes.c_double)
        momp = (P * 3)(*[_ptr(a) for a in arrs[1:4]])
        magp = (P * 3)(*[_ptr(a) for a in arrs[5:8]])
        self._check(self.L.orc_set_state(self.h, _ptr(arrs[0]), momp, _ptr(arrs[4]), magp))

    def get(self, field):
        idx = FIELDS.index(field) if isinstance(field, str) else int(field)
        out = np.empty(self.shape, dtype=np.float64)
        self._check(self.L.orc_get_field(self.h, idx, _ptr(out)))
        return out

    def set(self, field, value):
        idx = FIELDS.index(field) if isinstance(field, str) else int(field)
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(self.shape)
        self._check(self.L.orc_set_field(self.h, idx, _ptr(v)))

    def state(self):
        return {f: self.get(f) for f in FIELDS}

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError("oracle: " + self.L.orc_last_error(self.h).decode())

    def time_update(self, i):
        self._check(self.L.orc_time_update(self.h, int(i)))

    def flux(self, order):
        self._check(self.L.orc_flux(self.h, int(order)))

    def relaxing_fluid(self, X):
        self._check(self.L.orc_relaxing_fluid(self.h, int(X)))

    def magnet_flux(self, X, I2):
        self._check(self.L.orc_magnet_flux(self.h, int(X), (ctypes.c_int * 2)(*I2)))

    def iteration(self):
        self._check(self.L.orc_iteration(self.h))

    def set_gravity(self, potential, mass_threshold=0.0):
        """GravityManager with the 'empty' solver: grav = fixedPotential (None: off)."""
        if potential is None:
            self._check(self.L.orc_set_gravity(self.h, None, float(mass_threshold)))
            return
        v = np.ascontiguousarray(potential, dtype=np.float64).reshape(self.shape)
        self._check(self.L.orc_set_gravity(self.h, _ptr(v), float(mass_threshold)))

    def set_spacing(self, dim, line):
        """Non-uniform DGRID{dim} (1..3): a line of spacings along that dimension, or None
        for the configuration's uniform value."""
        if line is None:
            self._check(self.L.orc_set_spacing(self.h, int(dim), None, 0))
            return
        v = np.ascontiguousarray(line, dtype=np.float64).ravel()
        self._keep_dx = getattr(self, "_keep_dx", {})
        self._keep_dx[dim] = v
        self._check(self.L.orc_set_spacing(self.h, int(dim), _ptr(v), int(v.size)))

    def set_radiation(self, kind, exponent=0.5, initial_maximum=1.0):
        """Radiation type ('empty' / 'optically_thin'), strength from the current state."""
        t = {"empty": 0, "optically_thin": 1}[kind] if isinstance(kind, str) else int(kind)
        self._check(self.L.orc_set_radiation(self.h, t, float(exponent), float(initial_maximum)))

    @property
    def radiation_strength(self):
        return self.L.orc_radiation_strength(self.h)

    def set_viscosity(self, kind, linear=0.0, quadratic=0.0):
        """Artificial viscosity: 'empty', 'neumann_richtmyer' or 'caramana_shashkov_whalen'."""
        t = ({"empty": 0, "neumann_richtmyer": 1, "caramana_shashkov_whalen": 2}[kind]
             if isinstance(kind, str) else int(kind))
        self._check(self.L.orc_set_viscosity(self.h, t, float(linear), float(quadratic)))

    def viscosity_term(self, X):
        """The pressure term of the configured viscosity for the current state along X."""
        out = np.empty(self.shape, dtype=np.float64)
        self._check(self.L.orc_viscosity_term(self.h, int(X), _ptr(out)))
        return out

    def source(self):
        self._check(self.L.orc_source(self.h))

    @property
    def running(self):
        return bool(self.L.orc_running(self.h))

    @property
    def time(self):
        return self.L.orc_time(self.h)

    @property
    def dtime(self):
        return self.L.orc_dtime(self.h)

    @dtime.setter
    def dtime(self, v):
        self.L.orc_set_dtime(self.h, float(v))

    @property
    def iteration_count(self):
        return self.L.orc_iteration_count(self.h)

    @property
    def grid_index(self):
        return self.L.orc_last_grid_index(self.h)


def sweep_order(iteration, order, perms_order=0):
    dirs = (ctypes.c_int * 3)()
    mi = ((ctypes.c_int * 2) * 3)()
    lib().orc_sweep_order(int(iteration), int(order), int(perms_order), dirs, mi)
    return list(dirs), [list(r) for r in mi]


def pchip5(x, y, xi):
    return lib().orc_pchip5((ctypes.c_double * 5)(*x), (ctypes.c_double * 5)(*y), float(xi))


def fade_coeffs(f):
    out = np.empty(f, dtype=np.float64)
    lib().orc_fade_coeffs(int(f), _ptr(out))
    return out


def pressure(mode, gamma, rho, mom, ener, bc, use_mag=0):
    """mode 'totsnd' -> (P_total, c_fast); 'sound' -> c_sound."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).ravel()
    n = rho.size
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (*mom, ener, *bc)]
    res = np.empty(n)
    aux = np.empty(n)
    lib().orc_pressure(0 if mode == "totsnd" else 1, n, float(gamma), _ptr(rho),
                       *[_ptr(a) for a in arrs], int(use_mag), _ptr(res), _ptr(aux))
    return (res, aux) if mode == "totsnd" else res


def limiter(name, flux, dl, dr):
    f = np.array(flux, dtype=np.float64).ravel().copy()
    dl = np.ascontiguousarray(dl, dtype=np.float64).ravel()
    dr = np.ascontiguousarray(dr, dtype=np.float64).ravel()
    lib().orc_limiter(LIMITER[name], f.size, _ptr(f), _ptr(dl), _ptr(dr))
    return f


def shift_test(array, bc, D, n, component=0, is_flux=False, flux_type=0, bc_infinity=20,
               edge_lo=None, edge_hi=None, active_linear=1):
    """Shift a standalone array (shape (nz,ny,nx)) as ImogenArray.shift would.  Returns
    (result, edge_lo, edge_hi) with the tracker state after the call."""
    a = np.ascontiguousarray(array, dtype=np.float64)
    nz, ny, nx = a.shape
    grid = (ctypes.c_int * 3)(nx, ny, nz)
    modes = normalize_bc(bc)
    bcarr = ((ctypes.c_int * 3) * 2)()
    for s in range(2):
        for d in range(3):
            bcarr[s][d] = BC[modes[s][d]]
    out = np.empty_like(a)
    nface = a.size // (nx, ny, nz)[D - 1]
    lo = np.ascontiguousarray(edge_lo if edge_lo is not None else np.zeros(nface), dtype=np.float64).copy()
    hi = np.ascontiguousarray(edge_hi if edge_hi is not None else np.zeros(nface), dtype=np.float64).copy()
    rc = lib().orc_shift_test(grid, bcarr, int(bc_infinity), int(component), int(is_flux),
                              int(flux_type), int(D), int(n), _ptr(a), _ptr(out), _ptr(lo),
                              _ptr(hi), int(active_linear))
    if rc:
        raise ValueError("shift_test rejected the request")
    return out, lo, hi
