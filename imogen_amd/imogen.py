"""Host-side mirror of the reference's call surface for the relaxing-TVD path.

The reference drives the fluid/magnet update through manager singletons and array
objects (``initialize.m``, ``imogen.m:34-72``, ``flux/flux.m``,
``fluid/relaxingFluid.m``, ``magnet/magnetFlux.m``, ``time/TimeManager.m``).  This module
keeps those names, argument meanings and error behaviour; the state lives on the GPU
behind the C ABI (include/imogen_hip.h) and every hot-path call is one ABI call.

    run = initialize(ini)                            # initialize.m
    mass, mom, ener, mag = create_arrays(run, IC)    # imogen.m:34-42 constructors
    run.time.update(mass, mom, ener, mag, 1)         # TimeManager.update
    flux(run, mass, mom, ener, mag, None, 1)         # flux.m
    relaxingFluid(run, mass, mom, ener, mag, None, X)
    magnetFlux(run, mass, mom, mag, X, [I1, I2])
    imogen(run, mass, mom, ener, mag)                # the main loop
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD_IDS, ImhConfig, as_field, check, dptr, lib
from .config import BC_CODES, LIMITER_CODES, ENUM, RunConfig, from_ini


def make_imh_config(rc: RunConfig, precision=_lib.FP64, device=0):
    c = ImhConfig()
    for d in range(3):
        c.grid[d] = int(rc.grid[d])
        c.dgrid[d] = float(rc.dgrid[d])
    c.gamma, c.minmass, c.cfl = rc.gamma, rc.minmass, rc.cfl
    c.bc_infinity = int(rc.bc_infinity)
    for s in range(2):
        for d in range(3):
            c.bc[s][d] = BC_CODES[rc.bc[s][d]]
    for d in range(3):
        c.limiter[d] = LIMITER_CODES[rc.limiter[d]]
    c.fluid_active, c.magnet_active = int(rc.fluid_active), int(rc.magnet_active)
    c.update_mode, c.iter_max, c.time_max = int(rc.update_mode), int(rc.iter_max), float(rc.time_max)
    c.perms_order = int(rc.perms_order)
    c.precision = int(precision)
    c.device = int(device)
    return c


class FluidManager:
    """fluid/FluidManager.m: ACTIVE, MINMASS, limiter{3}."""

    def __init__(self, rc):
        self.ACTIVE = rc.fluid_active
        self.MINMASS = rc.minmass
        self.limiter = list(rc.limiter)


class MagnetManager:
    """magnet/MagnetManager.m: ACTIVE and the limiters (the fluid's, initialize.m:51)."""

    def __init__(self, rc):
        self.ACTIVE = rc.magnet_active
        self.limiter = list(rc.limiter)


class BCManager:
    """shifting/BCManager.m: modes and infinity."""

    def __init__(self, rc):
        self.modes = [list(rc.bc[0]), list(rc.bc[1])]
        self.infinity = rc.bc_infinity


class TimeManager:
    """time/TimeManager.m.  dTime and time live on the device; reading them syncs."""

    def __init__(self, run, rc):
        self._run = run
        self.CFL = rc.cfl
        self.ITERMAX = rc.iter_max
        self.TIMEMAX = rc.time_max
        self.updateMode = rc.update_mode

    def _state(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        it, gi = ctypes.c_int(), ctypes.c_int()
        check(lib().imh_get_time(self._run._ctx, ctypes.byref(t), ctypes.byref(dt),
                                 ctypes.byref(it), ctypes.byref(gi)))
        return t.value, dt.value, it.value, gi.value

    @property
    def time(self):
        return self._state()[0]

    @property
    def dTime(self):
        return self._state()[1]

    @dTime.setter
    def dTime(self, value):
        check(lib().imh_set_dtime(self._run._ctx, float(value)))

    @property
    def iteration(self):
        return self._state()[2]

    @iteration.setter
    def iteration(self, value):
        check(lib().imh_set_iteration(self._run._ctx, int(value)))

    @property
    def gridIndex(self):
        return self._state()[3]

    @property
    def running(self):
        """TimeManager.running (TimeManager.m:36-39), without the wall-time clause."""
        t, _, it, _ = self._state()
        return it <= self.ITERMAX and t < self.TIMEMAX

    def update(self, mass, mom, ener, mag, fluxDirection):
        """TimeManager.update: the CFL reduction and dt, on the device."""
        check(lib().imh_time_update(self._run._ctx, int(fluxDirection)))

    def step(self):
        """TimeManager.step: iteration = iteration + 1."""
        self.iteration = self.iteration + 1


class ImogenManager:
    """The run object handed to every hot-path function (classes/ImogenManager.m)."""

    def __init__(self, rc: RunConfig, precision=_lib.FP64, device=0):
        self.config = rc
        self.gridSize = tuple(rc.grid)
        self.DGRID = tuple(rc.dgrid)
        self.MINDGRID = tuple(rc.dgrid)
        self.GAMMA = rc.gamma
        self.precision = precision
        self.fluid = FluidManager(rc)
        self.magnet = MagnetManager(rc)
        self.bc = BCManager(rc)
        self.time = TimeManager(self, rc)
        h = ctypes.c_void_p()
        check(lib().imh_create(ctypes.byref(make_imh_config(rc, precision, device)),
                               ctypes.byref(h)))
        self._ctx = h
        self._arrays = None
        self.gravity_potential = None
        # ImogenManager.setGridSpacing: a non-uniform DGRID{i} is a line along i
        self._dlines = []
        for d, line in enumerate(rc.dgrid_lines):
            if line is not None:
                self.set_spacing(d + 1, line)

    def set_spacing(self, dim, line):
        """DGRID{dim} as a line of spacings along that dimension (None: the uniform
        configured value); MINDGRID(dim) becomes its minimum."""
        if line is None:
            check(lib().imh_set_spacing(self._ctx, int(dim), None, 0))
            return
        arr = np.ascontiguousarray(line, dtype=np.float64).ravel()
        self._dlines.append(arr)
        check(lib().imh_set_spacing(self._ctx, int(dim), dptr(arr), arr.size))
        dg = list(self.DGRID)
        dg[dim - 1] = arr.copy()
        self.DGRID = tuple(dg)
        md = list(self.MINDGRID)
        md[dim - 1] = float(np.min(arr))
        self.MINDGRID = tuple(md)

    def close(self):
        if getattr(self, "_ctx", None):
            lib().imh_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ----------------------------------------------------------------------
    def upload_state(self, mass, mom, ener, mag):
        n = self.config.ncell
        arrs = [as_field(a, n) for a in (mass, *mom, ener, *mag)]
        P = ctypes.POINTER(ctypes.c_double)
        check(lib().imh_upload_state(self._ctx, dptr(arrs[0]), (P * 3)(*[dptr(a) for a in arrs[1:4]]),
                                     dptr(arrs[4]), (P * 3)(*[dptr(a) for a in arrs[5:8]]), n))

    def download(self, field):
        out = np.empty(self.config.ncell, dtype=np.float64)
        check(lib().imh_download_field(self._ctx, FIELD_IDS[field], dptr(out), out.size))
        return out.reshape(self.config.shape)

    def upload(self, field, value):
        arr = as_field(value, self.config.ncell)
        check(lib().imh_upload_field(self._ctx, FIELD_IDS[field], dptr(arr), arr.size))

    def state(self):
        return {f: self.download(f) for f in FIELD_IDS}

    def conservation_sums(self):
        out = np.empty(8, dtype=np.float64)
        check(lib().imh_conservation_sums(self._ctx, dptr(out)))
        return out

    def synchronize(self):
        check(lib().imh_synchronize(self._ctx))

    def run_iterations(self, n):
        """Advance up to n iterations while TimeManager.running holds (graph replay)."""
        done = ctypes.c_int()
        check(lib().imh_run(self._ctx, int(n), ctypes.byref(done)))
        return done.value

    def memory_bytes(self):
        b = ctypes.c_size_t()
        check(lib().imh_memory_bytes(self._ctx, ctypes.byref(b)))
        return b.value

    IC_KINDS = {"sod": 1, "orszag_tang": 2, "sedov": 3, "riemann": 4, "kelvin_helmholtz": 5,
                "orszag_tang_reference": 6, "sedov_reference": 7, "magnetic_shock_tube": 8}

    def init_problem(self, kind, params=()):
        """Generate the initial state of a benchmark problem in device memory (with the
        constructor side effects of create_arrays); params as imh_init_problem takes them
        (include/imogen_hip.h)."""
        arr = (ctypes.c_double * max(1, len(params)))(*[float(p) for p in params])
        check(lib().imh_init_problem(self._ctx, self.IC_KINDS[kind], arr, len(params)))

    # -- source terms (source.m) ------------------------------------------------------
    def set_gravity(self, potential, mass_threshold=0.0):
        """run.gravity active with the 'empty' solver: grav.array = fixedPotential (a host
        array of the grid's shape; None turns gravity off) and run.fluid.MASS_THRESHOLD."""
        if potential is None:
            check(lib().imh_set_gravity(self._ctx, None, 0, float(mass_threshold)))
            self.gravity_potential = None
            return
        arr = as_field(potential, self.config.ncell)
        check(lib().imh_set_gravity(self._ctx, dptr(arr), arr.size, float(mass_threshold)))
        # grav.array as the slice saves write it (resultsHandler.m:81-84); the device copy
        # is the one the source step reads
        self.gravity_potential = arr.reshape(self.config.shape).copy()

    RADIATION = {"empty": 0, "optically_thin": 1}

    def set_radiation(self, kind, exponent=0.5, initial_maximum=1.0):
        """run.fluid.radiation type, exponent and initialMaximum; the strength follows from
        the current state (Radiation.initialize), so call it after the state is set."""
        check(lib().imh_set_radiation(self._ctx, self.RADIATION[kind], float(exponent),
                                      float(initial_maximum)))

    VISCOSITY = {"empty": 0, "neumann_richtmyer": 1, "caramana_shashkov_whalen": 2}

    def set_viscosity(self, kind, linear=0.0, quadratic=0.0):
        """run.fluid.viscosity type with linear / quadratic strengths (ENUM.m:53-55)."""
        check(lib().imh_set_viscosity(self._ctx, self.VISCOSITY[kind], float(linear),
                                      float(quadratic)))

    @property
    def radiation_strength(self):
        v = ctypes.c_double()
        check(lib().imh_get_radiation_strength(self._ctx, ctypes.byref(v)))
        return v.value

    def checkpoint(self):
        """The whole resumable state (fields, edge trackers, flux-object flags, time) as
        bytes; `restore` on a run of the same configuration continues bit for bit."""
        n = ctypes.c_size_t()
        check(lib().imh_checkpoint_bytes(self._ctx, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(lib().imh_checkpoint_save(self._ctx, buf, n.value))
        return buf.raw

    def restore(self, data):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        check(lib().imh_checkpoint_load(self._ctx, buf, len(data)))

    def set_fluid_tile(self, X, bs, kmax, px):
        check(lib().imh_set_fluid_tile(self._ctx, int(X), int(bs), int(kmax), int(px)))

    SWEEP_KINDS = {"general": 0, "wave": 1, "wave_fast": 2, "auto_fast": 3}

    def set_fluid_kernel(self, kind):
        """The fluid sweep kernel: 'general', 'wave' (wave-spread, the precision's own
        arithmetic), 'wave_fast' (wave-spread, fast FP64 arithmetic) or 'auto_fast'
        (wave_fast where its tiles fill the device, general elsewhere)."""
        check(lib().imh_set_fluid_kernel(self._ctx, self.SWEEP_KINDS[kind]))

    def get_fluid_tile(self, X):
        bs, km, px = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().imh_get_fluid_tile(self._ctx, int(X), ctypes.byref(bs), ctypes.byref(km),
                                       ctypes.byref(px)))
        return bs.value, km.value, px.value

    def set_graphs(self, enable):
        check(lib().imh_set_graphs(self._ctx, int(bool(enable))))

    def profile_iterations(self, n):
        """Run n iterations eagerly with HIP events around every kernel; returns
        {kind: (ms, algorithmic bytes, launches)}."""
        return _lib.run_profile(lib().imh_profile_iterations, self._ctx, n)


class _DeviceArray:
    """A reference array object whose data lives on the device (ImogenArray.array)."""

    def __init__(self, run, field, component, ident):
        self._run = run
        self._field = field
        self.component = component
        self.id = ident

    @property
    def array(self):
        return self._run.download(self._field)

    @array.setter
    def array(self, value):
        self._run.upload(self._field, value)

    @property
    def gridSize(self):
        return self._run.gridSize


class FluidArray(_DeviceArray):
    """fluid/FluidArray.m (mass, mom(i), ener)."""


class MagnetArray(_DeviceArray):
    """magnet/MagnetArray.m; cellMag is the cell-centred field."""

    def __init__(self, run, field, component):
        super().__init__(run, field, component, ENUM.MAG)
        self.cellMag = _DeviceArray(run, "cmag" + "xyz"[component - 1], component, ENUM.MAG)


def initialize(ini, precision="fp64", device=0, perms_order=0):
    """initialize.m for this path: parse ini and create the device context."""
    prec = {"fp64": _lib.FP64, "fp32": _lib.FP32_BULK, _lib.FP64: _lib.FP64,
            _lib.FP32_BULK: _lib.FP32_BULK}[precision]
    return ImogenManager(from_ini(ini, perms_order=perms_order), prec, device)


def create_arrays(run, state):
    """imogen.m:34-42: the FluidArray/MagnetArray constructors (uploads the IC)."""
    run.upload_state(state["mass"], state["mom"], state["ener"], state["mag"])
    mass = FluidArray(run, "mass", 0, ENUM.MASS)
    ener = FluidArray(run, "ener", 0, ENUM.ENER)
    mom = [FluidArray(run, "mom" + "xyz"[i], i + 1, ENUM.MOM) for i in range(3)]
    mag = [MagnetArray(run, "mag" + "xyz"[i], i + 1) for i in range(3)]
    run._arrays = (mass, mom, ener, mag)
    return mass, mom, ener, mag


def _check_run(run, *arrays):
    for a in arrays:
        if isinstance(a, _DeviceArray) and a._run is not run:
            raise ValueError("array belongs to another run")


def relaxingFluid(run, mass, mom, ener, mag, grav, X):
    """fluid/relaxingFluid.m: one fused fluid sweep along X (1..3)."""
    _check_run(run, mass, ener)
    check(lib().imh_fluid_sweep(run._ctx, int(X)))


def magnetFlux(run, mass, mom, mag, X, magneticIndices):
    """magnet/magnetFlux.m: CT sweep of the two components magneticIndices along X."""
    _check_run(run, mass)
    I2 = (ctypes.c_int * 2)(*[int(i) for i in magneticIndices])
    check(lib().imh_magnet_sweep(run._ctx, int(X), I2))


def flux(run, mass, mom, ener, mag, grav, order):
    """flux/flux.m: the sweeps of one timestep in the order set by the iteration."""
    if order not in (1, -1):
        raise ValueError("%g is not a recognized direction. Fluxing aborted." % order)
    check(lib().imh_flux(run._ctx, int(order)))


def source(run, mass, mom, ener, mag, grav):
    """source.m: gravity from the fixed potential and the radiation sink."""
    _check_run(run, mass, ener)
    check(lib().imh_source(run._ctx))


def iteration(run):
    """One pass of the imogen.m:61-71 loop body (device-side dt)."""
    check(lib().imh_iteration(run._ctx))


def imogen(run, mass=None, mom=None, ener=None, mag=None, max_iterations=None):
    """The main loop of imogen.m:58-72: iterate while TimeManager.running holds.
    Returns the number of iterations run."""
    n = run.time.ITERMAX if max_iterations is None else int(max_iterations)
    return run.run_iterations(n)
