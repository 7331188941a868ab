"""Shared test helpers: build the oracle and the GPU context from the same problem, and
compare their states."""
import numpy as np

from imogen_amd import config as cfgmod
from oracle import oracle as O

ORACLE_LIMITER = {"vanleer": "vanleer", "superbee": "superbee", "minmond": "minmod"}
FLUID_FIELDS = ["mass", "momx", "momy", "momz", "ener"]
MAG_FIELDS = ["magx", "magy", "magz", "cmagx", "cmagy", "cmagz"]


def oracle_config(rc):
    return O.make_config(rc.grid, rc.dgrid, rc.gamma, rc.minmass, rc.cfl, rc.bc_infinity,
                         rc.bc, [ORACLE_LIMITER[l] for l in rc.limiter], rc.fluid_active,
                         rc.magnet_active, rc.update_mode, rc.iter_max, rc.time_max,
                         rc.perms_order)


def make_oracle(ini, state, threads=1):
    rc = cfgmod.from_ini(ini)
    o = O.Oracle(oracle_config(rc), threads)
    for d, line in enumerate(rc.dgrid_lines):
        if line is not None:
            o.set_spacing(d + 1, line)
    o.set_state(state["mass"], state["mom"], state["ener"], state["mag"])
    return o


def make_gpu(ini, state, precision="fp64"):
    import imogen_amd
    run = imogen_amd.initialize(ini, precision)
    imogen_amd.create_arrays(run, state)
    return run


def rel_linf(a, ref):
    """Normwise relative L-infinity: max|a-ref| / max|ref| (absolute when ref == 0)."""
    d = float(np.max(np.abs(np.asarray(a) - np.asarray(ref))))
    m = float(np.max(np.abs(ref)))
    return d / m if m > 0 else d


def rel_l1(a, ref):
    d = float(np.sum(np.abs(np.asarray(a) - np.asarray(ref))))
    m = float(np.sum(np.abs(ref)))
    return d / m if m > 0 else d


def compare(run, orc, fields, tol):
    """Largest relative L-inf over `fields`; asserts each is <= tol."""
    worst = 0.0
    for f in fields:
        e = rel_linf(run.download(f), orc.get(f))
        assert e <= tol, "field %s: relative L-inf %.3e > %.1e" % (f, e, tol)
        worst = max(worst, e)
    return worst
