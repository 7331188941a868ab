"""The wormhole boundary (shifting/wormhole_shift.m) on the GPU against the oracle.

The lower ghosts of a pencil along D come from the pencil whose next-dimension coordinate
is mirrored (x->y, y->z, z->x), negated for the momentum along D; the upper side is the
constant shift.  The general sweep holds each such pair in one tile.  Lockstep with the
oracle at <= 1e-12 relative L-inf per iteration (observed 0) on 3-D hydro grids with the
wormhole on each axis, even and odd partner extents (an odd one pairs its middle pencil
with itself), the HachisuDisk set-up (z wormhole on both sides, gravity from a fixed
potential), the exact wave-spread sweep on the other axes, Neumann-Richtmyer viscosity,
FP32 bulk within 1e-5 of FP64, and a batch member equal to the same simulation alone.
"""
import numpy as np
import pytest

from imogen_amd import iteration, problems
from imogen_amd.ensemble import Ensemble

from .helpers import FLUID_FIELDS, compare, make_gpu, make_oracle, rel_l1, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12


def box(grid, bc, seed, limiter="vanleer", gamma=1.4, cfl=0.6):
    st = problems.random_state(grid, seed=seed, magnet=False)
    h = 1.0 / max(grid)
    ini = {"grid": grid, "dGrid": (h, h, h), "gamma": gamma, "cfl": cfl, "bcMode": bc,
           "bcInfinity": 20, "fluxLimiter": limiter, "minMass": 1e-5,
           "mode": {"fluid": True, "magnet": False}, "iterMax": 1000, "timeMax": 1e9}
    return ini, st


def potential(grid):
    nx, ny, nz = grid
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return 0.6 * np.sin(2 * np.pi * X) * np.cos(np.pi * Y) + 0.4 * (Z - 0.3) ** 2


def lockstep(ini, st, iterations, kind=None, grav=False, visc=None):
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    if kind:
        run.set_fluid_kernel(kind)
    if visc:
        orc.set_viscosity(*visc)
        run.set_viscosity(*visc)
    if grav:
        phi = potential(tuple(ini["grid"]))
        orc.set_gravity(phi, 0.0)
        run.set_gravity(phi, 0.0)
    worst = 0.0
    for it in range(iterations):
        orc.iteration()
        iteration(run)
        try:
            worst = max(worst, compare(run, orc, FLUID_FIELDS, TOL))
        except AssertionError as e:
            raise AssertionError("%s iteration %d: %s" % (ini["grid"], it + 1, e))
    print("wormhole %s %s: max relative L-inf over %d iterations %.3e" % (
        ini["grid"], ini["bcMode"], iterations, worst), flush=True)
    run.close()
    orc.close()


@pytest.mark.parametrize("grid,bc", [
    ((24, 12, 10), {"x": ["wormhole", "trans"], "y": "circ", "z": "circ"}),   # partner y, even
    ((24, 13, 10), {"x": ["wormhole", "const"], "y": "circ", "z": "mirror"}),  # odd: middle pencil
    ((16, 20, 9), {"x": "circ", "y": "wormhole", "z": "circ"}),               # partner z, odd
    ((15, 12, 16), {"x": "const", "y": "circ", "z": "wormhole"}),             # partner x, odd
    ((16, 12, 14), {"x": "trans", "y": ["wormhole", "wall"], "z": "const"}),  # partner z, even
])
def test_wormhole_lockstep(grid, bc):
    ini, st = box(grid, bc, seed=sum(grid))
    lockstep(ini, st, 8)


def test_wormhole_with_gravity_hachisu_setup():
    """HachisuDiskInitializer.m:123-127: z wormhole on both sides with gravity on."""
    ini, st = box((16, 16, 12), {"x": "const", "y": "const", "z": "wormhole"}, seed=77)
    lockstep(ini, st, 8, grav=True)


def test_wormhole_next_to_the_wave_sweep():
    """The exact wave-spread sweep takes the axes without a wormhole."""
    ini, st = box((32, 16, 16), {"x": "wormhole", "y": "circ", "z": "trans"}, seed=5,
                  limiter="superbee")
    lockstep(ini, st, 6, kind="wave")


def test_wormhole_with_neumann_richtmyer_viscosity():
    """The Neumann-Richtmyer term shifts mom(X) by +1 only: the wormhole's upper side
    (the constant shift).  Caramana-Shashkov-Whalen, which shifts both ways, is refused."""
    ini, st = box((20, 12, 10), {"x": "wormhole", "y": "circ", "z": "trans"}, seed=21)
    lockstep(ini, st, 6, visc=("neumann_richtmyer", 0.3, 0.6))
    run = make_gpu(ini, st)
    with pytest.raises(Exception, match="wormhole"):
        run.set_viscosity("caramana_shashkov_whalen", 0.3, 0.6)
    run.close()


def test_wormhole_fp32_close_to_fp64():
    ini, st = box((32, 24, 16), {"x": "circ", "y": "circ", "z": "wormhole"}, seed=9)
    r64 = make_gpu(ini, st, "fp64")
    r32 = make_gpu(ini, st, "fp32")
    r64.run_iterations(20)
    r32.run_iterations(20)
    e = rel_l1(r32.download("mass"), r64.download("mass"))
    print("wormhole FP32 vs FP64 density L1 after 20 iterations: %.3e" % e, flush=True)
    assert e <= 1e-5, e
    r64.close()
    r32.close()


def test_wormhole_in_a_batch_equals_alone():
    sims = [box((20, 14, 12), {"x": "wormhole", "y": "circ", "z": "const"}, seed=3),
            box((16, 16, 9), {"x": "circ", "y": "wormhole", "z": "wormhole"}, seed=4),
            problems.sod(256)]
    ens = Ensemble(sims, "fp64")
    ens.step(6)
    for i, (ini, st) in enumerate(sims):
        run = make_gpu(ini, st, "fp64")
        run.run_iterations(7)
        for f in FLUID_FIELDS:
            e = rel_linf(ens.download(i, f), run.download(f))
            assert e <= 1e-12, (i, f, e)
        run.close()
    ens.close()


def test_wormhole_long_pencils_pair_in_a_larger_tile():
    """FP64 pencils of 1024 cells fill a tile alone by default; with a wormhole two of
    them share a 512-thread block (four cells per thread)."""
    ini, st = box((1024, 4, 4), {"x": "wormhole", "y": "circ", "z": "circ"}, seed=2)
    lockstep(ini, st, 5)


def test_wormhole_refuses_pencils_too_long_to_pair():
    ini, st = box((4096, 4, 4), {"x": "wormhole", "y": "circ", "z": "circ"}, seed=1)
    with pytest.raises(Exception, match="wormhole"):
        make_gpu(ini, st, "fp64")
