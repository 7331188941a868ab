"""Pencils longer than a workgroup holds (fluid_long.hip) against the oracle.

The fused sweep keeps a pencil in one workgroup (4096 cells); longer ones take the
multi-launch sweep with device-wide pencil maxima of the freezing speed.  Lockstep with
the oracle at <= 1e-12 relative L-inf per iteration (observed 0): Sod on 8192 cells, every
boundary mode on 5000-cell pencils, 2-D grids whose x pencils are long (the first sweep's
edge records on the y faces included), fade ends (the shifted fluxes blended over
bcInfinity cells), the fluid's B terms (magnet off, field on), both artificial
viscosities (fade ends included); a
graph replay equals the eager iterations; FP32 bulk within 1e-5 of FP64; the refusals
(a lower wormhole end, magnet pencils past 4096 cells, 3-D grids, batches).
"""
import numpy as np
import pytest

from imogen_amd import iteration, problems
from imogen_amd.ensemble import Ensemble

from .helpers import FLUID_FIELDS, compare, make_gpu, make_oracle, rel_l1, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12


def box(grid, bc, seed, magnet_field=False, limiter="vanleer"):
    st = problems.random_state(grid, seed=seed, magnet=magnet_field)
    h = 1.0 / max(grid)
    ini = {"grid": grid, "dGrid": (h, h, h), "gamma": 1.4, "cfl": 0.6, "bcMode": bc,
           "bcInfinity": 20, "fluxLimiter": limiter, "minMass": 1e-5,
           "mode": {"fluid": True, "magnet": False}, "iterMax": 1000, "timeMax": 1e9}
    return ini, st


def lockstep(ini, st, iterations):
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    worst = 0.0
    for it in range(iterations):
        orc.iteration()
        iteration(run)
        try:
            worst = max(worst, compare(run, orc, FLUID_FIELDS, TOL))
        except AssertionError as e:
            raise AssertionError("%s iteration %d: %s" % (ini["grid"], it + 1, e))
        assert abs(run.time.time - orc.time) <= TOL * abs(orc.time), it
    print("long pencils %s %s: max relative L-inf over %d iterations %.3e" % (
        ini["grid"], ini["bcMode"], iterations, worst), flush=True)
    run.close()
    orc.close()


def test_sod_8192():
    ini, st = problems.sod(8192)
    lockstep(ini, st, 20)


@pytest.mark.parametrize("bc", ["trans", "const", "mirror", "flip", "wall", "zero", "circ",
                                ["trans", "wall"], ["zero", "mirror"], ["circ", "const"],
                                "fade", ["fade", "trans"], ["mirror", "fade"]])
def test_boundary_modes_5000(bc):
    ini, st = box((5000, 1, 1), {"x": bc, "y": "circ", "z": "circ"}, seed=3,
                  limiter="superbee")
    lockstep(ini, st, 6)


def test_2d_long_x_pencils_with_transparent_y():
    ini, st = box((6000, 6, 1), {"x": ["const", "trans"], "y": "trans", "z": "circ"}, seed=5,
                  limiter="minmond")
    lockstep(ini, st, 5)


def test_fluid_b_terms_on_long_pencils():
    ini, st = box((4500, 1, 1), {"x": "trans", "y": "circ", "z": "circ"}, seed=7,
                  magnet_field=True)
    lockstep(ini, st, 6)


@pytest.mark.parametrize("kind,bc", [
    ("neumann_richtmyer", "trans"), ("neumann_richtmyer", ["const", "fade"]),
    ("caramana_shashkov_whalen", "const"), ("caramana_shashkov_whalen", ["trans", "wall"]),
    ("caramana_shashkov_whalen", ["fade", "mirror"]),
])
def test_viscosity_on_long_pencils(kind, bc):
    ini, st = box((5000, 1, 1), {"x": bc, "y": "circ", "z": "circ"}, seed=9,
                  magnet_field=True)
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    orc.set_viscosity(kind, 0.2, 0.5)
    run.set_viscosity(kind, 0.2, 0.5)
    for it in range(6):
        orc.iteration()
        iteration(run)
        compare(run, orc, FLUID_FIELDS, TOL)
    run.close()
    orc.close()


def test_graph_replay_equals_eager():
    ini, st = problems.sod(8192)
    a = make_gpu(ini, st)
    b = make_gpu(ini, st)
    a.run_iterations(12)
    for _ in range(12):
        iteration(b)
    for f in FLUID_FIELDS:
        np.testing.assert_array_equal(a.download(f), b.download(f), err_msg=f)
    a.close()
    b.close()


def test_fp32_close_to_fp64():
    ini, st = problems.sod(8192)
    r64 = make_gpu(ini, st, "fp64")
    r32 = make_gpu(ini, st, "fp32")
    r64.run_iterations(100)
    r32.run_iterations(100)
    e = rel_l1(r32.download("mass"), r64.download("mass"))
    print("long pencils FP32 vs FP64 density L1 after 100 iterations: %.3e" % e, flush=True)
    assert e <= 1e-5, e
    r64.close()
    r32.close()


def test_refusals():
    ini, st = box((5000, 8, 4), {"x": "wormhole", "y": "circ", "z": "circ"}, seed=1)
    with pytest.raises(Exception, match="wormhole"):
        make_gpu(ini, st)
    ini, st = box((5000, 4, 3), {"x": "trans", "y": "circ", "z": "circ"}, seed=1)
    with pytest.raises(Exception, match="1-D and 2-D grids"):
        make_gpu(ini, st)
    ini, st = problems.mhd_riemann(256, seed=1)
    ini = dict(ini, grid=(5000, 1, 1))
    st = problems.random_state((5000, 1, 1), seed=2, magnet=True)
    with pytest.raises(Exception, match="magnet pencils"):
        make_gpu(ini, st)
    with pytest.raises(Exception, match="run on their own"):
        Ensemble([problems.sod(8192), problems.sod(256)], "fp64")
