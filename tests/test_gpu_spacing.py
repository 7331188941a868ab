"""Non-uniform grid spacing on the GPU against the oracle (ImogenManager.setGridSpacing:
DGRID{i} a line along i, MINDGRID(i) its minimum; fluxFactor = dt ./ DGRID per cell).

Lockstep at <= 1e-12 relative L-inf per iteration (observed 0) for hydro and MHD in 2-D
and 3-D (the fluid sweep's per-cell factors, the magnet advection's per-cell factors and
the CT's per-pencil DGRID{I}), gravity (the potential derivative per cell), the exact
wave-spread sweep on the uniform axes; a batch member equals the same simulation alone;
FP32 bulk within 1e-5 of FP64.
"""
import numpy as np
import pytest

from imogen_amd import iteration, problems
from imogen_amd.ensemble import Ensemble

from .helpers import FLUID_FIELDS, MAG_FIELDS, compare, make_gpu, make_oracle, rel_l1, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12
CORRUGATION = [[0, 5], [33.3, 1], [66.6, 1], [100, 5]]   # CorrugationShockInitializer.m:85


def box(grid, dgrid, bc="circ", magnet=False, seed=1):
    st = problems.random_state(grid, seed=seed, magnet=magnet)
    ini = {"grid": grid, "dGrid": dgrid, "gamma": 5.0 / 3.0, "cfl": 0.5, "bcMode": bc,
           "bcInfinity": 20, "fluxLimiter": {"x": "vanleer", "y": "superbee", "z": "minmond"},
           "minMass": 1e-5, "mode": {"fluid": True, "magnet": magnet}, "iterMax": 1000,
           "timeMax": 1e9}
    return ini, st


def potential(grid):
    nx, ny, nz = grid
    X = (np.arange(nx) + 0.5) / nx
    Y = (np.arange(ny) + 0.5) / ny
    Z = (np.arange(nz) + 0.5) / nz
    z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
    return 0.5 * np.sin(2 * np.pi * x) + 0.2 * y * y + 0.1 * z


def lockstep(ini, st, iterations, fields, kind=None, grav=False):
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    if kind:
        run.set_fluid_kernel(kind)
    if grav:
        phi = potential(tuple(ini["grid"]))
        orc.set_gravity(phi, 0.0)
        run.set_gravity(phi, 0.0)
    worst = 0.0
    for it in range(iterations):
        orc.iteration()
        iteration(run)
        try:
            worst = max(worst, compare(run, orc, fields, TOL))
        except AssertionError as e:
            raise AssertionError("%s iteration %d: %s" % (ini["grid"], it + 1, e))
        assert abs(run.time.time - orc.time) <= TOL * abs(orc.time), it
    print("spacing %s: max relative L-inf over %d iterations %.3e" % (
        ini["grid"], iterations, worst), flush=True)
    run.close()
    orc.close()


def test_hydro_2d_corrugation_spacing():
    ini, st = box((64, 24, 1), {"x": {"points": CORRUGATION}, "value": 0.02},
                  bc={"x": "trans", "y": "circ", "z": "circ"})
    lockstep(ini, st, 10, FLUID_FIELDS)


def test_mhd_2d_both_axes():
    ini, st = box((32, 40, 1), {"x": {"points": [[0, 2], [50, 1], [100, 3]]},
                                "y": {"points": [[0, 1], [100, 2]], "value": 0.03},
                                "value": 0.025}, magnet=True, seed=2)
    lockstep(ini, st, 8, FLUID_FIELDS + MAG_FIELDS)


def test_mhd_3d_every_axis():
    ini, st = box((16, 12, 10), {"x": {"points": [[0, 1.5], [100, 1]]},
                                 "y": {"points": [[0, 1], [40, 2], [100, 1]]},
                                 "z": {"points": [[0, 1], [100, 1.8]]}, "value": 0.05},
                  bc={"x": "trans", "y": "mirror", "z": "circ"}, magnet=True, seed=3)
    lockstep(ini, st, 5, FLUID_FIELDS + MAG_FIELDS)


def test_gravity_with_spacing():
    ini, st = box((40, 32, 1), {"x": {"points": CORRUGATION}, "value": 0.025},
                  bc={"x": "const", "y": "circ", "z": "circ"}, seed=4)
    lockstep(ini, st, 8, FLUID_FIELDS, grav=True)


def test_wave_sweep_on_the_uniform_axes():
    ini, st = box((32, 64, 1), {"y": {"points": [[0, 1], [100, 2]]}, "value": 0.02},
                  bc={"x": "circ", "y": "trans", "z": "circ"}, seed=5)
    lockstep(ini, st, 8, FLUID_FIELDS, kind="wave")


def test_batch_member_equals_alone():
    sims = [box((48, 20, 1), {"x": {"points": CORRUGATION}, "value": 0.02}, seed=6),
            box((24, 24, 1), {"y": {"points": [[0, 1], [100, 2]]}, "value": 0.03},
                magnet=True, seed=7),
            problems.sod(256)]
    ens = Ensemble(sims, "fp64")
    ens.step(6)
    for i, (ini, st) in enumerate(sims):
        run = make_gpu(ini, st, "fp64")
        run.run_iterations(7)
        t_b, _, _ = ens.time(i)
        assert abs(t_b - run.time.time) <= 1e-12 * abs(run.time.time), i
        fields = FLUID_FIELDS + (MAG_FIELDS if ini["mode"]["magnet"] else [])
        for f in fields:
            e = rel_linf(ens.download(i, f), run.download(f))
            assert e <= 1e-12, (i, f, e)
        run.close()
    ens.close()


def test_fp32_close_to_fp64():
    ini, st = box((64, 48, 1), {"x": {"points": CORRUGATION}, "value": 0.02}, seed=8)
    r64 = make_gpu(ini, st, "fp64")
    r32 = make_gpu(ini, st, "fp32")
    r64.run_iterations(20)
    r32.run_iterations(20)
    e = rel_l1(r32.download("mass"), r64.download("mass"))
    print("spacing FP32 vs FP64 density L1 after 20 iterations: %.3e" % e, flush=True)
    assert e <= 1e-5, e
    r64.close()
    r32.close()
