"""Every tile the autotuner may pick gives the oracle's result.

The sweep kernel is compiled for a few (block size, KMAX) variants and takes the pencils
per block at run time; y / z sweeps with circular or constant boundaries can also run as
a column march.  Each candidate tile of `autotune.candidates` is forced on a plain FP64
context and the run is compared with the oracle in lockstep at <= 1e-12 (expected 0):
1-D MHD Riemann pencils (one pencil per simulation, tiles far longer than the pencil
included), 2-D y-sweeps (several pencils per block, column march), and the FP32 batch
with the single-pencil tiles of one launch shared (the rule of `autotune.apply`).
"""
import pytest

from imogen_amd import autotune, problems

from .helpers import FLUID_FIELDS, MAG_FIELDS, compare, make_gpu, make_oracle, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12


def lockstep_tile(ini, st, X, tile, iterations, fields):
    from imogen_amd import iteration
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    run.set_fluid_tile(X, *tile)
    worst = 0.0
    for it in range(iterations):
        orc.iteration()
        iteration(run)
        try:
            worst = max(worst, compare(run, orc, fields, TOL))
        except AssertionError as e:
            raise AssertionError("tile %s, iteration %d: %s" % (tile, it + 1, e))
    run.close()
    orc.close()
    return worst


@pytest.mark.parametrize("n", [300, 1000, 2900])
def test_riemann_every_candidate_tile(n):
    ini, st = problems.mhd_riemann(n, seed=n)
    tiles = autotune.candidates(autotune._pow2(n), 8, False, 1)
    # the longest variants hold pencils far shorter than themselves, as in a shared launch
    tiles += [(bs, k, 1) for esz, bs, k in autotune._build.fluid_variants()
              if esz == 8 and bs * k >= n and (bs, k, 1) not in tiles]
    assert tiles
    for tile in tiles:
        lockstep_tile(ini, st, 1, tile, 6, FLUID_FIELDS + MAG_FIELDS)


def test_2d_y_sweep_every_candidate_tile():
    """Hydro 48 x 40 with circular x and constant y boundaries: every y-sweep tile,
    the column march included."""
    grid = (48, 40, 1)
    st = problems.random_state(grid, seed=4, magnet=False)
    ini = {"grid": grid, "dGrid": (1.0 / 48, 1.0 / 40, 1.0 / 48), "gamma": 1.4, "cfl": 0.5,
           "bcMode": {"x": "circ", "y": "const", "z": "circ"}, "bcInfinity": 20,
           "fluxLimiter": "superbee", "minMass": 1e-5,
           "mode": {"fluid": True, "magnet": False}, "iterMax": 100, "timeMax": 1e9}
    tiles = autotune.candidates(autotune._pow2(40), 8, True, 48)
    assert (0, 0, 0) in tiles  # the column march is a candidate here
    for tile in tiles:
        lockstep_tile(ini, st, 2, tile, 6, FLUID_FIELDS)


def test_batch_shared_single_pencil_tile_fp32():
    """FP32 batch: short 1-D pencils on the tile of the longest one (one shared launch)
    against their default tiles; the per-cell arithmetic is the same, so the results
    agree to FP32 rounding of the operation order."""
    from imogen_amd.ensemble import Ensemble
    sims = [problems.mhd_riemann(n, seed=n, iter_max=10 ** 8) for n in (260, 700, 1500, 4000)]
    a = Ensemble(sims, "fp32")
    b = Ensemble(sims, "fp32")
    for i in range(len(sims)):
        b.set_fluid_tile(i, 1, 1024, 4, 1)
    a.step(8)
    b.step(8)
    for i in range(len(sims)):
        for f in ("mass", "ener", "magy"):
            assert rel_linf(b.download(i, f), a.download(i, f)) <= 1e-5, (i, f)
    a.close()
    b.close()
