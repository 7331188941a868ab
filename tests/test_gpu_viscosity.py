"""Neumann-Richtmyer artificial viscosity in the fused sweep against the oracle, in
lockstep at <= 1e-12 relative L-inf per iteration: a Sod tube with the reference's
debug-run strengths (run_Debug.m:17-19), 2-D hydro with transparent ends (the momentum
FluidArray's and StorageArray's trackers), 2-D MHD with mirror / wall ends (sign flips of
the shifted momentum, stored wall edges), and 3-D MHD with gravity and radiation on too.
Graph replay equals eager iterations, a checkpoint resumes bit for bit, FP32-bulk stays
close to FP64.  The Caramana-Shashkov-Whalen form in lockstep too: a Sod tube, 2-D MHD
with circ / const / mirror / flip / zero ends, 3-D MHD with gravity and radiation, and a
1-D problem on every candidate tile, and with transparent and wall ends (the trackers of
every shifted object); with fade ends on 1-D and 2-D grids (the multi-launch sweep,
cell by cell), refused on 3-D grids and in a batch."""
import numpy as np
import pytest

from imogen_amd import ImogenError, iteration, problems

from .helpers import FLUID_FIELDS, MAG_FIELDS, compare, make_gpu, make_oracle, rel_l1
from .test_gpu_source import box, potential

pytestmark = pytest.mark.gpu

TOL = 1e-12
NR = (0.1, 0.25)


def lockstep(ini, st, iterations, fields, visc=NR, grav=False, rad=None):
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    orc.set_viscosity("neumann_richtmyer", *visc)
    run.set_viscosity("neumann_richtmyer", *visc)
    if grav:
        phi = potential(tuple(ini["grid"]))
        orc.set_gravity(phi)
        run.set_gravity(phi)
    if rad:
        orc.set_radiation("optically_thin", *rad)
        run.set_radiation("optically_thin", *rad)
    for it in range(iterations):
        print("iteration", it + 1, flush=True)
        orc.iteration()
        iteration(run)
        try:
            compare(run, orc, fields, TOL)
        except AssertionError as e:
            raise AssertionError("iteration %d: %s" % (it + 1, e))
    run.close()
    orc.close()


def test_sod_lockstep():
    ini, st = problems.sod(1024, iter_max=10 ** 6)
    lockstep(ini, st, 30, FLUID_FIELDS)


def test_2d_hydro_transparent_lockstep():
    ini, st = box((48, 40, 1), {"x": "trans", "y": "circ", "z": "circ"})
    lockstep(ini, st, 6, FLUID_FIELDS, visc=(0.3, 0.6))


def test_2d_mhd_mirror_wall_lockstep():
    ini, st = box((40, 32, 1), {"x": "mirror", "y": "wall", "z": "circ"}, magnet=True)
    lockstep(ini, st, 5, FLUID_FIELDS + MAG_FIELDS, visc=(0.3, 0.6))


def test_3d_mhd_with_sources_lockstep():
    ini, st = box((24, 20, 16), {"x": "trans", "y": "circ", "z": "const"}, magnet=True)
    lockstep(ini, st, 3, FLUID_FIELDS + MAG_FIELDS, visc=(0.3, 0.6), grav=True, rad=(0.5, 1.0))


def test_graph_replay_and_checkpoint_with_viscosity():
    ini, st = box((48, 40, 1), {"x": "trans", "y": "trans", "z": "circ"})
    a = make_gpu(ini, st)
    b = make_gpu(ini, st)
    for r in (a, b):
        r.set_viscosity("neumann_richtmyer", 0.3, 0.6)
    for _ in range(8):
        iteration(a)
    assert b.run_iterations(4) == 4
    ck = b.checkpoint()
    b.close()
    c = make_gpu(ini, st)
    c.set_viscosity("neumann_richtmyer", 0.3, 0.6)
    c.restore(ck)
    assert c.run_iterations(4) == 4
    for f in FLUID_FIELDS:
        np.testing.assert_array_equal(c.download(f), a.download(f), err_msg=f)
    a.close()
    c.close()


def test_fp32_viscosity_close_to_fp64():
    ini, st = box((48, 40, 1), {"x": "trans", "y": "circ", "z": "circ"})
    runs = []
    for prec in ("fp64", "fp32"):
        r = make_gpu(ini, st, prec)
        r.set_viscosity("neumann_richtmyer", 0.3, 0.6)
        for _ in range(10):
            iteration(r)
        runs.append(r)
    for f in ("mass", "momx", "ener"):
        assert rel_l1(runs[1].download(f), runs[0].download(f)) < 1e-5, f
    for r in runs:
        r.close()


def csw_lockstep(ini, st, iterations, fields, visc=(0.3, 0.8), tile=None, **kw):
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    if tile:
        run.set_fluid_tile(*tile)
    orc.set_viscosity("caramana_shashkov_whalen", *visc)
    run.set_viscosity("caramana_shashkov_whalen", *visc)
    if kw.get("grav"):
        phi = potential(tuple(ini["grid"]))
        orc.set_gravity(phi)
        run.set_gravity(phi)
    if kw.get("rad"):
        orc.set_radiation("optically_thin", *kw["rad"])
        run.set_radiation("optically_thin", *kw["rad"])
    for it in range(iterations):
        print("iteration", it + 1, "tile", tile, flush=True)
        orc.iteration()
        iteration(run)
        try:
            compare(run, orc, fields, TOL)
        except AssertionError as e:
            raise AssertionError("tile %s iteration %d: %s" % (tile, it + 1, e))
    run.close()
    orc.close()


def test_csw_sod_lockstep():
    ini, st = problems.sod(1024, bc="const", iter_max=10 ** 6)
    csw_lockstep(ini, st, 30, FLUID_FIELDS, visc=(0.1, 0.25))


@pytest.mark.parametrize("bc", [
    {"x": "circ", "y": "const", "z": "circ"},
    {"x": "mirror", "y": "flip", "z": "circ"},
    {"x": "zero", "y": "circ", "z": "circ"},
])
def test_csw_2d_lockstep(bc):
    ini, st = box((48, 40, 1), bc, magnet=True)
    csw_lockstep(ini, st, 5, FLUID_FIELDS + MAG_FIELDS)


def test_csw_3d_mhd_with_sources_lockstep():
    ini, st = box((24, 20, 16), {"x": "circ", "y": "mirror", "z": "const"}, magnet=True)
    csw_lockstep(ini, st, 3, FLUID_FIELDS + MAG_FIELDS, grav=True, rad=(0.5, 1.0))


def test_csw_every_candidate_tile():
    """1-D pencils of 300 cells on every tile the autotuner may pick (K = 1..8 cells per
    thread, several pencils per block): the neighbour rows cross chunk edges."""
    from imogen_amd import autotune
    ini, st = problems.mhd_riemann(300, seed=3)
    ini = dict(ini, bcMode={"x": "const", "y": "circ", "z": "circ"})
    tiles = [t for t in autotune.candidates(autotune._pow2(300), 8, False, 1) if t[0] > 0]
    assert tiles
    for bs, k, px in tiles:
        csw_lockstep(ini, st, 3, FLUID_FIELDS + MAG_FIELDS, tile=(1, bs, k, px))


@pytest.mark.parametrize("bc", [
    {"x": "trans", "y": "circ", "z": "circ"},
    {"x": "wall", "y": "trans", "z": "circ"},
    {"x": ["trans", "mirror"], "y": ["const", "wall"], "z": "circ"},
])
def test_csw_tracked_ends_lockstep(bc):
    """Transparent and wall ends: every shift call of the CSW term has its own ghost from
    its object's tracker (mass, momentum, energy and cell-centred B, FluidArrays in the
    predictor and StorageArrays in the corrector)."""
    ini, st = box((48, 40, 1), bc, magnet=True)
    csw_lockstep(ini, st, 6, FLUID_FIELDS + MAG_FIELDS)


def test_csw_transparent_sod_and_3d():
    ini, st = problems.sod(1024, iter_max=10 ** 6)  # transparent ends
    csw_lockstep(ini, st, 30, FLUID_FIELDS, visc=(0.1, 0.25))
    ini, st = box((24, 20, 16), {"x": "trans", "y": "wall", "z": "trans"}, magnet=True)
    csw_lockstep(ini, st, 3, FLUID_FIELDS + MAG_FIELDS, grav=True, rad=(0.5, 1.0))


def test_csw_transparent_checkpoint_resume():
    ini, st = box((48, 40, 1), {"x": "trans", "y": "trans", "z": "circ"}, magnet=True)
    a = make_gpu(ini, st)
    a.set_viscosity("caramana_shashkov_whalen", 0.3, 0.8)
    for _ in range(8):
        iteration(a)
    b = make_gpu(ini, st)
    b.set_viscosity("caramana_shashkov_whalen", 0.3, 0.8)
    assert b.run_iterations(4) == 4
    ck = b.checkpoint()
    b.close()
    c = make_gpu(ini, st)
    c.set_viscosity("caramana_shashkov_whalen", 0.3, 0.8)
    c.restore(ck)
    assert c.run_iterations(4) == 4
    for f in FLUID_FIELDS + MAG_FIELDS:
        np.testing.assert_array_equal(c.download(f), a.download(f), err_msg=f)
    a.close()
    c.close()


def test_neumann_richtmyer_fade_lockstep():
    """Fade ends: the momentum shift is blended toward its end values (fade_shift.m)."""
    ini, st = box((48, 40, 1), {"x": "fade", "y": ["const", "fade"], "z": "circ"})
    lockstep(ini, st, 6, FLUID_FIELDS, visc=(0.3, 0.6))


@pytest.mark.parametrize("bc", [
    {"x": "fade", "y": "circ", "z": "circ"},
    {"x": ["fade", "trans"], "y": ["const", "fade"], "z": "circ"},
    {"x": ["mirror", "fade"], "y": ["fade", "wall"], "z": "circ"},
])
def test_csw_fade_lockstep(bc):
    """Caramana-Shashkov-Whalen at fade ends: every array it shifts (mass, momentum,
    energy, cell-centred B) is blended near both ends of the shifted array, so its sound
    speed and velocities there come from blended arrays; these sweeps take the multi-launch
    form, cell by cell."""
    ini, st = box((48, 40, 1), bc, magnet=True)
    ini = dict(ini, bcInfinity=12)
    csw_lockstep(ini, st, 6, FLUID_FIELDS + MAG_FIELDS)


@pytest.mark.skip(reason="3-D grids refuse the multi-launch sweep: this case faulted the "
                         "device once while a serial host run of the same kernels kept every "
                         "access in bounds; cause not found")
def test_csw_fade_3d_with_sources_lockstep():
    ini, st = box((24, 20, 16), {"x": "fade", "y": "mirror", "z": ["trans", "fade"]},
                  magnet=True)
    ini = dict(ini, bcInfinity=8)
    csw_lockstep(ini, st, 3, FLUID_FIELDS + MAG_FIELDS, grav=True, rad=(0.5, 1.0))


def test_csw_fade_3d_is_refused():
    ini, st = box((24, 20, 16), {"x": "fade", "y": "mirror", "z": "circ"}, magnet=True)
    ini = dict(ini, bcInfinity=8)
    run = make_gpu(ini, st)
    with pytest.raises(ImogenError, match="1-D and 2-D grids"):
        run.set_viscosity("caramana_shashkov_whalen", 0.3, 0.8)
    run.set_viscosity("neumann_richtmyer", 0.3, 0.8)  # the fused sweep takes it
    run.close()


def test_csw_fade_refused_in_a_batch():
    from imogen_amd.ensemble import Ensemble
    ini, st = box((32, 24, 1), {"x": "fade", "y": "circ", "z": "circ"})
    ini = dict(ini, bcInfinity=8)
    ens = Ensemble([(ini, st), problems.sod(256)], "fp64")
    with pytest.raises(ImogenError):
        ens.set_viscosity(0, "caramana_shashkov_whalen", 0.1, 0.2)
    ens.close()
