"""The wave-spread fluid sweep (fluid_wave_impl.h) against the oracle.

* Exact arithmetic (sweep kind 'wave'): every division of the .m text in its order, no
  contraction, so the run matches the oracle in lockstep at <= 1e-12 relative L-inf per
  iteration (observed 0), over the layouts the kernel has: x pencils inside one wave
  (several per wave, or one of 256 cells), x pencils spread over waves (1024 cells; a
  512-cell 1-D pencil in a block with a padding pencil), y / z tiles of interleaved
  x-adjacent pencils (several lanes apart, crossing wave edges), and every boundary mode
  it takes (fade ends and circular-against-other ends go to the general sweep).  MHD
  sweeps are switched off in the wave path (see MHD_OFF): their cases are skipped.
* Fast FP64 arithmetic (sweep kind 'wave_fast', the plain path's sweep): one step from
  the same state as the exact sweep (the state, edge trackers and time copied over by a
  checkpoint) agrees at <= 1e-12 relative L-inf per field, every step of a run.
"""
import numpy as np
import pytest

from imogen_amd import problems

from .helpers import FLUID_FIELDS, MAG_FIELDS, compare, make_gpu, make_oracle, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12

# A run of these tests left the GPU faulted at the Orszag-Tang (MHD) case; the cause has
# not been found.  MHD sweeps are switched off in the wave-spread path (they take the
# general sweep) and the MHD cases stay off the GPU until the cause is shown.
MHD_OFF = pytest.mark.skip(reason="wave-spread MHD sweeps are switched off: a run of the "
                                  "MHD cases faulted the GPU, cause not found")


def lockstep_wave(ini, st, iterations, fields, kind="wave"):
    from imogen_amd import iteration
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    run.set_fluid_kernel(kind)
    worst = 0.0
    for it in range(iterations):
        orc.iteration()
        iteration(run)
        try:
            worst = max(worst, compare(run, orc, fields, TOL))
        except AssertionError as e:
            raise AssertionError("%s iteration %d: %s" % (ini["grid"], it + 1, e))
        assert abs(run.time.time - orc.time) <= TOL * abs(orc.time), it
    print("wave sweep %s: max relative L-inf over %d iterations %.3e" % (
        ini["grid"], iterations, worst), flush=True)
    run.close()
    orc.close()
    return worst


def box(grid, bc, limiter, magnet, seed, gamma=5.0 / 3.0, cfl=0.35):
    st = problems.random_state(grid, seed=seed, magnet=magnet)
    # one spacing for every direction: TimeManager.update takes dt from the spacing of the
    # direction with the largest speed (z included), so a coarser z would break the CFL
    # condition of a finer direction
    h = 1.0 / max(grid)
    ini = {"grid": grid, "dGrid": (h, h, h),
           "gamma": gamma, "cfl": cfl, "bcMode": bc, "bcInfinity": 20,
           "fluxLimiter": limiter, "minMass": 1e-5,
           "mode": {"fluid": True, "magnet": magnet}, "iterMax": 100, "timeMax": 1e9}
    return ini, st


def test_sod_one_wave_per_pencil():
    ini, st = problems.sod(256, limiter="minmond", bc="const")
    lockstep_wave(ini, st, 30, FLUID_FIELDS)


def test_sod_1024_spread_over_waves():
    ini, st = problems.sod(1024)
    lockstep_wave(ini, st, 40, FLUID_FIELDS)


@MHD_OFF
def test_mhd_riemann_512_padding_pencil():
    ini, st = problems.mhd_riemann(512, seed=5)
    lockstep_wave(ini, st, 15, FLUID_FIELDS + MAG_FIELDS)


def test_kelvin_helmholtz_64():
    ini, st = problems.kelvin_helmholtz(64, seed=3)
    lockstep_wave(ini, st, 15, FLUID_FIELDS)


@MHD_OFF
def test_orszag_tang_64():
    ini, st = problems.orszag_tang(64)
    lockstep_wave(ini, st, 12, FLUID_FIELDS + MAG_FIELDS)


def test_sedov_32_transparent():
    ini, st = problems.sedov(32, time_max=1.0)
    lockstep_wave(ini, st, 8, FLUID_FIELDS)


def test_y_pencils_across_wave_edges():
    """y pencils of 512 cells (two pencils per tile, lanes two apart, 32 lanes of a
    pencil per wave) and x pencils of 16 cells (sixteen per wave)."""
    ini, st = box((16, 512, 1), {"x": "circ", "y": "trans"}, "vanleer", False, 21, 1.4, 0.5)
    lockstep_wave(ini, st, 10, FLUID_FIELDS)


BC_CASES = [
    ("trans", "trans"), ("const", "trans"), ("mirror", "mirror"), ("flip", "flip"),
    ("wall", "wall"), ("zero", "zero"), ("circ", "circ"),
    (["zero", "mirror"], ["flip", "wall"]), (["trans", "wall"], ["const", "mirror"]),
]


@MHD_OFF
@pytest.mark.parametrize("bcx,bcy", BC_CASES)
def test_bc_modes_2d_mhd(bcx, bcy):
    ini, st = box((32, 64, 1), {"x": bcx, "y": bcy, "z": "circ"},
                  {"x": "vanleer", "y": "superbee"}, True, 11)
    lockstep_wave(ini, st, 8, FLUID_FIELDS + MAG_FIELDS)


@pytest.mark.parametrize("bcx,bcy", BC_CASES)
def test_bc_modes_2d_hydro(bcx, bcy):
    ini, st = box((32, 64, 1), {"x": bcx, "y": bcy, "z": "circ"},
                  {"x": "vanleer", "y": "superbee"}, False, 12, 1.4, 0.6)
    lockstep_wave(ini, st, 8, FLUID_FIELDS)


@pytest.mark.parametrize("bc", ["trans", "mirror", "const", "wall"])
def test_3d_hydro_bc(bc):
    ini, st = box((32, 64, 32), bc, "minmond", False, 7, 1.4, 0.7)
    lockstep_wave(ini, st, 4, FLUID_FIELDS)


@MHD_OFF
def test_3d_mhd_trans():
    ini, st = box((32, 32, 32), "trans", "vanleer", True, 9)
    lockstep_wave(ini, st, 4, FLUID_FIELDS + MAG_FIELDS)


def step_from_same_state(ini, st, steps, fields):
    """Fast FP64 sweep against the exact sweep, each step from the same state."""
    from imogen_amd import iteration
    ref = make_gpu(ini, st)
    ref.set_fluid_kernel("wave")
    fast = make_gpu(ini, st)
    fast.set_fluid_kernel("wave_fast")
    worst = 0.0
    iteration(ref)  # the first iteration (general form) is shared
    for _ in range(steps):
        fast.restore(ref.checkpoint())
        iteration(ref)
        iteration(fast)
        for f in fields:
            e = rel_linf(fast.download(f), ref.download(f))
            assert e <= TOL, (ini["grid"], f, e)
            worst = max(worst, e)
    print("fast FP64 sweep %s: max relative L-inf per step %.3e" % (ini["grid"], worst),
          flush=True)
    ref.close()
    fast.close()
    return worst


@pytest.mark.parametrize("case", ["sod", "kh", "sedov",
                                  pytest.param("ot", marks=MHD_OFF),
                                  pytest.param("riemann", marks=MHD_OFF)])
def test_fast_fp64_step_from_same_state(case):
    ini, st = {"sod": lambda: problems.sod(1024),
               "kh": lambda: problems.kelvin_helmholtz(128, seed=4),
               "ot": lambda: problems.orszag_tang(128),
               "sedov": lambda: problems.sedov(64, time_max=1.0),
               "riemann": lambda: problems.mhd_riemann(2048, seed=8)}[case]()
    mag = ini["mode"]["magnet"]
    step_from_same_state(ini, st, 6, FLUID_FIELDS + (MAG_FIELDS if mag else []))


def test_fast_fp64_run_tracks_the_oracle():
    """Over a whole Sod run the fast sweep's differences from the oracle stay at rounding
    level (no accumulation to visible error)."""
    from imogen_amd import iteration
    ini, st = problems.sod(1024)
    orc = make_oracle(ini, st)
    run = make_gpu(ini, st)
    run.set_fluid_kernel("wave_fast")
    for _ in range(200):
        orc.iteration()
        iteration(run)
    e = max(rel_linf(run.download(f), orc.get(f)) for f in ("mass", "momx", "ener"))
    print("fast FP64 Sod 1024 after 200 iterations: relative L-inf %.3e" % e, flush=True)
    assert e <= 1e-10, e
    run.close()
    orc.close()
