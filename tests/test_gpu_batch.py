"""The batched ensemble path: a simulation advanced inside a batch of uneven simulations
gives the same result as the same simulation on its own, and the FP32-bulk path stays
within the stated error of the FP64 oracle.

Tolerances:
  * FP64 batch vs FP64 alone: <= 1e-12 relative L-inf per field (the same device bodies
    and argument blocks run in both, so differences are expected to be exactly 0);
  * FP32 bulk vs the FP64 oracle: L1 density error <= 1e-4 after 40 iterations.
"""
import numpy as np
import pytest

from imogen_amd import problems
from imogen_amd.ensemble import Ensemble

from .helpers import FLUID_FIELDS, MAG_FIELDS, make_gpu, make_oracle, rel_l1, rel_linf

pytestmark = pytest.mark.gpu


def mixed_sims():
    """Uneven grids, dimensions, boundary conditions and limiters in one batch."""
    sims = [problems.mhd_riemann(n, seed=s) for s, n in enumerate([256, 700, 1500, 4096])]
    sims += [problems.kelvin_helmholtz(n, seed=10 + n) for n in (64, 96)]
    sims.append(problems.orszag_tang(48))
    sims.append(problems.sod(512, limiter="minmond"))
    sims.append(problems.sedov(16, time_max=1.0))
    ini, st = problems.sod(200, ny=8, bc="trans", limiter="superbee")
    sims.append((ini, st))
    return sims


def fields_of(ini):
    return FLUID_FIELDS + (MAG_FIELDS if ini["mode"]["magnet"] else [])


@pytest.mark.parametrize("iters", [1, 13])
def test_batch_fp64_matches_alone(iters):
    sims = mixed_sims()
    ens = Ensemble(sims, "fp64")
    ens.step(iters)  # the first iteration runs alone, then `iters` batched ones
    worst = 0.0
    for i, (ini, st) in enumerate(sims):
        run = make_gpu(ini, st, "fp64")
        run.run_iterations(iters + 1)
        t_b, dt_b, it_b = ens.time(i)
        assert it_b == run.time.iteration, (i, it_b, run.time.iteration)
        assert abs(t_b - run.time.time) <= 1e-12 * abs(run.time.time), i
        for f in fields_of(ini):
            e = rel_linf(ens.download(i, f), run.download(f))
            assert e <= 1e-12, (i, f, e)
            worst = max(worst, e)
        run.close()
    print("batch vs alone worst relative L-inf: %.3e" % worst, flush=True)
    ens.close()


def test_batch_fp32_error_vs_oracle():
    sims = [problems.mhd_riemann(n, seed=100 + n) for n in (512, 2048)]
    sims += [problems.kelvin_helmholtz(128, seed=5), problems.orszag_tang(64)]
    iters = 40
    ens = Ensemble(sims, "fp32")
    ens.step(iters - 1)
    for i, (ini, st) in enumerate(sims):
        orc = make_oracle(ini, st)
        for _ in range(iters):
            orc.iteration()
        assert ens.time(i)[2] == orc.iteration_count
        e1 = rel_l1(ens.download(i, "mass"), orc.get("mass"))
        einf = rel_linf(ens.download(i, "mass"), orc.get("mass"))
        print("sim %d %s: density L1 %.2e L-inf %.2e" % (i, ini["grid"], e1, einf), flush=True)
        assert e1 <= 1e-4, (i, e1)
    ens.close()


def test_fused_1d_lane_equals_per_sweep_launches():
    """The 1-D MHD simulations of an FP32 batch run their whole flux call in one launch
    per iteration (lane1d.hip); with that launch switched off they take the per-sweep
    launches.  Same bodies, same order: the fields agree bit for bit, at pencil lengths
    from 256 to 4096 cells, next to 2-D hydro and 2-D MHD simulations (not fused), and a
    simulation that ends early (TIMEMAX) stops in both."""
    sims = [problems.mhd_riemann(n, seed=30 + n) for n in (256, 700, 1500, 3000, 4096)]
    ini, st = problems.mhd_riemann(1024, seed=7)
    ini = dict(ini, timeMax=0.02)
    sims.append((ini, st))
    sims += [problems.kelvin_helmholtz(64, seed=8), problems.orszag_tang(48)]
    fused = Ensemble(sims, "fp32")
    plain = Ensemble(sims, "fp32")
    plain.set_fused_1d(False)
    fused.step(3)
    plain.step(3)
    pf = fused.profile_step(2)
    pp = plain.profile_step(2)
    assert pf["mhd_1d"][2] > 0 and pp["mhd_1d"][2] == 0, (pf["mhd_1d"], pp["mhd_1d"])
    fused.step(10)
    plain.step(10)
    for i, (ini_i, _) in enumerate(sims):
        assert fused.time(i) == plain.time(i), i
        for f in fields_of(ini_i):
            np.testing.assert_array_equal(fused.download(i, f), plain.download(i, f),
                                          err_msg="sim %d %s" % (i, f))
    print("1-D lane: profile fused %s, per-sweep fluid %s" % (pf["mhd_1d"], pp["fluid"]),
          flush=True)
    fused.close()
    plain.close()


def _as_fp32(st):
    """The state rounded to FP32; B on a grid of 2^-12, so that the cell-centred averages
    of neighbouring face values are exact in FP32 as in FP64 (the CFL reads them)."""
    f = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    q = lambda b: np.round(np.asarray(b, dtype=np.float64) * 4096.0) / 4096.0
    return {"mass": f(st["mass"]), "mom": [f(m) for m in st["mom"]], "ener": f(st["ener"]),
            "mag": [q(b) for b in st["mag"]]}


def _cold_gas(grid, seed):
    """Pressure a millionth of the kinetic and magnetic energy: E - rho v^2 / 2 - B^2 / 2
    cancels to six digits."""
    st = problems.random_state(grid, seed=seed, magnet=True)
    v2 = sum(m * m for m in st["mom"]) / st["mass"] ** 2
    b2 = sum(b * b for b in st["mag"])
    st["ener"] = (0.5 * st["mass"] * v2 + 0.5 * b2) * (1.0 + 1e-6)
    return st


@pytest.mark.parametrize("case", ["random_mhd", "cold_gas", "kh", "ot", "sedov"])
def test_fp32_cfl_equals_the_fp64_evaluation(case):
    """The FP32-bulk CFL reduction evaluates the speeds in FP64 from the FP32 state: the
    maxima, hence dt, equal the oracle's time update on the same FP32-rounded state (whose
    cell-centred B is then the same in both) bit for bit, a cold, nearly pressureless gas
    included."""
    if case in ("random_mhd", "cold_gas"):
        grid = (40, 33, 5)
        st = problems.random_state(grid, seed=3) if case == "random_mhd" else _cold_gas(grid, 4)
        ini = {"grid": grid, "dGrid": (0.05, 0.06, 0.07), "gamma": 5.0 / 3.0, "cfl": 0.5,
               "bcMode": "circ", "bcInfinity": 20, "fluxLimiter": "vanleer", "minMass": 1e-5,
               "mode": {"fluid": True, "magnet": True}, "iterMax": 100, "timeMax": 1e9}
    else:
        ini, st = {"kh": lambda: problems.kelvin_helmholtz(96, seed=2),
                   "ot": lambda: problems.orszag_tang(64),
                   "sedov": lambda: problems.sedov(24, time_max=1.0)}[case]()
    st = _as_fp32(st)
    run = make_gpu(ini, st, "fp32")
    orc = make_oracle(ini, st)
    run.time.update(None, None, None, None, 1)
    orc.time_update(1)
    assert run.time.dTime == orc.dtime, (run.time.dTime, orc.dtime)
    run.close()
    orc.close()


def test_batch_handles_finished_simulations():
    """A simulation whose TIMEMAX is reached stops advancing while the others go on."""
    a = problems.sod(256, time_max=0.01)
    b = problems.sod(256)
    ens = Ensemble([a, b], "fp64")
    ens.step(60)
    t_a, _, it_a = ens.time(0)
    t_b, _, it_b = ens.time(1)
    assert t_a == pytest.approx(0.01, rel=0, abs=1e-15)
    assert it_b == 62 and it_a < it_b
    run = make_gpu(*a, "fp64")
    run.run_iterations(1000)
    assert run.time.iteration == it_a
    np.testing.assert_array_equal(ens.download(0, "mass"), run.download("mass"))
    ens.close()
