"""The serving loop: requests admitted while others run give the same result as the same
request run alone, and admission respects the slot count and the HBM budget.

Tolerance: the result in the mix equals the result alone exactly (relative L-inf 0):
the same kernels run on the same per-simulation argument blocks, and the CFL maximum is
an exact (order-independent) reduction.
"""
import numpy as np
import pytest

from imogen_amd import problems
from imogen_amd.serve import DONE, Server, run_stream

from .helpers import make_gpu, rel_linf

pytestmark = pytest.mark.gpu


def requests():
    return [
        problems.sod(256, time_max=0.05),
        problems.mhd_riemann(512, seed=4),
        problems.kelvin_helmholtz(64, seed=2),
        problems.orszag_tang(48),
        problems.sedov(16, time_max=0.02),
        problems.mhd_riemann(1000, seed=9),
    ]


def limited(reqs, iters):
    out = []
    for ini, st in reqs:
        ini = dict(ini)
        ini["iterMax"] = min(ini["iterMax"], iters)
        out.append((ini, st))
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_result_alone_equals_result_in_mix(precision):
    reqs = limited(requests(), 40)
    srv = Server(precision, max_active=4)
    ids = [srv.submit(*reqs[i]) for i in range(3)]
    srv.step(5)
    srv.step(5)
    ids += [srv.submit(*reqs[i]) for i in range(3, len(reqs))]  # join while others run
    for _ in range(200):
        running, queued = srv.step(3)
        if running == 0 and queued == 0:
            break
    for rid, (ini, st) in zip(ids, reqs):
        status, t, it, ev = srv.status(rid)
        assert status == DONE, (rid, status)
        assert ev[0] <= ev[1] <= ev[2]
        run = make_gpu(ini, st, precision)
        run.run_iterations(10 ** 6)
        assert it == run.time.iteration and t == run.time.time
        for f in ("mass", "momx", "ener", "magy"):
            e = rel_linf(srv.download(rid, f), run.download(f))
            assert e == 0.0, (rid, f, e)
        run.close()
        srv.release(rid)
    assert srv.stats()["used_bytes"] == 0
    srv.close()


def test_admission_by_budget_and_slots():
    reqs = limited([problems.kelvin_helmholtz(64, seed=s) for s in range(5)], 12)
    from imogen_amd import _lib
    from imogen_amd.config import from_ini
    from imogen_amd.imogen import make_imh_config
    import ctypes
    b = ctypes.c_size_t()
    _lib.check(_lib.lib().imh_memory_estimate(
        ctypes.byref(make_imh_config(from_ini(reqs[0][0]), _lib.FP32_BULK)), ctypes.byref(b)))
    srv = Server("fp32", hbm_budget=2 * b.value + 1, max_active=8)  # two fit at a time
    rep = run_stream(srv, reqs, np.zeros(len(reqs)), iterations_per_round=4)
    assert not rep["failed"]
    assert rep["cell_updates"] == 5 * 64 * 64 * 2 * 12
    srv.close()

    srv = Server("fp32", max_active=1)  # one slot: strictly one after the other
    ids = [srv.submit(*r) for r in reqs[:3]]
    for _ in range(100):
        if srv.step(6) == (0, 0):
            break
    ev = [srv.status(i)[3] for i in ids]
    for k in range(2):
        assert ev[k + 1][1] >= ev[k][2]  # admitted after the previous one finished
    srv.close()


def test_device_generated_requests():
    """Requests whose initial state is generated on the device at admission give the
    result of a run initialised with init_problem alone (exactly), mixed with host-state
    requests."""
    import imogen_amd
    sod_ini, _ = problems.sod(300, time_max=0.05)
    ot_ini, _ = problems.orszag_tang(40)
    sed_ini, _ = problems.sedov(16, time_max=0.02)
    dev = [(sod_ini, "sod", ()), (ot_ini, "orszag_tang", ()), (sed_ini, "sedov", (3.0, 1e-5, 1.0))]
    dev = [(dict(ini, iterMax=30), k, p) for ini, k, p in dev]
    host = limited([problems.mhd_riemann(400, seed=1)], 30)
    srv = Server("fp32", max_active=3)
    ids = [srv.submit_problem(ini, k, p) for ini, k, p in dev]
    hid = srv.submit(*host[0])
    for _ in range(300):
        running, queued = srv.step(4)
        if running == 0 and queued == 0:
            break
    for rid, (ini, k, p) in zip(ids, dev):
        status, t, it, ev = srv.status(rid)
        assert status == DONE, (rid, status)
        run = imogen_amd.initialize(ini, "fp32")
        run.init_problem(k, p)
        run.run_iterations(10 ** 6)
        assert it == run.time.iteration and t == run.time.time
        for f in ("mass", "momx", "ener", "magy"):
            assert rel_linf(srv.download(rid, f), run.download(f)) == 0.0, (k, f)
        run.close()
        srv.release(rid)
    assert srv.status(hid)[0] == DONE
    srv.release(hid)
    srv.close()


def test_stated_sizes_alone_equals_mix():
    """The serving stream's size range: 2-D grids of 64^2, 128^2 and 1024^2 and Sedov
    blasts of 64^3 and 192^3, admitted while others run (four slots), each equal to the
    same request run alone (FP32 bulk, exactly)."""
    reqs = limited([problems.sedov(64, time_max=1.0),
                    problems.kelvin_helmholtz(128, seed=3),
                    problems.orszag_tang(128),
                    problems.kelvin_helmholtz(1024, seed=8),
                    problems.sedov(192, time_max=1.0),
                    problems.kelvin_helmholtz(64, seed=1)], 12)
    srv = Server("fp32", max_active=4)
    ids = [srv.submit(*reqs[i]) for i in range(2)]
    srv.step(3)
    ids += [srv.submit(*r) for r in reqs[2:]]
    for _ in range(200):
        running, queued = srv.step(2)
        if running == 0 and queued == 0:
            break
    for rid, (ini, st) in zip(ids, reqs):
        status, t, it, ev = srv.status(rid)
        assert status == DONE, (ini["grid"], status)
        run = make_gpu(ini, st, "fp32")
        run.run_iterations(10 ** 6)
        assert it == run.time.iteration and t == run.time.time, ini["grid"]
        for f in ("mass", "momx", "momy", "ener", "magx"):
            e = rel_linf(srv.download(rid, f), run.download(f))
            assert e == 0.0, (ini["grid"], f, e)
        run.close()
        srv.release(rid)
    assert srv.stats()["used_bytes"] == 0
    srv.close()


def test_recycled_contexts_give_the_results_of_fresh_ones():
    """Released requests leave their contexts in the server's pool and later admissions of
    the same cell count reuse that device memory (imh_ctx_recycle).  A recycled context
    starts from a fresh state: with one slot, each request (differing in boundaries, so in
    the layout of the edge stores, in limiters, magnet on and off) runs on the memory of
    the one before it and equals the same request run alone exactly, in both precisions."""
    base = [
        problems.sod(256, time_max=0.03),
        problems.sod(256, limiter="minmond", bc="const", time_max=0.03),
        problems.mhd_riemann(256, seed=11),
        problems.sod(256, bc="trans", time_max=0.03),
        problems.mhd_riemann(256, seed=12),
    ]
    reqs = limited(base, 30)
    for precision in ("fp32", "fp64"):
        srv = Server(precision, max_active=1)
        got = []
        for ini, st in reqs:
            rid = srv.submit(ini, st)
            for _ in range(200):
                running, queued = srv.step(4)
                if running == 0 and queued == 0:
                    break
            assert srv.status(rid)[0] == DONE
            got.append({f: srv.download(rid, f) for f in ("mass", "momx", "ener", "magy")})
            srv.release(rid)
        srv.close()
        for (ini, st), g in zip(reqs, got):
            run = make_gpu(ini, st, precision)
            run.run_iterations(10 ** 6)
            for f, v in g.items():
                e = rel_linf(v, run.download(f))
                assert e == 0.0, (precision, ini["grid"], f, e)
            run.close()
