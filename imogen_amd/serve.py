"""The serving loop: solve requests under continuous batching on one device.

Wraps the imh_serve_* entry points (include/imogen_hip.h).  A request is one simulation
(ini, state) run while TimeManager.running holds; requests are admitted while slots and
the HBM budget allow, advance together, and keep their results on the device until
released.

``run_stream`` drives the server with requests that arrive over wall-clock time (for
example Poisson arrivals) and reports per-request latency (completion - arrival) with its
p50/p99 and the aggregate cell-update rate.
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import FIELD_IDS, ImhConfig, as_field, check, dptr, lib
from .config import from_ini
from .imogen import make_imh_config

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3
PRECISIONS = {"fp64": _lib.FP64, "fp32": _lib.FP32_BULK}


class Server:
    def __init__(self, precision="fp32", hbm_budget=0, max_active=256, device=0):
        self.precision = PRECISIONS[precision]
        h = ctypes.c_void_p()
        check(lib().imh_serve_create(int(device), self.precision, int(hbm_budget),
                                     int(max_active), ctypes.byref(h)))
        self._s = h
        self._cfg = {}

    def close(self):
        if getattr(self, "_s", None):
            lib().imh_serve_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _uniform(rc):
        if any(line is not None for line in rc.dgrid_lines):
            raise ValueError("serving takes uniform grid spacing (a request's configuration "
                             "carries no spacing lines)")
        return rc

    def submit(self, ini, st):
        rc = self._uniform(from_ini(ini))
        c = make_imh_config(rc, self.precision)
        n = rc.ncell
        arrs = [as_field(a, n) for a in (st["mass"], *st["mom"], st["ener"], *st["mag"])]
        P = ctypes.POINTER(ctypes.c_double)
        rid = ctypes.c_longlong()
        check(lib().imh_serve_submit(self._s, ctypes.byref(c), dptr(arrs[0]),
                                     (P * 3)(*[dptr(a) for a in arrs[1:4]]), dptr(arrs[4]),
                                     (P * 3)(*[dptr(a) for a in arrs[5:8]]), n,
                                     ctypes.byref(rid)))
        self._cfg[rid.value] = rc
        return rid.value

    def submit_problem(self, ini, kind, params=()):
        """A request whose initial state is generated on the device at admission (kind:
        sod, orszag_tang, sedov, riemann, kelvin_helmholtz; params as imh_init_problem)."""
        from .imogen import ImogenManager
        rc = self._uniform(from_ini(ini))
        c = make_imh_config(rc, self.precision)
        arr = (ctypes.c_double * max(1, len(params)))(*[float(p) for p in params])
        rid = ctypes.c_longlong()
        check(lib().imh_serve_submit_problem(self._s, ctypes.byref(c), ImogenManager.IC_KINDS[kind], arr,
                                             len(params), ctypes.byref(rid)))
        self._cfg[rid.value] = rc
        return rid.value

    def step(self, iterations):
        """One scheduling round; returns (running, queued)."""
        a, q = ctypes.c_int(), ctypes.c_int()
        check(lib().imh_serve_step(self._s, int(iterations), ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def status(self, rid):
        """(status, sim time, iteration, (t_submit, t_admit, t_done)) in server seconds."""
        st, it = ctypes.c_int(), ctypes.c_int()
        t = ctypes.c_double()
        ev = (ctypes.c_double * 3)()
        check(lib().imh_serve_status(self._s, int(rid), ctypes.byref(st), ctypes.byref(t),
                                     ctypes.byref(it), ev))
        return st.value, t.value, it.value, tuple(ev)

    def download(self, rid, field):
        rc = self._cfg[rid]
        out = np.empty(rc.ncell, dtype=np.float64)
        check(lib().imh_serve_download(self._s, int(rid), FIELD_IDS[field], dptr(out), out.size))
        return out.reshape(rc.shape)

    def release(self, rid):
        check(lib().imh_serve_release(self._s, int(rid)))
        self._cfg.pop(rid, None)

    def stats(self):
        u, b = ctypes.c_size_t(), ctypes.c_size_t()
        cu = ctypes.c_longlong()
        check(lib().imh_serve_stats(self._s, ctypes.byref(u), ctypes.byref(b), ctypes.byref(cu)))
        return {"used_bytes": u.value, "budget_bytes": b.value, "cell_updates": cu.value}


def run_stream(server, requests, arrivals, iterations_per_round=4, keep=None, timeout=600.0):
    """Submit requests[i] at wall-clock offset arrivals[i] (seconds from the start), run
    scheduling rounds until every request has finished, and return a report.  A request
    is (ini, state) for a host initial state, or {"ini", "kind", "params"} for one
    generated on the device at admission.

    ``keep``: optional set of request indices whose final density is downloaded (into
    report["results"]) before release; all other results are released on completion."""
    order = np.argsort(arrivals, kind="stable")
    ids = [None] * len(requests)
    done_at, submitted_at = {}, {}
    results = {}
    t0 = time.perf_counter()
    nxt = 0
    pending = set()
    while nxt < len(order) or pending:
        now = time.perf_counter() - t0
        if now > timeout:
            raise TimeoutError("serving stream did not finish in %.0f s" % timeout)
        while nxt < len(order) and arrivals[order[nxt]] <= now:
            i = int(order[nxt])
            r = requests[i]
            ids[i] = (server.submit_problem(r["ini"], r["kind"], r.get("params", ()))
                      if isinstance(r, dict) else server.submit(*r))
            submitted_at[i] = time.perf_counter() - t0
            pending.add(i)
            nxt += 1
        if not pending:
            time.sleep(max(0.0, min(arrivals[order[nxt]] - now, 0.01)))
            continue
        server.step(iterations_per_round)
        t_round = time.perf_counter() - t0
        for i in list(pending):
            st = server.status(ids[i])[0]
            if st in (DONE, FAILED):
                done_at[i] = (t_round, st)
                if keep is not None and i in keep and st == DONE:
                    results[i] = server.download(ids[i], "mass")
                server.release(ids[i])
                pending.discard(i)
    wall = time.perf_counter() - t0
    lat = np.array([done_at[i][0] - arrivals[i] for i in range(len(requests))])
    failed = [i for i in range(len(requests)) if done_at[i][1] == FAILED]
    cu = server.stats()["cell_updates"]
    return {
        "requests": len(requests), "failed": failed, "wall_s": wall,
        "latency_s": lat, "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)), "cell_updates": cu,
        "mcell_updates_per_s": cu / wall / 1e6, "results": results,
    }
