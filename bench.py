#!/usr/bin/env python3
"""Throughput of the relaxing-TVD fluid/magnet update on one MI355X.

Workload (the batched ensemble): 512 one-dimensional MHD Riemann problems with N drawn
uniformly from [256, 4096] plus 64 two-dimensional Kelvin-Helmholtz boxes with side
drawn from {128, 192, 256, 384, 512}, seeded.  A step is one batched iteration (two
timesteps of every simulation, as one pass of the imogen.m loop body), replayed from
captured hipGraphs.  One cell-update = one cell advanced by one timestep, so a step is
2 * sum(cells) cell-updates.

Printed as one JSON line:
  value            batched FP32-bulk Mcell-updates/s over the K timed steps
  plain_fp64       the same ensemble through the plain FP64 path, one simulation after
                   the other (imh_run), and the ratio batched / plain
  error            L1 / L-inf of the FP32-bulk density against the FP64 path after the
                   same number of iterations (per simulation, max and median), and
                   against the CPU oracle on the cpu_baseline sample
  roofline         the dominant kernel of the batched step, timed with HIP events
                   around each launch on the batch stream (an eager pass of the same
                   launch plan), algorithmic bytes = compulsory field reads + writes
  plain_roofline   the same for the plain FP64 path on a 256^3 hydro grid
  cpu_baseline     the CPU oracle (oracle/, C FP64 restatement), single thread, on a
                   bounded sample of the ensemble
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
KH_SIDES = [128, 192, 256, 384, 512]
BYTES_NOTE = ("fluid sweep: (5 reads + 5 writes + 3 cell-centred B reads if MHD) x cells x "
              "element size per launch")


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--riemann", type=int, default=512, help="1-D MHD Riemann problems")
    ap.add_argument("--kh", type=int, default=64, help="2-D Kelvin-Helmholtz boxes")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 256^3 plain roofline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device copy of this many MiB (the reference kernel "
                         "for scaling HBM counters)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default tiles instead of the tuned table")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-rates", default="10,40,160", help="Poisson arrival rates (req/s)")
    ap.add_argument("--profile-steps", type=int, default=6)
    return ap.parse_args()


def ensemble(seed, n_riemann, n_kh):
    """(ini, state) of every simulation of the batched ensemble."""
    from imogen_amd import problems
    rng = np.random.default_rng(seed)
    sims = []
    for i in range(n_riemann):
        n = int(rng.integers(256, 4097))
        sims.append(problems.mhd_riemann(n, seed=seed * 1000003 + i, iter_max=10 ** 8))
    for j in range(n_kh):
        side = int(rng.choice(KH_SIDES))
        sims.append(problems.kelvin_helmholtz(side, seed=seed * 1000003 + 500000 + j,
                                              iter_max=10 ** 8))
    return sims


def ncells(ini):
    g = ini["grid"]
    return int(g[0]) * int(g[1]) * int(g[2])


def rel_l1(a, ref):
    return float(np.sum(np.abs(a - ref)) / np.sum(np.abs(ref)))


def rel_linf(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


def roofline(prof, label):
    """Dominant kernel kind of a profile {kind: (ms, bytes, launches)}."""
    kind = max((k for k in prof if k != "bookkeeping"), key=lambda k: prof[k][0])
    ms, by, cnt = prof[kind]
    avg_s = ms / 1e3 / max(cnt, 1)
    achieved = by / max(cnt, 1) / avg_s / 1e9
    total_ms = sum(v[0] for v in prof.values())
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "kernel": kind, "workload": label, "launches": cnt,
        "avg_launch_us": round(avg_s * 1e6, 2),
        "alg_bytes_per_launch": int(by / max(cnt, 1)),
        "share_of_step_time": round(ms / total_ms, 3) if total_ms > 0 else None,
        "by_kind_ms": {k: round(v[0], 3) for k, v in prof.items()},
    }


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batched_leg(sims, args, cells):
    from imogen_amd import autotune
    from imogen_amd.ensemble import Ensemble
    tuned = 0
    if not args.no_autotune:
        table = autotune.load_table()
        table = autotune.tune(sims, "fp32", table=table, log=log)
        autotune.save_table(table, os.path.join(HERE, "gpurun_out", "tuned_tiles.json")
                            if os.path.isdir(os.path.join(HERE, "gpurun_out")) else autotune.TABLE)
    ens = Ensemble(sims, "fp32")
    if not args.no_autotune:
        tuned = autotune.apply(ens, sims, table, "fp32")
    log("batched: uploaded, %d sweeps with tuned tiles" % tuned)
    ens.step(args.warmup)          # first iteration of each simulation + graph capture
    ens.synchronize()
    log("batched: warm")
    t0 = time.perf_counter()
    ens.step(args.steps)
    ens.synchronize()
    dt = time.perf_counter() - t0
    rate = 2.0 * cells.sum() * args.steps / dt / 1e6
    return ens, rate, dt


def plain_leg(sims, args):
    """Every simulation through the plain FP64 path on its own (imh_run); returns the
    rate over the timed K iterations and the final densities."""
    import imogen_amd
    total_t, total_cu, rho = 0.0, 0.0, []
    for ini, st in sims:
        run = imogen_amd.initialize(ini, "fp64")
        imogen_amd.create_arrays(run, st)
        run.run_iterations(args.warmup + 1)
        run.synchronize()
        t0 = time.perf_counter()
        run.run_iterations(args.steps)
        run.synchronize()
        total_t += time.perf_counter() - t0
        total_cu += 2.0 * ncells(ini) * args.steps
        rho.append(run.download("mass"))
        run.close()
    return total_cu / total_t / 1e6, rho, total_t


def cpu_leg(sims, args):
    """The CPU oracle, single thread, on a bounded sample: the ensemble's Riemann problems
    and its KH boxes of side <= 192, interleaved 8:1 in ensemble order, until cpu_seconds
    of work is done, each advanced warmup + 1 + steps iterations like the GPU legs."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from helpers import make_oracle
    n_iter = args.warmup + 1 + args.steps
    rie = [i for i in range(len(sims)) if sims[i][0]["grid"][1] == 1]
    kh = [i for i in range(len(sims)) if 1 < sims[i][0]["grid"][1] <= 192]
    order = []
    for k in range(max(len(rie) // 8, len(kh))):
        order += rie[8 * k:8 * k + 8] + kh[k:k + 1]
    done, t_used, cu = [], 0.0, 0.0
    for i in order:
        ini, st = sims[i]
        orc = make_oracle(ini, st, threads=1)
        t0 = time.perf_counter()
        for _ in range(n_iter):
            orc.iteration()
        t_used += time.perf_counter() - t0
        cu += 2.0 * ncells(ini) * n_iter
        done.append((i, orc.get("mass")))
        orc.close()
        if t_used >= args.cpu_seconds:
            break
    return cu / t_used / 1e6, done, t_used, n_iter


SERVE_2D = [64, 128, 256, 512, 1024]
SERVE_3D = [64, 96, 128, 192]


def serve_requests(n, seed):
    """Solve requests of the serving stream: Sod, Orszag-Tang and Kelvin-Helmholtz on
    64^2..1024^2 and Sedov on 64^3..192^3, each run to a fixed end time, seeded ICs."""
    from imogen_amd import problems
    rng = np.random.default_rng(seed)
    reqs, kinds = [], []
    for i in range(n):
        kind = rng.choice(["sod", "ot", "kh", "sedov"], p=[0.3, 0.3, 0.3, 0.1])
        big = 10 ** 8
        if kind == "sedov":
            m = int(rng.choice(SERVE_3D, p=[0.4, 0.3, 0.2, 0.1]))
            ini, st = problems.sedov(m, time_max=5e-3, iter_max=big)
        else:
            m = int(rng.choice(SERVE_2D, p=[0.3, 0.3, 0.2, 0.15, 0.05]))
            if kind == "sod":
                ini, st = problems.sod(m, ny=m, bc="const", iter_max=big, time_max=0.2)
            elif kind == "ot":
                ini, st = problems.orszag_tang(m, iter_max=big)
                ini["timeMax"] = 0.5
            else:
                ini, st = problems.kelvin_helmholtz(m, seed=int(rng.integers(1 << 30)),
                                                    iter_max=big)
                ini["timeMax"] = 0.2
        reqs.append((ini, st))
        kinds.append("%s%d" % (kind, m))
    return reqs, kinds


def serving_leg(args):
    """The serving stream at each arrival rate: latency p50 / p99 and Mcell-updates/s."""
    from imogen_amd.serve import Server, run_stream
    out = []
    for k, rate in enumerate(float(r) for r in args.serve_rates.split(",")):
        reqs, kinds = serve_requests(args.serve_requests, args.seed * 7919 + k)
        rng = np.random.default_rng(args.seed * 104729 + k)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
        srv = Server("fp32", max_active=256)
        rep = run_stream(srv, reqs, arrivals, iterations_per_round=4)
        srv.close()
        out.append({"rate_per_s": rate, "requests": len(reqs), "failed": len(rep["failed"]),
                    "p50_s": round(rep["p50_s"], 4), "p99_s": round(rep["p99_s"], 4),
                    "wall_s": round(rep["wall_s"], 3),
                    "mcell_updates_per_s": round(rep["mcell_updates_per_s"], 1)})
        log("serve rate %.0f/s: p50 %.3f s p99 %.3f s, %.0f Mcell-updates/s" % (
            rate, rep["p50_s"], rep["p99_s"], rep["mcell_updates_per_s"]))
    return out


def large_plain_roofline(args):
    import imogen_amd
    from imogen_amd import problems
    ini, st = problems.sedov(256)
    run = imogen_amd.initialize(ini, "fp64")
    imogen_amd.create_arrays(run, st)
    del st
    run.run_iterations(1)
    prof = run.profile_iterations(2)
    run.close()
    return roofline(prof, "sedov 256^3 hydro FP64, plain path")


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d "
              "requested); run with --gpus 1" % args.gpus, file=sys.stderr)
        return 2
    import imogen_amd._lib as L
    L.lib()  # fails loudly if the HIP library is missing

    sims = ensemble(args.seed, args.riemann, args.kh)
    cells = np.array([ncells(ini) for ini, _ in sims], dtype=np.int64)
    print("ensemble: %d simulations, %d cells" % (len(sims), cells.sum()), file=sys.stderr,
          flush=True)

    ens, rate, t_batched = batched_leg(sims, args, cells)
    tiles = "autotuned (imogen_amd/autotune.py)" if not args.no_autotune else "default"
    print("batched fp32: %.1f Mcell-updates/s (%.3f s)" % (rate, t_batched), file=sys.stderr,
          flush=True)
    rho32 = [ens.download(i, "mass") for i in range(len(sims))]

    out = {
        "metric": "Mcell-updates/s batched FP32 vs plain FP64 @L1<=1e-4; p50/p99 solve latency",
        "value": round(rate, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_batched / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched ensemble: %d 1-D MHD Riemann N~U[256,4096] + %d 2-D "
                               "KH side in %s, FP32 bulk (FP64 dt, pressure floor, sums), "
                               "seed %d" % (args.riemann, args.kh, KH_SIDES, args.seed),
                   "simulations": len(sims), "cells": int(cells.sum()),
                   "cell_updates_per_step": int(2 * cells.sum()),
                   "parallelism": "single GPU, one persistent launch per kernel kind",
                   "tiles": tiles},
    }

    if not args.no_plain:
        prate, rho64, t_plain = plain_leg(sims, args)
        l1 = [rel_l1(a, b) for a, b in zip(rho32, rho64)]
        li = [rel_linf(a, b) for a, b in zip(rho32, rho64)]
        out["plain_fp64"] = {"value": round(prate, 2), "unit": "Mcell-updates/s",
                             "seconds": round(t_plain, 3),
                             "ratio_batched_over_plain": round(rate / prate, 2)}
        out["error"] = {"density_l1_vs_fp64_max": float("%.3e" % max(l1)),
                        "density_l1_vs_fp64_median": float("%.3e" % np.median(l1)),
                        "density_linf_vs_fp64_max": float("%.3e" % max(li)),
                        "iterations": args.warmup + 1 + args.steps}
        print("plain fp64: %.1f Mcell-updates/s; max L1 %.2e" % (prate, max(l1)),
              file=sys.stderr, flush=True)

    log("profiling the batched step")
    prof = ens.profile_step(args.profile_steps)
    out["roofline"] = roofline(prof, "batched ensemble FP32, persistent launches")
    out["roofline"]["bytes_rule"] = BYTES_NOTE
    ens.close()

    if not args.no_large:
        log("plain fp64 sedov 256^3")
        out["plain_roofline"] = large_plain_roofline(args)

    if not args.no_serve:
        log("serving stream")
        out["serving"] = {
            "workload": "Poisson arrivals of Sod / Orszag-Tang / KH on 64^2..1024^2 and "
                        "Sedov on 64^3..192^3, each to a fixed end time; FP32 bulk, "
                        "HBM-budget admission, continuous batching",
            "latency": "completion - arrival, host wall clock",
            "runs": serving_leg(args)}

    if args.copy_probe:
        import ctypes
        nbytes = args.copy_probe << 20
        ms = ctypes.c_double()
        L.check(L.lib().imh_copy_probe(nbytes, 10, ctypes.byref(ms)))
        out["copy_probe"] = {"bytes_read": nbytes, "bytes_written": nbytes,
                             "ms": round(ms.value, 4),
                             "GB/s": round(2.0 * nbytes / (ms.value * 1e-3) / 1e9, 1)}

    if not args.no_cpu:
        log("cpu oracle sample")
        crate, done, t_cpu, n_iter = cpu_leg(sims, args)
        out["cpu_baseline"] = {
            "value": round(crate, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d of the ensemble's simulations (%d cells), %d iterations each, "
                      "%.1f s on one host core" % (len(done), sum(ncells(sims[i][0]) for i, _ in done),
                                                   n_iter, t_cpu)}
        l1o = [rel_l1(rho32[i], r) for i, r in done]
        out.setdefault("error", {})["density_l1_vs_oracle_max"] = float("%.3e" % max(l1o))
        out["error"]["oracle_sample"] = len(done)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
