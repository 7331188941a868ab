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
  error            L1 / L-inf of every FP32-bulk field against the FP64 path after the
                   same number of iterations (per simulation, max and median), and
                   against the CPU oracle on the cpu_baseline sample
  roofline         the dominant kernel of the batched step, timed with HIP events
                   around each launch on the batch stream (an eager pass of the same
                   launch plan), algorithmic bytes = compulsory field reads + writes
  plain_roofline   the same for the plain FP64 path on a 256^3 hydro grid
  source_roofline  the source step (gravity + radiation) on that grid, FP64
  cpu_baseline     the CPU oracle (oracle/, C FP64 restatement), single thread, on a
                   bounded sample of the ensemble
  configs          Sod 1024, Orszag-Tang 512^2 and Sedov 256^3 one at a time on the plain
                   path: FP64 and FP32-bulk Mcell-updates/s, every field's FP32 error
                   against FP64 after the same iterations, the FP64 conservation sums
                   before / after, and the CPU oracle on one core and on all cores

roofline.traffic is the HBM traffic per launch of the dominant kernel measured by two
rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over this same workload, read from
profiles/fluid_sweep_traffic.json (null when that file has no entry for the workload).
`python bench.py --pmc-traffic FETCH.csv WRITE.csv BENCH.log` computes those entries from
the counter_collection files of the two passes and the bench output of either pass.
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
    ap.add_argument("--copy-probe", type=int, default=1024, metavar="MIB",
                    help="device copy of this many MiB (measured copy bandwidth beside the "
                         "8 TB/s peak, and the reference kernel for scaling HBM counters; "
                         "0 skips it)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the Sod / Orszag-Tang / Sedov single-simulation leg")
    ap.add_argument("--save-tiles", metavar="PATH",
                    help="write the tile table (with the classes tuned in this run) here")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default tiles instead of the tuned table")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-rates", default="10,40,160", help="Poisson arrival rates (req/s)")
    ap.add_argument("--profile-steps", type=int, default=6)
    ap.add_argument("--only-plain", action="store_true",
                    help="run only the plain FP64 256^3 roofline leg (a short program for "
                         "counter passes) and print its JSON")
    ap.add_argument("--pmc-traffic", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "BENCH_LOG"),
                    help="offline: HBM bytes per launch from two rocprofv3 --pmc passes")
    ap.add_argument("--trace-summary", nargs=2, metavar=("KERNEL_TRACE_CSV", "BENCH_LOG"),
                    help="offline: the profiled passes of a bench run made under "
                         "rocprofv3 --kernel-trace, as markdown tables")
    return ap.parse_args()


def _bench_json(path):
    out = None
    for line in open(path):
        if line.startswith("{"):
            out = json.loads(line)
    if out is None:
        raise SystemExit("no bench JSON line in %s" % path)
    return out


def trace_summary(trace_csv, bench_log):
    """Markdown tables of the profiled passes in a rocprofv3 kernel trace of a bench run:
    the batched step (the last `roofline.launches` float fluid launches and every launch
    between the iteration start before the first of them and the next double-precision
    launch) and the plain 256^3 grid (the last `plain_roofline.launches` double fluid
    launches and what runs between them), with the average fluid launch duration beside
    the one bench.py measured with HIP events.  The traced run must have no FP32 work after
    the profiled pass: run it with --no-serve --no-configs."""
    import collections
    import csv
    out = _bench_json(bench_log)
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))

    def short(n):
        n = n.replace("(anonymous namespace)::", "")
        return n.split("(")[0].replace("void ", "").replace("imh::", "")

    def table(sel, title, ref):
        agg = collections.defaultdict(lambda: [0, 0])
        for r in sel:
            a = agg[short(r["Kernel_Name"])]
            a[0] += 1
            a[1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        tot = sum(v[1] for v in agg.values())
        print("### %s\n" % title)
        print("| kernel | launches | total us | avg us | share |\n|---|---|---|---|---|")
        for k, (n, ns) in sorted(agg.items(), key=lambda t: -t[1][1]):
            print("| %s | %d | %.1f | %.1f | %.1f%% |" % (k, n, ns / 1e3, ns / n / 1e3,
                                                         100.0 * ns / tot))
        fl = [(n, ns) for k, (n, ns) in agg.items() if k.startswith(("fluid_sweep", "fluid_wave"))]
        n, ns = sum(x[0] for x in fl), sum(x[1] for x in fl)
        print("\nfluid sweep: %d launches, average %.2f us per launch in the trace; "
              "bench.py (HIP events): %.2f us over %d launches\n" % (
                  n, ns / n / 1e3, ref["avg_launch_us"], ref["launches"]))

    def is_fluid(r, prec):
        return (("fluid_sweep" in r["Kernel_Name"] or "fluid_wave" in r["Kernel_Name"])
                and ("<%s" % prec) in r["Kernel_Name"])

    if "roofline" not in out:  # a --only-plain run
        n = out["plain_roofline"]["launches"]
        fl = [i for i, r in enumerate(rows) if is_fluid(r, "double")][-n:]
        table(rows[fl[0]:fl[-1] + 1], "plain path, profiled pass: " +
              out["plain_roofline"]["workload"], out["plain_roofline"])
        return 0
    n = out["roofline"]["launches"]
    fl = [i for i, r in enumerate(rows) if is_fluid(r, "float")][-n:]
    lo, hi = fl[0], fl[-1] + 1
    while lo > 0 and "begin_iteration" not in rows[lo]["Kernel_Name"]:
        lo -= 1
    while hi < len(rows) and "<double" not in rows[hi]["Kernel_Name"] \
            and "copy_probe" not in rows[hi]["Kernel_Name"] and not is_fluid(rows[hi], "float"):
        hi += 1
    table(rows[lo:hi], "batched step, profiled pass: " + out["config"]["workload"],
          out["roofline"])
    if "plain_roofline" in out:
        n = out["plain_roofline"]["launches"]
        fl = [i for i, r in enumerate(rows) if is_fluid(r, "double")][-n:]
        table(rows[fl[0]:fl[-1] + 1], "plain path, profiled pass: " +
              out["plain_roofline"]["workload"], out["plain_roofline"])
    return 0


TRAFFIC_FILE = os.path.join(HERE, "profiles", "fluid_sweep_traffic.json")


def measured_traffic(workload):
    """HBM bytes per launch of the dominant kernel for `workload`, if measured."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        e = json.load(f).get(workload)
    return None if e is None else int(e["hbm_bytes_per_launch"])


def pmc_traffic(fetch_csv, write_csv, bench_log):
    """HBM bytes per launch of the fluid sweep in the profiled passes of a bench run made
    under `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` (run with --copy-probe MIB).

    FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are scaled by what they read for the copy
    probe, which reads and writes a known number of bytes.  The profiled pass of the
    batched step is the last `roofline.launches` float fluid launches; that of the plain
    256^3 grid the last `plain_roofline.launches` double ones (so the counted run has no
    FP32 work after the profiled pass: run it with --no-serve --no-configs)."""
    import csv
    out = _bench_json(bench_log)
    probe = out["copy_probe"]["bytes_read"]

    def load(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        return rows

    res = {}
    for path, counter in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        rows = load(path, counter)
        cp = [r for r in rows if "copy_probe" in r["Kernel_Name"]]
        scale = probe / (float(cp[-1]["Counter_Value"]) * 1024.0)
        for key, prec, n in (("roofline", "float", out.get("roofline", {}).get("launches", 0)),
                             ("plain_roofline", "double",
                              out.get("plain_roofline", {}).get("launches", 0))):
            if not n:
                continue
            fl = [r for r in rows if ("fluid_sweep" in r["Kernel_Name"] or
                                      "fluid_wave" in r["Kernel_Name"])
                  and ("<%s" % prec) in r["Kernel_Name"]][-n:]
            kib = sum(float(r["Counter_Value"]) for r in fl)
            e = res.setdefault(out[key]["workload"] if key == "plain_roofline"
                               else out["config"]["workload"], {"launches": n})
            e[counter.lower() + "_scale"] = round(scale, 4)
            e[counter.lower() + "_bytes_per_launch"] = int(kib * 1024.0 * scale / n)
            e["alg_bytes_per_launch"] = out[key]["alg_bytes_per_launch"]
    for e in res.values():
        e["hbm_bytes_per_launch"] = e["fetch_size_bytes_per_launch"] + e["write_size_bytes_per_launch"]
        e["traffic_over_alg"] = round(e["hbm_bytes_per_launch"] / e["alg_bytes_per_launch"], 3)
    print(json.dumps(res, indent=1, sort_keys=True))
    return 0


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


FIELDS = ["mass", "momx", "momy", "momz", "ener", "magx", "magy", "magz"]


def field_errors(pairs):
    """Normwise relative L1 and L-inf error of every field over (result, reference) pairs
    of per-simulation {field: array}: the largest and the median over the simulations in
    which the reference field is not identically zero; where it is, the largest |result|
    is reported as `abs_where_ref_zero`."""
    out = {}
    for f in FIELDS:
        l1, li, az = [], [], 0.0
        for a, b in pairs:
            x, y = a[f], b[f]
            if np.any(y != 0):
                l1.append(rel_l1(x, y))
                li.append(rel_linf(x, y))
            else:
                az = max(az, float(np.max(np.abs(x))))
        e = {"sims": len(l1)}
        if l1:
            e.update(l1_max=float("%.3e" % max(l1)), l1_median=float("%.3e" % np.median(l1)),
                     linf_max=float("%.3e" % max(li)))
        if len(l1) < len(pairs):
            e["abs_where_ref_zero"] = float("%.3e" % az)
        out[f] = e
    return out


def roofline(prof, label, kind=None):
    """Dominant kernel kind of a profile {kind: (ms, bytes, launches)}, or `kind`."""
    if kind is None:
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
        # tiles tuned in this run are kept where asked (an IMH_TILE_TABLE file is updated)
        if os.environ.get("IMH_TILE_TABLE"):
            autotune.save_table(table, autotune.TABLE)
        elif args.save_tiles:
            autotune.save_table(table, args.save_tiles)
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
        run.set_fluid_kernel(PLAIN_SWEEP)
        run.run_iterations(args.warmup + 1)
        run.synchronize()
        t0 = time.perf_counter()
        run.run_iterations(args.steps)
        run.synchronize()
        total_t += time.perf_counter() - t0
        total_cu += 2.0 * ncells(ini) * args.steps
        rho.append({f: run.download(f) for f in FIELDS})
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
        done.append((i, {f: orc.get(f) for f in FIELDS}))
        orc.close()
        if t_used >= args.cpu_seconds:
            break
    return cu / t_used / 1e6, done, t_used, n_iter


SERVE_2D = [64, 128, 256, 512, 1024]
SERVE_3D = [64, 96, 128, 192]


def serve_requests(n, seed):
    """Solve requests of the serving stream: Sod, Orszag-Tang and Kelvin-Helmholtz on
    64^2..1024^2 and Sedov on 64^3..192^3, each run to a fixed end time, seeded; the
    initial states are generated on the device at admission (imh_serve_submit_problem),
    so a request carries only its configuration."""
    from imogen_amd import problems
    rng = np.random.default_rng(seed)
    reqs, kinds = [], []
    for i in range(n):
        kind = rng.choice(["sod", "ot", "kh", "sedov"], p=[0.3, 0.3, 0.3, 0.1])
        big = 10 ** 8
        if kind == "sedov":
            m = int(rng.choice(SERVE_3D, p=[0.4, 0.3, 0.2, 0.1]))
            ini, _ = problems.sedov(1, time_max=5e-3, iter_max=big)
            r = {"kind": "sedov", "params": (3.0, 1e-5, 1.0)}
            dims = (m, m, m)
        else:
            m = int(rng.choice(SERVE_2D, p=[0.3, 0.3, 0.2, 0.15, 0.05]))
            dims = (m, m, 1)
            if kind == "sod":
                ini, _ = problems.sod(1, ny=1, bc="const", iter_max=big, time_max=0.2)
                r = {"kind": "sod"}
            elif kind == "ot":
                ini, _ = problems.orszag_tang(1, iter_max=big)
                ini["timeMax"] = 0.5
                r = {"kind": "orszag_tang"}
            else:
                ini, _ = problems.kelvin_helmholtz(1, iter_max=big)
                ini["timeMax"] = 0.2
                r = {"kind": "kelvin_helmholtz",
                     "params": (0.5, 8.0, 0.01, int(rng.integers(1 << 30)))}
        ini = problems.resize(ini, dims)
        r["ini"] = ini
        reqs.append(r)
        kinds.append("%s%d" % (kind, m))
    return reqs, kinds


def serve_plain_fp64(reqs):
    """The requests of a serving stream through the plain FP64 path, one after another
    (imh_run to each request's end time, initial state generated on the device as at
    admission): the equal-work reference for the serving rate.  Cell updates are counted
    as the server counts them (2 x cells per completed iteration)."""
    import imogen_amd
    cu, t = 0.0, 0.0
    for r in reqs:
        run = imogen_amd.initialize(r["ini"], "fp64")
        run.set_fluid_kernel(PLAIN_SWEEP)
        run.synchronize()
        t0 = time.perf_counter()
        run.init_problem(r["kind"], r.get("params", ()))
        while run.run_iterations(1000) == 1000:
            pass
        run.synchronize()
        t += time.perf_counter() - t0
        cu += 2.0 * ncells(r["ini"]) * (run.time.iteration - 1)
        run.close()
    return cu / t / 1e6, t


def serving_leg(args):
    """The serving stream at each arrival rate: latency p50 / p99 and Mcell-updates/s;
    the last rate's requests also through the plain FP64 path one by one (equal work)."""
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
    last = reqs
    plain, secs = serve_plain_fp64(last)
    out[-1]["plain_fp64_sequential"] = {
        "mcell_updates_per_s": round(plain, 1), "seconds": round(secs, 3),
        "ratio_serving_over_plain": round(out[-1]["mcell_updates_per_s"] / plain, 3)}
    log("serve rate %.0f/s requests through plain FP64 one by one: %.0f Mcell-updates/s "
        "(%.2f s); serving / plain %.2f" % (out[-1]["rate_per_s"], plain, secs,
                                           out[-1]["mcell_updates_per_s"] / plain))
    return out


# (name, generator, GPU run: iterations or ("t_end", t), CPU iterations from the IC)
CONFIGS = [("sod 1024 (vanLeer, trans)", lambda p: p.sod(1024), 200, 200),
           ("orszag-tang 512^2 MHD (superbee, periodic)", lambda p: p.orszag_tang(512), 1000, 2),
           ("sedov 256^3 hydro (minmod, trans)", lambda p: p.sedov(256), ("t_end", 0.05), 1)]


def host_cpu():
    """The host CPU: model name (from /proc/cpuinfo) and the logical cores this process
    may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "nproc": usable, "cpu_count": os.cpu_count()}


def _advance(run, plan):
    """Advance a run by a fixed number of iterations, or to its end time (timeMax set in
    its settings); returns the iterations done."""
    if not isinstance(plan, tuple):
        run.run_iterations(plan)
        return plan
    done = 0
    while True:
        d = run.run_iterations(1000)
        done += d
        if d < 1000:
            return done


def configs_leg(args):
    """Each config on the plain path as stated: Sod 200 iterations, Orszag-Tang 1000,
    Sedov to t = 0.05.  FP64 and FP32-bulk rates over the whole run (after one warm
    iteration from the same initial state, which restarts from the IC), every field's
    FP32 error against FP64 at the end, FP64 conservation sums before / after (for Sedov
    of both precisions at t = 0.05), and the oracle on 1 core and on all cores (at most
    16; a bounded number of iterations from the initial state)."""
    import imogen_amd
    from imogen_amd import problems
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from helpers import make_oracle
    threads = max(1, min(16, host_cpu()["nproc"] or 1))
    out = []
    for name, gen, plan, n_cpu in CONFIGS:
        ini, st = gen(problems)
        if isinstance(plan, tuple):
            ini = dict(ini, iterMax=10 ** 8, timeMax=plan[1])
        else:
            ini = dict(ini, iterMax=10 ** 8, timeMax=1e9)
        cells = ncells(ini)
        e = {"config": name, "cells": cells}
        if isinstance(plan, tuple):
            e["t_end"] = plan[1]
        res = {}
        for prec in ("fp64", "fp32"):
            # warm-up (graph capture, first-assignment iteration) on a throw-away context
            run = imogen_amd.initialize(ini, prec)
            imogen_amd.create_arrays(run, st)
            if prec == "fp64":
                run.set_fluid_kernel(PLAIN_SWEEP)
            run.run_iterations(2)
            run.synchronize()
            run.close()
            run = imogen_amd.initialize(ini, prec)
            imogen_amd.create_arrays(run, st)
            if prec == "fp64":
                run.set_fluid_kernel(PLAIN_SWEEP)
            s0 = run.conservation_sums()
            run.synchronize()
            t0 = time.perf_counter()
            n_it = _advance(run, plan)
            run.synchronize()
            dt = time.perf_counter() - t0
            t_now, it_now = run.time.time, run.time.iteration
            e["gpu_%s_iterations" % prec] = int(it_now)
            e["gpu_%s_seconds" % prec] = round(dt, 3)
            e["gpu_%s_mcell_updates_per_s" % prec] = round(2.0 * cells * n_it / dt / 1e6, 1)
            if isinstance(plan, tuple):
                e["gpu_%s_time_reached" % prec] = float("%.12g" % t_now)
            res[prec] = {f: run.download(f) for f in FIELDS}
            s1 = run.conservation_sums()
            if prec == "fp64" or isinstance(plan, tuple):
                e["%s_sums" % prec] = {"before": [float("%.12e" % v) for v in s0[:5]],
                                       "after": [float("%.12e" % v) for v in s1[:5]],
                                       "order": "mass, mom x/y/z, energy (FP64, compensated)"}
            run.close()
        e["fp32_error_vs_fp64"] = field_errors([(res["fp32"], res["fp64"])])
        del res
        log("config %s: gpu fp64 %.1f (%d it), fp32 %.1f (%d it) Mcell-updates/s" % (
            name, e["gpu_fp64_mcell_updates_per_s"], e["gpu_fp64_iterations"],
            e["gpu_fp32_mcell_updates_per_s"], e["gpu_fp32_iterations"]))
        if not args.no_cpu:
            for label, th in (("cpu_1core", 1), ("cpu_all_cores", threads)):
                orc = make_oracle(ini, st, threads=th)  # no warm-up: plain C, nothing to JIT
                t0 = time.perf_counter()
                for _ in range(n_cpu):
                    orc.iteration()
                dt = time.perf_counter() - t0
                orc.close()
                e[label] = {"mcell_updates_per_s": round(2.0 * cells * n_cpu / dt / 1e6, 3),
                            "threads": th, "iterations": n_cpu}
            log("config %s: cpu %.2f (1 core), %.2f (%d threads) Mcell-updates/s" % (
                name, e["cpu_1core"]["mcell_updates_per_s"],
                e["cpu_all_cores"]["mcell_updates_per_s"], threads))
        out.append(e)
    return out


# The plain FP64 path's fluid sweep: the fast FP64 wave-spread sweep where its tiles fill
# the device, the general sweep elsewhere (include/imogen_hip_ext.h, IMH_SWEEP_AUTO_FAST).
PLAIN_SWEEP = "auto_fast"


def large_plain_roofline(args):
    """Plain FP64 path on Sedov 256^3: HIP-event time of every launch over two profiled
    iterations, with the plain path's sweep and, beside it, with the general sweep."""
    import imogen_amd
    from imogen_amd import problems
    ini, st = problems.sedov(256)
    res = {}
    # the plain path's sweep last: trace and counter summaries take the last launches
    for kind in ("general", PLAIN_SWEEP):
        run = imogen_amd.initialize(ini, "fp64")
        imogen_amd.create_arrays(run, st)
        run.set_fluid_kernel(kind)
        run.run_iterations(1)
        prof = run.profile_iterations(2)
        run.close()
        res[kind] = roofline(prof, "sedov 256^3 hydro FP64, plain path")
    del st
    out = res[PLAIN_SWEEP]
    out["sweep"] = "wave-spread sweep, fast FP64 arithmetic (%s)" % PLAIN_SWEEP
    out["general_sweep"] = {k: res["general"][k] for k in ("achieved", "frac", "avg_launch_us")}
    return out


def source_roofline(args):
    """The source step (gravity from a fixed potential + optically thin radiation) on the
    plain FP64 path, Sedov 256^3: HIP-event time of its launches (ghost + cell kernels)."""
    import numpy as np

    import imogen_amd
    from imogen_amd import problems
    ini, st = problems.sedov(256)
    run = imogen_amd.initialize(ini, "fp64")
    imogen_amd.create_arrays(run, st)
    del st
    x = (np.arange(256) + 0.5) / 256
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    run.set_gravity(((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2).reshape(-1))
    del X, Y, Z
    run.set_radiation("optically_thin", 0.5, 1e-3)
    run.run_iterations(1)
    prof = run.profile_iterations(2)
    run.close()
    out = roofline(prof, "sedov 256^3 FP64 with gravity and radiation, source step",
                   kind="source")
    out["bytes_rule"] = ("per cell: mass, momentum, energy, potential (8 B) and cell-centred "
                         "B read; momentum and energy written")
    return out


def copy_probe(mib):
    """A plain device copy of `mib` MiB (measured copy bandwidth; the reference kernel of
    known traffic for scaling the HBM counters)."""
    import ctypes
    import imogen_amd._lib as L
    nbytes = mib << 20
    ms = ctypes.c_double()
    L.check(L.lib().imh_copy_probe(nbytes, 10, ctypes.byref(ms)))
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": round(ms.value, 4),
            "GB/s": round(2.0 * nbytes / (ms.value * 1e-3) / 1e9, 1)}


def main():
    args = parse()
    if args.pmc_traffic:
        return pmc_traffic(*args.pmc_traffic)
    if args.trace_summary:
        return trace_summary(*args.trace_summary)
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs each simulation on one device and "
              "has no multi-GPU path (--gpus %d requested); run with --gpus 1" % args.gpus,
              file=sys.stderr)
        return 2
    import imogen_amd._lib as L
    L.lib()  # fails loudly if the HIP library is missing
    if args.only_plain:
        res = {"plain_roofline": large_plain_roofline(args)}
        if args.copy_probe:
            res["copy_probe"] = copy_probe(args.copy_probe)
        print(json.dumps(res))
        return 0

    sims = ensemble(args.seed, args.riemann, args.kh)
    cells = np.array([ncells(ini) for ini, _ in sims], dtype=np.int64)
    print("ensemble: %d simulations, %d cells" % (len(sims), cells.sum()), file=sys.stderr,
          flush=True)

    ens, rate, t_batched = batched_leg(sims, args, cells)
    tiles = "autotuned (imogen_amd/autotune.py)" if not args.no_autotune else "default"
    print("batched fp32: %.1f Mcell-updates/s (%.3f s)" % (rate, t_batched), file=sys.stderr,
          flush=True)
    res32 = [{f: ens.download(i, f) for f in FIELDS} for i in range(len(sims))]

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
        prate, res64, t_plain = plain_leg(sims, args)
        l1 = [rel_l1(a["mass"], b["mass"]) for a, b in zip(res32, res64)]
        li = [rel_linf(a["mass"], b["mass"]) for a, b in zip(res32, res64)]
        out["plain_fp64"] = {"value": round(prate, 2), "unit": "Mcell-updates/s",
                             "seconds": round(t_plain, 3),
                             "ratio_batched_over_plain": round(rate / prate, 2)}
        out["error"] = {"density_l1_vs_fp64_max": float("%.3e" % max(l1)),
                        "density_l1_vs_fp64_median": float("%.3e" % np.median(l1)),
                        "density_linf_vs_fp64_max": float("%.3e" % max(li)),
                        "iterations": args.warmup + 1 + args.steps,
                        "fields_vs_fp64": field_errors(list(zip(res32, res64)))}
        del res64
        print("plain fp64: %.1f Mcell-updates/s; max L1 %.2e" % (prate, max(l1)),
              file=sys.stderr, flush=True)

    log("profiling the batched step")
    prof = ens.profile_step(args.profile_steps)
    out["roofline"] = roofline(prof, "batched ensemble FP32, persistent launches")
    out["roofline"]["bytes_rule"] = BYTES_NOTE
    out["roofline"]["traffic"] = measured_traffic(out["config"]["workload"])
    ens.close()

    if not args.no_large:
        log("plain fp64 sedov 256^3")
        out["plain_roofline"] = large_plain_roofline(args)
        out["plain_roofline"]["traffic"] = measured_traffic(out["plain_roofline"]["workload"])
        log("source step, sedov 256^3")
        out["source_roofline"] = source_roofline(args)

    if not args.no_serve:
        log("serving stream")
        out["serving"] = {
            "workload": "Poisson arrivals of Sod / Orszag-Tang / KH on 64^2..1024^2 and "
                        "Sedov on 64^3..192^3, each to a fixed end time; FP32 bulk, "
                        "initial states generated on the device at admission, "
                        "HBM-budget admission, continuous batching",
            "latency": "completion - arrival, host wall clock",
            "runs": serving_leg(args)}

    if not args.no_configs:
        log("configs: sod / orszag-tang / sedov")
        out["configs"] = configs_leg(args)

    if args.copy_probe:
        out["copy_probe"] = copy_probe(args.copy_probe)

    if not args.no_cpu:
        log("cpu oracle sample")
        crate, done, t_cpu, n_iter = cpu_leg(sims, args)
        out["cpu_baseline"] = {
            "value": round(crate, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "host_cpu": host_cpu(),
            "sample": "%d of the ensemble's simulations (%d cells), %d iterations each, "
                      "%.1f s on one host core" % (len(done), sum(ncells(sims[i][0]) for i, _ in done),
                                                   n_iter, t_cpu)}
        l1o = [rel_l1(res32[i]["mass"], r["mass"]) for i, r in done]
        out.setdefault("error", {})["density_l1_vs_oracle_max"] = float("%.3e" % max(l1o))
        out["error"]["oracle_sample"] = len(done)
        out["error"]["fields_vs_oracle"] = field_errors([(res32[i], r) for i, r in done])
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
