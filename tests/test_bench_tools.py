"""CPU tests of the measurement tooling: bench.py's offline summaries of rocprofv3 output
(kernel trace, FETCH_SIZE / WRITE_SIZE passes) and the autotuner's tile rules."""
import csv
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

WORKLOAD = "toy workload"


def bench_line(launches=2, alg=1000):
    return {"config": {"workload": WORKLOAD},
            "roofline": {"launches": launches, "alg_bytes_per_launch": alg,
                         "avg_launch_us": 2.0, "workload": "batched"},
            "copy_probe": {"bytes_read": 1 << 20, "bytes_written": 1 << 20}}


def write_counters(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, (name, v) in enumerate(rows):
            w.writerow([i + 1, name, counter, v])


def run_bench(*args):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + list(args),
                       capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_pmc_traffic_scales_by_copy_probe(tmp_path):
    log = tmp_path / "bench.log"
    log.write_text("noise\n" + json.dumps(bench_line()) + "\n")
    fluid = "void imh::fluid_sweep_batch_kernel<float, 512, 4, false, false>(args)"
    # copy probe: FETCH reads half the bytes (scale 2), WRITE all (scale 1)
    write_counters(tmp_path / "f.csv", "FETCH_SIZE",
                   [(fluid, 100.0), ("imh::copy_probe_kernel(x)", 512.0), (fluid, 3.0), (fluid, 5.0)])
    write_counters(tmp_path / "w.csv", "WRITE_SIZE",
                   [(fluid, 100.0), ("imh::copy_probe_kernel(x)", 1024.0), (fluid, 1.0), (fluid, 2.0)])
    out = json.loads(run_bench("--pmc-traffic", str(tmp_path / "f.csv"), str(tmp_path / "w.csv"),
                               str(log)))
    e = out[WORKLOAD]
    assert e["fetch_size_scale"] == 2.0 and e["write_size_scale"] == 1.0
    # the last `launches` (2) fluid dispatches: fetch (3+5) KiB x2 / 2, write (1+2) KiB / 2
    assert e["fetch_size_bytes_per_launch"] == int(8 * 1024 * 2.0 / 2)
    assert e["write_size_bytes_per_launch"] == int(3 * 1024 / 2)
    assert e["hbm_bytes_per_launch"] == e["fetch_size_bytes_per_launch"] + e["write_size_bytes_per_launch"]
    assert e["traffic_over_alg"] == pytest.approx(e["hbm_bytes_per_launch"] / 1000, abs=1e-3)


def test_trace_summary_picks_the_profiled_pass(tmp_path):
    log = tmp_path / "bench.log"
    log.write_text(json.dumps(bench_line(launches=2)) + "\n")
    fl = "void imh::fluid_sweep_batch_kernel<float, 512, 4, false, false>(args)"
    rows = [("imh::begin_iteration_kernel(x)", 0, 1000),     # an earlier pass
            (fl, 1000, 9000),
            ("imh::begin_iteration_kernel(x)", 10000, 11000),  # the profiled pass
            (fl, 11000, 13000),
            ("void imh::magnet_advect_batch_kernel<float, 16>(x)", 13000, 14000),
            (fl, 14000, 16000),
            ("imh::copy_probe_kernel(x)", 20000, 30000)]
    with open(tmp_path / "trace.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for r in rows:
            w.writerow(r)
    out = run_bench("--trace-summary", str(tmp_path / "trace.csv"), str(log))
    assert "fluid sweep: 2 launches, average 2.00 us per launch in the trace" in out
    assert "magnet_advect_batch_kernel<float, 16> | 1 |" in out
    assert "copy_probe" not in out


def test_measured_traffic_lookup():
    import bench
    e = json.load(open(bench.TRAFFIC_FILE))
    for workload, entry in e.items():
        assert bench.measured_traffic(workload) == entry["hbm_bytes_per_launch"]
    assert bench.measured_traffic("no such workload") is None


def test_autotune_candidates_hold_the_pencil():
    from imogen_amd import autotune
    for n in (64, 200, 512, 1000, 4096):
        p2 = autotune._pow2(n)
        for bs, k, px in autotune.candidates(p2, 4, True, 1 << 20):
            if bs == 0:
                continue  # the column march
            assert bs % px == 0 and (bs // px) * k >= p2, (n, bs, k, px)
            assert px == 1 or (bs // px) * k < 2 * p2


def test_shared_single_pencil_tiles():
    """1-D simulations of one family all take the tuned tile of the longest class."""
    from imogen_amd import autotune, problems
    sims = [problems.mhd_riemann(n, seed=n) for n in (300, 900, 3000)]
    sims.append(problems.kelvin_helmholtz(64, seed=1))
    table = {}
    for ini, _ in sims:
        for key, X, n, col, nface in autotune.sweep_classes(ini, "fp32"):
            table[key] = [256, 4, 1, 1.0] if nface == 1 else [512, 4, 8, 1.0]
    longest = [k for k in table if "_n4096_" in k][0]
    table[longest] = [1024, 4, 1, 2.0]
    shared = autotune._shared_single_pencil_tiles(sims, table, "fp32")
    single = {k for ini, _ in sims for k, X, n, c, nf in autotune.sweep_classes(ini, "fp32")
              if nf == 1}
    assert set(shared) == single
    assert all(v[:3] == [1024, 4, 1] for v in shared.values())
