"""Build libimogen_hip.so (gfx950) in-tree with hipcc.

Every HIP source is compiled for gfx950 only, with packed-FP32 instructions disabled and
floating-point contraction off (the FP64 path reproduces the reference's rounding), then
linked into one shared library next to this file.
"""
import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libimogen_hip.so")
ARCH = os.environ.get("IMH_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

SOURCES = ["magnet_dispatch.hip", "cfl.hip", "fluid_dispatch.cpp", "fluid_column.hip",
           "initial.hip", "source.hip", "imh_api.cpp", "batch.cpp", "serve.cpp", "noalloc.hip",
           "lane1d.hip", "fluid_long.hip"]
FP32_ONLY = ("lane1d.hip",)
VARIANT_SRC = "fluid_sweep_variant.hip"
WAVE_VARIANT_SRC = "fluid_wave_variant.hip"
MAG_VARIANT_SRC = "magnet_sweep_variant.hip"


def fluid_variants():
    """(element size, block size, KMAX) of every fluid sweep variant, read from the
    IMH_FLUID_VARIANTS list in imh_kernels.h so the build and the dispatcher agree."""
    text = open(os.path.join(CSRC, "imh_kernels.h")).read()
    m = re.search(r"#define IMH_FLUID_VARIANTS\(X\)((?:[^\n]*\\\n)*[^\n]*)", text)
    return [tuple(int(v) for v in t) for t in re.findall(r"X\((\d+), (\d+), (\d+)\)", m.group(1))]


def magnet_variants():
    """(element size, KMAX) of every magnet variant (IMH_MAG_VARIANTS in imh_kernels.h)."""
    text = open(os.path.join(CSRC, "imh_kernels.h")).read()
    m = re.search(r"#define IMH_MAG_VARIANTS\(X\)((?:[^\n]*\\\n)*[^\n]*)", text)
    return [tuple(int(v) for v in t) for t in re.findall(r"X\((\d+), (\d+)\)", m.group(1))]


def wave_variants():
    """(element size, block size, fast FP64) of every wave-spread sweep variant
    (IMH_WAVE_VARIANTS in imh_kernels.h)."""
    text = open(os.path.join(CSRC, "imh_kernels.h")).read()
    m = re.search(r"#define IMH_WAVE_VARIANTS\(X\)((?:[^\n]*\\\n)*[^\n]*)", text)
    return [tuple(int(v) for v in t) for t in re.findall(r"X\((\d+), (\d+), (\d+)\)", m.group(1))]


def units():
    """(source, object name, extra flags) of every translation unit."""
    # FP64 units keep -ffp-contract=off (the FP64 kernels reproduce the oracle's rounding);
    # FP32-bulk units fuse the multiply-adds written within one expression (the FP32-only
    # 1-D lane too).  Not "fast": that also fuses across statements wherever the
    # surrounding code allows, so one body could round differently in two kernels that
    # inline it (the batched and single-simulation results must agree bit for bit).
    fp32 = ["-ffp-contract=on"]
    out = [(s, os.path.splitext(s)[0], fp32 if s in FP32_ONLY else []) for s in SOURCES]
    for e, b, k in fluid_variants():
        out.append((VARIANT_SRC, "fluid_sweep_v%d_%d_%d" % (e, b, k),
                    ["-DIMH_FLUID_ESZ=%d" % e, "-DIMH_FLUID_BS=%d" % b, "-DIMH_FLUID_KMAX=%d" % k]
                    + (fp32 if e == 4 else [])))
    for e, b, f in wave_variants():
        # FP64 units keep contraction off: the exact ones for bit parity with the
        # restatement, the fast ones because contracted builds of this kernel gave wrong
        # MHD sweeps along y (their arithmetic gain is the reciprocals)
        out.append((WAVE_VARIANT_SRC, "fluid_wave_v%d_%d_%d" % (e, b, f),
                    ["-DIMH_WAVE_ESZ=%d" % e, "-DIMH_WAVE_BS=%d" % b, "-DIMH_WAVE_FAST=%d" % f]
                    + (fp32 if e == 4 else [])))
    for e, k in magnet_variants():
        out.append((MAG_VARIANT_SRC, "magnet_sweep_v%d_%d" % (e, k),
                    ["-DIMH_MAG_ESZ=%d" % e, "-DIMH_MAG_KMAX=%d" % k] + (fp32 if e == 4 else [])))
    return out


def _deps_mtime():
    names = [f for f in os.listdir(CSRC) if f.endswith((".h", ".hip", ".cpp"))]
    inc = os.path.join(HERE, "..", "include")
    paths = [os.path.join(CSRC, f) for f in names] + [os.path.join(inc, f) for f in os.listdir(inc)]
    return max(os.path.getmtime(p) for p in paths)


def _compile(unit, verbose=False, incremental=False, defines=(), build_dir=None):
    src, name, flags = unit
    path = os.path.join(CSRC, src)
    obj = os.path.join(build_dir or BUILD, name + ".o")
    flags = list(flags) + ["-D" + d for d in defines]
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) >= _deps_mtime():
        return obj, ""
    cmd = [HIPCC, "-x", "hip", "--offload-arch=" + ARCH] + COMMON + flags + ["-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (name, r.stdout, r.stderr))
    return obj, r.stderr


def build(verbose=False, jobs=None, incremental=False, out=None, defines=()):
    """Compile every translation unit and link libimogen_hip.so.  Recompiles everything
    unless `incremental` (then only objects older than some source or header).

    `out` links the library elsewhere (loaded when IMH_LIB_PATH names it) and `defines`
    (NAME=VALUE) set the compile-time tuning knobs of the kernels, e.g.
    IMH_FLUID_WAVES_PER_EU or IMH_FLUID_GROUP (fluid_sweep_impl.h)."""
    lib_path = os.path.abspath(out) if out else LIB
    build_dir = BUILD if not out else os.path.join(
        BUILD, "alt_" + os.path.splitext(os.path.basename(lib_path))[0])
    os.makedirs(build_dir, exist_ok=True)
    us = units()
    # the long fluid variants first, so they overlap with the short units
    us.sort(key=lambda u: 0 if u[0] in (VARIANT_SRC, MAG_VARIANT_SRC, WAVE_VARIANT_SRC) else 1)
    jobs = jobs or min(len(us), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, verbose, incremental, defines, build_dir), us))
    objs = []
    for obj, warn in results:
        objs.append(obj)
        if warn.strip() and verbose:
            print(warn, file=sys.stderr)
    tmp = lib_path + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, lib_path)
    return lib_path


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="build libimogen_hip.so for " + ARCH)
    ap.add_argument("-v", action="store_true", help="print the compiler commands")
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--out", help="library path (default: in-tree libimogen_hip.so)")
    ap.add_argument("-D", dest="defines", action="append", default=[], metavar="NAME=VALUE")
    a = ap.parse_args()
    print(build(verbose=a.v, incremental=a.incremental, out=a.out, defines=a.defines))
