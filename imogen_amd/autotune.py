"""Tile autotuner for the fused fluid sweep.

The sweep kernel is compiled for a few (block size, registers per thread) variants and
takes the number of pencils per block at run time; sweeps along y / z with circular or
constant boundaries can also run as a column march (one thread per pencil).  For every
class of sweep in a workload -- precision, sweep dimension, pencil length (rounded up to a
power of two), pencil count (power-of-two bucket), MHD or not -- the tuner times each
candidate on a batch of copies of a representative simulation (HIP events around the
batched launches) and keeps the fastest.  Results are cached in a JSON table.
"""
import json
import math
import os

from . import build as _build
from .config import from_ini
from .ensemble import Ensemble

# IMH_TILE_TABLE selects another table (e.g. for a build with other kernel knobs)
TABLE = os.environ.get("IMH_TILE_TABLE") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json")


def _pow2(n):
    return 1 << max(0, int(math.ceil(math.log2(max(1, n)))))


def sweep_classes(ini, precision):
    """(class key, sweep dim X, N) of every sweep of a simulation."""
    rc = from_ini(ini)
    esz = 8 if precision == "fp64" else 4
    mhd = bool(rc.magnet_active)
    out = []
    for d in range(3):
        n = rc.grid[d]
        if n < 3:
            continue
        nface = rc.ncell // n
        simple = all(rc.bc[s][d] in ("circ", "const") for s in range(2))
        key = "e%d_d%d_n%d_f%d_m%d_s%d" % (esz, d, _pow2(n), _pow2(nface), int(mhd), int(simple))
        out.append((key, d + 1, n, simple and d > 0, nface))
    return out


def candidates(n, esz, column_ok, npencil=1 << 30):
    """Tiles (bs, kmax, px) that hold pencils of n cells with at most half padding."""
    out = []
    for e, bs, k in _build.fluid_variants():
        if e != esz:
            continue
        for px in (1, 2, 4, 8, 16, 32):
            if bs % px:
                continue
            tl = bs // px
            cells = tl * k
            if cells < n or (cells >= 2 * n and px > 1) or tl < 8 or px > npencil:
                continue
            out.append((bs, k, px))
    if column_ok:
        out.append((0, 0, 0))
    return out


def load_table(path=TABLE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tune(sims, precision="fp32", copies_cells=2_000_000, steps=2, table=None, log=None,
         max_candidates=12):
    """Time the candidates of every sweep class of `sims` (a list of (ini, state)) that
    the table lacks; returns the updated table {class: [bs, kmax, px, us]}."""
    table = dict(table or {})
    esz = 8 if precision == "fp64" else 4
    reps = {}
    for ini, st in sims:
        for key, X, n, col, nface in sweep_classes(ini, precision):
            if key not in table and key not in reps:
                reps[key] = (ini, st, X, n, col, nface)
    for key, (ini, st, X, n, col, nface) in reps.items():
        rc = from_ini(ini)
        m = int(max(1, min(64, copies_cells // max(1, rc.ncell))))
        ens = Ensemble([(ini, st)] * m, precision)
        ens.step(1)
        best = None
        # tiles that hold every pencil length of the class (rounded up to a power of 2)
        for bs, k, px in candidates(_pow2(n), esz, col, nface)[:max_candidates]:
            try:
                for i in range(m):
                    ens.set_fluid_tile(i, X, bs, k, px)
                ens.profile_step(1)  # plans and graphs for this tile
                prof = ens.profile_step(steps)
            except Exception as e:  # a tile the sweep rejects
                if log:
                    log("autotune %s: tile %s rejected (%s)" % (key, (bs, k, px), e))
                continue
            us = prof["fluid"][0] * 1e3 / max(prof["fluid"][2], 1)
            if best is None or us < best[3]:
                best = (bs, k, px, us)
        ens.close()
        if best is not None:
            table[key] = list(best)
            if log:
                log("autotune %s: bs=%d kmax=%d px=%d (%.1f us per launch)" % ((key,) + best))
    return table


def _shared_single_pencil_tiles(sims, table, precision):
    """Single-pencil sweeps (1-D simulations: one pencil per simulation along x) are
    latency-bound: a launch over them lasts as long as its longest pencil, and every
    distinct tile is a launch of its own.  So all single-pencil classes of a family
    (precision, dimension, MHD, boundary kind) take the tuned tile of the family's
    longest class present, and run as one launch.  Returns {class key: tile}."""
    longest = {}
    for ini, _ in sims:
        for key, X, n, _, nface in sweep_classes(ini, precision):
            if nface != 1 or key not in table:
                continue
            fam = key.split("_n")[0] + "_" + key.split("_", 3)[3].split("_", 1)[1]
            if fam not in longest or n > longest[fam][0]:
                longest[fam] = (n, key)
    out = {}
    for ini, _ in sims:
        for key, X, n, _, nface in sweep_classes(ini, precision):
            if nface != 1:
                continue
            fam = key.split("_n")[0] + "_" + key.split("_", 3)[3].split("_", 1)[1]
            if fam in longest:
                t = table[longest[fam][1]]
                if int(t[0]) * int(t[1]) >= n:
                    out[key] = t
    return out


def apply(ens, sims, table, precision="fp32", share_single_pencil=True):
    """Set the tuned tile of every sweep of every simulation of an Ensemble."""
    shared = _shared_single_pencil_tiles(sims, table, precision) if share_single_pencil else {}
    n = 0
    for i, (ini, _) in enumerate(sims):
        for key, X, _, _, _ in sweep_classes(ini, precision):
            t = shared.get(key, table.get(key))
            if t is None:
                continue
            ens.set_fluid_tile(i, X, int(t[0]), int(t[1]), int(t[2]))
            n += 1
    return n


def save_table(table, path=TABLE):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
