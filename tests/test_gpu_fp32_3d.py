"""The FP32-bulk path on 3-D grids against the FP64 oracle.

Sedov-Taylor blast on 32^3 (the hydro 3-D configuration at a size the oracle runs in
seconds): with transparent ends (the fused sweep on every direction, edge trackers), and
with constant ends and the y / z sweeps forced onto the column march.  Fields are FP32
(pressure, floor, dt and sums FP64), so the result differs from the oracle by FP32
rounding: tolerance density L1 <= 1e-4 (the stated FP32-bulk bar), and every other field
within 1e-4 L1 as well."""
import pytest

from imogen_amd import problems

from .helpers import FLUID_FIELDS, make_gpu, make_oracle, rel_l1, rel_linf

pytestmark = pytest.mark.gpu

ITERS = 30


def sedov32(bc):
    ini, st = problems.sedov(32, time_max=1.0)
    ini = dict(ini, bcMode=bc)
    return ini, st


@pytest.mark.parametrize("bc,column", [("trans", False), ("const", True)])
def test_sedov_32_fp32_bulk_vs_oracle(bc, column):
    ini, st = sedov32(bc)
    run = make_gpu(ini, st, "fp32")
    if column:
        run.set_fluid_tile(2, 0, 0, 0)
        run.set_fluid_tile(3, 0, 0, 0)
        assert run.get_fluid_tile(2)[0] == 0 and run.get_fluid_tile(3)[0] == 0
    orc = make_oracle(ini, st)
    run.run_iterations(ITERS)
    for _ in range(ITERS):
        orc.iteration()
    assert run.time.iteration == orc.iteration_count
    assert abs(run.time.time - orc.time) <= 1e-5 * orc.time
    for f in FLUID_FIELDS:
        e1 = rel_l1(run.download(f), orc.get(f))
        einf = rel_linf(run.download(f), orc.get(f))
        print("sedov 32^3 %s column=%s %s: L1 %.2e L-inf %.2e" % (bc, column, f, e1, einf),
              flush=True)
        assert e1 <= 1e-4, (f, e1)
    run.close()
    orc.close()
