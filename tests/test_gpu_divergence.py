"""Discrete divergence of B on the device: Orszag-Tang 512^2 (the MHD configuration at its
stated size), 1000 iterations, FP64 and FP32 bulk.

The constrained-transport step (relaxingMagnet.m:51,56-60) changes B_I by
-(dt/dx_X)(F(i) - F(i-e_X)) and B_X by -(dt/dx_I)(F(i-e_X) - F(i-e_X+e_I)), so the
forward-difference divergence sum_d (B_d(i+e_d) - B_d(i)) / dx_d of a periodic grid keeps
its initial value, 0 for Orszag-Tang, up to the rounding of the updates (which adds up
over the run).  A sign or index error in the CT lines breaks it at O(1) of max|B|/dx.
Bounds, relative to max|B|/dx: FP64 1e-11 (2000 timesteps of rounding at 1e-16), FP32
1e-4 (the same at 6e-8)."""
import numpy as np
import pytest

from imogen_amd import problems

from .helpers import make_gpu

pytestmark = pytest.mark.gpu


def fwd_div(run, dg):
    bx, by = run.download("magx"), run.download("magy")
    return (np.roll(bx, -1, axis=2) - bx) / dg[0] + (np.roll(by, -1, axis=1) - by) / dg[1]


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-11), ("fp32", 1e-4)])
def test_orszag_tang_512_keeps_divergence(precision, bound):
    ini, st = problems.orszag_tang(512, iter_max=1000)
    dg = ini["dGrid"]
    scale = max(np.abs(b).max() for b in st["mag"]) / dg[0]
    run = make_gpu(ini, st, precision)
    d0 = np.abs(fwd_div(run, dg)).max() / scale
    assert run.run_iterations(1000) == 1000
    d1 = np.abs(fwd_div(run, dg)).max() / scale
    print("OT 512^2 %s: max |div B| / (max|B|/dx) %.2e at start, %.2e after 1000 iterations"
          % (precision, d0, d1), flush=True)
    assert d0 <= 1e-15 if precision == "fp64" else d0 <= 1e-6
    assert d1 <= bound, d1
    run.close()
