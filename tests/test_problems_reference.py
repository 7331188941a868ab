"""The reference initializers restated (problems.orszag_tang_reference, sedov_reference,
magnetic_shock_tube; experiment/*Initializer.m) checked against what the mathematics of
their construction fixes, on the CPU:

* Orszag-Tang (OrszagTangVortexInitializer.m:72-90): B_x depends on y alone and B_y on
  x alone, so the forward-difference divergence of the periodic face field is exactly 0;
  the sines have period N - 1 in cell index (first and last columns equal to rounding);
  the gas pressure recovered from the energy is 5/(12 pi) in every cell; the spacing is
  1 ./ grid.
* Sedov-Taylor (SedovTaylorBlastWaveInitializer.m:71-86): the blast cells are the
  integer lattice points within the radius of the centre cell (515 for radius 5, OEIS
  A000605), the radius is max(5, round(0.06 min(grid))), the deposited energy is the
  closed form count * 3 / (4 pi r^3) (times the cell volume over (dx r)^3), and an even
  cell count is refused as the reference's ndgrid makes it fail.
* Magnetic shock tube (MagneticShockTubeInitializer.m:79-117): pressure 20 on the left
  and 1 on the right of MATLAB round(N/2) (1-based), momentum +10 / -10 along x, and for a
  y shock -0.0 on the right (the .m text negates the x momentum there).
"""
import math

import numpy as np
import pytest

from imogen_amd import problems


def test_orszag_tang_reference():
    n = 40
    ini, st = problems.orszag_tang_reference(n)
    assert ini["dGrid"] == (1.0 / n, 1.0 / n, 1.0) and ini["gamma"] == 5.0 / 3.0
    bx, by = st["mag"][0], st["mag"][1]
    assert np.all(bx == bx[:, :, :1]) and np.all(by == by[:, :1, :])
    div = (np.roll(bx, -1, axis=2) - bx) + (np.roll(by, -1, axis=1) - by)
    assert np.abs(div).max() == 0.0
    # period N - 1 in cell index: the last column repeats the first (sin(2 pi) ~ 1e-16)
    assert np.abs(st["mom"][1][:, :, -1] - st["mom"][1][:, :, 0]).max() <= 1e-15
    assert np.abs(st["mom"][0][:, -1, :] - st["mom"][0][:, 0, :]).max() <= 1e-15
    g = ini["gamma"]
    ke = 0.5 * (st["mom"][0] ** 2 + st["mom"][1] ** 2) / st["mass"]
    me = 0.5 * (bx ** 2 + by ** 2)
    p = (g - 1.0) * (st["ener"] - ke - me)
    assert np.abs(p - 5.0 / (12.0 * math.pi)).max() <= 1e-15
    assert np.all(st["mass"] == 25.0 / (36.0 * math.pi))


@pytest.mark.parametrize("n,count", [(11, 515), (65, 515), (167, 4169)])
def test_sedov_reference(n, count):
    ini, st = problems.sedov_reference(n)
    g = ini["gamma"]
    radius = max(5, int(math.floor(0.06 * n + 0.5)))
    p = st["ener"] * (g - 1.0)
    hot = p > 1e-5 * (1 + 1e-12)
    assert int(hot.sum()) == count
    # the hot cells are the lattice points within the radius of the centre cell
    c = n // 2
    k, j, i = np.nonzero(hot)
    assert np.all((i - c) ** 2 + (j - c) ** 2 + (k - c) ** 2 <= radius ** 2)
    inside = 3.0 * (g - 1.0) / (4.0 * math.pi * (radius / n) ** 3)
    assert np.allclose(p[hot], inside, rtol=1e-15, atol=0)
    # deposited energy: count cells of volume dx^3 at density inside / (g - 1)
    e = float((p[hot] / (g - 1.0)).sum() / n ** 3)
    assert e == pytest.approx(count * 3.0 / (4.0 * math.pi * radius ** 3), rel=1e-10)
    assert ini["bcMode"] == "fade" and ini["cfl"] == 0.7


def test_sedov_reference_2d_radius():
    # a 2-D grid: min(grid) = 1, so the radius is max(5, round(0.06)) = 5 and the blast
    # holds the 81 lattice points of a disc of radius 5 (OEIS A000328)
    ini, st = problems.sedov_reference(167, nz=1)
    assert int((st["ener"] * 0.4 > 1e-5 * (1 + 1e-12)).sum()) == 81


def test_sedov_reference_refuses_even_grids():
    with pytest.raises(ValueError, match="odd"):
        problems.sedov_reference(64)


@pytest.mark.parametrize("direction", ["X", "Y"])
def test_magnetic_shock_tube(direction):
    grid = (20, 3, 3) if direction == "X" else (3, 20, 3)
    ini, st = problems.magnetic_shock_tube(*grid, direction=direction)
    g = ini["gamma"]
    assert ini["dGrid"] == (1.0, 1.0, 1.0) and g == 1.4
    d = "XYZ".index(direction)
    axis = 2 - d
    line = [slice(0, 1)] * 3
    line[axis] = slice(None)
    line = tuple(line)
    mom = st["mom"][d][line].ravel()
    # MATLAB round(20 / 2) = 10 (1-based): 0-based cells 9.. are the right side
    assert np.all(mom[:9] == 10.0)
    if direction == "X":
        assert np.all(mom[9:] == -10.0)
    else:  # -1.0 * mom(1, right) with mom(1) zero: -0.0
        assert np.all(mom[9:] == 0.0) and np.all(np.signbit(mom[9:]))
    b = 5.0 / math.sqrt(4.0 * math.pi)
    assert np.all(st["mag"][0] == b) and np.all(st["mag"][1] == b) and np.all(st["mag"][2] == 0)
    ke = 0.5 * sum(m * m for m in st["mom"]) / st["mass"]
    me = 0.5 * sum(x * x for x in st["mag"])
    p = ((st["ener"] - ke - me) * (g - 1.0))[line].ravel()
    assert np.allclose(p[:9], 20.0, rtol=1e-14) and np.allclose(p[9:], 1.0, rtol=1e-13)
