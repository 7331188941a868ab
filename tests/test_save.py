"""Slice saves in the reference's format (save/resultsHandler.m:26-113,
save/SaveManager.m:230-253): file and variable names, struct fields, slice extraction
and MATLAB squeeze, read back with scipy's MATLAB reader."""
import numpy as np
import pytest

from imogen_amd import save


def state_of(shape_zyx, seed=0):
    rng = np.random.default_rng(seed)
    return {f: rng.standard_normal(shape_zyx) for f in
            ("mass", "momx", "momy", "momz", "ener", "magx", "magy", "magz")}


def meta(**kw):
    m = dict(gamma=5.0 / 3.0, dgrid=(0.1, 0.2, 0.3), time=0.25, dt=1e-3, iteration=12,
             iter_max=1000, time_max=1.0, magnet=True)
    m.update(kw)
    return m


def test_padded_number_matches_paths_m():
    assert save.padded_number(12, 4) == "0012"
    assert save.padded_number(12345, 3) == "12345"


def test_slices_follow_get_save_slice():
    nx, ny, nz = 5, 4, 3
    a = np.arange(nx * ny * nz, dtype=float).reshape(nz, ny, nx)  # (z, y, x), x fastest
    xyz = save.to_xyz(a)
    assert xyz[2, 1, 0] == a[0, 1, 2]
    si = (2, 3, 1)  # 1-based SLICEINDEX
    line_x = save.get_save_slice(xyz, 1, si)
    assert line_x.shape == (nx, 1) and np.array_equal(line_x[:, 0], a[0, 2, :])
    line_y = save.get_save_slice(xyz, 2, si)
    assert line_y.shape == (ny, 1) and np.array_equal(line_y[:, 0], a[0, :, 1])
    line_z = save.get_save_slice(xyz, 3, si)
    assert line_z.shape == (nz, 1) and np.array_equal(line_z[:, 0], a[:, 2, 1])
    plane_xy = save.get_save_slice(xyz, 4, si)
    assert plane_xy.shape == (nx, ny) and np.array_equal(plane_xy, a[0].T)
    plane_yz = save.get_save_slice(xyz, 6, si)
    assert plane_yz.shape == (ny, nz) and np.array_equal(plane_yz, a[:, :, 1].T)
    whole = save.get_save_slice(xyz, 7, si)
    assert whole.shape == (nx, ny, nz) and whole[4, 3, 2] == a[2, 3, 4]


def test_squeeze_like_matlab():
    assert save.matlab_squeeze(np.zeros((1, 5, 1))).shape == (5, 1)
    assert save.matlab_squeeze(np.zeros((3, 1, 2))).shape == (3, 2)
    assert save.matlab_squeeze(np.zeros((1, 1, 1))).shape == (1, 1)


def test_saved_file_reads_back(tmp_path):
    from scipy.io import loadmat
    st = state_of((1, 6, 8))
    paths = save.save_state_slices(str(tmp_path), st, kinds=(1, 4, 7), slice_index=(3, 2, 1),
                                   suffix="0012", **meta())
    assert [p.split("/")[-1] for p in paths] == ["1D_X_0012.mat", "2D_XY_0012.mat",
                                                  "3D_XYZ_0012.mat"]
    m = loadmat(paths[1], squeeze_me=False)
    assert "sx_XY_0012" in m
    s = m["sx_XY_0012"][0, 0]
    assert list(s.dtype.names) == ["gamma", "time", "about", "ver", "iter", "mass", "momX",
                                   "momY", "momZ", "ener", "magX", "magY", "magZ", "grav",
                                   "dGrid", "dim"]
    assert np.allclose(s["mass"], st["mass"][0].T)          # [x, y] plane
    assert np.allclose(s["magZ"], st["magz"][0].T)
    assert s["dim"][0] == "2D"
    assert float(s["time"][0, 0]["time"][0, 0]) == 0.25
    assert float(s["iter"][0, 0]) == 12
    assert [float(s["dGrid"][0, d][0, 0]) for d in range(3)] == [0.1, 0.2, 0.3]


def test_hydro_saves_empty_magnet_fields(tmp_path):
    from scipy.io import loadmat
    st = state_of((1, 4, 4))
    p = save.save_state_slices(str(tmp_path), st, kinds=(7,), slice_index=(1, 1, 1),
                               suffix="FINAL", **meta(magnet=False))[0]
    s = loadmat(p)["sx_XYZ_FINAL"][0, 0]
    assert s["magX"].size == 0 and s["grav"].size == 0


def test_gravity_potential_slice_reads_back(tmp_path):
    """With gravity on, grav holds the potential's slice (resultsHandler.m:81-84)."""
    from scipy.io import loadmat
    st = state_of((3, 4, 5))
    st["grav"] = np.arange(60, dtype=float).reshape(3, 4, 5)
    paths = save.save_state_slices(str(tmp_path), st, kinds=(2, 7), slice_index=(2, 1, 3),
                                   suffix="0012", **meta())
    line_y = loadmat(paths[0])["sx_Y_0012"][0, 0]["grav"]
    assert line_y.shape == (4, 1) and np.array_equal(line_y[:, 0], st["grav"][2, :, 1])
    whole = loadmat(paths[1])["sx_XYZ_0012"][0, 0]["grav"]
    assert whole.shape == (5, 4, 3) and whole[4, 3, 2] == st["grav"][2, 3, 4]


@pytest.mark.gpu
def test_run_slices_from_the_device(tmp_path):
    """resultsHandler-format slices of a GPU run hold the downloaded fields."""
    from scipy.io import loadmat

    import imogen_amd
    from imogen_amd import problems
    ini, st = problems.orszag_tang(32)
    run = imogen_amd.initialize(ini, "fp64")
    imogen_amd.create_arrays(run, st)
    run.run_iterations(3)
    paths = save.save_run_slices(run, str(tmp_path), kinds=(4, 7))
    it = run.time.iteration
    s = loadmat(paths[0])["sx_XY_%s" % save.padded_number(it, 4)][0, 0]
    assert np.array_equal(s["mass"], run.download("mass")[0].T)
    assert np.array_equal(s["magY"], run.download("magy")[0].T)
    assert float(s["time"][0, 0]["time"][0, 0]) == run.time.time
    run.close()
