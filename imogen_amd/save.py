"""Slice saves in the reference's format, and checkpoints for a bit-exact resume.

Slices follow save/resultsHandler.m:26-113: one MATLAB v5 file per slice type,
`<dir>/<dim>_<label>_<suffix>.mat`, holding one struct variable `sx_<label>_<suffix>` with
the fields gamma, time (TimeManager.toStruct, time/TimeManager.m:183-191), about, ver, iter,
mass, momX, momY, momZ, ener, magX, magY, magZ (empty without a magnetic field), grav
(the slice of the fixed potential when gravity is on, else empty: resultsHandler.m:81-84),
dGrid (a 1x3 cell) and dim.  The slice of an array is
SaveManager.getSaveSlice (save/SaveManager.m:230-253): types 1-3 are the x, y, z lines
through SLICEINDEX, 4-6 the xy, xz, yz planes, 7 the whole grid, each squeezed as MATLAB
squeezes.  The suffix is 'START' for iteration 0, 'FINAL' for the end of a run, else the
zero-padded iteration (Paths.paddedNumber, save/Paths.m:159-166).

The device keeps no per-iteration dt record, so time.history holds the last dt only.

Checkpoints (Run.checkpoint / Run.restore, imh_checkpoint_* in include/imogen_hip.h) hold
what the reference's saves lack -- the edge trackers of the flux objects and their
first-assignment state -- so a resumed run continues bit for bit.
"""
import datetime
import os

import numpy as np

SLICE_LABELS = ("X", "Y", "Z", "XY", "XZ", "YZ", "XYZ")
SLICE_DIM = ("1D", "1D", "1D", "2D", "2D", "2D", "3D")


def padded_number(number, pad_length):
    """Paths.paddedNumber: num2str with leading zeros up to pad_length digits."""
    s = str(int(number))
    return "0" * max(0, pad_length - len(s)) + s


def matlab_squeeze(a):
    """MATLAB squeeze: drop singleton dimensions, but keep at least two (a vector is a
    column)."""
    shape = [n for n in a.shape if n != 1]
    if len(shape) == 0:
        shape = [1, 1]
    elif len(shape) == 1:
        shape = [shape[0], 1]
    return np.asfortranarray(a.reshape(shape, order="F"))


def get_save_slice(a_xyz, kind, slice_index):
    """SaveManager.getSaveSlice of an array indexed [x, y, z] (the reference's layout),
    slice type kind in 1..7, slice_index the 1-based SLICEINDEX."""
    i, j, k = (int(v) - 1 for v in slice_index)
    sel = {1: (slice(None), slice(j, j + 1), slice(k, k + 1)),
           2: (slice(i, i + 1), slice(None), slice(k, k + 1)),
           3: (slice(i, i + 1), slice(j, j + 1), slice(None)),
           4: (slice(None), slice(None), slice(k, k + 1)),
           5: (slice(None), slice(j, j + 1), slice(None)),
           6: (slice(i, i + 1), slice(None), slice(None)),
           7: (slice(None), slice(None), slice(None))}[kind]
    return matlab_squeeze(a_xyz[sel])


def to_xyz(a):
    """A field as this package stores it on the host ((z, y, x), x fastest) -> [x, y, z]."""
    return np.asarray(a).transpose(2, 1, 0)


def slice_struct(state, kind, slice_index, gamma, dgrid, time, dt, iteration, iter_max,
                 time_max, magnet=True, about="", ver=""):
    """The struct resultsHandler.m saves for one slice type (a dict; field order kept)."""
    def sl(name):
        return get_save_slice(to_xyz(state[name]), kind, slice_index)

    now = datetime.datetime.now().strftime("%d-%b-%Y %H:%M:%S")
    t = {"time": float(time), "history": np.array([[float(dt)]]), "iterMax": float(iter_max),
         "timeMax": float(time_max), "wallMax": float("inf"), "started": now,
         "iteration": float(iteration)}
    out = {"gamma": float(gamma), "time": t, "about": about, "ver": ver,
           "iter": float(iteration)}
    for key, name in (("mass", "mass"), ("momX", "momx"), ("momY", "momy"), ("momZ", "momz"),
                      ("ener", "ener")):
        out[key] = sl(name)
    for key, name in (("magX", "magx"), ("magY", "magy"), ("magZ", "magz")):
        out[key] = sl(name) if magnet else np.zeros((0, 0))
    # run.gravity.ACTIVE && numel(grav.array) > 1
    grav = state.get("grav")
    out["grav"] = (get_save_slice(to_xyz(grav), kind, slice_index)
                   if grav is not None and np.size(grav) > 1 else np.zeros((0, 0)))
    cell = np.empty((1, 3), dtype=object)
    for d in range(3):
        cell[0, d] = float(dgrid[d])
    out["dGrid"] = cell
    out["dim"] = SLICE_DIM[kind - 1]
    return out


def slice_file(directory, kind, suffix):
    label = SLICE_LABELS[kind - 1]
    name = "sx_%s_%s" % (label, suffix)
    path = os.path.join(directory, "%s_%s_%s.mat" % (SLICE_DIM[kind - 1], label, suffix))
    return path, name


def save_state_slices(directory, state, kinds, slice_index, suffix, **meta):
    """Write the slice files of a host state {field: array (z, y, x)}; returns the paths."""
    from scipy.io import savemat
    os.makedirs(directory, exist_ok=True)
    paths = []
    for kind in kinds:
        path, name = slice_file(directory, kind, suffix)
        savemat(path, {name: slice_struct(state, kind, slice_index, **meta)},
                long_field_names=True, do_compression=False)
        paths.append(path)
    return paths


def save_run_slices(run, directory, kinds=(7,), slice_index=None, final=False, about="",
                    ver=""):
    """resultsHandler.m for a GPU run: download the fields and write the slice files.
    slice_index defaults to the grid centre (1-based)."""
    rc = run.config
    grid = rc.grid
    if slice_index is None:
        slice_index = tuple(max(1, (n + 1) // 2) for n in grid)
    iteration = run.time.iteration
    suffix = "FINAL" if final else ("START" if iteration == 0 else
                                    padded_number(iteration, len(str(rc.iter_max))))
    names = ["mass", "momx", "momy", "momz", "ener", "magx", "magy", "magz"]
    state = {f: run.download(f) for f in names}
    if getattr(run, "gravity_potential", None) is not None:
        state["grav"] = run.gravity_potential
    return save_state_slices(directory, state, kinds, slice_index, suffix,
                             gamma=rc.gamma, dgrid=rc.dgrid, time=run.time.time,
                             dt=run.time.dTime, iteration=iteration, iter_max=rc.iter_max,
                             time_max=rc.time_max, magnet=bool(rc.magnet_active),
                             about=about, ver=ver)
