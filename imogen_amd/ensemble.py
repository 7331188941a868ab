"""Batched ensembles: many independent simulations advanced by one set of launches.

Wraps the imh_batch_* entry points (include/imogen_hip.h).  Every simulation keeps its
own grid, spacing, gamma, boundary conditions, limiter and time limits; all share the
precision and the device.  Each batched iteration is a replay of a captured hipGraph in
which one persistent launch per kernel kind walks the tiles of every simulation.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD_IDS, ImhConfig, as_field, check, dptr, lib
from .config import from_ini
from .imogen import make_imh_config

PRECISIONS = {"fp64": _lib.FP64, "fp32": _lib.FP32_BULK}


class Ensemble:
    """A batch of simulations, each given as (ini, state) like the problem generators."""

    def __init__(self, sims, precision="fp32", device=0, perms_order=0):
        self.precision = PRECISIONS[precision]
        self.configs = [from_ini(ini, perms_order=perms_order) for ini, _ in sims]
        cfgs = (ImhConfig * len(sims))(*[make_imh_config(rc, self.precision, device)
                                         for rc in self.configs])
        h = ctypes.c_void_p()
        check(lib().imh_batch_create(cfgs, len(sims), ctypes.byref(h)))
        self._b = h
        self._dlines = []
        for i, rc in enumerate(self.configs):
            for d, line in enumerate(rc.dgrid_lines):
                if line is not None:
                    arr = np.ascontiguousarray(line, dtype=np.float64)
                    self._dlines.append(arr)
                    check(lib().imh_batch_set_spacing(self._b, i, d + 1, dptr(arr), arr.size))
        for i, (_, st) in enumerate(sims):
            self.upload_state(i, st)

    def __len__(self):
        return len(self.configs)

    def close(self):
        if getattr(self, "_b", None):
            lib().imh_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cells(self):
        """Cells per simulation."""
        return np.array([rc.ncell for rc in self.configs], dtype=np.int64)

    def upload_state(self, i, st):
        n = self.configs[i].ncell
        arrs = [as_field(a, n) for a in (st["mass"], *st["mom"], st["ener"], *st["mag"])]
        P = ctypes.POINTER(ctypes.c_double)
        check(lib().imh_batch_upload_state(self._b, int(i), dptr(arrs[0]),
                                           (P * 3)(*[dptr(a) for a in arrs[1:4]]), dptr(arrs[4]),
                                           (P * 3)(*[dptr(a) for a in arrs[5:8]]), n))

    def download(self, i, field):
        rc = self.configs[i]
        out = np.empty(rc.ncell, dtype=np.float64)
        check(lib().imh_batch_download_field(self._b, int(i), FIELD_IDS[field], dptr(out),
                                             out.size))
        return out.reshape(rc.shape)

    def step(self, iterations):
        """Advance every running simulation by `iterations` batched iterations (a fresh
        simulation first runs one iteration on its own)."""
        check(lib().imh_batch_step(self._b, int(iterations)))

    def profile_step(self, iterations):
        """Batched iterations run eagerly with HIP events around every launch; returns
        {kind: (ms, algorithmic bytes, launches)}."""
        return _lib.run_profile(lib().imh_batch_profile_step, self._b, iterations)

    def set_fluid_tile(self, i, X, bs, kmax, px):
        """Pin the fluid sweep tile of simulation i along X (1..3)."""
        check(lib().imh_batch_set_fluid_tile(self._b, int(i), int(X), int(bs), int(kmax),
                                             int(px)))

    def set_fused_1d(self, enable):
        """One launch per iteration for the 1-D MHD simulations of an FP32 batch (the
        default) or the per-sweep launches; the results are the same bit for bit."""
        check(lib().imh_batch_set_fused_1d(self._b, 1 if enable else 0))

    def set_gravity(self, i, potential, mass_threshold=0.0):
        """Fixed gravitational potential of simulation i (None turns it off)."""
        if potential is None:
            check(lib().imh_batch_set_gravity(self._b, int(i), None, 0, float(mass_threshold)))
            return
        arr = as_field(potential, self.configs[i].ncell)
        check(lib().imh_batch_set_gravity(self._b, int(i), dptr(arr), arr.size,
                                          float(mass_threshold)))

    def set_radiation(self, i, kind, exponent=0.5, initial_maximum=1.0):
        """Radiation of simulation i ('empty' / 'optically_thin'), after its state upload."""
        t = {"empty": 0, "optically_thin": 1}[kind]
        check(lib().imh_batch_set_radiation(self._b, int(i), t, float(exponent),
                                            float(initial_maximum)))

    def set_viscosity(self, i, kind, linear=0.0, quadratic=0.0):
        """Artificial viscosity of simulation i."""
        t = {"empty": 0, "neumann_richtmyer": 1, "caramana_shashkov_whalen": 2}[kind]
        check(lib().imh_batch_set_viscosity(self._b, int(i), t, float(linear), float(quadratic)))

    def synchronize(self):
        check(lib().imh_batch_synchronize(self._b))

    def time(self, i):
        t, dt, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(lib().imh_batch_get_time(self._b, int(i), ctypes.byref(t), ctypes.byref(dt),
                                       ctypes.byref(it)))
        return t.value, dt.value, it.value

    def memory_bytes(self):
        b = ctypes.c_size_t()
        check(lib().imh_batch_memory_bytes(self._b, ctypes.byref(b)))
        return b.value
