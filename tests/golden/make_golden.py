"""Independent NumPy restatement of the reference's relaxing-TVD path; writes the small
golden fixtures in this directory.

This is a second, separately written restatement of the same .m files as the C oracle
(oracle/imogen_oracle.c), kept close to the MATLAB text: whole arrays indexed with
MATLAB's index vectors (the *_shift.m files), one object per reference array object
with its own Edges state, and scipy's PchipInterpolator for MATLAB's pchip.  The C
oracle is checked against these fixtures in tests/test_oracle.py and
tests/test_oracle_source.py; agreement pins
the oracle's transcription (index maps, evaluation order, edge-tracker sequence).

Arrays here use the reference's dimension order: shape (nx, ny) or (nx, ny, nz).
Run:  python tests/golden/make_golden.py   (writes tests/golden/*.npz)
"""
import os
import sys

import numpy as np
from scipy.interpolate import PchipInterpolator

HERE = os.path.dirname(os.path.abspath(__file__))
FLUXL, FLUXR, FLUX, OTHER = "fluxl", "fluxr", "flux", "unspecified"


# ------------------------------------------------------------------------------------
# shifting/*_shift.m with MATLAB index vectors (0-based here)
# ------------------------------------------------------------------------------------
def _idx(shape):
    return [np.arange(s) for s in shape]


def _take(a, idx):
    return a[np.ix_(*idx)]


def circ_shift(a, D, n, obj):
    N = a.shape
    NDim = N[D]
    idx = _idx(N)
    d = -np.sign(n)
    n = abs(n)
    base = idx[D].copy()
    if d > 0:
        idx[D][n:] = base[:NDim - n]
        idx[D][:n] = np.arange(NDim - n, NDim)
    else:
        idx[D][:NDim - n] = base[n:]
        idx[D][NDim - n:] = np.arange(n)
    return _take(a, idx)


def constant_shift(a, D, n, obj):
    N = a.shape
    NDim = N[D]
    idx = _idx(N)
    d = -np.sign(n)
    n = abs(n)
    base = idx[D].copy()
    if d > 0:
        idx[D][n:] = base[:NDim - n]
        idx[D][:n] = 0
    else:
        idx[D][:NDim - n] = base[n:]
        idx[D][NDim - n:] = NDim - 1
    return _take(a, idx)


def _pchip_along(vals, D, B, xi):
    y = np.moveaxis(vals, D, -1)
    p = PchipInterpolator(np.array([1.0, 2.0, 3.0, B, B + 1.0]), y, axis=-1)
    out = p(np.asarray(xi, dtype=float))
    return np.moveaxis(out, -1, D)


def transparent_shift(a, D, n, obj):
    N = a.shape
    NDim = N[D]
    idx = _idx(N)
    d = -np.sign(n)
    n = abs(n)
    base = idx[D].copy()
    B = float(obj.sim.bc_infinity)
    if d > 0:
        idx[D][n:] = base[:NDim - n]
        idx[D][:n] = 0
        result = _take(a, idx)
        inSize = list(N)
        inSize[D] = 5
        inVals = np.zeros(inSize)
        inIdx = [i.copy() for i in idx]
        inIdx[D] = np.arange(3)
        idx[D] = np.minimum(np.array([3, 2, 1]) + n, NDim) - 1
        inVals[np.ix_(*inIdx)] = _take(result, idx)
        if obj.active_linear(D):
            e = obj.transparent_edge(D, False)
            sl = [slice(None)] * len(N)
            sl[D] = 3
            inVals[tuple(sl)] = e
            sl[D] = 4
            inVals[tuple(sl)] = e
        outVals = _pchip_along(inVals, D, B, 3 + np.arange(n, 0, -1))
        idx[D] = np.arange(n)
    else:
        idx[D][:NDim - n] = base[n:]
        idx[D][NDim - n:] = NDim - 1
        result = _take(a, idx)
        inSize = list(N)
        inSize[D] = 5
        inVals = np.zeros(inSize)
        inIdx = [i.copy() for i in idx]
        inIdx[D] = np.arange(3)
        idx[D] = np.maximum(NDim - np.array([2, 1, 0]) - n, 1) - 1
        inVals[np.ix_(*inIdx)] = _take(result, idx)
        if obj.active_linear(D):
            e = obj.transparent_edge(D, True)
            sl = [slice(None)] * len(N)
            sl[D] = 3
            inVals[tuple(sl)] = e
            sl[D] = 4
            inVals[tuple(sl)] = e
        outVals = _pchip_along(inVals, D, B, 3 + np.arange(1, n + 1))
        idx[D] = np.arange(NDim - n, NDim)
    orig = _take(a, idx)
    test = np.abs(orig - outVals) < 1e-4
    outVals = (~test) * outVals + test * orig
    result[np.ix_(*idx)] = outVals
    return result


def fade_shift(a, D, n, obj):
    N = a.shape
    f = obj.sim.bc_infinity
    result = constant_shift(a, D, n, obj)
    coeff = obj.sim.fade
    cshape = [1] * len(N)
    cshape[D] = f
    c = coeff.reshape(cshape)
    sl = [slice(None)] * len(N)
    sl[D] = slice(0, 1)
    fadeArray = result[tuple(sl)]
    sl[D] = slice(0, f)
    result[tuple(sl)] = (1 - c) * result[tuple(sl)] + c * fadeArray
    sl[D] = slice(N[D] - 1, N[D])
    fadeArray = result[tuple(sl)]
    sl[D] = slice(N[D] - f, N[D])
    cf = np.flip(c, axis=D)
    result[tuple(sl)] = (1 - cf) * result[tuple(sl)] + cf * fadeArray
    return result


def mirror_shift(a, D, n, obj):
    N = a.shape
    NDim = N[D]
    idx = _idx(N)
    d = -np.sign(n)
    n = abs(n)
    base = idx[D].copy()
    if d > 0:
        idx[D][n:] = base[:NDim - n]
        flip = np.arange(n)
        idx[D][flip] = np.arange(n + 1, 1, -1) - 1
    else:
        idx[D][:NDim - n] = base[n:]
        flip = np.arange(NDim - n, NDim)
        idx[D][flip] = np.arange(NDim - 1, NDim - n - 1, -1) - 1
    result = _take(a, idx)
    if obj.component == D + 1:
        sl = [slice(None)] * len(N)
        sl[D] = flip
        result[tuple(sl)] = -result[tuple(sl)]
    return result


def flip_shift(a, D, n, obj):
    N = a.shape
    result = constant_shift(a, D, n, obj)
    if obj.component == D + 1:
        m = abs(n)
        sl = [slice(None)] * len(N)
        sl[D] = slice(0, m) if n < 0 else slice(N[D] - m, N[D])
        result[tuple(sl)] = -result[tuple(sl)]
    return result


def wall_shift(a, D, n, obj):
    N = a.shape
    result = constant_shift(a, D, n, obj)
    m = abs(n)
    sl = [slice(None)] * len(N)
    if n < 0:
        sl[D] = slice(0, m)
        edge = obj.transparent_edge(D, False)
    else:
        sl[D] = slice(N[D] - m, N[D])
        edge = obj.transparent_edge(D, True)
    result[tuple(sl)] = np.expand_dims(edge, D)
    return result


def zero_shift(a, D, n, obj):
    if obj.kind == "flux":
        N = a.shape
        m = abs(n)
        if n < 0:
            if obj.flux_type == FLUXL:
                return constant_shift(a, D, n, obj)
            result = np.zeros(N)
            sl_out = [slice(None)] * len(N)
            sl_in = [slice(None)] * len(N)
            sl_out[D] = slice(m, N[D])
            sl_in[D] = slice(0, N[D] - m)
        else:
            if obj.flux_type == FLUXR:
                return constant_shift(a, D, n, obj)
            result = np.zeros(N)
            sl_out = [slice(None)] * len(N)
            sl_in = [slice(None)] * len(N)
            sl_out[D] = slice(0, N[D] - m)
            sl_in[D] = slice(m, N[D])
        result[tuple(sl_out)] = a[tuple(sl_in)]
        return result
    return constant_shift(a, D, n, obj)


SHIFTS = {"circ": circ_shift, "const": constant_shift, "trans": transparent_shift,
          "fade": fade_shift, "mirror": mirror_shift, "flip": flip_shift, "wall": wall_shift,
          "zero": zero_shift}


# ------------------------------------------------------------------------------------
# Objects (ImogenArray / Edges)
# ------------------------------------------------------------------------------------
class Obj:
    def __init__(self, sim, kind, component, flux_type=OTHER):
        self.sim, self.kind, self.component, self.flux_type = sim, kind, component, flux_type
        self.a = None
        self.edges = None
        self.active = np.zeros((2, 3), dtype=bool)

    def make_edges(self):
        a = self.a
        self.edges = {}
        self.active[:] = False
        for side in range(2):
            for i in range(a.ndim):
                mode = self.sim.bc[side][i]
                if mode in ("trans", "fade", "wall"):
                    self.active[side, i] = True
                if mode in ("trans", "wall"):
                    sl = [slice(None)] * a.ndim
                    sl[i] = 0 if side == 0 else a.shape[i] - 1
                    self.edges[(side, i)] = a[tuple(sl)].copy()

    def assign(self, value):
        self.a = value
        if self.kind in ("init", "store", "flux") and self.edges is None:
            self.make_edges()

    def active_linear(self, D):
        return bool(self.active.flatten(order="F")[D])

    def transparent_edge(self, D, upper):
        side = 1 if upper else 0
        mode = self.sim.bc[side][D]
        old = self.edges[(side, D)]
        if mode == "trans":
            sl = [slice(None)] * self.a.ndim
            sl[D] = 0 if side == 0 else self.a.shape[D] - 1
            new = self.a[tuple(sl)]
            delta = np.fmin(np.abs(new - old), 0.005)
            sgn = (new - old) > 0
            res = old + (sgn.astype(float) - (~sgn).astype(float)) * delta
            self.edges[(side, D)] = res
            return res
        return old

    def shift(self, D1, n):
        D = D1 - 1
        if D >= self.a.ndim:
            return self.a.copy()
        side = 1 if n > 0 else 0
        return SHIFTS[self.sim.bc[side][D]](self.a, D, n, self)


# ------------------------------------------------------------------------------------
# pressure.m, limiters, wFluidFlux.m, relaxingFluid.m, magnetFlux.m, relaxingMagnet.m
# ------------------------------------------------------------------------------------
def msign(x):
    return np.sign(x)  # NaN -> NaN, 0 -> 0


def vanleer(flux, dl, dr):
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.fmax(dl * dr, 0) / (dl + dr)
    t[np.isnan(t)] = 0
    flux.assign(flux.a + t)


def minmod(flux, dl, dr):
    flux.assign(flux.a + 0.125 * (msign(dl) + msign(dr)) * np.fmin(np.abs(dl), np.abs(dr)))


def superbee(flux, dl, dr):
    c = (np.abs(dl) < np.abs(dr)).astype(float)
    minmod(flux, (1 + c) * dl, (2 - c) * dr)


LIMITERS = {"vanleer": vanleer, "superbee": superbee, "minmod": minmod}


class Sim:
    def __init__(self, grid, dgrid, gamma, bc, limiter, magnet, cfl, minmass=1e-5,
                 bc_infinity=20):
        nx, ny, nz = grid
        self.shape = (nx, ny, nz) if nz > 1 else (nx, ny)
        self.grid = grid
        self.dgrid = dgrid
        self.gamma = gamma
        self.bc = bc
        self.limiter = limiter
        self.magnet_active = magnet
        self.cfl = cfl
        self.minmass = minmass
        self.bc_infinity = bc_infinity
        f = bc_infinity
        x = np.array([0, np.ceil(f / 4), np.floor(f / 2 + 0.5), f - 1, f], dtype=float)
        self.fade = PchipInterpolator(x, [1, 1, 0.01, 0, 0])(np.arange(1, f + 1))
        self.time = 0.0
        self.dtime = 0.0
        self.iteration = 1
        self.grav = None          # GravityArray with the fixed potential, or None
        self.threshold = 0.0      # run.fluid.MASS_THRESHOLD
        self.rad = None           # (exponent, strength) of the optically thin sink
        self.visc = None          # ("nr" | "csw", linear, quadratic)

    def set_state(self, mass, mom, ener, mag):
        comp = [0, 1, 2, 3, 0]
        self.prim = []
        for i, a in enumerate([mass, *mom, ener]):
            o = Obj(self, "fluid", comp[i])
            o.a = a.copy()
            o.make_edges()
            self.prim.append(o)
        self.store = [Obj(self, "store", c) for c in comp]
        self.sfr = [Obj(self, "flux", c, FLUXR) for c in comp]
        self.sfl = [Obj(self, "flux", c, FLUXL) for c in comp]
        self.fr = [Obj(self, "flux", c, FLUXR) for c in comp]
        self.fl = [Obj(self, "flux", c, FLUXL) for c in comp]
        self.freeze = [Obj(self, "flux", k + 1) for k in range(3)]
        self.freeze_tvd = [Obj(self, "flux", k + 1) for k in range(3)]
        self.mag, self.cmag = [], []
        self.is_zero = []
        self.vg = {}
        self.mstore = {}
        self.wmag = {}
        self.mflux = {}
        for k in range(3):
            o = Obj(self, "magnet", k + 1)
            o.a = mag[k].copy()
            o.make_edges()
            self.mag.append(o)
            self.is_zero.append(bool(np.sum(mag[k]) == 0) and not self.magnet_active)
            for X in range(3):
                self.vg[(k, X)] = Obj(self, "init", k + 1)
                st = Obj(self, "store", k + 1)
                st.fluxR = Obj(self, "flux", k + 1, FLUXR)
                self.mstore[(k, X)] = st
                self.wmag[(k, X)] = Obj(self, "init", k + 1)
                self.mflux[(k, X)] = Obj(self, "flux", k + 1, FLUX)
            self.cmag.append(Obj(self, "init", k + 1))
            self.update_cell_centered(k)

    def update_cell_centered(self, k):
        self.cmag[k].assign(0.5 * (self.mag[k].a + self.mag[k].shift(k + 1, 1)))

    def pressure(self, mode, rho, mom, ener):
        g = self.gamma
        v2 = (mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2]) / (rho * rho)
        b2 = 0
        for i in range(3):
            if not self.is_zero[i]:
                b2 = b2 + self.cmag[i].a * self.cmag[i].a
        r = (g - 1.0) * (ener - 0.5 * rho * v2)
        aux = None
        if mode == "gas":
            r = r - (g - 1.0) * 0.5 * b2
            return np.where(r < 0, 0.0, r), None
        if mode == "totsnd":
            aux = np.sqrt(np.abs((g * (r - (g - 1.0) * 0.5 * b2) + 2.0 * b2) / rho))
            r = r + 0.5 * (2.0 - g) * b2
        else:
            r = r - (g - 1.0) * 0.5 * b2
            r = np.sqrt(np.abs((g * r + 2.0 * b2) / rho))
        r = np.where(r < 0, 0.0, r)
        return r, aux

    def viscosity(self, objs, snd, X):
        """ArtificialViscosity.m: Neumann-Richtmyer (:73-77) or Caramana-Shashkov-Whalen
        (:80-148, cswSoundSpeed :155-190) on the objects of the five fields."""
        kind, c1, c2 = self.visc
        X1 = X + 1
        mass, mom, ener = objs[0], objs[1:4], objs[4]
        if kind == "nr":
            vm = np.fmin(mom[X].shift(X1, 1) - mom[X].a, 0)
            return vm * (c1 * snd + c2 * vm / mass.a)
        g = self.gamma
        massB = mass.shift(X1, -1)
        massC = mass.shift(X1, 1)
        rho = 2 * massC * massB / (massC + massB)
        sides = []
        for n in (1, -1):
            v2 = (mom[0].shift(X1, n) ** 2 + mom[1].shift(X1, n) ** 2
                  + mom[2].shift(X1, n) ** 2 / mass.shift(X1, n) ** 2)
            b2 = (self.cmag[0].shift(X1, n) ** 2 + self.cmag[1].shift(X1, n) ** 2
                  + self.cmag[2].shift(X1, n) ** 2)
            sp = (g - 1.0) * (ener.shift(X1, n) - 0.5 * mass.shift(X1, n) * v2)
            sp = sp - (g - 1.0) * 0.5 * b2
            sides.append(np.sqrt(np.abs((g * sp + 2.0 * b2) / mass.shift(X1, n))))
        cs = np.fmax(np.fmin(sides[1], sides[0]), 0)
        v0 = mom[X].shift(X1, -2) / mass.shift(X1, -2)
        v1 = mom[X].shift(X1, -1) / mass.shift(X1, -1)
        v2 = mom[X].shift(X1, 1) / mass.shift(X1, 1)
        v3 = mom[X].shift(X1, 2) / mass.shift(X1, 2)
        dv = v2 - v1
        rL = (v3 - v2) / (2 * dv)
        rR = (v1 - v0) / (2 * dv)
        lim = np.fmax(0, np.fmin(np.fmin(0.5 * (rL + rR), np.fmin(2 * rL, 2 * rR)), 1))
        return ((dv < 0) * rho * (c2 * (g + 1) * np.abs(dv) / 4
                                  + np.sqrt(c2 ** 2 * ((g + 1) / 4) ** 2 * dv ** 2
                                            + c1 ** 2 * cs ** 2))
                * (1 - lim) * np.abs(dv))

    def wfluidflux(self, q, freeze, X, objs=None):
        rho, mom, E = q[0], q[1:4], q[4]
        press, snd = self.pressure("totsnd", rho, mom, E)
        vel = mom[X] / rho
        spd = np.fmax.reduce(np.abs(vel) + snd, axis=X, keepdims=True)
        freeze.assign(np.broadcast_to(spd, rho.shape).copy())
        press = press + (self.viscosity(objs, snd, X) if self.visc else 0)
        C = freeze.a
        bc = [c.a for c in self.cmag]
        w = [None] * 5
        w[0] = mom[X] / C
        w[4] = (vel * (E + press) - bc[X] * (bc[0] * mom[0] + bc[1] * mom[1] + bc[2] * mom[2])
                / rho) / C
        for i in range(3):
            dv = 1.0 if i == X else 0.0
            w[i + 1] = (vel * mom[i] + press * dv - bc[X] * bc[i]) / C
        return w

    def relaxing_fluid(self, X1):
        X = X1 - 1
        ff = self.dtime / self.dgrid[X]
        q = [o.a for o in self.prim]
        w = self.wfluidflux(q, self.freeze[X], X, self.prim)
        C = self.freeze[X].a
        for i in range(5):
            self.sfr[i].assign(0.5 * (q[i] + w[i]))
            self.sfl[i].assign(0.5 * (q[i] - w[i]))
            self.store[i].assign(q[i] - 0.5 * ff * C * (self.sfr[i].a - self.sfr[i].shift(X1, -1)
                                                        + self.sfl[i].a - self.sfl[i].shift(X1, 1)))
        self.store[0].assign(np.fmax(self.store[0].a, self.minmass))
        qs = [o.a for o in self.store]
        w = self.wfluidflux(qs, self.freeze_tvd[X], X, self.store)
        C = self.freeze_tvd[X].a
        lim = LIMITERS[self.limiter[X]]
        for i in range(5):
            fr, fl = self.fr[i], self.fl[i]
            fr.assign(0.5 * (qs[i] + w[i]))
            fl.assign(0.5 * (qs[i] - w[i]))
            lim(fr, fr.a - fr.shift(X1, -1), fr.shift(X1, 1) - fr.a)
            lim(fl, fl.shift(X1, -1) - fl.a, fl.a - fl.shift(X1, 1))
            self.prim[i].a = self.prim[i].a - ff * C * (fr.a - fr.shift(X1, -1) + fl.a - fl.shift(X1, 1))
        self.prim[0].a = np.fmax(self.prim[0].a, self.minmass)

    def relaxing_magnet(self, vg, X1, I1):
        X, I = X1 - 1, I1 - 1
        ff = 0.5 * self.dtime / self.dgrid[X]
        vf = ((vg.a + vg.shift(X1, 1)) < 0.0).astype(float)
        B = self.mag[I]
        st = self.mstore[(I, X)]
        st.fluxR.assign(B.a * vg.a)
        st.fluxR.assign(st.fluxR.a * (1 - vf) + st.fluxR.shift(X1, 1) * vf)
        st.assign(B.a - ff * (st.fluxR.a - st.fluxR.shift(X1, -1)))
        ff = 2 * ff
        wm = self.wmag[(I, X)]
        fl = self.mflux[(I, X)]
        wm.assign(st.a * vg.a)
        fl.assign(wm.a * (1 - vf) + wm.shift(X1, 1) * vf)
        dR = wm.shift(X1, 1) - (vf * wm.shift(X1, 2) + (1 - vf) * wm.a)
        dL = wm.a - (vf * wm.shift(X1, 1) + (1 - vf) * wm.shift(X1, -1))
        LIMITERS[self.limiter[X]](fl, dL, dR)
        B.a = B.a - ff * (fl.a - fl.shift(X1, -1))
        ffi = self.dtime / self.dgrid[I]
        fl.assign(fl.a - fl.shift(I1, 1))
        fl.assign(fl.shift(X1, -1))
        self.mag[X].a = self.mag[X].a - ffi * fl.a

    def magnet_flux(self, X1, I2):
        X = X1 - 1
        for I1 in I2:
            I = I1 - 1
            vg = self.vg[(I, X)]
            vg.assign((self.prim[X1].shift(I1, -1) + self.prim[X1].a)
                      / (self.prim[0].shift(I1, -1) + self.prim[0].a))
            vg.assign(0.25 * (vg.shift(X1, -1) + vg.shift(X1, 1) + 2 * vg.a))
            self.relaxing_magnet(vg, X1, I1)
        self.update_cell_centered(I1 - 1)
        self.update_cell_centered(X)

    def time_update(self):
        cs, _ = self.pressure("sound", self.prim[0].a, [o.a for o in self.prim[1:4]], self.prim[4].a)
        cmax, gi = 1e-2, 0
        for i in range(3):
            c = np.fmax.reduce((np.abs(self.prim[i + 1].a / self.prim[0].a) + cs).ravel())
            if c > cmax:
                cmax, gi = c, i + 1
        assert gi > 0
        self.dtime = self.cfl * self.dgrid[gi - 1] / cmax
        self.time = self.time + 2 * self.dtime

    def flux(self, order):
        P = [[3, 2, 1], [3, 1, 2], [2, 3, 1], [2, 1, 3], [1, 2, 3], [1, 3, 2]]  # R2010 perms
        row = (self.iteration - 1) % 6
        if order > 0:
            dv = P[row]
            mi = {1: [2, 3], 2: [1, 3], 3: [1, 2]}
        else:
            dv = [4 - v for v in P[row]]
            mi = {1: [3, 2], 2: [3, 1], 3: [2, 1]}
        for X in dv:
            if self.grid[X - 1] < 3:
                continue
            if order > 0:
                self.relaxing_fluid(X)
                if self.magnet_active:
                    self.magnet_flux(X, mi[X])
            else:
                if self.magnet_active:
                    self.magnet_flux(X, mi[X])
                self.relaxing_fluid(X)

    def set_gravity(self, potential, threshold):
        """GravityArray: zeros first (its edge stores), then the fixed potential."""
        self.grav = Obj(self, "init", 0)
        self.grav.assign(np.zeros(self.shape))
        self.grav.assign(potential.copy())
        self.threshold = threshold

    def set_radiation(self, exponent, initial_maximum):
        """Radiation.initialize with the constructor's strength of 1."""
        rho = self.prim[0].a
        mom = [o.a for o in self.prim[1:4]]
        pg, _ = self.pressure("gas", rho, mom, self.prim[4].a)
        unscaled = 1 * rho ** (2 - exponent) * pg ** exponent
        ke = 0.5 * (mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2]) / rho
        me = 0.5 * (self.mag[0].a * self.mag[0].a + self.mag[1].a * self.mag[1].a
                    + self.mag[2].a * self.mag[2].a)
        self.rad = (exponent, (initial_maximum / 100) * np.nanmin((self.prim[4].a - ke - me) / unscaled))

    def source(self):
        """source.m: gravity from the fixed potential, then radiation."""
        dt = self.dtime
        rho = self.prim[0].a
        if self.grav is not None:
            thr = rho * (rho > self.threshold) if self.threshold > 0 else rho
            es = np.zeros(self.shape)
            for i in range(3):
                D1 = i + 1
                d5 = (-self.grav.shift(D1, 2) + 8 * self.grav.shift(D1, 1)
                      - 8 * self.grav.shift(D1, -1) + self.grav.shift(D1, -2)) / (12 * self.dgrid[i])
                ms = dt * thr * d5
                es = es + ms * self.prim[i + 1].a / rho
                self.prim[i + 1].a = self.prim[i + 1].a - ms
            self.prim[4].a = self.prim[4].a - es
        if self.rad is not None:
            a, strength = self.rad
            pg, _ = self.pressure("gas", rho, [o.a for o in self.prim[1:4]], self.prim[4].a)
            self.prim[4].a = self.prim[4].a - dt * (strength * rho ** (2 - a) * pg ** a)
        else:
            self.prim[4].a = self.prim[4].a - dt * 0

    def iterate(self):
        self.time_update()
        self.flux(1)
        self.source()
        self.flux(-1)
        self.source()
        self.iteration += 1


# ------------------------------------------------------------------------------------
# fixtures
# ------------------------------------------------------------------------------------
def to_ref(a, shape):
    """(nz, ny, nx) C-order -> reference dims (nx, ny[, nz])."""
    return np.ascontiguousarray(np.transpose(a, (2, 1, 0))).reshape(shape)


def from_ref(a):
    if a.ndim == 2:
        a = a[:, :, None]
    return np.ascontiguousarray(np.transpose(a, (2, 1, 0)))


CASES = [
    # name, grid, dgrid, bc modes [[lo x,y,z],[hi x,y,z]], limiters, magnet, iterations.
    # dGrid(3) is physical even for 2-D grids: TimeManager takes MINDGRID(gridIndex) of
    # the fastest direction, which can be z (TimeManager.m:90).
    ("hydro1d_trans_vanleer", (16, 1, 1), (1 / 16, 1 / 16, 1 / 16), [["trans"] * 3, ["trans"] * 3],
     ["vanleer"] * 3, False, 6),
    ("hydro2d_mixed_superbee", (12, 10, 1), (0.1, 0.12, 0.1),
     [["trans", "mirror", "circ"], ["const", "flip", "circ"]], ["superbee", "minmod", "vanleer"],
     False, 4),
    ("mhd2d_circ_vanleer", (10, 9, 1), (0.1, 0.1, 0.1), [["circ"] * 3, ["circ"] * 3],
     ["vanleer"] * 3, True, 4),
    ("mhd2d_trans_minmod", (11, 10, 1), (0.1, 0.09, 0.09), [["trans"] * 3, ["trans"] * 3],
     ["minmod"] * 3, True, 3),
    ("mhd2d_zero_wall_superbee", (10, 12, 1), (0.1, 0.1, 0.1),
     [["zero", "wall", "circ"], ["wall", "zero", "circ"]], ["superbee"] * 3, True, 3),
    ("mhd2d_fade", (22, 21, 1), (0.05, 0.05, 0.05), [["fade"] * 3, ["fade"] * 3], ["vanleer"] * 3,
     True, 2),
    ("hydro3d_trans", (8, 7, 6), (0.1, 0.1, 0.1), [["trans"] * 3, ["trans"] * 3], ["minmod"] * 3,
     False, 3),
    ("mhd3d_trans_mirror", (7, 6, 6), (0.1, 0.1, 0.1), [["trans", "mirror", "trans"],
                                                        ["mirror", "trans", "const"]],
     ["vanleer"] * 3, True, 2),
]


# source terms and viscosity: name, grid, dgrid, bc, limiters, magnet, iterations, extras
# (gravity: potential of the cell centres with a mass threshold; radiation: exponent,
# initialMaximum; viscosity: "nr" / "csw" with linear, quadratic strengths)
EXTRA_CASES = [
    ("hydro2d_grav_rad_nr", (12, 10, 1), (1 / 12, 1 / 10, 1 / 12),
     [["trans", "wall", "circ"], ["const", "trans", "circ"]], ["vanleer", "superbee", "vanleer"],
     False, 4, {"grav": 0.9, "rad": (0.5, 1.0), "visc": ("nr", 0.3, 0.6)}),
    ("mhd2d_csw_trans_mirror", (11, 10, 1), (0.09, 0.1, 0.09),
     [["trans", "mirror", "circ"], ["wall", "trans", "circ"]], ["minmod"] * 3,
     True, 3, {"grav": 0.0, "visc": ("csw", 0.3, 0.8)}),
    ("hydro2d_two_cell_nr_grav_fade", (24, 2, 1), (1 / 24, 1 / 24, 1 / 24),
     [["fade", "circ", "circ"], ["trans", "circ", "circ"]], ["vanleer"] * 3,
     False, 5, {"grav": 0.0, "visc": ("nr", 0.1, 0.25), "bc_infinity": 8}),
    ("hydro3d_csw_grav", (8, 7, 6), (0.1, 0.1, 0.1),
     [["trans", "circ", "mirror"], ["wall", "circ", "const"]], ["minmod"] * 3,
     False, 2, {"grav": 0.0, "rad": (0.5, 2.0), "visc": ("csw", 0.2, 0.5)}),
]


def potential(grid):
    """The fixed potential of the extra cases, (nz, ny, nx), from the cell centres."""
    nx, ny, nz = grid
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny,
                          (np.arange(nx) + 0.5) / nx, indexing="ij")
    return 0.8 * np.sin(2 * np.pi * x) * np.cos(np.pi * y) + 0.3 * (x - 0.5) ** 2 + 0.1 * z


def random_state(grid, seed, magnet):
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid
    shape = (nz, ny, nx)
    mass = 1.0 + 0.3 * rng.uniform(-1, 1, shape)
    vel = [0.3 * rng.uniform(-1, 1, shape) for _ in range(3)]
    mag = [(0.5 + 0.3 * rng.uniform(-1, 1, shape)) if magnet else np.zeros(shape) for _ in range(3)]
    P = 1.0 + 0.3 * rng.uniform(-1, 1, shape)
    g = 5.0 / 3.0
    mom = [mass * v for v in vel]
    ener = P / (g - 1) + 0.5 * mass * sum(v * v for v in vel) + 0.5 * sum(b * b for b in mag)
    return mass, mom, ener, mag


def main():
    out = {}
    cases = [c + ({},) for c in CASES] + EXTRA_CASES
    for k, (name, grid, dgrid, bc, lims, magnet, iters, extra) in enumerate(cases):
        mass, mom, ener, mag = random_state(grid, 100 + k, magnet)
        sim = Sim(grid, dgrid, 5.0 / 3.0, bc, lims, magnet, 0.35 if magnet else 0.7,
                  bc_infinity=extra.get("bc_infinity", 20))
        shp = sim.shape
        sim.set_state(to_ref(mass, shp), [to_ref(m, shp) for m in mom], to_ref(ener, shp),
                      [to_ref(b, shp) for b in mag])
        keep = {}
        if "grav" in extra:
            phi = potential(grid)
            sim.set_gravity(to_ref(phi, shp), extra["grav"])
            keep.update(potential=phi, threshold=np.array(extra["grav"]))
        if "rad" in extra:
            sim.set_radiation(*extra["rad"])
            keep.update(rad=np.array(extra["rad"], dtype=float))
        if "visc" in extra:
            sim.visc = extra["visc"]
            keep.update(visc_kind=np.array(extra["visc"][0]),
                        visc=np.array(extra["visc"][1:], dtype=float))
        if "bc_infinity" in extra:
            keep.update(bc_infinity=np.array(extra["bc_infinity"]))
        states = []
        times = []
        for _ in range(iters):
            sim.iterate()
            st = [from_ref(o.a) for o in sim.prim] + [from_ref(o.a) for o in sim.mag] + \
                 [from_ref(o.a) for o in sim.cmag]
            states.append(np.stack(st))
            times.append((sim.time, sim.dtime))
        np.savez_compressed(os.path.join(HERE, "golden_%s.npz" % name),
                            grid=np.array(grid), dgrid=np.array(dgrid, dtype=float),
                            bc=np.array(bc), limiter=np.array(lims), magnet=np.array(magnet),
                            mass=mass, mom=np.stack(mom), ener=ener, mag=np.stack(mag),
                            states=np.stack(states), times=np.array(times), **keep)
        print("wrote", name, flush=True)
    # pchip and fade vectors
    x = np.array([1.0, 2.0, 3.0, 20.0, 21.0])
    rng = np.random.default_rng(7)
    ys = rng.normal(size=(64, 5))
    ys[:, 4] = ys[:, 3]
    xi = np.array([0.5, 1.0, 1.5, 2.5, 4.0, 5.0, 10.0, 19.5, 20.5, 21.0])
    vals = np.stack([PchipInterpolator(x, y)(xi) for y in ys])
    fade = {}
    for f in (4, 5, 8, 12, 20, 25, 32):
        xx = np.array([0, np.ceil(f / 4), np.floor(f / 2 + 0.5), f - 1, f], dtype=float)
        fade["fade_%d" % f] = PchipInterpolator(xx, [1, 1, 0.01, 0, 0])(np.arange(1, f + 1))
    np.savez_compressed(os.path.join(HERE, "golden_pchip.npz"), x=x, ys=ys, xi=xi, vals=vals,
                        **fade)
    print("wrote pchip/fade vectors")


if __name__ == "__main__":
    sys.exit(main())
