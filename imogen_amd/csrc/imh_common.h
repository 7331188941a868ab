// imh_common.h -- shared definitions of the gfx950 relaxing-TVD implementation:
// object numbering for the boundary-condition edge trackers, per-simulation device
// state, and the scalar semantics (MATLAB sign/max/min, pchip, BC ghost cells, pressure,
// limiters) used by every kernel.  All functions are __host__ __device__ so the host
// set-up code and the kernels evaluate them identically.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/imogen_hip.h"

namespace imh {

// ------------------------------------------------------------------------------------
// Reference array objects that carry edge state (shifting/Edges.m).  Each object of the
// reference (FluidArray, its StorageArray/FluxArray children, MagnetArray and the
// magnet's per-(I,X) velGrid/store.fluxR/wMag/flux arrays) keeps its own stored edges,
// nudged on every transparent shift (Edges.m:115-140).
// ------------------------------------------------------------------------------------
enum ObjId : int {
    OBJ_SFR = 0,     // v(i).store.fluxR, i = 0..4 (mass, mom1..3, ener)
    OBJ_SFL = 5,     // v(i).store.fluxL
    OBJ_FR = 10,     // v(i).fluxR
    OBJ_FL = 15,     // v(i).fluxL
    OBJ_PRIM = 20,   // mass, mom(1..3), ener (FluidArray)
    OBJ_MAG = 25,    // mag(k) (MagnetArray)
    OBJ_VG = 28,     // mag(I).velGrid(X):   + 3*I + X (0-based I, X)
    OBJ_MSFR = 37,   // mag(I).store(X).fluxR
    OBJ_WMAG = 46,   // mag(I).wMag(X)
    OBJ_MFLUX = 55,  // mag(I).flux(X)
    OBJ_GRAV = 64,   // grav (GravityArray)
    OBJ_FSTORE = 65, // mass, mom(1..3), ener .store (StorageArray), shifted by the viscosity
    OBJ_CMAG = 70,   // mag(1..3).cellMag (InitializedArray), shifted by the CSW viscosity
    NOBJ = 73
};

// FluxArray.fluxType for zero_shift.m.
enum FluxType : int { FT_OTHER = 0, FT_FLUXL = 1, FT_FLUXR = 2, FT_FLUX = 3 };

// Per-simulation scalars that live on the device (TimeManager state and the CFL slots).
struct SimDev {
    double dt;                 // run.time.dTime
    double time;               // run.time.time
    unsigned long long cmax_bits[3];  // max(|p_i/rho| + c_s) as ordered bits (>= 0)
    int iteration;             // run.time.iteration
    int grid_index;            // last CFL direction
    int error;                 // IMH_ECFL when no speed exceeds 1e-2
    int active;                // TimeManager.running at the start of this iteration
    int pad[4];
};

// Static per-simulation parameters used by the device-side time update.
struct SimParams {
    double cfl;
    double dgrid[3];
    double time_max;
    int iter_max;
    int update_mode;
    int unconditional;         // 1: run the loop body even if not running (imh_iteration)
    int pad;
};

// ------------------------------------------------------------------------------------
// MATLAB scalar semantics
// ------------------------------------------------------------------------------------
template <typename T>
__host__ __device__ inline T msign(T x) {
    return x > T(0) ? T(1) : (x < T(0) ? T(-1) : (x == T(0) ? T(0) : x));
}

// Edges.getEdge transparent branch: move the stored edge toward the current edge cell by
// at most 0.005 (MATLAB min ignores NaN; signTest - ~signTest is +1 or -1).
__host__ __device__ inline double edge_nudge(double& stored, double newedge) {
    double old = stored;
    double diff = newedge - old;
    double delta = fmin(fabs(diff), 0.005);
    double sgn = (diff > 0.0) ? 1.0 : -1.0;
    double r = old + sgn * delta;
    stored = r;
    return r;
}

// pchip through 5 nodes evaluated at one point: Fritsch-Butland interior slopes,
// shape-preserving three-point end slopes, Hermite cubic in Horner form (the published
// pchip/pwch/ppval algorithm; Moler, Numerical Computing with MATLAB, sec. 3.4).
// Written with scalars only (no dynamically indexed arrays, so no scratch memory on the
// GPU); the operations and their order are those of the array form.
__host__ __device__ inline double pchip_end(double h1, double h2, double del1, double del2) {
    double d = ((2.0 * h1 + h2) * del1 - h1 * del2) / (h1 + h2);
    if (msign(d) != msign(del1)) {
        d = 0.0;
    } else if ((msign(del1) != msign(del2)) && (fabs(d) > fabs(3.0 * del1))) {
        d = 3.0 * del1;
    }
    return d;
}

__host__ __device__ inline double pchip_interior(double hl, double hr, double dell, double delr) {
    double d = 0.0;
    if (msign(dell) * msign(delr) > 0.0) {
        double w1 = 2.0 * hr + hl;
        double w2 = hr + 2.0 * hl;
        d = (w1 + w2) / (w1 / dell + w2 / delr);
    }
    return d;
}

__host__ __device__ inline double pchip5s(double x0, double x1, double x2, double x3, double x4,
                                          double y0, double y1, double y2, double y3, double y4,
                                          double xi) {
    const double h0 = x1 - x0, h1 = x2 - x1, h2 = x3 - x2, h3 = x4 - x3;
    const double e0 = (y1 - y0) / h0, e1 = (y2 - y1) / h1, e2 = (y3 - y2) / h2,
                 e3 = (y4 - y3) / h3;
    const double d1 = pchip_interior(h0, h1, e0, e1);
    const double d2 = pchip_interior(h1, h2, e1, e2);
    const double d3 = pchip_interior(h2, h3, e2, e3);
    const int j = (xi >= x3) ? 3 : ((xi >= x2) ? 2 : ((xi >= x1) ? 1 : 0));
    double hj, delj, dj, dj1, yj, xj;
    if (j == 0) {
        hj = h0; delj = e0; dj = pchip_end(h0, h1, e0, e1); dj1 = d1; yj = y0; xj = x0;
    } else if (j == 1) {
        hj = h1; delj = e1; dj = d1; dj1 = d2; yj = y1; xj = x1;
    } else if (j == 2) {
        hj = h2; delj = e2; dj = d2; dj1 = d3; yj = y2; xj = x2;
    } else {
        hj = h3; delj = e3; dj = d3; dj1 = pchip_end(h3, h2, e3, e2); yj = y3; xj = x3;
    }
    double dzzdx = (delj - dj) / hj;
    double dzdxdx = (dj1 - delj) / hj;
    double c1 = (dzdxdx - dzzdx) / hj;
    double c2 = 2.0 * dzzdx - dzdxdx;
    double s = xi - xj;
    double v = c1;
    v = s * v + c2;
    v = s * v + dj;
    v = s * v + yj;
    return v;
}

__host__ __device__ inline double pchip5(const double x[5], const double y[5], double xi) {
    return pchip5s(x[0], x[1], x[2], x[3], x[4], y[0], y[1], y[2], y[3], y[4], xi);
}

// fade_shift.m:27 blending coefficients for fadeSize = f.
inline void fade_coefficients(int f, double* c) {
    double x[5] = {0.0, ceil(f / 4.0), floor(f / 2.0 + 0.5), double(f - 1), double(f)};
    double y[5] = {1.0, 1.0, 0.01, 0.0, 0.0};
    for (int i = 1; i <= f; ++i) c[i - 1] = pchip5(x, y, double(i));
}

// ------------------------------------------------------------------------------------
// Boundary ghosts for a shift by m cells (1 or 2) of one pencil of length N >= 1.
//
// Ends holds a[min(k,N-1)] (a0..a2) and a[max(N-3+k,0)] (b0..b2) for k = 0..2.  The
// results g0 (and g1 for m = 2) are the values the shifted array takes at positions
// t = 0..m-1 (lower side, shift(D,-m)) or N-m+t (upper side, shift(D,+m)): the ghost
// cells t-m or N+t.  Fade mode returns the constant-shift value; the caller applies the
// fade blend.
// ------------------------------------------------------------------------------------
struct GhostSpec {
    int mode;          // IMH_BC_*
    int upper;         // 0 lower (n < 0), 1 upper (n > 0)
    int flip_sign;     // component == DIRECT (mirror / flip)
    int zero_ghost;    // zero mode on a FluxArray whose type does not keep this side
    int active_linear; // Edges.ACTIVE(DIRECT) for transparent
    int bc_infinity;
};

struct Ends {
    double a0, a1, a2, b0, b1, b2;
};

__host__ __device__ inline double pick3(int k, double v0, double v1, double v2) {
    return k <= 0 ? v0 : (k == 1 ? v1 : v2);
}

// a[i] for i in [0,2] or [N-3,N-1]
__host__ __device__ inline double end_at(const Ends& e, int i, int N) {
    if (i <= 2 && i < N) return pick3(i < N - 1 ? i : N - 1, e.a0, e.a1, e.a2);
    int k = i - (N - 3);
    k = k < 0 ? 0 : (k > 2 ? 2 : k);
    return pick3(k, e.b0, e.b1, e.b2);
}

// One ghost value at position t (0 <= t < m).  `tracker` is nudged by the caller once
// per shift call (see ghost_values); here ev is the nudged edge (or 0).
// M > 0 fixes the shift length at compile time: a separate instantiation, so callers of
// the run-time form (all with m = 1 in a translation unit) keep their code.
template <int M = 0>
__host__ __device__ inline double ghost_at(const GhostSpec& s, const Ends& e, int N, int m_, int t,
                                           double stored, double ev) {
    const int m = M > 0 ? M : m_;
    const int up = s.upper;
    switch (s.mode) {
    case IMH_BC_CIRC: {
        int src = up ? (t % N) : ((N - m + t) % N + N) % N;
        return end_at(e, src, N);
    }
    case IMH_BC_MIRROR: {
        // lower: position t takes a[m-t]; upper: position N-m+t takes a[N-2-t]
        int src = up ? (N - 2 - t) : (m - t);
        double v = end_at(e, src, N);
        return s.flip_sign ? -v : v;
    }
    case IMH_BC_WALL:
        return stored;
    case IMH_BC_TRANS: {
        double in0, in1, in2;
        for (int k = 0; k < 3; ++k) {
            int src;
            if (up) {
                int j1 = N - (2 - k) - m;  // 1-based index into the shifted array
                if (j1 < 1) j1 = 1;
                src = j1 - 1 + m;          // shifted[j] = a[min(j+m, N-1)]
                if (src > N - 1) src = N - 1;
            } else {
                int j1 = 3 - k + m;
                if (j1 > N) j1 = N;
                src = j1 - 1 - m;          // shifted[j] = a[max(j-m, 0)]
                if (src < 0) src = 0;
            }
            double v = end_at(e, src, N);
            if (k == 0) in0 = v;
            else if (k == 1) in1 = v;
            else in2 = v;
        }
        const double B = double(s.bc_infinity);
        double xi = up ? 3.0 + double(t + 1) : 3.0 + double(m - t);
        int pos = up ? N - m + t : t;
        double outv = pchip5s(1.0, 2.0, 3.0, B, B + 1.0, in0, in1, in2, ev, ev, xi);
        double orig = end_at(e, pos, N);
        double test = (fabs(orig - outv) < 1e-4) ? 1.0 : 0.0;
        return (1.0 - test) * outv + test * orig;
    }
    default: {  // const, fade (before blending), flip, zero
        double v = up ? end_at(e, N - 1, N) : end_at(e, 0, N);
        if (s.mode == IMH_BC_FLIP && s.flip_sign) v = -v;
        if (s.mode == IMH_BC_ZERO && s.zero_ghost) v = 0.0;
        return v;
    }
    }
}

// The m ghost values of one shift call; nudges the transparent tracker once
// (Edges.getEdge is called once per transparent_shift call).
template <int M = 0>
__host__ __device__ inline void ghost_values(const GhostSpec& s, const Ends& e, int N, int m_, double* tracker, double& g0,
                  double& g1) {
    const int m = M > 0 ? M : m_;
    double ev = 0.0;
    if (s.mode == IMH_BC_TRANS && s.active_linear)
        ev = edge_nudge(*tracker, s.upper ? end_at(e, N - 1, N) : end_at(e, 0, N));
    double stored = (s.mode == IMH_BC_WALL) ? *tracker : 0.0;
    g0 = ghost_at<M>(s, e, N, m, 0, stored, ev);
    g1 = (m > 1) ? ghost_at<M>(s, e, N, m, 1, stored, ev) : 0.0;
}

// Fade blend of a constant-shifted value r at position i (fade_shift.m:31-45).
// v1 = shifted[0], vN = shifted[N-1] before the blends.
template <typename T>
__host__ __device__ inline T fade_blend(T r, int i, int N, T v1, T vN, const double* c, int f) {
    // vN after the first blend when N-1 falls inside the first fade window
    T vNb = vN;
    if (N - 1 < f) vNb = T(1.0 - c[N - 1]) * vN + T(c[N - 1]) * v1;
    T out = r;
    if (i < f) out = T(1.0 - c[i]) * out + T(c[i]) * v1;
    if (i >= N - f) {
        double cf = c[f - 1 - (i - (N - f))];
        out = T(1.0 - cf) * out + T(cf) * vNb;
    }
    return out;
}

// ------------------------------------------------------------------------------------
// pressure.m ('totsnd' and 'sound'), evaluated in FP64 in both precisions.
// ------------------------------------------------------------------------------------
struct MagUse {
    int x, y, z;  // !MagnetArray.isZero per component
};

__host__ __device__ inline void pressure_totsnd(double rho, double px, double py, double pz,
                                                double E, double bx, double by, double bz,
                                                MagUse u, double g, double& P, double& c) {
    double v2 = (px * px + py * py + pz * pz) / (rho * rho);
    double b2 = 0.0;
    bool any = false;
    if (u.x) { b2 = 0.0 + bx * bx; any = true; }
    if (u.y) { b2 = any ? b2 + by * by : 0.0 + by * by; any = true; }
    if (u.z) { b2 = any ? b2 + bz * bz : 0.0 + bz * bz; any = true; }
    double r = (g - 1.0) * (E - 0.5 * rho * v2);
    c = sqrt(fabs((g * (r - (g - 1.0) * 0.5 * b2) + 2.0 * b2) / rho));
    r = r + 0.5 * (2.0 - g) * b2;
    if (r < 0.0) r = 0.0;
    P = r;
}

__host__ __device__ inline double pressure_sound(double rho, double px, double py, double pz,
                                                 double E, double bx, double by, double bz,
                                                 MagUse u, double g) {
    double v2 = (px * px + py * py + pz * pz) / (rho * rho);
    double b2 = 0.0;
    bool any = false;
    if (u.x) { b2 = 0.0 + bx * bx; any = true; }
    if (u.y) { b2 = any ? b2 + by * by : 0.0 + by * by; any = true; }
    if (u.z) { b2 = any ? b2 + bz * bz : 0.0 + bz * bz; any = true; }
    double r = (g - 1.0) * (E - 0.5 * rho * v2);
    r = r - (g - 1.0) * 0.5 * b2;
    r = sqrt(fabs((g * r + 2.0 * b2) / rho));
    if (r < 0.0) r = 0.0;
    return r;
}

// pressure.m 'gas': thermal pressure, floored at 0.
__host__ __device__ inline double pressure_gas(double rho, double px, double py, double pz,
                                               double E, double bx, double by, double bz,
                                               MagUse u, double g) {
    double v2 = (px * px + py * py + pz * pz) / (rho * rho);
    double b2 = 0.0;
    bool any = false;
    if (u.x) { b2 = 0.0 + bx * bx; any = true; }
    if (u.y) { b2 = any ? b2 + by * by : 0.0 + by * by; any = true; }
    if (u.z) { b2 = any ? b2 + bz * bz : 0.0 + bz * bz; any = true; }
    double r = (g - 1.0) * (E - 0.5 * rho * v2);
    r = r - (g - 1.0) * 0.5 * b2;
    if (r < 0.0) r = 0.0;
    return r;
}

// ------------------------------------------------------------------------------------
// Limiters (flux/vanleerLimiter.m, superbeeLimiter.m, minmodLimiter.m): the increment
// added to the flux.
// ------------------------------------------------------------------------------------
template <typename T>
__host__ __device__ inline T lim_minmod(T dl, T dr) {
    return T(0.125) * (msign(dl) + msign(dr)) * fmin(fabs(dl), fabs(dr));
}

template <typename T>
__host__ __device__ inline T lim_delta(int lim, T dl, T dr) {
    if (lim == IMH_LIM_VANLEER) {
        T t = fmax(dl * dr, T(0)) / (dl + dr);
        return isnan(t) ? T(0) : t;
    }
    if (lim == IMH_LIM_SUPERBEE) {
        T cf = (fabs(dl) < fabs(dr)) ? T(1) : T(0);
        return lim_minmod((T(1) + cf) * dl, (T(2) - cf) * dr);
    }
    return lim_minmod(dl, dr);
}

// Pencil geometry: pencil f along dim d -> base offset and stride, with the face index
// conventions d=0: f = j + ny*k, d=1: f = i + nx*k, d=2: f = i + nx*j.
__host__ __device__ inline void pencil_base(const int n[3], int d, long long f, long long& base,
                                            long long& stride) {
    // a simulation holds fewer than 2^31 cells: 32-bit divisions (64-bit ones are long
    // instruction sequences on the GPU)
    const unsigned nx = unsigned(n[0]), ny = unsigned(n[1]), uf = unsigned(f);
    if (d == 0) {
        const unsigned j = uf % ny, k = uf / ny;
        base = (long long)nx * (j + (long long)ny * k);
        stride = 1;
    } else if (d == 1) {
        const unsigned i = uf % nx, k = uf / nx;
        base = i + (long long)nx * ny * k;
        stride = nx;
    } else {
        const unsigned i = uf % nx, j = uf / nx;
        base = i + (long long)nx * j;
        stride = (long long)nx * ny;
    }
}

// Coordinates (x, y, z) of the first cell of pencil f along d.
__host__ __device__ inline void pencil_coords(const int n[3], int d, long long f, int c[3]) {
    const unsigned nx = unsigned(n[0]), ny = unsigned(n[1]), uf = unsigned(f);
    if (d == 0) {
        c[0] = 0; c[1] = int(uf % ny); c[2] = int(uf / ny);
    } else if (d == 1) {
        c[0] = int(uf % nx); c[1] = 0; c[2] = int(uf / nx);
    } else {
        c[0] = int(uf % nx); c[1] = int(uf / nx); c[2] = 0;
    }
}

// Pointers into device memory taken from an argument block are generic (flat) to the
// compiler; flat loads and stores count against the LDS wait counter too, so every LDS
// wait would also wait for them.  gmem() makes them global-memory accesses (the host
// pass of the compiler sees plain pointers).
#if defined(__HIP_DEVICE_COMPILE__)
#define IMH_GLOBAL_AS __attribute__((address_space(1)))
#else
#define IMH_GLOBAL_AS
#endif
template <typename T>
__device__ __forceinline__ const IMH_GLOBAL_AS T* gmem(const T* p) {
    return (const IMH_GLOBAL_AS T*)p;
}
template <typename T>
__device__ __forceinline__ IMH_GLOBAL_AS T* gmem(T* p) {
    return (IMH_GLOBAL_AS T*)p;
}

// Face index of cell (i,j,k) for dim d (same conventions).
// Pencil slots of a sweep along d whose lower boundary is a wormhole (wormhole_shift.m):
// the lower ghosts of the pencil at next-dimension coordinate c (next = x->y, y->z, z->x)
// come from the pencil at N_next-1-c with the same previous-dimension coordinate, so the
// two share a tile.  Slots 2P and 2P+1 hold such a pair; the middle pencil of an odd
// N_next pairs with itself and its odd slot is padding.  Pairs run along the x-fastest
// coordinate of the face, so neighbouring pairs are neighbours in memory.
__host__ __device__ inline long long worm_slots(const int n[3], int d) {
    const int nq = n[(d + 1) % 3], np = n[(d + 2) % 3];
    return 2LL * ((nq + 1) / 2) * np;
}
// Face index (face_index order) of slot s, or -1 for a padding slot.
__host__ __device__ inline long long worm_face(const int n[3], int d, long long s) {
    const int nq = n[(d + 1) % 3], np = n[(d + 2) % 3];
    const unsigned h = unsigned(nq + 1) / 2;
    const unsigned P = unsigned(s >> 1);
    unsigned cq, cp;
    if (d != 1) {  // the next dimension is the face's fast one (d = 0: y, d = 2: x)
        cq = P % h;
        cp = P / h;
    } else {       // d = 1: the next dimension (z) is the slow one, x the fast one
        cp = P % unsigned(np);
        cq = P / unsigned(np);
    }
    if (cp >= unsigned(np) || cq >= h) return -1;
    if (s & 1) {
        const unsigned m = unsigned(nq) - 1 - cq;
        if (m == cq) return -1;
        cq = m;
    }
    return d != 1 ? (long long)cq + (long long)nq * cp : (long long)cp + (long long)np * cq;
}

// Face index of the pencil whose lower wormhole ghosts face f takes: the next-dimension
// coordinate mirrored, the previous one kept.
__host__ __device__ inline long long worm_partner_face(const int n[3], int d, long long f) {
    const long long nq = n[(d + 1) % 3];
    if (d != 1) {
        const long long cq = f % nq, cp = f / nq;
        return (nq - 1 - cq) + nq * cp;
    }
    const long long np = n[0];
    const long long cp = f % np, cq = f / np;
    return cp + np * (nq - 1 - cq);
}

__host__ __device__ inline long long face_index(const int n[3], int d, int i, int j, int k) {
    if (d == 0) return j + (long long)n[1] * k;
    if (d == 1) return i + (long long)n[0] * k;
    return i + (long long)n[0] * j;
}

}  // namespace imh
