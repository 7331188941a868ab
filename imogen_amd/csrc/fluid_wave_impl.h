// fluid_wave_impl.h -- the fluid sweep with each pencil spread over the lanes of waves
// (relaxingFluid.m:16-73, wFluidFlux.m:14-44, pressure.m), four cells per lane.
//
// The general sweep (fluid_sweep_impl.h) exchanges every chunk-boundary value through LDS
// behind a workgroup barrier, several times per field.  Here neighbouring chunks trade
// values with cross-lane shuffles: a pencil of up to 256 cells along x lives in one wave
// and its sweep needs no workgroup barrier at all; a pencil spread over several waves (x
// pencils longer than 256 cells, and every y / z tile, whose PX x-adjacent pencils are
// interleaved across the lanes so that loads are coalesced) takes two barriers per tile,
// one per freezing-speed reduction, each also carrying the two-cell halos of the lanes at
// a wave edge.  The halos are published as the flux numerators (q and w*C) before the
// freezing speed is known, so one barrier serves both.
//
// Boundary ghosts: the wave holding a pencil end evaluates the ghosts of all five fields
// of that end in parallel (one lane per field and flux object, wave-local LDS staging),
// in three passes per tile -- predictor (store.fluxR lower / store.fluxL upper),
// corrector (fluxR and fluxL, both sides), final (fluxR lower / fluxL upper on the
// limited fluxes) -- so each object's edge tracker is nudged once per shift call in the
// .m text's order.  With four cells per lane the pencil-end lanes hold the three end
// cells every ghost rule reads.
//
// Arithmetic: FAST = false keeps every division and the .m evaluation order (built
// without contraction it reproduces the CPU restatement bit for bit); FAST = true (FP64
// plain path) divides by one reciprocal of rho per cell state and of the freezing speed
// per pencil (hardware estimate + two Newton steps), writes its multiply-adds as fma and
// uses branch-light limiters, so a step agrees with the restatement to a few units in
// the last place.  Both FP64 forms are built without automatic contraction.
//
// Applies when every pencil length N along the sweep is 4 * TL with TL a power of two,
// 2 <= TL <= 512, both ends circular or neither, and the sweep is in its lean form (no
// fade boundary, no first assignment of the flux objects, no artificial viscosity);
// other sweeps take the general kernel.
#pragma once

#include "fluid_math.h"
#include "imh_common.h"
#include "imh_kernels.h"

#include <type_traits>

namespace imh {
namespace wv {

constexpr int K = 4;       // cells per lane
constexpr int PXW = WAVE_PXMAX;

template <typename T, bool FAST>
struct Arith : FMath<T> {};

// 1/y from the hardware estimate and two Newton steps (relative error near the last
// place; y is a density or a freezing speed, finite and positive).
__device__ __forceinline__ double rcp_nr(double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = fma(-y, r, 1.0);
    r = fma(r, e, r);
    e = fma(-y, r, 1.0);
    return fma(r, e, r);
}

template <>
struct Arith<double, true> {
    static __device__ __forceinline__ double div(double x, double, double r) { return x * r; }
    static __device__ __forceinline__ double rcp(double y) { return rcp_nr(y); }
    static __device__ __forceinline__ void press(double rho, double px, double py, double pz,
                                                 double E, double bx, double by, double bz,
                                                 MagUse u, double g, double& P, double& c,
                                                 double& inv) {
        inv = rcp_nr(rho);
        const double m2 = fma(px, px, fma(py, py, pz * pz));
        double b2 = 0.0;
        if (u.x) b2 = fma(bx, bx, b2);
        if (u.y) b2 = fma(by, by, b2);
        if (u.z) b2 = fma(bz, bz, b2);
        double r = (g - 1.0) * fma(-0.5 * m2, inv, E);   // (g-1)(E - rho v^2 / 2)
        c = sqrt(fabs(fma(g, fma(-(g - 1.0) * 0.5, b2, r), 2.0 * b2) * inv));
        r = fma(0.5 * (2.0 - g), b2, r);
        P = r < 0.0 ? 0.0 : r;
    }
    static __device__ __forceinline__ double divrho(double x, double inv) { return x * inv; }
    static __device__ __forceinline__ double vanleer(double dl, double dr) {
        const double p = dl * dr;
        return p > 0.0 ? p * rcp_nr(dl + dr) : 0.0;
    }
};

// Limiters (flux/*Limiter.m).  The exact forms are those of limit() (MATLAB sign, min
// and 0/0 handling); the fast forms give the same values for finite inputs: minmod is
// +-min(|a|,|b|)/4 when a and b share a strict sign, else 0.
template <typename T, bool FAST>
__device__ __forceinline__ T wlimit(int lim, T dl, T dr) {
    if constexpr (!FAST) {
        return limit(lim, dl, dr);
    } else {
        auto mm = [](T a, T b) __attribute__((always_inline)) -> T {
            const T m = T(0.25) * fmin(fabs(a), fabs(b));
            return a * b > T(0) ? copysign(m, a) : T(0);
        };
        if (lim == IMH_LIM_VANLEER) return Arith<T, true>::vanleer(dl, dr);
        if (lim == IMH_LIM_SUPERBEE) {
            const bool c = fabs(dl) < fabs(dr);
            return mm(c ? T(2) * dl : dl, c ? dr : T(2) * dr);
        }
        return mm(dl, dr);
    }
}

using Lds = WaveLds;

template <int V>
using Fld = std::integral_constant<int, V>;
// fn(Fld<v>) for v = 0..4 (the field index is a constant in each body, so register
// arrays indexed by it stay in registers)
template <typename F>
__device__ __forceinline__ void for_fields(F&& fn) {
    fn(Fld<0>{});
    fn(Fld<1>{});
    fn(Fld<2>{});
    fn(Fld<3>{});
    fn(Fld<4>{});
}

// Value of the lane dl below (lane_up) or above (lane_down) in the wave; lanes without
// one get their own value.  x tiles (dl = 1) use DPP whole-wave shifts (VALU moves, no
// LDS round trip); interleaved y / z tiles use LDS permutes.
template <typename T>
__device__ __forceinline__ T dpp_shift(T v, int ctrl) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        const int x = __builtin_bit_cast(int, v);
        return __builtin_bit_cast(T, ctrl == 0x138 ? __builtin_amdgcn_update_dpp(x, x, 0x138, 0xF, 0xF, false)
                                                   : __builtin_amdgcn_update_dpp(x, x, 0x130, 0xF, 0xF, false));
    } else {
        typedef int I2 __attribute__((ext_vector_type(2)));
        I2 x = __builtin_bit_cast(I2, v), r;
        if (ctrl == 0x138) {
            r.x = __builtin_amdgcn_update_dpp(x.x, x.x, 0x138, 0xF, 0xF, false);
            r.y = __builtin_amdgcn_update_dpp(x.y, x.y, 0x138, 0xF, 0xF, false);
        } else {
            r.x = __builtin_amdgcn_update_dpp(x.x, x.x, 0x130, 0xF, 0xF, false);
            r.y = __builtin_amdgcn_update_dpp(x.y, x.y, 0x130, 0xF, 0xF, false);
        }
        return __builtin_bit_cast(T, r);
    }
}
template <bool XM, typename T>
__device__ __forceinline__ T lane_up(T v, int dl) {
    if constexpr (XM) return dpp_shift(v, 0x138);  // wave_shr:1
    else return __shfl_up(v, dl, 64);
}
template <bool XM, typename T>
__device__ __forceinline__ T lane_down(T v, int dl) {
    if constexpr (XM) return dpp_shift(v, 0x130);  // wave_shl:1
    else return __shfl_down(v, dl, 64);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ void load4(const T* p, unsigned stride, bool contig, T (&out)[K]) {
    if (contig) {
        typedef T V __attribute__((ext_vector_type(16 / sizeof(T))));
        constexpr int L = 16 / int(sizeof(T));
#pragma unroll
        for (int j = 0; j < K / L; ++j) {
            const V x = *gmem(reinterpret_cast<const V*>(p + j * L));
#pragma unroll
            for (int l = 0; l < L; ++l) out[j * L + l] = x[l];
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = gmem(p)[k * stride];
    }
}

template <typename T>
__device__ __forceinline__ void store4(T* p, unsigned stride, bool contig, const T (&in)[K]) {
    if (contig) {
        typedef T V __attribute__((ext_vector_type(16 / sizeof(T))));
        constexpr int L = 16 / int(sizeof(T));
#pragma unroll
        for (int j = 0; j < K / L; ++j) {
            V x;
#pragma unroll
            for (int l = 0; l < L; ++l) x[l] = in[j * L + l];
            *gmem(reinterpret_cast<V*>(p + j * L)) = x;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) gmem(p)[k * stride] = in[k];
    }
}

// The (pencil, side) pairs whose end lane lies in wave w: returns their count; pair i is
// (pencil p, side s).
__device__ __forceinline__ int end_pairs(int w, int nw, int PX, int TL, bool xm) {
    if (!xm) return (w == 0 || w == nw - 1) ? PX : 0;
    if (TL <= 64) return 2 * (64 / TL);
    const int W = TL / 64;
    return (w % W == 0 || w % W == W - 1) ? 1 : 0;
}
__device__ __forceinline__ void end_pair(int w, int nw, int PX, int TL, bool xm, int i, int& p,
                                         int& s) {
    if (!xm) {
        p = i;
        s = (w == 0) ? 0 : 1;
    } else if (TL <= 64) {
        p = w * (64 / TL) + (i >> 1);
        s = i & 1;
    } else {
        const int W = TL / 64;
        p = w / W;
        s = (w % W == 0) ? 0 : 1;
    }
}

template <typename T, int BS, bool MHD, bool FAST, int D>
__device__ __forceinline__ void wave_sweep_body(const FluidArgs<T>& ap, long long tile, char* lds) {
    const FluidArgs<T> a = ap;
    if (!a.force && !a.sim->active) return;
    using A = Arith<T, FAST>;
    constexpr int NW = BS / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int PX = a.PX, TL = a.TL;
    constexpr bool xm = D == 0;   // x pencils are contiguous; y / z tiles interleave pencils
    const int px = xm ? tid / TL : tid % PX;
    const int ty = xm ? tid % TL : tid / PX;
    const int dl = xm ? 1 : PX;
    const bool multi = !xm || TL > 64;
    const Lds L = wave_lds_layout(NW, PX, dl, multi, int(sizeof(T)));
    T* s_red = reinterpret_cast<T*>(lds + L.red);
    T* s_e1 = reinterpret_cast<T*>(lds + L.e1);
    T* s_e2 = reinterpret_cast<T*>(lds + L.e2);
    T* s_end1 = reinterpret_cast<T*>(lds + L.end1);
    T* s_end2 = reinterpret_cast<T*>(lds + L.end2);
    double* s_trk = reinterpret_cast<double*>(lds + L.trk);
    double* s_gin = reinterpret_cast<double*>(lds + L.gin);
    double* s_gout = reinterpret_cast<double*>(lds + L.gout);

    const long long pen0 = tile * PX;
    const bool pen_ok = pen0 + px < a.npencil;
    const long long pen = pen_ok ? pen0 + px : a.npencil - 1;
    constexpr int d = D;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int N = K * TL;
    const int i0 = ty * K;
    const bool first = ty == 0, last = ty == TL - 1;
    long long b64 = 0, s64 = 1;
    pencil_base(nn, d, pen, b64, s64);
    const unsigned stride = unsigned(s64);
    const unsigned cb = unsigned(b64) + unsigned(i0) * stride;
    // chunk neighbours: in this wave (shuffle) or across a wave edge (LDS slots)
    const bool lx = !first && lane < dl;        // left neighbour in the previous wave
    const bool rx = !last && lane >= 64 - dl;   // right neighbour in the next wave

    const double dt = a.sim->dt;
    const double ff = dt / a.dx;
    const double hf = 0.5 * ff;
    constexpr int X = d + 1;
    const double g = a.gamma;
    const int mlo = mode_at(a.modes, 0, d), mhi = mode_at(a.modes, 1, d);
    const bool circ = mlo == IMH_BC_CIRC;
    const bool trk_lo = (mlo == IMH_BC_TRANS || mlo == IMH_BC_WALL);
    const bool trk_hi = (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL);

    // ---- edge trackers of this wave's pencil ends -> LDS [p][side][role][field] ------
    // roles: 0 store.fluxR (lower) / store.fluxL (upper), 1 fluxR, 2 fluxL
    const int npair = end_pairs(wave, NW, PX, TL, xm);
    if (trk_lo || trk_hi) {
        for (int j = lane; j < npair * 15; j += 64) {
            int p, s;
            end_pair(wave, NW, PX, TL, xm, j / 15, p, s);
            const int r = (j % 15) / 5, v = j % 5;
            if (pen0 + p >= a.npencil || (s ? !trk_hi : !trk_lo)) continue;
            const int obj = r == 0 ? (s ? OBJ_SFL : OBJ_SFR) : (r == 1 ? OBJ_FR : OBJ_FL);
            s_trk[(p * 2 + s) * 15 + r * 5 + v] =
                a.edges[a.eoff[((obj + v) * 2 + s) * 3 + d] + pen0 + p];
        }
    }

    // ---- load the chunk ----------------------------------------------------------------
    T q[5][K], bm[3][K];
#pragma unroll
    for (int v = 0; v < 5; ++v) load4<T>(a.q[v] + cb, stride, xm, q[v]);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        if constexpr (MHD) {
            load4<T>(a.bc[v] + cb, stride, xm, bm[v]);
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) bm[v][k] = T(0);
        }
    }

    T P[K], U[K], IR[K];
    auto speeds = [&](const T (&Q)[5][K]) __attribute__((always_inline)) -> T {
        T lmax = T(NAN);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            T Pk, ck, ir;
            A::press(Q[0][k], Q[1][k], Q[2][k], Q[3][k], Q[4][k], bm[0][k], bm[1][k], bm[2][k],
                     a.use, g, Pk, ck, ir);
            P[k] = Pk;
            IR[k] = ir;
            U[k] = A::divrho(sel3(d, Q[1][k], Q[2][k], Q[3][k]), ir);
            lmax = fmax(lmax, fabs(U[k]) + ck);  // NaN-ignoring, like MATLAB max
        }
        return lmax;
    };
    // wFluidFlux.m numerator of w for field v at cell k (w = numer / C)
    // (hydro: the B terms of a zero field are left out; they add exact zeros)
    auto numer = [&](auto VC, const T (&Q)[5][K], int k) __attribute__((always_inline)) -> T {
        constexpr int v = decltype(VC)::value;
        if constexpr (v == 0) {
            return sel3(d, Q[1][k], Q[2][k], Q[3][k]);
        } else if constexpr (v == 4) {
            if constexpr (MHD) {
                const T bX = sel3(d, bm[0][k], bm[1][k], bm[2][k]);
                T bdotp = bm[0][k] * Q[1][k] + bm[1][k] * Q[2][k] + bm[2][k] * Q[3][k];
                return U[k] * (Q[4][k] + P[k]) - A::divrho(bX * bdotp, IR[k]);
            } else {
                return U[k] * (Q[4][k] + P[k]);
            }
        } else {
            if constexpr (MHD) {
                const T bX = sel3(d, bm[0][k], bm[1][k], bm[2][k]);
                const T dv = (v == X) ? T(1) : T(0);
                return U[k] * Q[v][k] + P[k] * dv - bX * bm[v - 1][k];
            } else {
                return v == X ? U[k] * Q[v][k] + P[k] : U[k] * Q[v][k];
            }
        }
    };
    // wave part of a pencil maximum (lanes of the same pencil in this wave)
    auto wave_max = [&](T v) __attribute__((always_inline)) -> T {
        if (xm) {
            const int span = TL < 64 ? TL : 64;
            for (int off = span >> 1; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
        } else {
            for (int off = PX; off < 64; off <<= 1) v = fmax(v, __shfl_xor(v, off, 64));
        }
        return v;
    };
    auto publish_max = [&](int stage, T v) __attribute__((always_inline)) {
        if (!multi) return;
        if (xm ? lane == 0 : lane < PX) s_red[(stage * NW + wave) * PXW + (xm ? 0 : px)] = v;
    };
    auto combine_max = [&](int stage, T v) __attribute__((always_inline)) -> T {
        if (!multi) return v;
        T m = v;
        if (xm) {
            const int w0 = (px * TL) >> 6, w1 = (px * TL + TL - 1) >> 6;
            for (int w = w0; w <= w1; ++w) m = fmax(m, s_red[(stage * NW + w) * PXW]);
        } else {
            for (int w = 0; w < NW; ++w) m = fmax(m, s_red[(stage * NW + w) * PXW + px]);
        }
        return m;
    };

    // A ghost pass: every (pencil end in this wave, field, role) job evaluates one shift
    // call's ghost from gin[p][s][v][role][0..2] (the side's three end cells) and [3] (the
    // far end cell, for circular ends) into gout[p][s][v][slot].
    auto ghost_pass = [&](int kind) __attribute__((always_inline)) {
        const int R = kind == 1 ? 2 : 1;
        wave_sync();
        for (int j = lane; j < npair * 5 * R; j += 64) {
            int p, s;
            end_pair(wave, NW, PX, TL, xm, j / (5 * R), p, s);
            const int v = (j / R) % 5, role = j % R;
            // flux object of this shift call: 0 fluxR-like, 1 fluxL-like
            const int which = kind == 1 ? role : s;
            const int trole = kind == 0 ? 0 : 1 + which;
            const int slot = kind == 0 ? 0 : (kind == 1 ? 1 + role : 3);
            const double* in = s_gin + ((((p * 2 + s) * 5 + v) * 2 + role) * 4);
            GhostSpec sp;
            sp.mode = s ? mhi : mlo;
            sp.upper = s;
            sp.flip_sign = (v == X);
            const int ft = which ? FT_FLUXL : FT_FLUXR;
            sp.zero_ghost = s ? (ft != FT_FLUXR) : (ft != FT_FLUXL);
            sp.active_linear = a.active_lin;
            sp.bc_infinity = a.bc_inf;
            Ends e;
            if (s) e = {in[3], in[3], in[3], in[0], in[1], in[2]};
            else e = {in[0], in[1], in[2], in[3], in[3], in[3]};
            double g0, g1;
            ghost_values(sp, e, N, 1, &s_trk[(p * 2 + s) * 15 + trole * 5 + v], g0, g1);
            s_gout[((p * 2 + s) * 5 + v) * 4 + slot] = g0;
        }
        wave_sync();
    };
    auto gin_at = [&](int s, int v, int role) __attribute__((always_inline)) -> double* {
        return s_gin + ((((px * 2 + s) * 5 + v) * 2 + role) * 4);
    };
    auto gout_at = [&](int s, int v, int slot) __attribute__((always_inline)) -> T {
        return T(s_gout[((px * 2 + s) * 5 + v) * 4 + slot]);
    };

    // ================= predictor: W and C from U^n (wFluidFlux.m) ====================
    const T m1 = wave_max(speeds(q));
    publish_max(0, m1);
    if (multi) {
        // halo: flux numerators of this chunk's first cell (lanes < dl) and last cell
        // (lanes >= 64 - dl) for the lanes across the wave edge
        if (lane < dl) {
            T* e = s_e1 + ((wave * 2 + 0) * dl + lane) * 10;
            for_fields([&](auto VC) __attribute__((always_inline)) {
                constexpr int v = decltype(VC)::value;
                e[2 * v] = q[v][0];
                e[2 * v + 1] = numer(VC, q, 0);
            });
        }
        if (lane >= 64 - dl) {
            T* e = s_e1 + ((wave * 2 + 1) * dl + lane - (64 - dl)) * 10;
            for_fields([&](auto VC) __attribute__((always_inline)) {
                constexpr int v = decltype(VC)::value;
                e[2 * v] = q[v][K - 1];
                e[2 * v + 1] = numer(VC, q, K - 1);
            });
        }
    }
    // circular ends: each end's cell for the other end's ghosts (indices are constants
    // in each branch: a select between register-array elements would put the array in
    // scratch memory)
    auto publish_end1 = [&](auto SC) __attribute__((always_inline)) {
        constexpr int s = decltype(SC)::value;
        constexpr int k = s ? K - 1 : 0;
        T* e = s_end1 + (px * 2 + s) * 10;
        for_fields([&](auto VC) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            e[2 * v] = q[v][k];
            e[2 * v + 1] = numer(VC, q, k);
        });
    };
    if (circ && first) publish_end1(Fld<0>{});
    if (circ && last) publish_end1(Fld<1>{});
    if (multi) __syncthreads();
    else wave_sync();
    const T C = combine_max(0, m1);
    const T rC = A::rcp(C);
    const T hfC = T(hf) * C;

    // predictor ghosts: store.fluxR at the lower end, store.fluxL at the upper end
    auto stage_pred = [&](auto SC) __attribute__((always_inline)) {
        constexpr int s = decltype(SC)::value;
        for_fields([&](auto VC) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            double* in = gin_at(s, v, 0);
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                constexpr int k0 = s ? K - 3 : 0;
                const T w = A::div(numer(VC, q, k0 + t), C, rC);
                in[t] = double(s ? T(0.5) * (q[v][k0 + t] - w) : T(0.5) * (q[v][k0 + t] + w));
            }
            if (circ) {  // lower: FR(N-1); upper: FL(0)
                const T* e = s_end1 + (px * 2 + (1 - s)) * 10;
                const T w = A::div(e[2 * v + 1], C, rC);
                in[3] = double(s ? T(0.5) * (e[2 * v] - w) : T(0.5) * (e[2 * v] + w));
            } else {
                in[3] = in[0];
            }
        });
    };
    if (first) stage_pred(Fld<0>{});
    if (last) stage_pred(Fld<1>{});
    ghost_pass(0);

    T qh[5][K];
    for_fields([&](auto VC) __attribute__((always_inline)) {
        constexpr int v = decltype(VC)::value;
        T FR[K], FL[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T w = A::div(numer(VC, q, k), C, rC);
            FR[k] = T(0.5) * (q[v][k] + w);
            FL[k] = T(0.5) * (q[v][k] - w);
        }
        T sR0 = lane_up<xm>(FR[K - 1], dl);   // FR of the left neighbour's last cell
        T sLK = lane_down<xm>(FL[0], dl);     // FL of the right neighbour's first cell
        if (multi && lx) {
            const T* e = s_e1 + (((wave - 1) * 2 + 1) * dl + lane) * 10;
            sR0 = T(0.5) * (e[2 * v] + A::div(e[2 * v + 1], C, rC));
        }
        if (multi && rx) {
            const T* e = s_e1 + (((wave + 1) * 2 + 0) * dl + lane - (64 - dl)) * 10;
            sLK = T(0.5) * (e[2 * v] - A::div(e[2 * v + 1], C, rC));
        }
        if (first) sR0 = gout_at(0, v, 0);
        if (last) sLK = gout_at(1, v, 0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T sR = k > 0 ? FR[k > 0 ? k - 1 : 0] : sR0;
            const T sL = k < K - 1 ? FL[k < K - 1 ? k + 1 : 0] : sLK;
            qh[v][k] = q[v][k] - hfC * (FR[k] - sR + FL[k] - sL);
        }
    });
    // mass.store.array = max(mass.store.array, MINMASS)
#pragma unroll
    for (int k = 0; k < K; ++k) qh[0][k] = fmax(qh[0][k], T(a.minmass));

    // ================= corrector: W' and C' from U^(n+1/2) ===========================
    const T m2 = wave_max(speeds(qh));
    publish_max(1, m2);
    if (multi) {
        if (lane < dl) {
            T* e = s_e2 + ((wave * 2 + 0) * dl + lane) * 20;
            for_fields([&](auto VC) __attribute__((always_inline)) {
                constexpr int v = decltype(VC)::value;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    e[4 * v + 2 * t] = qh[v][t];
                    e[4 * v + 2 * t + 1] = numer(VC, qh, t);
                }
            });
        }
        if (lane >= 64 - dl) {
            T* e = s_e2 + ((wave * 2 + 1) * dl + lane - (64 - dl)) * 20;
            for_fields([&](auto VC) __attribute__((always_inline)) {
                constexpr int v = decltype(VC)::value;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    e[4 * v + 2 * t] = qh[v][K - 2 + t];
                    e[4 * v + 2 * t + 1] = numer(VC, qh, K - 2 + t);
                }
            });
        }
    }
    auto publish_end2 = [&](auto SC) __attribute__((always_inline)) {
        constexpr int s = decltype(SC)::value;
        constexpr int k0 = s ? K - 2 : 0;
        T* e = s_end2 + (px * 2 + s) * 20;
        for_fields([&](auto VC) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                e[4 * v + 2 * t] = qh[v][k0 + t];
                e[4 * v + 2 * t + 1] = numer(VC, qh, k0 + t);
            }
        });
    };
    if (circ && first) publish_end2(Fld<0>{});
    if (circ && last) publish_end2(Fld<1>{});
    if (multi) __syncthreads();
    else wave_sync();
    const T C2 = combine_max(1, m2);
    const T rC2 = A::rcp(C2);
    const T ffC = T(ff) * C2;
    const int lim = a.limiter;
    auto fr2 = [&](T qv, T nu) __attribute__((always_inline)) -> T {
        return T(0.5) * (qv + A::div(nu, C2, rC2));
    };
    auto fl2 = [&](T qv, T nu) __attribute__((always_inline)) -> T {
        return T(0.5) * (qv - A::div(nu, C2, rC2));
    };

    // corrector ghosts: fluxR and fluxL at both ends, from the half-step fluxes
    auto stage_corr = [&](auto SC) __attribute__((always_inline)) {
        constexpr int s = decltype(SC)::value;
        for_fields([&](auto VC) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            double* inR = gin_at(s, v, 0);
            double* inL = gin_at(s, v, 1);
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                constexpr int k0 = s ? K - 3 : 0;
                const T nu = numer(VC, qh, k0 + t);
                inR[t] = double(fr2(qh[v][k0 + t], nu));
                inL[t] = double(fl2(qh[v][k0 + t], nu));
            }
            if (circ) {  // lower: FR', FL' of cell N-1; upper: of cell 0
                const T* e = s_end2 + (px * 2 + (1 - s)) * 20 + 4 * v + (s ? 0 : 2);
                inR[3] = double(fr2(e[0], e[1]));
                inL[3] = double(fl2(e[0], e[1]));
            } else {
                inR[3] = inR[0];
                inL[3] = inL[0];
            }
        });
    };
    if (first) stage_corr(Fld<0>{});
    if (last) stage_corr(Fld<1>{});
    ghost_pass(1);
    // final ghosts: the limited fluxR at the lower end, the limited fluxL at the upper end
    auto stage_final = [&](auto SC) __attribute__((always_inline)) {
        constexpr int s = decltype(SC)::value;
        for_fields([&](auto VC) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            double* in = gin_at(s, v, 0);
            T F[K], Fl[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const T nu = numer(VC, qh, k);
                F[k] = s ? fl2(qh[v][k], nu) : fr2(qh[v][k], nu);
            }
            if constexpr (!s) {  // limited fluxR at cells 0..2: FR'(-1) is the corrector ghost
                const T gm = gout_at(0, v, 1);
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const T rm = t > 0 ? F[t > 0 ? t - 1 : 0] : gm;
                    Fl[t] = F[t] + wlimit<T, FAST>(lim, F[t] - rm, F[t + 1] - F[t]);
                }
                in[0] = double(Fl[0]);
                in[1] = double(Fl[1]);
                in[2] = double(Fl[2]);
                if (circ) {  // limFR(N-1) from FR'(N-2), FR'(N-1) and FR'(0)
                    const T* e = s_end2 + (px * 2 + 1) * 20 + 4 * v;
                    const T fa = fr2(e[0], e[1]), fb = fr2(e[2], e[3]);
                    in[3] = double(fb + wlimit<T, FAST>(lim, fb - fa, F[0] - fb));
                } else {
                    in[3] = in[0];
                }
            } else {  // limited fluxL at cells N-3..N-1: FL'(N) is the corrector ghost
                const T gp = gout_at(1, v, 2);
#pragma unroll
                for (int t = 1; t < K; ++t) {
                    const T lp = t < K - 1 ? F[t < K - 1 ? t + 1 : 0] : gp;
                    Fl[t] = F[t] + wlimit<T, FAST>(lim, F[t - 1] - F[t], F[t] - lp);
                }
                in[0] = double(Fl[K - 3]);
                in[1] = double(Fl[K - 2]);
                in[2] = double(Fl[K - 1]);
                if (circ) {  // limFL(0) from FL'(N-1), FL'(0) and FL'(1)
                    const T* e = s_end2 + (px * 2 + 0) * 20 + 4 * v;
                    const T fa = fl2(e[0], e[1]), fb = fl2(e[2], e[3]);
                    in[3] = double(fa + wlimit<T, FAST>(lim, F[K - 1] - fa, fa - fb));
                } else {
                    in[3] = in[0];
                }
            }
        });
    };
    if (first) stage_final(Fld<0>{});
    if (last) stage_final(Fld<1>{});
    ghost_pass(2);

    for_fields([&](auto VC) __attribute__((always_inline)) {
        constexpr int v = decltype(VC)::value;
        T un[K];
        load4<T>(a.q[v] + cb, stride, xm, un);  // U^n (cache-resident)
        T FR[K + 4], FL[K + 4];                 // cells i0-2 .. i0+K+1
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T nu = numer(VC, qh, k);
            FR[k + 2] = fr2(qh[v][k], nu);
            FL[k + 2] = fl2(qh[v][k], nu);
        }
        FR[0] = lane_up<xm>(FR[K], dl);       // FR'(i0-2)
        FR[1] = lane_up<xm>(FR[K + 1], dl);   // FR'(i0-1)
        FL[1] = lane_up<xm>(FL[K + 1], dl);   // FL'(i0-1)
        FR[K + 2] = lane_down<xm>(FR[2], dl); // FR'(i0+K)
        FL[K + 2] = lane_down<xm>(FL[2], dl); // FL'(i0+K)
        FL[K + 3] = lane_down<xm>(FL[3], dl); // FL'(i0+K+1)
        if (multi && lx) {
            const T* e = s_e2 + (((wave - 1) * 2 + 1) * dl + lane) * 20 + 4 * v;
            FR[0] = fr2(e[0], e[1]);
            FR[1] = fr2(e[2], e[3]);
            FL[1] = fl2(e[2], e[3]);
        }
        if (multi && rx) {
            const T* e = s_e2 + (((wave + 1) * 2 + 0) * dl + lane - (64 - dl)) * 20 + 4 * v;
            FR[K + 2] = fr2(e[0], e[1]);
            FL[K + 2] = fl2(e[0], e[1]);
            FL[K + 3] = fl2(e[2], e[3]);
        }
        if (first) {
            FR[1] = gout_at(0, v, 1);
            FL[1] = gout_at(0, v, 2);
            FR[0] = FR[1];
        }
        if (last) {
            FR[K + 2] = gout_at(1, v, 1);
            FL[K + 2] = gout_at(1, v, 2);
            FL[K + 3] = FL[K + 2];
        }
        // limited fluxes: R at cells i0-1 .. i0+K-1 (index j -> cell i0-1+j), L at cells
        // i0 .. i0+K (index j -> cell i0+j)
        T R2[K + 1], L2[K + 1];
#pragma unroll
        for (int j = 0; j <= K; ++j) {
            const T fr = FR[j + 1];
            R2[j] = fr + wlimit<T, FAST>(lim, fr - FR[j], FR[j + 2] - fr);
            const T fl = FL[j + 2];
            L2[j] = fl + wlimit<T, FAST>(lim, FL[j + 1] - fl, fl - FL[j + 3]);
        }
        if (first) R2[0] = gout_at(0, v, 3);
        if (last) L2[K] = gout_at(1, v, 3);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T dq = ffC * (R2[k + 1] - R2[k] + L2[k] - L2[k + 1]);
            un[k] = un[k] - dq;
            if (v == 0) un[k] = fmax(un[k], T(a.minmass));  // mass.array = max(mass, MINMASS)
        }
        if (pen_ok) store4<T>(a.q[v] + cb, stride, xm, un);
    });

    // ---- edge trackers back to memory ----------------------------------------------
    if (trk_lo || trk_hi) {
        for (int j = lane; j < npair * 15; j += 64) {
            int p, s;
            end_pair(wave, NW, PX, TL, xm, j / 15, p, s);
            const int r = (j % 15) / 5, v = j % 5;
            if (pen0 + p >= a.npencil || (s ? !trk_hi : !trk_lo)) continue;
            const int obj = r == 0 ? (s ? OBJ_SFL : OBJ_SFR) : (r == 1 ? OBJ_FR : OBJ_FL);
            a.edges[a.eoff[((obj + v) * 2 + s) * 3 + d] + pen0 + p] =
                s_trk[(p * 2 + s) * 15 + r * 5 + v];
        }
    }
}

}  // namespace wv

// Waves per SIMD the register allocator aims for (a build-time knob; with 256-thread
// blocks the launch bound alone lets it take every register of the SIMD).
#ifndef IMH_WAVE_WAVES_PER_EU
#define IMH_WAVE_WAVES_PER_EU 2
#endif

// Persistent batched form (tiles != nullptr) or one block per tile of args[0].
template <typename T, int BS, bool MHD, bool FAST, int D>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(IMH_WAVE_WAVES_PER_EU)))
fluid_wave_kernel(const FluidArgs<T>* args,
                                                        const BatchTile* tiles, int ntiles) {
    extern __shared__ __attribute__((aligned(16))) char wv_lds[];
    for (int r = 0;; ++r) {
        int arg = 0;
        long long tile;
        if (tiles) {
            const int t = xcd_round_tile(r, ntiles);
            if (t < 0) break;
            uniform_tile(tiles, t, arg, tile);
        } else {
            if (r > 0 || (int)blockIdx.x >= ntiles) break;
            // y / z tiles of neighbouring x pencils share 128-byte lines: same XCD, same L2
            tile = args[0].xcd_remap ? xcd_tile(blockIdx.x, ntiles) : (long long)blockIdx.x;
        }
        wv::wave_sweep_body<T, BS, MHD, FAST, D>(args[arg], tile, wv_lds);
        __syncthreads();
    }
}

}  // namespace imh
