// fluid_long.hip -- the fluid sweep (relaxingFluid.m:16-73 with wFluidFlux.m:14-44) along
// pencils longer than one workgroup holds.
//
// The fused sweep keeps a pencil in one workgroup (up to 4096 cells); the reference sweeps
// any length.  Here a sweep is five launches over all cells of the dimension, each cell
// one thread, with the pencil-wide quantities passed through device memory:
//
//   speeds     freezeSpd(X): the pencil maximum of |v_X| + c_fast of U^n (wFluidFlux.m:20),
//              a device-wide maximum per pencil (atomic on the value's bits: the maximum
//              of non-negative numbers does not depend on the order)
//   ghosts 0   per pencil end: store.fluxR's lower and store.fluxL's upper ghost
//              (shift(X,-1) / shift(X,+1)), the edge trackers nudged once per call
//   predictor  U^{n+1/2} of every cell from U^n and its neighbours' half-step fluxes;
//              the mass floor; the pencil maximum of U^{n+1/2}'s speeds (freezeSpdTVD)
//   ghosts 1/2 per pencil end: fluxR / fluxL ghosts on both sides (the limiter's
//              shifts), then the limited fluxes at the ends and their ghosts (the final
//              update's shifts)
//   corrector  the limited fluxes of a cell and its neighbours from U^{n+1/2}, U^{n+1}
//
// Every cell evaluates the same expressions, in the same order, as the fused sweep
// (fluid_sweep_impl.h), so the two agree bit for bit; FP64 matches the restatement.  The
// boundary ghosts come from the same rule (ghost_values) with the same per-object
// trackers, and a simulation's first sweep records the flux objects' edges as the fused
// form does (InitializedArray.m:43-47).  Fade ends blend the shifted rows near both ends
// from per-pencil end values (fade_at); artificial viscosity (ArtificialViscosity.m) adds
// a per-pencil ghost replay and a per-cell term before the predictor and the corrector.
// A lower wormhole boundary is not taken here.
#define IMH_FLUID_NO_LAUNCHER
#include "fluid_sweep_impl.h"

namespace imh {

namespace {

template <typename T>
struct LongCell {
    T q[5];
    T bm[3];
};

template <typename T, bool MHD>
__device__ __forceinline__ void load_cell(const FluidArgs<T>& a, long long c, LongCell<T>& x,
                                          const T* const* src) {
#pragma unroll
    for (int v = 0; v < 5; ++v) x.q[v] = src[v][c];
#pragma unroll
    for (int v = 0; v < 3; ++v) x.bm[v] = MHD ? a.bc[v][c] : T(0);
}

// cell_speeds of the fused sweep for one cell: P, c, 1/rho (or rho), U_X and |U_X| + c
template <typename T>
__device__ __forceinline__ T cell_speed(const FluidArgs<T>& a, const LongCell<T>& x, T& P, T& U,
                                        T& IR) {
    using M = FMath<T>;
    T ck;
    M::press(x.q[0], x.q[1], x.q[2], x.q[3], x.q[4], x.bm[0], x.bm[1], x.bm[2], a.use, a.gamma,
             P, ck, IR);
    U = M::divrho(sel3(a.d, x.q[1], x.q[2], x.q[3]), IR);
    return fabs(U) + ck;
}

// wFluidFlux.m: FR / FL of every field of one cell (the fused sweep's `fluxes`)
template <typename T>
__device__ __forceinline__ void cell_fluxes(const FluidArgs<T>& a, const LongCell<T>& x, T C,
                                            T rC, T (&FR)[5], T (&FL)[5], T visc) {
    using M = FMath<T>;
    T P, U, IR;
    (void)cell_speed(a, x, P, U, IR);
    if (a.visc_type) P = P + visc;  // press = press + viscosity.solve(...) (wFluidFlux.m:24)
    const int d = a.d, X = d + 1;
    const T bX = sel3(d, x.bm[0], x.bm[1], x.bm[2]);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        T w;
        if (v == 0) {
            w = M::div(sel3(d, x.q[1], x.q[2], x.q[3]), C, rC);
        } else if (v == 4) {
            T bdotp = x.bm[0] * x.q[1] + x.bm[1] * x.q[2] + x.bm[2] * x.q[3];
            T we = U * (x.q[4] + P) - M::divrho(bX * bdotp, IR);
            w = M::div(we, C, rC);
        } else {
            T dv = (v == X) ? T(1) : T(0);
            w = M::div(U * x.q[v] + P * dv - bX * x.bm[v - 1], C, rC);
        }
        FR[v] = T(0.5) * (x.q[v] + w);
        FL[v] = T(0.5) * (x.q[v] - w);
    }
}

__device__ __forceinline__ void max_bits(unsigned long long* p, double v) {
    if (!(v == v)) return;  // MATLAB max ignores NaN
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__device__ __forceinline__ T from_bits(unsigned long long b) {
    return T(__longlong_as_double((long long)b));
}

template <typename T>
__device__ __forceinline__ long long cell_index(const FluidArgs<T>& a, long long p, int i) {
    const int nn[3] = {a.n0, a.n1, a.n2};
    long long base, stride;
    pencil_base(nn, a.d, p, base, stride);
    return base + (long long)i * stride;
}

// One ghost of a pencil end: the fused sweep's ghost_phase for job (field v, row which
// (0 fluxR, 1 fluxL), side upper), role 0 (store.flux*) or 1 (fluxR / fluxL);
// `first_assign` records the end value in the tracker first (the object's first array).
template <typename T>
__device__ double end_ghost(const FluidArgs<T>& a, long long p, int v, int which, int upper,
                            int obj, const Ends& e, bool first_assign) {
    const int d = a.d;
    const int mode = mode_at(a.modes, upper, d);
    const int N = sel3i(d, a.n0, a.n1, a.n2);
    double* tr = nullptr;
    double scratch = 0.0;
    if (mode == IMH_BC_TRANS || mode == IMH_BC_WALL) {
        tr = a.edges + a.eoff[(obj * 2 + upper) * 3 + d] + p;
        if (first_assign) *tr = upper ? e.b2 : e.a0;
    }
    if (mode == IMH_BC_CIRC) return upper ? e.a0 : e.b2;
    if (mode == IMH_BC_CONST) return upper ? e.b2 : e.a0;
    const int ft = which ? FT_FLUXL : FT_FLUXR;
    GhostSpec sp;
    sp.mode = mode;
    sp.upper = upper;
    sp.flip_sign = (v == d + 1);
    sp.zero_ghost = upper ? (ft != FT_FLUXR) : (ft != FT_FLUXL);
    sp.active_linear = a.active_lin;
    sp.bc_infinity = a.bc_inf;
    double g0, g1;
    ghost_values(sp, e, N, 1, tr ? tr : &scratch, g0, g1);
    return g0;
}

// ghost slots per pencil: [v][0..7] = store.fluxR lower, store.fluxL upper (predictor),
// fluxR lower, fluxL lower, fluxR upper, fluxL upper (limiter), limited fluxR lower,
// limited fluxL upper (final update)
constexpr int GH = FLUID_LONG_GH;

template <typename T>
struct LongArgs {
    FluidArgs<T> a;
    T* half[5];                 // U^{n+1/2}
    unsigned long long* cmax;   // [2][npencil]: freezeSpd, freezeSpdTVD (bits of doubles)
    double* gh;                 // [npencil][5][GH]
    double* en;                 // [npencil][5][3 stages][2 rows][6]: row ends (fade)
    T* vis;                     // the artificial viscosity term of every cell
    double* vg;                 // [npencil][VG]: its ghosts (see long_visc_ghosts)
};
constexpr int VG = FLUID_LONG_VG;

// the row ends (cells 0..2, N-3..N-1) of stage s (0 store.flux*, 1 raw flux*, 2 limited)
// and row w (0 fluxR, 1 fluxL) of field v of pencil p
__device__ __forceinline__ double* row_ends(double* en, long long p, int v, int s, int w) {
    return en + ((p * 5 + v) * 3 + s) * 12 + w * 6;
}
__device__ __forceinline__ void put_row_ends(double* en, long long p, int v, int s, int w,
                                             const Ends& e) {
    double* o = row_ends(en, p, v, s, w);
    o[0] = e.a0;
    o[1] = e.a1;
    o[2] = e.a2;
    o[3] = e.b0;
    o[4] = e.b1;
    o[5] = e.b2;
}

// fade_shift.m: the constant shift, then both ends of the shifted array blended toward
// its end values over bcInfinity cells (lower: shift(X,-1), whose ends are a[0] and
// a[N-2]; upper: shift(X,+1), a[1] and a[N-1]); the fused sweep's fade_fix
template <typename T>
__device__ __forceinline__ T fade_at(const FluidArgs<T>& a, bool lower, int i, int N,
                                     const double* e, T r) {
    if (mode_at(a.modes, lower ? 0 : 1, a.d) != IMH_BC_FADE) return r;
    const int f = a.bc_inf;
    if (!((i < f || i >= N - f) && i < N)) return r;
    const double v1 = lower ? e[0] : e[1];
    const double vN = lower ? e[4] : e[5];
    return T(fade_blend<double>(double(r), i, N, v1, vN, a.fade, f));
}


// ---- artificial viscosity (ArtificialViscosity.m), added to the pressure of the fluxes --
// The arrays it shifts: rows 0 mass, 1..3 momentum, 4 energy (the FluidArrays in the
// predictor, their StorageArrays in the corrector), 5..7 cell-centred B.
template <typename T, bool MHD>
__device__ __forceinline__ T prim_at(const FluidArgs<T>& a, const T* const* src, int r,
                                     long long c) {
    if (r < 5) return src[r][c];
    return MHD ? a.bc[r - 5][c] : T(0);
}

// Row r of the array shifted by n (|n| <= 2; shift(X, n) takes cell i + n) at cell i of
// pencil p, for the index-map modes (circ, const, mirror, flip, zero on a non-flux array,
// fade after its blend of both ends); component flips for mirror and flip.
template <typename T, bool MHD>
__device__ T shifted_prim(const FluidArgs<T>& a, const T* const* src, int r, long long p, int i,
                          int n, int N) {
    const int d = a.d;
    const int side = n > 0 ? 1 : 0;
    const int mode = mode_at(a.modes, side, d);
    const bool comp = (r == 1 + d) || (r == 5 + d);
    auto at = [&](int j) -> T { return prim_at<T, MHD>(a, src, r, cell_index(a, p, j)); };
    int j = i + n;
    T v;
    bool neg = false;
    if (j >= 0 && j < N) {
        v = at(j);
    } else if (mode == IMH_BC_CIRC) {
        v = at(j < 0 ? j + N : j - N);
    } else if (mode == IMH_BC_MIRROR) {
        v = at(j < 0 ? -j : 2 * (N - 1) - j);
        neg = comp;
    } else {  // const, fade, zero (not a flux array), flip
        v = at(j < 0 ? 0 : N - 1);
        neg = (mode == IMH_BC_FLIP) && comp;
    }
    if (neg) v = -v;
    if (mode == IMH_BC_FADE) {
        const int f = a.bc_inf;
        if (i < f || i >= N - f) {
            // ends of the constant-shifted array: r[0] = a[max(n, 0)], r[N-1] = a[min(N-1+n, N-1)]
            const T v1 = at(n > 0 ? n : 0), vN = at(n < 0 ? N - 1 + n : N - 1);
            v = fade_blend<T>(v, i, N, v1, vN, a.fade, f);
        }
    }
    return v;
}

// Ghosts the viscosity's shifts take past a transparent or wall end, one thread per
// pencil (each shift call nudges its object's tracker once): [0] the Neumann-Richtmyer
// upper ghost of mom(X); [1 + 5*side ..] Caramana-Shashkov-Whalen's values of the end
// cells (csw_side_ghosts).
template <typename T, bool MHD>
__global__ void __launch_bounds__(256) long_visc_ghosts(LongArgs<T> L, int half, int obj,
                                                        int first) {
    const FluidArgs<T>& a = L.a;
    if (!a.force && !a.sim->active) return;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npencil) return;
    const int d = a.d, X = d + 1;
    const int N = sel3i(d, a.n0, a.n1, a.n2);
    const T* const* src = half ? (const T* const*)L.half : (const T* const*)a.q;
    double* vg = L.vg + p * VG;
    auto ends = [&](int r, double* e) {
        for (int s = 0; s < 6; ++s) {
            const int i = s < 3 ? s : N - 6 + s;
            e[s] = double(prim_at<T, MHD>(a, src, r, cell_index(a, p, i)));
        }
    };
    if (a.visc_type == 1) {
        const int mhi = mode_at(a.modes, 1, d);
        double e[6];
        ends(X, e);
        const bool trk = (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL);
        double* tp = trk ? a.edges + a.eoff[((obj + X) * 2 + 1) * 3 + d] + p : nullptr;
        double tv = trk ? (first ? e[5] : *tp) : 0.0;
        GhostSpec sp;
        sp.mode = mhi;
        sp.upper = 1;
        sp.flip_sign = 1;  // the shifted array is the momentum component X
        sp.zero_ghost = 0;
        sp.active_linear = a.active_lin;
        sp.bc_infinity = a.bc_inf;
        const Ends en = {e[0], e[1], e[2], e[3], e[4], e[5]};
        double g0, g1;
        ghost_values(sp, en, N, 1, &tv, g0, g1);
        if (trk) *tp = tv;
        vg[0] = g0;
    } else if (a.visc_type == 2) {
        double e[8 * 6];
        bool any = false;
        for (int up = 0; up < 2; ++up) {
            const int mode = mode_at(a.modes, up, d);
            if (mode == IMH_BC_TRANS || mode == IMH_BC_WALL) any = true;
        }
        if (!any) return;
        for (int r = 0; r < 8; ++r) ends(r, e + r * 6);
        for (int up = 0; up < 2; ++up) {
            const int mode = mode_at(a.modes, up, d);
            if (mode != IMH_BC_TRANS && mode != IMH_BC_WALL) continue;
            double* tp[8];
            for (int r = 0; r < 8; ++r) {
                const int o = r < 5 ? obj + r : OBJ_CMAG + (r - 5);
                tp[r] = a.edges + a.eoff[(o * 2 + up) * 3 + d] + p;
            }
            csw_side_ghosts(mode, up, d, N, a.active_lin, a.bc_inf, a.gamma, tp, e,
                            vg + 1 + up * 5);
        }
    }
}

// The viscosity term of every cell (ArtificialViscosity.m:73-77 and 80-148 with
// cswSoundSpeed :155-190), in the fused sweep's expressions and order
template <typename T, bool MHD>
__global__ void __launch_bounds__(256) long_visc_cells(LongArgs<T> L, int half) {
    const FluidArgs<T>& a = L.a;
    if (!a.force && !a.sim->active) return;
    using M = FMath<T>;
    const int N = sel3i(a.d, a.n0, a.n1, a.n2);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npencil * (long long)N) return;
    const long long p = t / N;
    const int i = int(t - p * N);
    const int d = a.d, X = d + 1;
    const T* const* src = half ? (const T* const*)L.half : (const T* const*)a.q;
    const long long c = cell_index(a, p, i);
    const double* vg = L.vg + p * VG;
    LongCell<T> x;
    load_cell<T, MHD>(a, c, x, src);
    T term;
    if (a.visc_type == 1) {
        T P, U, IR, ck;
        M::press(x.q[0], x.q[1], x.q[2], x.q[3], x.q[4], x.bm[0], x.bm[1], x.bm[2], a.use,
                 a.gamma, P, ck, IR);
        const int mhi = mode_at(a.modes, 1, d);
        const T mX = x.q[X];
        T sh = i < N - 1 ? src[X][cell_index(a, p, i + 1)] : T(vg[0]);
        if (mhi == IMH_BC_FADE)  // fade_shift.m: both ends blended, v1 = a[1], vN = a[N-1]
            sh = fade_blend(sh, i, N, src[X][cell_index(a, p, 1)],
                            src[X][cell_index(a, p, N - 1)], a.fade, a.bc_inf);
        const T vm = fmin(sh - mX, T(0));  // MATLAB min(.., 0) skips NaN
        term = vm * (T(a.visc_lin) * ck + T(a.visc_quad) * vm / x.q[0]);
    } else {
        const T G = T(a.gamma);
        // a shifted row past a transparent or wall end comes from the ghost replay
        auto replay = [&](int n) -> bool {
            const int j = i + n;
            if (j >= 0 && j < N) return false;
            const int mode = mode_at(a.modes, j < 0 ? 0 : 1, d);
            return mode == IMH_BC_TRANS || mode == IMH_BC_WALL;
        };
        auto ov = [&](int n) -> const double* { return vg + 1 + (i + n < 0 ? 0 : 5); };
        // cswSoundSpeed of the arrays shifted by n (n = +-1)
        auto sound = [&](int n) -> T {
            if (replay(n)) return T(ov(n)[1]);
            const T t1 = shifted_prim<T, MHD>(a, src, 1, p, i, n, N);
            const T t2 = shifted_prim<T, MHD>(a, src, 2, p, i, n, N);
            const T t3 = shifted_prim<T, MHD>(a, src, 3, p, i, n, N);
            const T tm = shifted_prim<T, MHD>(a, src, 0, p, i, n, N);
            const T vel2 = t1 * t1 + t2 * t2 + t3 * t3 / (tm * tm);
            const T b1 = shifted_prim<T, MHD>(a, src, 5, p, i, n, N);
            const T b2 = shifted_prim<T, MHD>(a, src, 6, p, i, n, N);
            const T b3 = shifted_prim<T, MHD>(a, src, 7, p, i, n, N);
            const T mag2 = b1 * b1 + b2 * b2 + b3 * b3;
            const T te = shifted_prim<T, MHD>(a, src, 4, p, i, n, N);
            T sp = (G - T(1)) * (te - T(0.5) * tm * vel2);
            sp = sp - (G - T(1)) * T(0.5) * mag2;
            return sqrt(fabs((G * sp + T(2) * mag2) / tm));
        };
        auto mass_at = [&](int n) -> T {
            if (replay(n)) return T(ov(n)[0]);
            return shifted_prim<T, MHD>(a, src, 0, p, i, n, N);
        };
        auto vel_at = [&](int n) -> T {
            if (replay(n)) {
                const double* o = ov(n);
                if (n == -1) return T(o[4]);
                if (n == -2) return T(i == 0 ? o[2] : o[3]);
                if (n == 1) return T(o[2]);
                return T(i == N - 2 ? o[3] : o[4]);
            }
            return shifted_prim<T, MHD>(a, src, X, p, i, n, N) /
                   shifted_prim<T, MHD>(a, src, 0, p, i, n, N);
        };
        const T mB = mass_at(-1), mC = mass_at(1);
        const T rho = T(2) * mC * mB / (mC + mB);
        const T snd = fmax(fmin(sound(-1), sound(1)), T(0));
        const T v0 = vel_at(-2), v1 = vel_at(-1), v2 = vel_at(1), v3 = vel_at(2);
        const T c1 = T(a.visc_lin), c2 = T(a.visc_quad);
        const T q4 = (G + T(1)) / T(4);
        const T dv = v2 - v1;
        const T rL = (v3 - v2) / (T(2) * dv);
        const T rR = (v1 - v0) / (T(2) * dv);
        const T lim = fmax(T(0), fmin(fmin(T(0.5) * (rL + rR), fmin(T(2) * rL, T(2) * rR)), T(1)));
        const T A = c2 * (G + T(1)) * fabs(dv) / T(4);
        const T B = sqrt(c2 * c2 * (q4 * q4) * (dv * dv) + c1 * c1 * (snd * snd));
        term = (dv < T(0) ? T(1) : T(0)) * rho * (A + B) * (T(1) - lim) * fabs(dv);
    }
    L.vis[c] = term;
}

template <typename T, bool MHD>
__global__ void __launch_bounds__(256) long_speeds(LongArgs<T> L) {
    const FluidArgs<T>& a = L.a;
    if (!a.force && !a.sim->active) return;
    const int N = sel3i(a.d, a.n0, a.n1, a.n2);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npencil * (long long)N) return;
    const long long p = t / N;
    const long long c = cell_index(a, p, int(t - p * N));
    LongCell<T> x;
    load_cell<T, MHD>(a, c, x, a.q);
    T P, U, IR;
    max_bits(&L.cmax[p], double(cell_speed(a, x, P, U, IR)));
}

// Row values (fluxes) of the first and last three cells of pencil p
template <typename T, bool MHD>
__device__ void end_fluxes(const LongArgs<T>& L, const T* const* src, long long p, T C,
                           T (&FR)[6][5], T (&FL)[6][5]) {
    const FluidArgs<T>& a = L.a;
    const int N = sel3i(a.d, a.n0, a.n1, a.n2);
    const T rC = FMath<T>::rcp(C);
    for (int s = 0; s < 6; ++s) {
        const int i = s < 3 ? s : N - 6 + s;
        LongCell<T> x;
        const long long c = cell_index(a, p, i);
        load_cell<T, MHD>(a, c, x, src);
        cell_fluxes(a, x, C, rC, FR[s], FL[s], a.visc_type ? L.vis[c] : T(0));
    }
}

template <typename T, bool MHD>
__global__ void __launch_bounds__(256) long_ghosts_pred(LongArgs<T> L) {
    const FluidArgs<T>& a = L.a;
    if (!a.force && !a.sim->active) return;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npencil) return;
    const T C = from_bits<T>(L.cmax[p]);
    T FR[6][5], FL[6][5];
    end_fluxes<T, MHD>(L, a.q, p, C, FR, FL);
    double* g = L.gh + p * 5 * GH;
    for (int v = 0; v < 5; ++v) {
        const Ends eR = {double(FR[0][v]), double(FR[1][v]), double(FR[2][v]),
                         double(FR[3][v]), double(FR[4][v]), double(FR[5][v])};
        const Ends eL = {double(FL[0][v]), double(FL[1][v]), double(FL[2][v]),
                         double(FL[3][v]), double(FL[4][v]), double(FL[5][v])};
        g[v * GH + 0] = end_ghost(a, p, v, 0, 0, OBJ_SFR + v, eR, a.init_edges != 0);
        g[v * GH + 1] = end_ghost(a, p, v, 1, 1, OBJ_SFL + v, eL, a.init_edges != 0);
        put_row_ends(L.en, p, v, 0, 0, eR);
        put_row_ends(L.en, p, v, 0, 1, eL);
    }
}

template <typename T, bool MHD>
__global__ void __launch_bounds__(256) long_predictor(LongArgs<T> L) {
    const FluidArgs<T>& a = L.a;
    if (!a.force && !a.sim->active) return;
    using M = FMath<T>;
    const int N = sel3i(a.d, a.n0, a.n1, a.n2);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npencil * (long long)N) return;
    const long long p = t / N;
    const int i = int(t - p * N);
    const T C = from_bits<T>(L.cmax[p]);
    const T rC = M::rcp(C);
    const double dt = a.sim->dt;
    const double hf = 0.5 * (dt / a.dx);
    T hk = T(hf) * C;
    if (a.dxl) hk = T(0.5 * (dt / a.dxl[i])) * C;
    T FR[3][5], FL[3][5];  // cells i-1, i, i+1
    LongCell<T> x;
    const long long c = cell_index(a, p, i);
    for (int s = 0; s < 3; ++s) {
        const int j = i - 1 + s;
        if (j < 0 || j >= N) continue;
        LongCell<T> y;
        const long long cj = cell_index(a, p, j);
        load_cell<T, MHD>(a, cj, y, a.q);
        if (s == 1) x = y;
        cell_fluxes(a, y, C, rC, FR[s], FL[s], a.visc_type ? L.vis[cj] : T(0));
    }
    const double* g = L.gh + p * 5 * GH;
    T qh[5];
    for (int v = 0; v < 5; ++v) {
        T sR = i > 0 ? FR[0][v] : T(g[v * GH + 0]);
        T sL = i < N - 1 ? FL[2][v] : T(g[v * GH + 1]);
        sR = fade_at(a, true, i, N, row_ends(L.en, p, v, 0, 0), sR);
        sL = fade_at(a, false, i, N, row_ends(L.en, p, v, 0, 1), sL);
        qh[v] = x.q[v] - hk * (FR[1][v] - sR + FL[1][v] - sL);
    }
    if (a.init_edges) {
        // first assignment of store.fluxR / store.fluxL and of the StorageArrays (the
        // mass store's before its MINMASS floor): their edges on the other faces
        for (int v = 0; v < 5; ++v) {
            if (a.ndims >= 2) {
                fluid_init_face(a.edges, a.eoff, a.modes, a.ndims, a.d, OBJ_SFR + v, i, p,
                                a.n0, a.n1, a.n2, double(FR[1][v]));
                fluid_init_face(a.edges, a.eoff, a.modes, a.ndims, a.d, OBJ_SFL + v, i, p,
                                a.n0, a.n1, a.n2, double(FL[1][v]));
                fluid_init_face(a.edges, a.eoff, a.modes, a.ndims, a.d, OBJ_FSTORE + v, i, p,
                                a.n0, a.n1, a.n2, double(qh[v]));
            }
            const int mlo = mode_at(a.modes, 0, a.d), mhi = mode_at(a.modes, 1, a.d);
            if (i == 0 && (mlo == IMH_BC_TRANS || mlo == IMH_BC_WALL))
                a.edges[a.eoff[((OBJ_FSTORE + v) * 2 + 0) * 3 + a.d] + p] = double(qh[v]);
            if (i == N - 1 && (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL))
                a.edges[a.eoff[((OBJ_FSTORE + v) * 2 + 1) * 3 + a.d] + p] = double(qh[v]);
        }
    }
    qh[0] = fmax(qh[0], T(a.minmass));
    for (int v = 0; v < 5; ++v) L.half[v][c] = qh[v];
    LongCell<T> h;
    for (int v = 0; v < 5; ++v) h.q[v] = qh[v];
    for (int v = 0; v < 3; ++v) h.bm[v] = x.bm[v];
    T P, U, IR;
    max_bits(&L.cmax[a.npencil + p], double(cell_speed(a, h, P, U, IR)));
}

template <typename T, bool MHD>
__global__ void __launch_bounds__(256) long_ghosts_corr(LongArgs<T> L) {
    const FluidArgs<T>& a = L.a;
    if (!a.force && !a.sim->active) return;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npencil) return;
    const int N = sel3i(a.d, a.n0, a.n1, a.n2);
    const T C2 = from_bits<T>(L.cmax[a.npencil + p]);
    const T* half[5] = {L.half[0], L.half[1], L.half[2], L.half[3], L.half[4]};
    // raw fluxes of cells 0..3 and N-4..N-1 (rows r = 0..7)
    T FR[8][5], FL[8][5];
    const T rC2 = FMath<T>::rcp(C2);
    for (int r = 0; r < 8; ++r) {
        const int i = r < 4 ? r : N - 8 + r;
        LongCell<T> x;
        const long long c = cell_index(a, p, i);
        load_cell<T, MHD>(a, c, x, half);
        cell_fluxes(a, x, C2, rC2, FR[r], FL[r], a.visc_type ? L.vis[c] : T(0));
    }
    double* g = L.gh + p * 5 * GH;
    const int lim = a.limiter;
    for (int v = 0; v < 5; ++v) {
        // the rows' ends: cells 0..2 (r 0..2) and N-3..N-1 (r 5..7)
        const Ends eR = {double(FR[0][v]), double(FR[1][v]), double(FR[2][v]),
                         double(FR[5][v]), double(FR[6][v]), double(FR[7][v])};
        const Ends eL = {double(FL[0][v]), double(FL[1][v]), double(FL[2][v]),
                         double(FL[5][v]), double(FL[6][v]), double(FL[7][v])};
        const bool first = a.init_edges != 0;
        const T gRlo = T(end_ghost(a, p, v, 0, 0, OBJ_FR + v, eR, first));
        const T gLlo = T(end_ghost(a, p, v, 1, 0, OBJ_FL + v, eL, first));
        const T gRhi = T(end_ghost(a, p, v, 0, 1, OBJ_FR + v, eR, first));
        const T gLhi = T(end_ghost(a, p, v, 1, 1, OBJ_FL + v, eL, first));
        g[v * GH + 2] = double(gRlo);
        g[v * GH + 3] = double(gLlo);
        g[v * GH + 4] = double(gRhi);
        g[v * GH + 5] = double(gLhi);
        put_row_ends(L.en, p, v, 1, 0, eR);
        put_row_ends(L.en, p, v, 1, 1, eL);
        const double* e1R = row_ends(L.en, p, v, 1, 0);
        const double* e1L = row_ends(L.en, p, v, 1, 1);
        // limited fluxes of cells 0..2 and N-3..N-1
        auto rowR = [&](int i) -> T {  // raw fluxR of cell i in [-1, 3] or [N-4, N]
            if (i < 0) return gRlo;
            if (i >= N) return gRhi;
            return i < 4 ? FR[i][v] : FR[i - (N - 8)][v];
        };
        auto rowL = [&](int i) -> T {
            if (i < 0) return gLlo;
            if (i >= N) return gLhi;
            return i < 4 ? FL[i][v] : FL[i - (N - 8)][v];
        };
        T R2[6], L2[6];
        for (int s = 0; s < 6; ++s) {
            const int i = s < 3 ? s : N - 6 + s;
            const T fr = rowR(i), fl = rowL(i);
            const T rm = fade_at(a, true, i, N, e1R, rowR(i - 1));
            const T rp = fade_at(a, false, i, N, e1R, rowR(i + 1));
            const T lm = fade_at(a, true, i, N, e1L, rowL(i - 1));
            const T lp = fade_at(a, false, i, N, e1L, rowL(i + 1));
            R2[s] = fr + limit(lim, fr - rm, rp - fr);
            L2[s] = fl + limit(lim, lm - fl, fl - lp);
        }
        const Ends e2R = {double(R2[0]), double(R2[1]), double(R2[2]),
                          double(R2[3]), double(R2[4]), double(R2[5])};
        const Ends e2L = {double(L2[0]), double(L2[1]), double(L2[2]),
                          double(L2[3]), double(L2[4]), double(L2[5])};
        g[v * GH + 6] = end_ghost(a, p, v, 0, 0, OBJ_FR + v, e2R, false);
        g[v * GH + 7] = end_ghost(a, p, v, 1, 1, OBJ_FL + v, e2L, false);
        put_row_ends(L.en, p, v, 2, 0, e2R);
        put_row_ends(L.en, p, v, 2, 1, e2L);
    }
}

template <typename T, bool MHD>
__global__ void __launch_bounds__(256) long_corrector(LongArgs<T> L) {
    const FluidArgs<T>& a = L.a;
    if (!a.force && !a.sim->active) return;
    using M = FMath<T>;
    const int N = sel3i(a.d, a.n0, a.n1, a.n2);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npencil * (long long)N) return;
    const long long p = t / N;
    const int i = int(t - p * N);
    const T C2 = from_bits<T>(L.cmax[a.npencil + p]);
    const T rC2 = M::rcp(C2);
    const T* half[5] = {L.half[0], L.half[1], L.half[2], L.half[3], L.half[4]};
    const double* g = L.gh + p * 5 * GH;
    // raw fluxes of cells i-2 .. i+2
    T FR[5][5], FL[5][5];
    for (int s = 0; s < 5; ++s) {
        const int j = i - 2 + s;
        if (j < 0 || j >= N) continue;
        LongCell<T> x;
        const long long cj = cell_index(a, p, j);
        load_cell<T, MHD>(a, cj, x, half);
        cell_fluxes(a, x, C2, rC2, FR[s], FL[s], a.visc_type ? L.vis[cj] : T(0));
    }
    const double dt = a.sim->dt;
    const double ff = dt / a.dx;
    T fk = T(ff) * C2;
    if (a.dxl) fk = T(dt / a.dxl[i]) * C2;
    const int lim = a.limiter;
    const long long c = cell_index(a, p, i);
    for (int v = 0; v < 5; ++v) {
        const T gRlo = T(g[v * GH + 2]), gLlo = T(g[v * GH + 3]);
        const T gRhi = T(g[v * GH + 4]), gLhi = T(g[v * GH + 5]);
        auto rowR = [&](int s) -> T {  // raw fluxR of cell i - 2 + s
            const int j = i - 2 + s;
            return j < 0 ? gRlo : (j >= N ? gRhi : FR[s][v]);
        };
        auto rowL = [&](int s) -> T {
            const int j = i - 2 + s;
            return j < 0 ? gLlo : (j >= N ? gLhi : FL[s][v]);
        };
        if (a.init_edges && a.ndims >= 2) {
            fluid_init_face(a.edges, a.eoff, a.modes, a.ndims, a.d, OBJ_FR + v, i, p, a.n0,
                            a.n1, a.n2, double(FR[2][v]));
            fluid_init_face(a.edges, a.eoff, a.modes, a.ndims, a.d, OBJ_FL + v, i, p, a.n0,
                            a.n1, a.n2, double(FL[2][v]));
        }
        // limited fluxR of cells i-1, i and fluxL of cells i, i+1 (the shifted neighbours
        // fade-blended near the ends where the side is a fade one)
        const double* e1R = row_ends(L.en, p, v, 1, 0);
        const double* e1L = row_ends(L.en, p, v, 1, 1);
        auto R2 = [&](int s) -> T {
            const int j = i - 2 + s;
            const T fr = rowR(s);
            const T rm = fade_at(a, true, j, N, e1R, rowR(s - 1));
            const T rp = fade_at(a, false, j, N, e1R, rowR(s + 1));
            return fr + limit(lim, fr - rm, rp - fr);
        };
        auto L2 = [&](int s) -> T {
            const int j = i - 2 + s;
            const T fl = rowL(s);
            const T lm = fade_at(a, true, j, N, e1L, rowL(s - 1));
            const T lp = fade_at(a, false, j, N, e1L, rowL(s + 1));
            return fl + limit(lim, lm - fl, fl - lp);
        };
        const T r0 = R2(2), l0 = L2(2);
        T sR = i > 0 ? R2(1) : T(g[v * GH + 6]);
        T sL = i < N - 1 ? L2(3) : T(g[v * GH + 7]);
        sR = fade_at(a, true, i, N, row_ends(L.en, p, v, 2, 0), sR);
        sL = fade_at(a, false, i, N, row_ends(L.en, p, v, 2, 1), sL);
        const T dq = fk * (r0 - sR + l0 - sL);
        T un = a.q[v][c] - dq;
        if (v == 0) un = fmax(un, T(a.minmass));
        a.q[v][c] = un;
    }
}

// the pencil maxima start at +0 (a fill kernel: the sweep may be captured in a graph)
__global__ void __launch_bounds__(256) long_clear(unsigned long long* p, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

template <typename T, bool MHD>
hipError_t launch_long_t(const LongArgs<T>& L, long long ncell, hipStream_t st) {
    const unsigned nb = unsigned((ncell + 255) / 256);
    const unsigned pb = unsigned((L.a.npencil + 255) / 256);
    const long long nm = 2 * L.a.npencil;
    hipLaunchKernelGGL(long_clear, dim3(unsigned((nm + 255) / 256)), dim3(256), 0, st, L.cmax, nm);
    hipLaunchKernelGGL((long_speeds<T, MHD>), dim3(nb), dim3(256), 0, st, L);
    if (L.a.visc_type) {  // the viscosity of U^n (FluidArrays)
        hipLaunchKernelGGL((long_visc_ghosts<T, MHD>), dim3(pb), dim3(256), 0, st, L, 0,
                           (int)OBJ_PRIM, 0);
        hipLaunchKernelGGL((long_visc_cells<T, MHD>), dim3(nb), dim3(256), 0, st, L, 0);
    }
    hipLaunchKernelGGL((long_ghosts_pred<T, MHD>), dim3(pb), dim3(256), 0, st, L);
    hipLaunchKernelGGL((long_predictor<T, MHD>), dim3(nb), dim3(256), 0, st, L);
    if (L.a.visc_type) {  // the viscosity of U^{n+1/2} (StorageArrays, first assignment)
        hipLaunchKernelGGL((long_visc_ghosts<T, MHD>), dim3(pb), dim3(256), 0, st, L, 1,
                           (int)OBJ_FSTORE, L.a.init_edges ? 1 : 0);
        hipLaunchKernelGGL((long_visc_cells<T, MHD>), dim3(nb), dim3(256), 0, st, L, 1);
    }
    hipLaunchKernelGGL((long_ghosts_corr<T, MHD>), dim3(pb), dim3(256), 0, st, L);
    hipLaunchKernelGGL((long_corrector<T, MHD>), dim3(nb), dim3(256), 0, st, L);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_fluid_long(const FluidArgs<T>& a, void* scratch, long long ncell, bool mhd,
                             hipStream_t st) {
    LongArgs<T> L;
    L.a = a;
    char* s = (char*)scratch;
    for (int v = 0; v < 5; ++v) {
        L.half[v] = (T*)s;
        s += fluid_long_align(sizeof(T) * (size_t)ncell);
    }
    L.cmax = (unsigned long long*)s;
    s += fluid_long_align(sizeof(unsigned long long) * 2 * (size_t)a.npencil);
    L.gh = (double*)s;
    s += fluid_long_align(sizeof(double) * 5 * GH * (size_t)a.npencil);
    L.en = (double*)s;
    s += fluid_long_align(sizeof(double) * 5 * FLUID_LONG_EN * (size_t)a.npencil);
    L.vis = (T*)s;
    s += fluid_long_align(sizeof(T) * (size_t)ncell);
    L.vg = (double*)s;
    return mhd ? launch_long_t<T, true>(L, ncell, st) : launch_long_t<T, false>(L, ncell, st);
}

template hipError_t launch_fluid_long<double>(const FluidArgs<double>&, void*, long long, bool,
                                              hipStream_t);
template hipError_t launch_fluid_long<float>(const FluidArgs<float>&, void*, long long, bool,
                                             hipStream_t);

}  // namespace imh
