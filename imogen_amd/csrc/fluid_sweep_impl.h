// fluid_sweep_impl.h -- one fused kernel per directional fluid sweep (relaxingFluid.m).
//
// A workgroup of BS threads owns PX whole pencils along the sweep dimension; each pencil
// is split over TL = BS/PX threads and each thread keeps a contiguous chunk of K <= KMAX
// cells of all conserved fields in registers.  One HBM read of the 5 fluid fields (+3
// cell-centred B for MHD) and one write of the 5 updated fields per cell per sweep.
//
// Inside the kernel, per pencil (fluid/relaxingFluid.m:16-73, fluid/wFluidFlux.m:14-44):
//   1. pressure 'totsnd' and u = p_X / rho per cell; C = max(|u| + c_f) over the pencil
//      (wavefront shuffles + one LDS pass);
//   2. predictor, field by field: FR/FL = (q +- w)/2, neighbour values across chunk
//      boundaries through LDS, boundary ghosts from the BC rule (with the per-object
//      transparent edge trackers), q_half = q - (dt/dx/2) C (FR - FR(-1) + FL - FL(+1));
//   3. pressure and C' from q_half; corrector with the limiter; q_new stored.
// Evaluation order of every expression follows the .m text, so the FP64 instantiation
// reproduces the CPU restatement of the reference when built without FP contraction.
#include "imh_common.h"
#include "imh_kernels.h"

#include <type_traits>

namespace imh {

namespace {

constexpr int PXMAX = 32;

// Compile-time field index (0 mass, 1..3 momentum, 4 energy).  The fields are processed
// energy first and mass last; each field's arrays are independent, so the order does
// not change any value.
template <int V>
using FieldIdx = std::integral_constant<int, V>;

// Register-array element chosen by a run-time index in 0..2 (selects, not indexing, so
// the arrays stay in registers).
template <typename T>
__device__ inline T sel3(int i, T a, T b, T c) {
    return i == 0 ? a : (i == 1 ? b : c);
}

// Block index -> tile index so that consecutive tiles run on the same XCD (blocks are
// dispatched round-robin over the 8 XCDs; neighbouring pencils then share an L2).
__device__ inline long long xcd_tile(long long b, long long nblocks) {
    long long q = nblocks / 8, r = nblocks % 8;
    long long x = b % 8, t = b / 8;
    return x * q + (x < r ? x : r) + t;
}

// Ghost value of one shift call by one cell at a pencil end (the BC rule of that flux
// object, with its tracker).  Out of line: only pencil-end threads need it, and keeping
// the BC code out of every call site keeps the kernel inside the instruction cache.
__device__ __noinline__ double fluid_ghost(int mode, int upper, int flip_sign, int zero_ghost,
                                           int active_lin, int bc_inf, double a0, double a1,
                                           double a2, double b0, double b1, double b2, int N,
                                           double* tracker) {
    GhostSpec s;
    s.mode = mode;
    s.upper = upper;
    s.flip_sign = flip_sign;
    s.zero_ghost = zero_ghost;
    s.active_linear = active_lin;
    s.bc_infinity = bc_inf;
    const Ends e = {a0, a1, a2, b0, b1, b2};
    double g0, g1;
    ghost_values(s, e, N, 1, tracker, g0, g1);
    return g0;
}

// Fade blend (fade_shift.m:31-45) of a constant-shifted value r at cell gi inside one of
// the fade windows; v1 / vN are the shifted array's first and last values.  Out of line
// for the same reason.
__device__ __noinline__ double fluid_fade(double r, int gi, int N, double v1, double vN,
                                          const double* c, int f) {
    return fade_blend(r, gi, N, v1, vN, c, f);
}

// First assignment of a flux object (InitializedArray.m:43-47): cell gi of pencil fi
// along d records its value in the object's edge stores of the other dimensions when it
// lies on one of their faces.  Runs in a simulation's first sweep only.
__device__ __noinline__ void fluid_init_face(double* edges, const long long* eoff, int modes,
                                             int ndims, int d, int obj, int gi, long long fi,
                                             int n0, int n1, int n2, double val) {
    const int nn[3] = {n0, n1, n2};
    int c0, c1, c2;
    if (d == 0) { c0 = gi; c1 = int(fi % n1); c2 = int(fi / n1); }
    else if (d == 1) { c0 = int(fi % n0); c1 = gi; c2 = int(fi / n0); }
    else { c0 = int(fi % n0); c1 = int(fi / n0); c2 = gi; }
    for (int dd = 0; dd < ndims; ++dd) {
        if (dd == d) continue;
        for (int s = 0; s < 2; ++s) {
            int mode = mode_at(modes, s, dd);
            if (mode != IMH_BC_TRANS && mode != IMH_BC_WALL) continue;
            const int cdd = sel3i(dd, c0, c1, c2);
            if (cdd != (s ? sel3i(dd, n0, n1, n2) - 1 : 0)) continue;  // off this face
            edges[eoff[(obj * 2 + s) * 3 + dd] + face_index(nn, dd, c0, c1, c2)] = val;
        }
    }
}

template <typename T>
__device__ inline T pencil_reduce_max(T v, const FluidArgs<T>& a, int tid, int px, T* s_red) {
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (a.xmajor) {
        int g = a.TL < 64 ? a.TL : 64;
        for (int off = g >> 1; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    } else {
        for (int off = a.PX; off < 64; off <<= 1) v = fmax(v, __shfl_xor(v, off, 64));
    }
    (void)lane;
    s_red[wave * PXMAX + px] = v;
    __syncthreads();
    T m = v;
    const int nw = blockDim.x >> 6;
    if (a.xmajor) {
        int w0 = (px * a.TL) >> 6, w1 = (px * a.TL + a.TL - 1) >> 6;
        for (int w = w0; w <= w1; ++w) m = fmax(m, s_red[w * PXMAX + px]);
    } else {
        for (int w = 0; w < nw; ++w) m = fmax(m, s_red[w * PXMAX + px]);
    }
    return m;
}

}  // namespace

// Arithmetic of the two precisions.  FP64 keeps every division and square root of the
// .m text, in the same order, so it reproduces the CPU restatement; the FP32 bulk path
// divides by multiplying with the hardware reciprocal and takes the FP32 square root,
// while the pressure floor (E - KE, max(P, 0)) stays FP64.
template <typename T>
struct FMath;

template <>
struct FMath<double> {
    static __device__ __forceinline__ double div(double x, double y, double) { return x / y; }
    static __device__ __forceinline__ double rcp(double y) { return y; }  // divisor kept
    // pressure.m 'totsnd': P and the fast speed c
    static __device__ __forceinline__ void press(double rho, double px, double py, double pz,
                                                 double E, double bx, double by, double bz,
                                                 MagUse u, double g, double& P, double& c,
                                                 double& inv) {
        pressure_totsnd(rho, px, py, pz, E, bx, by, bz, u, g, P, c);
        inv = rho;
    }
    static __device__ __forceinline__ double divrho(double x, double rho_or_inv) {
        return x / rho_or_inv;
    }
    static __device__ __forceinline__ double vanleer(double dl, double dr) {
        double t = fmax(dl * dr, 0.0) / (dl + dr);
        return isnan(t) ? 0.0 : t;
    }
};

template <>
struct FMath<float> {
    static __device__ __forceinline__ float div(float x, float, float r) { return x * r; }
    static __device__ __forceinline__ float rcp(float y) { return __builtin_amdgcn_rcpf(y); }
    static __device__ __forceinline__ void press(float rho, float px, float py, float pz, float E,
                                                 float bx, float by, float bz, MagUse u, double g,
                                                 float& P, float& c, float& inv) {
        inv = __builtin_amdgcn_rcpf(rho);
        const float v2 = (px * px + py * py + pz * pz) * inv * inv;
        float b2 = 0.0f;
        if (u.x) b2 += bx * bx;
        if (u.y) b2 += by * by;
        if (u.z) b2 += bz * bz;
        const float gf = float(g);
        // positivity-critical difference and floor in FP64
        const double r = (g - 1.0) * (double(E) - 0.5 * double(rho * v2));
        const float rf = float(r);
        c = __builtin_amdgcn_sqrtf(fabsf((gf * (rf - (gf - 1.0f) * 0.5f * b2) + 2.0f * b2) * inv));
        const double Pd = r + 0.5 * (2.0 - g) * double(b2);
        P = float(Pd < 0.0 ? 0.0 : Pd);
    }
    static __device__ __forceinline__ float divrho(float x, float inv) { return x * inv; }
    static __device__ __forceinline__ float vanleer(float dl, float dr) {
        float t = fmaxf(dl * dr, 0.0f) * __builtin_amdgcn_rcpf(dl + dr);
        return isnan(t) ? 0.0f : t;
    }
};

template <typename T>
__device__ __forceinline__ T limit(int lim, T dl, T dr) {
    if (lim == IMH_LIM_VANLEER) return FMath<T>::vanleer(dl, dr);
    if (lim == IMH_LIM_SUPERBEE) {
        T cf = (fabs(dl) < fabs(dr)) ? T(1) : T(0);
        return lim_minmod((T(1) + cf) * dl, (T(2) - cf) * dr);
    }
    return lim_minmod(dl, dr);
}

// Fields handled per LDS pass: the FR / FL rows of G fields of the block's pencils live in
// LDS at once (at most about 48 KiB of rows).
template <typename T, int BS, int KMAX>
constexpr int fluid_group() {
    constexpr int per_field = 2 * BS * KMAX * int(sizeof(T));
    constexpr int g = (48 * 1024) / per_field;
    return g < 1 ? 1 : (g > 5 ? 5 : g);
}

// The sweep of one tile (PX pencils) of one simulation.
//
// Cells are held in registers (K per thread).  Each phase of relaxingFluid.m works on a
// group of G fields: the threads write that phase's FR / FL of their cells into LDS rows,
// barrier; one lane per (pencil, field, shift call) evaluates the pencil-end ghost values
// (BC rule + edge tracker) in parallel, barrier; every thread reads its cells' neighbours
// from the rows (or the ghosts) and updates its registers.  The row layout puts cell k of
// thread ty at (k*TL + ty), so consecutive lanes touch consecutive LDS words.
template <typename T, int BS, int KMAX, bool MHD>
__device__ __forceinline__ void fluid_sweep_body(const FluidArgs<T>& ap, long long tile) {
    const FluidArgs<T> a = ap;  // private copy: stores to the fields cannot alias it
    if (!a.force && !a.sim->active) return;
    using M = FMath<T>;
    constexpr int G = fluid_group<T, BS, KMAX>();
    constexpr int CELLS = BS * KMAX;
    __shared__ T s_row[2][G][CELLS];        // FR (0) and FL (1) rows of G fields
    __shared__ double s_gh[PXMAX][5][4];    // ghosts: FR lower, FL lower, FR upper, FL upper
    __shared__ T s_red[2][(BS / 64) * PXMAX];
    // edge trackers of the block's pencils: [px][side*16 + role*5 + field], roles lower:
    // store.fluxR, fluxR, fluxL; upper: store.fluxL, fluxR, fluxL
    __shared__ double s_trk[PXMAX][32];

    const int tid = threadIdx.x;
    const int PX = a.PX, TL = a.TL;
    const int px = a.xmajor ? tid / TL : tid % PX;
    const int ty = a.xmajor ? tid % TL : tid / PX;
    const long long pen0 = tile * PX;
    const long long pen = pen0 + px;
    const bool pen_ok = pen < a.npencil;
    const int d = a.d;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int N = sel3i(d, a.n0, a.n1, a.n2);
    const int K = a.K;
    const int L = K * TL;                   // row slots per pencil
    const int i0 = ty * K;
    int kc = N - i0;
    kc = kc < 0 ? 0 : (kc > K ? K : kc);
    if (!pen_ok) kc = 0;

    // LDS index of cell i of pencil p
    auto idx = [&](int p, int i) __attribute__((always_inline)) -> int {
        const int slot = (i % K) * TL + i / K;
        return a.xmajor ? p * L + slot : slot * PX + p;
    };
    auto at_slot = [&](int slot) __attribute__((always_inline)) -> int {
        return a.xmajor ? px * L + slot : slot * PX + px;
    };
    auto my_idx = [&](int k) __attribute__((always_inline)) -> int { return at_slot(k * TL + ty); };

    // 32-bit cell offsets (a simulation holds fewer than 2^31 cells): one register each
    long long base64 = 0, stride64 = 1;
    if (pen_ok) pencil_base(nn, d, pen, base64, stride64);
    const int base = int(base64), stride = int(stride64);

    const double dt = a.sim->dt;
    const double ff = dt / a.dx;          // fluxFactor = dTime ./ DGRID{X}
    const double hf = 0.5 * ff;
    const int X = d + 1;                  // momentum field of the sweep direction
    const double g = a.gamma;
    const int f = a.bc_inf;
    const int mlo = mode_at(a.modes, 0, d), mhi = mode_at(a.modes, 1, d);
    const bool trk_lo = (mlo == IMH_BC_TRANS || mlo == IMH_BC_WALL);
    const bool trk_hi = (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL);
    const bool fade = (mlo == IMH_BC_FADE || mhi == IMH_BC_FADE);

    // ---- tracker prefetch (cooperative) ----------------------------------------------
    if ((trk_lo || trk_hi) && !a.init_edges) {
        for (int e = tid; e < PX * 30; e += BS) {
            int p = e / 30, r = e % 30;
            int side = r / 15, role = (r % 15) / 5, v = r % 5;
            long long pp = pen0 + p;
            if (pp >= a.npencil || (side ? !trk_hi : !trk_lo)) continue;
            int obj = side ? (role == 0 ? OBJ_SFL : (role == 1 ? OBJ_FR : OBJ_FL))
                           : (role == 0 ? OBJ_SFR : (role == 1 ? OBJ_FR : OBJ_FL));
            s_trk[p][side * 16 + role * 5 + v] = a.edges[a.eoff[((obj + v) * 2 + side) * 3 + d] + pp];
        }
    }

    // ---- load the pencil chunk ------------------------------------------------------
    T q[5][KMAX];
    T bm[3][KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            const int o = base + (i0 + k) * stride;
#pragma unroll
            for (int v = 0; v < 5; ++v) q[v][k] = a.q[v][o];
#pragma unroll
            for (int v = 0; v < 3; ++v) bm[v][k] = MHD ? a.bc[v][o] : T(0);
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) q[v][k] = T(1);
#pragma unroll
            for (int v = 0; v < 3; ++v) bm[v][k] = T(0);
        }
    }

    // First assignment of a flux object (InitializedArray.m:43-47): record its edges on
    // the faces of the other dimensions.
    auto init_faces = [&](int obj, const T (&vals)[KMAX]) __attribute__((always_inline)) {
        if (!a.init_edges || kc == 0 || a.ndims < 2) return;
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            if (k < kc)
                fluid_init_face(a.edges, a.eoff, a.modes, a.ndims, d, obj, i0 + k, pen, a.n0,
                                a.n1, a.n2, double(vals[k]));
    };
    auto put = [&](int which, int gv, const T (&A)[KMAX]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            if (k < kc) s_row[which][gv][my_idx(k)] = A[k];
    };

    // Ghost values of one phase for the fields [v0, v0 + ng): kind 0 predictor
    // (store.fluxR lower, store.fluxL upper), 1 corrector (fluxR and fluxL, both sides),
    // 2 final (fluxR lower, fluxL upper).  One lane per (pencil, field, shift call).
    auto ghost_phase = [&](int kind, int v0, int ng) __attribute__((always_inline)) {
        const int J = kind == 1 ? 4 : 2;
        for (int job = tid; job < PX * ng * J; job += BS) {
            const int p = job / (ng * J);
            const int r = job % (ng * J);
            const int gv = r / J;
            const int v = v0 + gv;
            const int j = (J == 4) ? (r % J) : ((r % J) ? 3 : 0);  // 0 FR lo, 1 FL lo, 2 FR hi, 3 FL hi
            if (pen0 + p >= a.npencil) continue;
            const int which = j & 1;           // 0 fluxR, 1 fluxL
            const int upper = j >> 1;
            const int role = kind == 0 ? 0 : 1 + which;
            const int slot = upper * 16 + role * 5 + v;
            const T* row = s_row[which][gv];
            const double e0 = double(row[idx(p, 0)]), e1 = double(row[idx(p, N > 1 ? 1 : 0)]);
            const double e2 = double(row[idx(p, N > 2 ? 2 : N - 1)]);
            const double e3 = double(row[idx(p, N >= 3 ? N - 3 : 0)]);
            const double e4 = double(row[idx(p, N >= 2 ? N - 2 : 0)]);
            const double e5 = double(row[idx(p, N - 1)]);
            if (a.init_edges && kind < 2) s_trk[p][slot] = upper ? e5 : e0;
            const int mode = upper ? mhi : mlo;
            double gval;
            if (mode == IMH_BC_CIRC) gval = upper ? e0 : e5;
            else if (mode == IMH_BC_CONST || mode == IMH_BC_FADE) gval = upper ? e5 : e0;
            else {
                const int ft = which ? FT_FLUXL : FT_FLUXR;
                const int zero_ghost = upper ? (ft != FT_FLUXR) : (ft != FT_FLUXL);
                gval = fluid_ghost(mode, upper, v == X, zero_ghost, a.active_lin, f, e0, e1, e2,
                                   e3, e4, e5, N, &s_trk[p][slot]);
            }
            s_gh[p][v][j] = gval;
        }
    };

    // shift(X,-1) (lower) / shift(X,+1) of row (which, gv) of field v at cell i0+k
    auto shifted = [&](bool lower, int which, int gv, int v, int k) __attribute__((always_inline)) -> T {
        const int i = i0 + k;
        const T* row = s_row[which][gv];
        T r;
        // the neighbour is this thread's cell k -+ 1, or the last / first cell of the
        // neighbouring thread's chunk
        if (lower) r = (i > 0) ? row[at_slot(k > 0 ? (k - 1) * TL + ty : (K - 1) * TL + ty - 1)]
                               : T(s_gh[px][v][which]);
        else r = (i < N - 1) ? row[at_slot(k + 1 < K ? (k + 1) * TL + ty : ty + 1)]
                             : T(s_gh[px][v][2 + which]);
        if (fade) {
            const int mode = lower ? mlo : mhi;
            if (mode == IMH_BC_FADE && (i < f || i >= N - f)) {
                // constant-shifted end values: lower r[0]=a[0], r[N-1]=a[N-2];
                // upper r[0]=a[1], r[N-1]=a[N-1]
                const double v1 = double(row[idx(px, lower ? 0 : 1)]);
                const double vN = double(row[idx(px, lower ? N - 2 : N - 1)]);
                r = T(fluid_fade(double(r), i, N, v1, vN, a.fade, f));
            }
        }
        return r;
    };

    T P[KMAX], U[KMAX], IR[KMAX];
    // pressure, velocity and the freezing-speed candidate of every cell of state Q
    auto cell_speeds = [&](const T (&Q)[5][KMAX]) __attribute__((always_inline)) -> T {
        T lmax = T(NAN);
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            P[k] = T(0);
            U[k] = T(0);
            IR[k] = T(1);
            if (k < kc) {
                T Pk, ck, ir;
                M::press(Q[0][k], Q[1][k], Q[2][k], Q[3][k], Q[4][k], bm[0][k], bm[1][k],
                         bm[2][k], a.use, g, Pk, ck, ir);
                P[k] = Pk;
                IR[k] = ir;
                U[k] = M::divrho(sel3(d, Q[1][k], Q[2][k], Q[3][k]), ir);
                T sp = fabs(U[k]) + ck;
                lmax = (k == 0) ? sp : fmax(lmax, sp);
            }
        }
        return lmax;
    };

    // wFluidFlux.m: FR / FL of field v from state Q, freezing speed C
    auto fluxes = [&](auto VC, const T (&Q)[5][KMAX], T C, T rC, T (&FR)[KMAX], T (&FL)[KMAX])
        __attribute__((always_inline)) {
        constexpr int v = decltype(VC)::value;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const T bX = sel3(d, bm[0][k], bm[1][k], bm[2][k]);
            T w;
            if constexpr (v == 0) {
                w = M::div(sel3(d, Q[1][k], Q[2][k], Q[3][k]), C, rC);
            } else if constexpr (v == 4) {
                T bdotp = bm[0][k] * Q[1][k] + bm[1][k] * Q[2][k] + bm[2][k] * Q[3][k];
                T we = U[k] * (Q[4][k] + P[k]) - M::divrho(bX * bdotp, IR[k]);
                w = M::div(we, C, rC);
            } else {
                T dv = (v == X) ? T(1) : T(0);
                w = M::div(U[k] * Q[v][k] + P[k] * dv - bX * bm[v - 1][k], C, rC);
            }
            FR[k] = T(0.5) * (Q[v][k] + w);
            FL[k] = T(0.5) * (Q[v][k] - w);
        }
    };
    // run fn(FieldIdx<v>, gv) for the fields of group [v0, v0 + G)
    auto for_group = [&](int v0, auto fn) __attribute__((always_inline)) {
        auto one = [&](auto VC) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            if (v >= v0 && v < v0 + G) fn(VC, v - v0);
        };
        one(FieldIdx<0>{});
        one(FieldIdx<1>{});
        one(FieldIdx<2>{});
        one(FieldIdx<3>{});
        one(FieldIdx<4>{});
    };

    // ================= predictor: W and C from U^n (wFluidFlux.m) ====================
    const T C = pencil_reduce_max(cell_speeds(q), a, tid, px, s_red[0]);
    const T rC = M::rcp(C);
    const T hfC = T(hf) * C;
    T qh[5][KMAX];
#pragma unroll
    for (int v0 = 0; v0 < 5; v0 += G) {
        const int ng = (5 - v0) < G ? (5 - v0) : G;
        if (v0 > 0) __syncthreads();  // the previous group's rows have been read
        for_group(v0, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            T FR[KMAX], FL[KMAX];
            fluxes(VC, q, C, rC, FR, FL);
            init_faces(OBJ_SFR + v, FR);
            init_faces(OBJ_SFL + v, FL);
            put(0, gv, FR);
            put(1, gv, FL);
        });
        __syncthreads();
        ghost_phase(0, v0, ng);
        __syncthreads();
        for_group(v0, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                qh[v][k] = T(1);
                if (k < kc) {
                    const T fr = s_row[0][gv][my_idx(k)], fl = s_row[1][gv][my_idx(k)];
                    const T sR = shifted(true, 0, gv, v, k);
                    const T sL = shifted(false, 1, gv, v, k);
                    qh[v][k] = q[v][k] - hfC * (fr - sR + fl - sL);
                }
            }
        });
    }
    // mass.store.array = max(mass.store.array, MINMASS)
#pragma unroll
    for (int k = 0; k < KMAX; ++k) qh[0][k] = fmax(qh[0][k], T(a.minmass));

    // ================= corrector: W' and C' from U^(n+1/2) ===========================
    const T C2 = pencil_reduce_max(cell_speeds(qh), a, tid, px, s_red[1]);
    const T rC2 = M::rcp(C2);
    const T ffC = T(ff) * C2;
    const int lim = a.limiter;
#pragma unroll
    for (int v0 = 0; v0 < 5; v0 += G) {
        const int ng = (5 - v0) < G ? (5 - v0) : G;
        if (v0 > 0) __syncthreads();
        for_group(v0, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            T FR[KMAX], FL[KMAX];
            fluxes(VC, qh, C2, rC2, FR, FL);
            init_faces(OBJ_FR + v, FR);
            init_faces(OBJ_FL + v, FL);
            put(0, gv, FR);
            put(1, gv, FL);
        });
        __syncthreads();
        ghost_phase(1, v0, ng);
        __syncthreads();
        T FR2[G][KMAX], FL2[G][KMAX];
        for_group(v0, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                FR2[gv][k] = FL2[gv][k] = T(0);
                if (k < kc) {
                    const T fr = s_row[0][gv][my_idx(k)], fl = s_row[1][gv][my_idx(k)];
                    const T rm = shifted(true, 0, gv, v, k);
                    const T rp = shifted(false, 0, gv, v, k);
                    const T lm = shifted(true, 1, gv, v, k);
                    const T lp = shifted(false, 1, gv, v, k);
                    FR2[gv][k] = fr + limit(lim, fr - rm, rp - fr);
                    FL2[gv][k] = fl + limit(lim, lm - fl, fl - lp);
                }
            }
        });
        __syncthreads();  // every thread has read the FR / FL rows
        for_group(v0, [&](auto VC, int gv) __attribute__((always_inline)) {
            put(0, gv, FR2[gv]);
            put(1, gv, FL2[gv]);
        });
        __syncthreads();
        ghost_phase(2, v0, ng);
        __syncthreads();
        for_group(v0, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                if (k < kc) {
                    const T sR = shifted(true, 0, gv, v, k);
                    const T sL = shifted(false, 1, gv, v, k);
                    const int o = base + (i0 + k) * stride;
                    // U^n is re-read here (it is cache-resident) rather than kept live
                    T qn = a.q[v][o] - ffC * (FR2[gv][k] - sR + FL2[gv][k] - sL);
                    if (v == 0) qn = fmax(qn, T(a.minmass));  // mass.array = max(mass, MINMASS)
                    a.q[v][o] = qn;
                }
            }
        });
    }

    // ---- write the edge trackers back -----------------------------------------------
    if (trk_lo || trk_hi) {
        __syncthreads();
        for (int e = tid; e < PX * 30; e += BS) {
            int p = e / 30, r = e % 30;
            int side = r / 15, role = (r % 15) / 5, v = r % 5;
            long long pp = pen0 + p;
            if (pp >= a.npencil || (side ? !trk_hi : !trk_lo)) continue;
            int obj = side ? (role == 0 ? OBJ_SFL : (role == 1 ? OBJ_FR : OBJ_FL))
                           : (role == 0 ? OBJ_SFR : (role == 1 ? OBJ_FR : OBJ_FL));
            a.edges[a.eoff[((obj + v) * 2 + side) * 3 + d] + pp] = s_trk[p][side * 16 + role * 5 + v];
        }
    }
}

// Persistent batched sweep: the grid is sized to the device and every block walks the
// tile list (largest tiles first), so hundreds of uneven simulations share one launch.
// With tiles == nullptr it is the sweep of the single simulation args[0], one block per
// tile, with an XCD-aware block-to-tile map (neighbouring pencils share an L2).
template <typename T, int BS, int KMAX, bool MHD>
__global__ void __launch_bounds__(BS) fluid_sweep_batch_kernel(const FluidArgs<T>* args,
                                                               const BatchTile* tiles,
                                                               int ntiles) {
    // one call site of the body (it is large)
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int arg = 0;
        long long tile;
        if (tiles) {
            const BatchTile bt = tiles[t];
            arg = bt.arg;
            tile = bt.tile;
        } else {
            tile = args[0].xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (long long)blockIdx.x;
        }
        fluid_sweep_body<T, BS, KMAX, MHD>(args[arg], tile);
        if (!tiles) break;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// Launcher of this translation unit's variant.  fluid_sweep_variant.hip is compiled once
// per entry of IMH_FLUID_VARIANTS (imh_kernels.h) with IMH_FLUID_ESZ / IMH_FLUID_BS /
// IMH_FLUID_KMAX defined, so the variants build in parallel; launch_fluid_batch
// (fluid_dispatch.cpp) picks one per sweep dimension.
// ------------------------------------------------------------------------------------
#if IMH_FLUID_ESZ == 8
using FluidT = double;
#else
using FluidT = float;
#endif

hipError_t IMH_FLUID_LAUNCHER(IMH_FLUID_ESZ, IMH_FLUID_BS, IMH_FLUID_KMAX)(
    const void* args, const BatchTile* tiles, int ntiles, int nblocks, bool mhd, hipStream_t st) {
    const FluidArgs<FluidT>* a = (const FluidArgs<FluidT>*)args;
    if (mhd)
        hipLaunchKernelGGL((fluid_sweep_batch_kernel<FluidT, IMH_FLUID_BS, IMH_FLUID_KMAX, true>),
                           dim3(nblocks), dim3(IMH_FLUID_BS), 0, st, a, tiles, ntiles);
    else
        hipLaunchKernelGGL((fluid_sweep_batch_kernel<FluidT, IMH_FLUID_BS, IMH_FLUID_KMAX, false>),
                           dim3(nblocks), dim3(IMH_FLUID_BS), 0, st, a, tiles, ntiles);
    return hipGetLastError();
}

}  // namespace imh
