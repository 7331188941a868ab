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
                                           int active_lin, int bc_inf, const double* ends, int N,
                                           double* tracker) {
    GhostSpec s;
    s.mode = mode;
    s.upper = upper;
    s.flip_sign = flip_sign;
    s.zero_ghost = zero_ghost;
    s.active_linear = active_lin;
    s.bc_infinity = bc_inf;
    const Ends e = {ends[0], ends[1], ends[2], ends[3], ends[4], ends[5]};
    double g0, g1;
    ghost_values(s, e, N, 1, tracker, g0, g1);
    return g0;
}

// Fade blend of a constant-shifted value (fade_shift.m:31-45) for a cell inside one of
// the fade windows; out of line for the same reason.
__device__ __noinline__ double fluid_fade(double r, int lower, int gi, int N, const double* ends,
                                          const double* c, int f) {
    const Ends e = {ends[0], ends[1], ends[2], ends[3], ends[4], ends[5]};
    // constant-shifted end values: lower shift r[0]=a[0], r[N-1]=a[N-2];
    // upper shift r[0]=a[1], r[N-1]=a[N-1]
    const double v1 = lower ? e.a0 : end_at(e, 1, N);
    const double vN = lower ? end_at(e, N - 2, N) : e.b2;
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

// The sweep of one tile (PX pencils) of one simulation.  Called by the per-simulation
// kernel (one block per tile) and by the batched persistent kernel (a block loops over
// a tile list that mixes simulations of different sizes).
template <typename T, int BS, int KMAX, bool MHD>
__device__ __forceinline__ void fluid_sweep_body(const FluidArgs<T>& a, long long tile) {
    if (!a.force && !a.sim->active) return;
    __shared__ T s_x[2][4][BS];
    __shared__ double s_ends[2][2][PXMAX][6];
    __shared__ T s_red[2][(BS / 64) * PXMAX];
    // edge trackers of the block's pencils: [px][side*16 + role*5 + field], roles lower:
    // store.fluxR, fluxR, fluxL; upper: store.fluxL, fluxR, fluxL
    __shared__ double s_trk[PXMAX][32];

    const int tid = threadIdx.x;
    const int px = a.xmajor ? tid / a.TL : tid % a.PX;
    const int ty = a.xmajor ? tid % a.TL : tid / a.PX;
    const int dstep = a.xmajor ? 1 : a.PX;  // thread distance between neighbouring chunks
    const long long pen0 = tile * a.PX;
    const long long pen = pen0 + px;
    const bool pen_ok = pen < a.npencil;
    const int d = a.d;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int N = sel3i(d, a.n0, a.n1, a.n2);
    const int K = a.K;
    const int i0 = ty * K;
    int kc = N - i0;
    kc = kc < 0 ? 0 : (kc > K ? K : kc);
    if (!pen_ok) kc = 0;
    const bool has_left = ty > 0;
    const bool has_right = i0 + K < N;
    const bool own_first = pen_ok && ty == 0;
    const bool own_last = pen_ok && kc > 0 && i0 + kc == N;

    long long base = 0, stride = 1;
    if (pen_ok) pencil_base(nn, d, pen, base, stride);

    const double dt = a.sim->dt;
    const double ff = dt / a.dx;          // fluxFactor = dTime ./ DGRID{X}
    const double hf = 0.5 * ff;
    const int X = d + 1;                  // momentum field of the sweep direction
    const double g = a.gamma;
    const int f = a.bc_inf;
    const int mlo = mode_at(a.modes, 0, d), mhi = mode_at(a.modes, 1, d);
    const bool trk_lo = (mlo == IMH_BC_TRANS || mlo == IMH_BC_WALL);
    const bool trk_hi = (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL);

    // ---- tracker prefetch (cooperative) ----------------------------------------------
    if ((trk_lo || trk_hi) && !a.init_edges) {
        for (int e = tid; e < a.PX * 30; e += BS) {
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
            long long o = base + (long long)(i0 + k) * stride;
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

    // Ghost-cell spec of one side for an object of field v and flux type ft.
    auto spec = [&](int upper, int v, int ft) __attribute__((always_inline)) {
        GhostSpec s;
        s.mode = upper ? mhi : mlo;
        s.upper = upper;
        s.flip_sign = (v == X);  // component of mom(X) equals the shift direction
        // zero mode: a FLUXL object keeps its lower ghost, FLUXR its upper ghost
        s.zero_ghost = upper ? (ft != FT_FLUXR) : (ft != FT_FLUXL);
        s.active_linear = a.active_lin;
        s.bc_infinity = f;
        return s;
    };
    // ghost of one shift call by one cell, from the staged pencil ends of quantity
    // `which` in buffer `buf`; slot = tracker slot in s_trk
    auto ghost1 = [&](int upper, int v, int ft, int buf, int which, int slot) __attribute__((always_inline)) -> T {
        const GhostSpec sp = spec(upper, v, ft);
        return T(fluid_ghost(sp.mode, sp.upper, sp.flip_sign, sp.zero_ghost, sp.active_linear,
                             sp.bc_infinity, s_ends[buf][which][px], N, &s_trk[px][slot]));
    };

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

    // Stage the chunk boundary values and the pencil ends of two quantities.
    // Register arrays are indexed with compile-time k only (no scratch spills).
    auto stage = [&](int buf, const T (&A)[KMAX], const T (&B)[KMAX]) __attribute__((always_inline)) {
        if (kc > 0) {
            s_x[buf][0][tid] = A[0];
            s_x[buf][2][tid] = B[0];
            T la = A[0], lb = B[0];
#pragma unroll
            for (int k = 1; k < KMAX; ++k)
                if (k == kc - 1) { la = A[k]; lb = B[k]; }
            s_x[buf][1][tid] = la;
            s_x[buf][3][tid] = lb;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                if (k < kc) {
                    int gi = i0 + k;
                    if (gi < 3) {
                        s_ends[buf][0][px][gi] = double(A[k]);
                        s_ends[buf][1][px][gi] = double(B[k]);
                    }
                    if (gi >= N - 3) {
                        s_ends[buf][0][px][3 + gi - (N - 3)] = double(A[k]);
                        s_ends[buf][1][px][3 + gi - (N - 3)] = double(B[k]);
                    }
                }
            }
        }
    };
    auto last_of = [&](const T (&A)[KMAX]) __attribute__((always_inline)) {
        T r = A[0];
#pragma unroll
        for (int k = 1; k < KMAX; ++k)
            if (k == kc - 1) r = A[k];
        return r;
    };

    // Value of shift(X,-1) (lower=true) or shift(X,+1) at chunk cell k (k compile-time).
    //   nb: the neighbouring chunk's boundary value; ghost: the BC ghost.
    auto shifted = [&](bool lower, const T (&A)[KMAX], int k, T nb, T ghost, int mode,
                       int buf, int which) __attribute__((always_inline)) -> T {
        int gi = i0 + k;
        T r;
        if (lower) r = (k > 0) ? A[k > 0 ? k - 1 : 0] : (has_left ? nb : ghost);
        else r = (k < kc - 1) ? A[k + 1 < KMAX ? k + 1 : k] : (has_right ? nb : ghost);
        if (mode == IMH_BC_FADE && (gi < f || gi >= N - f))
            r = T(fluid_fade(double(r), lower, gi, N, s_ends[buf][which][px], a.fade, f));
        return r;
    };

    int buf = 0;
    T P[KMAX], U[KMAX];

    // ================= predictor: W and C from U^n (wFluidFlux.m) ====================
    T lmax = T(NAN);
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        P[k] = T(0);
        U[k] = T(0);
        if (k < kc) {
            double Pd, cd;
            pressure_totsnd(double(q[0][k]), double(q[1][k]), double(q[2][k]), double(q[3][k]),
                            double(q[4][k]), double(bm[0][k]), double(bm[1][k]), double(bm[2][k]),
                            a.use, g, Pd, cd);
            P[k] = T(Pd);
            U[k] = sel3(d, q[1][k], q[2][k], q[3][k]) / q[0][k];
            T s = fabs(U[k]) + T(cd);
            lmax = (k == 0) ? s : fmax(lmax, s);
        }
    }
    const T C = pencil_reduce_max(lmax, a, tid, px, s_red[0]);

    T qh[5][KMAX];
    const T hfC = T(hf) * C;
    auto predictor_field = [&](auto VC) __attribute__((always_inline)) {
        constexpr int v = decltype(VC)::value;
        T FR[KMAX], FL[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            T w;
            const T bX = sel3(d, bm[0][k], bm[1][k], bm[2][k]);
            if constexpr (v == 0) {
                w = sel3(d, q[1][k], q[2][k], q[3][k]) / C;
            } else if constexpr (v == 4) {
                T bdotp = bm[0][k] * q[1][k] + bm[1][k] * q[2][k] + bm[2][k] * q[3][k];
                T we = U[k] * (q[4][k] + P[k]) - bX * bdotp / q[0][k];
                w = we / C;
            } else {
                T dv = (v == X) ? T(1) : T(0);
                w = (U[k] * q[v][k] + P[k] * dv - bX * bm[v - 1][k]) / C;
            }
            FR[k] = T(0.5) * (q[v][k] + w);
            FL[k] = T(0.5) * (q[v][k] - w);
        }
        init_faces(OBJ_SFR + v, FR);
        init_faces(OBJ_SFL + v, FL);
        stage(buf, FR, FL);
        if (a.init_edges) {
            if (own_first) s_trk[px][0 * 16 + 0 * 5 + v] = double(FR[0]);
            if (own_last) s_trk[px][1 * 16 + 0 * 5 + v] = double(last_of(FL));
        }
        __syncthreads();
        T gR = T(0), gL = T(0);
        if (own_first) gR = ghost1(0, v, FT_FLUXR, buf, 0, 0 * 16 + 0 * 5 + v);
        if (own_last) gL = ghost1(1, v, FT_FLUXL, buf, 1, 1 * 16 + 0 * 5 + v);
        const T nbR = has_left ? s_x[buf][1][tid - dstep] : T(0);
        const T nbL = has_right ? s_x[buf][2][tid + dstep] : T(0);
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            qh[v][k] = T(1);
            if (k < kc) {
                T sR = shifted(true, FR, k, nbR, gR, mlo, buf, 0);
                T sL = shifted(false, FL, k, nbL, gL, mhi, buf, 1);
                qh[v][k] = q[v][k] - hfC * (FR[k] - sR + FL[k] - sL);
            }
        }
        buf ^= 1;
    };
    predictor_field(FieldIdx<4>{});
    predictor_field(FieldIdx<1>{});
    predictor_field(FieldIdx<2>{});
    predictor_field(FieldIdx<3>{});
    predictor_field(FieldIdx<0>{});
    // mass.store.array = max(mass.store.array, MINMASS)
#pragma unroll
    for (int k = 0; k < KMAX; ++k) qh[0][k] = fmax(qh[0][k], T(a.minmass));

    // ================= corrector: W' and C' from U^(n+1/2) ===========================
    lmax = T(NAN);
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            double Pd, cd;
            pressure_totsnd(double(qh[0][k]), double(qh[1][k]), double(qh[2][k]),
                            double(qh[3][k]), double(qh[4][k]), double(bm[0][k]),
                            double(bm[1][k]), double(bm[2][k]), a.use, g, Pd, cd);
            P[k] = T(Pd);
            U[k] = sel3(d, qh[1][k], qh[2][k], qh[3][k]) / qh[0][k];
            T s = fabs(U[k]) + T(cd);
            lmax = (k == 0) ? s : fmax(lmax, s);
        }
    }
    const T C2 = pencil_reduce_max(lmax, a, tid, px, s_red[1]);
    const T ffC = T(ff) * C2;
    const int lim = a.limiter;

    auto corrector_field = [&](auto VC) __attribute__((always_inline)) {
        constexpr int v = decltype(VC)::value;
        T FR[KMAX], FL[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            T w;
            const T bX = sel3(d, bm[0][k], bm[1][k], bm[2][k]);
            if constexpr (v == 0) {
                w = sel3(d, qh[1][k], qh[2][k], qh[3][k]) / C2;
            } else if constexpr (v == 4) {
                T bdotp = bm[0][k] * qh[1][k] + bm[1][k] * qh[2][k] + bm[2][k] * qh[3][k];
                T we = U[k] * (qh[4][k] + P[k]) - bX * bdotp / qh[0][k];
                w = we / C2;
            } else {
                T dv = (v == X) ? T(1) : T(0);
                w = (U[k] * qh[v][k] + P[k] * dv - bX * bm[v - 1][k]) / C2;
            }
            FR[k] = T(0.5) * (qh[v][k] + w);
            FL[k] = T(0.5) * (qh[v][k] - w);
        }
        init_faces(OBJ_FR + v, FR);
        init_faces(OBJ_FL + v, FL);
        stage(buf, FR, FL);
        if (a.init_edges) {
            if (own_first) {
                s_trk[px][0 * 16 + 1 * 5 + v] = double(FR[0]);
                s_trk[px][0 * 16 + 2 * 5 + v] = double(FL[0]);
            }
            if (own_last) {
                s_trk[px][1 * 16 + 1 * 5 + v] = double(last_of(FR));
                s_trk[px][1 * 16 + 2 * 5 + v] = double(last_of(FL));
            }
        }
        __syncthreads();
        T FR2[KMAX], FL2[KMAX];
        {
            // ghosts: fluxR lower (1st nudge), fluxR upper, fluxL lower, fluxL upper (1st)
            T gRlo = T(0), gRhi = T(0), gLlo = T(0), gLhi = T(0);
            if (own_first) {
                gRlo = ghost1(0, v, FT_FLUXR, buf, 0, 0 * 16 + 1 * 5 + v);
                gLlo = ghost1(0, v, FT_FLUXL, buf, 1, 0 * 16 + 2 * 5 + v);
            }
            if (own_last) {
                gRhi = ghost1(1, v, FT_FLUXR, buf, 0, 1 * 16 + 1 * 5 + v);
                gLhi = ghost1(1, v, FT_FLUXL, buf, 1, 1 * 16 + 2 * 5 + v);
            }
            const T nbRl = has_left ? s_x[buf][1][tid - dstep] : T(0);   // FR of cell i0-1
            const T nbRr = has_right ? s_x[buf][0][tid + dstep] : T(0);  // FR of cell i0+kc
            const T nbLl = has_left ? s_x[buf][3][tid - dstep] : T(0);
            const T nbLr = has_right ? s_x[buf][2][tid + dstep] : T(0);
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                FR2[k] = FL2[k] = T(0);
                if (k < kc) {
                    T rm = shifted(true, FR, k, nbRl, gRlo, mlo, buf, 0);
                    T rp = shifted(false, FR, k, nbRr, gRhi, mhi, buf, 0);
                    T lm = shifted(true, FL, k, nbLl, gLlo, mlo, buf, 1);
                    T lp = shifted(false, FL, k, nbLr, gLhi, mhi, buf, 1);
                    FR2[k] = FR[k] + lim_delta(lim, FR[k] - rm, rp - FR[k]);
                    FL2[k] = FL[k] + lim_delta(lim, lm - FL[k], FL[k] - lp);
                }
            }
        }
        buf ^= 1;
        stage(buf, FR2, FL2);
        __syncthreads();
        {
            T g2R = T(0), g2L = T(0);
            if (own_first) g2R = ghost1(0, v, FT_FLUXR, buf, 0, 0 * 16 + 1 * 5 + v);
            if (own_last) g2L = ghost1(1, v, FT_FLUXL, buf, 1, 1 * 16 + 2 * 5 + v);
            const T nb2R = has_left ? s_x[buf][1][tid - dstep] : T(0);
            const T nb2L = has_right ? s_x[buf][2][tid + dstep] : T(0);
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                if (k < kc) {
                    T sR = shifted(true, FR2, k, nb2R, g2R, mlo, buf, 0);
                    T sL = shifted(false, FL2, k, nb2L, g2L, mhi, buf, 1);
                    const long long o = base + (long long)(i0 + k) * stride;
                    // U^n is re-read here (it is L2-resident) rather than kept live
                    T qn = a.q[v][o] - ffC * (FR2[k] - sR + FL2[k] - sL);
                    if (v == 0) qn = fmax(qn, T(a.minmass));  // mass.array = max(mass, MINMASS)
                    a.q[v][o] = qn;
                }
            }
        }
        buf ^= 1;
    };
    corrector_field(FieldIdx<4>{});
    corrector_field(FieldIdx<1>{});
    corrector_field(FieldIdx<2>{});
    corrector_field(FieldIdx<3>{});
    corrector_field(FieldIdx<0>{});

    // ---- write the edge trackers back -----------------------------------------------
    if (trk_lo || trk_hi) {
        __syncthreads();
        for (int e = tid; e < a.PX * 30; e += BS) {
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
