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
#include "fluid_math.h"
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

// Caramana-Shashkov-Whalen past a transparent or wall end (side `up` of a pencil): the
// side's 15 shift calls in the .m text's order (ArtificialViscosity.m:80-148, 155-190),
// each a ghost of its object (rows 0 mass, 1..3 momentum, 4 energy, 5..7 cell-centred B;
// ends e[row * 6 + 0..5] = a[0..2], a[N-3..N-1]) with that object's tracker *tp[row];
// out = massB/C, sndB/C, then v0 (two cells), v1 on the lower side or v2, v3 (two cells)
// on the upper.  Out of line: it runs on one lane per pencil.
__device__ __noinline__ void csw_side_ghosts(int mode, int up, int d, int N, int active_lin,
                                             int f, double G, double* const tp[8],
                                             const double* e, double* out) {
    double tv[8];
    for (int r = 0; r < 8; ++r) tv[r] = *tp[r];
    auto call = [&](int r, int m, double& g0, double& g1) {
        GhostSpec sp;
        sp.mode = mode;
        sp.upper = up;
        sp.flip_sign = (r == 1 + d) || (r == 5 + d);
        sp.zero_ghost = 0;
        sp.active_linear = active_lin;
        sp.bc_infinity = f;
        const double* x = e + r * 6;
        const Ends en = {x[0], x[1], x[2], x[3], x[4], x[5]};
        if (m == 2)
            ghost_values<2>(sp, en, N, 2, &tv[r], g0, g1);
        else
            ghost_values<1>(sp, en, N, 1, &tv[r], g0, g1);
    };
    double g0, g1, u0, u1;
    call(0, 1, g0, g1);  // massB / massC
    const double mEnd = g0;
    double t1, t2, t3, tm;  // cswSoundSpeed, this side
    call(1, 1, t1, g1);
    call(2, 1, t2, g1);
    call(3, 1, t3, g1);
    call(0, 1, tm, g1);
    const double vel2 = t1 * t1 + t2 * t2 + t3 * t3 / (tm * tm);
    double b1, b2, b3;
    call(5, 1, b1, g1);
    call(6, 1, b2, g1);
    call(7, 1, b3, g1);
    const double mag2 = b1 * b1 + b2 * b2 + b3 * b3;
    double te, tm2, tm3;
    call(4, 1, te, g1);
    call(0, 1, tm2, g1);
    double sp = (G - 1.0) * (te - 0.5 * tm2 * vel2);
    sp = sp - (G - 1.0) * 0.5 * mag2;
    call(0, 1, tm3, g1);
    out[0] = mEnd;
    out[1] = sqrt(fabs((G * sp + 2.0 * mag2) / tm3));
    if (!up) {
        call(1 + d, 2, g0, g1);  // v0 = mom.shift(-2) ./ mass.shift(-2)
        call(0, 2, u0, u1);
        out[2] = g0 / u0;
        out[3] = g1 / u1;
        call(1 + d, 1, g0, g1);  // v1
        call(0, 1, u0, u1);
        out[4] = g0 / u0;
    } else {
        call(1 + d, 1, g0, g1);  // v2
        call(0, 1, u0, u1);
        out[2] = g0 / u0;
        call(1 + d, 2, g0, g1);  // v3 = mom.shift(2) ./ mass.shift(2)
        call(0, 2, u0, u1);
        out[3] = g0 / u0;
        out[4] = g1 / u1;
    }
    if (mode == IMH_BC_TRANS)
        for (int r = 0; r < 8; ++r) *tp[r] = tv[r];
}

// Max over each pencil of the block (NaN-ignoring): wave shuffles, then one LDS pass.
// Takes the tile geometry by value: a reference to the argument block here would keep
// the block's private copy in scratch memory.
template <typename T, int PXM>
__device__ __forceinline__ T pencil_reduce_max(T v, int xmajor, int TL, int PX, int tid, int px,
                                               T* s_red) {
    const int wave = tid >> 6;
    if (xmajor) {
        int g = TL < 64 ? TL : 64;
        for (int off = g >> 1; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    } else {
        for (int off = PX; off < 64; off <<= 1) v = fmax(v, __shfl_xor(v, off, 64));
    }
    s_red[wave * PXM + px] = v;
    __syncthreads();
    T m = v;
    const int nw = blockDim.x >> 6;
    if (xmajor) {
        int w0 = (px * TL) >> 6, w1 = (px * TL + TL - 1) >> 6;
        for (int w = w0; w <= w1; ++w) m = fmax(m, s_red[w * PXM + px]);
    } else {
        for (int w = 0; w < nw; ++w) m = fmax(m, s_red[w * PXM + px]);
    }
    return m;
}

}  // namespace

// K consecutive cells as 16-byte (or one 8-byte) vector accesses; p is aligned to
// K * sizeof(T) bytes.
template <typename T, int K>
__device__ __forceinline__ void chunk_load(const T* p, T (&out)[K]) {
    constexpr int B = K * int(sizeof(T));
    if constexpr (B % 16 == 0) {
        constexpr int L = 16 / int(sizeof(T));
        typedef T V __attribute__((ext_vector_type(L)));
#pragma unroll
        for (int j = 0; j < K / L; ++j) {
            const V x = *gmem(reinterpret_cast<const V*>(p + j * L));
#pragma unroll
            for (int l = 0; l < L; ++l) out[j * L + l] = x[l];
        }
    } else if constexpr (B == 8 && K == 2) {
        typedef T V __attribute__((ext_vector_type(2)));
        const V x = *gmem(reinterpret_cast<const V*>(p));
        out[0] = x[0];
        out[1] = x[1];
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = gmem(p)[k];
    }
}

template <typename T, int K>
__device__ __forceinline__ void chunk_store(T* p, const T (&in)[K]) {
    constexpr int B = K * int(sizeof(T));
    if constexpr (B % 16 == 0) {
        constexpr int L = 16 / int(sizeof(T));
        typedef T V __attribute__((ext_vector_type(L)));
#pragma unroll
        for (int j = 0; j < K / L; ++j) {
            V x;
#pragma unroll
            for (int l = 0; l < L; ++l) x[l] = in[j * L + l];
            *gmem(reinterpret_cast<V*>(p + j * L)) = x;
        }
    } else if constexpr (B == 8 && K == 2) {
        typedef T V __attribute__((ext_vector_type(2)));
        V x;
        x[0] = in[0];
        x[1] = in[1];
        *gmem(reinterpret_cast<V*>(p)) = x;
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) gmem(p)[k] = in[k];
    }
}

// Fields handled per exchange phase: their FR / FL live in registers across the phase's
// barriers, so FP32 chunks of 8 cells (and FP64 of 4) take them in smaller groups.
// IMH_FLUID_GROUP (1, 2 or 5) overrides the choice at build time.
#ifndef IMH_FLUID_GROUP
#define IMH_FLUID_GROUP 0
#endif
template <typename T, int KMAX>
constexpr int fluid_group() {
    if (IMH_FLUID_GROUP > 0) return IMH_FLUID_GROUP;
    return (int(sizeof(T)) * KMAX <= 16) ? 5 : (int(sizeof(T)) * KMAX <= 32 ? 2 : 1);
}

// Waves per SIMD the register allocator aims for (0: as many registers as the block size
// allows); a build-time knob.
#ifndef IMH_FLUID_WAVES_PER_EU
#define IMH_FLUID_WAVES_PER_EU 0
#endif
#if IMH_FLUID_WAVES_PER_EU > 0
#define IMH_FLUID_OCCUPANCY __attribute__((amdgpu_waves_per_eu(IMH_FLUID_WAVES_PER_EU)))
#else
#define IMH_FLUID_OCCUPANCY
#endif

// The sweep of one tile (PX pencils) of one simulation.
//
// A pencil is split over TL threads, each keeping KMAX consecutive cells in registers
// (cells past the pencil end are padding: loaded clamped, computed, never stored or
// counted).  Each phase of relaxingFluid.m works on a group of fields: compute FR / FL
// in registers, stage each chunk's first and last values (for the neighbouring threads)
// and the pencil ends (for the BC) in LDS, barrier, evaluate all pencil-end ghosts of
// the group in parallel (one lane per pencil, field and shift call), barrier, update.
//
// GEN = false is the common case and keeps the kernel free of calls (calls cost
// registers): no fade boundary and no first assignment of the flux objects.  GEN = true
// handles those too, with their code out of line.
//
// PXM: the most pencils a tile holds (the size of the per-pencil LDS arrays; the 1-D
// lane's tiles hold one).  GF > 0: fields per group instead of fluid_group's choice.
template <typename T, int BS, int KMAX, bool MHD, bool GEN, int PXM = PXMAX, int GF = 0>
__device__ __forceinline__ void fluid_sweep_body(const FluidArgs<T>& ap, long long tile) {
    const FluidArgs<T> a = ap;  // private copy: stores to the fields cannot alias it
    if (!a.force && !a.sim->active) return;
    using M = FMath<T>;
    constexpr int G = GF > 0 ? GF : fluid_group<T, KMAX>();
    constexpr int K = KMAX;
    // chunk boundary values of the group: [field*4 + q], q = 0 first FR, 1 last FR,
    // 2 first FL, 3 last FL
    __shared__ __attribute__((aligned(16))) T s_x[4 * G][BS];
    __shared__ double s_ends[2][G][PXM][6];     // FR / FL pencil ends (a[0..2], a[N-3..N-1])
    __shared__ double s_gh[PXM][G][4];          // ghosts: FR lower, FL lower, FR upper, FL upper
    __shared__ T s_red[2][(BS / 64) * PXM];
    // edge trackers of the block's pencils: [px][side*16 + role*5 + field], roles lower:
    // store.fluxR, fluxR, fluxL; upper: store.fluxL, fluxR, fluxL
    __shared__ double s_trk[PXM][32];
    // artificial viscosity (general form only): first momentum value of each chunk, the
    // pencil ends of the momentum row and its upper ghost
    __shared__ T s_mv[GEN ? BS : 1];
    __shared__ double s_vend[GEN ? PXM : 1][6];
    __shared__ double s_vgh[GEN ? PXM : 1];
    // Caramana-Shashkov-Whalen rows at the chunk edges (S first/last, m first/last, V at
    // k = 0, 1, K-2, K-1) and at the pencil ends (cells 0..2, N-3..N-1)
    __shared__ T s_cw[GEN ? 8 : 1][GEN ? BS : 1];
    __shared__ double s_cend[GEN ? PXM : 1][3][6];

    const int tid = threadIdx.x;
    const int PX = a.PX, TL = a.TL;
    const int px = a.xmajor ? tid / TL : tid % PX;
    const int ty = a.xmajor ? tid % TL : tid / PX;
    const int dstep = a.xmajor ? 1 : PX;
    const long long pen0 = tile * PX;
    const int d = a.d;
    const int nn[3] = {a.n0, a.n1, a.n2};
    // face of slot p of the tile, -1 for padding (with a lower wormhole boundary the slots
    // hold pencil pairs, worm_face)
    auto slot_face = [&](int p) __attribute__((always_inline)) -> long long {
        const long long sl = pen0 + p;
        if (sl >= a.npencil) return -1;
        return a.worm ? worm_face(nn, d, sl) : sl;
    };
    // the slot whose cell 0 gives slot p's lower wormhole ghosts
    auto partner_slot = [&](int p) __attribute__((always_inline)) -> int {
        return slot_face(p ^ 1) >= 0 ? (p ^ 1) : p;
    };
    const long long pface = slot_face(px);
    const bool pen_ok = pface >= 0;
    // padding pencils read a real one (the last, or face 0 with pairs)
    const long long pen = pen_ok ? pface : (a.worm ? 0 : a.npencil - 1);
    const int N = sel3i(d, a.n0, a.n1, a.n2);
    const int i0 = ty * K;
    const bool has_left = ty > 0;
    const int tl_ = tid - dstep < 0 ? 0 : tid - dstep;      // neighbour threads (clamped)
    const int tr_ = tid + dstep >= BS ? BS - 1 : tid + dstep;
    // 32-bit cell offsets (a simulation holds fewer than 2^31 cells)
    long long base64 = 0, stride64 = 1;
    pencil_base(nn, d, pen, base64, stride64);
    const unsigned base = unsigned(base64), stride = unsigned(stride64);
    auto off = [&](int k) __attribute__((always_inline)) -> unsigned {
        const int i = i0 + k < N ? i0 + k : N - 1;
        return base + unsigned(i) * stride;
    };
    auto valid = [&](int k) __attribute__((always_inline)) -> bool { return pen_ok && i0 + k < N; };

    const double dt = a.sim->dt;
    const double ff = dt / a.dx;          // fluxFactor = dTime ./ DGRID{X}
    const double hf = 0.5 * ff;
    const int X = d + 1;                  // momentum field of the sweep direction
    const double g = a.gamma;
    const int f = a.bc_inf;
    const int mlo = mode_at(a.modes, 0, d), mhi = mode_at(a.modes, 1, d);
    const bool trk_lo = (mlo == IMH_BC_TRANS || mlo == IMH_BC_WALL);
    const bool trk_hi = (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL);
    const bool fade = GEN && (mlo == IMH_BC_FADE || mhi == IMH_BC_FADE);
    // circular and constant boundaries need only a[0] and a[N-1] of a row: the first value
    // of the first chunk and the last value of the last chunk
    const bool simple_bc = (mlo == IMH_BC_CIRC || mlo == IMH_BC_CONST) &&
                           (mhi == IMH_BC_CIRC || mhi == IMH_BC_CONST);
    const bool owns_end = !simple_bc && pen_ok && (i0 < 3 || i0 + K > N - 3) && i0 < N;
    const int ty_last = (N - 1) / K;
    // padding cells past the pencil end repeat the last cell (loads are clamped, and the
    // half step and limited fluxes are copied forward), so a chunk's last value is a[N-1]
    auto pad_fill = [&](T (&A)[K]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 1; k < K; ++k) A[k] = (i0 + k < N) ? A[k] : A[k - 1];
    };

    // ---- tracker prefetch (cooperative) ----------------------------------------------
    if ((trk_lo || trk_hi) && !a.init_edges) {
        for (int e = tid; e < PX * 30; e += BS) {
            int p = e / 30, r = e % 30;
            int side = r / 15, role = (r % 15) / 5, v = r % 5;
            const long long pp = slot_face(p);
            if (pp < 0 || (side ? !trk_hi : !trk_lo)) continue;
            int obj = side ? (role == 0 ? OBJ_SFL : (role == 1 ? OBJ_FR : OBJ_FL))
                           : (role == 0 ? OBJ_SFR : (role == 1 ? OBJ_FR : OBJ_FL));
            s_trk[p][side * 16 + role * 5 + v] = a.edges[a.eoff[((obj + v) * 2 + side) * 3 + d] + pp];
        }
    }

    // ---- load the chunk (clamped at the pencil end: no branches) ----------------------
    // x pencils whose length is a multiple of K move whole chunks as vectors; threads past
    // the pencil end then read the last chunk (their values are never used)
    const bool vec = a.xmajor && (N % K == 0);
    const unsigned vbase = base + unsigned(i0 < N ? i0 : N - K);
    T q[5][K];
    T bm[3][K];
    if (vec) {
#pragma unroll
        for (int v = 0; v < 5; ++v) chunk_load<T, K>(a.q[v] + vbase, q[v]);
        if constexpr (MHD) {
#pragma unroll
            for (int v = 0; v < 3; ++v) chunk_load<T, K>(a.bc[v] + vbase, bm[v]);
        } else {
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int k = 0; k < K; ++k) bm[v][k] = T(0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const unsigned o = off(k);
#pragma unroll
            for (int v = 0; v < 5; ++v) q[v][k] = gmem(a.q[v])[o];
#pragma unroll
            for (int v = 0; v < 3; ++v) bm[v][k] = MHD ? gmem(a.bc[v])[o] : T(0);
        }
    }

    // First assignment of a flux object (InitializedArray.m:43-47): record its edges on
    // the faces of the other dimensions.
    auto init_faces = [&](int obj, const T (&vals)[K]) __attribute__((always_inline)) {
        if constexpr (!GEN) return;
        if (!a.init_edges || a.ndims < 2) return;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (valid(k))
                fluid_init_face(a.edges, a.eoff, a.modes, a.ndims, d, obj, i0 + k, pen, a.n0,
                                a.n1, a.n2, double(vals[k]));
    };
    // pencil ends of a row held by this thread's cells
    auto put_ends = [&](int which, int gv, const T (&A)[K]) __attribute__((always_inline)) {
        if (simple_bc || !owns_end) return;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int gi = i0 + k;
            if (gi < 3 && gi < N) s_ends[which][gv][px][gi] = double(A[k]);
            if (gi >= N - 3 && gi < N) s_ends[which][gv][px][3 + gi - (N - 3)] = double(A[k]);
        }
    };

    // Ghosts of one phase for the group [v0, v0 + ng): kind 0 predictor (store.fluxR
    // lower, store.fluxL upper), 1 corrector (fluxR and fluxL, both sides), 2 final
    // (fluxR lower, fluxL upper).
    auto ghost_phase = [&](int kind, int v0, int ng) __attribute__((always_inline)) {
        const int J = kind == 1 ? 4 : 2;
        for (int job = tid; job < PX * ng * J; job += BS) {
            const int p = job / (ng * J);
            const int r = job - p * (ng * J);
            const int gv = r / J;
            const int v = v0 + gv;
            const int jj = r - gv * J;
            const int j = (J == 4) ? jj : (jj ? 3 : 0);  // 0 FR lo, 1 FL lo, 2 FR hi, 3 FL hi
            if (slot_face(p) < 0) continue;
            const int which = j & 1;           // 0 fluxR, 1 fluxL
            const int upper = j >> 1;
            const int mode = upper ? mhi : mlo;
            if (simple_bc) {
                // circ: lower ghost a[N-1], upper a[0]; const: lower a[0], upper a[N-1]
                const bool first = (mode == IMH_BC_CIRC) == (upper == 1);
                const int t = first ? (a.xmajor ? p * TL : p)
                                    : (a.xmajor ? p * TL + ty_last : ty_last * PX + p);
                s_gh[p][gv][j] = double(s_x[gv * 4 + which * 2 + (first ? 0 : 1)][t]);
                continue;
            }
            const int role = kind == 0 ? 0 : 1 + which;
            const int slot = upper * 16 + role * 5 + v;
            const double* e = s_ends[which][gv][p];
            if (mode == IMH_BC_WORMHOLE && !upper) {
                // wormhole_shift.m:21-57: cell 0 of the partner pencil, negated for the
                // momentum along the sweep (isVecNorm)
                const double w = s_ends[which][gv][partner_slot(p)][0];
                s_gh[p][gv][j] = (v == X) ? -w : w;
                continue;
            }
            if (GEN && a.init_edges && kind < 2) s_trk[p][slot] = upper ? e[5] : e[0];
            double gval;
            if (mode == IMH_BC_CIRC) gval = upper ? e[0] : e[5];
            else if (mode == IMH_BC_CONST || mode == IMH_BC_FADE) gval = upper ? e[5] : e[0];
            else {
                const int ft = which ? FT_FLUXL : FT_FLUXR;
                const int zero_ghost = upper ? (ft != FT_FLUXR) : (ft != FT_FLUXL);
                if constexpr (GEN) {
                    gval = fluid_ghost(mode, upper, v == X, zero_ghost, a.active_lin, f, e[0],
                                       e[1], e[2], e[3], e[4], e[5], N, &s_trk[p][slot]);
                } else {
                    GhostSpec sp;
                    sp.mode = mode;
                    sp.upper = upper;
                    sp.flip_sign = (v == X);
                    sp.zero_ghost = zero_ghost;
                    sp.active_linear = a.active_lin;
                    sp.bc_infinity = f;
                    const Ends en = {e[0], e[1], e[2], e[3], e[4], e[5]};
                    double g1;
                    ghost_values(sp, en, N, 1, &s_trk[p][slot], gval, g1);
                }
            }
            s_gh[p][gv][j] = gval;
        }
    };

    // Ghost j (0 FR lower, 1 FL lower, 2 FR upper, 3 FL upper) of field gg of this
    // thread's pencil after ghost_phase.  Circular and constant boundaries read it straight
    // from the chunk-boundary exchange (ghost_phase and its barrier are skipped for them):
    // circ lower a[N-1], upper a[0]; const lower a[0], upper a[N-1].
    auto ghost_of = [&](int gg, int j) __attribute__((always_inline)) -> T {
        if (simple_bc) {
            const int which = j & 1, upper = j >> 1;
            const int mode = upper ? mhi : mlo;
            const bool first = (mode == IMH_BC_CIRC) == (upper == 1);
            const int t = first ? (a.xmajor ? px * TL : px)
                                : (a.xmajor ? px * TL + ty_last : ty_last * PX + px);
            return s_x[gg * 4 + which * 2 + (first ? 0 : 1)][t];
        }
        return T(s_gh[px][gg][j]);
    };
    // the ghost phase of one exchange and its barrier (none for circular / constant)
    auto ghosts = [&](int kind, int v0, int ng) __attribute__((always_inline)) {
        if (simple_bc) return;
        ghost_phase(kind, v0, ng);
        __syncthreads();
    };

    // fade_shift.m blend of a shifted value at cell i (lower: shift(X,-1))
    auto fade_fix = [&](bool lower, int which, int gv, int i, T r) __attribute__((always_inline)) -> T {
        if constexpr (!GEN) return r;
        const int mode = lower ? mlo : mhi;
        if (mode == IMH_BC_FADE && (i < f || i >= N - f) && i < N) {
            const double* e = s_ends[which][gv][px];
            // constant-shifted end values: lower r[0]=a[0], r[N-1]=a[N-2];
            // upper r[0]=a[1], r[N-1]=a[N-1]
            const double v1 = lower ? e[0] : e[1];
            const double vN = lower ? e[4] : e[5];
            r = T(fluid_fade(double(r), i, N, v1, vN, a.fade, f));
        }
        return r;
    };
    // lower / upper shifted values of a register row A at chunk cell k, from the
    // neighbouring chunk's boundary value nb or the pencil-end ghost
    auto lower_of = [&](const T (&A)[K], int k, T nb, T ghost) __attribute__((always_inline)) -> T {
        if (k > 0) return A[k > 0 ? k - 1 : 0];
        return has_left ? nb : ghost;
    };
    auto upper_of = [&](const T (&A)[K], int k, T nb, T ghost) __attribute__((always_inline)) -> T {
        const int i = i0 + k;
        const T in = (k + 1 < K) ? A[k + 1 < K ? k + 1 : k] : nb;
        return i < N - 1 ? in : ghost;
    };

    T P[K], U[K], IR[K], CS[K];
    // pressure, velocity and the freezing-speed candidate of every cell of state Q
    auto cell_speeds = [&](const T (&Q)[5][K]) __attribute__((always_inline)) -> T {
        T lmax = T(NAN);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            T Pk, ck, ir;
            M::press(Q[0][k], Q[1][k], Q[2][k], Q[3][k], Q[4][k], bm[0][k], bm[1][k], bm[2][k],
                     a.use, g, Pk, ck, ir);
            P[k] = Pk;
            CS[k] = ck;
            IR[k] = ir;
            U[k] = M::divrho(sel3(d, Q[1][k], Q[2][k], Q[3][k]), ir);
            const T sp = fabs(U[k]) + ck;
            lmax = fmax(lmax, valid(k) ? sp : T(NAN));  // NaN-ignoring, like MATLAB max
        }
        return lmax;
    };

    // wFluidFlux.m: FR / FL of field v from state Q and freezing speed C
    auto fluxes = [&](auto VC, const T (&Q)[5][K], T C, T rC, T (&FR)[K], T (&FL)[K])
        __attribute__((always_inline)) {
        constexpr int v = decltype(VC)::value;
#pragma unroll
        for (int k = 0; k < K; ++k) {
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
    // fn(FieldIdx<v>, gv) for the fields of group [V0, V0 + G)
    auto for_group = [&](auto V0C, auto fn) __attribute__((always_inline)) {
        constexpr int V0 = decltype(V0C)::value;
        auto one = [&](auto VC) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            if constexpr (v >= V0 && v < V0 + G) fn(VC, v - V0);
        };
        one(FieldIdx<0>{});
        one(FieldIdx<1>{});
        one(FieldIdx<2>{});
        one(FieldIdx<3>{});
        one(FieldIdx<4>{});
    };
    // fn(FieldIdx<v0>) for every group start v0 = 0, G, 2G, ...
    auto for_groups = [&](auto fn) __attribute__((always_inline)) {
        fn(FieldIdx<0>{});
        if constexpr (G < 5) fn(FieldIdx<(G < 5 ? G : 0)>{});
        if constexpr (2 * G < 5) fn(FieldIdx<(2 * G < 5 ? 2 * G : 0)>{});
        if constexpr (3 * G < 5) fn(FieldIdx<(3 * G < 5 ? 3 * G : 0)>{});
        if constexpr (4 * G < 5) fn(FieldIdx<(4 * G < 5 ? 4 * G : 0)>{});
    };

    // press = press + viscosity.solve(...) (wFluidFlux.m:24).  The objects of the five
    // fields are obj + 0..4: the FluidArrays (OBJ_PRIM) in the predictor, their
    // StorageArrays (OBJ_FSTORE) in the corrector; `first` = the corrector of the run's
    // first sweep, the StorageArrays' first assignment.
    // Neumann-Richtmyer (ArtificialViscosity.m:73-77): vm = min(p(i+1) - p(i), 0) of the
    // sweep-direction momentum, result vm .* (c1 .* c_f + c2 .* vm ./ rho); the shift's
    // upper ghost follows that momentum object's BC and nudges its tracker.
    // Caramana-Shashkov-Whalen (ArtificialViscosity.m:80-148 with cswSoundSpeed,
    // :155-190): inside the pencil every shifted array is a cell row, so three per-cell
    // rows carry it: S (cswSoundSpeed of one cell, whose velSquared divides only the third
    // momentum by mass^2, as the .m text does), m (mass) and V (mom(X) ./ mass).  Past a
    // circ / const / zero / mirror / flip end the shifts are index maps (mom(X) changing
    // sign at mirror and flip ends).  Past a transparent or wall end every shift call has
    // its own ghost from its object's tracker (about 30 calls per wFluidFlux, mass,
    // momentum, energy and cell-centred B objects): one lane per pencil replays the calls
    // of each side in the .m text's order and leaves the five values the cells read there
    // (massB/C, sndB/C, v0, v1 / v2, v3), staged in the exchange rows s_x (unused here).
    auto csw = [&](const T (&Q)[5][K], int obj) __attribute__((always_inline)) {
      // the body exists in the general form only: left in the common form it changes that
      // form's register allocation although it is never called there
      if constexpr (GEN) {
        const T G = T(g);
        T Sr[K], Mr[K], Vr[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T rho = Q[0][k], m1 = Q[1][k], m2 = Q[2][k], m3 = Q[3][k];
            const T vel2 = m1 * m1 + m2 * m2 + m3 * m3 / (rho * rho);
            const T mag2 = bm[0][k] * bm[0][k] + bm[1][k] * bm[1][k] + bm[2][k] * bm[2][k];
            T sp = (G - T(1)) * (Q[4][k] - T(0.5) * rho * vel2);
            sp = sp - (G - T(1)) * T(0.5) * mag2;
            Sr[k] = sqrt(fabs((G * sp + T(2) * mag2) / rho));
            Mr[k] = rho;
            Vr[k] = sel3(d, m1, m2, m3) / rho;
        }
        s_cw[0][tid] = Sr[0];
        s_cw[1][tid] = Sr[K - 1];
        s_cw[2][tid] = Mr[0];
        s_cw[3][tid] = Mr[K - 1];
        s_cw[4][tid] = Vr[0];
        s_cw[5][tid] = Vr[K > 1 ? 1 : 0];
        s_cw[6][tid] = Vr[K > 1 ? K - 2 : 0];
        s_cw[7][tid] = Vr[K - 1];
        // tracked ends: the 8 objects' end cells (mass, mom 1..3, ener, cellMag 1..3) and
        // the lane's results, in the exchange rows
        const bool trk_any = trk_lo || trk_hi;
        double* const xe = reinterpret_cast<double*>(&s_x[0][0]);
        constexpr int XE_ENDS = 48, XE_OUT = 10;
        static_assert(sizeof(s_x) >= sizeof(double) * PXM * (XE_ENDS + XE_OUT),
                      "exchange rows too small for the CSW pencil ends");
        if (pen_ok) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int gi = i0 + k;
                int slot = -1;
                if (gi < 3 && gi < N) slot = gi;
                else if (gi >= N - 3 && gi < N) slot = 3 + gi - (N - 3);
                if (slot >= 0) {
                    s_cend[px][0][slot] = double(Sr[k]);
                    s_cend[px][1][slot] = double(Mr[k]);
                    s_cend[px][2][slot] = double(Vr[k]);
                    if (trk_any) {
                        double* e = xe + px * XE_ENDS;
#pragma unroll
                        for (int v = 0; v < 5; ++v) e[v * 6 + slot] = double(Q[v][k]);
#pragma unroll
                        for (int b = 0; b < 3; ++b) e[(5 + b) * 6 + slot] = double(bm[b][k]);
                    }
                }
            }
        }
        __syncthreads();
        if (trk_any && tid < PX && slot_face(tid) >= 0) {
            const int p = tid;
            for (int up = 0; up < 2; ++up) {
                const int mode = up ? mhi : mlo;
                if (mode != IMH_BC_TRANS && mode != IMH_BC_WALL) continue;
                double* tp[8];
                for (int r = 0; r < 8; ++r) {
                    const int o = r < 5 ? obj + r : OBJ_CMAG + (r - 5);
                    tp[r] = a.edges + a.eoff[(o * 2 + up) * 3 + d] + slot_face(p);
                }
                csw_side_ghosts(mode, up, d, N, a.active_lin, f, g, tp, xe + p * XE_ENDS,
                                xe + PXM * XE_ENDS + p * XE_OUT + up * 5);
            }
        }
        if (trk_any) __syncthreads();
        // row value (0 S, 1 m, 2 V) of the array shifted by -o at cell i0 + k, |o| <= 2
        auto at = [&](int row, const T (&R)[K], int k, int o) __attribute__((always_inline)) -> T {
            const int i = i0 + k;
            const int j = i + o;
            if (j < 0 || j >= N) {
                const int mode = j < 0 ? mlo : mhi;
                if (mode == IMH_BC_TRANS || mode == IMH_BC_WALL) {
                    const double* ov = xe + PXM * XE_ENDS + px * XE_OUT + (j < 0 ? 0 : 5);
                    if (row == 1) return T(ov[0]);
                    if (row == 0) return T(ov[1]);
                    if (j < 0) return T(o == -1 ? ov[4] : (i == 0 ? ov[2] : ov[3]));
                    return T(o == 1 ? ov[2] : (i == N - 2 ? ov[3] : ov[4]));
                }
                int jj;
                bool neg = false;
                if (mode == IMH_BC_CIRC) {
                    jj = j < 0 ? j + N : j - N;
                } else if (mode == IMH_BC_MIRROR) {
                    jj = j < 0 ? -j : 2 * (N - 1) - j;
                    neg = true;
                } else {  // const, zero (not a flux array), flip
                    jj = j < 0 ? 0 : N - 1;
                    neg = (mode == IMH_BC_FLIP);
                }
                const int slot = jj <= 2 ? jj : 3 + jj - (N - 3);
                const T v = T(s_cend[px][row][slot]);
                return (neg && row == 2) ? -v : v;
            }
            const int kk = k + o;
            if (kk >= 0 && kk < K) return R[kk >= 0 && kk < K ? kk : 0];
            const int q = kk < 0 ? -((K - 1 - kk) / K) : kk / K;  // floor(kk / K)
            const int r = kk - q * K;
            const int t = tid + q * dstep;  // the thread holding cell j of this pencil
            if (row == 0) return s_cw[r == 0 ? 0 : 1][t];
            if (row == 1) return s_cw[r == 0 ? 2 : 3][t];
            return s_cw[r == 0 ? 4 : (r == 1 ? 5 : (r == K - 1 ? 7 : 6))][t];
        };
        const T c1 = T(a.visc_lin), c2 = T(a.visc_quad);
        const T q4 = (G + T(1)) / T(4);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T mB = at(1, Mr, k, -1), mC = at(1, Mr, k, 1);
            const T rho = T(2) * mC * mB / (mC + mB);
            const T snd = fmax(fmin(at(0, Sr, k, -1), at(0, Sr, k, 1)), T(0));
            const T v0 = at(2, Vr, k, -2), v1 = at(2, Vr, k, -1);
            const T v2 = at(2, Vr, k, 1), v3 = at(2, Vr, k, 2);
            const T dv = v2 - v1;
            const T rL = (v3 - v2) / (T(2) * dv);
            const T rR = (v1 - v0) / (T(2) * dv);
            const T lim = fmax(T(0), fmin(fmin(T(0.5) * (rL + rR), fmin(T(2) * rL, T(2) * rR)), T(1)));
            const T A = c2 * (G + T(1)) * fabs(dv) / T(4);
            const T B = sqrt(c2 * c2 * (q4 * q4) * (dv * dv) + c1 * c1 * (snd * snd));
            P[k] = P[k] + (dv < T(0) ? T(1) : T(0)) * rho * (A + B) * (T(1) - lim) * fabs(dv);
        }
        // the group loop that follows writes the exchange rows
        if (trk_any) __syncthreads();
      }
    };
    auto viscosity = [&](const T (&Q)[5][K], int obj, bool first) __attribute__((always_inline)) {
        if constexpr (GEN) {
            if (a.visc_type == 0) return;
            if (a.visc_type == 2) {
                csw(Q, obj);
                return;
            }
            T mX[K];
#pragma unroll
            for (int k = 0; k < K; ++k) mX[k] = sel3(d, Q[1][k], Q[2][k], Q[3][k]);
            s_mv[tid] = mX[0];
            if (pen_ok) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int gi = i0 + k;
                    if (gi < 3 && gi < N) s_vend[px][gi] = double(mX[k]);
                    if (gi >= N - 3 && gi < N) s_vend[px][3 + gi - (N - 3)] = double(mX[k]);
                }
            }
            __syncthreads();
            if (tid < PX && slot_face(tid) >= 0) {
                const double* e = s_vend[tid];
                const bool trk = (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL);
                double* tp = trk ? a.edges + a.eoff[((obj + X) * 2 + 1) * 3 + d] + slot_face(tid) : nullptr;
                double tv = trk ? (first ? e[5] : *tp) : 0.0;
                GhostSpec sp;
                sp.mode = mhi;
                sp.upper = 1;
                sp.flip_sign = 1;  // the shifted array is the momentum component X
                sp.zero_ghost = 0;
                sp.active_linear = a.active_lin;
                sp.bc_infinity = f;
                const Ends en = {e[0], e[1], e[2], e[3], e[4], e[5]};
                double g0, g1;
                ghost_values(sp, en, N, 1, &tv, g0, g1);
                if (trk) *tp = tv;
                s_vgh[tid] = g0;
            }
            __syncthreads();
            const T nb = s_mv[tr_];
            const T gh = T(s_vgh[px]);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                T sh = upper_of(mX, k, nb, gh);
                if (mhi == IMH_BC_FADE && i0 + k < N)  // fade_shift.m: both ends blended,
                                                       // v1 = a[1], vN = a[N-1]
                    sh = fade_blend(sh, i0 + k, N, T(s_vend[px][1]), T(s_vend[px][5]), a.fade, f);
                const T vm = fmin(sh - mX[k], T(0));  // MATLAB min(.., 0) skips NaN
                P[k] = P[k] + vm * (T(a.visc_lin) * CS[k] + T(a.visc_quad) * vm / Q[0][k]);
            }
        }
    };

    // ================= predictor: W and C from U^n (wFluidFlux.m) ====================
    const T C = pencil_reduce_max<T, PXM>(cell_speeds(q), a.xmajor, TL, PX, tid, px, s_red[0]);
    viscosity(q, OBJ_PRIM, false);
    const T rC = M::rcp(C);
    const T hfC = T(hf) * C;
    T qh[5][K];
    for_groups([&](auto V0C) __attribute__((always_inline)) {
        constexpr int V0 = decltype(V0C)::value;
        constexpr int NG = (5 - V0) < G ? (5 - V0) : G;
        if constexpr (V0 > 0) __syncthreads();  // the previous group's values have been read
        T FR[G][K], FL[G][K];
        for_group(V0C, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            fluxes(VC, q, C, rC, FR[v - V0], FL[v - V0]);
            init_faces(OBJ_SFR + v, FR[v - V0]);
            init_faces(OBJ_SFL + v, FL[v - V0]);
            s_x[(v - V0) * 4 + 0][tid] = FR[v - V0][0];
            s_x[(v - V0) * 4 + 1][tid] = FR[v - V0][K - 1];
            s_x[(v - V0) * 4 + 2][tid] = FL[v - V0][0];
            s_x[(v - V0) * 4 + 3][tid] = FL[v - V0][K - 1];
            put_ends(0, v - V0, FR[v - V0]);
            put_ends(1, v - V0, FL[v - V0]);
        });
        __syncthreads();
        ghosts(0, V0, NG);
        for_group(V0C, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            constexpr int gg = v - V0;
            const T nbR = s_x[gg * 4 + 1][tl_];
            const T nbL = s_x[gg * 4 + 2][tr_];
            const T gR = ghost_of(gg, 0), gL = ghost_of(gg, 3);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                T sR = lower_of(FR[gg], k, nbR, gR);
                T sL = upper_of(FL[gg], k, nbL, gL);
                if (fade) {
                    sR = fade_fix(true, 0, gg, i0 + k, sR);
                    sL = fade_fix(false, 1, gg, i0 + k, sL);
                }
                // 0.5*fluxFactor .* tempFreeze; fluxFactor = dTime ./ DGRID{X} per cell
                // on a non-uniform grid (general form only)
                T hk = hfC;
                if constexpr (GEN)
                    if (a.dxl) hk = T(0.5 * (dt / a.dxl[i0 + k < N ? i0 + k : N - 1])) * C;
                qh[v][k] = q[v][k] - hk * (FR[gg][k] - sR + FL[gg][k] - sL);
            }
        });
    });
    // First assignment of the StorageArrays (InitializedArray.m:43-47): their edges, read
    // only by the viscosity's shifts, come from this half step (the mass store's from the
    // array before its MINMASS floor, which is its second assignment).
    if constexpr (GEN) {
        if (a.init_edges) {
            auto init_store = [&](auto VC) __attribute__((always_inline)) {
                constexpr int v = decltype(VC)::value;
                init_faces(OBJ_FSTORE + v, qh[v]);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int gi = i0 + k;
                    if (!valid(k)) continue;
                    if (gi == 0 && trk_lo)
                        a.edges[a.eoff[((OBJ_FSTORE + v) * 2 + 0) * 3 + d] + pen] = double(qh[v][k]);
                    if (gi == N - 1 && trk_hi)
                        a.edges[a.eoff[((OBJ_FSTORE + v) * 2 + 1) * 3 + d] + pen] = double(qh[v][k]);
                }
            };
            init_store(FieldIdx<0>{});
            init_store(FieldIdx<1>{});
            init_store(FieldIdx<2>{});
            init_store(FieldIdx<3>{});
            init_store(FieldIdx<4>{});
        }
    }
    // mass.store.array = max(mass.store.array, MINMASS)
#pragma unroll
    for (int k = 0; k < K; ++k) qh[0][k] = fmax(qh[0][k], T(a.minmass));
#pragma unroll
    for (int v = 0; v < 5; ++v) pad_fill(qh[v]);

    // ================= corrector: W' and C' from U^(n+1/2) ===========================
    const T C2 = pencil_reduce_max<T, PXM>(cell_speeds(qh), a.xmajor, TL, PX, tid, px, s_red[1]);
    viscosity(qh, OBJ_FSTORE, a.init_edges != 0);
    const T rC2 = M::rcp(C2);
    const T ffC = T(ff) * C2;
    const int lim = a.limiter;
    for_groups([&](auto V0C) __attribute__((always_inline)) {
        constexpr int V0 = decltype(V0C)::value;
        constexpr int NG = (5 - V0) < G ? (5 - V0) : G;
        if constexpr (V0 > 0) __syncthreads();
        T FR[G][K], FL[G][K];
        for_group(V0C, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            constexpr int gg = v - V0;
            fluxes(VC, qh, C2, rC2, FR[gg], FL[gg]);
            init_faces(OBJ_FR + v, FR[gg]);
            init_faces(OBJ_FL + v, FL[gg]);
            s_x[gg * 4 + 0][tid] = FR[gg][0];
            s_x[gg * 4 + 1][tid] = FR[gg][K - 1];
            s_x[gg * 4 + 2][tid] = FL[gg][0];
            s_x[gg * 4 + 3][tid] = FL[gg][K - 1];
            put_ends(0, gg, FR[gg]);
            put_ends(1, gg, FL[gg]);
        });
        __syncthreads();
        ghosts(1, V0, NG);
        // limited fluxes, in place (each cell's update uses the values before it)
        for_group(V0C, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            constexpr int gg = v - V0;
            const T nbRl = s_x[gg * 4 + 1][tl_], nbRr = s_x[gg * 4 + 0][tr_];
            const T nbLl = s_x[gg * 4 + 3][tl_], nbLr = s_x[gg * 4 + 2][tr_];
            const T gRlo = ghost_of(gg, 0), gLlo = ghost_of(gg, 1);
            const T gRhi = ghost_of(gg, 2), gLhi = ghost_of(gg, 3);
            T R2[K], L2[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                T rm = lower_of(FR[gg], k, nbRl, gRlo), rp = upper_of(FR[gg], k, nbRr, gRhi);
                T lm = lower_of(FL[gg], k, nbLl, gLlo), lp = upper_of(FL[gg], k, nbLr, gLhi);
                if (fade) {
                    rm = fade_fix(true, 0, gg, i0 + k, rm);
                    rp = fade_fix(false, 0, gg, i0 + k, rp);
                    lm = fade_fix(true, 1, gg, i0 + k, lm);
                    lp = fade_fix(false, 1, gg, i0 + k, lp);
                }
                const T fr = FR[gg][k], fl = FL[gg][k];
                R2[k] = fr + limit(lim, fr - rm, rp - fr);
                L2[k] = fl + limit(lim, lm - fl, fl - lp);
            }
            pad_fill(R2);
            pad_fill(L2);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                FR[gg][k] = R2[k];
                FL[gg][k] = L2[k];
            }
        });
        __syncthreads();  // every thread has read this phase's boundary values and ends
        for_group(V0C, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int gg = decltype(VC)::value - V0;
            s_x[gg * 4 + 0][tid] = FR[gg][0];
            s_x[gg * 4 + 1][tid] = FR[gg][K - 1];
            s_x[gg * 4 + 2][tid] = FL[gg][0];
            s_x[gg * 4 + 3][tid] = FL[gg][K - 1];
            put_ends(0, gg, FR[gg]);
            put_ends(1, gg, FL[gg]);
        });
        // U^n of the group's fields.  One group of all five fields (the common form with
        // few cells per thread): the registers loaded at the start, kept live (measured:
        // the re-read, 16-byte row segments in y sweeps and past the XCD's L2 by now,
        // cost more than the spills of keeping them: 2-D sweep 202 -> 194 us).  Else
        // re-read, all loads issued before the exchange: the addresses are clamped into
        // the pencil and the pencil range, so the loads need no guard (only the stores do)
        T un_g[G][K];
        for_group(V0C, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            if constexpr (G == 5 && !GEN) {
#pragma unroll
                for (int k = 0; k < K; ++k) un_g[v - V0][k] = q[v][k];
            } else if (vec) {
                chunk_load<T, K>(a.q[v] + vbase, un_g[v - V0]);
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k) un_g[v - V0][k] = gmem(a.q[v])[off(k)];
            }
        });
        __syncthreads();
        ghosts(2, V0, NG);
        for_group(V0C, [&](auto VC, int gv) __attribute__((always_inline)) {
            constexpr int v = decltype(VC)::value;
            constexpr int gg = v - V0;
            const T nb2R = s_x[gg * 4 + 1][tl_];
            const T nb2L = s_x[gg * 4 + 2][tr_];
            const T g2R = ghost_of(gg, 0), g2L = ghost_of(gg, 3);
            T dq[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                T sR = lower_of(FR[gg], k, nb2R, g2R);
                T sL = upper_of(FL[gg], k, nb2L, g2L);
                if (fade) {
                    sR = fade_fix(true, 0, gg, i0 + k, sR);
                    sL = fade_fix(false, 1, gg, i0 + k, sL);
                }
                T fk = ffC;
                if constexpr (GEN)
                    if (a.dxl) fk = T(dt / a.dxl[i0 + k < N ? i0 + k : N - 1]) * C2;
                dq[k] = fk * (FR[gg][k] - sR + FL[gg][k] - sL);
            }
            T un[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                un[k] = un_g[gg][k] - dq[k];
                if (v == 0) un[k] = fmax(un[k], T(a.minmass));  // mass.array = max(mass, MINMASS)
            }
            if (vec) {
                if (pen_ok && i0 < N) chunk_store<T, K>(a.q[v] + vbase, un);
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (valid(k)) gmem(a.q[v])[off(k)] = un[k];
            }
        });
    });

    // ---- write the edge trackers back -----------------------------------------------
    if (trk_lo || trk_hi) {
        __syncthreads();
        for (int e = tid; e < PX * 30; e += BS) {
            int p = e / 30, r = e % 30;
            int side = r / 15, role = (r % 15) / 5, v = r % 5;
            const long long pp = slot_face(p);
            if (pp < 0 || (side ? !trk_hi : !trk_lo)) continue;
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
template <typename T, int BS, int KMAX, bool MHD, bool GEN>
__global__ void __launch_bounds__(BS) IMH_FLUID_OCCUPANCY fluid_sweep_batch_kernel(const FluidArgs<T>* args,
                                                               const BatchTile* tiles,
                                                               int ntiles) {
    // one call site of the body (it is large)
    for (int r = 0;; ++r) {
        int arg = 0;
        long long tile;
        if (tiles) {
            const int t = xcd_round_tile(r, ntiles);
            if (t < 0) break;
            uniform_tile(tiles, t, arg, tile);
        } else {
            if (r > 0 || (int)blockIdx.x >= ntiles) break;
            tile = args[0].xcd_remap ? xcd_tile(blockIdx.x, ntiles) : (long long)blockIdx.x;
        }
        fluid_sweep_body<T, BS, KMAX, MHD, GEN>(args[arg], tile);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// Launcher of this translation unit's variant.  fluid_sweep_variant.hip is compiled once
// per entry of IMH_FLUID_VARIANTS (imh_kernels.h) with IMH_FLUID_ESZ / IMH_FLUID_BS /
// IMH_FLUID_KMAX defined, so the variants build in parallel; launch_fluid_batch
// (fluid_dispatch.cpp) picks one per sweep dimension.
// ------------------------------------------------------------------------------------
#ifndef IMH_FLUID_NO_LAUNCHER
#if IMH_FLUID_ESZ == 8
using FluidT = double;
#else
using FluidT = float;
#endif

hipError_t IMH_FLUID_LAUNCHER(IMH_FLUID_ESZ, IMH_FLUID_BS, IMH_FLUID_KMAX)(
    const void* args, const BatchTile* tiles, int ntiles, int nblocks, bool mhd, bool general,
    hipStream_t st) {
    const FluidArgs<FluidT>* a = (const FluidArgs<FluidT>*)args;
#define IMH_FL(M, G)                                                                          \
    if (mhd == M && general == G) {                                                           \
        hipLaunchKernelGGL((fluid_sweep_batch_kernel<FluidT, IMH_FLUID_BS, IMH_FLUID_KMAX, M, G>), \
                           dim3(nblocks), dim3(IMH_FLUID_BS), 0, st, a, tiles, ntiles);          \
        return hipGetLastError();                                                             \
    }
    IMH_FL(false, false) IMH_FL(true, false) IMH_FL(false, true) IMH_FL(true, true)
#undef IMH_FL
    return hipErrorInvalidValue;
}
#endif

}  // namespace imh
