// magnet_sweep_impl.h -- constrained-transport magnetic sweep (magnet/magnetFlux.m,
// magnet/relaxingMagnet.m, MagnetArray.updateCellCentered).
//
// One magnetFlux(X, [I1 I2]) call is, per component I:
//   magnet_advect_kernel  (pencils along X): grid-aligned velocity (shift of mass and
//                         mom(X) along I, smoothing along X), upwind predictor, TVD
//                         corrector with the limiter, B_I update; the limited flux is kept
//                         in a scratch array for the CT step;
//   magnet_ct_kernel      (pencils along X): G = F - shift_I(F,+1), B_X -= dt/dx_I *
//                         shift_X(G,-1);
// followed by cell_centered_kernel for the last I and for X.
//
// Each pencil lives in LDS with two ghost slots per side; every shift of the .m text is
// one exchange of the whole pencil (write, barrier, ghost cells from the BC rule with the
// object's edge tracker, barrier, read).  The magnet sweeps are a smaller share of the
// MHD step than the fluid sweep, so this kernel favours generality (all BC modes, +-2
// shifts, fade windows) over the register-chunk layout of the fluid sweep.
#include "imh_common.h"
#include "imh_kernels.h"

namespace imh {

namespace {

constexpr int MAG_BS = 256;
constexpr int MAG_PXMAX = 16;
constexpr int MAG_NMAX = 4096;   // LDS holds 2 pencils of up to this many cells in total

// Index helpers for a pencil of length N at LDS row `row`: slot 2+i holds cell i, slots
// 0..1 hold ghost cells -2..-1 and N+2..N+3 ghost cells N..N+1.
template <typename T>
struct Pencil {
    T* row;
    int N;
    __device__ T& at(int i) { return row[2 + i]; }
};

// Ends (a[0..2], a[N-3..N-1], clamped) of an LDS pencil, in FP64.
template <typename T>
__device__ inline Ends lds_ends(const T* row, int N) {
    auto lo = [&](int k) __attribute__((always_inline)) { return double(row[2 + (k < N - 1 ? k : N - 1)]); };
    auto hi = [&](int k) __attribute__((always_inline)) { int j = N - 3 + k; return double(row[2 + (j > 0 ? j : 0)]); };
    Ends e = {lo(0), lo(1), lo(2), hi(0), hi(1), hi(2)};
    return e;
}

// The BC code below runs only at pencil ends (or fade windows) and is kept out of line
// so the kernels stay small enough for the instruction cache.

// Ghost values of one shift call of an LDS pencil (m = 1 or 2 cells), written to out[].
template <typename T>
__device__ __noinline__ void mag_ghost_lds(int mode, int upper, int flip_sign, int zero_ghost,
                                           int active_lin, int bc_inf, const T* row, int N,
                                           int m, double* tracker, double* out) {
    GhostSpec s;
    s.mode = mode;
    s.upper = upper;
    s.flip_sign = flip_sign;
    s.zero_ghost = zero_ghost;
    s.active_linear = active_lin;
    s.bc_infinity = bc_inf;
    const Ends e = lds_ends(row, N);
    double g0, g1;
    ghost_values(s, e, N, m, tracker, g0, g1);
    out[0] = g0;
    if (m > 1) out[1] = g1;
}

template <typename T>
__device__ __noinline__ T mag_fade_lds(T r, const T* row, int N, int i, int m, const double* fade,
                                       int f) {
    int up = m > 0;
    int am = up ? m : -m;
    T v1 = up ? row[2 + (am < N - 1 ? am : N - 1)] : row[2];
    int jn = N - 1 - am;
    T vN = up ? row[2 + N - 1] : row[2 + (jn > 0 ? jn : 0)];
    return fade_blend(r, i, N, v1, vN, fade, f);
}

// shift(D, m) (m = +-1, +-2) of an LDS pencil at cell i, with ghost values g (the m
// values ghost_values() returned for this call) and the fade window when mode is fade.
template <typename T>
__device__ inline T lds_shift(const T* row, int N, int i, int m, const double* g, int mode,
                              const double* fade, int f) {
    int up = m > 0;
    int idx = i + m;
    T r;
    if (idx >= 0 && idx <= N - 1) r = row[2 + idx];
    else r = T(up ? (idx - N == 0 ? g[0] : g[1]) : (i == 0 ? g[0] : g[1]));
    if (mode == IMH_BC_FADE) r = mag_fade_lds(r, row, N, i, m, fade, f);
    return r;
}

// Ghost of a one-cell shift along dI of global array `arr` for the cell on the face of
// the line starting at row0.
template <typename T>
__device__ __noinline__ double transverse_ghost(const T* arr, long long row0, int NI, long long sI,
                                                int mode, int upper, int flip_sign,
                                                int zero_ghost, int active_lin, int bc_inf,
                                                double* tracker) {
    GhostSpec s;
    s.mode = mode;
    s.upper = upper;
    s.flip_sign = flip_sign;
    s.zero_ghost = zero_ghost;
    s.active_linear = active_lin;
    s.bc_infinity = bc_inf;
    auto lo = [&](int k) __attribute__((always_inline)) { return double(arr[row0 + (long long)(k < NI - 1 ? k : NI - 1) * sI]); };
    auto hi = [&](int k) __attribute__((always_inline)) { int j = NI - 3 + k; return double(arr[row0 + (long long)(j > 0 ? j : 0) * sI]); };
    Ends e = {lo(0), lo(1), lo(2), hi(0), hi(1), hi(2)};
    double g0, g1;
    ghost_values(s, e, NI, 1, tracker, g0, g1);
    return g0;
}

template <typename T>
__device__ __noinline__ T transverse_fade(T r, const T* arr, long long row0, int jI, int NI,
                                          long long sI, int upper, const double* fade, int f) {
    int i1 = 1 < NI - 1 ? 1 : NI - 1;
    int jn = NI - 2 > 0 ? NI - 2 : 0;
    T v1 = upper ? arr[row0 + i1 * sI] : arr[row0];
    T vN = upper ? arr[row0 + (long long)(NI - 1) * sI] : arr[row0 + jn * sI];
    return fade_blend(r, jI, NI, v1, vN, fade, f);
}

// The ghost of a one-cell transverse shift for the modes whose ghost is an element of
// the array (circ, const, mirror, flip, zero): the element ghost_values() picks, read
// directly (no call, one load; none when it is the cell itself), the same value.  Returns
// false for the modes that need the edge tracker or a blend (trans, wall, fade).
// row0 is the cell of the line at coordinate 0 along dI; self the line's cell at jI.
template <typename T>
__device__ __forceinline__ bool transverse_index_ghost(const T* arr, long long row0, int jI, int NI,
                                                       long long sI, const GhostSpec& s, T self,
                                                       T& out) {
    int src;
    bool neg = false;
    switch (s.mode) {
    case IMH_BC_CIRC: src = s.upper ? 0 : NI - 1; break;
    case IMH_BC_CONST: src = s.upper ? NI - 1 : 0; break;
    case IMH_BC_MIRROR:  // lower a[min(1, N-1)], upper a[max(N-2, 0)], DIRECT negated
        src = s.upper ? (NI - 2 > 0 ? NI - 2 : 0) : (NI - 1 < 1 ? NI - 1 : 1);
        neg = s.flip_sign != 0;
        break;
    case IMH_BC_FLIP: src = s.upper ? NI - 1 : 0; neg = s.flip_sign != 0; break;
    case IMH_BC_ZERO:
        if (s.zero_ghost) {
            out = T(0);
            return true;
        }
        src = s.upper ? NI - 1 : 0;
        break;
    default: return false;
    }
    const T v = src == jI ? self : arr[row0 + (long long)src * sI];
    out = neg ? -v : v;
    return true;
}

// Transverse shift along dimension dI of global array `arr` at the cell whose
// coordinate along dI is jI: shift(I,-1) (upper=0) or shift(I,+1) (upper=1), one cell.
// tracker is this cell's stored edge (used when the cell lies on that face).
template <typename T>
__device__ inline T transverse_shift(const T* arr, long long cell, int jI, int NI,
                                     long long sI, bool enabled, const GhostSpec& s,
                                     double* tracker, const double* fade, int f) {
    if (!enabled) return arr[cell];
    T r;
    bool ghost = s.upper ? (jI >= NI - 1) : (jI <= 0);
    if (!ghost) r = arr[cell + (s.upper ? sI : -sI)];
    else if (!transverse_index_ghost(arr, cell - (long long)jI * sI, jI, NI, sI, s, arr[cell], r))
        r = T(transverse_ghost(arr, cell - (long long)jI * sI, NI, sI, s.mode, s.upper,
                               s.flip_sign, s.zero_ghost, s.active_linear, s.bc_infinity,
                               tracker));
    if (s.mode == IMH_BC_FADE)
        r = transverse_fade(r, arr, cell - (long long)jI * sI, jI, NI, sI, s.upper, fade, f);
    return r;
}

// transverse_shift with arr[cell] (self) and the interior neighbour arr[cell -+ sI] (nb,
// any value where the shift needs a ghost) already loaded: the callers issue those loads
// for all their cells at once and use them unconditionally, so the compiler cannot sink
// them into per-cell branches (where they would run one memory round trip at a time).
template <typename T>
__device__ inline T transverse_shift_pre(const T* arr, long long cell, int jI, int NI,
                                         long long sI, bool enabled, const GhostSpec& s,
                                         double* tracker, const double* fade, int f, T self,
                                         T nb) {
    if (!enabled) return self;
    T r;
    bool ghost = s.upper ? (jI >= NI - 1) : (jI <= 0);
    if (!ghost) r = nb;
    else if (!transverse_index_ghost(arr, cell - (long long)jI * sI, jI, NI, sI, s, self, r))
        r = T(transverse_ghost(arr, cell - (long long)jI * sI, NI, sI, s.mode, s.upper,
                               s.flip_sign, s.zero_ghost, s.active_linear, s.bc_infinity,
                               tracker));
    if (s.mode == IMH_BC_FADE)
        r = transverse_fade(r, arr, cell - (long long)jI * sI, jI, NI, sI, s.upper, fade, f);
    return r;
}

__device__ inline long long coord_of(long long cell, const int n[3], int d) {
    // fewer than 2^31 cells: 32-bit divisions
    const unsigned uc = unsigned(cell), n0 = unsigned(n[0]), n1 = unsigned(n[1]);
    const unsigned c0 = uc % n0;
    const unsigned c1 = (uc / n0) % n1;
    const unsigned c2 = uc / (n0 * n1);
    return d == 0 ? c0 : (d == 1 ? c1 : c2);
}

// First assignment of an object (InitializedArray.m:43-47): `cell` records `val` in the
// object's edge stores of the dims other than dX whose face it lies on.
__device__ __noinline__ void mag_init_face(double* edges, const long long* eoff, int modes,
                                           int ndims, int dX, int obj, long long cell, int n0,
                                           int n1, int n2, double val) {
    const int nn[3] = {n0, n1, n2};
    const int c0 = int(coord_of(cell, nn, 0));
    const int c1 = int(coord_of(cell, nn, 1));
    const int c2 = int(coord_of(cell, nn, 2));
    for (int dd = 0; dd < ndims; ++dd) {
        if (dd == dX) continue;
        for (int s = 0; s < 2; ++s) {
            int mode = mode_at(modes, s, dd);
            if (mode != IMH_BC_TRANS && mode != IMH_BC_WALL) continue;
            if (sel3i(dd, c0, c1, c2) != (s ? sel3i(dd, n0, n1, n2) - 1 : 0)) continue;
            edges[eoff[(obj * 2 + s) * 3 + dd] + face_index(nn, dd, c0, c1, c2)] = val;
        }
    }
}

}  // namespace

// ------------------------------------------------------------------------------------
// Advection sub-step for component I along X (one pencil per PX group of threads).
// ------------------------------------------------------------------------------------
template <typename T, int KMAX>
__device__ __forceinline__ void magnet_advect_body(const MagnetArgs<T>& ap, long long tile,
                                                   unsigned char* smem) {
    // private copy: one read of the argument block per tile; stores to the fields
    // cannot alias it, so its fields are not re-read after every store
    const MagnetArgs<T> a = ap;
    if (!a.force && !a.sim->active) return;
    const int dX = a.dX, dI = a.dI;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int N = sel3i(dX, a.n0, a.n1, a.n2);
    const int rowlen = N + 4;
    T* lds = reinterpret_cast<T*>(smem);                // [2][PX][rowlen]
    __shared__ double s_g[MAG_PXMAX][8];                 // ghost values per call

    const int tid = threadIdx.x;
    const int px = tid / a.TL;
    const int ty = tid % a.TL;
    const long long pen = tile * a.PX + px;
    const bool pen_ok = pen < a.npencil && px < a.PX;
    const int K = (N + a.TL - 1) / a.TL;
    // cells are dealt to the lanes cyclically (cell ty + k * TL): loads, stores and LDS
    // rows are touched at consecutive addresses by consecutive lanes
    const int TLc = a.TL;
    int kc = ty < N ? (N - 1 - ty) / TLc + 1 : 0;
    kc = kc < 0 ? 0 : (kc > K ? K : kc);
    if (!pen_ok) kc = 0;
    const bool ghost_thread = pen_ok && ty == 0;
    const int f = a.bc_inf;

    long long base = 0, stride = 1;
    int pc[3] = {0, 0, 0};  // coordinates of the pencil's first cell
    if (pen_ok) {
        pencil_base(nn, dX, pen, base, stride);
        pencil_coords(nn, dX, pen, pc);
    }
    const double dt = a.sim->dt;
    const double ff = 0.5 * dt / a.dxX;       // relaxingMagnet.m:15
    const double ff2 = 2.0 * ff;              // relaxingMagnet.m:32

    const int mlo = mode_at(a.modes, 0, dX), mhi = mode_at(a.modes, 1, dX);
    const int oVG = OBJ_VG + 3 * dI + dX, oSFR = OBJ_MSFR + 3 * dI + dX;
    const int oWM = OBJ_WMAG + 3 * dI + dX, oFL = OBJ_MFLUX + 3 * dI + dX;
    const int Xc = dX + 1;                    // field index of mom(X)
    const bool trk_lo = (mlo == IMH_BC_TRANS || mlo == IMH_BC_WALL);
    const bool trk_hi = (mhi == IMH_BC_TRANS || mhi == IMH_BC_WALL);
    const bool I_enabled = dI < a.ndims;
    const long long sI = dI == 0 ? 1 : (dI == 1 ? (long long)a.n0 : (long long)a.n0 * a.n1);
    const int NI = sel3i(dI, a.n0, a.n1, a.n2);

    T* row0 = lds + (size_t)(0 * a.PX + (pen_ok ? px : 0)) * rowlen;
    T* row1 = lds + (size_t)(1 * a.PX + (pen_ok ? px : 0)) * rowlen;

    // edge trackers of this pencil along X (ghost thread only)
    auto eidx = [&](int obj, int side) __attribute__((always_inline)) -> long long {
        return a.eoff[(obj * 2 + side) * 3 + dX] + pen;
    };
    double evg_lo = 0, evg_hi = 0, esfr_lo = 0, esfr_hi = 0, ewm_lo = 0, ewm_hi = 0, efl_lo = 0;
    if (ghost_thread) {
        if (trk_lo) {
            if (!a.init_vg) evg_lo = a.edges[eidx(oVG, 0)];
            if (!a.init_store) esfr_lo = a.edges[eidx(oSFR, 0)];
            if (!a.init_wmag) ewm_lo = a.edges[eidx(oWM, 0)];
            if (!a.init_flux) efl_lo = a.edges[eidx(oFL, 0)];
        }
        if (trk_hi) {
            if (!a.init_vg) evg_hi = a.edges[eidx(oVG, 1)];
            if (!a.init_store) esfr_hi = a.edges[eidx(oSFR, 1)];
            if (!a.init_wmag) ewm_hi = a.edges[eidx(oWM, 1)];
        }
    }
    // ghosts of one shift call along X (component I != X: no sign flips) into out[]
    auto gh = [&](int upper, int zero_ghost, const T* row, int m, double* trk, double* out)
        __attribute__((always_inline)) {
        const int mode = upper ? mhi : mlo;
        if (mode == IMH_BC_CONST || mode == IMH_BC_CIRC) {
            // ghost_values() of these modes (N >= 3), without the out-of-line call, whose
            // register saves around it cost the waiting block a memory round trip:
            // const repeats the end cell, circ wraps (lower N-m+t, upper t)
            for (int t = 0; t < m; ++t) {
                const int src = mode == IMH_BC_CONST ? (upper ? N - 1 : 0) : (upper ? t : N - m + t);
                out[t] = double(row[2 + src]);
            }
            return;
        }
        mag_ghost_lds<T>(mode, upper, 0, zero_ghost, a.active_linX, f, row, N, m, trk, out);
    };
    // record the edges of a freshly assigned object on all faces (InitializedArray.m:43-47)
    auto init_faces = [&](int obj, const T (&vals)[KMAX]) __attribute__((always_inline)) {
        if (a.ndims < 2) return;
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            if (k < kc)
                mag_init_face(a.edges, a.eoff, a.modes, a.ndims, dX, obj,
                              base + (long long)((ty + k * TLc)) * stride, a.n0, a.n1, a.n2,
                              double(vals[k]));
    };
    auto put = [&](T* row, const T (&v)[KMAX]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            if (k < kc) row[2 + (ty + k * TLc)] = v[k];
    };

    // ---- grid-aligned velocity: (shift_I(p_X,-1) + p_X) ./ (shift_I(rho,-1) + rho) ----
    T vg[KMAX], bI[KMAX];
    {
        const int mIlo = mode_at(a.modes, 0, dI);
        GhostSpec sm = {mIlo, 0, 0, 0, a.active_linI, f};  // FluidArray: no flips
        GhostSpec sr = sm;
        sm.flip_sign = (Xc == dI + 1);  // mom(X).component == I ?  (X != I, so never)
        // every load first, at addresses clamped into the pencil (padding cells read the
        // pencil's last cell; threads without a pencil read from cell 0)
        const int jI = pc[dI];  // I != X: constant along the pencil
        const long long nbo = (I_enabled && jI > 0) ? -sI : 0;  // interior lower neighbour
        T m0[KMAX], r0[KMAX], mn[KMAX], rn[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const int i = (ty + k * TLc) < N - 1 ? (ty + k * TLc) : N - 1;
            const long long cell = base + (long long)i * stride;
            m0[k] = gmem(a.mom)[cell];
            r0[k] = gmem(a.rho)[cell];
            mn[k] = gmem(a.mom)[cell + nbo];
            rn[k] = gmem(a.rho)[cell + nbo];
            bI[k] = gmem(a.bI)[cell];
        }
        // the shifted values: interior neighbours unless this pencil's row needs the BC
        // rule (its I-coordinate on the lower face, or a fade window)
        const bool slow = I_enabled && (jI <= 0 || mIlo == IMH_BC_FADE);
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            T sm_v = I_enabled ? mn[k] : m0[k];
            T sr_v = I_enabled ? rn[k] : r0[k];
            if (slow && k < kc) {
                long long cell = base + (long long)((ty + k * TLc)) * stride;
                double* tm = nullptr;
                double* tr = nullptr;
                double em = 0, er = 0;
                long long fI = 0;
                bool trkI = (mIlo == IMH_BC_TRANS || mIlo == IMH_BC_WALL) &&
                            jI == 0 && I_enabled;
                if (trkI) {
                    fI = face_index(nn, dI, dX == 0 ? (ty + k * TLc) : pc[0], dX == 1 ? (ty + k * TLc) : pc[1],
                                    dX == 2 ? (ty + k * TLc) : pc[2]);
                    em = a.edges[a.eoff[((OBJ_PRIM + Xc) * 2 + 0) * 3 + dI] + fI];
                    er = a.edges[a.eoff[((OBJ_PRIM + 0) * 2 + 0) * 3 + dI] + fI];
                    tm = &em;
                    tr = &er;
                }
                sm_v = transverse_shift_pre(a.mom, cell, jI, NI, sI, I_enabled, sm, tm, a.fade, f,
                                            m0[k], mn[k]);
                sr_v = transverse_shift_pre(a.rho, cell, jI, NI, sI, I_enabled, sr, tr, a.fade, f,
                                            r0[k], rn[k]);
                if (trkI) {
                    a.edges[a.eoff[((OBJ_PRIM + Xc) * 2 + 0) * 3 + dI] + fI] = em;
                    a.edges[a.eoff[((OBJ_PRIM + 0) * 2 + 0) * 3 + dI] + fI] = er;
                }
            }
            // evaluated for every k (padding cells hold clamped copies), masked after
            const T v = (sm_v + m0[k]) / (sr_v + r0[k]);
            vg[k] = k < kc ? v : T(0);
            bI[k] = k < kc ? bI[k] : T(0);
        }
    }
    if (a.init_vg) init_faces(oVG, vg);

    // ---- smoothing: vg = 0.25*(vg.shift(X,-1) + vg.shift(X,1) + 2*vg) -----------------
    put(row0, vg);
    __syncthreads();
    if (ghost_thread) {
        if (a.init_vg) { evg_lo = double(row0[2]); evg_hi = double(row0[2 + N - 1]); }
        gh(0, 0, row0, 1, &evg_lo, &s_g[px][0]);
        gh(1, 0, row0, 1, &evg_hi, &s_g[px][1]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            int i = (ty + k * TLc);
            T s1 = lds_shift(row0, N, i, -1, &s_g[px][0], mlo, a.fade, f);
            T s2 = lds_shift(row0, N, i, 1, &s_g[px][1], mhi, a.fade, f);
            vg[k] = T(0.25) * (s1 + s2 + T(2) * vg[k]);
        }
    }

    // ---- velocityFlow = (vg + vg.shift(X,1)) < 0 -------------------------------------
    T vf[KMAX];
    put(row1, vg);
    __syncthreads();
    if (ghost_thread) gh(1, 0, row1, 1, &evg_hi, &s_g[px][2]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            T s = lds_shift(row1, N, (ty + k * TLc), 1, &s_g[px][2], mhi, a.fade, f);
            vf[k] = (vg[k] + s < T(0)) ? T(1) : T(0);
        } else {
            vf[k] = T(0);
        }
    }

    // ---- predictor: store.fluxR = B.*vg; upwind select; B_half ----------------------
    T w[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) w[k] = bI[k] * vg[k];
    if (a.init_store) init_faces(oSFR, w);
    put(row0, w);
    __syncthreads();
    if (ghost_thread) {
        if (a.init_store) { esfr_lo = double(row0[2]); esfr_hi = double(row0[2 + N - 1]); }
        gh(1, 0, row0, 1, &esfr_hi, &s_g[px][3]);  // FLUXR keeps the upper side
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            T s = lds_shift(row0, N, (ty + k * TLc), 1, &s_g[px][3], mhi, a.fade, f);
            w[k] = w[k] * (T(1) - vf[k]) + s * vf[k];
        }
    }
    put(row1, w);
    __syncthreads();
    if (ghost_thread) gh(0, 1, row1, 1, &esfr_lo, &s_g[px][4]);  // FLUXR: zero-mode lower ghost 0
    __syncthreads();
    T st[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            T s = lds_shift(row1, N, (ty + k * TLc), -1, &s_g[px][4], mlo, a.fade, f);
            // fluxFactor = 0.5*dTime ./ DGRID{X}: per cell on a non-uniform grid
            const double fk = a.dxlX ? 0.5 * dt / a.dxlX[ty + k * TLc] : ff;
            st[k] = bI[k] - T(fk) * (w[k] - s);
        } else {
            st[k] = T(0);
        }
    }

    // ---- corrector: wMag = store.*vg; flux; dFluxR, dFluxL; limiter ------------------
#pragma unroll
    for (int k = 0; k < KMAX; ++k) w[k] = st[k] * vg[k];
    if (a.init_wmag) init_faces(oWM, w);
    put(row0, w);
    __syncthreads();
    if (ghost_thread) {
        if (a.init_wmag) { ewm_lo = double(row0[2]); ewm_hi = double(row0[2 + N - 1]); }
        // wMag is an InitializedArray (not a FluxArray): zero mode keeps constant ghosts
        gh(1, 0, row0, 1, &ewm_hi, &s_g[px][0]);  // flux = ... shift(X,1)
        gh(1, 0, row0, 1, &ewm_hi, &s_g[px][1]);  // dFluxR: shift(X,1)
        gh(1, 0, row0, 2, &ewm_hi, &s_g[px][2]);  // dFluxR: shift(X,2) (slots 2, 3)
        gh(1, 0, row0, 1, &ewm_hi, &s_g[px][4]);  // dFluxL: shift(X,1)
        gh(0, 0, row0, 1, &ewm_lo, &s_g[px][5]);  // dFluxL: shift(X,-1)
    }
    __syncthreads();
    T F[KMAX];
    const int lim = a.limiter;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            int i = (ty + k * TLc);
            T a1 = lds_shift(row0, N, i, 1, &s_g[px][0], mhi, a.fade, f);
            F[k] = w[k] * (T(1) - vf[k]) + a1 * vf[k];
        } else {
            F[k] = T(0);
        }
    }
    // flux(X) is first assigned here (pre-limiter): its edges come from this array.
    if (a.init_flux) {
        init_faces(oFL, F);
        if (ghost_thread) efl_lo = double(F[0]);
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            int i = (ty + k * TLc);
            T b1 = lds_shift(row0, N, i, 1, &s_g[px][1], mhi, a.fade, f);
            T b2 = lds_shift(row0, N, i, 2, &s_g[px][2], mhi, a.fade, f);
            T dR = b1 - (vf[k] * b2 + (T(1) - vf[k]) * w[k]);
            T c1 = lds_shift(row0, N, i, 1, &s_g[px][4], mhi, a.fade, f);
            T c2 = lds_shift(row0, N, i, -1, &s_g[px][5], mlo, a.fade, f);
            T dL = w[k] - (vf[k] * c1 + (T(1) - vf[k]) * c2);
            F[k] = F[k] + lim_delta(lim, dL, dR);
        }
    }

    // ---- B_I = B_I - 2*fluxFactor .* (flux - flux.shift(X,-1)) -------------------------
    put(row1, F);
    __syncthreads();
    if (ghost_thread) gh(0, 1, row1, 1, &efl_lo, &s_g[px][6]);  // 'flux' type: zero ghost 0
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            long long cell = base + (long long)((ty + k * TLc)) * stride;
            T s = lds_shift(row1, N, (ty + k * TLc), -1, &s_g[px][6], mlo, a.fade, f);
            const double f2k = a.dxlX ? 2.0 * (0.5 * dt / a.dxlX[ty + k * TLc]) : ff2;
            gmem(a.bI)[cell] = bI[k] - T(f2k) * (F[k] - s);
            gmem(a.flux)[cell] = F[k];
        }
    }

    if (ghost_thread) {
        if (trk_lo) {
            a.edges[eidx(oVG, 0)] = evg_lo;
            a.edges[eidx(oSFR, 0)] = esfr_lo;
            a.edges[eidx(oWM, 0)] = ewm_lo;
            a.edges[eidx(oFL, 0)] = efl_lo;
        }
        if (trk_hi) {
            a.edges[eidx(oVG, 1)] = evg_hi;
            a.edges[eidx(oSFR, 1)] = esfr_hi;
            a.edges[eidx(oWM, 1)] = ewm_hi;
        }
    }
}

// ------------------------------------------------------------------------------------
// Constrained transport: flux = flux - flux.shift(I,1); flux = flux.shift(X,-1);
// mag(X) = mag(X) - dt/DGRID{I} .* flux   (relaxingMagnet.m:56-60).
// ------------------------------------------------------------------------------------
template <typename T, int KMAX>
__device__ __forceinline__ void magnet_ct_body(const MagnetArgs<T>& ap, long long tile,
                                               unsigned char* smem) {
    const MagnetArgs<T> a = ap;  // private copy (see magnet_advect_body)
    if (!a.force && !a.sim->active) return;
    const int dX = a.dX, dI = a.dI;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int N = sel3i(dX, a.n0, a.n1, a.n2);
    const int rowlen = N + 4;
    T* lds = reinterpret_cast<T*>(smem);
    __shared__ double s_g[MAG_PXMAX][2];

    const int tid = threadIdx.x;
    const int px = tid / a.TL;
    const int ty = tid % a.TL;
    const long long pen = tile * a.PX + px;
    const bool pen_ok = pen < a.npencil && px < a.PX;
    const int K = (N + a.TL - 1) / a.TL;
    // cells are dealt to the lanes cyclically (cell ty + k * TL): loads, stores and LDS
    // rows are touched at consecutive addresses by consecutive lanes
    const int TLc = a.TL;
    int kc = ty < N ? (N - 1 - ty) / TLc + 1 : 0;
    kc = kc < 0 ? 0 : (kc > K ? K : kc);
    if (!pen_ok) kc = 0;
    const bool ghost_thread = pen_ok && ty == 0;
    const int f = a.bc_inf;
    long long base = 0, stride = 1;
    int pc[3] = {0, 0, 0};  // coordinates of the pencil's first cell
    if (pen_ok) {
        pencil_base(nn, dX, pen, base, stride);
        pencil_coords(nn, dX, pen, pc);
    }
    // fluxFactor = dTime ./ DGRID{I}: the pencil's own I coordinate on a non-uniform grid
    const double ffI = a.sim->dt / (a.dxlI ? a.dxlI[pc[dI]] : a.dxI);
    const int oFL = OBJ_MFLUX + 3 * dI + dX;
    const int mlo = mode_at(a.modes, 0, dX);
    const bool I_enabled = dI < a.ndims;
    const long long sI = dI == 0 ? 1 : (dI == 1 ? (long long)a.n0 : (long long)a.n0 * a.n1);
    const int NI = sel3i(dI, a.n0, a.n1, a.n2);
    T* row = lds + (size_t)(pen_ok ? px : 0) * rowlen;

    // G = F - shift_I(F, +1): flux object, component I == DIRECT (flips for mirror/flip),
    // type 'flux' (zero mode: ghost 0)
    const int mIhi = mode_at(a.modes, 1, dI);
    GhostSpec sI_spec = {mIhi, 1, 1, 1, a.active_linI, f};
    const bool trkI = (mIhi == IMH_BC_TRANS || mIhi == IMH_BC_WALL) && I_enabled;
    T G[KMAX];
    // loads first, clamped into the pencil and used unconditionally (see the advection)
    const int jI = pc[dI];  // I != X: constant along the pencil
    const long long nbo = (I_enabled && jI < NI - 1) ? sI : 0;  // interior upper neighbour
    T f0[KMAX], fn[KMAX], bx0[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int i = (ty + k * TLc) < N - 1 ? (ty + k * TLc) : N - 1;
        const long long cell = base + (long long)i * stride;
        f0[k] = gmem(a.flux)[cell];
        fn[k] = gmem(a.flux)[cell + nbo];
        bx0[k] = gmem(a.bX)[cell];
    }
    const bool slow = I_enabled && (jI >= NI - 1 || mIhi == IMH_BC_FADE);
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        T s = I_enabled ? fn[k] : f0[k];
        if (slow && k < kc) {
            long long cell = base + (long long)((ty + k * TLc)) * stride;
            double et = 0;
            double* tp = nullptr;
            long long fI = 0;
            bool onface = trkI && jI == NI - 1;
            if (onface) {
                fI = face_index(nn, dI, dX == 0 ? (ty + k * TLc) : pc[0], dX == 1 ? (ty + k * TLc) : pc[1],
                                dX == 2 ? (ty + k * TLc) : pc[2]);
                et = a.edges[a.eoff[(oFL * 2 + 1) * 3 + dI] + fI];
                tp = &et;
            }
            s = transverse_shift_pre((const T*)a.flux, cell, jI, NI, sI, I_enabled, sI_spec, tp,
                                     a.fade, f, f0[k], fn[k]);
            if (onface) a.edges[a.eoff[(oFL * 2 + 1) * 3 + dI] + fI] = et;
        }
        G[k] = k < kc ? f0[k] - s : T(0);
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
        if (k < kc) row[2 + (ty + k * TLc)] = G[k];
    __syncthreads();
    if (ghost_thread) {
        double tr = 0;
        bool trk = (mlo == IMH_BC_TRANS || mlo == IMH_BC_WALL);
        long long ei = a.eoff[(oFL * 2 + 0) * 3 + dX] + pen;
        if (trk) tr = a.edges[ei];
        if (mlo == IMH_BC_CONST || mlo == IMH_BC_CIRC)  // ghost_values() of these modes, inline
            s_g[px][0] = double(row[2 + (mlo == IMH_BC_CONST ? 0 : N - 1)]);
        else
            mag_ghost_lds<T>(mlo, 0, 0, 1, a.active_linX, f, row, N, 1, &tr, &s_g[px][0]);
        if (trk) a.edges[ei] = tr;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < kc) {
            long long cell = base + (long long)((ty + k * TLc)) * stride;
            T h = lds_shift(row, N, (ty + k * TLc), -1, &s_g[px][0], mlo, a.fade, f);
            gmem(a.bX)[cell] = bx0[k] - T(ffI) * h;
        }
    }
}

// ------------------------------------------------------------------------------------
// cellMag_k = 0.5*(B_k + B_k.shift(k, 1))  (MagnetArray.m:36-38), element-wise with the
// mag(k) edge tracker on the upper k face.
// ------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cell_centered_cell(const CellMagArgs<T>& a, long long cell) {
    const int d = a.k;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const bool enabled = d < a.ndims;
    const long long sd = d == 0 ? 1 : (d == 1 ? (long long)a.n0 : (long long)a.n0 * a.n1);
    const int Nd = sel3i(d, a.n0, a.n1, a.n2);
    int jd = int(coord_of(cell, nn, d));
    GhostSpec s = {mode_at(a.modes, 1, d), 1, 1, 0, a.active_lin, a.bc_inf};  // comp k == DIRECT
    bool trk = enabled && (s.mode == IMH_BC_TRANS || s.mode == IMH_BC_WALL) && jd == Nd - 1;
    double et = 0;
    long long fi = 0;
    if (trk) {
        fi = face_index(nn, d, int(coord_of(cell, nn, 0)), int(coord_of(cell, nn, 1)),
                        int(coord_of(cell, nn, 2)));
        et = a.edges[a.eoff[((OBJ_MAG + d) * 2 + 1) * 3 + d] + fi];
    }
    T sh = transverse_shift(a.b, cell, jd, Nd, sd, enabled, s, trk ? &et : nullptr, a.fade,
                            a.bc_inf);
    if (trk) a.edges[a.eoff[((OBJ_MAG + d) * 2 + 1) * 3 + d] + fi] = et;
    gmem(a.cb)[cell] = T(0.5) * (gmem(a.b)[cell] + sh);
}

template <typename T>
__device__ __forceinline__ void cell_centered_body(const CellMagArgs<T>& ap, long long chunk) {
    const CellMagArgs<T> a = ap;  // private copy (see magnet_advect_body)
    if (!a.force && !a.sim->active) return;
    const long long ncell = (long long)a.n0 * a.n1 * a.n2;
    const long long cell = chunk * 256 + threadIdx.x;
    if (cell < ncell) cell_centered_cell(a, cell);
}

// Persistent batched forms: blocks walk a tile list spanning many simulations.
// tiles == nullptr: the single simulation args[0], one block per tile.
// The (argument block, tile) a block works on in round t of a persistent launch; with
// tiles == nullptr the launch is one simulation, args[0], one block per tile.
__device__ inline bool next_tile(const BatchTile* tiles, int ntiles, int r, int& arg,
                                 long long& tile) {
    if (!tiles) {
        if (r > 0 || (int)blockIdx.x >= ntiles) return false;
        arg = 0;
        tile = xcd_tile(blockIdx.x, ntiles);  // neighbouring tiles on one XCD
        return true;
    }
    const int t = xcd_round_tile(r, ntiles);
    if (t < 0) return false;
    uniform_tile(tiles, t, arg, tile);
    return true;
}

// Persistent batched forms: blocks walk a tile list spanning many simulations (one call
// site of each body: they are large).
template <typename T, int KMAX>
__global__ void __launch_bounds__(MAG_BS) magnet_advect_batch_kernel(const MagnetArgs<T>* args,
                                                                     const BatchTile* tiles,
                                                                     int ntiles) {
    extern __shared__ unsigned char smem_raw[];
    int arg;
    long long tile;
    for (int r = 0; next_tile(tiles, ntiles, r, arg, tile); ++r) {
        magnet_advect_body<T, KMAX>(args[arg], tile, smem_raw);
        __syncthreads();
    }
}

template <typename T, int KMAX>
__global__ void __launch_bounds__(MAG_BS) magnet_ct_batch_kernel(const MagnetArgs<T>* args,
                                                                 const BatchTile* tiles,
                                                                 int ntiles) {
    extern __shared__ unsigned char smem_raw[];
    int arg;
    long long tile;
    for (int r = 0; next_tile(tiles, ntiles, r, arg, tile); ++r) {
        magnet_ct_body<T, KMAX>(args[arg], tile, smem_raw);
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) cell_centered_batch_kernel(const CellMagArgs<T>* args,
                                                                  const BatchTile* tiles,
                                                                  int ntiles) {
    int arg;
    long long tile;
    for (int r = 0; next_tile(tiles, ntiles, r, arg, tile); ++r)
        cell_centered_body<T>(args[arg], tile);
}

// ------------------------------------------------------------------------------------
// Launchers.  magnet_sweep_variant.hip is compiled once per entry of IMH_MAG_VARIANTS
// (imh_kernels.h) with IMH_MAG_ESZ / IMH_MAG_KMAX defined and holds that variant's
// advection and CT kernels; magnet_dispatch.hip holds the cell-centred kernels and picks
// the variant of a launch.
// ------------------------------------------------------------------------------------
#ifdef IMH_MAG_KMAX
#if IMH_MAG_ESZ == 8
using MagT = double;
#else
using MagT = float;
#endif

hipError_t IMH_MAG_LAUNCHER(IMH_MAG_ESZ, IMH_MAG_KMAX)(int which, const void* args,
                                                      const BatchTile* tiles, int ntiles,
                                                      int nblocks, size_t smem, hipStream_t st) {
    const MagnetArgs<MagT>* a = (const MagnetArgs<MagT>*)args;
    const void* fn = which == 0 ? (const void*)magnet_advect_batch_kernel<MagT, IMH_MAG_KMAX>
                                : (const void*)magnet_ct_batch_kernel<MagT, IMH_MAG_KMAX>;
    if (smem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           150 * 1024);
        if (e != hipSuccess) return e;
    }
    if (which == 0)
        hipLaunchKernelGGL((magnet_advect_batch_kernel<MagT, IMH_MAG_KMAX>), dim3(nblocks),
                           dim3(MAG_BS), smem, st, a, tiles, ntiles);
    else
        hipLaunchKernelGGL((magnet_ct_batch_kernel<MagT, IMH_MAG_KMAX>), dim3(nblocks),
                           dim3(MAG_BS), smem, st, a, tiles, ntiles);
    return hipGetLastError();
}
#endif

}  // namespace imh
