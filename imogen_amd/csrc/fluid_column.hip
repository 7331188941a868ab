// fluid_column.hip -- the fused fluid sweep (relaxingFluid.m) along y or z as a column
// march: one thread per pencil, consecutive lanes on consecutive x, so every load and
// store of a wave is one contiguous run (a pencil along y or z is strided in memory, and
// the chunked block layout of fluid_sweep_impl.h can only give short runs there).
//
// Each thread walks its pencil three times, keeping a few cells of every field in
// registers: (1) the freezing speed C = max(|u| + c_f) of U^n; (2) the half step U^(n+1/2)
// and its freezing speed C'; (3) the corrector, the limited fluxes and the new state.
// The expressions and their evaluation order are those of the chunked kernel, so both
// give the same values (the FP64 form reproduces the CPU restatement).  Pass 3 re-reads
// what passes 1 and 2 read (from the caches or HBM).
//
// Used for the sweeps along y and z whose both boundaries are circular or constant
// (their ghost values are a[0] or a[N-1] of the shifted array) and whose flux objects
// already hold their edges; other cases run the chunked kernel.
#include "fluid_math.h"
#include "imh_common.h"
#include "imh_kernels.h"

namespace imh {

namespace {

constexpr int COL_BS = 256;

template <typename T, bool MHD>
struct Column {
    const FluidArgs<T>& ap;
    T* q[5];
    const T* b[3];
    unsigned base, stride;
    int N, d, X;
    MagUse use;
    double g;
    int lim;

    __device__ __forceinline__ void load(int i, T (&Q)[5], T (&B)[3]) const {
        const int ic = i < 0 ? 0 : (i > N - 1 ? N - 1 : i);
        const unsigned o = base + unsigned(ic) * stride;
#pragma unroll
        for (int v = 0; v < 5; ++v) Q[v] = gmem(q[v])[o];
#pragma unroll
        for (int v = 0; v < 3; ++v) B[v] = MHD ? gmem(b[v])[o] : T(0);
    }
    // pressure.m 'totsnd', velocity along the sweep and the freezing-speed candidate
    __device__ __forceinline__ T speeds(const T (&Q)[5], const T (&B)[3], T& P, T& U,
                                        T& IR) const {
        T c;
        FMath<T>::press(Q[0], Q[1], Q[2], Q[3], Q[4], B[0], B[1], B[2], use, g, P, c, IR);
        U = FMath<T>::divrho(sel3(d, Q[1], Q[2], Q[3]), IR);
        return fabs(U) + c;
    }
    // wFluidFlux.m: FR / FL of every field of one cell
    __device__ __forceinline__ void fluxes(const T (&Q)[5], const T (&B)[3], T C, T rC,
                                           T (&FR)[5], T (&FL)[5]) const {
        T P, U, IR;
        (void)speeds(Q, B, P, U, IR);
        const T bX = sel3(d, B[0], B[1], B[2]);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            T w;
            if (v == 0) {
                w = FMath<T>::div(sel3(d, Q[1], Q[2], Q[3]), C, rC);
            } else if (v == 4) {
                T bdotp = B[0] * Q[1] + B[1] * Q[2] + B[2] * Q[3];
                T we = U * (Q[4] + P) - FMath<T>::divrho(bX * bdotp, IR);
                w = FMath<T>::div(we, C, rC);
            } else {
                T dv = (v == X) ? T(1) : T(0);
                w = FMath<T>::div(U * Q[v] + P * dv - bX * B[v - 1], C, rC);
            }
            FR[v] = T(0.5) * (Q[v] + w);
            FL[v] = T(0.5) * (Q[v] - w);
        }
    }
    // the half step of one cell from its fluxes and its neighbours'
    __device__ __forceinline__ void half(const T (&Q)[5], const T (&FRm)[5], const T (&FR)[5],
                                         const T (&FL)[5], const T (&FLp)[5], T hfC, T minmass,
                                         T (&QH)[5]) const {
#pragma unroll
        for (int v = 0; v < 5; ++v) QH[v] = Q[v] - hfC * (FR[v] - FRm[v] + FL[v] - FLp[v]);
        QH[0] = fmax(QH[0], minmass);  // mass.store.array = max(mass.store.array, MINMASS)
    }
    // limited fluxes of one cell from the corrector fluxes of it and its neighbours
    __device__ __forceinline__ void limited(const T (&Rm)[5], const T (&R)[5], const T (&Rp)[5],
                                            const T (&Lm)[5], const T (&L)[5], const T (&Lp)[5],
                                            T (&R2)[5], T (&L2)[5]) const {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            R2[v] = R[v] + limit(lim, R[v] - Rm[v], Rp[v] - R[v]);
            L2[v] = L[v] + limit(lim, Lm[v] - L[v], L[v] - Lp[v]);
        }
    }
};

template <typename T>
__device__ __forceinline__ void copy5(T (&dst)[5], const T (&src)[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) dst[v] = src[v];
}
template <typename T>
__device__ __forceinline__ void pick5(T (&dst)[5], bool c, const T (&a)[5], const T (&b)[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) dst[v] = c ? a[v] : b[v];
}

}  // namespace

template <typename T, bool MHD>
__device__ __forceinline__ void fluid_column_body(const FluidArgs<T>& ap, long long tile) {
    if (!ap.force && !ap.sim->active) return;
    const long long pen = tile * COL_BS + threadIdx.x;
    if (pen >= ap.npencil) return;
    Column<T, MHD> c{ap};
#pragma unroll
    for (int v = 0; v < 5; ++v) c.q[v] = ap.q[v];
#pragma unroll
    for (int v = 0; v < 3; ++v) c.b[v] = ap.bc[v];
    const int nn[3] = {ap.n0, ap.n1, ap.n2};
    c.d = ap.d;
    c.X = ap.d + 1;
    c.N = sel3i(ap.d, ap.n0, ap.n1, ap.n2);
    long long b64, s64;
    pencil_base(nn, ap.d, pen, b64, s64);
    c.base = unsigned(b64);
    c.stride = unsigned(s64);
    c.use = ap.use;
    c.g = ap.gamma;
    c.lim = ap.limiter;
    const int N = c.N;
    const bool circ_lo = mode_at(ap.modes, 0, ap.d) == IMH_BC_CIRC;
    const bool circ_hi = mode_at(ap.modes, 1, ap.d) == IMH_BC_CIRC;
    const double dt = ap.sim->dt;
    const double ff = dt / ap.dx;
    const T minmass = T(ap.minmass);

    T Q[5], B[3], P, U, IR;

    // ---- pass 1: C = max(|u| + c_f) over the pencil (NaN-ignoring, as MATLAB max) ----
    T C = T(NAN);
    for (int i = 0; i < N; ++i) {
        c.load(i, Q, B);
        C = fmax(C, c.speeds(Q, B, P, U, IR));
    }
    const T rC = FMath<T>::rcp(C);
    const T hfC = T(0.5 * ff) * C;

    // predictor fluxes of one cell
    auto pflux = [&](int i, T (&FR)[5], T (&FL)[5]) __attribute__((always_inline)) {
        T Qi[5], Bi[3];
        c.load(i, Qi, Bi);
        c.fluxes(Qi, Bi, C, rC, FR, FL);
    };
    // ghosts of the predictor: store.fluxR shifted down at cell 0, store.fluxL up at N-1
    T gpR[5], gpL[5];
    {
        T R0[5], L0[5], RN[5], LN[5];
        pflux(0, R0, L0);
        pflux(N - 1, RN, LN);
        pick5(gpR, circ_lo, RN, R0);
        pick5(gpL, circ_hi, L0, LN);
    }

    // ---- pass 2: the half step and C' -------------------------------------------------
    T C2 = T(NAN);
    {
        T FRm[5], FRc[5], FLc[5], FRn[5], FLn[5], Qc[5], Bc[3];
        copy5(FRm, gpR);
        c.load(0, Qc, Bc);
        c.fluxes(Qc, Bc, C, rC, FRc, FLc);
        for (int i = 0; i < N; ++i) {
            T Qn[5], Bn[3];
            c.load(i + 1, Qn, Bn);
            c.fluxes(Qn, Bn, C, rC, FRn, FLn);
            if (i + 1 >= N) copy5(FLn, gpL);
            T QH[5];
            c.half(Qc, FRm, FRc, FLc, FLn, hfC, minmass, QH);
            C2 = fmax(C2, c.speeds(QH, Bc, P, U, IR));
            copy5(FRm, FRc);
            copy5(FRc, FRn);
            copy5(FLc, FLn);
            copy5(Qc, Qn);
#pragma unroll
            for (int v = 0; v < 3; ++v) Bc[v] = Bn[v];
        }
    }
    const T rC2 = FMath<T>::rcp(C2);
    const T ffC = T(ff) * C2;

    // half step of cell i (its own loads and fluxes)
    auto hstep = [&](int i, T (&QH)[5], T (&Bi)[3]) __attribute__((always_inline)) {
        T Qi[5], Rm[5], Lm[5], R[5], L[5], Rp[5], Lp[5];
        pflux(i - 1, Rm, Lm);
        pflux(i + 1, Rp, Lp);
        c.load(i, Qi, Bi);
        c.fluxes(Qi, Bi, C, rC, R, L);
        if (i == 0) copy5(Rm, gpR);
        if (i == N - 1) copy5(Lp, gpL);
        c.half(Qi, Rm, R, L, Lp, hfC, minmass, QH);
    };
    // corrector fluxes of cell i
    auto cflux = [&](int i, T (&FR)[5], T (&FL)[5]) __attribute__((always_inline)) {
        T QH[5], Bi[3];
        hstep(i, QH, Bi);
        c.fluxes(QH, Bi, C2, rC2, FR, FL);
    };

    // ---- ghosts of the corrector and of the final update ------------------------------
    T gcRlo[5], gcLlo[5], gcRhi[5], gcLhi[5], gfR[5], gfL[5];
    T R0[5], L0[5], R1[5], L1[5];  // corrector fluxes of cells 0 and 1 (the march's start)
    {
        T RN[5], LN[5], RN1[5], LN1[5];
        cflux(0, R0, L0);
        cflux(N - 1, RN, LN);
        pick5(gcRlo, circ_lo, RN, R0);
        pick5(gcLlo, circ_lo, LN, L0);
        pick5(gcRhi, circ_hi, R0, RN);
        pick5(gcLhi, circ_hi, L0, LN);
        cflux(1, R1, L1);
        cflux(N - 2, RN1, LN1);
        T A2[5], B2[5], C2R[5], D2[5];
        // limited fluxes of cell 0 (neighbours: ghost below, cell 1 above) and of N-1
        c.limited(gcRlo, R0, R1, gcLlo, L0, L1, A2, B2);
        c.limited(RN1, RN, gcRhi, LN1, LN, gcLhi, C2R, D2);
        pick5(gfR, circ_lo, C2R, A2);
        pick5(gfL, circ_hi, B2, D2);
    }

    // ---- pass 3: corrector, limiter and the new state ---------------------------------
    {
        // carried: FRp(i+1), FRp(i+2), FLp(i+2), Q(i+2), B(i+2), FRc/FLc(i), (i+1),
        // FR2(i-1), FR2(i), FL2(i)
        T Rp1[5], Rp2[5], Lp2[5], Q2[5], B2v[3];
        T Rc0[5], Rc1[5], Lc0[5], Lc1[5], F2m[5], F2c[5], G2c[5];
        {
            T Lp1[5];
            pflux(1, Rp1, Lp1);
            pflux(2, Rp2, Lp2);
            c.load(2, Q2, B2v);
        }
        copy5(Rc0, R0);
        copy5(Lc0, L0);
        copy5(Rc1, R1);
        copy5(Lc1, L1);
        copy5(F2m, gfR);
        // FR2 / FL2 of cell 0
        c.limited(gcRlo, Rc0, Rc1, gcLlo, Lc0, Lc1, F2c, G2c);
        for (int i = 0; i < N; ++i) {
            // predictor fluxes of cell i+3 (FL past the end: the upper ghost)
            T Rp3[5], Lp3[5], Q3[5], B3[3];
            c.load(i + 3, Q3, B3);
            c.fluxes(Q3, B3, C, rC, Rp3, Lp3);
            if (i + 3 >= N) copy5(Lp3, gpL);
            // half step and corrector fluxes of cell i+2 (past the end: the upper ghosts)
            T QH[5], Rc2[5], Lc2[5];
            c.half(Q2, Rp1, Rp2, Lp2, Lp3, hfC, minmass, QH);
            c.fluxes(QH, B2v, C2, rC2, Rc2, Lc2);
            if (i + 2 >= N) {
                copy5(Rc2, gcRhi);
                copy5(Lc2, gcLhi);
            }
            // limited fluxes of cell i+1 (past the end: the final ghost of FL)
            T F2n[5], G2n[5];
            c.limited(Rc0, Rc1, Rc2, Lc0, Lc1, Lc2, F2n, G2n);
            if (i + 1 >= N) copy5(G2n, gfL);
            // new state of cell i: U^n - dt/dx C' (FR2 - FR2(-1) + FL2 - FL2(+1))
            {
                const unsigned o = c.base + unsigned(i) * c.stride;
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    T qn = gmem(c.q[v])[o] - ffC * (F2c[v] - F2m[v] + G2c[v] - G2n[v]);
                    if (v == 0) qn = fmax(qn, minmass);  // mass.array = max(mass, MINMASS)
                    gmem(c.q[v])[o] = qn;
                }
            }
            copy5(F2m, F2c);
            copy5(F2c, F2n);
            copy5(G2c, G2n);
            copy5(Rc0, Rc1);
            copy5(Rc1, Rc2);
            copy5(Lc0, Lc1);
            copy5(Lc1, Lc2);
            copy5(Rp1, Rp2);
            copy5(Rp2, Rp3);
            copy5(Lp2, Lp3);
            copy5(Q2, Q3);
#pragma unroll
            for (int v = 0; v < 3; ++v) B2v[v] = B3[v];
        }
    }
}

template <typename T, bool MHD>
__global__ void __launch_bounds__(COL_BS) fluid_column_kernel(const FluidArgs<T>* args,
                                                              const BatchTile* tiles,
                                                              int ntiles) {
    for (int r = 0;; ++r) {
        int arg = 0;
        long long tile;
        if (tiles) {
            const int t = xcd_round_tile(r, ntiles);
            if (t < 0) break;
            uniform_tile(tiles, t, arg, tile);
        } else {
            if (r > 0 || (int)blockIdx.x >= ntiles) break;
            tile = xcd_tile(blockIdx.x, ntiles);  // neighbouring tiles on one XCD
        }
        fluid_column_body<T, MHD>(args[arg], tile);
    }
}

template <typename T>
hipError_t launch_fluid_column(const FluidArgs<T>* args, const BatchTile* tiles, int ntiles,
                               int nblocks, bool mhd, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    if (mhd)
        hipLaunchKernelGGL((fluid_column_kernel<T, true>), dim3(nblocks), dim3(COL_BS), 0, st,
                           args, tiles, ntiles);
    else
        hipLaunchKernelGGL((fluid_column_kernel<T, false>), dim3(nblocks), dim3(COL_BS), 0, st,
                           args, tiles, ntiles);
    return hipGetLastError();
}

template hipError_t launch_fluid_column<double>(const FluidArgs<double>*, const BatchTile*, int,
                                                int, bool, hipStream_t);
template hipError_t launch_fluid_column<float>(const FluidArgs<float>*, const BatchTile*, int,
                                               int, bool, hipStream_t);

int fluid_column_pencils() { return COL_BS; }

}  // namespace imh
