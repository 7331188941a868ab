// source.hip -- the source step that follows every flux call (source.m:16-31): gravity
// from a fixed potential and the optically thin radiation sink.
//
// Gravity: mom(i) -= dt * thresholdArray(mass) * d(phi)/dx_i with the 5-point centred
// derivative of the potential object (ImogenArray.calculate5PtDerivative,
// ImogenArray.m:189-192), and ener -= sum_i momSource_i * mom_i / mass with the momentum
// before its own update.  The potential's shifts follow its boundary conditions; a
// transparent shift nudges the potential's edge tracker once per call (Edges.getEdge),
// four calls per dimension per step, so the ghosts are produced by one thread per pencil
// end first (source_ghost_kernel), in the order the expression evaluates the shifts,
// and the per-cell kernel reads them.  Every operation keeps the .m text's order; this
// unit is built without FP contraction.  Fields may be FP32 (the FP32-bulk path); the
// arithmetic is FP64 in both precisions.
#include "imh_common.h"
#include "imh_host.h"
#include "imh_kernels.h"

namespace imh {

namespace {

// Ghost values of the potential's shifts at the ends of pencil f along d, per face:
// [0] shift(+2) at N-2, [1] shift(+2) at N-1, [2] shift(+1) at N-1,
// [3] shift(-1) at 0,   [4] shift(-2) at 0,   [5] shift(-2) at 1.
__global__ void __launch_bounds__(256) source_ghost_kernel(SourceArgs a) {
    if (!a.sim->active) return;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int d = -1;
    long long f = 0;
    for (int k = 0; k < a.ndims; ++k) {
        if (t >= a.goff[k] && t < a.goff[k] + a.nface[k]) {
            d = k;
            f = t - a.goff[k];
        }
    }
    if (d < 0) return;
    long long base, stride;
    pencil_base(a.n, d, f, base, stride);
    const int N = a.n[d];
    const double* p = a.phi + base;
    Ends e;
    e.a0 = p[0];
    e.a1 = p[(long long)(N > 1 ? 1 : N - 1) * stride];
    e.a2 = p[(long long)(N > 2 ? 2 : N - 1) * stride];
    e.b0 = p[(long long)(N >= 3 ? N - 3 : 0) * stride];
    e.b1 = p[(long long)(N >= 2 ? N - 2 : 0) * stride];
    e.b2 = p[(long long)(N - 1) * stride];
    double scratch[2] = {0.0, 0.0};
    double* tr[2];
    for (int s = 0; s < 2; ++s)
        tr[s] = a.eoff[s][d] >= 0 ? a.edges + a.eoff[s][d] + f : &scratch[s];
    GhostSpec up{a.modes[1][d], 1, 0, 0, a.active_lin[d], a.bc_inf};
    GhostSpec lo{a.modes[0][d], 0, 0, 0, a.active_lin[d], a.bc_inf};
    double g[6], unused;
    ghost_values(up, e, N, 2, tr[1], g[0], g[1]);
    ghost_values(up, e, N, 1, tr[1], g[2], unused);
    ghost_values(lo, e, N, 1, tr[0], g[3], unused);
    ghost_values(lo, e, N, 2, tr[0], g[4], g[5]);
    if (a.modes[0][d] == IMH_BC_WORMHOLE) {
        // wormhole_shift.m:21-57 (the potential has no vector component): shift(-1) at 0
        // takes cell 0 of the partner pencil, shift(-2) at 0 and 1 its cells 1 and 0
        long long pb, ps;
        pencil_base(a.n, d, worm_partner_face(a.n, d, f), pb, ps);
        const double* q = a.phi + pb;
        g[3] = q[0];
        g[4] = q[(long long)(N > 1 ? 1 : 0) * ps];
        g[5] = q[0];
    }
    double* out = a.ghosts + 6 * t;
    for (int k = 0; k < 6; ++k) out[k] = g[k];
}

// FADE: some boundary of the potential is a fade one (the blend is compiled only then)
template <typename T, bool FADE>
__global__ void __launch_bounds__(256) source_cell_kernel(SourceArgs a, T* mass, T* mx, T* my,
                                                          T* mz, T* ener, const T* bx,
                                                          const T* by, const T* bz) {
    if (!a.sim->active) return;
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const unsigned nx = unsigned(a.n[0]), ny = unsigned(a.n[1]);
    const unsigned uc = unsigned(c);
    const int co[3] = {int(uc % nx), int((uc / nx) % ny), int(uc / (nx * ny))};
    const long long face[3] = {co[1] + (long long)ny * co[2], co[0] + (long long)nx * co[2],
                               co[0] + (long long)nx * co[1]};
    const long long strd[3] = {1, (long long)nx, (long long)nx * ny};
    const double dt = a.sim->dt;
    const double rho = double(mass[c]);
    double m[3] = {double(mx[c]), double(my[c]), double(mz[c])};
    double E = double(ener[c]);
    if (a.phi) {
        const double th = a.mass_threshold;
        const double thr = th > 0 ? rho * ((rho > th) ? 1.0 : 0.0) : rho;
        double es = 0.0;
        const double p0 = a.phi[c];
        for (int d = 0; d < 3; ++d) {
            double s2 = p0, s1 = p0, b1 = p0, b2 = p0;
            if (d < a.ndims) {
                const int N = a.n[d], p = co[d];
                const long long st = strd[d];
                const double* g = a.ghosts + 6 * (a.goff[d] + face[d]);
                s2 = p < N - 2 ? a.phi[c + 2 * st] : g[p - (N - 2)];
                s1 = p < N - 1 ? a.phi[c + st] : g[2];
                b1 = p >= 1 ? a.phi[c - st] : g[3];
                b2 = p >= 2 ? a.phi[c - 2 * st] : g[4 + p];
                if (FADE && (a.modes[0][d] == IMH_BC_FADE || a.modes[1][d] == IMH_BC_FADE)) {
                    // fade_shift.m: the constant shift, then both ends of the shifted array
                    // blended toward its end values over bcInfinity cells
                    const double* row = a.phi + (c - (long long)p * st);
                    auto at = [&](int i) { return row[(long long)(i < 0 ? 0 : (i > N - 1 ? N - 1 : i)) * st]; };
                    auto faded = [&](int m, int upper, double r) {
                        if (a.modes[upper][d] != IMH_BC_FADE) return r;
                        const double v1 = at(upper ? m : -m), vN = at(upper ? N - 1 + m : N - 1 - m);
                        return fade_blend(at(upper ? p + m : p - m), p, N, v1, vN, a.fade, a.bc_inf);
                    };
                    s2 = faded(2, 1, s2);
                    s1 = faded(1, 1, s1);
                    b1 = faded(1, 0, b1);
                    b2 = faded(2, 0, b2);
                }
            }
            const double dg = a.dxl[d] ? a.dxl[d][co[d]] : a.dgrid[d];  // DGRID{d} here
            const double d5 = (((-s2 + 8.0 * s1) - 8.0 * b1) + b2) / (12.0 * dg);
            const double ms = (dt * thr) * d5;
            es = es + ms * m[d] / rho;
            m[d] = m[d] - ms;
        }
        E = E - es;
        mx[c] = T(m[0]);
        my[c] = T(m[1]);
        mz[c] = T(m[2]);
    }
    if (a.rad_type == 1) {
        // pressure(ENUM.PRESSURE_GAS, ...) with the momentum and energy after gravity
        const double pg = pressure_gas(rho, m[0], m[1], m[2], E, double(bx[c]), double(by[c]),
                                       double(bz[c]), a.use, a.gamma);
        const double rad = a.rad_strength * pow(rho, 2.0 - a.rad_exponent) * pow(pg, a.rad_exponent);
        E = E - dt * rad;
    }
    ener[c] = T(E);
}

template <typename T>
__global__ void __launch_bounds__(256) edges_from_field_kernel(const T* field, int n0, int n1,
                                                               int n2, int d, int side,
                                                               double* edges, long long nface) {
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int n[3] = {n0, n1, n2};
    long long base, stride;
    pencil_base(n, d, f, base, stride);
    edges[f] = double(field[base + (side ? (long long)(n[d] - 1) : 0) * stride]);
}

}  // namespace

hipError_t launch_edges_from_field(const void* field, bool fp64, const int n[3], int d,
                                   int side, double* edges, long long nface, hipStream_t st) {
    if (nface <= 0) return hipSuccess;
    const dim3 grid(unsigned((nface + 255) / 256));
    if (fp64)
        hipLaunchKernelGGL(edges_from_field_kernel<double>, grid, dim3(256), 0, st,
                           (const double*)field, n[0], n[1], n[2], d, side, edges, nface);
    else
        hipLaunchKernelGGL(edges_from_field_kernel<float>, grid, dim3(256), 0, st,
                           (const float*)field, n[0], n[1], n[2], d, side, edges, nface);
    return hipGetLastError();
}

hipError_t launch_source(const SourceArgs& a, void* const f[IMH_NFIELDS], bool fp64,
                         hipStream_t st) {
    const long long nghost = a.ndims > 0 ? a.goff[a.ndims - 1] + a.nface[a.ndims - 1] : 0;
    if (a.phi && nghost > 0)
        hipLaunchKernelGGL(source_ghost_kernel, dim3(unsigned((nghost + 255) / 256)), dim3(256),
                           0, st, a);
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    const dim3 grid(unsigned((ncell + 255) / 256));
    bool fade = false;
    for (int d = 0; d < a.ndims; ++d)
        fade = fade || a.modes[0][d] == IMH_BC_FADE || a.modes[1][d] == IMH_BC_FADE;
    if (fp64 && fade)
        hipLaunchKernelGGL((source_cell_kernel<double, true>), grid, dim3(256), 0, st, a,
                           (double*)f[IMH_F_MASS], (double*)f[IMH_F_MOMX], (double*)f[IMH_F_MOMY],
                           (double*)f[IMH_F_MOMZ], (double*)f[IMH_F_ENER],
                           (const double*)f[IMH_F_CMAGX], (const double*)f[IMH_F_CMAGY],
                           (const double*)f[IMH_F_CMAGZ]);
    else if (fp64)
        hipLaunchKernelGGL((source_cell_kernel<double, false>), grid, dim3(256), 0, st, a,
                           (double*)f[IMH_F_MASS], (double*)f[IMH_F_MOMX], (double*)f[IMH_F_MOMY],
                           (double*)f[IMH_F_MOMZ], (double*)f[IMH_F_ENER],
                           (const double*)f[IMH_F_CMAGX], (const double*)f[IMH_F_CMAGY],
                           (const double*)f[IMH_F_CMAGZ]);
    else if (fade)
        hipLaunchKernelGGL((source_cell_kernel<float, true>), grid, dim3(256), 0, st, a,
                           (float*)f[IMH_F_MASS], (float*)f[IMH_F_MOMX], (float*)f[IMH_F_MOMY],
                           (float*)f[IMH_F_MOMZ], (float*)f[IMH_F_ENER],
                           (const float*)f[IMH_F_CMAGX], (const float*)f[IMH_F_CMAGY],
                           (const float*)f[IMH_F_CMAGZ]);
    else
        hipLaunchKernelGGL((source_cell_kernel<float, false>), grid, dim3(256), 0, st, a,
                           (float*)f[IMH_F_MASS], (float*)f[IMH_F_MOMX], (float*)f[IMH_F_MOMY],
                           (float*)f[IMH_F_MOMZ], (float*)f[IMH_F_ENER],
                           (const float*)f[IMH_F_CMAGX], (const float*)f[IMH_F_CMAGY],
                           (const float*)f[IMH_F_CMAGZ]);
    return hipGetLastError();
}

}  // namespace imh
