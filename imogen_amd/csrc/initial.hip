// initial.hip -- initial conditions generated on the device (the benchmark and test
// problems), so a simulation starts without a host-side state or a host-to-device copy.
//
// One cell's conserved state is a pure function of its coordinates and the problem
// parameters (ic_cell); the fill kernel writes it in the context's precision, and the
// edge kernel writes the FP64 face values that the FluidArray / MagnetArray constructors
// record as the transparent / wall edge stores (the values the host upload takes from its
// FP64 arrays).  The formulas and their evaluation order are those of the host
// generators in imogen_amd/problems.py (which follow the reference's initializers,
// experiment/*Initializer.m); the KH perturbation draws from a counter-based generator
// instead of numpy's.
#include "imh_common.h"
#include "imh_host.h"
#include "imh_kernels.h"

namespace imh {

namespace {

// splitmix64 of (seed, counter): a uniform double in [-1, 1)
__device__ inline double uniform_pm1(unsigned long long seed, unsigned long long ctr) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (ctr + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// mass, mom x/y/z, ener, mag x/y/z of cell (i, j, k)
__device__ void ic_cell(const IcParams& P, int i, int j, int k, double out[8]) {
    const double g = P.gamma;
    for (int v = 0; v < 8; ++v) out[v] = 0.0;
    switch (P.kind) {
    case IMH_IC_SOD: {
        const bool right = i >= P.ip[0] - 1;  // MATLAB round(N/2), 1-based
        out[0] = right ? 1.0 * 0.125 : 1.0;
        out[4] = right ? (1.0 / (g - 1.0)) * 0.1 : 1.0 / (g - 1.0);
        break;
    }
    case IMH_IC_ORSZAG_TANG: {
        const double dx = P.dx[0];
        const double x = double(i) * dx, y = double(j) * dx;
        const double rho = 25.0 / (36.0 * M_PI);
        const double Pr = 5.0 / (12.0 * M_PI);
        const double b0 = 1.0 / sqrt(4.0 * M_PI);
        out[0] = rho;
        out[1] = -rho * sin(y);
        out[2] = rho * sin(x);
        out[5] = -b0 * sin(y);
        out[6] = b0 * sin(2.0 * x);
        out[4] = Pr / (g - 1.0) + 0.5 * (out[1] * out[1] + out[2] * out[2] + out[3] * out[3]) / rho +
                 0.5 * (out[5] * out[5] + out[6] * out[6] + out[7] * out[7]);
        break;
    }
    case IMH_IC_SEDOV: {
        // p[0] radius in cells, p[1] ambient pressure, p[2] energy; ip[0] cells inside
        const double dx = P.dx[0];
        const double cx = (double(i) + 0.5) * dx - 0.5, cy = (double(j) + 0.5) * dx - 0.5,
                     cz = (double(k) + 0.5) * dx - 0.5;
        const double r2 = cz * cz + cy * cy + cx * cx;
        const double rr = P.p[0] * dx;
        out[0] = 1.0;
        out[4] = P.p[1] / (g - 1.0);
        if (r2 <= rr * rr) out[4] += P.p[2] / (double(P.ip[0]) * pow(dx, 3.0));
        break;
    }
    case IMH_IC_RIEMANN: {
        // p[0..7] left rho, P, v x/y/z, B y/z; p[8..15] right; p[16] B x; ip[0] split
        const double* s = i < P.ip[0] ? P.p : P.p + 8;
        const double rho = s[0], pr = s[1];
        const double v[3] = {s[2], s[3], s[4]};
        const double b[3] = {P.p[16], s[5], s[6]};
        out[0] = rho;
        for (int c = 0; c < 3; ++c) {
            out[1 + c] = rho * v[c];
            out[5 + c] = b[c];
        }
        out[4] = pr / (g - 1.0) + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) +
                 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        break;
    }
    case IMH_IC_KELVIN_HELMHOLTZ: {
        // p[0] shear speed, p[1] density ratio, p[2] perturbation; ip[0] ceil(ny/2), seed
        const bool low = j >= P.ip[0] - 1;
        const double mass = low ? 1.0 / P.p[1] : 1.0;
        const unsigned long long cell =
            (unsigned long long)i + (unsigned long long)P.n[0] * ((unsigned long long)j +
                                                                  (unsigned long long)P.n[1] * k);
        const double vx = (low ? -P.p[0] : P.p[0]) * (1.0 + P.p[2] * uniform_pm1(P.seed, 2 * cell));
        const double vy = P.p[2] * P.p[0] * uniform_pm1(P.seed, 2 * cell + 1);
        out[0] = mass;
        out[1] = mass * vx;
        out[2] = mass * vy;
        out[4] = 1.0 / (g - 1.0) + 0.5 * (out[1] * out[1] + out[2] * out[2]) / mass;
        break;
    }
    case IMH_IC_ORSZAG_TANG_REF: {
        // OrszagTangVortexInitializer.m:72-90: x, y are cell indices, ip[0..1] = grid - 1
        const double mass = 25.0 / (36.0 * M_PI);
        const double x = double(i), y = double(j);
        const double sy = sin(2.0 * M_PI * y / double(P.ip[1]));
        const double sx = sin(2.0 * M_PI * x / double(P.ip[0]));
        const double mag0 = 1.0 / sqrt(4.0 * M_PI);
        out[0] = mass;
        out[1] = -mass * sy;
        out[2] = mass * sx;
        out[5] = -mag0 * sy;
        out[6] = mag0 * sin(4.0 * M_PI * x / double(P.ip[0]));
        const double S = (out[1] * out[1] + out[2] * out[2]) + out[3] * out[3];
        const double B = (out[5] * out[5] + out[6] * out[6]) + out[7] * out[7];
        out[4] = (5.0 / (12.0 * M_PI) / (g - 1.0) + 0.5 * S / mass) + 0.5 * B;
        break;
    }
    case IMH_IC_SEDOV_REF: {
        // SedovTaylorBlastWaveInitializer.m:77-86: integer offsets from the centre cell
        // (ip[0..2] = floor(grid / 2)), ip[3] radius in cells, p[0] the inside pressure
        const long long X = i - P.ip[0], Y = j - P.ip[1], Z = k - P.ip[2];
        const long long r = P.ip[3];
        const double pr = (X * X + Y * Y + Z * Z > r * r) ? 1e-5 : P.p[0];
        out[0] = 1.0;
        out[4] = pr / (g - 1.0);
        break;
    }
    case IMH_IC_MAGNETIC_SHOCK_TUBE: {
        // MagneticShockTubeInitializer.m:79-117: ip[0] direction, ip[1] MATLAB round(N/2)
        // (1-based first cell of the right side), ip[2..4] field components on
        const int c[3] = {i, j, k};
        const int dir = P.ip[0];
        const bool right = c[dir] >= P.ip[1] - 1;
        const double b = 5.0 / sqrt(4.0 * M_PI);
        out[0] = 1.0;
        for (int q = 0; q < 3; ++q) out[5 + q] = P.ip[2 + q] ? b : 0.0;
        // mom(dir) = 10; on the right mom(dir) = -1.0 * mom(1) (the x component's value,
        // 0 unless dir is x)
        const double m1 = dir == 0 ? 10.0 : 0.0;
        out[1 + dir] = right ? -1.0 * m1 : 10.0;
        const double S = (out[1] * out[1] + out[2] * out[2]) + out[3] * out[3];
        const double B = (out[5] * out[5] + out[6] * out[6]) + out[7] * out[7];
        double e = 0.5 * S / out[0] + 0.5 * B;
        e = right ? e + 1.0 / (g - 1.0) : e + 20.0 / (g - 1.0);
        out[4] = e;
        break;
    }
    default: break;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) ic_fill_kernel(IcParams P, T* f0, T* f1, T* f2, T* f3,
                                                      T* f4, T* f5, T* f6, T* f7) {
    const long long ncell = (long long)P.n[0] * P.n[1] * P.n[2];
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int i = int(c % P.n[0]);
    const int j = int((c / P.n[0]) % P.n[1]);
    const int k = int(c / ((long long)P.n[0] * P.n[1]));
    double v[8];
    ic_cell(P, i, j, k, v);
    T* f[8] = {f0, f1, f2, f3, f4, f5, f6, f7};
#pragma unroll
    for (int q = 0; q < 8; ++q) f[q][c] = T(v[q]);
}

// the face values of field `field` on side `side` of dimension d into edges[off + face]
__global__ void __launch_bounds__(256) ic_edges_kernel(IcParams P, int field, int side, int d,
                                                       double* edges, long long off,
                                                       long long nface) {
    const long long fi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= nface) return;
    long long base, stride;
    pencil_base(P.n, d, fi, base, stride);
    const long long cell = base + (side ? (long long)(P.n[d] - 1) : 0) * stride;
    const int i = int(cell % P.n[0]);
    const int j = int((cell / P.n[0]) % P.n[1]);
    const int k = int(cell / ((long long)P.n[0] * P.n[1]));
    double v[8];
    ic_cell(P, i, j, k, v);
    edges[off + fi] = v[field];
}

}  // namespace

hipError_t launch_ic_fill(const IcParams& P, void* const f[8], bool fp64, hipStream_t st) {
    const long long ncell = (long long)P.n[0] * P.n[1] * P.n[2];
    const unsigned nb = unsigned((ncell + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(ic_fill_kernel<double>, dim3(nb), dim3(256), 0, st, P,
                           (double*)f[0], (double*)f[1], (double*)f[2], (double*)f[3],
                           (double*)f[4], (double*)f[5], (double*)f[6], (double*)f[7]);
    else
        hipLaunchKernelGGL(ic_fill_kernel<float>, dim3(nb), dim3(256), 0, st, P, (float*)f[0],
                           (float*)f[1], (float*)f[2], (float*)f[3], (float*)f[4], (float*)f[5],
                           (float*)f[6], (float*)f[7]);
    return hipGetLastError();
}

hipError_t launch_ic_edges(const IcParams& P, int field, int side, int d, double* edges,
                           long long off, long long nface, hipStream_t st) {
    if (nface <= 0) return hipSuccess;
    hipLaunchKernelGGL(ic_edges_kernel, dim3(unsigned((nface + 255) / 256)), dim3(256), 0, st,
                       P, field, side, d, edges, off, nface);
    return hipGetLastError();
}

}  // namespace imh
