// fluid_math.h -- arithmetic shared by the fluid sweep kernels (fluid_sweep_impl.h and
// fluid_column.hip).
#pragma once

#include "imh_common.h"

namespace imh {

// Register-array element chosen by a run-time index in 0..2 (selects, not indexing, so
// the arrays stay in registers).
template <typename T>
__device__ inline T sel3(int i, T a, T b, T c) {
    return i == 0 ? a : (i == 1 ? b : c);
}

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

}  // namespace imh
