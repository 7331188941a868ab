// cfl.hip -- CFL timestep reduction (time/TimeManager.m:55-99), device-side time
// bookkeeping, and the compensated conservation totals.
//
// The CFL maximum of |p_i/rho| + c_s for i = 1..3 is reduced per block with wavefront
// shuffles and one LDS pass, then merged with one 64-bit atomicMax per block and
// direction (non-negative doubles order like their bit patterns).  dt is computed on the
// device, so an iteration needs no host round trip and can be replayed as a hipGraph.
#include "imh_common.h"
#include "imh_host.h"
#include "imh_kernels.h"

namespace imh {

namespace {

__device__ inline double wave_max(double v) {
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ inline void atomic_max_nonneg(unsigned long long* slot, double v) {
    if (!(v >= 0.0)) return;  // NaN: MATLAB max ignores it
    atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

}  // namespace

// Max of |p_i/rho| + c_s over cells [c0, c1) of one simulation (grid-strided by
// `stride`), merged into the simulation's slots with one atomic per block and direction.
template <typename T>
__device__ __forceinline__ void cfl_body(const CflArgs<T>& a, long long c0, long long c1,
                                         long long stride) {
    __shared__ double s_red[3][4];
    if (!a.sim->active) return;
    double m[3] = {NAN, NAN, NAN};
    // U cells per thread and round, every load of the round issued before any use (a
    // loop of one cell at a time waits a memory round trip per cell); cells past c1 read
    // cell c0 and are left out of the maxima
    constexpr int U = 4;
    for (long long cb = c0 + threadIdx.x; cb < c1; cb += stride * U) {
        T q[5][U], b[3][U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long c = cb + u * stride < c1 ? cb + u * stride : c0;
#pragma unroll
            for (int v = 0; v < 5; ++v) q[v][u] = gmem(a.q[v])[c];
#pragma unroll
            for (int k = 0; k < 3; ++k) b[k][u] = a.bc[k] ? gmem(a.bc[k])[c] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (cb + u * stride >= c1) continue;
            const double rho = double(q[0][u]);
            const double px = double(q[1][u]), py = double(q[2][u]), pz = double(q[3][u]);
            const double cs = pressure_sound(rho, px, py, pz, double(q[4][u]), double(b[0][u]),
                                             double(b[1][u]), double(b[2][u]), a.use, a.gamma);
            m[0] = fmax(m[0], fabs(px / rho) + cs);
            m[1] = fmax(m[1], fabs(py / rho) + cs);
            m[2] = fmax(m[2], fabs(pz / rho) + cs);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        double v = wave_max(m[k]);
        if (lane == 0) s_red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int k = threadIdx.x;
        double v = fmax(fmax(s_red[k][0], s_red[k][1]), fmax(s_red[k][2], s_red[k][3]));
        atomic_max_nonneg(&a.sim->cmax_bits[k], v);
    }
    __syncthreads();
}

constexpr long long CFL_CHUNK = 256 * 32;  // cells per tile of a batched launch

// dt of one simulation from its reduced speeds cmax_bits[3] (TimeManager.m:65-98).
__device__ inline void finalize_sim(SimDev& d, const SimParams& q, int step,
                                    const unsigned long long (&cmax_bits)[3]) {
    if (!d.active) return;
    if (step > 1 && q.update_mode != 0) return;
    double cmax = 1e-2;
    int gi = 0;
    for (int k = 0; k < 3; ++k) {
        double c = __longlong_as_double((long long)cmax_bits[k]);
        if (c > cmax) {
            cmax = c;
            gi = k + 1;
        }
    }
    if (gi < 1) {
        d.error = IMH_ECFL;
        d.active = 0;
        return;
    }
    d.grid_index = gi;
    double upd = double(q.update_mode + 1);
    d.dt = q.cfl * q.dgrid[gi - 1] / cmax;
    double newtime = d.time + upd * d.dt;
    if (newtime > q.time_max) {
        d.dt = (q.time_max - d.time) / upd;
        newtime = q.time_max;
    }
    d.time = newtime;
}

template <typename T>
__global__ void __launch_bounds__(256) cfl_batch_kernel(const CflArgs<T>* args,
                                                        const BatchTile* tiles, int ntiles) {
    for (int r = 0;; ++r) {
        const int t = tiles ? xcd_round_tile(r, ntiles) : int(blockIdx.x + r * gridDim.x);
        if (t < 0 || t >= ntiles) break;
        int arg = 0;
        long long tile = t;
        if (tiles) uniform_tile(tiles, t, arg, tile);
        const CflArgs<T> a = args[arg];  // private copy, read once per tile
        long long c0 = tile * a.chunk;
        long long c1 = c0 + a.chunk < a.ncell ? c0 + a.chunk : a.ncell;
        cfl_body(a, c0, c1, 256);
    }
}

template <typename T>
hipError_t launch_cfl_batch(const CflArgs<T>* args, const BatchTile* tiles, int ntiles,
                            int nblocks, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL((cfl_batch_kernel<T>), dim3(nblocks), dim3(256), 0, st, args, tiles, ntiles);
    return hipGetLastError();
}

template hipError_t launch_cfl_batch<double>(const CflArgs<double>*, const BatchTile*, int, int,
                                             hipStream_t);
template hipError_t launch_cfl_batch<float>(const CflArgs<float>*, const BatchTile*, int, int,
                                            hipStream_t);

long long cfl_chunk() { return CFL_CHUNK; }

// The kernels below run one thread per simulation over arrays of per-simulation
// pointers (a batch is a view over independently allocated simulations).

// TimeManager.running (TimeManager.m:36-39), evaluated at the top of the loop body.
// reset: also clear the CFL slots of the active simulations (the reset before the CFL
// reduction, done here when that reduction follows).
__global__ void begin_iteration_kernel(SimDev* const* sim, const SimParams* const* p, int nsims,
                                       int unconditional, int reset) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsims) return;
    SimDev& d = *sim[s];
    const SimParams& q = *p[s];
    int running = (d.iteration <= q.iter_max) && (d.time < q.time_max) && d.error == 0;
    d.active = (unconditional || running) ? 1 : 0;
    if (reset && d.active)
        for (int k = 0; k < 3; ++k) d.cmax_bits[k] = 0ull;
}

__global__ void cfl_reset_kernel(SimDev* const* sim, int nsims) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsims) return;
    SimDev& d = *sim[s];
    if (!d.active) return;
    for (int k = 0; k < 3; ++k) d.cmax_bits[k] = 0ull;
}

// dt from the reduced speeds (TimeManager.m:65-98).
__global__ void time_finalize_kernel(SimDev* const* sim, const SimParams* const* p, int nsims,
                                     int step) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsims) return;
    SimDev& d = *sim[s];
    const unsigned long long cm[3] = {d.cmax_bits[0], d.cmax_bits[1], d.cmax_bits[2]};
    finalize_sim(d, *p[s], step, cm);
}

// TimeManager.step: iteration = iteration + 1.
__global__ void end_iteration_kernel(SimDev* const* sim, int nsims) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsims) return;
    if (sim[s]->active) sim[s]->iteration += 1;
}

static inline int nblk(int n) { return (n + 255) / 256; }

hipError_t launch_begin_iteration(SimDev* const* sim, const SimParams* const* p, int nsims,
                                  int unconditional, hipStream_t st, int reset) {
    hipLaunchKernelGGL(begin_iteration_kernel, dim3(nblk(nsims)), dim3(256), 0, st, sim, p, nsims,
                       unconditional, reset);
    return hipGetLastError();
}
hipError_t launch_cfl_reset(SimDev* const* sim, int nsims, hipStream_t st) {
    hipLaunchKernelGGL(cfl_reset_kernel, dim3(nblk(nsims)), dim3(256), 0, st, sim, nsims);
    return hipGetLastError();
}
hipError_t launch_time_finalize(SimDev* const* sim, const SimParams* const* p, int nsims,
                                int step, hipStream_t st) {
    hipLaunchKernelGGL(time_finalize_kernel, dim3(nblk(nsims)), dim3(256), 0, st, sim, p, nsims,
                       step);
    return hipGetLastError();
}
hipError_t launch_end_iteration(SimDev* const* sim, int nsims, hipStream_t st) {
    hipLaunchKernelGGL(end_iteration_kernel, dim3(nblk(nsims)), dim3(256), 0, st, sim, nsims);
    return hipGetLastError();
}

// ---- conservation totals -------------------------------------------------------------
// Neumaier-compensated sums; every block owns a fixed contiguous range and combines its
// 256 partials in a fixed binary tree, the block partials are combined the same way, so
// the result is independent of timing and of the device.

struct CSum {
    double s, c;
};

__device__ inline CSum csum_add(CSum a, double x) {
    double t = a.s + x;
    double e = (fabs(a.s) >= fabs(x)) ? (a.s - t) + x : (x - t) + a.s;
    return {t, a.c + e};
}

__device__ inline CSum csum_merge(CSum a, CSum b) {
    CSum r = csum_add(a, b.s);
    r.c += b.c;
    return r;
}

template <typename T>
__global__ void __launch_bounds__(256) csum_blocks_kernel(const T* x, long long n, long long chunk,
                                                          double* partials) {
    __shared__ double ss[256], sc[256];
    const long long b0 = (long long)blockIdx.x * chunk;
    long long b1 = b0 + chunk;
    if (b1 > n) b1 = n;
    CSum acc = {0.0, 0.0};
    const long long per = (chunk + 255) / 256;
    long long s0 = b0 + (long long)threadIdx.x * per;
    long long s1 = s0 + per;
    if (s1 > b1) s1 = b1;
    for (long long i = s0; i < s1; ++i) acc = csum_add(acc, double(x[i]));
    ss[threadIdx.x] = acc.s;
    sc[threadIdx.x] = acc.c;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            CSum r = csum_merge({ss[threadIdx.x], sc[threadIdx.x]},
                                {ss[threadIdx.x + w], sc[threadIdx.x + w]});
            ss[threadIdx.x] = r.s;
            sc[threadIdx.x] = r.c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = ss[0];
        partials[2 * blockIdx.x + 1] = sc[0];
    }
}

__global__ void csum_final_kernel(const double* partials, int nb, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    CSum acc = {0.0, 0.0};
    for (int b = 0; b < nb; ++b) acc = csum_merge(acc, {partials[2 * b], partials[2 * b + 1]});
    *out = acc.s + acc.c;
}

template <typename T>
hipError_t launch_conservation(const T* const* fields, int nfields, long long ncell,
                               double* partials, int nblocks, double* out, hipStream_t st) {
    long long chunk = (ncell + nblocks - 1) / nblocks;
    if (chunk < 1) chunk = 1;
    int nb = int((ncell + chunk - 1) / chunk);
    for (int f = 0; f < nfields; ++f) {
        hipLaunchKernelGGL((csum_blocks_kernel<T>), dim3(nb), dim3(256), 0, st, fields[f], ncell,
                           chunk, partials + (size_t)f * 2 * nblocks);
        hipLaunchKernelGGL(csum_final_kernel, dim3(1), dim3(1), 0, st,
                           partials + (size_t)f * 2 * nblocks, nb, out + f);
    }
    return hipGetLastError();
}

template hipError_t launch_conservation<double>(const double* const*, int, long long, double*, int,
                                                double*, hipStream_t);
template hipError_t launch_conservation<float>(const float* const*, int, long long, double*, int,
                                               double*, hipStream_t);

// ---- device copy probe ----------------------------------------------------------------
// A plain streaming copy (16-byte vectors, grid-stride): the reference kernel of known
// traffic for scaling HBM counters and the measured copy bandwidth beside the roofline.
__global__ void __launch_bounds__(256) copy_probe_kernel(const float4* __restrict__ src,
                                                         float4* __restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

}  // namespace imh

extern "C" int imh_copy_probe(size_t bytes, int reps, double* ms_per_copy) {
    using namespace imh;
    if (bytes < 16 || reps < 1 || !ms_per_copy) return fail(IMH_EINVAL, "bad argument");
    const long long n = (long long)(bytes / 16);
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t st = nullptr;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMemset(a, 0, n * 16);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int nb = 0;
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, 0);
        nb = prop.multiProcessorCount * 8;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(copy_probe_kernel, dim3(nb), dim3(256), 0, st, (const float4*)a,
                           (float4*)b, n);  // warm-up
        e = hipEventRecord(e0, st);
    }
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(copy_probe_kernel, dim3(nb), dim3(256), 0, st, (const float4*)a,
                           (float4*)b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(e, "copy probe");
    *ms_per_copy = ms / reps;
    return IMH_OK;
}
