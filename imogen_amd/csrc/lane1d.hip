// lane1d.hip -- one launch per batched iteration for the one-dimensional MHD simulations
// of a batch (FP32 bulk).
//
// A 1-D simulation (grid N x 1 x 1) is one pencil: every sweep of its iteration is a
// single tile, so the per-launch plan of the batch spends about a dozen grid-wide
// launches per iteration on work that one workgroup holds.  Here one workgroup runs the simulation's whole flux call (flux.m:37-62 for X = 1, the only
// direction with >= 3 cells) back to back, with workgroup barriers between the steps
// (one workgroup per simulation):
//
//   order +1: relaxingFluid(X=1); magnetFlux(X=1, I=[2 3]): advection + CT for I = 2, then
//             I = 3 (relaxingMagnet.m:15-60); updateCellCentered of mag(3) and mag(1)
//   order -1: magnetFlux(X=1, I=[3 2]); updateCellCentered of mag(2) and mag(1);
//             relaxingFluid(X=1)
//
// The bodies are those of the per-launch kernels (fluid_sweep_body, magnet_advect_body,
// magnet_ct_body, cell_centered_cell), on argument blocks whose tile is one pencil over
// the workgroup (LANE1D_BS threads of LANE1D_KMAX, 3/4, 1/2 or 1/4 of it cells each: the
// bodies are instantiated for all four, their per-pencil LDS arrays sized for the one
// pencil, and a simulation runs the smallest that holds its pencil), so the results are
// those of the per-launch plan bit for bit.  (Each instantiation keeps LDS arrays of its
// own: the four take 152 KB of the 160.)  The time step comes from the batch-wide CFL launch before, the iteration
// counters from the batch-wide launch after.  The state stays in the workgroup's L2
// between the steps.
#define IMH_FLUID_ESZ 4
#define IMH_FLUID_NO_LAUNCHER
#include "fluid_sweep_impl.h"
#include "magnet_sweep_impl.h"

namespace imh {

namespace {

// The flux call of one simulation with KC cells per thread (the fluid and magnet bodies
// instantiated for that many): a short pencil does not run the longest pencil's body.
template <typename T, int KC>
__device__ __forceinline__ void lane1d_run(const Lane1dArgs<T>& a, unsigned char* s_mag) {
    auto cellmag = [&](int k) {
        const CellMagArgs<T> ca = a.ca[k];
        if (!ca.force && !ca.sim->active) return;
        const long long ncell = (long long)ca.n0 * ca.n1 * ca.n2;
        for (long long c = threadIdx.x; c < ncell; c += LANE1D_BS) cell_centered_cell(ca, c);
    };
    // steps 0 and 3: the fluid sweep; 1: magnetFlux(X=1, I=[2 3]) (order +1); 2:
    // magnetFlux(X=1, I=[3 2]) (order -1).  One call site per body (they are large).
    for (int step = 0; step < 4; ++step) {
        if (step == 0 || step == 3) {
            // one pencil; one field per group for the longer chunks (fewer registers)
            fluid_sweep_body<T, LANE1D_BS, KC, true, false, 1, (KC > 4 ? 1 : 0)>(a.fa, 0);
        } else {
            const int d1 = step == 1 ? 1 : 2, d2 = 3 - d1;
            for (int j = 0; j < 4; ++j) {
                const int dI = j < 2 ? d1 : d2;
                if ((j & 1) == 0)
                    magnet_advect_body<T, KC>(a.ma[dI], 0, s_mag);
                else
                    magnet_ct_body<T, KC>(a.ma[dI], 0, s_mag);
                __syncthreads();
            }
            // updateCellCentered of the last I and of mag(X)
            cellmag(d2);
            cellmag(0);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(LANE1D_BS) lane1d_kernel(const Lane1dArgs<T>* args,
                                                           int nsims) {
    // the magnet rows [2][rowlen] of the longest pencil
    __shared__ __attribute__((aligned(16))) unsigned char s_mag[2 * (LANE1D_NMAX + 4) * sizeof(T)];
    const int s = blockIdx.x;
    if (s >= nsims) return;
    const Lane1dArgs<T>& a = args[s];
    // cells per thread of this simulation (its fluid tile's K), uniform over the block
    const int kc = __builtin_amdgcn_readfirstlane(a.fa.K);
    if (kc <= LANE1D_KMAX / 4)
        lane1d_run<T, LANE1D_KMAX / 4>(a, s_mag);
    else if (kc <= LANE1D_KMAX / 2)
        lane1d_run<T, LANE1D_KMAX / 2>(a, s_mag);
    else if (kc <= 3 * LANE1D_KMAX / 4)
        lane1d_run<T, 3 * LANE1D_KMAX / 4>(a, s_mag);
    else
        lane1d_run<T, LANE1D_KMAX>(a, s_mag);
}

}  // namespace

hipError_t launch_lane1d(const void* args, int nsims, hipStream_t st) {
    if (nsims <= 0) return hipSuccess;
    hipLaunchKernelGGL(lane1d_kernel<float>, dim3(nsims), dim3(LANE1D_BS), 0, st,
                       (const Lane1dArgs<float>*)args, nsims);
    return hipGetLastError();
}

}  // namespace imh
