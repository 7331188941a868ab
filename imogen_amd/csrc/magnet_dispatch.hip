// magnet_dispatch.hip -- cell-centred B kernels and the choice of the magnet variant
// for a launch.
#include "magnet_sweep_impl.h"

namespace imh {

template <typename T>
hipError_t launch_magnet_batch(int which, const MagnetArgs<T>* args, const BatchTile* tiles,
                               int ntiles, int nblocks, int kmax, size_t smem, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    if (smem > 150 * 1024) return hipErrorInvalidValue;
    constexpr int esz = (int)sizeof(T);
#define IMH_MAG_PICK(E, K) \
    if (esz == E && kmax == K) \
        return IMH_MAG_LAUNCHER(E, K)(which, args, tiles, ntiles, nblocks, smem, st);
    IMH_MAG_VARIANTS(IMH_MAG_PICK)
#undef IMH_MAG_PICK
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_cell_centered_batch(const CellMagArgs<T>* args, const BatchTile* tiles,
                                      int ntiles, int nblocks, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL((cell_centered_batch_kernel<T>), dim3(nblocks), dim3(256), 0, st, args,
                       tiles, ntiles);
    return hipGetLastError();
}

int magnet_kmax_for(int N, int TL) {
    int K = (N + TL - 1) / TL;
    return K <= 1 ? 1 : K <= 2 ? 2 : K <= 4 ? 4 : K <= 8 ? 8 : K <= 16 ? 16 : -1;
}

template hipError_t launch_magnet_batch<double>(int, const MagnetArgs<double>*, const BatchTile*,
                                                int, int, int, size_t, hipStream_t);
template hipError_t launch_magnet_batch<float>(int, const MagnetArgs<float>*, const BatchTile*,
                                               int, int, int, size_t, hipStream_t);
template hipError_t launch_cell_centered_batch<double>(const CellMagArgs<double>*,
                                                       const BatchTile*, int, int, hipStream_t);
template hipError_t launch_cell_centered_batch<float>(const CellMagArgs<float>*,
                                                      const BatchTile*, int, int, hipStream_t);

}  // namespace imh
