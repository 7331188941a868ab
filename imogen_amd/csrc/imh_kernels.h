// imh_kernels.h -- kernel argument blocks and launcher declarations.
#pragma once

#include "imh_common.h"

namespace imh {

constexpr int MAX_FADE = 64;

// BC modes [side][dim] packed 4 bits each (kernel arguments are never indexed at run
// time: that would copy the whole argument block to scratch memory).
__host__ __device__ inline int mode_at(int packed, int side, int dim) {
    return (packed >> (4 * (3 * side + dim))) & 15;
}
inline int pack_modes(const int m[2][3]) {
    int p = 0;
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < 3; ++d) p |= (m[s][d] & 15) << (4 * (3 * s + d));
    return p;
}
__host__ __device__ inline int sel3i(int i, int a, int b, int c) {
    return i == 0 ? a : (i == 1 ? b : c);
}

// One fluid sweep of one simulation (fluid_sweep_impl.h).
template <typename T>
struct FluidArgs {
    T* q[5];                 // mass, mom(1..3), ener -- updated in place
    const T* bc[3];          // cell-centred B (MagnetArray.cellMag)
    double* edges;           // edge-tracker storage
    const long long* eoff;   // [NOBJ][2][3] offsets into edges
    const SimDev* sim;       // device time state (dt, active)
    int n0, n1, n2;
    int d;                   // sweep dimension 0..2
    int ndims;
    int modes;               // BC modes [side][dim], packed (mode_at)
    int active_lin;          // Edges.ACTIVE(d+1) read as a linear index
    int limiter;
    MagUse use;              // !MagnetArray.isZero per component
    double gamma, minmass, dx;
    const double* dxl;       // non-uniform DGRID{X}: spacing of cell i along d, or nullptr
    int init_edges;          // first assignment of the flux objects (first fluid sweep)
    int PX, TL, K, xmajor;   // tile: PX pencils x TL threads, K cells per thread
    long long npencil;       // pencil slots (worm_slots with a lower wormhole boundary)
    int worm;                // lower boundary along d is a wormhole: slots hold pencil pairs
    int bc_inf;
    int xcd_remap;
    int force;               // run even when the loop's running flag is off (direct calls)
    const double* fade;      // fade_shift.m coefficients (device)
    int visc_type;           // run.fluid.viscosity: 0 'empty', 1 'neumann_richtmyer'
    double visc_lin, visc_quad;  // linear / quadraticViscousStrength
};

// One entry of a batched launch: argument block `arg` (one simulation and sweep) and the
// tile (pencil group, or cell chunk) within it.  Every launcher below also takes
// tiles == nullptr: then args[0] is one simulation and ntiles its tile count.
struct BatchTile {
    int arg;
    int pad;
    long long tile;
};

// A tile-list entry read so that the compiler knows it is the same in every lane (the
// block works on one tile): the argument-block index and tile go to scalar registers,
// and the argument block's fields are then read as uniform values rather than as
// per-lane loads that each cost a memory round trip.
__device__ __forceinline__ void uniform_tile(const BatchTile* tiles, int t, int& arg,
                                             long long& tile) {
    const BatchTile bt = tiles[t];
    arg = __builtin_amdgcn_readfirstlane(bt.arg);
    const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(bt.tile & 0xffffffffll));
    const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(bt.tile >> 32));
    tile = (long long)(((unsigned long long)hi << 32) | lo);
}

// Block index -> tile index so that consecutive tiles run on the same XCD (blocks are
// dispatched round-robin over the 8 XCDs; neighbouring pencils then share an L2).  A
// bijection of 0..nblocks-1; launches of one block per tile pass their tile count.
__device__ inline long long xcd_tile(long long b, long long nblocks) {
    long long q = nblocks / 8, r = nblocks % 8;
    long long x = b % 8, t = b / 8;
    return x * q + (x < r ? x : r) + t;
}

// Schedule of a persistent launch over a tile list: the list is cut into 8 contiguous
// ranges, one per XCD (blocks are dispatched round-robin over the XCDs, block b runs on
// XCD b % 8), and the blocks of an XCD walk its range together, so tiles that neighbour
// in memory (adjacent pencils of one simulation) are in flight on the same L2.  Returns
// the tile of round r of this block, or -1 when the block is done.  A grid that is not a
// multiple of 8 falls back to the grid-stride order.
__device__ __forceinline__ int xcd_round_tile(int r, int ntiles) {
    const int G = gridDim.x, b = blockIdx.x;
    if (G % 8 != 0) {
        const int t = b + r * G;
        return t < ntiles ? t : -1;
    }
    const int x = b & 7, j = b >> 3, per = G >> 3;
    const int lo = int((long long)ntiles * x / 8);
    const int hi = int((long long)ntiles * (x + 1) / 8);
    const int t = lo + j + r * per;
    return t < hi ? t : -1;
}

// general: the launch may include a fade boundary or the first sweep of a simulation
// (the kernel form that handles them out of line); otherwise the lean form.
template <typename T>
hipError_t launch_fluid_batch(const FluidArgs<T>* args, const BatchTile* tiles, int ntiles,
                              int nblocks, int bs, int kmax, bool mhd, bool general,
                              hipStream_t st);
// The column march for sweeps along y / z with circular or constant boundaries: one
// thread per pencil, fluid_column_pencils() pencils per tile.
template <typename T>
hipError_t launch_fluid_column(const FluidArgs<T>* args, const BatchTile* tiles, int ntiles,
                               int nblocks, bool mhd, hipStream_t st);
int fluid_column_pencils();

// whether a sweep of these arguments needs the general kernel form
template <typename T>
inline bool fluid_needs_general(const FluidArgs<T>& a) {
    return a.init_edges || a.visc_type || a.dxl || mode_at(a.modes, 0, a.d) == IMH_BC_FADE ||
           mode_at(a.modes, 1, a.d) == IMH_BC_FADE;
}

// Compiled fluid sweep variants X(element size, block size, KMAX): FP64 keeps K <= KMAX
// cells per thread in registers with KMAX 1..4 (8 for pencils of 2049..4096 cells, with
// register spills), FP32 with 2..8.  Pencils up to BS * KMAX cells fit one block.  Each
// variant is its own translation unit.
#define IMH_FLUID_VARIANTS(X) \
    X(8, 256, 1) X(8, 256, 2) X(8, 512, 2) X(8, 512, 4) X(8, 512, 8) \
    X(4, 256, 2) X(4, 256, 4) X(4, 512, 4) X(4, 512, 8) X(4, 1024, 4)
#define IMH_FLUID_LAUNCHER_(E, B, K) launch_fluid_v##E##_##B##_##K
#define IMH_FLUID_LAUNCHER(E, B, K) IMH_FLUID_LAUNCHER_(E, B, K)
#define IMH_FLUID_DECLARE(E, B, K)                                                         \
    hipError_t IMH_FLUID_LAUNCHER(E, B, K)(const void* args, const BatchTile* tiles, int ntiles, \
                                           int nblocks, bool mhd, bool general, hipStream_t st);
IMH_FLUID_VARIANTS(IMH_FLUID_DECLARE)
bool fluid_variant_compiled(int esz, int bs, int kmax);

// Wave-spread fluid sweep variants X(element size, block size, fast FP64 arithmetic)
// (fluid_wave_impl.h): four cells per lane, pencils of 4 * TL cells (TL a power of two).
#define IMH_WAVE_VARIANTS(X) \
    X(8, 256, 0) X(8, 512, 0) X(8, 256, 1) X(8, 512, 1) X(4, 256, 0) X(4, 512, 0)
#define IMH_WAVE_LAUNCHER_(E, B, F) launch_fluid_wave_v##E##_##B##_##F
#define IMH_WAVE_LAUNCHER(E, B, F) IMH_WAVE_LAUNCHER_(E, B, F)
#define IMH_WAVE_DECLARE(E, B, F)                                                          \
    hipError_t IMH_WAVE_LAUNCHER(E, B, F)(const void* args, const BatchTile* tiles, int ntiles, \
                                          int nblocks, int d, bool mhd, size_t lds,          \
                                          hipStream_t st);
IMH_WAVE_VARIANTS(IMH_WAVE_DECLARE)

// LDS of one block of the wave-spread sweep (dynamic shared memory), byte offsets: nw
// waves, px pencils per block, dl = lane distance of chunk neighbours, multi = a pencil
// spans waves (wave-edge halo slots needed), esz = element size.
constexpr int WAVE_PXMAX = 32;
struct WaveLds {
    int red, e1, e2, end1, end2, trk, gin, gout, total;
};
__host__ __device__ inline WaveLds wave_lds_layout(int nw, int px, int dl, int multi, int esz) {
    WaveLds L;
    int o = 0;
    auto take = [&](int bytes) {
        const int at = o;
        o += (bytes + 15) & ~15;
        return at;
    };
    L.red = take(2 * nw * WAVE_PXMAX * esz);          // pencil maxima per wave, 2 stages
    L.e1 = take(multi ? nw * 2 * dl * 10 * esz : 0);  // predictor halo: q, w*C of 1 cell
    L.e2 = take(multi ? nw * 2 * dl * 20 * esz : 0);  // corrector halo: of 2 cells
    L.end1 = take(px * 2 * 10 * esz);                 // circular ends, predictor
    L.end2 = take(px * 2 * 20 * esz);                 // circular ends, corrector
    L.trk = take(px * 2 * 15 * 8);                    // edge trackers of the pencil ends
    L.gin = take(px * 2 * 5 * 2 * 4 * 8);             // ghost inputs
    L.gout = take(px * 2 * 5 * 4 * 8);                // ghost values
    L.total = o;
    return L;
}

// Tile of the wave-spread sweep for pencils of N cells along dim d: block size, pencils
// per block and lanes per pencil, with its dynamic LDS bytes; false when the sweep does
// not fit it (N != 4 * TL with TL a power of two, too few lanes per pencil, or more LDS
// than max_lds, the device's per-workgroup limit).
bool wave_tile(int N, int d, int esz, size_t max_lds, int* bs, int* px, int* tl, size_t* lds);
// whether the boundaries of dim d allow the wave-spread sweep (both ends circular or
// neither)
inline bool wave_bc_ok(int modes, int d) {
    if (mode_at(modes, 0, d) == IMH_BC_WORMHOLE || mode_at(modes, 1, d) == IMH_BC_WORMHOLE)
        return false;  // pencil pairs: the general sweep
    return (mode_at(modes, 0, d) == IMH_BC_CIRC) == (mode_at(modes, 1, d) == IMH_BC_CIRC);
}
template <typename T>
hipError_t launch_fluid_wave(const FluidArgs<T>* args, const BatchTile* tiles, int ntiles,
                             int nblocks, int bs, bool fast, int d, bool mhd, size_t lds,
                             hipStream_t st);

// One magnetFlux sub-step (component I advected along X) of one simulation
// (magnet_sweep_impl.h).
template <typename T>
struct MagnetArgs {
    const T* rho;            // mass
    const T* mom;            // mom(X)
    T* bI;                   // mag(I), updated in place
    T* bX;                   // mag(X), updated by the CT step
    T* flux;                 // scratch: mag(I).flux(X) after the limiter
    double* edges;
    const long long* eoff;
    const SimDev* sim;
    int n0, n1, n2;
    int dX, dI;              // 0-based sweep and component dimensions
    int ndims;
    int modes;               // packed BC modes
    int active_linX, active_linI;  // Edges.ACTIVE(D) (linear index) for D = X, I
    int limiter;
    double dxX, dxI;
    const double* dxlX;      // non-uniform DGRID{X} / DGRID{I} lines, or nullptr
    const double* dxlI;
    int init_vg, init_store, init_wmag, init_flux;  // first assignments
    int PX, TL;              // tile for the pencil kernel
    long long npencil;
    int bc_inf;
    int force;
    const double* fade;
};


// updateCellCentered for component k: cmag_k = 0.5*(B_k + shift_k(B_k, +1)).
template <typename T>
struct CellMagArgs {
    const T* b;
    T* cb;
    double* edges;
    const long long* eoff;
    const SimDev* sim;
    int n0, n1, n2;
    int k;                   // component (0-based) = shift dim
    int ndims;
    int modes;
    int active_lin;
    int bc_inf;
    int force;               // run even if the simulation is not active (constructor)
    const double* fade;
};


// Batched magnet kernels: which = 0 advection sub-step, 1 CT step.  kmax from
// magnet_kmax_for(N, TL); smem = sizeof(T) * 2 * PX * (N + 4) of the largest tile.
template <typename T>
hipError_t launch_magnet_batch(int which, const MagnetArgs<T>* args, const BatchTile* tiles,
                               int ntiles, int nblocks, int kmax, size_t smem, hipStream_t st);
template <typename T>
hipError_t launch_cell_centered_batch(const CellMagArgs<T>* args, const BatchTile* tiles,
                                      int ntiles, int nblocks, hipStream_t st);
int magnet_kmax_for(int N, int TL);
constexpr int MAG_PENCIL_MAX = 4096;  // magnet pencils: 256 threads x 16 cells

// The fluid sweep along pencils longer than the fused sweep's tile (fluid_long.hip):
// five launches over the cells, the pencil-wide quantities in a scratch area of
// fluid_long_bytes (U^{n+1/2}, two pencil maxima, the end ghosts).
constexpr int FLUID_LONG_GH = 8;   // ghost values per pencil and field
constexpr int FLUID_LONG_EN = 36;  // row ends per pencil and field: 3 stages x 2 rows x 6
constexpr int FLUID_LONG_VG = 11;  // viscosity ghosts per pencil
inline size_t fluid_long_align(size_t b) { return (b + 255) & ~size_t(255); }
inline size_t fluid_long_bytes(size_t esz, long long ncell, long long npencil) {
    return 6 * fluid_long_align(esz * (size_t)ncell) +
           fluid_long_align(sizeof(unsigned long long) * 2 * (size_t)npencil) +
           fluid_long_align(sizeof(double) * 5 * FLUID_LONG_GH * (size_t)npencil) +
           fluid_long_align(sizeof(double) * 5 * FLUID_LONG_EN * (size_t)npencil) +
           sizeof(double) * FLUID_LONG_VG * (size_t)npencil;
}
template <typename T>
hipError_t launch_fluid_long(const FluidArgs<T>& a, void* scratch, long long ncell, bool mhd,
                             hipStream_t st);

// One launch per batched iteration for the 1-D MHD simulations of an FP32 batch
// (lane1d.hip): one workgroup per simulation runs its fluid sweeps, magnet sub-steps and
// cell-centred updates back to back.  The argument blocks carry a tile of one pencil
// over the whole workgroup.  IMH_LANE1D_BS (a build-time knob) is the workgroup size;
// each thread holds up to 4096 / IMH_LANE1D_BS cells.
#ifndef IMH_LANE1D_BS
#define IMH_LANE1D_BS 512
#endif
constexpr int LANE1D_BS = IMH_LANE1D_BS;
constexpr int LANE1D_NMAX = 4096;
constexpr int LANE1D_KMAX = LANE1D_NMAX / LANE1D_BS;
template <typename T>
struct Lane1dArgs {
    FluidArgs<T> fa;        // the sweep along x
    MagnetArgs<T> ma[3];    // [dI] for X = x (dI = 1, 2)
    CellMagArgs<T> ca[3];   // [k]
};
hipError_t launch_lane1d(const void* args, int nsims, hipStream_t st);

// Compiled magnet variants X(element size, KMAX): each a translation unit with the
// advection and CT kernels for pencils of up to KMAX * TL cells.
#define IMH_MAG_VARIANTS(X) \
    X(8, 1) X(8, 2) X(8, 4) X(8, 8) X(8, 16) X(4, 1) X(4, 2) X(4, 4) X(4, 8) X(4, 16)
#define IMH_MAG_LAUNCHER_(E, K) launch_magnet_v##E##_##K
#define IMH_MAG_LAUNCHER(E, K) IMH_MAG_LAUNCHER_(E, K)
#define IMH_MAG_DECLARE(E, K)                                                               \
    hipError_t IMH_MAG_LAUNCHER(E, K)(int which, const void* args, const BatchTile* tiles,  \
                                      int ntiles, int nblocks, size_t smem, hipStream_t st);
IMH_MAG_VARIANTS(IMH_MAG_DECLARE)

// CFL reduction (TimeManager.update) over one simulation.
template <typename T>
struct CflArgs {
    const T* q[5];
    const T* bc[3];
    SimDev* sim;
    long long ncell;
    long long chunk;         // cells per tile (a multiple of 256)
    MagUse use;
    double gamma;
};

template <typename T>
hipError_t launch_cfl_batch(const CflArgs<T>* args, const BatchTile* tiles, int ntiles,
                            int nblocks, hipStream_t st);
long long cfl_chunk();

// Device-side bookkeeping kernels (cfl.hip), one thread per simulation; `sim` and `p`
// are device arrays of per-simulation pointers.
hipError_t launch_begin_iteration(SimDev* const* sim, const SimParams* const* p, int nsims,
                                  int unconditional, hipStream_t st, int reset = 0);
hipError_t launch_cfl_reset(SimDev* const* sim, int nsims, hipStream_t st);
hipError_t launch_time_finalize(SimDev* const* sim, const SimParams* const* p, int nsims,
                                int step, hipStream_t st);
hipError_t launch_end_iteration(SimDev* const* sim, int nsims, hipStream_t st);

// Conservation totals (Neumaier-compensated FP64, fixed-order tree).
template <typename T>
hipError_t launch_conservation(const T* const* fields, int nfields, long long ncell,
                               double* partials, int nblocks, double* out, hipStream_t st);

}  // namespace imh
