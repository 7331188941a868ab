/*
 * imogen_hip_ext.h -- tuning and measurement entry points of libimogen_hip.so.
 *
 * None of these has a counterpart in the reference: they serve the tile autotuner, the
 * benchmark and the tests (kernel timings with HIP events, tile overrides, graph
 * replay on/off).  The drop-in boundary is include/imogen_hip.h.
 */
#ifndef IMOGEN_HIP_EXT_H
#define IMOGEN_HIP_EXT_H

#include "imogen_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel kinds reported by the profiling calls. */
#define IMH_K_CFL 0       /* CFL reduction (TimeManager.update) */
#define IMH_K_FLUID 1     /* fused relaxingFluid sweep */
#define IMH_K_MAG_ADV 2   /* relaxingMagnet advection sub-step */
#define IMH_K_MAG_CT 3    /* relaxingMagnet constrained-transport step */
#define IMH_K_CELLMAG 4   /* MagnetArray.updateCellCentered */
#define IMH_K_BOOK 5      /* per-iteration time bookkeeping */
#define IMH_K_SOURCE 6    /* source step (gravity, radiation) */
#define IMH_K_LANE1D 7    /* whole flux call of the 1-D MHD simulations of a batch */
#define IMH_NKIND 8

/* Pin the fluid sweep tile along dim X (1..3): block size, registers per thread (KMAX)
 * and pencils per block; px = 0 restores the automatic choice.  Drops captured graphs. */
int imh_set_fluid_tile(imh_ctx *ctx, int X, int bs, int kmax, int px);
int imh_get_fluid_tile(imh_ctx *ctx, int X, int *bs, int *kmax, int *px);
/* The same for simulation `sim` of a batch (its launch plans are rebuilt). */
int imh_batch_set_fluid_tile(imh_batch *b, int sim, int X, int bs, int kmax, int px);
/* 1 (the default): the 1-D MHD simulations of an FP32 batch run their whole flux call in
 * one launch per iteration (one workgroup each); 0: they take the per-sweep launches.
 * Both give the same results bit for bit. */
int imh_batch_set_fused_1d(imh_batch *b, int enable);
/* Fluid sweep kernel along every dimension: 0 the general sweep (LDS exchange), 1 the
 * wave-spread sweep with the arithmetic of the precision (FP64: every division, bit for
 * bit with the general sweep), 2 the wave-spread sweep with fast FP64 arithmetic (one
 * reciprocal of rho and of the freezing speed per cell state, hardware reciprocal with
 * Newton steps; within 1e-12 relative L-inf of the oracle per step), 3 the fast FP64
 * wave-spread sweep where its tiles fill the device (4 per CU) and the general sweep
 * elsewhere (the plain path's choice for large grids).  Sweeps the wave-spread form does
 * not cover (pencil lengths other than 4 * 2^j, a circular end opposite another mode,
 * fade ends, first sweeps, viscosity) use the general sweep.  Drops captured graphs. */
#define IMH_SWEEP_GENERAL 0
#define IMH_SWEEP_WAVE 1
#define IMH_SWEEP_WAVE_FAST 2
#define IMH_SWEEP_AUTO_FAST 3
int imh_set_fluid_kernel(imh_ctx *ctx, int kind);
int imh_get_fluid_kernel(imh_ctx *ctx, int *kind);
/* enable = 0 runs imh_run eagerly (one launch at a time) instead of replaying graphs. */
int imh_set_graphs(imh_ctx *ctx, int enable);

/* Run `iterations` loop bodies eagerly with HIP events around every kernel and return,
 * per kernel kind, the summed time (ms), the summed algorithmic HBM bytes and the number
 * of launches.  The simulation advances exactly as under imh_run. */
int imh_profile_iterations(imh_ctx *ctx, int iterations, double ms[IMH_NKIND],
                           double bytes[IMH_NKIND], int count[IMH_NKIND]);
/* The same for `iterations` batched iterations of a batch. */
int imh_batch_profile_step(imh_batch *b, int iterations, double ms[IMH_NKIND],
                           double bytes[IMH_NKIND], int count[IMH_NKIND]);

/* Time a plain device-to-device copy of `bytes` (16-byte vectors, grid-stride) over
 * `reps` launches; *ms_per_copy gets the mean time.  Reads and writes bytes each. */
int imh_copy_probe(size_t bytes, int reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
