/*
 * imogen_hip.h -- C ABI of the MI355X (gfx950) relaxing-TVD fluid/magnet update.
 *
 * This is the drop-in boundary: the MATLAB functions of the reference keep their
 * signatures and their bodies become calls of these entry points through the MEX
 * gateway (imogen_amd/mex/imogen_mex.cpp), exactly as the reference's own MEX modules
 * are called (gravity/utils/cudaCommon.cu:20-56 marshal GPUmat arrays; outputs are
 * allocated by the library or updated in place).  Plain C types only: pointers, sizes,
 * ints and doubles.
 *
 * Arrays crossing the boundary are column-major [nx ny nz] (x fastest), FP64, the
 * layout of the reference's FluidArray/MagnetArray data (imogen.m:34-42).  The library
 * owns all device memory and the boundary-condition edge trackers; host buffers are
 * borrowed for the duration of a call only.  A context is used by one thread at a time
 * and owns one HIP stream; separate contexts are independent.
 *
 * Every call returns IMH_OK (0) or a negative IMH_E* status; imh_last_error() returns
 * the message of the last failure on the calling thread.
 */
#ifndef IMOGEN_HIP_H
#define IMOGEN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define IMH_OK 0
#define IMH_EINVAL -1     /* bad argument or configuration */
#define IMH_EHIP -2       /* a HIP runtime call failed */
#define IMH_ENOMEM -3     /* device memory exhausted or HBM budget exceeded */
#define IMH_ECFL -4       /* no signal speed above 1e-2 (TimeManager.m:82-85) */
#define IMH_ESTATE -5     /* call out of order (e.g. a sweep before the state upload) */
#define IMH_ENOTSUP -6    /* configuration the reference itself rejects or errors on */

/* ---- boundary-condition modes (enums/ENUM.m:23-32) --------------------------------- */
#define IMH_BC_CIRC 0
#define IMH_BC_CONST 1
#define IMH_BC_TRANS 2
#define IMH_BC_FADE 3
#define IMH_BC_MIRROR 4
#define IMH_BC_FLIP 5
#define IMH_BC_WALL 6
#define IMH_BC_ZERO 7
/* shifting/wormhole_shift.m: the lower ghosts of a pencil are the first cells of the pencil
   whose next-dimension coordinate is mirrored (x->y, y->z, z->x), in reverse order and
   negated for the component along the shift; the upper side shifts as const.  3-D grids
   only (the reference indexes a third dimension); fluid sweeps, the gravity potential and
   the Neumann-Richtmyer viscosity (a magnet or the Caramana-Shashkov-Whalen viscosity
   with it is refused). */
#define IMH_BC_WORMHOLE 8

/* ---- flux limiters (enums/FluxLimiterEnum.m) ---------------------------------------- */
#define IMH_LIM_VANLEER 0
#define IMH_LIM_SUPERBEE 1
#define IMH_LIM_MINMOD 2

/* ---- precision --------------------------------------------------------------------- */
#define IMH_FP64 0        /* plain path: every operation in FP64 */
#define IMH_FP32_BULK 1   /* FP32 storage and flux arithmetic; FP64 pressure floor, dt and
                             compensated conservation sums */

/* ---- field ids ---------------------------------------------------------------------- */
#define IMH_F_MASS 0
#define IMH_F_MOMX 1
#define IMH_F_MOMY 2
#define IMH_F_MOMZ 3
#define IMH_F_ENER 4
#define IMH_F_MAGX 5
#define IMH_F_MAGY 6
#define IMH_F_MAGZ 7
#define IMH_F_CMAGX 8     /* cell-centred B (MagnetArray.cellMag) */
#define IMH_F_CMAGY 9
#define IMH_F_CMAGZ 10
#define IMH_NFIELDS 11

/* Run settings: the fields of the reference's managers that reach this path
 * (ImogenManager.DGRID/GAMMA, FluidManager.MINMASS/limiter/ACTIVE,
 * MagnetManager.ACTIVE, TimeManager.CFL/updateMode/ITERMAX/TIMEMAX,
 * BCManager.modes/infinity). */
typedef struct {
    int grid[3];          /* ini.grid */
    double dgrid[3];      /* run.DGRID (uniform spacing) */
    double gamma;         /* run.GAMMA */
    double minmass;       /* run.fluid.MINMASS */
    double cfl;           /* run.time.CFL */
    int bc_infinity;      /* run.bc.infinity */
    int bc[2][3];         /* run.bc.modes: [side 0 lower / 1 upper][dim] */
    int limiter[3];       /* run.fluid.limiter{dim}; run.magnet.limiter is the same */
    int fluid_active;     /* run.fluid.ACTIVE */
    int magnet_active;    /* run.magnet.ACTIVE */
    int update_mode;      /* run.time.updateMode: 0 per step, 1 per iteration */
    int iter_max;         /* run.time.ITERMAX */
    double time_max;      /* run.time.TIMEMAX */
    int perms_order;      /* 0: perms() row order of MATLAB R2009b-R2010b; 1: reverse lexicographic */
    int precision;        /* IMH_FP64 or IMH_FP32_BULK */
    int device;           /* HIP device ordinal */
} imh_config;

typedef struct imh_ctx imh_ctx;

/* Message of the last failing call on this thread ("" if none). */
const char *imh_last_error(void);

/* Library version string. */
const char *imh_version(void);

/* ---- lifecycle --------------------------------------------------------------------- */
/* Replaces the manager set-up of initialize.m:24-190 for this path. */
int imh_create(const imh_config *cfg, imh_ctx **out);
int imh_destroy(imh_ctx *ctx);

/* ---- state ------------------------------------------------------------------------- */
/* Initial state; builds what the FluidArray/MagnetArray constructors build
 * (imogen.m:34-42, FluidArray.m:44-54, MagnetArray.m:41-46,88-104): the data, the edge
 * stores taken from the IC, MagnetArray.isZero and the initial cell-centred B.
 * n = grid[0]*grid[1]*grid[2] elements per array. */
int imh_upload_state(imh_ctx *ctx, const double *mass, const double *const mom[3],
                     const double *ener, const double *const mag[3], size_t n);
/* Overwrite one field's data (no constructor side effects). */
int imh_upload_field(imh_ctx *ctx, int field, const double *colmajor, size_t n);
/* Copy one field to the host (converted to FP64). */
int imh_download_field(imh_ctx *ctx, int field, double *out, size_t n);

/* ---- hot path ---------------------------------------------------------------------- */
/* relaxingFluid(run,mass,mom,ener,mag,grav,X) -- fluid/relaxingFluid.m:1-76.
 * X in 1..3; uses the context's current dTime (imh_set_dtime or imh_time_update). */
int imh_fluid_sweep(imh_ctx *ctx, int X);
/* magnetFlux(run,mass,mom,mag,X,magneticIndices) -- magnet/magnetFlux.m:1-32. */
int imh_magnet_sweep(imh_ctx *ctx, int X, const int I[2]);
/* run.time.update(mass,mom,ener,mag,i) -- time/TimeManager.m:55-99.  dt stays on the
 * device; nothing is copied back. */
int imh_time_update(imh_ctx *ctx, int step_in_iteration);
/* flux(run,...,order) -- flux/flux.m:1-64 (order +1 / -1). */
int imh_flux(imh_ctx *ctx, int order);
/* One pass of the imogen.m:61-71 loop body: two timesteps with device-side dt and the
 * iteration counter advanced (TimeManager.step). */
int imh_iteration(imh_ctx *ctx);
/* Run iterations while TimeManager.running holds (iteration <= ITERMAX and
 * time < TIMEMAX), at most max_iterations; *done gets the number run.  The per-iteration
 * launch sequence is captured once into a hipGraph and replayed. */
int imh_run(imh_ctx *ctx, int max_iterations, int *done);

/* ---- source terms ------------------------------------------------------------------
 * source(run,mass,mom,ener,mag,grav) -- source.m:1-31, run after every flux call of
 * imh_iteration / imh_run (imogen.m:63-66) and callable on its own as imh_source.
 * Gravity: GravityManager active with the 'empty' solver, whose potential is the fixed
 * one (GravityManager.m:104-116; initialize.m:343-354 ini.gravity.fixedPotential):
 * mom(i) -= dt * mass.thresholdArray .* grav.calculate5PtDerivative(i, DGRID{i}) and the
 * matching energy change (source.m:16-25).  imh_set_gravity copies the potential
 * (column-major, n cells, FP64; NULL turns gravity off) and sets run.fluid.MASS_THRESHOLD
 * (initialize.m:72-77); the potential's edge stores start at zero, as a new GravityArray's
 * do.  A one-cell dimension below ndims (a 1-D grid's y) is refused, as the reference's
 * shift by two cells errors there; two-cell dimensions need circ, const, flip or zero.  The Poisson solvers stay out (a solved potential is uploaded
 * with imh_set_gravity).
 * Radiation: run.fluid.radiation (Radiation.m): type 0 'empty', 1 'optically_thin'
 * with exponent and initialMaximum (initialize.m:403-409); imh_set_radiation computes
 * the strength from the current state as Radiation.initialize does (Radiation.m:61-77),
 * so it is called after the state is in place.  Sources run on single contexts; a batch
 * holding a context with sources runs its source step inside the batched iteration. */
int imh_source(imh_ctx *ctx);
int imh_set_gravity(imh_ctx *ctx, const double *potential, size_t n, double mass_threshold);
int imh_set_radiation(imh_ctx *ctx, int type, double exponent, double initial_maximum);
int imh_get_radiation_strength(imh_ctx *ctx, double *strength);

/* Non-uniform grid spacing: ImogenManager.setGridSpacing (classes/ImogenManager.m:190-254)
 * with a struct dgrid builds DGRID{dim} as one interpolated line of grid(dim) spacings
 * replicated over the other dimensions, and MINDGRID(dim) as its minimum.  This sets that
 * line (n == the context's extent along dim, every value > 0; line == NULL restores the
 * configuration's uniform dgrid[dim-1]).  fluxFactor = dt ./ DGRID then acts per cell in
 * the fluid sweep (relaxingFluid.m:16), the magnet advection and CT (relaxingMagnet.m:15,
 * 56) and the potential derivative (source.m:20); the CFL step uses MINDGRID
 * (TimeManager.m:90).  Sweeps along such a dimension take the general fluid kernel. */
int imh_set_spacing(imh_ctx *ctx, int dim, const double *line, size_t n);

/* run.fluid.viscosity -- fluid/ArtificialViscosity.m (FluidManager.m:42, initialize.m:
 * 395-400): type 0 'empty', 1 'neumann_richtmyer' (ArtificialViscosity.m:73-77), 2
 * 'caramana_shashkov_whalen' (ArtificialViscosity.m:80-148, 155-190) with
 * linearViscousStrength and quadraticViscousStrength; its term is added to the pressure
 * of both wFluidFlux calls of every fluid sweep (wFluidFlux.m:24).  Type 2 at a fade
 * end takes the multi-launch sweep: 1-D and 2-D grids only (IMH_ENOTSUP on 3-D grids,
 * with a wormhole boundary, and in a batch). */
int imh_set_viscosity(imh_ctx *ctx, int type, double linear, double quadratic);

/* ---- time state -------------------------------------------------------------------- */
int imh_set_dtime(imh_ctx *ctx, double dt);
/* Blocks until the stream is idle, then reads the device time state. */
int imh_get_time(imh_ctx *ctx, double *time, double *dtime, int *iteration, int *grid_index);
int imh_set_iteration(imh_ctx *ctx, int iteration);
int imh_synchronize(imh_ctx *ctx);

/* Totals of mass, momentum (3), energy and B (3) in FP64 with Neumaier compensation,
 * reduced in a fixed order (the result does not depend on launch shape). */
int imh_conservation_sums(imh_ctx *ctx, double out[8]);

/* Allocation-free form of imh_conservation_sums.  It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).  fields[0..nfields) are device arrays
 * of ncell elements (double for IMH_FP64, float for IMH_FP32_BULK) and out is device
 * memory for nfields doubles; the sums are queued on `stream` (a hipStream_t, NULL for the
 * default stream) and the call returns without waiting.  Same kernels and reduction order
 * as imh_conservation_sums, so the same bits.  The block partials live in one fixed
 * 128 KiB device array per GPU, held until the process ends: at most
 * IMH_NOALLOC_MAX_FIELDS fields per call (more is refused with IMH_EINVAL), and calls that
 * overlap on one GPU race on that array (order them on one stream). */
#define IMH_NOALLOC_MAX_FIELDS 8
int imh_conservation_sums_noalloc(const void *const *fields, int nfields, int precision,
                                  size_t ncell, double *out, void *stream);

/* Device bytes held by the context. */
int imh_memory_bytes(imh_ctx *ctx, size_t *bytes);
/* Device bytes a context with this configuration would hold (validates cfg first). */
int imh_memory_estimate(const imh_config *cfg, size_t *bytes);

/* ---- initial conditions on the device ---------------------------------------------------
 * No counterpart in the reference (its experiment/ initializer classes build the state
 * on the host): the state of one of these problems is generated in device memory, with
 * the constructor side effects of imh_upload_state (edge stores from the initial arrays,
 * isZero, cell-centred B, a new run's time state), so a simulation starts without a host
 * state or a host-to-device copy.  params (nparams values):
 *   SOD              none (Sod shock tube along x, split at MATLAB round(nx/2))
 *   ORSZAG_TANG      none (periodic [0, 2pi)^2 with the x spacing of the grid)
 *   SEDOV            radius in cells, ambient pressure, blast energy
 *   RIEMANN          left rho, P, vx, vy, vz, By, Bz; right the same; Bx; split cell
 *                    (< 0: nx / 2)
 *   KELVIN_HELMHOLTZ Mach number, density ratio, perturbation amplitude, seed
 * The formulas are those of the benchmark generators (imogen_amd/problems.py); the KH
 * perturbation comes from a counter-based generator of the seed. */
#define IMH_IC_SOD 1
#define IMH_IC_ORSZAG_TANG 2
#define IMH_IC_SEDOV 3
#define IMH_IC_RIEMANN 4
#define IMH_IC_KELVIN_HELMHOLTZ 5
/* The reference initializers themselves (the experiment/ initializer classes), cell for cell:
 *   ORSZAG_TANG_REF  none: OrszagTangVortexInitializer.m:72-90 (sines of 2 pi i / (N - 1)
 *                    over cell indices; the run's spacing is 1 ./ grid)
 *   SEDOV_REF        none: SedovTaylorBlastWaveInitializer.m:77-86 (needs an odd cell
 *                    count in every dimension with more than one cell, as the
 *                    reference's ndgrid does; radius max(5, round(0.06 min(grid))) cells,
 *                    pressure 3 (gamma-1) / (4 pi (radius / nx)^3) inside, 1e-5 outside)
 *   MAGNETIC_SHOCK_TUBE  direction (0 x, 1 y, 2 z), x / y / z field on (1) or off (0):
 *                    MagneticShockTubeInitializer.m:79-117 (field 5 / sqrt(4 pi), momentum
 *                    10 on the left, the x momentum negated on the right, pressure 20 left
 *                    and 1 right, split at MATLAB round(N / 2)) */
#define IMH_IC_ORSZAG_TANG_REF 6
#define IMH_IC_SEDOV_REF 7
#define IMH_IC_MAGNETIC_SHOCK_TUBE 8
int imh_init_problem(imh_ctx *ctx, int kind, const double *params, int nparams);

/* ---- checkpoint / resume --------------------------------------------------------------
 * No counterpart in the reference: its saves (save/resultsHandler.m:26-113) hold the
 * fields and the time only, so a restarted run loses the per-object edge trackers of
 * the transparent and wall boundaries (Edges.m:121-136) and the first-assignment state
 * of the flux objects (InitializedArray.m:43-47).  A checkpoint holds those as well, so a
 * run resumed from it continues bit for bit.  Layout: a fixed header (magic, version,
 * precision, grid, sizes), the device time state, the flux-object flags, the 11 field
 * arrays in the context's precision, the edge stores (FP64; the trackers of the
 * potential, the StorageArrays and the cell-centred B included).  Load into a context created with the same
 * configuration, with the same imh_set_gravity / imh_set_radiation / imh_set_viscosity
 * settings applied first (settings are not part of a checkpoint); its captured
 * iterations are dropped. */
int imh_checkpoint_bytes(imh_ctx *ctx, size_t *bytes);
int imh_checkpoint_save(imh_ctx *ctx, void *buf, size_t bytes);
int imh_checkpoint_load(imh_ctx *ctx, const void *buf, size_t bytes);

/* ---- batched ensemble (many independent simulations, one launch per sweep phase) -- */
typedef struct imh_batch imh_batch;

/* All simulations share precision and device; grids, spacings, gamma, BCs, limiters,
 * CFL and time limits are per simulation. */
int imh_batch_create(const imh_config *cfgs, int nsims, imh_batch **out);
int imh_batch_destroy(imh_batch *b);
int imh_batch_upload_state(imh_batch *b, int sim, const double *mass,
                           const double *const mom[3], const double *ener,
                           const double *const mag[3], size_t n);
int imh_batch_download_field(imh_batch *b, int sim, int field, double *out, size_t n);
/* A batch over existing contexts (same device and precision), which stay owned by the
 * caller.  The serving loop re-forms its batch this way as requests come and go. */
int imh_batch_create_view(imh_ctx *const *ctxs, int n, imh_batch **out);
/* Advance every simulation that is still running by `iterations` batched iterations.  A
 * simulation that has not run an iteration yet first runs one on its own (its flux
 * objects create their edge stores there), so it advances iterations + 1. */
int imh_batch_step(imh_batch *b, int iterations);
int imh_batch_get_time(imh_batch *b, int sim, double *time, double *dtime, int *iteration);
int imh_batch_synchronize(imh_batch *b);
int imh_batch_memory_bytes(imh_batch *b, size_t *bytes);
/* imh_set_gravity / imh_set_spacing / imh_set_radiation / imh_set_viscosity for
 * simulation `sim` of the batch: its sweeps then run in the general kernel form and its
 * source step follows each flux call inside the batched iteration. */
int imh_batch_set_gravity(imh_batch *b, int sim, const double *potential, size_t n,
                          double mass_threshold);
int imh_batch_set_spacing(imh_batch *b, int sim, int dim, const double *line, size_t n);
int imh_batch_set_radiation(imh_batch *b, int sim, int type, double exponent,
                            double initial_maximum);
int imh_batch_set_viscosity(imh_batch *b, int sim, int type, double linear, double quadratic);

/* ---- serving loop (continuous batching of solve requests on one device) ------------ */
/* A request is one simulation run while TimeManager.running holds (to its TIMEMAX or
 * ITERMAX).  Requests queue in FIFO order and are admitted while fewer than max_active
 * run and their device memory (imh_memory_estimate) fits the HBM budget; admitted
 * requests stay resident and advance together with the running ones; finished ones leave
 * the batch at the end of the round and keep their result on the device until
 * imh_serve_release.  No counterpart in the reference (its main loop runs one simulation). */
typedef struct imh_server imh_server;

#define IMH_REQ_QUEUED 0
#define IMH_REQ_RUNNING 1
#define IMH_REQ_DONE 2
#define IMH_REQ_FAILED 3

/* hbm_budget 0: 90% of the device memory free at creation (also the upper limit). */
int imh_serve_create(int device, int precision, size_t hbm_budget, int max_active,
                     imh_server **out);
int imh_serve_destroy(imh_server *s);
/* Copies the initial state (host FP64, n cells per field); cfg.precision and cfg.device
 * are taken from the server. */
int imh_serve_submit(imh_server *s, const imh_config *cfg, const double *mass,
                     const double *const mom[3], const double *ener,
                     const double *const mag[3], size_t n, long long *id);
/* A request whose initial state is generated on the device at admission (kind and params
 * as imh_init_problem takes them): no host state, no host-to-device copy. */
int imh_serve_submit_problem(imh_server *s, const imh_config *cfg, int kind,
                             const double *params, int nparams, long long *id);
/* One scheduling round: admit, advance the running batch by `iterations` iterations,
 * retire the finished requests and admit into the freed slots. */
int imh_serve_step(imh_server *s, int iterations, int *active, int *queued);
/* Status and host wall-clock seconds (since imh_serve_create) of submission, admission
 * and completion (-1 while not reached); time and iteration as of the last round. */
int imh_serve_status(imh_server *s, long long id, int *status, double *time, int *iteration,
                     double t_events[3]);
int imh_serve_download(imh_server *s, long long id, int field, double *out, size_t n);
int imh_serve_release(imh_server *s, long long id);
/* Device bytes held by admitted and finished requests, the budget, and the cell-updates
 * (2 x cells per iteration) of every finished request. */
int imh_serve_stats(imh_server *s, size_t *used_bytes, size_t *budget_bytes,
                    long long *cell_updates);

#ifdef __cplusplus
}
#endif
#endif
