/*
 * imogen_oracle.h -- CPU FP64 restatement of Imogen's relaxing-TVD fluid/magnet update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the HIP
 * implementation in imogen_amd/ and the CPU baseline that bench.py reports beside it.
 * The product never links, loads or calls it: only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it, and only as the checker.
 *
 * What it restates (reference files, read as text; MATLAB cannot run here):
 *   flux/flux.m:1-64                 sweep order (perms rows), fluid/magnet order per sweep
 *   fluid/relaxingFluid.m:16-73      predictor / corrector
 *   fluid/wFluidFlux.m:14-44         freezing speed and relaxation (W) arrays
 *   pressure/pressure.m:16-64        'totsnd' and 'sound' modes, floor at 0
 *   flux/vanleerLimiter.m:9-11, superbeeLimiter.m:9-10, minmodLimiter.m:9-10
 *   magnet/magnetFlux.m:14-31, magnet/relaxingMagnet.m:15-60, magnet/MagnetArray.m:36-38
 *   classes/ImogenArray.m:156-166,292-306  shift dispatch, shift enable per dim
 *   classes/InitializedArray.m:43-47        edge store created at first assignment
 *   shifting/{circ,constant,transparent,fade,mirror,flip,wall,zero}_shift.m, Edges.m:37-142
 *   time/TimeManager.m:55-99, utils/maxFinderND.m:15-19   CFL reduction
 *   source.m:16-31, classes/ImogenArray.m:189-192, fluid/FluidArray.m:29-34,
 *   gravity/GravityManager.m:104-116, fluid/Radiation.m:61-89   source step
 *   fluid/ArtificialViscosity.m:73-77, fluid/wFluidFlux.m:24   Neumann-Richtmyer viscosity
 *   fluid/ArtificialViscosity.m:80-148,155-190   Caramana-Shashkov-Whalen viscosity
 *
 * Every array operation keeps MATLAB's left-to-right evaluation order so that the
 * rounding is the one the .m text prescribes.  Build without FP contraction
 * (-ffp-contract=off): a fused multiply-add would change that rounding.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   pressure       closed form c = sqrt(gamma P / rho) at B = 0; MHD fast-speed closed form;
 *                  floor of negative pressure.
 *   limiters       dL*dR <= 0 -> 0; dL = dR = d gives vanLeer d/2, minmod d/4, superbee d/4;
 *                  0/0 -> 0; brute-force table against the .m formulas.
 *   pchip          scipy.interpolate.PchipInterpolator on the same 5 nodes (interior and
 *                  end intervals), hand-computed vectors; fade coefficients fixture.
 *   BC shifts      index tables on small arrays (circ / const / mirror / flip / zero / wall),
 *                  transparent ghost closed forms, edge-tracker recurrence.
 *   sweeps         uniform state preserved exactly; periodic BCs conserve sum(q) to rounding;
 *                  mirror symmetry; Sod exact solution L1 convergence; golden vectors from
 *                  the independent NumPy restatement in tests/golden/make_golden.py.
 *   source         5-point derivative exact on cubic potentials (interior cells);
 *                  momentum change -dt*rho*dphi/dx and energy change from the old momentum;
 *                  mass threshold; dimensions beyond ndims; radiation strength and sink
 *                  closed forms on uniform states.
 *   viscosity      zero on expanding or uniform flow; on a linear compression the added
 *                  pressure is vm*(c1*c_f + c2*vm/rho) with vm = p(i+1) - p(i) (NR) and the
 *                  published Caramana-Shashkov-Whalen form with limiter 1/4 (CSW); a run
 *                  with zero strengths equals the inviscid run bit for bit.
 *   both           golden vectors of the NumPy restatement with gravity, radiation and
 *                  both viscosities over transparent, wall, mirror and fade ends.
 *   CFL            closed form on a uniform state.
 * Parity unpinned (no MATLAB anywhere): MATLAB's own perms row order (two orders are
 * offered, see orc_config.perms_order), MATLAB pchip's last-bit rounding, and the last bit
 * of MATLAB's .^ in the radiation sink (C pow here).
 */
#ifndef IMOGEN_ORACLE_H
#define IMOGEN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary-condition modes (enums/ENUM.m:23-32). */
enum {
    ORC_BC_CIRC = 0, ORC_BC_CONST = 1, ORC_BC_TRANS = 2, ORC_BC_FADE = 3,
    ORC_BC_MIRROR = 4, ORC_BC_FLIP = 5, ORC_BC_WALL = 6, ORC_BC_ZERO = 7,
    ORC_BC_WORMHOLE = 8
};

/* Flux limiters (enums/FluxLimiterEnum.m). */
enum { ORC_LIM_VANLEER = 0, ORC_LIM_SUPERBEE = 1, ORC_LIM_MINMOD = 2 };

/* Field ids used by orc_get_field / orc_set_field. */
enum {
    ORC_F_MASS = 0, ORC_F_MOMX = 1, ORC_F_MOMY = 2, ORC_F_MOMZ = 3, ORC_F_ENER = 4,
    ORC_F_MAGX = 5, ORC_F_MAGY = 6, ORC_F_MAGZ = 7,
    ORC_F_CMAGX = 8, ORC_F_CMAGY = 9, ORC_F_CMAGZ = 10,
    ORC_NFIELDS = 11
};

typedef struct {
    int grid[3];          /* cells per dimension (ini.grid) */
    double dgrid[3];      /* uniform spacing (ini.dGrid) */
    double gamma;
    double minmass;       /* run.fluid.MINMASS */
    double cfl;           /* run.time.CFL */
    int bc_infinity;      /* run.bc.infinity */
    int bc[2][3];         /* [side][dim]: side 0 = lower, 1 = upper */
    int limiter[3];       /* run.fluid.limiter{dim} (magnet uses the same) */
    int fluid_active;     /* run.fluid.ACTIVE */
    int magnet_active;    /* run.magnet.ACTIVE */
    int update_mode;      /* 0 = TIMEUPDATE_PER_STEP, 1 = TIMEUPDATE_PER_ITERATION */
    int iter_max;         /* run.time.ITERMAX */
    double time_max;      /* run.time.TIMEMAX */
    int perms_order;      /* 0 = perms() of MATLAB R2009b-R2010b, 1 = reverse lexicographic */
} orc_config;

typedef struct orc_sim orc_sim;

/* Lifecycle.  orc_create validates the configuration the way the reference would fail
 * (unknown mode, fade/mirror on a too-short dimension, ...): returns NULL and fills err. */
orc_sim *orc_create(const orc_config *cfg, char *err, int errlen);
/* ImogenManager.setGridSpacing with a non-uniform dimension: DGRID{dim} is a line of
 * grid[dim-1] spacings (replicated over the other dimensions), MINDGRID(dim) its minimum.
 * line == NULL restores the uniform spacing of the configuration. */
int orc_set_spacing(orc_sim *s, int dim, const double *line, int n);
void orc_destroy(orc_sim *s);

/* Initial state: arrays are column-major [nx ny nz] (x fastest), the layout of imogen.m:34-42.
 * mag is face-centred.  Builds every FluidArray/MagnetArray object as the constructors do
 * (edge stores from the IC, MagnetArray.isZero, initial cell-centred B). */
int orc_set_state(orc_sim *s, const double *mass, const double *const mom[3],
                  const double *ener, const double *const mag[3]);
int orc_get_field(const orc_sim *s, int field, double *out);
int orc_set_field(orc_sim *s, int field, const double *in);   /* overwrite data only */

/* Hot-path operators (same argument meaning as the .m functions). X, I in 1..3. */
int orc_time_update(orc_sim *s, int step_in_iteration);       /* TimeManager.update(...,i) */
int orc_flux(orc_sim *s, int order);                          /* flux(run,...,order) */
int orc_relaxing_fluid(orc_sim *s, int X);                    /* relaxingFluid(...,X) */
int orc_magnet_flux(orc_sim *s, int X, const int I[2]);       /* magnetFlux(...,X,I2) */
int orc_iteration(orc_sim *s);                                /* imogen.m:61-71 body */
int orc_running(const orc_sim *s);                            /* TimeManager.running */

/* Source terms (source.m:16-31).  orc_set_gravity: GravityManager active with the
 * 'empty' solver, grav.array = fixedPotential (column-major [nx ny nz]; NULL turns gravity
 * off) and run.fluid.MASS_THRESHOLD.  orc_set_radiation: run.fluid.radiation type (0
 * 'empty', 1 'optically_thin'), exponent, initialMaximum; the strength is computed from
 * the current state as Radiation.initialize does.  orc_source runs after every flux call
 * inside orc_iteration. */
int orc_set_gravity(orc_sim *s, const double *potential, double mass_threshold);
int orc_set_radiation(orc_sim *s, int type, double exponent, double initial_maximum);
double orc_radiation_strength(const orc_sim *s);
/* Artificial viscosity (fluid/ArtificialViscosity.m): type 0 'empty', 1
 * 'neumann_richtmyer', 2 'caramana_shashkov_whalen' (the averaging form the .m file
 * enables, with cswSoundSpeed) with linear / quadratic strengths; the term is added to the
 * pressure of both wFluidFlux calls of every fluid sweep (wFluidFlux.m:24). */
int orc_set_viscosity(orc_sim *s, int type, double linear, double quadratic);
/* The term the viscosity adds to the pressure for the current state along X (unit tests). */
int orc_viscosity_term(orc_sim *s, int X, double *out);
int orc_source(orc_sim *s);

/* Time state. */
double orc_time(const orc_sim *s);
double orc_dtime(const orc_sim *s);
int orc_iteration_count(const orc_sim *s);
void orc_set_dtime(orc_sim *s, double dt);
int orc_last_grid_index(const orc_sim *s);
const char *orc_last_error(const orc_sim *s);

/* Sweep order helpers (flux.m:17-31). dirs[3] gets directVec, mi[3][2] magneticIndices. */
void orc_sweep_order(int iteration, int order, int perms_order, int dirs[3], int mi[3][2]);

/* Unit-level entry points for the tests. */
void orc_pressure(int mode, int n, double gamma, const double *rho, const double *px,
                  const double *py, const double *pz, const double *ener,
                  const double *bx, const double *by, const double *bz, int use_mag,
                  double *result, double *aux);   /* mode 0 = 'totsnd', 1 = 'sound' */
void orc_limiter(int lim, int n, double *flux, const double *dl, const double *dr);
double orc_pchip5(const double x[5], const double y[5], double xi);
void orc_fade_coeffs(int fade_size, double *coeff);
/* Apply one shift of a standalone object: shape grid[3], BC modes bc[2][3], object
 * component and flux type (0 other, 1 fluxl, 2 fluxr, 3 'flux'), is_flux_array flag.
 * edge[side][dim] arrays (face sized) hold the tracker state in/out; may be NULL for
 * untracked sides.  Returns 0 or -1 on an unsupported request. */
int orc_shift_test(const int grid[3], const int bc[2][3], int bc_infinity, int component,
                   int is_flux_array, int flux_type, int D, int n, const double *in,
                   double *out, double *edge_lo, double *edge_hi, int active_linear);

/* OpenMP thread count for the element loops (1 = scalar CPU baseline). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
