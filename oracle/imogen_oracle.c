/*
 * imogen_oracle.c -- CPU FP64 restatement of Imogen's relaxing-TVD update.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle and CPU baseline for imogen_amd/.  Never part of
 * the product path.  The header imogen_oracle.h lists the reference files restated here
 * and what pins each function.
 *
 * Structure: every reference array object (FluidArray, MagnetArray, StorageArray,
 * FluxArray, InitializedArray) is an obj_t with its own data array and its own edge
 * store (shifting/Edges.m).  Every `obj.shift(D,n)` of the .m text is one call of
 * do_shift(), which dispatches on the object's BC mode exactly as ImogenArray.shift
 * (classes/ImogenArray.m:156-166) and, for the transparent mode, nudges the object's
 * stored edge as Edges.getEdge does (shifting/Edges.m:115-140).  Array expressions are
 * evaluated whole-array, left to right, as MATLAB does.
 */
#include "imogen_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EDGE_TOLERANCE 0.005 /* classes/ImogenArray.m:305 Edges(...,0.005) */
#define SNAP_TOLERANCE 1e-4  /* shifting/transparent_shift.m:92 */

/* Element loops go parallel above 16384 cells (OpenMP over cells / pencils).  The
 * argument must be a variable named n: the pragma text refers to it by name. */
#define EW_LOOP(n) _Pragma("omp parallel for schedule(static) if(n > 16384)") \
    for (long c = 0; c < (n); ++c)

enum { K_FLUID, K_MAGNET, K_INIT, K_STORE, K_FLUX };
enum { FT_OTHER = 0, FT_FLUXL = 1, FT_FLUXR = 2, FT_FLUX = 3 };

typedef struct {
    int kind;            /* class of the reference object */
    int component;       /* ImogenArray.component: 0 scalar, 1..3 vector component */
    int flux_type;       /* FluxArray.fluxType for zero_shift.m */
    int has_edges;       /* Edges object created */
    int active[2][3];    /* Edges.ACTIVE */
    double *edge[2][3];  /* Edges.lower/upper.(x|y|z): one value per face cell */
    double *a;           /* pArray; NULL after cleanup() */
} obj_t;

struct orc_sim {
    orc_config cfg;
    int n[3];
    int ndims;
    int squeezed;                  /* grid relabelled by squeeze (singleton x or y) */
    double *dline[3];              /* non-uniform DGRID{i}: one line along i, or NULL */
    double mindgrid[3];            /* MINDGRID(i) = minFinderND(DGRID{i}) */
    long ncell;
    long nface[3];
    obj_t prim[5];                 /* mass, mom(1..3), ener  (FluidArray) */
    obj_t store[5], sfr[5], sfl[5];/* v(i).store, .store.fluxR, .store.fluxL */
    obj_t fr[5], fl[5];            /* v(i).fluxR, v(i).fluxL */
    obj_t freeze[3], freeze_tvd[3];/* run.fluid.freezeSpd(X), freezeSpdTVD(X) */
    obj_t mag[3], cmag[3];         /* mag(k), mag(k).cellMag */
    obj_t vg[3][3], mstore[3][3], msfr[3][3], wmag[3][3], mflux[3][3]; /* [I][X] */
    int mag_is_zero[3];
    double time, dtime;
    int iteration;
    int grid_index;
    double *fade;                  /* fade_shift.m:27 coefficients, bc_infinity long */
    obj_t grav;                    /* GravityArray (InitializedArray) */
    double *grav_potential;        /* run.gravity.fixedPotential, NULL: gravity inactive */
    double mass_threshold;         /* run.fluid.MASS_THRESHOLD */
    int rad_type;                  /* 0 'empty', 1 'optically_thin' */
    int visc_type;                 /* 0 'empty', 1 'neumann_richtmyer' */
    double visc_lin, visc_quad;    /* linear / quadraticViscousStrength */
    double rad_exponent, rad_strength;
    int have_state;
    char err[256];
};

static int g_threads = 1;

void orc_set_threads(int n) {
    g_threads = n < 1 ? 1 : n;
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
}
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------------------ */
/* MATLAB scalar semantics                                                               */
/* ------------------------------------------------------------------------------------ */

/* MATLAB sign(): NaN for NaN, 0 for +-0. */
/* DGRID{d} at cell c (the cell's coordinate along d indexes the line; ImogenManager.m:
 * 226-254 replicates one interpolated line over the other dimensions) */
static inline double dx_cell(const orc_sim *s, int d, long c) {
    if (!s->dline[d]) return s->cfg.dgrid[d];
    long i;
    if (d == 0) i = c % s->n[0];
    else if (d == 1) i = (c / s->n[0]) % s->n[1];
    else i = c / ((long)s->n[0] * s->n[1]);
    return s->dline[d][i];
}

static inline double msign(double x) {
    if (x > 0) return 1.0;
    if (x < 0) return -1.0;
    if (x == 0) return 0.0;
    return x; /* NaN */
}

/* ------------------------------------------------------------------------------------ */
/* pchip (MATLAB R2010 pchip.m / pchipslopes / pchipend / pwch / ppval), 5 nodes.        */
/* Restated from the published algorithm: Fritsch-Butland weighted harmonic mean for the */
/* interior slopes, Moler's shape-preserving three-point end rule, Hermite cubic in      */
/* Horner form (Moler, Numerical Computing with MATLAB, sec. 3.4).                       */
/* ------------------------------------------------------------------------------------ */

static double pchip_end(double h1, double h2, double del1, double del2) {
    double d = ((2.0 * h1 + h2) * del1 - h1 * del2) / (h1 + h2);
    if (msign(d) != msign(del1)) {
        d = 0.0;
    } else if ((msign(del1) != msign(del2)) && (fabs(d) > fabs(3.0 * del1))) {
        d = 3.0 * del1;
    }
    return d;
}

double orc_pchip5(const double x[5], const double y[5], double xi) {
    double h[4], del[4], d[5];
    for (int j = 0; j < 4; ++j) {
        h[j] = x[j + 1] - x[j];
        del[j] = (y[j + 1] - y[j]) / h[j];
    }
    for (int k = 1; k < 4; ++k) {
        d[k] = 0.0;
        if (msign(del[k - 1]) * msign(del[k]) > 0) {
            double w1 = 2.0 * h[k] + h[k - 1];
            double w2 = h[k] + 2.0 * h[k - 1];
            d[k] = (w1 + w2) / (w1 / del[k - 1] + w2 / del[k]);
        }
    }
    d[0] = pchip_end(h[0], h[1], del[0], del[1]);
    d[4] = pchip_end(h[3], h[2], del[3], del[2]);
    /* ppval interval: breaks x(1..5), last interval extended on both sides. */
    int j = 0;
    for (int t = 1; t < 4; ++t)
        if (xi >= x[t]) j = t;
    double dzzdx = (del[j] - d[j]) / h[j];
    double dzdxdx = (d[j + 1] - del[j]) / h[j];
    double c1 = (dzdxdx - dzzdx) / h[j];
    double c2 = 2.0 * dzzdx - dzdxdx;
    double c3 = d[j];
    double c4 = y[j];
    double s = xi - x[j];
    double v = c1;
    v = s * v + c2;
    v = s * v + c3;
    v = s * v + c4;
    return v;
}

/* fade_shift.m:27: coeff = pchip([0 ceil(f/4) round(f/2) f-1 f],[1 1 .01 0 0],1:f) */
void orc_fade_coeffs(int f, double *coeff) {
    double x[5] = {0.0, ceil(f / 4.0), floor(f / 2.0 + 0.5), (double)(f - 1), (double)f};
    double y[5] = {1.0, 1.0, 0.01, 0.0, 0.0};
    for (int i = 1; i <= f; ++i) coeff[i - 1] = orc_pchip5(x, y, (double)i);
}

/* ------------------------------------------------------------------------------------ */
/* Geometry helpers                                                                      */
/* ------------------------------------------------------------------------------------ */

/* Pencil f along dim d: base offset and stride.  Face index conventions:
 *   d=0: f = j + ny*k,  d=1: f = i + nx*k,  d=2: f = i + nx*j.                          */
static inline void pencil(const int n[3], int d, long f, long *base, long *stride) {
    long nx = n[0], ny = n[1];
    if (d == 0) {
        long j = f % ny, k = f / ny;
        *base = nx * (j + ny * k);
        *stride = 1;
    } else if (d == 1) {
        long i = f % nx, k = f / nx;
        *base = i + nx * ny * k;
        *stride = nx;
    } else {
        long i = f % nx, j = f / nx;
        *base = i + nx * j;
        *stride = nx * ny;
    }
}

static double *alloc_arr(long n) {
    double *p = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

static double *dup_arr(const double *src, long n) {
    double *p = alloc_arr(n);
    memcpy(p, src, sizeof(double) * (size_t)n);
    return p;
}

/* ------------------------------------------------------------------------------------ */
/* Objects, edges (shifting/Edges.m:37-142, InitializedArray.m:43-47)                    */
/* ------------------------------------------------------------------------------------ */

static void obj_init(obj_t *o, int kind, int component, int flux_type) {
    memset(o, 0, sizeof(*o));
    o->kind = kind;
    o->component = component;
    o->flux_type = flux_type;
}

static void obj_free(obj_t *o) {
    free(o->a);
    o->a = NULL;
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < 3; ++d) {
            free(o->edge[s][d]);
            o->edge[s][d] = NULL;
        }
}

/* Edges(bcModes, array, 0.005): store the face slice for trans and wall sides; mark
 * ACTIVE for trans, fade and wall on dims <= ndims(array) (Edges.m:53-108). */
static void edges_create(orc_sim *s, obj_t *o) {
    memset(o->active, 0, sizeof(o->active));
    for (int side = 0; side < 2; ++side)
        for (int d = 0; d < s->ndims; ++d) {
            int mode = s->cfg.bc[side][d];
            if (mode != ORC_BC_TRANS && mode != ORC_BC_FADE && mode != ORC_BC_WALL) continue;
            o->active[side][d] = 1;
            if (mode == ORC_BC_FADE) continue; /* fade's stored slab is never read on the path */
            free(o->edge[side][d]);
            o->edge[side][d] = alloc_arr(s->nface[d]);
            long N = s->n[d];
            for (long f = 0; f < s->nface[d]; ++f) {
                long base, stride;
                pencil(s->n, d, f, &base, &stride);
                o->edge[side][d][f] = o->a[base + (side ? (N - 1) : 0) * stride];
            }
        }
    o->has_edges = 1;
}

/* obj.array = value (ImogenArray.m:82-89): fades/statics are empty on this path; an
 * InitializedArray builds its edges from its first array. Takes ownership of value. */
static void obj_assign(orc_sim *s, obj_t *o, double *value) {
    free(o->a);
    o->a = value;
    if ((o->kind == K_INIT || o->kind == K_STORE || o->kind == K_FLUX) && !o->has_edges)
        edges_create(s, o);
}

static void obj_cleanup(obj_t *o) {
    free(o->a);
    o->a = NULL;
}

/* Edges.getEdge for the transparent mode: nudge the stored edge toward the current edge
 * cell by at most TOLERANCE and return it (Edges.m:121-136). */
static inline double edge_nudge(double *stored, double newedge) {
    double old = *stored;
    double diff = newedge - old;
    double delta = fmin(fabs(diff), EDGE_TOLERANCE); /* MATLAB min ignores NaN */
    double sgn = (diff > 0) ? 1.0 : -1.0;            /* signTest - ~signTest */
    double r = old + sgn * delta;
    *stored = r;
    return r;
}

/* ------------------------------------------------------------------------------------ */
/* Shift (ImogenArray.shift + shifting/X_shift.m)                                        */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    const int *n;
    int ndims;
    const int (*bc)[3];
    int bc_infinity;
    const double *fade;
} shift_ctx;

/* One pencil of length N, values a[0..N-1] with stride; result written to out with the
 * same stride.  m = |n| cells, upper = (n > 0).  Implements the index maps of the
 * *_shift.m files; the transparent and wall modes read/write the tracker value *edge. */
static void shift_pencil(const shift_ctx *c, const obj_t *o, int d, int upper, int m,
                         const double *a, double *out, long N, long stride, double *edge,
                         int active_linear) {
#define A(i) a[(long)(i) * stride]
#define R(i) out[(long)(i) * stride]
    int mode = c->bc[upper][d];
    switch (mode) {
    case ORC_BC_CIRC: /* circ_shift.m:21-29 */
        for (long i = 0; i < N; ++i) {
            long src = upper ? (i + m) % N : ((i - m) % N + N) % N;
            R(i) = A(src);
        }
        break;
    case ORC_BC_CONST:                 /* constant_shift.m:21-29 */
    case ORC_BC_FLIP:                  /* flip_shift.m:24-40 */
    case ORC_BC_WALL:                  /* wall_shift.m:12-33 */
    case ORC_BC_FADE:                  /* fade_shift.m:21 then blend */
    case ORC_BC_ZERO:                  /* zero_shift.m */
    case ORC_BC_WORMHOLE:              /* wormhole_shift.m:61: constant_shift first */
        for (long i = 0; i < N; ++i) {
            long src = upper ? (i + m < N ? i + m : N - 1) : (i - m >= 0 ? i - m : 0);
            R(i) = A(src);
        }
        if (mode == ORC_BC_FLIP && o->component == d + 1) {
            for (int t = 0; t < m; ++t) {
                long i = upper ? N - m + t : t;
                R(i) = -R(i);
            }
        } else if (mode == ORC_BC_WALL) {
            for (int t = 0; t < m; ++t) {
                long i = upper ? N - m + t : t;
                R(i) = *edge;
            }
        } else if (mode == ORC_BC_ZERO && o->kind == K_FLUX) {
            /* lower ghosts stay constant for a FLUXL object, upper ones for FLUXR */
            int keep = upper ? (o->flux_type == FT_FLUXR) : (o->flux_type == FT_FLUXL);
            if (!keep)
                for (int t = 0; t < m; ++t) {
                    long i = upper ? N - m + t : t;
                    R(i) = 0.0;
                }
        } else if (mode == ORC_BC_FADE) {
            int f = c->bc_infinity;
            double v1 = R(0);
            for (int i = 0; i < f; ++i) R(i) = (1.0 - c->fade[i]) * R(i) + c->fade[i] * v1;
            double vN = R(N - 1);
            for (int t = 0; t < f; ++t) {
                double cf = c->fade[f - 1 - t]; /* flipdim(coeff) */
                long i = N - f + t;
                R(i) = (1.0 - cf) * R(i) + cf * vN;
            }
        }
        break;
    case ORC_BC_MIRROR: /* mirror_shift.m:23-39 */
        for (long i = 0; i < N; ++i) {
            long src;
            if (upper)
                src = (i < N - m) ? i + m : (N - 2) - (i - (N - m));
            else
                src = (i >= m) ? i - m : m - i;
            R(i) = A(src);
        }
        if (o->component == d + 1)
            for (int t = 0; t < m; ++t) {
                long i = upper ? N - m + t : t;
                R(i) = -R(i);
            }
        break;
    case ORC_BC_TRANS: { /* transparent_shift.m:23-99 */
        double inv[5], x[5];
        for (long i = 0; i < N; ++i) {
            long src = upper ? (i + m < N ? i + m : N - 1) : (i - m >= 0 ? i - m : 0);
            R(i) = A(src);
        }
        for (int k = 0; k < 3; ++k) {
            /* lower: min((3:-1:1)+n, NDim); upper: max(NDim-(2:-1:0)-n, 1)  (1-based) */
            long idx1 = upper ? (N - (2 - k) - m) : (3 - k + m);
            if (upper) idx1 = idx1 < 1 ? 1 : idx1;
            else idx1 = idx1 > N ? N : idx1;
            inv[k] = R(idx1 - 1);
        }
        double e = 0.0;
        if (active_linear) e = edge_nudge(edge, upper ? A(N - 1) : A(0));
        inv[3] = e;
        inv[4] = e;
        x[0] = 1.0; x[1] = 2.0; x[2] = 3.0;
        x[3] = (double)c->bc_infinity;
        x[4] = (double)c->bc_infinity + 1.0;
        for (int t = 0; t < m; ++t) {
            long i;
            double xi;
            if (upper) {
                i = N - m + t;              /* index (NDim-n+1):NDim */
                xi = 3.0 + (double)(t + 1); /* 3+(1:n) */
            } else {
                i = t;                      /* index 1:n */
                xi = 3.0 + (double)(m - t); /* 3+(n:-1:1) */
            }
            double outv = orc_pchip5(x, inv, xi);
            double orig = A(i);
            double test = (fabs(orig - outv) < SNAP_TOLERANCE) ? 1.0 : 0.0;
            R(i) = (1.0 - test) * outv + test * orig; /* not(test).*outVals + test.*array */
        }
        break;
    }
    default:
        break;
    }
#undef A
#undef R
}

/* wormhole_shift.m:21-57 (n < 0, after the constant shift): with P = previousDir and
 * Q = nextDir of D ([3 1 2] and [2 3 1]),
 *   result(amt:-1:1, all of P, N_Q:-1:1) = isVecNorm * array(1:amt, all of P, 1:N_Q)
 * along D, i.e. cell t < m of the pencil at Q-coordinate q takes cell m-1-t of the
 * unshifted pencil at N_Q-1-q (same P-coordinate), negated when the object's component
 * is D.  Written with explicit coordinates (x fastest in memory). */
static void wormhole_lower(const orc_sim *s, const obj_t *o, int d, int m, double *out) {
    const int q = (d + 1) % 3; /* nextDir(D) - 1 */
    const double sgn = (o->component == d + 1) ? -1.0 : 1.0;
    const long sx = 1, sy = s->n[0], sz = (long)s->n[0] * s->n[1];
    const long st[3] = {sx, sy, sz};
    for (long k = 0; k < s->n[2]; ++k)
        for (long j = 0; j < s->n[1]; ++j)
            for (long i = 0; i < s->n[0]; ++i) {
                long c[3] = {i, j, k};
                if (c[d] >= m) continue;
                long src[3] = {i, j, k};
                src[d] = m - 1 - c[d];
                src[q] = s->n[q] - 1 - c[q];
                out[c[0] * st[0] + c[1] * st[1] + c[2] * st[2]] =
                    sgn * o->a[src[0] * st[0] + src[1] * st[1] + src[2] * st[2]];
            }
}

/* result = obj.shift(D, n)  (D 1-based).  Writes a full-size array into out. */
static void do_shift(orc_sim *s, obj_t *o, int D, int n, double *out) {
    int d = D - 1;
    int upper = n > 0;
    int m = n > 0 ? n : -n;
    if (D > s->ndims) { /* pShiftEnabled false: returns the array unchanged */
        memcpy(out, o->a, sizeof(double) * (size_t)s->ncell);
        return;
    }
    shift_ctx c = {s->n, s->ndims, (const int(*)[3])s->cfg.bc, s->cfg.bc_infinity, s->fade};
    /* Edges.ACTIVE(DIRECT): linear index into the 2x3 logical array (column-major). */
    int active_linear = o->has_edges ? o->active[(D - 1) % 2][(D - 1) / 2] : 0;
    long N = s->n[d];
    long nf = s->nface[d];
    double *edge = o->edge[upper][d];
#pragma omp parallel for schedule(static) if (nf > 64)
    for (long f = 0; f < nf; ++f) {
        long base, stride;
        pencil(s->n, d, f, &base, &stride);
        shift_pencil(&c, o, d, upper, m, o->a + base, out + base, N, stride,
                     edge ? &edge[f] : NULL, active_linear);
    }
    if (!upper && s->cfg.bc[0][d] == ORC_BC_WORMHOLE) wormhole_lower(s, o, d, m, out);
}

/* ------------------------------------------------------------------------------------ */
/* pressure.m                                                                            */
/* ------------------------------------------------------------------------------------ */

/* mode 0 = 'totsnd' (result = total pressure, aux = fast speed), 1 = 'sound', 2 = 'gas'. */
static void pressure_arr(int mode, long n, double g, const double *rho, const double *px,
                         const double *py, const double *pz, const double *ener,
                         const double *bx, const double *by, const double *bz,
                         const int use[3], double *result, double *aux) {
    EW_LOOP(n) {
        double v2 = (px[c] * px[c] + py[c] * py[c] + pz[c] * pz[c]) / (rho[c] * rho[c]);
        double b2 = 0.0;
        int any = 0;
        if (use[0]) { b2 = any ? b2 + bx[c] * bx[c] : 0.0 + bx[c] * bx[c]; any = 1; }
        if (use[1]) { b2 = any ? b2 + by[c] * by[c] : 0.0 + by[c] * by[c]; any = 1; }
        if (use[2]) { b2 = any ? b2 + bz[c] * bz[c] : 0.0 + bz[c] * bz[c]; any = 1; }
        double r = (g - 1.0) * (ener[c] - 0.5 * rho[c] * v2);
        if (mode == 0) {
            aux[c] = sqrt(fabs((g * (r - (g - 1.0) * 0.5 * b2) + 2.0 * b2) / rho[c]));
            r = r + 0.5 * (2.0 - g) * b2;
        } else if (mode == 2) { /* 'gas' */
            r = r - (g - 1.0) * 0.5 * b2;
        } else {
            r = r - (g - 1.0) * 0.5 * b2;
            r = sqrt(fabs((g * r + 2.0 * b2) / rho[c]));
        }
        if (r < 0) r = 0.0; /* result(result < 0) = 0 */
        result[c] = r;
    }
}

void orc_pressure(int mode, int n, double gamma, const double *rho, const double *px,
                  const double *py, const double *pz, const double *ener, const double *bx,
                  const double *by, const double *bz, int use_mag, double *result,
                  double *aux) {
    int use[3] = {use_mag & 1, (use_mag >> 1) & 1, (use_mag >> 2) & 1};
    pressure_arr(mode, n, gamma, rho, px, py, pz, ener, bx, by, bz, use, result, aux);
}

/* ------------------------------------------------------------------------------------ */
/* Limiters (flux/XLimiter.m): flux.array = flux.array + delta                           */
/* ------------------------------------------------------------------------------------ */

static inline double lim_minmod(double dl, double dr) {
    return 0.125 * (msign(dl) + msign(dr)) * fmin(fabs(dl), fabs(dr));
}

static inline double lim_delta(int lim, double dl, double dr) {
    if (lim == ORC_LIM_VANLEER) {
        double t = fmax(dl * dr, 0.0) / (dl + dr);
        return isnan(t) ? 0.0 : t;
    }
    if (lim == ORC_LIM_SUPERBEE) {
        double cf = (fabs(dl) < fabs(dr)) ? 1.0 : 0.0;
        return lim_minmod((1.0 + cf) * dl, (2.0 - cf) * dr);
    }
    return lim_minmod(dl, dr);
}

void orc_limiter(int lim, int n, double *flux, const double *dl, const double *dr) {
    for (int c = 0; c < n; ++c) flux[c] = flux[c] + lim_delta(lim, dl[c], dr[c]);
}

static void apply_limiter(orc_sim *s, int lim, obj_t *flux, const double *dl,
                          const double *dr) {
    long n = s->ncell;
    double *res = alloc_arr(n);
    const double *f = flux->a;
    EW_LOOP(n) res[c] = f[c] + lim_delta(lim, dl[c], dr[c]);
    obj_assign(s, flux, res);
}

/* ------------------------------------------------------------------------------------ */
/* wFluidFlux.m                                                                           */
/* ------------------------------------------------------------------------------------ */

/* ArtificialViscosity.cswSoundSpeed (ArtificialViscosity.m:155-190): the smaller of the
 * sound speeds of the two neighbours, every operand a shift of its own object, in the
 * order the .m text evaluates them. */
static void csw_sound_speed(orc_sim *s, obj_t *const o[5], int X, double *result) {
    long n = s->ncell;
    double G = s->cfg.gamma;
    double *t1 = alloc_arr(n), *t2 = alloc_arr(n), *t3 = alloc_arr(n), *tm = alloc_arr(n);
    double *b1 = alloc_arr(n), *b2 = alloc_arr(n), *b3 = alloc_arr(n), *te = alloc_arr(n);
    double *vel2 = alloc_arr(n), *mag2 = alloc_arr(n), *sp = alloc_arr(n);
    for (int side = 1; side >= 0; --side) { /* C (shift +1) first, then B (shift -1) */
        int sh = side ? 1 : -1;
        do_shift(s, o[1], X, sh, t1);
        do_shift(s, o[2], X, sh, t2);
        do_shift(s, o[3], X, sh, t3);
        do_shift(s, o[0], X, sh, tm);
        EW_LOOP(n) vel2[c] = t1[c] * t1[c] + t2[c] * t2[c] + t3[c] * t3[c] / (tm[c] * tm[c]);
        do_shift(s, &s->cmag[0], X, sh, b1);
        do_shift(s, &s->cmag[1], X, sh, b2);
        do_shift(s, &s->cmag[2], X, sh, b3);
        EW_LOOP(n) mag2[c] = b1[c] * b1[c] + b2[c] * b2[c] + b3[c] * b3[c];
        do_shift(s, o[4], X, sh, te);
        do_shift(s, o[0], X, sh, tm);
        EW_LOOP(n) sp[c] = (G - 1.0) * (te[c] - 0.5 * tm[c] * vel2[c]);
        EW_LOOP(n) sp[c] = sp[c] - (G - 1.0) * 0.5 * mag2[c];
        do_shift(s, o[0], X, sh, tm);
        if (side) {
            EW_LOOP(n) result[c] = sqrt(fabs((G * sp[c] + 2.0 * mag2[c]) / tm[c]));
        } else {
            EW_LOOP(n) {
                double sb = sqrt(fabs((G * sp[c] + 2.0 * mag2[c]) / tm[c]));
                double r = fmin(sb, result[c]); /* min(sndSpdB, sndSpdC), NaN skipped */
                result[c] = fmax(r, 0.0);
            }
        }
    }
    free(t1); free(t2); free(t3); free(tm); free(b1); free(b2); free(b3); free(te);
    free(vel2); free(mag2); free(sp);
}

/* ArtificialViscosity.caramanaShashkovWhalenViscosity (ArtificialViscosity.m:80-148), the
 * averaging form the .m file enables; the result is added to press. */
static void csw_viscosity(orc_sim *s, obj_t *const o[5], int X, double *press) {
    long n = s->ncell;
    double G = s->cfg.gamma, c1 = s->visc_lin, c2 = s->visc_quad;
    double *mB = alloc_arr(n), *mC = alloc_arr(n), *rho = alloc_arr(n), *snd = alloc_arr(n);
    double *ta = alloc_arr(n), *tb = alloc_arr(n), *v[4];
    for (int k = 0; k < 4; ++k) v[k] = alloc_arr(n);
    do_shift(s, o[0], X, -1, mB);
    do_shift(s, o[0], X, 1, mC);
    EW_LOOP(n) rho[c] = 2.0 * mC[c] * mB[c] / (mC[c] + mB[c]);
    csw_sound_speed(s, o, X, snd);
    static const int sh[4] = {-2, -1, 1, 2};
    for (int k = 0; k < 4; ++k) { /* v0..v3 = mom(X).shift ./ mass.shift */
        do_shift(s, o[X], X, sh[k], ta);
        do_shift(s, o[0], X, sh[k], tb);
        double *vk = v[k];
        EW_LOOP(n) vk[c] = ta[c] / tb[c];
    }
    const double q = (G + 1.0) / 4.0;
    EW_LOOP(n) {
        double dv = v[2][c] - v[1][c];
        double rL = (v[3][c] - v[2][c]) / (2.0 * dv);
        double rR = (v[1][c] - v[0][c]) / (2.0 * dv);
        double lim = fmax(0.0, fmin(fmin(0.5 * (rL + rR), fmin(2.0 * rL, 2.0 * rR)), 1.0));
        double A = c2 * (G + 1.0) * fabs(dv) / 4.0;
        double B = sqrt(c2 * c2 * (q * q) * (dv * dv) + c1 * c1 * (snd[c] * snd[c]));
        double r = (dv < 0 ? 1.0 : 0.0) * rho[c] * (A + B) * (1.0 - lim) * fabs(dv);
        press[c] = press[c] + r;
    }
    free(mB); free(mC); free(rho); free(snd); free(ta); free(tb);
    for (int k = 0; k < 4; ++k) free(v[k]);
}

static void wfluidflux(orc_sim *s, double *const q[5], obj_t *const obj[5], obj_t *freeze,
                       int X, double *w[5]) {
    obj_t *momX = obj[X];
    long n = s->ncell;
    int d = X - 1;
    double g = s->cfg.gamma;
    const double *bx = s->cmag[0].a, *by = s->cmag[1].a, *bz = s->cmag[2].a;
    const double *bc[3] = {bx, by, bz};
    const double *bX = bc[d];
    int use[3] = {!s->mag_is_zero[0], !s->mag_is_zero[1], !s->mag_is_zero[2]};
    double *press = alloc_arr(n), *snd = alloc_arr(n), *vel = alloc_arr(n);
    pressure_arr(0, n, g, q[0], q[1], q[2], q[3], q[4], bx, by, bz, use, press, snd);
    const double *qx = q[X];
    const double *rho = q[0];
    EW_LOOP(n) vel[c] = qx[c] / rho[c];
    /* spd = max(abs(velocity)+soundSpd,[],X); freezeSpd.array = repmat(spd) */
    double *C = alloc_arr(n);
    long nf = s->nface[d], N = s->n[d];
#pragma omp parallel for schedule(static) if (nf > 64)
    for (long f = 0; f < nf; ++f) {
        long base, stride;
        pencil(s->n, d, f, &base, &stride);
        double m = fabs(vel[base]) + snd[base];
        for (long i = 1; i < N; ++i) {
            double v = fabs(vel[base + i * stride]) + snd[base + i * stride];
            m = fmax(m, v);
        }
        for (long i = 0; i < N; ++i) C[base + i * stride] = m;
    }
    obj_assign(s, freeze, C);
    /* press = press + viscosity.solve(run, mass, mom, ener, mag, soundSpd, X) */
    if (s->visc_type == 1) {
        /* ArtificialViscosity.neumannRichtmyerViscosity (ArtificialViscosity.m:73-77):
         * mom(X) is the FluidArray in the predictor and its StorageArray in the corrector */
        double *sh = alloc_arr(n);
        do_shift(s, momX, X, 1, sh);
        const double lin = s->visc_lin, quad = s->visc_quad;
        EW_LOOP(n) {
            double vm = fmin(sh[c] - qx[c], 0.0); /* MATLAB min(.., 0) skips NaN */
            press[c] = press[c] + vm * (lin * snd[c] + quad * vm / rho[c]);
        }
        free(sh);
    } else if (s->visc_type == 2) {
        csw_viscosity(s, obj, X, press);
    } else {
        EW_LOOP(n) press[c] = press[c] + 0.0; /* 'empty' viscosity returns 0 */
    }
    const double *Cf = freeze->a;
    const double *E = q[4];
    EW_LOOP(n) {
        w[0][c] = qx[c] / Cf[c];
        double bdotp = bx[c] * q[1][c] + by[c] * q[2][c] + bz[c] * q[3][c];
        double we = vel[c] * (E[c] + press[c]) - bX[c] * bdotp / rho[c];
        w[4][c] = we / Cf[c];
        for (int i = 1; i <= 3; ++i) {
            double dv = (i == X) ? 1.0 : 0.0;
            w[i][c] = (vel[c] * q[i][c] + press[c] * dv - bX[c] * bc[i - 1][c]) / Cf[c];
        }
    }
    free(press);
    free(snd);
    free(vel);
}

/* ------------------------------------------------------------------------------------ */
/* relaxingFluid.m                                                                        */
/* ------------------------------------------------------------------------------------ */

int orc_relaxing_fluid(orc_sim *s, int X) {
    if (!s->have_state) return -1;
    long n = s->ncell;
    int d = X - 1;
    const double dtime = s->dtime; /* fluxFactor = dTime ./ DGRID{X}, per cell */
    double *w[5];
    for (int i = 0; i < 5; ++i) w[i] = alloc_arr(n);
    double *q[5];
    for (int i = 0; i < 5; ++i) q[i] = s->prim[i].a;

    /* Half-timestep predictor */
    obj_t *pobj[5] = {&s->prim[0], &s->prim[1], &s->prim[2], &s->prim[3], &s->prim[4]};
    wfluidflux(s, q, pobj, &s->freeze[d], X, w);
    const double *C = s->freeze[d].a;
    double *s1 = alloc_arr(n), *s2 = alloc_arr(n);
    for (int i = 0; i < 5; ++i) {
        const double *v = q[i], *wi = w[i];
        double *fr = alloc_arr(n), *fl = alloc_arr(n);
        EW_LOOP(n) {
            fr[c] = 0.5 * (v[c] + wi[c]);
            fl[c] = 0.5 * (v[c] - wi[c]);
        }
        obj_assign(s, &s->sfr[i], fr);
        obj_assign(s, &s->sfl[i], fl);
        do_shift(s, &s->sfr[i], X, -1, s1);
        do_shift(s, &s->sfl[i], X, 1, s2);
        double *st = alloc_arr(n);
        EW_LOOP(n) st[c] = v[c] - (0.5 * (dtime / dx_cell(s, d, c))) * C[c] *
                                     (fr[c] - s1[c] + fl[c] - s2[c]);
        obj_assign(s, &s->store[i], st);
        obj_cleanup(&s->sfr[i]);
        obj_cleanup(&s->sfl[i]);
    }
    {
        double *m = s->store[0].a, mm = s->cfg.minmass;
        double *r = alloc_arr(n);
        EW_LOOP(n) r[c] = fmax(m[c], mm);
        obj_assign(s, &s->store[0], r);
    }
    obj_cleanup(&s->freeze[d]);

    /* Full-timestep corrector */
    double *qs[5];
    for (int i = 0; i < 5; ++i) qs[i] = s->store[i].a;
    obj_t *sobj[5] = {&s->store[0], &s->store[1], &s->store[2], &s->store[3], &s->store[4]};
    wfluidflux(s, qs, sobj, &s->freeze_tvd[d], X, w);
    const double *C2 = s->freeze_tvd[d].a;
    double *s3 = alloc_arr(n), *dl = alloc_arr(n), *dr = alloc_arr(n);
    int lim = s->cfg.limiter[d];
    for (int i = 0; i < 5; ++i) {
        const double *st = qs[i], *wi = w[i];
        double *fr = alloc_arr(n), *fl = alloc_arr(n);
        EW_LOOP(n) {
            fr[c] = 0.5 * (st[c] + wi[c]);
            fl[c] = 0.5 * (st[c] - wi[c]);
        }
        obj_assign(s, &s->fr[i], fr);
        obj_assign(s, &s->fl[i], fl);
        /* limiter{X}(fluxR, fluxR - fluxR.shift(X,-1), fluxR.shift(X,1) - fluxR) */
        do_shift(s, &s->fr[i], X, -1, s1);
        do_shift(s, &s->fr[i], X, 1, s2);
        {
            const double *a = s->fr[i].a;
            EW_LOOP(n) {
                dl[c] = a[c] - s1[c];
                dr[c] = s2[c] - a[c];
            }
        }
        apply_limiter(s, lim, &s->fr[i], dl, dr);
        /* limiter{X}(fluxL, fluxL.shift(X,-1) - fluxL, fluxL - fluxL.shift(X,1)) */
        do_shift(s, &s->fl[i], X, -1, s1);
        do_shift(s, &s->fl[i], X, 1, s2);
        {
            const double *a = s->fl[i].a;
            EW_LOOP(n) {
                dl[c] = s1[c] - a[c];
                dr[c] = a[c] - s2[c];
            }
        }
        apply_limiter(s, lim, &s->fl[i], dl, dr);
        /* v = v - fluxFactor .* tempFreeze .* (fR - fR.shift(X,-1) + fL - fL.shift(X,1)) */
        do_shift(s, &s->fr[i], X, -1, s1);
        do_shift(s, &s->fl[i], X, 1, s3);
        {
            const double *v = s->prim[i].a, *a = s->fr[i].a, *b = s->fl[i].a;
            double *nv = alloc_arr(n);
            EW_LOOP(n) nv[c] = v[c] - (dtime / dx_cell(s, d, c)) * C2[c] *
                                          (a[c] - s1[c] + b[c] - s3[c]);
            obj_assign(s, &s->prim[i], nv);
        }
        obj_cleanup(&s->fr[i]);
        obj_cleanup(&s->fl[i]);
    }
    {
        double *m = s->prim[0].a, mm = s->cfg.minmass;
        double *r = alloc_arr(n);
        EW_LOOP(n) r[c] = fmax(m[c], mm);
        obj_assign(s, &s->prim[0], r);
    }
    obj_cleanup(&s->freeze_tvd[d]);
    for (int i = 0; i < 5; ++i) free(w[i]);
    free(s1);
    free(s2);
    free(s3);
    free(dl);
    free(dr);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* relaxingMagnet.m, magnetFlux.m, MagnetArray.updateCellCentered                        */
/* ------------------------------------------------------------------------------------ */

static void update_cell_centered(orc_sim *s, int k) {
    long n = s->ncell;
    double *sh = alloc_arr(n), *r = alloc_arr(n);
    do_shift(s, &s->mag[k], k + 1, 1, sh);
    const double *b = s->mag[k].a;
    EW_LOOP(n) r[c] = 0.5 * (b[c] + sh[c]);
    obj_assign(s, &s->cmag[k], r);
    free(sh);
}

static void relaxing_magnet(orc_sim *s, obj_t *vg, int X, int I) {
    long n = s->ncell;
    int dx = X - 1, di = I - 1;
    const double hdt = 0.5 * s->dtime; /* fluxFactor = 0.5*dTime ./ DGRID{X}, per cell */
    double *sh = alloc_arr(n), *sh2 = alloc_arr(n), *vf = alloc_arr(n);
    /* velocityFlow = ((velGrid.array + velGrid.shift(X,1)) < 0) */
    do_shift(s, vg, X, 1, sh);
    {
        const double *v = vg->a;
        EW_LOOP(n) vf[c] = (v[c] + sh[c] < 0.0) ? 1.0 : 0.0;
    }
    obj_t *B = &s->mag[di];
    obj_t *st = &s->mstore[di][dx], *sfr = &s->msfr[di][dx];
    obj_t *wm = &s->wmag[di][dx], *fl = &s->mflux[di][dx];
    /* predictor */
    {
        double *a = alloc_arr(n);
        const double *b = B->a, *v = vg->a;
        EW_LOOP(n) a[c] = b[c] * v[c];
        obj_assign(s, sfr, a);
    }
    do_shift(s, sfr, X, 1, sh);
    {
        double *a = alloc_arr(n);
        const double *f = sfr->a;
        EW_LOOP(n) a[c] = f[c] * (1.0 - vf[c]) + sh[c] * vf[c];
        obj_assign(s, sfr, a);
    }
    do_shift(s, sfr, X, -1, sh);
    {
        double *a = alloc_arr(n);
        const double *b = B->a, *f = sfr->a;
        EW_LOOP(n) a[c] = b[c] - (hdt / dx_cell(s, dx, c)) * (f[c] - sh[c]);
        obj_assign(s, st, a);
    }
    /* corrector: fluxFactor = 2*fluxFactor */
    {
        double *a = alloc_arr(n);
        const double *t = st->a, *v = vg->a;
        EW_LOOP(n) a[c] = t[c] * v[c];
        obj_assign(s, wm, a);
    }
    do_shift(s, wm, X, 1, sh);
    {
        double *a = alloc_arr(n);
        const double *w = wm->a;
        EW_LOOP(n) a[c] = w[c] * (1.0 - vf[c]) + sh[c] * vf[c];
        obj_assign(s, fl, a);
    }
    double *dR = alloc_arr(n), *dL = alloc_arr(n);
    /* dFluxR = wMag.shift(X,1) - (vf.*wMag.shift(X,2) + (1-vf).*wMag) */
    do_shift(s, wm, X, 1, sh);
    do_shift(s, wm, X, 2, sh2);
    {
        const double *w = wm->a;
        EW_LOOP(n) dR[c] = sh[c] - (vf[c] * sh2[c] + (1.0 - vf[c]) * w[c]);
    }
    /* dFluxL = wMag - (vf.*wMag.shift(X,1) + (1-vf).*wMag.shift(X,-1)) */
    do_shift(s, wm, X, 1, sh);
    do_shift(s, wm, X, -1, sh2);
    {
        const double *w = wm->a;
        EW_LOOP(n) dL[c] = w[c] - (vf[c] * sh[c] + (1.0 - vf[c]) * sh2[c]);
    }
    apply_limiter(s, s->cfg.limiter[dx], fl, dL, dR);
    do_shift(s, fl, X, -1, sh);
    {
        double *a = alloc_arr(n);
        const double *b = B->a, *f = fl->a;
        EW_LOOP(n) a[c] = b[c] - (2.0 * (hdt / dx_cell(s, dx, c))) * (f[c] - sh[c]);
        obj_assign(s, B, a);
    }
    /* constrained transport */
    const double dtime = s->dtime; /* fluxFactor = dTime ./ DGRID{I}, per cell */
    do_shift(s, fl, I, 1, sh);
    {
        double *a = alloc_arr(n);
        const double *f = fl->a;
        EW_LOOP(n) a[c] = f[c] - sh[c];
        obj_assign(s, fl, a);
    }
    do_shift(s, fl, X, -1, sh);
    obj_assign(s, fl, dup_arr(sh, n));
    {
        obj_t *BX = &s->mag[dx];
        double *a = alloc_arr(n);
        const double *b = BX->a, *f = fl->a;
        EW_LOOP(n) a[c] = b[c] - (dtime / dx_cell(s, di, c)) * f[c];
        obj_assign(s, BX, a);
    }
    free(sh);
    free(sh2);
    free(vf);
    free(dR);
    free(dL);
}

static void mag_cleanup(orc_sim *s, int k) {
    for (int x = 0; x < 3; ++x) {
        obj_cleanup(&s->vg[k][x]);
        obj_cleanup(&s->mstore[k][x]);
        obj_cleanup(&s->msfr[k][x]);
        obj_cleanup(&s->wmag[k][x]);
        obj_cleanup(&s->mflux[k][x]);
    }
}

int orc_magnet_flux(orc_sim *s, int X, const int I2[2]) {
    if (!s->have_state) return -1;
    long n = s->ncell;
    int dx = X - 1;
    int I = I2[0];
    double *a = alloc_arr(n), *b = alloc_arr(n), *s1 = alloc_arr(n), *s2 = alloc_arr(n);
    for (int t = 0; t < 2; ++t) {
        I = I2[t];
        int di = I - 1;
        /* velGrid = (mom(X).shift(I,-1) + mom(X)) ./ (mass.shift(I,-1) + mass) */
        do_shift(s, &s->prim[X], I, -1, a);
        do_shift(s, &s->prim[0], I, -1, b);
        obj_t *vg = &s->vg[di][dx];
        {
            double *v = alloc_arr(n);
            const double *p = s->prim[X].a, *r = s->prim[0].a;
            EW_LOOP(n) v[c] = (a[c] + p[c]) / (b[c] + r[c]);
            obj_assign(s, vg, v);
        }
        /* velGrid = 0.25*(vg.shift(X,-1) + vg.shift(X,1) + 2*vg) */
        do_shift(s, vg, X, -1, s1);
        do_shift(s, vg, X, 1, s2);
        {
            double *v = alloc_arr(n);
            const double *g = vg->a;
            EW_LOOP(n) v[c] = 0.25 * (s1[c] + s2[c] + 2.0 * g[c]);
            obj_assign(s, vg, v);
        }
        relaxing_magnet(s, vg, X, I);
        mag_cleanup(s, di);
    }
    free(a);
    free(b);
    free(s1);
    free(s2);
    update_cell_centered(s, I - 1); /* only the last I (magnetFlux.m:30) */
    update_cell_centered(s, dx);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* flux.m, TimeManager.update, imogen.m loop body                                        */
/* ------------------------------------------------------------------------------------ */

/* perms([1 2 3]) rows.  Order 0: MATLAB R2009b-R2010b perms.m (recursive construction
 * P = [n q; ... ; i t ...]); order 1: reverse lexicographic (later MATLAB releases). */
static const int PERMS_R2010[6][3] = {{3, 2, 1}, {3, 1, 2}, {2, 3, 1},
                                      {2, 1, 3}, {1, 2, 3}, {1, 3, 2}};
static const int PERMS_LEXREV[6][3] = {{3, 2, 1}, {3, 1, 2}, {2, 3, 1},
                                       {2, 1, 3}, {1, 3, 2}, {1, 2, 3}};

void orc_sweep_order(int iteration, int order, int perms_order, int dirs[3], int mi[3][2]) {
    const int(*P)[3] = perms_order ? PERMS_LEXREV : PERMS_R2010;
    int row = ((iteration - 1) % 6 + 6) % 6;
    static const int FWD[3][2] = {{2, 3}, {1, 3}, {1, 2}};
    static const int BWD[3][2] = {{3, 2}, {3, 1}, {2, 1}};
    for (int t = 0; t < 3; ++t) {
        if (order > 0) {
            dirs[t] = P[row][t];
            mi[t][0] = FWD[dirs[t] - 1][0];
            mi[t][1] = FWD[dirs[t] - 1][1];
        } else {
            dirs[t] = 4 - P[row][t]; /* perms([3 2 1]) = V(perms(1:3)), V = [3 2 1] */
            mi[t][0] = BWD[dirs[t] - 1][0];
            mi[t][1] = BWD[dirs[t] - 1][1];
        }
    }
}

int orc_flux(orc_sim *s, int order) {
    if (order != 1 && order != -1) return -1;
    int dirs[3], mi[3][2];
    orc_sweep_order(s->iteration, order, s->cfg.perms_order, dirs, mi);
    for (int t = 0; t < 3; ++t) {
        int X = dirs[t];
        if (s->n[X - 1] < 3) continue;
        if (order > 0) {
            if (s->cfg.fluid_active && orc_relaxing_fluid(s, X)) return -1;
            if (s->cfg.magnet_active && orc_magnet_flux(s, X, mi[t])) return -1;
        } else {
            if (s->cfg.magnet_active && orc_magnet_flux(s, X, mi[t])) return -1;
            if (s->cfg.fluid_active && orc_relaxing_fluid(s, X)) return -1;
        }
    }
    return 0;
}

int orc_time_update(orc_sim *s, int i) {
    if (!s->have_state) return -1;
    if (i > 1 && s->cfg.update_mode != 0) return 0;
    long n = s->ncell;
    double cmax = 1e-2;
    int gi = 0;
    double *cs = alloc_arr(n);
    int use[3] = {!s->mag_is_zero[0], !s->mag_is_zero[1], !s->mag_is_zero[2]};
    pressure_arr(1, n, s->cfg.gamma, s->prim[0].a, s->prim[1].a, s->prim[2].a, s->prim[3].a,
                 s->prim[4].a, s->cmag[0].a, s->cmag[1].a, s->cmag[2].a, use, cs, NULL);
    for (int k = 1; k <= 3; ++k) {
        const double *p = s->prim[k].a, *r = s->prim[0].a;
        double m = NAN;
        for (long c = 0; c < n; ++c) {
            double v = fabs(p[c] / r[c]) + cs[c];
            m = (c == 0) ? v : fmax(m, v);
        }
        if (m > cmax) {
            cmax = m;
            gi = k;
        }
    }
    free(cs);
    if (gi < 1) {
        snprintf(s->err, sizeof(s->err),
                 "CFL error: no signal speed above 1e-2 (fireCFLError, TimeManager.m:82-85)");
        return -1;
    }
    s->grid_index = gi;
    double upd = (double)(s->cfg.update_mode + 1);
    /* MINDGRID(gridIndex): the spacing itself, or the minimum of its line */
    const double md = s->dline[gi - 1] ? s->mindgrid[gi - 1] : s->cfg.dgrid[gi - 1];
    s->dtime = s->cfg.cfl * md / cmax;
    double newtime = s->time + upd * s->dtime;
    if (newtime > s->cfg.time_max) {
        s->dtime = (s->cfg.time_max - s->time) / upd;
        newtime = s->cfg.time_max;
    }
    s->time = newtime;
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* source.m: gravity (fixed potential) and radiation                                     */
/* ------------------------------------------------------------------------------------ */

/* GravityArray(ENUM.GRAV, run) assigns zeros first, so its edge stores are zero
 * (GravityArray.m:31-35, InitializedArray.m); GravityManager.solvePotential then sets
 * grav.array = fixedPotential at iteration 0 ('empty' solver, GravityManager.m:104-116). */
static void grav_create(orc_sim *s) {
    obj_free(&s->grav);
    obj_init(&s->grav, K_INIT, 0, FT_OTHER);
    double *zeros = alloc_arr(s->ncell);
    memset(zeros, 0, sizeof(double) * (size_t)s->ncell);
    obj_assign(s, &s->grav, zeros);
    obj_assign(s, &s->grav, dup_arr(s->grav_potential, s->ncell));
}

int orc_set_gravity(orc_sim *s, const double *potential, double mass_threshold) {
    free(s->grav_potential);
    s->grav_potential = NULL;
    obj_free(&s->grav);
    s->mass_threshold = mass_threshold;
    if (!potential) return 0;
    if (s->squeezed) {
        /* GravityArray.m:34 sizes the potential by run.gridSize, unsqueezed, while the
         * fluid arrays are squeezed: the source terms would mix the two shapes */
        snprintf(s->err, sizeof(s->err), "gravity on a grid with a singleton x or y");
        return -1;
    }
    for (int d = 0; d < s->ndims; ++d)
        for (int side = 0; side < 2; ++side) {
            /* the 5-point derivative shifts by two cells: on a one-cell dimension every
             * *_shift.m indexes element 0 (MATLAB errors); on two cells only the index
             * maps of circ / const / flip / zero stay inside the array */
            const int m = s->cfg.bc[side][d];
            if (s->n[d] == 1 || (s->n[d] == 2 && !(m == ORC_BC_CIRC || m == ORC_BC_CONST ||
                                                  m == ORC_BC_FLIP || m == ORC_BC_ZERO ||
                                                  m == ORC_BC_WORMHOLE))) {
                snprintf(s->err, sizeof(s->err),
                         "gravity along dimension %d of %d cells: the reference's shift by two "
                         "cells indexes outside the array", d + 1, s->n[d]);
                return -1;
            }
        }
    s->grav_potential = dup_arr(potential, s->ncell);
    grav_create(s);
    return 0;
}

/* ImogenManager.setGridSpacing (ImogenManager.m:190-254) with a non-uniform dimension:
 * DGRID{dim} is one line of n[dim-1] spacings along that dimension (replicated over the
 * others) and MINDGRID(dim) its minimum (minFinderND).  line == NULL: uniform again. */
int orc_set_spacing(orc_sim *s, int dim, const double *line, int n) {
    if (dim < 1 || dim > 3) return -1;
    const int d = dim - 1;
    free(s->dline[d]);
    s->dline[d] = NULL;
    if (!line) return 0;
    if (n != s->n[d]) {
        snprintf(s->err, sizeof(s->err), "spacing line of %d values along a dimension of %d "
                 "cells", n, s->n[d]);
        return -1;
    }
    double mn = line[0];
    for (int i = 0; i < n; ++i) {
        if (!(line[i] > 0)) {
            snprintf(s->err, sizeof(s->err), "spacing must be > 0");
            return -1;
        }
        mn = fmin(mn, line[i]);
    }
    s->dline[d] = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(s->dline[d], line, sizeof(double) * (size_t)n);
    s->mindgrid[d] = mn;
    return 0;
}

/* Radiation.initialize (Radiation.m:61-77): strength from the current state, the
 * unscaled radiation evaluated with the constructor's strength of 1. */
int orc_set_radiation(orc_sim *s, int type, double exponent, double initial_maximum) {
    s->rad_type = type;
    s->rad_exponent = exponent;
    if (type == 0) {
        s->rad_strength = 0.0;
        return 0;
    }
    if (!s->have_state) {
        snprintf(s->err, sizeof(s->err), "radiation needs the state first");
        return -1;
    }
    long n = s->ncell;
    double g = s->cfg.gamma;
    int use[3] = {!s->mag_is_zero[0], !s->mag_is_zero[1], !s->mag_is_zero[2]};
    double *pg = alloc_arr(n);
    pressure_arr(2, n, g, s->prim[0].a, s->prim[1].a, s->prim[2].a, s->prim[3].a, s->prim[4].a,
                 s->cmag[0].a, s->cmag[1].a, s->cmag[2].a, use, pg, NULL);
    double mn = NAN;
    for (long c = 0; c < n; ++c) {
        const double rho = s->prim[0].a[c];
        const double *m1 = s->prim[1].a, *m2 = s->prim[2].a, *m3 = s->prim[3].a;
        const double *b1 = s->mag[0].a, *b2 = s->mag[1].a, *b3 = s->mag[2].a;
        double unscaled = 1.0 * pow(rho, 2.0 - exponent) * pow(pg[c], exponent);
        double ke = 0.5 * (m1[c] * m1[c] + m2[c] * m2[c] + m3[c] * m3[c]) / rho;
        double me = 0.5 * (b1[c] * b1[c] + b2[c] * b2[c] + b3[c] * b3[c]);
        double v = (s->prim[4].a[c] - ke - me) / unscaled;
        if (!isnan(v) && (isnan(mn) || v < mn)) mn = v; /* MATLAB min ignores NaN */
    }
    free(pg);
    s->rad_strength = (initial_maximum / 100.0) * mn;
    return 0;
}

double orc_radiation_strength(const orc_sim *s) { return s->rad_strength; }

/* The viscosity term wFluidFlux adds to the pressure for the current state along X
 * (the predictor's call: FluidArray objects), written to out; trackers are nudged as in
 * the sweep. */
int orc_viscosity_term(orc_sim *s, int X, double *out) {
    if (!s->have_state || X < 1 || X > 3) return -1;
    long n = s->ncell;
    obj_t *pobj[5] = {&s->prim[0], &s->prim[1], &s->prim[2], &s->prim[3], &s->prim[4]};
    EW_LOOP(n) out[c] = 0.0;
    if (s->visc_type == 1) {
        double *press = alloc_arr(n), *snd = alloc_arr(n), *sh = alloc_arr(n);
        int use[3] = {!s->mag_is_zero[0], !s->mag_is_zero[1], !s->mag_is_zero[2]};
        pressure_arr(0, n, s->cfg.gamma, s->prim[0].a, s->prim[1].a, s->prim[2].a,
                     s->prim[3].a, s->prim[4].a, s->cmag[0].a, s->cmag[1].a, s->cmag[2].a, use,
                     press, snd);
        do_shift(s, pobj[X], X, 1, sh);
        const double *qx = pobj[X]->a, *rho = s->prim[0].a;
        EW_LOOP(n) {
            double vm = fmin(sh[c] - qx[c], 0.0);
            out[c] = vm * (s->visc_lin * snd[c] + s->visc_quad * vm / rho[c]);
        }
        free(press);
        free(snd);
        free(sh);
    } else if (s->visc_type == 2) {
        csw_viscosity(s, pobj, X, out);
    }
    return 0;
}

/* run.fluid.viscosity (FluidManager.m:42, initialize.m:395-400): 0 'empty',
 * 1 'neumann_richtmyer', 2 'caramana_shashkov_whalen'. */
int orc_set_viscosity(orc_sim *s, int type, double linear, double quadratic) {
    if (type < 0 || type > 2) {
        snprintf(s->err, sizeof(s->err), "viscosity type %d is not supported", type);
        return -1;
    }
    s->visc_type = type;
    s->visc_lin = linear;
    s->visc_quad = quadratic;
    return 0;
}

/* source(run, mass, mom, ener, mag, grav) (source.m:16-31). */
int orc_source(orc_sim *s) {
    long n = s->ncell;
    double dt = s->dtime;
    if (s->grav_potential) {
        double *a2 = alloc_arr(n), *a1 = alloc_arr(n), *b1 = alloc_arr(n), *b2 = alloc_arr(n);
        double *es = alloc_arr(n), *thr = alloc_arr(n);
        const double *rho = s->prim[0].a;
        double th = s->mass_threshold;
        /* mass.thresholdArray (FluidArray.m:29-34) */
        EW_LOOP(n) {
            es[c] = 0.0;
            thr[c] = th > 0 ? rho[c] * ((rho[c] > th) ? 1.0 : 0.0) : rho[c];
        }
        for (int i = 1; i <= 3; ++i) {
            /* grav.calculate5PtDerivative(i, DGRID{i}) (ImogenArray.m:189-192): the
             * shifts in the order the expression evaluates them */
            do_shift(s, &s->grav, i, 2, a2);
            do_shift(s, &s->grav, i, 1, a1);
            do_shift(s, &s->grav, i, -1, b1);
            do_shift(s, &s->grav, i, -2, b2);
            const double *mom = s->prim[i].a;
            double *nm = alloc_arr(n);
            EW_LOOP(n) {
                const double den = 12.0 * dx_cell(s, i - 1, c); /* ./ (12*dGrid) */
                double d5 = (((-a2[c] + 8.0 * a1[c]) - 8.0 * b1[c]) + b2[c]) / den;
                double ms = (dt * thr[c]) * d5;
                es[c] = es[c] + ms * mom[c] / rho[c];
                nm[c] = mom[c] - ms;
            }
            obj_assign(s, &s->prim[i], nm);
        }
        const double *e = s->prim[4].a;
        double *ne = alloc_arr(n);
        EW_LOOP(n) ne[c] = e[c] - es[c];
        obj_assign(s, &s->prim[4], ne);
        free(a2);
        free(a1);
        free(b1);
        free(b2);
        free(es);
        free(thr);
    }
    double *e = s->prim[4].a;
    if (s->rad_type == 1) {
        /* Radiation.opticallyThinSolver (Radiation.m:86-89) */
        int use[3] = {!s->mag_is_zero[0], !s->mag_is_zero[1], !s->mag_is_zero[2]};
        double *pg = alloc_arr(n);
        pressure_arr(2, n, s->cfg.gamma, s->prim[0].a, s->prim[1].a, s->prim[2].a,
                     s->prim[3].a, e, s->cmag[0].a, s->cmag[1].a, s->cmag[2].a, use, pg, NULL);
        const double *rho = s->prim[0].a;
        double st = s->rad_strength, ex = s->rad_exponent;
        EW_LOOP(n) {
            double rad = st * pow(rho[c], 2.0 - ex) * pow(pg[c], ex);
            e[c] = e[c] - dt * rad;
        }
        free(pg);
    } else {
        /* 'empty' radiation subtracts dt*0 */
        EW_LOOP(n) e[c] = e[c] - dt * 0.0;
    }
    return 0;
}

int orc_iteration(orc_sim *s) {
    static const int direction[2] = {1, -1};
    for (int i = 1; i <= 2; ++i) {
        if (orc_time_update(s, i)) return -1;
        if (orc_flux(s, direction[i - 1])) return -1;
        /* treadmillGrid: inactive; gravity.solvePotential: the fixed potential is set once */
        if (orc_source(s)) return -1;
    }
    s->iteration += 1;
    return 0;
}

int orc_running(const orc_sim *s) {
    return (s->iteration <= s->cfg.iter_max) && (s->time < s->cfg.time_max);
}

double orc_time(const orc_sim *s) { return s->time; }
double orc_dtime(const orc_sim *s) { return s->dtime; }
int orc_iteration_count(const orc_sim *s) { return s->iteration; }
void orc_set_dtime(orc_sim *s, double dt) { s->dtime = dt; }
int orc_last_grid_index(const orc_sim *s) { return s->grid_index; }
const char *orc_last_error(const orc_sim *s) { return s->err; }

/* ------------------------------------------------------------------------------------ */
/* Lifecycle and state                                                                   */
/* ------------------------------------------------------------------------------------ */

orc_sim *orc_create(const orc_config *cfg_in, char *err, int errlen) {
#define FAIL(...) do { if (err) snprintf(err, (size_t)errlen, __VA_ARGS__); return NULL; } while (0)
    for (int d = 0; d < 3; ++d)
        if (cfg_in->grid[d] < 1) FAIL("grid[%d] must be >= 1", d);
    /* FluidArray.m:48 / MagnetArray.m:43 keep squeeze(array): with nz > 1 and a singleton
     * x or y the array is 2-D over the non-singleton dimensions in order, and gridSize
     * (ImogenArray.m:93-96), the shift switches (:292-297), the BC modes, limiters and
     * DGRID, all indexed by dimension number, then act on the squeezed dimensions.  Same
     * cells in the same memory order: the grid is relabelled. */
    orc_config sq = *cfg_in;
    int squeezed = 0;
    if (cfg_in->grid[2] > 1 && (cfg_in->grid[0] == 1 || cfg_in->grid[1] == 1)) {
        int k = 0;
        for (int d = 0; d < 3; ++d)
            if (cfg_in->grid[d] > 1) sq.grid[k++] = cfg_in->grid[d];
        while (k < 3) sq.grid[k++] = 1;
        squeezed = 1;
    }
    const orc_config *cfg = &sq;
    int ndims = cfg->grid[2] > 1 ? 3 : 2;
    for (int sd = 0; sd < 2; ++sd)
        for (int d = 0; d < 3; ++d) {
            int m = cfg->bc[sd][d];
            if (m < ORC_BC_CIRC || m > ORC_BC_WORMHOLE) FAIL("unknown BC mode %d", m);
            /* Every dim <= ndims is shifted at least once: MagnetArray's constructor
             * calls updateCellCentered, which shifts mag(k) along k (MagnetArray.m:36-37,103). */
            if (d >= ndims) continue;
            if (m == ORC_BC_FADE && (cfg->bc_infinity < 4 || cfg->grid[d] < cfg->bc_infinity))
                FAIL("fade BC needs bcInfinity >= 4 and >= bcInfinity cells along dim %d", d + 1);
            if (m == ORC_BC_TRANS && cfg->bc_infinity <= 3)
                FAIL("transparent BC needs bcInfinity > 3");
            if (m == ORC_BC_MIRROR && cfg->grid[d] < 2)
                FAIL("mirror BC needs >= 2 cells along dim %d", d + 1);
            /* wormhole_shift.m:16-27 indexes N(nextDir) and index{previousDir}: a shifted
             * dimension with this mode needs a third dimension in the array */
            if (m == ORC_BC_WORMHOLE && ndims < 3)
                FAIL("wormhole BC along dim %d needs a 3-D grid (wormhole_shift.m indexes a "
                     "third dimension)", d + 1);
        }
    for (int d = 0; d < 3; ++d)
        if (cfg->limiter[d] < 0 || cfg->limiter[d] > 2) FAIL("unknown limiter %d", cfg->limiter[d]);
#undef FAIL
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->cfg = *cfg;
    s->squeezed = squeezed;
    for (int d = 0; d < 3; ++d) s->n[d] = cfg->grid[d];
    s->ndims = ndims;
    s->ncell = (long)s->n[0] * s->n[1] * s->n[2];
    for (int d = 0; d < 3; ++d) s->nface[d] = s->ncell / s->n[d];
    int fsz = cfg->bc_infinity > 0 ? cfg->bc_infinity : 1;
    s->fade = alloc_arr(fsz);
    if (cfg->bc_infinity >= 4) orc_fade_coeffs(cfg->bc_infinity, s->fade);
    s->iteration = 1; /* imogen.m:50 */
    return s;
}

static void all_objs(orc_sim *s, void (*fn)(obj_t *)) {
    for (int i = 0; i < 5; ++i) {
        fn(&s->prim[i]); fn(&s->store[i]); fn(&s->sfr[i]); fn(&s->sfl[i]);
        fn(&s->fr[i]); fn(&s->fl[i]);
    }
    for (int k = 0; k < 3; ++k) {
        fn(&s->freeze[k]); fn(&s->freeze_tvd[k]); fn(&s->mag[k]); fn(&s->cmag[k]);
        for (int x = 0; x < 3; ++x) {
            fn(&s->vg[k][x]); fn(&s->mstore[k][x]); fn(&s->msfr[k][x]);
            fn(&s->wmag[k][x]); fn(&s->mflux[k][x]);
        }
    }
}

void orc_destroy(orc_sim *s) {
    if (!s) return;
    all_objs(s, obj_free);
    obj_free(&s->grav);
    free(s->grav_potential);
    free(s->fade);
    for (int d = 0; d < 3; ++d) free(s->dline[d]);
    free(s);
}

/* sumND (utils/sumND.m): sum along the last dim first, down to dim 1. */
static double sum_nd(const orc_sim *s, const double *a) {
    long nx = s->n[0], ny = s->n[1], nz = s->n[2];
    double *plane = alloc_arr(nx * ny);
    for (long ij = 0; ij < nx * ny; ++ij) {
        double t = 0.0;
        for (long k = 0; k < nz; ++k) t += a[ij + nx * ny * k];
        plane[ij] = t;
    }
    double *col = alloc_arr(nx);
    for (long i = 0; i < nx; ++i) {
        double t = 0.0;
        for (long j = 0; j < ny; ++j) t += plane[i + nx * j];
        col[i] = t;
    }
    double tot = 0.0;
    for (long i = 0; i < nx; ++i) tot += col[i];
    free(plane);
    free(col);
    return tot;
}

int orc_set_state(orc_sim *s, const double *mass, const double *const mom[3],
                  const double *ener, const double *const mag[3]) {
    all_objs(s, obj_free);
    long n = s->ncell;
    /* FluidArray(ENUM.SCALAR/VECTOR(i), id, array, run): edges from the IC */
    const double *src[5] = {mass, mom[0], mom[1], mom[2], ener};
    int comp[5] = {0, 1, 2, 3, 0};
    for (int i = 0; i < 5; ++i) {
        obj_init(&s->prim[i], K_FLUID, comp[i], FT_OTHER);
        s->prim[i].a = dup_arr(src[i], n);
        edges_create(s, &s->prim[i]);
        obj_init(&s->store[i], K_STORE, comp[i], FT_OTHER);
        obj_init(&s->sfr[i], K_FLUX, comp[i], FT_FLUXR);
        obj_init(&s->sfl[i], K_FLUX, comp[i], FT_FLUXL);
        obj_init(&s->fr[i], K_FLUX, comp[i], FT_FLUXR);
        obj_init(&s->fl[i], K_FLUX, comp[i], FT_FLUXL);
    }
    for (int k = 0; k < 3; ++k) {
        obj_init(&s->freeze[k], K_FLUX, k + 1, FT_OTHER);
        obj_init(&s->freeze_tvd[k], K_FLUX, k + 1, FT_OTHER);
        obj_init(&s->mag[k], K_MAGNET, k + 1, FT_OTHER);
        s->mag[k].a = dup_arr(mag[k], n);
        s->mag_is_zero[k] = (sum_nd(s, s->mag[k].a) == 0.0) && !s->cfg.magnet_active;
        edges_create(s, &s->mag[k]);
        for (int x = 0; x < 3; ++x) {
            obj_init(&s->vg[k][x], K_INIT, k + 1, FT_OTHER);
            obj_init(&s->mstore[k][x], K_STORE, k + 1, FT_OTHER);
            obj_init(&s->msfr[k][x], K_FLUX, k + 1, FT_FLUXR);
            obj_init(&s->wmag[k][x], K_INIT, k + 1, FT_OTHER);
            obj_init(&s->mflux[k][x], K_FLUX, k + 1, FT_FLUX);
        }
        obj_init(&s->cmag[k], K_INIT, k + 1, FT_OTHER);
        update_cell_centered(s, k);
    }
    if (s->grav_potential) grav_create(s); /* a new run builds a new GravityArray */
    s->have_state = 1;
    return 0;
}

static const double *field_ptr(const orc_sim *s, int field) {
    if (field >= ORC_F_MASS && field <= ORC_F_ENER) return s->prim[field].a;
    if (field >= ORC_F_MAGX && field <= ORC_F_MAGZ) return s->mag[field - ORC_F_MAGX].a;
    if (field >= ORC_F_CMAGX && field <= ORC_F_CMAGZ) return s->cmag[field - ORC_F_CMAGX].a;
    return NULL;
}

int orc_get_field(const orc_sim *s, int field, double *out) {
    const double *p = field_ptr(s, field);
    if (!p) return -1;
    memcpy(out, p, sizeof(double) * (size_t)s->ncell);
    return 0;
}

int orc_set_field(orc_sim *s, int field, const double *in) {
    double *p = (double *)field_ptr(s, field);
    if (!p) return -1;
    memcpy(p, in, sizeof(double) * (size_t)s->ncell);
    return 0;
}

/* Standalone shift of one object for the unit tests. */
int orc_shift_test(const int grid[3], const int bc[2][3], int bc_infinity, int component,
                   int is_flux_array, int flux_type, int D, int n, const double *in,
                   double *out, double *edge_lo, double *edge_hi, int active_linear) {
    orc_sim s;
    memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) s.n[d] = grid[d];
    s.ndims = grid[2] > 1 ? 3 : 2;
    s.ncell = (long)grid[0] * grid[1] * grid[2];
    for (int d = 0; d < 3; ++d) s.nface[d] = s.ncell / s.n[d];
    memcpy(s.cfg.bc, bc, sizeof(s.cfg.bc));
    s.cfg.bc_infinity = bc_infinity;
    double fade[256];
    if (bc_infinity > 256) return -1;
    if (bc_infinity >= 4) orc_fade_coeffs(bc_infinity, fade);
    s.fade = fade;
    obj_t o;
    obj_init(&o, is_flux_array ? K_FLUX : K_INIT, component, flux_type);
    o.a = (double *)in;
    o.has_edges = 1;
    int d = D - 1;
    o.edge[0][d] = edge_lo;
    o.edge[1][d] = edge_hi;
    /* the linear-index ACTIVE flag is supplied by the caller */
    o.active[(D - 1) % 2][(D - 1) / 2] = active_linear;
    do_shift(&s, &o, D, n, out);
    return 0;
}
