// imh_api.cpp -- host side of the C ABI (include/imogen_hip.h) for one simulation
// context: device memory, the boundary-condition edge-tracker store, the sweep order of
// flux.m, the time loop of imogen.m and the hipGraph replay of an iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "imh_common.h"
#include "imh_host.h"
#include "imh_kernels.h"

using namespace imh;

namespace imh {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(IMH_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// perms([1 2 3]) rows (flux.m:19).  Order 0: the recursive construction of MATLAB
// R2009b-R2010b perms.m; order 1: reverse lexicographic (later releases).
static const int PERMS_R2010[6][3] = {{3, 2, 1}, {3, 1, 2}, {2, 3, 1},
                                      {2, 1, 3}, {1, 2, 3}, {1, 3, 2}};
static const int PERMS_LEXREV[6][3] = {{3, 2, 1}, {3, 1, 2}, {2, 3, 1},
                                       {2, 1, 3}, {1, 3, 2}, {1, 2, 3}};

void sweep_order(int iteration, int order, int perms_order, int dirs[3], int mi[3][2]) {
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
            dirs[t] = 4 - P[row][t];  // perms([3 2 1]) = V(perms(1:3)) with V = [3 2 1]
            mi[t][0] = BWD[dirs[t] - 1][0];
            mi[t][1] = BWD[dirs[t] - 1][1];
        }
    }
}

// FluidArray.m:48 and MagnetArray.m:43 store squeeze(array): a 3-D grid with a singleton
// x or y becomes a 2-D array whose dimensions are the non-singleton ones in order, and
// everything per dimension (gridSize, ImogenArray.m:93-96; the shift switches,
// :292-297; the BC modes, limiters and DGRID, indexed by dimension number) then applies
// to the squeezed array's dimensions.  The cells keep their memory order, so the grid is
// relabelled and nothing moves.  (MATLAB's squeeze leaves 2-D arrays alone.)
imh_config squeeze_grid(const imh_config& in, bool* squeezed) {
    imh_config c = in;
    *squeezed = false;
    if (in.grid[2] > 1 && (in.grid[0] == 1 || in.grid[1] == 1)) {
        int k = 0;
        for (int d = 0; d < 3; ++d)
            if (in.grid[d] > 1) c.grid[k++] = in.grid[d];
        while (k < 3) c.grid[k++] = 1;
        *squeezed = true;
    }
    return c;
}

int validate_config(const imh_config* c) {
    for (int d = 0; d < 3; ++d)
        if (c->grid[d] < 1) return fail(IMH_EINVAL, "grid[%d] must be >= 1", d);
    for (int d = 0; d < 3; ++d)
        if (c->dgrid[d] <= 0) return fail(IMH_EINVAL, "dgrid[%d] must be > 0", d);
    if (c->precision != IMH_FP64 && c->precision != IMH_FP32_BULK)
        return fail(IMH_EINVAL, "unknown precision %d", c->precision);
    const int ndims = c->grid[2] > 1 ? 3 : 2;
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < 3; ++d) {
            int m = c->bc[s][d];
            if (m < IMH_BC_CIRC || m > IMH_BC_WORMHOLE)
                return fail(IMH_EINVAL, "Imogen:UnknownType: Unknown BC mode %d", m);
            if (d >= ndims) continue;  // every dim <= ndims is shifted (MagnetArray.m:36-37,103)
            if (m == IMH_BC_FADE &&
                (c->bc_infinity < 4 || c->bc_infinity > MAX_FADE || c->grid[d] < c->bc_infinity))
                return fail(IMH_ENOTSUP,
                            "fade BC needs 4 <= bcInfinity <= %d and >= bcInfinity cells "
                            "along dim %d (fade_shift.m:36)", MAX_FADE, d + 1);
            if (m == IMH_BC_TRANS && c->bc_infinity <= 3)
                return fail(IMH_EINVAL, "transparent BC needs bcInfinity > 3 (pchip nodes)");
            if (m == IMH_BC_MIRROR && c->grid[d] < 2)
                return fail(IMH_ENOTSUP, "mirror BC needs >= 2 cells along dim %d", d + 1);
            if (m == IMH_BC_WORMHOLE && ndims < 3)
                return fail(IMH_ENOTSUP, "wormhole BC along dim %d needs a 3-D grid "
                            "(wormhole_shift.m indexes a third dimension)", d + 1);
            if (m == IMH_BC_WORMHOLE && c->magnet_active)
                return fail(IMH_ENOTSUP, "wormhole BC with the magnetic field evolved: the "
                            "magnet sweeps take no wormhole ghosts");
        }
    for (int d = 0; d < 3; ++d)
        if (c->limiter[d] < 0 || c->limiter[d] > 2)
            return fail(IMH_EINVAL, "Imogen:UnknownType: Unknown flux limiter type %d",
                        c->limiter[d]);
    if (c->update_mode != 0 && c->update_mode != 1)
        return fail(IMH_EINVAL, "timeUpdateMode must be 0 or 1");
    if (c->cfl <= 0) return fail(IMH_EINVAL, "cfl must be > 0");
    return IMH_OK;
}

static int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Tile of the fused fluid sweep for pencils of length N (fluid_sweep_impl.h).  Compiled
// variants: FP64 (bs, kmax) in {(256,1), (256,2), (512,2), (512,4)}; FP32 in {(256,2),
// (256,4), (512,4), (512,8)}.  Pencils along x are contiguous (threads of a pencil are
// adjacent lanes); along y/z several adjacent pencils share a block so that each row
// access is a contiguous run.
bool fluid_variant_exists(int esz, int bs, int kmax) { return fluid_variant_compiled(esz, bs, kmax); }

FluidTile choose_fluid_tile(int N, int d, int esz, const TileOverride* ov) {
    FluidTile t;
    t.xmajor = (d == 0);
    t.column = 0;
    if (ov && ov->valid) {
        t.bs = ov->bs;
        t.kmax = ov->kmax;
        t.PX = ov->px;
    } else {
        const int kpref = esz == 8 ? 2 : 4;
        int tl = next_pow2((N + kpref - 1) / kpref);
        if (tl < 8) tl = 8;
        if (tl <= 256 && t.xmajor) {
            t.bs = 256;
            t.kmax = kpref;
        } else if (tl <= 512) {
            t.bs = 512;
            t.kmax = kpref;
        } else {
            t.bs = 512;
            t.kmax = 2 * kpref;
            tl = next_pow2((N + t.kmax - 1) / t.kmax);
            if (tl < 8) tl = 8;
        }
        t.PX = t.bs / (tl < t.bs ? tl : t.bs);
        // pencils longer than BS * KMAX of the preferred variant: one pencil per block,
        // the compiled variant with the fewest cells per thread that holds it (the
        // smaller block on a tie)
        if ((N + t.bs - 1) / t.bs > t.kmax) {
            int best_k = 1 << 30;
            for (int bs : {512, 1024}) {
                const int need = (N + bs - 1) / bs;
                for (int k = need; k <= 64; ++k)
                    if (fluid_variant_exists(esz, bs, k)) {
                        if (k < best_k) {
                            best_k = k;
                            t.bs = bs;
                            t.kmax = k;
                            t.PX = 1;
                        }
                        break;
                    }
            }
        }
    }
    if (t.PX > 32) t.PX = 32;
    if (t.PX < 1) t.PX = 1;
    t.TL = t.bs / t.PX;
    t.K = (N + t.TL - 1) / t.TL;
    if (t.K < 1) t.K = 1;
    return t;
}

// The column march applies along y / z when both boundaries are circular or constant.
int column_possible(const imh_config* cfg, int d) {
    if (d == 0) return 0;
    for (int s = 0; s < 2; ++s)
        if (cfg->bc[s][d] != IMH_BC_CIRC && cfg->bc[s][d] != IMH_BC_CONST) return 0;
    return 1;
}

MagTile choose_mag_tile(int N) {
    MagTile t;
    int tl = next_pow2((N + 3) / 4);
    if (tl > 256) tl = 256;
    if (tl < 16) tl = 16;
    t.TL = tl;
    t.PX = 256 / tl;
    return t;
}

}  // namespace imh

// ======================================================================================
// context
// ======================================================================================

static size_t esize(const imh_ctx* c) { return c->cfg.precision == IMH_FP64 ? 8 : 4; }

static void ctx_free(imh_ctx* c) {
    if (!c) return;
    for (int i = 0; i < 6; ++i)
        if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
    for (int i = 0; i < IMH_NFIELDS; ++i)
        if (c->f[i]) hipFree(c->f[i]);
    if (c->scratch) hipFree(c->scratch);
    for (int d = 0; d < 3; ++d)
        if (c->dline_dev[d]) hipFree(c->dline_dev[d]);
    if (c->edges) hipFree(c->edges);
    if (c->eoff_dev) hipFree(c->eoff_dev);
    if (c->sim) hipFree(c->sim);
    if (c->par) hipFree(c->par);
    if (c->arg_host) hipHostFree(c->arg_host);
    if (c->arg_dev) hipFree(c->arg_dev);
    for (int i = 0; i < 6; ++i)
        if (c->cap_dev[i]) hipFree(c->cap_dev[i]);
    delete[] c->cap_host;
    if (c->sim_ptr) hipFree(c->sim_ptr);
    if (c->par_ptr) hipFree(c->par_ptr);
    if (c->fade_dev) hipFree(c->fade_dev);
    if (c->csum_part) hipFree(c->csum_part);
    if (c->csum_out) hipFree(c->csum_out);
    if (c->long_scratch) hipFree(c->long_scratch);
    if (c->grav) hipFree(c->grav);
    if (c->grav_ghost) hipFree(c->grav_ghost);
    if (c->st && !c->external_stream) hipStreamDestroy(c->st);
    delete c;
}

void imh_ctx_free(imh_ctx* c) { ctx_free(c); }

namespace imh {

void prof_begin(Profiler* p, hipStream_t st, int kind, double bytes) {
    ProfRec r{kind, bytes, nullptr, nullptr};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess ||
        hipEventRecord(r.a, st) != hipSuccess)
        p->failed = true;
    p->recs.push_back(r);
}

void prof_end(Profiler* p, hipStream_t st) {
    if (p->recs.empty() || hipEventRecord(p->recs.back().b, st) != hipSuccess) p->failed = true;
}

int prof_collect(Profiler* p, double* ms, double* bytes, int* count) {
    for (int k = 0; k < IMH_NKIND; ++k) {
        if (ms) ms[k] = 0;
        if (bytes) bytes[k] = 0;
        if (count) count[k] = 0;
    }
    hipError_t err = hipSuccess;
    for (ProfRec& r : p->recs) {
        float t = 0;
        hipError_t e = (r.a && r.b) ? hipEventSynchronize(r.b) : hipErrorInvalidValue;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
        if (e != hipSuccess && err == hipSuccess) err = e;
        if (e == hipSuccess && r.kind >= 0 && r.kind < IMH_NKIND) {
            if (ms) ms[r.kind] += t;
            if (bytes) bytes[r.kind] += r.bytes;
            if (count) count[r.kind] += 1;
        }
        if (r.a) hipEventDestroy(r.a);
        if (r.b) hipEventDestroy(r.b);
    }
    p->recs.clear();
    if (p->failed) return fail(IMH_EHIP, "event creation or recording failed while profiling");
    return err == hipSuccess ? IMH_OK : hip_fail(err, "profile events");
}

}  // namespace imh

// Event pair around one launch on c->st while a profile runs.
struct ProfScope {
    imh_ctx* c;
    ProfScope(imh_ctx* ctx, int kind, double bytes) : c(ctx) {
        if (c->prof) prof_begin(c->prof, c->st, kind, bytes);
    }
    ~ProfScope() {
        if (c->prof) prof_end(c->prof, c->st);
    }
};

const void* imh_stage_args(imh_ctx* c, const void* a, size_t n) {
    if (n > ARG_SLOT) {
        fail(IMH_EINVAL, "argument block of %zu bytes exceeds a slot", n);
        return nullptr;
    }
    if (c->capture_row >= 0) {
        if (c->cap_count >= CAP_SLOTS) {
            fail(IMH_EINVAL, "more than %d launches in one captured iteration", CAP_SLOTS);
            return nullptr;
        }
        const int k = c->cap_count++;
        memcpy(c->cap_host + k * ARG_SLOT, a, n);
        return c->cap_dev[c->capture_row] + k * ARG_SLOT;
    }
    if (c->arg_next == ARG_RING) {
        // every earlier copy out of the ring has to be done before a slot is rewritten
        hipError_t e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) {
            hip_fail(e, "argument ring");
            return nullptr;
        }
        c->arg_next = 0;
    }
    const int k = c->arg_next++;
    char* h = c->arg_host + k * ARG_SLOT;
    memcpy(h, a, n);
    hipError_t e = hipMemcpyAsync(c->arg_dev + k * ARG_SLOT, h, n, hipMemcpyHostToDevice, c->st);
    if (e != hipSuccess) {
        hip_fail(e, "argument copy");
        return nullptr;
    }
    return c->arg_dev + k * ARG_SLOT;
}

// Offsets of the edge stores (objects x sides x dims) for trans/wall sides; returns the
// number of doubles they need.
static long long edges_layout(const imh_config& cfg, long long eoff[NOBJ][2][3]) {
    const int ndims = cfg.grid[2] > 1 ? 3 : 2;
    const long long ncell = (long long)cfg.grid[0] * cfg.grid[1] * cfg.grid[2];
    long long off = 0;
    for (int o = 0; o < NOBJ; ++o)
        for (int s = 0; s < 2; ++s)
            for (int d = 0; d < 3; ++d) eoff[o][s][d] = 0;
    auto exists = [&](int o) {
        if (o < OBJ_MAG + 3 || o >= OBJ_GRAV) return true;
        if (!cfg.magnet_active) return false;
        int k = (o - OBJ_VG) % 9;
        return (k / 3) != (k % 3);  // (I, X) pairs with I != X
    };
    for (int o = 0; o < NOBJ; ++o) {
        if (o >= OBJ_MFLUX + 9 && o < OBJ_GRAV) continue;
        if (!exists(o)) continue;
        for (int s = 0; s < 2; ++s)
            for (int d = 0; d < ndims; ++d) {
                int m = cfg.bc[s][d];
                if (m != IMH_BC_TRANS && m != IMH_BC_WALL) continue;
                eoff[o][s][d] = off;
                off += ncell / cfg.grid[d];
            }
    }
    return off > 0 ? off : 1;
}

static void layout_edges(imh_ctx* c) { c->edges_len = (size_t)edges_layout(c->cfg, c->eoff); }

// Device bytes of a context (fields, CT scratch, edge stores, small bookkeeping).
static size_t ctx_bytes(const imh_config& cfg, size_t edges_len) {
    const size_t es = cfg.precision == IMH_FP64 ? 8 : 4;
    const size_t ncell = (size_t)cfg.grid[0] * cfg.grid[1] * cfg.grid[2];
    size_t b = ncell * es * (IMH_NFIELDS + (cfg.magnet_active ? 1 : 0));
    b += edges_len * sizeof(double) + sizeof(long long) * NOBJ * 6 + 2 * sizeof(SimDev);
    b += ARG_SLOT * ARG_RING + 6 * ARG_SLOT * CAP_SLOTS + sizeof(double) * (2 * 1024 * 8 + 8) +
         sizeof(double) * MAX_FADE;
    return b;
}

extern "C" {

const char* imh_last_error(void) { return imh::g_last_error.c_str(); }

const char* imh_version(void) { return "imogen-hip 0.1 (gfx950)"; }

}  // extern "C"

// Host-side fields of a context from its (squeezed, validated) configuration.
static int ctx_init_host(imh_ctx* c, const imh_config* cfg, bool squeezed) {
    c->cfg = *cfg;
    c->squeezed = squeezed;
    for (int d = 0; d < 3; ++d) c->n[d] = cfg->grid[d];
    c->ndims = cfg->grid[2] > 1 ? 3 : 2;
    c->ncell = (long long)c->n[0] * c->n[1] * c->n[2];
    if (c->ncell >= (1LL << 31))
        return fail(IMH_ENOTSUP, "a simulation of %lld cells (the kernels index cells with 32 "
                    "bits; split it into grids below 2^31 cells)", c->ncell);
    for (int d = 0; d < 3; ++d) c->nface[d] = c->ncell / c->n[d];
    for (int d = 0; d < 3; ++d) {
        int L = d;  // Edges.ACTIVE(D) with linear index D = d+1 into the 2x3 array
        int row = L % 2, col = L / 2;
        int m = cfg->bc[row][col];
        c->active_lin[d] = (col < c->ndims) &&
                           (m == IMH_BC_TRANS || m == IMH_BC_FADE || m == IMH_BC_WALL);
    }
    if (cfg->bc_infinity >= 4 && cfg->bc_infinity <= MAX_FADE)
        fade_coefficients(cfg->bc_infinity, c->fade);
    layout_edges(c);
    return IMH_OK;
}

#define CHK(x, what)                                                                       \
    do {                                                                                   \
        hipError_t _e = (x);                                                               \
        if (_e != hipSuccess)                                                              \
            return (_e == hipErrorOutOfMemory) ? fail(IMH_ENOMEM, "%s: out of device memory", \
                                                      what)                                \
                                               : hip_fail(_e, what);                       \
    } while (0)

// Device memory of a context: its stream, the fields, edge stores and bookkeeping.
static int ctx_alloc(imh_ctx* c) {
    const imh_config* cfg = &c->cfg;
    const size_t bytes = (size_t)c->ncell * esize(c);
    CHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking), "hipStreamCreate");
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess && prop.multiProcessorCount > 0) {
            c->num_cu = prop.multiProcessorCount;
            // dynamic LDS above 64 KiB would need a per-kernel attribute: the wave-spread
            // sweep stays within 64 KiB
            c->max_lds = std::min<size_t>(prop.sharedMemPerBlock, 64 * 1024);
        }
    }
    for (int i = 0; i < IMH_NFIELDS; ++i) CHK(hipMalloc(&c->f[i], bytes), "hipMalloc(field)");
    if (cfg->magnet_active) CHK(hipMalloc(&c->scratch, bytes), "hipMalloc(scratch)");
    c->edges_cap = c->edges_len;
    CHK(hipMalloc(&c->edges, c->edges_cap * sizeof(double)), "hipMalloc(edges)");
    CHK(hipMalloc(&c->eoff_dev, sizeof(c->eoff)), "hipMalloc(eoff)");
    CHK(hipMalloc(&c->fade_dev, sizeof(c->fade)), "hipMalloc(fade)");
    CHK(hipMalloc(&c->sim, sizeof(SimDev)), "hipMalloc(sim)");
    CHK(hipMalloc(&c->par, sizeof(SimParams)), "hipMalloc(par)");
    CHK(hipMalloc(&c->sim_ptr, sizeof(SimDev*)), "hipMalloc(sim_ptr)");
    CHK(hipMalloc(&c->par_ptr, sizeof(SimParams*)), "hipMalloc(par_ptr)");
    CHK(hipMemcpy(c->sim_ptr, &c->sim, sizeof(SimDev*), hipMemcpyHostToDevice), "sim_ptr");
    CHK(hipMemcpy(c->par_ptr, &c->par, sizeof(SimParams*), hipMemcpyHostToDevice), "par_ptr");
    CHK(hipMalloc(&c->csum_part, sizeof(double) * 2 * 1024 * 8), "hipMalloc(csum)");
    CHK(hipMalloc(&c->csum_out, sizeof(double) * 8), "hipMalloc(csum_out)");
    CHK(hipHostMalloc((void**)&c->arg_host, ARG_SLOT * ARG_RING, hipHostMallocDefault),
        "hipHostMalloc(args)");
    CHK(hipMalloc(&c->arg_dev, ARG_SLOT * ARG_RING), "hipMalloc(args)");
    c->cap_host = new char[ARG_SLOT * CAP_SLOTS];
    return IMH_OK;
}

// 3-D grids do not take the multi-launch sweep (fluid_long.hip): its one 3-D run (MHD,
// Caramana-Shashkov-Whalen viscosity at fade ends, gravity and radiation) faulted the
// device while a serial host run of the same kernels on the same configuration kept every
// access in bounds; until that is explained, such sweeps are refused.
static const char* const LONG_3D_REFUSAL =
    "the multi-launch sweep (pencils longer than a workgroup holds, or this viscosity at a "
    "fade end) runs on 1-D and 2-D grids only";

// The scratch area of the multi-launch sweep (fluid_long.hip) along d
static hipError_t ensure_long_scratch(imh_ctx* c, int d) {
    const size_t lb = fluid_long_bytes(esize(c), c->ncell, c->nface[d]);
    if (lb <= c->long_bytes) return hipSuccess;
    if (c->long_scratch) hipFree(c->long_scratch);
    c->long_scratch = nullptr;
    c->long_bytes = 0;
    hipError_t e = hipMalloc(&c->long_scratch, lb);
    if (e == hipSuccess) c->long_bytes = lb;
    return e;
}

// Configuration-dependent device data and tiles of a context whose memory is in place
// (fresh or recycled): zeroed edge stores, offsets, fade coefficients, time state and
// parameters, sweep tiles.
static int ctx_configure(imh_ctx* c) {
    const imh_config* cfg = &c->cfg;
    const size_t es = esize(c);
    CHK(hipMemset(c->edges, 0, c->edges_cap * sizeof(double)), "hipMemset(edges)");
    CHK(hipMemcpy(c->eoff_dev, c->eoff, sizeof(c->eoff), hipMemcpyHostToDevice), "eoff upload");
    CHK(hipMemcpy(c->fade_dev, c->fade, sizeof(c->fade), hipMemcpyHostToDevice), "fade upload");
    SimDev sd;
    memset(&sd, 0, sizeof(sd));
    sd.iteration = 1;  // imogen.m:50
    sd.active = 1;
    CHK(hipMemcpy(c->sim, &sd, sizeof(sd), hipMemcpyHostToDevice), "sim upload");
    c->host_par.cfl = cfg->cfl;
    for (int d = 0; d < 3; ++d) c->host_par.dgrid[d] = cfg->dgrid[d];
    c->host_par.time_max = cfg->time_max;
    c->host_par.iter_max = cfg->iter_max;
    c->host_par.update_mode = cfg->update_mode;
    c->host_par.unconditional = 0;
    CHK(hipMemcpy(c->par, &c->host_par, sizeof(SimParams), hipMemcpyHostToDevice), "par upload");
    for (int d = 0; d < 3; ++d) {
        c->ftile[d] = choose_fluid_tile(c->n[d], d, (int)es, nullptr);
        // the column march has one thread per pencil: it pays where a sweep has many
        // pencils (3-D grids); the autotuner can switch it on for a dimension
        c->ftile[d].column = column_possible(cfg, d) && c->nface[d] >= (1 << 16);
        c->mtile[d] = choose_mag_tile(c->n[d]);
        c->long_dim[d] = false;
        if (c->n[d] >= 3 && c->ftile[d].K > c->ftile[d].kmax) {
            // longer than a workgroup holds: the multi-launch sweep (fluid_long.hip)
            if (cfg->bc[0][d] == IMH_BC_WORMHOLE)
                return fail(IMH_ENOTSUP, "pencils of %d cells along dim %d with a wormhole "
                            "boundary (the sweep of pencils longer than %d cells does not pair "
                            "pencils)", cfg->grid[d], d + 1, c->ftile[d].kmax * c->ftile[d].bs);
            if (cfg->magnet_active && c->n[d] > MAG_PENCIL_MAX)
                return fail(IMH_ENOTSUP, "magnet pencils of %d cells along dim %d (at most %d)",
                            cfg->grid[d], d + 1, MAG_PENCIL_MAX);
            if (c->ndims == 3) return fail(IMH_ENOTSUP, "%s", LONG_3D_REFUSAL);
            c->long_dim[d] = true;
            CHK(ensure_long_scratch(c, d), "hipMalloc(long sweep)");
        }
        // a lower wormhole boundary pairs pencils inside a tile: long pencils take two
        // pencils of 256 threads per 512-thread block (more cells per thread)
        if (c->n[d] >= 3 && d < c->ndims && cfg->bc[0][d] == IMH_BC_WORMHOLE &&
            (c->ftile[d].PX & 1)) {
            bool ok = false;
            for (int k = (c->n[d] + 255) / 256; k <= 64 && !ok; ++k) {
                if (!fluid_variant_exists((int)es, 512, k)) continue;
                const TileOverride ov = {1, 512, k, 2};
                const FluidTile t = choose_fluid_tile(c->n[d], d, (int)es, &ov);
                if (t.K <= t.kmax) {
                    c->ftile[d] = t;
                    ok = true;
                }
            }
            if (!ok)
                return fail(IMH_ENOTSUP, "wormhole BC along dim %d: pencils of %d cells need "
                            "a sweep tile of their own, and a wormhole pairs two pencils per "
                            "tile", d + 1, cfg->grid[d]);
        }
    }
    c->host_iteration = 1;
    return IMH_OK;
}
#undef CHK

extern "C" {

int imh_create(const imh_config* cfg_in, imh_ctx** out) {
    if (!cfg_in || !out) return fail(IMH_EINVAL, "null argument");
    *out = nullptr;
    bool squeezed = false;
    const imh_config sq = squeeze_grid(*cfg_in, &squeezed);
    const imh_config* cfg = &sq;
    int rc = validate_config(cfg);
    if (rc) return rc;
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    imh_ctx* c = new imh_ctx();
    if ((rc = ctx_init_host(c, cfg, squeezed)) || (rc = ctx_alloc(c)) || (rc = ctx_configure(c))) {
        ctx_free(c);
        return rc;
    }
    *out = c;
    return IMH_OK;
}

}  // extern "C"

// A new context for cfg on the device memory of `old` (a finished context that is no
// longer used): same cell count, precision, field set (the magnet's scratch) and room
// for the edge stores.  Everything else starts fresh, as from imh_create; `old` is
// consumed whether or not the call succeeds.  IMH_ENOTSUP (old untouched and not
// consumed) when the memory does not fit.
int imh_ctx_recycle(imh_ctx* old, const imh_config* cfg_in, imh_ctx** out) {
    if (!old || !cfg_in || !out) return fail(IMH_EINVAL, "null argument");
    *out = nullptr;
    bool squeezed = false;
    const imh_config sq = squeeze_grid(*cfg_in, &squeezed);
    int rc = validate_config(&sq);
    if (rc) return rc;
    imh_ctx* c = new imh_ctx();
    if ((rc = ctx_init_host(c, &sq, squeezed))) {
        delete c;
        return rc;
    }
    if (c->ncell != old->ncell || sq.precision != old->cfg.precision ||
        sq.device != old->cfg.device || (sq.magnet_active && !old->scratch) ||
        c->edges_len > old->edges_cap || old->external_stream) {
        delete c;
        return fail(IMH_ENOTSUP, "the context's memory does not fit this configuration");
    }
    hipError_t e = hipStreamSynchronize(old->st);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "sync");
    }
    // move the device memory over; the old context keeps (and frees) everything else
    c->st = old->st;
    old->st = nullptr;
    c->num_cu = old->num_cu;
    c->max_lds = old->max_lds;
    for (int i = 0; i < IMH_NFIELDS; ++i) {
        c->f[i] = old->f[i];
        old->f[i] = nullptr;
    }
    c->scratch = old->scratch;
    old->scratch = nullptr;
    c->edges = old->edges;
    c->edges_cap = old->edges_cap;
    old->edges = nullptr;
#define IMH_MOVE(field) \
    c->field = old->field; \
    old->field = nullptr;
    IMH_MOVE(eoff_dev)
    IMH_MOVE(fade_dev)
    IMH_MOVE(sim)
    IMH_MOVE(par)
    IMH_MOVE(sim_ptr)
    IMH_MOVE(par_ptr)
    IMH_MOVE(csum_part)
    IMH_MOVE(csum_out)
    IMH_MOVE(arg_host)
    IMH_MOVE(arg_dev)
    IMH_MOVE(cap_host)
#undef IMH_MOVE
    ctx_free(old);
    if ((rc = ctx_configure(c))) {
        ctx_free(c);
        return rc;
    }
    *out = c;
    return IMH_OK;
}

extern "C" {

int imh_destroy(imh_ctx* c) {
    if (!c) return IMH_OK;
    hipStreamSynchronize(c->st);
    ctx_free(c);
    return IMH_OK;
}

}  // extern "C"

// ---- state --------------------------------------------------------------------------

static int upload_array(imh_ctx* c, void* dst, const double* src) {
    hipError_t e;
    if (c->cfg.precision == IMH_FP64) {
        e = hipMemcpyAsync(dst, src, sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    } else {
        std::vector<float> tmp(c->ncell);
        for (long long i = 0; i < c->ncell; ++i) tmp[i] = (float)src[i];
        e = hipMemcpyAsync(dst, tmp.data(), sizeof(float) * c->ncell, hipMemcpyHostToDevice, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    return e == hipSuccess ? IMH_OK : hip_fail(e, "upload");
}

// sumND (utils/sumND.m): sum along the last dim first, down to dim 1.
static double sum_nd(const imh_ctx* c, const double* a) {
    long long nx = c->n[0], ny = c->n[1], nz = c->n[2];
    std::vector<double> plane(nx * ny), col(nx);
    for (long long ij = 0; ij < nx * ny; ++ij) {
        double t = 0.0;
        for (long long k = 0; k < nz; ++k) t += a[ij + nx * ny * k];
        plane[ij] = t;
    }
    for (long long i = 0; i < nx; ++i) {
        double t = 0.0;
        for (long long j = 0; j < ny; ++j) t += plane[i + nx * j];
        col[i] = t;
    }
    double tot = 0.0;
    for (long long i = 0; i < nx; ++i) tot += col[i];
    return tot;
}

// Edge stores of an object built from an array (Edges.m:37-108): face slices of the
// trans/wall sides.  `h` is the host copy of the edges buffer.
static void host_edges_from(imh_ctx* c, std::vector<double>& h, int obj, const double* a) {
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < c->ndims; ++d) {
            int m = c->cfg.bc[s][d];
            if (m != IMH_BC_TRANS && m != IMH_BC_WALL) continue;
            long long off = c->eoff[obj][s][d];
            long long N = c->n[d];
            for (long long fi = 0; fi < c->nface[d]; ++fi) {
                long long base, stride;
                pencil_base(c->n, d, fi, base, stride);
                h[off + fi] = a[base + (s ? (N - 1) : 0) * stride];
            }
        }
}

template <typename T>
CellMagArgs<T> make_cellmag_args(const imh_ctx* c, int k, int force) {
    CellMagArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.b = (const T*)c->f[IMH_F_MAGX + k];
    a.cb = (T*)c->f[IMH_F_CMAGX + k];
    a.edges = c->edges;
    a.eoff = c->eoff_dev;
    a.sim = c->sim;
    a.n0 = c->n[0];
    a.n1 = c->n[1];
    a.n2 = c->n[2];
    a.k = k;
    a.ndims = c->ndims;
    a.modes = pack_modes(c->cfg.bc);
    a.active_lin = c->active_lin[k];
    a.bc_inf = c->cfg.bc_infinity;
    a.force = force;
    a.fade = c->fade_dev;
    return a;
}
template CellMagArgs<double> make_cellmag_args<double>(const imh_ctx*, int, int);
template CellMagArgs<float> make_cellmag_args<float>(const imh_ctx*, int, int);

template <typename T>
static int launch_cellmag(imh_ctx* c, int k, int force) {
    CellMagArgs<T> a = make_cellmag_args<T>(c, k, force);
    const void* da = imh_stage_args(c, &a, sizeof(a));
    if (!da) return IMH_EHIP;
    const int nt = int((c->ncell + 255) / 256);
    ProfScope ps(c, IMH_K_CELLMAG, 2.0 * c->ncell * sizeof(T));
    hipError_t e = launch_cell_centered_batch((const CellMagArgs<T>*)da, nullptr, nt, nt, c->st);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "cell_centered launch");
}

static int cellmag(imh_ctx* c, int k, int force) {
    return c->cfg.precision == IMH_FP64 ? launch_cellmag<double>(c, k, force)
                                        : launch_cellmag<float>(c, k, force);
}

// The cellMag InitializedArrays take their edge stores from their first array, the one
// the MagnetArray constructor computes (InitializedArray.m:43-47).
static int cellmag_edges(imh_ctx* c) {
    for (int k = 0; k < 3; ++k)
        for (int s = 0; s < 2; ++s)
            for (int d = 0; d < c->ndims; ++d) {
                const int m = c->cfg.bc[s][d];
                if (m != IMH_BC_TRANS && m != IMH_BC_WALL) continue;
                hipError_t e = launch_edges_from_field(c->f[IMH_F_CMAGX + k],
                                                       c->cfg.precision == IMH_FP64, c->n, d, s,
                                                       c->edges + c->eoff[OBJ_CMAG + k][s][d],
                                                       c->nface[d], c->st);
                if (e != hipSuccess) return hip_fail(e, "cellMag edges");
            }
    return IMH_OK;
}

extern "C" int imh_upload_state(imh_ctx* c, const double* mass, const double* const mom[3],
                                const double* ener, const double* const mag[3], size_t n) {
    if (!c || !mass || !mom || !ener || !mag) return fail(IMH_EINVAL, "null argument");
    if ((long long)n != c->ncell)
        return fail(IMH_EINVAL, "array length %zu does not match the grid (%lld cells)", n,
                    c->ncell);
    for (int k = 0; k < 3; ++k)
        if (!mom[k] || !mag[k]) return fail(IMH_EINVAL, "null component array");
    const double* src[8] = {mass, mom[0], mom[1], mom[2], ener, mag[0], mag[1], mag[2]};
    for (int i = 0; i < 8; ++i) {
        int rc = upload_array(c, c->f[i], src[i]);
        if (rc) return rc;
    }
    // edge stores of the FluidArray and MagnetArray objects come from the IC
    std::vector<double> h(c->edges_len, 0.0);
    for (int i = 0; i < 5; ++i) host_edges_from(c, h, OBJ_PRIM + i, src[i]);
    for (int k = 0; k < 3; ++k) host_edges_from(c, h, OBJ_MAG + k, src[5 + k]);
    hipError_t e = hipMemcpy(c->edges, h.data(), sizeof(double) * c->edges_len,
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "edges upload");
    // MagnetArray.isZero = (sumND(array) == 0) && ~run.magnet.ACTIVE
    bool anyB = false;
    for (int k = 0; k < 3; ++k) {
        c->mag_is_zero[k] = (sum_nd(c, src[5 + k]) == 0.0) && !c->cfg.magnet_active;
        for (long long i = 0; i < c->ncell && !anyB; ++i) anyB = src[5 + k][i] != 0.0;
    }
    c->mhd_kernel = c->cfg.magnet_active || anyB;
    // time state of a new run (imogen.m:50)
    SimDev sd;
    memset(&sd, 0, sizeof(sd));
    sd.iteration = 1;
    sd.active = 1;
    e = hipMemcpy(c->sim, &sd, sizeof(sd), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "sim reset");
    c->host_iteration = 1;
    c->fluid_init = false;
    for (int o = 0; o < NOBJ; ++o) c->obj_init[o] = false;
    for (int i = 0; i < 6; ++i) {
        if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
        c->graph[i] = nullptr;
    }
    // cellMag from updateCellCentered in the MagnetArray constructor (MagnetArray.m:102-103)
    for (int k = 0; k < 3; ++k) {
        int rc = cellmag(c, k, 1);
        if (rc) return rc;
    }
    if (int rc = cellmag_edges(c)) return rc;
    e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "cellMag");
    c->have_state = true;
    return IMH_OK;
}

extern "C" int imh_upload_field(imh_ctx* c, int field, const double* data, size_t n) {
    if (!c || !data) return fail(IMH_EINVAL, "null argument");
    if (field < 0 || field >= IMH_NFIELDS) return fail(IMH_EINVAL, "bad field id %d", field);
    if ((long long)n != c->ncell) return fail(IMH_EINVAL, "array length mismatch");
    return upload_array(c, c->f[field], data);
}

extern "C" int imh_download_field(imh_ctx* c, int field, double* out, size_t n) {
    if (!c || !out) return fail(IMH_EINVAL, "null argument");
    if (field < 0 || field >= IMH_NFIELDS) return fail(IMH_EINVAL, "bad field id %d", field);
    if ((long long)n != c->ncell) return fail(IMH_EINVAL, "array length mismatch");
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "stream sync");
    if (c->cfg.precision == IMH_FP64) {
        e = hipMemcpy(out, c->f[field], sizeof(double) * n, hipMemcpyDeviceToHost);
    } else {
        std::vector<float> tmp(n);
        e = hipMemcpy(tmp.data(), c->f[field], sizeof(float) * n, hipMemcpyDeviceToHost);
        for (size_t i = 0; i < n; ++i) out[i] = tmp[i];
    }
    return e == hipSuccess ? IMH_OK : hip_fail(e, "download");
}

// ---- sweeps -------------------------------------------------------------------------

template <typename T>
FluidArgs<T> make_fluid_args(const imh_ctx* c, int d, int force) {
    FluidArgs<T> a;
    memset(&a, 0, sizeof(a));
    for (int v = 0; v < 5; ++v) a.q[v] = (T*)c->f[v];
    for (int k = 0; k < 3; ++k) a.bc[k] = (const T*)c->f[IMH_F_CMAGX + k];
    a.edges = c->edges;
    a.eoff = c->eoff_dev;
    a.sim = c->sim;
    a.n0 = c->n[0];
    a.n1 = c->n[1];
    a.n2 = c->n[2];
    a.d = d;
    a.ndims = c->ndims;
    a.modes = pack_modes(c->cfg.bc);
    a.active_lin = c->active_lin[d];
    a.limiter = c->cfg.limiter[d];
    a.use = {!c->mag_is_zero[0], !c->mag_is_zero[1], !c->mag_is_zero[2]};
    a.gamma = c->cfg.gamma;
    a.minmass = c->cfg.minmass;
    a.dx = c->cfg.dgrid[d];
    a.dxl = c->dline_dev[d];
    a.init_edges = c->fluid_init ? 0 : 1;
    const FluidTile& t = c->ftile[d];
    a.PX = t.PX;
    a.TL = t.TL;
    a.K = t.K;
    a.xmajor = t.xmajor;
    a.worm = c->cfg.bc[0][d] == IMH_BC_WORMHOLE;
    a.npencil = a.worm ? worm_slots(c->n, d) : c->nface[d];
    a.bc_inf = c->cfg.bc_infinity;
    a.xcd_remap = 1;
    a.force = force;
    a.fade = c->fade_dev;
    a.visc_type = c->visc_type;
    a.visc_lin = c->visc_lin;
    a.visc_quad = c->visc_quad;
    return a;
}
template FluidArgs<double> make_fluid_args<double>(const imh_ctx*, int, int);
template FluidArgs<float> make_fluid_args<float>(const imh_ctx*, int, int);

template <typename T>
static int fluid_sweep_t(imh_ctx* c, int d, int force) {
    FluidArgs<T> a = make_fluid_args<T>(c, d, force);
    const FluidTile& t = c->ftile[d];
    const void* da = imh_stage_args(c, &a, sizeof(a));
    if (!da) return IMH_EHIP;
    ProfScope ps(c, IMH_K_FLUID, double(10 + (c->mhd_kernel ? 3 : 0)) * c->ncell * sizeof(T));
    hipError_t e;
    int wbs = 0, wpx = 0, wtl = 0;
    size_t wlds = 0;
    // the wave-spread sweep where it applies (with IMH_SWEEP_AUTO_FAST only where its
    // tiles fill the device, at least 4 per CU: short sweeps keep the general form's
    // smaller tiles)
    // Hydro sweeps only: a run of the MHD instantiations (the only ones that spill
    // registers to scratch) left the GPU faulted once, cause not found; MHD sweeps take the
    // general form until it is.
    if (long_sweep(c, d)) {
        e = launch_fluid_long<T>(a, c->long_scratch, c->ncell, c->mhd_kernel, c->st);
    } else if (c->sweep_kind != IMH_SWEEP_GENERAL && !fluid_needs_general(a) && !c->mhd_kernel && !a.dxl &&
        wave_bc_ok(a.modes, d) && wave_tile(c->n[d], d, (int)sizeof(T), c->max_lds, &wbs, &wpx, &wtl, &wlds) &&
        (c->sweep_kind != IMH_SWEEP_AUTO_FAST || (a.npencil + wpx - 1) / wpx >= 4LL * c->num_cu)) {
        // the wave-spread sweep: its own tile in the argument block
        a.PX = wpx;
        a.TL = wtl;
        a.K = 4;
        a.xmajor = d == 0;
        const void* dw = imh_stage_args(c, &a, sizeof(a));
        if (!dw) return IMH_EHIP;
        const int nt = int((a.npencil + wpx - 1) / wpx);
        e = launch_fluid_wave((const FluidArgs<T>*)dw, nullptr, nt, nt, wbs,
                              c->sweep_kind != IMH_SWEEP_WAVE, d, c->mhd_kernel, wlds, c->st);
    } else if (t.column && !fluid_needs_general(a) && !a.dxl) {
        const int nt = int((a.npencil + fluid_column_pencils() - 1) / fluid_column_pencils());
        e = launch_fluid_column((const FluidArgs<T>*)da, nullptr, nt, nt, c->mhd_kernel, c->st);
    } else {
        const int nt = int((a.npencil + t.PX - 1) / t.PX);
        e = launch_fluid_batch((const FluidArgs<T>*)da, nullptr, nt, nt, t.bs, t.kmax,
                               c->mhd_kernel, fluid_needs_general(a), c->st);
    }
    if (e != hipSuccess) return hip_fail(e, "fluid sweep launch");
    c->fluid_init = true;
    return IMH_OK;
}

int imh_fluid_sweep_internal(imh_ctx* c, int X, int force) {
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    if (X < 1 || X > 3) return fail(IMH_EINVAL, "X must be 1, 2 or 3");
    if (c->n[X - 1] < 3)
        return fail(IMH_ENOTSUP, "sweep along a dimension with < 3 cells (flux.m:39 skips it)");
    return c->cfg.precision == IMH_FP64 ? fluid_sweep_t<double>(c, X - 1, force)
                                        : fluid_sweep_t<float>(c, X - 1, force);
}

template <typename T>
MagnetArgs<T> make_magnet_args(const imh_ctx* c, int dX, int dI, int force) {
    MagnetArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.rho = (const T*)c->f[IMH_F_MASS];
    a.mom = (const T*)c->f[IMH_F_MOMX + dX];
    a.bI = (T*)c->f[IMH_F_MAGX + dI];
    a.bX = (T*)c->f[IMH_F_MAGX + dX];
    a.flux = (T*)c->scratch;
    a.edges = c->edges;
    a.eoff = c->eoff_dev;
    a.sim = c->sim;
    a.n0 = c->n[0];
    a.n1 = c->n[1];
    a.n2 = c->n[2];
    a.dX = dX;
    a.dI = dI;
    a.ndims = c->ndims;
    a.modes = pack_modes(c->cfg.bc);
    a.active_linX = c->active_lin[dX];
    a.active_linI = c->active_lin[dI];
    a.limiter = c->cfg.limiter[dX];
    a.dxX = c->cfg.dgrid[dX];
    a.dxI = c->cfg.dgrid[dI];
    a.dxlX = c->dline_dev[dX];
    a.dxlI = c->dline_dev[dI];
    int pair = 3 * dI + dX;
    a.init_vg = !c->obj_init[OBJ_VG + pair];
    a.init_store = !c->obj_init[OBJ_MSFR + pair];
    a.init_wmag = !c->obj_init[OBJ_WMAG + pair];
    a.init_flux = !c->obj_init[OBJ_MFLUX + pair];
    a.PX = c->mtile[dX].PX;
    a.TL = c->mtile[dX].TL;
    a.npencil = c->nface[dX];
    a.bc_inf = c->cfg.bc_infinity;
    a.force = force;
    a.fade = c->fade_dev;
    return a;
}
template MagnetArgs<double> make_magnet_args<double>(const imh_ctx*, int, int, int);
template MagnetArgs<float> make_magnet_args<float>(const imh_ctx*, int, int, int);

template <typename T>
static int magnet_sweep_t(imh_ctx* c, int dX, const int I[2], int force) {
    int lastI = I[1] - 1;
    for (int t = 0; t < 2; ++t) {
        int dI = I[t] - 1;
        MagnetArgs<T> a = make_magnet_args<T>(c, dX, dI, force);
        const void* da = imh_stage_args(c, &a, sizeof(a));
        if (!da) return IMH_EHIP;
        const int N = c->n[dX];
        const int kmax = magnet_kmax_for(N, a.TL);
        const size_t smem = sizeof(T) * 2 * a.PX * (size_t)(N + 4);
        const int nt = int((a.npencil + a.PX - 1) / a.PX);
        if (kmax < 0) return fail(IMH_ENOTSUP, "magnet pencil of %d cells is too long", N);
        hipError_t e;
        {
            ProfScope ps(c, IMH_K_MAG_ADV, 5.0 * c->ncell * sizeof(T));
            e = launch_magnet_batch(0, (const MagnetArgs<T>*)da, nullptr, nt, nt, kmax, smem, c->st);
        }
        if (e != hipSuccess) return hip_fail(e, "magnet advect launch");
        {
            ProfScope ps(c, IMH_K_MAG_CT, 3.0 * c->ncell * sizeof(T));
            e = launch_magnet_batch(1, (const MagnetArgs<T>*)da, nullptr, nt, nt, kmax, smem / 2,
                                    c->st);
        }
        if (e != hipSuccess) return hip_fail(e, "magnet ct launch");
        int pair = 3 * dI + dX;
        c->obj_init[OBJ_VG + pair] = c->obj_init[OBJ_MSFR + pair] = true;
        c->obj_init[OBJ_WMAG + pair] = c->obj_init[OBJ_MFLUX + pair] = true;
    }
    // mag(I).updateCellCentered() for the last I only, then mag(X) (magnetFlux.m:30-31)
    int rc = launch_cellmag<T>(c, lastI, force);
    if (rc) return rc;
    return launch_cellmag<T>(c, dX, force);
}

int imh_magnet_sweep_internal(imh_ctx* c, int X, const int I[2], int force) {
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    if (!c->cfg.magnet_active || !c->scratch)
        return fail(IMH_ESTATE, "magnet sweep on a context created with magnet_active = 0");
    if (X < 1 || X > 3 || !I) return fail(IMH_EINVAL, "bad direction");
    for (int t = 0; t < 2; ++t)
        if (I[t] < 1 || I[t] > 3 || I[t] == X) return fail(IMH_EINVAL, "bad magnetic index");
    if (c->n[X - 1] < 3) return fail(IMH_ENOTSUP, "magnet sweep along a dimension with < 3 cells");
    return c->cfg.precision == IMH_FP64 ? magnet_sweep_t<double>(c, X - 1, I, force)
                                        : magnet_sweep_t<float>(c, X - 1, I, force);
}

template <typename T>
CflArgs<T> make_cfl_args(const imh_ctx* c) {
    CflArgs<T> a;
    for (int v = 0; v < 5; ++v) a.q[v] = (const T*)c->f[v];
    // components that pressure.m leaves out (MagnetArray.isZero) are not read at all
    for (int k = 0; k < 3; ++k)
        a.bc[k] = c->mag_is_zero[k] ? nullptr : (const T*)c->f[IMH_F_CMAGX + k];
    a.sim = c->sim;
    a.ncell = c->ncell;
    a.chunk = cfl_chunk();
    a.use = {!c->mag_is_zero[0], !c->mag_is_zero[1], !c->mag_is_zero[2]};
    a.gamma = c->cfg.gamma;
    return a;
}
template CflArgs<double> make_cfl_args<double>(const imh_ctx*);
template CflArgs<float> make_cfl_args<float>(const imh_ctx*);

template <typename T>
static int cfl_t(imh_ctx* c) {
    CflArgs<T> a = make_cfl_args<T>(c);
    // one simulation: at most 2048 tiles of a multiple of 256 cells
    a.chunk = ((c->ncell + 2048LL * 256 - 1) / (2048LL * 256)) * 256;
    const void* da = imh_stage_args(c, &a, sizeof(a));
    if (!da) return IMH_EHIP;
    const int nt = int((c->ncell + a.chunk - 1) / a.chunk);
    int nb = 5;
    for (int k = 0; k < 3; ++k) nb += a.bc[k] ? 1 : 0;
    ProfScope ps(c, IMH_K_CFL, double(nb) * c->ncell * sizeof(T));
    hipError_t e = launch_cfl_batch((const CflArgs<T>*)da, nullptr, nt, nt, c->st);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "cfl launch");
}

int imh_time_update_internal(imh_ctx* c, int step) {
    if (step > 1 && c->cfg.update_mode != 0) return IMH_OK;  // TimeManager.m:61-63
    hipError_t e;
    {
        ProfScope ps(c, IMH_K_BOOK, 0.0);
        e = launch_cfl_reset(c->sim_ptr, 1, c->st);
    }
    if (e != hipSuccess) return hip_fail(e, "cfl reset");
    int rc = c->cfg.precision == IMH_FP64 ? cfl_t<double>(c) : cfl_t<float>(c);
    if (rc) return rc;
    {
        ProfScope ps(c, IMH_K_BOOK, 0.0);
        e = launch_time_finalize(c->sim_ptr, c->par_ptr, 1, step, c->st);
    }
    return e == hipSuccess ? IMH_OK : hip_fail(e, "time finalize");
}

int imh_flux_internal(imh_ctx* c, int order, int force) {
    if (order != 1 && order != -1) return fail(IMH_EINVAL, "%d is not a recognized direction", order);
    int dirs[3], mi[3][2];
    sweep_order(c->host_iteration, order, c->cfg.perms_order, dirs, mi);
    for (int t = 0; t < 3; ++t) {
        int X = dirs[t];
        if (c->n[X - 1] < 3) continue;  // flux.m:39,53
        int rc;
        if (order > 0) {
            if (c->cfg.fluid_active && (rc = imh_fluid_sweep_internal(c, X, force))) return rc;
            if (c->cfg.magnet_active && (rc = imh_magnet_sweep_internal(c, X, mi[t], force)))
                return rc;
        } else {
            if (c->cfg.magnet_active && (rc = imh_magnet_sweep_internal(c, X, mi[t], force)))
                return rc;
            if (c->cfg.fluid_active && (rc = imh_fluid_sweep_internal(c, X, force))) return rc;
        }
    }
    return IMH_OK;
}

// ---- source terms (source.m:16-31) ----------------------------------------------------

static void drop_graphs(imh_ctx* c) {
    for (int i = 0; i < 6; ++i) {
        if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
        c->graph[i] = nullptr;
    }
}

hipError_t imh_launch_source(imh_ctx* c, hipStream_t st) {
    SourceArgs a;
    memset(&a, 0, sizeof(a));
    a.phi = c->grav;
    a.ghosts = c->grav_ghost;
    a.edges = c->edges;
    long long g = 0;
    for (int d = 0; d < 3; ++d) {
        a.n[d] = c->n[d];
        a.nface[d] = c->nface[d];
        a.dgrid[d] = c->cfg.dgrid[d];
        a.dxl[d] = c->dline_dev[d];
        a.active_lin[d] = c->active_lin[d];
        a.goff[d] = g;
        if (d < c->ndims) g += c->nface[d];
        for (int s = 0; s < 2; ++s) {
            const int m = c->cfg.bc[s][d];
            a.modes[s][d] = m;
            a.eoff[s][d] = (d < c->ndims && (m == IMH_BC_TRANS || m == IMH_BC_WALL))
                               ? c->eoff[OBJ_GRAV][s][d] : -1;
        }
    }
    a.ndims = c->ndims;
    a.bc_inf = c->cfg.bc_infinity;
    a.rad_type = c->rad_type;
    a.sim = c->sim;
    a.fade = c->fade_dev;
    a.mass_threshold = c->mass_threshold;
    a.gamma = c->cfg.gamma;
    a.rad_strength = c->rad_strength;
    a.rad_exponent = c->rad_exponent;
    a.use = {!c->mag_is_zero[0], !c->mag_is_zero[1], !c->mag_is_zero[2]};
    return launch_source(a, c->f, c->cfg.precision == IMH_FP64, st);
}

// Algorithmic HBM bytes of one source step: mass, momentum and energy read, the
// potential read with gravity, cell-centred B read with radiation; momentum and energy
// written with gravity, energy alone otherwise.
double source_bytes(const imh_ctx* c) {
    const double es = c->cfg.precision == IMH_FP64 ? 8.0 : 4.0;
    double per = 5 * es + (c->grav ? 8.0 + 4 * es : es) + (c->rad_type ? 3 * es : 0.0);
    return per * double(c->ncell);
}

int imh_source_internal(imh_ctx* c) {
    if (!c->grav && c->rad_type == 0) return IMH_OK;
    ProfScope ps(c, IMH_K_SOURCE, source_bytes(c));
    hipError_t e = imh_launch_source(c, c->st);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "source launch");
}

// One imogen.m:61-71 loop body; `conditional` = 0 runs it unconditionally (imh_iteration).
int imh_iteration_body(imh_ctx* c, int conditional) {
    int rc;
    hipError_t e;
    {
        ProfScope ps(c, IMH_K_BOOK, 0.0);
        e = launch_begin_iteration(c->sim_ptr, c->par_ptr, 1, conditional ? 0 : 1, c->st);
    }
    if (e != hipSuccess) return hip_fail(e, "begin iteration");
    for (int i = 1; i <= 2; ++i) {
        if ((rc = imh_time_update_internal(c, i))) return rc;
        if ((rc = imh_flux_internal(c, i == 1 ? 1 : -1, 0))) return rc;
        // treadmill: inactive; gravity.solvePotential: the fixed potential is set once;
        // source(): without gravity and radiation, ener - dt*0 changes no value
        if ((rc = imh_source_internal(c))) return rc;
    }
    {
        ProfScope ps(c, IMH_K_BOOK, 0.0);
        e = launch_end_iteration(c->sim_ptr, 1, c->st);
    }
    if (e != hipSuccess) return hip_fail(e, "end iteration");
    return IMH_OK;
}

extern "C" {

int imh_fluid_sweep(imh_ctx* c, int X) {
    if (!c) return fail(IMH_EINVAL, "null context");
    return imh_fluid_sweep_internal(c, X, 1);
}

int imh_magnet_sweep(imh_ctx* c, int X, const int I[2]) {
    if (!c) return fail(IMH_EINVAL, "null context");
    return imh_magnet_sweep_internal(c, X, I, 1);
}

int imh_time_update(imh_ctx* c, int step) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    // a direct call behaves as inside an unconditional loop body
    hipError_t e = launch_begin_iteration(c->sim_ptr, c->par_ptr, 1, 1, c->st);
    if (e != hipSuccess) return hip_fail(e, "begin");
    return imh_time_update_internal(c, step);
}

int imh_flux(imh_ctx* c, int order) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    return imh_flux_internal(c, order, 1);
}

int imh_source(imh_ctx* c) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    // a direct call behaves as inside an unconditional loop body
    hipError_t e = launch_begin_iteration(c->sim_ptr, c->par_ptr, 1, 1, c->st);
    if (e != hipSuccess) return hip_fail(e, "begin");
    return imh_source_internal(c);
}

int imh_set_gravity(imh_ctx* c, const double* potential, size_t n, double mass_threshold) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (potential && n != (size_t)c->ncell)
        return fail(IMH_EINVAL, "potential holds %zu values, the grid %lld", n, c->ncell);
    if (potential && c->squeezed)
        return fail(IMH_ENOTSUP, "gravity on a grid with a singleton x or y: the potential "
                    "keeps run.gridSize (GravityArray.m:34) while the fluid arrays are "
                    "squeezed, and the source terms mix the two shapes");
    if (potential) {
        for (int d = 0; d < c->ndims; ++d)
            for (int s = 0; s < 2; ++s) {
                const int m = c->cfg.bc[s][d];
                // the 5-point derivative shifts by two cells: on a one-cell dimension (a
                // 1-D grid's y) every shift rule of the reference indexes element 0 and
                // MATLAB errors; on two cells only circ / const / flip / zero stay inside
                if (c->n[d] == 1 || (c->n[d] == 2 && !(m == IMH_BC_CIRC || m == IMH_BC_CONST ||
                                                       m == IMH_BC_FLIP || m == IMH_BC_ZERO ||
                                                       m == IMH_BC_WORMHOLE)))
                    return fail(IMH_ENOTSUP, "gravity along dimension %d of %d cells: the "
                                "reference's shift by two cells indexes outside the array",
                                d + 1, c->n[d]);
            }
    }
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "sync");
    drop_graphs(c);
    c->mass_threshold = mass_threshold;
    if (!potential) {
        if (c->grav) hipFree(c->grav);
        if (c->grav_ghost) hipFree(c->grav_ghost);
        c->grav = c->grav_ghost = nullptr;
        return IMH_OK;
    }
    long long nghost = 0;
    for (int d = 0; d < c->ndims; ++d) nghost += c->nface[d];
    if (!c->grav && (e = hipMalloc(&c->grav, sizeof(double) * c->ncell)) != hipSuccess)
        return hip_fail(e, "hipMalloc(potential)");
    if (!c->grav_ghost &&
        (e = hipMalloc(&c->grav_ghost, sizeof(double) * 6 * (nghost > 0 ? nghost : 1))) != hipSuccess)
        return hip_fail(e, "hipMalloc(potential ghosts)");
    e = hipMemcpy(c->grav, potential, sizeof(double) * c->ncell, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "potential upload");
    // a new GravityArray: zeros are its first array, so its edge stores hold zeros
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < c->ndims; ++d) {
            const int m = c->cfg.bc[s][d];
            if (m != IMH_BC_TRANS && m != IMH_BC_WALL) continue;
            e = hipMemset(c->edges + c->eoff[OBJ_GRAV][s][d], 0, sizeof(double) * c->nface[d]);
            if (e != hipSuccess) return hip_fail(e, "potential edges");
        }
    return IMH_OK;
}

int imh_set_spacing(imh_ctx* c, int dim, const double* line, size_t n) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (dim < 1 || dim > 3) return fail(IMH_EINVAL, "dim must be 1, 2 or 3");
    const int d = dim - 1;
    if (line && c->squeezed)
        return fail(IMH_ENOTSUP, "non-uniform spacing on a grid with a singleton x or y: "
                    "DGRID keeps run.gridSize while the arrays are squeezed");
    if (line) {
        if (n != (size_t)c->n[d])
            return fail(IMH_EINVAL, "spacing line of %zu values along a dimension of %d cells",
                        n, c->n[d]);
        for (size_t i = 0; i < n; ++i)
            if (!(line[i] > 0)) return fail(IMH_EINVAL, "spacing must be > 0");
    }
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "sync");
    drop_graphs(c);
    if (!line) {
        if (c->dline_dev[d]) hipFree(c->dline_dev[d]);
        c->dline_dev[d] = nullptr;
        c->host_par.dgrid[d] = c->cfg.dgrid[d];
    } else {
        if (!c->dline_dev[d] &&
            (e = hipMalloc(&c->dline_dev[d], sizeof(double) * n)) != hipSuccess)
            return hip_fail(e, "hipMalloc(spacing)");
        e = hipMemcpy(c->dline_dev[d], line, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "spacing upload");
        double mn = line[0];
        for (size_t i = 0; i < n; ++i) mn = std::min(mn, line[i]);
        c->host_par.dgrid[d] = mn;  // MINDGRID(d) = minFinderND(DGRID{d})
    }
    e = hipMemcpy(c->par, &c->host_par, sizeof(SimParams), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "par upload");
    return IMH_OK;
}

int imh_set_radiation(imh_ctx* c, int type, double exponent, double initial_maximum) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (type != 0 && type != 1) return fail(IMH_EINVAL, "radiation type %d", type);
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "sync");
    drop_graphs(c);
    c->rad_type = type;
    c->rad_exponent = exponent;
    if (type == 0) {
        c->rad_strength = 0.0;
        return IMH_OK;
    }
    if (!c->have_state) return fail(IMH_ESTATE, "radiation needs the state first");
    // Radiation.initialize (Radiation.m:61-77), from the current state with strength 1
    const size_t n = (size_t)c->ncell;
    std::vector<double> q[9];
    const int fid[9] = {IMH_F_MASS, IMH_F_MOMX, IMH_F_MOMY, IMH_F_MOMZ, IMH_F_ENER,
                        IMH_F_MAGX, IMH_F_MAGY, IMH_F_MAGZ, IMH_F_CMAGX};
    for (int v = 0; v < 9; ++v) {
        q[v].resize(n);
        int rc = imh_download_field(c, fid[v], q[v].data(), n);
        if (rc) return rc;
    }
    std::vector<double> cm[2];
    for (int k = 0; k < 2; ++k) {
        cm[k].resize(n);
        int rc = imh_download_field(c, IMH_F_CMAGY + k, cm[k].data(), n);
        if (rc) return rc;
    }
    const MagUse u = {!c->mag_is_zero[0], !c->mag_is_zero[1], !c->mag_is_zero[2]};
    double mn = NAN;
    for (size_t i = 0; i < n; ++i) {
        const double rho = q[0][i];
        const double pg = pressure_gas(rho, q[1][i], q[2][i], q[3][i], q[4][i], q[8][i], cm[0][i],
                                       cm[1][i], u, c->cfg.gamma);
        const double unscaled = 1.0 * pow(rho, 2.0 - exponent) * pow(pg, exponent);
        const double ke = 0.5 * (q[1][i] * q[1][i] + q[2][i] * q[2][i] + q[3][i] * q[3][i]) / rho;
        const double me = 0.5 * (q[5][i] * q[5][i] + q[6][i] * q[6][i] + q[7][i] * q[7][i]);
        const double v = (q[4][i] - ke - me) / unscaled;
        if (!std::isnan(v) && (std::isnan(mn) || v < mn)) mn = v;  // MATLAB min skips NaN
    }
    c->rad_strength = (initial_maximum / 100.0) * mn;
    return IMH_OK;
}

int imh_set_viscosity(imh_ctx* c, int type, double linear, double quadratic) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (type < 0 || type > 2) return fail(IMH_EINVAL, "viscosity type %d", type);
    if (type == 2)
        for (int d = 0; d < c->ndims; ++d)
            for (int s = 0; s < 2; ++s) {
                if (c->cfg.bc[s][d] == IMH_BC_WORMHOLE)
                    return fail(IMH_ENOTSUP, "the Caramana-Shashkov-Whalen viscosity with a "
                                "wormhole boundary is not supported");
                if (c->cfg.bc[s][d] == IMH_BC_FADE && c->ndims == 3)
                    return fail(IMH_ENOTSUP, "the Caramana-Shashkov-Whalen viscosity at a fade "
                                "end: %s", LONG_3D_REFUSAL);
            }
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "sync");
    drop_graphs(c);
    // Caramana-Shashkov-Whalen with fade ends: every shifted array of the viscosity is
    // blended near both ends, which the multi-launch sweep evaluates cell by cell
    if (type == 2)
        for (int d = 0; d < c->ndims; ++d)
            if ((c->cfg.bc[0][d] == IMH_BC_FADE || c->cfg.bc[1][d] == IMH_BC_FADE) &&
                c->n[d] >= 3 && (e = ensure_long_scratch(c, d)) != hipSuccess)
                return hip_fail(e, "hipMalloc(long sweep)");
    c->visc_type = type;
    c->visc_lin = linear;
    c->visc_quad = quadratic;
    return IMH_OK;
}

int imh_get_radiation_strength(imh_ctx* c, double* strength) {
    if (!c || !strength) return fail(IMH_EINVAL, "bad argument");
    *strength = c->rad_strength;
    return IMH_OK;
}

int imh_iteration(imh_ctx* c) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    int rc = imh_iteration_body(c, 0);
    if (rc) return rc;
    c->host_iteration += 1;
    return IMH_OK;
}

int imh_run(imh_ctx* c, int max_iterations, int* done) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    int rc;
    SimDev sd;
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(&sd, c->sim, sizeof(sd), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "time state");
    c->host_iteration = sd.iteration;
    const int start = sd.iteration;
    int todo = std::min(max_iterations, c->cfg.iter_max - start + 1);
    for (int it = 0; it < todo; ++it) {
        bool inited = c->fluid_init || !c->cfg.fluid_active;
        if (!inited || c->no_graphs) {
            if ((rc = imh_iteration_body(c, 1))) return rc;
        } else {
            int row = ((c->host_iteration - 1) % 6 + 6) % 6;
            if (!c->graph[row]) {
                if (!c->cap_dev[row]) {
                    e = hipMalloc(&c->cap_dev[row], ARG_SLOT * CAP_SLOTS);
                    if (e != hipSuccess) return hip_fail(e, "hipMalloc(graph args)");
                }
                hipGraph_t g;
                e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return hip_fail(e, "begin capture");
                c->capture_row = row;
                c->cap_count = 0;
                rc = imh_iteration_body(c, 1);
                c->capture_row = -1;
                e = hipStreamEndCapture(c->st, &g);
                if (rc) return rc;
                if (e != hipSuccess) return hip_fail(e, "end capture");
                // the graph's kernels read their argument blocks from cap_dev[row]
                e = hipMemcpy(c->cap_dev[row], c->cap_host, ARG_SLOT * c->cap_count,
                              hipMemcpyHostToDevice);
                if (e == hipSuccess) e = hipGraphInstantiate(&c->graph[row], g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (e != hipSuccess) return hip_fail(e, "graph instantiate");
            }
            e = hipGraphLaunch(c->graph[row], c->st);
            if (e != hipSuccess) return hip_fail(e, "graph launch");
        }
        c->host_iteration += 1;
        if ((it & 63) == 63) {
            // stop early once the run has ended on the device
            e = hipMemcpyAsync(&sd, c->sim, sizeof(sd), hipMemcpyDeviceToHost, c->st);
            if (e == hipSuccess) e = hipStreamSynchronize(c->st);
            if (e != hipSuccess) return hip_fail(e, "time state");
            if (sd.error || sd.iteration > c->cfg.iter_max || sd.time >= c->cfg.time_max) break;
        }
    }
    e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(&sd, c->sim, sizeof(sd), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "time state");
    c->host_iteration = sd.iteration;
    if (done) *done = sd.iteration - start;
    if (sd.error == IMH_ECFL)
        return fail(IMH_ECFL, "CFL error: no signal speed above 1e-2 (TimeManager.m:82-85)");
    return IMH_OK;
}

int imh_set_dtime(imh_ctx* c, double dt) {
    if (!c) return fail(IMH_EINVAL, "null context");
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess)
        e = hipMemcpy((char*)c->sim + offsetof(SimDev, dt), &dt, sizeof(double), hipMemcpyHostToDevice);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "set dt");
}

int imh_set_iteration(imh_ctx* c, int iteration) {
    if (!c) return fail(IMH_EINVAL, "null context");
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess)
        e = hipMemcpy((char*)c->sim + offsetof(SimDev, iteration), &iteration, sizeof(int),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "set iteration");
    c->host_iteration = iteration;
    return IMH_OK;
}

int imh_get_time(imh_ctx* c, double* time, double* dtime, int* iteration, int* grid_index) {
    if (!c) return fail(IMH_EINVAL, "null context");
    SimDev sd;
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(&sd, c->sim, sizeof(sd), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "time state");
    if (time) *time = sd.time;
    if (dtime) *dtime = sd.dt;
    if (iteration) *iteration = sd.iteration;
    if (grid_index) *grid_index = sd.grid_index;
    if (sd.error == IMH_ECFL)
        return fail(IMH_ECFL, "CFL error: no signal speed above 1e-2 (TimeManager.m:82-85)");
    return IMH_OK;
}

int imh_synchronize(imh_ctx* c) {
    if (!c) return fail(IMH_EINVAL, "null context");
    hipError_t e = hipStreamSynchronize(c->st);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "synchronize");
}

// ---- initial conditions on the device ---------------------------------------------------

extern "C" int imh_init_problem(imh_ctx* c, int kind, const double* params, int nparams) {
    if (!c) return fail(IMH_EINVAL, "null context");
    static const int need[9] = {0, 0, 0, 3, 15, 4, 0, 0, 4};
    if (kind < IMH_IC_SOD || kind > IMH_IC_MAGNETIC_SHOCK_TUBE)
        return fail(IMH_EINVAL, "unknown initial-condition kind %d", kind);
    if (c->squeezed)
        return fail(IMH_ENOTSUP, "initial conditions on the device are built on the grid as "
                    "given; this context's grid was squeezed (singleton x or y)");
    if (nparams < need[kind] || (need[kind] > 0 && !params))
        return fail(IMH_EINVAL, "initial condition %d takes %d parameters", kind, need[kind]);
    IcParams P;
    memset(&P, 0, sizeof(P));
    P.kind = kind;
    for (int d = 0; d < 3; ++d) {
        P.n[d] = c->n[d];
        P.dx[d] = c->cfg.dgrid[d];
    }
    P.gamma = c->cfg.gamma;
    for (int i = 0; i < need[kind] && i < 24; ++i) P.p[i] = params[i];
    bool anyB = false;
    const double g = c->cfg.gamma;
    switch (kind) {
    case IMH_IC_SOD:
        P.ip[0] = int(std::floor(c->n[0] / 2.0 + 0.5));  // MATLAB round(N/2)
        break;
    case IMH_IC_ORSZAG_TANG:
        anyB = true;
        break;
    case IMH_IC_SEDOV: {
        // cells inside the blast radius, with the kernel's expression (a small box around
        // the centre holds them all)
        const int n = c->n[0];
        const double dx = c->cfg.dgrid[0], rr = params[0] * dx;
        if (c->n[1] != n || c->n[2] != n) return fail(IMH_EINVAL, "Sedov needs a cubic grid");
        long long inside = 0;
        const int h = int(std::ceil(params[0])) + 2;
        for (int k = n / 2 - h; k <= n / 2 + h; ++k)
            for (int j = n / 2 - h; j <= n / 2 + h; ++j)
                for (int i = n / 2 - h; i <= n / 2 + h; ++i) {
                    if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) continue;
                    const double cx = (double(i) + 0.5) * dx - 0.5, cy = (double(j) + 0.5) * dx - 0.5,
                                 cz = (double(k) + 0.5) * dx - 0.5;
                    if (cz * cz + cy * cy + cx * cx <= rr * rr) ++inside;
                }
        if (inside == 0) return fail(IMH_EINVAL, "Sedov blast radius holds no cell");
        P.ip[0] = int(inside);
        break;
    }
    case IMH_IC_RIEMANN:
        // the kernel reads the left state at p[0..6], the right at p[8..14], Bx at p[16]
        for (int i = 0; i < 7; ++i) {
            P.p[i] = params[i];
            P.p[8 + i] = params[7 + i];
        }
        P.p[16] = params[14];
        P.ip[0] = (nparams > 15 && params[15] >= 0) ? int(params[15]) : c->n[0] / 2;
        for (int i : {5, 6, 13, 14, 16}) anyB = anyB || P.p[i] != 0.0;
        break;
    case IMH_IC_ORSZAG_TANG_REF:
        if (c->n[0] < 2 || c->n[1] < 2)
            return fail(IMH_EINVAL, "the reference Orszag-Tang grid needs 2+ cells in x and y "
                        "(its sines divide by grid - 1)");
        P.ip[0] = c->n[0] - 1;
        P.ip[1] = c->n[1] - 1;
        anyB = true;
        break;
    case IMH_IC_SEDOV_REF: {
        // ndgrid(-half:half) has 2 * half + 1 points: the reference needs odd counts
        for (int d = 0; d < 3; ++d) {
            if (c->n[d] > 1 && c->n[d] % 2 == 0)
                return fail(IMH_EINVAL, "the reference Sedov-Taylor initializer needs an odd "
                            "cell count along every dimension (dim %d has %d): its "
                            "ndgrid(-half:half) has 2*floor(n/2)+1 points", d + 1, c->n[d]);
            P.ip[d] = c->n[d] / 2;
        }
        const int nmin = std::min(c->n[0], std::min(c->n[1], c->n[2]));
        const int radius = std::max(5, int(std::floor(0.06 * nmin + 0.5)));  // MATLAB round
        const double dx = 1.0 / double(c->n[0]);                               // obj.dGrid(1)
        P.ip[3] = radius;
        P.p[0] = 3.0 * (g - 1.0) / (4.0 * M_PI * std::pow(dx * radius, 3.0));
        break;
    }
    case IMH_IC_MAGNETIC_SHOCK_TUBE: {
        const int dir = int(params[0]);
        if (dir < 0 || dir > 2) return fail(IMH_EINVAL, "shock direction must be 0, 1 or 2");
        P.ip[0] = dir;
        P.ip[1] = int(std::floor(c->n[dir] / 2.0 + 0.5));  // MATLAB round(grid / 2)
        for (int q = 0; q < 3; ++q) {
            P.ip[2 + q] = params[1 + q] != 0.0;
            anyB = anyB || P.ip[2 + q];
        }
        break;
    }
    case IMH_IC_KELVIN_HELMHOLTZ: {
        // external/speedFromMach.m with rho = 1, E = 1 / (gamma - 1), B = 0
        const double mach = params[0];
        const double ener = 1.0 / (g - 1.0);
        P.p[0] = mach * std::sqrt((g * ener) / (1.0 / (g - 1.0) + g * mach * mach / 2.0));
        P.p[1] = params[1];
        P.p[2] = params[2];
        P.seed = (unsigned long long)(long long)params[3];
        P.ip[0] = int(std::ceil(c->n[1] / 2.0));
        break;
    }
    }
    hipError_t e = hipStreamSynchronize(c->st);
    void* const f[8] = {c->f[0], c->f[1], c->f[2], c->f[3], c->f[4], c->f[5], c->f[6], c->f[7]};
    if (e == hipSuccess) e = launch_ic_fill(P, f, c->cfg.precision == IMH_FP64, c->st);
    // edge stores of the FluidArray and MagnetArray objects come from the initial state
    if (e == hipSuccess && c->edges_len)
        e = hipMemsetAsync(c->edges, 0, sizeof(double) * c->edges_len, c->st);
    for (int fld = 0; fld < 8 && e == hipSuccess; ++fld) {
        const int obj = fld < 5 ? OBJ_PRIM + fld : OBJ_MAG + (fld - 5);
        for (int s = 0; s < 2 && e == hipSuccess; ++s)
            for (int d = 0; d < c->ndims && e == hipSuccess; ++d) {
                const int m = c->cfg.bc[s][d];
                if (m != IMH_BC_TRANS && m != IMH_BC_WALL) continue;
                e = launch_ic_edges(P, fld, s, d, c->edges, c->eoff[obj][s][d], c->nface[d], c->st);
            }
    }
    if (e != hipSuccess) return hip_fail(e, "initial condition");
    // MagnetArray.isZero = (sumND(array) == 0) && ~run.magnet.ACTIVE: the B of these
    // problems is zero exactly when none of its parameters is nonzero
    for (int k = 0; k < 3; ++k) c->mag_is_zero[k] = !anyB && !c->cfg.magnet_active;
    c->mhd_kernel = c->cfg.magnet_active || anyB;
    SimDev sd;  // time state of a new run (imogen.m:50)
    memset(&sd, 0, sizeof(sd));
    sd.iteration = 1;
    sd.active = 1;
    e = hipMemcpyAsync(c->sim, &sd, sizeof(sd), hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "sim reset");
    c->host_iteration = 1;
    c->fluid_init = false;
    for (int o = 0; o < NOBJ; ++o) c->obj_init[o] = false;
    for (int i = 0; i < 6; ++i) {
        if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
        c->graph[i] = nullptr;
    }
    for (int k = 0; k < 3; ++k) {  // MagnetArray.m:102-103
        int rc = cellmag(c, k, 1);
        if (rc) return rc;
    }
    if (int rc = cellmag_edges(c)) return rc;
    e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "cellMag");
    c->have_state = true;
    return IMH_OK;
}

// ---- checkpoint / resume --------------------------------------------------------------

namespace {
struct CkptHeader {
    char magic[8];            // "IMHCKPT\0"
    int version;              // 2
    int precision;
    int grid[3];
    int nfields;
    int nobj;
    long long ncell;
    long long edges_len;
    int host_iteration;
    int fluid_init;
    int have_state;
    int mhd_kernel;
    int mag_is_zero[3];
    int pad;
    unsigned long long cfg_hash;  // run settings that shape the evolution (ckpt_cfg_hash)
};
const char kCkptMagic[8] = {'I', 'M', 'H', 'C', 'K', 'P', 'T', 0};
constexpr int kCkptVersion = 2;

// FNV-1a over the settings a resumed run must share with the saved one to continue bit
// for bit: boundary modes, limiters, gamma, spacing, bcInfinity, MINMASS, CFL, the
// fluid / magnet flags, the time-update mode and the perms order (the iteration and time
// limits may change between the save and the resume).
unsigned long long ckpt_cfg_hash(const imh_config& c) {
    unsigned long long h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    };
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < 3; ++d) mix(&c.bc[s][d], sizeof(int));
    for (int d = 0; d < 3; ++d) mix(&c.limiter[d], sizeof(int));
    for (int d = 0; d < 3; ++d) mix(&c.dgrid[d], sizeof(double));
    mix(&c.gamma, sizeof(double));
    mix(&c.minmass, sizeof(double));
    mix(&c.cfl, sizeof(double));
    mix(&c.bc_infinity, sizeof(int));
    mix(&c.fluid_active, sizeof(int));
    mix(&c.magnet_active, sizeof(int));
    mix(&c.update_mode, sizeof(int));
    mix(&c.perms_order, sizeof(int));
    return h;
}

size_t ckpt_bytes(const imh_ctx* c) {
    return sizeof(CkptHeader) + sizeof(SimDev) + NOBJ + (size_t)IMH_NFIELDS * c->ncell * esize(c) +
           c->edges_len * sizeof(double);
}
}  // namespace

extern "C" int imh_checkpoint_bytes(imh_ctx* c, size_t* bytes) {
    if (!c || !bytes) return fail(IMH_EINVAL, "null argument");
    *bytes = ckpt_bytes(c);
    return IMH_OK;
}

extern "C" int imh_checkpoint_save(imh_ctx* c, void* buf, size_t bytes) {
    if (!c || !buf) return fail(IMH_EINVAL, "null argument");
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    if (bytes < ckpt_bytes(c)) return fail(IMH_EINVAL, "checkpoint buffer too small");
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "stream sync");
    char* p = (char*)buf;
    CkptHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.precision = c->cfg.precision;
    h.cfg_hash = ckpt_cfg_hash(c->cfg);
    for (int d = 0; d < 3; ++d) h.grid[d] = c->n[d];
    h.nfields = IMH_NFIELDS;
    h.nobj = NOBJ;
    h.ncell = c->ncell;
    h.edges_len = (long long)c->edges_len;
    h.host_iteration = c->host_iteration;
    h.fluid_init = c->fluid_init;
    h.have_state = c->have_state;
    h.mhd_kernel = c->mhd_kernel;
    for (int k = 0; k < 3; ++k) h.mag_is_zero[k] = c->mag_is_zero[k];
    memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    e = hipMemcpy(p, c->sim, sizeof(SimDev), hipMemcpyDeviceToHost);
    p += sizeof(SimDev);
    for (int o = 0; o < NOBJ; ++o) *p++ = c->obj_init[o] ? 1 : 0;
    const size_t fb = (size_t)c->ncell * esize(c);
    for (int f = 0; f < IMH_NFIELDS && e == hipSuccess; ++f, p += fb)
        e = hipMemcpy(p, c->f[f], fb, hipMemcpyDeviceToHost);
    if (e == hipSuccess && c->edges_len)
        e = hipMemcpy(p, c->edges, c->edges_len * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "checkpoint save");
}

extern "C" int imh_checkpoint_load(imh_ctx* c, const void* buf, size_t bytes) {
    if (!c || !buf) return fail(IMH_EINVAL, "null argument");
    if (bytes < sizeof(CkptHeader)) return fail(IMH_EINVAL, "checkpoint too short");
    const char* p = (const char*)buf;
    CkptHeader h;
    memcpy(&h, p, sizeof(h));
    if (memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != kCkptVersion)
        return fail(IMH_EINVAL, "not a checkpoint of this library");
    if (h.precision != c->cfg.precision || h.grid[0] != c->n[0] || h.grid[1] != c->n[1] ||
        h.grid[2] != c->n[2] || h.nfields != IMH_NFIELDS || h.nobj != NOBJ ||
        h.ncell != c->ncell || h.edges_len != (long long)c->edges_len)
        return fail(IMH_EINVAL, "checkpoint of another configuration (precision, grid or "
                    "boundary layout differ)");
    if (h.cfg_hash != ckpt_cfg_hash(c->cfg))
        return fail(IMH_EINVAL, "checkpoint of another configuration (boundary modes, "
                    "limiters, gamma, spacing, bcInfinity, MINMASS, CFL, fluid / magnet flags, "
                    "update mode or perms order differ)");
    if (bytes < ckpt_bytes(c)) return fail(IMH_EINVAL, "checkpoint truncated");
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "stream sync");
    p += sizeof(h);
    e = hipMemcpy(c->sim, p, sizeof(SimDev), hipMemcpyHostToDevice);
    p += sizeof(SimDev);
    for (int o = 0; o < NOBJ; ++o) c->obj_init[o] = *p++ != 0;
    const size_t fb = (size_t)c->ncell * esize(c);
    for (int f = 0; f < IMH_NFIELDS && e == hipSuccess; ++f, p += fb)
        e = hipMemcpy(c->f[f], p, fb, hipMemcpyHostToDevice);
    if (e == hipSuccess && c->edges_len)
        e = hipMemcpy(c->edges, p, c->edges_len * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "checkpoint load");
    c->host_iteration = h.host_iteration;
    c->fluid_init = h.fluid_init != 0;
    c->have_state = h.have_state != 0;
    c->mhd_kernel = h.mhd_kernel != 0;
    for (int k = 0; k < 3; ++k) c->mag_is_zero[k] = h.mag_is_zero[k];
    for (int i = 0; i < 6; ++i) {  // captured iterations hold argument blocks of the old state
        if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
        c->graph[i] = nullptr;
    }
    return IMH_OK;
}

int imh_conservation_sums(imh_ctx* c, double out[8]) {
    if (!c || !out) return fail(IMH_EINVAL, "null argument");
    hipError_t e;
    if (c->cfg.precision == IMH_FP64) {
        const double* f[8];
        for (int i = 0; i < 8; ++i) f[i] = (const double*)c->f[i];
        e = launch_conservation<double>(f, 8, c->ncell, c->csum_part, 1024, c->csum_out, c->st);
    } else {
        const float* f[8];
        for (int i = 0; i < 8; ++i) f[i] = (const float*)c->f[i];
        e = launch_conservation<float>(f, 8, c->ncell, c->csum_part, 1024, c->csum_out, c->st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(out, c->csum_out, sizeof(double) * 8, hipMemcpyDeviceToHost);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "conservation sums");
}

int imh_memory_bytes(imh_ctx* c, size_t* bytes) {
    if (!c || !bytes) return fail(IMH_EINVAL, "null argument");
    *bytes = ctx_bytes(c->cfg, c->edges_len);
    return IMH_OK;
}

int imh_memory_estimate(const imh_config* cfg, size_t* bytes) {
    if (!cfg || !bytes) return fail(IMH_EINVAL, "null argument");
    bool squeezed = false;
    const imh_config sq = squeeze_grid(*cfg, &squeezed);
    int rc = validate_config(&sq);
    if (rc) return rc;
    long long eoff[NOBJ][2][3];
    *bytes = ctx_bytes(sq, (size_t)edges_layout(sq, eoff));
    return IMH_OK;
}

// Tuning hook (used by the autotuner): override the fluid tile of one sweep dim.
int imh_set_fluid_tile(imh_ctx* c, int X, int bs, int kmax, int px) {
    if (!c || X < 1 || X > 3) return fail(IMH_EINVAL, "bad argument");
    if (bs == 0) {  // the column march
        if (!column_possible(&c->cfg, X - 1))
            return fail(IMH_EINVAL, "the column march needs circular or constant boundaries "
                        "along y or z");
        c->ftile[X - 1].column = 1;
        for (int i = 0; i < 6; ++i) {
            if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
            c->graph[i] = nullptr;
        }
        return IMH_OK;
    }
    TileOverride ov = {1, bs, kmax, px};
    FluidTile t = choose_fluid_tile(c->n[X - 1], X - 1, (int)esize(c), &ov);
    if (t.K > t.kmax || !fluid_variant_exists((int)esize(c), t.bs, t.kmax) || (t.bs % t.PX) != 0 ||
        t.PX > 32 || (t.PX & (t.PX - 1)) != 0 ||
        (c->cfg.bc[0][X - 1] == IMH_BC_WORMHOLE && (t.PX & 1)))
        return fail(IMH_EINVAL, "tile (bs=%d kmax=%d px=%d) does not fit pencils of %d cells",
                    bs, kmax, px, c->n[X - 1]);
    c->ftile[X - 1] = t;
    for (int i = 0; i < 6; ++i) {
        if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
        c->graph[i] = nullptr;
    }
    return IMH_OK;
}

int imh_set_fluid_kernel(imh_ctx* c, int kind) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (kind != IMH_SWEEP_GENERAL && kind != IMH_SWEEP_WAVE && kind != IMH_SWEEP_WAVE_FAST &&
        kind != IMH_SWEEP_AUTO_FAST)
        return fail(IMH_EINVAL, "unknown sweep kind %d", kind);
    if ((kind == IMH_SWEEP_WAVE_FAST || kind == IMH_SWEEP_AUTO_FAST) &&
        c->cfg.precision != IMH_FP64)
        return fail(IMH_EINVAL, "the fast FP64 sweep needs an FP64 context");
    c->sweep_kind = kind;
    for (int i = 0; i < 6; ++i) {
        if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
        c->graph[i] = nullptr;
    }
    return IMH_OK;
}

int imh_get_fluid_kernel(imh_ctx* c, int* kind) {
    if (!c || !kind) return fail(IMH_EINVAL, "null argument");
    *kind = c->sweep_kind;
    return IMH_OK;
}

int imh_get_fluid_tile(imh_ctx* c, int X, int* bs, int* kmax, int* px) {
    if (!c || X < 1 || X > 3) return fail(IMH_EINVAL, "bad argument");
    *bs = c->ftile[X - 1].column ? 0 : c->ftile[X - 1].bs;  // 0: the column march
    *kmax = c->ftile[X - 1].kmax;
    *px = c->ftile[X - 1].PX;
    return IMH_OK;
}

int imh_set_graphs(imh_ctx* c, int enable) {
    if (!c) return fail(IMH_EINVAL, "null context");
    c->no_graphs = !enable;
    return IMH_OK;
}

int imh_profile_iterations(imh_ctx* c, int iterations, double* ms, double* bytes, int* count) {
    if (!c) return fail(IMH_EINVAL, "null context");
    if (!c->have_state) return fail(IMH_ESTATE, "state not uploaded");
    SimDev sd;
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(&sd, c->sim, sizeof(sd), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "time state");
    c->host_iteration = sd.iteration;
    Profiler p;
    c->prof = &p;
    int rc = IMH_OK;
    for (int it = 0; it < iterations && rc == IMH_OK; ++it) {
        rc = imh_iteration_body(c, 1);
        c->host_iteration += 1;
    }
    c->prof = nullptr;
    int rc2 = prof_collect(&p, ms, bytes, count);
    return rc ? rc : rc2;
}

hipStream_t imh_stream(imh_ctx* c) { return c ? c->st : nullptr; }

}  // extern "C"
