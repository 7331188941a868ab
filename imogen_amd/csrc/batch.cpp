// batch.cpp -- many independent simulations advanced together (the batched ensemble
// and the serving loop's step).
//
// A batch is a view over simulation contexts (imh_ctx) sharing one HIP stream.  One
// iteration of every simulation is a fixed sequence of launches:
//
//   begin  -> CFL reset / CFL reduction (tiles = (sim, cell chunk)) / dt finalize
//   for order in (+1, -1), for sweep slot t = 0..2, op o = 0..1:
//       fluid sweeps      one persistent launch per (block size, KMAX, MHD) group over the
//                         pencil tiles of every simulation whose op (t, o) is a fluid sweep
//       magnet sweeps     advection and CT launches per component, then cell-centred B
//   end    -> iteration counters
//
// Every simulation keeps its own sweep order (flux.m: the perms row follows its own
// iteration number), so the launch plan repeats with period 6 in the batch step; the six
// plans are captured once as hipGraphs and replayed.  Simulations that have ended
// (TimeManager.running false) turn into no-ops inside the kernels.  The first iteration of
// a simulation (which creates the edge stores of its flux objects, InitializedArray.m:43-47)
// runs on its own before it joins the replayed plans.
//
// Lanes (batches from imh_batch_create; not the serving loop's short-lived views):
// simulations with a magnetic field and pure-hydro ones share no data, so their
// sweeps run on two streams, forked from the batch stream after the time step is known
// and joined before the next batch-wide kernel (captured into the same graph as parallel
// branches).  The 1-D MHD sweeps of an ensemble are latency-bound and leave most of the
// device idle; the hydro sweeps fill it meanwhile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "imh_common.h"
#include "imh_host.h"
#include "imh_kernels.h"

using namespace imh;

namespace {

enum LaunchKind { L_BEGIN, L_CFL_RESET, L_CFL, L_FINALIZE, L_FLUID, L_MAG_ADV, L_MAG_CT, L_CELLMAG, L_END,
                  L_FORK, L_JOIN, L_SOURCE, L_LANE1D };

constexpr int MAX_LANES = 2;

struct Launch {
    LaunchKind kind;
    int step;       // finalize: 1 or 2; source: the simulation
    int bs, kmax;   // fluid / magnet variant
    bool mhd;
    bool general;   // fluid: the kernel form with fade boundaries / first assignments
    size_t smem;
    long long off;  // into the tile array
    int ntiles;
    double bytes;   // algorithmic HBM bytes of the launch
    int lane;       // -1: the batch stream; else the lane's stream
    bool reset;     // begin: also clear the CFL slots (the reduction follows)
};

int launch_kind_class(LaunchKind k) {
    switch (k) {
    case L_CFL: return IMH_K_CFL;
    case L_FLUID: return IMH_K_FLUID;
    case L_MAG_ADV: return IMH_K_MAG_ADV;
    case L_MAG_CT: return IMH_K_MAG_CT;
    case L_CELLMAG: return IMH_K_CELLMAG;
    case L_SOURCE: return IMH_K_SOURCE;
    case L_LANE1D: return IMH_K_LANE1D;
    default: return IMH_K_BOOK;
    }
}

}  // namespace

struct imh_batch {
    int n = 0;
    int prec = IMH_FP64;
    int device = 0;
    bool owns = false;
    hipStream_t st = nullptr;
    std::vector<imh_ctx*> ctx;
    int ncu = 256;
    SimDev** sims_dev = nullptr;
    SimParams** pars_dev = nullptr;
    void* fargs = nullptr;    // FluidArgs<T>[n*3]
    void* margs = nullptr;    // MagnetArgs<T>[n*9], index sim*9 + 3*dI + dX
    void* cargs = nullptr;    // CellMagArgs<T>[n*3]
    void* cflargs = nullptr;  // CflArgs<T>[n]
    BatchTile* tiles = nullptr;
    size_t tiles_cap = 0;
    std::vector<Launch> plan[6];
    hipGraphExec_t graph[6] = {};
    int uses[6] = {};                  // eager runs of each plan since it was built
    bool planned = false;
    long long step = 0;
    bool any_step_mode = false;
    Profiler* prof = nullptr;
    bool lanes_ok = true;              // views of the serving loop run on one stream
    int nlanes = 1;                    // lanes used by the plans (1: no fork)
    hipStream_t lane_st[MAX_LANES] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[MAX_LANES] = {};
    bool fuse_1d = true;               // one launch per iteration for the 1-D MHD lane
    std::vector<char> fused;           // per simulation: runs in that launch
    std::vector<int> fused_list;       // those simulations, longest first
    void* l1args = nullptr;            // Lane1dArgs<float>[fused_list.size()]
    size_t l1cap = 0;
};

// A simulation whose whole flux call runs in the 1-D lane launch (lane1d.hip): FP32
// bulk, one pencil (N x 1 x 1, 3 <= N <= 4096), fluid and magnet both active, no source
// terms, viscosity, fade ends or spacing lines (those take the general sweep), the time
// step once per iteration in every simulation of the batch.
static bool lane1d_ok(const imh_batch* b, const imh_ctx* c) {
    if (!b->fuse_1d || b->prec != IMH_FP32_BULK || b->any_step_mode) return false;
    if (c->n[1] != 1 || c->n[2] != 1 || c->n[0] < 3 || c->n[0] > LANE1D_NMAX) return false;
    if (!c->cfg.fluid_active || !c->cfg.magnet_active || !c->mhd_kernel) return false;
    if (c->grav || c->rad_type || c->visc_type || c->dline_dev[0]) return false;
    for (int s = 0; s < 2; ++s)
        if (c->cfg.bc[s][0] == IMH_BC_FADE) return false;
    return true;
}

// Wait until nothing of the batch is in flight: the batch stream (graph replays join
// their lane branches back into it) and the lane streams.  Every call that changes a
// simulation's settings or state, or drops the captured graphs, waits here first.
static int batch_quiesce(imh_batch* b) {
    hipError_t e = b->st ? hipStreamSynchronize(b->st) : hipSuccess;
    for (int l = 0; l < MAX_LANES && e == hipSuccess; ++l)
        if (b->lane_st[l]) e = hipStreamSynchronize(b->lane_st[l]);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "batch sync");
}

static void free_plans(imh_batch* b) {
    for (int p = 0; p < 6; ++p) {
        if (b->graph[p]) hipGraphExecDestroy(b->graph[p]);
        b->graph[p] = nullptr;
        b->plan[p].clear();
        b->uses[p] = 0;
    }
    b->planned = false;
}

static void batch_free(imh_batch* b) {
    if (!b) return;
    if (b->st) hipStreamSynchronize(b->st);
    free_plans(b);
    hipFree(b->sims_dev);
    hipFree(b->pars_dev);
    hipFree(b->fargs);
    hipFree(b->margs);
    hipFree(b->cargs);
    hipFree(b->cflargs);
    hipFree(b->tiles);
    hipFree(b->l1args);
    if (b->owns)
        for (imh_ctx* c : b->ctx) imh_ctx_free(c);
    else
        for (imh_ctx* c : b->ctx) (void)c;
    for (int l = 0; l < MAX_LANES; ++l) {
        if (b->lane_st[l]) {
            hipStreamSynchronize(b->lane_st[l]);
            hipStreamDestroy(b->lane_st[l]);
        }
        if (b->join_ev[l]) hipEventDestroy(b->join_ev[l]);
    }
    if (b->fork_ev) hipEventDestroy(b->fork_ev);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

template <typename T>
static int upload_args_t(imh_batch* b) {
    const int n = b->n;
    std::vector<FluidArgs<T>> fa(n * 3);
    std::vector<MagnetArgs<T>> ma(n * 9);
    std::vector<CellMagArgs<T>> ca(n * 3);
    std::vector<CflArgs<T>> cf(n);
    memset(ma.data(), 0, sizeof(MagnetArgs<T>) * ma.size());
    for (int i = 0; i < n; ++i) {
        const imh_ctx* c = b->ctx[i];
        for (int d = 0; d < 3; ++d) {
            fa[i * 3 + d] = make_fluid_args<T>(c, d, 0);
            fa[i * 3 + d].xcd_remap = 0;
            ca[i * 3 + d] = make_cellmag_args<T>(c, d, 0);
        }
        if (c->cfg.magnet_active)
            for (int dI = 0; dI < 3; ++dI)
                for (int dX = 0; dX < 3; ++dX)
                    if (dI != dX) ma[i * 9 + 3 * dI + dX] = make_magnet_args<T>(c, dX, dI, 0);
        cf[i] = make_cfl_args<T>(c);
    }
    hipError_t e = hipMemcpyAsync(b->fargs, fa.data(), sizeof(FluidArgs<T>) * fa.size(),
                                  hipMemcpyHostToDevice, b->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->margs, ma.data(), sizeof(MagnetArgs<T>) * ma.size(),
                           hipMemcpyHostToDevice, b->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->cargs, ca.data(), sizeof(CellMagArgs<T>) * ca.size(),
                           hipMemcpyHostToDevice, b->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->cflargs, cf.data(), sizeof(CflArgs<T>) * cf.size(),
                           hipMemcpyHostToDevice, b->st);
    // the 1-D lane: argument blocks with one pencil over the whole workgroup
    b->fused.assign(n, 0);
    b->fused_list.clear();
    if constexpr (std::is_same<T, float>::value) {
        for (int i = 0; i < n; ++i)
            if (lane1d_ok(b, b->ctx[i])) {
                b->fused[i] = 1;
                b->fused_list.push_back(i);
            }
        std::stable_sort(b->fused_list.begin(), b->fused_list.end(), [&](int x, int y) {
            return b->ctx[x]->n[0] > b->ctx[y]->n[0];
        });
        std::vector<Lane1dArgs<T>> la(b->fused_list.size());
        for (size_t j = 0; j < la.size(); ++j) {
            const int i = b->fused_list[j];
            Lane1dArgs<T>& l = la[j];
            memset(&l, 0, sizeof(l));
            l.fa = fa[i * 3 + 0];
            l.fa.PX = 1;
            l.fa.TL = LANE1D_BS;
            // cells per thread: the smallest of a quarter, a half, three quarters and all
            // of the longest pencil's that holds the pencil (the lane kernel has a body
            // for each)
            const int N = b->ctx[i]->n[0];
            int kc = LANE1D_KMAX;
            for (int q = 3; q >= 1; --q)
                if (N <= LANE1D_BS * (q * LANE1D_KMAX / 4)) kc = q * LANE1D_KMAX / 4;
            l.fa.K = kc;
            l.fa.xmajor = 1;
            for (int dI = 1; dI < 3; ++dI) {
                l.ma[dI] = ma[i * 9 + 3 * dI + 0];
                l.ma[dI].PX = 1;
                l.ma[dI].TL = LANE1D_BS;
            }
            for (int k = 0; k < 3; ++k) l.ca[k] = ca[i * 3 + k];
        }
        if (e == hipSuccess && la.size() > b->l1cap) {
            hipFree(b->l1args);
            b->l1args = nullptr;
            b->l1cap = 0;
            e = hipMalloc(&b->l1args, sizeof(Lane1dArgs<T>) * la.size());
            if (e == hipSuccess) b->l1cap = la.size();
        }
        if (e == hipSuccess && !la.empty())
            e = hipMemcpyAsync(b->l1args, la.data(), sizeof(Lane1dArgs<T>) * la.size(),
                               hipMemcpyHostToDevice, b->st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "batch args upload");
}

static int upload_args(imh_batch* b) {
    return b->prec == IMH_FP64 ? upload_args_t<double>(b) : upload_args_t<float>(b);
}

static int batch_setup(imh_batch* b) {
    const int n = b->n;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->ncu = prop.multiProcessorCount;
    size_t es = b->prec == IMH_FP64 ? 8 : 4;
    (void)es;
    hipError_t e = hipMalloc(&b->sims_dev, sizeof(SimDev*) * n);
    if (e == hipSuccess) e = hipMalloc(&b->pars_dev, sizeof(SimParams*) * n);
    size_t fsz = b->prec == IMH_FP64 ? sizeof(FluidArgs<double>) : sizeof(FluidArgs<float>);
    size_t msz = b->prec == IMH_FP64 ? sizeof(MagnetArgs<double>) : sizeof(MagnetArgs<float>);
    size_t csz = b->prec == IMH_FP64 ? sizeof(CellMagArgs<double>) : sizeof(CellMagArgs<float>);
    size_t lsz = b->prec == IMH_FP64 ? sizeof(CflArgs<double>) : sizeof(CflArgs<float>);
    if (e == hipSuccess) e = hipMalloc(&b->fargs, fsz * n * 3);
    if (e == hipSuccess) e = hipMalloc(&b->margs, msz * n * 9);
    if (e == hipSuccess) e = hipMalloc(&b->cargs, csz * n * 3);
    if (e == hipSuccess) e = hipMalloc(&b->cflargs, lsz * n);
    if (e != hipSuccess) return hip_fail(e, "batch allocation");
    std::vector<SimDev*> sp(n);
    std::vector<SimParams*> pp(n);
    b->any_step_mode = false;
    for (int i = 0; i < n; ++i) {
        sp[i] = b->ctx[i]->sim;
        pp[i] = b->ctx[i]->par;
        if (b->ctx[i]->cfg.update_mode == 0) b->any_step_mode = true;
    }
    e = hipMemcpy(b->sims_dev, sp.data(), sizeof(SimDev*) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->pars_dev, pp.data(), sizeof(SimParams*) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "batch pointers");
    return IMH_OK;
}

// The first iteration of simulations that have not run one yet (edge stores of the flux
// objects are created there), each on its own, on the batch stream.
static int run_first_iterations(imh_batch* b, bool* changed) {
    *changed = false;
    for (imh_ctx* c : b->ctx) {
        if (!c->have_state) return fail(IMH_ESTATE, "batch simulation without uploaded state");
        bool fresh = (c->cfg.fluid_active && !c->fluid_init);
        if (!fresh && c->cfg.magnet_active) {
            for (int dI = 0; dI < 3 && !fresh; ++dI)
                for (int dX = 0; dX < 3 && !fresh; ++dX)
                    if (dI != dX && c->n[dX] >= 3 && !c->obj_init[OBJ_VG + 3 * dI + dX]) fresh = true;
        }
        if (!fresh) continue;
        hipStream_t own = c->st;
        c->st = b->st;
        int rc = imh_iteration_body(c, 1);
        c->st = own;
        if (rc) return rc;
        c->host_iteration += 1;
        *changed = true;
    }
    return IMH_OK;
}

// Build the six launch plans and their tile lists from every simulation's current
// iteration number.
static int build_plans(imh_batch* b) {
    free_plans(b);
    std::vector<BatchTile> all;
    const int n = b->n;
    std::vector<SimDev> sd(n);
    for (int i = 0; i < n; ++i) {
        hipError_t e = hipMemcpy(&sd[i], b->ctx[i]->sim, sizeof(SimDev), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "batch time state");
        b->ctx[i]->host_iteration = sd[i].iteration;
    }
    auto add_launch = [&](int p, Launch L, std::vector<BatchTile>& tl) {
        L.off = (long long)all.size();
        L.ntiles = (int)tl.size();
        all.insert(all.end(), tl.begin(), tl.end());
        b->plan[p].push_back(L);
    };
    const size_t es = b->prec == IMH_FP64 ? 8 : 4;
    // lane of each simulation: those with a magnetic field, then the pure-hydro ones
    std::vector<int> lane(n);
    bool has[2] = {false, false};
    for (int i = 0; i < n; ++i) {
        const imh_ctx* c = b->ctx[i];
        lane[i] = (c->cfg.magnet_active || c->mhd_kernel) ? 0 : 1;
        has[lane[i]] = true;
    }
    b->nlanes = (has[0] && has[1] && b->lanes_ok) ? 2 : 1;
    if (b->nlanes == 1)
        for (int i = 0; i < n; ++i) lane[i] = 0;
    for (int p = 0; p < 6; ++p) {
        Launch L{};
        L.lane = -1;
        std::vector<BatchTile> none;
        L.kind = L_BEGIN;
        L.reset = true;  // the first CFL reduction follows
        add_launch(p, L, none);
        auto cfl_block = [&](int step) {
            std::vector<BatchTile> tl;
            double by = 0;
            for (int i = 0; i < n; ++i) {
                const imh_ctx* c = b->ctx[i];
                long long nch = (c->ncell + cfl_chunk() - 1) / cfl_chunk();
                for (long long t = 0; t < nch; ++t) tl.push_back({i, 0, t});
                int nf = 5;
                for (int k = 0; k < 3; ++k) nf += c->mag_is_zero[k] ? 0 : 1;
                by += double(nf) * c->ncell * es;
            }
            if (step > 1) {  // the begin launch cleared the slots for step 1
                Launch r{};
                r.kind = L_CFL_RESET;
                r.lane = -1;
                add_launch(p, r, none);
            }
            Launch c{};
            c.kind = L_CFL;
            c.lane = -1;
            c.bytes = by;
            add_launch(p, c, tl);
            Launch f{};
            f.kind = L_FINALIZE;
            f.lane = -1;
            f.step = step;
            add_launch(p, f, none);
        };
        auto fork_join = [&](LaunchKind k) {
            if (b->nlanes < 2) return;
            Launch x{};
            x.kind = k;
            x.lane = -1;
            add_launch(p, x, none);
        };
        cfl_block(1);
        fork_join(L_FORK);
        if (!b->fused_list.empty()) {
            Launch f{};
            f.kind = L_LANE1D;
            f.lane = 0;
            double cells = 0;
            for (int i : b->fused_list) cells += double(b->ctx[i]->ncell);
            // one read and one write of the state (mass, momentum, energy, B) and the
            // cell-centred B written
            f.bytes = 19.0 * cells * es;
            add_launch(p, f, none);
        }
        for (int oi = 0; oi < 2; ++oi) {
            const int order = oi == 0 ? 1 : -1;
            if (oi == 1 && b->any_step_mode) {
                fork_join(L_JOIN);
                cfl_block(2);
                fork_join(L_FORK);
            }
            for (int t = 0; t < 3; ++t)
                for (int o = 0; o < 2; ++o) {
                    // fluid groups keyed by (bs, kmax, mhd, general)
                    struct G {
                        int lane;
                        int bs, kmax;
                        bool mhd, general;
                        double bytes;
                        std::vector<std::pair<int, BatchTile>> tl;
                    };
                    std::vector<G> groups;
                    std::vector<std::pair<int, BatchTile>> adv[2], ct[2];
                    std::vector<BatchTile> cm;
                    double mag_cells = 0;
                    int magkmax[2] = {1, 1};
                    size_t magsmem[2] = {0, 0};
                    for (int i = 0; i < n; ++i) {
                        imh_ctx* c = b->ctx[i];
                        if (b->fused[i]) continue;  // the 1-D lane launch runs it
                        int it = c->host_iteration + p;
                        int dirs[3], mi[3][2];
                        sweep_order(it, order, c->cfg.perms_order, dirs, mi);
                        const int X = dirs[t];
                        if (c->n[X - 1] < 3) continue;
                        const bool fluid_op = (order > 0) ? (o == 0) : (o == 1);
                        if (fluid_op && c->cfg.fluid_active) {
                            const FluidTile& ft = c->ftile[X - 1];
                            const int dX = X - 1;
                            const bool gen = c->cfg.bc[0][dX] == IMH_BC_FADE ||
                                             c->cfg.bc[1][dX] == IMH_BC_FADE || c->visc_type ||
                                             c->dline_dev[dX];
                            // bs 0: the column march (one thread per pencil)
                            const bool col = ft.column && !gen && !c->dline_dev[dX];
                            const int gbs = col ? 0 : ft.bs, gkm = col ? 0 : ft.kmax;
                            G* g = nullptr;
                            for (G& gg : groups)
                                if (gg.lane == lane[i] && gg.bs == gbs && gg.kmax == gkm &&
                                    gg.mhd == c->mhd_kernel &&
                                    gg.general == gen)
                                    g = &gg;
                            if (!g) {
                                groups.push_back({lane[i], gbs, gkm, c->mhd_kernel, gen, 0.0, {}});
                                g = &groups.back();
                            }
                            g->bytes += double(10 + (c->mhd_kernel ? 3 : 0)) * c->ncell * es;
                            const long long per = col ? fluid_column_pencils() : ft.PX;
                            // pencil slots: pairs with a lower wormhole boundary
                            const long long nslot = c->cfg.bc[0][X - 1] == IMH_BC_WORMHOLE
                                                        ? worm_slots(c->n, X - 1)
                                                        : c->nface[X - 1];
                            long long ntile = (nslot + per - 1) / per;
                            for (long long tt = 0; tt < ntile; ++tt)
                                g->tl.push_back({c->n[X - 1], {i * 3 + (X - 1), 0, tt}});
                        } else if (!fluid_op && c->cfg.magnet_active) {
                            const int dX = X - 1;
                            const MagTile& mt = c->mtile[dX];
                            mag_cells += double(c->ncell);
                            long long ntile = (c->nface[dX] + mt.PX - 1) / mt.PX;
                            int km = magnet_kmax_for(c->n[dX], mt.TL);
                            size_t sm = es * 2 * mt.PX * (size_t)(c->n[dX] + 4);
                            for (int s2 = 0; s2 < 2; ++s2) {
                                const int dI = mi[t][s2] - 1;
                                for (long long tt = 0; tt < ntile; ++tt) {
                                    BatchTile bt = {i * 9 + 3 * dI + dX, 0, tt};
                                    adv[s2].push_back({c->n[dX], bt});
                                    ct[s2].push_back({c->n[dX], bt});
                                }
                                magkmax[s2] = std::max(magkmax[s2], km);
                                magsmem[s2] = std::max(magsmem[s2], sm);
                            }
                            const int lastI = mi[t][1] - 1;
                            long long nch = (c->ncell + 255) / 256;
                            for (long long ch = 0; ch < nch; ++ch) {
                                cm.push_back({i * 3 + lastI, 0, ch});
                                cm.push_back({i * 3 + dX, 0, ch});
                            }
                        }
                    }
                    auto sorted = [](std::vector<std::pair<int, BatchTile>>& v) {
                        std::stable_sort(v.begin(), v.end(),
                                         [](const std::pair<int, BatchTile>& x,
                                            const std::pair<int, BatchTile>& y) { return x.first > y.first; });
                        std::vector<BatchTile> out;
                        out.reserve(v.size());
                        for (auto& e : v) out.push_back(e.second);
                        return out;
                    };
                    for (G& g : groups) {
                        Launch f{};
                        f.kind = L_FLUID;
                        f.lane = g.lane;
                        f.bs = g.bs;
                        f.kmax = g.kmax;
                        f.mhd = g.mhd;
                        f.general = g.general;
                        f.bytes = g.bytes;
                        std::vector<BatchTile> tl = sorted(g.tl);
                        add_launch(p, f, tl);
                    }
                    if (!adv[0].empty()) {
                        for (int s2 = 0; s2 < 2; ++s2) {
                            Launch a1{};
                            a1.kind = L_MAG_ADV;
                            a1.lane = 0;  // simulations with a magnetic field
                            a1.kmax = magkmax[s2];
                            a1.smem = magsmem[s2];
                            a1.bytes = 5.0 * mag_cells * es;
                            std::vector<BatchTile> tl = sorted(adv[s2]);
                            add_launch(p, a1, tl);
                            Launch c1{};
                            c1.kind = L_MAG_CT;
                            c1.lane = 0;
                            c1.kmax = magkmax[s2];
                            c1.smem = magsmem[s2] / 2;
                            c1.bytes = 3.0 * mag_cells * es;
                            std::vector<BatchTile> tl2 = sorted(ct[s2]);
                            add_launch(p, c1, tl2);
                        }
                        Launch c2{};
                        c2.kind = L_CELLMAG;
                        c2.lane = 0;
                        c2.bytes = 4.0 * mag_cells * es;
                        add_launch(p, c2, cm);
                    }
                }
            // source(): after each flux call, on the simulation's lane
            for (int i = 0; i < n; ++i) {
                const imh_ctx* c = b->ctx[i];
                if (!c->grav && !c->rad_type) continue;
                Launch s{};
                s.kind = L_SOURCE;
                s.lane = lane[i];
                s.step = i;
                s.bytes = source_bytes(c);
                add_launch(p, s, none);
            }
        }
        fork_join(L_JOIN);
        Launch E{};
        E.lane = -1;
        E.kind = L_END;
        add_launch(p, E, none);
    }
    if (all.size() > b->tiles_cap) {
        hipFree(b->tiles);
        b->tiles = nullptr;
        hipError_t e = hipMalloc(&b->tiles, sizeof(BatchTile) * std::max<size_t>(all.size(), 1));
        if (e != hipSuccess) return hip_fail(e, "batch tiles");
        b->tiles_cap = all.size();
    }
    if (!all.empty()) {
        hipError_t e = hipMemcpy(b->tiles, all.data(), sizeof(BatchTile) * all.size(),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "batch tiles upload");
    }
    b->planned = true;
    return IMH_OK;
}

template <typename T>
static hipError_t run_launch(imh_batch* b, const Launch& L, hipStream_t st) {
    const int n = b->n;
    const BatchTile* tl = b->tiles + L.off;
    // a multiple of 8 blocks: the XCD-range schedule (xcd_round_tile)
    const int nb = ((std::max(1, std::min(L.ntiles, b->ncu * 4)) + 7) / 8) * 8;
    switch (L.kind) {
    case L_BEGIN: return launch_begin_iteration(b->sims_dev, b->pars_dev, n, 0, st, L.reset ? 1 : 0);
    case L_CFL_RESET: return launch_cfl_reset(b->sims_dev, n, st);
    case L_CFL: return launch_cfl_batch<T>((const CflArgs<T>*)b->cflargs, tl, L.ntiles, nb, st);
    case L_FINALIZE: return launch_time_finalize(b->sims_dev, b->pars_dev, n, L.step, st);
    case L_FLUID:
        if (L.bs == 0)
            return launch_fluid_column<T>((const FluidArgs<T>*)b->fargs, tl, L.ntiles, nb, L.mhd,
                                          st);
        return launch_fluid_batch<T>((const FluidArgs<T>*)b->fargs, tl, L.ntiles, nb, L.bs, L.kmax,
                                     L.mhd, L.general, st);
    case L_MAG_ADV:
        return launch_magnet_batch<T>(0, (const MagnetArgs<T>*)b->margs, tl, L.ntiles, nb, L.kmax,
                                      L.smem, st);
    case L_MAG_CT:
        return launch_magnet_batch<T>(1, (const MagnetArgs<T>*)b->margs, tl, L.ntiles, nb, L.kmax,
                                      L.smem, st);
    case L_CELLMAG:
        return launch_cell_centered_batch<T>((const CellMagArgs<T>*)b->cargs, tl, L.ntiles, nb, st);
    case L_END: return launch_end_iteration(b->sims_dev, n, st);
    case L_SOURCE: return imh_launch_source(b->ctx[L.step], st);
    case L_LANE1D: return launch_lane1d(b->l1args, (int)b->fused_list.size(), st);
    case L_FORK:
    case L_JOIN: return hipSuccess;  // handled by run_plan
    }
    return hipErrorInvalidValue;
}

// Lane streams and events (default priority), made once the plans need them.
static int ensure_lanes(imh_batch* b) {
    if (b->nlanes < 2 || b->fork_ev) return IMH_OK;
    hipError_t e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
    for (int l = 0; l < MAX_LANES && e == hipSuccess; ++l) {
        e = hipStreamCreateWithFlags(&b->lane_st[l], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join_ev[l], hipEventDisableTiming);
    }
    return e == hipSuccess ? IMH_OK : hip_fail(e, "batch lanes");
}

// The launches of plan p.  With a profiler attached every launch goes to the batch
// stream (one at a time, so each launch's events bracket it alone).
static int run_plan(imh_batch* b, int p) {
    const bool lanes = b->nlanes > 1 && !b->prof && b->fork_ev;
    for (const Launch& L : b->plan[p]) {
        hipError_t e = hipSuccess;
        if (L.kind == L_FORK || L.kind == L_JOIN) {
            if (!lanes) continue;
            if (L.kind == L_FORK) {
                e = hipEventRecord(b->fork_ev, b->st);
                for (int l = 0; l < b->nlanes && e == hipSuccess; ++l)
                    e = hipStreamWaitEvent(b->lane_st[l], b->fork_ev, 0);
            } else {
                for (int l = 0; l < b->nlanes && e == hipSuccess; ++l) {
                    e = hipEventRecord(b->join_ev[l], b->lane_st[l]);
                    if (e == hipSuccess) e = hipStreamWaitEvent(b->st, b->join_ev[l], 0);
                }
            }
            if (e != hipSuccess) return hip_fail(e, "batch lane fork/join");
            continue;
        }
        hipStream_t st = (lanes && L.lane >= 0) ? b->lane_st[L.lane] : b->st;
        if (b->prof) prof_begin(b->prof, b->st, launch_kind_class(L.kind), L.bytes);
        e = b->prec == IMH_FP64 ? run_launch<double>(b, L, st) : run_launch<float>(b, L, st);
        if (b->prof) prof_end(b->prof, b->st);
        if (e != hipSuccess) return hip_fail(e, "batch launch");
    }
    return IMH_OK;
}

extern "C" {

static int batch_create_common(imh_batch* b) {
    for (const imh_ctx* c : b->ctx)
        for (int d = 0; d < 3; ++d)
            if (long_sweep(c, d))
                return fail(IMH_ENOTSUP, "pencils of %d cells along dim %d: simulations with "
                            "pencils longer than a workgroup holds, or Caramana-Shashkov-Whalen "
                            "viscosity at a fade end, run on their own (imh_run)", c->n[d], d + 1);
    int rc = batch_setup(b);
    if (rc) return rc;
    return IMH_OK;
}

int imh_batch_create(const imh_config* cfgs, int nsims, imh_batch** out) {
    if (!cfgs || !out || nsims < 1) return fail(IMH_EINVAL, "bad argument");
    *out = nullptr;
    for (int i = 0; i < nsims; ++i) {
        if (cfgs[i].precision != cfgs[0].precision || cfgs[i].device != cfgs[0].device)
            return fail(IMH_EINVAL, "batch simulations must share precision and device");
    }
    imh_batch* b = new imh_batch();
    b->n = nsims;
    b->prec = cfgs[0].precision;
    b->device = cfgs[0].device;
    b->owns = true;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "batch stream");
    }
    for (int i = 0; i < nsims; ++i) {
        imh_ctx* c = nullptr;
        int rc = imh_create(&cfgs[i], &c);
        if (rc) {
            batch_free(b);
            return rc;
        }
        hipStreamSynchronize(c->st);
        hipStreamDestroy(c->st);
        c->st = b->st;
        c->external_stream = true;
        b->ctx.push_back(c);
    }
    int rc = batch_create_common(b);
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return IMH_OK;
}

// A batch over existing contexts (the serving loop re-forms its batch this way when
// requests are admitted or retired).  The contexts stay owned by the caller; their data
// must not be touched through imh_* single-context calls while the view steps.
int imh_batch_create_view(imh_ctx* const* ctxs, int n, imh_batch** out) {
    if (!ctxs || !out || n < 1) return fail(IMH_EINVAL, "bad argument");
    *out = nullptr;
    imh_batch* b = new imh_batch();
    b->n = n;
    b->prec = ctxs[0]->cfg.precision;
    b->device = ctxs[0]->cfg.device;
    b->owns = false;
    // a view lives only until the next admission or retirement: the lane streams, events
    // and larger graphs cost more there than the overlap returns (measured on the
    // serving stream: p50 at 40 requests/s 0.13 s on one stream, 0.22 s with lanes)
    b->lanes_ok = false;
    for (int i = 0; i < n; ++i) {
        if (ctxs[i]->cfg.precision != b->prec || ctxs[i]->cfg.device != b->device) {
            delete b;
            return fail(IMH_EINVAL, "batch simulations must share precision and device");
        }
        hipStreamSynchronize(ctxs[i]->st);
        b->ctx.push_back(ctxs[i]);
    }
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "batch stream");
    }
    int rc = batch_create_common(b);
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return IMH_OK;
}

int imh_batch_destroy(imh_batch* b) {
    batch_free(b);
    return IMH_OK;
}

int imh_batch_upload_state(imh_batch* b, int sim, const double* mass, const double* const mom[3],
                           const double* ener, const double* const mag[3], size_t n) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    int rc = batch_quiesce(b);
    if (rc) return rc;
    rc = imh_upload_state(b->ctx[sim], mass, mom, ener, mag, n);
    free_plans(b);
    return rc;
}

int imh_batch_download_field(imh_batch* b, int sim, int field, double* out, size_t n) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    hipError_t e = hipStreamSynchronize(b->st);
    if (e != hipSuccess) return hip_fail(e, "batch sync");
    return imh_download_field(b->ctx[sim], field, out, n);
}

// First iterations of fresh simulations, then (re)build the argument blocks and plans.
static int batch_prepare(imh_batch* b) {
    bool changed = false;
    int rc = run_first_iterations(b, &changed);
    if (rc) return rc;
    if (changed || !b->planned) {
        hipError_t e = hipStreamSynchronize(b->st);
        if (e != hipSuccess) return hip_fail(e, "batch sync");
        if ((rc = upload_args(b))) return rc;
        if ((rc = build_plans(b))) return rc;
        // lane streams and events exist before any capture (no creation while capturing)
        if ((rc = ensure_lanes(b))) return rc;
        b->step = 0;
    }
    return IMH_OK;
}

int imh_batch_profile_step(imh_batch* b, int iterations, double* ms, double* bytes, int* count) {
    if (!b) return fail(IMH_EINVAL, "null batch");
    int rc = batch_prepare(b);
    if (rc) return rc;
    Profiler p;
    b->prof = &p;
    for (int it = 0; it < iterations && rc == IMH_OK; ++it) {
        rc = run_plan(b, (int)(b->step % 6));
        b->step += 1;
        for (imh_ctx* c : b->ctx) c->host_iteration += 1;
    }
    b->prof = nullptr;
    int rc2 = prof_collect(&p, ms, bytes, count);
    return rc ? rc : rc2;
}

int imh_batch_step(imh_batch* b, int iterations) {
    if (!b) return fail(IMH_EINVAL, "null batch");
    int rc = batch_prepare(b);
    if (rc) return rc;
    for (int it = 0; it < iterations; ++it) {
        const int p = (int)(b->step % 6);
        // a plan's first run is eager: short-lived batches (the serving loop's views,
        // re-formed at every admission or retirement) would capture graphs they replay
        // once; from the second run on the plan is captured and replayed
        if (!b->graph[p] && b->uses[p] == 0) {
            b->uses[p] = 1;
            if ((rc = run_plan(b, p))) return rc;
            b->step += 1;
            for (imh_ctx* c : b->ctx) c->host_iteration += 1;
            continue;
        }
        if (!b->graph[p]) {
            hipGraph_t g;
            hipError_t e = hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(e, "batch capture");
            rc = run_plan(b, p);
            e = hipStreamEndCapture(b->st, &g);
            if (rc) return rc;
            if (e != hipSuccess) return hip_fail(e, "batch end capture");
            e = hipGraphInstantiate(&b->graph[p], g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return hip_fail(e, "batch graph instantiate");
        }
        hipError_t e = hipGraphLaunch(b->graph[p], b->st);
        if (e != hipSuccess) return hip_fail(e, "batch graph launch");
        b->step += 1;
        for (imh_ctx* c : b->ctx) c->host_iteration += 1;
    }
    return IMH_OK;
}

int imh_batch_get_time(imh_batch* b, int sim, double* time, double* dtime, int* iteration) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    hipError_t e = hipStreamSynchronize(b->st);
    if (e != hipSuccess) return hip_fail(e, "batch sync");
    SimDev sd;
    e = hipMemcpy(&sd, b->ctx[sim]->sim, sizeof(sd), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "batch time state");
    if (time) *time = sd.time;
    if (dtime) *dtime = sd.dt;
    if (iteration) *iteration = sd.iteration;
    if (sd.error == IMH_ECFL) return fail(IMH_ECFL, "CFL error in simulation %d", sim);
    return IMH_OK;
}

int imh_batch_set_fluid_tile(imh_batch* b, int sim, int X, int bs, int kmax, int px) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    int rc = batch_quiesce(b);
    if (rc) return rc;
    rc = imh_set_fluid_tile(b->ctx[sim], X, bs, kmax, px);
    if (rc) return rc;
    free_plans(b);  // argument blocks and plans are rebuilt before the next step
    return IMH_OK;
}

// Source terms and viscosity of one simulation of the batch (the single-context calls,
// after the batch's work in flight has finished); the plans are rebuilt before the next
// step.
int imh_batch_set_gravity(imh_batch* b, int sim, const double* potential, size_t n,
                          double mass_threshold) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    // the setters write device memory that in-flight source and sweep kernels read, and
    // free_plans destroys graphs that may still run
    int rc = batch_quiesce(b);
    if (rc) return rc;
    rc = imh_set_gravity(b->ctx[sim], potential, n, mass_threshold);
    free_plans(b);
    return rc;
}

int imh_batch_set_fused_1d(imh_batch* b, int enable) {
    if (!b) return fail(IMH_EINVAL, "null batch");
    int rc = batch_quiesce(b);
    if (rc) return rc;
    b->fuse_1d = enable != 0;
    free_plans(b);
    return IMH_OK;
}

int imh_batch_set_spacing(imh_batch* b, int sim, int dim, const double* line, size_t n) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    int rc = batch_quiesce(b);
    if (rc) return rc;
    rc = imh_set_spacing(b->ctx[sim], dim, line, n);
    free_plans(b);
    return rc;
}

int imh_batch_set_radiation(imh_batch* b, int sim, int type, double exponent,
                            double initial_maximum) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    // the setters write device memory that in-flight source and sweep kernels read, and
    // free_plans destroys graphs that may still run
    int rc = batch_quiesce(b);
    if (rc) return rc;
    rc = imh_set_radiation(b->ctx[sim], type, exponent, initial_maximum);
    free_plans(b);
    return rc;
}

int imh_batch_set_viscosity(imh_batch* b, int sim, int type, double linear, double quadratic) {
    if (!b || sim < 0 || sim >= b->n) return fail(IMH_EINVAL, "bad simulation index");
    // the setters write device memory that in-flight source and sweep kernels read, and
    // free_plans destroys graphs that may still run
    int rc = batch_quiesce(b);
    if (rc) return rc;
    const imh_ctx* c = b->ctx[sim];
    if (type == 2)
        for (int d = 0; d < c->ndims; ++d)
            if (c->cfg.bc[0][d] == IMH_BC_FADE || c->cfg.bc[1][d] == IMH_BC_FADE)
                return fail(IMH_ENOTSUP, "Caramana-Shashkov-Whalen viscosity at a fade end runs "
                            "on single contexts (imh_run), not in a batch");
    rc = imh_set_viscosity(b->ctx[sim], type, linear, quadratic);
    free_plans(b);
    return rc;
}

int imh_batch_synchronize(imh_batch* b) {
    if (!b) return fail(IMH_EINVAL, "null batch");
    hipError_t e = hipStreamSynchronize(b->st);
    return e == hipSuccess ? IMH_OK : hip_fail(e, "batch sync");
}

int imh_batch_memory_bytes(imh_batch* b, size_t* bytes) {
    if (!b || !bytes) return fail(IMH_EINVAL, "null argument");
    size_t t = 0;
    for (imh_ctx* c : b->ctx) {
        size_t x = 0;
        imh_memory_bytes(c, &x);
        t += x;
    }
    *bytes = t;
    return IMH_OK;
}

imh_ctx* imh_batch_context(imh_batch* b, int sim) {
    return (b && sim >= 0 && sim < b->n) ? b->ctx[sim] : nullptr;
}

}  // extern "C"

hipStream_t imh_batch_stream(imh_batch* b) { return b ? b->st : nullptr; }
