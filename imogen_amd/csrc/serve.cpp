// serve.cpp -- the serving loop: continuous batching of independent solve requests on one
// device.
//
// A request is one simulation (its configuration and initial state) run until
// TimeManager.running turns false (TIMEMAX reached, ITERMAX passed, or a CFL error).
// Requests wait in a FIFO queue and are admitted while a slot is free and their device
// memory fits the HBM budget; admitted requests keep their state on the device, run
// their first iteration on their own and then join the batch that every scheduling round
// advances with replayed launch plans.  Finished requests leave the batch at the next
// round (their result stays on the device until released), so requests that arrive while
// others run overlap with them on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "imh_common.h"
#include "imh_host.h"
#include "imogen_hip.h"

using namespace imh;

namespace {

struct Request {
    imh_config cfg;
    std::vector<double> state;  // mass, mom x3, ener, mag x3 (8 x n), until admission
    size_t n = 0;
    int ic_kind = 0;                  // > 0: the state is generated on the device
    std::vector<double> ic_params;
    int status = IMH_REQ_QUEUED;
    imh_ctx* ctx = nullptr;
    size_t bytes = 0;
    double t_submit = 0, t_admit = -1, t_done = -1;
    SimDev last{};
};

}  // namespace

struct imh_server {
    int device = 0;
    int precision = IMH_FP32_BULK;
    size_t budget = 0;
    size_t used = 0;
    int max_active = 0;
    long long next_id = 1;
    std::deque<long long> queue;
    std::map<long long, Request> reqs;
    std::vector<long long> active;
    imh_batch* view = nullptr;
    bool dirty = false;
    SimDev* host_sd = nullptr;  // pinned, max_active entries
    std::chrono::steady_clock::time_point t0;
    long long cell_updates = 0;  // 2 * cells per completed iteration, finished requests
    // contexts of released requests, kept for the next admission of the same cell count
    // (their device memory is reused instead of allocated), oldest first
    struct Pooled {
        imh_ctx* ctx;
        size_t bytes;
    };
    std::deque<Pooled> pool;
    size_t pooled = 0;           // device bytes held by the pool
};

// Released contexts kept for reuse at most (each holds its grid's fields).
constexpr size_t POOL_MAX = 64;

extern "C" imh_ctx* imh_batch_context(imh_batch* b, int sim);
hipStream_t imh_batch_stream(imh_batch* b);

static double now_s(const imh_server* s) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

static void release_request(imh_server* s, Request& r) {
    if (r.ctx) {
        hipStreamSynchronize(r.ctx->st);
        if (s->pool.size() < POOL_MAX && !r.ctx->external_stream) {
            s->pool.push_back({r.ctx, r.bytes});
            s->pooled += r.bytes;
        } else {
            imh_destroy(r.ctx);
        }
        r.ctx = nullptr;
        s->used -= std::min(s->used, r.bytes);
    }
}

// A context for request r: a pooled one of the same cell count, precision and field set
// when there is one (its memory reused), else a new one.  Pooled contexts are freed,
// oldest first, while the budget needs their memory.
static int context_for(imh_server* s, Request& r) {
    const long long ncell = (long long)r.n;
    for (size_t i = 0; i < s->pool.size(); ++i) {
        imh_ctx* p = s->pool[i].ctx;
        if (p->ncell != ncell || p->cfg.precision != r.cfg.precision ||
            (r.cfg.magnet_active && !p->cfg.magnet_active))
            continue;
        const size_t pb = s->pool[i].bytes;
        const int rc = imh_ctx_recycle(p, &r.cfg, &r.ctx);
        if (rc == IMH_ENOTSUP) continue;  // edge stores do not fit: untouched, kept
        s->pool.erase(s->pool.begin() + (long)i);
        s->pooled -= std::min(s->pooled, pb);
        return rc;
    }
    while (!s->pool.empty() && s->used + s->pooled + r.bytes > s->budget) {
        imh_destroy(s->pool.front().ctx);
        s->pooled -= std::min(s->pooled, s->pool.front().bytes);
        s->pool.pop_front();
    }
    return imh_create(&r.cfg, &r.ctx);
}

// Admit queued requests in order while a slot is free and the budget allows.
static int admit(imh_server* s) {
    while (!s->queue.empty() && (int)s->active.size() < s->max_active) {
        Request& r = s->reqs[s->queue.front()];
        if (s->used + r.bytes > s->budget) {
            if (s->active.empty() && s->used == 0) {
                // can never fit: fail it rather than block the queue forever
                r.status = IMH_REQ_FAILED;
                r.t_done = now_s(s);
                s->queue.pop_front();
                continue;
            }
            break;
        }
        const long long id = s->queue.front();
        s->queue.pop_front();
        int rc = context_for(s, r);
        if (rc == IMH_OK && r.ic_kind > 0) {
            rc = imh_init_problem(r.ctx, r.ic_kind, r.ic_params.data(), (int)r.ic_params.size());
        } else if (rc == IMH_OK) {
            const double* d = r.state.data();
            const size_t n = r.n;
            const double* mom[3] = {d + n, d + 2 * n, d + 3 * n};
            const double* mag[3] = {d + 5 * n, d + 6 * n, d + 7 * n};
            rc = imh_upload_state(r.ctx, d, mom, d + 4 * n, mag, n);
        }
        if (rc != IMH_OK) {
            if (r.ctx) imh_destroy(r.ctx);
            r.ctx = nullptr;
            r.status = IMH_REQ_FAILED;
            r.t_done = now_s(s);
            if (rc == IMH_ENOMEM) return rc;
            continue;
        }
        r.state.clear();
        r.state.shrink_to_fit();
        r.status = IMH_REQ_RUNNING;
        r.t_admit = now_s(s);
        s->used += r.bytes;
        s->active.push_back(id);
        s->dirty = true;
    }
    return IMH_OK;
}

static int reform(imh_server* s) {
    if (s->view) imh_batch_destroy(s->view);
    s->view = nullptr;
    s->dirty = false;
    if (s->active.empty()) return IMH_OK;
    std::vector<imh_ctx*> ctxs;
    for (long long id : s->active) ctxs.push_back(s->reqs[id].ctx);
    return imh_batch_create_view(ctxs.data(), (int)ctxs.size(), &s->view);
}

// Read every active simulation's time state; retire the finished ones.
static int poll(imh_server* s) {
    if (!s->view || s->active.empty()) return IMH_OK;
    hipStream_t st = imh_batch_stream(s->view);
    for (size_t i = 0; i < s->active.size(); ++i) {
        imh_ctx* c = s->reqs[s->active[i]].ctx;
        hipError_t e = hipMemcpyAsync(&s->host_sd[i], c->sim, sizeof(SimDev),
                                      hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(e, "serve poll");
    }
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "serve poll");
    const double t = now_s(s);
    std::vector<long long> still;
    for (size_t i = 0; i < s->active.size(); ++i) {
        Request& r = s->reqs[s->active[i]];
        const SimDev& d = s->host_sd[i];
        r.last = d;
        const bool ended = d.error != 0 || d.iteration > r.cfg.iter_max || d.time >= r.cfg.time_max;
        if (ended) {
            r.status = d.error ? IMH_REQ_FAILED : IMH_REQ_DONE;
            r.t_done = t;
            s->cell_updates += 2LL * (long long)r.n * (d.iteration - 1);
            s->dirty = true;
        } else {
            still.push_back(s->active[i]);
        }
    }
    s->active.swap(still);
    return IMH_OK;
}

extern "C" {

int imh_serve_create(int device, int precision, size_t hbm_budget, int max_active,
                     imh_server** out) {
    if (!out || max_active < 1 || (precision != IMH_FP64 && precision != IMH_FP32_BULK))
        return fail(IMH_EINVAL, "bad argument");
    *out = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    imh_server* s = new imh_server();
    s->device = device;
    s->precision = precision;
    s->max_active = max_active;
    size_t free_b = 0, total_b = 0;
    e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipMemGetInfo");
    }
    // the budget never exceeds 90% of what is free now
    const size_t cap = free_b / 10 * 9;
    s->budget = hbm_budget == 0 ? cap : std::min(hbm_budget, cap);
    e = hipHostMalloc((void**)&s->host_sd, sizeof(SimDev) * max_active, hipHostMallocDefault);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipHostMalloc");
    }
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return IMH_OK;
}

int imh_serve_destroy(imh_server* s) {
    if (!s) return IMH_OK;
    if (s->view) imh_batch_destroy(s->view);
    for (auto& kv : s->reqs) {
        if (kv.second.ctx) imh_destroy(kv.second.ctx);
        kv.second.ctx = nullptr;
    }
    for (auto& p : s->pool) imh_destroy(p.ctx);
    s->pool.clear();
    if (s->host_sd) (void)hipHostFree(s->host_sd);
    delete s;
    return IMH_OK;
}

int imh_serve_submit(imh_server* s, const imh_config* cfg, const double* mass,
                     const double* const mom[3], const double* ener, const double* const mag[3],
                     size_t n, long long* id) {
    if (!s || !cfg || !mass || !mom || !ener || !mag || !id) return fail(IMH_EINVAL, "null argument");
    imh_config c = *cfg;
    c.precision = s->precision;
    c.device = s->device;
    size_t bytes = 0;
    int rc = imh_memory_estimate(&c, &bytes);
    if (rc) return rc;
    const size_t ncell = (size_t)c.grid[0] * c.grid[1] * c.grid[2];
    if (n != ncell) return fail(IMH_EINVAL, "state has %zu cells, the grid %zu", n, ncell);
    Request r;
    r.cfg = c;
    r.n = n;
    r.bytes = bytes;
    r.state.resize(8 * n);
    const double* src[8] = {mass, mom[0], mom[1], mom[2], ener, mag[0], mag[1], mag[2]};
    for (int f = 0; f < 8; ++f) {
        if (!src[f]) return fail(IMH_EINVAL, "null field %d", f);
        std::memcpy(r.state.data() + f * n, src[f], n * sizeof(double));
    }
    r.t_submit = now_s(s);
    *id = s->next_id++;
    s->reqs[*id] = std::move(r);
    s->queue.push_back(*id);
    return IMH_OK;
}

int imh_serve_submit_problem(imh_server* s, const imh_config* cfg, int kind,
                             const double* params, int nparams, long long* id) {
    if (!s || !cfg || !id || nparams < 0 || (nparams > 0 && !params))
        return fail(IMH_EINVAL, "bad argument");
    if (kind < IMH_IC_SOD || kind > IMH_IC_KELVIN_HELMHOLTZ)
        return fail(IMH_EINVAL, "unknown initial-condition kind %d", kind);
    imh_config c = *cfg;
    c.precision = s->precision;
    c.device = s->device;
    size_t bytes = 0;
    int rc = imh_memory_estimate(&c, &bytes);
    if (rc) return rc;
    Request r;
    r.cfg = c;
    r.n = (size_t)c.grid[0] * c.grid[1] * c.grid[2];
    r.bytes = bytes;
    r.ic_kind = kind;
    r.ic_params.assign(params, params + nparams);
    r.t_submit = now_s(s);
    *id = s->next_id++;
    s->reqs[*id] = std::move(r);
    s->queue.push_back(*id);
    return IMH_OK;
}

int imh_serve_step(imh_server* s, int iterations, int* active, int* queued) {
    if (!s || iterations < 0) return fail(IMH_EINVAL, "bad argument");
    int rc = admit(s);
    if (rc) return rc;
    if (s->dirty && (rc = reform(s))) return rc;
    if (s->view && iterations > 0) {
        if ((rc = imh_batch_step(s->view, iterations))) return rc;
        if ((rc = poll(s))) return rc;
    }
    if (s->dirty) {
        // retire finished requests from the batch now, so the next round only plans
        // the running ones (and admits into the freed slots)
        if ((rc = admit(s))) return rc;
        if ((rc = reform(s))) return rc;
    }
    if (active) *active = (int)s->active.size();
    if (queued) *queued = (int)s->queue.size();
    return IMH_OK;
}

int imh_serve_status(imh_server* s, long long id, int* status, double* time, int* iteration,
                     double t_events[3]) {
    if (!s) return fail(IMH_EINVAL, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(IMH_EINVAL, "unknown request %lld", id);
    const Request& r = it->second;
    if (status) *status = r.status;
    if (time) *time = r.last.time;
    if (iteration) *iteration = r.last.iteration;
    if (t_events) {
        t_events[0] = r.t_submit;
        t_events[1] = r.t_admit;
        t_events[2] = r.t_done;
    }
    return IMH_OK;
}

int imh_serve_download(imh_server* s, long long id, int field, double* out, size_t n) {
    if (!s) return fail(IMH_EINVAL, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(IMH_EINVAL, "unknown request %lld", id);
    Request& r = it->second;
    if (!r.ctx) return fail(IMH_ESTATE, "request %lld has no device state", id);
    if (s->view) {
        hipError_t e = hipStreamSynchronize(imh_batch_stream(s->view));
        if (e != hipSuccess) return hip_fail(e, "serve sync");
    }
    return imh_download_field(r.ctx, field, out, n);
}

int imh_serve_release(imh_server* s, long long id) {
    if (!s) return fail(IMH_EINVAL, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(IMH_EINVAL, "unknown request %lld", id);
    if (it->second.status == IMH_REQ_RUNNING) return fail(IMH_ESTATE, "request %lld is running", id);
    if (it->second.status == IMH_REQ_QUEUED) {
        s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), id), s->queue.end());
    }
    release_request(s, it->second);
    s->reqs.erase(it);
    return IMH_OK;
}

int imh_serve_stats(imh_server* s, size_t* used_bytes, size_t* budget_bytes,
                    long long* cell_updates) {
    if (!s) return fail(IMH_EINVAL, "null server");
    if (used_bytes) *used_bytes = s->used;
    if (budget_bytes) *budget_bytes = s->budget;
    if (cell_updates) *cell_updates = s->cell_updates;
    return IMH_OK;
}

}  // extern "C"
