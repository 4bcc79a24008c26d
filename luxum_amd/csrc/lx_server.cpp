// lx_server.cpp -- continuous batching of simulation requests.
//
// Requests carry their own arrival time.  The run loop admits every arrived
// request that fits (first fit in arrival order: the sum of reserved device
// footprints stays within the HBM cap, the number of active requests within
// the slot limit), advances all active requests together by one tick, and
// retires the requests that reached their step count at the end of the tick
// (probe series and, when asked, final fields copied to the host; device
// memory released).  Requests that arrive while a tick runs join at the next
// tick boundary, so a long request never blocks a short one behind it.
//
// 3D requests also wait for the device's 3D capacity: they are admitted in
// arrival order while the 3D cells already active stay below k3dCells (a 3D
// grid of that size fills the device on its own), so large 3D requests run
// one after another at full speed instead of sharing every tick with each
// other, which would stretch the latency of every early arrival; small 3D
// grids still batch together, and 2D requests fill in first fit around them.
// LX_SERVE_3D_CELLS overrides the limit (0: none).
//
// A tick is the smallest remaining step count of the active requests, capped
// at kTick: every request ends exactly on a tick boundary.  Batched mode
// steps one lx_batch (planar requests in one tiled launch per chunk, the
// others on their own streams, concurrently); plain mode runs one request at
// a time, start to end, with the single-simulation kernels.  Requests that
// arrived are set up (allocation, tables, material upload) while the tick
// runs on the device, and join the next one.  The server fills the
// context's stream pool once, so admitting a request creates no stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <new>
#include <thread>
#include <vector>

#include "../../include/luxum.h"
#include "lx_sim.hpp"
#include "lx_util.hpp"

using lx::fail;

namespace {

constexpr int kTick = 128;
constexpr int64_t k3dCells = int64_t(8) << 20;
constexpr int kPrewarm = 64;   // pooled streams created with the server

struct Req {
    lx_params p{};
    std::vector<uint8_t> ids;
    std::vector<double> eps;
    std::vector<int32_t> probes;
    int32_t n_probes = 0;
    int32_t steps = 0;
    double arrival = 0;
    int keep_fields = 0;
    int state = LX_REQ_QUEUED;
    lx_sim *sim = nullptr;
    int32_t done = 0;
    int64_t footprint = 0;
    int64_t cells = 0;
    int64_t bytes_per_step = 0;
    double admitted = 0, finished = 0;
    std::vector<double> probe_data;
    std::vector<std::vector<double>> fields;   // per component, when kept
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct lx_server {
    lx_ctx *ctx = nullptr;
    int mode = LX_SERVE_BATCHED;
    int64_t cap = 0;
    int max_active = 1;
    lx_batch *batch = nullptr;
    std::vector<Req *> reqs;      // by id
    std::deque<Req *> queue;      // not yet admitted, arrival order
    std::vector<Req *> active;
    std::vector<Req *> retiring;  // finished on the device, results not yet read back
    int64_t used = 0;
    int64_t cells3d_cap = 0;      // 3D cells that may be active together (0: no limit)
    lx_server_stats st{};
    Clock::time_point t0;
};

namespace {

double since(const lx_server *s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

lx_status admit(lx_server *s, Req *r) {
    lx_status st = lx_sim_create(s->ctx, &r->p, &r->sim);
    if (st) return st;
    if ((st = lx_sim_set_materials(r->sim, r->ids.data(), r->eps.data(), (int32_t)r->eps.size()))) return st;
    if (r->n_probes && (st = lx_sim_set_probes(r->sim, r->probes.data(), r->n_probes))) return st;
    lx_sim_info info{};
    lx_sim_get_info(r->sim, &info);
    r->bytes_per_step = info.bytes_per_step;
    if (s->mode == LX_SERVE_BATCHED && (st = lx_batch_add(s->batch, r->sim))) return st;
    r->ids.clear();
    r->ids.shrink_to_fit();
    r->state = LX_REQ_ACTIVE;
    r->admitted = since(s);
    return LX_OK;
}

lx_status retire(lx_server *s, Req *r) {
    lx_status st;
    if (r->n_probes) {
        r->probe_data.assign((size_t)r->steps * r->n_probes, 0.0);
        int32_t ns = 0;
        if ((st = lx_sim_read_probes(r->sim, r->probe_data.data(), r->probe_data.size(), &ns))) return st;
    }
    if (r->keep_fields) {
        const int N[3] = {r->p.nx, r->p.is_planar ? 1 : r->p.ny, r->p.nz};
        r->fields.assign(6, {});
        for (int c = 0; c < 6; ++c) {
            int ext[3];
            if (!lx::comp_extent(N, r->p.is_planar, c, ext)) continue;
            r->fields[c].assign((size_t)ext[0] * ext[1] * ext[2], 0.0);
            if ((st = lx_sim_read_field(r->sim, c, r->fields[c].data(), r->fields[c].size()))) return st;
        }
    }
    r->finished = since(s);
    lx_sim_destroy(r->sim);   // leaves the batch
    r->sim = nullptr;
    r->state = LX_REQ_DONE;
    s->used -= r->footprint;
    s->st.completed++;
    s->st.makespan = std::max(s->st.makespan, r->finished);
    s->st.cell_updates += (double)r->cells * r->steps;
    s->st.alg_bytes += (double)r->bytes_per_step * r->steps;
    return LX_OK;
}

void fail_request(lx_server *s, Req *r) {
    if (r->sim) lx_sim_destroy(r->sim);
    r->sim = nullptr;
    if (r->state == LX_REQ_ACTIVE) s->used -= r->footprint;
    r->state = LX_REQ_FAILED;
    s->st.failed++;
}

// admit arrived requests that fit, first fit in arrival order; 3D requests
// in strict arrival order within the 3D cell limit
lx_status admit_arrived(lx_server *s) {
    const double now = since(s);
    int64_t cells3d = 0;
    for (const Req *r : s->active)
        if (!r->p.is_planar) cells3d += r->cells;
    bool hold3d = false;   // an earlier 3D request waits: later ones wait behind it
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Req *r = *it;
        if (r->arrival > now) break;
        if ((int)s->active.size() >= s->max_active) break;
        const bool is3d = !r->p.is_planar;
        if (is3d && s->mode == LX_SERVE_BATCHED &&
            (hold3d || (s->cells3d_cap > 0 && cells3d > 0 && cells3d + r->cells > s->cells3d_cap))) {
            hold3d = true;
            ++it;
            continue;
        }
        if (s->used + r->footprint > s->cap) {
            if (is3d) hold3d = true;
            ++it;
            continue;
        }
        if (is3d) cells3d += r->cells;
        s->used += r->footprint;
        r->state = LX_REQ_ACTIVE;   // reserved
        lx_status st = admit(s, r);
        if (st == LX_ENOMEM && !s->active.empty()) {
            // the device is fuller than the estimates say: retry after a retirement
            if (r->sim) lx_sim_destroy(r->sim);
            r->sim = nullptr;
            r->state = LX_REQ_QUEUED;
            s->used -= r->footprint;
            break;
        }
        it = s->queue.erase(it);
        if (st) {
            fail_request(s, r);
            continue;
        }
        s->active.push_back(r);
        s->st.peak_active = std::max<int32_t>(s->st.peak_active, (int32_t)s->active.size());
        s->st.peak_bytes = std::max(s->st.peak_bytes, s->used);
        if (s->mode == LX_SERVE_PLAIN) break;
    }
    return LX_OK;
}

void wait_for_arrival(lx_server *s) {
    double next = 1e300;
    for (Req *r : s->queue) next = std::min(next, r->arrival);
    const double dt = next - since(s);
    if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(dt, 0.5)));
}

lx_status retire_pending(lx_server *s) {
    for (Req *r : s->retiring) {
        lx_status st = retire(s, r);
        if (st) return st;
    }
    s->retiring.clear();
    return LX_OK;
}

lx_status tick(lx_server *s) {
    int32_t k = kTick;
    for (Req *r : s->active) k = std::min(k, r->steps - r->done);
    const std::vector<Req *> stepped = s->active;   // the members of this tick
    lx_status st;
    if (s->mode == LX_SERVE_BATCHED) {
        if ((st = lx_batch_step(s->batch, k))) return st;
        // while the tick runs on the device: read back the requests that
        // finished in the previous tick (their memory returns to the pool)
        // and set up the requests that arrived meanwhile
        if ((st = retire_pending(s))) return st;
        if ((st = admit_arrived(s))) return st;
        if ((st = lx_batch_sync(s->batch))) return st;
    } else {
        Req *r = s->active.front();
        k = r->steps - r->done;
        if ((st = lx_sim_step(r->sim, k))) return st;
        if ((st = lx_sim_sync(r->sim))) return st;
    }
    s->st.ticks++;
    std::vector<Req *> keep;
    for (Req *r : stepped) {
        r->done += k;
        if (r->done < r->steps) {
            keep.push_back(r);
        } else if (s->mode == LX_SERVE_BATCHED) {
            // leave the batch now (it is idle); results are read during the next tick
            if ((st = lx_batch_remove(s->batch, r->sim))) return st;
            s->retiring.push_back(r);
        } else if ((st = retire(s, r))) {
            return st;
        }
    }
    for (size_t q = stepped.size(); q < s->active.size(); ++q) keep.push_back(s->active[q]);   // admitted during the tick
    s->active.swap(keep);
    return LX_OK;
}

}  // namespace

extern "C" {

lx_status lx_server_create(lx_ctx *ctx, int32_t mode, int64_t hbm_cap_bytes, int32_t max_active, lx_server **out) {
    if (!ctx || !out) return fail(LX_EINVAL, "lx_server_create: null pointer");
    if (mode != LX_SERVE_BATCHED && mode != LX_SERVE_PLAIN) return fail(LX_EINVAL, "unknown serving mode %d", mode);
    if (hbm_cap_bytes <= 0 || max_active < 1) return fail(LX_EINVAL, "cap and slot count must be positive");
    lx_server *s = new (std::nothrow) lx_server;
    if (!s) return fail(LX_ENOMEM, "out of host memory");
    s->ctx = ctx;
    s->mode = mode;
    s->cap = hbm_cap_bytes;
    s->max_active = mode == LX_SERVE_PLAIN ? 1 : max_active;
    s->cells3d_cap = k3dCells;
    if (const char *e = std::getenv("LX_SERVE_3D_CELLS")) s->cells3d_cap = std::max<int64_t>(0, std::atoll(e));
    if (hipSetDevice(ctx->device) != hipSuccess) {
        delete s;
        return fail(LX_EDEVICE, "hipSetDevice(%d) failed", ctx->device);
    }
    lx::prewarm_streams(ctx, std::min(s->max_active, kPrewarm));
    if (mode == LX_SERVE_BATCHED) {
        lx_status st = lx_batch_create(ctx, &s->batch);
        if (st) {
            lx_server_destroy(s);
            return st;
        }
    }
    *out = s;
    return LX_OK;
}

void lx_server_destroy(lx_server *s) {
    if (!s) return;
    for (Req *r : s->reqs) {
        if (r->sim) lx_sim_destroy(r->sim);
        delete r;
    }
    if (s->batch) lx_batch_destroy(s->batch);
    delete s;
}

lx_status lx_server_submit(lx_server *s, const lx_request *q, int64_t *id) {
    if (!s || !q || !q->params || !q->ids || !q->eps_r) return fail(LX_EINVAL, "lx_server_submit: null pointer");
    if (q->steps < 1) return fail(LX_EINVAL, "a request needs at least one step");
    if (q->n_ids < 1 || q->n_ids > 256) return fail(LX_EINVAL, "n_ids must be in [1, 256]");
    if (q->n_probes < 0 || (q->n_probes && !q->probes)) return fail(LX_EINVAL, "bad probe list");
    lx_status st = lx_params_check(q->params);
    if (st) return st;
    const lx_params &p = *q->params;
    const int64_t cells = (int64_t)p.nx * (p.is_planar ? 1 : p.ny) * p.nz;
    if ((int64_t)q->n_cells != cells) return fail(LX_EINVAL, "ids holds %zu cells, the grid %lld", q->n_cells, (long long)cells);
    Req *r = new (std::nothrow) Req;
    if (!r) return fail(LX_ENOMEM, "out of host memory");
    r->p = p;
    r->p.num_timesteps = q->steps;
    r->p.monitor_every = 0;
    r->ids.assign(q->ids, q->ids + q->n_cells);
    r->eps.assign(q->eps_r, q->eps_r + q->n_ids);
    r->n_probes = q->n_probes;
    if (q->n_probes) r->probes.assign(q->probes, q->probes + 4 * (size_t)q->n_probes);
    r->steps = q->steps;
    r->arrival = q->arrival;
    r->keep_fields = q->keep_fields;
    r->cells = cells;
    if ((st = lx_params_footprint(&r->p, s->mode == LX_SERVE_BATCHED, q->steps, q->n_probes, &r->footprint))) {
        delete r;
        return st;
    }
    if (r->footprint > s->cap) {
        delete r;
        return fail(LX_ENOMEM, "request footprint %lld bytes exceeds the cap", (long long)r->footprint);
    }
    const int64_t n = (int64_t)s->reqs.size();
    s->reqs.push_back(r);
    // keep the queue in arrival order
    auto it = std::upper_bound(s->queue.begin(), s->queue.end(), r,
                               [](const Req *a, const Req *b) { return a->arrival < b->arrival; });
    s->queue.insert(it, r);
    s->st.submitted++;
    if (id) *id = n;
    return LX_OK;
}

lx_status lx_server_run(lx_server *s) {
    if (!s) return fail(LX_EINVAL, "null server");
    s->t0 = Clock::now();
    while (!s->queue.empty() || !s->active.empty() || !s->retiring.empty()) {
        lx_status st;
        if (s->active.empty() && (st = retire_pending(s))) return st;
        if ((st = admit_arrived(s))) return st;
        if (s->active.empty()) {
            if (!s->queue.empty()) wait_for_arrival(s);
            continue;
        }
        if ((st = tick(s))) return st;
    }
    return LX_OK;
}

lx_status lx_server_info(lx_server *s, int64_t id, lx_request_info *info) {
    if (!s || !info || id < 0 || id >= (int64_t)s->reqs.size()) return fail(LX_EINVAL, "lx_server_info: bad request id");
    const Req *r = s->reqs[id];
    std::memset(info, 0, sizeof(*info));
    info->state = r->state;
    info->steps_done = r->done;
    info->arrival = r->arrival;
    info->admitted = r->admitted;
    info->finished = r->finished;
    info->footprint = r->footprint;
    info->cells = r->cells;
    info->n_probes = r->n_probes;
    return LX_OK;
}

lx_status lx_server_probes(lx_server *s, int64_t id, double *dst, size_t n) {
    if (!s || !dst || id < 0 || id >= (int64_t)s->reqs.size()) return fail(LX_EINVAL, "lx_server_probes: bad arguments");
    const Req *r = s->reqs[id];
    if (r->state != LX_REQ_DONE) return fail(LX_EINVAL, "request %lld is not done", (long long)id);
    if (n != r->probe_data.size()) return fail(LX_EINVAL, "probe buffer holds %zu values, expected %zu", n, r->probe_data.size());
    std::memcpy(dst, r->probe_data.data(), n * sizeof(double));
    return LX_OK;
}

lx_status lx_server_field(lx_server *s, int64_t id, int32_t comp, double *dst, size_t n) {
    if (!s || !dst || id < 0 || id >= (int64_t)s->reqs.size() || comp < 0 || comp > 5)
        return fail(LX_EINVAL, "lx_server_field: bad arguments");
    const Req *r = s->reqs[id];
    if (r->state != LX_REQ_DONE || r->fields.empty()) return fail(LX_EINVAL, "request %lld kept no fields", (long long)id);
    const std::vector<double> &f = r->fields[comp];
    if (f.empty()) return fail(LX_EINVAL, "component %d does not exist on this grid", comp);
    if (n != f.size()) return fail(LX_EINVAL, "field buffer holds %zu values, expected %zu", n, f.size());
    std::memcpy(dst, f.data(), n * sizeof(double));
    return LX_OK;
}

lx_status lx_server_get_stats(lx_server *s, lx_server_stats *st) {
    if (!s || !st) return fail(LX_EINVAL, "null pointer");
    *st = s->st;
    return LX_OK;
}

}  // extern "C"
