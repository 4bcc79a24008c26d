// lx_batch.cpp -- batches of independent simulations.
//
// A batch advances its member simulations together: every member's steps are
// enqueued on the member's own compute stream, so the device overlaps their
// kernels; lx_batch_sync waits for all of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/luxum.h"
#include "lx_util.hpp"

struct lx_batch {
    lx_ctx *ctx;
    std::vector<lx_sim *> sims;
};

extern "C" {

lx_status lx_batch_create(lx_ctx *ctx, lx_batch **out) {
    if (!ctx || !out) return lx::fail(LX_EINVAL, "lx_batch_create: null pointer");
    lx_batch *b = new (std::nothrow) lx_batch;
    if (!b) return lx::fail(LX_ENOMEM, "out of host memory");
    b->ctx = ctx;
    *out = b;
    return LX_OK;
}

void lx_batch_destroy(lx_batch *b) { delete b; }

lx_status lx_batch_add(lx_batch *b, lx_sim *s) {
    if (!b || !s) return lx::fail(LX_EINVAL, "lx_batch_add: null pointer");
    if (std::find(b->sims.begin(), b->sims.end(), s) != b->sims.end()) return lx::fail(LX_EINVAL, "already in batch");
    b->sims.push_back(s);
    return LX_OK;
}

lx_status lx_batch_remove(lx_batch *b, lx_sim *s) {
    if (!b || !s) return lx::fail(LX_EINVAL, "lx_batch_remove: null pointer");
    auto it = std::find(b->sims.begin(), b->sims.end(), s);
    if (it == b->sims.end()) return lx::fail(LX_EINVAL, "not in batch");
    b->sims.erase(it);
    return LX_OK;
}

lx_status lx_batch_step(lx_batch *b, int32_t nsteps) {
    if (!b) return lx::fail(LX_EINVAL, "null batch");
    for (lx_sim *s : b->sims) {
        lx_status st = lx_sim_step(s, nsteps);
        if (st) return st;
    }
    return LX_OK;
}

lx_status lx_batch_sync(lx_batch *b) {
    if (!b) return lx::fail(LX_EINVAL, "null batch");
    for (lx_sim *s : b->sims) {
        lx_status st = lx_sim_sync(s);
        if (st) return st;
    }
    return LX_OK;
}

}  // extern "C"
