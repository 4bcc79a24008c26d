// lx_batch.cpp -- batches of independent simulations.
//
// Planar members that the temporally tiled kernel supports (no point source,
// aux-1D or no plane wave, at most 256 probes) advance together: per chunk of up to 8
// steps, one launch of k_aux_chunk (incident line of every member), two
// launches of k_tile2d over the flattened tile list of every member (jobs
// whose region touches a CPML slab, then the rest), one launch of
// k_batch_finish.
// Pairs of full chunks are captured once as a hipGraph and replayed.
// 3D members that the fused step supports (no point source, no periodic
// monitor) advance together too: per step, the H TF/SF launch, the fused
// step over the interior boxes and over the CPML shells, and the E TF/SF
// launch, each over a flattened block list of every member; one sequence-end
// launch per chunk.  Full chunks are captured as hipGraphs.  Other members
// step on their own streams with the single-simulation kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/luxum.h"
#include "lx_batch.hpp"
#include "lx_sim.hpp"
#include "lx_util.hpp"

using namespace lx;

#define HIP_TRY(x)                                                                                         \
    do {                                                                                                   \
        hipError_t _e = (x);                                                                               \
        if (_e != hipSuccess) return fail(LX_EDEVICE, "%s failed: %s", #x, hipGetErrorString(_e));          \
    } while (0)

namespace {

struct Group {
    int fp16 = 0;
    std::vector<lx_sim *> sims;
    BSim *d_sims = nullptr;
    BJob *d_jobs = nullptr;
    BAux *d_aux = nullptr;
    int njobs[kJobClasses] = {};              // jobs per class, stored in class order (general 0..3, inner, inner unit)
    int naux = 0;
    int par = 0;                              // buffer index holding the current state
    hipGraphExec_t gexec[2] = {nullptr, nullptr};   // two chunks starting at parity 0 / 1
    int stable = 0;                           // stepped at least once since the last rebuild
};

// 3D members of one storage precision.
struct Group3 {
    int fp16 = 0;
    int vec = 1;                              // x cells per lane of the fused step
    std::vector<lx_sim *> sims;
    void *d_sims = nullptr;                   // B3Sim<T>[sims]
    int2 *d_map = nullptr;                    // block maps: interior, shell, H TF/SF, E TF/SF, between steps,
                                              // between the chunk's last step and the next chunk
    int nblk[6] = {0, 0, 0, 0, 0, 0};
    int off[6] = {0, 0, 0, 0, 0, 0};          // first entry of each map in d_map
    int merge = 0;                            // one TF/SF launch between steps (every member allows it)
    int par = 0;                              // buffer set holding the current state
    hipGraphExec_t gexec[2] = {nullptr, nullptr};   // one chunk starting at parity 0 / 1
    int stable = 0;
};

}  // namespace

struct lx_batch {
    lx_ctx *ctx = nullptr;
    std::vector<lx_sim *> sims;
    hipStream_t stream = nullptr, side = nullptr;   // general-path jobs run on `side`, concurrently
    hipEvent_t done = nullptr, pre = nullptr, fork = nullptr, join = nullptr;
    bool dirty = true;
    Group grp[2];                   // planar members, FP32 and FP16 storage
    Group3 grp3[2];                 // 3D members, FP32 and FP16 storage
    std::vector<lx_sim *> others;
    std::vector<void *> allocs;     // snapshot buffers
    // descriptor arrays per precision (planar sims, jobs, aux lines; 3D sims,
    // block maps), kept across rebuilds and grown when a membership needs more
    void *desc[2][5] = {};
    size_t desc_cap[2][5] = {};
    int graphs = 1;
    int shell_side = 1;             // 3D: shell launch on `side` beside the interior launch (LX_SHELL_STREAM)
};

namespace {

bool tiled_ok(const lx_sim *s) {
    const lx_params &p = s->p;
    if (!p.is_planar || p.use_ps) return false;
    if (p.use_pw && p.incident != LX_INC_AUX1D) return false;
    if (s->np > 256) return false;
    if (s->tfH.n > 4 || s->tfE.n > 4) return false;
    return true;
}

// Class of the job starting at (i0, k0): kInnerJobs when its region lies
// inside [1, N-2] on both axes and inside the uniform-coefficient range (no
// slab, one z coefficient); otherwise bit 0 = the region meets an x slab,
// bit 1 = it meets a z slab.  Every x or z index with a nonzero CPML c lies
// in a slab, and the kernel's coefficients outside the domain are zero or
// taken from its edge index, so a class without a bit has zero psi there.
int job_class(const lx_sim *s, int i0, int k0) {
    const int x0 = i0 - kTT, x1 = x0 + 64, z0 = k0 - kTT, z1 = z0 + kRH;
    if (x0 >= std::max(1, s->ulo[0]) && x1 <= std::min(s->N[0] - 1, s->uhi[0]) && z0 >= std::max(1, s->ulo[2]) &&
        z1 <= std::min(s->N[2] - 1, s->uhi[2]))
        return s->unit_ke ? kInnerUnitJobs : kInnerJobs;
    if (!s->pml) return 0;
    return (x0 < s->slo[0] || x1 > s->shi[0] ? 1 : 0) | (z0 < s->slo[2] || z1 > s->shi[2] ? 2 : 0);
}

void free_group(Group &g) {
    for (hipGraphExec_t &x : g.gexec) {
        if (x) (void)hipGraphExecDestroy(x);
        x = nullptr;
    }
    g.par = 0;
    g.stable = 0;
    g.sims.clear();
    g.d_sims = nullptr;
    g.d_jobs = nullptr;
    g.d_aux = nullptr;
    for (int &n : g.njobs) n = 0;
    g.naux = 0;
}

void free_group3(Group3 &g) {
    for (hipGraphExec_t &x : g.gexec) {
        if (x) (void)hipGraphExecDestroy(x);
        x = nullptr;
    }
    g.par = 0;
    g.stable = 0;
    g.sims.clear();
    g.d_sims = nullptr;
    g.d_map = nullptr;
    for (int k = 0; k < 6; ++k) g.nblk[k] = g.off[k] = 0;
    g.merge = 0;
}

bool batch3_ok(const lx_sim *s) {
    return !s->planar && fused_ok(s) && !s->p.use_ps && s->p.monitor_every <= 0;
}

void free_allocs(lx_batch *b) {
    for (void *p : b->allocs) (void)hipFree(p);
    b->allocs.clear();
}

// Descriptor array `kind` of group `gi`, at least `bytes` long (the batch
// stream is idle when this runs: rebuild synchronises it first).
lx_status desc_buf(lx_batch *b, int gi, int kind, size_t bytes, void **out) {
    void *&p = b->desc[gi][kind];
    size_t &cap = b->desc_cap[gi][kind];
    if (cap < bytes) {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 2, 4096);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fail(LX_ENOMEM, "batch allocation of %zu bytes failed", want);
        }
        cap = want;
    }
    *out = p;
    return LX_OK;
}

lx_status ensure_alt(lx_sim *s) {
    if (s->F_alt) return LX_OK;
    void *p = nullptr;
    const size_t fbytes = (size_t)s->fstride * 6 * s->esize;
    lx_status st = dalloc(s, &p, fbytes);
    if (st) return st;
    launch_zero_bytes(p, fbytes, s->cs);
    s->F_alt = p;
    st = dalloc(s, &p, psi_total(s) * sizeof(float));
    if (st) return st;
    launch_zero_bytes(p, psi_total(s) * sizeof(float), s->cs);
    s->psi_alt = (float *)p;
    s->gen++;
    HIP_TRY(hipStreamSynchronize(s->cs));
    return LX_OK;
}

// Descriptors and block maps of the 3D members (large members first: their
// long-running blocks start early, small ones fill the tail).
template <typename T>
lx_status build_group3(lx_batch *b, Group3 &g) {
    const size_t n = g.sims.size();
    // the widest cells-per-lane width every member's row pitch allows (FP16:
    // 4 or 2; FP32: 1, the faster one at every size measured)
    g.vec = 1;
    if (g.fp16)
        for (int v : {4, 2}) {
            bool ok = true;
            for (const lx_sim *s : g.sims) ok = ok && vec_ok(s, v);
            if (ok) {
                g.vec = v;
                break;
            }
        }
    std::vector<int> order(n);
    for (size_t q = 0; q < n; ++q) order[q] = (int)q;
    auto cells = [&](int q) { return (int64_t)g.sims[q]->N[0] * g.sims[q]->N[1] * g.sims[q]->N[2]; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cells(x) > cells(y); });
    std::vector<B3Sim<T>> hs(n);
    std::vector<int2> map[6];
    g.merge = 1;
    for (const lx_sim *s : g.sims) g.merge = g.merge && tf_merge_ok(s);
    for (int q : order) {
        lx_sim *s = g.sims[q];
        B3Sim<T> &d = hs[q];
        std::memset(&d, 0, sizeof(d));
        d.g = make_grid<T>(s);
        d.g.F = (T *)s->F;
        d.g.Fn = (T *)s->F_alt;
        d.g.psi = s->psi;
        d.g.psin = s->psi_alt;
        if (!s->fused) d.g.kz = 8;   // the half-step kernels' z-block depth does not apply
        step3_unit_maps(s->N[0], s->N[1], s->N[2], d.g.kz, s->fused ? s->kzs : 0, s->pml, s->slo, s->shi, g.vec,
                        d.um, d.nunits);
        d.tfH = s->tfH.dev;
        d.tfE = s->tfE.dev;
        d.ntfH = s->tfH.n;
        d.ntfE = s->tfE.n;
        d.totH = s->tfH.total;
        d.totE = s->tfE.total;
        d.inc = s->inc;
        d.aux = s->aux;
        d.pr = probe_args(s);
        d.use_aux = s->use_aux;
        d.probes = s->np > 0;
        const int extra = (d.use_aux || d.probes) ? 1 : 0;
        const int bx = between_extra_blocks(d.use_aux ? d.aux.L : 0, d.probes);
        const int nb[6] = {(d.nunits[0] + kB3Waves - 1) / kB3Waves, (d.nunits[1] + kB3Waves - 1) / kB3Waves,
                           (d.totH + 255) / 256 + extra, (d.totE + 255) / 256 + extra,
                           (d.totE + 255) / 256 + (d.totH + 255) / 256 + bx, (d.totE + 255) / 256 + bx};
        for (int k = 0; k < 6; ++k)
            for (int j = 0; j < nb[k]; ++j) map[k].push_back(make_int2(q, j));
    }
    size_t total = 0;
    for (int k = 0; k < 6; ++k) {
        g.off[k] = (int)total;
        g.nblk[k] = (int)map[k].size();
        total += map[k].size();
    }
    const int gi = g.fp16;
    void *p = nullptr;
    lx_status st = desc_buf(b, gi, 3, sizeof(B3Sim<T>) * n, &p);
    if (st) return st;
    g.d_sims = p;
    HIP_TRY(hipMemcpy(p, hs.data(), sizeof(B3Sim<T>) * n, hipMemcpyHostToDevice));
    st = desc_buf(b, gi, 4, sizeof(int2) * std::max<size_t>(total, 1), &p);
    if (st) return st;
    g.d_map = (int2 *)p;
    for (int k = 0; k < 6; ++k)
        if (!map[k].empty())
            HIP_TRY(hipMemcpy(g.d_map + g.off[k], map[k].data(), sizeof(int2) * map[k].size(), hipMemcpyHostToDevice));
    return LX_OK;
}

lx_status rebuild(lx_batch *b) {
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (Group &g : b->grp) free_group(g);
    for (Group3 &g : b->grp3) free_group3(g);
    free_allocs(b);
    b->others.clear();
    b->grp[0].fp16 = 0;
    b->grp[1].fp16 = 1;
    b->grp3[0].fp16 = 0;
    b->grp3[1].fp16 = 1;
    for (lx_sim *s : b->sims) {
        if (tiled_ok(s)) {
            b->grp[s->fp16].sims.push_back(s);
        } else if (batch3_ok(s)) {
            lx_status st = ensure_fused_buffers(s);
            if (st) return st;
            b->grp3[s->fp16].sims.push_back(s);
        } else {
            b->others.push_back(s);
        }
    }
    for (Group3 &g : b->grp3) {
        if (g.sims.empty()) continue;
        lx_status st = g.fp16 ? build_group3<__half>(b, g) : build_group3<float>(b, g);
        if (st) return st;
    }
    for (Group &g : b->grp) {
        if (g.sims.empty()) continue;
        std::vector<BSim> hs(g.sims.size());
        std::vector<BJob> cjobs[kJobClasses];
        std::vector<BAux> aux;
        for (size_t q = 0; q < g.sims.size(); ++q) {
            lx_sim *s = g.sims[q];
            lx_status st = ensure_alt(s);
            if (st) return st;
            BSim &d = hs[q];
            std::memset(&d, 0, sizeof(d));
            d.nx = s->N[0];
            d.nz = s->N[2];
            d.px = s->px;
            d.fstride = s->fstride;
            d.F[0] = s->F;
            d.F[1] = s->F_alt;
            d.psi[0] = s->psi;
            d.psi[1] = s->psi_alt;
            d.mat = s->mat;
            d.cb = s->cb;
            d.tab = s->tab;
            d.tn = s->tn;
            for (int a = 0; a < 3; ++a) {
                d.slo[a] = s->slo[a];
                d.shi[a] = s->shi[a];
                d.sw[a] = s->sw[a];
                d.scells[a] = s->slab_cells[a];
            }
            d.pml = s->pml;
            d.counter = s->counter;
            d.ntfH = s->tfH.n;
            d.ntfE = s->tfE.n;
            for (int k = 0; k < s->tfH.n; ++k) d.tfH[k] = s->tfH.b[k];
            for (int k = 0; k < s->tfE.n; ++k) d.tfE[k] = s->tfE.b[k];
            if (s->use_aux) {
                const int ax = s->inc.axis;
                const int win = std::min(s->aux.L, s->p.tf_hi[ax] + 2 - s->inc.x0 + 1);
                if (!s->bsnap || s->bsnap_win != win) {
                    void *snap = nullptr;
                    st = dalloc(s, &snap, sizeof(double) * 2 * kTT * (size_t)win);
                    if (st) return st;
                    s->bsnap = (double *)snap;
                    s->bsnap_win = win;
                }
                d.esnap = s->bsnap;
                d.hsnap = s->bsnap + (size_t)kTT * win;
                d.win = win;
                d.inc_axis = ax;
                d.inc_x0 = s->inc.x0;
                BAux a{};
                a.e = s->aux.e;
                a.h = s->aux.h;
                a.L = s->aux.L;
                a.S = s->aux.S;
                a.wave = s->aux.wave;
                a.c = s->aux.c;
                a.dt = s->aux.dt;
                a.lambda0 = s->aux.lambda0;
                a.fwhm = s->aux.fwhm;
                a.delay = s->aux.delay;
                a.esnap = (double *)d.esnap;
                a.hsnap = (double *)d.hsnap;
                a.win = win;
                a.counter = s->counter;
                aux.push_back(a);
            } else {
                d.ntfH = d.ntfE = 0;
            }
            d.np = s->np;
            d.pcomp = s->pcomp;
            d.pidx = s->pidx;
            d.prec = s->prec;
            d.pcap = s->pcap;
            d.hscale = 1.0 / s->eta0;
        }
        // large simulations first: long jobs start early, small ones fill the
        // tail (members ordered by size, each member's tiles in order)
        std::vector<int> order(g.sims.size());
        for (size_t q = 0; q < order.size(); ++q) order[q] = (int)q;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return (int64_t)g.sims[x]->N[0] * g.sims[x]->N[2] > (int64_t)g.sims[y]->N[0] * g.sims[y]->N[2];
        });
        for (int q : order) {
            const lx_sim *s = g.sims[q];
            for (int tz = 0; tz < (s->N[2] + kTH - 1) / kTH; ++tz)
                for (int tx = 0; tx < (s->N[0] + kTW - 1) / kTW; ++tx) {
                    const BJob j{q, tx * kTW, tz * kTH};
                    cjobs[job_class(s, j.i0, j.k0)].push_back(j);
                }
        }
        std::vector<BJob> jobs;
        for (int c = 0; c < kJobClasses; ++c) {
            g.njobs[c] = (int)cjobs[c].size();
            jobs.insert(jobs.end(), cjobs[c].begin(), cjobs[c].end());
        }
        void *p = nullptr;
        const int gi = g.fp16;
        lx_status st = desc_buf(b, gi, 0, sizeof(BSim) * hs.size(), &p);
        if (st) return st;
        g.d_sims = (BSim *)p;
        HIP_TRY(hipMemcpy(g.d_sims, hs.data(), sizeof(BSim) * hs.size(), hipMemcpyHostToDevice));
        st = desc_buf(b, gi, 1, sizeof(BJob) * jobs.size(), &p);
        if (st) return st;
        g.d_jobs = (BJob *)p;
        HIP_TRY(hipMemcpy(g.d_jobs, jobs.data(), sizeof(BJob) * jobs.size(), hipMemcpyHostToDevice));
        if (!aux.empty()) {
            st = desc_buf(b, gi, 2, sizeof(BAux) * aux.size(), &p);
            if (st) return st;
            g.d_aux = (BAux *)p;
            HIP_TRY(hipMemcpy(g.d_aux, aux.data(), sizeof(BAux) * aux.size(), hipMemcpyHostToDevice));
            g.naux = (int)aux.size();
        }
    }
    for (lx_sim *s : b->sims) s->batch_gen = s->gen;
    b->dirty = false;
    return LX_OK;
}

// One chunk: incident lines, then the tile jobs on two streams at once: the
// inner jobs and then the corner class (x and z slabs, the fewest jobs) on
// the main stream, the other general classes on the side stream; then the
// step counters.
void enqueue_chunk(const lx_batch *b, const Group &g, int nt, int par) {
    hipStream_t s = b->stream;
    if (g.naux) launch_aux_chunk(g.d_aux, g.naux, nt, s);
    const bool fork =
        (g.njobs[0] || g.njobs[1] || g.njobs[2]) && (g.njobs[3] || g.njobs[kInnerJobs] || g.njobs[kInnerUnitJobs]);
    if (fork) {
        (void)hipEventRecord(b->fork, s);
        (void)hipStreamWaitEvent(b->side, b->fork, 0);
    }
    hipStream_t so = fork ? b->side : s;
    int off[kJobClasses + 1] = {0};
    for (int c = 0; c < kJobClasses; ++c) off[c + 1] = off[c] + g.njobs[c];
    static const int order[kJobClasses] = {kInnerUnitJobs, kInnerJobs, 3, 0, 1, 2};
    for (int c : order) {
        hipStream_t sc = (c == kInnerJobs || c == kInnerUnitJobs || c == 3) ? s : so;
        if (g.fp16) launch_tile2d<__half>(g.d_sims, g.d_jobs + off[c], g.njobs[c], c, nt, par, sc);
        else launch_tile2d<float>(g.d_sims, g.d_jobs + off[c], g.njobs[c], c, nt, par, sc);
    }
    if (fork) {
        (void)hipEventRecord(b->join, b->side);
        (void)hipStreamWaitEvent(s, b->join, 0);
    }
    launch_batch_finish(g.d_sims, (int)g.sims.size(), nt, s);
}

// nt steps of the 3D members from parity par: per step the H TF/SF
// corrections on the read buffers, the fused step (interior boxes, then
// shells), the E TF/SF corrections on the written buffers; then the
// sequence end.
template <typename T>
void enqueue3_t(const lx_batch *b, const Group3 &g, int nt, int par) {
    hipStream_t s = b->stream;
    const B3Sim<T> *sims = (const B3Sim<T> *)g.d_sims;
    // the interior boxes and the CPML shells read one buffer set and write
    // disjoint cells of the other: the shells may run beside on `side`
    const bool fork = b->shell_side && g.nblk[0] && g.nblk[1];
    for (int q = 0; q < nt; ++q) {
        const int pq = (par + q) & 1;
        if (!g.merge || q == 0) launch_tfsf_batch<T>(sims, g.d_map + g.off[2], g.nblk[2], 1, pq, q, s);
        if (fork) {
            (void)hipEventRecord(b->fork, s);
            (void)hipStreamWaitEvent(b->side, b->fork, 0);
            launch_step3_batch<T>(sims, g.d_map + g.off[1], g.nblk[1], 1, pq, g.vec, b->side);
            launch_step3_batch<T>(sims, g.d_map + g.off[0], g.nblk[0], 0, pq, g.vec, s);
            (void)hipEventRecord(b->join, b->side);
            (void)hipStreamWaitEvent(s, b->join, 0);
        } else {
            launch_step3_batch<T>(sims, g.d_map + g.off[0], g.nblk[0], 0, pq, g.vec, s);
            launch_step3_batch<T>(sims, g.d_map + g.off[1], g.nblk[1], 1, pq, g.vec, s);
        }
        if (!g.merge) {
            launch_tfsf_batch<T>(sims, g.d_map + g.off[3], g.nblk[3], 0, pq, q, s);
        } else if (q == nt - 1) {
            launch_tfsf_between_batch<T>(sims, g.d_map + g.off[5], g.nblk[5], pq, q, 1, s);
        } else {
            launch_tfsf_between_batch<T>(sims, g.d_map + g.off[4], g.nblk[4], pq, q, 0, s);
        }
    }
    launch_finish_batch<T>(sims, (int)g.sims.size(), nt, (par + nt) & 1, s);
}

void enqueue3(const lx_batch *b, const Group3 &g, int nt, int par) {
    if (g.fp16) enqueue3_t<__half>(b, g, nt, par);
    else enqueue3_t<float>(b, g, nt, par);
}

lx_status ensure_graph3(lx_batch *b, Group3 &g, int par) {
    if (g.gexec[par] || !b->graphs) return LX_OK;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue3(b, g, kTT, par);
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess || !graph) {
        (void)hipGetLastError();
        b->graphs = 0;
        return LX_OK;
    }
    e = hipGraphInstantiate(&g.gexec[par], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g.gexec[par] = nullptr;
        b->graphs = 0;
    }
    return LX_OK;
}

lx_status ensure_pair_graph(lx_batch *b, Group &g, int par) {
    if (g.gexec[par] || !b->graphs) return LX_OK;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_chunk(b, g, kTT, par);
    enqueue_chunk(b, g, kTT, par ^ 1);
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess || !graph) {
        (void)hipGetLastError();
        b->graphs = 0;
        return LX_OK;
    }
    e = hipGraphInstantiate(&g.gexec[par], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g.gexec[par] = nullptr;
        b->graphs = 0;
    }
    return LX_OK;
}

}  // namespace

extern "C" {

lx_status lx_batch_create(lx_ctx *ctx, lx_batch **out) {
    if (!ctx || !out) return fail(LX_EINVAL, "lx_batch_create: null pointer");
    lx_batch *b = new (std::nothrow) lx_batch;
    if (!b) return fail(LX_ENOMEM, "out of host memory");
    b->ctx = ctx;
    const char *ks = std::getenv("LX_SHELL_STREAM");
    b->shell_side = !(ks && *ks == '0');
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->pre, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        lx_batch_destroy(b);
        return fail(LX_EDEVICE, "batch stream creation failed");
    }
    *out = b;
    return LX_OK;
}

void lx_batch_destroy(lx_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (Group &g : b->grp) free_group(g);
    for (Group3 &g : b->grp3) free_group3(g);
    free_allocs(b);
    for (auto &row : b->desc)
        for (void *p : row)
            if (p) (void)hipFree(p);
    // the batch's last step is complete (stream synchronised above): members
    // no longer wait for it, and its event is about to be destroyed
    for (lx_sim *s : b->sims) {
        s->batch = nullptr;
        s->batch_wait = nullptr;
    }
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->pre) (void)hipEventDestroy(b->pre);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

lx_status lx_batch_add(lx_batch *b, lx_sim *s) {
    if (!b || !s) return fail(LX_EINVAL, "lx_batch_add: null pointer");
    if (s->batch) return fail(LX_EINVAL, "simulation already belongs to a batch");
    if (s->ctx != b->ctx) return fail(LX_EINVAL, "simulation and batch use different contexts");
    b->sims.push_back(s);
    s->batch = b;
    b->dirty = true;
    return LX_OK;
}

lx_status lx_batch_remove(lx_batch *b, lx_sim *s) {
    if (!b || !s) return fail(LX_EINVAL, "lx_batch_remove: null pointer");
    auto it = std::find(b->sims.begin(), b->sims.end(), s);
    if (it == b->sims.end()) return fail(LX_EINVAL, "not in batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->sims.erase(it);
    s->batch = nullptr;
    s->batch_wait = nullptr;   // the batch's steps so far are complete; later ones are not the member's
    b->dirty = true;
    return LX_OK;
}

lx_status lx_batch_step(lx_batch *b, int32_t nsteps) {
    if (!b || nsteps < 0) return fail(LX_EINVAL, "lx_batch_step: bad arguments");
    if (nsteps == 0 || b->sims.empty()) return LX_OK;
    HIP_TRY(hipSetDevice(b->ctx->device));
    // probe buffers must hold the new steps before the descriptors are built;
    // descriptors are rebuilt when a member's buffers or settings changed
    // since the last rebuild (its own steps, probes, autotuning)
    for (lx_sim *s : b->sims) {
        lx_status st = ensure_probe_cap(s, s->steps + nsteps);
        if (st) return st;
        if (s->gen != s->batch_gen) b->dirty = true;
    }
    if (b->dirty) {
        lx_status st = rebuild(b);
        if (st) return st;
    }
    // order after each member's pending work (members whose stream is idle
    // need no event: a served batch steps right after a sync)
    std::vector<lx_sim *> grouped;
    for (Group &g : b->grp) grouped.insert(grouped.end(), g.sims.begin(), g.sims.end());
    for (Group3 &g : b->grp3) grouped.insert(grouped.end(), g.sims.begin(), g.sims.end());
    for (lx_sim *s : grouped) {
        if (s->mguard) {   // a started monitor still reads the member's fields
            HIP_TRY(hipStreamWaitEvent(b->stream, s->mdone, 0));
            s->mguard = nullptr;
        }
        if (!s->cs_dirty) continue;
        s->cs_dirty = false;
        const hipError_t q = hipStreamQuery(s->cs);
        if (q == hipSuccess) continue;
        if (q != hipErrorNotReady) HIP_TRY(q);
        HIP_TRY(hipEventRecord(b->pre, s->cs));
        HIP_TRY(hipStreamWaitEvent(b->stream, b->pre, 0));
    }
    for (Group &g : b->grp) {
        if (g.sims.empty()) continue;
        int par = g.par;
        int full = nsteps / kTT;
        const int rem = nsteps % kTT;
        // a graph pays for its capture only when it is replayed: capture once
        // the membership has survived a step, or for a long first step (a
        // served batch changes members at most ticks and launches directly)
        if (full >= 2 && (g.stable || full >= 32)) {
            lx_status st = ensure_pair_graph(b, g, par);
            if (st) return st;
            if (g.gexec[par])
                for (; full >= 2; full -= 2) HIP_TRY(hipGraphLaunch(g.gexec[par], b->stream));
        }
        for (; full > 0; --full) {
            enqueue_chunk(b, g, kTT, par);
            par ^= 1;
        }
        if (rem) {
            enqueue_chunk(b, g, rem, par);
            par ^= 1;
        }
        lx_status st = check_launch();
        if (st) return st;
        if (par != g.par) {
            // the current state moved to the other buffer of each member
            for (lx_sim *s : g.sims) {
                std::swap(s->F, s->F_alt);
                std::swap(s->psi, s->psi_alt);
                drop_graph(s);
            }
            g.par = par;
        }
        for (lx_sim *s : g.sims) s->steps += nsteps;
        g.stable = 1;
    }
    for (Group3 &g : b->grp3) {
        if (g.sims.empty()) continue;
        int par = g.par;
        int full = nsteps / kTT;
        const int rem = nsteps % kTT;
        // a chunk of kTT (even) steps keeps the parity: one graph per parity
        if (full >= 1 && (g.stable || full >= 16)) {
            lx_status st = ensure_graph3(b, g, par);
            if (st) return st;
            if (g.gexec[par])
                for (; full >= 1; --full) HIP_TRY(hipGraphLaunch(g.gexec[par], b->stream));
        }
        for (; full > 0; --full) enqueue3(b, g, kTT, par);
        if (rem) {
            enqueue3(b, g, rem, par);
            par ^= rem & 1;
        }
        lx_status st = check_launch();
        if (st) return st;
        if (par != g.par) {
            for (lx_sim *s : g.sims) {
                std::swap(s->F, s->F_alt);
                std::swap(s->psi, s->psi_alt);
                drop_graph(s);
            }
            g.par = par;
        }
        for (lx_sim *s : g.sims) s->steps += nsteps;
        g.stable = 1;
    }
    HIP_TRY(hipEventRecord(b->done, b->stream));
    for (lx_sim *s : grouped) s->batch_wait = b->done;   // the member's own next work waits for this step
    for (lx_sim *s : b->others) {
        lx_status st = lx_sim_step(s, nsteps);
        if (st) return st;
    }
    return LX_OK;
}

lx_status lx_batch_sync(lx_batch *b) {
    if (!b) return fail(LX_EINVAL, "null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (lx_sim *s : b->sims) {
        s->batch_wait = nullptr;   // that step is complete
        if (!s->cs_dirty && !s->ss && dma_idle(&s->snap_dma[0]) && dma_idle(&s->snap_dma[1])) continue;
        lx_status st = lx_sim_sync(s);
        if (st) return st;
    }
    return LX_OK;
}

}  // extern "C"
