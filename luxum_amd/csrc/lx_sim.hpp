// lx_sim.hpp -- host-side state of a simulation (shared by lx_api.cpp and lx_batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/luxum.h"
#include "lx_dma.hpp"
#include "lx_internal.hpp"

struct lx_ctx {
    int device;
    int64_t total_mem;
    std::mutex mu;
    std::vector<hipStream_t> streams;   // idle streams of destroyed simulations
    // device blocks of destroyed simulations, by size, reused by exact size
    // (a server creates and destroys simulations of a few grid sizes)
    std::multimap<size_t, void *> blocks;
    int64_t cached = 0, cache_limit = 0;
};

struct lx_sim {
    // batch membership: alternate field/psi buffers for the ping-pong tiles
    void *F_alt = nullptr;
    float *psi_alt = nullptr;
    double *bsnap = nullptr;    // incident line snapshots of one chunk: E then H, [8][win] each
    int bsnap_win = 0;
    lx_batch *batch = nullptr;
    // ordering with a batch without per-member stream operations: the batch
    // waits for a member's stream only if the member enqueued work of its own
    // (cs_dirty); a member's own work waits for the batch's last step
    // (batch_wait, set by the batch) only when there is such work
    hipEvent_t batch_wait = nullptr;
    bool cs_dirty = false;
    // bumped whenever device buffers or step settings that a batch's
    // descriptors copy change (probe buffers, the buffer pair's order, the
    // second buffer set, the autotuned settings); a batch rebuilds its
    // descriptors when a member's gen differs from the one it recorded
    uint32_t gen = 0, batch_gen = 0;
    lx_ctx *ctx = nullptr;
    lx_params p{};
    int planar = 0, fp16 = 0, pml = 0;
    int N[3] = {0, 0, 0};
    int px = 0;
    long long plane = 0;
    size_t esize = 4;
    // constants (maxwell.cc:63-84)
    double eps0 = 0, mu0 = 0, c0 = 0, dt = 0, eta0 = 0, sigma_max = 0;
    // fields are stored times sscale (a power of two): sources inject sscale*amplitude, outputs divide it out
    double sscale = 1;
    hipStream_t cs = nullptr, ss = nullptr;
    // fused step: the CPML shell launch runs on ks beside the interior launch
    // on cs (both read one buffer set and write disjoint cells of the other)
    hipStream_t ks = nullptr;
    hipEvent_t kfork = nullptr, kjoin = nullptr;
    int kconc = 1;              // use ks (when it exists); autotuned
    int tf_merge = 1;           // fused step: one TF/SF launch between steps (LX_TFSF_MERGE=0: two per step)
    int rows = 0;               // fused step: rows per block of the interior launch sharing neighbours
                                // through LDS (3 or 7; 0: one unit per wavefront); autotuned
    int shell_pipe = 1;         // fused step: shell launch pipelined past the TF/SF launch (LX_SHELL_PIPE=0: not)
    std::vector<std::pair<void *, size_t>> allocs;   // device blocks and their sizes
    int64_t bytes = 0;
    void *F = nullptr;          // six components, component c at F + c*fstride elements
    long long fstride = 0;
    uint8_t *mat = nullptr;
    float *cb = nullptr;
    float *tab = nullptr;
    float *psi = nullptr;
    int *counter = nullptr;
    int tn = 0;
    int slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0}, sw[3] = {0, 0, 0};
    int ulo[3] = {0, 0, 0}, uhi[3] = {0, 0, 0};   // uniform-coefficient range per axis, inside [slo, shi)
    int unit_ke = 0;            // the E coefficients of x and z in that range are exactly 1 (batched 2D tiles)
    int64_t slab_cells[3] = {0, 0, 0};
    lx::TfList tfH{}, tfE{};
    int kz = 32;
    int kzs = 0;                // fused step: z-block depth of the CPML shell launch (0: kz)
    int fused = 0;              // 3D: one fused launch per step over ping-pong buffers (F/psi and F_alt/psi_alt)
    int nw = 8;                 // fused step: wavefronts per block
    int vec = 1;                // fused step: x cells per lane (1, 2 or 4)
    lx::IncParams inc{};
    lx::AuxLine aux{};
    int use_aux = 0;
    lx::PointSrc ps{};
    // probes
    int np = 0, pcap = 0;
    long long *pidx = nullptr;
    int *pcomp = nullptr;
    double *prec = nullptr;
    std::vector<int32_t> probe_host;
    // snapshots
    double *stage[2] = {nullptr, nullptr};
    size_t stage_n[2] = {0, 0};
    hipEvent_t ready[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
    lx::DmaSlot snap_dma[2];    // copies into host memory allocated by the runtime (lx_dma.hpp)
    int snap_next = 0;
    double *rd_stage = nullptr;
    size_t rd_n = 0;
    double *mon = nullptr;      // monitor scratch: the synchronous monitor's, then the started one's
    // the E-only monitor started by lx_sim_max_e2_start: on stream ms after
    // event mstart (recorded on cs), complete at event mdone; its result
    // triple lands in pinned host memory mres.  mguard is the field buffer it
    // reads, while work on cs that may write that buffer must still wait for
    // mdone (nullptr: nothing to wait for).
    hipStream_t ms = nullptr;
    hipEvent_t mstart = nullptr, mdone = nullptr;
    double *mres = nullptr;
    int mpend = 0;
    const void *mguard = nullptr;
    // graphs of kGraphSteps[i]-step sequences, one per buffer state of the
    // fused step (keyed by the field buffer read first, gF), kept while the
    // fused buffer pair swaps after odd step counts
    hipGraphExec_t gexec[6][2] = {};
    const void *gF[6][2] = {};
    int graphs = 1;
    int steps = 0;
    float cb_host[3 * 256];   // Ex, Ey, Ez coefficient tables
};


namespace lx {
// helpers implemented in lx_api.cpp
template <typename T>
Grid<T> make_grid(const lx_sim *s);
lx_status ensure_probe_cap(lx_sim *s, int need);
lx_status check_launch();
lx_status dalloc(lx_sim *s, void **out, size_t bytes);
size_t psi_total(const lx_sim *s);
ProbeArgs probe_args(const lx_sim *s);
void drop_graph(lx_sim *s);
bool fused_ok(const lx_sim *s);
bool vec_ok(const lx_sim *s, int vec);
// fused sequences may apply one TF/SF launch between steps (no H probe on an H face cell)
bool tf_merge_ok(const lx_sim *s);
// the CPML shell launch may overlap the TF/SF launch before it (disjoint cells)
bool shell_pipe_ok(const lx_sim *s);
// The side stream (snapshot copies) is taken from the context's pool on
// first use.
lx_status ensure_side_stream(lx_sim *s);
// Called by every simulation entry point before it enqueues work on s->cs.
void member_enter(lx_sim *s);
// Fill the context's stream pool up to n idle streams (servers call it once,
// so that admitting a request does not create streams).
void prewarm_streams(lx_ctx *c, int n);
lx_status ensure_fused_buffers(lx_sim *s);
// enqueue on the compute stream: one component (what == LX_OUT_FIELD) or
// |E|^2 (LX_OUT_E2) into dst_dev in the FP64 host layout of extents ext
lx_status gather_output(lx_sim *s, int what, int comp, double *dst_dev, const int ext[3]);
void release_fused_buffers(lx_sim *s);
// Look up the device's allocation-free readback scratch (lx_noalloc.hip);
// called with the device current.
lx_status noalloc_init(int device);
}  // namespace lx
