// lx_batch.hpp -- device descriptors of the batched, temporally tiled 2D kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>

#include "lx_internal.hpp"

namespace lx {

// Tile geometry of the batched 2D kernel.  Measured on 256 PhC members
// (Gcell-updates/s): kTT 4/6/8/10/12/16 -> 526/690/802/686/617/507; kRW
// 4/8/16/32 rows per wavefront (7-8, 7, 4, 2 waves per SIMD) -> 519/728/
// 802/598.  Fewer rows per wavefront raise occupancy but add boundary-row
// exchanges, barriers and per-wavefront work per cell; a shallower halo adds
// region loads per step, a deeper one recomputes more halo cells.
constexpr int kTT = 8;               // timesteps per chunk = halo width
constexpr int kRW = 16;              // region rows (z) per wavefront
constexpr int kNW = 64 / kRW;        // wavefronts per block
constexpr int kRH = kRW * kNW;       // region rows (z): 64
constexpr int kTW = 64 - 2 * kTT;    // tile interior columns (x): region of 64 columns, one per lane
constexpr int kTH = kRH - 2 * kTT;   // tile interior rows (z)

// One planar simulation of a batch.
struct BSim {
    int nx, nz, px;
    long long fstride;
    void *F[2];                       // field buffers (storage type), ping-pong
    float *psi[2];                    // CPML psi, ping-pong
    const uint8_t *mat;
    const float *cb;                  // [3][256]; planar members use the Ey table
    const float *tab;
    int tn;
    int slo[3], shi[3], sw[3];
    long long scells[3];
    int pml;
    int *counter;
    int ntfH, ntfE;
    TfBox tfH[4], tfE[4];
    const double *esnap, *hsnap;      // incident line snapshots of the chunk [s][win]
    int win, inc_axis, inc_x0;
    int np;
    const int *pcomp;
    const long long *pidx;
    double *prec;
    int pcap;
    double hscale;
};

struct BJob {
    int sim, i0, k0;
};

struct BAux {
    double *e, *h;
    int L;
    double S;
    int wave;
    double c, dt, lambda0, fwhm, delay;
    double *esnap, *hsnap;
    int win;
    int *counter;
};

// One 3D simulation of a batch, advanced by the fused step (the single-
// simulation kernels' bodies over a flattened block list of every member).
template <typename T>
struct B3Sim {
    Grid<T> g;                 // F/psi: buffer set 0, Fn/psin: buffer set 1
    UnitMap um[2];             // fused step: interior box, CPML shell
    int nunits[2];
    const TfBox *tfH, *tfE;    // device face boxes
    int ntfH, ntfE, totH, totE;
    IncParams inc;
    AuxLine aux;
    ProbeArgs pr;
    int use_aux, probes;
};
constexpr int kB3Waves = 4;    // wavefronts per block of the batched fused step

// Batched fused step, part 0 (interior boxes, whole grids without CPML) or 1
// (CPML shells); bmap[b] = (member, member-local block) of launch block b;
// vec cells per lane (4 with FP16 storage only).
template <typename T>
void launch_step3_batch(const B3Sim<T> *sims, const int2 *bmap, int nblocks, int part, int par, int vec,
                        hipStream_t s);
// Batched TF/SF launch of one step: pre = 1 the H corrections on the read
// buffer (k_tfsf_hpre's work), pre = 0 the E corrections on the written
// buffer (k_tfsf<E>'s work); each member's blocks end with its extra block.
template <typename T>
void launch_tfsf_batch(const B3Sim<T> *sims, const int2 *bmap, int nblocks, int pre, int par, int toff,
                       hipStream_t s);
// Sequence end of every member after n steps: last E probes, step counter.
// One TF/SF launch between steps toff and toff+1 of every member (the E
// corrections of step toff, the H corrections of step toff+1 unless last);
// bmap lists each member's E and H correction blocks and its extra block.
template <typename T>
void launch_tfsf_between_batch(const B3Sim<T> *sims, const int2 *bmap, int nblocks, int par, int toff, int last,
                               hipStream_t s);
template <typename T>
void launch_finish_batch(const B3Sim<T> *sims, int nsims, int n, int par_end, hipStream_t s);

// Job class: kInnerJobs = every job's region lies inside the domain, outside
// the CPML slabs, on rows with one z coefficient; kInnerUnitJobs the same for
// members whose E coefficients there are exactly 1 (dz = dx: the E update's
// curl terms need no multiplication); 0..3 = the general path for regions
// that meet an x slab (bit 0) and/or a z slab (bit 1).
constexpr int kInnerJobs = 4;
constexpr int kInnerUnitJobs = 5;
constexpr int kJobClasses = 6;
template <typename T>
void launch_tile2d(const BSim *sims, const BJob *jobs, int njobs, int cls, int nt, int par, hipStream_t s);
void launch_aux_chunk(const BAux *aux, int naux, int nt, hipStream_t s);
void launch_batch_finish(const BSim *sims, int nsims, int nt, hipStream_t s);

}  // namespace lx
