// lx_batch.hpp -- device descriptors of the batched, temporally tiled 2D kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>

#include "lx_internal.hpp"

namespace lx {

constexpr int kTW = 48;              // tile interior (cells per side)
constexpr int kTT = 8;               // timesteps per chunk = halo width
constexpr int kRW = kTW + 2 * kTT;   // region side in LDS (64)
constexpr int kLP = kRW + 1;         // LDS row pitch
constexpr int kPSW = 24;             // max CPML slab cells of one region per axis

// One planar simulation of a batch.
struct BSim {
    int nx, nz, px;
    long long fstride;
    void *F[2];                       // field buffers (storage type), ping-pong
    float *psi[2];                    // CPML psi, ping-pong
    const uint8_t *mat;
    const float *cb;
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

template <typename T>
void launch_tile2d(const BSim *sims, const BJob *jobs, int njobs, int nt, int par, hipStream_t s);
void launch_aux_chunk(const BAux *aux, int naux, int nt, hipStream_t s);
void launch_batch_finish(const BSim *sims, int nsims, int nt, hipStream_t s);

}  // namespace lx
