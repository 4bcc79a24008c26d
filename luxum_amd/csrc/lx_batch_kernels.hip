// lx_batch_kernels.hip -- batched 2D (x-z plane, Ey Hx Hz) timesteps with
// temporal tiling.
//
// Many planar simulations of uneven sizes advance together: one launch covers
// a flattened list of tile jobs from every simulation.  A job is a 48 x 48
// tile of one simulation; its block loads the tile plus an 8-cell halo
// (a 64 x 64 region of Ey, Hx, Hz and the cell's material coefficient) into
// registers, advances up to 8 timesteps there, and writes the 48 x 48
// interior back.  After s steps the cells within s of the region edge are
// stale (the Yee dependency cone grows one cell per half step), so the
// interior is exact after 8.  Neighbouring jobs read the halo of step t while
// others write step t+8, so fields and psi ping-pong between two buffers.
//
// Per step the block does what the single-simulation kernels do
// (maxwell_two_dim_pml.cc:144-176): H update (+psi), TF/SF face corrections
// of the H half step, E update (+psi), TF/SF corrections of the E half step,
// probe records.  The incident values of the aux-1D line for the chunk's 8
// steps are precomputed per simulation by k_aux_chunk.
//
// HBM traffic per cell-step is (64^2*13 + 48^2*12)/(48^2*8) = 4.4 bytes
// (FP32), against 37 for the half-step kernels.
#include "lx_batch.hpp"
#include "../../include/luxum.h"

namespace lx {

template <typename T>
__device__ __forceinline__ float bld(const T *p);
template <>
__device__ __forceinline__ float bld<float>(const float *p) { return *p; }
template <>
__device__ __forceinline__ float bld<__half>(const __half *p) { return __half2float(*p); }
// the interior a job writes back is read only by the next chunk: non-temporal stores
template <typename T>
__device__ __forceinline__ void bst(T *p, float v);
template <>
__device__ __forceinline__ void bst<float>(float *p, float v) { __builtin_nontemporal_store(v, p); }
template <>
__device__ __forceinline__ void bst<__half>(__half *p, float v) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half(v)), reinterpret_cast<unsigned short *>(p));
}

__device__ __forceinline__ int bslab(int x, int lo, int hi) { return x < lo ? x : lo + (x - hi); }

__device__ double bwave(int kind, double t, double c, double lambda0, double fwhm, double delay) {
    const double f0 = c / lambda0, T0 = 1.0 / f0;
    const double tau = fwhm * T0 / (2.0 * sqrt(log(2.0)));
    const double t0 = delay * tau;
    const double x = (t - t0) / tau;
    const double env = exp(-x * x);
    if (kind == LX_WAVE_GAUSSIAN) return env;
    return env * sin(2.0 * M_PI * f0 * (t - t0));
}

// One block per simulation: advance its aux line `nt` steps and keep, for
// every step s, the line E before the H half step (esnap) and the line H
// before the E half step (hsnap).
__global__ __launch_bounds__(256) void k_aux_chunk(const BAux *__restrict__ aux, int nt) {
    const BAux a = aux[blockIdx.x];
    const int t0 = a.counter[0];
    for (int s = 0; s < nt; ++s) {
        for (int m = threadIdx.x; m < a.win; m += blockDim.x) a.esnap[(long long)s * a.win + m] = a.e[m];
        __syncthreads();
        for (int m = threadIdx.x; m < a.L - 1; m += blockDim.x) a.h[m] = a.h[m] - a.S * (a.e[m + 1] - a.e[m]);
        __syncthreads();
        for (int m = threadIdx.x; m < a.win; m += blockDim.x) a.hsnap[(long long)s * a.win + m] = a.h[m];
        __syncthreads();
        for (int m = 1 + threadIdx.x; m < a.L - 1; m += blockDim.x) a.e[m] = a.e[m] - a.S * (a.h[m] - a.h[m - 1]);
        if (threadIdx.x == 0) {
            a.e[0] = bwave(a.wave, (t0 + s + 1) * a.dt, a.c, a.lambda0, a.fwhm, a.delay);
            a.e[a.L - 1] = 0.0;
        }
        __syncthreads();
    }
}

__global__ void k_batch_finish(const BSim *__restrict__ sims, int nsims, int nt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nsims) sims[q].counter[0] += nt;
}

// Incident line values of a region: for each step of the chunk, the line
// nodes from 2 before to 2 after the region along the incidence axis (box
// node offsets are within half a cell, so every face cell of the region
// reads inside this window).
constexpr int kGW = (kRH > 64 ? kRH : 64) + 4;

// incident value of a TF/SF box cell at global (i, k) from the region's window
__device__ __forceinline__ double snap_value(const BSim &S, const TfBox &b, int i, int k, const double *win_s,
                                             int mbase) {
    const double cc = S.inc_axis == 0 ? i + (double)b.off[0] : k + (double)b.off[2];
    const int j = (int)floor(cc) - S.inc_x0 - mbase;
    return win_s[j];
}

__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Region of a job: 64 columns (one per lane) x kRH rows; wave w owns rows
// kRW*w .. kRW*w+kRW-1.  Fields Ey, Hx, Hz, the cell's cb and (general path) the
// CPML psi stay in VGPRs for the whole chunk; x coefficients are per lane,
// z coefficients per row.  Per half step the waves exchange one boundary row
// through LDS (Ey row kRW*w before the H update, Hx row kRW*w+kRW-1 before the E
// update); x neighbours come from DPP wavefront shifts.
//
// No cell update is masked and no psi update is conditional: the CPML terms
// run on every cell with c = 0 (psi stays 0 and adds nothing) off the slabs,
// which is exactly the single-simulation result there.  Cells that the
// reference never updates (the PEC walls: E on the first and last index of an
// axis, H on the last) and cells outside the domain get zero coefficients:
// cb = 0 for such E cells, kh = b = c = 0 for such H rows and columns,
// b = c = 0 for E psi, so fields and psi keep their values.  Cells near the
// region edge, whose neighbours lie outside the region, take finite garbage
// values; they are stale either way and never written back.
//
// INNER jobs have their region inside [1, N-2] on both axes, outside the
// CPML slabs and on rows with one z coefficient (the host checks): no psi,
// one z coefficient per wave.  For INNER jobs, XZ = 1 marks members whose E
// coefficients there are exactly 1 (dz = dx, kappa 1): their E update forms
// d1 - d2 without the two unit products, which rounds exactly as
// d1*kez0 - d2*kex does with kez0 = kex = 1.  The other jobs come in four classes XZ by the
// slabs their region meets (bit 0: an x slab, bit 1: a z slab); the psi of a
// slab the region does not meet is zero throughout (c = 0 there), so those
// classes carry only the psi they need.
template <typename T, bool INNER, int XZ>
__global__ __launch_bounds__(64 * kNW) void k_tile2d(const BSim *__restrict__ sims, const BJob *__restrict__ jobs,
                                                     int nt, int par) {
    constexpr int NT = 64 * kNW;
    constexpr bool XP = !INNER && (XZ & 1), ZP = !INNER && (XZ & 2);   // x / z psi carried
    constexpr bool UE = INNER && (XZ & 1);                              // unit E coefficients
    __shared__ float scb[256];
    __shared__ float sEtop[kNW][64], sHbot[kNW][64];
    __shared__ int sprb[256];                 // probe index
    __shared__ int sprd[256];                 // packed component << 16 | region row << 8 | region column
    __shared__ int snprb;
    __shared__ float4 szh[INNER ? 1 : kRH], sze[INNER ? 1 : kRH];   // z {kh, b_h, c_h}, {ke, b_e, c_e} per row
    __shared__ double sg[2][kTT][kGW];        // incident E (before H half step) / H (before E half step)

    const BJob job = jobs[xcd_remap(blockIdx.x, gridDim.x)];
    const BSim &S = sims[job.sim];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nx = S.nx, nz = S.nz;
    const int gi0 = job.i0 - kTT, gk0 = job.k0 - kTT;   // region origin
    const int i = gi0 + lane;
    const int kb = gk0 + kRW * w;                        // first row of this wave
    const long long px = S.px, fs = S.fstride;
    const float *tab = S.tab;
    const int tn = S.tn;
    const T *Fin = (const T *)S.F[par];
    T *Fout = (T *)S.F[par ^ 1];
    const bool colin = i >= 0 && i < nx;
    const bool pml = !INNER && S.pml;

    for (int q = tid; q < 256; q += NT) scb[q] = S.cb[256 + q];   // the Ey table
    if (tid == 0) snprb = 0;
    if (!INNER) {
        for (int q = tid; q < kRH; q += NT) {
            const int k = gk0 + q, kc = min(max(k, 0), tn - 1);
            const bool hok = k >= 0 && k <= nz - 2, eok = k >= 1 && k <= nz - 2;
            szh[q] = hok ? make_float4(tab[(2 * 6 + TK_KH) * tn + kc], tab[(2 * 6 + TK_BH) * tn + kc],
                                       tab[(2 * 6 + TK_CH) * tn + kc], 0.f)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
            sze[q] = make_float4(tab[(2 * 6 + TK_KE) * tn + kc], eok ? tab[(2 * 6 + TK_BE) * tn + kc] : 0.f,
                                 eok ? tab[(2 * 6 + TK_CE) * tn + kc] : 0.f, 0.f);
        }
    }
    // incident line window of the region, when a TF/SF box touches it
    bool tf_any = false;
    for (int bq = 0; bq < S.ntfH; ++bq) {
        const TfBox &bx = S.tfH[bq];
        tf_any |= bx.hi[0] >= gi0 && bx.lo[0] < gi0 + 64 && bx.hi[2] >= gk0 && bx.lo[2] < gk0 + kRH;
    }
    for (int bq = 0; bq < S.ntfE; ++bq) {
        const TfBox &bx = S.tfE[bq];
        tf_any |= bx.hi[0] >= gi0 && bx.lo[0] < gi0 + 64 && bx.hi[2] >= gk0 && bx.lo[2] < gk0 + kRH;
    }
    const int mbase = (S.inc_axis == 0 ? gi0 : gk0) - 2 - S.inc_x0;
    if (tf_any) {
        for (int q = tid; q < nt * kGW; q += NT) {
            const int st = q / kGW, j = q % kGW, m = mbase + j;
            const bool in = m >= 0 && m < S.win;
            sg[0][st][j] = in ? S.esnap[(long long)st * S.win + m] : 0.0;
            sg[1][st][j] = in ? S.hsnap[(long long)st * S.win + m] : 0.0;
        }
    }
    __syncthreads();
    // probes on interior cells of this job
    for (int q = tid; q < S.np; q += NT) {
        const long long o = S.pidx[q];
        const int rI = (int)(o % px) - gi0, rK = (int)(o / px) - gk0;
        if (rI >= kTT && rI < 64 - kTT && rK >= kTT && rK < kRH - kTT) {
            const int slot = atomicAdd(&snprb, 1);
            sprb[slot] = q;
            sprd[slot] = S.pcomp[q] << 16 | rK << 8 | rI;
        }
    }

    // coefficients: x per lane, zeroed where the cell is never updated
    const int ic = min(max(i, 0), tn - 1);
    const bool hxok = INNER || (i >= 0 && i <= nx - 2), exok = INNER || (i >= 1 && i <= nx - 2);
    const float khx = hxok ? tab[(0 * 6 + TK_KH) * tn + ic] : 0.f, kex = tab[(0 * 6 + TK_KE) * tn + ic];
    float bhx = 0.f, chx = 0.f, bex = 0.f, cex = 0.f;
    if (!INNER) {
        bhx = hxok ? tab[(0 * 6 + TK_BH) * tn + ic] : 0.f;
        chx = hxok ? tab[(0 * 6 + TK_CH) * tn + ic] : 0.f;
        bex = exok ? tab[(0 * 6 + TK_BE) * tn + ic] : 0.f;
        cex = exok ? tab[(0 * 6 + TK_CE) * tn + ic] : 0.f;
    }
    const int kc0 = min(max(kb, 0), tn - 1);
    const float khz0 = tab[(2 * 6 + TK_KH) * tn + kc0], kez0 = tab[(2 * 6 + TK_KE) * tn + kc0];

    // fields, cb and psi of this lane's kRW cells
    const int slo0 = S.slo[0], shi0 = S.shi[0], slo2 = S.slo[2], shi2 = S.shi[2];
    const bool xslab = pml && colin && (i < slo0 || i >= shi0);
    const float *px0 = S.psi[par], *pz0 = S.psi[par] + 4 * (S.scells[0] + S.scells[1]);
    float ey[kRW], hx[kRW], hz[kRW], cb[kRW];
    float pxh[kRW], pxe[kRW], pzh[kRW], pze[kRW];   // x psi (Hz, Ey), z psi (Hx, Ey); unused when INNER
#pragma unroll
    for (int m = 0; m < kRW; ++m) {
        const int k = kb + m;
        ey[m] = hx[m] = hz[m] = cb[m] = 0.f;
        pxh[m] = pxe[m] = pzh[m] = pze[m] = 0.f;
        if (INNER || (colin && k >= 0 && k < nz)) {
            const long long o = (long long)k * px + i;
            ey[m] = bld(Fin + fs + o);
            hx[m] = bld(Fin + 3 * fs + o);
            hz[m] = bld(Fin + 5 * fs + o);
            cb[m] = scb[S.mat[o]];
            if (!INNER && !(exok && k >= 1 && k <= nz - 2)) cb[m] = 0.f;   // PEC wall
            if (!INNER && pml) {
                if (XP && xslab) {
                    const long long off = (long long)k * S.sw[0] + bslab(i, slo0, shi0);
                    pxh[m] = px0[off];                         // H slot0 (Hz)
                    pxe[m] = px0[3 * S.scells[0] + off];       // E slot1 (Ey)
                }
                if (ZP && (k < slo2 || k >= shi2)) {
                    const long long off = (long long)bslab(k, slo2, shi2) * px + i;
                    pzh[m] = pz0[S.scells[2] + off];           // H slot1 (Hx)
                    pze[m] = pz0[2 * S.scells[2] + off];       // E slot0 (Ey)
                }
            }
        }
    }
    __syncthreads();   // probe list complete
    const int nprb = snprb;
    // TF/SF boxes that touch this wave's kRW x 64 cells (wave-uniform bit masks)
    unsigned hmask = 0, emask = 0;
    for (int bq = 0; bq < S.ntfH; ++bq) {
        const TfBox &bx = S.tfH[bq];
        if (bx.hi[0] >= gi0 && bx.lo[0] < gi0 + 64 && bx.hi[2] >= kb && bx.lo[2] < kb + kRW) hmask |= 1u << bq;
    }
    for (int bq = 0; bq < S.ntfE; ++bq) {
        const TfBox &bx = S.tfE[bq];
        if (bx.hi[0] >= gi0 && bx.lo[0] < gi0 + 64 && bx.hi[2] >= kb && bx.lo[2] < kb + kRW) emask |= 1u << bq;
    }
    hmask = __builtin_amdgcn_readfirstlane(hmask);
    emask = __builtin_amdgcn_readfirstlane(emask);
    const int pcap = S.pcap, np = S.np;
    double *prec = S.prec;

    const int t0 = S.counter[0];
    for (int s = 0; s < nt; ++s) {
        const int t = t0 + s;
        // ---------------- H half step
        sEtop[w][lane] = ey[0];
        __syncthreads();
        const float eyn = (w + 1 < kNW) ? sEtop[w + 1][lane] : 0.f;
#pragma unroll
        for (int m = 0; m < kRW; ++m) {
            const float e0 = ey[m];
            const float eu = m < kRW - 1 ? ey[m + 1] : eyn;
            const float er = wave_next(e0);
            const float d2 = eu - e0, d1 = er - e0;
            if (INNER) {
                hx[m] -= -d2 * khz0;
                hz[m] -= d1 * khx;
            } else {
                const float4 z = szh[kRW * w + m];
                float acc = -d2 * z.x;
                if (ZP) {
                    pzh[m] = z.y * pzh[m] + z.z * d2;
                    acc -= pzh[m];
                }
                hx[m] -= acc;
                float acc2 = d1 * khx;
                if (XP) {
                    pxh[m] = bhx * pxh[m] + chx * d1;
                    acc2 += pxh[m];
                }
                hz[m] -= acc2;
            }
        }
        // TF/SF corrections of the H half step (incident E of this step).
        // One incident value per box and lane: for an x face, lane l reads
        // the value of row l and the face lane takes it with readlane.
        for (unsigned mk = hmask; mk; mk &= mk - 1) {
            const TfBox &bx = S.tfH[__builtin_ctz(mk)];
            const double *snap = sg[0][s];
            const int a = (bx.comp + 1 + bx.curl) % 3;
            const float sgn = bx.curl == 0 ? 1.f : -1.f;
            const bool xf = bx.fa == 0;
            const int kl = kb + (lane & (kRW - 1));
            const bool ld = xf ? (lane < kRW && kl >= bx.lo[2] && kl <= bx.hi[2])
                               : (i >= bx.lo[0] && i <= bx.hi[0]);
            const float dl = ld ? (float)(bx.dir * snap_value(S, bx, xf ? bx.lo[0] : i, xf ? kl : bx.lo[2], snap, mbase)) : 0.f;
            const bool face_lane = xf ? i == bx.lo[0] : ld;
#pragma unroll
            for (int m = 0; m < kRW; ++m) {
                const int k = kb + m;
                if (!(xf ? (k >= bx.lo[2] && k <= bx.hi[2]) : k == bx.lo[2])) continue;   // wave-uniform
                const float delta = xf ? rdlane(dl, m) : dl;
                float coef;
                if (a == 0) {
                    coef = khx;
                    if (XP && face_lane) {
                        pxh[m] += chx * delta;
                        coef += chx;
                    }
                } else if (INNER) {
                    coef = khz0;
                } else {
                    const float4 z = szh[kRW * w + m];
                    coef = z.x;
                    if (ZP) {
                        if (face_lane) pzh[m] += z.z * delta;
                        coef += z.z;
                    }
                }
                const float v = sgn * coef * delta;
                if (face_lane) {
                    if (bx.comp == 2) hz[m] -= v;
                    else hx[m] -= v;
                }
            }
        }
        // H probes of step t
        for (int q = 0; q < nprb; ++q) {
            const int d = sprd[q], c = d >> 16, rK = (d >> 8) & 255, rI = d & 255;
            if (c >= 3 && t < pcap && lane == rI && rK / kRW == w) {
                float v = 0.f;
#pragma unroll
                for (int m = 0; m < kRW; ++m)
                    if (m == (rK & (kRW - 1))) v = c == 3 ? hx[m] : (c == 5 ? hz[m] : 0.f);
                prec[(long long)t * np + sprb[q]] = (double)v * S.hscale;
            }
        }
        // ---------------- E half step
        sHbot[w][lane] = hx[kRW - 1];
        __syncthreads();
        const float hxp = w > 0 ? sHbot[w - 1][lane] : 0.f;
#pragma unroll
        for (int m = 0; m < kRW; ++m) {
            const float hd = m > 0 ? hx[m - 1] : hxp;
            const float hl = wave_prev(hz[m]);
            const float d1 = hx[m] - hd;
            const float d2 = hz[m] - hl;
            if (UE) {
                ey[m] += cb[m] * (d1 - d2);
            } else if (INNER) {
                ey[m] += cb[m] * (d1 * kez0 - d2 * kex);
            } else {
                const float4 z = sze[kRW * w + m];
                float acc = d1 * z.x - d2 * kex;
                if (ZP) {
                    pze[m] = z.y * pze[m] + z.z * d1;
                    acc += pze[m];
                }
                if (XP) {
                    pxe[m] = bex * pxe[m] + cex * d2;
                    acc -= pxe[m];
                }
                ey[m] += cb[m] * acc;
            }
        }
        // TF/SF corrections of the E half step (incident H of this step)
        for (unsigned mk = emask; mk; mk &= mk - 1) {
            const TfBox &bx = S.tfE[__builtin_ctz(mk)];
            const double *snap = sg[1][s];
            const int a = (bx.comp + 1 + bx.curl) % 3;
            const float sgn = bx.curl == 0 ? 1.f : -1.f;
            const bool xf = bx.fa == 0;
            const int kl = kb + (lane & (kRW - 1));
            const bool ld = xf ? (lane < kRW && kl >= bx.lo[2] && kl <= bx.hi[2])
                               : (i >= bx.lo[0] && i <= bx.hi[0]);
            const float dl = ld ? (float)(bx.dir * snap_value(S, bx, xf ? bx.lo[0] : i, xf ? kl : bx.lo[2], snap, mbase)) : 0.f;
            const bool face_lane = xf ? i == bx.lo[0] : ld;
#pragma unroll
            for (int m = 0; m < kRW; ++m) {
                const int k = kb + m;
                if (!(xf ? (k >= bx.lo[2] && k <= bx.hi[2]) : k == bx.lo[2])) continue;   // wave-uniform
                const float delta = xf ? rdlane(dl, m) : dl;
                float coef;
                if (a == 0) {
                    coef = kex;
                    if (XP && face_lane) {
                        pxe[m] += cex * delta;
                        coef += cex;
                    }
                } else if (INNER) {
                    coef = kez0;
                } else {
                    const float4 z = sze[kRW * w + m];
                    coef = z.x;
                    if (ZP) {
                        if (face_lane) pze[m] += z.z * delta;
                        coef += z.z;
                    }
                }
                if (face_lane) ey[m] += cb[m] * (sgn * coef * delta);
            }
        }
        // E probes of step t
        for (int q = 0; q < nprb; ++q) {
            const int d = sprd[q], c = d >> 16, rK = (d >> 8) & 255, rI = d & 255;
            if (c < 3 && t < pcap && lane == rI && rK / kRW == w) {
                float v = 0.f;
#pragma unroll
                for (int m = 0; m < kRW; ++m)
                    if (m == (rK & (kRW - 1)) && c == 1) v = ey[m];
                prec[(long long)t * np + sprb[q]] = (double)v;
            }
        }
    }

    // ---------------- write the interior back into the other buffers
    float *qx0 = S.psi[par ^ 1], *qz0 = S.psi[par ^ 1] + 4 * (S.scells[0] + S.scells[1]);
#pragma unroll
    for (int m = 0; m < kRW; ++m) {
        const int k = kb + m, rk = kRW * w + m;
        if (rk >= kTT && rk < kRH - kTT && lane >= kTT && lane < 64 - kTT && (INNER || (i < nx && k < nz))) {
            const long long o = (long long)k * px + i;
            bst(Fout + fs + o, ey[m]);
            bst(Fout + 3 * fs + o, hx[m]);
            bst(Fout + 5 * fs + o, hz[m]);
            if (!INNER && pml) {
                if (XP && xslab) {
                    const long long off = (long long)k * S.sw[0] + bslab(i, slo0, shi0);
                    __builtin_nontemporal_store(pxh[m], qx0 + off);
                    __builtin_nontemporal_store(pxe[m], qx0 + 3 * S.scells[0] + off);
                }
                if (ZP && (k < slo2 || k >= shi2)) {
                    const long long off = (long long)bslab(k, slo2, shi2) * px + i;
                    __builtin_nontemporal_store(pzh[m], qz0 + S.scells[2] + off);
                    __builtin_nontemporal_store(pze[m], qz0 + 2 * S.scells[2] + off);
                }
            }
        }
    }
}

template <typename T>
void launch_tile2d(const BSim *sims, const BJob *jobs, int njobs, int cls, int nt, int par, hipStream_t s) {
    if (!njobs) return;
    const dim3 grd((unsigned)njobs), blk(64 * kNW);
    switch (cls) {
    case kInnerJobs: hipLaunchKernelGGL((k_tile2d<T, true, 0>), grd, blk, 0, s, sims, jobs, nt, par); break;
    case kInnerUnitJobs: hipLaunchKernelGGL((k_tile2d<T, true, 1>), grd, blk, 0, s, sims, jobs, nt, par); break;
    case 0: hipLaunchKernelGGL((k_tile2d<T, false, 0>), grd, blk, 0, s, sims, jobs, nt, par); break;
    case 1: hipLaunchKernelGGL((k_tile2d<T, false, 1>), grd, blk, 0, s, sims, jobs, nt, par); break;
    case 2: hipLaunchKernelGGL((k_tile2d<T, false, 2>), grd, blk, 0, s, sims, jobs, nt, par); break;
    default: hipLaunchKernelGGL((k_tile2d<T, false, 3>), grd, blk, 0, s, sims, jobs, nt, par); break;
    }
}
void launch_aux_chunk(const BAux *aux, int naux, int nt, hipStream_t s) {
    hipLaunchKernelGGL(k_aux_chunk, dim3((unsigned)naux), dim3(256), 0, s, aux, nt);
}
void launch_batch_finish(const BSim *sims, int nsims, int nt, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_finish, dim3((unsigned)((nsims + 255) / 256)), dim3(256), 0, s, sims, nsims, nt);
}
template void launch_tile2d<float>(const BSim *, const BJob *, int, int, int, int, hipStream_t);
template void launch_tile2d<__half>(const BSim *, const BJob *, int, int, int, int, hipStream_t);

}  // namespace lx
