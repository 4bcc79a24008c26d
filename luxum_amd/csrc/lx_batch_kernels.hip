// lx_batch_kernels.hip -- batched 2D (x-z plane, Ey Hx Hz) timesteps with
// temporal tiling in LDS.
//
// Many planar simulations of uneven sizes advance together: one launch covers
// a flattened list of tile jobs from every simulation.  A job is a 48 x 48
// tile of one simulation; its block loads the tile plus an 8-cell halo
// (64 x 64 cells of Ey, Hx, Hz, material ids and the CPML psi of slab cells)
// into LDS, advances up to 8 timesteps there, and writes the 48 x 48 interior
// back.  After s steps the cells within s of the halo edge are stale (the
// Yee dependency cone grows one cell per half step), so the interior is exact
// after 8.  Neighbouring jobs read the halo of step t while others write step
// t+8, so fields and psi ping-pong between two buffers.
//
// Per step the block does what the single-simulation kernels do
// (maxwell_two_dim_pml.cc:144-176): H update (+psi), TF/SF face corrections
// of the H half step, E update (+psi), TF/SF corrections of the E half step,
// probe records.  The incident values of the aux-1D line for the chunk's 8
// steps are precomputed per simulation by k_aux_chunk.
//
// HBM traffic per cell-step is ~(64^2*13 + 48^2*12)/(48^2*8) = 4.4 bytes,
// against 37 for the half-step kernels.
#include "lx_batch.hpp"
#include "../../include/luxum.h"

namespace lx {

template <typename T>
__device__ __forceinline__ float bld(const T *p);
template <>
__device__ __forceinline__ float bld<float>(const float *p) { return *p; }
template <>
__device__ __forceinline__ float bld<__half>(const __half *p) { return __half2float(*p); }
template <typename T>
__device__ __forceinline__ void bst(T *p, float v);
template <>
__device__ __forceinline__ void bst<float>(float *p, float v) { *p = v; }
template <>
__device__ __forceinline__ void bst<__half>(__half *p, float v) { *p = __float2half(v); }

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

// incident value of a TF/SF box cell at region-independent global (i, k)
__device__ __forceinline__ double snap_value(const BSim &S, const TfBox &b, int i, int k, const double *snap) {
    const double cc = S.inc_axis == 0 ? i + (double)b.off[0] : k + (double)b.off[2];
    const int m = (int)floor(cc) - S.inc_x0;
    return (m >= 0 && m < S.win) ? snap[m] : 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_tile2d(const BSim *__restrict__ sims, const BJob *__restrict__ jobs, int nt,
                                                int par) {
    __shared__ float sEy[kRW * kLP], sHx[kRW * kLP], sHz[kRW * kLP];
    __shared__ uint8_t sid[kRW * kRW];
    __shared__ float scb[256];
    __shared__ float spx[2][kRW * kPSW];   // x-slab psi: [0] Hz (H slot0), [1] Ey (E slot1); [rk][col]
    __shared__ float spz[2][kPSW * kRW];   // z-slab psi: [0] Hx (H slot1), [1] Ey (E slot0); [row][ri]

    const BJob job = jobs[blockIdx.x];
    const BSim &S = sims[job.sim];
    const int tid = threadIdx.x;
    const int nx = S.nx, nz = S.nz;
    const int gi0 = job.i0 - kTT, gk0 = job.k0 - kTT;   // region origin
    const T *Fin = (const T *)S.F[par];
    T *Fout = (T *)S.F[par ^ 1];
    const float *psin = S.psi[par];
    float *psout = S.psi[par ^ 1];
    const long long px = S.px, fs = S.fstride;
    const float *tab = S.tab;
    const int tn = S.tn;
    const bool pml = S.pml;
    // x-slab columns and z-slab rows of the region (one contiguous range each)
    int xs0 = 0, xs1 = 0, zs0 = 0, zs1 = 0;
    if (pml) {
        const int lo0 = max(0, -gi0), hi0 = min(kRW, nx - gi0);
        int a = min(hi0, max(lo0, S.slo[0] - gi0)), b = max(lo0, min(hi0, S.shi[0] - gi0));
        if (a > lo0) { xs0 = lo0; xs1 = a; } else if (b < hi0) { xs0 = b; xs1 = hi0; }
        const int lz = max(0, -gk0), hz = min(kRW, nz - gk0);
        int c = min(hz, max(lz, S.slo[2] - gk0)), d = max(lz, min(hz, S.shi[2] - gk0));
        if (c > lz) { zs0 = lz; zs1 = c; } else if (d < hz) { zs0 = d; zs1 = hz; }
    }
    const float *px0 = psin, *pz0 = psin + 4 * (S.scells[0] + S.scells[1]);
    float *qx0 = psout, *qz0 = psout + 4 * (S.scells[0] + S.scells[1]);

    // ---- load the region
    scb[tid] = S.cb[tid];
    for (int q = tid; q < kRW * kRW; q += 256) {
        const int ri = q & (kRW - 1), rk = q >> 6;
        const int i = gi0 + ri, k = gk0 + rk;
        float ey = 0.f, hx = 0.f, hz = 0.f;
        uint8_t id = 0;
        if (i >= 0 && i < nx && k >= 0 && k < nz) {
            const long long o = (long long)k * px + i;
            ey = bld(Fin + fs + o);
            hx = bld(Fin + 3 * fs + o);
            hz = bld(Fin + 5 * fs + o);
            id = S.mat[o];
        }
        sEy[rk * kLP + ri] = ey;
        sHx[rk * kLP + ri] = hx;
        sHz[rk * kLP + ri] = hz;
        sid[q] = id;
    }
    if (pml) {
        for (int q = tid; q < kRW * kPSW; q += 256) {
            const int rk = q / kPSW, col = q % kPSW;
            const int ri = xs0 + col, i = gi0 + ri, k = gk0 + rk;
            float a = 0.f, b = 0.f;
            if (ri < xs1 && k >= 0 && k < nz) {
                const long long off = (long long)k * S.sw[0] + bslab(i, S.slo[0], S.shi[0]);
                a = px0[off];                          // H slot0 (Hz)
                b = px0[3 * S.scells[0] + off];        // E slot1 (Ey)
            }
            spx[0][q] = a;
            spx[1][q] = b;
        }
        for (int q = tid; q < kPSW * kRW; q += 256) {
            const int row = q >> 6, ri = q & (kRW - 1);
            const int rk = zs0 + row, i = gi0 + ri, k = gk0 + rk;
            float a = 0.f, b = 0.f;
            if (rk < zs1 && i >= 0 && i < nx) {
                const long long off = (long long)bslab(k, S.slo[2], S.shi[2]) * px + i;
                a = pz0[S.scells[2] + off];            // H slot1 (Hx)
                b = pz0[2 * S.scells[2] + off];        // E slot0 (Ey)
            }
            spz[0][q] = a;
            spz[1][q] = b;
        }
    }
    __syncthreads();

    const int ri = tid & (kRW - 1);
    const int i = gi0 + ri;
    const int t0 = S.counter[0];
    for (int s = 0; s < nt; ++s) {
        const int t = t0 + s;
        // ---- H half step
#pragma unroll 4
        for (int m = 0; m < kRW / 4; ++m) {
            const int rk = (tid >> 6) + 4 * m, k = gk0 + rk;
            const float e0 = sEy[rk * kLP + ri];
            if (rk + 1 < kRW && i >= 0 && i < nx && k >= 0 && k < nz - 1) {   // Hx
                const float d2 = sEy[(rk + 1) * kLP + ri] - e0;
                float acc = -d2 * tab[(2 * 6 + TK_KH) * tn + k];
                if (pml && rk >= zs0 && rk < zs1) {
                    float &p = spz[0][(rk - zs0) * kRW + ri];
                    p = tab[(2 * 6 + TK_BH) * tn + k] * p + tab[(2 * 6 + TK_CH) * tn + k] * d2;
                    acc -= p;
                }
                sHx[rk * kLP + ri] -= acc;
            }
            if (ri + 1 < kRW && i >= 0 && i < nx - 1 && k >= 0 && k < nz) {   // Hz
                const float d1 = sEy[rk * kLP + ri + 1] - e0;
                float acc = d1 * tab[(0 * 6 + TK_KH) * tn + i];
                if (pml && ri >= xs0 && ri < xs1) {
                    float &p = spx[0][rk * kPSW + (ri - xs0)];
                    p = tab[(0 * 6 + TK_BH) * tn + i] * p + tab[(0 * 6 + TK_CH) * tn + i] * d1;
                    acc += p;
                }
                sHz[rk * kLP + ri] -= acc;
            }
        }
        __syncthreads();
        // TF/SF corrections of the H half step: box b = tid/64, cell along the in-plane transverse axis
        if (tid < 64 * S.ntfH) {
            const TfBox &bx = S.tfH[tid >> 6];
            const int u = tid & 63;
            const int fa = bx.fa;                 // 0 (x-face) or 2 (z-face)
            int gi, gk;
            if (fa == 0) { gi = bx.lo[0]; gk = gk0 + u; } else { gk = bx.lo[2]; gi = gi0 + u; }
            const int rI = gi - gi0, rK = gk - gk0;
            if (rI >= 0 && rI < kRW && rK >= 0 && rK < kRW && gi >= bx.lo[0] && gi <= bx.hi[0] && gk >= bx.lo[2] &&
                gk <= bx.hi[2]) {
                const float delta = (float)(bx.dir * snap_value(S, bx, gi, gk, S.esnap + (long long)s * S.win));
                const int a = (bx.comp + 1 + bx.curl) % 3;   // 0 (Hz curl 0) or 2 (Hx curl 1)
                const int ia = a == 0 ? gi : gk;
                float coef = tab[(a * 6 + TK_KH) * tn + ia];
                if (pml) {
                    if (a == 0 && rI >= xs0 && rI < xs1) {
                        const float cp = tab[(0 * 6 + TK_CH) * tn + ia];
                        spx[0][rK * kPSW + (rI - xs0)] += cp * delta;
                        coef += cp;
                    } else if (a == 2 && rK >= zs0 && rK < zs1) {
                        const float cp = tab[(2 * 6 + TK_CH) * tn + ia];
                        spz[0][(rK - zs0) * kRW + rI] += cp * delta;
                        coef += cp;
                    }
                }
                const float sgn = bx.curl == 0 ? 1.f : -1.f;
                if (bx.comp == 2) sHz[rK * kLP + rI] -= sgn * coef * delta;
                else sHx[rK * kLP + rI] -= sgn * coef * delta;
            }
        }
        __syncthreads();
        // H probes of step t (owned interior cells)
        if (tid < S.np && t < S.pcap) {
            const int c = S.pcomp[tid];
            const long long o = S.pidx[tid];
            const int pi = (int)(o % px), pk = (int)(o / px);
            const int rI = pi - gi0, rK = pk - gk0;
            if (c >= 3 && rI >= kTT && rI < kTT + kTW && rK >= kTT && rK < kTT + kTW)
                S.prec[(long long)t * S.np + tid] = (double)(c == 3 ? sHx[rK * kLP + rI] : sHz[rK * kLP + rI]) * S.hscale;
        }
        // ---- E half step
#pragma unroll 4
        for (int m = 0; m < kRW / 4; ++m) {
            const int rk = (tid >> 6) + 4 * m, k = gk0 + rk;
            if (ri >= 1 && rk >= 1 && i >= 1 && i <= nx - 2 && k >= 1 && k <= nz - 2) {
                const int o = rk * kLP + ri;
                const float d1 = sHx[o] - sHx[o - kLP];
                const float d2 = sHz[o] - sHz[o - 1];
                float acc = d1 * tab[(2 * 6 + TK_KE) * tn + k] - d2 * tab[(0 * 6 + TK_KE) * tn + i];
                if (pml) {
                    if (rk >= zs0 && rk < zs1) {
                        float &p = spz[1][(rk - zs0) * kRW + ri];
                        p = tab[(2 * 6 + TK_BE) * tn + k] * p + tab[(2 * 6 + TK_CE) * tn + k] * d1;
                        acc += p;
                    }
                    if (ri >= xs0 && ri < xs1) {
                        float &p = spx[1][rk * kPSW + (ri - xs0)];
                        p = tab[(0 * 6 + TK_BE) * tn + i] * p + tab[(0 * 6 + TK_CE) * tn + i] * d2;
                        acc -= p;
                    }
                }
                sEy[o] += scb[sid[rk * kRW + ri]] * acc;
            }
        }
        __syncthreads();
        // TF/SF corrections of the E half step: first-curl boxes, then second-curl
        // boxes (they share the corner cells of the TF box)
        for (int pass = 0; pass < 2; ++pass) {
            if (tid < 64 * S.ntfE) {
                const TfBox &bx = S.tfE[tid >> 6];
                const int u = tid & 63;
                if (bx.curl == pass) {
                    const int fa = bx.fa;
                    int gi, gk;
                    if (fa == 0) { gi = bx.lo[0]; gk = gk0 + u; } else { gk = bx.lo[2]; gi = gi0 + u; }
                    const int rI = gi - gi0, rK = gk - gk0;
                    if (rI >= 0 && rI < kRW && rK >= 0 && rK < kRW && gi >= bx.lo[0] && gi <= bx.hi[0] &&
                        gk >= bx.lo[2] && gk <= bx.hi[2]) {
                        const float delta =
                            (float)(bx.dir * snap_value(S, bx, gi, gk, S.hsnap + (long long)s * S.win));
                        const int a = (bx.comp + 1 + bx.curl) % 3;   // 2 (curl 0) or 0 (curl 1)
                        const int ia = a == 0 ? gi : gk;
                        float coef = tab[(a * 6 + TK_KE) * tn + ia];
                        if (pml) {
                            if (a == 0 && rI >= xs0 && rI < xs1) {
                                const float cp = tab[(0 * 6 + TK_CE) * tn + ia];
                                spx[1][rK * kPSW + (rI - xs0)] += cp * delta;
                                coef += cp;
                            } else if (a == 2 && rK >= zs0 && rK < zs1) {
                                const float cp = tab[(2 * 6 + TK_CE) * tn + ia];
                                spz[1][(rK - zs0) * kRW + rI] += cp * delta;
                                coef += cp;
                            }
                        }
                        const float sgn = bx.curl == 0 ? 1.f : -1.f;
                        sEy[rK * kLP + rI] += scb[sid[rK * kRW + rI]] * (sgn * coef * delta);
                    }
                }
            }
            __syncthreads();
        }
        // E probes of step t
        if (tid < S.np && t < S.pcap) {
            const int c = S.pcomp[tid];
            const long long o = S.pidx[tid];
            const int pi = (int)(o % px), pk = (int)(o / px);
            const int rI = pi - gi0, rK = pk - gk0;
            if (c == 1 && rI >= kTT && rI < kTT + kTW && rK >= kTT && rK < kTT + kTW)
                S.prec[(long long)t * S.np + tid] = (double)sEy[rK * kLP + rI];
        }
    }

    // ---- write the interior back into the other buffer
    for (int q = tid; q < kTW * kTW; q += 256) {
        const int ri2 = kTT + q % kTW, rk2 = kTT + q / kTW;
        const int gi = gi0 + ri2, gk = gk0 + rk2;
        if (gi < nx && gk < nz) {
            const long long o = (long long)gk * px + gi;
            bst(Fout + fs + o, sEy[rk2 * kLP + ri2]);
            bst(Fout + 3 * fs + o, sHx[rk2 * kLP + ri2]);
            bst(Fout + 5 * fs + o, sHz[rk2 * kLP + ri2]);
        }
    }
    if (pml) {
        for (int q = tid; q < kTW * kTW; q += 256) {
            const int ri2 = kTT + q % kTW, rk2 = kTT + q / kTW;
            const int gi = gi0 + ri2, gk = gk0 + rk2;
            if (gi >= nx || gk >= nz) continue;
            if (ri2 >= xs0 && ri2 < xs1) {
                const long long off = (long long)gk * S.sw[0] + bslab(gi, S.slo[0], S.shi[0]);
                qx0[off] = spx[0][rk2 * kPSW + (ri2 - xs0)];
                qx0[3 * S.scells[0] + off] = spx[1][rk2 * kPSW + (ri2 - xs0)];
            }
            if (rk2 >= zs0 && rk2 < zs1) {
                const long long off = (long long)bslab(gk, S.slo[2], S.shi[2]) * px + gi;
                qz0[S.scells[2] + off] = spz[0][(rk2 - zs0) * kRW + ri2];
                qz0[2 * S.scells[2] + off] = spz[1][(rk2 - zs0) * kRW + ri2];
            }
        }
    }
}

template <typename T>
void launch_tile2d(const BSim *sims, const BJob *jobs, int njobs, int nt, int par, hipStream_t s) {
    hipLaunchKernelGGL(k_tile2d<T>, dim3((unsigned)njobs), dim3(256), 0, s, sims, jobs, nt, par);
}
void launch_aux_chunk(const BAux *aux, int naux, int nt, hipStream_t s) {
    hipLaunchKernelGGL(k_aux_chunk, dim3((unsigned)naux), dim3(256), 0, s, aux, nt);
}
void launch_batch_finish(const BSim *sims, int nsims, int nt, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_finish, dim3((unsigned)((nsims + 255) / 256)), dim3(256), 0, s, sims, nsims, nt);
}
template void launch_tile2d<float>(const BSim *, const BJob *, int, int, int, hipStream_t);
template void launch_tile2d<__half>(const BSim *, const BJob *, int, int, int, hipStream_t);

}  // namespace lx
