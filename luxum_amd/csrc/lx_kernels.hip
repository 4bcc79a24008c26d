// lx_kernels.hip -- gfx950 kernels of the Yee timestep.
//
// One launch per half step.  Each thread owns one (i, j) column of the
// x-fastest device layout and marches along z over kz planes, keeping the
// plane above (H update) or below (E update) in registers, so the z
// neighbour is loaded once per cell.  A thread updates all three components
// at its node: the split curl differences (lib/field/vfield.cc:196-275), the
// TF/SF corrections of that half step (src/solvers/maxwell.cc:512-812), the
// CPML psi recursion on the slab cells (maxwell_three_dim_pml.cc:182-255) and
// the field update, with no curl temporaries in memory.
#include "lx_internal.hpp"
#include "../../include/luxum.h"

namespace lx {

template <typename T>
__device__ __forceinline__ float ld(const T *p, long long i);
template <>
__device__ __forceinline__ float ld<float>(const float *p, long long i) { return p[i]; }
template <>
__device__ __forceinline__ float ld<__half>(const __half *p, long long i) { return __half2float(p[i]); }

template <typename T>
__device__ __forceinline__ void st(T *p, long long i, float v);
template <>
__device__ __forceinline__ void st<float>(float *p, long long i, float v) { p[i] = v; }
template <>
__device__ __forceinline__ void st<__half>(__half *p, long long i, float v) { p[i] = __float2half(v); }

__device__ __forceinline__ bool in_box(const TfBox &b, int i, int j, int k) {
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] && k <= b.hi[2];
}

// Incident scalar at node coordinates (cells) and time level tp (steps).
__device__ double inc_value(const IncParams &in, double cx, double cy, double cz, double tp, bool e_node) {
    if (in.mode == LX_INC_AUX1D) {
        double cc = in.axis == 0 ? cx : (in.axis == 1 ? cy : cz);
        int m = (int)floor(cc) - in.x0;
        if (m < 0 || m >= in.L) return 0.0;
        if (e_node) return in.e[m];
        return m < in.L - 1 ? in.h[m] : 0.0;
    }
    double pos = in.planar ? (cx * in.dx * in.kx + cz * in.dz * in.kz)
                           : (cx * in.dx * in.kx + cy * in.dy * in.ky + cz * in.dz * in.kz);
    double arg = pos + in.init_dx - tp * in.cdt;
    return exp(-(arg * arg) * in.inv_sigma);
}

// Sum of the corrections of one split curl at (i,j,k).
__device__ __forceinline__ float tf_delta(const TfBox *b2, const IncParams &inc, int i, int j, int k, double tp,
                                          bool e_node) {
    float d = 0.f;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        const TfBox &b = b2[f];
        if (in_box(b, i, j, k)) {
            double g = inc_value(inc, i + (double)b.off[0], j + (double)b.off[1], k + (double)b.off[2], tp, e_node);
            d += (float)(b.dir * g);
        }
    }
    return d;
}

__device__ __forceinline__ int slab_index(int x, int lo, int hi) { return x < lo ? x : lo + (x - hi); }

// ---------------------------------------------------------------------------
// 3D H update: Hs -= (d1*kh_a1 - d2*kh_a2) + (psi1 - psi2)
// comp c: curl1 along (c+1)%3, curl2 along (c+2)%3.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBX *kBY) void k_h3(Grid<T> g, TfSet tf, IncParams inc) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int t = g.counter[0];
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) g.counter[1] = t;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= nx || j >= ny) return;
    const int k0 = blockIdx.z * g.kz;
    const int k1 = min(k0 + g.kz, nz);
    const int ip = min(i + 1, nx - 1), jp = min(j + 1, ny - 1);
    const bool vhx = j < ny - 1;
    const bool vhy = i < nx - 1;
    const bool vhz = (i < nx - 1) && (j < ny - 1);
    const long long px = g.px, plane = g.plane;
    const T *Ex = g.f.E[0], *Ey = g.f.E[1], *Ez = g.f.E[2];
    T *Hx = g.f.H[0], *Hy = g.f.H[1], *Hz = g.f.H[2];
    const double tp = t - 0.5;

    const int ix = min(i, nx - 2 < 0 ? 0 : nx - 2), jy = min(j, ny - 2 < 0 ? 0 : ny - 2);
    const float khx = g.tab.kh[0][ix], khy = g.tab.kh[1][jy];
    const bool sx = g.pml && (i < g.slab.lo[0] || i >= g.slab.hi[0]);
    const bool sy = g.pml && (j < g.slab.lo[1] || j >= g.slab.hi[1]);
    const int six = slab_index(i, g.slab.lo[0], g.slab.hi[0]);
    const int siy = slab_index(j, g.slab.lo[1], g.slab.hi[1]);
    const float bhx = sx ? g.tab.bh[0][ix] : 0.f, chx = sx ? g.tab.chh[0][ix] : 0.f;
    const float bhy = sy ? g.tab.bh[1][jy] : 0.f, chy = sy ? g.tab.chh[1][jy] : 0.f;

    long long b = (long long)k0 * plane + (long long)j * px;
    float ex0 = ld(Ex, b + i), ey0 = ld(Ey, b + i), ez0 = ld(Ez, b + i);
    for (int k = k0; k < k1; ++k, b += plane) {
        const long long bn = (k + 1 < nz) ? b + plane : b;
        const float ex1 = ld(Ex, bn + i), ey1 = ld(Ey, bn + i), ez1 = ld(Ez, bn + i);
        const long long bj = (long long)k * plane + (long long)jp * px;
        const float ezj = ld(Ez, bj + i), exj = ld(Ex, bj + i);
        const float ezi = ld(Ez, b + ip), eyi = ld(Ey, b + ip);
        const int kk = min(k, nz - 2 < 0 ? 0 : nz - 2);
        const float khz = g.tab.kh[2][kk];
        const bool sz = g.pml && (k < g.slab.lo[2] || k >= g.slab.hi[2]);
        const int siz = slab_index(k, g.slab.lo[2], g.slab.hi[2]);
        const float bhz = sz ? g.tab.bh[2][kk] : 0.f, chz = sz ? g.tab.chh[2][kk] : 0.f;
        const long long ixs = ((long long)k * ny + j) * g.slab.w[0] + six;   // x slab
        const long long iys = ((long long)k * g.slab.w[1] + siy) * nx + i;   // y slab
        const long long izs = ((long long)siz * ny + j) * nx + i;            // z slab

        // Hx: d1 = dEz/dy (axis y), d2 = dEy/dz (axis z)
        if (vhx && k < nz - 1) {
            float d1 = ezj - ez0, d2 = ey1 - ey0;
            if (tf.any) {
                d1 += tf_delta(tf.box[0][0], inc, i, j, k, tp, true);
                d2 += tf_delta(tf.box[0][1], inc, i, j, k, tp, true);
            }
            float acc = d1 * khy - d2 * khz, pp = 0.f;
            if (sy) { float *q = g.slab.psiH[1][0] + iys; float v = bhy * *q + chy * d1; *q = v; pp += v; }
            if (sz) { float *q = g.slab.psiH[2][1] + izs; float v = bhz * *q + chz * d2; *q = v; pp -= v; }
            const long long o = b + i;
            st(Hx, o, ld(Hx, o) - acc - pp);
        }
        // Hy: d1 = dEx/dz (axis z), d2 = dEz/dx (axis x)
        if (vhy && k < nz - 1) {
            float d1 = ex1 - ex0, d2 = ezi - ez0;
            if (tf.any) {
                d1 += tf_delta(tf.box[1][0], inc, i, j, k, tp, true);
                d2 += tf_delta(tf.box[1][1], inc, i, j, k, tp, true);
            }
            float acc = d1 * khz - d2 * khx, pp = 0.f;
            if (sz) { float *q = g.slab.psiH[2][0] + izs; float v = bhz * *q + chz * d1; *q = v; pp += v; }
            if (sx) { float *q = g.slab.psiH[0][1] + ixs; float v = bhx * *q + chx * d2; *q = v; pp -= v; }
            const long long o = b + i;
            st(Hy, o, ld(Hy, o) - acc - pp);
        }
        // Hz: d1 = dEy/dx (axis x), d2 = dEx/dy (axis y)
        if (vhz) {
            float d1 = eyi - ey0, d2 = exj - ex0;
            if (tf.any) {
                d1 += tf_delta(tf.box[2][0], inc, i, j, k, tp, true);
                d2 += tf_delta(tf.box[2][1], inc, i, j, k, tp, true);
            }
            float acc = d1 * khx - d2 * khy, pp = 0.f;
            if (sx) { float *q = g.slab.psiH[0][0] + ixs; float v = bhx * *q + chx * d1; *q = v; pp += v; }
            if (sy) { float *q = g.slab.psiH[1][1] + iys; float v = bhy * *q + chy * d2; *q = v; pp -= v; }
            const long long o = b + i;
            st(Hz, o, ld(Hz, o) - acc - pp);
        }
        ex0 = ex1; ey0 = ey1; ez0 = ez1;
    }
}

// ---------------------------------------------------------------------------
// 3D E update: E += cb[id] * ((d1*ke_a1 - d2*ke_a2) + (psi1 - psi2)),
// interior only (PEC walls come from the curl ranges, vfield.cc:242-268).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBX *kBY) void k_e3(Grid<T> g, TfSet tf, IncParams inc) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int t = g.counter[1];
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) g.counter[0] = t + 1;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= nx || j >= ny) return;
    const int k0 = blockIdx.z * g.kz;
    const int k1 = min(k0 + g.kz, nz);
    const int im = max(i - 1, 0), jm = max(j - 1, 0);
    const bool vex = (i < nx - 1) && (j >= 1) && (j <= ny - 2);
    const bool vey = (i >= 1) && (i <= nx - 2) && (j < ny - 1);
    const bool vez = (i >= 1) && (i <= nx - 2) && (j >= 1) && (j <= ny - 2);
    const long long px = g.px, plane = g.plane;
    const T *Hx = g.f.H[0], *Hy = g.f.H[1], *Hz = g.f.H[2];
    T *Ex = g.f.E[0], *Ey = g.f.E[1], *Ez = g.f.E[2];
    const double tp = (double)t;

    const float kex = g.tab.ke[0][i], key = g.tab.ke[1][j];
    const bool sx = g.pml && (i < g.slab.lo[0] || i >= g.slab.hi[0]);
    const bool sy = g.pml && (j < g.slab.lo[1] || j >= g.slab.hi[1]);
    const int six = slab_index(i, g.slab.lo[0], g.slab.hi[0]);
    const int siy = slab_index(j, g.slab.lo[1], g.slab.hi[1]);
    const float bex = sx ? g.tab.be[0][i] : 0.f, cex = sx ? g.tab.ce[0][i] : 0.f;
    const float bey = sy ? g.tab.be[1][j] : 0.f, cey = sy ? g.tab.ce[1][j] : 0.f;

    long long b = (long long)k0 * plane + (long long)j * px;
    const long long bprev = k0 > 0 ? b - plane : b;
    float hxm = ld(Hx, bprev + i), hym = ld(Hy, bprev + i);
    for (int k = k0; k < k1; ++k, b += plane) {
        const float hx0 = ld(Hx, b + i), hy0 = ld(Hy, b + i), hz0 = ld(Hz, b + i);
        const long long bj = (long long)k * plane + (long long)jm * px;
        const float hzj = ld(Hz, bj + i), hxj = ld(Hx, bj + i);
        const float hzi = ld(Hz, b + im), hyi = ld(Hy, b + im);
        const float kez = g.tab.ke[2][k];
        const bool sz = g.pml && (k < g.slab.lo[2] || k >= g.slab.hi[2]);
        const int siz = slab_index(k, g.slab.lo[2], g.slab.hi[2]);
        const float bez = sz ? g.tab.be[2][k] : 0.f, cez = sz ? g.tab.ce[2][k] : 0.f;
        const long long ixs = ((long long)k * ny + j) * g.slab.w[0] + six;
        const long long iys = ((long long)k * g.slab.w[1] + siy) * nx + i;
        const long long izs = ((long long)siz * ny + j) * nx + i;
        const long long o = b + i;
        const bool kin = (k >= 1) && (k <= nz - 2);
        const bool any = (vex && kin) || (vey && kin) || (vez && k < nz - 1);
        const float cb = any ? g.cb[g.mat[o]] : 0.f;

        // Ex: d1 = dHz/dy (axis y), d2 = dHy/dz (axis z)
        if (vex && kin) {
            float d1 = hz0 - hzj, d2 = hy0 - hym;
            if (tf.any) {
                d1 += tf_delta(tf.box[0][0], inc, i, j, k, tp, false);
                d2 += tf_delta(tf.box[0][1], inc, i, j, k, tp, false);
            }
            float acc = d1 * key - d2 * kez, pp = 0.f;
            if (sy) { float *q = g.slab.psiE[1][0] + iys; float v = bey * *q + cey * d1; *q = v; pp += v; }
            if (sz) { float *q = g.slab.psiE[2][1] + izs; float v = bez * *q + cez * d2; *q = v; pp -= v; }
            st(Ex, o, ld(Ex, o) + cb * acc + cb * pp);
        }
        // Ey: d1 = dHx/dz (axis z), d2 = dHz/dx (axis x)
        if (vey && kin) {
            float d1 = hx0 - hxm, d2 = hz0 - hzi;
            if (tf.any) {
                d1 += tf_delta(tf.box[1][0], inc, i, j, k, tp, false);
                d2 += tf_delta(tf.box[1][1], inc, i, j, k, tp, false);
            }
            float acc = d1 * kez - d2 * kex, pp = 0.f;
            if (sz) { float *q = g.slab.psiE[2][0] + izs; float v = bez * *q + cez * d1; *q = v; pp += v; }
            if (sx) { float *q = g.slab.psiE[0][1] + ixs; float v = bex * *q + cex * d2; *q = v; pp -= v; }
            st(Ey, o, ld(Ey, o) + cb * acc + cb * pp);
        }
        // Ez: d1 = dHy/dx (axis x), d2 = dHx/dy (axis y)
        if (vez && k < nz - 1) {
            float d1 = hy0 - hyi, d2 = hx0 - hxj;
            if (tf.any) {
                d1 += tf_delta(tf.box[2][0], inc, i, j, k, tp, false);
                d2 += tf_delta(tf.box[2][1], inc, i, j, k, tp, false);
            }
            float acc = d1 * kex - d2 * key, pp = 0.f;
            if (sx) { float *q = g.slab.psiE[0][0] + ixs; float v = bex * *q + cex * d1; *q = v; pp += v; }
            if (sy) { float *q = g.slab.psiE[1][1] + iys; float v = bey * *q + cey * d2; *q = v; pp -= v; }
            st(Ez, o, ld(Ez, o) + cb * acc + cb * pp);
        }
        hxm = hx0; hym = hy0;
    }
}

// ---------------------------------------------------------------------------
// 2D (x-z plane): Ey, Hx, Hz (maxwell_two_dim_pml.cc:144-176)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBX *kBY) void k_h2(Grid<T> g, TfSet tf, IncParams inc) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int k = blockIdx.y * kBY + threadIdx.y;
    const int t = g.counter[0];
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) g.counter[1] = t;
    const int nx = g.nx, nz = g.nz;
    if (i >= nx || k >= nz) return;
    const long long px = g.px;
    const T *Ey = g.f.E[1];
    T *Hx = g.f.H[0], *Hz = g.f.H[2];
    const double tp = t - 0.5;
    const long long o = (long long)k * px + i;
    const float e0 = ld(Ey, o);
    // Hx: d2 = dEy/dz (curl2, axis z); Hs_x = Hs_x - (-d2*kh_z) - (0 - psi2)
    if (k < nz - 1) {
        float d2 = ld(Ey, o + px) - e0;
        if (tf.any) d2 += tf_delta(tf.box[0][1], inc, i, 0, k, tp, true);
        const float khz = g.tab.kh[2][k];
        float pp = 0.f;
        if (g.pml && (k < g.slab.lo[2] || k >= g.slab.hi[2])) {
            const int siz = slab_index(k, g.slab.lo[2], g.slab.hi[2]);
            float *q = g.slab.psiH[2][1] + (long long)siz * nx + i;
            float v = g.tab.bh[2][k] * *q + g.tab.chh[2][k] * d2;
            *q = v;
            pp = -v;
        }
        st(Hx, o, ld(Hx, o) + d2 * khz - pp);
    }
    // Hz: d1 = dEy/dx (curl1, axis x)
    if (i < nx - 1) {
        float d1 = ld(Ey, o + 1) - e0;
        if (tf.any) d1 += tf_delta(tf.box[2][0], inc, i, 0, k, tp, true);
        const float khx = g.tab.kh[0][i];
        float pp = 0.f;
        if (g.pml && (i < g.slab.lo[0] || i >= g.slab.hi[0])) {
            const int six = slab_index(i, g.slab.lo[0], g.slab.hi[0]);
            float *q = g.slab.psiH[0][0] + (long long)k * g.slab.w[0] + six;
            float v = g.tab.bh[0][i] * *q + g.tab.chh[0][i] * d1;
            *q = v;
            pp = v;
        }
        st(Hz, o, ld(Hz, o) - d1 * khx - pp);
    }
}

template <typename T>
__global__ __launch_bounds__(kBX *kBY) void k_e2(Grid<T> g, TfSet tf, IncParams inc) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int k = blockIdx.y * kBY + threadIdx.y;
    const int t = g.counter[1];
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) g.counter[0] = t + 1;
    const int nx = g.nx, nz = g.nz;
    if (i < 1 || i > nx - 2 || k < 1 || k > nz - 2) return;
    const long long px = g.px;
    const T *Hx = g.f.H[0], *Hz = g.f.H[2];
    T *Ey = g.f.E[1];
    const double tp = (double)t;
    const long long o = (long long)k * px + i;
    // Ey: d1 = dHx/dz (curl1, axis z), d2 = dHz/dx (curl2, axis x)
    float d1 = ld(Hx, o) - ld(Hx, o - px);
    float d2 = ld(Hz, o) - ld(Hz, o - 1);
    if (tf.any) {
        d1 += tf_delta(tf.box[1][0], inc, i, 0, k, tp, false);
        d2 += tf_delta(tf.box[1][1], inc, i, 0, k, tp, false);
    }
    float acc = d1 * g.tab.ke[2][k] - d2 * g.tab.ke[0][i], pp = 0.f;
    if (g.pml) {
        if (k < g.slab.lo[2] || k >= g.slab.hi[2]) {
            const int siz = slab_index(k, g.slab.lo[2], g.slab.hi[2]);
            float *q = g.slab.psiE[2][0] + (long long)siz * nx + i;
            float v = g.tab.be[2][k] * *q + g.tab.ce[2][k] * d1;
            *q = v;
            pp += v;
        }
        if (i < g.slab.lo[0] || i >= g.slab.hi[0]) {
            const int six = slab_index(i, g.slab.lo[0], g.slab.hi[0]);
            float *q = g.slab.psiE[0][1] + (long long)k * g.slab.w[0] + six;
            float v = g.tab.be[0][i] * *q + g.tab.ce[0][i] * d2;
            *q = v;
            pp -= v;
        }
    }
    const float cb = g.cb[g.mat[o]];
    st(Ey, o, ld(Ey, o) + cb * acc + cb * pp);
}

// ---------------------------------------------------------------------------
// aux-1D incident line (FP64), point source, probes
// ---------------------------------------------------------------------------
__device__ double waveform(int kind, double t, double c, double lambda0, double fwhm, double delay) {
    const double f0 = c / lambda0, T0 = 1.0 / f0;
    const double tau = fwhm * T0 / (2.0 * sqrt(log(2.0)));
    const double t0 = delay * tau;
    const double x = (t - t0) / tau;
    const double env = exp(-x * x);
    if (kind == LX_WAVE_GAUSSIAN) return env;
    return env * sin(2.0 * M_PI * f0 * (t - t0));
}

__global__ void k_aux_h(AuxLine a) {
    for (int m = threadIdx.x; m < a.L - 1; m += blockDim.x) a.h[m] = a.h[m] - a.S * (a.e[m + 1] - a.e[m]);
}

__global__ void k_aux_e(AuxLine a, const int *counter) {
    const int tn = counter[0];   // = t + 1 after the E update of step t
    for (int m = 1 + threadIdx.x; m < a.L - 1; m += blockDim.x) a.e[m] = a.e[m] - a.S * (a.h[m] - a.h[m - 1]);
    if (threadIdx.x == 0) {
        a.e[0] = waveform(a.wave, tn * a.dt, a.c, a.lambda0, a.fwhm, a.delay);
        a.e[a.L - 1] = 0.0;
    }
}

template <typename T>
__global__ void k_point_source(Grid<T> g, PointSrc ps) {
    const int t = g.counter[0];
    T *E = g.f.E[ps.comp];
    if (ps.kind == LX_SRC_HARD_CW) {
        const int prod = (int)((unsigned)ps.omega * (unsigned)t);
        st(E, ps.idx, (float)(ps.amp * sin(prod * ps.dt)));
    } else {
        const double w = waveform(ps.wave, t * ps.dt, ps.c, ps.lambda0, ps.fwhm, ps.delay);
        st(E, ps.idx, (float)((double)ld(E, ps.idx) + ps.amp * w));
    }
}

template <typename T>
__global__ void k_probes(Grid<T> g, ProbeSet p) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    const int t = g.counter[0] - 1;
    if (t < 0 || t >= p.cap) return;
    const int c = p.comp[q];
    const T *F = c < 3 ? g.f.E[c] : g.f.H[c - 3];
    double v = (double)ld(F, p.idx[q]);
    if (c >= 3) v *= p.hscale;
    p.rec[(long long)t * p.n + q] = v;
}

// ---------------------------------------------------------------------------
// gather / monitor / fills
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_gather(Grid<T> g, const T *F, int ex, int ey, int ez, double scale, double *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i >= ex) return;
    const double v = (double)ld(F, (long long)k * g.plane + (long long)j * g.px + i) * scale;
    dst[((long long)i * ey + j) * ez + k] = v;
}

__device__ __forceinline__ void atomic_max_pos(double *addr, double v) {
    // non-negative doubles order like their bit patterns as unsigned integers
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__global__ __launch_bounds__(256) void k_monitor(Grid<T> g, int planar, double *out) {
    __shared__ double smax[256];
    __shared__ double ssum[256];
    __shared__ int sbad[256];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 256 + tid;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    double mx = 0.0, sum = 0.0;
    int bad = 0;
    if (i < g.nx) {
        const long long o = (long long)k * g.plane + (long long)j * g.px + i;
        if (planar) {
            const double e = ld(g.f.E[1], o);
            if (i < g.nx && k < g.nz) { mx = e * e; sum = e * e; }
            const double hx = ld(g.f.H[0], o), hz = ld(g.f.H[2], o);
            sum += hx * hx + hz * hz;
            bad = !isfinite(e) || !isfinite(hx) || !isfinite(hz);
        } else {
            const double ex = ld(g.f.E[0], o), ey = ld(g.f.E[1], o), ez = ld(g.f.E[2], o);
            const double hx = ld(g.f.H[0], o), hy = ld(g.f.H[1], o), hz = ld(g.f.H[2], o);
            if (i < g.nx - 1 && j < g.ny - 1 && k < g.nz - 1) mx = ex * ex + ey * ey + ez * ez;
            sum = ex * ex + ey * ey + ez * ez + hx * hx + hy * hy + hz * hz;
            bad = !isfinite(ex) || !isfinite(ey) || !isfinite(ez) || !isfinite(hx) || !isfinite(hy) || !isfinite(hz);
        }
    }
    smax[tid] = mx; ssum[tid] = sum; sbad[tid] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            smax[tid] = fmax(smax[tid], smax[tid + s]);
            ssum[tid] += ssum[tid + s];
            sbad[tid] += sbad[tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        atomic_max_pos(&out[0], smax[0]);
        atomicAdd(&out[1], ssum[0]);
        if (sbad[0]) atomicAdd(&out[2], (double)sbad[0]);
    }
}

__global__ void k_fill_u8(uint8_t *p, uint8_t v, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = v;
}

__global__ void k_zero_u32(uint32_t *p, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = 0u;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
void launch_h3(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s) {
    dim3 blk(kBX, kBY), grd((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, (g.nz + g.kz - 1) / g.kz);
    hipLaunchKernelGGL(k_h3<T>, grd, blk, 0, s, g, tf, inc);
}
template <typename T>
void launch_e3(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s) {
    dim3 blk(kBX, kBY), grd((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, (g.nz + g.kz - 1) / g.kz);
    hipLaunchKernelGGL(k_e3<T>, grd, blk, 0, s, g, tf, inc);
}
template <typename T>
void launch_h2(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s) {
    dim3 blk(kBX, kBY), grd((g.nx + kBX - 1) / kBX, (g.nz + kBY - 1) / kBY);
    hipLaunchKernelGGL(k_h2<T>, grd, blk, 0, s, g, tf, inc);
}
template <typename T>
void launch_e2(const Grid<T> &g, const TfSet &tf, const IncParams &inc, hipStream_t s) {
    dim3 blk(kBX, kBY), grd((g.nx + kBX - 1) / kBX, (g.nz + kBY - 1) / kBY);
    hipLaunchKernelGGL(k_e2<T>, grd, blk, 0, s, g, tf, inc);
}
void launch_aux_h(const AuxLine &a, hipStream_t s) { hipLaunchKernelGGL(k_aux_h, dim3(1), dim3(1024), 0, s, a); }
void launch_aux_e(const AuxLine &a, const int *counter, hipStream_t s) {
    hipLaunchKernelGGL(k_aux_e, dim3(1), dim3(1024), 0, s, a, counter);
}
template <typename T>
void launch_point_source(const Grid<T> &g, const PointSrc &ps, hipStream_t s) {
    hipLaunchKernelGGL(k_point_source<T>, dim3(1), dim3(1), 0, s, g, ps);
}
template <typename T>
void launch_probes(const Grid<T> &g, const ProbeSet &p, hipStream_t s) {
    hipLaunchKernelGGL(k_probes<T>, dim3((p.n + 63) / 64), dim3(64), 0, s, g, p);
}
template <typename T>
void launch_gather(const Grid<T> &g, int comp, int ex, int ey, int ez, double scale, double *dst, int planar,
                   hipStream_t s) {
    const T *F = comp < 3 ? g.f.E[comp] : g.f.H[comp - 3];
    (void)planar;
    dim3 blk(256), grd((ex + 255) / 256, ey, ez);
    hipLaunchKernelGGL(k_gather<T>, grd, blk, 0, s, g, F, ex, ey, ez, scale, dst);
}
template <typename T>
void launch_monitor(const Grid<T> &g, int planar, double *out, hipStream_t s) {
    dim3 grd((g.nx + 255) / 256, g.ny, g.nz);
    hipLaunchKernelGGL(k_monitor<T>, grd, dim3(256), 0, s, g, planar, out);
}
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}
void launch_zero_bytes(void *p, size_t n, hipStream_t s) {
    size_t words = n / 4;
    size_t blocks = (words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t *)p, words);
}

#define LX_INST(T)                                                                                            \
    template void launch_h3<T>(const Grid<T> &, const TfSet &, const IncParams &, hipStream_t);              \
    template void launch_e3<T>(const Grid<T> &, const TfSet &, const IncParams &, hipStream_t);              \
    template void launch_h2<T>(const Grid<T> &, const TfSet &, const IncParams &, hipStream_t);              \
    template void launch_e2<T>(const Grid<T> &, const TfSet &, const IncParams &, hipStream_t);              \
    template void launch_point_source<T>(const Grid<T> &, const PointSrc &, hipStream_t);                    \
    template void launch_probes<T>(const Grid<T> &, const ProbeSet &, hipStream_t);                          \
    template void launch_gather<T>(const Grid<T> &, int, int, int, int, double, double *, int, hipStream_t); \
    template void launch_monitor<T>(const Grid<T> &, int, double *, hipStream_t);
LX_INST(float)
LX_INST(__half)

}  // namespace lx
