// lx_device.hpp -- device helpers shared by the kernel sources: FP32/FP16
// loads and stores, float4 arithmetic, element masks, CPML psi addressing,
// the incident waveform and field, and the extra block that advances the
// aux-1D line and records probes.
#pragma once
#include "lx_internal.hpp"
#include "../../include/luxum.h"

namespace lx {

// ---------------------------------------------------------------------------
// loads and stores
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float ld1(const T *p);
template <>
__device__ __forceinline__ float ld1<float>(const float *p) { return *p; }
template <>
__device__ __forceinline__ float ld1<__half>(const __half *p) { return __half2float(*p); }

template <typename T>
__device__ __forceinline__ void st1(T *p, float v);
template <>
__device__ __forceinline__ void st1<float>(float *p, float v) { *p = v; }
template <>
__device__ __forceinline__ void st1<__half>(__half *p, float v) { *p = __float2half(v); }

template <typename T>
__device__ __forceinline__ float4 ld4(const T *p);
template <>
__device__ __forceinline__ float4 ld4<float>(const float *p) { return *reinterpret_cast<const float4 *>(p); }
template <>
__device__ __forceinline__ float4 ld4<__half>(const __half *p) {
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    const __half2 a = *reinterpret_cast<const __half2 *>(&u.x), b = *reinterpret_cast<const __half2 *>(&u.y);
    const float2 fa = __half22float2(a), fb = __half22float2(b);
    return make_float4(fa.x, fa.y, fb.x, fb.y);
}

template <typename T>
__device__ __forceinline__ void st4(T *p, float4 v);
template <>
__device__ __forceinline__ void st4<float>(float *p, float4 v) { *reinterpret_cast<float4 *>(p) = v; }
template <>
__device__ __forceinline__ void st4<__half>(__half *p, float4 v) {
    const __half2 a = __floats2half2_rn(v.x, v.y), b = __floats2half2_rn(v.z, v.w);
    uint2 u;
    u.x = *reinterpret_cast<const unsigned *>(&a);
    u.y = *reinterpret_cast<const unsigned *>(&b);
    *reinterpret_cast<uint2 *>(p) = u;
}

__device__ __forceinline__ float4 f4(float v) { return make_float4(v, v, v, v); }
__device__ __forceinline__ float4 operator-(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ float4 operator+(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float4 operator*(float4 a, float4 b) { return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
__device__ __forceinline__ float4 operator*(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ float get(const float4 &v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
__device__ __forceinline__ void set(float4 &v, int e, float s) {
    if (e == 0) v.x = s; else if (e == 1) v.y = s; else if (e == 2) v.z = s; else v.w = s;
}
// keep the old value where the cell is outside the component's update range
__device__ __forceinline__ float4 sel(unsigned m, float4 nw, float4 od) {
    return make_float4((m & 1) ? nw.x : od.x, (m & 2) ? nw.y : od.y, (m & 4) ? nw.z : od.z, (m & 8) ? nw.w : od.w);
}
// 4-bit mask of elements e (x = i0+e) with lo <= x <= hi
__device__ __forceinline__ unsigned xmask(int i0, int lo, int hi) {
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) m |= (i0 + e >= lo && i0 + e <= hi) ? (1u << e) : 0u;
    return m;
}


__device__ __forceinline__ int slab_index(int x, int lo, int hi) { return x < lo ? x : lo + (x - hi); }

// psi arrays: axis a holds [H slot0, H slot1, E slot0, E slot1]; slot w of
// axis a belongs to split curl w of component (a + 2 - w) % 3.
template <typename T>
__device__ __forceinline__ float *psi_base(const Grid<T> &g, int a, int type_slot) {
    long long base = a == 0 ? 0 : (a == 1 ? 4 * g.scells[0] : 4 * (g.scells[0] + g.scells[1]));
    return g.psi + base + (long long)type_slot * g.scells[a];
}
// psi cell offset inside one array of axis a
template <typename T>
__device__ __forceinline__ long long psi_off(const Grid<T> &g, int a, int i, int j, int k) {
    if (a == 0) return ((long long)k * g.ny + j) * g.sw[0] + slab_index(i, g.slo[0], g.shi[0]);
    if (a == 1) return ((long long)k * g.sw[1] + slab_index(j, g.slo[1], g.shi[1])) * g.px + i;
    return ((long long)slab_index(k, g.slo[2], g.shi[2]) * g.ny + j) * g.px + i;
}

__device__ inline double waveform(int kind, double t, double c, double lambda0, double fwhm, double delay) {
    const double f0 = c / lambda0, T0 = 1.0 / f0;
    const double tau = fwhm * T0 / (2.0 * sqrt(log(2.0)));
    const double t0 = delay * tau;
    const double x = (t - t0) / tau;
    const double env = exp(-x * x);
    if (kind == LX_WAVE_GAUSSIAN) return env;
    return env * sin(2.0 * M_PI * f0 * (t - t0));
}

// Incident scalar at node coordinates (cells) and time level tp (steps).
__device__ inline double inc_value(const IncParams &in, double cx, double cy, double cz, double tp, bool e_node) {
    if (in.mode == LX_INC_AUX1D) {
        const double cc = in.axis == 0 ? cx : (in.axis == 1 ? cy : cz);
        const int m = (int)floor(cc) - in.x0;
        if (m < 0 || m >= in.L) return 0.0;
        if (e_node) return in.e[m];
        return m < in.L - 1 ? in.h[m] : 0.0;
    }
    const double pos = in.planar ? (cx * in.dx * in.kx + cz * in.dz * in.kz)
                                 : (cx * in.dx * in.kx + cy * in.dy * in.ky + cz * in.dz * in.kz);
    const double arg = pos + in.init_dx - tp * in.cdt;
    return exp(-(arg * arg) * in.inv_sigma);
}


// The aux line's E phase at node m: e_{t+1} from e_t and h_{t+1/2} (the
// source sample at node 0, a zero far end).  Written as explicit fused
// multiply-adds so that every caller rounds alike: the launch between two
// fused steps computes the nodes its face cells need with it while its extra
// block writes the whole line with it.
__device__ __forceinline__ double aux_e_next(const AuxLine &a, int m, int t) {
    if (m == 0) return waveform(a.wave, (t + 1) * a.dt, a.c, a.lambda0, a.fwhm, a.delay);
    if (m == a.L - 1) return 0.0;
    return fma(-a.S, a.h[m] - a.h[m - 1], a.e[m]);
}
// The H phase at node m (m < L-1): h_{t+1/2} from h_{t-1/2} and e_t.
__device__ __forceinline__ double aux_h_next(const double *h, const double *e, double S, int m) {
    return fma(-S, e[m + 1] - e[m], h[m]);
}

// H half (IS_H): aux line H phase, E probes of step t-1.
// E half:        aux line E phase (source sample t+1), H probes of step t.
template <typename T, bool IS_H>
__device__ inline void extra_block(const Grid<T> &g, const AuxLine &aux, const ProbeArgs &pr, Launch l, int t) {
    if (IS_H) {
        if (l.aux)
            for (int m = threadIdx.x; m < aux.L - 1; m += blockDim.x) aux.h[m] = aux_h_next(aux.h, aux.e, aux.S, m);
        if (l.eprobes && t >= 1 && t - 1 < pr.cap)
            for (int q = threadIdx.x; q < pr.n; q += blockDim.x)
                if (pr.comp[q] < 3)
                    pr.rec[(long long)(t - 1) * pr.n + q] = (double)ld1(g.F + pr.comp[q] * g.fstride + pr.idx[q]);
    } else {
        if (l.aux)
            for (int m = threadIdx.x; m < aux.L; m += blockDim.x) aux.e[m] = aux_e_next(aux, m, t);
        if (l.probes && t < pr.cap)
            for (int q = threadIdx.x; q < pr.n; q += blockDim.x)
                if (pr.comp[q] >= 3)
                    pr.rec[(long long)t * pr.n + q] = (double)ld1(g.F + pr.comp[q] * g.fstride + pr.idx[q]) * pr.hscale;
    }
}


}  // namespace lx
