// lx_step3.hip -- fused 3D Yee timestep: the H and the E half step of one
// timestep in one z-march, reading one buffer set (fields, psi) and writing
// the other.
//
// The two half-step kernels read every E value twice per step (once by the
// H update, once by the E update) and every H value twice; here each old
// value is read once and each new value written once, 12*s + 1 bytes per
// cell and step plus the slab psi, against 18*s + 1 for the half-step pair
// (maxwell_three_dim_pml.cc:182-255 is the update order followed).
//
// Work unit: one wavefront per (x chunk of 64*V cells, grid row j, block of
// kz planes), V consecutive cells per lane, marching along z.  Units are
// numbered chunk first, so the chunks of a row share a block (and an L2) and
// read each other's edge cells while those are cached.  Per plane k:
//   loads: E_old at (j, k+1) (the z queue), at (j+1, k) and (j-1, k), H_old
//     at (j, k) and (j-1, k), the material ids, the slab psi and the edge
//     cells x0-1 and x0+64V.  All loads of a plane are issued before its
//     arithmetic and none sits in a branch that depends on loaded data, so a
//     plane costs one memory round trip.
//   H: H_new(j, k) of the three components; Hx, Hz of row j-1 (needed by Ex,
//     Ez of row j) and Hy, Hz of cell x0-1 (needed by Ey, Ez of the chunk's
//     first cell) are recomputed here and stored by their own wavefront.
//   E: E_new(j, k) from H_new at (j, k), (j, k-1) (registers), (j-1, k) and
//     x-1 (previous element or lane).
// The plane below the block (k0-1) is recomputed for Hx and Hy only.  Cells
// outside a component's update range are written with their old value, so
// both buffers always hold a complete state.  Blocks only read the old
// buffer: there is no ordering between blocks and no barrier after the
// coefficient load.  The H TF/SF corrections are applied to H_old beforehand
// (k_tfsf_hpre), the E ones to E_new afterwards (k_tfsf): this is the bulk
// update only.  V = 2 (two cells per lane, 8- or 4-byte loads) halves the
// memory instructions and address arithmetic per cell; V = 4 (FP16 storage
// only) gives 8-byte field loads.
#include <type_traits>

#include "lx_batch.hpp"
#include "lx_device.hpp"

namespace lx {

template <typename T>
__device__ __forceinline__ float *psi_new(const Grid<T> &g, int a, int slot) {
    return g.psin + (psi_base(g, a, slot) - g.psi);
}

// Accesses at a wave-uniform base plus a 32-bit per-lane byte offset.  uo()
// passes a per-plane element offset through readfirstlane, so that the
// compiler does not fold the lane offset into one 64-bit induction variable
// per access (the base pointer keeps its global address space).
__device__ __forceinline__ long long uo(long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ const T *at(const T *base, unsigned boff) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
template <typename T>
__device__ __forceinline__ T *at(T *base, unsigned boff) {
    return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff);
}
template <typename T>
__device__ __forceinline__ float ldb(const T *base, unsigned boff) {
    return ld1(at(base, boff));
}
__device__ __forceinline__ float lane0(float v) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
}

// V consecutive x cells of one lane
template <int V>
struct vf {
    float v[V];
};
template <int V>
__device__ __forceinline__ vf<V> vs(float a) {
    vf<V> r;
#pragma unroll
    for (int e = 0; e < V; ++e) r.v[e] = a;
    return r;
}
#define LX_VOP(op)                                                                         \
    template <int V>                                                                       \
    __device__ __forceinline__ vf<V> operator op(const vf<V> &a, const vf<V> &b) {         \
        vf<V> r;                                                                           \
        _Pragma("unroll") for (int e = 0; e < V; ++e) r.v[e] = a.v[e] op b.v[e];           \
        return r;                                                                          \
    }
LX_VOP(+)
LX_VOP(-)
LX_VOP(*)
#undef LX_VOP
// keep the old value where element e is outside the update range (bit e of m clear)
template <int V>
__device__ __forceinline__ vf<V> vsel(unsigned m, const vf<V> &nw, const vf<V> &od) {
    vf<V> r;
#pragma unroll
    for (int e = 0; e < V; ++e) r.v[e] = (m >> e) & 1u ? nw.v[e] : od.v[e];
    return r;
}
// bit e: lo <= i0 + e <= hi
template <int V>
__device__ __forceinline__ unsigned vmask(int i0, int lo, int hi) {
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < V; ++e) m |= (i0 + e >= lo && i0 + e <= hi) ? 1u << e : 0u;
    return m;
}

// loads and stores of V consecutive values at a byte offset
template <typename T, int V>
__device__ __forceinline__ vf<V> ldv(const T *base, unsigned boff);
template <>
__device__ __forceinline__ vf<1> ldv<float, 1>(const float *b, unsigned o) { return vf<1>{{*at(b, o)}}; }
template <>
__device__ __forceinline__ vf<2> ldv<float, 2>(const float *b, unsigned o) {
    const float2 x = *reinterpret_cast<const float2 *>(at(b, o));
    return vf<2>{{x.x, x.y}};
}
template <>
__device__ __forceinline__ vf<1> ldv<__half, 1>(const __half *b, unsigned o) { return vf<1>{{__half2float(*at(b, o))}}; }
template <>
__device__ __forceinline__ vf<2> ldv<__half, 2>(const __half *b, unsigned o) {
    const float2 x = __half22float2(*reinterpret_cast<const __half2 *>(at(b, o)));
    return vf<2>{{x.x, x.y}};
}
template <>
__device__ __forceinline__ vf<4> ldv<float, 4>(const float *b, unsigned o) {
    const float4 x = *reinterpret_cast<const float4 *>(at(b, o));
    return vf<4>{{x.x, x.y, x.z, x.w}};
}
template <>
__device__ __forceinline__ vf<4> ldv<__half, 4>(const __half *b, unsigned o) {
    const uint2 u = *reinterpret_cast<const uint2 *>(at(b, o));
    const float2 lo = __half22float2(*reinterpret_cast<const __half2 *>(&u.x));
    const float2 hi = __half22float2(*reinterpret_cast<const __half2 *>(&u.y));
    return vf<4>{{lo.x, lo.y, hi.x, hi.y}};
}
template <typename T, int V>
__device__ __forceinline__ void stv(T *base, unsigned boff, const vf<V> &v);
// stores of the written buffer are non-temporal: nothing reads them in the same step
typedef float f2nt __attribute__((ext_vector_type(2)));
typedef float f4nt __attribute__((ext_vector_type(4)));
typedef unsigned u2nt __attribute__((ext_vector_type(2)));
template <>
__device__ __forceinline__ void stv<float, 1>(float *b, unsigned o, const vf<1> &v) { __builtin_nontemporal_store(v.v[0], at(b, o)); }
template <>
__device__ __forceinline__ void stv<float, 2>(float *b, unsigned o, const vf<2> &v) {
    const f2nt x = {v.v[0], v.v[1]};
    __builtin_nontemporal_store(x, reinterpret_cast<f2nt *>(at(b, o)));
}
template <>
__device__ __forceinline__ void stv<__half, 1>(__half *b, unsigned o, const vf<1> &v) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half(v.v[0])), reinterpret_cast<unsigned short *>(at(b, o)));
}
template <>
__device__ __forceinline__ void stv<__half, 2>(__half *b, unsigned o, const vf<2> &v) {
    const __half2 h = __floats2half2_rn(v.v[0], v.v[1]);
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned *>(&h), reinterpret_cast<unsigned *>(at(b, o)));
}
template <>
__device__ __forceinline__ void stv<float, 4>(float *b, unsigned o, const vf<4> &v) {
    const f4nt x = {v.v[0], v.v[1], v.v[2], v.v[3]};
    __builtin_nontemporal_store(x, reinterpret_cast<f4nt *>(at(b, o)));
}
template <>
__device__ __forceinline__ void stv<__half, 4>(__half *b, unsigned o, const vf<4> &v) {
    const __half2 lo = __floats2half2_rn(v.v[0], v.v[1]), hi = __floats2half2_rn(v.v[2], v.v[3]);
    u2nt u;
    u.x = *reinterpret_cast<const unsigned *>(&lo);
    u.y = *reinterpret_cast<const unsigned *>(&hi);
    __builtin_nontemporal_store(u, reinterpret_cast<u2nt *>(at(b, o)));
}
// coefficient row of the x axis (float table) at x index i0
template <int V>
__device__ __forceinline__ vf<V> ldtab(const float *row, int i0) {
    return ldv<float, V>(row, (unsigned)i0 * 4u);
}

// x+1 neighbour of every element; the last element of lane 63 takes `edge`
template <int V>
__device__ __forceinline__ vf<V> xnext(const vf<V> &a, float edge, int lane) {
    vf<V> r;
#pragma unroll
    for (int e = 0; e + 1 < V; ++e) r.v[e] = a.v[e + 1];
    const float n = wave_next(a.v[0]);
    r.v[V - 1] = lane == 63 ? edge : n;
    return r;
}
// x-1 neighbour of every element; the first element of lane 0 takes `edge`
template <int V>
__device__ __forceinline__ vf<V> xprev(const vf<V> &a, float edge, int lane) {
    vf<V> r;
    const float p = wave_prev(a.v[V - 1]);
    r.v[0] = lane == 0 ? edge : p;
#pragma unroll
    for (int e = 1; e < V; ++e) r.v[e] = a.v[e - 1];
    return r;
}

// Split curl of one component, d1*k1 - d2*k2, with the CPML recursion of
// each term on slab cells (p = b*p + c*d, added for term 1, subtracted for
// term 2).  Every update of the kernels goes through it, so a recomputed
// neighbour equals the value its owner stores.  The fused multiply-adds are
// written out and contraction is off here, so that every kernel shape (one
// unit per wavefront, rows sharing neighbours, the batched step) rounds the
// same way whatever the surrounding code.
template <bool PML>
__device__ __forceinline__ float curl2(float d1, float k1, bool s1, float &p1, float b1, float c1, float d2, float k2,
                                       bool s2, float &p2, float b2, float c2) {
#pragma clang fp contract(off)
    float acc = __builtin_fmaf(d1, k1, -(d2 * k2));
    if (PML && s1) { p1 = __builtin_fmaf(p1, b1, d1 * c1); acc = acc + p1; }
    if (PML && s2) { p2 = __builtin_fmaf(p2, b2, d2 * c2); acc = acc - p2; }
    return acc;
}
template <bool PML, int V>
__device__ __forceinline__ vf<V> curl2(const vf<V> &d1, const vf<V> &k1, bool s1, vf<V> &p1, const vf<V> &b1,
                                       const vf<V> &c1, const vf<V> &d2, const vf<V> &k2, bool s2, vf<V> &p2,
                                       const vf<V> &b2, const vf<V> &c2) {
    vf<V> r;
#pragma unroll
    for (int e = 0; e < V; ++e)
        r.v[e] = curl2<PML>(d1.v[e], k1.v[e], s1, p1.v[e], b1.v[e], c1.v[e], d2.v[e], k2.v[e], s2, p2.v[e], b2.v[e],
                            c2.v[e]);
    return r;
}

// E update e + cb*acc of the lane's V cells (one fused multiply-add each)
template <int V>
__device__ __forceinline__ vf<V> eupd(const vf<V> &e0, const vf<V> &cb, const vf<V> &acc) {
    vf<V> r;
#pragma unroll
    for (int e = 0; e < V; ++e) r.v[e] = __builtin_fmaf(cb.v[e], acc.v[e], e0.v[e]);
    return r;
}

// Coefficient of one E component for the lane's V cells: tab is that
// component's 256-entry table in LDS (the reference keeps one Cb array per
// component, maxwell_three_dim_pml.cc:109-128).
template <int V>
__device__ __forceinline__ vf<V> cbv(const float *tab, const unsigned (&id)[V]) {
    vf<V> r;
#pragma unroll
    for (int e = 0; e < V; ++e) r.v[e] = tab[id[e]];
    return r;
}

// One unit (x chunk, row, z block) of the step; cbs holds the coefficient
// tables of Ex, Ey and Ez (256 entries each) of the grid's materials in LDS.  PM: CPML axes the launch handles (bit 0:
// x slab, 1: y, 2: z).  A CPML grid runs two launches: the units whose rows
// and planes lie outside the y and z slabs (PM = 1, x slab only, far fewer
// registers) and the rest (PM = 7).
template <typename T, int PM, int V>
__device__ __forceinline__ void step3_unit(const Grid<T> &g, const UnitMap &um, int unit, const float *cbs) {
    constexpr bool PML = PM != 0;
    using F = vf<V>;
    const int lane = threadIdx.x & 63;
    const int nx = g.nx, ny = g.ny, nz = g.nz, px = g.px;
    const int nch = (nx + 64 * V - 1) / (64 * V);
    const int ch = unit % nch;
    int jr, k0, k1;
    unit_span(um, ny, nz, unit / nch, jr, k0, k1);
    const int x0 = ch * 64 * V;
    const long long plane = g.plane, fs = g.fstride;
    const int tn = g.tn;
    const float *tab = g.tab;
    const bool has_m = jr >= 1;                    // row j-1 exists (its Hx, Hz feed Ex, Ez)
    const int jm = has_m ? jr - 1 : 0;
    const int jp = min(jr + 1, ny - 1);
    const bool rowx = jr < ny - 1;                 // Hx, Hz, Ey rows
    const bool jin = jr >= 1 && jr <= ny - 2;      // Ex, Ez rows
    const float khy = tab[(1 * 6 + TK_KH) * tn + min(jr, ny - 2)], khym = tab[(1 * 6 + TK_KH) * tn + jm];
    const float key = tab[(1 * 6 + TK_KE) * tn + jr];
    const bool sy = (PM & 2) && (jr < g.slo[1] || jr >= g.shi[1]);
    const bool sym = (PM & 2) && has_m && (jm < g.slo[1] || jm >= g.shi[1]);
    const float bhy = sy ? tab[(1 * 6 + TK_BH) * tn + jr] : 0.f, chy = sy ? tab[(1 * 6 + TK_CH) * tn + jr] : 0.f;
    const float bhym = sym ? tab[(1 * 6 + TK_BH) * tn + jm] : 0.f, chym = sym ? tab[(1 * 6 + TK_CH) * tn + jm] : 0.f;
    const float bey = sy ? tab[(1 * 6 + TK_BE) * tn + jr] : 0.f, cey = sy ? tab[(1 * 6 + TK_CE) * tn + jr] : 0.f;
    const int siy = sy ? slab_index(jr, g.slo[1], g.shi[1]) : 0;
    const int siym = sym ? slab_index(jm, g.slo[1], g.shi[1]) : 0;
    const F vkhy = vs<V>(khy), vkhym = vs<V>(khym), vkey = vs<V>(key);
    const F vbhy = vs<V>(bhy), vchy = vs<V>(chy), vbhym = vs<V>(bhym), vchym = vs<V>(chym);
    const F vbey = vs<V>(bey), vcey = vs<V>(cey);

    // every lane is inside the padded row (x0 + 64V - 1 < px); lanes past the
    // grid read the zero padding and write it back unchanged
    const int i0 = x0 + lane * V;                  // the lane's first cell
    const bool nxt = x0 + 64 * V < nx;             // x+1 of the chunk's last cell lies in the next chunk
    const bool lft = x0 > 0;                       // x-1 of the chunk's first cell lies in the previous chunk
    const unsigned m_hx = vmask<V>(i0, 0, nx - 1), m_in = vmask<V>(i0, 0, nx - 2);
    const unsigned m_ex = jin ? vmask<V>(i0, 0, nx - 2) : 0u;
    const unsigned m_ey = rowx ? vmask<V>(i0, 1, nx - 2) : 0u;
    const unsigned m_ez = jin ? vmask<V>(i0, 1, nx - 2) : 0u;
    const F khx = ldtab<V>(tab + (0 * 6 + TK_KH) * tn, i0), kex = ldtab<V>(tab + (0 * 6 + TK_KE) * tn, i0);
    // the x slab bounds are multiples of 4, so a lane's V cells are wholly in or out
    const bool sxw = (PM & 1) && g.pml && (x0 < g.slo[0] || x0 + 64 * V > g.shi[0]);   // chunk meets the x slab
    const bool sx = sxw && (i0 < g.slo[0] || i0 >= g.shi[0]);
    const int xs = sx ? (i0 < g.slo[0] ? i0 : g.slo[0] + (i0 - g.shi[0])) : 0;
    F bhx = vs<V>(0.f), chx = vs<V>(0.f), bex = vs<V>(0.f), cex = vs<V>(0.f);
    if (sx) {
        bhx = ldtab<V>(tab + (0 * 6 + TK_BH) * tn, i0);
        chx = ldtab<V>(tab + (0 * 6 + TK_CH) * tn, i0);
        bex = ldtab<V>(tab + (0 * 6 + TK_BE) * tn, i0);
        cex = ldtab<V>(tab + (0 * 6 + TK_CE) * tn, i0);
    }
    // cell x0-1 (wave-uniform): its x coefficients and slab
    const int xl = lft ? x0 - 1 : 0;
    const float khxl = tab[(0 * 6 + TK_KH) * tn + xl];
    const bool sxl = (PM & 1) && lft && (xl < g.slo[0] || xl >= g.shi[0]);
    const int xsl = sxl ? (xl < g.slo[0] ? xl : g.slo[0] + (xl - g.shi[0])) : 0;
    const float bhxl = sxl ? tab[(0 * 6 + TK_BH) * tn + xl] : 0.f, chxl = sxl ? tab[(0 * 6 + TK_CH) * tn + xl] : 0.f;
    // per-lane byte offsets inside a plane (rows j, j-1, j+1) and inside a psi
    // row (y and z slabs: x index; x slab: slab x index)
    constexpr unsigned S = sizeof(T);
    const unsigned rj = (unsigned)(jr * px + i0) * S, rjm = (unsigned)(jm * px + i0) * S;
    const unsigned rjp = (unsigned)(jp * px + i0) * S, ri = (unsigned)i0 * 4u, rx = (unsigned)xs * 4u;
    const unsigned re = (unsigned)(jr * px + i0);  // material ids (one byte per cell)
    // byte offsets of the edge cells x0+64V (rows j, j-1) and x0-1 (rows j,
    // j+1; psi rows): equal in every lane, derived from the lane offsets so
    // that they share their base registers
    const unsigned bn = rj + (unsigned)(64 * V - lane * V) * S, bnm = rjm + (unsigned)(64 * V - lane * V) * S;
    const unsigned bl = rj - (unsigned)(lane * V + 1) * S, blp = rjp - (unsigned)(lane * V + 1) * S;
    const unsigned ril = ri - (unsigned)(lane * V + 1) * 4u;
    const unsigned rxl = (unsigned)xsl * 4u;

    // z queue at k0 (row j: Ex, Ey, Ez; row j-1: Ey; cell x0-1: Ex) and H_new(j, k0-1)
    F ex0, ey0, ez0, eym0;
    float exl0 = 0.f;
    F hxm = vs<V>(0.f), hym = vs<V>(0.f);
    if (k0 > 0) {
        const int kh = k0 - 1;
        const T *Fk = g.F + kh * plane;
        const F exh = ldv<T, V>(Fk, rj), eyh = ldv<T, V>(Fk + fs, rj), ezh = ldv<T, V>(Fk + 2 * fs, rj);
        ex0 = ldv<T, V>(Fk + plane, rj);
        ey0 = ldv<T, V>(Fk + plane + fs, rj);
        ez0 = ldv<T, V>(Fk + plane + 2 * fs, rj);
        eym0 = ldv<T, V>(Fk + plane + fs, rjm);
        if (lft) exl0 = ldb(Fk + plane, bl);
        const F ezph = ldv<T, V>(Fk + 2 * fs, rjp);
        const float ezNh = nxt ? ldb(Fk + 2 * fs, bn) : 0.f;
        const F hxh = ldv<T, V>(Fk + 3 * fs, rj), hyh = ldv<T, V>(Fk + 4 * fs, rj);
        const bool sz = (PM & 4) && (kh < g.slo[2] || kh >= g.shi[2]);
        F py0 = vs<V>(0.f), pz1 = vs<V>(0.f), pz0 = vs<V>(0.f), px1 = vs<V>(0.f);
        if (sy) py0 = ldv<float, V>(psi_base(g, 1, 0) + ((long long)kh * g.sw[1] + siy) * px, ri);
        if (sz) {
            const long long zb = ((long long)slab_index(kh, g.slo[2], g.shi[2]) * ny + jr) * px;
            pz1 = ldv<float, V>(psi_base(g, 2, 1) + zb, ri);
            pz0 = ldv<float, V>(psi_base(g, 2, 0) + zb, ri);
        }
        if (sx) px1 = ldv<float, V>(psi_base(g, 0, 1) + ((long long)kh * ny + jr) * g.sw[0], rx);
        const float khz = tab[(2 * 6 + TK_KH) * tn + kh];
        const float bhz = sz ? tab[(2 * 6 + TK_BH) * tn + kh] : 0.f, chz = sz ? tab[(2 * 6 + TK_CH) * tn + kh] : 0.f;
        const F vkhz = vs<V>(khz), vbhz = vs<V>(bhz), vchz = vs<V>(chz);
        const F ezi = xnext<V>(ezh, ezNh, lane);
        hxm = hxh;
        if (rowx) {   // Hx: d1 = dEz/dy, d2 = dEy/dz
            const F acc = curl2<PML, V>(ezph - ezh, vkhy, sy, py0, vbhy, vchy, ey0 - eyh, vkhz, sz, pz1, vbhz, vchz);
            hxm = vsel<V>(m_hx, hxh - acc, hxh);
        }
        {             // Hy: d1 = dEx/dz, d2 = dEz/dx
            const F acc = curl2<PML, V>(ex0 - exh, vkhz, sz, pz0, vbhz, vchz, ezi - ezh, khx, sx, px1, bhx, chx);
            hym = vsel<V>(m_in, hyh - acc, hyh);
        }
    } else {
        ex0 = ldv<T, V>(g.F, rj);
        ey0 = ldv<T, V>(g.F + fs, rj);
        ez0 = ldv<T, V>(g.F + 2 * fs, rj);
        eym0 = ldv<T, V>(g.F + fs, rjm);
        if (lft) exl0 = ldb(g.F, bl);
    }

    for (int k = k0; k < k1; ++k) {
        const bool kin = k < nz - 1;
        const bool kine = k >= 1 && kin;
        const long long kp = uo(k * plane);
        const T *Fk = g.F + kp;
        T *Fnk = g.Fn + kp;
        // ---------------- loads of the plane
        // E of plane k+1 (the z queue); the grid's last plane reloads its own
        // plane instead (nothing reads the values there), so no branch and no
        // register clearing precede the loads
        const long long kn = kin ? plane : 0;
        const F ex1 = ldv<T, V>(Fk + kn, rj), ey1 = ldv<T, V>(Fk + kn + fs, rj);
        const F ez1 = ldv<T, V>(Fk + kn + 2 * fs, rj), eym1 = ldv<T, V>(Fk + kn + fs, rjm);
        float exl1 = 0.f;
        const F ezp = ldv<T, V>(Fk + 2 * fs, rjp), exq = ldv<T, V>(Fk, rjp);
        F hx = ldv<T, V>(Fk + 3 * fs, rj), hy = ldv<T, V>(Fk + 4 * fs, rj), hz = ldv<T, V>(Fk + 5 * fs, rj);
        F hxj = ldv<T, V>(Fk + 3 * fs, rjm), hzj = ldv<T, V>(Fk + 5 * fs, rjm);
        const F ezm = ldv<T, V>(Fk + 2 * fs, rjm), exm = ldv<T, V>(Fk, rjm);
        unsigned id[V];
        if constexpr (V == 1) {
            id[0] = g.mat[kp + re];
        } else if constexpr (V == 2) {
            const uchar2 u = *reinterpret_cast<const uchar2 *>(g.mat + kp + re);
            id[0] = u.x;
            id[1] = u.y;
        } else {
            const unsigned u = *reinterpret_cast<const unsigned *>(g.mat + kp + re);
#pragma unroll
            for (int e = 0; e < V; ++e) id[e] = (u >> (8 * e)) & 0xffu;
        }
        float ezN = 0.f, eyN = 0.f, eymN = 0.f;
        if (nxt) {
            ezN = ldb(Fk + 2 * fs, bn);
            eyN = ldb(Fk + fs, bn);
            eymN = ldb(Fk + fs, bnm);
        }
        float hyl = 0.f, hzl = 0.f, ezl = 0.f, eyl = 0.f, exlp = 0.f;
        if (lft) {
            exl1 = ldb(Fk + kn, bl);
            hyl = ldb(Fk + 4 * fs, bl);
            hzl = ldb(Fk + 5 * fs, bl);
            ezl = ldb(Fk + 2 * fs, bl);
            eyl = ldb(Fk + fs, bl);
            exlp = ldb(Fk, blp);
        }
        const bool sz = (PM & 4) && (k < g.slo[2] || k >= g.shi[2]);
        // psi: H of row j (p*), H of row j-1 (q*), E of row j (e*), H of cell
        // x0-1 (l*); row bases of the read (P*) and written (Q*) arrays, slot s
        // of axis a lying s*scells[a] further on
        F py0 = vs<V>(0.f), py1 = vs<V>(0.f), pz0 = vs<V>(0.f), pz1 = vs<V>(0.f);
        F qy0 = vs<V>(0.f), qy1 = vs<V>(0.f), qz1 = vs<V>(0.f);
        F ey2 = vs<V>(0.f), ey3 = vs<V>(0.f), ez2 = vs<V>(0.f), ez3 = vs<V>(0.f);
        F px0, px1, ex2, ex3, qx0;   // x-slab psi, defined below on every path
        float ly1 = 0.f, lz0 = 0.f, lx0 = 0.f, lx1 = 0.f;
        const float *PY = nullptr, *PZ = nullptr, *PX = nullptr;
        float *QY = nullptr, *QZ = nullptr, *QX = nullptr;
        const long long cy = g.scells[1], cz = g.scells[2], cx = g.scells[0];
        if (sy) {
            const long long yb = uo(((long long)k * g.sw[1] + siy) * px);
            PY = psi_base(g, 1, 0) + yb;
            QY = psi_new(g, 1, 0) + yb;
            py0 = ldv<float, V>(PY, ri);
            py1 = ldv<float, V>(PY + cy, ri);
            ey2 = ldv<float, V>(PY + 2 * cy, ri);
            ey3 = ldv<float, V>(PY + 3 * cy, ri);
            if (lft) ly1 = ldb(PY + cy, ril);
        }
        if (sym) {
            const float *PYm = psi_base(g, 1, 0) + uo(((long long)k * g.sw[1] + siym) * px);
            qy0 = ldv<float, V>(PYm, ri);
            qy1 = ldv<float, V>(PYm + cy, ri);
        }
        if (sz) {
            const long long zs = (long long)slab_index(k, g.slo[2], g.shi[2]) * ny;
            const long long zb = uo((zs + jr) * px);
            PZ = psi_base(g, 2, 0) + zb;
            QZ = psi_new(g, 2, 0) + zb;
            pz0 = ldv<float, V>(PZ, ri);
            pz1 = ldv<float, V>(PZ + cz, ri);
            ez2 = ldv<float, V>(PZ + 2 * cz, ri);
            ez3 = ldv<float, V>(PZ + 3 * cz, ri);
            qz1 = ldv<float, V>(psi_base(g, 2, 1) + uo((zs + jm) * px), ri);
            if (lft) lz0 = ldb(PZ, ril);
        }
        if (sxw || sxl) {
            const long long xb = uo(((long long)k * ny + jr) * g.sw[0]);
            PX = psi_base(g, 0, 0) + xb;
            QX = psi_new(g, 0, 0) + xb;
            if (sxl) {
                lx0 = ldb(PX, rxl);
                lx1 = ldb(PX + cx, rxl);
            }
        }
        // interior launch: every lane of a chunk that meets the slab loads
        // (lanes outside the slab read its first cell and never use the
        // value), so no per-lane branch precedes the loads; the shell keeps
        // the per-lane branch (one register fewer holds it at 5 waves/SIMD)
        const bool xload = PM == 1 ? sxw : sx;
        if (xload) {
            px0 = ldv<float, V>(PX, rx);
            px1 = ldv<float, V>(PX + cx, rx);
            ex2 = ldv<float, V>(PX + 2 * cx, rx);
            ex3 = ldv<float, V>(PX + 3 * cx, rx);
            qx0 = ldv<float, V>(psi_base(g, 0, 0) + uo(((long long)k * ny + jm) * g.sw[0]), rx);
        } else {
            px0 = px1 = ex2 = ex3 = qx0 = vs<V>(0.f);
        }
        const float khz = tab[(2 * 6 + TK_KH) * tn + min(k, nz - 2)];
        const float bhz = sz ? tab[(2 * 6 + TK_BH) * tn + k] : 0.f, chz = sz ? tab[(2 * 6 + TK_CH) * tn + k] : 0.f;
        const float kez = tab[(2 * 6 + TK_KE) * tn + k];
        const float bez = sz ? tab[(2 * 6 + TK_BE) * tn + k] : 0.f, cez = sz ? tab[(2 * 6 + TK_CE) * tn + k] : 0.f;
        const F vkhz = vs<V>(khz), vbhz = vs<V>(bhz), vchz = vs<V>(chz);
        const F vkez = vs<V>(kez), vbez = vs<V>(bez), vcez = vs<V>(cez);
        const F ezi = xnext<V>(ez0, ezN, lane), eyi = xnext<V>(ey0, eyN, lane), eymi = xnext<V>(eym0, eymN, lane);

        // ---------------- H half step, row j
        if (rowx && kin) {   // Hx: d1 = dEz/dy, d2 = dEy/dz
            const F acc = curl2<PML, V>(ezp - ez0, vkhy, sy, py0, vbhy, vchy, ey1 - ey0, vkhz, sz, pz1, vbhz, vchz);
            hx = vsel<V>(m_hx, hx - acc, hx);
        }
        if (kin) {           // Hy: d1 = dEx/dz, d2 = dEz/dx
            const F acc = curl2<PML, V>(ex1 - ex0, vkhz, sz, pz0, vbhz, vchz, ezi - ez0, khx, sx, px1, bhx, chx);
            hy = vsel<V>(m_in, hy - acc, hy);
        }
        if (rowx) {          // Hz: d1 = dEy/dx, d2 = dEx/dy
            const F acc = curl2<PML, V>(eyi - ey0, khx, sx, px0, bhx, chx, exq - ex0, vkhy, sy, py1, vbhy, vchy);
            hz = vsel<V>(m_in, hz - acc, hz);
        }
        // ---------------- Hx, Hz of row j-1 (only Ex, Ez of row j >= 1 use them)
        if (has_m) {
            if (kin) {       // Hx(j-1): d1 = Ez(j) - Ez(j-1), d2 = dEy(j-1)/dz
                const F acc =
                    curl2<PML, V>(ez0 - ezm, vkhym, sym, qy0, vbhym, vchym, eym1 - eym0, vkhz, sz, qz1, vbhz, vchz);
                hxj = vsel<V>(m_hx, hxj - acc, hxj);
            }
            {                // Hz(j-1): d1 = dEy(j-1)/dx, d2 = Ex(j) - Ex(j-1)
                const F acc =
                    curl2<PML, V>(eymi - eym0, khx, sx, qx0, bhx, chx, ex0 - exm, vkhym, sym, qy1, vbhym, vchym);
                hzj = vsel<V>(m_in, hzj - acc, hzj);
            }
        }
        // ---------------- Hy, Hz of cell x0-1 (only the chunk's first Ey, Ez use them)
        if (lft) {
            if (kin) {       // Hy(x0-1): d1 = dEx/dz, d2 = Ez(x0) - Ez(x0-1)
                const float acc = curl2<PML>(exl1 - exl0, khz, sz, lz0, bhz, chz, lane0(ez0.v[0]) - ezl, khxl, sxl, lx1,
                                             bhxl, chxl);
                hyl = hyl - acc;
            }
            if (rowx) {      // Hz(x0-1): d1 = Ey(x0) - Ey(x0-1), d2 = dEx/dy
                const float acc = curl2<PML>(lane0(ey0.v[0]) - eyl, khxl, sxl, lx0, bhxl, chxl, exlp - exl0, khy, sy, ly1,
                                             bhy, chy);
                hzl = hzl - acc;
            }
        }
        stv<T, V>(Fnk + 3 * fs, rj, hx);
        stv<T, V>(Fnk + 4 * fs, rj, hy);
        stv<T, V>(Fnk + 5 * fs, rj, hz);
        if (PML) {
            if (sy && rowx && kin) stv<float, V>(QY, ri, py0);
            if (sz && rowx && kin) stv<float, V>(QZ + cz, ri, pz1);
            if (sz && kin) stv<float, V>(QZ, ri, pz0);
            if (sx && kin) stv<float, V>(QX + cx, rx, px1);
            if (sx && rowx) stv<float, V>(QX, rx, px0);
            if (sy && rowx) stv<float, V>(QY + cy, ri, py1);
        }

        // ---------------- E half step, row j
        const F hyp = xprev<V>(hy, hyl, lane), hzp = xprev<V>(hz, hzl, lane);
        F ex = ex0, ey = ey0, ez = ez0;
        if (kine) {
            {                // Ex: d1 = dHz/dy, d2 = dHy/dz
                const F acc = curl2<PML, V>(hz - hzj, vkey, sy, ey2, vbey, vcey, hy - hym, vkez, sz, ez3, vbez, vcez);
                ex = vsel<V>(m_ex, eupd<V>(ex0, cbv<V>(cbs, id), acc), ex0);
            }
            {                // Ey: d1 = dHx/dz, d2 = dHz/dx
                const F acc = curl2<PML, V>(hx - hxm, vkez, sz, ez2, vbez, vcez, hz - hzp, kex, sx, ex3, bex, cex);
                ey = vsel<V>(m_ey, eupd<V>(ey0, cbv<V>(cbs + 256, id), acc), ey0);
            }
        }
        if (kin) {           // Ez: d1 = dHy/dx, d2 = dHx/dy
            const F acc = curl2<PML, V>(hy - hyp, kex, sx, ex2, bex, cex, hx - hxj, vkey, sy, ey3, vbey, vcey);
            ez = vsel<V>(m_ez, eupd<V>(ez0, cbv<V>(cbs + 512, id), acc), ez0);
        }
        stv<T, V>(Fnk, rj, ex);
        stv<T, V>(Fnk + fs, rj, ey);
        stv<T, V>(Fnk + 2 * fs, rj, ez);
        if (PML) {
            if (sy && kine && jin) stv<float, V>(QY + 2 * cy, ri, ey2);
            if (sz && kine && jin) stv<float, V>(QZ + 3 * cz, ri, ez3);
            if (sz && kine && rowx) stv<float, V>(QZ + 2 * cz, ri, ez2);
            if (sx && kine && rowx) stv<float, V>(QX + 3 * cx, rx, ex3);
            if (sx && kin && jin) stv<float, V>(QX + 2 * cx, rx, ex2);
            if (sy && kin && jin) stv<float, V>(QY + 3 * cy, ri, ey3);
        }
        hxm = hx;
        hym = hy;
        ex0 = ex1;
        ey0 = ey1;
        ez0 = ez1;
        eym0 = eym1;
        exl0 = exl1;
    }
}

// Interior launch with R rows per block sharing their neighbours through
// LDS (PM = 0 or 1: no y- or z-slab psi).  A block holds R consecutive rows
// of one x chunk and one z block, a wavefront per row, plus a halo wavefront
// for row j0-1 below the block.  Per plane:
//   1. every row wavefront publishes E(k) of its row (Ex, Ez), which the
//      row below needs for its Hx and Hz; barrier;
//   2. each row wavefront computes H_new of its row (row j+1's Ex, Ez from
//      LDS, from memory for the block's top row) and publishes Hx, Hz; the
//      halo wavefront computes Hx, Hz of row j0-1 with the same curl2 calls
//      as the row's own unit and publishes them; barrier;
//   3. each row wavefront computes E_new of its row (row j-1's new Hx, Hz
//      from LDS).
// Rows j-1 and j+1 are then neither read from memory nor recomputed inside a
// block; every value goes through the same arithmetic as in step3_unit, so
// the fields are bit-identical.  Waves of rows past the box (the last group)
// only take part in the barriers.
template <typename T, int PM, int V, int R>
__device__ __forceinline__ void step3_rows(const Grid<T> &g, const UnitMap &um, int blk, const float *cbs, float *xe,
                                           float *xh) {
    static_assert(PM == 0 || PM == 1, "interior launch only");
    constexpr bool PML = PM != 0;
    using F = vf<V>;
    constexpr int SL = 64 * V;                     // floats per slot array
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // 0..R-1 rows, R halo
    const int nx = g.nx, ny = g.ny, nz = g.nz, px = g.px;
    const int nch = (nx + 64 * V - 1) / (64 * V);
    const int ch = blk % nch;
    const int ng = (um.jhi - um.jlo + R - 1) / R;
    const int rest = blk / nch;
    const int j0 = um.jlo + (rest % ng) * R;
    const int k0 = (um.tlo + rest / ng) * um.kz;
    const int k1 = min(k0 + um.kz, nz);
    const int x0 = ch * 64 * V;
    const long long plane = g.plane, fs = g.fstride;
    const int tn = g.tn;
    const float *tab = g.tab;
    const bool halo = w == R;
    const int jr = halo ? j0 - 1 : j0 + w;         // this wavefront's row
    const bool act = halo ? jr >= 0 : jr < um.jhi;
    // row j+1 comes from LDS when its wavefront is active, else from memory
    const bool up_lds = !halo && w + 1 < R && jr + 1 < um.jhi;
    const int jp = min(jr + 1, ny - 1);
    const bool rowx = jr < ny - 1;                 // Hx, Hz, Ey rows
    const bool jin = jr >= 1 && jr <= ny - 2;      // Ex, Ez rows
    const float khy = tab[(1 * 6 + TK_KH) * tn + min(max(jr, 0), ny - 2)];
    const float key = tab[(1 * 6 + TK_KE) * tn + min(max(jr, 0), ny - 1)];
    const F vkhy = vs<V>(khy), vkey = vs<V>(key);
    const F zero = vs<V>(0.f);

    const int i0 = x0 + lane * V;
    const bool nxt = x0 + 64 * V < nx;
    const bool lft = x0 > 0;
    const unsigned m_hx = vmask<V>(i0, 0, nx - 1), m_in = vmask<V>(i0, 0, nx - 2);
    const unsigned m_ex = jin ? vmask<V>(i0, 0, nx - 2) : 0u;
    const unsigned m_ey = rowx ? vmask<V>(i0, 1, nx - 2) : 0u;
    const unsigned m_ez = jin ? vmask<V>(i0, 1, nx - 2) : 0u;
    const F khx = ldtab<V>(tab + (0 * 6 + TK_KH) * tn, i0), kex = ldtab<V>(tab + (0 * 6 + TK_KE) * tn, i0);
    const bool sxw = (PM & 1) && g.pml && (x0 < g.slo[0] || x0 + 64 * V > g.shi[0]);
    const bool sx = sxw && (i0 < g.slo[0] || i0 >= g.shi[0]);
    const int xs = sx ? (i0 < g.slo[0] ? i0 : g.slo[0] + (i0 - g.shi[0])) : 0;
    F bhx = zero, chx = zero, bex = zero, cex = zero;
    if (sx) {
        bhx = ldtab<V>(tab + (0 * 6 + TK_BH) * tn, i0);
        chx = ldtab<V>(tab + (0 * 6 + TK_CH) * tn, i0);
        bex = ldtab<V>(tab + (0 * 6 + TK_BE) * tn, i0);
        cex = ldtab<V>(tab + (0 * 6 + TK_CE) * tn, i0);
    }
    const int xl = lft ? x0 - 1 : 0;
    const float khxl = tab[(0 * 6 + TK_KH) * tn + xl];
    const bool sxl = (PM & 1) && lft && (xl < g.slo[0] || xl >= g.shi[0]);
    const int xsl = sxl ? (xl < g.slo[0] ? xl : g.slo[0] + (xl - g.shi[0])) : 0;
    const float bhxl = sxl ? tab[(0 * 6 + TK_BH) * tn + xl] : 0.f, chxl = sxl ? tab[(0 * 6 + TK_CH) * tn + xl] : 0.f;
    constexpr unsigned S = sizeof(T);
    const int jrr = max(jr, 0);
    const unsigned rj = (unsigned)(jrr * px + i0) * S, rjp = (unsigned)(jp * px + i0) * S;
    const unsigned ri = (unsigned)i0 * 4u, rx = (unsigned)xs * 4u;
    const unsigned re = (unsigned)(jrr * px + i0);
    const unsigned bn = rj + (unsigned)(64 * V - lane * V) * S;
    const unsigned bl = rj - (unsigned)(lane * V + 1) * S, blp = rjp - (unsigned)(lane * V + 1) * S;
    const unsigned ril = ri - (unsigned)(lane * V + 1) * 4u;
    const unsigned rxl = (unsigned)xsl * 4u;
    (void)ril;
    // LDS slots: xe[w] holds row j0+w's Ex, Ez of the plane; xh[w+1] holds row
    // j0+w's new Hx, Hz, xh[0] the halo row's
    auto put = [&](float *base, int slot, const F &a, const F &b) {
#pragma unroll
        for (int e = 0; e < V; ++e) {
            base[(slot * 2 + 0) * SL + e * 64 + lane] = a.v[e];
            base[(slot * 2 + 1) * SL + e * 64 + lane] = b.v[e];
        }
    };
    auto get = [&](const float *base, int slot, F &a, F &b) {
#pragma unroll
        for (int e = 0; e < V; ++e) {
            a.v[e] = base[(slot * 2 + 0) * SL + e * 64 + lane];
            b.v[e] = base[(slot * 2 + 1) * SL + e * 64 + lane];
        }
    };

    if (halo) {
        // row jr = j0-1: Hx and Hz of each plane, the values row j0's unit
        // recomputes in step3_unit (its hxj, hzj), from Ey of the row (z queue)
        F eym0 = zero;
        if (act) eym0 = ldv<T, V>(g.F + (long long)k0 * plane + fs, rj);
        for (int k = k0; k < k1; ++k) {
            const bool kin = k < nz - 1;
            const long long kp = uo(k * plane);
            const T *Fk = g.F + kp;
            const long long kn = kin ? plane : 0;
            F eym1 = zero, hxj = zero, hzj = zero, ezm = zero, exm = zero, qx0 = zero;
            float eymN = 0.f;
            if (act) {
                eym1 = ldv<T, V>(Fk + kn + fs, rj);
                hxj = ldv<T, V>(Fk + 3 * fs, rj);
                hzj = ldv<T, V>(Fk + 5 * fs, rj);
                ezm = ldv<T, V>(Fk + 2 * fs, rj);
                exm = ldv<T, V>(Fk, rj);
                if (nxt) eymN = ldb(Fk + fs, bn);
                if (PM == 1 ? sxw : false)
                    qx0 = ldv<float, V>(psi_base(g, 0, 0) + uo(((long long)k * ny + jrr) * g.sw[0]), rx);
            }
            const float khz = tab[(2 * 6 + TK_KH) * tn + min(k, nz - 2)];
            const F vkhz = vs<V>(khz);
            __syncthreads();                       // row j0's E(k) published
            F ex0, ez0;
            get(xe, 0, ex0, ez0);
            if (act) {
                const F eymi = xnext<V>(eym0, eymN, lane);
                F qy0 = zero, qy1 = zero, qz1 = zero;
                if (kin) {   // Hx(j0-1): d1 = Ez(j0) - Ez(j0-1), d2 = dEy(j0-1)/dz
                    const F acc = curl2<PML, V>(ez0 - ezm, vkhy, false, qy0, zero, zero, eym1 - eym0, vkhz, false, qz1,
                                                zero, zero);
                    hxj = vsel<V>(m_hx, hxj - acc, hxj);
                }
                {            // Hz(j0-1): d1 = dEy(j0-1)/dx, d2 = Ex(j0) - Ex(j0-1)
                    const F acc = curl2<PML, V>(eymi - eym0, khx, sx, qx0, bhx, chx, ex0 - exm, vkhy, false, qy1,
                                                zero, zero);
                    hzj = vsel<V>(m_in, hzj - acc, hzj);
                }
            }
            put(xh, 0, hxj, hzj);
            __syncthreads();                       // new H of every row published
            eym0 = eym1;
        }
        return;
    }

    F ex0 = zero, ey0 = zero, ez0 = zero;
    float exl0 = 0.f;
    F hxm = zero, hym = zero;
    if (act) {
        if (k0 > 0) {
            const int kh = k0 - 1;
            const T *Fk = g.F + kh * plane;
            const F exh = ldv<T, V>(Fk, rj), eyh = ldv<T, V>(Fk + fs, rj), ezh = ldv<T, V>(Fk + 2 * fs, rj);
            ex0 = ldv<T, V>(Fk + plane, rj);
            ey0 = ldv<T, V>(Fk + plane + fs, rj);
            ez0 = ldv<T, V>(Fk + plane + 2 * fs, rj);
            if (lft) exl0 = ldb(Fk + plane, bl);
            const F ezph = ldv<T, V>(Fk + 2 * fs, rjp);
            const float ezNh = nxt ? ldb(Fk + 2 * fs, bn) : 0.f;
            const F hxh = ldv<T, V>(Fk + 3 * fs, rj), hyh = ldv<T, V>(Fk + 4 * fs, rj);
            F py0 = zero, pz1 = zero, pz0 = zero, px1 = zero;
            if (sx) px1 = ldv<float, V>(psi_base(g, 0, 1) + ((long long)kh * ny + jr) * g.sw[0], rx);
            const float khz = tab[(2 * 6 + TK_KH) * tn + kh];
            const F vkhz = vs<V>(khz);
            const F ezi = xnext<V>(ezh, ezNh, lane);
            hxm = hxh;
            if (rowx) {
                const F acc = curl2<PML, V>(ezph - ezh, vkhy, false, py0, zero, zero, ey0 - eyh, vkhz, false, pz1, zero,
                                            zero);
                hxm = vsel<V>(m_hx, hxh - acc, hxh);
            }
            {
                const F acc = curl2<PML, V>(ex0 - exh, vkhz, false, pz0, zero, zero, ezi - ezh, khx, sx, px1, bhx, chx);
                hym = vsel<V>(m_in, hyh - acc, hyh);
            }
        } else {
            ex0 = ldv<T, V>(g.F, rj);
            ey0 = ldv<T, V>(g.F + fs, rj);
            ez0 = ldv<T, V>(g.F + 2 * fs, rj);
            if (lft) exl0 = ldb(g.F, bl);
        }
    }

    for (int k = k0; k < k1; ++k) {
        const bool kin = k < nz - 1;
        const bool kine = k >= 1 && kin;
        const long long kp = uo(k * plane);
        const T *Fk = g.F + kp;
        T *Fnk = g.Fn + kp;
        const long long kn = kin ? plane : 0;
        F ex1 = zero, ey1 = zero, ez1 = zero, ezp = zero, exq = zero, hx = zero, hy = zero, hz = zero;
        unsigned id[V];
#pragma unroll
        for (int e = 0; e < V; ++e) id[e] = 0;
        float ezN = 0.f, eyN = 0.f, exl1 = 0.f, hyl = 0.f, hzl = 0.f, ezl = 0.f, eyl = 0.f, exlp = 0.f;
        F px0 = zero, px1 = zero, ex2 = zero, ex3 = zero;
        float lx0 = 0.f, lx1 = 0.f;
        const float *PX = nullptr;
        float *QX = nullptr;
        const long long cx = g.scells[0];
        if (act) {
            ex1 = ldv<T, V>(Fk + kn, rj);
            ey1 = ldv<T, V>(Fk + kn + fs, rj);
            ez1 = ldv<T, V>(Fk + kn + 2 * fs, rj);
            if (!up_lds) {
                ezp = ldv<T, V>(Fk + 2 * fs, rjp);
                exq = ldv<T, V>(Fk, rjp);
            }
            hx = ldv<T, V>(Fk + 3 * fs, rj);
            hy = ldv<T, V>(Fk + 4 * fs, rj);
            hz = ldv<T, V>(Fk + 5 * fs, rj);
            if constexpr (V == 1) {
                id[0] = g.mat[kp + re];
            } else if constexpr (V == 2) {
                const uchar2 u = *reinterpret_cast<const uchar2 *>(g.mat + kp + re);
                id[0] = u.x;
                id[1] = u.y;
            } else {
                const unsigned u = *reinterpret_cast<const unsigned *>(g.mat + kp + re);
#pragma unroll
                for (int e = 0; e < V; ++e) id[e] = (u >> (8 * e)) & 0xffu;
            }
            if (nxt) {
                ezN = ldb(Fk + 2 * fs, bn);
                eyN = ldb(Fk + fs, bn);
            }
            if (lft) {
                exl1 = ldb(Fk + kn, bl);
                hyl = ldb(Fk + 4 * fs, bl);
                hzl = ldb(Fk + 5 * fs, bl);
                ezl = ldb(Fk + 2 * fs, bl);
                eyl = ldb(Fk + fs, bl);
                exlp = ldb(Fk, blp);
            }
            if (sxw || sxl) {
                const long long xb = uo(((long long)k * ny + jr) * g.sw[0]);
                PX = psi_base(g, 0, 0) + xb;
                QX = psi_new(g, 0, 0) + xb;
                if (sxl) {
                    lx0 = ldb(PX, rxl);
                    lx1 = ldb(PX + cx, rxl);
                }
            }
            if (PM == 1 && sxw) {
                px0 = ldv<float, V>(PX, rx);
                px1 = ldv<float, V>(PX + cx, rx);
                ex2 = ldv<float, V>(PX + 2 * cx, rx);
                ex3 = ldv<float, V>(PX + 3 * cx, rx);
            }
        }
        const float khz = tab[(2 * 6 + TK_KH) * tn + min(k, nz - 2)];
        const float kez = tab[(2 * 6 + TK_KE) * tn + k];
        const F vkhz = vs<V>(khz), vkez = vs<V>(kez);
        put(xe, w, ex0, ez0);
        __syncthreads();                           // E(k) of every row published
        if (up_lds) get(xe, w + 1, exq, ezp);
        const F ezi = xnext<V>(ez0, ezN, lane), eyi = xnext<V>(ey0, eyN, lane);
        F py0 = zero, py1 = zero, pz0 = zero, pz1 = zero;
        // ---------------- H half step, row j
        if (rowx && kin) {   // Hx: d1 = dEz/dy, d2 = dEy/dz
            const F acc = curl2<PML, V>(ezp - ez0, vkhy, false, py0, zero, zero, ey1 - ey0, vkhz, false, pz1, zero, zero);
            hx = vsel<V>(m_hx, hx - acc, hx);
        }
        if (kin) {           // Hy: d1 = dEx/dz, d2 = dEz/dx
            const F acc = curl2<PML, V>(ex1 - ex0, vkhz, false, pz0, zero, zero, ezi - ez0, khx, sx, px1, bhx, chx);
            hy = vsel<V>(m_in, hy - acc, hy);
        }
        if (rowx) {          // Hz: d1 = dEy/dx, d2 = dEx/dy
            const F acc = curl2<PML, V>(eyi - ey0, khx, sx, px0, bhx, chx, exq - ex0, vkhy, false, py1, zero, zero);
            hz = vsel<V>(m_in, hz - acc, hz);
        }
        // ---------------- Hy, Hz of cell x0-1 (only the chunk's first Ey, Ez use them)
        if (lft) {
            float lz0 = 0.f, ly1 = 0.f;
            if (kin) {
                const float acc = curl2<PML>(exl1 - exl0, khz, false, lz0, 0.f, 0.f, lane0(ez0.v[0]) - ezl, khxl, sxl,
                                             lx1, bhxl, chxl);
                hyl = hyl - acc;
            }
            if (rowx) {
                const float acc = curl2<PML>(lane0(ey0.v[0]) - eyl, khxl, sxl, lx0, bhxl, chxl, exlp - exl0, khy,
                                             false, ly1, 0.f, 0.f);
                hzl = hzl - acc;
            }
        }
        if (act) {
            stv<T, V>(Fnk + 3 * fs, rj, hx);
            stv<T, V>(Fnk + 4 * fs, rj, hy);
            stv<T, V>(Fnk + 5 * fs, rj, hz);
            if (PML) {
                if (sx && kin) stv<float, V>(QX + cx, rx, px1);
                if (sx && rowx) stv<float, V>(QX, rx, px0);
            }
        }
        put(xh, w + 1, hx, hz);
        __syncthreads();                           // new H of every row published
        F hxj, hzj;
        get(xh, w, hxj, hzj);
        // ---------------- E half step, row j
        if (act) {
            const F hyp = xprev<V>(hy, hyl, lane), hzp = xprev<V>(hz, hzl, lane);
            F ex = ex0, ey = ey0, ez = ez0;
            F ey2 = zero, ez3 = zero, ez2 = zero, ey3 = zero;
            if (kine) {
                {            // Ex: d1 = dHz/dy, d2 = dHy/dz
                    const F acc = curl2<PML, V>(hz - hzj, vkey, false, ey2, zero, zero, hy - hym, vkez, false, ez3, zero,
                                                zero);
                    ex = vsel<V>(m_ex, eupd<V>(ex0, cbv<V>(cbs, id), acc), ex0);
                }
                {            // Ey: d1 = dHx/dz, d2 = dHz/dx
                    const F acc = curl2<PML, V>(hx - hxm, vkez, false, ez2, zero, zero, hz - hzp, kex, sx, ex3, bex,
                                                cex);
                    ey = vsel<V>(m_ey, eupd<V>(ey0, cbv<V>(cbs + 256, id), acc), ey0);
                }
            }
            if (kin) {       // Ez: d1 = dHy/dx, d2 = dHx/dy
                const F acc = curl2<PML, V>(hy - hyp, kex, sx, ex2, bex, cex, hx - hxj, vkey, false, ey3, zero, zero);
                ez = vsel<V>(m_ez, eupd<V>(ez0, cbv<V>(cbs + 512, id), acc), ez0);
            }
            stv<T, V>(Fnk, rj, ex);
            stv<T, V>(Fnk + fs, rj, ey);
            stv<T, V>(Fnk + 2 * fs, rj, ez);
            if (PML) {
                if (sx && kine && rowx) stv<float, V>(QX + 3 * cx, rx, ex3);
                if (sx && kin && jin) stv<float, V>(QX + 2 * cx, rx, ex2);
            }
        }
        hxm = hx;
        hym = hy;
        ex0 = ex1;
        ey0 = ey1;
        ez0 = ez1;
        exl0 = exl1;
    }
}

template <typename T, int PM, int R, int V>
__global__ __launch_bounds__(64 * (R + 1)) void k_step3r(Grid<T> g, UnitMap um, int nblocks) {
    __shared__ float cbs[3 * 256];
    __shared__ float xe[R * 2 * 64 * V];
    __shared__ float xh[(R + 1) * 2 * 64 * V];
    for (int q = threadIdx.x; q < 3 * 256; q += 64 * (R + 1)) cbs[q] = g.cb[q];
    __syncthreads();
    step3_rows<T, PM, V, R>(g, um, xcd_remap(blockIdx.x, nblocks), cbs, xe, xh);
}

template <typename T, int PM, int NW, int V>
__global__ __launch_bounds__(64 * NW) void k_step3(Grid<T> g, UnitMap um, int nunits) {
    __shared__ float cbs[3 * 256];
    for (int q = threadIdx.x; q < 3 * 256; q += 64 * NW) cbs[q] = g.cb[q];
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int unit = xcd_remap(blockIdx.x, (nunits + NW - 1) / NW) * NW + w;
    if (unit >= nunits) return;                    // whole wavefront; no barrier follows
    step3_unit<T, PM, V>(g, um, unit, cbs);
}

// Batched step: block b of the launch runs NW units of member bmap[b].x, its
// units bmap[b].y*NW on (a member's blocks never hold another's units, so
// the block's LDS table is that member's).  Part 0: interior boxes (and whole
// grids without CPML), PM = 1; part 1: CPML shells, PM = 7.  par selects the
// direction: 0 reads buffer set 0 and writes set 1, 1 the reverse.
template <typename T, int PM, int NW, int V>
__global__ __launch_bounds__(64 * NW) void k_step3_batch(const B3Sim<T> *__restrict__ sims,
                                                         const int2 *__restrict__ bmap, int nblocks, int part,
                                                         int par) {
    __shared__ float cbs[3 * 256];
    const int2 mb = bmap[xcd_remap(blockIdx.x, nblocks)];
    const B3Sim<T> *d = sims + mb.x;
    Grid<T> g = d->g;
    if (par) {
        T *f = g.F;
        g.F = g.Fn;
        g.Fn = f;
        float *q = g.psi;
        g.psi = g.psin;
        g.psin = q;
    }
    for (int q = threadIdx.x; q < 3 * 256; q += 64 * NW) cbs[q] = g.cb[q];
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int unit = mb.y * NW + w;
    if (unit >= d->nunits[part]) return;
    step3_unit<T, PM, V>(g, d->um[part], unit, cbs);
}

template <typename T, int PM, int V>
static void launch_units(const Grid<T> &g, int nw, const UnitMap &um, int nunits, hipStream_t s) {
    if (nunits <= 0) return;
    if (nw == 8)
        hipLaunchKernelGGL((k_step3<T, PM, 8, V>), dim3((nunits + 7) / 8), dim3(512), 0, s, g, um, nunits);
    else
        hipLaunchKernelGGL((k_step3<T, PM, 4, V>), dim3((nunits + 3) / 4), dim3(256), 0, s, g, um, nunits);
}

// Rows [jlo, jhi) and z blocks [tlo, thi) of the interior box: rows j with j
// and j-1 outside the y slabs, z blocks whose planes (k0-1 included) lie
// outside the z slabs.
static UnitMap core_box(int ny, int nz, int kz, const int slo[3], const int shi[3]) {
    const int ntz = (nz + kz - 1) / kz;
    UnitMap um{};
    um.mode = 0;
    um.ntz = ntz;
    um.kz = kz;
    um.jlo = slo[1] > 0 ? slo[1] + 1 : 0;
    um.jhi = std::max(um.jlo, shi[1]);
    um.tlo = std::min(slo[2] > 0 ? (slo[2] + 1 + kz - 1) / kz : 0, ntz);
    um.thi = std::max(shi[2] >= nz ? ntz : shi[2] / kz, um.tlo);
    return um;
}

void step3_core_extent(int ny, int nz, int kz, const int slo[3], const int shi[3], int *rows, int *planes) {
    const UnitMap um = core_box(ny, nz, kz, slo, shi);
    *rows = um.jhi - um.jlo;
    *planes = std::min(um.thi * kz, nz) - um.tlo * kz;
}

void step3_unit_maps(int nx, int ny, int nz, int kz, int kzs, int pml, const int slo[3], const int shi[3], int vec,
                     UnitMap um[2], int nunits[2]) {
    const int ntz = (nz + kz - 1) / kz;
    const int nch = (nx + 64 * vec - 1) / (64 * vec);   // one wavefront per (x chunk, row, z block)
    if (!pml) {
        um[0] = UnitMap{};
        um[0].jhi = ny;
        um[0].thi = um[0].ntz = ntz;
        um[0].kz = kz;
        um[1] = um[0];
        nunits[0] = nch * ny * ntz;
        nunits[1] = 0;
        return;
    }
    um[0] = core_box(ny, nz, kz, slo, shi);
    nunits[0] = nch * (um[0].jhi - um[0].jlo) * (um[0].thi - um[0].tlo);
    // the shell in blocks of its own depth: rows outside the box over the
    // whole z range, rows inside below and above the box's planes
    if (kzs <= 0) kzs = kz;
    UnitMap &sh = um[1];
    sh = um[0];
    sh.mode = 1;
    sh.kz = kzs;
    sh.ntz = (nz + kzs - 1) / kzs;
    sh.zlo = std::min(um[0].tlo * kz, nz);
    sh.zhi = std::min(um[0].thi * kz, nz);
    sh.nlo = (sh.zlo + kzs - 1) / kzs;
    const int nhi = (nz - sh.zhi + kzs - 1) / kzs;
    const int nro = um[0].jlo + (ny - um[0].jhi), nr = um[0].jhi - um[0].jlo;
    nunits[1] = nch * (nro * sh.ntz + nr * (sh.nlo + nhi));
}

// Interior launch with rows sharing neighbours: blocks of R rows (+ the
// halo wavefront) of one x chunk and one z block.
template <typename T, int PM, int V>
static void launch_rows(const Grid<T> &g, const UnitMap &um, hipStream_t s) {
    const int nch = (g.nx + 64 * V - 1) / (64 * V);
    const int R = g.rows;
    const int ng = (um.jhi - um.jlo + R - 1) / R;
    const int nb = nch * ng * (um.thi - um.tlo);
    if (nb <= 0) return;
    if (R == 7) hipLaunchKernelGGL((k_step3r<T, PM, 7, V>), dim3(nb), dim3(512), 0, s, g, um, nb);
    else hipLaunchKernelGGL((k_step3r<T, PM, 3, V>), dim3(nb), dim3(256), 0, s, g, um, nb);
}

template <typename T, int V>
static void launch_parts(const Grid<T> &g, int nw, int part, hipStream_t s) {
    UnitMap um[2];
    int nu[2];
    step3_unit_maps(g.nx, g.ny, g.nz, g.kz, g.kzs, g.pml, g.slo, g.shi, V, um, nu);
    if (part == 0 && g.rows > 0 && nu[0] > 0) {
        if (g.pml) launch_rows<T, 1, V>(g, um[0], s);
        else launch_rows<T, 0, V>(g, um[0], s);
        return;
    }
    if (!g.pml) {
        if (part == 0) launch_units<T, 0, V>(g, nw, um[0], nu[0], s);
    } else if (part == 0) {
        launch_units<T, 1, V>(g, nw, um[0], nu[0], s);
    } else {
        launch_units<T, 7, V>(g, nw, um[1], nu[1], s);
    }
}

template <typename T>
void launch_step3(const Grid<T> &g, int nw, int vec, int part, hipStream_t s) {
    if constexpr (std::is_same<T, __half>::value) {
        if (vec == 4) {
            launch_parts<T, 4>(g, nw, part, s);
            return;
        }
    }
    if (vec == 2) launch_parts<T, 2>(g, nw, part, s);
    else launch_parts<T, 1>(g, nw, part, s);
}

template void launch_step3<float>(const Grid<float> &, int, int, int, hipStream_t);
template void launch_step3<__half>(const Grid<__half> &, int, int, int, hipStream_t);

template <typename T>
void launch_step3_batch(const B3Sim<T> *sims, const int2 *bmap, int nblocks, int part, int par, int vec,
                        hipStream_t s) {
    if (nblocks <= 0) return;
    constexpr int NW = kB3Waves;
    if constexpr (std::is_same<T, __half>::value) {
        if (vec == 4) {
            if (part == 0)
                hipLaunchKernelGGL((k_step3_batch<T, 1, NW, 4>), dim3(nblocks), dim3(64 * NW), 0, s, sims, bmap, nblocks, 0, par);
            else
                hipLaunchKernelGGL((k_step3_batch<T, 7, NW, 4>), dim3(nblocks), dim3(64 * NW), 0, s, sims, bmap, nblocks, 1, par);
            return;
        }
    }
    if (part == 0) {
        if (vec == 2)
            hipLaunchKernelGGL((k_step3_batch<T, 1, NW, 2>), dim3(nblocks), dim3(64 * NW), 0, s, sims, bmap, nblocks, 0, par);
        else
            hipLaunchKernelGGL((k_step3_batch<T, 1, NW, 1>), dim3(nblocks), dim3(64 * NW), 0, s, sims, bmap, nblocks, 0, par);
    } else {
        if (vec == 2)
            hipLaunchKernelGGL((k_step3_batch<T, 7, NW, 2>), dim3(nblocks), dim3(64 * NW), 0, s, sims, bmap, nblocks, 1, par);
        else
            hipLaunchKernelGGL((k_step3_batch<T, 7, NW, 1>), dim3(nblocks), dim3(64 * NW), 0, s, sims, bmap, nblocks, 1, par);
    }
}

template void launch_step3_batch<float>(const B3Sim<float> *, const int2 *, int, int, int, int, hipStream_t);
template void launch_step3_batch<__half>(const B3Sim<__half> *, const int2 *, int, int, int, int, hipStream_t);

}  // namespace lx
