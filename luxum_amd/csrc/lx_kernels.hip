// lx_kernels.hip -- gfx950 kernels of the Yee timestep.
//
// Bulk half-step kernels (one launch per half step).  A block of 256 threads
// (4 wavefronts) owns a 256 (x) x 4 (y) tile: each lane holds 4 consecutive x
// cells (one 16-byte load per component), each wavefront one y row.  The
// block marches along z over kz planes; every lane keeps the plane above (H
// update) or below (E update) in registers, so the z neighbour is read once.
// x neighbours come from the adjacent lane by a cross-lane shuffle, y
// neighbours from the adjacent row through the cache.  A lane updates the
// three components of its cells: split-curl differences
// (lib/field/vfield.cc:196-275), the CPML psi recursion on slab cells
// (maxwell_three_dim_pml.cc:182-255) and the field update, with no curl
// temporaries in memory.
//
// TF/SF (src/solvers/maxwell.cc:512-812) is applied by a second, small launch
// over the face cells only: the corrections are linear in the curl, so adding
// dir*g to a curl difference after the bulk update equals adding it before.
// One extra block of that launch advances the FP64 aux-1D incident line and
// records probes.
//
// Block ids are remapped so that each of the 8 XCDs (which receive block ids
// round robin) works on a contiguous range of tiles: tiles adjacent in y and z
// share their halo rows and planes in that XCD's L2.
#include "lx_batch.hpp"
#include "lx_device.hpp"

namespace lx {

// ---------------------------------------------------------------------------
// 3D H update:  Hs -= (d1*kh_a1 - d2*kh_a2) + (psi1 - psi2)
// component c: first split curl along (c+1)%3, second along (c+2)%3.
// ---------------------------------------------------------------------------
template <typename T, bool PML>
__global__ __launch_bounds__(256) void k_h3(Grid<T> g) {
    const int tile = xcd_remap(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, rr = tile / g.ntx, ty = rr % g.nty, tz = rr / g.nty;
    const int lane = threadIdx.x & 63;
    const int i0 = tx * 256 + lane * 4;
    const int j = ty * kBY + (threadIdx.x >> 6);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (j >= ny) return;                       // whole wavefront
    const bool on = i0 < nx;
    const int k0 = tz * g.kz;
    const int k1 = min(k0 + g.kz, nz);
    const long long px = g.px, plane = g.plane, fs = g.fstride;
    const T *Ex = g.F, *Ey = g.F + fs, *Ez = g.F + 2 * fs;
    T *Hx = g.F + 3 * fs, *Hy = g.F + 4 * fs, *Hz = g.F + 5 * fs;
    const int tn = g.tn;
    const float *tab = g.tab;
    const int jp = min(j + 1, ny - 1);
    const bool last_lane = (lane == 63) || (i0 + 4 >= nx);
    const bool nxt = i0 + 4 < nx;
    // element masks
    const unsigned mx_all = on ? xmask(i0, 0, nx - 1) : 0u;      // Hx: i < nx
    const unsigned mx_in = on ? xmask(i0, 0, nx - 2) : 0u;       // Hy, Hz: i < nx-1
    const bool rowx = j < ny - 1;
    const float4 khx = on ? *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_KH) * tn + i0) : f4(0.f);
    const float khy = tab[(1 * 6 + TK_KH) * tn + min(j, ny - 2)];
    // CPML (slabs)
    const bool sy = PML && (j < g.slo[1] || j >= g.shi[1]);
    const float bhy = sy ? tab[(1 * 6 + TK_BH) * tn + j] : 0.f, chy = sy ? tab[(1 * 6 + TK_CH) * tn + j] : 0.f;
    // x slab: 4-aligned, so a lane's 4 cells are wholly in or out of it
    const bool sx = PML && on && (i0 < g.slo[0] || i0 >= g.shi[0]);
    const int xs = sx ? (i0 < g.slo[0] ? i0 : g.slo[0] + (i0 - g.shi[0])) : 0;
    float *psy0 = PML ? psi_base(g, 1, 0) : nullptr, *psy1 = PML ? psi_base(g, 1, 1) : nullptr;
    float *psz0 = PML ? psi_base(g, 2, 0) : nullptr, *psz1 = PML ? psi_base(g, 2, 1) : nullptr;
    float *psx0 = PML ? psi_base(g, 0, 0) : nullptr, *psx1 = PML ? psi_base(g, 0, 1) : nullptr;
    const int siy = PML && sy ? slab_index(j, g.slo[1], g.shi[1]) : 0;

    long long o = (long long)k0 * plane + (long long)j * px + i0;
    const long long dj = (long long)(jp - j) * px;
    float4 ex0 = f4(0.f), ey0 = f4(0.f), ez0 = f4(0.f);
    if (on) { ex0 = ld4(Ex + o); ey0 = ld4(Ey + o); ez0 = ld4(Ez + o); }
    for (int k = k0; k < k1; ++k, o += plane) {
        const bool kin = k < nz - 1;
        float4 ex1 = f4(0.f), ey1 = f4(0.f), ez1 = f4(0.f);
        if (on && kin) { ex1 = ld4(Ex + o + plane); ey1 = ld4(Ey + o + plane); ez1 = ld4(Ez + o + plane); }
        // x+1 neighbours of element 3 from the next lane (or memory at the tile edge)
        float ezn = __shfl_down(ez0.x, 1), eyn = __shfl_down(ey0.x, 1);
        if (last_lane) {
            ezn = nxt ? ld1(Ez + o + 4) : 0.f;
            eyn = nxt ? ld1(Ey + o + 4) : 0.f;
        }
        const int kk = min(k, nz - 2);
        const float khz = tab[(2 * 6 + TK_KH) * tn + kk];
        const bool sz = PML && (k < g.slo[2] || k >= g.shi[2]);
        const int zoff = sz ? (slab_index(k, g.slo[2], g.shi[2]) * ny + j) * (int)px + i0 : 0;
        const int yoff = sy ? (k * g.sw[1] + siy) * (int)px + i0 : 0;
        const int xoff = sx ? (k * ny + j) * g.sw[0] + xs : 0;
        const float bhz = sz ? tab[(2 * 6 + TK_BH) * tn + k] : 0.f, chz = sz ? tab[(2 * 6 + TK_CH) * tn + k] : 0.f;
        if (on) {
            // Hx: d1 = dEz/dy (y, psi y slot0), d2 = dEy/dz (z, psi z slot1)
            if (rowx && kin) {
                const float4 ezj = ld4(Ez + o + dj);
                const float4 d1 = ezj - ez0, d2 = ey1 - ey0;
                float4 acc = d1 * khy - d2 * khz;
                if (sy) {
                    float4 p = ld4(psy0 + yoff);
                    p = p * bhy + d1 * chy;
                    st4(psy0 + yoff, p);
                    acc = acc + p;
                }
                if (sz) {
                    float4 p = ld4(psz1 + zoff);
                    p = p * bhz + d2 * chz;
                    st4(psz1 + zoff, p);
                    acc = acc - p;
                }
                const float4 h = ld4(Hx + o);
                st4(Hx + o, sel(mx_all, h - acc, h));
            }
            // Hy: d1 = dEx/dz (z, psi z slot0), d2 = dEz/dx (x, psi x slot1)
            if (kin) {
                const float4 ezi = make_float4(ez0.y, ez0.z, ez0.w, ezn);
                const float4 d1 = ex1 - ex0, d2 = ezi - ez0;
                float4 acc = d1 * khz - d2 * khx;
                if (sz) {
                    float4 p = ld4(psz0 + zoff);
                    p = p * bhz + d1 * chz;
                    st4(psz0 + zoff, p);
                    acc = acc + p;
                }
                if (sx) {
                    float4 p = ld4(psx1 + xoff);
                    p = p * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BH) * tn + i0) +
                        d2 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CH) * tn + i0);
                    st4(psx1 + xoff, p);
                    acc = acc - p;
                }
                const float4 h = ld4(Hy + o);
                st4(Hy + o, sel(mx_in, h - acc, h));
            }
            // Hz: d1 = dEy/dx (x, psi x slot0), d2 = dEx/dy (y, psi y slot1)
            if (rowx) {
                const float4 eyi = make_float4(ey0.y, ey0.z, ey0.w, eyn);
                const float4 exj = ld4(Ex + o + dj);
                const float4 d1 = eyi - ey0, d2 = exj - ex0;
                float4 acc = d1 * khx - d2 * khy;
                if (sx) {
                    float4 p = ld4(psx0 + xoff);
                    p = p * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BH) * tn + i0) +
                        d1 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CH) * tn + i0);
                    st4(psx0 + xoff, p);
                    acc = acc + p;
                }
                if (sy) {
                    float4 p = ld4(psy1 + yoff);
                    p = p * bhy + d2 * chy;
                    st4(psy1 + yoff, p);
                    acc = acc - p;
                }
                const float4 h = ld4(Hz + o);
                st4(Hz + o, sel(mx_in, h - acc, h));
            }
        }
        ex0 = ex1; ey0 = ey1; ez0 = ez1;
    }
}

// ---------------------------------------------------------------------------
// 3D E update:  E_c += cb[c][id] * ((d1*ke_a1 - d2*ke_a2) + (psi1 - psi2)),
// interior only (PEC walls come from the curl ranges, vfield.cc:242-268).
// Each E component has its own coefficient table (the reference's per-
// component Cb arrays, maxwell_three_dim_pml.cc:109-128).
// ---------------------------------------------------------------------------
template <typename T, bool PML>
__global__ __launch_bounds__(256) void k_e3(Grid<T> g) {
    __shared__ float cbs[3 * 256];
    for (int q = threadIdx.x; q < 3 * 256; q += 256) cbs[q] = g.cb[q];
    __syncthreads();
    const int tile = xcd_remap(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, rr = tile / g.ntx, ty = rr % g.nty, tz = rr / g.nty;
    const int lane = threadIdx.x & 63;
    const int i0 = tx * 256 + lane * 4;
    const int j = ty * kBY + (threadIdx.x >> 6);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (j >= ny) return;
    const bool on = i0 < nx;
    const int k0 = tz * g.kz;
    const int k1 = min(k0 + g.kz, nz);
    const long long px = g.px, plane = g.plane, fs = g.fstride;
    T *Ex = g.F, *Ey = g.F + fs, *Ez = g.F + 2 * fs;
    const T *Hx = g.F + 3 * fs, *Hy = g.F + 4 * fs, *Hz = g.F + 5 * fs;
    const int tn = g.tn;
    const float *tab = g.tab;
    const int jm = max(j - 1, 0);
    const bool first_lane = (lane == 0);
    const bool prv = i0 >= 1;
    const bool jin = (j >= 1) && (j <= ny - 2);
    const unsigned m_ex = on && jin ? xmask(i0, 0, nx - 2) : 0u;            // Ex: i < nx-1, 1 <= j <= ny-2
    const unsigned m_ey = on && (j < ny - 1) ? xmask(i0, 1, nx - 2) : 0u;   // Ey: 1 <= i <= nx-2, j < ny-1
    const unsigned m_ez = on && jin ? xmask(i0, 1, nx - 2) : 0u;            // Ez: 1 <= i <= nx-2, 1 <= j <= ny-2
    if (!__any(m_ex | m_ey | m_ez)) return;
    const float4 kex = on ? *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_KE) * tn + i0) : f4(0.f);
    const float key = tab[(1 * 6 + TK_KE) * tn + j];
    const bool sy = PML && (j < g.slo[1] || j >= g.shi[1]);
    const float bey = sy ? tab[(1 * 6 + TK_BE) * tn + j] : 0.f, cey = sy ? tab[(1 * 6 + TK_CE) * tn + j] : 0.f;
    const bool sx = PML && on && (i0 < g.slo[0] || i0 >= g.shi[0]);
    const int xs = sx ? (i0 < g.slo[0] ? i0 : g.slo[0] + (i0 - g.shi[0])) : 0;
    float *psy0 = PML ? psi_base(g, 1, 2) : nullptr, *psy1 = PML ? psi_base(g, 1, 3) : nullptr;
    float *psz0 = PML ? psi_base(g, 2, 2) : nullptr, *psz1 = PML ? psi_base(g, 2, 3) : nullptr;
    float *psx0 = PML ? psi_base(g, 0, 2) : nullptr, *psx1 = PML ? psi_base(g, 0, 3) : nullptr;
    const int siy = PML && sy ? slab_index(j, g.slo[1], g.shi[1]) : 0;

    long long o = (long long)k0 * plane + (long long)j * px + i0;
    const long long dj = (long long)(j - jm) * px;
    float4 hxm = f4(0.f), hym = f4(0.f);
    if (on && k0 > 0) { hxm = ld4(Hx + o - plane); hym = ld4(Hy + o - plane); }
    for (int k = k0; k < k1; ++k, o += plane) {
        float4 hx0 = f4(0.f), hy0 = f4(0.f), hz0 = f4(0.f);
        if (on) { hx0 = ld4(Hx + o); hy0 = ld4(Hy + o); hz0 = ld4(Hz + o); }
        // x-1 neighbours of element 0 from the previous lane (or memory at the tile edge)
        float hzp = __shfl_up(hz0.w, 1), hyp = __shfl_up(hy0.w, 1);
        if (first_lane) {
            hzp = (on && prv) ? ld1(Hz + o - 1) : 0.f;
            hyp = (on && prv) ? ld1(Hy + o - 1) : 0.f;
        }
        const bool kin = (k >= 1) && (k <= nz - 2);
        const float kez = tab[(2 * 6 + TK_KE) * tn + k];
        const bool sz = PML && (k < g.slo[2] || k >= g.shi[2]);
        const int zoff = sz ? (slab_index(k, g.slo[2], g.shi[2]) * ny + j) * (int)px + i0 : 0;
        const int yoff = sy ? (k * g.sw[1] + siy) * (int)px + i0 : 0;
        const int xoff = sx ? (k * ny + j) * g.sw[0] + xs : 0;
        const float bez = sz ? tab[(2 * 6 + TK_BE) * tn + k] : 0.f, cez = sz ? tab[(2 * 6 + TK_CE) * tn + k] : 0.f;
        if (on) {
            const uchar4 id = *reinterpret_cast<const uchar4 *>(g.mat + o);
            // Ex: d1 = dHz/dy (y, psi y slot0), d2 = dHy/dz (z, psi z slot1)
            if (kin && m_ex) {
                const float4 hzj = ld4(Hz + o - dj);
                const float4 d1 = hz0 - hzj, d2 = hy0 - hym;
                float4 acc = d1 * key - d2 * kez;
                if (sy) {
                    float4 p = ld4(psy0 + yoff);
                    p = p * bey + d1 * cey;
                    st4(psy0 + yoff, p);
                    acc = acc + p;
                }
                if (sz) {
                    float4 p = ld4(psz1 + zoff);
                    p = p * bez + d2 * cez;
                    st4(psz1 + zoff, p);
                    acc = acc - p;
                }
                const float4 cb = make_float4(cbs[id.x], cbs[id.y], cbs[id.z], cbs[id.w]);
                const float4 e = ld4(Ex + o);
                st4(Ex + o, sel(m_ex, e + cb * acc, e));
            }
            // Ey: d1 = dHx/dz (z, psi z slot0), d2 = dHz/dx (x, psi x slot1)
            if (kin && m_ey) {
                const float4 hzi = make_float4(hzp, hz0.x, hz0.y, hz0.z);
                const float4 d1 = hx0 - hxm, d2 = hz0 - hzi;
                float4 acc = d1 * kez - d2 * kex;
                if (sz) {
                    float4 p = ld4(psz0 + zoff);
                    p = p * bez + d1 * cez;
                    st4(psz0 + zoff, p);
                    acc = acc + p;
                }
                if (sx) {
                    float4 p = ld4(psx1 + xoff);
                    p = p * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BE) * tn + i0) +
                        d2 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CE) * tn + i0);
                    st4(psx1 + xoff, p);
                    acc = acc - p;
                }
                const float4 cb = make_float4(cbs[256 + id.x], cbs[256 + id.y], cbs[256 + id.z], cbs[256 + id.w]);
                const float4 e = ld4(Ey + o);
                st4(Ey + o, sel(m_ey, e + cb * acc, e));
            }
            // Ez: d1 = dHy/dx (x, psi x slot0), d2 = dHx/dy (y, psi y slot1)
            if (k < nz - 1 && m_ez) {
                const float4 hyi = make_float4(hyp, hy0.x, hy0.y, hy0.z);
                const float4 hxj = ld4(Hx + o - dj);
                const float4 d1 = hy0 - hyi, d2 = hx0 - hxj;
                float4 acc = d1 * kex - d2 * key;
                if (sx) {
                    float4 p = ld4(psx0 + xoff);
                    p = p * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BE) * tn + i0) +
                        d1 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CE) * tn + i0);
                    st4(psx0 + xoff, p);
                    acc = acc + p;
                }
                if (sy) {
                    float4 p = ld4(psy1 + yoff);
                    p = p * bey + d2 * cey;
                    st4(psy1 + yoff, p);
                    acc = acc - p;
                }
                const float4 cb = make_float4(cbs[512 + id.x], cbs[512 + id.y], cbs[512 + id.z], cbs[512 + id.w]);
                const float4 e = ld4(Ez + o);
                st4(Ez + o, sel(m_ez, e + cb * acc, e));
            }
        }
        hxm = hx0; hym = hy0;
    }
}

// ---------------------------------------------------------------------------
// 2D (x-z plane): Ey, Hx, Hz (maxwell_two_dim_pml.cc:144-176).  Tiles of
// 256 (x) x 4 (z) cells, 4 cells per lane, one z row per wavefront.
// ---------------------------------------------------------------------------
template <typename T, bool PML>
__global__ __launch_bounds__(256) void k_h2(Grid<T> g) {
    const int tile = xcd_remap(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, tz = tile / g.ntx;
    const int lane = threadIdx.x & 63;
    const int i0 = tx * 256 + lane * 4;
    const int k = tz * kBY + (threadIdx.x >> 6);
    const int nx = g.nx, nz = g.nz;
    if (k >= nz) return;
    const bool on = i0 < nx;
    const long long px = g.px, fs = g.fstride;
    const T *Ey = g.F + fs;
    T *Hx = g.F + 3 * fs, *Hz = g.F + 5 * fs;
    const int tn = g.tn;
    const float *tab = g.tab;
    const long long o = (long long)k * px + i0;
    float4 e0 = f4(0.f);
    if (on) e0 = ld4(Ey + o);
    float en = __shfl_down(e0.x, 1);
    if (lane == 63 || i0 + 4 >= nx) en = (i0 + 4 < nx) ? ld1(Ey + o + 4) : 0.f;
    if (!on) return;
    // Hx: second split curl d2 = dEy/dz (z, psi z slot1):  Hs_x -= -d2*kh_z - psi
    if (k < nz - 1) {
        const float4 d2 = ld4(Ey + o + px) - e0;
        float4 acc = d2 * (-tab[(2 * 6 + TK_KH) * tn + k]);
        if (PML && (k < g.slo[2] || k >= g.shi[2])) {
            float *q = psi_base(g, 2, 1) + psi_off(g, 2, i0, 0, k);
            float4 p = ld4(q);
            p = p * tab[(2 * 6 + TK_BH) * tn + k] + d2 * tab[(2 * 6 + TK_CH) * tn + k];
            st4(q, p);
            acc = acc - p;
        }
        const float4 h = ld4(Hx + o);
        st4(Hx + o, sel(xmask(i0, 0, nx - 1), h - acc, h));
    }
    // Hz: first split curl d1 = dEy/dx (x, psi x slot0)
    {
        const float4 ei = make_float4(e0.y, e0.z, e0.w, en);
        const float4 d1 = ei - e0;
        float4 acc = d1 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_KH) * tn + i0);
        const unsigned mi = xmask(i0, 0, nx - 2);
        if (PML && (i0 < g.slo[0] || i0 >= g.shi[0])) {
            float *q = psi_base(g, 0, 0) + k * g.sw[0] + (i0 < g.slo[0] ? i0 : g.slo[0] + (i0 - g.shi[0]));
            float4 p = ld4(q);
            p = p * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BH) * tn + i0) +
                d1 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CH) * tn + i0);
            st4(q, p);
            acc = acc + p;
        }
        const float4 h = ld4(Hz + o);
        st4(Hz + o, sel(mi, h - acc, h));
    }
}

template <typename T, bool PML>
__global__ __launch_bounds__(256) void k_e2(Grid<T> g) {
    __shared__ float cbs[256];
    cbs[threadIdx.x] = g.cb[256 + threadIdx.x];   // the Ey table
    __syncthreads();
    const int tile = xcd_remap(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, tz = tile / g.ntx;
    const int lane = threadIdx.x & 63;
    const int i0 = tx * 256 + lane * 4;
    const int k = tz * kBY + (threadIdx.x >> 6);
    const int nx = g.nx, nz = g.nz;
    if (k < 1 || k > nz - 2) return;
    const bool on = i0 < nx;
    const long long px = g.px, fs = g.fstride;
    const T *Hx = g.F + 3 * fs, *Hz = g.F + 5 * fs;
    T *Ey = g.F + fs;
    const int tn = g.tn;
    const float *tab = g.tab;
    const long long o = (long long)k * px + i0;
    float4 hz0 = f4(0.f);
    if (on) hz0 = ld4(Hz + o);
    float hp = __shfl_up(hz0.w, 1);
    if (lane == 0) hp = (on && i0 >= 1) ? ld1(Hz + o - 1) : 0.f;
    if (!on) return;
    const unsigned m = xmask(i0, 1, nx - 2);
    if (!m) return;
    // Ey: d1 = dHx/dz (z, psi z slot0), d2 = dHz/dx (x, psi x slot1)
    const float4 d1 = ld4(Hx + o) - ld4(Hx + o - px);
    const float4 d2 = hz0 - make_float4(hp, hz0.x, hz0.y, hz0.z);
    float4 acc = d1 * tab[(2 * 6 + TK_KE) * tn + k] - d2 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_KE) * tn + i0);
    if (PML) {
        if (k < g.slo[2] || k >= g.shi[2]) {
            float *q = psi_base(g, 2, 2) + psi_off(g, 2, i0, 0, k);
            float4 p = ld4(q);
            p = p * tab[(2 * 6 + TK_BE) * tn + k] + d1 * tab[(2 * 6 + TK_CE) * tn + k];
            st4(q, p);
            acc = acc + p;
        }
        if (i0 < g.slo[0] || i0 >= g.shi[0]) {
            float *q = psi_base(g, 0, 3) + k * g.sw[0] + (i0 < g.slo[0] ? i0 : g.slo[0] + (i0 - g.shi[0]));
            float4 p = ld4(q);
            p = p * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BE) * tn + i0) +
                d2 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CE) * tn + i0);
            st4(q, p);
            acc = acc - p;
        }
    }
    const uchar4 id = *reinterpret_cast<const uchar4 *>(g.mat + o);
    const float4 cb = make_float4(cbs[id.x], cbs[id.y], cbs[id.z], cbs[id.w]);
    const float4 e = ld4(Ey + o);
    st4(Ey + o, sel(m, e + cb * acc, e));
}

// ---------------------------------------------------------------------------
// TF/SF face corrections + aux line + probes (one launch per half step)
// ---------------------------------------------------------------------------
// One thread per face cell.  A cell of component c gets, for split curl w
// (axis a = (c+1+w)%3, sign s = +1 for w = 0, -1 for w = 1) and delta = dir*g:
//   H:  Hs -= s*(kh_a + c'h_a)*delta,     psiH_a[w] += c'h_a*delta
//   E:  E  += cb*s*(ke_a + c'e_a)*delta,  psiE_a[w] += c'e_a*delta
// (c' terms only on slab cells), which is exactly the bulk update with the
// curl difference d replaced by d + delta.  Boxes of one component and one
// split curl never overlap; a first-curl and a second-curl box of the same E
// component share the cells on the TF box edges (the reference applies both
// corrections there), so the first-curl thread applies both and the
// second-curl thread skips those cells.
// Copy the box table to LDS (the per-thread box search and the edge checks
// then read LDS instead of chaining dependent global loads).  Called by every
// thread of a face-cell block before any of them returns.
__device__ __forceinline__ void stage_boxes(TfBox *sb, const TfBox *__restrict__ tb, int ntf) {
    constexpr int W = (int)(sizeof(TfBox) / sizeof(int));
    const int *src = reinterpret_cast<const int *>(tb);
    int *dst = reinterpret_cast<int *>(sb);
    for (int q = threadIdx.x; q < ntf * W; q += blockDim.x) dst[q] = src[q];
    __syncthreads();
}

__device__ __forceinline__ bool box_has(const TfBox &b, const int ijk[3]) {
    return ijk[0] >= b.lo[0] && ijk[0] <= b.hi[0] && ijk[1] >= b.lo[1] && ijk[1] <= b.hi[1] && ijk[2] >= b.lo[2] &&
           ijk[2] <= b.hi[2];
}

template <typename T, bool IS_H>
__device__ __forceinline__ float tf_term(const Grid<T> &g, const TfBox &bx, const IncParams &inc, const int ijk[3],
                                         double tp) {
    const double gi = inc_value(inc, ijk[0] + (double)bx.off[0], ijk[1] + (double)bx.off[1], ijk[2] + (double)bx.off[2],
                                tp, IS_H);
    const float delta = (float)(bx.dir * gi);
    const int a = (bx.comp + 1 + bx.curl) % 3;
    const int ia = ijk[a];
    const int tn = g.tn;
    float coef = g.tab[(a * 6 + (IS_H ? TK_KH : TK_KE)) * tn + ia];
    if (g.pml && (ia < g.slo[a] || ia >= g.shi[a])) {
        const float cp = g.tab[(a * 6 + (IS_H ? TK_CH : TK_CE)) * tn + ia];
        float *q = psi_base(g, a, (IS_H ? 0 : 2) + bx.curl) + psi_off(g, a, ijk[0], ijk[1], ijk[2]);
        *q = *q + cp * delta;
        coef += cp;
    }
    return (bx.curl == 0 ? 1.f : -1.f) * coef * delta;
}

// Block bid of a TF/SF launch over the face cells of one grid (blocks past
// the face cells run the extra block).
template <typename T, bool IS_H>
__device__ __forceinline__ void tfsf_block(const Grid<T> &g, const TfBox *__restrict__ tb_g, int ntf, int total,
                                           const IncParams &inc, const AuxLine &aux, const ProbeArgs &pr, Launch l,
                                           int bid) {
    const int t = g.counter[0] + l.toff;
    const int nb = (total + 255) / 256;
    if (bid >= nb) {
        extra_block<T, IS_H>(g, aux, pr, l, t);
        return;
    }
    __shared__ TfBox sb[kMaxTf];
    stage_boxes(sb, tb_g, ntf);
    const TfBox *tb = sb;
    const int gid = bid * 256 + threadIdx.x;
    if (gid >= total) return;
    int b = 0;
    while (b + 1 < ntf && gid >= tb[b + 1].start) ++b;
    const TfBox &bx = tb[b];
    // face-local index: x fastest where the face contains x (coalesced rows)
    const int fa = bx.fa, u = fa == 0 ? 1 : 0, v = fa == 2 ? 1 : 2;
    const int nu = bx.hi[u] - bx.lo[u] + 1;
    const int loc = gid - bx.start;
    int ijk[3];
    ijk[fa] = bx.lo[fa];
    ijk[u] = bx.lo[u] + loc % nu;
    ijk[v] = bx.lo[v] + loc / nu;
    const double tp = IS_H ? t - 0.5 : (double)t;
    const int c = bx.comp;
    if (!IS_H && bx.curl == 1) {
        for (int q = 0; q < ntf; ++q)
            if (tb[q].comp == c && tb[q].curl == 0 && box_has(tb[q], ijk)) return;   // owned by that thread
    }
    float upd = tf_term<T, IS_H>(g, bx, inc, ijk, tp);
    if (!IS_H && bx.curl == 0) {
        for (int q = 0; q < ntf; ++q)
            if (tb[q].comp == c && tb[q].curl == 1 && box_has(tb[q], ijk)) upd += tf_term<T, IS_H>(g, tb[q], inc, ijk, tp);
    }
    const long long o = (long long)ijk[2] * g.plane + (long long)ijk[1] * g.px + ijk[0];
    if (IS_H) {
        T *F = g.F + (3 + c) * g.fstride;
        st1(F + o, ld1(F + o) - upd);
    } else {
        T *F = g.F + c * g.fstride;
        const float cb = g.cb[c * 256 + g.mat[o]];
        st1(F + o, ld1(F + o) + cb * upd);
    }
}

template <typename T, bool IS_H>
__global__ __launch_bounds__(256) void k_tfsf(Grid<T> g, const TfBox *__restrict__ tb_g, int ntf, int total,
                                              IncParams inc, AuxLine aux, ProbeArgs pr, Launch l) {
    tfsf_block<T, IS_H>(g, tb_g, ntf, total, inc, aux, pr, l, blockIdx.x);
}

// Fused 3D step: the H corrections of the step applied to H of the buffer
// the fused step reads, H -= s*kh*delta per face cell (same delta and
// coefficient as k_tfsf<H>; the H face boxes of one component are disjoint,
// so every cell has one owner thread).  Every face lies outside the CPML
// slabs (checked on the host), so there is no psi term.  The extra block
// runs the aux line's H phase (it writes h; the face cells read e) and
// records the E probes of the previous step (E is not written here).
// NEXT: the corrections of step t+1 instead, inside the launch between steps
// t and t+1 (`aux` then holds e_t and h_{t+1/2}; the incident E of step t+1
// at a face node is computed here with aux_e_next, as the launch's extra
// block computes the whole line).
template <typename T, bool NEXT = false>
__device__ __forceinline__ void hpre_block(const Grid<T> &g, const TfBox *__restrict__ tb_g, int ntf, int total,
                                           const IncParams &inc, const AuxLine &aux, const ProbeArgs &pr, Launch l,
                                           int bid) {
    const int t = g.counter[0] + l.toff;
    const int nb = (total + 255) / 256;
    if (bid >= nb) {
        if (!NEXT) extra_block<T, true>(g, aux, pr, l, t);
        return;
    }
    __shared__ TfBox sb[kMaxTf];
    stage_boxes(sb, tb_g, ntf);
    const TfBox *tb = sb;
    const int gid = bid * 256 + threadIdx.x;
    if (gid >= total) return;
    int b = 0;
    while (b + 1 < ntf && gid >= tb[b + 1].start) ++b;
    const TfBox &bx = tb[b];
    // face-local index: x fastest where the face contains x (coalesced rows)
    const int fa = bx.fa, u = fa == 0 ? 1 : 0, v = fa == 2 ? 1 : 2;
    const int nu = bx.hi[u] - bx.lo[u] + 1;
    const int loc = gid - bx.start;
    int ijk[3];
    ijk[fa] = bx.lo[fa];
    ijk[u] = bx.lo[u] + loc % nu;
    ijk[v] = bx.lo[v] + loc / nu;
    const double cx = ijk[0] + (double)bx.off[0], cy = ijk[1] + (double)bx.off[1], cz = ijk[2] + (double)bx.off[2];
    double gi;
    if (NEXT && inc.mode == LX_INC_AUX1D) {
        const double cc = inc.axis == 0 ? cx : (inc.axis == 1 ? cy : cz);
        const int m = (int)floor(cc) - inc.x0;
        gi = (m < 0 || m >= inc.L) ? 0.0 : aux_e_next(aux, m, t);
    } else {
        gi = inc_value(inc, cx, cy, cz, NEXT ? t + 0.5 : t - 0.5, true);
    }
    const float delta = (float)(bx.dir * gi);
    const int a = (bx.comp + 1 + bx.curl) % 3;
    const float upd = (bx.curl == 0 ? 1.f : -1.f) * g.tab[(a * 6 + TK_KH) * g.tn + ijk[a]] * delta;
    const long long o = (long long)ijk[2] * g.plane + (long long)ijk[1] * g.px + ijk[0];
    T *H = g.F + (3 + bx.comp) * g.fstride;
    st1(H + o, ld1(H + o) - upd);
}

template <typename T>
__global__ __launch_bounds__(256) void k_tfsf_hpre(Grid<T> g, const TfBox *__restrict__ tb_g, int ntf, int total,
                                                   IncParams inc, AuxLine aux, ProbeArgs pr, Launch l) {
    hpre_block<T>(g, tb_g, ntf, total, inc, aux, pr, l, blockIdx.x);
}

// Fused 3D step, between steps t and t+1: one launch on the buffer step t
// wrote does what k_tfsf<E> of step t and k_tfsf_hpre of step t+1 do there.
// Blocks [0, nbE) apply the E corrections of step t, blocks [nbE, nbE+nbH)
// the H corrections of step t+1 (H there is what step t+1 reads), the last
// block moves the aux line from set ain (e_t, h_{t+1/2}) to set aout (e_{t+1}
// and h_{t+3/2}) and records probes.  No thread reads what another writes:
// the E corrections read h of ain, the H corrections compute their e_{t+1}
// nodes from ain themselves, E and H cells are distinct, the extra block
// writes only aout; the host allows the launch only when no H probe lies on
// an H face cell.  H probes are those of step t; E probes those of step t-1,
// read from the other buffer (final there, and nothing writes it now).
// last: the sequence's final step: no H corrections, aout gets e_{t+1} and
// h_{t+1/2} (the state the next sequence's k_tfsf_hpre expects).
template <typename T>
__device__ __forceinline__ void between_block(const Grid<T> &g, const T *__restrict__ Fprev,
                                              const TfBox *__restrict__ tbE, int ntfE, int totE,
                                              const TfBox *__restrict__ tbH, int ntfH, int totH, const IncParams &inc,
                                              const AuxLine &ain, const AuxLine &aout, const ProbeArgs &pr, Launch l,
                                              int last, int bid) {
    const int nbE = (totE + 255) / 256, nbH = last ? 0 : (totH + 255) / 256;
    if (bid < nbE) {
        IncParams ie = inc;
        ie.e = ain.e;
        ie.h = ain.h;
        tfsf_block<T, false>(g, tbE, ntfE, totE, ie, ain, pr, l, bid);
        return;
    }
    if (bid < nbE + nbH) {
        hpre_block<T, true>(g, tbH, ntfH, totH, inc, ain, pr, l, bid - nbE);
        return;
    }
    // aux line: one thread per node over between_aux_blocks blocks; the H
    // phase's e_{t+1}[m+1] is recomputed by the node's own thread (the same
    // routine and inputs as its owner), so no block waits for another
    const int t = g.counter[0] + l.toff;
    const int xb = bid - nbE - nbH;
    if (l.aux && xb < (ain.L + 255) / 256) {
        const int m = xb * 256 + threadIdx.x;
        if (m < ain.L) {
            const double e0 = aux_e_next(ain, m, t);
            aout.e[m] = e0;
            if (m < ain.L - 1) {
                if (!last) aout.h[m] = fma(-ain.S, aux_e_next(ain, m + 1, t) - e0, ain.h[m]);   // aux_h_next
                else if (aout.h != ain.h) aout.h[m] = ain.h[m];
            }
        }
    }
    if (xb != 0) return;
    if (l.probes && t < pr.cap)
        for (int q = threadIdx.x; q < pr.n; q += blockDim.x)
            if (pr.comp[q] >= 3)
                pr.rec[(long long)t * pr.n + q] = (double)ld1(g.F + pr.comp[q] * g.fstride + pr.idx[q]) * pr.hscale;
    if (l.eprobes && t >= 1 && t - 1 < pr.cap)
        for (int q = threadIdx.x; q < pr.n; q += blockDim.x)
            if (pr.comp[q] < 3)
                pr.rec[(long long)(t - 1) * pr.n + q] = (double)ld1(Fprev + pr.comp[q] * g.fstride + pr.idx[q]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_tfsf_between(Grid<T> g, const T *__restrict__ Fprev,
                                                      const TfBox *__restrict__ tbE, int ntfE, int totE,
                                                      const TfBox *__restrict__ tbH, int ntfH, int totH, IncParams inc,
                                                      AuxLine ain, AuxLine aout, ProbeArgs pr, Launch l, int last) {
    between_block<T>(g, Fprev, tbE, ntfE, totE, tbH, ntfH, totH, inc, ain, aout, pr, l, last, blockIdx.x);
}

// Batched TF/SF launch of one step over every member of a 3D batch: block b
// runs block bmap[b].y of member bmap[b].x.  pre: the H corrections on the
// buffer the step reads (buffer set par); else the E corrections on the
// buffer it wrote (set par ^ 1).
template <typename T, bool PRE>
__global__ __launch_bounds__(256) void k_tfsf_batch(const B3Sim<T> *__restrict__ sims, const int2 *__restrict__ bmap,
                                                    int par, int toff) {
    const int2 mb = bmap[blockIdx.x];
    const B3Sim<T> *d = sims + mb.x;
    Grid<T> g = d->g;
    const bool second = PRE ? par != 0 : par == 0;   // the buffer set these corrections apply to
    if (second) {
        g.F = g.Fn;
        g.psi = g.psin;
    }
    const Launch l{toff, d->use_aux, d->probes, d->probes};   // members have no point source
    if (PRE) hpre_block<T>(g, d->tfH, d->ntfH, d->totH, d->inc, d->aux, d->pr, l, mb.y);
    else tfsf_block<T, false>(g, d->tfE, d->ntfE, d->totE, d->inc, d->aux, d->pr, l, mb.y);
}

// The launch between steps toff and toff+1 of every member of a 3D batch
// (k_tfsf_between per member): step toff read buffer set par and wrote set
// par ^ 1; the member's aux line goes from set (toff & 1) to set
// ((toff + 1) & 1), or to set 0 at the chunk's last step.
template <typename T>
__global__ __launch_bounds__(256) void k_tfsf_between_batch(const B3Sim<T> *__restrict__ sims,
                                                            const int2 *__restrict__ bmap, int par, int toff, int last) {
    const int2 mb = bmap[blockIdx.x];
    const B3Sim<T> *d = sims + mb.x;
    Grid<T> g = d->g;
    const T *Fprev = g.Fn;
    if (par == 0) {
        Fprev = g.F;
        g.F = g.Fn;
        g.psi = g.psin;
    }
    const Launch l{toff, d->use_aux, d->probes, d->probes && toff >= 1};
    AuxLine ain = d->aux, aout = d->aux;
    if (toff & 1) {
        ain.e = d->aux.e2;
        ain.h = d->aux.h2;
    }
    if (!last && !(toff & 1)) {
        aout.e = d->aux.e2;
        aout.h = d->aux.h2;
    }
    between_block<T>(g, Fprev, d->tfE, d->ntfE, d->totE, d->tfH, d->ntfH, d->totH, d->inc, ain, aout, d->pr, l, last,
                     mb.y);
}

// ---------------------------------------------------------------------------
// point source, sequence end, gather, monitor, fills
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_point_source(Grid<T> g, PointSrc ps, ProbeArgs pr, int toff) {
    const int t = g.counter[0] + toff;
    // the E probes of step t-1 first: the source below changes one E cell
    // before step t's update, and the probes record E as step t-1 left it
    if (t >= 1 && t - 1 < pr.cap)
        for (int q = threadIdx.x; q < pr.n; q += blockDim.x)
            if (pr.comp[q] < 3)
                pr.rec[(long long)(t - 1) * pr.n + q] = (double)ld1(g.F + pr.comp[q] * g.fstride + pr.idx[q]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    T *E = g.F + ps.comp * g.fstride;
    if (ps.kind == LX_SRC_HARD_CW) {
        const int prod = (int)((unsigned)ps.omega * (unsigned)t);
        st1(E + ps.idx, (float)(ps.amp * sin(prod * ps.dt)));
    } else {
        const double w = waveform(ps.wave, t * ps.dt, ps.c, ps.lambda0, ps.fwhm, ps.delay);
        st1(E + ps.idx, (float)((double)ld1(E + ps.idx) + ps.amp * w));
    }
}

template <typename T>
__device__ __forceinline__ void finish_block(const Grid<T> &g, const ProbeArgs &pr, int n) {
    const int tlast = g.counter[0] + n - 1;
    for (int q = threadIdx.x; q < pr.n; q += blockDim.x)
        if (pr.comp[q] < 3 && tlast < pr.cap)
            pr.rec[(long long)tlast * pr.n + q] = (double)ld1(g.F + pr.comp[q] * g.fstride + pr.idx[q]);
    __syncthreads();
    if (threadIdx.x == 0) g.counter[0] = tlast + 1;
}

template <typename T>
__global__ void k_finish(Grid<T> g, ProbeArgs pr, int n) {
    finish_block<T>(g, pr, n);
}

// one block per member of a 3D batch; the state is in buffer set par_end
template <typename T>
__global__ void k_finish_batch(const B3Sim<T> *__restrict__ sims, int n, int par_end) {
    const B3Sim<T> *d = sims + blockIdx.x;
    Grid<T> g = d->g;
    if (par_end) g.F = g.Fn;
    finish_block<T>(g, d->pr, n);
}

// Device layout (k, j, i), i fastest, to the host layout (i, j, k), k
// fastest, in FP64: a transpose of each (k, i) plane of row j through a
// 64 x 64 LDS tile, so that the loads run along i and the stores along k.
// E2: |E|^2 at equal storage indices into the Ez-shaped layout, zero outside
// i < nx-1, j < ny-1, k < nz-1 (maxwell_three_dim_pml.cc:264-272); else one
// component (F) times scale.  Block: 256 threads, tile (i0, j, k0).
template <typename T, bool E2>
__global__ __launch_bounds__(256) void k_gather(Grid<T> g, const T *F, int ex, int ey, int ez, double scale,
                                               double *dst) {
    __shared__ double tile[64][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = blockIdx.x * 64, j = blockIdx.y, k0 = blockIdx.z * 64;
    const int i = i0 + lane;
    for (int r = w; r < 64; r += 4) {
        const int k = k0 + r;
        double v = 0.0;
        if (i < ex && k < ez) {
            const long long o = (long long)k * g.plane + (long long)j * g.px + i;
            if (E2) {
                if (i < g.nx - 1 && j < g.ny - 1 && k < g.nz - 1) {
                    const double a = ld1(g.F + o), b = ld1(g.F + g.fstride + o), c = ld1(g.F + 2 * g.fstride + o);
                    v = (c * c + b * b + a * a) * scale;
                }
            } else {
                v = (double)ld1(F + o) * scale;
            }
        }
        tile[r][lane] = v;
    }
    __syncthreads();
    const int k = k0 + lane;
    for (int r = w; r < 64; r += 4) {
        const int ii = i0 + r;
        if (ii < ex && k < ez) dst[((long long)ii * ey + j) * ez + k] = tile[lane][r];
    }
}

// Monitor, pass 1: max|E|^2 at equal indices (over the reference's index
// range), the discrete energy sum of eps_r*E^2 + (eta0*H)^2 and the count of
// non-finite values, one partial triple per block.  eps_r of each E
// component and material is cb0/cb (cb0 = c*dt/dx), computed once per block
// into LDS.  Block b takes rows b, b + gridDim.x, ... (a row: px contiguous
// cells of one (j, k)), two rows per iteration so that each thread has two
// rows' loads in flight; the threads of a block stride along x.  The cell
// partition and every reduction order are fixed, so the result does not
// change from run to run.
template <typename T, bool EONLY>
__device__ __forceinline__ void monitor_cell(const Grid<T> &g, int planar, const double (*seps)[256], long long o,
                                             bool in_range, double &mx, double &sum, int &bad) {
    const long long fs = g.fstride;
    if (EONLY) {   // max|E|^2 and the E components' non-finite count only
        if (planar) {
            const double e = ld1(g.F + fs + o);
            mx = fmax(mx, e * e);
            bad += !isfinite(e);
        } else {
            const double ex = ld1(g.F + o), ey = ld1(g.F + fs + o), ez = ld1(g.F + 2 * fs + o);
            if (in_range) mx = fmax(mx, ex * ex + ey * ey + ez * ez);
            bad += !isfinite(ex) || !isfinite(ey) || !isfinite(ez);
        }
        return;
    }
    if (planar) {
        const double e = ld1(g.F + fs + o);
        const double hx = ld1(g.F + 3 * fs + o), hz = ld1(g.F + 5 * fs + o);
        mx = fmax(mx, e * e);
        sum += e * e * seps[1][g.mat[o]] + hx * hx + hz * hz;
        bad += !isfinite(e) || !isfinite(hx) || !isfinite(hz);
    } else {
        const double ex = ld1(g.F + o), ey = ld1(g.F + fs + o), ez = ld1(g.F + 2 * fs + o);
        const double hx = ld1(g.F + 3 * fs + o), hy = ld1(g.F + 4 * fs + o), hz = ld1(g.F + 5 * fs + o);
        const int id = g.mat[o];
        if (in_range) mx = fmax(mx, ex * ex + ey * ey + ez * ez);
        sum += ex * ex * seps[0][id] + ey * ey * seps[1][id] + ez * ez * seps[2][id] + hx * hx + hy * hy + hz * hz;
        bad += !isfinite(ex) || !isfinite(ey) || !isfinite(ez) || !isfinite(hx) || !isfinite(hy) || !isfinite(hz);
    }
}

// block-wide (256 threads) max, sum and count; thread 0 holds the result
__device__ __forceinline__ void block_reduce3(double &mx, double &sum, int &bad) {
    __shared__ double smax[4], ssum[4];
    __shared__ int sbad[4];
    for (int d = 32; d >= 1; d >>= 1) {
        mx = fmax(mx, __shfl_down(mx, d, 64));
        sum += __shfl_down(sum, d, 64);
        bad += __shfl_down(bad, d, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smax[w] = mx;
        ssum[w] = sum;
        sbad[w] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        mx = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
        sum = (ssum[0] + ssum[1]) + (ssum[2] + ssum[3]);
        bad = (sbad[0] + sbad[1]) + (sbad[2] + sbad[3]);
    }
}

// EONLY: the reference's monitor alone (max|E|^2 and the E components'
// non-finite count; 3*s B per cell instead of 6*s+1), energy left 0.
template <typename T, bool EONLY>
__global__ __launch_bounds__(256) void k_monitor(Grid<T> g, int planar, double cb0, double *part) {
    __shared__ double seps[3][256];
    if (!EONLY) {
        for (int q = threadIdx.x; q < 3 * 256; q += 256) seps[q >> 8][q & 255] = cb0 / (double)g.cb[q];
        __syncthreads();
    }
    const int ny = planar ? 1 : g.ny, nz = g.nz, nx = g.nx;
    const long long rows = (long long)ny * nz;
    double mx = 0.0, sum = 0.0;
    int bad = 0;
    for (long long r = blockIdx.x; r < rows; r += 2 * (long long)gridDim.x) {
        const long long r2 = r + gridDim.x;
        const bool two = r2 < rows;
        const int j = (int)(r % ny), k = (int)(r / ny);
        const int j2 = two ? (int)(r2 % ny) : 0, k2 = two ? (int)(r2 / ny) : 0;
        const long long base = (long long)k * g.plane + (long long)j * g.px;
        const long long base2 = (long long)k2 * g.plane + (long long)j2 * g.px;
        const bool rin = j < ny - 1 && k < nz - 1, rin2 = j2 < ny - 1 && k2 < nz - 1;
        for (int i = threadIdx.x; i < nx; i += 256) {
            monitor_cell<T, EONLY>(g, planar, seps, base + i, rin && i < nx - 1, mx, sum, bad);
            if (two) monitor_cell<T, EONLY>(g, planar, seps, base2 + i, rin2 && i < nx - 1, mx, sum, bad);
        }
    }
    block_reduce3(mx, sum, bad);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = mx;
        part[3 * blockIdx.x + 1] = sum;
        part[3 * blockIdx.x + 2] = (double)bad;
    }
}

// Monitor, pass 2 (one block): the partial triples of pass 1 into out[0..2],
// each thread over a fixed subset in a fixed order.
__global__ __launch_bounds__(256) void k_monitor_final(const double *__restrict__ part, int nb, double *out) {
    double mx = 0.0, sum = 0.0;
    int bad = 0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        mx = fmax(mx, part[3 * b]);
        sum += part[3 * b + 1];
        bad += (int)part[3 * b + 2];
    }
    block_reduce3(mx, sum, bad);
    if (threadIdx.x == 0) {
        out[0] = mx;
        out[1] = sum;
        out[2] = (double)bad;
    }
}

// Material ids from the host's C order (i, j, k; k fastest) to the device
// layout (k, j, x; x fastest, rows px long).  One thread per (i, j, k),
// threads along i so that the stores are contiguous.
__global__ void k_ids_to_device(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, int nx, int ny, int nz,
                                int px, long long plane) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, k = blockIdx.z;
    if (i >= nx) return;
    dst[(long long)k * plane + (long long)j * px + i] = src[((long long)i * ny + j) * nz + k];
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
void launch_bulk(const Grid<T> &g, int planar, int half, hipStream_t s) {
    const dim3 grd((unsigned)g.ntiles), blk(256);
    if (planar) {
        if (half == 0) {
            if (g.pml) hipLaunchKernelGGL((k_h2<T, true>), grd, blk, 0, s, g);
            else hipLaunchKernelGGL((k_h2<T, false>), grd, blk, 0, s, g);
        } else {
            if (g.pml) hipLaunchKernelGGL((k_e2<T, true>), grd, blk, 0, s, g);
            else hipLaunchKernelGGL((k_e2<T, false>), grd, blk, 0, s, g);
        }
    } else {
        if (half == 0) {
            if (g.pml) hipLaunchKernelGGL((k_h3<T, true>), grd, blk, 0, s, g);
            else hipLaunchKernelGGL((k_h3<T, false>), grd, blk, 0, s, g);
        } else {
            if (g.pml) hipLaunchKernelGGL((k_e3<T, true>), grd, blk, 0, s, g);
            else hipLaunchKernelGGL((k_e3<T, false>), grd, blk, 0, s, g);
        }
    }
}

template <typename T>
void launch_tfsf(const Grid<T> &g, const TfList &tf, const IncParams &inc, const AuxLine &aux, const ProbeArgs &pr,
                 Launch l, int half, hipStream_t s) {
    const int nb = (tf.total + 255) / 256;
    const unsigned blocks = (unsigned)nb + ((l.aux || l.probes) ? 1u : 0u);
    if (blocks == 0) return;
    if (half == 0) hipLaunchKernelGGL((k_tfsf<T, true>), dim3(blocks), dim3(256), 0, s, g, tf.dev, tf.n, tf.total, inc, aux, pr, l);
    else hipLaunchKernelGGL((k_tfsf<T, false>), dim3(blocks), dim3(256), 0, s, g, tf.dev, tf.n, tf.total, inc, aux, pr, l);
}

template <typename T>
void launch_tfsf_hpre(const Grid<T> &g, const TfList &tf, const IncParams &inc, const AuxLine &aux,
                      const ProbeArgs &pr, Launch l, hipStream_t s) {
    const int nb = (tf.total + 255) / 256;
    const unsigned blocks = (unsigned)nb + ((l.aux || l.probes) ? 1u : 0u);
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_tfsf_hpre<T>, dim3(blocks), dim3(256), 0, s, g, tf.dev, tf.n, tf.total, inc, aux, pr, l);
}

template <typename T>
void launch_tfsf_between(const Grid<T> &g, const T *Fprev, const TfList &tfE, const TfList &tfH, const IncParams &inc,
                         const AuxLine &ain, const AuxLine &aout, const ProbeArgs &pr, Launch l, int last,
                         hipStream_t s) {
    const int nbE = (tfE.total + 255) / 256, nbH = last ? 0 : (tfH.total + 255) / 256;
    const unsigned blocks = (unsigned)(nbE + nbH + between_extra_blocks(l.aux ? ain.L : 0, l.probes || l.eprobes));
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_tfsf_between<T>, dim3(blocks), dim3(256), 0, s, g, Fprev, tfE.dev, tfE.n, tfE.total, tfH.dev,
                       tfH.n, tfH.total, inc, ain, aout, pr, l, last);
}

template <typename T>
void launch_finish(const Grid<T> &g, const ProbeArgs &pr, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_finish<T>, dim3(1), dim3(64), 0, s, g, pr, n);
}
template <typename T>
void launch_tfsf_batch(const B3Sim<T> *sims, const int2 *bmap, int nblocks, int pre, int par, int toff,
                       hipStream_t s) {
    if (nblocks <= 0) return;
    if (pre) hipLaunchKernelGGL((k_tfsf_batch<T, true>), dim3(nblocks), dim3(256), 0, s, sims, bmap, par, toff);
    else hipLaunchKernelGGL((k_tfsf_batch<T, false>), dim3(nblocks), dim3(256), 0, s, sims, bmap, par, toff);
}
template <typename T>
void launch_tfsf_between_batch(const B3Sim<T> *sims, const int2 *bmap, int nblocks, int par, int toff, int last,
                               hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_tfsf_between_batch<T>, dim3(nblocks), dim3(256), 0, s, sims, bmap, par, toff, last);
}
template <typename T>
void launch_finish_batch(const B3Sim<T> *sims, int nsims, int n, int par_end, hipStream_t s) {
    if (nsims <= 0) return;
    hipLaunchKernelGGL(k_finish_batch<T>, dim3(nsims), dim3(64), 0, s, sims, n, par_end);
}
template <typename T>
void launch_point_source(const Grid<T> &g, const PointSrc &ps, const ProbeArgs &pr, int toff, hipStream_t s) {
    hipLaunchKernelGGL(k_point_source<T>, dim3(1), dim3(64), 0, s, g, ps, pr, toff);
}
template <typename T>
void launch_gather(const Grid<T> &g, int comp, int ex, int ey, int ez, double scale, double *dst, hipStream_t s) {
    const T *F = g.F + comp * g.fstride;
    dim3 blk(256), grd((ex + 63) / 64, ey, (ez + 63) / 64);
    hipLaunchKernelGGL((k_gather<T, false>), grd, blk, 0, s, g, F, ex, ey, ez, scale, dst);
}
template <typename T>
void launch_gather_e2(const Grid<T> &g, int ex, int ey, int ez, double scale, double *dst, hipStream_t s) {
    dim3 blk(256), grd((ex + 63) / 64, ey, (ez + 63) / 64);
    hipLaunchKernelGGL((k_gather<T, true>), grd, blk, 0, s, g, (const T *)g.F, ex, ey, ez, scale, dst);
}
template <typename T>
void launch_monitor(const Grid<T> &g, int planar, double cb0, double *out, bool e_only, hipStream_t s, double *res) {
    const long long rows = (long long)(planar ? 1 : g.ny) * g.nz;
    const int nb = (int)std::min<long long>(kMonBlocks, std::max<long long>(rows, 1));
    double *part = out + 4;
    if (e_only) hipLaunchKernelGGL((k_monitor<T, true>), dim3(nb), dim3(256), 0, s, g, planar, cb0, part);
    else hipLaunchKernelGGL((k_monitor<T, false>), dim3(nb), dim3(256), 0, s, g, planar, cb0, part);
    hipLaunchKernelGGL(k_monitor_final, dim3(1), dim3(256), 0, s, part, nb, res ? res : out);
}
void launch_ids_to_device(const uint8_t *src, uint8_t *dst, int nx, int ny, int nz, int px, long long plane,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_ids_to_device, dim3((nx + 255) / 256, ny, nz), dim3(256), 0, s, src, dst, nx, ny, nz, px, plane);
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
    template void launch_bulk<T>(const Grid<T> &, int, int, hipStream_t);                                    \
    template void launch_tfsf<T>(const Grid<T> &, const TfList &, const IncParams &, const AuxLine &,        \
                                 const ProbeArgs &, Launch, int, hipStream_t);                               \
    template void launch_finish<T>(const Grid<T> &, const ProbeArgs &, int, hipStream_t);                    \
    template void launch_tfsf_between<T>(const Grid<T> &, const T *, const TfList &, const TfList &,           \
                                         const IncParams &, const AuxLine &, const AuxLine &, const ProbeArgs &, \
                                         Launch, int, hipStream_t);                                           \
    template void launch_tfsf_hpre<T>(const Grid<T> &, const TfList &, const IncParams &, const AuxLine &,   \
                                      const ProbeArgs &, Launch, hipStream_t);                              \
    template void launch_point_source<T>(const Grid<T> &, const PointSrc &, const ProbeArgs &, int, hipStream_t); \
    template void launch_gather<T>(const Grid<T> &, int, int, int, int, double, double *, hipStream_t);      \
    template void launch_monitor<T>(const Grid<T> &, int, double, double *, bool, hipStream_t, double *);                  \
    template void launch_gather_e2<T>(const Grid<T> &, int, int, int, double, double *, hipStream_t);              \
    template void launch_tfsf_batch<T>(const B3Sim<T> *, const int2 *, int, int, int, int, hipStream_t);   \
    template void launch_finish_batch<T>(const B3Sim<T> *, int, int, int, hipStream_t);                   \
    template void launch_tfsf_between_batch<T>(const B3Sim<T> *, const int2 *, int, int, int, int, hipStream_t);
LX_INST(float)
LX_INST(__half)

}  // namespace lx
