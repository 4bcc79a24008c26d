// lx_step3.hip -- fused 3D Yee timestep: the H half step and the E half step
// of one timestep in a single z-march, reading one buffer set (fields, psi)
// and writing the other.
//
// The two half-step kernels read every E value twice per step (once by the
// H update, once by the E update) and every H value twice; here each old
// value is read once and each new value written once, 12*s + 1 bytes per
// cell and step plus the slab psi, against 18*s + 1 for the half-step pair
// (maxwell_three_dim_pml.cc:182-255 is the update order followed).
//
// Block: NW wavefronts, wave w owns grid row jr = j0 - 1 + w of its tile
// (j0 = first of the tile's NW-1 output rows), every x cell of that row (in
// chunks of 256, 4 cells per lane) and the planes [k0, k1).  Per plane k:
//   H: every wave computes H_new(jr, k) from the old buffer: E_old at k and
//      k+1 (a register queue along z), at jr+1 (load) and at x+1 (next lane).
//      Wave 0's row j0-1 and the plane k0-1 are halo: computed, because the E
//      update needs them, but not stored (their owner stores them).  TF/SF
//      corrections of the H half step, precomputed per face cell by
//      k_tfsf_hprep, are added to every computed cell.
//   exchange: Hx and Hz of every wave's row go through LDS.
//   E: waves 1..NW-1 compute E_new(jr, k) from H_new(jr, k), H_new(jr, k-1)
//      (registers), H_new(jr-1, k) (wave w-1 through LDS) and x-1 (previous
//      lane; lane 0 takes the last cell of the previous x chunk from LDS).
// Every cell of a stored row is written, non-updated cells with their old
// value, so both buffers always hold a complete state.  Neighbouring tiles
// only read the old buffer, so there is no ordering between blocks.
#include "lx_device.hpp"

namespace lx {

constexpr int kMaxKz = 64;

template <typename T>
__device__ __forceinline__ float *psi_new(const Grid<T> &g, int a, int slot) {
    return g.psin + (psi_base(g, a, slot) - g.psi);
}

// H_new(row, k) of the three components for the 4 cells of a lane, with the
// row's psi (returned in p*), from E_old at (row, k), (row, k+1), (row+1, k).
struct HRow {
    float4 hx, hy, hz;
    float4 py0, pz1, pz0, px1, px0, py1;   // psi: y0 z1 (Hx), z0 x1 (Hy), x0 y1 (Hz)
};

template <typename T, bool PML, int NW>
__global__ __launch_bounds__(64 * NW) void k_step3(Grid<T> g, const TfBox *__restrict__ tbH, int ntfH,
                                                   const float *__restrict__ dh, const float *__restrict__ dp,
                                                   int nunits) {
    __shared__ float cbs[256];
    __shared__ float xedge[NW][kMaxKz + 1][2];     // per wave and plane: Hy, Hz of the previous chunk's last cell
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int q = threadIdx.x; q < 256; q += 64 * NW) cbs[q] = g.cb[q];
    __syncthreads();
    const int unit = xcd_remap(blockIdx.x, (nunits + NW - 1) / NW) * NW + w;
    if (unit >= nunits) return;                    // whole wavefront; no barrier follows
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int jr = unit % ny, tz = unit / ny;
    const int k0 = tz * g.kz, k1 = min(k0 + g.kz, nz);
    const int ks = k0 > 0 ? k0 - 1 : 0;
    const long long px = g.px, plane = g.plane, fs = g.fstride;
    const T *Ex = g.F, *Ey = g.F + fs, *Ez = g.F + 2 * fs;
    const T *Hx = g.F + 3 * fs, *Hy = g.F + 4 * fs, *Hz = g.F + 5 * fs;
    T *Exn = g.Fn, *Eyn = g.Fn + fs, *Ezn = g.Fn + 2 * fs;
    T *Hxn = g.Fn + 3 * fs, *Hyn = g.Fn + 4 * fs, *Hzn = g.Fn + 5 * fs;
    const int tn = g.tn;
    const float *tab = g.tab;
    const bool has_m = jr >= 1;                    // row jr-1 exists (its Hx, Hz feed Ex, Ez)
    const int jm = has_m ? jr - 1 : 0;
    const int jp = min(jr + 1, ny - 1);
    const bool rowx = jr < ny - 1;                 // Hx, Hz rows of jr; Ey rows
    const bool jin = jr >= 1 && jr <= ny - 2;      // Ex, Ez rows
    // row coefficients of jr and jr-1
    const float khy = tab[(1 * 6 + TK_KH) * tn + min(jr, ny - 2)], khym = tab[(1 * 6 + TK_KH) * tn + jm];
    const float key = tab[(1 * 6 + TK_KE) * tn + jr];
    const bool sy = PML && (jr < g.slo[1] || jr >= g.shi[1]);
    const bool sym = PML && has_m && (jm < g.slo[1] || jm >= g.shi[1]);
    const float bhy = sy ? tab[(1 * 6 + TK_BH) * tn + jr] : 0.f, chy = sy ? tab[(1 * 6 + TK_CH) * tn + jr] : 0.f;
    const float bhym = sym ? tab[(1 * 6 + TK_BH) * tn + jm] : 0.f, chym = sym ? tab[(1 * 6 + TK_CH) * tn + jm] : 0.f;
    const float bey = sy ? tab[(1 * 6 + TK_BE) * tn + jr] : 0.f, cey = sy ? tab[(1 * 6 + TK_CE) * tn + jr] : 0.f;
    const int siy = sy ? slab_index(jr, g.slo[1], g.shi[1]) : 0;
    const int siym = sym ? slab_index(jm, g.slo[1], g.shi[1]) : 0;

    for (int x0 = 0; x0 < nx; x0 += 256) {
        const int i0 = x0 + lane * 4;
        const bool on = i0 < nx;
        const bool onm = on && has_m;
        const unsigned mx_all = on ? xmask(i0, 0, nx - 1) : 0u;
        const unsigned mx_in = on ? xmask(i0, 0, nx - 2) : 0u;
        const unsigned m_ex = on && jin ? xmask(i0, 0, nx - 2) : 0u;
        const unsigned m_ey = on && rowx ? xmask(i0, 1, nx - 2) : 0u;
        const unsigned m_ez = on && jin ? xmask(i0, 1, nx - 2) : 0u;
        const int ic = min(i0, g.px - 4);
        const float4 khx = *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_KH) * tn + ic);
        const float4 kex = *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_KE) * tn + ic);
        const bool sx = PML && on && (i0 < g.slo[0] || i0 >= g.shi[0]);
        const int xs = sx ? (i0 < g.slo[0] ? i0 : g.slo[0] + (i0 - g.shi[0])) : 0;
        const bool last_lane = (lane == 63) || (i0 + 4 >= nx);
        const bool nxt = i0 + 4 < nx;
        long long o = (long long)ks * plane + (long long)jr * px + i0;
        const long long djp = (long long)(jp - jr) * px, djm = (long long)(jr - jm) * px;
        // E_old queue along z: row jr all components, row jr-1 Ey
        float4 ex0 = f4(0.f), ey0 = f4(0.f), ez0 = f4(0.f), eym0 = f4(0.f);
        if (on) { ex0 = ld4(Ex + o); ey0 = ld4(Ey + o); ez0 = ld4(Ez + o); }
        if (onm) eym0 = ld4(Ey + o - djm);
        float4 hxm = f4(0.f), hym = f4(0.f);       // H_new(jr, k-1): Hx, Hy

        for (int k = ks; k < k1; ++k, o += plane) {
            const bool halo_plane = k < k0;
            const bool kin = k < nz - 1;
            float4 ex1 = f4(0.f), ey1 = f4(0.f), ez1 = f4(0.f), eym1 = f4(0.f);
            if (on && kin) { ex1 = ld4(Ex + o + plane); ey1 = ld4(Ey + o + plane); ez1 = ld4(Ez + o + plane); }
            if (onm && kin) eym1 = ld4(Ey + o - djm + plane);
            float ezn = __shfl_down(ez0.x, 1), eyn = __shfl_down(ey0.x, 1);
            if (last_lane) {
                ezn = (on && nxt) ? ld1(Ez + o + 4) : 0.f;
                eyn = (on && nxt) ? ld1(Ey + o + 4) : 0.f;
            }
            const bool sz = PML && (k < g.slo[2] || k >= g.shi[2]);
            const long long zs = sz ? (long long)slab_index(k, g.slo[2], g.shi[2]) * ny : 0;
            const long long zoff = (zs + jr) * px + i0, zoffm = (zs + jm) * px + i0;
            const long long yoff = sy ? ((long long)k * g.sw[1] + siy) * px + i0 : 0;
            const long long yoffm = sym ? ((long long)k * g.sw[1] + siym) * px + i0 : 0;
            const long long xoff = sx ? ((long long)k * ny + jr) * g.sw[0] + xs : 0;
            const long long xoffm = sx ? ((long long)k * ny + jm) * g.sw[0] + xs : 0;
            const float khz = tab[(2 * 6 + TK_KH) * tn + min(k, nz - 2)];
            const float bhz = sz ? tab[(2 * 6 + TK_BH) * tn + k] : 0.f, chz = sz ? tab[(2 * 6 + TK_CH) * tn + k] : 0.f;

            // ---------------- H half step, row jr
            float4 hx = f4(0.f), hy = f4(0.f), hz = f4(0.f);
            if (on) { hx = ld4(Hx + o); hy = ld4(Hy + o); hz = ld4(Hz + o); }
            float4 py0 = f4(0.f), pz1 = f4(0.f), pz0 = f4(0.f), px1 = f4(0.f), px0 = f4(0.f), py1 = f4(0.f);
            const bool dhx = on && rowx && kin, dhy = on && kin, dhz = on && rowx;
            if (dhx) {   // Hx: d1 = dEz/dy, d2 = dEy/dz
                const float4 d1 = ld4(Ez + o + djp) - ez0, d2 = ey1 - ey0;
                float4 acc = d1 * khy - d2 * khz;
                if (sy) { py0 = ld4(psi_base(g, 1, 0) + yoff) * bhy + d1 * chy; acc = acc + py0; }
                if (sz) { pz1 = ld4(psi_base(g, 2, 1) + zoff) * bhz + d2 * chz; acc = acc - pz1; }
                hx = sel(mx_all, hx - acc, hx);
            }
            if (dhy) {   // Hy: d1 = dEx/dz, d2 = dEz/dx
                const float4 ezi = make_float4(ez0.y, ez0.z, ez0.w, ezn);
                const float4 d1 = ex1 - ex0, d2 = ezi - ez0;
                float4 acc = d1 * khz - d2 * khx;
                if (sz) { pz0 = ld4(psi_base(g, 2, 0) + zoff) * bhz + d1 * chz; acc = acc + pz0; }
                if (sx) {
                    px1 = ld4(psi_base(g, 0, 1) + xoff) * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BH) * tn + i0) +
                          d2 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CH) * tn + i0);
                    acc = acc - px1;
                }
                hy = sel(mx_in, hy - acc, hy);
            }
            if (dhz) {   // Hz: d1 = dEy/dx, d2 = dEx/dy
                const float4 eyi = make_float4(ey0.y, ey0.z, ey0.w, eyn);
                const float4 d1 = eyi - ey0, d2 = ld4(Ex + o + djp) - ex0;
                float4 acc = d1 * khx - d2 * khy;
                if (sx) {
                    px0 = ld4(psi_base(g, 0, 0) + xoff) * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BH) * tn + i0) +
                          d1 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CH) * tn + i0);
                    acc = acc + px0;
                }
                if (sy) { py1 = ld4(psi_base(g, 1, 1) + yoff) * bhy + d2 * chy; acc = acc - py1; }
                hz = sel(mx_in, hz - acc, hz);
            }
            // ---------------- Hx, Hz of row jr-1 (needed by Ex, Ez of row jr)
            float4 hxj = f4(0.f), hzj = f4(0.f);
            float4 qy0 = f4(0.f), qz1 = f4(0.f), qx0 = f4(0.f), qy1 = f4(0.f);
            float eynm = __shfl_down(eym0.x, 1);
            if (onm) {
                hxj = ld4(Hx + o - djm);
                hzj = ld4(Hz + o - djm);
                const float4 ezm = ld4(Ez + o - djm), exm = ld4(Ex + o - djm);
                if (kin) {   // Hx(jr-1): d1 = Ez(jr) - Ez(jr-1), d2 = Ey(jr-1, k+1) - Ey(jr-1, k)
                    const float4 d1 = ez0 - ezm, d2 = eym1 - eym0;
                    float4 acc = d1 * khym - d2 * khz;
                    if (sym) { qy0 = ld4(psi_base(g, 1, 0) + yoffm) * bhym + d1 * chym; acc = acc + qy0; }
                    if (sz) { qz1 = ld4(psi_base(g, 2, 1) + zoffm) * bhz + d2 * chz; acc = acc - qz1; }
                    hxj = sel(mx_all, hxj - acc, hxj);
                }
                {            // Hz(jr-1): d1 = dEy(jr-1)/dx, d2 = Ex(jr) - Ex(jr-1)
                    if (last_lane) eynm = nxt ? ld1(Ey + o - djm + 4) : 0.f;
                    const float4 eyi = make_float4(eym0.y, eym0.z, eym0.w, eynm);
                    const float4 d1 = eyi - eym0, d2 = ex0 - exm;
                    float4 acc = d1 * khx - d2 * khym;
                    if (sx) {
                        qx0 = ld4(psi_base(g, 0, 0) + xoffm) * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BH) * tn + i0) +
                              d1 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CH) * tn + i0);
                        acc = acc + qx0;
                    }
                    if (sym) { qy1 = ld4(psi_base(g, 1, 1) + yoffm) * bhym + d2 * chym; acc = acc - qy1; }
                    hzj = sel(mx_in, hzj - acc, hzj);
                }
            }
            // TF/SF corrections of the H half step (maxwell.cc:512-594 face
            // sets): H -= dh, psi += dp, with dh = s*(kh + c'h)*delta and
            // dp = c'h*delta per face cell, computed by k_tfsf_hprep.  Rows jr
            // (all components) and jr-1 (Hx, Hz) both get theirs.
            for (int b = 0; b < ntfH; ++b) {
                const TfBox &bx = tbH[b];
                if (k < bx.lo[2] || k > bx.hi[2]) continue;   // wave-uniform
                const bool on_j = jr >= bx.lo[1] && jr <= bx.hi[1];
                const bool on_jm = has_m && bx.comp != 1 && jm >= bx.lo[1] && jm <= bx.hi[1];
                if (!on_j && !on_jm) continue;
                const int a = (bx.comp + 1 + bx.curl) % 3;
                if (bx.fa == 0) {
                    // x face: one cell per row, at a fixed lane and element;
                    // its values are wave-uniform (scalar loads)
                    const int fx = bx.lo[0] - x0;
                    if (fx < 0 || fx >= 256) continue;
                    const int fl = fx >> 2, fe = fx & 3;
                    const int nu = bx.hi[1] - bx.lo[1] + 1;
                    const int base = bx.start + nu * (k - bx.lo[2]) - bx.lo[1];
                    const float vj = on_j ? dh[base + jr] : 0.f, vm = on_jm ? dh[base + jm] : 0.f;
                    const float qj = (PML && on_j) ? dp[base + jr] : 0.f;
                    if (lane == fl) {
                        if (bx.comp == 0) {
                            set(hx, fe, get(hx, fe) - vj);
                            set(hxj, fe, get(hxj, fe) - vm);
                        } else if (bx.comp == 1) {
                            set(hy, fe, get(hy, fe) - vj);
                        } else {
                            set(hz, fe, get(hz, fe) - vj);
                            set(hzj, fe, get(hzj, fe) - vm);
                        }
                        if (PML) {   // an x face carries the x-derivative correction (a = 0)
                            if (bx.curl == 0) set(px0, fe, get(px0, fe) + qj);
                            else set(px1, fe, get(px1, fe) + qj);
                        }
                    }
                    continue;
                }
                const unsigned m = on ? xmask(i0, bx.lo[0], bx.hi[0]) : 0u;
                if (!m) continue;
                const int fa = bx.fa, u = fa == 2 ? 0 : fa + 1, v = fa == 0 ? 2 : fa - 1;
                const int nu = bx.hi[u] - bx.lo[u] + 1;
                for (int e = 0; e < 4; ++e) {
                    if (!(m & (1u << e))) continue;
                    for (int rr = 0; rr < 2; ++rr) {
                        if (rr == 0 ? !on_j : !on_jm) continue;
                        const int jj = rr == 0 ? jr : jm;
                        const int cu = u == 0 ? i0 + e : (u == 1 ? jj : k);
                        const int cv = v == 0 ? i0 + e : (v == 1 ? jj : k);
                        const int idx = bx.start + (cu - bx.lo[u]) + nu * (cv - bx.lo[v]);
                        const float vh = dh[idx];
                        const float q = PML ? dp[idx] : 0.f;
                        if (rr == 0) {
                            if (bx.comp == 0) set(hx, e, get(hx, e) - vh);
                            else if (bx.comp == 1) set(hy, e, get(hy, e) - vh);
                            else set(hz, e, get(hz, e) - vh);
                            if (PML) {
                                if (a == 0) {
                                    if (bx.curl == 0) set(px0, e, get(px0, e) + q);
                                    else set(px1, e, get(px1, e) + q);
                                } else if (a == 1) {
                                    if (bx.curl == 0) set(py0, e, get(py0, e) + q);
                                    else set(py1, e, get(py1, e) + q);
                                } else {
                                    if (bx.curl == 0) set(pz0, e, get(pz0, e) + q);
                                    else set(pz1, e, get(pz1, e) + q);
                                }
                            }
                        } else {
                            // halo row: only the field matters (its psi is not stored)
                            if (bx.comp == 0) set(hxj, e, get(hxj, e) - vh);
                            else set(hzj, e, get(hzj, e) - vh);
                        }
                    }
                }
            }
            (void)qy0; (void)qz1; (void)qx0; (void)qy1;
            if (!halo_plane && on) {
                st4(Hxn + o, hx);
                st4(Hyn + o, hy);
                st4(Hzn + o, hz);
                if (sy && dhx) st4(psi_new(g, 1, 0) + yoff, py0);
                if (sz && dhx) st4(psi_new(g, 2, 1) + zoff, pz1);
                if (sz && dhy) st4(psi_new(g, 2, 0) + zoff, pz0);
                if (sx && dhy) st4(psi_new(g, 0, 1) + xoff, px1);
                if (sx && dhz) st4(psi_new(g, 0, 0) + xoff, px0);
                if (sy && dhz) st4(psi_new(g, 1, 1) + yoff, py1);
            }

            // ---------------- E half step, row jr
            float hzp = __shfl_up(hz.w, 1), hyp = __shfl_up(hy.w, 1);
            if (lane == 0) {
                hyp = x0 > 0 ? xedge[w][k - ks][0] : 0.f;
                hzp = x0 > 0 ? xedge[w][k - ks][1] : 0.f;
            }
            if (!halo_plane && on) {
                const bool kine = k >= 1 && k <= nz - 2;
                const float kez = tab[(2 * 6 + TK_KE) * tn + k];
                const float bez = sz ? tab[(2 * 6 + TK_BE) * tn + k] : 0.f, cez = sz ? tab[(2 * 6 + TK_CE) * tn + k] : 0.f;
                float4 ex = ex0, ey = ey0, ez = ez0;
                const uchar4 id = *reinterpret_cast<const uchar4 *>(g.mat + o);
                const float4 cb = make_float4(cbs[id.x], cbs[id.y], cbs[id.z], cbs[id.w]);
                if (kine && m_ex) {   // Ex: d1 = dHz/dy, d2 = dHy/dz
                    const float4 d1 = hz - hzj, d2 = hy - hym;
                    float4 acc = d1 * key - d2 * kez;
                    if (sy) {
                        const float4 p = ld4(psi_base(g, 1, 2) + yoff) * bey + d1 * cey;
                        st4(psi_new(g, 1, 2) + yoff, p);
                        acc = acc + p;
                    }
                    if (sz) {
                        const float4 p = ld4(psi_base(g, 2, 3) + zoff) * bez + d2 * cez;
                        st4(psi_new(g, 2, 3) + zoff, p);
                        acc = acc - p;
                    }
                    ex = sel(m_ex, ex0 + cb * acc, ex0);
                }
                if (kine && m_ey) {   // Ey: d1 = dHx/dz, d2 = dHz/dx
                    const float4 hzi = make_float4(hzp, hz.x, hz.y, hz.z);
                    const float4 d1 = hx - hxm, d2 = hz - hzi;
                    float4 acc = d1 * kez - d2 * kex;
                    if (sz) {
                        const float4 p = ld4(psi_base(g, 2, 2) + zoff) * bez + d1 * cez;
                        st4(psi_new(g, 2, 2) + zoff, p);
                        acc = acc + p;
                    }
                    if (sx) {
                        const float4 p = ld4(psi_base(g, 0, 3) + xoff) *
                                             *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BE) * tn + i0) +
                                         d2 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CE) * tn + i0);
                        st4(psi_new(g, 0, 3) + xoff, p);
                        acc = acc - p;
                    }
                    ey = sel(m_ey, ey0 + cb * acc, ey0);
                }
                if (kin && m_ez) {   // Ez: d1 = dHy/dx, d2 = dHx/dy
                    const float4 hyi = make_float4(hyp, hy.x, hy.y, hy.z);
                    const float4 d1 = hy - hyi, d2 = hx - hxj;
                    float4 acc = d1 * kex - d2 * key;
                    if (sx) {
                        const float4 p = ld4(psi_base(g, 0, 2) + xoff) *
                                             *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_BE) * tn + i0) +
                                         d1 * *reinterpret_cast<const float4 *>(tab + (0 * 6 + TK_CE) * tn + i0);
                        st4(psi_new(g, 0, 2) + xoff, p);
                        acc = acc + p;
                    }
                    if (sy) {
                        const float4 p = ld4(psi_base(g, 1, 3) + yoff) * bey + d2 * cey;
                        st4(psi_new(g, 1, 3) + yoff, p);
                        acc = acc - p;
                    }
                    ez = sel(m_ez, ez0 + cb * acc, ez0);
                }
                st4(Exn + o, ex);
                st4(Eyn + o, ey);
                st4(Ezn + o, ez);
            }
            // the last cell of this chunk, for lane 0 of the next chunk (same wave)
            if (lane == 63) {
                xedge[w][k - ks][0] = hy.w;
                xedge[w][k - ks][1] = hz.w;
            }
            hxm = hx;
            hym = hy;
            ex0 = ex1;
            ey0 = ey1;
            ez0 = ez1;
            eym0 = eym1;
        }
    }
}

template <typename T>
void launch_step3(const Grid<T> &g, int nw, const TfList &tfH, const float *dh, const float *dp, hipStream_t s) {
    const int ntz = (g.nz + g.kz - 1) / g.kz;
    const int nunits = g.ny * ntz;    // one wavefront per (row, z block)
    const int ntf = dh ? tfH.n : 0;
    if (nw == 8) {
        const unsigned blocks = (unsigned)((nunits + 7) / 8);
        if (g.pml) hipLaunchKernelGGL((k_step3<T, true, 8>), dim3(blocks), dim3(512), 0, s, g, tfH.dev, ntf, dh, dp, nunits);
        else hipLaunchKernelGGL((k_step3<T, false, 8>), dim3(blocks), dim3(512), 0, s, g, tfH.dev, ntf, dh, dp, nunits);
    } else {
        const unsigned blocks = (unsigned)((nunits + 3) / 4);
        if (g.pml) hipLaunchKernelGGL((k_step3<T, true, 4>), dim3(blocks), dim3(256), 0, s, g, tfH.dev, ntf, dh, dp, nunits);
        else hipLaunchKernelGGL((k_step3<T, false, 4>), dim3(blocks), dim3(256), 0, s, g, tfH.dev, ntf, dh, dp, nunits);
    }
}

template void launch_step3<float>(const Grid<float> &, int, const TfList &, const float *, const float *, hipStream_t);
template void launch_step3<__half>(const Grid<__half> &, int, const TfList &, const float *, const float *,
                                   hipStream_t);

}  // namespace lx
