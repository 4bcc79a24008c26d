// lx_api.cpp -- C ABI of the MI355X Yee timestep (include/luxum.h).
//
// Device units.  The reference updates (maxwell_three_dim_pml.cc:182-255)
//     H -= Db*(c1/k1 - c2/k2) + Db*(psi1 - psi2),   Db = dt/mu0
//     E += Cb*(c1/k1 - c2/k2) + Cb*(psi1 - psi2),   Cb = dt/(eps0*eps_r)
// with c1 = dF/d_a (split curl) are stored here with Hs = eta0*H and the
// curl written as a plain difference dF (no 1/d_a):
//     Hs -= (dE1*kh_a1 - dE2*kh_a2) + (psiH1 - psiH2)
//     E  += cb[id]*((dHs1*ke_a1 - dHs2*ke_a2) + (psiE1 - psiE2))
// where ch = c*dt/dx (= eta0*Db/dx), kh_a = ch*(dx/d_a)/kappa_h,
// ke_a = (dx/d_a)/kappa_e, cb = c*dt/(dx*eps_r) (= Cb/(eta0*dx)), and the psi
// recursions psi = b*psi + c'*dF use c'_h = ch*c_h*(dx/d_a), c'_e = c_e*(dx/d_a).
// All of these are exact rewrites of the reference's algebra; they are
// evaluated in FP64 on the host and rounded once to FP32, so every device
// quantity is O(1) (FP16 storage needs that).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <new>
#include <vector>

#include "../../include/luxum.h"
#include "lx_internal.hpp"
#include "lx_util.hpp"
#include "lx_sim.hpp"

namespace lx {
char *error_buffer() {
    static thread_local char buf[512];
    return buf;
}
}  // namespace lx

using namespace lx;

#define HIP_TRY(x)                                                                                         \
    do {                                                                                                   \
        hipError_t _e = (x);                                                                               \
        if (_e != hipSuccess) return fail(LX_EDEVICE, "%s failed: %s", #x, hipGetErrorString(_e));          \
    } while (0)

namespace lx {

template <typename T>
Grid<T> make_grid(const lx_sim *s) {
    Grid<T> g{};
    g.nx = s->N[0];
    g.ny = s->N[1];
    g.nz = s->N[2];
    g.px = s->px;
    g.plane = s->plane;
    g.fstride = s->fstride;
    g.F = (T *)s->F;
    g.mat = s->mat;
    g.cb = s->cb;
    g.tab = s->tab;
    g.tn = s->tn;
    g.psi = s->psi;
    for (int a = 0; a < 3; ++a) {
        g.scells[a] = s->slab_cells[a];
        g.slo[a] = s->slo[a];
        g.shi[a] = s->shi[a];
        g.sw[a] = s->sw[a];
    }
    g.pml = s->pml;
    g.kz = s->kz;
    g.kzs = s->kzs;
    g.rows = s->rows;
    g.ntx = (g.nx + 255) / 256;
    if (s->planar) {
        g.nty = (g.nz + kBY - 1) / kBY;
        g.ntz = 1;
    } else {
        g.nty = (g.ny + kBY - 1) / kBY;
        g.ntz = (g.nz + g.kz - 1) / g.kz;
    }
    g.ntiles = g.ntx * g.nty * g.ntz;
    g.counter = s->counter;
    return g;
}

lx_status dalloc(lx_sim *s, void **out, size_t bytes) {
    void *p = nullptr;
    const size_t b = bytes ? (bytes + 255) / 256 * 256 : 256;
    {
        lx_ctx *c = s->ctx;
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->blocks.find(b);
        if (it != c->blocks.end()) {
            p = it->second;
            c->blocks.erase(it);
            c->cached -= (int64_t)b;
        }
    }
    if (!p) {
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            // blocks cached for other sizes may hold the memory: free them and retry once
            (void)hipGetLastError();
            std::multimap<size_t, void *> drop;
            {
                lx_ctx *c = s->ctx;
                std::lock_guard<std::mutex> lk(c->mu);
                drop.swap(c->blocks);
                c->cached = 0;
            }
            for (const auto &d : drop) (void)hipFree(d.second);
            e = drop.empty() ? e : hipMalloc(&p, b);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(LX_ENOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        }
    }
    s->allocs.emplace_back(p, b);
    s->bytes += (int64_t)b;
    *out = p;
    return LX_OK;
}

// Give a block back to the context's cache (or free it when the cache is
// full).  The caller has synchronised every stream that used the block.
static void release_block(lx_ctx *c, void *p, size_t b) {
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->cached + (int64_t)b <= c->cache_limit) {
            c->blocks.emplace(b, p);
            c->cached += (int64_t)b;
            return;
        }
    }
    (void)hipFree(p);
}

// ---- CPML profiles (maxwell.cc:997-1280), FP64 ---------------------------
// Same expressions as the reference; the piecewise form equals the
// reference's masked sums wherever those are finite.
struct Prof {
    std::vector<double> sig, kap, alp, b, c;
};

void build_profile(const lx_sim *s, int dim, int is_E, int n, Prof &pr) {
    const lx_params &p = s->p;
    const double delta = dim == 0 ? p.dx : dim == 1 ? p.dy : p.dz;
    const int N = s->N[dim];
    const double d = p.pml_thickness / delta, m = p.sigma_m, ma = p.a_m;
    pr.sig.assign(n, 0.0);
    pr.kap.assign(n, 1.0);
    pr.alp.assign(n, 0.0);
    pr.b.assign(n, 1.0);
    pr.c.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        const double x = is_E ? (double)i : i + 0.5;
        double sg = 0.0, ka = 1.0, al = 0.0;
        if (x <= (d - 1)) {
            sg = std::pow(((d - 1) - x) / (d - 1), m) * s->sigma_max;
            ka = std::pow(((d - 1) - x) / (d - 1), m) * (p.kappa_max - 1) + 1;
            al = std::pow((x / (d - 1)), ma) * p.a_max;
        }
        if (x >= (N - d)) {
            sg = std::pow((x - N + d) / (d - 1), m) * s->sigma_max;
            ka = std::pow((x - N + d) / (d - 1), m) * (p.kappa_max - 1) + 1;
            al = std::pow((N - 1 - x) / (d - 1), ma) * p.a_max;
        }
        pr.sig[i] = sg;
        pr.kap[i] = ka;
        pr.alp[i] = al;
        pr.b[i] = std::exp(-((sg / ka / s->eps0) - (al / s->eps0)) * s->dt);   // maxwell.cc:1215
        pr.c[i] = sg == 0 ? 0.0 : sg / (sg * ka + std::pow(ka, 2) * al) * (pr.b[i] - 1);   // :1271-1278
    }
}

lx_status build_tables(lx_sim *s) {
    const int axes = 3;
    // rows cover the padded x pitch (kernels read 4 x cells at a time) and are 16-byte aligned
    int Nmax = std::max(s->px, std::max(s->N[1], s->N[2]));
    Nmax = (Nmax + 3) / 4 * 4;
    std::vector<float> host((size_t)axes * 6 * Nmax, 0.f);
    const double d[3] = {s->p.dx, s->p.dy, s->p.dz};
    const double ch = s->c0 * s->dt / s->p.dx;
    for (int a = 0; a < 3; ++a) {
        const int N = s->N[a];
        Prof pe, ph;
        const bool on = s->pml && !(s->planar && a == 1);
        if (on) {
            build_profile(s, a, 1, N, pe);
            build_profile(s, a, 0, N - 1, ph);
        } else {
            pe.kap.assign(N, 1.0); pe.b.assign(N, 1.0); pe.c.assign(N, 0.0);
            ph.kap.assign(std::max(N - 1, 1), 1.0); ph.b.assign(std::max(N - 1, 1), 1.0); ph.c.assign(std::max(N - 1, 1), 0.0);
        }
        const double r = (s->planar && a == 1) ? 1.0 : s->p.dx / d[a];   // planar runs have no y derivative
        float *kh = &host[((size_t)a * 6 + 0) * Nmax], *bh = &host[((size_t)a * 6 + 1) * Nmax],
              *chh = &host[((size_t)a * 6 + 2) * Nmax], *ke = &host[((size_t)a * 6 + 3) * Nmax],
              *be = &host[((size_t)a * 6 + 4) * Nmax], *ce = &host[((size_t)a * 6 + 5) * Nmax];
        for (int i = 0; i < N; ++i) {
            const int ih = std::min(i, (int)ph.kap.size() - 1);   // the H table has N-1 entries; pad the last
            kh[i] = (float)(ch * r / ph.kap[ih]);
            bh[i] = (float)ph.b[ih];
            chh[i] = (float)(ch * ph.c[ih] * r);
            ke[i] = (float)(r / pe.kap[i]);
            be[i] = (float)pe.b[i];
            ce[i] = (float)(pe.c[i] * r);
        }
        // slab extents: every index with c != 0 (E or H node) lies in [0,lo) U [hi,N)
        int lo = 0, hi = N;
        if (on) {
            auto nz = [&](int i) { return pe.c[i] != 0.0 || (i < N - 1 && ph.c[i] != 0.0); };
            while (lo < N && nz(lo)) ++lo;
            while (hi > lo && nz(hi - 1)) --hi;
            for (int i = lo; i < hi; ++i)
                if (nz(i)) return fail(LX_EINVAL, "CPML profile along axis %d is not confined to the boundary slabs", a);
        }
        if (a == 0 && on) {
            // x slab: 4-aligned so that every 4-cell lane group is wholly in or out of it
            // (the extra cells have c = 0, so their psi stays zero)
            lo = (lo + 3) / 4 * 4;
            hi = std::max(lo, hi / 4 * 4);
            s->slo[a] = lo;
            s->shi[a] = hi;
            s->sw[a] = lo + (s->px - hi);
        } else {
            s->slo[a] = lo;
            s->shi[a] = hi;
            s->sw[a] = lo + (N - hi);
        }
        // widest index range around the centre, outside the slabs, whose
        // coefficients equal the centre's and carry no CPML term
        const int mid = N / 2;
        auto same = [&](int i) { return kh[i] == kh[mid] && ke[i] == ke[mid] && chh[i] == 0.f && ce[i] == 0.f; };
        int ul = mid, uh = mid;
        if (same(mid)) {
            uh = mid + 1;
            while (ul > 0 && same(ul - 1)) --ul;
            while (uh < N && same(uh)) ++uh;
        }
        s->ulo[a] = std::max(ul, s->slo[a]);
        s->uhi[a] = std::max(s->ulo[a], std::min(uh, s->shi[a]));
        if (a == 0) s->unit_ke = ke[mid] == 1.f;
        if (a == 2) s->unit_ke = s->unit_ke && ke[mid] == 1.f;
    }
    void *p = nullptr;
    lx_status st = dalloc(s, &p, host.size() * sizeof(float));
    if (st) return st;
    s->tab = (float *)p;
    s->tn = Nmax;
    HIP_TRY(hipMemcpy(s->tab, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice));
    // psi slabs: per axis [H slot0, H slot1, E slot0, E slot1], each w_a * (product of the other extents)
    // x slab: (k, j, slab x) ; y slab: (k, slab y, px) ; z slab: (slab z, j, px)
    size_t total = 0;
    s->slab_cells[0] = (int64_t)s->sw[0] * s->N[1] * s->N[2];
    s->slab_cells[1] = (int64_t)s->sw[1] * s->N[2] * s->px;
    s->slab_cells[2] = (int64_t)s->sw[2] * s->N[1] * s->px;
    for (int a = 0; a < 3; ++a) total += 4 * (size_t)s->slab_cells[a];
    st = dalloc(s, &p, std::max<size_t>(total, 1) * sizeof(float));
    if (st) return st;
    s->psi = (float *)p;
    launch_zero_bytes(s->psi, std::max<size_t>(total, 1) * sizeof(float), s->cs);
    return LX_OK;
}

// ---- TF/SF boxes (maxwell.cc:512-929) -----------------------------------
// For a face normal to axis a, the two split-curl entries it corrects:
// {target component, split curl (1|2), differentiated component}.
const int kHFace[3][2][3] = {
    {{LX_HZ, 1, LX_EY}, {LX_HY, 2, LX_EZ}},
    {{LX_HZ, 2, LX_EX}, {LX_HX, 1, LX_EZ}},
    {{LX_HY, 1, LX_EX}, {LX_HX, 2, LX_EY}},
};
const int kEFace[3][2][3] = {
    {{LX_EY, 2, LX_HZ}, {LX_EZ, 1, LX_HY}},
    {{LX_EX, 1, LX_HZ}, {LX_EZ, 2, LX_HX}},
    {{LX_EX, 2, LX_HY}, {LX_EY, 1, LX_HX}},
};

// plane_wave_initialise (maxwell.cc:825-929) with one rank: clip to [0, N-1].
void pw_clip(int lo, int hi, int N, int *st, int *en) {
    const int s0 = 0, e0 = N - 1;
    if (lo > e0) { *st = -1; *en = -1; }
    else if (lo >= s0) { *st = lo; *en = hi > e0 ? e0 : hi; }
    else if (hi > e0) { *st = s0; *en = e0; }
    else if (hi >= s0) { *st = s0; *en = hi; }
    else { *st = -1; *en = -1; }
}

lx_status build_tfsf(lx_sim *s) {
    s->tfH.n = s->tfE.n = 0;
    s->tfH.total = s->tfE.total = 0;
    s->tfH.dev = s->tfE.dev = nullptr;
    if (!s->p.use_pw) return LX_OK;
    const lx_params &p = s->p;
    const bool three = !s->planar;
    double Edir[3], Hdir[3];
    for (int a = 0; a < 3; ++a) Edir[a] = p.e_pw[a];
    const double *k = p.k_pw;
    double hx = k[1] * Edir[2] - k[2] * Edir[1], hy = k[2] * Edir[0] - k[0] * Edir[2], hz = k[0] * Edir[1] - k[1] * Edir[0];
    double ha = std::pow(std::pow(hx, 2) + std::pow(hy, 2) + std::pow(hz, 2), 0.5);   // maxwell.cc:95-102
    Hdir[0] = hx / ha; Hdir[1] = hy / ha; Hdir[2] = hz / ha;
    // E-update corrections in device units: (g*Hdir/(d_a*c*mu0)) * d_a * eta0
    const double hfac = s->eta0 / (s->c0 * s->mu0);
    int st[3], en[3];
    for (int a = 0; a < 3; ++a) {
        if (!three && a == 1) { st[a] = en[a] = -1; continue; }
        pw_clip(p.tf_lo[a], p.tf_hi[a], s->N[a], &st[a], &en[a]);
    }
    for (int half = 0; half < 2; ++half) {   // 0: H update (curl E), 1: E update (curl H)
        TfList &set = half == 0 ? s->tfH : s->tfE;
        for (int a = 0; a < 3; ++a) {
            if (!three && a == 1) continue;
            const int b0 = (a + 1) % 3, b1 = (a + 2) % 3;
            bool trans_ok = true;
            for (int b : {b0, b1})
                if ((three || b != 1) && st[b] == -1) trans_ok = false;
            if (!trans_ok) continue;
            for (int side = 0; side < 2; ++side) {
                int idx;
                double coord;
                if (half == 0) {
                    idx = side == 0 ? p.tf_lo[a] - 1 : p.tf_hi[a];
                    coord = side == 0 ? (double)p.tf_lo[a] : (double)p.tf_hi[a];
                } else {
                    idx = side == 0 ? p.tf_lo[a] : p.tf_hi[a];
                    coord = side == 0 ? p.tf_lo[a] - 0.5 : p.tf_hi[a] + 0.5;
                }
                if (idx < 0 || idx > s->N[a] - 1) continue;
                for (int e = 0; e < 2; ++e) {
                    const int *fd = half == 0 ? kHFace[a][e] : kEFace[a][e];
                    const int tgt = fd[0], which = fd[1], src = fd[2];
                    if (!three && !(tgt == LX_EY || tgt == LX_HX || tgt == LX_HZ)) continue;
                    if (!three && !(src == LX_EY || src == LX_HX || src == LX_HZ)) continue;
                    double sgn = side == 0 ? -1.0 : 1.0;   // low face subtracts, high face adds
                    double dir;
                    if (half == 0) {
                        if (!three) sgn = -sgn;            // 2D injects the negated wave (maxwell.cc:523)
                        dir = sgn * Edir[src];
                    } else {
                        dir = sgn * Hdir[src - 3] * hfac;
                        if (!three) dir = -dir;            // explicit (-1) of maxwell.cc:563-588
                    }
                    if (dir == 0.0) continue;
                    if (set.n >= kMaxTf) return fail(LX_EINVAL, "too many TF/SF boxes");
                    TfBox &bx = set.b[set.n++];
                    bx.comp = tgt % 3;
                    bx.curl = which - 1;
                    bx.fa = a;
                    bx.dir = dir * s->sscale;
                    bx.lo[a] = bx.hi[a] = idx;
                    bx.off[a] = (float)(coord - idx);
                    for (int b : {b0, b1}) {
                        if (!three && b == 1) { bx.lo[b] = bx.hi[b] = 0; bx.off[b] = 0.f; continue; }
                        const int sg = kStag[src][b];
                        bx.lo[b] = st[b];
                        bx.hi[b] = en[b] - sg;
                        bx.off[b] = sg ? 0.5f : 0.f;
                    }
                }
            }
        }
    }
    for (TfList *L : {&s->tfH, &s->tfE}) {
        int start = 0;
        for (int q = 0; q < L->n; ++q) {
            TfBox &bx = L->b[q];
            bx.start = start;
            int cnt = 1;
            for (int a = 0; a < 3; ++a) cnt *= std::max(0, bx.hi[a] - bx.lo[a] + 1);
            start += cnt;
        }
        L->total = start;
        void *q = nullptr;
        lx_status e = dalloc(s, &q, sizeof(TfBox) * kMaxTf);
        if (e) return e;
        HIP_TRY(hipMemcpy(q, L->b, sizeof(TfBox) * kMaxTf, hipMemcpyHostToDevice));
        L->dev = (const TfBox *)q;
    }
    return LX_OK;
}

double waveform_host(int kind, double t, double c, double lambda0, double fwhm, double delay) {
    const double f0 = c / lambda0, T0 = 1.0 / f0;
    const double tau = fwhm * T0 / (2.0 * std::sqrt(std::log(2.0)));
    const double t0 = delay * tau;
    const double x = (t - t0) / tau;
    const double env = std::exp(-x * x);
    if (kind == LX_WAVE_GAUSSIAN) return env;
    return env * std::sin(2.0 * M_PI * f0 * (t - t0));
}

lx_status build_incident(lx_sim *s) {
    const lx_params &p = s->p;
    IncParams &in = s->inc;
    std::memset(&in, 0, sizeof(in));
    in.mode = p.incident;
    in.planar = s->planar;
    in.dx = p.dx; in.dy = p.dy; in.dz = p.dz;
    in.kx = p.k_pw[0]; in.ky = p.k_pw[1]; in.kz = p.k_pw[2];
    in.init_dx = p.off_pw * p.dx;
    in.cdt = s->c0 * s->dt;
    in.inv_sigma = 1.0 / p.pw_sigma;
    if (!(p.use_pw && p.incident == LX_INC_AUX1D)) return LX_OK;
    int ax = 0;
    for (int a = 0; a < 3; ++a)
        if (p.k_pw[a] != 0.0) ax = a;
    const double d[3] = {p.dx, p.dy, p.dz};
    AuxLine &A = s->aux;
    A.S = s->c0 * s->dt / d[ax];
    const int x0 = p.tf_lo[ax] - 2;
    const int T = p.num_timesteps > 0 ? p.num_timesteps : 1;
    A.L = (p.tf_hi[ax] + 2 - x0) + (int)std::ceil(A.S * T / 2.0) + 8;
    A.wave = p.pw_wave;
    A.c = s->c0;
    A.dt = s->dt;
    A.lambda0 = p.pw_lambda0;
    A.fwhm = p.pw_fwhm;
    A.delay = p.pw_delay;
    void *q = nullptr;
    lx_status st = dalloc(s, &q, sizeof(double) * 4 * (size_t)A.L);   // two sets of (e, h)
    if (st) return st;
    A.e = (double *)q;
    A.h = A.e + A.L;
    A.e2 = A.e + 2 * (size_t)A.L;
    A.h2 = A.e2 + A.L;
    std::vector<double> init(2 * (size_t)A.L, 0.0);
    init[0] = waveform_host(A.wave, 0.0, A.c, A.lambda0, A.fwhm, A.delay);
    HIP_TRY(hipMemcpy(A.e, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice));
    in.axis = ax;
    in.x0 = x0;
    in.L = A.L;
    in.e = A.e;
    in.h = A.h;
    s->use_aux = 1;
    return LX_OK;
}

void fill_cb_vacuum(lx_sim *s) {
    const double v = s->c0 * s->dt / s->p.dx;
    for (int q = 0; q < 3 * 256; ++q) s->cb_host[q] = (float)v;
}

ProbeArgs probe_args(const lx_sim *s) {
    ProbeArgs pr{};
    pr.n = s->np;
    pr.comp = s->pcomp;
    pr.idx = s->pidx;
    pr.rec = s->prec;
    pr.cap = s->pcap;
    pr.hscale = 1.0 / s->eta0;
    return pr;
}

// The fused step's two launches: the interior box (part 0) and the CPML
// shell (part 1).  Both read one buffer set and write disjoint cells of the
// other, so the shell may run on the second stream ks at the same time, its
// blocks filling the interior launch's tail.
template <typename T>
void launch_fused_step(lx_sim *s, const Grid<T> &g) {
    if (s->ks && s->kconc) {
        (void)hipEventRecord(s->kfork, s->cs);
        (void)hipStreamWaitEvent(s->ks, s->kfork, 0);
        launch_step3(g, s->nw, s->vec, 1, s->ks);
        launch_step3(g, s->nw, s->vec, 0, s->cs);
        (void)hipEventRecord(s->kjoin, s->ks);
        (void)hipStreamWaitEvent(s->cs, s->kjoin, 0);
    } else {
        launch_step3(g, s->nw, s->vec, 0, s->cs);
        launch_step3(g, s->nw, s->vec, 1, s->cs);
    }
}

// Can a fused sequence of this simulation run one TF/SF launch between
// steps?  Not when an H probe lies on an H face cell (that launch records the
// H probes of a step while it applies the next step's H corrections).
bool tf_merge_ok(const lx_sim *s) {
    if (!s->tf_merge) return false;
    for (int q = 0; q < s->np; ++q) {
        const int *c = &s->probe_host[4 * (size_t)q];
        if (c[0] < 3) continue;
        const int ijk[3] = {c[1], c[2], c[3]};
        for (int b = 0; b < s->tfH.n; ++b) {
            const TfBox &x = s->tfH.b[b];
            if (x.comp == c[0] - 3 && ijk[0] >= x.lo[0] && ijk[0] <= x.hi[0] && ijk[1] >= x.lo[1] &&
                ijk[1] <= x.hi[1] && ijk[2] >= x.lo[2] && ijk[2] <= x.hi[2])
                return false;
        }
    }
    return true;
}

// Can the CPML shell launch of step q+1 start before the TF/SF launch that
// follows step q's interior launch?  The shell reads and writes only shell
// cells and the interior box's cells next to them (the rows and planes it
// recomputes); the launch between steps writes TF/SF face cells and reads the
// probe cells, and the point source writes its cell.  When all of these lie
// at least two cells inside the interior box, the two touch disjoint cells,
// and the shell's only orders are the interior launch before it (whose output
// it reads) and the interior launch after it (which reads its output).
// LX_SHELL_PIPE=0 keeps the shell launch between the same events as the
// interior launch.
bool shell_pipe_ok(const lx_sim *s) {
    if (!s->ks || !s->kconc || !s->pml || !s->shell_pipe) return false;
    UnitMap um[2];
    int nu[2];
    step3_unit_maps(s->N[0], s->N[1], s->N[2], s->kz, s->kzs, s->pml, s->slo, s->shi, s->vec, um, nu);
    const int jlo = um[0].jlo + 2, jhi = um[0].jhi - 3;
    const int zlo = um[0].tlo * um[0].kz + 2, zhi = std::min(um[0].thi * um[0].kz, s->N[2]) - 3;
    auto inside = [&](int j, int k) { return j >= jlo && j <= jhi && k >= zlo && k <= zhi; };
    for (const TfList *L : {&s->tfE, &s->tfH})
        for (int b = 0; b < L->n; ++b)
            if (!inside(L->b[b].lo[1], L->b[b].lo[2]) || !inside(L->b[b].hi[1], L->b[b].hi[2])) return false;
    for (int q = 0; q < s->np; ++q)
        if (!inside(s->probe_host[4 * (size_t)q + 2], s->probe_host[4 * (size_t)q + 3])) return false;
    if (s->p.use_ps && !inside(s->p.src_pos[1], s->p.src_pos[2])) return false;
    return true;
}

// One launch sequence of n timesteps: per step [point source] H launch,
// E launch; then the sequence-end kernel.  Kernels read the step index as
// counter[0] + toff, so the same sequence can be captured once and replayed.
// Fused 3D: per step [point source on the current buffer], the fused step
// (current -> other buffer), the E TF/SF corrections on the other buffer;
// the buffers alternate, so the state sits in F_alt after an odd count.
// Merged TF/SF (tf_merge_ok): the H corrections precede the first step
// only; each later step's ride in the launch between steps
// (k_tfsf_between), whose aux line alternates between its two sets and
// ends the sequence in the first.
template <typename T>
void enqueue_seq_t(lx_sim *s, int n, std::vector<hipEvent_t> *ev) {
    Grid<T> g = make_grid<T>(s);
    const ProbeArgs pr = probe_args(s);
    const bool corr = s->tfH.n > 0 || s->use_aux || s->np > 0;
    if (s->fused) {
        T *buf[2] = {(T *)s->F, (T *)s->F_alt};
        float *pbuf[2] = {s->psi, s->psi_alt};
        const bool merge = corr && !ev && tf_merge_ok(s);
        // pipelined shell: the shell of step q waits for the interior launch
        // of step q-1 (event kfork) and the interior launch of step q for the
        // shell of step q-1 (event kjoin), so a shell overlaps the interior
        // launch of its own step and the TF/SF launch before it
        const bool pipe = merge && shell_pipe_ok(s);
        if (pipe) {
            (void)hipEventRecord(s->kfork, s->cs);
            (void)hipStreamWaitEvent(s->ks, s->kfork, 0);
        }
        AuxLine aset[2] = {s->aux, s->aux};
        aset[1].e = s->aux.e2;
        aset[1].h = s->aux.h2;
        for (int q = 0; q < n; ++q) {
            Launch l{q, s->use_aux, s->np > 0, s->np > 0 && !s->p.use_ps};
            Grid<T> gq = g;
            gq.F = buf[q & 1];
            gq.psi = pbuf[q & 1];
            gq.Fn = buf[(q + 1) & 1];
            gq.psin = pbuf[(q + 1) & 1];
            if (s->p.use_ps) launch_point_source(gq, s->ps, pr, q, s->cs);
            if (corr && (!merge || q == 0)) launch_tfsf_hpre(gq, s->tfH, s->inc, s->aux, pr, l, s->cs);
            // timed parts: [0,1] interior box, [2,3] CPML shell
            if (pipe) {
                if (q >= 1) {
                    (void)hipStreamWaitEvent(s->cs, s->kjoin, 0);   // shell of step q-1
                    (void)hipStreamWaitEvent(s->ks, s->kfork, 0);   // interior of step q-1
                }
                launch_step3(gq, s->nw, s->vec, 1, s->ks);
                (void)hipEventRecord(s->kjoin, s->ks);
                launch_step3(gq, s->nw, s->vec, 0, s->cs);
                (void)hipEventRecord(s->kfork, s->cs);
            } else if (!ev) {
                launch_fused_step(s, gq);
            } else {
                if (ev) (void)hipEventRecord((*ev)[4 * q + 0], s->cs);
                launch_step3(gq, s->nw, s->vec, 0, s->cs);
                if (ev) (void)hipEventRecord((*ev)[4 * q + 1], s->cs);
                if (ev) (void)hipEventRecord((*ev)[4 * q + 2], s->cs);
                launch_step3(gq, s->nw, s->vec, 1, s->cs);
                if (ev) (void)hipEventRecord((*ev)[4 * q + 3], s->cs);
            }
            Grid<T> gn = g;
            gn.F = gq.Fn;
            gn.psi = gq.psin;
            if (merge) {
                const int last = q == n - 1;
                const Launch lb{q, s->use_aux, s->np > 0, s->np > 0 && !s->p.use_ps && q >= 1};
                launch_tfsf_between(gn, (const T *)gq.F, s->tfE, s->tfH, s->inc, aset[q & 1], aset[last ? 0 : (q + 1) & 1],
                                    pr, lb, last, s->cs);
            } else if (corr) {
                launch_tfsf(gn, s->tfE, s->inc, s->aux, pr, l, 1, s->cs);
            }
        }
        if (pipe) (void)hipStreamWaitEvent(s->cs, s->kjoin, 0);   // the last shell
        Grid<T> gf = g;
        gf.F = buf[n & 1];
        launch_finish(gf, pr, n, s->cs);
        return;
    }
    for (int q = 0; q < n; ++q) {
        Launch l{q, s->use_aux, s->np > 0, s->np > 0 && !s->p.use_ps};
        if (s->p.use_ps) launch_point_source(g, s->ps, pr, q, s->cs);
        if (ev) (void)hipEventRecord((*ev)[4 * q + 0], s->cs);
        launch_bulk(g, s->planar, 0, s->cs);
        if (ev) (void)hipEventRecord((*ev)[4 * q + 1], s->cs);
        if (corr) launch_tfsf(g, s->tfH, s->inc, s->aux, pr, l, 0, s->cs);
        if (ev) (void)hipEventRecord((*ev)[4 * q + 2], s->cs);
        launch_bulk(g, s->planar, 1, s->cs);
        if (ev) (void)hipEventRecord((*ev)[4 * q + 3], s->cs);
        if (corr) launch_tfsf(g, s->tfE, s->inc, s->aux, pr, l, 1, s->cs);
    }
    launch_finish(g, pr, n, s->cs);
}

// after an enqueued sequence of n fused steps: the current state is in the
// other buffer when n is odd (the graphs of both buffer states stay valid)
void settle_parity(lx_sim *s, int n) {
    if (!s->fused || !(n & 1)) return;
    std::swap(s->F, s->F_alt);
    std::swap(s->psi, s->psi_alt);
    s->gen++;
}

void enqueue_seq(lx_sim *s, int n, std::vector<hipEvent_t> *ev = nullptr) {
    if (n <= 0) return;
    if (s->fp16) enqueue_seq_t<__half>(s, n, ev);
    else enqueue_seq_t<float>(s, n, ev);
}

// Captured step sequences of 32, 16, 8, 4, 2 and 1 steps; a call of n steps
// replays the largest that fit (a one-step call, as a per-step time loop
// makes, replays the one-step graph; an odd count leaves the fused step's
// state in the other buffer set, as the eager sequence does)
constexpr int kGraphLens = 6;
constexpr int kGraphSteps[kGraphLens] = {32, 16, 8, 4, 2, 1};

void drop_graph(lx_sim *s) {
    for (int i = 0; i < kGraphLens; ++i)
        for (int q = 0; q < 2; ++q) {
            if (s->gexec[i][q]) (void)hipGraphExecDestroy(s->gexec[i][q]);
            s->gexec[i][q] = nullptr;
            s->gF[i][q] = nullptr;
        }
}

// The graph of kGraphSteps[i] steps for the current buffer state, captured
// on first use; nullptr when graphs are off or capture failed.
hipGraphExec_t ensure_graph(lx_sim *s, int i, lx_status *st) {
    *st = LX_OK;
    if (!s->graphs) return nullptr;
    int slot = -1;
    for (int q = 0; q < 2; ++q) {
        if (s->gexec[i][q] && s->gF[i][q] == s->F) return s->gexec[i][q];
        if (!s->gexec[i][q] && slot < 0) slot = q;
    }
    if (slot < 0) {   // both slots hold other buffer states: the buffers changed
        drop_graph(s);
        slot = 0;
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(s->cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        *st = fail(LX_EDEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(e));
        return nullptr;
    }
    enqueue_seq(s, kGraphSteps[i]);
    e = hipStreamEndCapture(s->cs, &graph);
    if (e != hipSuccess || !graph) {
        (void)hipGetLastError();
        s->graphs = 0;
        return nullptr;
    }
    hipGraphExec_t gx = nullptr;
    e = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        s->graphs = 0;
        return nullptr;
    }
    s->gexec[i][slot] = gx;
    s->gF[i][slot] = s->F;
    return gx;
}

// Capture every step-sequence graph for both buffer states of the fused
// step, so that no later lx_sim_step call captures (after autotuning: the
// configuration is final).
lx_status prepare_graphs(lx_sim *s) {
    if (!s->graphs) return LX_OK;
    for (int side = 0; side < (s->fused ? 2 : 1); ++side) {
        for (int i = 0; i < kGraphLens; ++i) {
            lx_status st;
            ensure_graph(s, i, &st);
            if (st) return st;
        }
        if (s->fused) {   // the other buffer state
            std::swap(s->F, s->F_alt);
            std::swap(s->psi, s->psi_alt);
        }
    }
    return LX_OK;
}

lx_status ensure_probe_cap(lx_sim *s, int need) {
    if (!s->np || need <= s->pcap) return LX_OK;
    int cap = std::max(need, std::max(2 * s->pcap, s->p.num_timesteps));
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(double) * (size_t)cap * s->np);
    if (e != hipSuccess) return fail(LX_ENOMEM, "probe buffer allocation failed");
    HIP_TRY(hipStreamSynchronize(s->cs));
    if (s->prec) {
        HIP_TRY(hipMemcpy(q, s->prec, sizeof(double) * (size_t)s->pcap * s->np, hipMemcpyDeviceToDevice));
        (void)hipFree(s->prec);
        s->bytes -= (int64_t)sizeof(double) * s->pcap * s->np;
    }
    s->prec = (double *)q;
    s->pcap = cap;
    s->bytes += (int64_t)sizeof(double) * cap * s->np;
    s->gen++;
    drop_graph(s);
    return LX_OK;
}

lx_status check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(LX_EDEVICE, "kernel launch failed: %s", hipGetErrorString(e));
    return LX_OK;
}

size_t psi_total(const lx_sim *s) {
    size_t total = 0;
    for (int a = 0; a < 3; ++a) total += 4 * (size_t)s->slab_cells[a];
    return std::max<size_t>(total, 1);
}

// The fused 3D step applies the H TF/SF corrections to the read buffer
// before the step, which leaves out the CPML psi term: it needs every H face
// outside the slabs along its face axis.
// x cells per lane of the fused step that the row pitch (and, for 4, the
// storage type) allows
bool vec_ok(const lx_sim *s, int vec) {
    if (vec == 4 && !s->fp16) return false;
    return (s->N[0] + 64 * vec - 1) / (64 * vec) * (64 * vec) <= s->px;
}

bool fused_ok(const lx_sim *s) {
    if (s->planar) return false;
    for (int b = 0; b < s->tfH.n; ++b) {
        const TfBox &bx = s->tfH.b[b];
        const int a = bx.fa, c = bx.lo[a];
        if (s->pml && (c < s->slo[a] || c >= s->shi[a])) return false;
    }
    return true;
}

// Second field/psi buffer set of the fused 3D step.
lx_status ensure_fused_buffers(lx_sim *s) {
    if (s->F_alt) return LX_OK;
    const size_t fbytes = (size_t)s->fstride * 6 * s->esize;
    lx_status st = dalloc(s, &s->F_alt, fbytes);
    if (st) return st;
    launch_zero_bytes(s->F_alt, fbytes, s->cs);
    void *pp = nullptr;
    if ((st = dalloc(s, &pp, psi_total(s) * sizeof(float)))) return st;
    launch_zero_bytes(pp, psi_total(s) * sizeof(float), s->cs);
    s->psi_alt = (float *)pp;
    s->gen++;
    HIP_TRY(hipStreamSynchronize(s->cs));
    return LX_OK;
}

static void dfree(lx_sim *s, void *p) {
    if (!p) return;
    auto it = std::find_if(s->allocs.begin(), s->allocs.end(), [&](const auto &a) { return a.first == p; });
    if (it == s->allocs.end()) return;
    const size_t bytes = it->second;
    s->allocs.erase(it);
    s->bytes -= (int64_t)bytes;
    release_block(s->ctx, p, bytes);
}

void release_fused_buffers(lx_sim *s) {
    (void)hipStreamSynchronize(s->cs);
    dfree(s, s->F_alt);
    dfree(s, s->psi_alt);
    s->F_alt = nullptr;
    s->psi_alt = nullptr;
    s->fused = 0;
    s->gen++;
}

template Grid<float> make_grid<float>(const lx_sim *s);
template Grid<__half> make_grid<__half>(const lx_sim *s);

}  // namespace lx

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *lx_last_error(void) { return lx::error_buffer(); }
const char *lx_version(void) { return "luxum-mi355x 0.1 (gfx950)"; }

}  // extern "C"

// Streams of destroyed simulations go back to their context for reuse
// (creating a stream costs milliseconds; a server creates two per request).
static lx_status take_stream(lx_ctx *c, hipStream_t *out) {
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (!c->streams.empty()) {
            *out = c->streams.back();
            c->streams.pop_back();
            return LX_OK;
        }
    }
    const hipError_t he = hipStreamCreateWithFlags(out, hipStreamNonBlocking);
    if (he != hipSuccess) {
        *out = nullptr;
        return fail(LX_EDEVICE, "stream: %s", hipGetErrorString(he));
    }
    return LX_OK;
}

static void give_stream(lx_ctx *c, hipStream_t st);

namespace lx {
lx_status ensure_side_stream(lx_sim *s) { return s->ss ? LX_OK : take_stream(s->ctx, &s->ss); }

void member_enter(lx_sim *s) {
    if (s->batch_wait) {
        (void)hipStreamWaitEvent(s->cs, s->batch_wait, 0);
        s->batch_wait = nullptr;
    }
    if (s->mguard) {   // a started monitor still reads the fields
        (void)hipStreamWaitEvent(s->cs, s->mdone, 0);
        s->mguard = nullptr;
    }
    s->cs_dirty = true;
}

void prewarm_streams(lx_ctx *c, int n) {
    std::lock_guard<std::mutex> lk(c->mu);
    while ((int)c->streams.size() < n) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        c->streams.push_back(st);
    }
}
}  // namespace lx

static void give_stream(lx_ctx *c, hipStream_t st) {
    if (!st) return;
    if (!c) {
        (void)hipStreamDestroy(st);
        return;
    }
    (void)hipStreamSynchronize(st);
    std::lock_guard<std::mutex> lk(c->mu);
    c->streams.push_back(st);
}

extern "C" {

lx_status lx_ctx_create(int32_t device, lx_ctx **out) {
    if (!out) return fail(LX_EINVAL, "lx_ctx_create: null pointer");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LX_EINVAL, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    lx_ctx *c = new (std::nothrow) lx_ctx;
    if (!c) return fail(LX_ENOMEM, "out of host memory");
    c->device = device;
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    c->total_mem = (int64_t)tot;
    c->cache_limit = std::min<int64_t>(16LL << 30, c->total_mem / 16);
    lx_status st = noalloc_init(device);   // once per device: the address of its readback scratch
    if (st) {
        delete c;
        return st;
    }
    *out = c;
    return LX_OK;
}

void lx_ctx_destroy(lx_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (hipStream_t st : ctx->streams) (void)hipStreamDestroy(st);
    for (const auto &b : ctx->blocks) (void)hipFree(b.second);
    delete ctx;
}

lx_status lx_ctx_mem_info(lx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes) {
    if (!ctx) return fail(LX_EINVAL, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    if (free_bytes) *free_bytes = (int64_t)fr;
    if (total_bytes) *total_bytes = (int64_t)tot;
    return LX_OK;
}

void lx_sim_destroy(lx_sim *s) {
    if (!s) return;
    if (s->batch) (void)lx_batch_remove(s->batch, s);
    (void)hipSetDevice(s->ctx->device);
    if (s->cs) (void)hipStreamSynchronize(s->cs);
    if (s->ss) (void)hipStreamSynchronize(s->ss);
    if (s->ks) (void)hipStreamSynchronize(s->ks);
    for (lx::DmaSlot &d : s->snap_dma) lx::dma_release(&d);
    if (s->ms) (void)hipStreamSynchronize(s->ms);
    drop_graph(s);
    for (const auto &a : s->allocs) release_block(s->ctx, a.first, a.second);
    if (s->prec) (void)hipFree(s->prec);
    for (int q = 0; q < 2; ++q) {
        if (s->stage[q]) (void)hipFree(s->stage[q]);
        if (s->ready[q]) (void)hipEventDestroy(s->ready[q]);
        if (s->done[q]) (void)hipEventDestroy(s->done[q]);
    }
    if (s->rd_stage) (void)hipFree(s->rd_stage);
    give_stream(s->ctx, s->cs);
    give_stream(s->ctx, s->ss);
    give_stream(s->ctx, s->ks);
    give_stream(s->ctx, s->ms);
    if (s->mstart) (void)hipEventDestroy(s->mstart);
    if (s->mdone) (void)hipEventDestroy(s->mdone);
    if (s->mres) (void)hipHostFree(s->mres);
    if (s->kfork) (void)hipEventDestroy(s->kfork);
    if (s->kjoin) (void)hipEventDestroy(s->kjoin);
    delete s;
}

lx_status lx_sim_create(lx_ctx *ctx, const lx_params *p, lx_sim **out) {
    if (!ctx || !p || !out) return fail(LX_EINVAL, "lx_sim_create: null pointer");
    lx_status st = lx_params_check(p);
    if (st) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    lx_sim *s = new (std::nothrow) lx_sim;
    if (!s) return fail(LX_ENOMEM, "out of host memory");
    s->ctx = ctx;
    s->p = *p;
    s->planar = p->is_planar;
    s->fp16 = p->precision == LX_FP16;
    s->sscale = p->storage_scale;
    if (s->sscale == 0.0) {
        // automatic: FP16 storage keeps the largest source amplitude near 1
        double amp = 0;
        if (p->use_pw) amp = std::max(amp, std::sqrt(p->e_pw[0] * p->e_pw[0] + p->e_pw[1] * p->e_pw[1] +
                                                     p->e_pw[2] * p->e_pw[2]));
        if (p->use_ps)
            amp = std::max(amp, std::fabs(p->src_kind == LX_SRC_HARD_CW ? (double)(int)p->src_amp : p->src_amp));
        s->sscale = 1.0;
        if (s->fp16 && amp > 0 && std::isfinite(amp))
            s->sscale = std::ldexp(1.0, std::max(-60, std::min(60, -(int)std::lround(std::log2(amp)))));
    }
    s->pml = p->use_pml;
    s->N[0] = p->nx; s->N[1] = p->is_planar ? 1 : p->ny; s->N[2] = p->nz;
    s->px = (s->N[0] + 63) / 64 * 64;
    s->plane = (long long)s->px * s->N[1];
    s->esize = s->fp16 ? 2 : 4;
    const char *ng = std::getenv("LX_NO_GRAPH");
    s->graphs = !(ng && *ng && *ng != '0');
    const char *mg = std::getenv("LX_TFSF_MERGE");   // 0: two TF/SF launches per fused step
    s->tf_merge = !(mg && *mg == '0');
    const char *sp = std::getenv("LX_SHELL_PIPE");   // 0: the shell launch between the interior launch's events
    s->shell_pipe = !(sp && *sp == '0');
    // constants exactly as maxwell.cc:63-84
    s->eps0 = 8.85 * std::pow(10, -12);
    s->mu0 = M_PI * 4 * std::pow(10, -7);
    s->c0 = 1 / std::pow(s->eps0 * s->mu0, 0.5);
    s->dt = p->courant * p->dx / s->c0;
    s->eta0 = std::sqrt(s->mu0 / s->eps0);
    if (p->use_pml) {
        double sopt = 0.8 * (p->sigma_m + 1) / p->dx / std::pow(s->mu0 / s->eps0, 0.5);
        s->sigma_max = sopt * p->sigma_ratio;
    }
    auto bail = [&](lx_status e) { lx_sim_destroy(s); return e; };
    if ((st = take_stream(ctx, &s->cs))) return bail(st);   // the side stream is taken on first use
    hipError_t he = hipSuccess;
    for (int q = 0; q < 2; ++q) {
        if (hipEventCreateWithFlags(&s->ready[q], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->done[q], hipEventDisableTiming) != hipSuccess)
            return bail(fail(LX_EDEVICE, "event creation failed"));
    }
    const size_t ncell = (size_t)s->plane * s->N[2];
    s->fstride = (long long)((ncell + 127) / 128 * 128);   // 256-byte aligned components
    const size_t fbytes = (size_t)s->fstride * 6 * s->esize;
    if ((st = dalloc(s, &s->F, fbytes))) return bail(st);
    launch_zero_bytes(s->F, fbytes, s->cs);
    void *q = nullptr;
    if ((st = dalloc(s, &q, ncell))) return bail(st);
    s->mat = (uint8_t *)q;
    launch_fill_u8(s->mat, 0, ncell, s->cs);
    if ((st = dalloc(s, &q, sizeof(s->cb_host)))) return bail(st);
    s->cb = (float *)q;
    fill_cb_vacuum(s);
    he = hipMemcpy(s->cb, s->cb_host, sizeof(s->cb_host), hipMemcpyHostToDevice);
    if (he != hipSuccess) return bail(fail(LX_EDEVICE, "copy: %s", hipGetErrorString(he)));
    if ((st = dalloc(s, &q, 64))) return bail(st);
    s->counter = (int *)q;
    launch_zero_bytes(s->counter, 64, s->cs);
    if ((st = dalloc(s, &q, 2 * sizeof(double) * (4 + 3 * (size_t)kMonBlocks)))) return bail(st);
    s->mon = (double *)q;
    if ((st = build_tables(s))) return bail(st);
    if ((st = build_tfsf(s))) return bail(st);
    if ((st = build_incident(s))) return bail(st);
    if (!s->planar) {
        // 3D: the fused step where it applies and its second buffer set fits
        // (faster than the half-step pair on every grid measured), else the
        // half-step pair.  LX_FUSED=0 selects the half-step pair, 1 the fused
        // step with one cell per lane, 2 with two and 4 with four (FP16
        // storage) where the row pitch allows it; unset: the widest of four
        // and two cells per lane for FP16 storage, one for FP32.
        // lx_sim_autotune times them all and keeps the fastest.
        const char *fz = std::getenv("LX_FUSED");
        const char f = fz && *fz ? *fz : 'd';
        if (f != '0' && fused_ok(s)) {
            st = ensure_fused_buffers(s);
            if (st == LX_OK) {
                s->fused = 1;
                s->kz = 8;
                s->nw = 4;
                s->vec = 1;
                // rows per block of the interior launch sharing neighbours through LDS: 7 for
                // FP32 storage (faster at every depth measured), 0 (one unit per wavefront) for
                // FP16, whose four-cells units hold too many registers for 8-wavefront blocks;
                // LX_ROWS overrides (3, 7 or 0)
                s->rows = s->fp16 ? 0 : 7;
                if (const char *rw = std::getenv("LX_ROWS"); rw && *rw) {
                    const int r = std::atoi(rw);
                    s->rows = (r == 3 || r == 7) ? r : 0;
                }
                if ((f == '4' || (f == 'd' && s->fp16)) && vec_ok(s, 4)) s->vec = 4;
                else if ((f == '2' || (f == 'd' && s->fp16)) && vec_ok(s, 2)) s->vec = 2;
                const char *sk = std::getenv("LX_SHELL_KZ");   // z-block depth of the CPML shell launch
                if (sk && *sk) s->kzs = std::max(0, std::atoi(sk));
                // LX_SHELL_STREAM=0 keeps the shell launch behind the interior launch on one stream
                const char *ks = std::getenv("LX_SHELL_STREAM");
                if (s->pml && !(ks && *ks == '0')) {
                    if ((st = take_stream(ctx, &s->ks))) return bail(st);
                    if (hipEventCreateWithFlags(&s->kfork, hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&s->kjoin, hipEventDisableTiming) != hipSuccess)
                        return bail(fail(LX_EDEVICE, "event creation failed"));
                }
            } else if (st == LX_ENOMEM && f == 'd') {
                (void)hipGetLastError();
                release_fused_buffers(s);
            } else {
                return bail(st);
            }
        }
    }
    if (p->use_ps) {
        PointSrc &ps = s->ps;
        ps.kind = p->src_kind;
        ps.comp = p->src_comp;
        ps.wave = p->src_wave;
        const int j = s->planar ? 0 : p->src_pos[1];
        ps.idx = (long long)p->src_pos[2] * s->plane + (long long)j * s->px + p->src_pos[0];
        if (p->src_kind == LX_SRC_HARD_CW) {
            ps.amp = (double)(int)p->src_amp;   // int amp_source (maxwell.h:203)
            ps.omega = (int)p->src_omega;       // int omega_source (maxwell.h:206)
        } else {
            ps.amp = p->src_amp;
            ps.omega = 0;
        }
        ps.amp *= s->sscale;
        ps.dt = s->dt;
        ps.c = s->c0;
        ps.lambda0 = p->src_lambda0;
        ps.fwhm = p->src_fwhm;
        ps.delay = p->src_delay;
    }
    if ((st = check_launch())) return bail(st);
    he = hipStreamSynchronize(s->cs);
    if (he != hipSuccess) return bail(fail(LX_EDEVICE, "set-up failed: %s", hipGetErrorString(he)));
    *out = s;
    return LX_OK;
}

}  // extern "C"

// Material ids plus per-id permittivities of the three E components
// (eps3[3*id + c]); the isotropic and the per-component entry points share it.
static lx_status set_materials(lx_sim *s, const uint8_t *ids, const double *eps3, int32_t n_ids, const char *fn) {
    if (!s || !ids || !eps3) return fail(LX_EINVAL, "%s: null pointer", fn);
    if (n_ids < 1 || n_ids > 256) return fail(LX_EINVAL, "%s: n_ids must be in [1, 256]", fn);
    for (int q = 0; q < 3 * n_ids; ++q)
        if (!(eps3[q] > 0)) return fail(LX_EINVAL, "%s: eps_r must be positive", fn);
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    const int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    const size_t n = (size_t)Nx * Ny * Nz;
    uint8_t top = 0;
    for (size_t q = 0; q < n; ++q) top = std::max(top, ids[q]);
    if (top >= n_ids) return fail(LX_EINVAL, "%s: material id %d >= n_ids", fn, (int)top);
    for (int c = 0; c < 3; ++c)
        for (int q = 0; q < 256; ++q) {
            const double er = q < n_ids ? eps3[3 * q + c] : 1.0;
            const double Cb = s->dt / (s->eps0 * er);             // maxwell_three_dim_pml.cc:122-128
            s->cb_host[c * 256 + q] = (float)(Cb / (s->eta0 * s->p.dx));
        }
    // upload in the caller's C order and transpose on the device into the
    // x-fastest layout (the padding columns keep id 0)
    void *raw = nullptr;
    lx_status st = dalloc(s, &raw, n);
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(s->cs));
    HIP_TRY(hipMemcpy(raw, ids, n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->cb, s->cb_host, sizeof(s->cb_host), hipMemcpyHostToDevice));
    launch_ids_to_device((const uint8_t *)raw, s->mat, Nx, Ny, Nz, s->px, s->plane, s->cs);
    st = check_launch();
    HIP_TRY(hipStreamSynchronize(s->cs));
    dfree(s, raw);
    return st;
}

extern "C" {

lx_status lx_sim_set_materials(lx_sim *s, const uint8_t *ids, const double *eps_r, int32_t n_ids) {
    if (!eps_r || n_ids < 1 || n_ids > 256) return set_materials(s, ids, eps_r, n_ids, "lx_sim_set_materials");
    std::vector<double> eps3(3 * (size_t)n_ids);
    for (int q = 0; q < n_ids; ++q) eps3[3 * q] = eps3[3 * q + 1] = eps3[3 * q + 2] = eps_r[q];
    return set_materials(s, ids, eps3.data(), n_ids, "lx_sim_set_materials");
}

lx_status lx_sim_set_materials3(lx_sim *s, const uint8_t *ids, const double *eps_r3, int32_t n_ids) {
    return set_materials(s, ids, eps_r3, n_ids, "lx_sim_set_materials3");
}

lx_status lx_sim_set_probes(lx_sim *s, const int32_t *cijk, int32_t n) {
    if (!s || (n > 0 && !cijk) || n < 0) return fail(LX_EINVAL, "lx_sim_set_probes: bad arguments");
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    std::vector<long long> idx(n);
    std::vector<int> comp(n);
    for (int q = 0; q < n; ++q) {
        const int c = cijk[4 * q];
        int ext[3];
        if (c < 0 || c > 5 || !comp_extent(s->N, s->planar, c, ext)) return fail(LX_EINVAL, "probe %d: bad component", q);
        for (int a = 0; a < 3; ++a)
            if (cijk[4 * q + 1 + a] < 0 || cijk[4 * q + 1 + a] >= ext[a]) return fail(LX_EINVAL, "probe %d outside the grid", q);
        idx[q] = (long long)cijk[4 * q + 3] * s->plane + (long long)cijk[4 * q + 2] * s->px + cijk[4 * q + 1];
        comp[q] = c;
    }
    HIP_TRY(hipStreamSynchronize(s->cs));
    drop_graph(s);
    s->gen++;
    if (s->pidx) { (void)hipFree(s->pidx); s->pidx = nullptr; }
    if (s->pcomp) { (void)hipFree(s->pcomp); s->pcomp = nullptr; }
    if (s->prec) { (void)hipFree(s->prec); s->prec = nullptr; s->bytes -= (int64_t)sizeof(double) * s->pcap * s->np; }
    s->np = 0;
    s->pcap = 0;
    s->probe_host.assign(cijk, cijk + 4 * (size_t)n);
    if (n == 0) return LX_OK;
    HIP_TRY(hipMalloc(&s->pidx, sizeof(long long) * n));
    HIP_TRY(hipMalloc(&s->pcomp, sizeof(int) * n));
    HIP_TRY(hipMemcpy(s->pidx, idx.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->pcomp, comp.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    s->np = n;
    return ensure_probe_cap(s, std::max(s->p.num_timesteps, 1) + s->steps);
}

lx_status lx_sim_step(lx_sim *s, int32_t nsteps) {
    if (!s || nsteps < 0) return fail(LX_EINVAL, "lx_sim_step: bad arguments");
    if (nsteps == 0) return LX_OK;
    HIP_TRY(hipSetDevice(s->ctx->device));
    // One fused step from the buffer a started monitor reads writes the other
    // buffer (its H corrections touch H, which the E-only monitor does not
    // read), so it may run beside the monitor; the next writer of that buffer
    // waits.  A point source writes E into the buffer read, and the other
    // paths update in place: they wait at once.
    const void *keep = nullptr;
    if (s->mguard && s->fused && nsteps == 1 && !s->p.use_ps && s->p.monitor_every <= 0 && s->F == s->mguard) {
        keep = s->mguard;
        s->mguard = nullptr;
    }
    member_enter(s);
    lx_status st = ensure_probe_cap(s, s->steps + nsteps);
    if (st) return st;
    int left = nsteps;
    const int chunk = s->p.monitor_every > 0 ? s->p.monitor_every : nsteps;
    while (left > 0) {
        const int n = std::min(left, chunk);
        int done = 0;
        for (int i = 0; i < kGraphLens; ++i) {
            const int K = kGraphSteps[i];
            while (n - done >= K) {
                hipGraphExec_t gx = ensure_graph(s, i, &st);   // per buffer state
                if (st) return st;
                if (!gx) break;
                HIP_TRY(hipGraphLaunch(gx, s->cs));
                done += K;
                settle_parity(s, K);
            }
        }
        enqueue_seq(s, n - done);
        settle_parity(s, n - done);
        if ((st = check_launch())) return st;
        s->steps += n;
        left -= n;
        if (keep) s->mguard = keep;
        if (s->p.monitor_every > 0) {   // the NaN guard: the E-only pass (H feeds E within a step)
            double mx;
            if ((st = lx_sim_max_e2(s, &mx))) return st;
        }
    }
    return LX_OK;
}

// Wait for a stream unless it is already idle (a query is far cheaper than
// a synchronisation call; batches sync every member after each step).
static lx_status drain(hipStream_t st) {
    const hipError_t q = hipStreamQuery(st);
    if (q == hipSuccess) return LX_OK;
    if (q != hipErrorNotReady) HIP_TRY(q);
    HIP_TRY(hipStreamSynchronize(st));
    return LX_OK;
}

lx_status lx_sim_sync(lx_sim *s) {
    if (!s) return fail(LX_EINVAL, "null simulation");
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (s->batch_wait) {
        HIP_TRY(hipEventSynchronize(s->batch_wait));
        s->batch_wait = nullptr;
    }
    lx_status st = drain(s->cs);
    if (st) return st;
    s->cs_dirty = false;
    for (lx::DmaSlot &d : s->snap_dma)
        if ((st = lx::dma_wait(&d))) return st;
    if (s->ms && (st = drain(s->ms))) return st;
    return s->ss ? drain(s->ss) : LX_OK;
}

static lx_status gather(lx_sim *s, int comp, double *dst_dev, int ext[3]) {
    const double scale = (comp >= 3 ? 1.0 / s->eta0 : 1.0) / s->sscale;
    if (s->fp16) launch_gather(make_grid<__half>(s), comp, ext[0], ext[1], ext[2], scale, dst_dev, s->cs);
    else launch_gather(make_grid<float>(s), comp, ext[0], ext[1], ext[2], scale, dst_dev, s->cs);
    return check_launch();
}

}  // extern "C"

namespace lx {
lx_status gather_output(lx_sim *s, int what, int comp, double *dst_dev, const int ext[3]) {
    if (what == LX_OUT_FIELD) {
        int e[3] = {ext[0], ext[1], ext[2]};
        return gather(s, comp, dst_dev, e);
    }
    const double e2scale = 1.0 / (s->sscale * s->sscale);
    if (s->fp16) launch_gather_e2(make_grid<__half>(s), ext[0], ext[1], ext[2], e2scale, dst_dev, s->cs);
    else launch_gather_e2(make_grid<float>(s), ext[0], ext[1], ext[2], e2scale, dst_dev, s->cs);
    return check_launch();
}
}  // namespace lx

extern "C" {

lx_status lx_sim_read_field(lx_sim *s, int32_t comp, double *dst, size_t n) {
    if (!s || !dst) return fail(LX_EINVAL, "lx_sim_read_field: null pointer");
    int ext[3];
    if (comp < 0 || comp > 5 || !comp_extent(s->N, s->planar, comp, ext)) return fail(LX_EINVAL, "bad component %d", comp);
    const size_t m = (size_t)ext[0] * ext[1] * ext[2];
    if (n != m) return fail(LX_EINVAL, "buffer holds %zu values, component has %zu", n, m);
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    if (s->rd_n < m) {
        if (s->rd_stage) (void)hipFree(s->rd_stage);
        s->rd_stage = nullptr;
        HIP_TRY(hipMalloc(&s->rd_stage, sizeof(double) * m));
        s->rd_n = m;
    }
    lx_status st = gather(s, comp, s->rd_stage, ext);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(dst, s->rd_stage, sizeof(double) * m, hipMemcpyDeviceToHost, s->cs));
    HIP_TRY(hipStreamSynchronize(s->cs));
    return LX_OK;
}

lx_status lx_sim_snapshot_async(lx_sim *s, int32_t comp, double *dst, size_t n) {
    if (!s || !dst) return fail(LX_EINVAL, "lx_sim_snapshot_async: null pointer");
    int ext[3];
    if (comp < 0 || comp > 5 || !comp_extent(s->N, s->planar, comp, ext)) return fail(LX_EINVAL, "bad component %d", comp);
    const size_t m = (size_t)ext[0] * ext[1] * ext[2];
    if (n != m) return fail(LX_EINVAL, "buffer holds %zu values, component has %zu", n, m);
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    const int slot = s->snap_next;
    s->snap_next ^= 1;
    // the previous copy out of this staging slot must be finished (a DMA
    // copy on the host, a stream copy on the device)
    lx_status st = dma_wait(&s->snap_dma[slot]);
    if (st) return st;
    if (s->stage_n[slot] < m) {
        HIP_TRY(hipEventSynchronize(s->done[slot]));
        if (s->stage[slot]) (void)hipFree(s->stage[slot]);
        s->stage[slot] = nullptr;
        HIP_TRY(hipMalloc(&s->stage[slot], sizeof(double) * m));
        s->stage_n[slot] = m;
    }
    HIP_TRY(hipStreamWaitEvent(s->cs, s->done[slot], 0));
    if ((st = gather(s, comp, s->stage[slot], ext))) return st;
    const size_t bytes = sizeof(double) * m;
    if (dma_eligible(dst, s->stage[slot], bytes)) {   // runtime-allocated (pinned) host memory: the DMA engine
        if ((st = dma_mark(&s->snap_dma[slot], s->cs))) return st;
        return dma_enqueue(&s->snap_dma[slot], dst, s->stage[slot], bytes);
    }
    // pageable or registered memory: a copy on the side stream
    if ((st = ensure_side_stream(s))) return st;
    HIP_TRY(hipEventRecord(s->ready[slot], s->cs));
    HIP_TRY(hipStreamWaitEvent(s->ss, s->ready[slot], 0));
    HIP_TRY(hipMemcpyAsync(dst, s->stage[slot], bytes, hipMemcpyDeviceToHost, s->ss));
    HIP_TRY(hipEventRecord(s->done[slot], s->ss));
    return LX_OK;
}

lx_status lx_sim_read_probes(lx_sim *s, double *dst, size_t n, int32_t *n_steps) {
    if (!s || !n_steps) return fail(LX_EINVAL, "lx_sim_read_probes: null pointer");
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    HIP_TRY(hipStreamSynchronize(s->cs));
    const int rec = std::min(s->steps, s->pcap);
    const size_t m = (size_t)rec * s->np;
    *n_steps = s->np ? rec : 0;
    if (!s->np) return LX_OK;
    if (n < m || !dst) return fail(LX_EINVAL, "probe buffer too small: need %zu", m);
    HIP_TRY(hipMemcpy(dst, s->prec, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (s->sscale != 1.0)
        for (size_t q = 0; q < m; ++q) dst[q] /= s->sscale;
    return LX_OK;
}

// The monitor's two passes on the compute stream and the result to the host
// (max|E|^2, the energy sum and the non-finite count, unscaled); e_only
// reads the E components alone (max and their non-finite count).
static lx_status run_monitor(lx_sim *s, bool e_only, double out[3]) {
    const double cb0 = s->c0 * s->dt / s->p.dx;   // vacuum coefficient: eps_r = cb0/cb
    if (s->fp16) launch_monitor(make_grid<__half>(s), s->planar, cb0, s->mon, e_only, s->cs);
    else launch_monitor(make_grid<float>(s), s->planar, cb0, s->mon, e_only, s->cs);
    lx_status st = check_launch();
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(out, s->mon, sizeof(double) * 3, hipMemcpyDeviceToHost, s->cs));
    HIP_TRY(hipStreamSynchronize(s->cs));
    const double s2 = s->sscale * s->sscale;
    out[0] /= s2;
    out[1] /= s2;
    return LX_OK;
}

lx_status lx_sim_monitor(lx_sim *s, double *max_e2, double *energy) {
    if (!s) return fail(LX_EINVAL, "null simulation");
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    double out[3] = {0, 0, 0};
    lx_status st = run_monitor(s, false, out);
    if (st) return st;
    const double vol = s->p.dx * s->p.dz * (s->planar ? 1.0 : s->p.dy);
    if (max_e2) *max_e2 = out[0];
    if (energy) *energy = 0.5 * s->eps0 * out[1] * vol;
    if (out[2] > 0) return fail(LX_EUNSTABLE, "non-finite field values after %d steps", s->steps);
    return LX_OK;
}

lx_status lx_sim_max_e2(lx_sim *s, double *max_e2) {
    if (!s || !max_e2) return fail(LX_EINVAL, "lx_sim_max_e2: null pointer");
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    double out[3] = {0, 0, 0};
    lx_status st = run_monitor(s, true, out);
    if (st) return st;
    *max_e2 = out[0];
    if (out[2] > 0) return fail(LX_EUNSTABLE, "non-finite E values after %d steps", s->steps);
    return LX_OK;
}

lx_status lx_sim_max_e2_start(lx_sim *s) {
    if (!s) return fail(LX_EINVAL, "null simulation");
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (s->mpend) HIP_TRY(hipEventSynchronize(s->mdone));   // its result is dropped
    s->mpend = 0;
    if (!s->ms) {
        lx_status st = take_stream(s->ctx, &s->ms);
        if (st) return st;
        HIP_TRY(hipEventCreateWithFlags(&s->mstart, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&s->mdone, hipEventDisableTiming));
        void *p = nullptr;
        HIP_TRY(hipHostMalloc(&p, 4 * sizeof(double), hipHostMallocCoherent | hipHostMallocMapped));
        s->mres = (double *)p;
    }
    member_enter(s);
    const double cb0 = s->c0 * s->dt / s->p.dx;
    double *scratch = s->mon + 4 + 3 * (size_t)kMonBlocks;   // the second half: not the synchronous monitor's
    double *res = nullptr;
    HIP_TRY(hipHostGetDevicePointer((void **)&res, s->mres, 0));
    HIP_TRY(hipEventRecord(s->mstart, s->cs));
    HIP_TRY(hipStreamWaitEvent(s->ms, s->mstart, 0));
    if (s->fp16) launch_monitor(make_grid<__half>(s), s->planar, cb0, scratch, true, s->ms, res);
    else launch_monitor(make_grid<float>(s), s->planar, cb0, scratch, true, s->ms, res);
    lx_status st = check_launch();
    if (st) return st;
    HIP_TRY(hipEventRecord(s->mdone, s->ms));
    s->mpend = 1;
    s->mguard = s->F;
    return LX_OK;
}

lx_status lx_sim_max_e2_finish(lx_sim *s, double *max_e2) {
    if (!s || !max_e2) return fail(LX_EINVAL, "lx_sim_max_e2_finish: null pointer");
    if (!s->mpend) return fail(LX_EINVAL, "lx_sim_max_e2_finish: no started monitor");
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipEventSynchronize(s->mdone));
    s->mpend = 0;
    s->mguard = nullptr;   // complete: nothing on cs needs to wait for it
    const double mx = s->mres[0], bad = s->mres[2];
    *max_e2 = mx / (s->sscale * s->sscale);
    if (bad > 0) return fail(LX_EUNSTABLE, "non-finite E values");
    return LX_OK;
}

lx_status lx_sim_get_info(lx_sim *s, lx_sim_info *info) {
    if (!s || !info) return fail(LX_EINVAL, "null pointer");
    std::memset(info, 0, sizeof(*info));
    info->dt = s->dt;
    info->c0 = s->c0;
    info->eps0 = s->eps0;
    info->mu0 = s->mu0;
    info->eta0 = s->eta0;
    info->cells = (int64_t)s->N[0] * s->N[1] * s->N[2];
    int64_t slab_bytes = 0;
    for (int a = 0; a < 3; ++a) {
        info->slab_cells[a] = s->slab_cells[a];
        info->slab_lo[a] = s->slo[a];
        info->slab_hi[a] = s->shi[a];
        slab_bytes += s->slab_cells[a] * (s->planar ? 16 : 32);
    }
    const int64_t per = s->planar ? (9 * (int64_t)s->esize + 1) : (18 * (int64_t)s->esize + 1);
    info->bytes_per_step = info->cells * per + slab_bytes;
    info->device_bytes = s->bytes;
    info->step = s->steps;
    info->fused_step = s->fused;
    info->kz = s->kz;
    info->waves = s->fused ? s->nw : 4;
    info->core_bytes = info->bytes_per_step;
    info->cells_per_lane = s->fused ? s->vec : 1;
    info->shell_kz = s->fused && s->pml && s->kzs > 0 ? s->kzs : s->kz;
    info->shell_concurrent = s->fused && s->ks && s->kconc ? 1 : 0;
    info->interior_rows = s->fused ? s->rows : 0;
    info->shell_pipelined = s->fused && shell_pipe_ok(s) && (s->tfH.n > 0 || s->use_aux || s->np > 0) &&
                            tf_merge_ok(s) ? 1 : 0;
    if (s->fused && s->pml) {
        int rows = 0, planes = 0;
        step3_core_extent(s->N[1], s->N[2], s->kz, s->slo, s->shi, &rows, &planes);
        const int64_t box = (int64_t)rows * planes;
        info->core_bytes = (int64_t)s->N[0] * box * per + (int64_t)s->sw[0] * box * 32;
    }
    return LX_OK;
}

void *lx_sim_stream(lx_sim *s) { return s ? (void *)s->cs : nullptr; }

static lx_status reset_state(lx_sim *s) {
    launch_zero_bytes(s->F, (size_t)s->fstride * 6 * s->esize, s->cs);
    launch_zero_bytes(s->psi, psi_total(s) * sizeof(float), s->cs);
    if (s->F_alt) launch_zero_bytes(s->F_alt, (size_t)s->fstride * 6 * s->esize, s->cs);
    if (s->psi_alt) launch_zero_bytes(s->psi_alt, psi_total(s) * sizeof(float), s->cs);
    launch_zero_bytes(s->counter, 64, s->cs);
    if (s->prec) launch_zero_bytes(s->prec, sizeof(double) * (size_t)s->pcap * s->np, s->cs);
    if (s->use_aux) {
        std::vector<double> init(2 * (size_t)s->aux.L, 0.0);
        init[0] = waveform_host(s->aux.wave, 0.0, s->aux.c, s->aux.lambda0, s->aux.fwhm, s->aux.delay);
        HIP_TRY(hipMemcpyAsync(s->aux.e, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice, s->cs));
    }
    HIP_TRY(hipStreamSynchronize(s->cs));
    s->steps = 0;
    return check_launch();
}

lx_status lx_sim_autotune(lx_sim *s, int32_t *kz_out) {
    if (!s) return fail(LX_EINVAL, "null simulation");
    if (s->steps != 0) return fail(LX_EINVAL, "lx_sim_autotune must run before the first step");
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    using Key = std::tuple<int, int, int, int, int, int, int>;
    static std::mutex mu;
    static std::map<Key, std::array<int, 7>> cache;   // (kz, nw, fused, vec, shell kz, shell stream, rows)
    const Key key{s->N[0], s->N[1], s->N[2], s->fp16, s->pml, s->planar, 0};
    // LX_TUNE=fused,nw,vec,kz,shell_concurrent[,rows] pins the result without
    // timing, ahead of any cached result (a profiling process repeats another
    // process's choice)
    if (const char *pin = std::getenv("LX_TUNE"); pin && !s->planar) {
        int f = 0, nw = 0, vec = 0, kz = 0, conc = 0, rows = 0;
        const int nf = std::sscanf(pin, "%d,%d,%d,%d,%d,%d", &f, &nw, &vec, &kz, &conc, &rows);
        if ((nf != 5 && nf != 6) || kz < 1 || (nw != 4 && nw != 8) || !vec_ok(s, vec) ||
            (rows != 0 && rows != 3 && rows != 7))
            return fail(LX_EINVAL,
                        "LX_TUNE: expected fused,nw,vec,kz,shell_concurrent[,rows] valid for this grid, got '%s'", pin);
        s->rows = rows;
        s->kz = kz;
        s->nw = nw;
        s->vec = vec;
        const char *sk = std::getenv("LX_SHELL_KZ");   // the shell depth stays pinnable beside the pin
        s->kzs = sk && *sk ? std::max(0, std::atoi(sk)) : 0;
        s->kconc = conc ? 1 : 0;
        if (f && fused_ok(s)) {
            lx_status st = ensure_fused_buffers(s);
            if (st) return st;
            s->fused = 1;
        } else {
            s->fused = 0;
            if (s->F_alt) release_fused_buffers(s);
        }
        s->gen++;
        drop_graph(s);
        if (kz_out) *kz_out = s->kz;
        return prepare_graphs(s);
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            s->kz = it->second[0];
            s->nw = it->second[1];
            s->vec = it->second[3];
            s->kzs = it->second[4];
            s->kconc = it->second[5];
            s->rows = it->second[6];
            if (it->second[2] && fused_ok(s)) {
                lx_status st = ensure_fused_buffers(s);
                if (st) return st;
                s->fused = 1;
            } else if (s->F_alt && !s->planar) {
                release_fused_buffers(s);
            }
            s->gen++;
            drop_graph(s);
            if (kz_out) *kz_out = s->kz;
            return prepare_graphs(s);
        }
    }
    if (!s->planar) {
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        // n steps of the current settings: the sequence lx_sim_step enqueues
        // (TF/SF launches, the shell pipelined where it may be, the
        // sequence-end kernel), so that every candidate is timed as the time
        // loop will run it; the state it leaves is reset below
        auto enq = [&](int n) { enqueue_seq(s, n); };
        // milliseconds per step of the current settings: the fastest of reps
        // repetitions of n steps, launched directly or, with graph, replayed
        // from a captured graph as the time loop runs them
        auto time_steps = [&](int n, int reps, float *out, bool graph = false) -> lx_status {
            *out = 1e30f;
            hipGraphExec_t gx = nullptr;
            if (graph) {
                hipGraph_t gr = nullptr;
                HIP_TRY(hipStreamBeginCapture(s->cs, hipStreamCaptureModeThreadLocal));
                enq(n);
                HIP_TRY(hipStreamEndCapture(s->cs, &gr));
                const hipError_t e = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) return fail(LX_EDEVICE, "autotune graph: %s", hipGetErrorString(e));
            }
            for (int rep = 0; rep < reps; ++rep) {
                HIP_TRY(hipEventRecord(e0, s->cs));
                if (gx) HIP_TRY(hipGraphLaunch(gx, s->cs));
                else enq(n);
                HIP_TRY(hipEventRecord(e1, s->cs));
                HIP_TRY(hipEventSynchronize(e1));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0 || reps == 1) *out = std::min(*out, ms / n);   // the first repetition warms up
            }
            if (gx) (void)hipGraphExecDestroy(gx);
            return check_launch();
        };
        const bool verbose = std::getenv("LX_VERBOSE") != nullptr;
        // candidates: the half-step pair (nw = 0) and the fused step with 8 or
        // 4 wavefronts per block and 1, 2 or 4 cells per lane (its second
        // buffer set is kept only if it wins); the CPML shell takes the
        // interior's z-block depth.  A first pass times every candidate over
        // 4 steps, best of 2 (the shell on the second stream where there is one); the
        // five fastest are timed again over longer runs replayed from graphs,
        // as the time loop runs them, which separates near-ties the short
        // pass cannot (below).
        const int can_fuse = fused_ok(s) && ensure_fused_buffers(s) == LX_OK;
        s->kzs = 0;
        s->kconc = 1;
        // rows: the interior launch with 3 or 7 rows per block sharing their
        // neighbours through LDS (0: one unit per wavefront)
        struct Cand { int nw, vec, kz; float ms; int conc; int rows; };
        std::vector<Cand> tried;
        for (Cand c : {Cand{0, 1, 0, 0, 1, 0}, Cand{8, 1, 0, 0, 1, 0}, Cand{4, 1, 0, 0, 1, 0}, Cand{4, 2, 0, 0, 1, 0},
                       Cand{8, 2, 0, 0, 1, 0}, Cand{4, 4, 0, 0, 1, 0}, Cand{4, 1, 0, 0, 1, 3}, Cand{4, 1, 0, 0, 1, 7},
                       Cand{4, 2, 0, 0, 1, 3}, Cand{4, 2, 0, 0, 1, 7}, Cand{4, 4, 0, 0, 1, 3},
                       Cand{4, 4, 0, 0, 1, 7}}) {
            if (c.nw && !can_fuse) continue;
            if (!vec_ok(s, c.vec)) continue;
            for (int kz : {4, 5, 6, 8, 12, 16, 32, 64}) {
                if (kz > s->N[2] && kz != 4) continue;
                s->fused = c.nw != 0;
                if (c.nw) s->nw = c.nw;
                s->vec = c.vec;
                s->rows = c.rows;
                s->kz = kz;
                lx_status st = time_steps(4, 3, &c.ms);
                if (st) return st;
                c.kz = kz;
                tried.push_back(c);
                if (verbose)
                    std::fprintf(stderr, "lx_sim_autotune: %s nw=%d vec=%d rows=%d kz=%d  %.4f ms per step\n",
                                 s->fused ? "fused" : "half-step", c.nw, c.vec, c.rows, kz, c.ms);
            }
        }
        std::sort(tried.begin(), tried.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
        // finalists: the five fastest, each with the shell beside and behind
        // the interior launch; timed in three interleaved rounds (so that a
        // drift of the device's clock falls on every candidate alike) of
        // about 25 ms of steps each, replayed from graphs, best round kept
        std::vector<Cand> fin;
        for (size_t q = 0; q < std::min<size_t>(5, tried.size()); ++q)
            for (int conc = 1; conc >= 0; --conc) {
                Cand c = tried[q];
                if (conc == 0 && !(c.nw && s->ks)) continue;   // only the fused step has a shell launch
                c.conc = conc;
                c.ms = 1e30f;
                fin.push_back(c);
            }
        const int nfin = (int)std::min(256.0, std::max(8.0, std::ceil(25.0 / std::max(tried.front().ms, 1e-3f))));
        for (int round = 0; round < 3; ++round) {
            for (Cand &c : fin) {
                s->fused = c.nw != 0;
                if (c.nw) s->nw = c.nw;
                s->vec = c.vec;
                s->rows = c.rows;
                s->kz = c.kz;
                s->kconc = c.conc;
                float ms = 0;
                lx_status st = time_steps(nfin, 2, &ms, s->graphs != 0);
                if (st) return st;
                c.ms = std::min(c.ms, ms);
                if (verbose)
                    std::fprintf(stderr,
                                 "lx_sim_autotune: finalist round %d %s nw=%d vec=%d rows=%d kz=%d shell %s  %.4f ms per "
                                 "step\n",
                                 round, s->fused ? "fused" : "half-step", c.nw, c.vec, c.rows, c.kz,
                                 c.conc ? "beside" : "behind", ms);
            }
        }
        Cand best = fin.front();
        for (const Cand &c : fin)
            if (c.ms < best.ms) best = c;
        s->fused = best.nw != 0;
        if (best.nw) s->nw = best.nw;
        s->vec = best.vec;
        s->rows = best.rows;
        s->kz = best.kz;
        s->kconc = best.conc;
        s->kzs = 0;   // the shell takes the interior's depth: a depth of its own was never faster when measured
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (!s->fused) release_fused_buffers(s);
    }
    drop_graph(s);
    s->gen++;
    lx_status st = reset_state(s);
    if (st) return st;
    if ((st = prepare_graphs(s))) return st;
    {
        std::lock_guard<std::mutex> lk(mu);
        cache[key] = {s->kz, s->nw, s->fused, s->vec, s->kzs, s->kconc, s->rows};
    }
    if (kz_out) *kz_out = s->kz;
    return LX_OK;
}

lx_status lx_sim_time_kernels(lx_sim *s, int32_t nsteps, double *ms_h, double *ms_e, double *ms_step) {
    if (!s || nsteps < 1) return fail(LX_EINVAL, "lx_sim_time_kernels: bad arguments");
    HIP_TRY(hipSetDevice(s->ctx->device));
    member_enter(s);
    lx_status st = ensure_probe_cap(s, s->steps + nsteps);
    if (st) return st;
    std::vector<hipEvent_t> ev(4 * (size_t)nsteps + 2);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(ev[4 * nsteps], s->cs));
    enqueue_seq(s, nsteps, &ev);
    settle_parity(s, nsteps);
    HIP_TRY(hipEventRecord(ev[4 * nsteps + 1], s->cs));
    if ((st = check_launch())) return st;
    HIP_TRY(hipStreamSynchronize(s->cs));
    double th = 0, te = 0;
    for (int q = 0; q < nsteps; ++q) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, ev[4 * q + 0], ev[4 * q + 1]));
        HIP_TRY(hipEventElapsedTime(&b, ev[4 * q + 2], ev[4 * q + 3]));
        th += a;
        te += b;
    }
    float tot = 0;
    HIP_TRY(hipEventElapsedTime(&tot, ev[4 * nsteps], ev[4 * nsteps + 1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    s->steps += nsteps;
    if (ms_h) *ms_h = th / nsteps;
    if (ms_e) *ms_e = te / nsteps;
    if (ms_step) *ms_step = tot / nsteps;
    return LX_OK;
}

}  // extern "C"

// ---- footprint estimate (admission control) --------------------------------
extern "C" lx_status lx_params_footprint(const lx_params *p, int32_t batched, int32_t steps, int32_t n_probes,
                                         int64_t *bytes) {
    if (!p || !bytes) return fail(LX_EINVAL, "lx_params_footprint: null pointer");
    lx_status st = lx_params_check(p);
    if (st) return st;
    const int64_t N[3] = {p->nx, p->is_planar ? 1 : p->ny, p->nz};
    const int64_t px = (N[0] + 63) / 64 * 64;
    const int64_t ncell = px * N[1] * N[2];
    const int64_t fstride = (ncell + 127) / 128 * 128;
    const int64_t esize = p->precision == LX_FP16 ? 2 : 4;
    const int64_t fields = fstride * 6 * esize;
    int64_t psi = 0;
    if (p->use_pml) {
        const double d[3] = {p->dx, p->dy, p->dz};
        const int64_t ext[3] = {px, N[1], N[2]};
        for (int a = 0; a < 3; ++a) {
            if (p->is_planar && a == 1) continue;
            const int64_t w = std::min<int64_t>(ext[a], 2 * ((int64_t)std::ceil(p->pml_thickness / d[a]) + 5));
            int64_t other = 1;
            for (int b = 0; b < 3; ++b)
                if (b != a) other *= ext[b];
            psi += 4 * w * other * (int64_t)sizeof(float);
        }
    }
    const int64_t nmax = std::max(px, std::max(N[1], N[2])) + 4;
    const int64_t tables = 3 * 6 * nmax * 4;
    const int64_t line = 2 * 8 * (std::max(N[0], N[2]) + (int64_t)std::ceil(p->courant * std::max(steps, p->num_timesteps)) + 64);
    const int64_t probes = (int64_t)n_probes * std::max(steps, p->num_timesteps) * 8;
    int64_t total = fields + ncell + psi + tables + line + probes + (2 << 20);   // 2 MiB of allocation slack
    if (batched && p->is_planar) total += fields + psi + 2 * 8 * 8 * (std::max(N[0], N[2]) + 8);
    if (!p->is_planar) total += fields + psi;   // second buffer set of the fused step
    *bytes = total;
    return LX_OK;
}
