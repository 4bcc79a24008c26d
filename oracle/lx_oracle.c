/*
 * lx_oracle.c -- FP64 CPU restatement of Luxum's Yee/CPML/TF-SF timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load liblxoracle.so.  Nothing under luxum_amd/ links, loads or calls it.
 *
 * It restates, in its own code, the reference's loop structure
 * ("reference-structured" mode): separate split-curl arrays, full-domain psi
 * arrays, per-cell Cb arrays, serial TF/SF corrections and OpenMP
 * `schedule(dynamic)` over the outermost index, all in IEEE FP64.  It is
 * compiled with -ffp-contract=off so that every multiply and add rounds as in
 * the reference's x86-64 build.  The reference cannot be compiled here
 * (Blitz++ and yaml-cpp are absent), so no reference-built vectors exist.
 *
 * Followed reference code (paths relative to the reference root):
 *   constants, PW directions ......... src/solvers/maxwell.cc:60-148
 *   split curls, PEC ranges .......... lib/field/vfield.cc:102-386
 *   3D CPML loop ..................... src/solvers/three_dim/maxwell_three_dim_pml.cc:52-287
 *   3D plain loop .................... src/solvers/three_dim/maxwell_three_dim.cc:52-201
 *   2D CPML loop ..................... src/solvers/two_dim/maxwell_two_dim_pml.cc:54-192
 *   2D plain loop .................... src/solvers/two_dim/maxwell_two_dim.cc:51-155
 *   TF/SF faces and signs ............ src/solvers/maxwell.cc:194-812
 *   TF/SF clipping ................... src/solvers/maxwell.cc:825-929
 *   CPML profiles .................... src/solvers/maxwell.cc:997-1280
 *   field extents .................... lib/field/field.cc:58-92, lib/field/vfield.cc:58-87
 *
 * Extensions that the reference lacks (same definitions as the product):
 *   uint8 material ids, the FP64 aux-1D incident line, Gaussian-modulated
 *   waveforms, the soft point source and probes.
 *
 * Pinning status (what checks each function; see tests/test_oracle.py):
 *   profiles      closed forms at hand-evaluated indices, symmetry, c==0 iff sigma==0
 *   curls/update  hand-computed 4^3 case, PEC cavity eigenfrequency (Yee dispersion),
 *                 discrete energy invariant, div E invariance
 *   CPML          reflection of a normally incident pulse below a bound, psi==0 off slabs
 *   TF/SF         vacuum TF region equals the incident wave; SF leakage bound;
 *                 face index sets vs the reference loop bounds
 *   aux-1D line   dispersion-matched propagation vs the 1D Yee closed form
 *   The exact FP64 bits of the reference's sign/offset conventions are parity
 *   unpinned (only running the reference could pin them).
 */
#include "lx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* storage                                                                 */
/* ---------------------------------------------------------------------- */

/* Storage type of fields, split curls and psi.  The checker is built with
 * double; a second build (-DLXO_STORAGE_FLOAT, liblxoracle_f32s.so) rounds
 * every stored value to float while keeping FP64 arithmetic, the precision
 * floor of any FP32-storage implementation of this algorithm, which tests
 * use to tell storage rounding from kernel error. */
#ifdef LXO_STORAGE_FLOAT
typedef float store_t;
#else
typedef double store_t;
#endif

typedef struct {
    int n[3];
    store_t *v;
} fld;

#define AT(f, i, j, k) ((f).v[((size_t)(i) * (size_t)(f).n[1] + (size_t)(j)) * (size_t)(f).n[2] + (size_t)(k)])

/* Yee stagger (x,y,z) per component Ex,Ey,Ez,Hx,Hy,Hz (vfield.cc:75-84). */
static const int STAG[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

struct lxo_sim {
    lx_params p;
    int N[3];
    int planar, pml, threads;
    double eps0, mu0, c, dt, d[3], sigma_pw, init, sigma_opt, sigma_max, eta0;
    double kv[3], Edir[3], Hdir[3];
    int lo[3], hi[3], st[3], en[3];
    fld E[3], H[3];
    fld Cb[3];
    double Db;
    fld c1E[3], c2E[3], c1H[3], c2H[3];
    fld pE1[3], pE2[3], pH1[3], pH2[3];
    double *ke[3], *kh[3], *be[3], *bh[3], *ce[3], *chh[3];
    /* aux-1D line */
    int aux_axis, aux_x0, aux_L;
    double aux_S;
    double *aux_e, *aux_h;
    /* probes */
    int nprobe;
    int32_t *probe;
    double *rec;
    int rec_steps, rec_cap;
    int t;
};

static int comp_ext(const int N[3], int planar, int comp, int ext[3]) {
    if (planar) {
        /* 2D: Ey (Nx,1,Nz), Hx (Nx,1,Nz-1), Hz (Nx-1,1,Nz) (vfield.cc:62-72) */
        if (comp == LX_EY) { ext[0] = N[0]; ext[1] = 1; ext[2] = N[2]; return 1; }
        if (comp == LX_HX) { ext[0] = N[0]; ext[1] = 1; ext[2] = N[2] - 1; return 1; }
        if (comp == LX_HZ) { ext[0] = N[0] - 1; ext[1] = 1; ext[2] = N[2]; return 1; }
        ext[0] = ext[1] = ext[2] = 0;
        return 0;
    }
    for (int a = 0; a < 3; ++a) ext[a] = N[a] - STAG[comp][a];
    return 1;
}

int lxo_field_size(const lx_params *p, int comp, int ext[3]) {
    int N[3] = {p->nx, p->ny, p->nz};
    if (comp < 0 || comp > 5) return -1;
    return comp_ext(N, p->is_planar, comp, ext) ? 0 : -1;
}

static int falloc(fld *f, const int ext[3]) {
    size_t n = (size_t)ext[0] * ext[1] * ext[2];
    f->n[0] = ext[0]; f->n[1] = ext[1]; f->n[2] = ext[2];
    f->v = (store_t *)calloc(n ? n : 1, sizeof(store_t));
    return f->v ? 0 : -1;
}

static int falloc_comp(lxo_sim *s, fld *f, int comp) {
    int ext[3];
    if (!comp_ext(s->N, s->planar, comp, ext)) { f->v = NULL; return 0; }
    return falloc(f, ext);
}

/* ---------------------------------------------------------------------- */
/* constants and CPML profiles (maxwell.cc:60-86, 997-1280)                 */
/* ---------------------------------------------------------------------- */

static void constants(const lx_params *p, double *eps0, double *mu0, double *c, double *dt,
                      double *sopt, double *smax) {
    *eps0 = 8.85 * pow(10, -12);
    *mu0 = M_PI * 4 * pow(10, -7);
    *c = 1 / pow(*eps0 * *mu0, 0.5);
    *dt = p->courant * p->dx / *c;
    *sopt = 0.8 * (p->sigma_m + 1) / p->dx / pow(*mu0 / *eps0, 0.5);
    *smax = *sopt * p->sigma_ratio;
}

void lxo_constants(const lx_params *p, double out[6]) {
    constants(p, &out[0], &out[1], &out[2], &out[3], &out[4], &out[5]);
}

/* Node coordinate (in cells) of index i: E nodes at i, H nodes at i+0.5. */
static double node(int i, int is_E) { return is_E ? (double)i : i + 0.5; }

static void prof_sigma(const lx_params *p, int dim, int is_E, double *w, int n) {
    double eps0, mu0, c, dt, sopt, smax;
    constants(p, &eps0, &mu0, &c, &dt, &sopt, &smax);
    double delta = dim == 0 ? p->dx : dim == 1 ? p->dy : p->dz;
    int N = dim == 0 ? p->nx : dim == 1 ? p->ny : p->nz;
    double d = p->pml_thickness / delta, m = p->sigma_m;
    for (int i = 0; i < n; ++i) {
        double x = node(i, is_E), v = 0.0;
        v = (x <= (d - 1) ? 1.0 : 0.0) * pow(((d - 1) - x) / (d - 1), m) * smax + (x > (d - 1) ? 1.0 : 0.0) * v;
        v = (x >= (N - d) ? 1.0 : 0.0) * pow((x - N + d) / (d - 1), m) * smax + (x < (N - d) ? 1.0 : 0.0) * v;
        w[i] = v;
    }
}

static void prof_kappa(const lx_params *p, int dim, int is_E, double *w, int n) {
    double delta = dim == 0 ? p->dx : dim == 1 ? p->dy : p->dz;
    int N = dim == 0 ? p->nx : dim == 1 ? p->ny : p->nz;
    double d = p->pml_thickness / delta, m = p->sigma_m, km = p->kappa_max;
    for (int i = 0; i < n; ++i) {
        double x = node(i, is_E), v = 1;
        v = (x <= (d - 1) ? 1.0 : 0.0) * (pow(((d - 1) - x) / (d - 1), m) * (km - 1) + 1) + (x > (d - 1) ? 1.0 : 0.0) * v;
        v = (x >= (N - d) ? 1.0 : 0.0) * (pow((x - N + d) / (d - 1), m) * (km - 1) + 1) + (x < (N - d) ? 1.0 : 0.0) * v;
        w[i] = v;
    }
}

static void prof_alpha(const lx_params *p, int dim, int is_E, double *w, int n) {
    double delta = dim == 0 ? p->dx : dim == 1 ? p->dy : p->dz;
    int N = dim == 0 ? p->nx : dim == 1 ? p->ny : p->nz;
    double d = p->pml_thickness / delta, ma = p->a_m, am = p->a_max;
    for (int i = 0; i < n; ++i) {
        double x = node(i, is_E), v = 0.0;
        v = (x <= (d - 1) ? 1.0 : 0.0) * pow((x / (d - 1)), ma) * am + (x > (d - 1) ? 1.0 : 0.0) * v;
        v = (x >= (N - d) ? 1.0 : 0.0) * pow((N - 1 - x) / (d - 1), ma) * am + (x < (N - d) ? 1.0 : 0.0) * v;
        w[i] = v;
    }
}

/* b = exp(-(sigma/kappa/eps0 - a/eps0)*dt): the alpha term enters with a
 * minus sign as in maxwell.cc:1215. */
static void prof_b(const lx_params *p, int dim, int is_E, double *w, int n) {
    double eps0, mu0, c, dt, sopt, smax;
    constants(p, &eps0, &mu0, &c, &dt, &sopt, &smax);
    double *sg = (double *)malloc(sizeof(double) * n), *al = (double *)malloc(sizeof(double) * n),
           *ka = (double *)malloc(sizeof(double) * n);
    prof_sigma(p, dim, is_E, sg, n);
    prof_alpha(p, dim, is_E, al, n);
    prof_kappa(p, dim, is_E, ka, n);
    for (int i = 0; i < n; ++i) w[i] = exp(-((sg[i] / ka[i] / eps0) - (al[i] / eps0)) * dt);
    free(sg); free(al); free(ka);
}

/* c = sigma/(sigma*kappa + kappa^2*a)*(b-1), zero where sigma == 0 (maxwell.cc:1266-1278). */
static void prof_c(const lx_params *p, int dim, int is_E, double *w, int n) {
    double *sg = (double *)malloc(sizeof(double) * n), *al = (double *)malloc(sizeof(double) * n),
           *ka = (double *)malloc(sizeof(double) * n);
    prof_sigma(p, dim, is_E, sg, n);
    prof_alpha(p, dim, is_E, al, n);
    prof_kappa(p, dim, is_E, ka, n);
    prof_b(p, dim, is_E, w, n);
    for (int i = 0; i < n; ++i) {
        if (sg[i] == 0) w[i] = 0.0;
        else w[i] = sg[i] / (sg[i] * ka[i] + pow(ka[i], 2) * al[i]) * (w[i] - 1);
    }
    free(sg); free(al); free(ka);
}

int lxo_profile(const lx_params *p, int dim, int is_E, int kind, double *out, int n) {
    switch (kind) {
    case LXO_SIGMA: prof_sigma(p, dim, is_E, out, n); return 0;
    case LXO_KAPPA: prof_kappa(p, dim, is_E, out, n); return 0;
    case LXO_ALPHA: prof_alpha(p, dim, is_E, out, n); return 0;
    case LXO_B: prof_b(p, dim, is_E, out, n); return 0;
    case LXO_C: prof_c(p, dim, is_E, out, n); return 0;
    }
    return -1;
}

/* ---------------------------------------------------------------------- */
/* waveforms (extension)                                                    */
/* ---------------------------------------------------------------------- */

static double waveform(int kind, double t, double c, double lambda0, double fwhm, double delay) {
    double f0 = c / lambda0, T0 = 1.0 / f0;
    double tau = fwhm * T0 / (2.0 * sqrt(log(2.0)));
    double t0 = delay * tau;
    double x = (t - t0) / tau;
    double env = exp(-x * x);
    if (kind == LX_WAVE_GAUSSIAN) return env;
    return env * sin(2.0 * M_PI * f0 * (t - t0));
}

/* ---------------------------------------------------------------------- */
/* TF/SF (maxwell.cc:194-929)                                               */
/* ---------------------------------------------------------------------- */

/* plane_wave_initialise for one axis: clip [lo,hi] to [0, N-1]. */
static void pw_clip(int lo, int hi, int N, int *st, int *en) {
    int s0 = 0, e0 = N - 1;
    if (lo > e0) { *st = -1; *en = -1; }
    else if (lo >= s0) { *st = lo; *en = hi > e0 ? e0 : hi; }
    else if (hi > e0) { *st = s0; *en = e0; }
    else if (hi >= s0) { *st = s0; *en = hi; }
    else { *st = -1; *en = -1; }
}

/* Incident scalar g at node coordinates (cx,cy,cz) (cells) and time level tp.
 * Analytic: the reference Gaussian (e.g. maxwell.cc:289, :397).
 * Aux-1D: the line's E (is_E_node) or eta0*H value at the coordinate along k. */
static double g_inc(const lxo_sim *s, double cx, double cy, double cz, double tp, int is_E_node) {
    if (s->p.incident == LX_INC_AUX1D) {
        double cc = s->aux_axis == 0 ? cx : s->aux_axis == 1 ? cy : cz;
        int m = (int)floor(cc) - s->aux_x0;
        if (m < 0 || m >= s->aux_L) return 0.0;
        return is_E_node ? s->aux_e[m] : (m < s->aux_L - 1 ? s->aux_h[m] : 0.0);
    }
    double pos;
    if (s->planar) pos = cx * s->d[0] * s->kv[0] + cz * s->d[2] * s->kv[2];
    else pos = cx * s->d[0] * s->kv[0] + cy * s->d[1] * s->kv[1] + cz * s->d[2] * s->kv[2];
    double arg = pos + s->init * s->d[0] - tp * s->c * s->dt;
    return exp(-pow(arg, 2) / s->sigma_pw);
}

/* Face tables.  For a face normal to axis a, the two curl entries corrected:
 * {target component, which split curl (1 or 2), differentiated component}. */
static const int HFACE[3][2][3] = {
    {{LX_HZ, 1, LX_EY}, {LX_HY, 2, LX_EZ}},
    {{LX_HZ, 2, LX_EX}, {LX_HX, 1, LX_EZ}},
    {{LX_HY, 1, LX_EX}, {LX_HX, 2, LX_EY}},
};
static const int EFACE[3][2][3] = {
    {{LX_EY, 2, LX_HZ}, {LX_EZ, 1, LX_HY}},
    {{LX_EX, 1, LX_HZ}, {LX_EZ, 2, LX_HX}},
    {{LX_EX, 2, LX_HY}, {LX_EY, 1, LX_HX}},
};

/* Apply the TF/SF corrections of one half step.
 * h_update: corrections of curl E (before the H update, time t-1/2),
 *           else of curl H (before the E update, time t).
 * c1/c2: split curl arrays indexed by component; single-curl solvers pass
 * c2 == NULL and get curl = curl1 - curl2 folded into c1. */
static void pw_apply(lxo_sim *s, fld *c1, fld *c2, int h_update, int t) {
    const int three = !s->planar;
    for (int a = 0; a < 3; ++a) {
        if (!three && a == 1) continue;
        int b0 = (a + 1) % 3, b1 = (a + 2) % 3;   /* transverse axes */
        /* transverse ranges must exist (start != -1) */
        if (s->st[b0] == -1 && (three || b0 != 1)) continue;
        if (s->st[b1] == -1 && (three || b1 != 1)) continue;
        for (int side = 0; side < 2; ++side) {
            int idx, present;
            double coord;
            if (h_update) {
                idx = side == 0 ? s->lo[a] - 1 : s->hi[a];
                coord = side == 0 ? (double)s->lo[a] : (double)s->hi[a];
            } else {
                idx = side == 0 ? s->lo[a] : s->hi[a];
                coord = side == 0 ? s->lo[a] - 0.5 : s->hi[a] + 0.5;
            }
            present = idx >= 0 && idx <= s->N[a] - 1;
            if (!present) continue;
            for (int f = 0; f < 2; ++f) {
                const int *fd = h_update ? HFACE[a][f] : EFACE[a][f];
                int tgt = fd[0], which = fd[1], src = fd[2];
                int tci = tgt >= 3 ? tgt - 3 : tgt;
                if (!three && tgt != LX_EY && tgt != LX_HX && tgt != LX_HZ) continue;
                if (!three && src != LX_EY && src != LX_HX && src != LX_HZ) continue;
                double dir = h_update ? s->Edir[src] : s->Hdir[src - 3];
                /* sign: 3D low face subtracts, high face adds; 2D injects the
                 * negated wave (maxwell.cc:523 vs :604, :563 vs :712). */
                double sgn = side == 0 ? -1.0 : 1.0;
                if (!three && h_update) sgn = -sgn;
                if (which == 2 && c2 == NULL) sgn = -sgn; /* curl = curl1 - curl2 */
                fld *arr = (which == 1 || c2 == NULL) ? &c1[tci] : &c2[tci];
                /* transverse loop ranges from the differentiated component's stagger */
                int lo0, hi0, lo1, hi1;
                int sb0 = STAG[src][b0], sb1 = STAG[src][b1];
                if (!three) {
                    /* 2D: the only transverse axis is the other of x/z; y is index 0 */
                    sb0 = (b0 == 1) ? 0 : STAG[src][b0];
                    sb1 = (b1 == 1) ? 0 : STAG[src][b1];
                }
                lo0 = (!three && b0 == 1) ? 0 : s->st[b0];
                hi0 = (!three && b0 == 1) ? 0 : s->en[b0] - sb0;
                lo1 = (!three && b1 == 1) ? 0 : s->st[b1];
                hi1 = (!three && b1 == 1) ? 0 : s->en[b1] - sb1;
                for (int u = lo0; u <= hi0; ++u) {
                    for (int v = lo1; v <= hi1; ++v) {
                        int ijk[3];
                        double cc[3];
                        ijk[a] = idx; cc[a] = coord;
                        ijk[b0] = u; cc[b0] = (sb0 && !(!three && b0 == 1)) ? u + 0.5 : (double)u;
                        ijk[b1] = v; cc[b1] = (sb1 && !(!three && b1 == 1)) ? v + 0.5 : (double)v;
                        double tp = h_update ? (t - 0.5) : (double)t;
                        double g = g_inc(s, cc[0], cc[1], cc[2], tp, h_update);
                        double val;
                        if (h_update) {
                            val = g * dir / s->d[a];
                        } else if (three) {
                            val = g * dir / s->d[a] / s->c / s->mu0;
                        } else {
                            /* 2D E terms carry an explicit (-1) (maxwell.cc:563-588) */
                            val = g * (-1) * dir / s->d[a] / s->c / s->mu0;
                        }
                        store_t *q = &AT(*arr, ijk[0], ijk[1], ijk[2]);
                        if (sgn < 0) *q = *q - val;
                        else *q = *q + val;
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* aux-1D incident line (extension)                                         */
/* ---------------------------------------------------------------------- */

static double aux_source(const lxo_sim *s, int t) {
    return waveform(s->p.pw_wave, t * s->dt, s->c, s->p.pw_lambda0, s->p.pw_fwhm, s->p.pw_delay);
}

static void aux_h_phase(lxo_sim *s) {
    double S = s->aux_S;
    for (int m = 0; m < s->aux_L - 1; ++m) s->aux_h[m] = s->aux_h[m] - S * (s->aux_e[m + 1] - s->aux_e[m]);
}

static void aux_e_phase(lxo_sim *s, int t) {
    double S = s->aux_S;
    for (int m = 1; m < s->aux_L - 1; ++m) s->aux_e[m] = s->aux_e[m] - S * (s->aux_h[m] - s->aux_h[m - 1]);
    s->aux_e[0] = aux_source(s, t + 1);
    s->aux_e[s->aux_L - 1] = 0.0;
}

int lxo_aux_line(lxo_sim *s, double *e, double *h, int n, int *x0) {
    if (s->p.incident != LX_INC_AUX1D || !s->aux_e) return -1;
    int L = s->aux_L < n ? s->aux_L : n;
    for (int m = 0; m < L; ++m) e[m] = s->aux_e[m];
    for (int m = 0; m < L - 1 && m < s->aux_L - 1; ++m) h[m] = s->aux_h[m];
    *x0 = s->aux_x0;
    return s->aux_L;
}

/* ---------------------------------------------------------------------- */
/* creation                                                                */
/* ---------------------------------------------------------------------- */

static void *xcalloc(size_t n) { return calloc(n ? n : 1, sizeof(double)); }

int lxo_create(const lx_params *p, int threads, lxo_sim **out) {
    lxo_sim *s = (lxo_sim *)calloc(1, sizeof(lxo_sim));
    if (!s) return -1;
    s->p = *p;
    s->N[0] = p->nx; s->N[1] = p->is_planar ? 1 : p->ny; s->N[2] = p->nz;
    s->planar = p->is_planar;
    s->pml = p->use_pml;
    s->threads = threads > 0 ? threads : 1;
    double sopt, smax;
    constants(p, &s->eps0, &s->mu0, &s->c, &s->dt, &sopt, &smax);
    s->sigma_opt = sopt; s->sigma_max = smax;
    s->eta0 = sqrt(s->mu0 / s->eps0);
    s->d[0] = p->dx; s->d[1] = p->dy; s->d[2] = p->dz;
    s->sigma_pw = p->pw_sigma;
    s->init = p->off_pw;
    if (p->use_pw) {
        for (int a = 0; a < 3; ++a) { s->kv[a] = p->k_pw[a]; s->Edir[a] = p->e_pw[a]; }
        /* H = normalize(k x E) (maxwell.cc:95-102) */
        double hx = s->kv[1] * s->Edir[2] - s->kv[2] * s->Edir[1];
        double hy = s->kv[2] * s->Edir[0] - s->kv[0] * s->Edir[2];
        double hz = s->kv[0] * s->Edir[1] - s->kv[1] * s->Edir[0];
        double ha = pow(pow(hx, 2) + pow(hy, 2) + pow(hz, 2), 0.5);
        s->Hdir[0] = hx / ha; s->Hdir[1] = hy / ha; s->Hdir[2] = hz / ha;
        for (int a = 0; a < 3; ++a) { s->lo[a] = p->tf_lo[a]; s->hi[a] = p->tf_hi[a]; }
        for (int a = 0; a < 3; ++a) {
            int Nend = s->N[a];
            if (s->planar && a == 1) { s->st[a] = -1; s->en[a] = -1; continue; }
            pw_clip(s->lo[a], s->hi[a], Nend, &s->st[a], &s->en[a]);
        }
    }
    int rc = 0;
    for (int c = 0; c < 3; ++c) {
        rc |= falloc_comp(s, &s->E[c], c);
        rc |= falloc_comp(s, &s->H[c], c + 3);
        rc |= falloc_comp(s, &s->Cb[c], c);
        rc |= falloc_comp(s, &s->c1E[c], c + 3);
        rc |= falloc_comp(s, &s->c1H[c], c);
        if (s->pml) {
            rc |= falloc_comp(s, &s->c2E[c], c + 3);
            rc |= falloc_comp(s, &s->c2H[c], c);
            rc |= falloc_comp(s, &s->pE1[c], c);
            rc |= falloc_comp(s, &s->pE2[c], c);
            rc |= falloc_comp(s, &s->pH1[c], c + 3);
            rc |= falloc_comp(s, &s->pH2[c], c + 3);
        }
    }
    if (rc) { lxo_destroy(s); return -1; }
    /* Cb = dt/eps0 everywhere (vacuum), Db = dt/mu0 */
    for (int c = 0; c < 3; ++c) {
        if (!s->Cb[c].v) continue;
        size_t n = (size_t)s->Cb[c].n[0] * s->Cb[c].n[1] * s->Cb[c].n[2];
        for (size_t q = 0; q < n; ++q) s->Cb[c].v[q] = s->dt / s->eps0;
    }
    s->Db = s->dt / s->mu0;
    if (s->pml) {
        for (int a = 0; a < 3; ++a) {
            if (s->planar && a == 1) continue;
            int n = s->N[a];
            s->ke[a] = xcalloc(n); s->be[a] = xcalloc(n); s->ce[a] = xcalloc(n);
            s->kh[a] = xcalloc(n); s->bh[a] = xcalloc(n); s->chh[a] = xcalloc(n);
            prof_kappa(p, a, 1, s->ke[a], n);
            prof_b(p, a, 1, s->be[a], n);
            prof_c(p, a, 1, s->ce[a], n);
            if (n > 1) {
                prof_kappa(p, a, 0, s->kh[a], n - 1);
                prof_b(p, a, 0, s->bh[a], n - 1);
                prof_c(p, a, 0, s->chh[a], n - 1);
            }
        }
    }
    if (p->use_pw && p->incident == LX_INC_AUX1D) {
        int ax = -1;
        for (int a = 0; a < 3; ++a)
            if (p->k_pw[a] != 0.0) { if (ax >= 0) { lxo_destroy(s); return -2; } ax = a; }
        if (ax < 0 || p->k_pw[ax] != 1.0 || p->e_pw[ax] != 0.0) { lxo_destroy(s); return -2; }
        s->aux_axis = ax;
        s->aux_S = s->c * s->dt / s->d[ax];
        s->aux_x0 = p->tf_lo[ax] - 2;
        int T = p->num_timesteps > 0 ? p->num_timesteps : 1;
        s->aux_L = (p->tf_hi[ax] + 2 - s->aux_x0) + (int)ceil(s->aux_S * T / 2.0) + 8;
        s->aux_e = xcalloc(s->aux_L);
        s->aux_h = xcalloc(s->aux_L);
        s->aux_e[0] = aux_source(s, 0);
    }
    s->t = 0;
    *out = s;
    return 0;
}

static void ffree(fld *f) { free(f->v); f->v = NULL; }

void lxo_destroy(lxo_sim *s) {
    if (!s) return;
    for (int c = 0; c < 3; ++c) {
        ffree(&s->E[c]); ffree(&s->H[c]); ffree(&s->Cb[c]);
        ffree(&s->c1E[c]); ffree(&s->c2E[c]); ffree(&s->c1H[c]); ffree(&s->c2H[c]);
        ffree(&s->pE1[c]); ffree(&s->pE2[c]); ffree(&s->pH1[c]); ffree(&s->pH2[c]);
        free(s->ke[c]); free(s->kh[c]); free(s->be[c]); free(s->bh[c]); free(s->ce[c]); free(s->chh[c]);
    }
    free(s->aux_e); free(s->aux_h);
    free(s->probe); free(s->rec);
    free(s);
}

/* Per-component Cb = dt/(eps0*eps_r) (maxwell_three_dim_pml.cc:121-128):
 * component c at array index (i,j,k) takes eps3[3*ids[i,j,k] + c]. */
int lxo_set_materials3(lxo_sim *s, const uint8_t *ids, const double *eps3, int n_ids) {
    int Ny = s->N[1], Nz = s->N[2];
    for (int c = 0; c < 3; ++c) {
        fld *f = &s->Cb[c];
        if (!f->v) continue;
        for (int i = 0; i < f->n[0]; ++i)
            for (int j = 0; j < f->n[1]; ++j)
                for (int k = 0; k < f->n[2]; ++k) {
                    int id = ids[((size_t)i * Ny + j) * Nz + k];
                    if (id >= n_ids) return -1;
                    AT(*f, i, j, k) = s->dt / (s->eps0 * eps3[3 * id + c]);
                }
    }
    return 0;
}

int lxo_set_materials(lxo_sim *s, const uint8_t *ids, const double *eps_r, int n_ids) {
    if (n_ids < 1) return -1;
    double *eps3 = (double *)malloc(sizeof(double) * 3 * (size_t)n_ids);
    if (!eps3) return -1;
    for (int q = 0; q < n_ids; ++q) eps3[3 * q] = eps3[3 * q + 1] = eps3[3 * q + 2] = eps_r[q];
    int rc = lxo_set_materials3(s, ids, eps3, n_ids);
    free(eps3);
    return rc;
}

int lxo_set_probes(lxo_sim *s, const int32_t *cijk, int n) {
    free(s->probe); free(s->rec);
    s->probe = NULL; s->rec = NULL; s->rec_steps = 0; s->rec_cap = 0; s->nprobe = 0;
    if (n <= 0) return 0;
    s->probe = (int32_t *)malloc(sizeof(int32_t) * 4 * n);
    memcpy(s->probe, cijk, sizeof(int32_t) * 4 * n);
    for (int q = 0; q < n; ++q) {
        int c = cijk[4 * q], ext[3];
        if (c < 0 || c > 5 || !comp_ext(s->N, s->planar, c, ext)) return -1;
        for (int a = 0; a < 3; ++a)
            if (cijk[4 * q + 1 + a] < 0 || cijk[4 * q + 1 + a] >= ext[a]) return -1;
    }
    s->nprobe = n;
    return 0;
}

static void record_probes(lxo_sim *s) {
    if (!s->nprobe) return;
    if (s->rec_steps == s->rec_cap) {
        s->rec_cap = s->rec_cap ? 2 * s->rec_cap : 256;
        s->rec = (double *)realloc(s->rec, sizeof(double) * s->rec_cap * s->nprobe);
    }
    for (int q = 0; q < s->nprobe; ++q) {
        int c = s->probe[4 * q];
        fld *f = c < 3 ? &s->E[c] : &s->H[c - 3];
        s->rec[(size_t)s->rec_steps * s->nprobe + q] = AT(*f, s->probe[4 * q + 1], s->probe[4 * q + 2], s->probe[4 * q + 3]);
    }
    s->rec_steps++;
}

/* ---------------------------------------------------------------------- */
/* point source (maxwell_three_dim_pml.cc:161-165; extension: soft source)  */
/* ---------------------------------------------------------------------- */

static void point_source(lxo_sim *s, int t) {
    if (!s->p.use_ps) return;
    int c = s->p.src_comp;
    if (c < 0 || c > 2 || !s->E[c].v) return;
    int i = s->p.src_pos[0], j = s->planar ? 0 : s->p.src_pos[1], k = s->p.src_pos[2];
    if (s->p.src_kind == LX_SRC_HARD_CW) {
        /* int amp_source, int omega_source (maxwell.h:203,206); omega*t is an int product */
        int amp = (int)s->p.src_amp, om = (int)s->p.src_omega;
        int prod = (int)((unsigned)om * (unsigned)t);
        AT(s->E[c], i, j, k) = amp * sin(prod * s->dt);
    } else {
        double w = waveform(s->p.src_wave, t * s->dt, s->c, s->p.src_lambda0, s->p.src_fwhm, s->p.src_delay);
        AT(s->E[c], i, j, k) = AT(s->E[c], i, j, k) + s->p.src_amp * w;
    }
}

/* ---------------------------------------------------------------------- */
/* 3D steps                                                                 */
/* ---------------------------------------------------------------------- */

/* curl_3d_adv of E into H-extent split curls (vfield.cc:212-239). */
static void curlE_split(lxo_sim *s) {
    int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    double dx = s->d[0], dy = s->d[1], dz = s->d[2];
    fld *E = s->E, *c1 = s->c1E, *c2 = s->c2E;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny - 1; j++)
            for (int k = 0; k < Nz - 1; k++) {
                AT(c1[0], i, j, k) = (AT(E[2], i, j + 1, k) - AT(E[2], i, j, k)) / dy;
                AT(c2[0], i, j, k) = (AT(E[1], i, j, k + 1) - AT(E[1], i, j, k)) / dz;
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx - 1; i++)
            for (int k = 0; k < Nz - 1; k++) {
                AT(c1[1], i, j, k) = (AT(E[0], i, j, k + 1) - AT(E[0], i, j, k)) / dz;
                AT(c2[1], i, j, k) = (AT(E[2], i + 1, j, k) - AT(E[2], i, j, k)) / dx;
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny - 1; j++)
            for (int i = 0; i < Nx - 1; i++) {
                AT(c1[2], i, j, k) = (AT(E[1], i + 1, j, k) - AT(E[1], i, j, k)) / dx;
                AT(c2[2], i, j, k) = (AT(E[0], i, j + 1, k) - AT(E[0], i, j, k)) / dy;
            }
}

/* curl_3d_adv of H into E-extent split curls; interior only (vfield.cc:242-268). */
static void curlH_split(lxo_sim *s) {
    int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    double dx = s->d[0], dy = s->d[1], dz = s->d[2];
    fld *H = s->H, *c1 = s->c1H, *c2 = s->c2H;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int j = 1; j < Ny - 1; j++)
            for (int k = 1; k < Nz - 1; k++) {
                AT(c1[0], i, j, k) = (AT(H[2], i, j, k) - AT(H[2], i, j - 1, k)) / dy;
                AT(c2[0], i, j, k) = (AT(H[1], i, j, k) - AT(H[1], i, j, k - 1)) / dz;
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int j = 0; j < Ny - 1; j++)
        for (int i = 1; i < Nx - 1; i++)
            for (int k = 1; k < Nz - 1; k++) {
                AT(c1[1], i, j, k) = (AT(H[0], i, j, k) - AT(H[0], i, j, k - 1)) / dz;
                AT(c2[1], i, j, k) = (AT(H[2], i, j, k) - AT(H[2], i - 1, j, k)) / dx;
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int k = 0; k < Nz - 1; k++)
        for (int j = 1; j < Ny - 1; j++)
            for (int i = 1; i < Nx - 1; i++) {
                AT(c1[2], i, j, k) = (AT(H[1], i, j, k) - AT(H[1], i - 1, j, k)) / dx;
                AT(c2[2], i, j, k) = (AT(H[0], i, j, k) - AT(H[0], i, j - 1, k)) / dy;
            }
}

/* curl_3d (single curl) of E (vfield.cc:117-141). */
static void curlE_single(lxo_sim *s) {
    int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    double dx = s->d[0], dy = s->d[1], dz = s->d[2];
    fld *E = s->E, *cu = s->c1E;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny - 1; j++)
            for (int k = 0; k < Nz - 1; k++)
                AT(cu[0], i, j, k) = (AT(E[2], i, j + 1, k) - AT(E[2], i, j, k)) / dy - (AT(E[1], i, j, k + 1) - AT(E[1], i, j, k)) / dz;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx - 1; i++)
            for (int k = 0; k < Nz - 1; k++)
                AT(cu[1], i, j, k) = (AT(E[0], i, j, k + 1) - AT(E[0], i, j, k)) / dz - (AT(E[2], i + 1, j, k) - AT(E[2], i, j, k)) / dx;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny - 1; j++)
            for (int i = 0; i < Nx - 1; i++)
                AT(cu[2], i, j, k) = (AT(E[1], i + 1, j, k) - AT(E[1], i, j, k)) / dx - (AT(E[0], i, j + 1, k) - AT(E[0], i, j, k)) / dy;
}

/* curl_3d (single curl) of H, interior only (vfield.cc:144-167). */
static void curlH_single(lxo_sim *s) {
    int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    double dx = s->d[0], dy = s->d[1], dz = s->d[2];
    fld *H = s->H, *cu = s->c1H;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int j = 1; j < Ny - 1; j++)
            for (int k = 1; k < Nz - 1; k++)
                AT(cu[0], i, j, k) = (AT(H[2], i, j, k) - AT(H[2], i, j - 1, k)) / dy - (AT(H[1], i, j, k) - AT(H[1], i, j, k - 1)) / dz;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int j = 0; j < Ny - 1; j++)
        for (int i = 1; i < Nx - 1; i++)
            for (int k = 1; k < Nz - 1; k++)
                AT(cu[1], i, j, k) = (AT(H[0], i, j, k) - AT(H[0], i, j, k - 1)) / dz - (AT(H[2], i, j, k) - AT(H[2], i - 1, j, k)) / dx;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int k = 0; k < Nz - 1; k++)
        for (int j = 1; j < Ny - 1; j++)
            for (int i = 1; i < Nx - 1; i++)
                AT(cu[2], i, j, k) = (AT(H[1], i, j, k) - AT(H[1], i - 1, j, k)) / dx - (AT(H[0], i, j, k) - AT(H[0], i, j - 1, k)) / dy;
}

/* psi_H and H update, three loops (maxwell_three_dim_pml.cc:182-213). */
static void update_H_pml(lxo_sim *s) {
    int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    const double Da = 1.0, Db = s->Db;
    fld *H = s->H, *c1 = s->c1E, *c2 = s->c2E, *p1 = s->pH1, *p2 = s->pH2;
    double *khx = s->kh[0], *khy = s->kh[1], *khz = s->kh[2];
    double *bhx = s->bh[0], *bhy = s->bh[1], *bhz = s->bh[2];
    double *chx = s->chh[0], *chy = s->chh[1], *chz = s->chh[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny - 1; j++)
            for (int k = 0; k < Nz - 1; k++) {
                AT(p1[0], i, j, k) = bhy[j] * AT(p1[0], i, j, k) + chy[j] * AT(c1[0], i, j, k);
                AT(p2[0], i, j, k) = bhz[k] * AT(p2[0], i, j, k) + chz[k] * AT(c2[0], i, j, k);
                AT(H[0], i, j, k) = Da * AT(H[0], i, j, k) - Db * (AT(c1[0], i, j, k) / khy[j] - AT(c2[0], i, j, k) / khz[k]) - Db * (AT(p1[0], i, j, k) - AT(p2[0], i, j, k));
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz - 1; k++) {
                AT(p1[1], i, j, k) = bhz[k] * AT(p1[1], i, j, k) + chz[k] * AT(c1[1], i, j, k);
                AT(p2[1], i, j, k) = bhx[i] * AT(p2[1], i, j, k) + chx[i] * AT(c2[1], i, j, k);
                AT(H[1], i, j, k) = Da * AT(H[1], i, j, k) - Db * (AT(c1[1], i, j, k) / khz[k] - AT(c2[1], i, j, k) / khx[i]) - Db * (AT(p1[1], i, j, k) - AT(p2[1], i, j, k));
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int j = 0; j < Ny - 1; j++)
            for (int k = 0; k < Nz; k++) {
                AT(p1[2], i, j, k) = bhx[i] * AT(p1[2], i, j, k) + chx[i] * AT(c1[2], i, j, k);
                AT(p2[2], i, j, k) = bhy[j] * AT(p2[2], i, j, k) + chy[j] * AT(c2[2], i, j, k);
                AT(H[2], i, j, k) = Da * AT(H[2], i, j, k) - Db * (AT(c1[2], i, j, k) / khx[i] - AT(c2[2], i, j, k) / khy[j]) - Db * (AT(p1[2], i, j, k) - AT(p2[2], i, j, k));
            }
}

/* psi_E and E update, three loops (maxwell_three_dim_pml.cc:224-255). */
static void update_E_pml(lxo_sim *s) {
    int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    const double Ca = 1.0;
    fld *E = s->E, *c1 = s->c1H, *c2 = s->c2H, *p1 = s->pE1, *p2 = s->pE2, *Cb = s->Cb;
    double *kex = s->ke[0], *key = s->ke[1], *kez = s->ke[2];
    double *bex = s->be[0], *bey = s->be[1], *bez = s->be[2];
    double *cex = s->ce[0], *cey = s->ce[1], *cez = s->ce[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                AT(p1[0], i, j, k) = bey[j] * AT(p1[0], i, j, k) + cey[j] * AT(c1[0], i, j, k);
                AT(p2[0], i, j, k) = bez[k] * AT(p2[0], i, j, k) + cez[k] * AT(c2[0], i, j, k);
                AT(E[0], i, j, k) = Ca * AT(E[0], i, j, k) + AT(Cb[0], i, j, k) * (AT(c1[0], i, j, k) / key[j] - AT(c2[0], i, j, k) / kez[k]) + AT(Cb[0], i, j, k) * (AT(p1[0], i, j, k) - AT(p2[0], i, j, k));
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny - 1; j++)
            for (int k = 0; k < Nz; k++) {
                AT(p1[1], i, j, k) = bez[k] * AT(p1[1], i, j, k) + cez[k] * AT(c1[1], i, j, k);
                AT(p2[1], i, j, k) = bex[i] * AT(p2[1], i, j, k) + cex[i] * AT(c2[1], i, j, k);
                AT(E[1], i, j, k) = Ca * AT(E[1], i, j, k) + AT(Cb[1], i, j, k) * (AT(c1[1], i, j, k) / kez[k] - AT(c2[1], i, j, k) / kex[i]) + AT(Cb[1], i, j, k) * (AT(p1[1], i, j, k) - AT(p2[1], i, j, k));
            }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz - 1; k++) {
                AT(p1[2], i, j, k) = bex[i] * AT(p1[2], i, j, k) + cex[i] * AT(c1[2], i, j, k);
                AT(p2[2], i, j, k) = bey[j] * AT(p2[2], i, j, k) + cey[j] * AT(c2[2], i, j, k);
                AT(E[2], i, j, k) = Ca * AT(E[2], i, j, k) + AT(Cb[2], i, j, k) * (AT(c1[2], i, j, k) / kex[i] - AT(c2[2], i, j, k) / key[j]) + AT(Cb[2], i, j, k) * (AT(p1[2], i, j, k) - AT(p2[2], i, j, k));
            }
}

/* Plain 3D H/E updates (maxwell_three_dim.cc:107-165); Cb per cell (vacuum
 * gives dt/eps0 everywhere, the reference's scalar). */
static void update_H_plain(lxo_sim *s) {
    const double Da = 1.0, Db = s->Db;
    for (int c = 0; c < 3; ++c) {
        fld *H = &s->H[c], *cu = &s->c1E[c];
        int n0 = H->n[0], n1 = H->n[1], n2 = H->n[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
        for (int i = 0; i < n0; i++)
            for (int j = 0; j < n1; j++)
                for (int k = 0; k < n2; k++) AT(*H, i, j, k) = Da * AT(*H, i, j, k) - Db * AT(*cu, i, j, k);
    }
}

static void update_E_plain(lxo_sim *s) {
    const double Ca = 1.0;
    for (int c = 0; c < 3; ++c) {
        fld *E = &s->E[c], *cu = &s->c1H[c], *Cb = &s->Cb[c];
        int n0 = E->n[0], n1 = E->n[1], n2 = E->n[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
        for (int i = 0; i < n0; i++)
            for (int j = 0; j < n1; j++)
                for (int k = 0; k < n2; k++) AT(*E, i, j, k) = Ca * AT(*E, i, j, k) + AT(*Cb, i, j, k) * AT(*cu, i, j, k);
    }
}

/* ---------------------------------------------------------------------- */
/* 2D steps (x-z plane, Ey Hx Hz)                                           */
/* ---------------------------------------------------------------------- */

/* curl_planar_adv of Ey (vfield.cc:360-372) */
static void curlE_2d_split(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    double dx = s->d[0], dz = s->d[2];
    fld *Ey = &s->E[1];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int k = 0; k < Nz - 1; k++) AT(s->c2E[0], i, 0, k) = (AT(*Ey, i, 0, k + 1) - AT(*Ey, i, 0, k)) / dz;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int k = 0; k < Nz; k++) AT(s->c1E[2], i, 0, k) = (AT(*Ey, i + 1, 0, k) - AT(*Ey, i, 0, k)) / dx;
}

/* curl_planar_adv of (Hx,Hz) (vfield.cc:374-381) */
static void curlH_2d_split(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    double dx = s->d[0], dz = s->d[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 1; i < Nx - 1; i++)
        for (int k = 1; k < Nz - 1; k++) {
            AT(s->c1H[1], i, 0, k) = (AT(s->H[0], i, 0, k) - AT(s->H[0], i, 0, k - 1)) / dz;
            AT(s->c2H[1], i, 0, k) = (AT(s->H[2], i, 0, k) - AT(s->H[2], i - 1, 0, k)) / dx;
        }
}

/* curl_planar (single) of Ey (vfield.cc:303-315) */
static void curlE_2d_single(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    double dx = s->d[0], dz = s->d[2];
    fld *Ey = &s->E[1];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int k = 0; k < Nz - 1; k++) AT(s->c1E[0], i, 0, k) = -(AT(*Ey, i, 0, k + 1) - AT(*Ey, i, 0, k)) / dz;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int k = 0; k < Nz; k++) AT(s->c1E[2], i, 0, k) = (AT(*Ey, i + 1, 0, k) - AT(*Ey, i, 0, k)) / dx;
}

/* curl_planar (single) of (Hx,Hz) (vfield.cc:317-323) */
static void curlH_2d_single(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    double dx = s->d[0], dz = s->d[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 1; i < Nx - 1; i++)
        for (int k = 1; k < Nz - 1; k++)
            AT(s->c1H[1], i, 0, k) = (AT(s->H[0], i, 0, k) - AT(s->H[0], i, 0, k - 1)) / dz - (AT(s->H[2], i, 0, k) - AT(s->H[2], i - 1, 0, k)) / dx;
}

/* 2D CPML H loops (maxwell_two_dim_pml.cc:149-162) */
static void update_H_2d_pml(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    const double Da = 1.0, Db = s->Db;
    double *khx = s->kh[0], *khz = s->kh[2], *bhx = s->bh[0], *bhz = s->bh[2], *chx = s->chh[0], *chz = s->chh[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int k = 0; k < Nz - 1; k++) {
            AT(s->pH2[0], i, 0, k) = bhz[k] * AT(s->pH2[0], i, 0, k) + chz[k] * AT(s->c2E[0], i, 0, k);
            AT(s->H[0], i, 0, k) = Da * AT(s->H[0], i, 0, k) - Db * (-AT(s->c2E[0], i, 0, k) / khz[k] - AT(s->pH2[0], i, 0, k));
        }
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int k = 0; k < Nz; k++) {
            AT(s->pH1[2], i, 0, k) = bhx[i] * AT(s->pH1[2], i, 0, k) + chx[i] * AT(s->c1E[2], i, 0, k);
            AT(s->H[2], i, 0, k) = Da * AT(s->H[2], i, 0, k) - Db * (AT(s->c1E[2], i, 0, k) / khx[i] + AT(s->pH1[2], i, 0, k));
        }
}

/* 2D CPML E loop (maxwell_two_dim_pml.cc:169-176) */
static void update_E_2d_pml(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    const double Ca = 1.0;
    double *kex = s->ke[0], *kez = s->ke[2], *bex = s->be[0], *bez = s->be[2], *cex = s->ce[0], *cez = s->ce[2];
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int k = 0; k < Nz - 1; k++) {
            AT(s->pE1[1], i, 0, k) = bez[k] * AT(s->pE1[1], i, 0, k) + cez[k] * AT(s->c1H[1], i, 0, k);
            AT(s->pE2[1], i, 0, k) = bex[i] * AT(s->pE2[1], i, 0, k) + cex[i] * AT(s->c2H[1], i, 0, k);
            AT(s->E[1], i, 0, k) = Ca * AT(s->E[1], i, 0, k) + AT(s->Cb[1], i, 0, k) * (AT(s->c1H[1], i, 0, k) / kez[k] - AT(s->c2H[1], i, 0, k) / kex[i] + AT(s->pE1[1], i, 0, k) - AT(s->pE2[1], i, 0, k));
        }
}

/* 2D plain loops (maxwell_two_dim.cc:100-123) */
static void update_H_2d_plain(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    const double Da = 1.0, Db = s->Db;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx; i++)
        for (int k = 0; k < Nz - 1; k++) AT(s->H[0], i, 0, k) = Da * AT(s->H[0], i, 0, k) - Db * AT(s->c1E[0], i, 0, k);
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int k = 0; k < Nz; k++) AT(s->H[2], i, 0, k) = Da * AT(s->H[2], i, 0, k) - Db * AT(s->c1E[2], i, 0, k);
}

static void update_E_2d_plain(lxo_sim *s) {
    int Nx = s->N[0], Nz = s->N[2];
    const double Ca = 1.0;
#pragma omp parallel for num_threads(s->threads) schedule(dynamic)
    for (int i = 0; i < Nx - 1; i++)
        for (int k = 0; k < Nz - 1; k++) AT(s->E[1], i, 0, k) = Ca * AT(s->E[1], i, 0, k) + AT(s->Cb[1], i, 0, k) * AT(s->c1H[1], i, 0, k);
}

/* ---------------------------------------------------------------------- */
/* time loop                                                               */
/* ---------------------------------------------------------------------- */

int lxo_step(lxo_sim *s, int nsteps) {
    const int aux = s->p.use_pw && s->p.incident == LX_INC_AUX1D;
    for (int n = 0; n < nsteps; ++n) {
        int t = s->t;
        point_source(s, t);
        if (s->planar) {
            if (s->pml) {
                curlE_2d_split(s);
                if (s->p.use_pw) pw_apply(s, s->c1E, s->c2E, 1, t);
                update_H_2d_pml(s);
                if (aux) aux_h_phase(s);
                curlH_2d_split(s);
                if (s->p.use_pw) pw_apply(s, s->c1H, s->c2H, 0, t);
                update_E_2d_pml(s);
            } else {
                curlE_2d_single(s);
                if (s->p.use_pw) pw_apply(s, s->c1E, NULL, 1, t);
                update_H_2d_plain(s);
                if (aux) aux_h_phase(s);
                curlH_2d_single(s);
                if (s->p.use_pw) pw_apply(s, s->c1H, NULL, 0, t);
                update_E_2d_plain(s);
            }
        } else {
            if (s->pml) {
                curlE_split(s);
                if (s->p.use_pw) pw_apply(s, s->c1E, s->c2E, 1, t);
                update_H_pml(s);
                if (aux) aux_h_phase(s);
                curlH_split(s);
                if (s->p.use_pw) pw_apply(s, s->c1H, s->c2H, 0, t);
                update_E_pml(s);
            } else {
                curlE_single(s);
                if (s->p.use_pw) pw_apply(s, s->c1E, NULL, 1, t);
                update_H_plain(s);
                if (aux) aux_h_phase(s);
                curlH_single(s);
                if (s->p.use_pw) pw_apply(s, s->c1H, NULL, 0, t);
                update_E_plain(s);
            }
        }
        if (aux) aux_e_phase(s, t);
        record_probes(s);
        s->t++;
    }
    return 0;
}

int lxo_read_field(lxo_sim *s, int comp, double *dst, size_t n) {
    if (comp < 0 || comp > 5) return -1;
    fld *f = comp < 3 ? &s->E[comp] : &s->H[comp - 3];
    if (!f->v) return -1;
    size_t m = (size_t)f->n[0] * f->n[1] * f->n[2];
    if (n != m) return -1;
    for (size_t q = 0; q < m; ++q) dst[q] = f->v[q];
    return 0;
}

int lxo_set_field(lxo_sim *s, int comp, const double *src, size_t n) {
    if (comp < 0 || comp > 5) return -1;
    fld *f = comp < 3 ? &s->E[comp] : &s->H[comp - 3];
    if (!f->v) return -1;
    size_t m = (size_t)f->n[0] * f->n[1] * f->n[2];
    if (n != m) return -1;
    for (size_t q = 0; q < m; ++q) f->v[q] = (store_t)src[q];
    return 0;
}

int lxo_read_probes(lxo_sim *s, double *dst, size_t n, int *n_steps) {
    size_t m = (size_t)s->rec_steps * s->nprobe;
    if (n < m) return -1;
    if (m) memcpy(dst, s->rec, sizeof(double) * m);
    *n_steps = s->rec_steps;
    return 0;
}

double lxo_max_e2(lxo_sim *s) {
    double mx = 0;
    if (s->planar) {
        fld *f = &s->E[1];
        size_t m = (size_t)f->n[0] * f->n[1] * f->n[2];
        for (size_t q = 0; q < m; ++q) mx = f->v[q] * f->v[q] > mx ? f->v[q] * f->v[q] : mx;
        return mx;
    }
    for (int i = 0; i < s->N[0] - 1; ++i)
        for (int j = 0; j < s->N[1] - 1; ++j)
            for (int k = 0; k < s->N[2] - 1; ++k) {
                double v = AT(s->E[2], i, j, k) * AT(s->E[2], i, j, k) + AT(s->E[1], i, j, k) * AT(s->E[1], i, j, k) +
                           AT(s->E[0], i, j, k) * AT(s->E[0], i, j, k);
                if (v > mx) mx = v;
            }
    return mx;
}

double lxo_energy(lxo_sim *s) {
    double we = 0, wh = 0, vol = s->d[0] * s->d[2] * (s->planar ? 1.0 : s->d[1]);
    for (int c = 0; c < 3; ++c) {
        if (s->E[c].v) {
            size_t m = (size_t)s->E[c].n[0] * s->E[c].n[1] * s->E[c].n[2];
            for (size_t q = 0; q < m; ++q) we += s->E[c].v[q] * s->E[c].v[q] * (s->dt / s->Cb[c].v[q]);
        }
        if (s->H[c].v) {
            size_t m = (size_t)s->H[c].n[0] * s->H[c].n[1] * s->H[c].n[2];
            for (size_t q = 0; q < m; ++q) wh += s->H[c].v[q] * s->H[c].v[q] * s->mu0;
        }
    }
    return 0.5 * (we + wh) * vol;
}
