/*
 * lx_oracle.h -- FP64 CPU restatement of Luxum's Yee/CPML/TF-SF timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 * The product (luxum_amd/, include/luxum.h) never links or calls it.
 */
#ifndef LX_ORACLE_H
#define LX_ORACLE_H
#include "../include/luxum.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lxo_sim lxo_sim;

/* Profile kinds for lxo_profile. */
enum { LXO_SIGMA = 0, LXO_KAPPA = 1, LXO_ALPHA = 2, LXO_B = 3, LXO_C = 4 };

int lxo_create(const lx_params *p, int threads, lxo_sim **out);
void lxo_destroy(lxo_sim *s);
int lxo_set_materials(lxo_sim *s, const uint8_t *ids, const double *eps_r, int n_ids);
/* Per-component permittivities: E component c takes eps3[3*id + c]. */
int lxo_set_materials3(lxo_sim *s, const uint8_t *ids, const double *eps3, int n_ids);
int lxo_set_probes(lxo_sim *s, const int32_t *comp_ijk, int n);
int lxo_step(lxo_sim *s, int nsteps);
int lxo_read_field(lxo_sim *s, int comp, double *dst, size_t n);
/* Overwrite a component (reference layout, SI units); test set-up only. */
int lxo_set_field(lxo_sim *s, int comp, const double *src, size_t n);
int lxo_read_probes(lxo_sim *s, double *dst, size_t n, int *n_steps);
int lxo_field_size(const lx_params *p, int comp, int ext[3]);
/* 1D CPML profile along `dim` at E nodes (is_E=1, n = N) or H nodes (n = N-1). */
int lxo_profile(const lx_params *p, int dim, int is_E, int kind, double *out, int n);
/* Reference constants: out[0..5] = eps0, mu0, c, dt, sigma_opt, sigma_max. */
void lxo_constants(const lx_params *p, double out[6]);
/* Monitor of the reference: max over [0,Nx-1)x[0,Ny-1)x[0,Nz-1) of Ex^2+Ey^2+Ez^2 (3D), max Ey^2 (2D). */
double lxo_max_e2(lxo_sim *s);
/* Discrete electromagnetic energy sum(eps*E^2 + mu0*H^2)/2 * cell volume. */
double lxo_energy(lxo_sim *s);
/* Aux-1D incident line (E at integer nodes, eta0*H at half nodes) for tests. */
int lxo_aux_line(lxo_sim *s, double *e, double *h, int n, int *x0);

#ifdef __cplusplus
}
#endif
#endif
