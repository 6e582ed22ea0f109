/* FP64 CPU restatement of the soap_turbo get_soap hot path -- TEST ORACLE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * this library, and only as the checker; the product never links it.
 *
 * Layouts are the reference's Fortran column-major arrays, indexed from 0:
 *   mask[k + s*n_pairs]                      mask(n_pairs, n_species)
 *   W[n + m*n_max], S[...]                   W(n_max, n_max), block-diagonal
 *   rad[k*n_max + n]                         radial_exp_coeff(n_max, n_pairs)
 *   ang[2*(k*k_max + kk) + {0,1}]            angular_exp_coeff(k_max, n_pairs), complex
 *   soap[i*n_soap + s]                       soap(n_soap, n_sites)
 *   cart[(k*n_soap + s)*3 + c]               soap_cart_der(3, n_soap, n_pairs)
 *   G[s*n_sites + i]                         G(n_sites, n_soap)
 *   pair[k*3 + c]                            pair(3, n_pairs)
 */
#ifndef SOAP_ORACLE_H
#define SOAP_ORACLE_H
#include <stdint.h>

typedef struct {
  int n_species;
  int l_max;
  int radial_enhancement;
  int basis;          /* 0 poly3, 1 poly3gauss, 2 poly3operator (polynomial scaling only) */
  int scaling_poly;   /* 1: scaling_mode == "polynomial" */
  int legacy_seed;    /* legacy_filter_seed */
  const int32_t* alpha_max;
  const double *rcut_hard, *rcut_soft, *nf, *global_scaling, *atom_sigma_r, *atom_sigma_r_scaling,
      *atom_sigma_t, *atom_sigma_t_scaling, *amplitude_scaling, *central_weight;
} so_params;

typedef struct {
  int n_sites;
  int n_pairs;
  const int32_t* n_neigh;
  const int32_t* species;          /* species(1, n_sites), 1-based */
  const int32_t* species_mult;     /* species_multiplicity(n_sites) */
  const int32_t* mask;
  const double *rjs, *thetas, *phis;
} so_inputs;

typedef struct {
  int compress;
  int P_nonzero;
  const int32_t *P_i, *P_j; /* 1-based */
  const double* P_el;
} so_compress;

int so_k_max(int l_max);
int so_n_soap(const so_params* p, const so_compress* c);
void so_preflm(int l_max, double* preflm);
void so_overlap_S(int basis, int alpha_max, double atom_sigma_r, double rcut_hard, double* S);
/* poly3operator, one pair of species s, before W (unscaled rjs): values and r-derivatives */
void so_op_pair(const so_params* p, int s, double rjs_in, int is_central, double* outv, double* outd);
void so_radial(const so_params* p, const so_inputs* in, const double* W, int do_der, double* rad, double* rad_der);
void so_radial2(const so_params* p, const so_inputs* in, const double* W, int do_der, double* rad, double* rad_der,
                double* rad_der2);
/* get_soap_central_hessian (src/soap_turbo.f90:1151-1700) after get_soap with derivatives:
 * vecs (n_sites, n_vec, n_soap) C, hess (n_pairs, n_vec, 3 a, 3 b) C, soap (n_sites, n_soap)
 * normalised; rad_der2 (n_pairs, n_max) may be NULL. Returns 1 for legacy_seed (the reference
 * stops there). */
int so_central_hessian(const so_params* p, const so_inputs* in, const double* W, const double* S,
                       const so_compress* c, int n_vec, const double* vecs, double* soap, double* hess,
                       double* rad_der2, double* selftest_err);
void so_angular(const so_params* p, const so_inputs* in, int do_der, double* ang, double* ang_r, double* ang_a,
                double* ang_p);
/* soap (unnormalised before return? no: normalised, as get_soap returns it);
 * cart may be NULL (no derivatives); sqrt_dot_p may be NULL. */
int so_get_soap(const so_params* p, const so_inputs* in, const double* W, const double* S, const so_compress* c,
                int do_der, double* soap, double* cart, double* sqrt_dot_p);
int so_contract(const so_params* p, const so_inputs* in, const double* W, const double* S, const so_compress* c,
                const double* G, double sgn, double* soap, double* pair);
#endif
