// Launchers of the HIP kernels (host-callable, stream-ordered).
#pragma once
#include <hip/hip_runtime.h>

#include "st_internal.h"

namespace st {

// FP64 path (kernels_f64.hip)
void launch_pack_pairs(hipStream_t st, int n_pairs, int n_species, const int32_t* mask, const int* site_ptr,
                       int n_sites, uint32_t* spmask, uint8_t* first);
void launch_fill(hipStream_t st, double* x, size_t n, double v);
// rad_der2 non-null: also the second derivative (the constants must be built for it)
bool launch_radial(hipStream_t st, const RadialConsts& c, int n_pairs, const double* rjs, const uint32_t* spmask,
                   const uint8_t* first, double* rad, double* rad_der, double* rad_der2 = nullptr);
bool launch_angular(hipStream_t st, const AngularConsts& c, int n_pairs, const double* rjs, const double* th,
                    const double* ph, const uint32_t* spmask, const double* preflm, double2* a, double2* ar,
                    double2* aa, double2* ap);
void launch_cnk(hipStream_t st, int n_sites, int n_max, int k_max, const int* site_ptr, const double* rad,
                const double2* ang, const int32_t* site_species, int max_mult, const int32_t* species_mult,
                const double* central, double2* cnk);
void launch_power(hipStream_t st, int n_sites, int n_max, int k_max, int n_soap, const int4* comp, const int* out_ptr,
                  const int* out_c, const double* out_el, const int4* ocomp, const double* oel, const double2* cnk,
                  double* soap, double* sqrt_dot_p);
void launch_normalise(hipStream_t st, int n_sites, int n_soap, const double* sqrt_dot_p, double* soap);
// launch variants of the Jacobian for n_soap outputs: the batched kernel's shapes (threads x outputs
// per thread; variant 0, or -1, is the default), then the wave-per-pair kernel when a compiled
// shape covers n_soap
int derivatives_variants(int n_soap);
size_t derivatives_lds(int n_max, int l_max, int k_max, int n_soap, int variant = -1);
bool launch_derivatives(hipStream_t st, int n_sites, int n_pairs, int n_max, int l_max, int k_max, int n_soap,
                        const int* site_ptr, const int4* comp, const int* out_ptr, const int* out_c,
                        const double* out_el, const int4* ocomp, const double* oel,
                        const double2* cnk, const double* rad, const double* rad_der, const double2* ang,
                        const double2* ang_r, const double2* ang_a, const double2* ang_p, const double* soap,
                        const double* sqrt_dot_p, const double* rjs, const double* thetas, const double* phis,
                        double* rotp, double* cart, int variant = -1);
void launch_contract_cart(hipStream_t st, int n_sites, int n_pairs, int n_soap, const int* site_ptr, const double* G,
                          int ldg, const double* cart, double* pair);
size_t contract_lds(int n_max, int k_max, int n_soap, int n_soap_unc);
void launch_contract(hipStream_t st, int n_sites, int n_max, int l_max, int k_max, int n_soap, int n_soap_unc,
                     const int* site_ptr, const uint8_t* skip, const int* unc_ptr, const int* unc_e,
                     const int* umap, const double* uel, const int32_t* P_i, const double* P_el, const double2* cnk, const double* rad,
                     const double* rad_der, const double2* ang, const double2* ang_r, const double2* ang_a,
                     const double2* ang_p, const double* soap, const double* G, double sgn, const double* sqrt_dot_p,
                     const double* rjs, const double* thetas, const double* phis, double* pair);

// Device neighbour lists (neighbours.hip)
// The periodic cell of a neighbour list: a cube of side L, or (general) the lattice vectors h[3c + v]
// (component c of vector v) with h^-1 in hi[3v + c]; n: cells of the cell list along each vector.
struct Lattice {
  double L;
  double h[9], hi[9];
  int n[3];
  int general;
};
Lattice cubic_lattice(double L, double rcut);
// false for a left-handed or degenerate cell; min_width: the smallest distance between opposite faces
bool general_lattice(const double* cell, double rcut, Lattice& g, double& min_width);
void launch_neighbour_bins(hipStream_t st, int n_atoms, const double* pos, const Lattice& g, int* cell_of,
                           int* cell_count, int* cell_start, int* cursor, int* cell_atoms);
void launch_neighbour_count(hipStream_t st, int n_sites, const int* site_atom, const double* pos, const Lattice& g,
                            double rcut, const int* cell_start, const int* cell_atoms, int* n_neigh, int* site_ptr);
void launch_neighbour_fill(hipStream_t st, int n_sites, int n_pairs, int n_species, const int* site_atom,
                           const int* species, const double* pos, const Lattice& g, double rcut,
                           const int* cell_start, const int* cell_atoms, const int* site_ptr, int* pair_atom,
                           double* vec, double* rjs, double* thetas, double* phis, int32_t* mask,
                           int32_t* site_species, int32_t* mult, int* overflow);
void launch_atom_pairs(hipStream_t st, int n_atoms, int n_pairs, const int* pair_atom, int* count, int* cursor,
                       int* atom_ptr, int* atom_pairs, int* overflow);

// central-atom Hessian (hessian.hip): scratch doubles per site, LDS bytes, launch
size_t hessian_scratch_per_site(int n_max, int k_max, int n_soap, int n_unc, int n_vec);
size_t hessian_lds(int n_max, int k_max, int n_unc);
size_t hessian_res_per_pair(int n_vec);  // doubles of per-pair contraction sums
// doubles into a site's scratch slot of its per-site scalars: |p|-rule dot products, then per weight
// vector v (w.p, w.d1 p (3)) at 3 + 4 v
size_t hessian_scal_offset(int n_max, int k_max, int n_soap, int n_unc, int n_vec);
size_t hessian_geom_head();  // doubles of a pair's geometry head (its first is 0 for an inactive pair)
size_t hessian_heads_per_pair();  // doubles of per-pair geometry scratch
bool launch_hessian(hipStream_t st, int n_sites, int n_pairs, int n_max, int l_max, int n_soap, int n_unc, int n_vec,
                    int compressed, int n_species, const double* sigma_t, const double* sigma_t_scaling,
                    double rcut_max, const int* site_ptr, const double* rjs, const double* thetas, const double* phis,
                    const uint32_t* spmask, const uint8_t* first, const double* Ck, const double* rad,
                    const double* rad_der, const double* rad_der2, const double2* cnk, const double* soap,
                    const double* sdp, const double* vecs, const int4* comp, const uint8_t* skip, const int* out_ptr,
                    const int* out_c, const double* out_el, const int* unc_ptr, const int* unc_e, const int32_t* P_i,
                    const double* P_el, double* scratch, double* heads, double* res, double* hess);

}  // namespace st
