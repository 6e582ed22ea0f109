/* C-ABI of the MI355X (gfx950) soap_turbo get_soap library, libsoapturbo_hip.so.
 *
 * Plain pointers and sizes; arrays are the reference's Fortran column-major
 * arrays (first index fastest). Every entry point returns 0 on success or one
 * of the ST_ERR_* codes; st_error() then gives the message, which for the
 * conditions the reference stops on is the reference's own message.
 *
 * Entry points and the reference interfaces they replace:
 *   st_get_soap             get_soap, src/soap_turbo.f90:139-169 (argument list, output layout and
 *                           caller contract: soap and soap_cart_der zero on entry)
 *   st_soap_contract        soap_contract, src/soap_turbo.f90:1723-1730
 *   st_get_compress_indices get_compress_indices, src/soap_turbo_compress.f90:44-298 ("trivial" and the
 *                           nu_R_nu_S modes "0_0" .. "2_2")
 *   st_set_basis            the save'd W/S cache of get_soap, src/soap_turbo.f90:294-388
 *                           (lets a caller inject the W the reference would have built)
 *   st_get_basis, st_set_basis_route, st_basis_matrices
 *                           get_orthonormalization_matrix_poly3[gauss], src/soap_turbo_radial.f90:1023, :1113
 *   st_upload_neighbours, st_compute, st_download, st_upload_dedq, st_pairs_full_jacobian,
 *   st_adjoint_f32, st_download_pairs, st_download_forces
 *                           the same get_soap / soap_contract split into upload, device-resident
 *                           compute and download, for callers that keep their data on the GPU
 *                           (get_soap's persistent work buffers, src/soap_turbo.f90:43-61, are the
 *                           reference's counterpart of the resident state)
 *   st_adjoint_f32          no reference counterpart: dE/dq -> pair gradients, forces and virial in FP32
 *                           without forming soap_cart_der (the FP32 form of soap_contract, :1703-1721)
 *   st_build_neighbours, st_build_neighbours_cell
 *                           no reference counterpart: the callers' neighbour list, built on the device
 *                           (cubic box; general lattice vectors)
 *   st_central_hessian      get_soap_central_hessian, src/soap_turbo.f90:1151-1700 (the caller calls
 *                           it after its get_soap; here it runs that pass itself)
 *   st_set_gap_model, st_gap_energy_forces, st_gap_site_energies
 *                           no reference counterpart: a GAP-form energy model on the descriptors
 *                           (what get_soap's callers feed back as dE/dq), evaluated on the device
 *
 * Threading: one context per host thread; calls on a context are ordered on
 * its HIP stream. Like the reference, a context is not reentrant.
 */
#ifndef SOAP_TURBO_HIP_H
#define SOAP_TURBO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct st_ctx st_ctx;

enum {
  ST_OK = 0,
  ST_ERR_ARG = 1,          /* invalid argument or size */
  ST_ERR_HIP = 2,          /* HIP runtime failure */
  ST_ERR_UNSUPPORTED = 3,  /* a case this build does not implement (l_max > 12, the poly3tab basis, ...) */
  ST_ERR_STATE = 4,        /* call out of order, e.g. soap_contract before get_soap(contract=.true.) */
  ST_ERR_STOP = 5          /* a condition the reference stops on (alpha_max limits, ...) */
};

enum { ST_BASIS_POLY3 = 0, ST_BASIS_POLY3GAUSS = 1, ST_BASIS_POLY3OPERATOR = 2 };
enum { ST_SCALING_POLYNOMIAL = 0, ST_SCALING_OTHER = 1 };

/* Descriptor hyper-parameters (get_soap arguments; one entry per species for the arrays). */
typedef struct {
  int32_t n_species;
  int32_t l_max;
  int32_t radial_enhancement;
  int32_t basis;               /* ST_BASIS_* */
  int32_t scaling_mode;        /* ST_SCALING_* */
  int32_t legacy_filter_seed;  /* soap_turbo_radial's legacy_filter_seed, src/soap_turbo_radial.f90:67 */
  const int32_t* alpha_max;
  const double* rcut_hard;
  const double* rcut_soft;
  const double* nf;
  const double* global_scaling;
  const double* atom_sigma_r;
  const double* atom_sigma_r_scaling;
  const double* atom_sigma_t;
  const double* atom_sigma_t_scaling;
  const double* amplitude_scaling;
  const double* central_weight;
} st_params;

/* The neighbour list (get_soap arguments n_sites .. phis). Host pointers. */
typedef struct {
  int32_t n_sites;
  int32_t n_pairs;
  int32_t n_species;                   /* columns of mask */
  int32_t max_mult;                    /* leading dimension of species */
  const int32_t* n_neigh;              /* n_neigh(n_sites), the site itself first */
  const int32_t* species;              /* species(max_mult, n_sites), 1-based */
  const int32_t* species_multiplicity; /* species_multiplicity(n_sites) */
  const int32_t* mask;                 /* mask(n_pairs, n_species), nonzero = .true. */
  const double* rjs;
  const double* thetas;
  const double* phis;
} st_neighbours;

/* Compression (compress_P_nonzero, compress_P_i, compress_P_j, compress_P_el); P_nonzero = 0: none. */
typedef struct {
  int32_t P_nonzero;
  const int32_t* P_i;
  const int32_t* P_j;
  const double* P_el;
} st_compression;

int st_create(st_ctx** ctx, int device);
void st_destroy(st_ctx* ctx);
const char* st_error(const st_ctx* ctx);

/* n_soap that get_soap will produce for these parameters (maxval(P_i) when compressed). */
int st_n_soap(const st_params* p, const st_compression* c);

/* W and S (n_max x n_max, block-diagonal, column-major) the library uses for p. */
int st_get_basis(st_ctx* ctx, const st_params* p, double* W, double* S);
/* Use these W and S instead of building them, while n_max matches; n_max = 0 clears. */
int st_set_basis(st_ctx* ctx, int32_t n_max, const double* W, const double* S);
/* How the context builds W = S^{-1/2} (get_orthonormalization_matrix_poly3[gauss],
 * src/soap_turbo_radial.f90:1023-1219):
 *   ST_BASIS_ROUTE_EXTENDED (the default): an eigendecomposition in x87 extended precision, accurate to
 *     about 1e-9 at cond(S) = 1e13 (poly3, alpha_max 8);
 *   ST_BASIS_ROUTE_LAPACK: the reference's own sequence, dgesvd -> U sqrt(Sigma) V^T (dgemm) ->
 *     dpotrf/dpotri, on the host's LAPACK (ST_LAPACK names the library, else the MKL runtime or a
 *     system liblapack is loaded): with the LAPACK the reference was linked with, the reference's W. */
enum { ST_BASIS_ROUTE_EXTENDED = 0, ST_BASIS_ROUTE_LAPACK = 1 };
int st_set_basis_route(st_ctx* ctx, int32_t route);
/* W and S of a route for p, on the host without a device context. */
int st_basis_matrices(int32_t route, const st_params* p, double* W, double* S);

/* get_soap: soap(n_soap, n_sites) and, with do_derivatives and not contract,
 * soap_cart_der(3, n_soap, n_pairs). soap_cart_der may be NULL otherwise. */
int st_get_soap(st_ctx* ctx, const st_params* p, const st_neighbours* nb, const st_compression* c,
                int32_t do_derivatives, int32_t contract, int32_t n_soap, double* soap, double* soap_cart_der);

/* soap_contract: pair(3, n_pairs) = sgn * sum_s G(i,s) soap_cart_der(:,s,k), G(n_sites, n_soap). */
int st_soap_contract(st_ctx* ctx, const double* soap, const double* G, double sgn, double* pair);

/* get_compress_indices (mode "trivial"/"Trivial"): set_indices = 0 only fills dim and P_nonzero. */
int st_get_compress_indices(const char* mode, int32_t n_species, const int32_t* alpha_max, int32_t l_max,
                            int32_t* dim, int32_t* P_nonzero, int32_t* P_i, int32_t* P_j, double* P_el,
                            int32_t set_indices);

/* ---- device-resident use (data uploaded once, compute repeated, results read back on demand) ---- */
int st_upload_neighbours(st_ctx* ctx, const st_neighbours* nb);
/* pair_atom(n_pairs): 0-based atom index of each pair's neighbour (the site's own atom for its first
 * pair), n_atoms atoms; needed only for forces. vec(3, n_pairs): r_j - r_i, needed only for the virial
 * (may be NULL: it is rebuilt from rjs, thetas, phis). */
int st_upload_atoms(st_ctx* ctx, int32_t n_atoms, const int32_t* pair_atom, const double* vec);
/* Device neighbour list of a periodic cubic box (no reference counterpart: the reference's callers
 * build it). Fills the same resident arrays as st_upload_neighbours + st_upload_atoms, in the order
 * get_soap expects (src/soap_turbo.f90:139-169): per site its own atom first (r = 0), then every other
 * atom j with |r_j - r_i| < rcut under the minimum image in increasing j; thetas = acos(z/r),
 * phis = atan2(y, x); mask(k, s) = (species of the neighbour == s), the mask semantics of
 * assign_species_multiplicity (src/soap_turbo.f90:2188-2214); one species per site. positions(3, n_atoms),
 * atom_species(n_atoms) 1-based, box side > 2 rcut, site_atoms(n_sites) 0-based. *n_pairs gets the pair
 * count (may be NULL). */
int st_build_neighbours(st_ctx* ctx, int32_t n_atoms, const double* positions, const int32_t* atom_species,
                        int32_t n_species, double box, double rcut, int32_t n_sites, const int32_t* site_atoms,
                        int32_t* n_pairs);
/* The same for a general periodic cell (no reference counterpart): cell(3, 3) holds the lattice vectors as
 * columns (cell(:, v), a right-handed cell); every width of the cell (the distance between opposite
 * faces) must exceed 2 rcut, so that each neighbour has one image inside the cutoff. The pair vector is
 * d - H rint(H^-1 d) for d = r_j - r_i, H the lattice vectors as columns. */
int st_build_neighbours_cell(st_ctx* ctx, int32_t n_atoms, const double* positions, const int32_t* atom_species,
                             int32_t n_species, const double* cell, double rcut, int32_t n_sites,
                             const int32_t* site_atoms, int32_t* n_pairs);
/* The resident neighbour list: n_neigh(n_sites), pair_atom(n_pairs), rjs, thetas, phis(n_pairs),
 * mask(n_pairs, n_species) int32, vec(3, n_pairs); any may be NULL. */
int st_download_neighbours(st_ctx* ctx, int32_t* n_neigh, int32_t* pair_atom, double* rjs, double* thetas,
                           double* phis, int32_t* mask, double* vec);
/* The pairs of each atom the force gather reads: atom_ptr(n_atoms + 1), atom_pairs(n_pairs). */
int st_atom_pairs(st_ctx* ctx, int32_t* atom_ptr, int32_t* atom_pairs);
/* dE/dq as G(n_sites, n_soap) */
int st_upload_dedq(st_ctx* ctx, int32_t n_soap, const double* G);
/* FP64 get_soap on the resident neighbour list. */
int st_compute(st_ctx* ctx, const st_params* p, const st_compression* c, int32_t do_derivatives, int32_t contract);
/* FP64 pair gradients from the resident soap_cart_der and dE/dq (what a full-Jacobian caller does). */
int st_pairs_full_jacobian(st_ctx* ctx);
/* Atoms st_pairs_full_jacobian_noalloc accepts: its scratch is a fixed 9 doubles per 16 atoms. */
#define ST_NOALLOC_MAX_ATOMS (1 << 21)
/* No reference counterpart: st_pairs_full_jacobian for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc). Every array is the caller's,
 * in device memory: site_ptr(n_sites + 1) with site_ptr(1) = 0, G(n_sites, n_soap), soap_cart_der(3,
 * n_soap, n_pairs), atom_ptr(n_atoms + 1) and atom_pairs(n_pairs) (the pairs of each atom, as
 * st_atom_pairs gives them), vec(3, n_pairs) in; pair(3, n_pairs), forces(3, n_atoms), virial(3, 3) out.
 * The call allocates nothing: the force gather's partial sums live in a static device array of
 * ST_NOALLOC_MAX_ATOMS / 16 * 9 doubles (9 MiB, held by each device the library runs on until the
 * process ends); larger n_atoms are refused (ST_ERR_ARG). Launches are ordered on the context's
 * stream and the call returns before they finish. All calls on one device share that array, so
 * calls that overlap on one device (other contexts' streams, other threads) race. */
int st_pairs_full_jacobian_noalloc(st_ctx* ctx, int32_t n_sites, int32_t n_pairs, int32_t n_soap, int32_t n_atoms,
                                   const int32_t* site_ptr, const double* G, const double* soap_cart_der,
                                   const int32_t* atom_ptr, const int32_t* atom_pairs, const double* vec, double* pair,
                                   double* forces, double* virial);
/* The FP64 full-Jacobian path in batches of sites: for each batch of consecutive sites, get_soap with
 * soap_cart_der on that batch, then its contraction with dE/dq to pair gradients; forces and virial at
 * the end (st_download_pairs / st_download_forces with which = 0). This is what the reference's callers
 * do when soap_cart_der of all sites exceeds memory: they call get_soap once per batch of sites, its
 * work buffers sized for one batch (src/soap_turbo.f90:43-57). max_pairs bounds the pairs of a batch
 * (0: as many as the free device memory holds); batches are as even as that bound allows.
 * *n_batches (may be NULL) gets the batch count. The normalised soap of every site is resident
 * afterwards (st_download); soap_cart_der is downloadable only when one batch held every pair. */
int st_full_jacobian_forces(st_ctx* ctx, const st_params* p, const st_compression* c, int64_t max_pairs,
                            int32_t* n_batches);
/* FP64 soap_contract on the resident state and dE/dq (normalised soap from st_compute). */
int st_contract_resident(st_ctx* ctx, double sgn);
/* FP32 adjoint: descriptors, pair gradients, forces and virial from the resident data and dE/dq. */
int st_adjoint_f32(st_ctx* ctx, const st_params* p, const st_compression* c);
/* Graph mode for st_adjoint_f32 (off by default; no reference counterpart): the second call with the
 * same parameters, sizes and resident buffers records the step's kernel launches in a HIP graph and
 * later calls replay it (new dE/dq or coordinates of the same sizes are read by the replay). Calls
 * made with kernel timing on run uncaptured. */
int st_set_graphs(st_ctx* ctx, int32_t on);
/* Launch-shape autotuning (off by default; no reference counterpart): the first call per shape of
 * st_compute's Jacobian kernel (threads x outputs per thread) and of st_adjoint_f32's site kernel
 * (workgroups per CU of its persistent grid) times each candidate on the current data and keeps the
 * fastest for later calls. Results do not depend on the choice beyond summation order. */
int st_set_autotune(st_ctx* ctx, int32_t on);
/* The Jacobian kernel's launch variants for n_soap outputs (the batched kernel's shapes, then the
 * wave-per-pair kernel where it applies), and a fixed choice among them (-1: the default, or the
 * tuned one); for tests and tuning studies. */
int st_jacobian_variants(st_ctx* ctx, int32_t n_soap);
int st_set_jacobian_variant(st_ctx* ctx, int32_t variant);
int st_download(st_ctx* ctx, double* soap, double* soap_cart_der);
/* pair(3, n_pairs) of the last st_pairs_full_jacobian / st_contract_resident (FP64) or st_adjoint_f32. */
int st_download_pairs(st_ctx* ctx, int32_t which, double* pair);   /* which: 0 FP64, 1 FP32 */
/* forces(3, n_atoms) = -sum over pairs with that neighbour atom; virial(3,3) = -sum_k vec_k (x) pair_k. */
int st_download_forces(st_ctx* ctx, int32_t which, double* forces, double* virial);
/* soap(n_soap, n_sites) of the last st_adjoint_f32, widened to FP64. */
int st_download_soap_f32(st_ctx* ctx, double* soap);

/* get_soap_central_hessian (src/soap_turbo.f90:1151-1700) on the resident neighbour list, FP64:
 * runs the descriptor pass itself (st_compute without derivatives, plus the radial second
 * derivatives), then hess(3, 3, n_vec, n_pairs) with hess(a, b, v, k) =
 * sum_d vecs(d, v, i) d^2 q_d / dr_ia dr_jb for pair k = (site i, neighbour j); the self pair's
 * block is d^2 / dr_ia dr_ib. vecs(n_soap, n_vec, n_sites) are weight vectors in the space of the
 * (normalised, compressed) descriptor. Needs legacy_filter_seed = 0, as the reference does. */
int st_central_hessian(st_ctx* ctx, const st_params* p, const st_compression* c, int32_t n_vec, const double* vecs,
                       double* hess);
/* soap_hessian_selftest of get_soap_central_hessian (src/soap_turbo.f90:1419-1451, :1498-1499): after
 * st_central_hessian with these vecs on the current list, the first derivatives that the Hessian's
 * contraction also produces (v.dq/dr_jb for each caller's vector v and non-self pair, from its
 * Omega/Q_l^m route) against the same quantity from the independent FP64 Jacobian (soap_cart_der,
 * computed here); *err = max |difference| / max |Jacobian value|. Off the normal path: it recomputes
 * soap_cart_der of the whole list and compares on the host. */
int st_hessian_selftest(st_ctx* ctx, const st_params* p, const st_compression* c, int32_t n_vec, const double* vecs,
                        double* err);

/* ---- GAP energy model (no reference counterpart: the callers' energy model that supplies get_soap
 * with dE/dq, here on the device) ----
 * E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta over the descriptors q_i of the sites and the
 * model's sparse points q_t, sparse_q(n_soap, n_sparse), alpha(n_sparse); zeta >= 1. Stored in FP32
 * on the device. */
int st_set_gap_model(st_ctx* ctx, int32_t n_sparse, int32_t n_soap, const double* sparse_q, const double* alpha,
                     int32_t zeta, double delta);
/* One energy-and-forces step on the resident neighbour list: FP32 descriptors, E and dE/dq on the
 * device, then the adjoint pass (pair gradients, forces and virial as after st_adjoint_f32, read with
 * st_download_forces(ctx, 1, ...)). energy may be NULL. The resident dE/dq is the model's afterwards:
 * upload one again before the next st_adjoint_f32. */
int st_gap_energy_forces(st_ctx* ctx, const st_params* p, const st_compression* c, double* energy);
/* e_site(n_sites): the site energies of the last st_gap_energy_forces. */
int st_gap_site_energies(st_ctx* ctx, double* e_site);

int st_synchronize(st_ctx* ctx);
/* Stream-ordered timing: record event slot (0..15) on the context's stream; elapsed ms between two. */
int st_event_record(st_ctx* ctx, int32_t slot);
int st_event_elapsed(st_ctx* ctx, int32_t a, int32_t b, float* ms);
/* Per-kernel timing with events around each launch (off by default). */
int st_kernel_timing(st_ctx* ctx, int32_t on);
int st_kernel_time(st_ctx* ctx, const char* name, double* total_ms, int32_t* launches);

#ifdef __cplusplus
}
#endif
#endif
