// Shared definitions of the soap_turbo HIP library (host and device side).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#define ST_MAX_ALPHA 11     // poly3gauss stops above 11 (src/soap_turbo_radial.f90:210-219)
#define ST_MAX_LMAX 12      // angular kernels are instantiated for l_max <= 12
#define ST_MAX_SPECIES 32   // species masks are packed into 32-bit words
#define ST_REC 20           // recursion slots: alpha_max + 6 (second derivatives), rounded up

namespace st {

// Per-species constants of the radial expansion, computed on the host in FP64
// with the operations of radial_expansion_range (src/soap_turbo_radial.f90:436-540).
struct RadialConsts {
  int alpha_max, n_poly, n_hi, off, n_max;
  int with_gauss, do_central, has_filter, unified, amp_mode, radial_enh, do_der, do_der2;
  uint32_t species_bit;
  double rcut_hard_in, rcut_soft, atom_sigma, dr, inv_rcut, sqrt_rcut, inv_sqrt_rcut, inv_Nm2, sigma_sq0, N_gauss;
  double q, drnf, scaling, amp_scaling, central_weight, global_scaling;
  double ra[ST_REC], rb[ST_REC], rc[ST_REC], R1[ST_REC], R2[ST_REC], R3[ST_REC], R4[ST_REC];
  double W[ST_MAX_ALPHA * ST_MAX_ALPHA];  // species block of W, column-major alpha x alpha
  // poly3operator (op = 1): 2 sqrt(2 ln 2), unscaled sigma and rcut_soft for the region tests, the
  // basis-side moments A (alpha x 7, row-major) and the filter's cubic at 0 (g_aux(0, rs, fw))
  int op;
  double op_K, op_sigma_in, op_rcut_soft_in;
  double op_A[ST_MAX_ALPHA * 7];
  double op_cf[4];
};

// Per-species constants of the angular expansion (src/soap_turbo_angular.f90:548-605).
struct AngularConsts {
  int n_species, l_max, k_max, do_der;
  double rcut_max, rcut2;
  double sigma_t[ST_MAX_SPECIES], sigma_t_scaling[ST_MAX_SPECIES];
};

void overlap_matrix(int basis, int a, double atom_sigma_r, double rcut_hard, double* S);
bool orthonormalization_matrix(int a, const double* S, double* W);
// the reference's LAPACK sequence on the host's LAPACK (false and msg if unavailable or failing)
bool orthonormalization_matrix_lapack(int a, const double* S, double* W, std::string& msg);

}  // namespace st
