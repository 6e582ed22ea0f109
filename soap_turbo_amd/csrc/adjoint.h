// FP32 adjoint path (descriptors + dE/dq -> pair gradients, forces, virial)
// and the force/virial reduction shared with the FP64 path.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/soap_turbo_hip.h"
#include "st_internal.h"

namespace st {

struct AdjBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

struct AdjointState {
  // tables of the radial and angular radial factors (rebuilt when the key changes)
  std::string table_key;
  AdjBuf rtab, atab, tabinfo;
  int n_seg = 0;
  int rt_off[8] = {}, at_off[8] = {};  // per species: where its tables start (equal tables shared)
  // per-call work and outputs
  AdjBuf soap, sdp, pair, forces, virial_part, virial, G32, order, tickets;
  bool order_ok = false;                 // order holds the sites of the current neighbour list
  int g32_sites = 0, g32_soap = 0;  // shape of the site-major FP32 dE/dq in G32 (0: none)
  AdjBuf el32, odir;
  AdjBuf cenv;                  // per-site central-atom term (poly3gauss)
  AdjBuf gapK;                  // GAP kernel values / their weights, n_sites x n_sparse
  bool cen_ok = false;          // cenv holds the sites of the current neighbour list ...
  std::string cen_key;          // ... for the tables of this key
  std::string atoms_key, el_key;
  int n_pf = 0;
  int n_pairs = 0, n_atoms = 0, n_sites = 0, n_soap = 0;
  bool have_out = false, have_forces = false;
  std::map<std::string, int> tuned;  // site kernel: workgroups per CU chosen per shape (0: occupancy)
};

struct AdjointInputs {
  const st_params* p;
  const std::vector<double>* W;
  const std::vector<double>* central;  // [species][n_max] FP64 central-atom term (poly3gauss)
  int n_sites, n_pairs, n_max, k_max, l_max, n_soap, n_soap_unc;
  const std::vector<int>* site_ptr_h;
  const int* site_ptr;
  const int32_t* species;
  int max_mult;
  const int32_t* mult;
  const uint32_t* spmask;
  const uint8_t* first;
  const double *rjs, *thetas, *phis, *G;
  const int4* comp;
  const int* out_ptr;   // null without compression
  const int* out_c;
  const int* unc_ptr;
  const int* unc_e;
  const int32_t* P_i;
  const std::vector<double>* out_el_h;  // compression elements by output (host)
  const std::vector<double>* P_el_h;    // compression elements in P order (host)
  const std::string* tables_key;
  const int4* ocomp;                     // direct maps (null unless single-entry compression or none)
  const int* umap;
  const std::vector<double>* oel_h;
  const std::vector<double>* uel_h;
  const std::vector<int32_t>* umap_h;
  std::vector<RadialConsts> radial_tab;  // per species, block offset 0, n_max = alpha_max, with derivatives
  const int32_t* pair_atom;
  const int* atom_ptr;                   // pairs of each atom (CSR over pair_atom)
  const int* atom_pairs;
  const double* vec;
  int n_atoms;
  size_t lds_max;                        // LDS bytes one workgroup may allocate
  int n_cu;                              // compute units of the device
  bool autotune = false;                 // choose the site kernel's grid by timing (first call per shape)
  bool forward_only = false;             // descriptors only (q, |p|): no backward pass, no forces
};

using Timer = std::function<void(const char*, const std::function<void()>&)>;

int adjoint_run(AdjointState& s, hipStream_t st, const AdjointInputs& in, const Timer& timer, std::string& msg);
// dE/dq G(n_sites, n_soap) (FP64, site fastest) -> the site-major FP32 copy the site kernel reads;
// run when dE/dq is uploaded.
int adjoint_set_dedq(AdjointState& s, hipStream_t st, int n_sites, int n_soap, const double* G, std::string& msg);
// the site-major FP32 dE/dq buffer of n_sites x n_soap, to be filled on the device (GAP gradient)
int adjoint_prepare_dedq(AdjointState& s, int n_sites, int n_soap, std::string& msg);
int adjoint_download_pairs(AdjointState& s, hipStream_t st, int n_pairs, double* pair, std::string& msg);
int adjoint_download_forces(AdjointState& s, hipStream_t st, int n_atoms, double* forces, double* virial,
                            std::string& msg);
int adjoint_download_soap(AdjointState& s, hipStream_t st, int n_sites, int n_soap, double* soap, std::string& msg);
// GAP-form energy of the FP32 descriptors in s.soap, E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta
// over sparse points q_t (T: n_sparse x n_soap, FP32, row t contiguous), and its gradient dE/dq
// written over s.G32 (site-major FP32, as the adjoint pass reads it); site energies in e_site, the
// total in *energy_dev (FP64, device). Two FP32 GEMMs (rocBLAS) and an elementwise kernel.
int gap_energy_grad(AdjointState& s, hipStream_t st, int n_sites, int n_soap, int n_sparse, const float* T,
                    const float* alpha, int zeta, double delta, double* e_site, double* energy_dev, std::string& msg);
// forces(:, a) = -sum of pair(:, k) over pairs with neighbour atom a (atom_ptr/atom_pairs: the
// pairs of each atom); virial = -sum_k vec_k (x) pair_k, reduced through per-workgroup partials
// (vpart holds gather_blocks(n_atoms) * 9 doubles).
inline int gather_blocks(int n_atoms) { return n_atoms > 0 ? (n_atoms + 15) / 16 : 1; }
// workgroups of the FP32 path's virial sum over the pairs (at most gather_blocks(n_atoms), so vpart fits)
inline int virial_blocks(int n_atoms, int n_pairs) {
  const int nb = (n_pairs + 255) / 256;
  return nb < 1 ? 1 : (nb > 1024 ? 1024 : (nb > gather_blocks(n_atoms) ? gather_blocks(n_atoms) : nb));
}
void launch_forces_f64(hipStream_t st, int n_atoms, const double* pair, const int* atom_ptr, const int* atom_pairs,
                       const double* vec, const double* rjs, const double* thetas, const double* phis,
                       double* forces, double* vpart, double* virial);
// (reset: a counter the last kernel of the step sets back to zero, or null)
void launch_forces_f32(hipStream_t st, int n_atoms, int n_pairs, const float* pair, const int* atom_ptr,
                       const int* atom_pairs, const double* vec, const double* rjs, const double* thetas,
                       const double* phis, double* forces, double* vpart, double* virial,
                       unsigned long long* reset);
// sets *counter to zero (a kernel, so that it can be part of a captured graph)
void launch_counter_reset(hipStream_t st, unsigned long long* counter);
void adjoint_invalidate(AdjointState& s);
void adjoint_release(AdjointState& s);

}  // namespace st
