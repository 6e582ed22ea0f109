// C-ABI of libsoapturbo_hip.so: device context, host-side set-up of the
// per-species constants and index tables, and the orchestration of the
// kernels on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>


#include "../../include/soap_turbo_hip.h"
#include "adjoint.h"
#include "kernels.h"
#include "st_internal.h"

// Scratch of st_pairs_full_jacobian_noalloc: the force gather's per-workgroup virial partials, 9 doubles
// per 16 atoms, for up to ST_NOALLOC_MAX_ATOMS atoms (9.0 MiB). Each device holds its own copy for the
// life of the process; host code uses the address hipGetSymbolAddress gives on that device (looked up
// at st_create), never the array's name.
static __device__ __align__(256) double scratch[(size_t)((ST_NOALLOC_MAX_ATOMS + 15) / 16) * 9];

namespace {

std::mutex scratch_mu;
std::vector<double*> scratch_dev;  // scratch's address on each device (null until looked up)

// scratch's address on `device` (the current device), looked up once per device.
double* scratch_on(int device) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  if ((int)scratch_dev.size() <= device) scratch_dev.resize(device + 1, nullptr);
  if (!scratch_dev[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    scratch_dev[device] = static_cast<double*>(p);
  }
  return scratch_dev[device];
}


struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct st_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t lds_max = 64 * 1024;  // LDS bytes one workgroup may allocate on this device
  int n_cu = 1;                // compute units
  double* scratch = nullptr;   // this device's copy of the no-alloc scratch array
  std::string err;

  // basis: injected or built (cache key = parameters that change W and S); route of W
  int basis_route = ST_BASIS_ROUTE_EXTENDED;
  int inj_n_max = 0;
  std::vector<double> injW, injS;
  std::string basis_key;
  std::vector<double> W, S;

  // resident neighbour list
  bool have_nb = false;
  int n_sites = 0, n_pairs = 0, n_species_mask = 0, max_mult = 1;
  int packed_species = 0;  // n_species the packed pair masks were made for (0: none since the upload)
  std::vector<int> site_ptr_h;
  DevBuf site_ptr, species, mult, mask, spmask, first, rjs, thetas, phis;
  // atoms (forces / virial)
  int n_atoms = 0;
  bool have_atoms = false, have_vec = false;
  DevBuf pair_atom, vec, atom_ptr, atom_pairs;
  // device neighbour-list builder: positions, atom species, sites and the cell list
  DevBuf nb_pos, nb_species, nb_sites, nb_cell_of, nb_cell_count, nb_cell_start, nb_cursor, nb_cell_atoms,
      nb_count, nb_flag;
  // dE/dq
  int G_n_soap = 0;
  bool have_G = false;
  DevBuf G;

  // FP64 work and outputs
  DevBuf rad, rad_der, ang, ang_r, ang_a, ang_p, cnk, soap, sdp, cart, pair64, preflm, central, rotp;
  DevBuf batch_ptr;  // per-batch site offsets of st_full_jacobian_forces
  DevBuf forces64, vpart64, virial64;
  bool have_forces64 = false;
  DevBuf comp, out_ptr, out_c, out_el, skip, unc_ptr, unc_e, P_i, P_el;
  // direct maps when every output has one P entry and every component at most one
  DevBuf ocomp, oel, umap, uel;
  bool single = false;
  std::vector<double> oel_h, uel_h;
  std::vector<int32_t> umap_h;
  std::string tables_key;
  std::vector<double> out_el_h, P_el_h;
  int n_max = 0, k_max = 0, l_max = -1, n_soap = 0, n_soap_unc = 0;
  bool compressed = false, have_soap = false, have_cart = false, ct_ready = false, have_pair64 = false;

  // FP32 path
  st::AdjointState adj;
  // captured launch sequence of the FP32 step (st_set_graphs): recorded on the second call with the
  // same key, replayed while the key holds
  bool graphs = false;
  uint64_t nb_gen = 0;  // bumped by every new neighbour list or atom upload
  // launch-shape autotuning (st_set_autotune): the fastest variant per shape key
  bool autotune = false;
  std::map<std::string, int> tuned;
  int jacobian_variant = -1;  // >= 0: the Jacobian kernel's launch variant, fixed (st_set_jacobian_variant)
  // GAP energy model (st_set_gap_model): sparse points and weights in FP32, site energies
  int gap_n_sparse = 0, gap_n_soap = 0, gap_zeta = 0;
  double gap_delta = 0.0;
  DevBuf gap_T, gap_alpha, gap_esite, gap_energy;
  // central-atom Hessian (st_central_hessian): second radial derivatives, C_k, weight vectors, scratch
  DevBuf rad_der2, hess_ck, hess_vecs, hess_scratch, hess_heads, hess_res, hess_out;
  int hess_n_vec = 0;      // n_vec of the last st_central_hessian on the current list (0: none)
  uint64_t hess_gen = 0;   // ... and the neighbour list it ran on
  std::string g32_key, g32_warm;
  hipGraph_t g32_graph = nullptr;
  hipGraphExec_t g32_exec = nullptr;

  // events
  hipEvent_t ev[16] = {};
  bool timing = false;
  struct Pending { std::string name; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<double, int>> ktime;
};

namespace {

int fail(st_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(ctx, ST_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int ensure(st_ctx* ctx, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return ST_OK;
  if (b.p) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(ctx, ST_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
  }
  b.bytes = bytes;
  return ST_OK;
}

int upload(st_ctx* ctx, DevBuf& b, const void* src, size_t bytes) {
  int rc = ensure(ctx, b, bytes);
  if (rc) return rc;
  if (bytes) HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return ST_OK;
}

// The fastest of n launch variants of one kernel, timed on the current data (a warm-up launch, then
// one timed with events, per variant), cached under key. launch(v) returns false for a variant that
// cannot run (it is skipped).
int tuned_variant(st_ctx* ctx, const std::string& key, int n, const std::function<bool(int)>& launch);

hipEvent_t get_event(st_ctx* ctx) {
  if (!ctx->free_events.empty()) {
    hipEvent_t e = ctx->free_events.back();
    ctx->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Runs f() on the stream, bracketed by events when per-kernel timing is on.
template <class F>
void timed(st_ctx* ctx, const char* name, F&& f) {
  if (!ctx->timing) {
    f();
    return;
  }
  hipEvent_t a = get_event(ctx), b = get_event(ctx);
  (void)hipEventRecord(a, ctx->stream);
  f();
  (void)hipEventRecord(b, ctx->stream);
  ctx->pending.push_back({name, a, b});
}

void resolve_timing(st_ctx* ctx) {
  if (ctx->pending.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& pd : ctx->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, pd.a, pd.b);
    auto& t = ctx->ktime[pd.name];
    t.first += ms;
    t.second += 1;
    ctx->free_events.push_back(pd.a);
    ctx->free_events.push_back(pd.b);
  }
  ctx->pending.clear();
}

int n_max_of(const st_params* p) {
  int n = 0;
  for (int s = 0; s < p->n_species; ++s) n += p->alpha_max[s];
  return n;
}

int k_max_of(int l) { return 1 + l * (l + 1) / 2 + l; }

// Argument checks, with the reference's stop conditions and warnings
// (src/soap_turbo_radial.f90:139-164, :205-230).
int check_params(st_ctx* ctx, const st_params* p) {
  if (!p || p->n_species < 1 || p->n_species > ST_MAX_SPECIES)
    return fail(ctx, ST_ERR_ARG, "n_species must be in 1..32");
  if (p->l_max < 0) return fail(ctx, ST_ERR_ARG, "l_max must be >= 0");
  if (p->l_max > ST_MAX_LMAX) return fail(ctx, ST_ERR_UNSUPPORTED, "l_max > 12 is not supported by this build");
  if (p->basis != ST_BASIS_POLY3 && p->basis != ST_BASIS_POLY3GAUSS && p->basis != ST_BASIS_POLY3OPERATOR)
    return fail(ctx, ST_ERR_UNSUPPORTED, "only the poly3, poly3gauss and poly3operator bases are supported");
  // the reference's poly3operator amplitude is defined for polynomial scaling only
  // (src/soap_turbo_radial.f90:1701-1727 leaves it unset otherwise)
  if (p->basis == ST_BASIS_POLY3OPERATOR && p->scaling_mode != ST_SCALING_POLYNOMIAL)
    return fail(ctx, ST_ERR_UNSUPPORTED, "the poly3operator basis needs scaling_mode polynomial");
  static bool warned_p3 = false, warned_p3g = false;
  for (int s = 0; s < p->n_species; ++s) {
    const int a = p->alpha_max[s];
    if (a < 1) return fail(ctx, ST_ERR_ARG, "alpha_max must be >= 1");
    if (p->basis == ST_BASIS_POLY3OPERATOR) {
      if (a > ST_MAX_ALPHA) return fail(ctx, ST_ERR_UNSUPPORTED, "alpha_max > 11 is not supported by this build");
    } else if (p->basis == ST_BASIS_POLY3) {
      if (a > 10)
        return fail(ctx, ST_ERR_STOP,
                    "ERROR: Due to numerical instabilities in the basis construction for the poly3 basis, it is "
                    "strongly recommended not to exceed alpha_max = 7. For alpha_max > 10 the instabilities are too "
                    "large to proceed.");
      if (a > 7 && !warned_p3) {
        warned_p3 = true;
        std::printf(" WARNING: Due to numerical instabilities in the basis construction for the poly3 basis, it is\n"
                    " strongly recommended not to exceed alpha_max = 7.\n");
      }
    } else {
      if (a > 11)
        return fail(ctx, ST_ERR_STOP,
                    "ERROR: Due to numerical instabilities in the basis construction for the poly3gauss basis, it is "
                    "strongly recommended not to exceed alpha_max = 8. For alpha_max > 11 the instabilities are too "
                    "large to proceed.");
      if (a > 8 && !warned_p3g) {
        warned_p3g = true;
        std::printf(" WARNING: Due to numerical instabilities in the basis construction for the poly3gauss basis, it "
                    "is\n strongly recommended not to exceed alpha_max = 8.\n");
      }
    }
  }
  return ST_OK;
}

// W and S for p: injected, cached, or built (block-diagonal, column-major).
int basis_for(st_ctx* ctx, const st_params* p) {
  const int n_max = n_max_of(p);
  if (ctx->inj_n_max) {
    if (ctx->inj_n_max != n_max) return fail(ctx, ST_ERR_ARG, "injected basis has a different n_max");
    ctx->W = ctx->injW;
    ctx->S = ctx->injS;
    ctx->basis_key = "injected";
    return ST_OK;
  }
  // the reference rebuilds W when the basis name or alpha_max change, and for
  // poly3gauss also atom_sigma_r and rcut_hard (src/soap_turbo.f90:294-354)
  std::string key(reinterpret_cast<const char*>(&p->basis), sizeof(int32_t));
  key.append(reinterpret_cast<const char*>(&ctx->basis_route), sizeof(int));
  key.append(reinterpret_cast<const char*>(p->alpha_max), sizeof(int32_t) * p->n_species);
  if (p->basis == ST_BASIS_POLY3GAUSS) {
    key.append(reinterpret_cast<const char*>(p->atom_sigma_r), sizeof(double) * p->n_species);
    key.append(reinterpret_cast<const char*>(p->rcut_hard), sizeof(double) * p->n_species);
  }
  if (key == ctx->basis_key && (int)ctx->W.size() == n_max * n_max) return ST_OK;
  ctx->W.assign((size_t)n_max * n_max, 0.0);
  ctx->S.assign((size_t)n_max * n_max, 0.0);
  int off = 0;
  for (int s = 0; s < p->n_species; ++s) {
    const int a = p->alpha_max[s];
    std::vector<double> Sb((size_t)a * a), Wb((size_t)a * a);
    st::overlap_matrix(p->basis, a, p->atom_sigma_r[s], p->rcut_hard[s], Sb.data());
    if (ctx->basis_route == ST_BASIS_ROUTE_LAPACK) {
      std::string msg;
      if (!st::orthonormalization_matrix_lapack(a, Sb.data(), Wb.data(), msg)) return fail(ctx, ST_ERR_STOP, msg);
    } else if (!st::orthonormalization_matrix(a, Sb.data(), Wb.data())) {
      return fail(ctx, ST_ERR_STOP, "the overlap matrix of the radial basis is not positive definite");
    }
    for (int j = 0; j < a; ++j)
      for (int i = 0; i < a; ++i) {
        ctx->W[(off + i) + (size_t)(off + j) * n_max] = Wb[i + j * a];
        ctx->S[(off + i) + (size_t)(off + j) * n_max] = Sb[i + j * a];
      }
    off += a;
  }
  ctx->basis_key = key;
  return ST_OK;
}

// Per-species radial constants: radial_expansion_range set-up,
// src/soap_turbo_radial.f90:436-540, computed with the same FP64 operations.
st::RadialConsts radial_consts(const st_params* p, int s, const std::vector<double>& W, int n_max, int off,
                               bool do_der, bool do_der2 = false) {
  st::RadialConsts c;
  std::memset(&c, 0, sizeof c);
  const int a = p->alpha_max[s];
  c.alpha_max = a;
  c.n_max = n_max;
  c.off = off;
  c.species_bit = 1u << s;
  c.with_gauss = p->basis == ST_BASIS_POLY3GAUSS;
  c.do_central = c.with_gauss ? 0 : (p->central_weight[s] != 0.0);
  c.unified = !p->legacy_filter_seed;
  c.radial_enh = p->radial_enhancement;
  c.do_der = do_der || do_der2;
  c.do_der2 = do_der2;
  c.scaling = p->atom_sigma_r_scaling[s];
  c.amp_scaling = p->amplitude_scaling[s];
  c.amp_mode = p->scaling_mode != ST_SCALING_POLYNOMIAL ? 0
               : c.amp_scaling == 0.0                    ? 1
               : c.amp_scaling == 1.0                    ? 2
                                                         : 3;
  c.central_weight = p->central_weight[s];
  c.global_scaling = p->global_scaling[s];
  c.rcut_hard_in = p->rcut_hard[s];
  c.inv_rcut = 1.0 / p->rcut_hard[s];
  c.rcut_soft = p->rcut_soft[s] * c.inv_rcut;
  c.atom_sigma = p->atom_sigma_r[s] * c.inv_rcut;
  c.dr = 1.0 - c.rcut_soft;
  c.sqrt_rcut = std::sqrt(p->rcut_hard[s]);
  c.inv_sqrt_rcut = 1.0 / c.sqrt_rcut;
  c.sigma_sq0 = c.atom_sigma * c.atom_sigma;
  c.N_gauss = std::sqrt(2.0 / c.atom_sigma) / std::pow(std::acos(-1.0), 0.25);
  c.has_filter = c.dr > 0.0;
  if (c.has_filter) {
    c.drnf = c.dr / p->nf[s];
    c.q = c.drnf * c.drnf;
  }
  c.n_poly = c.with_gauss ? a - 1 : a;
  c.n_hi = c.n_poly + (do_der2 ? 4 : do_der ? 2 : 0);
  if (p->basis == ST_BASIS_POLY3OPERATOR) {
    // get_constant_poly_coeff(alpha_max, 1, A) (src/soap_turbo_radial.f90:2276-2297) and the
    // filter's cubic g_aux(0, rcut_soft, filter_width, "right") (:2154-2171, :2320)
    c.op = 1;
    c.op_K = 2.0 * std::sqrt(2.0 * std::log(2.0));
    c.op_sigma_in = p->atom_sigma_r[s];
    c.op_rcut_soft_in = p->rcut_soft[s];
    for (int i = 1; i <= a; ++i) {
      c.op_A[(i - 1) * 7] = 1.0 / std::sqrt(1.0 / (double)(2 * i + 5)) / (double)(i + 3);
      for (int j = 2; j <= 7; ++j) c.op_A[(i - 1) * 7 + j - 1] = c.op_A[(i - 1) * 7 + j - 2] * 1.0 / (double)(i + 3 + j - 1);
    }
    const double fw = c.op_K * (1.0 - c.rcut_soft), x = (0.0 - c.rcut_soft) / fw;
    c.op_cf[0] = 1.0 - 3.0 * (x * x) + 2.0 * (x * x * x);
    c.op_cf[1] = 6.0 * (x * x - x) / fw;
    c.op_cf[2] = 3.0 * (2.0 * x - 1.0) / (fw * fw);
    c.op_cf[3] = 2.0 / (fw * fw * fw);
  }
  double Nv[ST_REC + 8];  // Nv(n) at n + 3
  Nv[0] = 1.0;
  for (int n = -2; n <= a + 4; ++n) Nv[n + 3] = std::sqrt(1.0 / (double)(2 * n + 5));
  for (int n = -1; n <= c.n_hi; ++n) {
    c.ra[n + 1] = (double)(n + 1) * Nv[n - 2 + 3] / Nv[n + 3];
    c.rb[n + 1] = Nv[n - 1 + 3] / Nv[n + 3];
    c.rc[n + 1] = 1.0 / Nv[n + 3];
  }
  c.ra[0] = 0.0;
  c.inv_Nm2 = 1.0 / Nv[1];
  for (int n = 1; n <= c.n_poly; ++n) {
    c.R1[n] = Nv[n + 1 + 3] / Nv[n + 3];
    c.R2[n] = Nv[n + 2 + 3] / Nv[n + 3];
    c.R3[n] = Nv[n + 3 + 3] / Nv[n + 3];
    c.R4[n] = Nv[n + 4 + 3] / Nv[n + 3];
  }
  for (int j = 0; j < a; ++j)
    for (int i = 0; i < a; ++i) c.W[i + j * a] = W[(off + i) + (size_t)(off + j) * n_max];
  return c;
}

// get_preflm, src/soap_turbo_functions.f90:215-241
void preflm_host(int l_max, std::vector<double>& out) {
  out.clear();
  for (int l = 0; l <= l_max; ++l) {
    double ppi = std::sqrt((double)(2 * l + 1) / 4.0 / std::acos(-1.0));
    double fact2 = 1.0;
    for (int i = 1; i <= l; ++i) fact2 = fact2 * (double)i;
    double fact1 = fact2;
    for (int m = 0; m <= l; ++m) {
      if (m > 0) {
        fact1 = fact1 / (double)(l + 1 - m);
        fact2 = fact2 * (double)(l + m);
      }
      out.push_back(ppi * std::sqrt(fact1 / fact2));
    }
  }
}

int n_soap_of(const st_params* p, const st_compression* c) {
  const int n_max = n_max_of(p);
  if (c && c->P_nonzero > 0) {
    int m = 0;
    for (int i = 0; i < c->P_nonzero; ++i) m = std::max(m, c->P_i[i]);
    return m;
  }
  return n_max * (n_max + 1) / 2 * (p->l_max + 1);
}

// Index tables of the spectrum and the compression (see power_k and
// contract_k), and the poly3gauss central-atom term.
int build_tables(st_ctx* ctx, const st_params* p, const st_compression* c) {
  const int n_max = n_max_of(p), L = p->l_max, L1 = L + 1;
  const int n_unc = n_max * (n_max + 1) / 2 * L1;
  const bool cmp = c && c->P_nonzero > 0;
  std::string key(reinterpret_cast<const char*>(&n_max), sizeof n_max);
  key.append(reinterpret_cast<const char*>(&L), sizeof L);
  if (cmp) {
    key.append(reinterpret_cast<const char*>(c->P_i), sizeof(int32_t) * c->P_nonzero);
    key.append(reinterpret_cast<const char*>(c->P_j), sizeof(int32_t) * c->P_nonzero);
    key.append(reinterpret_cast<const char*>(c->P_el), sizeof(double) * c->P_nonzero);
  }
  ctx->n_max = n_max;
  ctx->k_max = k_max_of(L);
  ctx->l_max = L;
  ctx->n_soap_unc = n_unc;
  ctx->n_soap = n_soap_of(p, c);
  ctx->compressed = cmp;
  int rc;
  if (key != ctx->tables_key) {
    std::vector<int32_t> comp((size_t)4 * n_unc);
    int cnt = 0;
    for (int n = 0; n < n_max; ++n)
      for (int np = n; np < n_max; ++np)
        for (int l = 0; l <= L; ++l, ++cnt) {
          comp[4 * cnt] = n; comp[4 * cnt + 1] = np; comp[4 * cnt + 2] = l; comp[4 * cnt + 3] = 0;
        }
    std::vector<uint8_t> skip(n_unc, cmp ? 1 : 0);
    if ((rc = upload(ctx, ctx->comp, comp.data(), comp.size() * 4))) return rc;
    if (cmp) {
      const int nz = c->P_nonzero, ns = ctx->n_soap;
      for (int k = 0; k < nz; ++k) {
        if (c->P_j[k] < 1 || c->P_j[k] > n_unc || c->P_i[k] < 1)
          return fail(ctx, ST_ERR_ARG, "compression index out of range");
        skip[c->P_j[k] - 1] = 0;
      }
      std::vector<int> order(nz);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return c->P_i[a] < c->P_i[b]; });
      std::vector<int32_t> optr(ns + 1, 0), oc(nz);
      std::vector<double> oel(nz);
      for (int k = 0; k < nz; ++k) optr[c->P_i[k]]++;
      for (int s = 0; s < ns; ++s) optr[s + 1] += optr[s];
      for (int e = 0; e < nz; ++e) { oc[e] = c->P_j[order[e]] - 1; oel[e] = c->P_el[order[e]]; }
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return c->P_j[a] < c->P_j[b]; });
      std::vector<int32_t> uptr(n_unc + 1, 0), ue(nz);
      for (int k = 0; k < nz; ++k) uptr[c->P_j[k]]++;
      for (int u = 0; u < n_unc; ++u) uptr[u + 1] += uptr[u];
      for (int e = 0; e < nz; ++e) ue[e] = order[e];
      if ((rc = upload(ctx, ctx->out_ptr, optr.data(), optr.size() * 4))) return rc;
      if ((rc = upload(ctx, ctx->out_c, oc.data(), oc.size() * 4))) return rc;
      if ((rc = upload(ctx, ctx->out_el, oel.data(), oel.size() * 8))) return rc;
      ctx->out_el_h = oel;
      ctx->P_el_h.assign(c->P_el, c->P_el + nz);
      if ((rc = upload(ctx, ctx->unc_ptr, uptr.data(), uptr.size() * 4))) return rc;
      if ((rc = upload(ctx, ctx->unc_e, ue.data(), ue.size() * 4))) return rc;
      if ((rc = upload(ctx, ctx->P_i, c->P_i, (size_t)nz * 4))) return rc;
      if ((rc = upload(ctx, ctx->P_el, c->P_el, (size_t)nz * 8))) return rc;
    }
    if ((rc = upload(ctx, ctx->skip, skip.data(), skip.size()))) return rc;
    // direct maps: output s <- one component, component c -> one output (or none)
    {
      const int ns = ctx->n_soap;
      std::vector<int32_t> oc4((size_t)4 * ns);
      std::vector<double> oel(ns, 1.0), uel(n_unc, 0.0);
      std::vector<int32_t> um(n_unc, -1);
      bool single = true;
      if (!cmp) {
        for (int q = 0; q < ns; ++q) {
          for (int j = 0; j < 4; ++j) oc4[4 * q + j] = comp[4 * q + j];
          um[q] = q;
          uel[q] = 1.0;
        }
      } else {
        std::vector<int> nout(ns, 0), nunc(n_unc, 0);
        for (int k = 0; k < c->P_nonzero; ++k) { nout[c->P_i[k] - 1]++; nunc[c->P_j[k] - 1]++; }
        for (int q = 0; q < ns && single; ++q) single = nout[q] == 1;
        for (int u = 0; u < n_unc && single; ++u) single = nunc[u] <= 1;
        if (single)
          for (int k = 0; k < c->P_nonzero; ++k) {
            const int q = c->P_i[k] - 1, u = c->P_j[k] - 1;
            for (int j = 0; j < 4; ++j) oc4[4 * q + j] = comp[4 * u + j];
            oel[q] = c->P_el[k];
            um[u] = q;
            uel[u] = c->P_el[k];
          }
      }
      ctx->single = single;
      if (single) {
        if ((rc = upload(ctx, ctx->ocomp, oc4.data(), oc4.size() * 4))) return rc;
        if ((rc = upload(ctx, ctx->oel, oel.data(), oel.size() * 8))) return rc;
        if ((rc = upload(ctx, ctx->umap, um.data(), um.size() * 4))) return rc;
        if ((rc = upload(ctx, ctx->uel, uel.data(), uel.size() * 8))) return rc;
        ctx->oel_h = oel;
        ctx->uel_h = uel;
        ctx->umap_h = um;
      }
    }
    std::vector<double> pf;
    preflm_host(L, pf);
    if ((rc = upload(ctx, ctx->preflm, pf.data(), pf.size() * 8))) return rc;
    ctx->tables_key = key;
  }
  return ST_OK;
}

// poly3gauss central-atom term per species (src/soap_turbo.f90:646-661):
// fac * (W S(:, last)) on the species block; zero for other bases.
int build_central(st_ctx* ctx, const st_params* p, std::vector<double>& cen) {
  const int n_max = n_max_of(p);
  cen.assign((size_t)p->n_species * n_max, 0.0);
  if (p->basis != ST_BASIS_POLY3GAUSS) return ST_OK;
  const double pi = std::acos(-1.0);
  int off = 0;
  for (int j = 0; j < p->n_species; ++j) {
    const int a = p->alpha_max[j];
    if (p->central_weight[j] != 0.0) {
      double amplitude;
      if (p->radial_enhancement == 1) amplitude = std::sqrt(2.0 / pi) * p->atom_sigma_r[j] / p->rcut_hard[j];
      else if (p->radial_enhancement == 2)
        amplitude = (p->atom_sigma_r[j] * p->atom_sigma_r[j]) / (p->rcut_hard[j] * p->rcut_hard[j]);
      else amplitude = 1.0;
      const double rc = p->rcut_hard[j];
      const double fac = amplitude * p->central_weight[j] * std::sqrt(4.0 * pi) * std::pow(pi, 0.25) *
                         std::sqrt(p->atom_sigma_r[j] / 2.0) * (rc * rc * rc) /
                         (p->atom_sigma_t[j] * p->atom_sigma_t[j]) / p->atom_sigma_r[j];
      const int col = off + a - 1;
      for (int n = 0; n < a; ++n) {
        double v = 0.0;
        for (int m = 0; m < a; ++m) v = v + ctx->W[(off + n) + (size_t)(off + m) * n_max] * ctx->S[(off + m) + (size_t)col * n_max];
        cen[(size_t)j * n_max + off + n] = fac * v;
      }
    }
    off += a;
  }
  return ST_OK;
}

int pack_mask(st_ctx* ctx, int n_species) {
  if (n_species > ctx->n_species_mask) return fail(ctx, ST_ERR_ARG, "mask has fewer columns than n_species");
  int rc;
  if ((rc = ensure(ctx, ctx->spmask, (size_t)ctx->n_pairs * 4))) return rc;
  if ((rc = ensure(ctx, ctx->first, (size_t)ctx->n_pairs))) return rc;
  if (ctx->packed_species == n_species) return ST_OK;
  st::launch_pack_pairs(ctx->stream, ctx->n_pairs, n_species, ctx->mask.as<int32_t>(), ctx->site_ptr.as<int>(),
                        ctx->n_sites, ctx->spmask.as<uint32_t>(), ctx->first.as<uint8_t>());
  ctx->packed_species = n_species;
  return ST_OK;
}

// Forces and virial of the FP64 pair gradients, when the pair atoms are known.
int forces64(st_ctx* ctx) {
  ctx->have_forces64 = false;
  if (!ctx->have_atoms) return ST_OK;
  int rc;
  if ((rc = ensure(ctx, ctx->forces64, (size_t)ctx->n_atoms * 24))) return rc;
  if ((rc = ensure(ctx, ctx->vpart64, (size_t)st::gather_blocks(ctx->n_atoms) * 72))) return rc;
  if ((rc = ensure(ctx, ctx->virial64, 72))) return rc;
  timed(ctx, "forces_f64", [&] {
    st::launch_forces_f64(ctx->stream, ctx->n_atoms, ctx->pair64.as<double>(), ctx->atom_ptr.as<int>(),
                          ctx->atom_pairs.as<int>(), ctx->have_vec ? ctx->vec.as<double>() : nullptr,
                          ctx->rjs.as<double>(), ctx->thetas.as<double>(), ctx->phis.as<double>(),
                          ctx->forces64.as<double>(), ctx->vpart64.as<double>(), ctx->virial64.as<double>());
  });
  ctx->have_forces64 = true;
  return ST_OK;
}

}  // namespace

// =====================================================================================================
extern "C" {

int st_create(st_ctx** out, int device) {
  if (!out) return ST_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return ST_ERR_HIP;
  if (device < 0 || device >= n) return ST_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return ST_ERR_HIP;
  st_ctx* c = new st_ctx;
  c->device = device;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->lds_max = (size_t)lds;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    c->n_cu = ncu;
  c->scratch = scratch_on(device);
  if (!c->scratch || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ST_ERR_HIP;
  }
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  *out = c;
  return ST_OK;
}

void st_destroy(st_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& e : c->ev) (void)hipEventDestroy(e);
  for (auto& pd : c->pending) { (void)hipEventDestroy(pd.a); (void)hipEventDestroy(pd.b); }
  for (auto e : c->free_events) (void)hipEventDestroy(e);
  if (c->g32_exec) (void)hipGraphExecDestroy(c->g32_exec);
  if (c->g32_graph) (void)hipGraphDestroy(c->g32_graph);
  st::adjoint_release(c->adj);
  hipStream_t s = c->stream;
  delete c;
  (void)hipStreamDestroy(s);
}

const char* st_error(const st_ctx* c) { return c ? c->err.c_str() : "no context"; }

int st_n_soap(const st_params* p, const st_compression* c) { return p ? n_soap_of(p, c) : -1; }

int st_get_basis(st_ctx* ctx, const st_params* p, double* W, double* S) {
  if (!ctx) {
    // host-only: build the basis without a device context
    st_ctx tmp;
    int rc = check_params(&tmp, p);
    if (!rc) rc = basis_for(&tmp, p);
    if (rc) return rc;
    if (W) std::memcpy(W, tmp.W.data(), tmp.W.size() * 8);
    if (S) std::memcpy(S, tmp.S.data(), tmp.S.size() * 8);
    return ST_OK;
  }
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = basis_for(ctx, p))) return rc;
  if (W) std::memcpy(W, ctx->W.data(), ctx->W.size() * 8);
  if (S) std::memcpy(S, ctx->S.data(), ctx->S.size() * 8);
  return ST_OK;
}

int st_basis_matrices(int32_t route, const st_params* p, double* W, double* S) {
  if (route != ST_BASIS_ROUTE_EXTENDED && route != ST_BASIS_ROUTE_LAPACK) return ST_ERR_ARG;
  st_ctx tmp;  // host-only: no device is touched
  tmp.basis_route = route;
  int rc = check_params(&tmp, p);
  if (!rc) rc = basis_for(&tmp, p);
  if (rc) return rc;
  if (W) std::memcpy(W, tmp.W.data(), tmp.W.size() * 8);
  if (S) std::memcpy(S, tmp.S.data(), tmp.S.size() * 8);
  return ST_OK;
}

int st_set_basis_route(st_ctx* ctx, int32_t route) {
  if (!ctx) return ST_ERR_ARG;
  if (route != ST_BASIS_ROUTE_EXTENDED && route != ST_BASIS_ROUTE_LAPACK)
    return fail(ctx, ST_ERR_ARG, "no such route for W");
  ctx->basis_route = route;
  return ST_OK;
}

int st_set_basis(st_ctx* ctx, int32_t n_max, const double* W, const double* S) {
  if (!ctx) return ST_ERR_ARG;
  if (n_max <= 0) {
    ctx->inj_n_max = 0;
    ctx->basis_key.clear();
    return ST_OK;
  }
  if (!W || !S) return fail(ctx, ST_ERR_ARG, "W and S are required");
  ctx->inj_n_max = n_max;
  ctx->injW.assign(W, W + (size_t)n_max * n_max);
  ctx->injS.assign(S, S + (size_t)n_max * n_max);
  ctx->basis_key.clear();
  return ST_OK;
}

int st_upload_neighbours(st_ctx* ctx, const st_neighbours* nb) {
  if (!ctx || !nb) return ST_ERR_ARG;
  if (nb->n_sites < 0 || nb->n_pairs < 0 || nb->n_species < 1 || nb->n_species > ST_MAX_SPECIES || nb->max_mult < 1)
    return fail(ctx, ST_ERR_ARG, "bad neighbour-list sizes");
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<int> ptr(nb->n_sites + 1, 0);
  for (int i = 0; i < nb->n_sites; ++i) {
    if (nb->n_neigh[i] < 1) return fail(ctx, ST_ERR_ARG, "every site needs at least its own pair (n_neigh >= 1)");
    ptr[i + 1] = ptr[i] + nb->n_neigh[i];
  }
  if (ptr[nb->n_sites] != nb->n_pairs) return fail(ctx, ST_ERR_ARG, "sum(n_neigh) /= n_atom_pairs");
  int rc;
  ctx->n_sites = nb->n_sites;
  ctx->n_pairs = nb->n_pairs;
  ctx->n_species_mask = nb->n_species;
  ctx->packed_species = 0;
  ctx->max_mult = nb->max_mult;
  ctx->site_ptr_h = ptr;
  const size_t P = nb->n_pairs;
  if ((rc = upload(ctx, ctx->site_ptr, ptr.data(), ptr.size() * 4))) return rc;
  if ((rc = upload(ctx, ctx->species, nb->species, (size_t)nb->max_mult * nb->n_sites * 4))) return rc;
  if ((rc = upload(ctx, ctx->mult, nb->species_multiplicity, (size_t)nb->n_sites * 4))) return rc;
  if ((rc = upload(ctx, ctx->mask, nb->mask, P * nb->n_species * 4))) return rc;
  if ((rc = upload(ctx, ctx->rjs, nb->rjs, P * 8))) return rc;
  if ((rc = upload(ctx, ctx->thetas, nb->thetas, P * 8))) return rc;
  if ((rc = upload(ctx, ctx->phis, nb->phis, P * 8))) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));  // the host arrays may go away after return
  ctx->have_nb = true;
  ctx->have_soap = ctx->have_cart = ctx->ct_ready = ctx->have_pair64 = false;
  // the pair -> atom map and pair vectors belong to the previous list: upload them again
  ctx->have_atoms = ctx->have_vec = false;
  ctx->have_G = false;  // dE/dq is per site: upload it again for the new sites
  st::adjoint_invalidate(ctx->adj);
  ++ctx->nb_gen;
  return ST_OK;
}

int st_upload_atoms(st_ctx* ctx, int32_t n_atoms, const int32_t* pair_atom, const double* vec) {
  if (!ctx || !ctx->have_nb) return fail(ctx, ST_ERR_STATE, "upload the neighbour list first");
  if (n_atoms < 1 || !pair_atom) return fail(ctx, ST_ERR_ARG, "pair_atom and n_atoms are required");
  for (int p = 0; p < ctx->n_pairs; ++p)
    if (pair_atom[p] < 0 || pair_atom[p] >= n_atoms) return fail(ctx, ST_ERR_ARG, "pair_atom out of range");
  int rc;
  HIPCHK(hipSetDevice(ctx->device));
  if ((rc = upload(ctx, ctx->pair_atom, pair_atom, (size_t)ctx->n_pairs * 4))) return rc;
  // the pairs of each atom, in increasing pair index (counting sort)
  std::vector<int> aptr((size_t)n_atoms + 1, 0), apairs(ctx->n_pairs);
  for (int p = 0; p < ctx->n_pairs; ++p) aptr[pair_atom[p] + 1]++;
  for (int a = 0; a < n_atoms; ++a) aptr[a + 1] += aptr[a];
  {
    std::vector<int> fill(aptr.begin(), aptr.end() - 1);
    for (int p = 0; p < ctx->n_pairs; ++p) apairs[fill[pair_atom[p]]++] = p;
  }
  if ((rc = upload(ctx, ctx->atom_ptr, aptr.data(), aptr.size() * 4))) return rc;
  if ((rc = upload(ctx, ctx->atom_pairs, apairs.data(), apairs.size() * 4))) return rc;
  ctx->have_vec = vec != nullptr;
  if (vec && (rc = upload(ctx, ctx->vec, vec, (size_t)ctx->n_pairs * 24))) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->n_atoms = n_atoms;
  ctx->have_atoms = true;
  ++ctx->nb_gen;
  return ST_OK;
}

int st_upload_dedq(st_ctx* ctx, int32_t n_soap, const double* G) {
  if (!ctx || !ctx->have_nb) return fail(ctx, ST_ERR_STATE, "upload the neighbour list first");
  if (n_soap < 1 || !G) return fail(ctx, ST_ERR_ARG, "bad dE/dq");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = upload(ctx, ctx->G, G, (size_t)n_soap * ctx->n_sites * 8);
  if (rc) return rc;
  // the FP32 adjoint path reads dE/dq site-major in FP32: converted here, once per upload
  std::string msg;
  if ((rc = st::adjoint_set_dedq(ctx->adj, ctx->stream, ctx->n_sites, n_soap, ctx->G.as<double>(), msg)))
    return fail(ctx, rc, msg);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->G_n_soap = n_soap;
  ctx->have_G = true;
  return ST_OK;
}

}  // extern "C"

namespace {

// One range of sites [s0, s1) of the resident list, whose pairs are [p0, p1). site_ptr holds the
// range's own offsets (site_ptr[0] = 0), so every per-pair array is addressed from p0 and every
// per-site array from s0; the whole list is the range (0, n_sites) with the resident offsets.
struct SiteRange {
  int s0, s1;
  size_t p0, p1;
  const int* site_ptr;
};

// Per-pair and per-site device bytes of the FP64 work arrays (radial values and derivatives, the four
// angular arrays, the rotation coefficients, soap_cart_der; cnk per site).
size_t fp64_pair_bytes(int n_max, int k_max, int n_soap) {
  return (size_t)16 * n_max + (size_t)64 * k_max + 64 + (size_t)24 * n_soap;
}
size_t fp64_site_bytes(int n_max, int k_max) { return (size_t)16 * n_max * k_max; }

// Grows the FP64 work arrays to hold ranges of up to np pairs and ns sites.
int fp64_work(st_ctx* ctx, bool der, bool ctr, size_t np, size_t ns) {
  const int n_max = ctx->n_max, k_max = ctx->k_max, n_soap = ctx->n_soap;
  int rc;
  if ((rc = ensure(ctx, ctx->rad, np * n_max * 8))) return rc;
  if ((rc = ensure(ctx, ctx->rad_der, np * n_max * 8))) return rc;
  if ((rc = ensure(ctx, ctx->ang, np * k_max * 16))) return rc;
  if (der) {
    if ((rc = ensure(ctx, ctx->ang_r, np * k_max * 16))) return rc;
    if ((rc = ensure(ctx, ctx->ang_a, np * k_max * 16))) return rc;
    if ((rc = ensure(ctx, ctx->ang_p, np * k_max * 16))) return rc;
  }
  if ((rc = ensure(ctx, ctx->cnk, ns * n_max * k_max * 16))) return rc;
  if (der && !ctr) {
    if ((rc = ensure(ctx, ctx->cart, np * n_soap * 24))) {
      ctx->err += " (soap_cart_der of " + std::to_string(np) + " pairs; st_full_jacobian_forces runs the "
                  "Jacobian in batches of sites that fit)";
      return rc;
    }
    if ((rc = ensure(ctx, ctx->rotp, np * 64))) return rc;
  }
  return ST_OK;
}

// get_soap's FP64 pass over one range of sites (src/soap_turbo.f90:489-909): radial and angular
// expansions of the range's pairs, cnk, the power spectrum, with der the Jacobian into ctx->cart
// (range-local rows), then the normalisation. Basis, tables, masks and the central term must be set.
int fp64_range(st_ctx* ctx, const st_params* p, bool der, bool ctr, const SiteRange& r) {
  const int n_max = ctx->n_max, k_max = ctx->k_max, L = ctx->l_max, n_soap = ctx->n_soap;
  const int NS = r.s1 - r.s0;
  const size_t P = r.p1 - r.p0;
  if (NS == 0) return ST_OK;
  hipStream_t st = ctx->stream;
  const double* rjs = ctx->rjs.as<double>() + r.p0;
  const double* th = ctx->thetas.as<double>() + r.p0;
  const double* ph = ctx->phis.as<double>() + r.p0;
  const uint32_t* spmask = ctx->spmask.as<uint32_t>() + r.p0;
  const uint8_t* first = ctx->first.as<uint8_t>() + r.p0;
  double* soap = ctx->soap.as<double>() + (size_t)r.s0 * n_soap;
  double* sdp = ctx->sdp.as<double>() + r.s0;

  timed(ctx, "fill_rad", [&] {
    st::launch_fill(st, ctx->rad.as<double>(), P * n_max, 0.0);
    if (der) st::launch_fill(st, ctx->rad_der.as<double>(), P * n_max, 0.0);
  });
  int off = 0;
  for (int s = 0; s < p->n_species; ++s) {
    st::RadialConsts rcst = radial_consts(p, s, ctx->W, n_max, off, der);
    bool ok = true;
    timed(ctx, "radial_f64", [&] {
      ok = st::launch_radial(st, rcst, (int)P, rjs, spmask, first, ctx->rad.as<double>(), ctx->rad_der.as<double>());
    });
    if (!ok) return fail(ctx, ST_ERR_UNSUPPORTED, "alpha_max out of the compiled range");
    off += p->alpha_max[s];
  }
  st::AngularConsts ac;
  std::memset(&ac, 0, sizeof ac);
  ac.n_species = p->n_species;
  ac.l_max = L;
  ac.k_max = k_max;
  ac.do_der = der;
  for (int s = 0; s < p->n_species; ++s) {
    ac.rcut_max = std::max(ac.rcut_max, p->rcut_hard[s]);
    ac.sigma_t[s] = p->atom_sigma_t[s];
    ac.sigma_t_scaling[s] = p->atom_sigma_t_scaling[s];
  }
  ac.rcut2 = ac.rcut_max * ac.rcut_max;
  bool ok = true;
  timed(ctx, "angular_f64", [&] {
    ok = st::launch_angular(st, ac, (int)P, rjs, th, ph, spmask, ctx->preflm.as<double>(), ctx->ang.as<double2>(),
                            der ? ctx->ang_r.as<double2>() : nullptr, der ? ctx->ang_a.as<double2>() : nullptr,
                            der ? ctx->ang_p.as<double2>() : nullptr);
  });
  if (!ok) return fail(ctx, ST_ERR_UNSUPPORTED, "l_max out of the compiled range");
  timed(ctx, "cnk_f64", [&] {
    st::launch_cnk(st, NS, n_max, k_max, r.site_ptr, ctx->rad.as<double>(), ctx->ang.as<double2>(),
                   ctx->species.as<int32_t>() + (size_t)r.s0 * ctx->max_mult, ctx->max_mult,
                   ctx->mult.as<int32_t>() + r.s0,
                   p->basis == ST_BASIS_POLY3GAUSS ? ctx->central.as<double>() : nullptr, ctx->cnk.as<double2>());
  });
  const int* optr = ctx->compressed ? ctx->out_ptr.as<int>() : nullptr;
  timed(ctx, "power_f64", [&] {
    st::launch_power(st, NS, n_max, k_max, n_soap, ctx->comp.as<int4>(), optr, ctx->out_c.as<int>(),
                     ctx->out_el.as<double>(), ctx->single ? ctx->ocomp.as<int4>() : nullptr, ctx->oel.as<double>(),
                     ctx->cnk.as<double2>(), soap, sdp);
  });
  if (der && !ctr) {
    auto jacobian = [&](int variant) {
      if (st::derivatives_lds(n_max, L, k_max, n_soap, variant) > ctx->lds_max) return false;
      return st::launch_derivatives(st, NS, (int)P, n_max, L, k_max, n_soap, r.site_ptr, ctx->comp.as<int4>(), optr,
                                    ctx->out_c.as<int>(), ctx->out_el.as<double>(),
                                    ctx->single ? ctx->ocomp.as<int4>() : nullptr, ctx->oel.as<double>(),
                                    ctx->cnk.as<double2>(), ctx->rad.as<double>(), ctx->rad_der.as<double>(),
                                    ctx->ang.as<double2>(), ctx->ang_r.as<double2>(), ctx->ang_a.as<double2>(),
                                    ctx->ang_p.as<double2>(), soap, sdp, rjs, th, ph, ctx->rotp.as<double>(),
                                    ctx->cart.as<double>(), variant);
    };
    int variant = 0;
    if (ctx->jacobian_variant >= 0) {
      if (ctx->jacobian_variant >= st::derivatives_variants(n_soap))
        return fail(ctx, ST_ERR_ARG, "no such Jacobian launch variant for this n_soap");
      variant = ctx->jacobian_variant;
    } else if (ctx->autotune && !ctx->timing) {
      // the kernel's shape (threads x outputs per thread) for this output count, basis and batch size
      const std::string key = "jacobian " + std::to_string(n_soap) + " " + std::to_string(n_max) + " " +
                              std::to_string(L) + " " + std::to_string(ctx->single) + " " +
                              std::to_string(NS / 256);
      variant = tuned_variant(ctx, key, st::derivatives_variants(n_soap), jacobian);
    }
    timed(ctx, "derivatives_f64", [&] { ok = jacobian(variant); });
    if (!ok) return fail(ctx, ST_ERR_UNSUPPORTED, "n_soap > 3072 is not supported with derivatives");
  }
  timed(ctx, "normalise_f64", [&] { st::launch_normalise(st, NS, n_soap, sdp, soap); });
  return ST_OK;
}

// Set-up shared by the FP64 entry points: parameters, basis, index tables, packed masks, central term,
// the per-site outputs, and the LDS checks of the spectrum and Jacobian kernels.
int fp64_setup(st_ctx* ctx, const st_params* p, const st_compression* c, bool der, bool ctr) {
  if (!ctx->have_nb) return fail(ctx, ST_ERR_STATE, "upload the neighbour list first");
  int rc = check_params(ctx, p);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  if ((rc = basis_for(ctx, p))) return rc;
  if ((rc = build_tables(ctx, p, c))) return rc;
  if ((rc = pack_mask(ctx, p->n_species))) return rc;
  std::vector<double> cen;
  build_central(ctx, p, cen);
  if ((rc = upload(ctx, ctx->central, cen.data(), cen.size() * 8))) return rc;
  if ((rc = ensure(ctx, ctx->soap, (size_t)ctx->n_sites * ctx->n_soap * 8))) return rc;
  if ((rc = ensure(ctx, ctx->sdp, (size_t)ctx->n_sites * 8))) return rc;
  if (sizeof(double2) * ctx->n_max * ctx->k_max > 64 * 1024)
    return fail(ctx, ST_ERR_UNSUPPORTED, "n_max * k_max too large for the spectrum kernel");
  if (der && !ctr && st::derivatives_lds(ctx->n_max, ctx->l_max, ctx->k_max, ctx->n_soap) > ctx->lds_max)
    return fail(ctx, ST_ERR_UNSUPPORTED, "the derivative working set exceeds the LDS");
  return ST_OK;
}

}  // namespace

extern "C" {

int st_compute(st_ctx* ctx, const st_params* p, const st_compression* c, int32_t do_derivatives, int32_t contract) {
  if (!ctx) return ST_ERR_ARG;
  const bool der = do_derivatives != 0, ctr = der && contract != 0;
  int rc = fp64_setup(ctx, p, c, der, ctr);
  if (rc) return rc;
  if ((rc = fp64_work(ctx, der, ctr, ctx->n_pairs, ctx->n_sites))) return rc;
  const SiteRange all{0, ctx->n_sites, 0, (size_t)ctx->n_pairs, ctx->site_ptr.as<int>()};
  if ((rc = fp64_range(ctx, p, der, ctr, all))) return rc;
  HIPCHK(hipGetLastError());
  ctx->have_soap = true;
  ctx->have_cart = der && !ctr;
  ctx->ct_ready = ctr;
  ctx->have_pair64 = false;
  return ST_OK;
}

int st_full_jacobian_forces(st_ctx* ctx, const st_params* p, const st_compression* c, int64_t max_pairs,
                            int32_t* n_batches) {
  if (!ctx) return ST_ERR_ARG;
  if (max_pairs < 0) return fail(ctx, ST_ERR_ARG, "max_pairs must be >= 0");
  int rc = fp64_setup(ctx, p, c, true, false);
  if (rc) return rc;
  if (!ctx->have_G || ctx->G_n_soap != ctx->n_soap) return fail(ctx, ST_ERR_STATE, "dE/dq missing or of another size");
  const int NS = ctx->n_sites;
  const std::vector<int>& ptr = ctx->site_ptr_h;
  const size_t pb = fp64_pair_bytes(ctx->n_max, ctx->k_max, ctx->n_soap);
  const size_t sb = fp64_site_bytes(ctx->n_max, ctx->k_max);
  int maxn = 0;
  for (int i = 0; i < NS; ++i) maxn = std::max(maxn, ptr[i + 1] - ptr[i]);
  size_t avail = 0;
  if (max_pairs == 0) {
    // automatic: what the device has free, plus what the work arrays already hold, less a margin
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    size_t held = 0;
    for (DevBuf* b : {&ctx->rad, &ctx->rad_der, &ctx->ang, &ctx->ang_r, &ctx->ang_a, &ctx->ang_p, &ctx->cnk,
                      &ctx->cart, &ctx->rotp, &ctx->pair64})
      held += b->bytes;
    const size_t margin = ((size_t)2 << 30) + (ctx->pair64.bytes >= (size_t)ctx->n_pairs * 24 ? 0 : (size_t)ctx->n_pairs * 24);
    avail = fr + held > margin ? fr + held - margin : 0;
  }
  size_t cap = max_pairs > 0 ? (size_t)max_pairs : avail / pb;
  // the batches: consecutive sites, as even in pairs as the cap allows; with an automatic cap, shrunk
  // until the largest batch's pairs and sites fit
  const size_t P = ctx->n_pairs;
  std::vector<int> cut;
  size_t big_p = 0, big_s = 0;
  for (;;) {
    if (cap < (size_t)maxn)
      return fail(ctx, ST_ERR_ARG, "a batch must hold the largest site's pairs (" + std::to_string(maxn) +
                                       (max_pairs > 0 ? ")" : "), more than the free device memory holds"));
    const size_t nb_min = std::max<size_t>(1, (P + cap - 1) / cap);
    const size_t target = (P + nb_min - 1) / nb_min;
    cut.assign(1, 0);
    for (int i = 0; i < NS; ++i) {
      if ((size_t)ptr[i + 1] - ptr[cut.back()] > cap) cut.push_back(i);                    // site i opens a batch
      if ((size_t)ptr[i + 1] - ptr[cut.back()] >= target && i + 1 < NS) cut.push_back(i + 1);  // the batch is full
    }
    if (cut.back() != NS) cut.push_back(NS);
    big_p = big_s = 0;
    for (size_t b = 0; b + 1 < cut.size(); ++b) {
      big_p = std::max(big_p, (size_t)ptr[cut[b + 1]] - ptr[cut[b]]);
      big_s = std::max(big_s, (size_t)(cut[b + 1] - cut[b]));
    }
    if (max_pairs > 0 || big_p * pb + big_s * sb <= avail) break;
    cap = cap - cap / 8;
  }
  const int nbat = (int)cut.size() - 1;
  if ((rc = fp64_work(ctx, true, false, big_p, big_s))) return rc;
  if ((rc = ensure(ctx, ctx->pair64, P * 24))) return rc;
  // the batches' own site offsets, one run of (sites + 1) entries per batch
  std::vector<int> bptr;
  std::vector<size_t> bofs;
  if (nbat > 1) {
    for (int b = 0; b < nbat; ++b) {
      bofs.push_back(bptr.size());
      for (int i = cut[b]; i <= cut[b + 1]; ++i) bptr.push_back(ptr[i] - ptr[cut[b]]);
    }
    if ((rc = upload(ctx, ctx->batch_ptr, bptr.data(), bptr.size() * 4))) return rc;
  }
  for (int b = 0; b < nbat; ++b) {
    const SiteRange r{cut[b], cut[b + 1], (size_t)ptr[cut[b]], (size_t)ptr[cut[b + 1]],
                      nbat > 1 ? ctx->batch_ptr.as<int>() + bofs[b] : ctx->site_ptr.as<int>()};
    if ((rc = fp64_range(ctx, p, true, false, r))) return rc;
    timed(ctx, "pairs_full_jacobian_f64", [&] {
      st::launch_contract_cart(ctx->stream, r.s1 - r.s0, (int)(r.p1 - r.p0), ctx->n_soap, r.site_ptr,
                               ctx->G.as<double>() + r.s0, NS, ctx->cart.as<double>(),
                               ctx->pair64.as<double>() + 3 * r.p0);
    });
  }
  HIPCHK(hipGetLastError());
  if (n_batches) *n_batches = nbat;
  ctx->have_soap = true;
  ctx->have_cart = nbat == 1;  // soap_cart_der is whole only when one batch held every pair
  ctx->ct_ready = false;
  ctx->have_pair64 = true;
  return forces64(ctx);
}

int st_pairs_full_jacobian(st_ctx* ctx) {
  if (!ctx || !ctx->have_cart) return fail(ctx, ST_ERR_STATE, "soap_cart_der has not been computed");
  if (!ctx->have_G || ctx->G_n_soap != ctx->n_soap) return fail(ctx, ST_ERR_STATE, "dE/dq missing or of another size");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = ensure(ctx, ctx->pair64, (size_t)ctx->n_pairs * 24);
  if (rc) return rc;
  timed(ctx, "pairs_full_jacobian_f64", [&] {
    st::launch_contract_cart(ctx->stream, ctx->n_sites, ctx->n_pairs, ctx->n_soap, ctx->site_ptr.as<int>(),
                             ctx->G.as<double>(), ctx->n_sites, ctx->cart.as<double>(), ctx->pair64.as<double>());
  });
  HIPCHK(hipGetLastError());
  ctx->have_pair64 = true;
  return forces64(ctx);
}

int st_pairs_full_jacobian_noalloc(st_ctx* ctx, int32_t n_sites, int32_t n_pairs, int32_t n_soap, int32_t n_atoms,
                                   const int32_t* site_ptr, const double* G, const double* soap_cart_der,
                                   const int32_t* atom_ptr, const int32_t* atom_pairs, const double* vec, double* pair,
                                   double* forces, double* virial) {
  if (!ctx) return ST_ERR_ARG;
  if (n_sites < 1 || n_pairs < n_sites || n_soap < 1 || n_atoms < 1)
    return fail(ctx, ST_ERR_ARG, "st_pairs_full_jacobian_noalloc: bad sizes");
  if (n_atoms > ST_NOALLOC_MAX_ATOMS)
    return fail(ctx, ST_ERR_ARG, "st_pairs_full_jacobian_noalloc: n_atoms = " + std::to_string(n_atoms) +
                                     " exceeds ST_NOALLOC_MAX_ATOMS = " + std::to_string(ST_NOALLOC_MAX_ATOMS) +
                                     "; use st_pairs_full_jacobian");
  if (!site_ptr || !G || !soap_cart_der || !atom_ptr || !atom_pairs || !vec || !pair || !forces || !virial)
    return fail(ctx, ST_ERR_ARG, "st_pairs_full_jacobian_noalloc: every array is required");
  if ((size_t)n_soap * 8 > ctx->lds_max)
    return fail(ctx, ST_ERR_UNSUPPORTED, "st_pairs_full_jacobian_noalloc: G(i, :) exceeds the LDS");
  if (!ctx->scratch) return fail(ctx, ST_ERR_STATE, "st_pairs_full_jacobian_noalloc: no scratch on this device");
  HIPCHK(hipSetDevice(ctx->device));
  st::launch_contract_cart(ctx->stream, n_sites, n_pairs, n_soap, site_ptr, G, n_sites, soap_cart_der, pair);
  st::launch_forces_f64(ctx->stream, n_atoms, pair, atom_ptr, atom_pairs, vec, nullptr, nullptr, nullptr, forces,
                        ctx->scratch, virial);
  HIPCHK(hipGetLastError());
  return ST_OK;
}

int st_contract_resident(st_ctx* ctx, double sgn) {
  if (!ctx || !ctx->ct_ready)
    return fail(ctx, ST_ERR_STATE, "ERROR: soap_contract needs get_soap called with contract = .true. first");
  if (!ctx->have_G || ctx->G_n_soap != ctx->n_soap) return fail(ctx, ST_ERR_STATE, "dE/dq missing or of another size");
  HIPCHK(hipSetDevice(ctx->device));
  size_t sh = st::contract_lds(ctx->n_max, ctx->k_max, ctx->n_soap, ctx->n_soap_unc);
  if (sh > 64 * 1024) return fail(ctx, ST_ERR_UNSUPPORTED, "soap_contract working set exceeds the LDS");
  int rc = ensure(ctx, ctx->pair64, (size_t)ctx->n_pairs * 24);
  if (rc) return rc;
  timed(ctx, "contract_f64", [&] {
    st::launch_contract(ctx->stream, ctx->n_sites, ctx->n_max, ctx->l_max, ctx->k_max, ctx->n_soap, ctx->n_soap_unc,
                        ctx->site_ptr.as<int>(), ctx->skip.as<uint8_t>(),
                        ctx->compressed ? ctx->unc_ptr.as<int>() : nullptr, ctx->unc_e.as<int>(),
                        ctx->single ? ctx->umap.as<int>() : nullptr, ctx->uel.as<double>(),
                        ctx->P_i.as<int32_t>(), ctx->P_el.as<double>(), ctx->cnk.as<double2>(), ctx->rad.as<double>(),
                        ctx->rad_der.as<double>(), ctx->ang.as<double2>(), ctx->ang_r.as<double2>(),
                        ctx->ang_a.as<double2>(), ctx->ang_p.as<double2>(), ctx->soap.as<double>(), ctx->G.as<double>(),
                        sgn, ctx->sdp.as<double>(), ctx->rjs.as<double>(), ctx->thetas.as<double>(),
                        ctx->phis.as<double>(), ctx->pair64.as<double>());
  });
  HIPCHK(hipGetLastError());
  ctx->have_pair64 = true;
  return forces64(ctx);
}

int st_download(st_ctx* ctx, double* soap, double* cart) {
  if (!ctx || !ctx->have_soap) return fail(ctx, ST_ERR_STATE, "nothing computed");
  HIPCHK(hipSetDevice(ctx->device));
  if (soap)
    HIPCHK(hipMemcpyAsync(soap, ctx->soap.p, (size_t)ctx->n_sites * ctx->n_soap * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (cart) {
    if (!ctx->have_cart) return fail(ctx, ST_ERR_STATE, "soap_cart_der was not computed");
    HIPCHK(hipMemcpyAsync(cart, ctx->cart.p, (size_t)ctx->n_pairs * ctx->n_soap * 24, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return ST_OK;
}

int st_download_pairs(st_ctx* ctx, int32_t which, double* pair) {
  if (!ctx || !pair) return ST_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  if (which == 0) {
    if (!ctx->have_pair64) return fail(ctx, ST_ERR_STATE, "no FP64 pair gradients computed");
    HIPCHK(hipMemcpyAsync(pair, ctx->pair64.p, (size_t)ctx->n_pairs * 24, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ST_OK;
  }
  std::string msg;
  int rc = st::adjoint_download_pairs(ctx->adj, ctx->stream, ctx->n_pairs, pair, msg);
  return rc ? fail(ctx, rc, msg) : ST_OK;
}

int st_download_forces(st_ctx* ctx, int32_t which, double* forces, double* virial) {
  if (!ctx) return ST_ERR_ARG;
  if (!ctx->have_atoms) return fail(ctx, ST_ERR_STATE, "upload the pair atoms first (st_upload_atoms)");
  HIPCHK(hipSetDevice(ctx->device));
  std::string msg;
  int rc;
  if (which == 0) {
    if (!ctx->have_forces64) return fail(ctx, ST_ERR_STATE, "no FP64 forces computed");
    if (forces) HIPCHK(hipMemcpyAsync(forces, ctx->forces64.p, (size_t)ctx->n_atoms * 24, hipMemcpyDeviceToHost, ctx->stream));
    if (virial) HIPCHK(hipMemcpyAsync(virial, ctx->virial64.p, 72, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ST_OK;
  } else {
    rc = st::adjoint_download_forces(ctx->adj, ctx->stream, ctx->n_atoms, forces, virial, msg);
  }
  return rc ? fail(ctx, rc, msg) : ST_OK;
}

namespace {

// Everything the FP32 step's kernels take by value or by address: a captured graph of the step is
// valid while this key is unchanged (buffer contents may change; addresses, sizes and parameters not).
std::string adjoint_graph_key(const st_ctx* ctx, const st_params* p) {
  std::string k;
  auto add = [&](const void* x, size_t n) { k.append(static_cast<const char*>(x), n); };
  add(p, 6 * sizeof(int32_t));
  add(p->alpha_max, 4 * p->n_species);
  for (const double* a : {p->rcut_hard, p->rcut_soft, p->nf, p->global_scaling, p->atom_sigma_r,
                          p->atom_sigma_r_scaling, p->atom_sigma_t, p->atom_sigma_t_scaling, p->amplitude_scaling,
                          p->central_weight})
    add(a, 8 * p->n_species);
  k += ctx->basis_key;
  k += ctx->tables_key;
  const uint64_t gen = ctx->nb_gen;
  const int sizes[] = {ctx->n_sites, ctx->n_pairs, ctx->n_atoms, ctx->n_soap, ctx->G_n_soap, ctx->have_atoms,
                       ctx->have_vec, ctx->packed_species};
  add(&gen, sizeof gen);
  add(sizes, sizeof sizes);
  const void* bufs[] = {ctx->site_ptr.p, ctx->species.p, ctx->mult.p, ctx->spmask.p, ctx->rjs.p, ctx->thetas.p,
                        ctx->phis.p, ctx->comp.p, ctx->out_ptr.p, ctx->out_c.p, ctx->unc_ptr.p, ctx->unc_e.p,
                        ctx->P_i.p, ctx->pair_atom.p, ctx->atom_ptr.p, ctx->atom_pairs.p, ctx->vec.p,
                        ctx->adj.rtab.p, ctx->adj.atab.p, ctx->adj.tabinfo.p, ctx->adj.soap.p, ctx->adj.sdp.p,
                        ctx->adj.pair.p, ctx->adj.forces.p, ctx->adj.virial_part.p, ctx->adj.virial.p,
                        ctx->adj.G32.p, ctx->adj.order.p, ctx->adj.tickets.p, ctx->adj.el32.p, ctx->adj.odir.p,
                        ctx->adj.cenv.p};
  add(bufs, sizeof bufs);
  return k;
}

void drop_graph(st_ctx* ctx) {
  if (ctx->g32_exec) (void)hipGraphExecDestroy(ctx->g32_exec);
  if (ctx->g32_graph) (void)hipGraphDestroy(ctx->g32_graph);
  ctx->g32_exec = nullptr;
  ctx->g32_graph = nullptr;
  ctx->g32_key.clear();
}

}  // namespace

namespace {

int tuned_variant(st_ctx* ctx, const std::string& key, int n, const std::function<bool(int)>& launch) {
  auto it = ctx->tuned.find(key);
  if (it != ctx->tuned.end()) return it->second;
  int best = 0;
  if (n > 1) {
    hipEvent_t a = get_event(ctx), b = get_event(ctx);
    float tbest = 0.f;
    bool any = false;
    for (int v = 0; v < n; ++v) {
      if (!launch(v)) continue;
      (void)hipEventRecord(a, ctx->stream);
      launch(v);
      (void)hipEventRecord(b, ctx->stream);
      float ms = 0.f;
      if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) continue;
      if (!any || ms < tbest) {
        tbest = ms;
        best = v;
        any = true;
      }
    }
    ctx->free_events.push_back(a);
    ctx->free_events.push_back(b);
  }
  ctx->tuned[key] = best;
  return best;
}

}  // namespace

int st_set_jacobian_variant(st_ctx* ctx, int32_t variant) {
  if (!ctx) return ST_ERR_ARG;
  ctx->jacobian_variant = variant < 0 ? -1 : variant;
  return ST_OK;
}

int st_jacobian_variants(st_ctx* ctx, int32_t n_soap) {
  if (!ctx) return -1;
  return st::derivatives_variants(n_soap);
}

int st_set_autotune(st_ctx* ctx, int32_t on) {
  if (!ctx) return ST_ERR_ARG;
  ctx->autotune = on != 0;
  return ST_OK;
}

int st_set_graphs(st_ctx* ctx, int32_t on) {
  if (!ctx) return ST_ERR_ARG;
  ctx->graphs = on != 0;
  if (!ctx->graphs) {
    drop_graph(ctx);
    ctx->g32_warm.clear();
  }
  return ST_OK;
}

static int run_adjoint(st_ctx* ctx, const st_params* p, bool forward_only = false);

int st_adjoint_f32(st_ctx* ctx, const st_params* p, const st_compression* c) {
  if (!ctx) return ST_ERR_ARG;
  if (!ctx->have_nb) return fail(ctx, ST_ERR_STATE, "upload the neighbour list first");
  int rc = check_params(ctx, p);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  if ((rc = basis_for(ctx, p))) return rc;
  if ((rc = build_tables(ctx, p, c))) return rc;
  if (!ctx->have_G || ctx->G_n_soap != ctx->n_soap) return fail(ctx, ST_ERR_STATE, "dE/dq missing or of another size");
  if ((rc = pack_mask(ctx, p->n_species))) return rc;
  if (!ctx->graphs || ctx->timing) return run_adjoint(ctx, p);
  // graph mode: the first call with a key runs as usual (it builds tables and buffers), the second
  // records the step's launches, later ones replay them
  const std::string key = adjoint_graph_key(ctx, p);
  if (ctx->g32_exec && key == ctx->g32_key) {
    HIPCHK(hipGraphLaunch(ctx->g32_exec, ctx->stream));
    return ST_OK;
  }
  if (key != ctx->g32_warm) {
    ctx->g32_warm = key;
    return run_adjoint(ctx, p);
  }
  drop_graph(ctx);
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  rc = run_adjoint(ctx, p);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc || e != hipSuccess || !g) {
    // nothing ran: record the failure and run the step uncaptured
    if (g) (void)hipGraphDestroy(g);
    ctx->g32_warm.clear();
    (void)hipGetLastError();
    return run_adjoint(ctx, p);
  }
  ctx->g32_graph = g;
  HIPCHK(hipGraphInstantiate(&ctx->g32_exec, g, nullptr, nullptr, 0));
  ctx->g32_key = key;
  HIPCHK(hipGraphLaunch(ctx->g32_exec, ctx->stream));
  return ST_OK;
}

static int run_adjoint(st_ctx* ctx, const st_params* p, bool forward_only) {
  int rc;
  std::vector<double> cen;
  build_central(ctx, p, cen);
  st::AdjointInputs in;
  in.p = p;
  in.W = &ctx->W;
  in.central = &cen;
  in.n_sites = ctx->n_sites;
  in.n_pairs = ctx->n_pairs;
  in.n_max = ctx->n_max;
  in.k_max = ctx->k_max;
  in.l_max = ctx->l_max;
  in.n_soap = ctx->n_soap;
  in.n_soap_unc = ctx->n_soap_unc;
  in.site_ptr_h = &ctx->site_ptr_h;
  in.site_ptr = ctx->site_ptr.as<int>();
  in.species = ctx->species.as<int32_t>();
  in.max_mult = ctx->max_mult;
  in.mult = ctx->mult.as<int32_t>();
  in.spmask = ctx->spmask.as<uint32_t>();
  in.first = ctx->first.as<uint8_t>();
  in.rjs = ctx->rjs.as<double>();
  in.thetas = ctx->thetas.as<double>();
  in.phis = ctx->phis.as<double>();
  in.G = ctx->G.as<double>();
  in.comp = ctx->comp.as<int4>();
  in.out_ptr = ctx->compressed ? ctx->out_ptr.as<int>() : nullptr;
  in.out_c = ctx->out_c.as<int>();
  in.unc_ptr = ctx->unc_ptr.as<int>();
  in.unc_e = ctx->unc_e.as<int>();
  in.P_i = ctx->P_i.as<int32_t>();
  in.out_el_h = &ctx->out_el_h;
  in.P_el_h = &ctx->P_el_h;
  in.ocomp = ctx->single ? ctx->ocomp.as<int4>() : nullptr;
  in.umap = ctx->umap.as<int>();
  in.oel_h = &ctx->oel_h;
  in.uel_h = &ctx->uel_h;
  in.umap_h = &ctx->umap_h;
  in.tables_key = &ctx->tables_key;
  in.lds_max = ctx->lds_max;
  in.n_cu = ctx->n_cu;
  in.autotune = ctx->autotune && !ctx->timing;
  in.forward_only = forward_only;
  {
    int off = 0;
    for (int s = 0; s < p->n_species; ++s) {
      st::RadialConsts c = radial_consts(p, s, ctx->W, n_max_of(p), off, true);
      // the table is one species block on its own: offset 0, leading dimension alpha_max
      c.off = 0;
      c.n_max = p->alpha_max[s];
      c.species_bit = 1u;
      in.radial_tab.push_back(c);
      off += p->alpha_max[s];
    }
  }
  in.pair_atom = ctx->have_atoms ? ctx->pair_atom.as<int32_t>() : nullptr;
  in.atom_ptr = ctx->have_atoms ? ctx->atom_ptr.as<int>() : nullptr;
  in.atom_pairs = ctx->have_atoms ? ctx->atom_pairs.as<int>() : nullptr;
  in.vec = ctx->have_vec ? ctx->vec.as<double>() : nullptr;
  in.n_atoms = ctx->have_atoms ? ctx->n_atoms : 0;
  std::string msg;
  auto timer = [ctx](const char* name, const std::function<void()>& f) { timed(ctx, name, f); };
  rc = st::adjoint_run(ctx->adj, ctx->stream, in, timer, msg);
  if (rc) return fail(ctx, rc, msg);
  HIPCHK(hipGetLastError());
  return ST_OK;
}

// the device neighbour list of either cell form (st_build_neighbours, st_build_neighbours_cell)
static int build_neighbours(st_ctx* ctx, int32_t n_atoms, const double* positions, const int32_t* atom_species,
                            int32_t n_species, const st::Lattice& g, double rcut, int32_t n_sites,
                            const int32_t* site_atoms, int32_t* n_pairs) {
  if (n_atoms < 1 || n_sites < 1 || !positions || !atom_species || !site_atoms)
    return fail(ctx, ST_ERR_ARG, "positions, species and sites are required");
  if (n_species < 1 || n_species > ST_MAX_SPECIES) return fail(ctx, ST_ERR_ARG, "bad number of species");
  for (int a = 0; a < n_atoms; ++a)
    if (atom_species[a] < 1 || atom_species[a] > n_species) return fail(ctx, ST_ERR_ARG, "atom species out of range");
  for (int s = 0; s < n_sites; ++s)
    if (site_atoms[s] < 0 || site_atoms[s] >= n_atoms) return fail(ctx, ST_ERR_ARG, "site atom out of range");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int rc;
  const size_t n_cells = (size_t)g.n[0] * g.n[1] * g.n[2];
  if ((rc = upload(ctx, ctx->nb_pos, positions, (size_t)n_atoms * 24))) return rc;
  if ((rc = upload(ctx, ctx->nb_species, atom_species, (size_t)n_atoms * 4))) return rc;
  if ((rc = upload(ctx, ctx->nb_sites, site_atoms, (size_t)n_sites * 4))) return rc;
  if ((rc = ensure(ctx, ctx->nb_cell_of, (size_t)n_atoms * 4))) return rc;
  if ((rc = ensure(ctx, ctx->nb_cell_count, n_cells * 4))) return rc;
  if ((rc = ensure(ctx, ctx->nb_cell_start, (n_cells + 1) * 4))) return rc;
  if ((rc = ensure(ctx, ctx->nb_cursor, std::max(n_cells, (size_t)n_atoms) * 4))) return rc;
  if ((rc = ensure(ctx, ctx->nb_cell_atoms, (size_t)n_atoms * 4))) return rc;
  if ((rc = ensure(ctx, ctx->nb_count, std::max((size_t)n_sites, (size_t)n_atoms) * 4))) return rc;
  if ((rc = ensure(ctx, ctx->nb_flag, 4))) return rc;
  if ((rc = ensure(ctx, ctx->site_ptr, ((size_t)n_sites + 1) * 4))) return rc;
  st::launch_neighbour_bins(st, n_atoms, ctx->nb_pos.as<double>(), g, ctx->nb_cell_of.as<int>(),
                            ctx->nb_cell_count.as<int>(), ctx->nb_cell_start.as<int>(), ctx->nb_cursor.as<int>(),
                            ctx->nb_cell_atoms.as<int>());
  st::launch_neighbour_count(st, n_sites, ctx->nb_sites.as<int>(), ctx->nb_pos.as<double>(), g, rcut,
                             ctx->nb_cell_start.as<int>(), ctx->nb_cell_atoms.as<int>(), ctx->nb_count.as<int>(),
                             ctx->site_ptr.as<int>());
  std::vector<int> ptr((size_t)n_sites + 1);
  HIPCHK(hipMemcpyAsync(ptr.data(), ctx->site_ptr.p, ptr.size() * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const size_t P = ptr[n_sites];
  if (P > (size_t)INT32_MAX / 4) return fail(ctx, ST_ERR_UNSUPPORTED, "too many pairs");
  if ((rc = ensure(ctx, ctx->pair_atom, P * 4))) return rc;
  if ((rc = ensure(ctx, ctx->vec, P * 24))) return rc;
  if ((rc = ensure(ctx, ctx->rjs, P * 8))) return rc;
  if ((rc = ensure(ctx, ctx->thetas, P * 8))) return rc;
  if ((rc = ensure(ctx, ctx->phis, P * 8))) return rc;
  if ((rc = ensure(ctx, ctx->mask, P * n_species * 4))) return rc;
  if ((rc = ensure(ctx, ctx->species, (size_t)n_sites * 4))) return rc;
  if ((rc = ensure(ctx, ctx->mult, (size_t)n_sites * 4))) return rc;
  if ((rc = ensure(ctx, ctx->atom_ptr, ((size_t)n_atoms + 1) * 4))) return rc;
  if ((rc = ensure(ctx, ctx->atom_pairs, P * 4))) return rc;
  HIPCHK(hipMemsetAsync(ctx->nb_flag.p, 0, 4, st));
  st::launch_neighbour_fill(st, n_sites, (int)P, n_species, ctx->nb_sites.as<int>(), ctx->nb_species.as<int>(),
                            ctx->nb_pos.as<double>(), g, rcut, ctx->nb_cell_start.as<int>(),
                            ctx->nb_cell_atoms.as<int>(), ctx->site_ptr.as<int>(), ctx->pair_atom.as<int>(),
                            ctx->vec.as<double>(), ctx->rjs.as<double>(), ctx->thetas.as<double>(),
                            ctx->phis.as<double>(), ctx->mask.as<int32_t>(), ctx->species.as<int32_t>(),
                            ctx->mult.as<int32_t>(), ctx->nb_flag.as<int>());
  st::launch_atom_pairs(st, n_atoms, (int)P, ctx->pair_atom.as<int>(), ctx->nb_count.as<int>(),
                        ctx->nb_cursor.as<int>(), ctx->atom_ptr.as<int>(), ctx->atom_pairs.as<int>(),
                        ctx->nb_flag.as<int>());
  int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, ctx->nb_flag.p, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  // the resident state now describes the new list (as after st_upload_neighbours + st_upload_atoms)
  ctx->have_nb = ctx->have_atoms = false;
  ctx->have_soap = ctx->have_cart = ctx->ct_ready = ctx->have_pair64 = false;
  ctx->have_G = false;
  st::adjoint_invalidate(ctx->adj);
  ++ctx->nb_gen;
  if (flag & 1) return fail(ctx, ST_ERR_UNSUPPORTED, "a site has more than 1024 neighbours");
  if (flag & 2) return fail(ctx, ST_ERR_UNSUPPORTED, "an atom is the neighbour in more than 1024 pairs");
  ctx->n_sites = n_sites;
  ctx->n_pairs = (int)P;
  ctx->n_species_mask = n_species;
  ctx->packed_species = 0;
  ctx->max_mult = 1;
  ctx->site_ptr_h = ptr;
  ctx->have_nb = true;
  ctx->n_atoms = n_atoms;
  ctx->have_atoms = ctx->have_vec = true;
  ++ctx->nb_gen;
  if (n_pairs) *n_pairs = (int32_t)P;
  return ST_OK;
}

int st_build_neighbours(st_ctx* ctx, int32_t n_atoms, const double* positions, const int32_t* atom_species,
                        int32_t n_species, double box, double rcut, int32_t n_sites, const int32_t* site_atoms,
                        int32_t* n_pairs) {
  if (!ctx) return ST_ERR_ARG;
  if (!(rcut > 0.0) || !(box > 2.0 * rcut))
    return fail(ctx, ST_ERR_ARG, "the box side must exceed twice the cutoff (minimum image)");
  return build_neighbours(ctx, n_atoms, positions, atom_species, n_species, st::cubic_lattice(box, rcut), rcut,
                          n_sites, site_atoms, n_pairs);
}

int st_build_neighbours_cell(st_ctx* ctx, int32_t n_atoms, const double* positions, const int32_t* atom_species,
                             int32_t n_species, const double* cell, double rcut, int32_t n_sites,
                             const int32_t* site_atoms, int32_t* n_pairs) {
  if (!ctx) return ST_ERR_ARG;
  if (!cell) return fail(ctx, ST_ERR_ARG, "the lattice vectors are required");
  if (!(rcut > 0.0)) return fail(ctx, ST_ERR_ARG, "the cutoff must be positive");
  st::Lattice g;
  double width = 0.0;
  if (!st::general_lattice(cell, rcut, g, width))
    return fail(ctx, ST_ERR_ARG, "the lattice vectors must span a right-handed cell of positive volume");
  if (!(width > 2.0 * rcut))
    return fail(ctx, ST_ERR_ARG, "every width of the cell must exceed twice the cutoff (minimum image)");
  return build_neighbours(ctx, n_atoms, positions, atom_species, n_species, g, rcut, n_sites, site_atoms, n_pairs);
}

int st_download_neighbours(st_ctx* ctx, int32_t* n_neigh, int32_t* pair_atom, double* rjs, double* thetas,
                           double* phis, int32_t* mask, double* vec) {
  if (!ctx) return ST_ERR_ARG;
  if (!ctx->have_nb) return fail(ctx, ST_ERR_STATE, "no neighbour list");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t P = ctx->n_pairs;
  if (n_neigh)
    for (int s = 0; s < ctx->n_sites; ++s) n_neigh[s] = ctx->site_ptr_h[s + 1] - ctx->site_ptr_h[s];
  if (pair_atom) {
    if (!ctx->have_atoms) return fail(ctx, ST_ERR_STATE, "no pair atoms");
    HIPCHK(hipMemcpyAsync(pair_atom, ctx->pair_atom.p, P * 4, hipMemcpyDeviceToHost, st));
  }
  if (rjs) HIPCHK(hipMemcpyAsync(rjs, ctx->rjs.p, P * 8, hipMemcpyDeviceToHost, st));
  if (thetas) HIPCHK(hipMemcpyAsync(thetas, ctx->thetas.p, P * 8, hipMemcpyDeviceToHost, st));
  if (phis) HIPCHK(hipMemcpyAsync(phis, ctx->phis.p, P * 8, hipMemcpyDeviceToHost, st));
  if (mask) HIPCHK(hipMemcpyAsync(mask, ctx->mask.p, P * ctx->n_species_mask * 4, hipMemcpyDeviceToHost, st));
  if (vec) {
    if (!ctx->have_vec) return fail(ctx, ST_ERR_STATE, "no pair vectors");
    HIPCHK(hipMemcpyAsync(vec, ctx->vec.p, P * 24, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return ST_OK;
}

int st_atom_pairs(st_ctx* ctx, int32_t* atom_ptr, int32_t* atom_pairs) {
  if (!ctx) return ST_ERR_ARG;
  if (!ctx->have_atoms) return fail(ctx, ST_ERR_STATE, "no pair atoms");
  HIPCHK(hipSetDevice(ctx->device));
  if (atom_ptr)
    HIPCHK(hipMemcpyAsync(atom_ptr, ctx->atom_ptr.p, ((size_t)ctx->n_atoms + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (atom_pairs)
    HIPCHK(hipMemcpyAsync(atom_pairs, ctx->atom_pairs.p, (size_t)ctx->n_pairs * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return ST_OK;
}

int st_set_gap_model(st_ctx* ctx, int32_t n_sparse, int32_t n_soap, const double* sparse_q, const double* alpha,
                     int32_t zeta, double delta) {
  if (!ctx) return ST_ERR_ARG;
  if (n_sparse < 1 || n_soap < 1 || !sparse_q || !alpha || zeta < 1)
    return fail(ctx, ST_ERR_ARG, "bad GAP model (n_sparse, n_soap >= 1, zeta >= 1 and the arrays are required)");
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<float> T((size_t)n_sparse * n_soap), a(n_sparse);
  for (size_t k = 0; k < T.size(); ++k) T[k] = (float)sparse_q[k];  // (n_soap, n_sparse), point t contiguous
  for (int t = 0; t < n_sparse; ++t) a[t] = (float)alpha[t];
  int rc;
  if ((rc = upload(ctx, ctx->gap_T, T.data(), T.size() * 4))) return rc;
  if ((rc = upload(ctx, ctx->gap_alpha, a.data(), a.size() * 4))) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->gap_n_sparse = n_sparse;
  ctx->gap_n_soap = n_soap;
  ctx->gap_zeta = zeta;
  ctx->gap_delta = delta;
  return ST_OK;
}

int st_gap_energy_forces(st_ctx* ctx, const st_params* p, const st_compression* c, double* energy) {
  if (!ctx) return ST_ERR_ARG;
  if (!ctx->have_nb) return fail(ctx, ST_ERR_STATE, "upload the neighbour list first");
  if (!ctx->gap_n_sparse) return fail(ctx, ST_ERR_STATE, "set the GAP model first");
  int rc = check_params(ctx, p);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  if ((rc = basis_for(ctx, p))) return rc;
  if ((rc = build_tables(ctx, p, c))) return rc;
  if (ctx->gap_n_soap != ctx->n_soap) return fail(ctx, ST_ERR_ARG, "the GAP model's descriptor size differs from n_soap");
  if ((rc = pack_mask(ctx, p->n_species))) return rc;
  std::string msg;
  if ((rc = st::adjoint_prepare_dedq(ctx->adj, ctx->n_sites, ctx->n_soap, msg))) return fail(ctx, rc, msg);
  if ((rc = ensure(ctx, ctx->gap_esite, (size_t)std::max(ctx->n_sites, 1) * 8))) return rc;
  if ((rc = ensure(ctx, ctx->gap_energy, 8))) return rc;
  // descriptors, the model's energy and dE/dq on the device, then the adjoint pass for the forces
  if ((rc = run_adjoint(ctx, p, true))) return rc;
  timed(ctx, "gap_model", [&] {
    rc = st::gap_energy_grad(ctx->adj, ctx->stream, ctx->n_sites, ctx->n_soap, ctx->gap_n_sparse,
                             ctx->gap_T.as<float>(), ctx->gap_alpha.as<float>(), ctx->gap_zeta, ctx->gap_delta,
                             ctx->gap_esite.as<double>(), ctx->gap_energy.as<double>(), msg);
  });
  if (rc) return fail(ctx, rc, msg);
  if ((rc = run_adjoint(ctx, p, false))) return rc;
  ctx->have_G = false;  // dE/dq now holds the model's gradient, not an uploaded one
  if (energy) {
    HIPCHK(hipMemcpyAsync(energy, ctx->gap_energy.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return ST_OK;
}

int st_gap_site_energies(st_ctx* ctx, double* e_site) {
  if (!ctx || !e_site) return ST_ERR_ARG;
  if (!ctx->gap_esite.p) return fail(ctx, ST_ERR_STATE, "no GAP energies computed");
  HIPCHK(hipMemcpyAsync(e_site, ctx->gap_esite.p, (size_t)ctx->n_sites * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return ST_OK;
}

int st_central_hessian(st_ctx* ctx, const st_params* p, const st_compression* c, int32_t n_vec, const double* vecs,
                       double* hess) {
  if (!ctx) return ST_ERR_ARG;
  if (!ctx->have_nb) return fail(ctx, ST_ERR_STATE, "upload the neighbour list first");
  if (n_vec < 1 || !vecs || !hess) return fail(ctx, ST_ERR_ARG, "n_vec >= 1, vecs and hess are required");
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (p->basis == ST_BASIS_POLY3OPERATOR)
    return fail(ctx, ST_ERR_STOP,
                "ERROR: soap_turbo: the central-atom Hessian needs second radial derivatives, which the "
                "poly3operator basis does not provide. Use poly3 or poly3gauss.");
  if (p->legacy_filter_seed)
    return fail(ctx, ST_ERR_STOP,
                "ERROR: second radial derivatives need soap_turbo_radial's legacy_filter_seed = .false. .");
  // the descriptors, cnk and norms of the FP64 pass, then the radial expansion with both derivatives
  if ((rc = st_compute(ctx, p, c, 0, 0))) return rc;
  const int n_max = ctx->n_max, k_max = ctx->k_max, L = ctx->l_max, n_soap = ctx->n_soap, n_unc = ctx->n_soap_unc;
  const int NS = ctx->n_sites;
  const size_t P = ctx->n_pairs;
  hipStream_t st = ctx->stream;
  if ((rc = ensure(ctx, ctx->rad_der, P * n_max * 8))) return rc;
  if ((rc = ensure(ctx, ctx->rad_der2, P * n_max * 8))) return rc;
  st::launch_fill(st, ctx->rad.as<double>(), P * n_max, 0.0);
  st::launch_fill(st, ctx->rad_der.as<double>(), P * n_max, 0.0);
  st::launch_fill(st, ctx->rad_der2.as<double>(), P * n_max, 0.0);
  int off = 0;
  for (int s = 0; s < p->n_species; ++s) {
    st::RadialConsts rcst = radial_consts(p, s, ctx->W, n_max, off, true, true);
    bool ok = true;
    timed(ctx, "radial_d2_f64", [&] {
      ok = st::launch_radial(st, rcst, (int)P, ctx->rjs.as<double>(), ctx->spmask.as<uint32_t>(),
                             ctx->first.as<uint8_t>(), ctx->rad.as<double>(), ctx->rad_der.as<double>(),
                             ctx->rad_der2.as<double>());
    });
    if (!ok) return fail(ctx, ST_ERR_UNSUPPORTED, "alpha_max out of the compiled range");
    off += p->alpha_max[s];
  }
  // Omega = C_k Q_l^m(t) w^m carries the (-1)^m between P_l^m and Q_l^m
  std::vector<double> pf, ck(k_max);
  preflm_host(L, pf);
  for (int l = 0, k = 0; l <= L; ++l)
    for (int m = 0; m <= l; ++m, ++k) ck[k] = m % 2 == 0 ? pf[k] : -pf[k];
  if ((rc = upload(ctx, ctx->hess_ck, ck.data(), ck.size() * 8))) return rc;
  if ((rc = upload(ctx, ctx->hess_vecs, vecs, (size_t)NS * n_vec * n_soap * 8))) return rc;
  const size_t stride = st::hessian_scratch_per_site(n_max, k_max, n_soap, n_unc, n_vec);
  if ((rc = ensure(ctx, ctx->hess_scratch, (size_t)std::max(NS, 1) * stride * 8))) return rc;
  if ((rc = ensure(ctx, ctx->hess_out, P * 9 * n_vec * 8))) return rc;
  if ((rc = ensure(ctx, ctx->hess_heads, P * st::hessian_heads_per_pair() * 8))) return rc;
  if ((rc = ensure(ctx, ctx->hess_res, P * st::hessian_res_per_pair(n_vec) * 8))) return rc;
  if (st::hessian_lds(n_max, k_max, n_unc) > ctx->lds_max)
    return fail(ctx, ST_ERR_UNSUPPORTED, "the Hessian's working set exceeds the LDS");
  std::vector<double> sig(p->n_species), sigs(p->n_species);
  double rcut_max = 0.0;
  for (int s = 0; s < p->n_species; ++s) {
    sig[s] = p->atom_sigma_t[s];
    sigs[s] = p->atom_sigma_t_scaling[s];
    rcut_max = std::max(rcut_max, p->rcut_hard[s]);
  }
  bool ok = true;
  timed(ctx, "hessian_f64", [&] {
    ok = st::launch_hessian(st, NS, (int)P, n_max, L, n_soap, n_unc, n_vec, ctx->compressed ? 1 : 0, p->n_species,
                            sig.data(), sigs.data(), rcut_max, ctx->site_ptr.as<int>(), ctx->rjs.as<double>(),
                            ctx->thetas.as<double>(), ctx->phis.as<double>(), ctx->spmask.as<uint32_t>(),
                            ctx->first.as<uint8_t>(), ctx->hess_ck.as<double>(), ctx->rad.as<double>(),
                            ctx->rad_der.as<double>(),
                            ctx->rad_der2.as<double>(), ctx->cnk.as<double2>(), ctx->soap.as<double>(),
                            ctx->sdp.as<double>(), ctx->hess_vecs.as<double>(), ctx->comp.as<int4>(),
                            ctx->skip.as<uint8_t>(), ctx->compressed ? ctx->out_ptr.as<int>() : nullptr,
                            ctx->out_c.as<int>(), ctx->out_el.as<double>(), ctx->unc_ptr.as<int>(),
                            ctx->unc_e.as<int>(), ctx->P_i.as<int32_t>(), ctx->P_el.as<double>(),
                            ctx->hess_scratch.as<double>(), ctx->hess_heads.as<double>(), ctx->hess_res.as<double>(),
                            ctx->hess_out.as<double>());
  });
  if (!ok) return fail(ctx, ST_ERR_UNSUPPORTED, "l_max or n_species out of the Hessian's compiled range, or n_sites * (n_vec + 4) above 2^31 - 1");
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(hess, ctx->hess_out.p, P * 9 * n_vec * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  ctx->hess_n_vec = n_vec;
  ctx->hess_gen = ctx->nb_gen;
  return ST_OK;
}

int st_hessian_selftest(st_ctx* ctx, const st_params* p, const st_compression* c, int32_t n_vec, const double* vecs,
                        double* err) {
  if (!ctx) return ST_ERR_ARG;
  if (!vecs || !err) return fail(ctx, ST_ERR_ARG, "vecs and err are required");
  if (ctx->hess_n_vec != n_vec || ctx->hess_gen != ctx->nb_gen || !ctx->have_nb)
    return fail(ctx, ST_ERR_STATE, "the self-test needs st_central_hessian with these vectors on this list first");
  HIPCHK(hipSetDevice(ctx->device));
  const int NS = ctx->n_sites, nvz = n_vec + 4;
  const size_t P = ctx->n_pairs;
  // the Hessian route's first derivatives: per pair, v.dp/dr_jb for the caller's vectors and for p
  // (res), per site w.p and |p|; and whether a pair took part (its geometry head)
  std::vector<double> res(P * nvz * 12), wpr((size_t)NS * (3 + 4 * n_vec)), sdp(NS), head(P);
  const size_t stride = st::hessian_scratch_per_site(ctx->n_max, ctx->k_max, ctx->n_soap, ctx->n_soap_unc, n_vec);
  const size_t o_scal = st::hessian_scal_offset(ctx->n_max, ctx->k_max, ctx->n_soap, ctx->n_soap_unc, n_vec);
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemcpyAsync(res.data(), ctx->hess_res.p, res.size() * 8, hipMemcpyDeviceToHost, st));
  if (NS > 0) {
    HIPCHK(hipMemcpy2DAsync(wpr.data(), (3 + 4 * n_vec) * 8, ctx->hess_scratch.as<double>() + o_scal, stride * 8,
                            (3 + 4 * n_vec) * 8, NS, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sdp.data(), ctx->sdp.p, (size_t)NS * 8, hipMemcpyDeviceToHost, st));
  }
  if (P > 0)
    HIPCHK(hipMemcpy2DAsync(head.data(), 8, ctx->hess_heads.p, st::hessian_geom_head() * 8, 8, P,
                            hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // the independent first derivatives: the FP64 Jacobian of the same descriptor (soap_cart_der)
  int rc = st_compute(ctx, p, c, 1, 0);
  if (rc) return rc;
  const int n_soap = ctx->n_soap;
  std::vector<double> cart(P * n_soap * 3);
  HIPCHK(hipMemcpyAsync(cart.data(), ctx->cart.p, cart.size() * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // v.dq/dr_jb = (v.dp)/s - (v.p)(p.dp)/s^3 from the Hessian route, against sum_d v_d soap_cart_der(b, d, k)
  double num = 0.0, den = 0.0;
  for (int i = 0; i < NS; ++i) {
    const double s = sdp[i], s3 = s * s * s;
    for (int k2 = ctx->site_ptr_h[i] + 1; k2 < ctx->site_ptr_h[i + 1]; ++k2) {
      if (head[k2] == 0.0) continue;
      const double* rw = res.data() + (size_t)k2 * nvz * 12;
      for (int v = 0; v < n_vec; ++v) {
        const double wp = wpr[(size_t)i * (3 + 4 * n_vec) + 3 + 4 * v];
        const double* w = vecs + ((size_t)i * n_vec + v) * n_soap;
        const double* ck = cart.data() + (size_t)k2 * n_soap * 3;
        for (int b = 0; b < 3; ++b) {
          double ref = 0.0;
          for (int d = 0; d < n_soap; ++d) ref += w[d] * ck[3 * d + b];
          const double rec = rw[v * 12 + b] / s - wp * rw[n_vec * 12 + b] / s3;
          num = std::max(num, std::abs(rec - ref));
          den = std::max(den, std::abs(ref));
        }
      }
    }
  }
  *err = num / std::max(den, 1e-300);
  return ST_OK;
}

int st_download_soap_f32(st_ctx* ctx, double* soap) {
  if (!ctx || !soap) return ST_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  std::string msg;
  int rc = st::adjoint_download_soap(ctx->adj, ctx->stream, ctx->n_sites, ctx->n_soap, soap, msg);
  return rc ? fail(ctx, rc, msg) : ST_OK;
}

int st_get_soap(st_ctx* ctx, const st_params* p, const st_neighbours* nb, const st_compression* c,
                int32_t do_derivatives, int32_t contract, int32_t n_soap, double* soap, double* cart) {
  if (!ctx || !p || !nb || !soap) return fail(ctx, ST_ERR_ARG, "missing argument");
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (n_soap != n_soap_of(p, c)) return fail(ctx, ST_ERR_ARG, "soap has the wrong leading dimension");
  if ((rc = st_upload_neighbours(ctx, nb))) return rc;
  if ((rc = st_compute(ctx, p, c, do_derivatives, contract))) return rc;
  const bool want_cart = do_derivatives && !contract;
  if (want_cart && !cart) return fail(ctx, ST_ERR_ARG, "soap_cart_der is required with derivatives");
  return st_download(ctx, soap, want_cart ? cart : nullptr);
}

int st_soap_contract(st_ctx* ctx, const double* soap, const double* G, double sgn, double* pair) {
  if (!ctx) return ST_ERR_ARG;
  if (!ctx->ct_ready)
    return fail(ctx, ST_ERR_STATE, "ERROR: soap_contract needs get_soap called with contract = .true. first");
  if (!soap || !G || !pair) return fail(ctx, ST_ERR_ARG, "missing argument");
  HIPCHK(hipSetDevice(ctx->device));
  int rc;
  // the reference contracts against the soap it is handed, not its own copy
  if ((rc = upload(ctx, ctx->soap, soap, (size_t)ctx->n_sites * ctx->n_soap * 8))) return rc;
  if ((rc = upload(ctx, ctx->G, G, (size_t)ctx->n_sites * ctx->n_soap * 8))) return rc;
  std::string msg;
  if ((rc = st::adjoint_set_dedq(ctx->adj, ctx->stream, ctx->n_sites, ctx->n_soap, ctx->G.as<double>(), msg)))
    return fail(ctx, rc, msg);
  ctx->G_n_soap = ctx->n_soap;
  ctx->have_G = true;
  if ((rc = st_contract_resident(ctx, sgn))) return rc;
  return st_download_pairs(ctx, 0, pair);
}

namespace {

// "trivial": keep (n, m) when either is the first radial index of a species
// (src/soap_turbo_compress.f90:86-119)
void compress_trivial(int n_species, const int32_t* alpha_max, int l_max, int32_t* dim, int32_t* P_nonzero,
                      int32_t* P_i, int32_t* P_j, double* P_el, bool set) {
  std::vector<int> pivot(n_species);
  pivot[0] = 1;
  for (int i = 0; i + 1 < n_species; ++i) pivot[i + 1] = pivot[i] + alpha_max[i];
  int n_max = 0;
  for (int i = 0; i < n_species; ++i) n_max += alpha_max[i];
  int counter = 0, k = 1;
  for (int n = 1; n <= n_max; ++n)
    for (int m = n; m <= n_max; ++m)
      for (int l = 0; l <= l_max; ++l) {
        bool keep = false;
        for (int q : pivot) keep = keep || n == q || m == q;
        if (keep) {
          ++counter;
          if (set) {
            P_i[counter - 1] = counter;
            P_j[counter - 1] = k;
            P_el[counter - 1] = 1.0;
          }
        }
        ++k;
      }
  *dim = counter;
  *P_nonzero = counter;
}

// The nu_R_nu_S modes (src/soap_turbo_compress.f90:120-298): the combined radial index of
// (species z, radial n) is z alpha + n (every species has the same alpha). Component (i, j >= i, l)
// of the power spectrum, enumerated in the uncompressed order, goes to compressed channel
// (I, J, l) with I = (z_i mod S1) N1 + (n_i mod N1) and J = (z_j mod S2) N2 + (n_j mod N2), where
// N1 = alpha if nu_R > 0, N2 = alpha if nu_R = 2, S1 = n_species if nu_S > 0, S2 = n_species if
// nu_S = 2 (else 1). With nu_R and nu_S both even the compressed spectrum is symmetric: the pair
// (min, max) of I, J in triangular order, weight sqrt 2 for i != j falling on the diagonal. Else the
// (I, J) grid is kept whole, and an off-diagonal component contributes to both (I, J) and its
// mirror image (I', J') of (j, i), weight 1/sqrt 2 each (sqrt 2 when they coincide).
void compress_darby(int nu_R, int nu_S, int n_species, int alpha, int l_max, int32_t* dim, int32_t* P_nonzero,
                    int32_t* P_i, int32_t* P_j, double* P_el, bool set) {
  const int L1 = l_max + 1, nab = n_species * alpha;
  const int N1 = nu_R > 0 ? alpha : 1, N2 = nu_R == 2 ? alpha : 1;
  const int S1 = nu_S > 0 ? n_species : 1, S2 = nu_S == 2 ? n_species : 1;
  const bool sym = nu_R % 2 == 0 && nu_S % 2 == 0;
  const double r2 = std::sqrt(2.0), ir2 = 1.0 / std::sqrt(2.0);
  int counter = 0, ind = 0;
  auto put = [&](int comp, double w) {
    ++counter;
    if (set) {
      P_i[counter - 1] = comp + 1;
      P_j[counter - 1] = ind + 1;
      P_el[counter - 1] = w;
    }
  };
  for (int i = 0; i < nab; ++i) {
    const int zi = i / alpha, ni = i % alpha;
    for (int j = i; j < nab; ++j) {
      const int zj = j / alpha, nj = j % alpha;
      const int ic = (zi % S1) * N1 + ni % N1, jc = (zj % S2) * N2 + nj % N2;
      for (int l = 0; l < L1; ++l, ++ind) {
        if (sym) {
          const int a = std::min(ic, jc), b = std::max(ic, jc), K = N1 * S1;
          put(a * (2 * K + 1 - a) * L1 / 2 + (b - a) * L1 + l, i != j && ic == jc ? r2 : 1.0);
        } else {
          const int ic2 = (zj % S1) * N1 + nj % N1, jc2 = (zi % S2) * N2 + ni % N2;
          const int c1 = (ic * S2 * N2 + jc) * L1 + l, c2 = (ic2 * S2 * N2 + jc2) * L1 + l;
          if (i == j) {
            put(c1, 1.0);
          } else if (c1 == c2) {
            put(c1, r2);
          } else {
            put(c1, ir2);
            put(c2, ir2);
          }
        }
      }
    }
  }
  *dim = sym ? N1 * S1 * (N1 * S1 + 1) / 2 * L1 : N1 * S1 * N2 * S2 * L1;
  *P_nonzero = counter;
}

}  // namespace

int st_get_compress_indices(const char* mode, int32_t n_species, const int32_t* alpha_max, int32_t l_max,
                            int32_t* dim, int32_t* P_nonzero, int32_t* P_i, int32_t* P_j, double* P_el,
                            int32_t set_indices) {
  if (!mode || !alpha_max || !dim || !P_nonzero || n_species < 1 || l_max < 0) return ST_ERR_ARG;
  if (set_indices && (!P_i || !P_j || !P_el)) return ST_ERR_ARG;
  // Fortran string equality: trailing blanks do not count
  std::string m(mode);
  while (!m.empty() && m.back() == ' ') m.pop_back();
  if (m == "trivial" || m == "Trivial") {
    compress_trivial(n_species, alpha_max, l_max, dim, P_nonzero, P_i, P_j, P_el, set_indices != 0);
    return ST_OK;
  }
  if (m.size() == 3 && m[1] == '_' && m[0] >= '0' && m[0] <= '2' && m[2] >= '0' && m[2] <= '2') {
    // every species needs the same number of radial functions, or the reference stops
    for (int i = 1; i < n_species; ++i)
      if (alpha_max[i] != alpha_max[0]) return ST_ERR_STOP;
    compress_darby(m[0] - '0', m[2] - '0', n_species, alpha_max[0], l_max, dim, P_nonzero, P_i, P_j, P_el,
                   set_indices != 0);
    return ST_OK;
  }
  return ST_ERR_STOP;  // a mode the reference does not know
}

int st_synchronize(st_ctx* ctx) {
  if (!ctx) return ST_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return ST_OK;
}

int st_event_record(st_ctx* ctx, int32_t slot) {
  if (!ctx || slot < 0 || slot >= 16) return ST_ERR_ARG;
  HIPCHK(hipEventRecord(ctx->ev[slot], ctx->stream));
  return ST_OK;
}

int st_event_elapsed(st_ctx* ctx, int32_t a, int32_t b, float* ms) {
  if (!ctx || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) return ST_ERR_ARG;
  HIPCHK(hipEventSynchronize(ctx->ev[b]));
  HIPCHK(hipEventElapsedTime(ms, ctx->ev[a], ctx->ev[b]));
  return ST_OK;
}

int st_kernel_timing(st_ctx* ctx, int32_t on) {
  if (!ctx) return ST_ERR_ARG;
  resolve_timing(ctx);
  ctx->timing = on != 0;
  ctx->ktime.clear();
  return ST_OK;
}

int st_kernel_time(st_ctx* ctx, const char* name, double* total_ms, int32_t* launches) {
  if (!ctx || !name) return ST_ERR_ARG;
  resolve_timing(ctx);
  auto it = ctx->ktime.find(name);
  if (total_ms) *total_ms = it == ctx->ktime.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == ctx->ktime.end() ? 0 : it->second.second;
  return ST_OK;
}

}  // extern "C"
