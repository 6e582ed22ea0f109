// FP32 adjoint path: descriptors plus dE/dq -> pair gradients without forming
// soap_cart_der (the FP32 form of soap_contract, src/soap_turbo.f90:1703-1884).
//
// Set-up, once per basis (FP64):
//   * radial table: for every species, the reference's own radial function
//     c_n(r) and dc_n/dr (the FP64 radial kernel evaluated on a grid), stored
//     FP32 for cubic Hermite interpolation; the grid is split where c_n(r)
//     jumps (the 4-sigma cut of the filter term, src/soap_turbo_radial.f90:733,
//     and of the origin Gaussian, :882) so no cell straddles a jump;
//   * angular radial table: g_l(r) = (rcut_max/sigma(r))^2 i_l(x^2) e^{-x^2},
//     x = r/sigma(r), and dg_l/dr (get_ilexp, get_ilexp_der and the amplitude,
//     src/soap_turbo_angular.f90:361-463, :576-605), same layout.
// Per step, one workgroup per site (FP32 arithmetic):
//   forward:  per pair R_n(r) (tables) and A_k = preflm P_l^m(cos t) g_l(r) e^{-i m phi};
//             cnk = sum_pairs 4 pi R_n A_k with Kahan-compensated accumulation;
//             p(n,n',l) = sum_m mult Re(c c*), compression, q = p/|p|;
//   backward: Gp = (G - q (q.G))/|p|, back through the compression,
//             H(k,n) = w(m) sum_n' Y(n,n',l) c(k,n') (Y doubled on the diagonal);
//             per pair f_rad = 4 pi sum_k Re(A conj(v) + A_r conj(u)), f_azi, f_pol with
//             u_k = sum_n R_n H(k,n), v_k = sum_n R'_n H(k,n); rotation to Cartesian;
//             the site's own pair gets minus the sum of the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.h"
#include "kernels.h"

namespace st {

namespace {

constexpr int NTAB = 2048;        // grid cells over rcut for the tables
constexpr int MAXSEG = 3;
constexpr int ADJ_MAX_SPECIES = 8;
constexpr int PT = 64;            // pairs per tile
constexpr int NTHREADS = 256;

struct SpeciesTab {
  int a, off, do_central, nseg;
  double rcut;
  float cw;
  double seg_lo[MAXSEG], seg_inv_h[MAXSEG];
  float seg_hi[MAXSEG], seg_ih[MAXSEG];
  int seg_base[MAXSEG], seg_n[MAXSEG];   // first cell and node count of each segment
  int rtab_off;   // floats into rtab: cell-major, 4*a floats per cell (cubic coefficients c0..c3)
  int atab_off;   // floats into atab: cell-major, 4*(L+1) floats per cell
};

struct AdjParams {
  int n_species, n_max, k_max, l_max, n_soap, n_soap_unc, n_sites, compressed, gauss;
  double rcut_max, ang_inv_h;
  float ang_ih;
  int ang_n;
  SpeciesTab sp[ADJ_MAX_SPECIES];
};

int ensure(AdjBuf& b, size_t bytes, std::string& msg) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return ST_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    msg = std::string("hipMalloc: ") + hipGetErrorString(e);
    return ST_ERR_HIP;
  }
  b.bytes = bytes;
  return ST_OK;
}

// ------------------------------------------------------------ FP64 table builders
// g_l(r), dg_l/dr on nodes for one species (the angular part of the reference
// with its ilexp branches; amplitude rcut_max^2/sigma^2).
__global__ void ang_table_k(int n_nodes, double h, int L, double rcut2, double sig0, double scal,
                            double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const double rj = i * h;
  const double sigma = sig0 + scal * rj;
  const double amplitude = rcut2 / (sigma * sigma);
  double il[ST_MAX_LMAX + 2], ild[ST_MAX_LMAX + 2];
  const double xx = rj / sigma, xcut = 1e-7;
  double x2 = xx * xx, x4 = x2 * x2;
  double inv_x2 = x2 > 0.0 ? 1.0 / x2 : 0.0;
  double x2pow = x2 * x2;
  double flm2, flm1;
  if (xx > 0.0) {
    double e = exp(-2.0 * x2);
    flm2 = fabs((1.0 - e) / 2.0 / x2);
    flm1 = fabs((x2 - 1.0 + e * (x2 + 1.0)) / 2.0 / x4);
  } else {
    flm2 = 1.0;
    flm1 = 0.0;
  }
  double fact = 3.0;
  il[0] = xx < xcut ? 1.0 - x2 : flm2;
  il[1] = x2 / 1000.0 < xcut ? (x2 - x4) / fact : flm1;
  for (int l = 2; l <= L + 1; ++l) {
    fact = fact * (2.0 * (double)l + 1.0);
    double ratio = x2pow / fact, fl;
    if (ratio * (double)l < xcut) fl = ratio;
    else fl = fabs(flm2 - (2.0 * (double)l - 1.0) * inv_x2 * flm1);
    x2pow = x2pow * x2;
    flm2 = flm1;
    flm1 = fl;
    il[l] = fl;
  }
  double coeff1 = 2.0 * rj / (sigma * sigma), coeff2 = 1.0 - scal * rj / sigma;
  for (int l = 0; l <= L; ++l) ild[l] = 0.0;
  if (!(rj < 1e-5)) {
    ild[0] = coeff1 * ((L >= 1 ? il[1] : 0.0) - il[0]);
    for (int l = 1; l <= L; ++l) ild[l] = (-coeff1 - (double)(2 * l + 2) / rj) * il[l] + coeff1 * il[l - 1];
    for (int l = 0; l <= L; ++l) ild[l] = ild[l] * coeff2;
  }
  const double sig_fac = 2.0 * scal / sigma;
  double* o = out + (size_t)i * 2 * (L + 1);
  for (int l = 0; l <= L; ++l) {
    o[l] = amplitude * il[l];
    o[L + 1 + l] = amplitude * (ild[l] - sig_fac * il[l]);
  }
}

// ------------------------------------------------------------ device helpers
// Tables hold, per grid cell and function, the cubic c0 + c1 t + c2 t^2 + c3 t^3
// (t in [0, 1) across the cell) of the Hermite interpolant of the FP64 values
// and derivatives at the cell's ends; the coefficients are formed in FP64, so
// the FP32 derivative (c1 + 2 c2 t + 3 c3 t^2)/h has no cancellation.
__device__ __forceinline__ int radial_cell(const SpeciesTab& s, double r, float& t, float& ih) {
  int j = 0;
  while (j + 1 < s.nseg && r >= s.seg_lo[j + 1]) ++j;
  const double x = (r - s.seg_lo[j]) * s.seg_inv_h[j];
  int i = (int)x;
  if (i < 0) i = 0;
  if (i > s.seg_n[j] - 2) i = s.seg_n[j] - 2;
  t = (float)(x - (double)i);
  ih = s.seg_ih[j];
  return s.seg_base[j] + i;
}

__device__ __forceinline__ float block_sum_f(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  float s = 0.f;
  for (int j = 0; j < (int)(blockDim.x >> 6); ++j) s += red[j];
  return s;
}

// Angular quantities of one pair, row by row in l. For every (l, m) it calls
// f(k, l, m, Plm, dms, dsin) with preflm folded in, where dms = sin(t) dP/dcos(t)
// and dsin = -m P/sin(t) (the reference's two derivative recursions,
// src/soap_turbo_angular.f90:181-248).
template <int L, class F>
__device__ __forceinline__ void angular_rows(float x, float sq, const float* __restrict__ preflm, F&& f) {
  float pm1[L + 3], p0[L + 3], p1[L + 3];  // rows l-1, l, l+1 (index m)
#pragma unroll
  for (int m = 0; m < L + 3; ++m) pm1[m] = p0[m] = p1[m] = 0.f;
  p0[0] = 1.f;           // l = 0
  p1[0] = x; p1[1] = -sq; // l = 1
#pragma unroll
  for (int l = 0; l <= L; ++l) {
    // rows: pm1 = P_{l-1}, p0 = P_l, p1 = P_{l+1}
    const int kb = l * (l + 1) / 2;
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      float part1 = m == 0 ? (l == 0 ? 0.f : -0.5f * p0[1]) : 0.5f * (float)((l + m) * (l - m + 1)) * p0[m - 1];
      float part2 = m == l ? 0.f : -0.5f * p0[m + 1];
      float dms = part1 + part2;
      float dsin = 0.f;
      if (m >= 1) dsin = 0.5f * (float)((l - m + 1) * (l - m + 2)) * p1[m - 1] + 0.5f * p1[m + 1];
      const float pf = preflm[kb + m];
      f(kb + m, l, m, pf * p0[m], pf * dms, pf * dsin);
    }
    // advance: P_{l+2} from P_{l+1}, P_l
    const int ln = l + 2;
    float pn[L + 3];
#pragma unroll
    for (int m = 0; m < L + 3; ++m) pn[m] = 0.f;
    if (ln <= L + 1) {
#pragma unroll
      for (int m = 0; m <= ln - 2; ++m)
        pn[m] = ((float)(2 * ln - 1) * x * p1[m] - (float)(ln - 1 + m) * p0[m]) / (float)(ln - m);
      pn[ln - 1] = x * (float)(2 * ln - 1) * p1[ln - 1];
      pn[ln] = -(float)(2 * ln - 1) * sq * p1[ln - 1];
    }
#pragma unroll
    for (int m = 0; m < L + 3; ++m) { pm1[m] = p0[m]; p0[m] = p1[m]; p1[m] = pn[m]; }
  }
}

// P_l^m(x) for l <= L + 1 into out[l (l + 1) / 2 + m] (Condon-Shortley phase,
// the recursion of get_plm_array, src/soap_turbo_angular.f90:108-160).
template <int L>
__device__ __forceinline__ void plm_rows(float x, float sq, float* out) {
  float pm[L + 3], pc[L + 3];
#pragma unroll
  for (int m = 0; m < L + 3; ++m) pm[m] = pc[m] = 0.f;
  pm[0] = 1.f;             // l = 0
  pc[0] = x; pc[1] = -sq;  // l = 1
  out[0] = 1.f;
  out[1] = x;
  out[2] = -sq;
#pragma unroll
  for (int l = 2; l <= L + 1; ++l) {
    float pn[L + 3];
#pragma unroll
    for (int m = 0; m < L + 3; ++m) pn[m] = 0.f;
#pragma unroll
    for (int m = 0; m <= l - 2; ++m)
      pn[m] = ((float)(2 * l - 1) * x * pc[m] - (float)(l - 1 + m) * pm[m]) / (float)(l - m);
    pn[l - 1] = x * (float)(2 * l - 1) * pc[l - 1];
    pn[l] = -(float)(2 * l - 1) * sq * pc[l - 1];
#pragma unroll
    for (int m = 0; m <= l; ++m) out[l * (l + 1) / 2 + m] = pn[m];
#pragma unroll
    for (int m = 0; m < L + 3; ++m) { pm[m] = pc[m]; pc[m] = pn[m]; }
  }
}

// Per-pair radial values R_n(r) for the species blocks the pair belongs to
// (n in [nlo, nhi) only), and optionally dR_n/dr; zero elsewhere.
template <int NR>
__device__ __forceinline__ void radial_eval(const AdjParams& P, const float* __restrict__ rtab, double rd,
                                            uint32_t msk, bool self, int nlo, int nhi, float* Rn, float* Rd) {
#pragma unroll
  for (int j = 0; j < NR; ++j) { Rn[j] = 0.f; if (Rd) Rd[j] = 0.f; }
  for (int s = 0; s < P.n_species; ++s) {
    const SpeciesTab& sp = P.sp[s];
    if (!(msk & (1u << s)) || !(rd < sp.rcut)) continue;
    if (self && !sp.do_central) continue;
    const int lo = max(sp.off, nlo), hi = min(sp.off + sp.a, nhi);
    if (lo >= hi) continue;
    float t, ih;
    const int cell = radial_cell(sp, rd, t, ih);
    const float* c = rtab + sp.rtab_off + (size_t)cell * 4 * sp.a;
    const float sc = self ? sp.cw : 1.f;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int n = nlo + j;
      if (n >= lo && n < hi) {
        const int q = n - sp.off;
        const float c0 = c[q], c1 = c[sp.a + q], c2 = c[2 * sp.a + q], c3 = c[3 * sp.a + q];
        Rn[j] = sc * (c0 + t * (c1 + t * (c2 + t * c3)));
        if (Rd) Rd[j] = (c1 + t * (2.f * c2 + 3.f * t * c3)) * ih;
      }
    }
  }
}

// The pair's whole radial row R_n(r) (times scale) and optionally dR_n/dr,
// written to R[0..n_max) / Rd; each species' table cell is located once.
__device__ __forceinline__ void radial_row(const AdjParams& P, const float* __restrict__ rtab, double rd, uint32_t msk,
                                           bool self, float scale, float* R, float* Rd) {
  for (int n = 0; n < P.n_max; ++n) {
    R[n] = 0.f;
    if (Rd) Rd[n] = 0.f;
  }
  for (int s = 0; s < P.n_species; ++s) {
    const SpeciesTab& sp = P.sp[s];
    if (!(msk & (1u << s)) || !(rd < sp.rcut)) continue;
    if (self && !sp.do_central) continue;
    float t, ih;
    const int cell = radial_cell(sp, rd, t, ih);
    const float* c = rtab + sp.rtab_off + (size_t)cell * 4 * sp.a;
    const float sc = (self ? sp.cw : 1.f) * scale;
    for (int q = 0; q < sp.a; ++q) {
      const float c0 = c[q], c1 = c[sp.a + q], c2 = c[2 * sp.a + q], c3 = c[3 * sp.a + q];
      R[sp.off + q] = sc * (c0 + t * (c1 + t * (c2 + t * c3)));
      if (Rd) Rd[sp.off + q] = (c1 + t * (2.f * c2 + 3.f * t * c3)) * ih;
    }
  }
}

// Angular radial factors g_l(r) and dg_l/dr of the pair's first species.
template <int L>
__device__ __forceinline__ void angular_factors(const AdjParams& P, const float* __restrict__ atab, double rd, int s0,
                                                float* g, float* gd) {
  const double xa = rd * P.ang_inv_h;
  int ia = (int)xa;
  if (ia > P.ang_n - 2) ia = P.ang_n - 2;
  const float ta = (float)(xa - (double)ia);
  const float* gc = atab + P.sp[s0].atab_off + (size_t)ia * 4 * (L + 1);
#pragma unroll
  for (int l = 0; l <= L; ++l) {
    const float c0 = gc[l], c1 = gc[L + 1 + l], c2 = gc[2 * (L + 1) + l], c3 = gc[3 * (L + 1) + l];
    g[l] = c0 + ta * (c1 + ta * (c2 + ta * c3));
    gd[l] = (c1 + ta * (2.f * c2 + 3.f * ta * c3)) * P.ang_ih;
  }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x4f32: D(16x16) += A(16x4) B(4x16). Lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16]; it holds D[4 (l / 16) + v][l % 16] in register v.
__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// hi + lo += x without losing the rounding error of the add (TwoSum).
__device__ __forceinline__ void two_sum_acc(floatx4& hi, floatx4& lo, floatx4 x) {
  const floatx4 s = hi + x;
  const floatx4 bp = s - hi;
  lo += (hi - (s - bp)) + (x - bp);
  hi = s;
}

constexpr int MAX_TILES_PER_WAVE = 9;  // forward cnk tiles per wavefront: n_max <= 48, l_max <= 12

// One site per workgroup of 256 threads (4 wavefronts). The GEMM-shaped steps
// run on FP32 MFMA: cnk = sum_pairs 4 pi R^T A, the per-l Hermitian products
// S_l = C_l W C_l^T of the spectrum, H_l = Y_l C_l, and U = R H, V = R' H.
template <int L>
__global__ __launch_bounds__(NTHREADS, 2) void adjoint_site_k(
    AdjParams P, const int* __restrict__ order, const int* __restrict__ site_ptr, const float* __restrict__ rtab,
    const float* __restrict__ atab, const float* __restrict__ preflm_g, const float* __restrict__ central,
    const int32_t* __restrict__ site_species, int max_mult, const int32_t* __restrict__ species_mult,
    const uint32_t* __restrict__ spmask, const double* __restrict__ rjs, const double* __restrict__ thetas,
    const double* __restrict__ phis, const float* __restrict__ G32, const int4* __restrict__ comp,
    const int* __restrict__ out_ptr, const int* __restrict__ out_c, const float* __restrict__ out_el,
    const int* __restrict__ unc_ptr, const int* __restrict__ unc_e, const int32_t* __restrict__ P_i,
    const float* __restrict__ P_el, const int4* __restrict__ ocomp, const float* __restrict__ oel,
    const int* __restrict__ umap, const float* __restrict__ uel, float* __restrict__ q_out,
    float* __restrict__ sdp_out, float* __restrict__ pair_out) {
  constexpr int KM = (L + 1) * (L + 2) / 2;
  constexpr int KM1 = (L + 2) * (L + 3) / 2;
  constexpr int KS = KM1 | 1;                    // odd row stride of the per-pair P_l^m rows
  constexpr int L1 = L + 1;
  constexpr int L1S = L1 | 1;
  constexpr int CSTR = ((2 * KM + 15) / 16) * 16; // columns of C and H: 2k + {re, im}, padded
  constexpr int CT = CSTR / 16;
  extern __shared__ float smem[];
  __shared__ float red[16];
  __shared__ float preflm[KM];
  __shared__ int klm[KM];
  const int i = order ? order[blockIdx.x] : blockIdx.x;
  const int n_max = P.n_max;
  const int NT = (n_max + 15) / 16;
  const int p0 = site_ptr[i], p1 = site_ptr[i + 1];
  const float fourpi = 12.566370614359172f;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  float* C = smem;                                 // n_max * CSTR
  float* H = C + n_max * CSTR;                     // n_max * CSTR
  float* region = H + n_max * CSTR;
  float* Rt = region;                              // forward: PT * n_max
  float* At = Rt + PT * n_max;                     //          PT * CSTR
  float* Sl = region;                              // spectrum: L1 * n_max * n_max (then Y_l)
  float* qv = Sl + L1 * n_max * n_max;             //           n_soap
  float* Gp = qv + P.n_soap;                       //           n_soap
  float* Gu = Gp + P.n_soap;                       //           n_soap_unc
  float* Rb = region;                              // backward: PT * n_max (R)
  float* Rdb = Rb + PT * n_max;                    //           PT * n_max (dR/dr)
  float* Fp = Rdb + PT * n_max;                    //           PT * KS (P_l^m, l <= L + 1)
  float* Gl = Fp + PT * KS;                        //           PT * L1S (g_l)
  float* Gd = Gl + PT * L1S;                       //           PT * L1S (dg_l/dr)
  float2* Em = reinterpret_cast<float2*>(Gd + PT * L1S);  //    PT * L1S (e^{-i m phi})
  float* fout = reinterpret_cast<float*>(Em + PT * L1S);  //    PT * 3

  for (int t = threadIdx.x; t < KM; t += blockDim.x) {
    preflm[t] = preflm_g[t];
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= t) ++l;
    klm[t] = l | ((t - l * (l + 1) / 2) << 8);
  }
  // --------------------------------------------------------------- forward
  const int n_tiles = NT * CT;
  floatx4 hi[MAX_TILES_PER_WAVE], lo[MAX_TILES_PER_WAVE];
#pragma unroll
  for (int j = 0; j < MAX_TILES_PER_WAVE; ++j) {
    hi[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    lo[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  for (int tb = p0; tb < p1; tb += PT) {
    const int np_ = min(PT, p1 - tb);
    __syncthreads();
    {
      const int p = tb + lane;
      const bool live = lane < np_ && spmask[p] != 0 && rjs[p] < P.rcut_max;
      if (wv == 0) {
        // radial values of the pair, 4 pi folded in
        float* R = Rt + lane * n_max;
        if (live) {
          radial_row(P, rtab, rjs[p], spmask[p], p == p0, fourpi, R, nullptr);
        } else {
          for (int n = 0; n < n_max; ++n) R[n] = 0.f;
        }
      } else if (wv == 1) {
        // angular coefficients A_k = preflm P_l^m g_l e^{-i m phi}, as (re, im) columns
        float* A = At + lane * CSTR;
        if (live) {
          float g[L1], gd[L1];
          angular_factors<L>(P, atab, rjs[p], __ffs(spmask[p]) - 1, g, gd);
          float st_, ct_, sf_, cf_;
          sincosf((float)thetas[p], &st_, &ct_);
          sincosf((float)phis[p], &sf_, &cf_);
          float er[L1], ei[L1];
          er[0] = 1.f; ei[0] = 0.f;
#pragma unroll
          for (int m = 1; m <= L; ++m) {
            er[m] = er[m - 1] * cf_ + ei[m - 1] * sf_;
            ei[m] = ei[m - 1] * cf_ - er[m - 1] * sf_;
          }
          angular_rows<L>(ct_, st_, preflm, [&](int k, int l, int m, float pl, float, float) {
            const float v = pl * g[l];
            A[2 * k] = v * er[m];
            A[2 * k + 1] = v * ei[m];
          });
          for (int c = 2 * KM; c < CSTR; ++c) A[c] = 0.f;
        } else {
          for (int c = 0; c < CSTR; ++c) A[c] = 0.f;
        }
      }
    }
    __syncthreads();
    // cnk: the tile's pairs in chunks of 16 neighbours, each chunk on MFMA, chunks
    // combined with TwoSum compensation
    const int npad = (np_ + 3) & ~3;
#pragma unroll
    for (int j = 0; j < MAX_TILES_PER_WAVE; ++j) {
      const int t = wv + 4 * j;
      if (t < n_tiles) {
        const int a = t / CT, b = t - a * CT;
        const int n = 16 * a + li;
        const int col = 16 * b + li;
        for (int c0 = 0; c0 < npad; c0 += 16) {
          floatx4 acc = {0.f, 0.f, 0.f, 0.f};
          const int c1 = min(npad, c0 + 16);
          for (int pk = c0; pk < c1; pk += 4) {
            const int pp = pk + lk;
            const float av = n < n_max ? Rt[pp * n_max + n] : 0.f;
            const float bv = At[pp * CSTR + col];
            acc = mfma4(av, bv, acc);
          }
          two_sum_acc(hi[j], lo[j], acc);
        }
      }
    }
  }
  // C = hi + lo, plus the poly3gauss central-atom term on k = 0 (real part)
#pragma unroll
  for (int j = 0; j < MAX_TILES_PER_WAVE; ++j) {
    const int t = wv + 4 * j;
    if (t < n_tiles) {
      const int a = t / CT, b = t - a * CT;
      const int col = 16 * b + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = 16 * a + 4 * lk + v;
        if (n < n_max) {
          float val = hi[j][v] + lo[j][v];
          if (P.gauss && col == 0)
            for (int m = 0; m < species_mult[i]; ++m)
              val += central[(size_t)(site_species[(size_t)i * max_mult + m] - 1) * n_max + n];
          C[n * CSTR + col] = val;
        }
      }
    }
  }
  __syncthreads();
  // --------------------------------------------------------------- spectrum
  // S_l(n, n') = sum over the l block of w(m) C(n, c) C(n', c), w = 1 (m = 0), 2 (m > 0)
  for (int job = wv; job < L1 * NT * NT; job += 4) {
    const int l = job / (NT * NT), ab = job - l * NT * NT, a = ab / NT, b = ab - a * NT;
    const int cb = l * (l + 1), cw = 2 * (l + 1);
    const int n = 16 * a + li, np = 16 * b + li;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int c4 = 0; c4 < cw; c4 += 4) {
      const int c = c4 + lk;
      const bool inb = c < cw;
      const float av = (inb && n < n_max) ? C[n * CSTR + cb + c] : 0.f;
      const float bv = (inb && np < n_max) ? (c >= 2 ? 2.f : 1.f) * C[np * CSTR + cb + c] : 0.f;
      acc = mfma4(av, bv, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * a + 4 * lk + v, cc = 16 * b + li;
      if (r < n_max && cc < n_max) Sl[(l * n_max + r) * n_max + cc] = acc[v];
    }
  }
  __syncthreads();
  const int n_soap = P.n_soap;
  float ss = 0.f;
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x) {
    float v = 0.f;
    if (ocomp) {
      const int4 cm = ocomp[s];
      v = oel[s] * (cm.x != cm.y ? 1.41421356f : 1.f) * Sl[(cm.z * n_max + cm.x) * n_max + cm.y];
    } else {
      const int e0 = P.compressed ? out_ptr[s] : s, e1 = P.compressed ? out_ptr[s + 1] : s + 1;
      for (int e = e0; e < e1; ++e) {
        const int4 cm = comp[P.compressed ? out_c[e] : e];
        const float base = cm.x != cm.y ? 1.41421356f : 1.f;
        v += (P.compressed ? out_el[e] : 1.f) * base * Sl[(cm.z * n_max + cm.x) * n_max + cm.y];
      }
    }
    qv[s] = v;
    ss += v * v;
  }
  const float tot = block_sum_f(ss, red);
  float sdp = sqrtf(tot);
  if (sdp < 1e-5f) sdp = 1.f;
  const float inv = 1.f / sdp;
  float qg_part = 0.f;
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x) {
    const float q = qv[s] * inv;
    qv[s] = q;
    q_out[(size_t)i * n_soap + s] = q;
    const float g = G32[(size_t)i * n_soap + s];
    Gp[s] = g;
    qg_part += q * g;
  }
  if (threadIdx.x == 0) sdp_out[i] = sdp;
  const float qg = block_sum_f(qg_part, red + 4);
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x) Gp[s] = (Gp[s] - qv[s] * qg) * inv;
  __syncthreads();
  for (int c = threadIdx.x; c < P.n_soap_unc; c += blockDim.x) {
    float v;
    if (umap) {
      const int q = umap[c];
      v = q >= 0 ? uel[c] * Gp[q] : 0.f;
    } else if (P.compressed) {
      v = 0.f;
      for (int e = unc_ptr[c]; e < unc_ptr[c + 1]; ++e) v += P_el[unc_e[e]] * Gp[P_i[unc_e[e]] - 1];
    } else {
      v = Gp[c];
    }
    Gu[c] = v;
  }
  __syncthreads();
  // Y_l(n, n') = Gu(c) x 2 on the diagonal, sqrt 2 off it; c(a, b, l) = (a n_max - a(a-1)/2 + b - a) L1 + l
  float* Y = Sl;
  for (int t = threadIdx.x; t < L1 * n_max * n_max; t += blockDim.x) {
    const int l = t / (n_max * n_max), r = t - l * n_max * n_max, n = r / n_max, np = r - n * n_max;
    const int a = n < np ? n : np, b = n < np ? np : n;
    const float g = Gu[(a * n_max - a * (a - 1) / 2 + (b - a)) * L1 + l];
    Y[t] = a == b ? 2.f * g : 1.41421356f * g;
  }
  for (int t = threadIdx.x; t < n_max * (CSTR - 2 * KM); t += blockDim.x) {
    const int n = t / (CSTR - 2 * KM);
    H[n * CSTR + 2 * KM + (t - n * (CSTR - 2 * KM))] = 0.f;
  }
  __syncthreads();
  // H_l = 4 pi w(m) Y_l C_l, on the columns of block l
  {
    int njobs = 0;
    for (int l = 0; l <= L; ++l) njobs += NT * ((2 * (l + 1) + 15) / 16);
    for (int job = wv; job < njobs; job += 4) {
      int l = 0, rem = job;
      while (rem >= NT * ((2 * (l + 1) + 15) / 16)) { rem -= NT * ((2 * (l + 1) + 15) / 16); ++l; }
      const int nb_ = (2 * (l + 1) + 15) / 16;
      const int a = rem / nb_, b = rem - a * nb_;
      const int cb = l * (l + 1), cw = 2 * (l + 1);
      const int n = 16 * a + li, c = 16 * b + li;
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int k4 = 0; k4 < n_max; k4 += 4) {
        const int np = k4 + lk;
        const float av = (n < n_max && np < n_max) ? Y[(l * n_max + n) * n_max + np] : 0.f;
        const float bv = (np < n_max && c < cw) ? C[np * CSTR + cb + c] : 0.f;
        acc = mfma4(av, bv, acc);
      }
      if (c < cw) {
        const float w = (c >= 2 ? 2.f : 1.f) * fourpi;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = 16 * a + 4 * lk + v;
          if (r < n_max) H[r * CSTR + cb + c] = w * acc[v];
        }
      }
    }
  }
  // --------------------------------------------------------------- backward
  float self0 = 0.f, self1 = 0.f, self2 = 0.f;
  for (int tb = p0; tb < p1; tb += PT) {
    __syncthreads();
    {
      const int p = tb + lane;
      const bool live = p < p1 && p != p0 && spmask[p] != 0 && rjs[p] < P.rcut_max;
      if (wv == 0) {
        float* R = Rb + lane * n_max;
        float* Rd = Rdb + lane * n_max;
        if (live) {
          radial_row(P, rtab, rjs[p], spmask[p], false, 1.f, R, Rd);
        } else {
          for (int n = 0; n < n_max; ++n) { R[n] = 0.f; Rd[n] = 0.f; }
        }
      } else if (wv == 1) {
        if (live) {
          float st_, ct_;
          sincosf((float)thetas[p], &st_, &ct_);
          plm_rows<L>(ct_, st_, Fp + lane * KS);
        } else {
          for (int k = 0; k < KS; ++k) Fp[lane * KS + k] = 0.f;
        }
      } else if (wv == 2) {
        if (live) {
          float g[L1], gd[L1];
          angular_factors<L>(P, atab, rjs[p], __ffs(spmask[p]) - 1, g, gd);
#pragma unroll
          for (int l = 0; l <= L; ++l) { Gl[lane * L1S + l] = g[l]; Gd[lane * L1S + l] = gd[l]; }
          float sf_, cf_;
          sincosf((float)phis[p], &sf_, &cf_);
          float er = 1.f, ei = 0.f;
          Em[lane * L1S] = make_float2(1.f, 0.f);
          for (int m = 1; m <= L; ++m) {
            const float nr = er * cf_ + ei * sf_, ni = ei * cf_ - er * sf_;
            er = nr; ei = ni;
            Em[lane * L1S + m] = make_float2(er, ei);
          }
        } else {
#pragma unroll
          for (int l = 0; l <= L; ++l) { Gl[lane * L1S + l] = 0.f; Gd[lane * L1S + l] = 0.f; }
          for (int m = 0; m <= L; ++m) Em[lane * L1S + m] = make_float2(0.f, 0.f);
        }
      }
    }
    __syncthreads();
    // wavefront wv: pairs 16 wv .. 16 wv + 15 of the tile
    float fr[4] = {0.f, 0.f, 0.f, 0.f}, fa[4] = {0.f, 0.f, 0.f, 0.f}, fo[4] = {0.f, 0.f, 0.f, 0.f};
    const int pa = 16 * wv + li;  // A-operand row (pair) of this lane
    for (int ct = 0; ct < CT; ++ct) {
      floatx4 U = {0.f, 0.f, 0.f, 0.f}, V = {0.f, 0.f, 0.f, 0.f};
      const int col = 16 * ct + li;
      for (int k4 = 0; k4 < n_max; k4 += 4) {
        const int n = k4 + lk;
        const bool inb = n < n_max;
        const float hv = inb ? H[n * CSTR + col] : 0.f;
        U = mfma4(inb ? Rb[pa * n_max + n] : 0.f, hv, U);
        V = mfma4(inb ? Rdb[pa * n_max + n] : 0.f, hv, V);
      }
      // lane holds columns col (re if even, im if odd) of rows 16 wv + 4 lk + v
      const int k = col >> 1;
      const bool act = (col & 1) == 0 && k < KM;
      const int lm = act ? klm[k] : 0;
      const int l = lm & 255, m = lm >> 8;
      const int kb = l * (l + 1) / 2, kb1 = (l + 1) * (l + 2) / 2;
      const float pf = act ? preflm[k] : 0.f;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float ur = U[v], vr = V[v];
        const float ui = __shfl_xor(ur, 1, 64), vi = __shfl_xor(vr, 1, 64);
        if (act) {
          const int lp = 16 * wv + 4 * lk + v;
          const float* pr = Fp + lp * KS;
          const float gl = Gl[lp * L1S + l], gdl = Gd[lp * L1S + l];
          const float2 e = Em[lp * L1S + m];
          const float eu = e.x * ur + e.y * ui;     // Re(E conj u)
          const float ev = e.x * vr + e.y * vi;     // Re(E conj v)
          const float ex = e.x * ui - e.y * ur;     // Re(i E conj u)
          const float pl = pf * pr[kb + m];
          const float pp1 = m < l ? pr[kb + m + 1] : 0.f;
          float dms, dsin = 0.f;
          if (m == 0) {
            dms = -pp1;
          } else {
            dms = 0.5f * (float)((l + m) * (l - m + 1)) * pr[kb + m - 1] - 0.5f * pp1;
            dsin = 0.5f * (float)((l - m + 1) * (l - m + 2)) * pr[kb1 + m - 1] + 0.5f * pr[kb1 + m + 1];
          }
          // A = pl g E, A_r = pl g' E, A_pol = dms g E, A_azi = i dsin g E
          fr[v] += pl * (gl * ev + gdl * eu);
          fo[v] += pf * dms * gl * eu;
          fa[v] += pf * dsin * gl * ex;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        fr[v] += __shfl_xor(fr[v], o, 64);
        fa[v] += __shfl_xor(fa[v], o, 64);
        fo[v] += __shfl_xor(fo[v], o, 64);
      }
      if (li == 0) {
        const int lp = 16 * wv + 4 * lk + v;
        fout[lp * 3] = fr[v];
        fout[lp * 3 + 1] = fa[v];
        fout[lp * 3 + 2] = fo[v];
      }
    }
    __syncthreads();
    const int p = tb + lane;
    if (wv == 0 && p < p1 && p != p0) {
      const float f_rad = fout[lane * 3], f_azi = fout[lane * 3 + 1], f_pol = fout[lane * 3 + 2];
      float st_, ct_, sf_, cf_;
      sincosf((float)thetas[p], &st_, &ct_);
      sincosf((float)phis[p], &sf_, &cf_);
      const float inv_r = 1.f / (float)rjs[p];
      const float x0 = st_ * cf_ * f_rad - ct_ * cf_ * inv_r * f_pol - sf_ * inv_r * f_azi;
      const float x1 = st_ * sf_ * f_rad - ct_ * sf_ * inv_r * f_pol + cf_ * inv_r * f_azi;
      const float x2 = ct_ * f_rad + st_ * inv_r * f_pol;
      pair_out[3 * (size_t)p] = x0;
      pair_out[3 * (size_t)p + 1] = x1;
      pair_out[3 * (size_t)p + 2] = x2;
      self0 -= x0; self1 -= x1; self2 -= x2;
    }
  }
  self0 = block_sum_f(wv == 0 ? self0 : 0.f, red);
  self1 = block_sum_f(wv == 0 ? self1 : 0.f, red + 4);
  self2 = block_sum_f(wv == 0 ? self2 : 0.f, red + 8);
  if (threadIdx.x == 0) {
    pair_out[3 * (size_t)p0] = self0;
    pair_out[3 * (size_t)p0 + 1] = self1;
    pair_out[3 * (size_t)p0 + 2] = self2;
  }
}

// G(n_sites, n_soap) FP64 (site fastest) -> site-major FP32
__global__ void transpose_g_k(int n_sites, int n_soap, const double* __restrict__ G, float* __restrict__ G32) {
  __shared__ float tile[32][33];
  const int s0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int s = s0 + j, i = i0 + tx;
    tile[j][tx] = (s < n_soap && i < n_sites) ? (float)G[(size_t)s * n_sites + i] : 0.f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int i = i0 + j, s = s0 + tx;
    if (s < n_soap && i < n_sites) G32[(size_t)i * n_soap + s] = tile[tx][j];
  }
}

__global__ void widen_k(size_t n, const float* __restrict__ a, double* __restrict__ b) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

size_t site_lds(int n_max, int l_max, int n_soap, int n_soap_unc) {
  const int KM = (l_max + 1) * (l_max + 2) / 2, KS = ((l_max + 2) * (l_max + 3) / 2) | 1, L1S = (l_max + 1) | 1;
  const int CSTR = ((2 * KM + 15) / 16) * 16;
  size_t c = sizeof(float) * 2 * n_max * CSTR;
  size_t fwd = sizeof(float) * PT * (n_max + CSTR);
  size_t spec = sizeof(float) * ((l_max + 1) * n_max * n_max + 2 * n_soap + n_soap_unc);
  size_t bwd = sizeof(float) * PT * (2 * n_max + KS + 2 * L1S + 2 * L1S + 3);
  return c + std::max(std::max(fwd, spec), bwd);
}

template <int L>
void launch_site(hipStream_t st, const AdjParams& P, size_t sh, const int* order, const int* site_ptr,
                 const float* rtab, const float* atab, const float* preflm, const float* central,
                 const int32_t* species, int max_mult, const int32_t* mult, const uint32_t* spmask, const double* rjs,
                 const double* th, const double* ph, const float* G32, const int4* comp, const int* out_ptr,
                 const int* out_c, const float* out_el, const int* unc_ptr, const int* unc_e, const int32_t* P_i,
                 const float* P_el, const int4* ocomp, const float* oel, const int* umap, const float* uel, float* q,
                 float* sdp, float* pair) {
  adjoint_site_k<L><<<P.n_sites, NTHREADS, sh, st>>>(P, order, site_ptr, rtab, atab, preflm, central, species, max_mult,
                                                     mult, spmask, rjs, th, ph, G32, comp, out_ptr, out_c, out_el,
                                                     unc_ptr, unc_e, P_i, P_el, ocomp, oel, umap, uel, q, sdp, pair);
}

// host copy of the per-call constant tables that live next to the FP64 ones
struct HostTables {
  std::vector<float> preflm, central, out_el, P_el;
};

void preflm_f(int L, std::vector<float>& out) {
  out.clear();
  for (int l = 0; l <= L; ++l) {
    double ppi = std::sqrt((double)(2 * l + 1) / 4.0 / std::acos(-1.0));
    double f2 = 1.0;
    for (int i = 1; i <= l; ++i) f2 *= i;
    double f1 = f2;
    for (int m = 0; m <= l; ++m) {
      if (m > 0) { f1 /= (double)(l + 1 - m); f2 *= (double)(l + m); }
      out.push_back((float)(ppi * std::sqrt(f1 / f2)));
    }
  }
}

}  // namespace

// Cubic coefficients of the Hermite interpolant on [x0, x0 + h] from the FP64
// values and derivatives at both ends: f(t) = c0 + c1 t + c2 t^2 + c3 t^3.
static void cell_cubic(double v0, double d0, double v1, double d1, double h, float* c0, float* c1, float* c2,
                       float* c3) {
  const double a = d0 * h, b = d1 * h;
  *c0 = (float)v0;
  *c1 = (float)a;
  *c2 = (float)(3.0 * (v1 - v0) - 2.0 * a - b);
  *c3 = (float)(2.0 * (v0 - v1) + a + b);
}

// Builds (or reuses) the FP32 tables; fills P.
static int build_tables(AdjointState& s, hipStream_t st, const AdjointInputs& in, AdjParams& P, std::string& msg) {
  const st_params* p = in.p;
  const int L = in.l_max, ns = p->n_species;
  if (ns > ADJ_MAX_SPECIES) {
    msg = "the FP32 path supports up to 8 species";
    return ST_ERR_UNSUPPORTED;
  }
  std::memset(&P, 0, sizeof P);
  P.n_species = ns;
  P.n_max = in.n_max;
  P.k_max = in.k_max;
  P.l_max = L;
  P.n_soap = in.n_soap;
  P.n_soap_unc = in.n_soap_unc;
  P.n_sites = in.n_sites;
  P.compressed = in.out_ptr != nullptr;
  P.gauss = p->basis == ST_BASIS_POLY3GAUSS;
  double rcut_max = 0.0;
  for (int q = 0; q < ns; ++q) rcut_max = std::max(rcut_max, p->rcut_hard[q]);
  P.rcut_max = rcut_max;
  // key: everything the tables depend on
  std::string key;
  auto add = [&](const void* x, size_t n) { key.append(static_cast<const char*>(x), n); };
  add(p, 6 * sizeof(int32_t));
  add(p->alpha_max, 4 * ns);
  for (const double* arr : {p->rcut_hard, p->rcut_soft, p->nf, p->global_scaling, p->atom_sigma_r,
                            p->atom_sigma_r_scaling, p->atom_sigma_t, p->atom_sigma_t_scaling, p->amplitude_scaling,
                            p->central_weight})
    add(arr, 8 * ns);
  add(in.W->data(), in.W->size() * 8);
  // radial segments: split where c_n(r) jumps (in A: the filter's 4-sigma cut at
  // rcut_soft - r = 4 sigma(r), the origin Gaussian's at r = 4 (sigma_0 + sigma(r)))
  std::vector<std::vector<double>> edges(ns);
  int rfloats = 0;
  for (int q = 0; q < ns; ++q) {
    SpeciesTab& sp = P.sp[q];
    const double rc = p->rcut_hard[q];
    const double rs_n = p->rcut_soft[q] / rc, s0 = p->atom_sigma_r[q] / rc, c = p->atom_sigma_r_scaling[q];
    std::vector<double> br;
    if (1.0 - rs_n > 0.0) {
      double b = rc * (rs_n - 4.0 * s0) / (1.0 + 4.0 * c);
      if (b > 0.0 && b < rc) br.push_back(b);
    }
    if (P.gauss && c < 0.25) {
      double b = rc * 8.0 * s0 / (1.0 - 4.0 * c);
      if (b > 0.0 && b < rc) br.push_back(b);
    }
    std::sort(br.begin(), br.end());
    std::vector<double>& e = edges[q];
    e.push_back(0.0);
    for (double b : br)
      if (b - e.back() > 1e-6 * rc) e.push_back(b);
    e.push_back(rc);
    sp.a = p->alpha_max[q];
    sp.rcut = rc;
    sp.cw = (float)p->central_weight[q];
    sp.do_central = P.gauss ? 0 : (p->central_weight[q] != 0.0);
    sp.nseg = (int)e.size() - 1;
    sp.rtab_off = rfloats;
    int cells = 0;
    for (int j = 0; j < sp.nseg; ++j) {
      const double lo = e[j], hi = e[j + 1];
      const int nn = std::max(8, (int)std::ceil((hi - lo) / rc * NTAB) + 1);
      sp.seg_lo[j] = lo;
      sp.seg_inv_h[j] = (nn - 1) / (hi - lo);
      sp.seg_ih[j] = (float)((nn - 1) / (hi - lo));
      sp.seg_hi[j] = (float)hi;
      sp.seg_base[j] = cells;
      sp.seg_n[j] = nn;
      cells += nn - 1;
    }
    rfloats += cells * 4 * sp.a;
  }
  int off = 0;
  for (int q = 0; q < ns; ++q) { P.sp[q].off = off; off += p->alpha_max[q]; }
  P.ang_n = NTAB + 1;
  P.ang_inv_h = NTAB / rcut_max;
  P.ang_ih = (float)(NTAB / rcut_max);
  for (int q = 0; q < ns; ++q) P.sp[q].atab_off = q * (P.ang_n - 1) * 4 * (L + 1);
  if (key == s.table_key) return ST_OK;

  int rc_;
  std::vector<float> rt(rfloats);
  for (int q = 0; q < ns; ++q) {
    SpeciesTab& sp = P.sp[q];
    const double rc = p->rcut_hard[q];
    const double eps = 1e-9 * rc;
    const int a = sp.a;
    int nn = 0;
    for (int j = 0; j < sp.nseg; ++j) nn += sp.seg_n[j];
    // nodes, evaluated strictly inside their segment (beyond the jump below, before the one above)
    std::vector<double> r(nn), hseg(sp.nseg);
    int k0 = 0;
    for (int j = 0; j < sp.nseg; ++j) {
      const double lo = edges[q][j], hi = edges[q][j + 1];
      hseg[j] = (hi - lo) / (sp.seg_n[j] - 1);
      for (int k = 0; k < sp.seg_n[j]; ++k) {
        double x = lo + (hi - lo) * k / (sp.seg_n[j] - 1);
        if (j > 0) x = std::max(x, lo + eps);
        x = std::min(x, hi - eps);
        r[k0 + k] = x;
      }
      k0 += sp.seg_n[j];
    }
    AdjBuf dr, dsp, dfirst, drad, dder;
    std::vector<uint32_t> m(nn, 1u);
    std::vector<uint8_t> f(nn, 0);
    bool ok = ensure(dr, nn * 8, msg) == 0 && ensure(dsp, nn * 4, msg) == 0 && ensure(dfirst, nn, msg) == 0 &&
              ensure(drad, (size_t)nn * a * 8, msg) == 0 && ensure(dder, (size_t)nn * a * 8, msg) == 0;
    std::vector<double> v((size_t)nn * a), d((size_t)nn * a);
    if (ok) {
      (void)hipMemcpyAsync(dr.p, r.data(), nn * 8, hipMemcpyHostToDevice, st);
      (void)hipMemcpyAsync(dsp.p, m.data(), nn * 4, hipMemcpyHostToDevice, st);
      (void)hipMemcpyAsync(dfirst.p, f.data(), nn, hipMemcpyHostToDevice, st);
      launch_fill(st, drad.as<double>(), (size_t)nn * a, 0.0);
      launch_fill(st, dder.as<double>(), (size_t)nn * a, 0.0);
      RadialConsts c = in.radial_tab[q];
      ok = launch_radial(st, c, nn, dr.as<double>(), dsp.as<uint32_t>(), dfirst.as<uint8_t>(), drad.as<double>(),
                         dder.as<double>());
      (void)hipMemcpyAsync(v.data(), drad.p, v.size() * 8, hipMemcpyDeviceToHost, st);
      (void)hipMemcpyAsync(d.data(), dder.p, d.size() * 8, hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) { msg = "radial table build failed"; ok = false; }
    }
    for (AdjBuf* b : {&dr, &dsp, &dfirst, &drad, &dder})
      if (b->p) (void)hipFree(b->p);
    if (!ok) {
      if (msg.empty()) msg = "radial table build failed";
      return ST_ERR_HIP;
    }
    k0 = 0;
    for (int j = 0; j < sp.nseg; ++j) {
      for (int k = 0; k + 1 < sp.seg_n[j]; ++k) {
        float* cc = rt.data() + sp.rtab_off + (size_t)(sp.seg_base[j] + k) * 4 * a;
        const size_t n0 = (size_t)(k0 + k) * a, n1 = n0 + a;
        for (int n = 0; n < a; ++n)
          cell_cubic(v[n0 + n], d[n0 + n], v[n1 + n], d[n1 + n], hseg[j], cc + n, cc + a + n, cc + 2 * a + n,
                     cc + 3 * a + n);
      }
      k0 += sp.seg_n[j];
    }
  }
  if ((rc_ = ensure(s.rtab, rt.size() * 4, msg))) return rc_;
  (void)hipMemcpyAsync(s.rtab.p, rt.data(), rt.size() * 4, hipMemcpyHostToDevice, st);
  // angular radial factors, uniform grid on [0, rcut_max]
  const int L1 = L + 1, ncell = P.ang_n - 1;
  std::vector<float> at((size_t)ns * ncell * 4 * L1);
  {
    AdjBuf dout;
    if ((rc_ = ensure(dout, (size_t)P.ang_n * 2 * L1 * 8, msg))) return rc_;
    std::vector<double> h((size_t)P.ang_n * 2 * L1);
    const double hh = rcut_max / NTAB;
    for (int q = 0; q < ns; ++q) {
      ang_table_k<<<(P.ang_n + 127) / 128, 128, 0, st>>>(P.ang_n, hh, L, rcut_max * rcut_max, p->atom_sigma_t[q],
                                                          p->atom_sigma_t_scaling[q], dout.as<double>());
      (void)hipMemcpyAsync(h.data(), dout.p, h.size() * 8, hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) {
        (void)hipFree(dout.p);
        msg = "angular table build failed";
        return ST_ERR_HIP;
      }
      for (int k = 0; k < ncell; ++k) {
        float* cc = at.data() + P.sp[q].atab_off + (size_t)k * 4 * L1;
        const double* a0 = h.data() + (size_t)k * 2 * L1;
        const double* a1 = a0 + 2 * L1;
        for (int l = 0; l < L1; ++l)
          cell_cubic(a0[l], a0[L1 + l], a1[l], a1[L1 + l], hh, cc + l, cc + L1 + l, cc + 2 * L1 + l, cc + 3 * L1 + l);
      }
    }
    (void)hipFree(dout.p);
  }
  if ((rc_ = ensure(s.atab, at.size() * 4, msg))) return rc_;
  (void)hipMemcpyAsync(s.atab.p, at.data(), at.size() * 4, hipMemcpyHostToDevice, st);
  // constant tables
  std::vector<float> pf;
  preflm_f(L, pf);
  std::vector<float> cen(in.central->begin(), in.central->end());
  if ((rc_ = ensure(s.tabinfo, (pf.size() + cen.size()) * 4, msg))) return rc_;
  (void)hipMemcpyAsync(s.tabinfo.p, pf.data(), pf.size() * 4, hipMemcpyHostToDevice, st);
  (void)hipMemcpyAsync(s.tabinfo.as<float>() + pf.size(), cen.data(), cen.size() * 4, hipMemcpyHostToDevice, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    msg = "table upload failed";
    return ST_ERR_HIP;
  }
  s.table_key = key;
  s.n_pf = (int)pf.size();
  return ST_OK;
}

int adjoint_run(AdjointState& s, hipStream_t st, const AdjointInputs& in, const Timer& timer, std::string& msg) {
  AdjParams P;
  int rc;
  timer("adj_tables", [&] { rc = build_tables(s, st, in, P, msg); });
  if (rc) return rc;
  const size_t sh = site_lds(in.n_max, in.l_max, in.n_soap, in.n_soap_unc);
  if (sh > 64 * 1024) {
    msg = "the FP32 site working set exceeds the LDS";
    return ST_ERR_UNSUPPORTED;
  }
  if (in.n_max > 48) {
    msg = "the FP32 path supports n_max <= 48";
    return ST_ERR_UNSUPPORTED;
  }
  // compression element tables in FP32
  // FP32 copies: compression elements by output and in P order, then the direct-map elements
  size_t n_oe = 0, n_pe = 0;
  if (in.out_ptr) { n_oe = in.out_el_h->size(); n_pe = in.P_el_h->size(); }
  const size_t n_o1 = in.ocomp ? in.oel_h->size() : 0, n_u1 = in.ocomp ? in.uel_h->size() : 0;
  if (s.el_key != *in.tables_key) {
    std::vector<float> all;
    if (in.out_ptr) {
      all.insert(all.end(), in.out_el_h->begin(), in.out_el_h->end());
      all.insert(all.end(), in.P_el_h->begin(), in.P_el_h->end());
    }
    if (in.ocomp) {
      all.insert(all.end(), in.oel_h->begin(), in.oel_h->end());
      all.insert(all.end(), in.uel_h->begin(), in.uel_h->end());
    }
    if ((rc = ensure(s.el32, (all.size() + 1) * 4, msg))) return rc;
    if (!all.empty()) (void)hipMemcpyAsync(s.el32.p, all.data(), all.size() * 4, hipMemcpyHostToDevice, st);
    (void)hipStreamSynchronize(st);
    s.el_key = *in.tables_key;
  }
  if ((rc = ensure(s.G32, (size_t)in.n_sites * in.n_soap * 4, msg))) return rc;
  if ((rc = ensure(s.soap, (size_t)in.n_sites * in.n_soap * 4, msg))) return rc;
  if ((rc = ensure(s.sdp, (size_t)in.n_sites * 4, msg))) return rc;
  if ((rc = ensure(s.pair, (size_t)in.n_pairs * 12, msg))) return rc;
  timer("adj_transpose_g", [&] {
    dim3 grid((in.n_soap + 31) / 32, (in.n_sites + 31) / 32);
    transpose_g_k<<<grid, 256, 0, st>>>(in.n_sites, in.n_soap, in.G, s.G32.as<float>());
  });
  const float* pf = s.tabinfo.as<float>();
  const float* cen = pf + s.n_pf;
  const float* oel = in.out_ptr ? s.el32.as<float>() : nullptr;
  const float* pel = in.out_ptr ? s.el32.as<float>() + n_oe : nullptr;
  const float* oel1 = in.ocomp ? s.el32.as<float>() + n_oe + n_pe : nullptr;
  const float* uel1 = in.ocomp ? oel1 + n_o1 : nullptr;
  (void)n_u1;
  bool ok = true;
  timer("adj_site_f32", [&] {
    switch (in.l_max) {
#define LS(l)                                                                                                      \
  case l:                                                                                                          \
    launch_site<l>(st, P, sh, nullptr, in.site_ptr, s.rtab.as<float>(), s.atab.as<float>(), pf, cen, in.species,   \
                   in.max_mult, in.mult, in.spmask, in.rjs, in.thetas, in.phis, s.G32.as<float>(), in.comp,        \
                   in.out_ptr, in.out_c, oel, in.unc_ptr, in.unc_e, in.P_i, pel, in.ocomp, oel1,                   \
                   in.ocomp ? in.umap : nullptr, uel1, s.soap.as<float>(), s.sdp.as<float>(), s.pair.as<float>());  \
    break;
      LS(0) LS(1) LS(2) LS(3) LS(4) LS(5) LS(6) LS(7) LS(8) LS(9) LS(10) LS(11) LS(12)
#undef LS
      default: ok = false;
    }
  });
  if (!ok) {
    msg = "l_max out of the compiled range";
    return ST_ERR_UNSUPPORTED;
  }
  s.have_out = true;
  s.n_pairs = in.n_pairs;
  s.n_sites = in.n_sites;
  s.n_soap = in.n_soap;
  s.have_forces = false;
  if (in.pair_atom) {
    if ((rc = ensure(s.forces, (size_t)in.n_atoms * 24, msg))) return rc;
    if ((rc = ensure(s.virial_part, (size_t)SCATTER_BLOCKS * 72, msg))) return rc;
    if ((rc = ensure(s.virial, 72, msg))) return rc;
    timer("adj_forces", [&] {
      launch_scatter_f32(st, in.n_pairs, in.n_atoms, s.pair.as<float>(), in.pair_atom, in.vec, in.rjs, in.thetas,
                         in.phis, s.forces.as<double>(), s.virial_part.as<double>(), s.virial.as<double>());
    });
    s.have_forces = true;
    s.n_atoms = in.n_atoms;
  }
  return ST_OK;
}

int adjoint_download_pairs(AdjointState& s, hipStream_t st, int n_pairs, double* pair, std::string& msg) {
  if (!s.have_out) {
    msg = "no FP32 pair gradients computed";
    return ST_ERR_STATE;
  }
  AdjBuf tmp;
  int rc = ensure(tmp, (size_t)n_pairs * 24, msg);
  if (rc) return rc;
  widen_k<<<(unsigned)((3 * (size_t)n_pairs + 255) / 256), 256, 0, st>>>(3 * (size_t)n_pairs, s.pair.as<float>(),
                                                                         tmp.as<double>());
  (void)hipMemcpyAsync(pair, tmp.p, (size_t)n_pairs * 24, hipMemcpyDeviceToHost, st);
  hipError_t e = hipStreamSynchronize(st);
  (void)hipFree(tmp.p);
  if (e != hipSuccess) {
    msg = "copy of pair gradients failed";
    return ST_ERR_HIP;
  }
  return ST_OK;
}

int adjoint_download_soap(AdjointState& s, hipStream_t st, int n_sites, int n_soap, double* soap, std::string& msg) {
  if (!s.have_out) {
    msg = "no FP32 descriptors computed";
    return ST_ERR_STATE;
  }
  const size_t n = (size_t)n_sites * n_soap;
  AdjBuf tmp;
  int rc = ensure(tmp, n * 8, msg);
  if (rc) return rc;
  widen_k<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, s.soap.as<float>(), tmp.as<double>());
  (void)hipMemcpyAsync(soap, tmp.p, n * 8, hipMemcpyDeviceToHost, st);
  hipError_t e = hipStreamSynchronize(st);
  (void)hipFree(tmp.p);
  if (e != hipSuccess) {
    msg = "copy of descriptors failed";
    return ST_ERR_HIP;
  }
  return ST_OK;
}

}  // namespace st
