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
#include <numeric>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.h"
#include "kernels.h"

namespace st {

namespace {

constexpr int NTAB = 2048;        // grid cells over rcut for the tables
constexpr int MAXSEG = 8;  // radial table segments per species (poly3operator: up to 7)
constexpr int ADJ_MAX_SPECIES = 8;
static_assert(sizeof(AdjointState::rt_off) / sizeof(int) >= ADJ_MAX_SPECIES, "table offsets per species");
constexpr int PT = 64;            // pairs per tile
constexpr int NTHREADS = 256;
// set-up items are (row, pair) = (wavefront + 4 u, lane): a tile's pairs are one wavefront's lanes
static_assert(PT == 64 && NTHREADS == 256, "tile pairs map to the lanes of a wavefront");

struct SpeciesTab {
  int a, off, do_central, nseg;
  double rcut;
  float cw;
  double seg_lo[MAXSEG], seg_inv_h[MAXSEG];
  float seg_hi[MAXSEG], seg_ih[MAXSEG];
  int seg_base[MAXSEG], seg_n[MAXSEG];   // first cell and node count of each segment
  int rtab_off;   // floats into rtab: per cell, per n, the cubic's c0..c3 (one float4; 4*a floats per cell)
  int atab_off;   // floats into atab: per cell, per l, c0..c3 (one float4; 4*(L+1) floats per cell)
};

struct AdjParams {
  int n_species, n_max, k_max, l_max, n_soap, n_soap_unc, n_sites, compressed, gauss;
  // radial set-up items: each species block split into nch chunks of at most qc functions;
  // keep: one item per thread and qc <= the kept-register count (the backward pass reuses them)
  int nch, qc, keep;
  double rcut_max;
  float ang_ih;
  int ang_n;
  SpeciesTab sp[ADJ_MAX_SPECIES];
};
static_assert(sizeof(AdjParams) <= 4096, "kernel arguments are limited to 4 KB");

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
// The segment is the last whose lower edge r reaches (edges increase). s is uniform over the
// wavefront: its fields are read with uniform indices and the lane's segment selected from them (a
// lane-indexed read of the kernel argument is a vector memory load, which waits behind every earlier
// load and store of the wavefront).
__device__ __forceinline__ int radial_cell(const SpeciesTab& s, double r, float& t, float& ih) {
  double lo = s.seg_lo[0], inv_h = s.seg_inv_h[0];
  int nn = s.seg_n[0], base = s.seg_base[0];
  float ihf = s.seg_ih[0];
#pragma unroll
  for (int q = 1; q < MAXSEG; ++q)
    if (q < s.nseg && r >= s.seg_lo[q]) {
      lo = s.seg_lo[q];
      inv_h = s.seg_inv_h[q];
      nn = s.seg_n[q];
      base = s.seg_base[q];
      ihf = s.seg_ih[q];
    }
  const double x = (r - lo) * inv_h;
  int i = (int)x;
  if (i < 0) i = 0;
  if (i > nn - 2) i = nn - 2;
  t = (float)(x - (double)i);
  ih = ihf;
  return base + i;
}

// e^{-i m phi} by m complex multiplications from (cos phi, sin phi).
__device__ __forceinline__ float2 exp_mphi(int m, float cf, float sf) {
  float er = 1.f, ei = 0.f;
  for (int j = 1; j <= m; ++j) {
    const float nr = er * cf + ei * sf, ni = ei * cf - er * sf;
    er = nr;
    ei = ni;
  }
  return make_float2(er, ei);
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

// Sums of NV values over the workgroup (every thread gets them).
template <int NV>
__device__ __forceinline__ void block_sums_f(float (&v)[NV], float* red) {
#pragma unroll
  for (int j = 0; j < NV; ++j)
    for (int o = 32; o > 0; o >>= 1) v[j] += __shfl_xor(v[j], o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NV; ++j) red[4 * j + w] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = red[4 * j] + red[4 * j + 1] + red[4 * j + 2] + red[4 * j + 3];
}

// x of the neighbouring lane (lanes 2j and 2j + 1 exchange), by DPP quad_perm [1, 0, 3, 2].
__device__ __forceinline__ float swap_adjacent(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}

// Sum of x over the 16 lanes of a DPP row, in the row's last lane: prefix sums by
// row_shr 1, 2, 4, 8 (lanes shifted in from outside the row read 0).
__device__ __forceinline__ float row_total(float x) {
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xF, 0xF, false));
  return x;
}

constexpr int PTS = PT + 1;  // LDS stride of pair-fastest rows (odd: adjacent rows on other banks)
constexpr int PTF = PT + 4;  // ... of the forward MFMA operand rows (4 mod 64 words)

// Columns m = 0 .. nm - 1 of P_l^m (l = m .. nm - 1, so column m holds nm - m entries) dealt to the
// four wavefronts longest first, each to the least loaded one. Word u: byte w is the m of
// wavefront w's item u (0xff: none). Items: the most columns any wavefront takes.
constexpr uint32_t m_deal(int nm, int u) {
  int load[4] = {0, 0, 0, 0}, cnt[4] = {0, 0, 0, 0};
  uint32_t word = 0xffffffffu;
  for (int m = 0; m < nm; ++m) {
    int b = 0;
    for (int w = 1; w < 4; ++w)
      if (load[w] < load[b]) b = w;
    if (cnt[b] == u) word = (word & ~(0xffu << (8 * b))) | ((uint32_t)m << (8 * b));
    ++cnt[b];
    load[b] += nm - m;
  }
  return word;
}
constexpr int m_deal_items(int nm) {
  int u = 0;
  while (m_deal(nm, u) != 0xffffffffu) ++u;
  return u;
}
static_assert(m_deal_items(9) == 3 && m_deal(9, 0) == 0x03020100u && m_deal(9, 1) == 0x04050607u &&
                  m_deal(9, 2) == 0xffffff08u,
              "l_max 8: m {0, 7, 8}, {1, 6}, {2, 5}, {3, 4} (12, 11, 11, 11 entries)");

// The m of wavefront w's item u (0xff: none), from the dealt words of nm columns.
template <int NM>
__device__ __forceinline__ int dealt_m(int u, int w) {
  constexpr uint32_t W0 = m_deal(NM, 0), W1 = m_deal(NM, 1), W2 = m_deal(NM, 2), W3 = m_deal(NM, 3);
  static_assert(m_deal_items(NM) <= 4, "at most four columns per wavefront");
  const uint32_t word = u == 0 ? W0 : u == 1 ? W1 : u == 2 ? W2 : W3;
  return (int)((word >> (8 * w)) & 0xffu);
}

// For a tile of at most 32 pairs: each wavefront takes two columns at a time, one per half (lanes
// 0-31 and 32-63), and a round costs the longer one. Columns are paired in order of length, (0, 1),
// (2, 3), ..., and the pairs dealt longest first to the least loaded wavefront. Word u: byte
// 2 w + h is the m of half h of wavefront w in round u (0xff: none).
constexpr unsigned long long m_deal2(int nm, int u) {
  int load[4] = {0, 0, 0, 0}, cnt[4] = {0, 0, 0, 0};
  unsigned long long word = ~0ull;
  for (int m = 0; m < nm; m += 2) {
    int b = 0;
    for (int w = 1; w < 4; ++w)
      if (load[w] < load[b]) b = w;
    if (cnt[b] == u) {
      word = (word & ~(0xffull << (16 * b))) | ((unsigned long long)m << (16 * b));
      if (m + 1 < nm) word = (word & ~(0xffull << (16 * b + 8))) | ((unsigned long long)(m + 1) << (16 * b + 8));
    }
    ++cnt[b];
    load[b] += nm - m;
  }
  return word;
}
constexpr int m_deal2_items(int nm) {
  int u = 0;
  while (m_deal2(nm, u) != ~0ull) ++u;
  return u;
}
static_assert(m_deal2_items(9) == 2 && m_deal2(9, 0) == 0x0706050403020100ull && m_deal2(9, 1) == 0xff08ffffffffffffull,
              "l_max 8: rounds {0|1}, {2|3}, {4|5}, {6|7} then {8} (9, 7, 5, 3 + 1 entries)");
template <int NM>
__device__ __forceinline__ int dealt2_m(int u, int w, int h) {
  constexpr unsigned long long W0 = m_deal2(NM, 0), W1 = m_deal2(NM, 1), W2 = m_deal2(NM, 2);
  static_assert(m_deal2_items(NM) <= 3, "at most three rounds per wavefront");
  const unsigned long long word = u == 0 ? W0 : u == 1 ? W1 : W2;
  return (int)((word >> (16 * w + 8 * h)) & 0xffu);
}

template <int L>
struct SiteShape {
  static constexpr int KM = (L + 1) * (L + 2) / 2;    // (l, m) with m <= l <= L
  static constexpr int KM1 = (L + 2) * (L + 3) / 2;   // ... l <= L + 1
  static constexpr int L1 = L + 1;
  static constexpr int CSTR = ((2 * KM + 15) / 16) * 16;  // columns of C: 2k + {re, im}, padded
  static constexpr int CT = CSTR / 16;
  static constexpr int CS = CSTR + 4;                     // C row stride
  static constexpr int KMP = ((KM + 15) / 16) * 16;       // k padded to tiles of 16
  static constexpr int HS = 2 * KMP + 16;                 // H row stride: re block, im block, pad
  // waves per SIMD the registers are held to: 3 (168 VGPRs) up to n_max 32 and l_max 9
  static constexpr int waves(int ntt) { return ntt >= 3 || L >= 10 ? 2 : 3; }
};

// One site per workgroup of 256 threads (4 wavefronts); NTT = ceil(n_max / 16).
// Per tile of up to 64 neighbours, the pairs' geometry and table cells are
// staged in LDS once; then every thread takes part in the set-up: (n, pair)
// radial items and (l, pair) angular-factor items, whose table loads are all
// issued before any is used, then (m, pair) columns of P_l^m. The GEMM-shaped
// steps run on FP32 MFMA with operands loaded ahead: cnk = sum_pairs 4 pi R^T A,
// the per-l Hermitian products S_l = C_l W C_l^T of the spectrum, H_l = Y_l C_l
// (H stored with the real and imaginary parts of k in separate column blocks),
// and U = R H, V = R' H, so that in the last step each lane owns one k.
//
// LDS row strides: PTF (= 4 mod 64 words) for the forward R and A rows, so the
// 16 x 4 MFMA operand reads hit 64 distinct banks; PTS (odd) elsewhere; C rows
// CSTR + 4 and H rows 2 KMP + 16 for the same reason.
// MODE: 0 the full kernel; 1 the full kernel for n_max <= 16 NTT - 8, whose products over n (H,
// U and V) take two K steps of 4 fewer; 2 (FO) descriptors only (q and |p|; no dE/dq read, no
// backward pass). Separate instantiations, so that each holds only the code it runs: with both K
// forms in one kernel, the registers the longer form needs spill in the shorter one's runs too.
template <int L, int NTT, int MODE>
__global__ __launch_bounds__(NTHREADS, SiteShape<L>::waves(NTT)) void adjoint_site_k(
    AdjParams P, const int* __restrict__ order, const int* __restrict__ site_ptr, const float* __restrict__ rtab,
    const float* __restrict__ atab, const float* __restrict__ preflm_g, const float* __restrict__ cenv,
    const uint32_t* __restrict__ spmask, const double* __restrict__ rjs, const double* __restrict__ thetas,
    const double* __restrict__ phis, const float* __restrict__ G32, const int4* __restrict__ comp,
    const int* __restrict__ out_ptr, const int* __restrict__ out_c, const float* __restrict__ out_el,
    const int* __restrict__ unc_ptr, const int* __restrict__ unc_e, const int32_t* __restrict__ P_i,
    const float* __restrict__ P_el, const int2* __restrict__ odir, float* __restrict__ q_out,
    float* __restrict__ sdp_out, float* __restrict__ pair_out, unsigned long long* __restrict__ tickets) {
  constexpr bool FO = MODE == 2;
  using Sh = SiteShape<L>;
  constexpr int KM = Sh::KM, KM1 = Sh::KM1, L1 = Sh::L1, CSTR = Sh::CSTR, CT = Sh::CT;
  constexpr int CS = Sh::CS, KMP = Sh::KMP, HS = Sh::HS, KT = KMP / 16;
  constexpr int NC = 2 * KM;                        // live columns of C
  constexpr int TPW = (NTT * CT + 3) / 4;           // cnk tiles per wavefront
  constexpr int KN = 4 * NTT;                       // K steps over n (padded to 16 NTT)
  constexpr int QB = 6;                             // radial functions per batch of table loads
  constexpr int QK = 2 * QB;                        // radial functions per item (alpha_max <= 11)
  constexpr int QKEEP = 8;                          // ... kept in registers for the backward pass
  constexpr int GIT = (L1 + 3) / 4;                 // l per (pair, l-chunk) item, l <= L
  constexpr int FIT = m_deal_items(L + 1);          // forward (m, pair) items per thread, m <= L
  constexpr int PIT = m_deal_items(L + 2);          // backward (m, pair) items per thread, m <= L + 1
  constexpr int FIT2 = m_deal2_items(L + 1);        // ... for a tile of at most 32 pairs (two columns
  constexpr int PIT2 = m_deal2_items(L + 2);        //     per wavefront at a time)
  constexpr uint32_t SELF = 1u << 31;               // pair flag: the site's own pair
  extern __shared__ float smem[];
  __shared__ float red[16];
  __shared__ float preflm[KM];
  __shared__ int klm[KM];
  __shared__ float rec_a[KM1], rec_b[KM1];         // P_l^m = rec_a x P_{l-1}^m - rec_b P_{l-2}^m
  __shared__ uint32_t pr_msk[PT];                   // species bits (0: no contribution) | SELF
  __shared__ float pr_r[PT], pr_st[PT], pr_ct[PT], pr_sf[PT], pr_cf[PT];
  __shared__ int pr_acell[PT];                      // angular table cell (floats into atab)
  __shared__ float pr_ta[PT];
  __shared__ double pr_rd[PT];                      // r in FP64 (the radial table cell coordinate)
  __shared__ float cen_s[64];                       // the site's central-atom term per n (poly3gauss)
  __shared__ int site_next;
  // each species' angular table offset, read per lane from here (a lane-indexed read of the kernel
  // argument is a vector memory load, which waits behind every earlier load and store)
  __shared__ int sp_atab[ADJ_MAX_SPECIES];
  const int n_max = P.n_max, n_species = P.n_species;
  int i = 0, p0 = 0, p1 = 0;  // the current site and its pairs
  // the backward pass takes the tiles last first and reuses the forward pass's set-up of the last
  // tile (kept in registers, affordable up to n_max = 32; the tile's staged pair data is still in LDS)
  const bool KEEP = NTT <= 2 && P.keep != 0;
  int last_tb = 0;
  const float fourpi = 12.566370614359172f;
  int tx = threadIdx.x;
  int wv = tx >> 6, lane = tx & 63;
  int li = lane & 15, lk = lane >> 4;
  float* C = smem;                                 // n_max * max(CS, HS) (C, then H)
  float* region = C + n_max * (CS > HS ? CS : HS);
  float* Rt = region;                              // forward: n_max * PTF (4 pi R, pair fastest)
  float* Gt = Rt + n_max * PTF;                    //          L1 * PTS (g_l)
  float* At = Gt + L1 * PTS;                       //          NC * PTF (A columns, pair fastest)
  float* Sl = region;                              // spectrum: L1 * n_max * n_max (S_l, then Y_l)
  float* qv = Sl + L1 * n_max * n_max;             //           n_soap
  float* Gp = qv + P.n_soap;                       //           n_soap
  float* Gu = Gp + P.n_soap;                       //           n_soap_unc
  float* Rb = region;                              // backward: n_max * PTS (R)
  float* Rdb = Rb + n_max * PTS;                   //           n_max * PTS (dR/dr)
  float* Fp = Rdb + n_max * PTS;                   //           KM1 * PTS (P_l^m, l <= L + 1)
  float* Gl = Fp + ((KM1 * PTS + 1) & ~1);         //           L1 * PTS (g_l)
  float* Gd = Gl + L1 * PTS;                       //           L1 * PTS (dg_l/dr)
  float2* Em = reinterpret_cast<float2*>(Gd + L1 * PTS);  //    L1 * PTS (e^{-i m phi})
  float* fout = reinterpret_cast<float*>(Em + L1 * PTS);  //    3 * PT

  // The next tile's pair data, loaded a phase or more before it is staged (the loads' latency passes
  // while the current tile is processed): wavefront 0 r and the species bits, wavefront 1 theta (and
  // the site's central-atom term with a site's first tile), wavefront 2 phi.
  double pf_d = 0.0;
  uint32_t pf_m = 0u;
  float pf_c = 0.f;
  auto fetch = [&](int site, int tb, int pend, bool first) {
    const int p = tb + (tx & (PT - 1));
    if (tx < PT) {
      pf_d = p < pend ? rjs[p] : 1.0;
      pf_m = p < pend ? spmask[p] : 0u;
    } else if (tx < 2 * PT) {
      pf_d = p < pend ? thetas[p] : 0.0;
      if (cenv && first && tx - PT < n_max) pf_c = cenv[(size_t)site * n_max + tx - PT];
    } else if (tx < 3 * PT) {
      pf_d = p < pend ? phis[p] : 0.0;
    }
  };
  // The tile's pairs (fetched before): r (FP32, and FP64 for the radial table cell coordinate),
  // species bits of contributing pairs, the angular table cell, sin/cos of theta and phi.
  // Wavefront 0 takes r, the species and the angular cell, wavefronts 1 and 2 the sines and cosines
  // of theta and phi; with the site's own pair (forward pass), wavefront 1 also stores the
  // central-atom term.
  auto stage = [&](int tb, bool with_self) {
    const int pl = tx & (PT - 1), p = tb + pl;
    if (tx < PT) {
      uint32_t msk = 0;
      const double rd = pf_d;
      if (p < p1) {
        msk = pf_m;
        if (!(rd < P.rcut_max) || (p == p0 && !with_self)) msk = 0;
      }
      const bool self = p == p0;
      int acell = 0;
      float ta = 0.f;
      if (msk != 0) {
        // FP32 cell coordinate: g_l(r) is smooth, so a pair that rounds into the next cell takes the
        // same value there (unlike the radial table's segments, whose jumps need the FP64 distance)
        const float xa = (float)rd * P.ang_ih;
        int ia = (int)xa;
        if (ia > P.ang_n - 2) ia = P.ang_n - 2;
        ta = xa - (float)ia;
        const int s0 = __ffs(msk) - 1;
        acell = sp_atab[s0] + ia * 4 * L1;
      }
      pr_msk[pl] = msk != 0 && self ? (msk | SELF) : msk;
      pr_r[pl] = (float)rd;
      pr_rd[pl] = rd;
      pr_acell[pl] = acell;
      pr_ta[pl] = ta;
    } else if (tx < 2 * PT) {
      const double th = pf_d;
      float st_, ct_;
      sincosf((float)th, &st_, &ct_);
      pr_st[pl] = st_;
      pr_ct[pl] = ct_;
      if (cenv && with_self && tb == p0 && tx - PT < n_max) cen_s[tx - PT] = pf_c;
    } else if (tx < 3 * PT) {
      const double ph = pf_d;
      float sf_, cf_;
      sincosf((float)ph, &sf_, &cf_);
      pr_sf[pl] = sf_;
      pr_cf[pl] = cf_;
    }
  };
  // Radial items (species block s, chunk ch, pair): the pair's table cell for species s (FP64
  // coordinate), then R_n times scale (and the central weight on the site's own pair), and dR_n/dr if
  // Rd is given, for the chunk's functions n, whose cubics lie side by side in the cell's table row.
  // (s, ch) is uniform over a wavefront. With keep, R_n and dR_n/dr of every pair but the site's
  // own are also kept in registers (kv, kd: the thread's one item), for the backward pass.
  float kv[QKEEP], kd[QKEEP], kg[GIT], kgd[GIT];
  auto radial_items = [&](int tid, float scale, float* R, float* Rd, int stride, bool keep) {
    const int n_items = n_species * P.nch * PT;
    for (int j = tid; j < n_items; j += NTHREADS) {
      const int pl = j & (PT - 1);
      const int sc_ = __builtin_amdgcn_readfirstlane(j >> 6);
      const int s = sc_ / P.nch, ch = sc_ - s * P.nch;
      const SpeciesTab& sp = P.sp[s];
      const int q0 = ch * P.qc, qn = min(sp.a - q0, P.qc);
      const uint32_t msk = pr_msk[pl];
      const double rd = pr_rd[pl];
      const bool self = (msk & SELF) != 0;
      int c = -1;
      float t = 0.f, ih = 0.f;
      if ((msk & (1u << s)) && rd < sp.rcut && !(self && !sp.do_central)) c = radial_cell(sp, rd, t, ih);
      const float sc = (self ? sp.cw : 1.f) * scale;
      const float4* row = reinterpret_cast<const float4*>(rtab + sp.rtab_off) + (size_t)(c < 0 ? 0 : c) * sp.a + q0;
      float* Rrow = R + (sp.off + q0) * stride + pl;
      float* Rdrow = Rd ? Rd + (sp.off + q0) * stride + pl : nullptr;
#pragma unroll
      for (int b = 0; b < QK; b += QB) {
        if (b < qn) {
          // the batch's table loads, all issued before any is used
          float4 cf[QB];
#pragma unroll
          for (int q = 0; q < QB; ++q)
            cf[q] = (c >= 0 && b + q < qn) ? row[b + q] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
          for (int q = 0; q < QB; ++q) {
            if (b + q < qn) {
              const float v = cf[q].x + t * (cf[q].y + t * (cf[q].z + t * cf[q].w));
              const float d = (cf[q].y + t * (2.f * cf[q].z + 3.f * t * cf[q].w)) * ih;
              Rrow[(b + q) * stride] = sc * v;
              if (Rd) Rdrow[(b + q) * stride] = d;
              if (keep && b + q < QKEEP) {
                kv[b + q] = self ? 0.f : v;
                kd[b + q] = self ? 0.f : d;
              }
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);  // one batch's loads live at a time
      }
    }
  };
  // The kept radial values back into LDS (the same items as radial_items).
  auto radial_kept = [&](int tid, float* R, float* Rd, int stride) {
    const int pl = tid & (PT - 1);
    const int sc_ = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (sc_ < n_species * P.nch) {
      const int s = sc_ / P.nch, ch = sc_ - s * P.nch;
      const int q0 = ch * P.qc, qn = min(P.sp[s].a - q0, P.qc);
      const int n0 = P.sp[s].off + q0;
#pragma unroll
      for (int q = 0; q < QKEEP; ++q)
        if (q < qn) {
          R[(n0 + q) * stride + pl] = kv[q];
          Rd[(n0 + q) * stride + pl] = kd[q];
        }
    }
  };
  // (pair, l-chunk) items: g_l(r), and dg_l/dr if Gdr is given (both kept in kg/kgd with keep), for
  // l = GIT wv .. GIT wv + GIT - 1, whose cubics lie side by side in the pair's table row
  // (gl_load issues the item's table loads; they are issued before the radial items', so that both
  // sets of loads are in flight together)
  auto gl_load = [&](int tid, float4 (&gc)[GIT]) {
    const int pl = tid & (PT - 1), l0 = GIT * __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool live = pr_msk[pl] != 0;
    const float4* row = reinterpret_cast<const float4*>(atab + pr_acell[pl]);
#pragma unroll
    for (int u = 0; u < GIT; ++u)
      gc[u] = (live && l0 + u < L1) ? row[l0 + u] : make_float4(0.f, 0.f, 0.f, 0.f);
  };
  auto gl_items = [&](int tid, const float4 (&gc)[GIT], float* Gr, float* Gdr, bool keep) {
    const int pl = tid & (PT - 1), l0 = GIT * __builtin_amdgcn_readfirstlane(tid >> 6);
    const float ta = pr_ta[pl];
#pragma unroll
    for (int u = 0; u < GIT; ++u) {
      const int l = l0 + u;
      if (l < L1) {
        const float g = gc[u].x + ta * (gc[u].y + ta * (gc[u].z + ta * gc[u].w));
        const float gd = (gc[u].y + ta * (2.f * gc[u].z + 3.f * ta * gc[u].w)) * P.ang_ih;
        Gr[l * PTS + pl] = g;
        if (Gdr) Gdr[l * PTS + pl] = gd;
        if (keep) {
          kg[u] = g;
          kgd[u] = gd;
        }
      }
    }
  };

  if (threadIdx.x < ADJ_MAX_SPECIES) sp_atab[threadIdx.x] = threadIdx.x < n_species ? P.sp[threadIdx.x].atab_off : 0;
  for (int t = threadIdx.x; t < KM; t += NTHREADS) {
    preflm[t] = preflm_g[t];
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= t) ++l;
    klm[t] = l | ((t - l * (l + 1) / 2) << 8);
  }
  // the three-term recursion's coefficients (2l - 1)/(l - m) and (l - 1 + m)/(l - m), l >= m + 2
  for (int t = threadIdx.x; t < KM1; t += NTHREADS) {
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= t) ++l;
    const int m = t - l * (l + 1) / 2;
    rec_a[t] = l >= m + 2 ? (float)(2 * l - 1) / (float)(l - m) : 0.f;
    rec_b[t] = l >= m + 2 ? (float)(l - 1 + m) / (float)(l - m) : 0.f;
  }
  // Sites in the order given (heaviest first): workgroup b starts on site b, and each workgroup
  // takes its next site from a queue (a ticket counter that is zero at launch; ticket t is site
  // gridDim.x + t) while it works on the current one, until the queue runs dry. The grid is one
  // resident wave of workgroups. The step's last kernel sets the counter back to zero.
  if (threadIdx.x == 0)
    site_next = (int)blockIdx.x < P.n_sites ? (order ? order[blockIdx.x] : (int)blockIdx.x) : -1;
  unsigned long long tk = 0;  // thread 0: the ticket for the next site
  int nxt = -1;               // thread 0: the next site (-1: none)
  bool pf_site = false;       // the fetch registers hold the first tile of the site taken next
  for (;;) {
  __syncthreads();
  i = site_next;
  if (i < 0) break;
  // lane coordinates made opaque per site: what derives from them is formed per site, not
  // hoisted out of the queue loop into long-lived registers
  asm volatile("" : "+v"(tx), "+v"(wv), "+v"(lane), "+v"(li), "+v"(lk));
  p0 = site_ptr[i];
  p1 = site_ptr[i + 1];
  last_tb = p0 + ((p1 - p0 - 1) / PT) * PT;
  if (!pf_site) fetch(i, p0, p1, true);
  pf_site = false;
  // --------------------------------------------------------------- forward
  // With a direct output map (every output one uncompressed component, every component at
  // most one output) and at most DOUT outputs per thread, each thread's outputs are computed from
  // C where they are used, without S_l.
  constexpr int DOUT = 4;
  const bool spec_direct = odir != nullptr && P.n_soap <= DOUT * NTHREADS;
  floatx4 hi[TPW], lo[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    hi[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    lo[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  for (int tb = p0; tb < p1; tb += PT) {
    const int np_ = min(PT, p1 - tb);
    __syncthreads();
    stage(tb, true);
    {
      // the next tile to stage: this pass's next, else the backward pass's first restaged one
      const int nt = tb + PT < p1 ? tb + PT : (KEEP ? last_tb - PT : last_tb);
      if (nt >= p0) fetch(i, nt, p1, false);
    }
    __syncthreads();
    // the thread index made opaque per tile: item addresses are formed here, not hoisted
    int tid = tx;
    asm volatile("" : "+v"(tid));
    float4 gc[GIT];
    gl_load(tid, gc);
    radial_items(tid, fourpi, Rt, nullptr, PTF, KEEP && tb == last_tb);
    gl_items(tid, gc, Gt, nullptr, KEEP && tb == last_tb);
    __syncthreads();
    // (m, pair) items: A_k = preflm P_l^m g_l e^{-i m phi} for l = m .. L, as (re, im) rows.
    // m is uniform over a wavefront (columns dealt by length, m_deal); l is unrolled, P_l^m by the
    // recursions of get_plm_array (src/soap_turbo_angular.f90:108-160): P_m^m, P_{m+1}^m =
    // (2m+1) x P_m^m, three-term in l.
    // A tile of at most 32 pairs takes two columns per wavefront at a time (m per half-wave).
    auto fwd_item = [&](int pl, int m) {
      {
        const bool live = pr_msk[pl] != 0;
        // g_l, the recursion's coefficients and the prefactors of the column, all read before the
        // recursion (k clamped into the column for l < m, whose values are not used), so that the
        // chain does not wait for LDS at every step
        float gv[L1], ca[L1], cb[L1], pf[L1];
#pragma unroll
        for (int l = 0; l < L1; ++l) {
          const int kl = l * (l + 1) / 2 + (m < l ? m : l);
          gv[l] = Gt[l * PTS + pl];
          ca[l] = rec_a[kl];
          cb[l] = rec_b[kl];
          pf[l] = preflm[kl];
        }
        const float x = pr_ct[pl], sq = pr_st[pl];
        const float2 e = exp_mphi(m, pr_cf[pl], pr_sf[pl]);
        float pmm = 1.f;
#pragma unroll
        for (int j = 1; j <= L; ++j)
          if (j <= m) pmm = -(float)(2 * j - 1) * sq * pmm;
        float p0 = 0.f, p1 = 0.f;
#pragma unroll
        for (int l = 0; l <= L; ++l) {
          if (l >= m) {
            float pc;
            if (l == m) pc = pmm;
            else if (l == m + 1) pc = x * (float)(2 * m + 1) * pmm;
            else pc = ca[l] * x * p1 - cb[l] * p0;
            p0 = p1;
            p1 = pc;
            const int k = l * (l + 1) / 2 + m;
            const float v = live ? pf[l] * pc * gv[l] : 0.f;
            At[(2 * k) * PTF + pl] = v * e.x;
            At[(2 * k + 1) * PTF + pl] = v * e.y;
          }
        }
      }
    };
    if (np_ <= 32) {
#pragma unroll
      for (int u = 0; u < FIT2; ++u) {
        const int m = dealt2_m<L1>(u, __builtin_amdgcn_readfirstlane(tid >> 6), (tid >> 5) & 1);
        if (m < L1) fwd_item(tid & 31, m);
      }
    } else {
#pragma unroll
      for (int u = 0; u < FIT; ++u) {
        const int m = dealt_m<L1>(u, __builtin_amdgcn_readfirstlane(tid >> 6));  // m uniform
        if (m < L1) fwd_item(tid & (PT - 1), m);
      }
    }
    __syncthreads();
    // cnk: the tile's pairs on MFMA, 16 neighbours per load batch, in one accumulation chain
    // per tile; tiles are combined with TwoSum compensation
    const int npad = (np_ + 3) & ~3;
    floatx4 acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < npad; c0 += 16) {
      const int ns = (min(npad, c0 + 16) - c0) >> 2;
      // every tile's operands of the batch loaded before any is used (rows and columns clamped into
      // range, the terms past them zeroed), so that the batch waits for LDS once
      float av[TPW][4], bv[TPW][4];
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        const int t = wv + 4 * j;
        const int a = t < NTT * CT ? t / CT : 0, b = t < NTT * CT ? t - a * CT : 0;
        const int n = 16 * a + li, col = 16 * b + li;
        const int nr = n < n_max ? n : 0, cr = col < NC ? col : 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int pp = c0 + 4 * s + lk;
          av[j][s] = Rt[nr * PTF + pp];
          bv[j][s] = At[cr * PTF + pp];
        }
      }
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        const int t = wv + 4 * j;
        const int a = t / CT, b = t - a * CT;
        const bool ok = t < NTT * CT;
        const int n = 16 * a + li, col = 16 * b + li;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          av[j][s] = ok && s < ns && n < n_max ? av[j][s] : 0.f;
          bv[j][s] = ok && s < ns && col < NC ? bv[j][s] : 0.f;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[j] = mfma4(av[j][s], bv[j][s], acc[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < TPW; ++j) two_sum_acc(hi[j], lo[j], acc[j]);
  }
  // C = hi + lo, plus the poly3gauss central-atom term on k = 0 (real part)
  // (the central terms read for every row first, index clamped: a read under the column's
  // condition was waited for before the next)
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = wv + 4 * j;
    if (t < NTT * CT) {
      const int a = t / CT, b = t - a * CT;
      const int col = 16 * b + li;
      float cen[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = 16 * a + 4 * lk + v;
        cen[v] = cen_s[n < 64 ? n : 63];
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = 16 * a + 4 * lk + v;
        if (n < n_max) {
          float val = hi[j][v] + lo[j][v];
          if (cenv && col == 0) val += cen[v];
          C[n * CS + col] = val;
        }
      }
    }
  }
  __syncthreads();
  // --------------------------------------------------------------- spectrum
  // S_l(n, n') = sum over the l block of w(m) C(n, c) C(n', c), w = 1 (m = 0), 2 (m > 0).
  // With a direct output map and at most DOUT outputs per thread, each output's entry is summed
  // where it is used, below. Otherwise S_l is formed on MFMA: tiles with a <= b (the outputs
  // read n <= n'), a wavefront's jobs unrolled so that their loads are in flight together.
  if (!spec_direct) {
    constexpr int NAB = NTT * (NTT + 1) / 2, NJ = L1 * NAB, JPW = (NJ + 3) / 4;
    constexpr int MS = (2 * L1 + 3) / 4;
    float av[JPW][MS], bv[JPW][MS];
#pragma unroll
    for (int jj = 0; jj < JPW; ++jj) {
      const int job = wv + 4 * jj;
      const int l = job / NAB;
      int ab = job - l * NAB, a = 0;
      while (ab >= NTT - a) { ab -= NTT - a; ++a; }
      const int b = a + ab;
      const int cb = l * (l + 1), cw = 2 * (l + 1);
      const int n = 16 * a + li, np = 16 * b + li;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        const int c = 4 * s + lk;
        const bool inb = job < NJ && c < cw;
        av[jj][s] = (inb && n < n_max) ? C[n * CS + cb + c] : 0.f;
        bv[jj][s] = (inb && np < n_max) ? (c >= 2 ? 2.f : 1.f) * C[np * CS + cb + c] : 0.f;
      }
    }
#pragma unroll
    for (int jj = 0; jj < JPW; ++jj) {
      const int job = wv + 4 * jj;
      if (job < NJ) {
        const int l = job / NAB;
        int ab = job - l * NAB, a = 0;
        while (ab >= NTT - a) { ab -= NTT - a; ++a; }
        const int b = a + ab;
        const int cw = 2 * (l + 1);
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (4 * s < cw) acc = mfma4(av[jj][s], bv[jj][s], acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = 16 * a + 4 * lk + v, cc = 16 * b + li;
          if (r < n_max && cc < n_max) Sl[(l * n_max + r) * n_max + cc] = acc[v];
        }
      }
    }
  }
  __syncthreads();
  // q, |p|, and Gp = (G - q (q.G)) / |p| from one pass: q.G = (p.G) / |p|. With a direct
  // output map, Y_l is then filled from the outputs: zero, and for output s of component
  // (n, n', l) with weight w, Y_l(n, n') = Y_l(n', n) = w Gp(s) (2 w on the diagonal).
  const int n_soap = P.n_soap;
  const int ny = L1 * n_max * n_max;
  float* Y = Sl;
  float sums[2] = {0.f, 0.f};
  if (spec_direct) {
    int2 od[DOUT];
    float gd[DOUT], vq[DOUT];
#pragma unroll
    for (int u = 0; u < DOUT; ++u) {
      const int s = tx + NTHREADS * u;
      od[u] = s < n_soap ? odir[s] : make_int2(0, 0);
      gd[u] = !FO && s < n_soap ? G32[(size_t)i * n_soap + s] : 0.f;
    }
    for (int t = tx; t < ny; t += NTHREADS) Y[t] = 0.f;
#pragma unroll
    for (int u = 0; u < DOUT; ++u) {
      const int s = tx + NTHREADS * u;
      vq[u] = 0.f;
      if (s < n_soap) {
        const int x = od[u].x & 255, y = (od[u].x >> 8) & 255, l = od[u].x >> 16;
        const int cb = l * (l + 1), cw = 2 * (l + 1);
        const float* ca = C + x * CS + cb;
        const float* cc = C + y * CS + cb;
        // all 2 (L + 1) columns loaded at once (in the row for any l: cb + 2 L + 1 < 2 KM), the ones
        // past block l masked out of the sum
        // (as float2: rows and cb are even, so 8-byte aligned; half the reads, issued together)
        static_assert(CS % 2 == 0, "C rows of whole float2");
        float va[2 * L1], vb[2 * L1];
        const float2* ca2 = reinterpret_cast<const float2*>(ca);
        const float2* cc2 = reinterpret_cast<const float2*>(cc);
#pragma unroll
        for (int c = 0; c < L1; ++c) {
          const float2 x2 = ca2[c], y2 = cc2[c];
          va[2 * c] = x2.x;
          va[2 * c + 1] = x2.y;
          vb[2 * c] = y2.x;
          vb[2 * c + 1] = y2.y;
        }
        float acc = va[0] * vb[0] + va[1] * vb[1];
#pragma unroll
        for (int c = 2; c < 2 * L1; ++c) acc += c < cw ? va[c] * (2.f * vb[c]) : 0.f;
        const float v = __int_as_float(od[u].y) * acc;
        vq[u] = v;
        sums[0] += v * v;
        sums[1] += v * gd[u];
      }
    }
    block_sums_f<2>(sums, red);
    float sdp = sqrtf(sums[0]);
    if (sdp < 1e-5f) sdp = 1.f;
    const float inv = 1.f / sdp;
    const float qg = sums[1] * inv;
#pragma unroll
    for (int u = 0; u < DOUT; ++u) {
      const int s = tx + NTHREADS * u;
      if (s < n_soap) {
        const float q = vq[u] * inv;
        q_out[(size_t)i * n_soap + s] = q;
        const int x = od[u].x & 255, y = (od[u].x >> 8) & 255, l = od[u].x >> 16;
        const float yv = (x == y ? 2.f : 1.f) * __int_as_float(od[u].y) * ((gd[u] - q * qg) * inv);
        Y[(l * n_max + x) * n_max + y] = yv;
        Y[(l * n_max + y) * n_max + x] = yv;
      }
    }
    if (tx == 0) sdp_out[i] = sdp;
    __syncthreads();
  } else {
    for (int s0 = tx; s0 < n_soap; s0 += 4 * NTHREADS) {
      float vq[4], gq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int s = s0 + NTHREADS * u;
        vq[u] = gq[u] = 0.f;
        if (s < n_soap) {
          float v = 0.f;
          if (odir) {
            const int2 o = odir[s];
            const int x = o.x & 255, y = (o.x >> 8) & 255, l = o.x >> 16;
            v = __int_as_float(o.y) * Sl[(l * n_max + x) * n_max + y];
          } else {
            const int e0 = P.compressed ? out_ptr[s] : s, e1 = P.compressed ? out_ptr[s + 1] : s + 1;
            for (int e = e0; e < e1; ++e) {
              const int4 cm = comp[P.compressed ? out_c[e] : e];
              const float base = cm.x != cm.y ? 1.41421356f : 1.f;
              v += (P.compressed ? out_el[e] : 1.f) * base * Sl[(cm.z * n_max + cm.x) * n_max + cm.y];
            }
          }
          const float g = FO ? 0.f : G32[(size_t)i * n_soap + s];
          vq[u] = v;
          gq[u] = g;
          sums[0] += v * v;
          sums[1] += v * g;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int s = s0 + NTHREADS * u;
        if (s < n_soap) {
          qv[s] = vq[u];
          Gp[s] = gq[u];
        }
      }
    }
    block_sums_f<2>(sums, red);  // its barriers also end the reads of S_l
    float sdp = sqrtf(sums[0]);
    if (sdp < 1e-5f) sdp = 1.f;
    const float inv = 1.f / sdp;
    const float qg = sums[1] * inv;
    if (odir) {
      for (int t = tx; t < ny; t += NTHREADS) Y[t] = 0.f;
      __syncthreads();
    }
    for (int s = tx; s < n_soap; s += NTHREADS) {
      const float q = qv[s] * inv;
      q_out[(size_t)i * n_soap + s] = q;
      const float gp = (Gp[s] - q * qg) * inv;
      if (odir) {
        const int2 o = odir[s];
        const int x = o.x & 255, y = (o.x >> 8) & 255, l = o.x >> 16;
        const float yv = (x == y ? 2.f : 1.f) * __int_as_float(o.y) * gp;
        Y[(l * n_max + x) * n_max + y] = yv;
        Y[(l * n_max + y) * n_max + x] = yv;
      } else {
        Gp[s] = gp;
      }
    }
    if (tx == 0) sdp_out[i] = sdp;
    __syncthreads();
    if (!odir) {
      // general compression: dE/dp per uncompressed component, then Y_l from it
      for (int c0 = tx; c0 < P.n_soap_unc; c0 += 4 * NTHREADS) {
        float gu[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int c = c0 + NTHREADS * u;
          gu[u] = 0.f;
          if (c < P.n_soap_unc) {
            float v;
            if (P.compressed) {
              v = 0.f;
              for (int e = unc_ptr[c]; e < unc_ptr[c + 1]; ++e) v += P_el[unc_e[e]] * Gp[P_i[unc_e[e]] - 1];
            } else {
              v = Gp[c];
            }
            gu[u] = v;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int c = c0 + NTHREADS * u;
          if (c < P.n_soap_unc) Gu[c] = gu[u];
        }
      }
      __syncthreads();
      // Y_l(n, n') = Gu(c) x 2 on the diagonal, sqrt 2 off it, over S_l; one (l, n) row per thread,
      // eight columns at a time (their loads ahead of their stores); c(a, b, l) for a <= b is the
      // uncompressed component (a n_max - a(a-1)/2 + b - a) L1 + l
      for (int row = tx; row < L1 * n_max; row += NTHREADS) {
        const int l = row / n_max, n = row - l * n_max;
        float* yr = Y + (size_t)row * n_max;
        for (int np0 = 0; np0 < n_max; np0 += 8) {
          float y[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int np = np0 + j;
            const int lo_ = n < np ? n : np, hi_ = n < np ? np : n;
            const int c = (((lo_ * (2 * n_max - lo_ - 1)) >> 1) + hi_) * L1 + l;  // (a n_max - a(a+1)/2 + b) L1 + l
            y[j] = np < n_max ? (lo_ == hi_ ? 2.f : 1.41421356f) * Gu[c] : 0.f;
          }
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (np0 + j < n_max) yr[np0 + j] = y[j];
        }
      }
      __syncthreads();
    }
  }
  if constexpr (FO) {
    // descriptors only: the next site now (the wait publishes it to the loop-top barrier)
    if (tx == 0) {
      const unsigned long long t = atomicAdd(tickets, 1ull) + gridDim.x;
      site_next = t < (unsigned long long)P.n_sites ? (order ? order[t] : (int)t) : -1;
      __builtin_amdgcn_s_waitcnt(0xc07f);
    }
    continue;
  }
  // the ticket for the next site, taken here so that its latency and the site's lookup below
  // pass while this site finishes
  if (tx == 0) tk = atomicAdd(tickets, 1ull);
  // H_l = 4 pi w(m) Y_l C_l over the columns of block l. A job is one column tile of one
  // block and every row tile; each wavefront holds its jobs' C operands in registers across a
  // barrier, after which H is written over C: real parts of k in columns k, imaginary
  // parts in columns KMP + k.
  {
    constexpr int NJH = [] {
      int n = 0;
      for (int l = 0; l <= L; ++l) n += (2 * (l + 1) + 15) / 16;
      return n;
    }();
    constexpr int JHW = (NJH + 3) / 4;
    // When the last round has fewer jobs than wavefronts and their row tiles number at most four, the
    // row tiles are dealt one per wavefront (wavefront w: job w / NTT of the round, row tile w % NTT)
    constexpr int RL = NJH - 4 * (JHW - 1);
    constexpr bool RSPLIT = RL < 4 && RL * NTT <= 4;
    const int wq = __builtin_amdgcn_readfirstlane(wv);
    auto job_of = [&](int jj) { return RSPLIT && jj == JHW - 1 ? 4 * jj + wq / NTT : wq + 4 * jj; };
    float bv[JHW][KN];
#pragma unroll
    for (int jj = 0; jj < JHW; ++jj) {
      const int job = job_of(jj);
      int l = 0, rem = job < NJH ? job : 0;
      while (rem >= (2 * (l + 1) + 15) / 16) { rem -= (2 * (l + 1) + 15) / 16; ++l; }
      const int cb = l * (l + 1), cw = 2 * (l + 1), c = 16 * rem + li;
#pragma unroll
      for (int s = 0; s < KN; ++s) {
        const int np = 4 * s + lk;
        bv[jj][s] = (job < NJH && np < n_max && c < cw) ? C[np * CS + cb + c] : 0.f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < JHW; ++jj) {
      const int job = job_of(jj);
      // row tiles of this job: all, or the one dealt to this wavefront
      const bool one = RSPLIT && jj == JHW - 1;
      const int arow = one ? wq % NTT : 0;
      if (job < NJH) {
        int l = 0, rem = job;
        while (rem >= (2 * (l + 1) + 15) / 16) { rem -= (2 * (l + 1) + 15) / 16; ++l; }
        const int cw = 2 * (l + 1), c = 16 * rem + li;
        floatx4 acc[NTT];
        // the row tiles' chains over n' < 4 K (K steps of 4; the steps past n_max add nothing)
        // the column's weight, with the k's prefactor of A folded in (read with the operands)
        const int kc = l * (l + 1) / 2 + (min(c, cw - 1) >> 1);
        const float pf = preflm[kc];
        auto chains = [&](auto kn, auto na_) {
          constexpr int K = decltype(kn)::value, NA = decltype(na_)::value;
          // every row tile's Y operands loaded before the first product (rows clamped into range,
          // the terms past n_max zeroed), so that the job waits for LDS once
          float av[NA][K];
#pragma unroll
          for (int a = 0; a < NA; ++a) {
            const int n = 16 * (a + arow) + li, nr = n < n_max ? n : 0;
#pragma unroll
            for (int s = 0; s < K; ++s) {
              const int np = 4 * s + lk, npr = np < n_max ? np : 0;
              av[a][s] = Y[(l * n_max + nr) * n_max + npr];
            }
          }
#pragma unroll
          for (int a = 0; a < NA; ++a) {
            const int n = 16 * (a + arow) + li;
#pragma unroll
            for (int s = 0; s < K; ++s) av[a][s] = n < n_max && 4 * s + lk < n_max ? av[a][s] : 0.f;
            acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < K; ++s) acc[a] = mfma4(av[a][s], bv[jj][s], acc[a]);
          }
        };
        auto rows = [&](auto na_) {
          constexpr int NA = decltype(na_)::value;
          if constexpr (MODE == 1) chains(std::integral_constant<int, KN - 2>{}, na_);
          else chains(std::integral_constant<int, KN>{}, na_);
          if (c < cw) {
            const int k = l * (l + 1) / 2 + (c >> 1), hc = (c & 1) * KMP + k;
            const float w = (c >= 2 ? 2.f : 1.f) * fourpi * pf;
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int r = 16 * (a + arow) + 4 * lk + v;
                if (r < n_max) C[r * HS + hc] = w * acc[a][v];
              }
          }
        };
        if (RSPLIT && one) rows(std::integral_constant<int, 1>{});
        else rows(std::integral_constant<int, NTT>{});
      }
      __builtin_amdgcn_sched_barrier(0);  // one job's operands live at a time
    }
    // H columns k in [KM, KMP) and [KMP + KM, 2 KMP): zero
    for (int t = tx; t < n_max * 2 * (KMP - KM); t += NTHREADS) {
      const int r = t / (2 * (KMP - KM)), j = t - r * 2 * (KMP - KM);
      const int hc = j < KMP - KM ? KM + j : KMP + KM + (j - (KMP - KM));
      C[r * HS + hc] = 0.f;
    }
  }
  const float* H = C;
  if (tx == 0) {
    // the next site, published now (no thread reads site_next before the loop-top barrier) so
    // that the backward pass can fetch its first tile
    const unsigned long long t = tk + gridDim.x;
    nxt = t < (unsigned long long)P.n_sites ? (order ? order[t] : (int)t) : -1;
    site_next = nxt;
    __builtin_amdgcn_s_waitcnt(0xc07f);
  }
  // the next site's first tile into the fetch registers (after a barrier that publishes site_next)
  auto fetch_next_site = [&]() {
    const int ns = site_next;
    if (ns >= 0) {
      fetch(ns, site_ptr[ns], site_ptr[ns + 1], true);
      pf_site = true;
    }
  };
  // --------------------------------------------------------------- backward
  float self[3] = {0.f, 0.f, 0.f};
  for (int tb = last_tb; tb >= p0; tb -= PT) {
    __syncthreads();
    int tid = tx;
    asm volatile("" : "+v"(tid));
    if (KEEP && tb == last_tb) {
      // the forward pass staged this tile last and kept R, R', g, g' of its items: no table loads
      // (the fetch registers hold the tile restaged next, if any)
      if (last_tb == p0) fetch_next_site();
      radial_kept(tid, Rb, Rdb, PTS);
      {
        const int pl = tid & (PT - 1), l0 = GIT * (tid >> 6);
#pragma unroll
        for (int u = 0; u < GIT; ++u)
          if (l0 + u < L1) {
            Gl[(l0 + u) * PTS + pl] = kg[u];
            Gd[(l0 + u) * PTS + pl] = kgd[u];
          }
      }
    } else {
      stage(tb, false);
      if (tb - PT >= p0) fetch(i, tb - PT, p1, false);
      else fetch_next_site();
      __syncthreads();
      float4 gc[GIT];
      gl_load(tid, gc);
      radial_items(tid, 1.f, Rb, Rdb, PTS, false);
      gl_items(tid, gc, Gl, Gd, false);
    }
    // (m, pair) items, m <= L + 1: column m of P_l^m (l <= L + 1); e^{-i m phi} for m <= L
    // (a tile of at most 32 pairs: two columns per wavefront at a time, as in the forward pass)
    auto bwd_item = [&](int pl, int m) {
      {
        // l unrolled: the recursion's coefficients are read ahead of the chain that uses them
        const bool live = pr_msk[pl] != 0 && tb + pl != p0;
        const float x = pr_ct[pl], sq = pr_st[pl];
        // the recursion's coefficients of the column, read before it (k clamped for l < m)
        float ca[L + 2], cb[L + 2];
#pragma unroll
        for (int l = 0; l <= L + 1; ++l) {
          const int kl = l * (l + 1) / 2 + (m < l ? m : l);
          ca[l] = rec_a[kl];
          cb[l] = rec_b[kl];
        }
        float pmm = 1.f;
#pragma unroll
        for (int j = 1; j <= L + 1; ++j)
          if (j <= m) pmm = -(float)(2 * j - 1) * sq * pmm;
        float q0 = 0.f, q1 = 0.f;
#pragma unroll
        for (int l = 0; l <= L + 1; ++l) {
          if (l >= m) {
            const int k = l * (l + 1) / 2 + m;
            float pc;
            if (l == m) pc = pmm;
            else if (l == m + 1) pc = x * (float)(2 * m + 1) * pmm;
            else pc = ca[l] * x * q1 - cb[l] * q0;
            q0 = q1;
            q1 = pc;
            Fp[k * PTS + pl] = live ? pc : 0.f;
          }
        }
        if (m <= L) Em[m * PTS + pl] = live ? exp_mphi(m, pr_cf[pl], pr_sf[pl]) : make_float2(0.f, 0.f);
      }
    };
    if (p1 - tb <= 32) {
#pragma unroll
      for (int u = 0; u < PIT2; ++u) {
        const int m = dealt2_m<L + 2>(u, __builtin_amdgcn_readfirstlane(tid >> 6), (tid >> 5) & 1);
        if (m <= L + 1) bwd_item(tid & 31, m);
      }
    } else {
#pragma unroll
      for (int u = 0; u < PIT; ++u) {
        const int m = dealt_m<L + 2>(u, __builtin_amdgcn_readfirstlane(tid >> 6));  // m uniform
        if (m <= L + 1) bwd_item(tid & (PT - 1), m);
      }
    }
    __syncthreads();
    // wavefront wv: pairs 16 wv .. 16 wv + 15 of the tile. Per tile of 16 k: U = R H and
    // V = R' H for the real and the imaginary columns; lane li then owns k = 16 kt + li. A
    // wavefront whose pairs all lie past the tile's last one has nothing to do.
    // A tile of at most 32 pairs (KT > 1): wavefronts w and w + 2 share the rows 16 w .. 16 w + 15,
    // w taking the k tiles below KS and w + 2 the rest; w + 2's sums go to the upper half of fout.
    constexpr int KS = KT / 2;
    const int wq = __builtin_amdgcn_readfirstlane(wv);
    const bool ksplit = KT > 1 && p1 - tb <= 32;
    const int rw = ksplit ? (wq & 1) : wq;  // this wavefront's row group
    if (16 * rw < p1 - tb) {
      float fr[4] = {0.f, 0.f, 0.f, 0.f}, fa[4] = {0.f, 0.f, 0.f, 0.f}, fo[4] = {0.f, 0.f, 0.f, 0.f};
      // lane coordinates made opaque here, so that the LDS addresses derived from them are
      // formed in the loop rather than hoisted out of it into long-lived registers
      int lrow = 16 * rw + 4 * lk, lcol = li;
      asm volatile("" : "+v"(lrow), "+v"(lcol));
      const int pa = 16 * rw + lcol;  // A-operand row (pair) of this lane
      // U and V over n < 4 K (K steps of 4; the steps past n_max add nothing) for the k tiles
      // KA .. KB - 1, then the contraction
      auto uv = [&](auto kn, auto ka_, auto kb_) {
        constexpr int K = decltype(kn)::value, KA = decltype(ka_)::value, KB = decltype(kb_)::value;
        float ra[K], rda[K];
#pragma unroll
        for (int s = 0; s < K; ++s) {
          const int n = 4 * s + lk;
          ra[s] = n < n_max ? Rb[n * PTS + pa] : 0.f;
          rda[s] = n < n_max ? Rdb[n * PTS + pa] : 0.f;
        }
        constexpr int KG = KB - KA < 3 ? KB - KA : 3;
#pragma unroll
        for (int kt0 = KA; kt0 < KB; kt0 += KG) {
          floatx4 Ur[KG], Ui[KG], Vr[KG], Vi[KG];
#pragma unroll
          for (int g = 0; g < KG; ++g) {
            Ur[g] = Ui[g] = Vr[g] = Vi[g] = floatx4{0.f, 0.f, 0.f, 0.f};
            if (kt0 + g < KB) {
              const int k = 16 * (kt0 + g) + lcol;
              float hr[K], hm[K];
#pragma unroll
              for (int s = 0; s < K; ++s) {
                const int n = 4 * s + lk;
                hr[s] = n < n_max ? H[n * HS + k] : 0.f;
                hm[s] = n < n_max ? H[n * HS + KMP + k] : 0.f;
              }
#pragma unroll
              for (int s = 0; s < K; ++s) {
                Ur[g] = mfma4(ra[s], hr[s], Ur[g]);
                Ui[g] = mfma4(ra[s], hm[s], Ui[g]);
                Vr[g] = mfma4(rda[s], hr[s], Vr[g]);
                Vi[g] = mfma4(rda[s], hm[s], Vi[g]);
              }
            }
          }
#pragma unroll
          for (int g = 0; g < KG; ++g) {
            const int k = 16 * (kt0 + g) + lcol;
            const bool act = kt0 + g < KB && k < KM;
            const int lm = act ? klm[k] : 0;
            const int l = lm & 255, m = lm >> 8;
            const int kb = l * (l + 1) / 2, kb1 = (l + 1) * (l + 2) / 2;
            // the P_l^m rows a k reads, clamped into range where it lacks them (m = 0, m = l), so that
            // every load of a row is issued before the branch on act and on m
            const int rp1 = kb + (m < l ? m + 1 : m), rm1 = kb + (m > 0 ? m - 1 : 0), r1m1 = kb1 + (m > 0 ? m - 1 : 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int lp = lrow + v;
              const float gl = Gl[l * PTS + lp], gdl = Gd[l * PTS + lp];
              const float2 e = Em[m * PTS + lp];
              const float pl = Fp[(kb + m) * PTS + lp];
              const float fp1 = Fp[rp1 * PTS + lp], fm1 = Fp[rm1 * PTS + lp];
              const float f1m1 = Fp[r1m1 * PTS + lp], f1p1 = Fp[(kb1 + m + 1) * PTS + lp];
              if (act) {
                const float ur = Ur[g][v], ui = Ui[g][v], vr = Vr[g][v], vi = Vi[g][v];
                const float eu = e.x * ur + e.y * ui;     // Re(E conj u)
                const float ev = e.x * vr + e.y * vi;     // Re(E conj v)
                const float ex = e.x * ui - e.y * ur;     // Re(i E conj u)
                const float pp1 = m < l ? fp1 : 0.f;
                float dms, dsin = 0.f;
                if (m == 0) {
                  dms = -pp1;
                } else {
                  dms = 0.5f * (float)((l + m) * (l - m + 1)) * fm1 - 0.5f * pp1;
                  dsin = 0.5f * (float)((l - m + 1) * (l - m + 2)) * f1m1 + 0.5f * f1p1;
                }
                // A = pl g E, A_r = pl g' E, A_pol = dms g E, A_azi = i dsin g E (the prefactor is in H)
                fr[v] += pl * (gl * ev + gdl * eu);
                fo[v] += dms * gl * eu;
                fa[v] += dsin * gl * ex;
              }
            }
          }
        }
      };
      auto uv_k = [&](auto kn) {
        if constexpr (KT > 1) {
          if (ksplit) {
            if (wq < 2) uv(kn, std::integral_constant<int, 0>{}, std::integral_constant<int, KS>{});
            else uv(kn, std::integral_constant<int, KS>{}, std::integral_constant<int, KT>{});
            return;
          }
        }
        uv(kn, std::integral_constant<int, 0>{}, std::integral_constant<int, KT>{});
      };
      if constexpr (MODE == 1) uv_k(std::integral_constant<int, KN - 2>{});
      else uv_k(std::integral_constant<int, KN>{});
      // sums over the 16 k of a row: prefix sums along the DPP row, total in lane 15
      const int fo0 = ksplit && wq >= 2 ? 32 : 0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float tr = row_total(fr[v]), ta = row_total(fa[v]), to = row_total(fo[v]);
        if (li == 15) {
          const int lp = fo0 + 16 * rw + 4 * lk + v;
          fout[lp] = tr;
          fout[PT + lp] = ta;
          fout[2 * PT + lp] = to;
        }
      }
    }
    __syncthreads();
    const int p = tb + lane;
    if (wv == 0 && p < p1 && p != p0) {
      float f_rad = fout[lane], f_azi = fout[PT + lane], f_pol = fout[2 * PT + lane];
      if (ksplit) {  // (lane < 32 here) the upper k tiles' sums
        f_rad += fout[32 + lane];
        f_azi += fout[PT + 32 + lane];
        f_pol += fout[2 * PT + 32 + lane];
      }
      const float st_ = pr_st[lane], ct_ = pr_ct[lane], sf_ = pr_sf[lane], cf_ = pr_cf[lane];
      const float inv_r = 1.f / pr_r[lane];
      const float x0 = st_ * cf_ * f_rad - ct_ * cf_ * inv_r * f_pol - sf_ * inv_r * f_azi;
      const float x1 = st_ * sf_ * f_rad - ct_ * sf_ * inv_r * f_pol + cf_ * inv_r * f_azi;
      const float x2 = ct_ * f_rad + st_ * inv_r * f_pol;
      pair_out[3 * (size_t)p] = x0;
      pair_out[3 * (size_t)p + 1] = x1;
      pair_out[3 * (size_t)p + 2] = x2;
      self[0] -= x0; self[1] -= x1; self[2] -= x2;
    }
  }
  block_sums_f<3>(self, red);
  if (tx == 0) {
    pair_out[3 * (size_t)p0] = self[0];
    pair_out[3 * (size_t)p0 + 1] = self[1];
    pair_out[3 * (size_t)p0 + 2] = self[2];
  }
  }  // site queue
}

// the central-atom term of each site: the sum of its species' vectors (species 1-based)
__global__ void site_central_k(int n_sites, int n_max, int max_mult, const int32_t* __restrict__ mult,
                               const int32_t* __restrict__ species, const float* __restrict__ central,
                               float* __restrict__ cenv) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n_sites * n_max) return;
  const int i = (int)(t / n_max), n = (int)(t - (size_t)i * n_max);
  float v = 0.f;
  for (int m = 0; m < mult[i]; ++m) v += central[(size_t)(species[(size_t)i * max_mult + m] - 1) * n_max + n];
  cenv[t] = v;
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

// Dynamic LDS of the site kernel. Its static arrays take about 3.4 KB more. Measured on MI355X:
// workgroups of 53,360 B (static + dynamic, l_max 8, n_max 24, 3 species) run three per CU, while
// 53,936 B ran only two at a time although the occupancy query still reported three, so the
// LDS budget of a three-per-CU shape is tight.
size_t site_lds(int n_species, int n_max, int l_max, int n_soap, int n_soap_unc) {
  const int KM = (l_max + 1) * (l_max + 2) / 2, KM1 = (l_max + 2) * (l_max + 3) / 2, L1 = l_max + 1;
  const int CSTR = ((2 * KM + 15) / 16) * 16, CS = CSTR + 4, KMP = ((KM + 15) / 16) * 16, HS = 2 * KMP + 16;
  const size_t c = (size_t)n_max * std::max(CS, HS);
  const size_t fwd = (size_t)n_max * PTF + (size_t)L1 * PTS + (size_t)2 * KM * PTF;
  const size_t spec = (size_t)L1 * n_max * n_max + 2 * n_soap + n_soap_unc;
  const size_t bwd = (size_t)2 * n_max * PTS + ((KM1 * PTS + 1) & ~1) + 4 * L1 * PTS + 3 * PT;
  (void)n_species;
  return sizeof(float) * (c + std::max(std::max(fwd, spec), bwd));
}

template <int L, int NTT>
void launch_site(bool fo, hipStream_t st, const AdjParams& P, size_t sh, const int* order, const int* site_ptr,
                 const float* rtab, const float* atab, const float* preflm, const float* cenv,
                 const uint32_t* spmask, const double* rjs, const double* th, const double* ph, const float* G32,
                 const int4* comp, const int* out_ptr, const int* out_c, const float* out_el, const int* unc_ptr,
                 const int* unc_e, const int32_t* P_i, const float* P_el, const int2* odir, float* q, float* sdp,
                 float* pair, unsigned long long* tickets, int n_cu, int per_cu_cap) {
  // the short K form when every site's n rows fit in 4 NTT - 2 steps of 4
  const int mode = fo ? 2 : P.n_max <= 4 * (4 * NTT - 2) ? 1 : 0;
  auto run = [&](auto mode_) {
    constexpr int M = decltype(mode_)::value;
    const void* fn = reinterpret_cast<const void*>(&adjoint_site_k<L, NTT, M>);
    if (sh > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NTHREADS, sh) != hipSuccess || per_cu < 1) per_cu = 1;
    if (per_cu_cap > 0 && per_cu > per_cu_cap) per_cu = per_cu_cap;
    const int grid = std::max(1, std::min(P.n_sites, per_cu * n_cu));
    adjoint_site_k<L, NTT, M><<<grid, NTHREADS, sh, st>>>(P, order, site_ptr, rtab, atab, preflm, cenv, spmask, rjs, th,
                                                         ph, G32, comp, out_ptr, out_c, out_el, unc_ptr, unc_e, P_i,
                                                         P_el, odir, q, sdp, pair, tickets);
  };
  if (mode == 2) run(std::integral_constant<int, 2>{});
  else if (mode == 1) run(std::integral_constant<int, 1>{});
  else run(std::integral_constant<int, 0>{});
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
  // radial set-up items: (species, chunk, pair) filling the 256 threads where the species are few
  P.nch = std::max(1, 256 / (ns * 64));
  P.qc = 0;
  for (int q = 0; q < ns; ++q) P.qc = std::max(P.qc, (p->alpha_max[q] + P.nch - 1) / P.nch);
  P.keep = ns * P.nch * 64 <= 256 && P.qc <= 8;  // the kernel keeps 8 radial values per thread
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
    // poly3operator has no cutoff test: its tables run to the largest cutoff, split where the
    // integration limits of the atom's cubic (half-width K sigma(r)) switch: r - w = 0, r - w =
    // rcut_soft, r + w = rcut_soft, r = rcut_soft, r + w = rcut_hard, r = rcut_hard
    const bool op = p->basis == ST_BASIS_POLY3OPERATOR;
    const double range = op ? rcut_max : rc;
    std::vector<double> br;
    if (op) {
      const double K = 2.0 * std::sqrt(2.0 * std::log(2.0));
      std::vector<double> cand = {(rs_n - K * s0) / (1.0 + K * c), rs_n, (1.0 - K * s0) / (1.0 + K * c), 1.0};
      if (1.0 - K * c > 0.0) {
        cand.push_back(K * s0 / (1.0 - K * c));
        cand.push_back((rs_n + K * s0) / (1.0 - K * c));
      }
      for (double b : cand)
        if (b * rc > 0.0 && b * rc < range) br.push_back(b * rc);
    } else {
      if (1.0 - rs_n > 0.0) {
        double b = rc * (rs_n - 4.0 * s0) / (1.0 + 4.0 * c);
        if (b > 0.0 && b < rc) br.push_back(b);
      }
      if (P.gauss && c < 0.25) {
        double b = rc * 8.0 * s0 / (1.0 - 4.0 * c);
        if (b > 0.0 && b < rc) br.push_back(b);
      }
    }
    std::sort(br.begin(), br.end());
    if (br.size() + 1 > (size_t)MAXSEG) {
      msg = "too many radial table segments";
      return ST_ERR_UNSUPPORTED;
    }
    std::vector<double>& e = edges[q];
    e.push_back(0.0);
    for (double b : br)
      if (b - e.back() > 1e-6 * rc) e.push_back(b);
    e.push_back(range);
    sp.a = p->alpha_max[q];
    sp.rcut = range;
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
  P.ang_ih = (float)(NTAB / rcut_max);
  for (int q = 0; q < ns; ++q) P.sp[q].atab_off = q * (P.ang_n - 1) * 4 * (L + 1);
  if (key == s.table_key) {
    for (int q = 0; q < ns; ++q) {
      P.sp[q].rtab_off = s.rt_off[q];
      P.sp[q].atab_off = s.at_off[q];
    }
    return ST_OK;
  }

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
          cell_cubic(v[n0 + n], d[n0 + n], v[n1 + n], d[n1 + n], hseg[j], cc + 4 * n, cc + 4 * n + 1,
                     cc + 4 * n + 2, cc + 4 * n + 3);
      }
      k0 += sp.seg_n[j];
    }
  }
  // Species with equal tables share one copy (equal parameters give equal functions), which keeps
  // the tables the site kernel reads per pair small enough to stay in L2.
  {
    std::vector<float> kept;
    int orig[ADJ_MAX_SPECIES];
    for (int q = 0; q < ns; ++q) orig[q] = P.sp[q].rtab_off;
    for (int q = 0; q < ns; ++q) {
      SpeciesTab& sp = P.sp[q];
      const size_t len = (size_t)(sp.seg_base[sp.nseg - 1] + sp.seg_n[sp.nseg - 1] - 1) * 4 * sp.a;
      int same = -1;
      for (int r = 0; r < q && same < 0; ++r) {
        const SpeciesTab& o = P.sp[r];
        bool eq = o.a == sp.a && o.nseg == sp.nseg;
        for (int j = 0; eq && j < sp.nseg; ++j)
          eq = o.seg_lo[j] == sp.seg_lo[j] && o.seg_inv_h[j] == sp.seg_inv_h[j] && o.seg_n[j] == sp.seg_n[j] &&
               o.seg_base[j] == sp.seg_base[j];
        if (eq && std::memcmp(rt.data() + orig[r], rt.data() + orig[q], len * 4) == 0) same = r;
      }
      if (same >= 0) {
        sp.rtab_off = P.sp[same].rtab_off;
      } else {
        sp.rtab_off = (int)kept.size();
        kept.insert(kept.end(), rt.begin() + orig[q], rt.begin() + orig[q] + len);
      }
    }
    rt.swap(kept);
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
          cell_cubic(a0[l], a0[L1 + l], a1[l], a1[L1 + l], hh, cc + 4 * l, cc + 4 * l + 1, cc + 4 * l + 2,
                     cc + 4 * l + 3);
      }
    }
    (void)hipFree(dout.p);
  }
  {
    // the same sharing for the angular tables
    const size_t len = (size_t)ncell * 4 * L1;
    std::vector<float> kept;
    for (int q = 0; q < ns; ++q) {
      int same = -1;
      for (int r = 0; r < q && same < 0; ++r)
        if (std::memcmp(at.data() + r * len, at.data() + q * len, len * 4) == 0) same = r;
      if (same >= 0) {
        P.sp[q].atab_off = P.sp[same].atab_off;
      } else {
        P.sp[q].atab_off = (int)kept.size();
        kept.insert(kept.end(), at.begin() + q * len, at.begin() + (q + 1) * len);
      }
    }
    at.swap(kept);
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
  for (int q = 0; q < ns; ++q) {
    s.rt_off[q] = P.sp[q].rtab_off;
    s.at_off[q] = P.sp[q].atab_off;
  }
  s.n_pf = (int)pf.size();
  return ST_OK;
}

int adjoint_run(AdjointState& s, hipStream_t st, const AdjointInputs& in, const Timer& timer, std::string& msg) {
  AdjParams P;
  int rc;
  timer("adj_tables", [&] { rc = build_tables(s, st, in, P, msg); });
  if (rc) return rc;
  // (the per-component dE/dp array, n_soap_unc, is needed only without a direct output map)
  const size_t sh = site_lds(in.p->n_species, in.n_max, in.l_max, in.n_soap, in.ocomp ? 0 : in.n_soap_unc);
  if (sh > in.lds_max) {
    msg = "the FP32 site working set exceeds the LDS";
    return ST_ERR_UNSUPPORTED;
  }
  if (in.n_max > 48) {
    msg = "the FP32 path supports n_max <= 48";
    return ST_ERR_UNSUPPORTED;
  }
  // compression element tables in FP32: by output, then in P order
  size_t n_oe = 0;
  if (in.out_ptr) n_oe = in.out_el_h->size();
  if (s.el_key != *in.tables_key) {
    std::vector<float> all;
    if (in.out_ptr) {
      all.insert(all.end(), in.out_el_h->begin(), in.out_el_h->end());
      all.insert(all.end(), in.P_el_h->begin(), in.P_el_h->end());
    }
    if ((rc = ensure(s.el32, (all.size() + 1) * 4, msg))) return rc;
    if (!all.empty()) (void)hipMemcpyAsync(s.el32.p, all.data(), all.size() * 4, hipMemcpyHostToDevice, st);
    // direct maps: output q of component (a, b, l), a <= b, packed as a | b << 8 | l << 16, with
    // the weight of S_l(a, b) in q (its element, times sqrt 2 off the diagonal) as float bits;
    // components follow the trivial order, c = (a n_max - a (a - 1) / 2 + b - a) (l_max + 1) + l
    if (in.ocomp) {
      const int nm = in.n_max, L1 = in.l_max + 1;
      std::vector<int2> od((size_t)in.n_soap, make_int2(0, 0));
      for (int a = 0; a < nm; ++a)
        for (int b = a; b < nm; ++b)
          for (int l = 0; l < L1; ++l) {
            const int c = (a * nm - a * (a - 1) / 2 + (b - a)) * L1 + l;
            const int q = (*in.umap_h)[c];
            if (q < 0) continue;
            const float w = (float)((a == b ? 1.0 : 1.4142135623730951) * (*in.uel_h)[c]);
            int wb;
            std::memcpy(&wb, &w, 4);
            od[q] = make_int2(a | b << 8 | l << 16, wb);
          }
      if ((rc = ensure(s.odir, od.size() * sizeof(int2), msg))) return rc;
      (void)hipMemcpyAsync(s.odir.p, od.data(), od.size() * sizeof(int2), hipMemcpyHostToDevice, st);
    }
    (void)hipStreamSynchronize(st);
    s.el_key = *in.tables_key;
  }
  if ((rc = ensure(s.soap, (size_t)in.n_sites * in.n_soap * 4, msg))) return rc;
  if ((rc = ensure(s.sdp, (size_t)in.n_sites * 4, msg))) return rc;
  if ((rc = ensure(s.pair, (size_t)in.n_pairs * 12, msg))) return rc;
  if (s.g32_sites != in.n_sites || s.g32_soap != in.n_soap) {
    msg = "dE/dq of this shape has not been uploaded";
    return ST_ERR_STATE;
  }
  const float* pf = s.tabinfo.as<float>();
  const float* cen = pf + s.n_pf;
  const float* oel = in.out_ptr ? s.el32.as<float>() : nullptr;
  const float* pel = in.out_ptr ? s.el32.as<float>() + n_oe : nullptr;
  const int2* odir = in.ocomp ? s.odir.as<int2>() : nullptr;
  // the central-atom term of each site (poly3gauss), once per neighbour list and table set
  if (P.gauss && (!s.cen_ok || s.cen_key != s.table_key)) {
    const size_t nc = (size_t)in.n_sites * in.n_max;
    if ((rc = ensure(s.cenv, nc * 4, msg))) return rc;
    if (nc > 0)
      site_central_k<<<(unsigned)((nc + 255) / 256), 256, 0, st>>>(in.n_sites, in.n_max, in.max_mult, in.mult,
                                                                 in.species, cen, s.cenv.as<float>());
    s.cen_ok = true;
    s.cen_key = s.table_key;
  }
  const float* cenv = P.gauss ? s.cenv.as<float>() : nullptr;
  // sites by falling neighbour count (longest first from the queue), once per neighbour list
  if (!s.order_ok) {
    std::vector<int> ord(in.n_sites);
    std::iota(ord.begin(), ord.end(), 0);
    const std::vector<int>& sp = *in.site_ptr_h;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sp[a + 1] - sp[a] > sp[b + 1] - sp[b]; });
    if ((rc = ensure(s.order, (size_t)in.n_sites * 4, msg))) return rc;
    if (in.n_sites > 0 &&
        hipMemcpyAsync(s.order.p, ord.data(), ord.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
      msg = "upload of the site order failed";
      return ST_ERR_HIP;
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
      msg = "upload of the site order failed";
      return ST_ERR_HIP;
    }
    s.order_ok = true;
  }
  if (!s.tickets.p) {
    if ((rc = ensure(s.tickets, 8, msg))) return rc;
    if (hipMemsetAsync(s.tickets.p, 0, 8, st) != hipSuccess) {
      msg = "clearing the site queue failed";
      return ST_ERR_HIP;
    }
  }
  const int ntt = (in.n_max + 15) / 16;
  // the site kernel with at most per_cu_cap workgroups per CU in its persistent grid (0: occupancy)
  auto site = [&](int per_cu_cap) {
    switch (in.l_max * 4 + ntt) {
#define LS1(l, t)                                                                                                  \
  case l * 4 + t:                                                                                                  \
    launch_site<l, t>(in.forward_only, st, P, sh, s.order.as<int>(), in.site_ptr, s.rtab.as<float>(), s.atab.as<float>(), pf, cenv, \
                      in.spmask, in.rjs, in.thetas, in.phis, s.G32.as<float>(), in.comp, in.out_ptr, in.out_c, oel, \
                      in.unc_ptr, in.unc_e, in.P_i, pel, odir, s.soap.as<float>(), s.sdp.as<float>(),              \
                      s.pair.as<float>(), s.tickets.as<unsigned long long>(), in.n_cu, per_cu_cap);              \
    return true;
#define LS(l) LS1(l, 1) LS1(l, 2) LS1(l, 3)
      LS(0) LS(1) LS(2) LS(3) LS(4) LS(5) LS(6) LS(7) LS(8) LS(9) LS(10) LS(11) LS(12)
#undef LS
#undef LS1
      default: return false;
    }
  };
  int cap = 0;
  if (in.autotune) {
    // the persistent grid's workgroups per CU: the occupancy limit or two, whichever runs faster
    const std::string key = std::to_string(in.l_max) + " " + std::to_string(ntt) + " " + std::to_string(sh) + " " +
                            std::to_string(in.n_soap) + " " + std::to_string(in.n_sites / 256) +
                            (in.forward_only ? " fo" : "");
    auto it = s.tuned.find(key);
    if (it != s.tuned.end()) {
      cap = it->second;
    } else {
      hipEvent_t a = nullptr, b = nullptr;
      float best = 0.f;
      bool any = false;
      if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
        for (int c : {0, 2}) {
          site(c);  // warm-up
          launch_counter_reset(st, s.tickets.as<unsigned long long>());
          (void)hipEventRecord(a, st);
          site(c);
          (void)hipEventRecord(b, st);
          launch_counter_reset(st, s.tickets.as<unsigned long long>());
          float ms = 0.f;
          if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) continue;
          if (!any || ms < best) {
            best = ms;
            cap = c;
            any = true;
          }
        }
      }
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      s.tuned[key] = cap;
    }
  }
  bool ok = true;
  timer("adj_site_f32", [&] { ok = site(cap); });
  if (!ok) {
    msg = "l_max out of the compiled range";
    return ST_ERR_UNSUPPORTED;
  }
  s.have_out = true;
  s.n_pairs = in.n_pairs;
  s.n_sites = in.n_sites;
  s.n_soap = in.n_soap;
  s.have_forces = false;
  if (in.pair_atom && !in.forward_only) {
    if ((rc = ensure(s.forces, (size_t)in.n_atoms * 24, msg)) ||
        (rc = ensure(s.virial_part, (size_t)gather_blocks(in.n_atoms) * 72, msg)) || (rc = ensure(s.virial, 72, msg))) {
      launch_counter_reset(st, s.tickets.as<unsigned long long>());  // the site queue starts at zero next time
      return rc;
    }
    timer("adj_forces", [&] {
      launch_forces_f32(st, in.n_atoms, in.n_pairs, s.pair.as<float>(), in.atom_ptr, in.atom_pairs, in.vec, in.rjs, in.thetas,
                        in.phis, s.forces.as<double>(), s.virial_part.as<double>(), s.virial.as<double>(),
                        s.tickets.as<unsigned long long>());
    });
    s.have_forces = true;
    s.n_atoms = in.n_atoms;
  } else {
    launch_counter_reset(st, s.tickets.as<unsigned long long>());
  }
  return ST_OK;
}

int adjoint_set_dedq(AdjointState& s, hipStream_t st, int n_sites, int n_soap, const double* G, std::string& msg) {
  int rc;
  if ((rc = ensure(s.G32, (size_t)n_sites * n_soap * 4 + 4, msg))) return rc;
  if (n_sites > 0) {
    dim3 grid((n_soap + 31) / 32, (n_sites + 31) / 32);
    transpose_g_k<<<grid, 256, 0, st>>>(n_sites, n_soap, G, s.G32.as<float>());
    if (hipGetLastError() != hipSuccess) {
      msg = "dE/dq conversion failed";
      return ST_ERR_HIP;
    }
  }
  s.g32_sites = n_sites;
  s.g32_soap = n_soap;
  return ST_OK;
}

int adjoint_prepare_dedq(AdjointState& s, int n_sites, int n_soap, std::string& msg) {
  int rc;
  if ((rc = ensure(s.G32, (size_t)n_sites * n_soap * 4 + 4, msg))) return rc;
  s.g32_sites = n_sites;
  s.g32_soap = n_soap;
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
