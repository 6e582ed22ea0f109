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
  float rcut, cw;
  float seg_lo[MAXSEG], seg_hi[MAXSEG], seg_inv_h[MAXSEG], seg_h[MAXSEG];
  int seg_base[MAXSEG], seg_n[MAXSEG];
  int rtab_off;   // floats into rtab: node-major, 2*a floats per node (values then derivatives)
  int atab_off;   // floats into atab: node-major, 2*(L+1) floats per node
};

struct AdjParams {
  int n_species, n_max, k_max, l_max, n_soap, n_soap_unc, n_sites, compressed, gauss;
  float rcut_max, ang_inv_h, ang_h;
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
__device__ __forceinline__ void hermite(float t, float h, float& h00, float& h10, float& h01, float& h11, float& d00,
                                        float& d10, float& d01, float& d11) {
  const float t2 = t * t, t3 = t2 * t;
  h00 = 2.f * t3 - 3.f * t2 + 1.f;
  h10 = (t3 - 2.f * t2 + t) * h;
  h01 = -2.f * t3 + 3.f * t2;
  h11 = (t3 - t2) * h;
  const float ih = 1.f / h;
  d00 = (6.f * t2 - 6.f * t) * ih;
  d10 = 3.f * t2 - 4.f * t + 1.f;
  d01 = (-6.f * t2 + 6.f * t) * ih;
  d11 = 3.f * t2 - 2.f * t;
}

// Locates r in the species' radial segments: returns the node row index and t.
__device__ __forceinline__ int radial_cell(const SpeciesTab& s, float r, float& t, float& h) {
  int j = 0;
  while (j + 1 < s.nseg && r >= s.seg_lo[j + 1]) ++j;
  float x = (r - s.seg_lo[j]) * s.seg_inv_h[j];
  int i = (int)x;
  if (i < 0) i = 0;
  if (i > s.seg_n[j] - 2) i = s.seg_n[j] - 2;
  t = x - (float)i;
  h = s.seg_h[j];
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

// One site per workgroup.
template <int L>
__global__ __launch_bounds__(NTHREADS) void adjoint_site_k(
    AdjParams P, const int* __restrict__ order, const int* __restrict__ site_ptr, const float* __restrict__ rtab,
    const float* __restrict__ atab, const float* __restrict__ preflm_g, const float* __restrict__ central,
    const int32_t* __restrict__ site_species, int max_mult, const int32_t* __restrict__ species_mult,
    const uint32_t* __restrict__ spmask, const double* __restrict__ rjs, const double* __restrict__ thetas,
    const double* __restrict__ phis, const float* __restrict__ G32, const int4* __restrict__ comp,
    const int* __restrict__ out_ptr, const int* __restrict__ out_c, const float* __restrict__ out_el,
    const int* __restrict__ unc_ptr, const int* __restrict__ unc_e, const int32_t* __restrict__ P_i,
    const float* __restrict__ P_el, float* __restrict__ q_out, float* __restrict__ sdp_out,
    float* __restrict__ pair_out) {
  constexpr int KM = (L + 1) * (L + 2) / 2;
  extern __shared__ float smem[];
  __shared__ float red[16];
  __shared__ float preflm[KM];
  const int i = order ? order[blockIdx.x] : blockIdx.x;
  const int n_max = P.n_max, k_max = KM;
  const int p0 = site_ptr[i], p1 = site_ptr[i + 1];
  const float fourpi = 12.566370614359172f;
  // LDS: C hi/lo, H (float2 [n][k]); tile region
  float2* C = reinterpret_cast<float2*>(smem);                 // n_max * k_max
  float2* Clo = C + n_max * k_max;                             // compensation
  float2* H = Clo + n_max * k_max;                             // n_max * k_max
  float* region = reinterpret_cast<float*>(H + n_max * k_max);
  float* Rt = region;                                          // PT * n_max
  float2* At = reinterpret_cast<float2*>(Rt + PT * n_max);     // PT * k_max

  for (int t = threadIdx.x; t < KM; t += blockDim.x) preflm[t] = preflm_g[t];
  for (int t = threadIdx.x; t < n_max * k_max; t += blockDim.x) {
    C[t] = make_float2(0.f, 0.f);
    Clo[t] = make_float2(0.f, 0.f);
  }
  // --------------------------------------------------------------- forward
  for (int tb = p0; tb < p1; tb += PT) {
    const int np_ = min(PT, p1 - tb);
    __syncthreads();
    if (threadIdx.x < PT) {
      const int lp = threadIdx.x;
      float* R = Rt + lp * n_max;
      float2* A = At + lp * k_max;
      for (int n = 0; n < n_max; ++n) R[n] = 0.f;
      for (int k = 0; k < k_max; ++k) A[k] = make_float2(0.f, 0.f);
      if (lp < np_) {
        const int p = tb + lp;
        const float r = (float)rjs[p];
        const uint32_t msk = spmask[p];
        const bool self = p == p0;
        if (msk != 0 && r < P.rcut_max) {
          for (int s = 0; s < P.n_species; ++s) {
            const SpeciesTab& sp = P.sp[s];
            if (!(msk & (1u << s)) || !(r < sp.rcut)) continue;
            if (self && !sp.do_central) continue;
            float t, h;
            const int row = radial_cell(sp, r, t, h);
            float h00, h10, h01, h11, d00, d10, d01, d11;
            hermite(t, h, h00, h10, h01, h11, d00, d10, d01, d11);
            const float* n0 = rtab + sp.rtab_off + (size_t)row * 2 * sp.a;
            const float* n1 = n0 + 2 * sp.a;
            const float sc = self ? sp.cw : 1.f;
            for (int n = 0; n < sp.a; ++n)
              R[sp.off + n] = sc * (h00 * n0[n] + h10 * n0[sp.a + n] + h01 * n1[n] + h11 * n1[sp.a + n]);
          }
          const int s0 = __ffs(msk) - 1;
          // angular radial factors g_l(r)
          float x = r * P.ang_inv_h;
          int ia = (int)x;
          if (ia > P.ang_n - 2) ia = P.ang_n - 2;
          const float ta = x - (float)ia;
          float h00, h10, h01, h11, d00, d10, d01, d11;
          hermite(ta, P.ang_h, h00, h10, h01, h11, d00, d10, d01, d11);
          const float* g0 = atab + P.sp[s0].atab_off + (size_t)ia * 2 * (L + 1);
          const float* g1 = g0 + 2 * (L + 1);
          float g[L + 1];
#pragma unroll
          for (int l = 0; l <= L; ++l) g[l] = h00 * g0[l] + h10 * g0[L + 1 + l] + h01 * g1[l] + h11 * g1[L + 1 + l];
          float st_, ct_, sf_, cf_;
          sincosf((float)thetas[p], &st_, &ct_);
          sincosf((float)phis[p], &sf_, &cf_);
          // e^{-i m phi}
          float er[L + 1], ei[L + 1];
          er[0] = 1.f; ei[0] = 0.f;
#pragma unroll
          for (int m = 1; m <= L; ++m) {
            er[m] = er[m - 1] * cf_ + ei[m - 1] * sf_;
            ei[m] = ei[m - 1] * cf_ - er[m - 1] * sf_;
          }
          angular_rows<L>(ct_, st_, preflm, [&](int k, int l, int m, float pl, float, float) {
            const float v = pl * g[l];
            A[k] = make_float2(v * er[m], v * ei[m]);
          });
        }
      }
    }
    __syncthreads();
    // cnk += sum over the tile of 4 pi R_n A_k (Kahan across pairs)
    for (int t = threadIdx.x; t < n_max * k_max; t += blockDim.x) {
      const int n = t / k_max, k = t - n * k_max;
      float2 s = C[t], c = Clo[t];
      for (int lp = 0; lp < np_; ++lp) {
        const float rr = fourpi * Rt[lp * n_max + n];
        const float2 a = At[lp * k_max + k];
        float y = rr * a.x - c.x, tt = s.x + y;
        c.x = (tt - s.x) - y; s.x = tt;
        y = rr * a.y - c.y; tt = s.y + y;
        c.y = (tt - s.y) - y; s.y = tt;
      }
      C[t] = s;
      Clo[t] = c;
    }
  }
  __syncthreads();
  // apply the Kahan compensation, then the poly3gauss central-atom term on k = 0
  for (int t = threadIdx.x; t < n_max * k_max; t += blockDim.x) {
    float2 s = C[t];
    const float2 c = Clo[t];
    s.x -= c.x;
    s.y -= c.y;
    C[t] = s;
  }
  __syncthreads();
  if (P.gauss) {
    for (int n = threadIdx.x; n < n_max; n += blockDim.x) {
      float add = 0.f;
      for (int m = 0; m < species_mult[i]; ++m) {
        const int s = site_species[(size_t)i * max_mult + m] - 1;
        add += central[(size_t)s * n_max + n];
      }
      C[n * k_max].x += add;
    }
  }
  __syncthreads();
  // --------------------------------------------------------------- spectrum
  const int n_soap = P.n_soap;
  float* qv = region;                 // n_soap: spectrum, then q
  float* Gp = qv + n_soap;            // n_soap
  float* Gu = Gp + n_soap;            // n_soap_unc
  float ss = 0.f;
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x) {
    float v = 0.f;
    const int e0 = P.compressed ? out_ptr[s] : s, e1 = P.compressed ? out_ptr[s + 1] : s + 1;
    for (int e = e0; e < e1; ++e) {
      const int4 cm = comp[P.compressed ? out_c[e] : e];
      const int n = cm.x, np = cm.y, l = cm.z, kb = l * (l + 1) / 2;
      const float base = n != np ? 1.41421356f : 1.f;
      float acc = 0.f;
      for (int m = 0; m <= l; ++m) {
        const float2 u = C[n * k_max + kb + m], w = C[np * k_max + kb + m];
        acc += (m > 0 ? 2.f * base : base) * (u.x * w.x + u.y * w.y);
      }
      v += (P.compressed ? out_el[e] : 1.f) * acc;
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
    if (P.compressed) {
      v = 0.f;
      for (int e = unc_ptr[c]; e < unc_ptr[c + 1]; ++e) v += P_el[unc_e[e]] * Gp[P_i[unc_e[e]] - 1];
    } else {
      v = Gp[c];
    }
    Gu[c] = v;
  }
  __syncthreads();
  // H(k,n) = w(m) sum_np Y(n,np,l) c(k,np)
  const int L1 = L + 1;
  for (int t = threadIdx.x; t < n_max * k_max; t += blockDim.x) {
    const int n = t / k_max, k = t - n * k_max;
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= k) ++l;
    const int m = k - l * (l + 1) / 2;
    float hr = 0.f, hi = 0.f;
    for (int np = 0; np < n_max; ++np) {
      const int a = n < np ? n : np, b = n < np ? np : n;
      const int c = (a * n_max - a * (a - 1) / 2 + (b - a)) * L1 + l;
      const float y = a == b ? 2.f * Gu[c] : 1.41421356f * Gu[c];
      const float2 cv = C[np * k_max + k];
      hr += y * cv.x;
      hi += y * cv.y;
    }
    const float w = m == 0 ? fourpi : 2.f * fourpi;
    H[n * k_max + k] = make_float2(w * hr, w * hi);
  }
  // --------------------------------------------------------------- backward
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ng = (n_max + 3) / 4;
  const int nlo = wv * ng, nhi = min(n_max, nlo + ng);
  float* fred = region;  // [4][PT][3]
  float self0 = 0.f, self1 = 0.f, self2 = 0.f;
  for (int tb = p0; tb < p1; tb += PT) {
    __syncthreads();
    const int p = tb + lane;
    float fr = 0.f, fa = 0.f, fo = 0.f;
    if (p < p1 && p != p0) {
      const float r = (float)rjs[p];
      const uint32_t msk = spmask[p];
      if (msk != 0 && r < P.rcut_max) {
        float Rn[12], Rd[12];
#pragma unroll
        for (int j = 0; j < 12; ++j) Rn[j] = Rd[j] = 0.f;
        for (int s = 0; s < P.n_species; ++s) {
          const SpeciesTab& sp = P.sp[s];
          if (!(msk & (1u << s)) || !(r < sp.rcut)) continue;
          const int lo = max(sp.off, nlo), hi = min(sp.off + sp.a, nhi);
          if (lo >= hi) continue;
          float t, h;
          const int row = radial_cell(sp, r, t, h);
          float h00, h10, h01, h11, d00, d10, d01, d11;
          hermite(t, h, h00, h10, h01, h11, d00, d10, d01, d11);
          const float* n0 = rtab + sp.rtab_off + (size_t)row * 2 * sp.a;
          const float* n1 = n0 + 2 * sp.a;
#pragma unroll
          for (int j = 0; j < 12; ++j) {
            const int n = nlo + j;
            if (n >= lo && n < hi) {
              const int q = n - sp.off;
              Rn[j] = h00 * n0[q] + h10 * n0[sp.a + q] + h01 * n1[q] + h11 * n1[sp.a + q];
              Rd[j] = d00 * n0[q] + d10 * n0[sp.a + q] + d01 * n1[q] + d11 * n1[sp.a + q];
            }
          }
        }
        const int s0 = __ffs(msk) - 1;
        float x = r * P.ang_inv_h;
        int ia = (int)x;
        if (ia > P.ang_n - 2) ia = P.ang_n - 2;
        const float ta = x - (float)ia;
        float h00, h10, h01, h11, d00, d10, d01, d11;
        hermite(ta, P.ang_h, h00, h10, h01, h11, d00, d10, d01, d11);
        const float* g0 = atab + P.sp[s0].atab_off + (size_t)ia * 2 * (L + 1);
        const float* g1 = g0 + 2 * (L + 1);
        float g[L + 1], gd[L + 1];
#pragma unroll
        for (int l = 0; l <= L; ++l) {
          g[l] = h00 * g0[l] + h10 * g0[L + 1 + l] + h01 * g1[l] + h11 * g1[L + 1 + l];
          gd[l] = d00 * g0[l] + d10 * g0[L + 1 + l] + d01 * g1[l] + d11 * g1[L + 1 + l];
        }
        float st_, ct_, sf_, cf_;
        sincosf((float)thetas[p], &st_, &ct_);
        sincosf((float)phis[p], &sf_, &cf_);
        float er[L + 1], ei[L + 1];
        er[0] = 1.f; ei[0] = 0.f;
#pragma unroll
        for (int m = 1; m <= L; ++m) {
          er[m] = er[m - 1] * cf_ + ei[m - 1] * sf_;
          ei[m] = ei[m - 1] * cf_ - er[m - 1] * sf_;
        }
        angular_rows<L>(ct_, st_, preflm, [&](int k, int l, int m, float pl, float dms, float dsin) {
          float ur = 0.f, ui = 0.f, vr = 0.f, vi = 0.f;
#pragma unroll
          for (int j = 0; j < 12; ++j) {
            if (nlo + j < nhi) {
              const float2 hh = H[(nlo + j) * k_max + k];
              ur += Rn[j] * hh.x; ui += Rn[j] * hh.y;
              vr += Rd[j] * hh.x; vi += Rd[j] * hh.y;
            }
          }
          const float e_r = er[m], e_i = ei[m];
          // A = pl g E, A_r = pl g' E, A_pol = dms g E, A_azi = i dsin g E
          const float a = pl * g[l], ar = pl * gd[l], ap = dms * g[l], az = dsin * g[l];
          const float Ar = a * e_r, Ai = a * e_i;
          fr += Ar * vr + Ai * vi + ar * (e_r * ur + e_i * ui);
          fo += ap * (e_r * ur + e_i * ui);
          // A_azi = i * az * E = (-az e_i, az e_r)
          fa += -az * e_i * ur + az * e_r * ui;
        });
      }
    }
    fred[(wv * PT + lane) * 3] = fr;
    fred[(wv * PT + lane) * 3 + 1] = fa;
    fred[(wv * PT + lane) * 3 + 2] = fo;
    __syncthreads();
    if (wv == 0 && p < p1 && p != p0) {
      float f_rad = 0.f, f_azi = 0.f, f_pol = 0.f;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        f_rad += fred[(w * PT + lane) * 3];
        f_azi += fred[(w * PT + lane) * 3 + 1];
        f_pol += fred[(w * PT + lane) * 3 + 2];
      }
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
    } else if (wv == 0 && p < p1) {
      // beyond the cutoff or of no species: zero gradient, as the reference's zero columns give
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

size_t site_lds(int n_max, int k_max, int n_soap, int n_soap_unc) {
  size_t c = sizeof(float2) * 3 * n_max * k_max;
  size_t fwd = sizeof(float) * PT * n_max + sizeof(float2) * PT * k_max;
  size_t bwd = sizeof(float) * (2 * n_soap + n_soap_unc);
  size_t red = sizeof(float) * 4 * PT * 3;
  return c + std::max(std::max(fwd, bwd), red);
}

template <int L>
void launch_site(hipStream_t st, const AdjParams& P, size_t sh, const int* order, const int* site_ptr,
                 const float* rtab, const float* atab, const float* preflm, const float* central,
                 const int32_t* species, int max_mult, const int32_t* mult, const uint32_t* spmask, const double* rjs,
                 const double* th, const double* ph, const float* G32, const int4* comp, const int* out_ptr,
                 const int* out_c, const float* out_el, const int* unc_ptr, const int* unc_e, const int32_t* P_i,
                 const float* P_el, float* q, float* sdp, float* pair) {
  adjoint_site_k<L><<<P.n_sites, NTHREADS, sh, st>>>(P, order, site_ptr, rtab, atab, preflm, central, species, max_mult,
                                                     mult, spmask, rjs, th, ph, G32, comp, out_ptr, out_c, out_el,
                                                     unc_ptr, unc_e, P_i, P_el, q, sdp, pair);
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
  P.rcut_max = (float)rcut_max;
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
  // radial segments
  int node_total = 0, rfloats = 0;
  double seg_edge_lo[ADJ_MAX_SPECIES][MAXSEG], seg_edge_hi[ADJ_MAX_SPECIES][MAXSEG];
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
    std::vector<double> edges{0.0};
    for (double b : br)
      if (b - edges.back() > 1e-6 * rc) edges.push_back(b);
    edges.push_back(rc);
    sp.a = p->alpha_max[q];
    sp.rcut = (float)rc;
    sp.cw = (float)p->central_weight[q];
    sp.do_central = P.gauss ? 0 : (p->central_weight[q] != 0.0);
    sp.nseg = (int)edges.size() - 1;
    sp.rtab_off = rfloats;
    int base = 0;
    for (int j = 0; j < sp.nseg; ++j) {
      // segment starts rounded up to float so that a float r >= seg_lo lies beyond the jump
      float lo_f = (float)edges[j];
      if ((double)lo_f < edges[j]) lo_f = std::nextafter(lo_f, 1e30f);
      float hi_f = (float)edges[j + 1];
      if (j + 1 < sp.nseg && (double)hi_f < edges[j + 1]) hi_f = std::nextafter(hi_f, 1e30f);
      const double lo = j == 0 ? 0.0 : (double)lo_f, hi = (double)hi_f;
      int nn = std::max(8, (int)std::ceil((hi - lo) / rc * NTAB) + 1);
      sp.seg_lo[j] = (float)lo;
      sp.seg_hi[j] = hi_f;
      sp.seg_h[j] = (float)((hi - lo) / (nn - 1));
      sp.seg_inv_h[j] = (float)((nn - 1) / (hi - lo));
      sp.seg_base[j] = base;
      sp.seg_n[j] = nn;
      seg_edge_lo[q][j] = edges[j];
      seg_edge_hi[q][j] = edges[j + 1];
      base += nn;
    }
    node_total += base;
    rfloats += base * 2 * sp.a;
  }
  int off = 0;
  for (int q = 0; q < ns; ++q) { P.sp[q].off = off; off += p->alpha_max[q]; }
  P.ang_n = NTAB + 1;
  P.ang_h = (float)(rcut_max / NTAB);
  P.ang_inv_h = (float)(NTAB / rcut_max);
  for (int q = 0; q < ns; ++q) P.sp[q].atab_off = q * P.ang_n * 2 * (L + 1);
  if (key == s.table_key) return ST_OK;

  int rc_;
  // radial: FP64 radial kernel on the nodes, one species at a time
  std::vector<float> rt(rfloats);
  for (int q = 0; q < ns; ++q) {
    SpeciesTab& sp = P.sp[q];
    const double rc = p->rcut_hard[q];
    int nn = 0;
    for (int j = 0; j < sp.nseg; ++j) nn += sp.seg_n[j];
    std::vector<double> r(nn);
    const double eps = 1e-9 * rc;
    for (int j = 0; j < sp.nseg; ++j) {
      const double lo = (double)sp.seg_lo[j];
      for (int k = 0; k < sp.seg_n[j]; ++k) {
        double x = lo + (double)sp.seg_h[j] * 0.0 + ((double)sp.seg_hi[j] - lo) * k / (sp.seg_n[j] - 1);
        // evaluate strictly inside the segment's true bounds: beyond the jump below, before the one above
        if (j > 0) x = std::max(x, seg_edge_lo[q][j] + eps);
        x = std::min(x, seg_edge_hi[q][j] - eps);
        r[sp.seg_base[j] + k] = x;
      }
    }
    AdjBuf dr, dsp, dfirst, drad, dder;
    std::vector<uint32_t> m(nn, 1u);
    std::vector<uint8_t> f(nn, 0);
    bool ok = ensure(dr, nn * 8, msg) == 0 && ensure(dsp, nn * 4, msg) == 0 && ensure(dfirst, nn, msg) == 0 &&
              ensure(drad, (size_t)nn * sp.a * 8, msg) == 0 && ensure(dder, (size_t)nn * sp.a * 8, msg) == 0;
    if (ok) {
      (void)hipMemcpyAsync(dr.p, r.data(), nn * 8, hipMemcpyHostToDevice, st);
      (void)hipMemcpyAsync(dsp.p, m.data(), nn * 4, hipMemcpyHostToDevice, st);
      (void)hipMemcpyAsync(dfirst.p, f.data(), nn, hipMemcpyHostToDevice, st);
      launch_fill(st, drad.as<double>(), (size_t)nn * sp.a, 0.0);
      launch_fill(st, dder.as<double>(), (size_t)nn * sp.a, 0.0);
      RadialConsts c = in.radial_tab[q];
      ok = launch_radial(st, c, nn, dr.as<double>(), dsp.as<uint32_t>(), dfirst.as<uint8_t>(), drad.as<double>(),
                         dder.as<double>());
      std::vector<double> v((size_t)nn * sp.a), d((size_t)nn * sp.a);
      (void)hipMemcpyAsync(v.data(), drad.p, v.size() * 8, hipMemcpyDeviceToHost, st);
      (void)hipMemcpyAsync(d.data(), dder.p, d.size() * 8, hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) { msg = "radial table build failed"; ok = false; }
      for (int k = 0; k < nn && ok; ++k)
        for (int n = 0; n < sp.a; ++n) {
          rt[sp.rtab_off + (size_t)k * 2 * sp.a + n] = (float)v[(size_t)k * sp.a + n];
          rt[sp.rtab_off + (size_t)k * 2 * sp.a + sp.a + n] = (float)d[(size_t)k * sp.a + n];
        }
    }
    for (AdjBuf* b : {&dr, &dsp, &dfirst, &drad, &dder})
      if (b->p) (void)hipFree(b->p);
    if (!ok) {
      if (msg.empty()) msg = "radial table build failed";
      return ST_ERR_HIP;
    }
  }
  if ((rc_ = ensure(s.rtab, rt.size() * 4, msg))) return rc_;
  (void)hipMemcpyAsync(s.rtab.p, rt.data(), rt.size() * 4, hipMemcpyHostToDevice, st);
  // angular radial factors
  const size_t af = (size_t)ns * P.ang_n * 2 * (L + 1);
  std::vector<float> at(af);
  {
    AdjBuf dout;
    if ((rc_ = ensure(dout, (size_t)P.ang_n * 2 * (L + 1) * 8, msg))) return rc_;
    std::vector<double> h((size_t)P.ang_n * 2 * (L + 1));
    for (int q = 0; q < ns; ++q) {
      ang_table_k<<<(P.ang_n + 127) / 128, 128, 0, st>>>(P.ang_n, rcut_max / NTAB, L, rcut_max * rcut_max,
                                                          p->atom_sigma_t[q], p->atom_sigma_t_scaling[q],
                                                          dout.as<double>());
      (void)hipMemcpyAsync(h.data(), dout.p, h.size() * 8, hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) {
        (void)hipFree(dout.p);
        msg = "angular table build failed";
        return ST_ERR_HIP;
      }
      for (size_t k = 0; k < h.size(); ++k) at[P.sp[q].atab_off + k] = (float)h[k];
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
  const size_t sh = site_lds(in.n_max, in.k_max, in.n_soap, in.n_soap_unc);
  if (sh > 64 * 1024) {
    msg = "the FP32 site working set exceeds the LDS";
    return ST_ERR_UNSUPPORTED;
  }
  if (in.n_max > 48) {
    msg = "the FP32 path supports n_max <= 48";
    return ST_ERR_UNSUPPORTED;
  }
  // compression element tables in FP32
  if (in.out_ptr) {
    if (s.el_key != *in.tables_key) {
      std::vector<float> oe(in.out_el_h->begin(), in.out_el_h->end()), pe(in.P_el_h->begin(), in.P_el_h->end());
      if ((rc = ensure(s.el32, (oe.size() + pe.size()) * 4, msg))) return rc;
      (void)hipMemcpyAsync(s.el32.p, oe.data(), oe.size() * 4, hipMemcpyHostToDevice, st);
      (void)hipMemcpyAsync(s.el32.as<float>() + oe.size(), pe.data(), pe.size() * 4, hipMemcpyHostToDevice, st);
      (void)hipStreamSynchronize(st);
      s.el_key = *in.tables_key;
    }
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
  const float* pel = in.out_ptr ? s.el32.as<float>() + in.out_el_h->size() : nullptr;
  bool ok = true;
  timer("adj_site_f32", [&] {
    switch (in.l_max) {
#define LS(l)                                                                                                      \
  case l:                                                                                                          \
    launch_site<l>(st, P, sh, nullptr, in.site_ptr, s.rtab.as<float>(), s.atab.as<float>(), pf, cen, in.species,   \
                   in.max_mult, in.mult, in.spmask, in.rjs, in.thetas, in.phis, s.G32.as<float>(), in.comp,        \
                   in.out_ptr, in.out_c, oel, in.unc_ptr, in.unc_e, in.P_i, pel, s.soap.as<float>(),               \
                   s.sdp.as<float>(), s.pair.as<float>());                                                         \
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
