// FP64 kernels of the get_soap hot path ("plain" path): radial and angular
// expansions per pair, the expansion coefficients cnk and the power spectrum
// per site, the full Jacobian soap_cart_der, and soap_contract.
//
// Every formula keeps the reference's operation order and this file is built
// without fused multiply-adds (the pragma below), so that the results round
// as the reference's x86-64 build does; what remains are the last-bit
// differences of the device's exp/erf/sin/cos and of parallel reductions.
//
// Device layouts are the reference's (column-major, pair- or site-major):
//   rad[p*n_max + n], ang[p*k_max + k] (double2), cnk[(i*n_max + n)*k_max + k],
//   soap[i*n_soap + s], cart[(p*n_soap + s)*3 + c], G[s*n_sites + i], pair[p*3 + c].
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernels.h"

namespace st {

// One of two or four arrays by a lane-dependent index, chosen on the bits of the addresses: a
// select among the pointers themselves is compiled to a table in private memory, read per lane.
// The result is a global-memory pointer (address space 1): a pointer formed from integers is
// otherwise generic, and its loads are flat loads, which also count against the LDS wait counter
// (an LDS wait then waits for them too).
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
__device__ inline double ld(gptr<double> p, size_t i) { return p[i]; }
__device__ inline double2 ld(gptr<double2> p, size_t i) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  const d2v v = reinterpret_cast<const __attribute__((address_space(1))) d2v*>(p)[i];
  return make_double2(v.x, v.y);
}
template <class T>
__device__ inline gptr<T> pick2(int which, const T* p0, const T* p1) {
  const unsigned long long m = which != 0 ? ~0ull : 0ull;
  return (gptr<T>)((reinterpret_cast<unsigned long long>(p0) & ~m) | (reinterpret_cast<unsigned long long>(p1) & m));
}
template <class T>
__device__ inline gptr<T> pick4(int which, const T* p0, const T* p1, const T* p2, const T* p3) {
  const unsigned long long m0 = which == 0 ? ~0ull : 0ull, m1 = which == 1 ? ~0ull : 0ull;
  const unsigned long long m2 = which == 2 ? ~0ull : 0ull, m3 = which == 3 ? ~0ull : 0ull;
  return (gptr<T>)((reinterpret_cast<unsigned long long>(p0) & m0) | (reinterpret_cast<unsigned long long>(p1) & m1) |
                   (reinterpret_cast<unsigned long long>(p2) & m2) | (reinterpret_cast<unsigned long long>(p3) & m3));
}

__device__ __forceinline__ int site_of(const int* site_ptr, int n_sites, int p) {
  int lo = 0, hi = n_sites;  // largest i with site_ptr[i] <= p
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (site_ptr[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------- set-up
// Packs mask(n_pairs, n_species) into one bit per species and marks the
// first pair of every site (the site itself, src/soap_turbo.f90:849-852).
__global__ void pack_pairs_k(int n_pairs, int n_species, const int32_t* __restrict__ mask, const int* __restrict__ site_ptr,
                             int n_sites, uint32_t* __restrict__ spmask, uint8_t* __restrict__ first) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  uint32_t b = 0;
  for (int s = 0; s < n_species; ++s)
    if (mask[(size_t)s * n_pairs + p] != 0) b |= 1u << s;
  spmask[p] = b;
  int i = site_of(site_ptr, n_sites, p);
  first[p] = site_ptr[i] == p;
}

__global__ void fill_k(double* __restrict__ x, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] = v;
}

// ---------------------------------------------------------------- radial
// One thread per pair: process_block of src/soap_turbo_radial.f90:550-1002
// for one species block, then global_scaling (src/soap_turbo.f90:510-511).
// D2: also the second derivative in r (unified form, :838-861, :895-906, :917-920, :993-999),
// which carries the recursions two steps further.
template <int A, bool D2>
__global__ __launch_bounds__(128) void radial_k(RadialConsts c, int n_pairs, const double* __restrict__ rjs,
                                                const uint32_t* __restrict__ spmask, const uint8_t* __restrict__ first,
                                                double* __restrict__ rad, double* __restrict__ rad_der,
                                                double* __restrict__ rad_der2) {
  constexpr int EX = D2 ? 4 : 2;  // recursion steps beyond the basis
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const bool is_central = first[p] != 0;
  if (is_central && !c.do_central) return;
  const double rjs_p = rjs[p];
  if (rjs_p >= c.rcut_hard_in) return;
  if (!(spmask[p] & c.species_bit)) return;

  const double pi = 3.141592653589793, sq2 = 1.4142135623730951;
  const double isq2 = 1.0 / sq2, sqpi2 = sqrt(pi / 2.0), sq2pi = sqrt(2.0 / pi), sq8pi = sqrt(8.0 / pi);
  const double sc = c.scaling, rs = c.rcut_soft;
  double T1[A + 4 + EX], T2[A + 4 + EX], pol[A + 1 + EX], DT[A + 1], D2T[D2 ? A + 1 : 1];  // T(n) at n+3
#pragma unroll
  for (int n = 0; n < A + 4 + EX; ++n) T2[n] = 0.0;
#pragma unroll
  for (int n = 0; n <= A; ++n) DT[n] = 0.0;
  if constexpr (D2) {
#pragma unroll
    for (int n = 0; n <= A; ++n) D2T[n] = 0.0;
  }

  const double rj = rjs_p * c.inv_rcut;
  const double sig = c.atom_sigma + sc * rj, s2 = sig * sig, isig = 1.0 / sig;
  const double is2 = isig * isig, is3 = is2 * isig;
  double amp, ampd, ampd2 = 0.0;
  if (c.amp_mode == 0) {
    amp = 1.0; ampd = 0.0;
  } else if (c.amp_mode == 1) {
    amp = isig; ampd = -sc * is2;
    if constexpr (D2) ampd2 = 2.0 * sc * sc * is3;
  } else if (c.amp_mode == 2) {
    double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
    double hp = 6.0 * (rj * rj) - 6.0 * rj;
    double g = isig * h;
    amp = h <= 1e-10 ? 0.0 : g;
    ampd = h <= 1e-10 ? 0.0 : hp * isig - sc * isig * g;
    if constexpr (D2) {
      double hpp = 12.0 * rj - 6.0;
      ampd2 = h <= 1e-10 ? 0.0 : hpp * isig - 2.0 * sc * hp * is2 + 2.0 * sc * sc * h * is3;
    }
  } else {
    double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
    if (h <= 1e-10) {
      amp = 0.0; ampd = 0.0;
    } else {
      double hp = 6.0 * (rj * rj) - 6.0 * rj, aa = c.amp_scaling;
      amp = isig * pow(h, aa);
      ampd = aa * pow(h, aa - 1.0) * hp * isig - sc * isig * amp;
      if constexpr (D2) {
        double hpp = 12.0 * rj - 6.0;
        ampd2 = (aa * (aa - 1.0) * pow(h, aa - 2.0) * hp * hp + aa * pow(h, aa - 1.0) * hpp) * isig -
                2.0 * sc * aa * pow(h, aa - 1.0) * hp * is2 + 2.0 * sc * sc * pow(h, aa) * is3;
      }
    }
  }
  if (c.do_central && c.central_weight != 1.0 && is_central) {
    amp = c.central_weight * amp;
    ampd = c.central_weight * ampd;
    if constexpr (D2) ampd2 = c.central_weight * ampd2;
  }
  if (c.radial_enh == 1) {
    double g = rj + sq2pi * sig, f = 1.0 + sq2pi * sc;
    if constexpr (D2) ampd2 = ampd2 * g + 2.0 * ampd * f;
    ampd = amp * f + ampd * g;
    amp = amp * g;
  } else if (c.radial_enh == 2) {
    double g = (rj * rj) + s2 + sq8pi * sig * rj;
    double f = 2.0 * rj + 2.0 * sig * sc + sq8pi * sig + sq8pi * rj * sc;
    if constexpr (D2) {
      double h = 2.0 + 2.0 * sc * sc + 2.0 * sq8pi * sc;
      ampd2 = ampd2 * g + 2.0 * ampd * f + amp * h;
    }
    ampd = amp * f + ampd * g;
    amp = amp * g;
  }
  // overlap on [0, rcut_soft]
  const double es = exp(-0.5 * ((rs - rj) * (rs - rj)) * is2);
  const double e0 = exp(-0.5 * (rj * rj) * is2);
  double C1 = s2 * es;
  const double C2 = s2 * e0;
  T1[0] = 0.0;
  T1[1] = sqpi2 * sig * (erf((rs - rj) * isq2 * isig) + erf(rj * isq2 * isig)) * c.inv_Nm2;
#pragma unroll
  for (int n = -1; n <= A + EX; ++n) {
    if (n > c.n_hi) break;
    if (n >= 0) C1 = C1 * c.dr;
    T1[n + 3] = s2 * c.ra[n + 1] * T1[n + 1] - c.rb[n + 1] * (rj - 1.0) * T1[n + 2] + (C1 - C2) * c.rc[n + 1];
  }
  const bool do_poly = c.n_poly >= 1;
  if (c.do_der && do_poly && !c.unified) {
    double f = rj - 1.0, g = sc * f * isig;
    double ca = f * is2 * (g - 1.0), cb = is2 * (2.0 * g - 1.0), cc = sc * is3;
#pragma unroll
    for (int n = 1; n <= A; ++n)
      if (n <= c.n_poly) DT[n] = ca * T1[n + 3] + c.R1[n] * cb * T1[n + 4] + c.R2[n] * cc * T1[n + 5];
  }
  double pref = 0.0;
  if (c.has_filter) {
    double den = s2 + c.q, iD = 1.0 / den, sqiD = sqrt(iD);
    double sgf = sig * c.drnf * sqiD;
    double rjf = (s2 * rs + c.q * rj) * iD;
    double sf2 = sgf * sgf;
    bool keep = (rs - rj) < 4.0 * sig;
    double isgf = 1.0 / sgf, isf2 = isgf * isgf;
    pref = keep ? exp(-0.5 * ((rs - rj) * (rs - rj)) * iD) : 0.0;
    double C2f = sf2 * exp(-0.5 * ((rs - rjf) * (rs - rjf)) * isf2);
    T2[0] = 0.0;
    T2[1] = sqpi2 * sgf * (erf((1.0 - rjf) * isq2 * isgf) - erf((rs - rjf) * isq2 * isgf)) * c.inv_Nm2;
#pragma unroll
    for (int n = -1; n <= A + EX; ++n) {
      if (n > c.n_hi) break;
      if (n >= 0) C2f = C2f * c.dr;
      T2[n + 3] = sf2 * c.ra[n + 1] * T2[n + 1] - c.rb[n + 1] * (rjf - 1.0) * T2[n + 2] - C2f * c.rc[n + 1];
      if (n == -1 && c.unified) T2[2] = T2[2] + sf2 * exp(-0.5 * ((1.0 - rjf) * (1.0 - rjf)) * isf2) * c.rc[0];
    }
    if (c.do_der && do_poly && !c.unified) {
      double f = rs - rj;
      double dpref = pref * (f * iD + f * f * iD * iD * sig * sc);
      double drjf = (2.0 * sig * rs * sc + c.q) * iD - (s2 * rs + c.q * rj) * 2.0 * sig * sc * iD * iD;
      double dsgf = sc * c.drnf * sqiD * (1.0 - s2 * iD);
      double ff = rjf - 1.0, g = dsgf * ff * isgf;
      double ca = pref * ff * isf2 * (g - drjf);
      double cb = pref * isf2 * (2.0 * g - drjf);
      double cc = pref * dsgf * isgf * isf2;
#pragma unroll
      for (int n = 1; n <= A; ++n)
        if (n <= c.n_poly)
          DT[n] = DT[n] + ca * T2[n + 3] + c.R1[n] * cb * T2[n + 4] + c.R2[n] * cc * T2[n + 5] + dpref * T2[n + 3];
    }
  }
#pragma unroll
  for (int n = 1; n <= A + EX; ++n)
    if (n <= c.n_hi) pol[n] = T1[n + 3] + pref * T2[n + 3];
  if (c.do_der && do_poly && c.unified) {
    double f = 1.0 - rj;
    double g0 = f * is2 + sc * f * f * is3;
    double g1 = -(is2 + 2.0 * sc * f * is3);
    double g2 = sc * is3;
#pragma unroll
    for (int n = 1; n <= A; ++n)
      if (n <= c.n_poly) DT[n] = g0 * pol[n] + c.R1[n] * g1 * pol[n + 1] + c.R2[n] * g2 * pol[n + 2];
    if constexpr (D2) {
      double is4 = is2 * is2, is5 = is4 * isig, is6 = is4 * is2;
      double k0 = -is2, k1 = -4.0 * sc * is3, k2 = (1.0 - 3.0 * sc * sc) * is4, k3 = 2.0 * sc * is5,
             k4 = sc * sc * is6;
      double B0 = k0 + k1 * f + k2 * f * f + k3 * (f * f * f) + k4 * (f * f * f * f);
      double B1 = -(k1 + 2.0 * k2 * f + 3.0 * k3 * f * f + 4.0 * k4 * (f * f * f));
      double B2 = k2 + 3.0 * k3 * f + 6.0 * k4 * f * f;
      double B3 = -(k3 + 4.0 * k4 * f);
      double B4 = k4;
#pragma unroll
      for (int n = 1; n <= A; ++n)
        if (n <= c.n_poly)
          D2T[n] = B0 * pol[n] + c.R1[n] * B1 * pol[n + 1] + c.R2[n] * B2 * pol[n + 2] + c.R3[n] * B3 * pol[n + 3] +
                   c.R4[n] * B4 * pol[n + 4];
    }
  }
  if (c.with_gauss) {
    double ss2 = c.sigma_sq0 + s2, iss = 1.0 / ss2, siss = sqrt(iss);
    bool gm = rj < 4.0 * (c.atom_sigma + sig);
    double ge = gm ? c.N_gauss * e0 : 0.0;
    double M0 = gm ? exp(-0.5 * (rj * rj) * iss) * sqpi2 * sig * c.atom_sigma * siss *
                         (1.0 + erf(c.atom_sigma * isig * rj * isq2 * siss)) * c.N_gauss
                   : 0.0;
    pol[A] = M0;
    if (c.do_der) {
      double sc0 = c.sigma_sq0 * s2 * iss, pp = s2 * rj * iss;
      double M1 = sc0 * ge - pp * M0;
      double M2 = sc0 * ge * (-rj) - pp * M1 + sc0 * M0;
      DT[A] = M1 * is2 + sc * M2 * is3;
      if constexpr (D2) {
        double M3 = sc0 * ge * (rj * rj) - pp * M2 + 2.0 * sc0 * M1;
        double M4 = sc0 * ge * (-(rj * rj * rj)) - pp * M3 + 3.0 * sc0 * M2;
        double is4 = is2 * is2, is5 = is4 * isig, is6 = is4 * is2;
        D2T[A] = -M0 * is2 - 4.0 * sc * M1 * is3 + (1.0 - 3.0 * sc * sc) * M2 * is4 + 2.0 * sc * M3 * is5 +
                 sc * sc * M4 * is6;
      }
    }
  }
  const double gs = c.global_scaling;
  double* out = rad + (size_t)p * c.n_max + c.off;
  if (c.do_der) {
    double dtot[A + 1];
#pragma unroll
    for (int n = 1; n <= A; ++n) dtot[n] = amp * DT[n] + ampd * pol[n];
    double* outd = rad_der + (size_t)p * c.n_max + c.off;
#pragma unroll
    for (int n = 1; n <= A; ++n) {
      double acc = c.W[n - 1] * dtot[1];
#pragma unroll
      for (int m = 2; m <= A; ++m) acc = acc + c.W[(n - 1) + (m - 1) * A] * dtot[m];
      outd[n - 1] = acc * c.inv_sqrt_rcut * gs;
    }
    if constexpr (D2) {
      double d2tot[A + 1];
#pragma unroll
      for (int n = 1; n <= A; ++n) d2tot[n] = amp * D2T[n] + 2.0 * ampd * DT[n] + ampd2 * pol[n];
      double* outh = rad_der2 + (size_t)p * c.n_max + c.off;
#pragma unroll
      for (int n = 1; n <= A; ++n) {
        double acc = c.W[n - 1] * d2tot[1];
#pragma unroll
        for (int m = 2; m <= A; ++m) acc = acc + c.W[(n - 1) + (m - 1) * A] * d2tot[m];
        outh[n - 1] = acc * c.inv_sqrt_rcut * c.inv_rcut * gs;
      }
    }
  }
#pragma unroll
  for (int n = 1; n <= A; ++n) {
    double acc = c.W[n - 1] * pol[1];
#pragma unroll
    for (int m = 2; m <= A; ++m) acc = acc + c.W[(n - 1) + (m - 1) * A] * pol[m];
    out[n - 1] = amp * acc * c.sqrt_rcut * gs;
  }
}

// ---------------------------------------------------------------- radial, poly3operator
// One thread per pair: get_radial_expansion_coefficients_poly3operator,
// src/soap_turbo_radial.f90:1549-2141, for one species block. The atom is a piecewise cubic of
// half-width w = 2 sqrt(2 ln 2) sigma(r_j) about r_j (g_aux, :2154-2217); inside rcut_soft it is
// integrated against the polynomial basis directly, between rcut_soft and rcut_hard times the
// cubic filter. Each integral is a polynomial moment between two limits: A (per species, on the
// host, get_constant_poly_coeff :2276-2297) holds the basis side, the atom and filter enter through
// their derivatives at the limits. No cutoff test (as the reference); the self pair is weighted by
// central_weight. Then W, sqrt(rcut_hard) (1/sqrt for the derivative) and global_scaling.
__device__ inline void op_g(double r, double r0, double winv, bool right, double* g) {
  const double x = (r - r0) * winv, x2 = x * x, x3 = x2 * x, w2 = winv * winv, w3 = w2 * winv;
  if (!right) {
    g[0] = 1.0 - 3.0 * x2 - 2.0 * x3; g[1] = -6.0 * (x2 + x) * winv; g[2] = -3.0 * (2.0 * x + 1.0) * w2; g[3] = -2.0 * w3;
  } else {
    g[0] = 1.0 - 3.0 * x2 + 2.0 * x3; g[1] = 6.0 * (x2 - x) * winv; g[2] = 3.0 * (2.0 * x - 1.0) * w2; g[3] = 2.0 * w3;
  }
}

__device__ inline void op_gd(double r, double r0, double winv, double ws, bool right, double* g) {
  const double x = (r - r0) * winv, x2 = x * x, x3 = x2 * x, w2 = winv * winv, w3 = w2 * winv, w4 = w3 * winv;
  if (!right) {
    g[0] = 6.0 * (x + (ws + 1.0) * x2 + ws * x3) * winv;
    g[1] = 6.0 * (1.0 + 2.0 * (ws + 1.0) * x + 3.0 * ws * x2) * w2;
    g[2] = 6.0 * (ws + 1.0 + 3.0 * ws * x) * w3;
    g[3] = 6.0 * ws * w4;
  } else {
    g[0] = 6.0 * (x + (ws - 1.0) * x2 - ws * x3) * winv;
    g[1] = 6.0 * (1.0 + 2.0 * (ws - 1.0) * x - 3.0 * ws * x2) * w2;
    g[2] = 6.0 * (ws - 1.0 - 3.0 * ws * x) * w3;
    g[3] = -6.0 * ws * w4;
  }
}

// the degree-6 product of the atom's cubic about 0 at -r_j and the filter's cubic (cf)
__device__ inline void op_filter_B(double rj, double winv, double ws, const double* cf, bool right, bool der,
                                   double* B) {
  double col[4];
  if (der) op_gd(-rj, 0.0, winv, ws, right, col);
  else op_g(-rj, 0.0, winv, right, col);
#pragma unroll
  for (int r = 0; r < 7; ++r) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc += (r - i >= 0 && r - i < 4 ? col[r - i] : 0.0) * cf[i];
    B[r] = acc;
  }
}

// v(q) = -sum_p B(p) (p-1)!/(p-q)! r^(p-q): minus the derivatives of sum_p B(p) r^(p-1) at r
__device__ inline void op_bm(const double* B, double r, double* v) {
  double rt[7];
  rt[0] = 1.0;
#pragma unroll
  for (int p = 1; p < 7; ++p) rt[p] = r * rt[p - 1];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    double acc = 0.0;
#pragma unroll
    for (int pp = q; pp < 7; ++pp) {
      double f = 1.0;
#pragma unroll
      for (int t = pp - q + 1; t <= pp; ++t) f *= (double)t;
      acc += -B[pp] * (rt[pp - q] * f);
    }
    v[q] = acc;
  }
}

template <int A>
__global__ __launch_bounds__(128) void radial_op_k(RadialConsts c, int n_pairs, const double* __restrict__ rjs,
                                                   const uint32_t* __restrict__ spmask,
                                                   const uint8_t* __restrict__ first, double* __restrict__ rad,
                                                   double* __restrict__ rad_der) {
  constexpr int NT = A + 4 > 7 ? A + 4 : 7;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  if (!(spmask[p] & c.species_bit)) return;
  const bool is_central = first[p] != 0;
  const double pi = 3.141592653589793;
  const double rjs_in = rjs[p], K = c.op_K, sc = c.scaling, rs = c.rcut_soft;
  const double win = K * (c.op_sigma_in + sc * rjs_in);
  const bool soft = rjs_in - win < c.op_rcut_soft_in;
  const bool buffer = c.op_rcut_soft_in < c.rcut_hard_in && rjs_in + win > c.op_rcut_soft_in;
  const double rj = rjs_in * c.inv_rcut;
  const double sg = c.atom_sigma + sc * rj, s2 = sg * sg, w = K * sg, winv = 1.0 / w, ws = K * sc;
  double amp, ampd;
  const double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
  if (c.amp_scaling == 0.0) {
    amp = 1.0 / sg; ampd = -sc / s2;
  } else if (c.amp_scaling == 1.0) {
    amp = 1.0 / sg * h;
    ampd = 6.0 / sg * (rj * rj - rj) - sc / sg * amp;
  } else {
    amp = 1.0 / sg * pow(h, c.amp_scaling);
    ampd = 6.0 * c.amp_scaling / sg * (rj * rj - rj) * pow(h, c.amp_scaling - 1.0) - sc / sg * amp;
  }
  if (is_central) { amp = c.central_weight * amp; ampd = c.central_weight * ampd; }
  if (c.radial_enh == 1) {
    ampd = amp * (1.0 + sqrt(2.0 / pi) * sc) + ampd * (rj + sqrt(2.0 / pi) * sg);
    amp = amp * (rj + sqrt(2.0 / pi) * sg);
  } else if (c.radial_enh == 2) {
    ampd = amp * (2.0 * rj + 2.0 * sg * sc + sqrt(8.0 / pi) * sg + sqrt(8.0 / pi) * rj * sc) +
           ampd * (rj * rj + s2 + sqrt(8.0 / pi) * sg * rj);
    amp = amp * (rj * rj + s2 + sqrt(8.0 / pi) * sg * rj);
  }
  const double vect[4] = {-1.0, -1.0, -2.0, -6.0};
  double ex[A], exd[A];
#pragma unroll
  for (int n = 0; n < A; ++n) ex[n] = exd[n] = 0.0;
  if (soft) {
    const double L[3] = {rj - w > 0.0 ? rj - w : 0.0, rj < rs ? rj : rs, rs < rj + w ? rs : rj + w};
    double I0[3][A + 4];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      I0[t][0] = 1.0;
#pragma unroll
      for (int i = 1; i < A + 4; ++i) I0[t][i] = I0[t][i - 1] * (1.0 - L[t]);
    }
    double sv[A], sd[A];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      double M[3][2][4];
#pragma unroll
      for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int pc = 0; pc < 2; ++pc) {
          double g[4];
          if (d == 0) op_g(L[t], rj, winv, pc == 1, g);
          else op_gd(L[t], rj, winv, ws, pc == 1, g);
#pragma unroll
          for (int j = 0; j < 4; ++j) M[t][pc][j] = I0[t][j] * g[j] * vect[j];
        }
#pragma unroll
      for (int n = 0; n < A; ++n) {
        double s12 = 0.0, s11 = 0.0, s23 = 0.0, s22 = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double an = c.op_A[n * 7 + j];
          s12 += M[1][0][j] * an; s11 += M[0][0][j] * an; s23 += M[2][1][j] * an; s22 += M[1][1][j] * an;
        }
        const double v = (s12 * I0[1][n + 4] - s11 * I0[0][n + 4]) + (s23 * I0[2][n + 4] - s22 * I0[1][n + 4]);
        if (d == 0) sv[n] = v;
        else sd[n] = v;
      }
    }
#pragma unroll
    for (int n = 0; n < A; ++n) {
      ex[n] = amp * sv[n];
      exd[n] = amp * sd[n] + ampd * sv[n];
    }
  }
  if (buffer) {
    const double L[3] = {rs > rj - w ? rs : rj - w, rj > rs ? rj : rs, 1.0 < rj + w ? 1.0 : rj + w};
    double I0[3][NT];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      I0[t][0] = 1.0;
#pragma unroll
      for (int i = 1; i < NT; ++i) I0[t][i] = I0[t][i - 1] * (1.0 - L[t]);
    }
    double bb[A], bd[A];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      double M[3][2][7];
#pragma unroll
      for (int pc = 0; pc < 2; ++pc) {
        double B[7];
        op_filter_B(rj, winv, ws, c.op_cf, pc == 1, d == 1, B);
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          double v[7];
          op_bm(B, L[t], v);
#pragma unroll
          for (int q = 0; q < 7; ++q) M[t][pc][q] = v[q] * I0[t][q];
        }
      }
#pragma unroll
      for (int n = 0; n < A; ++n) {
        double s12 = 0.0, s11 = 0.0, s23 = 0.0, s22 = 0.0;
#pragma unroll
        for (int q = 0; q < 7; ++q) {
          const double an = c.op_A[n * 7 + q];
          s12 += M[1][0][q] * an; s11 += M[0][0][q] * an; s23 += M[2][1][q] * an; s22 += M[1][1][q] * an;
        }
        const double v = (s12 * I0[1][n + 4] - s11 * I0[0][n + 4]) + (s23 * I0[2][n + 4] - s22 * I0[1][n + 4]);
        if (d == 0) bb[n] = v;
        else bd[n] = v;
      }
    }
#pragma unroll
    for (int n = 0; n < A; ++n) {
      ex[n] += amp * bb[n];
      exd[n] += amp * bd[n] + ampd * bb[n];
    }
  }
  const double gs = c.global_scaling;
  double* out = rad + (size_t)p * c.n_max + c.off;
#pragma unroll
  for (int n = 0; n < A; ++n) {
    double av = 0.0;
#pragma unroll
    for (int m = 0; m < A; ++m) av += c.W[n + m * A] * ex[m];
    out[n] = av * c.sqrt_rcut * gs;
  }
  if (c.do_der) {
    double* outd = rad_der + (size_t)p * c.n_max + c.off;
#pragma unroll
    for (int n = 0; n < A; ++n) {
      double ad = 0.0;
#pragma unroll
      for (int m = 0; m < A; ++m) ad += c.W[n + m * A] * exd[m];
      outd[n] = ad / c.sqrt_rcut * gs;
    }
  }
}

// ---------------------------------------------------------------- angular
// One thread per pair: get_angular_expansion_coefficients,
// src/soap_turbo_angular.f90:548-605, with get_eimphi_conjg (:267-329),
// get_ilexp (:361-422), get_ilexp_der (:436-463), get_plm_array (:108-160)
// and get_plm_array_der (:181-248).
template <int L>
//
// One wavefront per 64 pairs, a thread per pair. The rows of the block's pairs (KMAX values each,
// contiguous in the pair-major output) are staged in LDS one output array at a time and written out
// as one contiguous region, so that every store instruction covers consecutive addresses.
__global__ __launch_bounds__(64) void angular_k(AngularConsts c, int n_pairs, const double* __restrict__ rjs,
                                                const double* __restrict__ thetas, const double* __restrict__ phis,
                                                const uint32_t* __restrict__ spmask, const double* __restrict__ preflm,
                                                double2* __restrict__ ang, double2* __restrict__ ang_r,
                                                double2* __restrict__ ang_a, double2* __restrict__ ang_p) {
  constexpr int KMAX = 1 + L * (L + 1) / 2 + L;
  constexpr int KMAX1 = 1 + (L + 1) * (L + 2) / 2 + (L + 1);
  extern __shared__ double2 stage[];  // 64 * KMAX
  const int lane = threadIdx.x, base = blockIdx.x * 64;
  const int p = base + lane, nb = min(64, n_pairs - base);
  // pairs past the list or outside the cutoff are computed on stand-in geometry and written as zeros
  double rj = 1.0, th = 0.0, phi = 0.0;
  uint32_t msk = 1u;
  bool live = false;
  if (p < n_pairs) {
    const double r0 = rjs[p];
    const uint32_t m0 = spmask[p];
    if (r0 < c.rcut_max && m0 != 0) {
      live = true;
      rj = r0;
      msk = m0;
      th = thetas[p];
      phi = phis[p];
    }
  }
  // the staged rows out to the block's region of `out`
  auto flush = [&](double2* __restrict__ out) {
    __syncthreads();
    double2* o = out + (size_t)base * KMAX;
    for (int e = lane; e < nb * KMAX; e += 64) o[e] = stage[e];
    __syncthreads();
  };
  double2* row = stage + lane * KMAX;
  const int i_sp = __ffs(msk) - 1;
  const double x = cos(th);
  const double sigma = c.sigma_t[i_sp] + c.sigma_t_scaling[i_sp] * rj;
  const double amplitude = c.rcut2 / (sigma * sigma);
  // ilexp(rj / sigma)
  double il[L + 2];
  {
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
    double fact = 1.0;
    il[0] = xx < xcut ? 1.0 - x2 : flm2;
    if (L >= 1) {
      fact = fact * 3.0;
      il[1] = x2 / 1000.0 < xcut ? (x2 - x4) / fact : flm1;
    }
#pragma unroll
    for (int l = 2; l <= L; ++l) {
      fact = fact * (2.0 * (double)l + 1.0);
      double ratio = x2pow / fact, fl;
      if (ratio * (double)l < xcut)
        fl = ratio;
      else
        fl = fabs(flm2 - (2.0 * (double)l - 1.0) * inv_x2 * flm1);
      x2pow = x2pow * x2;
      flm2 = flm1;
      flm1 = fl;
      il[l] = fl;
    }
  }
  // e^{-i m phi} by the Chebyshev recursion
  double pm_re[L + 1], pm_im[L + 1];
  {
    double cosm2 = cos(phi), cosphi2 = 2.0 * cosm2, sinm2 = -sin(phi), cosm1 = 1.0, sinm1 = 0.0;
    pm_re[0] = 1.0; pm_im[0] = 0.0;
#pragma unroll
    for (int l = 1; l <= L; ++l) {
      double cos0 = cosphi2 * cosm1 - cosm2, sin0 = cosphi2 * sinm1 - sinm2;
      cosm2 = cosm1; sinm2 = sinm1; cosm1 = cos0; sinm1 = sin0;
      pm_re[l] = cos0; pm_im[l] = -sin0;
    }
  }
  // P_l^m(x) to l = L (+1 with derivatives)
  double plm[KMAX1];
  {
    const int LL = c.do_der ? L + 1 : L;
    double x2 = x * x, sq = sqrt(1.0 - x2);
    plm[0] = 1.0;
    if (LL >= 1) { plm[1] = x; plm[2] = -sq; }
    if (L >= 1 && LL >= 2) { plm[3] = 1.5 * x2 - 0.5; plm[4] = -3.0 * x * sq; plm[5] = 3.0 - 3.0 * x2; }
#pragma unroll
    for (int l = 3; l <= L + 1; ++l) {
      if (l > LL) break;
#pragma unroll
      for (int m = 0; m <= l - 2; ++m) {
        int k = 1 + l * (l + 1) / 2 + m;
        double a = (double)(2 * l - 1) / (double)(l - m), b = (double)(l - 1 + m) / (double)(l - m);
        plm[k - 1] = a * x * plm[k - l - 1] - b * plm[k - 2 * l];
      }
      int k = 1 + l * (l + 1) / 2 + l - 1;
      plm[k - 1] = x * (double)(2 * l - 1) * plm[k - l - 1];
      k = k + 1;
      plm[k - 1] = -((double)(2 * l - 1) * sq * plm[k - l - 2]);
    }
  }
  if (c.do_der) {
    const double scal = c.sigma_t_scaling[i_sp];
    double ild[L + 1];
    {
      double coeff1 = 2.0 * rj / (sigma * sigma);
      double coeff2 = 1.0 - scal * rj / sigma;
#pragma unroll
      for (int l = 0; l <= L; ++l) ild[l] = 0.0;
      if (!(rj < 1e-5)) {
        // il[L+1] is read only for L = 0, where the reference reads past its array
        ild[0] = coeff1 * ((L >= 1 ? il[1] : 0.0) - il[0]);
#pragma unroll
        for (int l = 1; l <= L; ++l) ild[l] = (-coeff1 - (double)(2 * l + 2) / rj) * il[l] + coeff1 * il[l - 1];
#pragma unroll
        for (int l = 0; l <= L; ++l) ild[l] = ild[l] * coeff2;
      }
    }
    const double sig_fac = 2.0 * scal / sigma;
    // the four arrays in turn (each entry formed as in one pass over (l, m))
#pragma unroll
    for (int arr = 0; arr < 4; ++arr) {
#pragma unroll
      for (int l = 0; l <= L; ++l) {
#pragma unroll
        for (int m = 0; m <= l; ++m) {
          const int kk = l * (l + 1) / 2 + m;
          const int k = kk + 1;
          const double e_re = il[l] * pm_re[m], e_im = il[l] * pm_im[m];
          const double u = amplitude * preflm[kk], v = u * plm[kk];
          double2 out;
          if (arr == 0) {
            out = make_double2(v * e_re, v * e_im);
          } else if (arr == 1) {
            const double er_re = ild[l] * pm_re[m], er_im = ild[l] * pm_im[m];
            const double ec_re = v * e_re, ec_im = v * e_im;
            out = make_double2(v * er_re - sig_fac * ec_re, v * er_im - sig_fac * ec_im);
          } else if (arr == 2) {
            double dsin = 0.0;
            if (m >= 1) {
              const int kp = 1 + (l + 1) * (l + 2) / 2 + m + 1, km = 1 + (l + 1) * (l + 2) / 2 + m - 1;
              dsin = (0.5 * (double)(l - m + 1) * (double)(l - m + 2)) * plm[km - 1] + 0.5 * plm[kp - 1];
            }
            const double t = u * dsin;
            out = make_double2(-(t * e_im), t * e_re);
          } else {
            // sin(t) dP/dcos(t) and -m P/sin(t)
            double part1, part2;
            if (m == 0) part1 = l == 0 ? 0.0 : -0.5 * plm[l * (l + 1) / 2 + 1];
            else part1 = (0.5 * (double)(l + m) * (double)(l - m + 1)) * plm[k - 2];
            part2 = m == l ? 0.0 : -0.5 * plm[k];
            const double dms = part1 + part2;
            const double t2 = u * dms;
            out = make_double2(t2 * e_re, t2 * e_im);
          }
          row[kk] = live ? out : make_double2(0.0, 0.0);
        }
      }
      flush(arr == 0 ? ang : arr == 1 ? ang_r : arr == 2 ? ang_a : ang_p);
    }
  } else {
#pragma unroll
    for (int l = 0; l <= L; ++l) {
#pragma unroll
      for (int m = 0; m <= l; ++m) {
        const int kk = l * (l + 1) / 2 + m;
        const double e_re = il[l] * pm_re[m], e_im = il[l] * pm_im[m];
        const double t = amplitude * preflm[kk] * plm[kk];
        row[kk] = live ? make_double2(t * e_re, t * e_im) : make_double2(0.0, 0.0);
      }
    }
    flush(ang);
  }
}

// ---------------------------------------------------------------- cnk
// One workgroup per site: cnk(k,n,i) = sum_j 4 pi R(n,j) A(k,j), then the
// poly3gauss central-atom term (src/soap_turbo.f90:626-664); central[s*n_max + n]
// holds that term per species (zero where the species has none). The sum over
// neighbours is a product on the FP64 matrix cores: rows n (tiles of 16),
// columns the real and imaginary parts of A_k (tiles of 16), K the neighbours
// four at a time; each wavefront takes every fourth tile.
typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void cnk_k(int n_sites, int n_max, int k_max, const int* __restrict__ site_ptr,
                                             const double* __restrict__ rad, const double2* __restrict__ ang,
                                             const int32_t* __restrict__ site_species, int max_mult,
                                             const int32_t* __restrict__ species_mult,
                                             const double* __restrict__ central, double2* __restrict__ cnk) {
  const int i = blockIdx.x;
  const int p0 = site_ptr[i], p1 = site_ptr[i + 1];
  const double fourpi = 4.0 * 3.141592653589793;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
  const int nt = (n_max + 15) >> 4, ct = (2 * k_max + 15) >> 4;
  const double* angd = reinterpret_cast<const double*>(ang);
  double* cnkd = reinterpret_cast<double*>(cnk);
  for (int t = w; t < nt * ct; t += 4) {
    const int a = t / ct, b = t - a * ct;
    const int n = 16 * a + li, col = 16 * b + li;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int q0 = p0; q0 < p1; q0 += 16) {
      double av[4], bv[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int p = q0 + 4 * s + g;
        const bool live = p < p1;
        av[s] = (live && n < n_max) ? fourpi * rad[(size_t)p * n_max + n] : 0.0;
        bv[s] = (live && col < 2 * k_max) ? angd[(size_t)p * 2 * k_max + col] : 0.0;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc, 0, 0, 0);
    }
    // D(row, col) sits in lane col + 16 (row % 4), element row / 4
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int nn = 16 * a + g + 4 * v;
      if (nn < n_max && col < 2 * k_max) {
        double val = acc[v];
        if (col == 0 && central)
          for (int m = 0; m < species_mult[i]; ++m) {
            const int sp = site_species[(size_t)i * max_mult + m] - 1;
            val = val + central[(size_t)sp * n_max + nn];
          }
        cnkd[((size_t)i * n_max + nn) * 2 * k_max + col] = val;
      }
    }
  }
}

// ---------------------------------------------------------------- power spectrum
// Component table of the uncompressed spectrum: comp[c] = (n, np, l, 0) for
// c in the reference's (n, np >= n, l) order. The compressed output s is
// sum over e in [out_ptr[s], out_ptr[s+1]) of P_el[e] * value(P_j[e]) in the
// reference's order; without compression out_ptr is null and s = c.
__device__ __forceinline__ double spectrum_value(const double2* cn, int k_max, int4 cmp) {
  const int n = cmp.x, np = cmp.y, l = cmp.z;
  const int kb = l * (l + 1) / 2;
  const double base = n != np ? 1.4142135623730951 : 1.0;
  double acc = 0.0;
  for (int m = 0; m <= l; ++m) {
    const double mult = m > 0 ? base * 2.0 : base;
    const double2 u = cn[n * k_max + kb + m], v = cn[np * k_max + kb + m];
    acc = acc + mult * (u.x * v.x + u.y * v.y);
  }
  return acc;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int j = 0; j < nw; ++j) s += red[j];
  return s;
}

// Unnormalised soap and sqrt_dot_p (src/soap_turbo.f90:750-813).
__global__ __launch_bounds__(256) void power_k(int n_sites, int n_max, int k_max, int n_soap,
                                               const int4* __restrict__ comp, const int* __restrict__ out_ptr,
                                               const int* __restrict__ out_c, const double* __restrict__ out_el,
                                               const int4* __restrict__ ocomp, const double* __restrict__ oel,
                                               const double2* __restrict__ cnk, double* __restrict__ soap,
                                               double* __restrict__ sqrt_dot_p) {
  extern __shared__ double2 smem2[];
  __shared__ double red[16];
  const int i = blockIdx.x;
  double2* cn = smem2;
  for (int t = threadIdx.x; t < n_max * k_max; t += blockDim.x) cn[t] = cnk[(size_t)i * n_max * k_max + t];
  __syncthreads();
  double ss = 0.0;
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x) {
    double v;
    if (ocomp) {
      // one entry per output: 0 + P_el value, as the reference's accumulation gives
      v = out_ptr ? 0.0 + oel[s] * spectrum_value(cn, k_max, ocomp[s]) : spectrum_value(cn, k_max, ocomp[s]);
    } else if (out_ptr) {
      v = 0.0;
      for (int e = out_ptr[s]; e < out_ptr[s + 1]; ++e) v = v + out_el[e] * spectrum_value(cn, k_max, comp[out_c[e]]);
    } else {
      v = spectrum_value(cn, k_max, comp[s]);
    }
    soap[(size_t)i * n_soap + s] = v;
    ss += v * v;
  }
  double tot = block_sum(ss, red);
  if (threadIdx.x == 0) {
    double d = sqrt(tot);
    if (d < 1e-5) d = 1.0;
    sqrt_dot_p[i] = d;
  }
}

__global__ void normalise_k(int n_sites, int n_soap, const double* __restrict__ sqrt_dot_p, double* __restrict__ soap) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n_sites * n_soap) return;
  soap[t] = soap[t] / sqrt_dot_p[t / n_soap];
}

// ---------------------------------------------------------------- full Jacobian
// One workgroup per site, neighbours in batches of PB: the four contractions
// over m (soap_derivatives_from_expansion, src/soap_turbo.f90:1996-2025), the
// assembly of the surviving components (:2027-2066), then normalisation and
// rotation to Cartesian with the self column accumulated as minus the sum of
// the others in neighbour order (:837-895). soap holds the unnormalised
// spectrum here. Each thread owns the outputs s = tid + u * NT for the whole
// site (256 or 512 threads). A batch computes each output's three components once, keeps them in
// registers across the block reduction of q . d, then normalises, rotates and
// stores them: three barriers per batch.
//
// LDS layout: the batch's P, Q, Qa, Qo tables are [n][l][pair][kind] and the
// radial values [n][pair][value], so that one output's reads for every pair
// and kind share one address register and differ by immediate offsets; rows
// are padded by one double to spread adjacent (n, l) over the banks. The
// angular rows are [k][pair][kind] (double2), padded likewise.

// tuning: pairs per batch and waves per SIMD for up to 3 outputs per thread,
// and the largest n_soap run with 256-thread blocks
#ifndef ST_DER_PB3
#define ST_DER_PB3 4
#endif
#ifndef ST_DER_WAVES3
#define ST_DER_WAVES3 2
#endif
#ifndef ST_DER_NT256_MAX
#define ST_DER_NT256_MAX 1024
#endif

template <int NSPT>
struct DerShape {
  // component registers per thread: PB * NSPT * 3 doubles, kept near 36
  static constexpr int PB = NSPT <= 2 ? 4 : NSPT <= 3 ? ST_DER_PB3 : NSPT <= 4 ? 3 : NSPT <= 6 ? 2 : 1;
  // waves per SIMD the register allocation is held to, for NT threads per block
  static constexpr int waves(int nt) { return nt == 256 && NSPT <= 1 ? 3 : NSPT == 3 ? ST_DER_WAVES3 : 2; }
  static constexpr int PQS = 4 * PB + 1;   // doubles per (n, l) row
  static constexpr int RS = 2 * PB + 1;    // doubles per n row of radial values
  static constexpr int AS = 4 * PB + 1;    // double2 per k row of angular values
  static constexpr int NRED = 3 * PB <= 4 ? 4 : 3 * PB <= 8 ? 8 : 16;
};

__device__ __forceinline__ int pack_component(const int4 c) { return c.x | (c.y << 10) | (c.z << 20); }

// Sum each of V values over the 64 lanes of a wave by recursive halving:
// every exchange moves half of the still-unreduced values, so the wave needs
// V/2 + V/4 + ... shuffles instead of 6 V. Lane j ends with the total of value
// idx(j) = bit5(j) V/2 + bit4(j) V/4 + ...; lanes with the low bits below that
// hold the same totals.
//
// The exchanges across 32 and 16 lanes use gfx950's lane swaps, which move no data through the LDS:
// permlane32_swap(a, b) exchanges lanes 32-63 of a with lanes 0-31 of b, permlane16_swap the odd
// 16-lane rows of a with the even rows of b. With a = x[j], b = x[j + H], a lane with the bit clear
// then holds its own x[j] in a and its partner's in b, a lane with the bit set its partner's x[j + H]
// in a and its own in b: a + b is the exchange's sum in both (the same sum as own + partner).
template <int BIT>
__device__ __forceinline__ void swap_halves(double& a, double& b) {
  const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  const unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32), blo = (unsigned)ub, bhi = (unsigned)(ub >> 32);
  unsigned lo0, lo1, hi0, hi1;
  if constexpr (BIT == 32) {
    const auto l = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    lo0 = l[0]; lo1 = l[1]; hi0 = h[0]; hi1 = h[1];
  } else {
    static_assert(BIT == 16, "BIT");
    const auto l = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    lo0 = l[0]; lo1 = l[1]; hi0 = h[0]; hi1 = h[1];
  }
  a = __longlong_as_double((long long)(((unsigned long long)hi0 << 32) | lo0));
  b = __longlong_as_double((long long)(((unsigned long long)hi1 << 32) | lo1));
}

// Within a 16-lane row the partner is reached by a DPP move, which also leaves the LDS alone: the row
// mirror (lane i <- 15 - i) and half-row mirror (i <- 7 - i) flip bit 3 and bit 2 with the bits below
// them, the quad permutations flip bit 1 or bit 0 alone. Each pairs every lane with one lane of the
// other half, so the sums cover every lane once as with exclusive-or partners.
template <int BIT>
__device__ __forceinline__ double dpp_partner(double v) {
  constexpr int CTRL = BIT == 8 ? 0x140 : BIT == 4 ? 0x141 : BIT == 2 ? 0x4e : 0xb1;
  static_assert(BIT == 8 || BIT == 4 || BIT == 2 || BIT == 1, "BIT");
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <int H, int BIT, int V>
__device__ __forceinline__ void halving_stage(double (&x)[V], int lane) {
  if constexpr (BIT >= 16) {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      double a = x[j], b = x[j + H];
      swap_halves<BIT>(a, b);
      x[j] = a + b;
    }
  } else {
    const bool hi = (lane & BIT) != 0;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const double keep = hi ? x[j + H] : x[j];
      const double send = hi ? x[j] : x[j + H];
      x[j] = keep + dpp_partner<BIT>(send);
    }
  }
  if constexpr (H > 1) halving_stage<H / 2, BIT / 2, V>(x, lane);
}

template <int V>
__device__ __forceinline__ double wave_multi_sum(double (&x)[V], int lane) {
  static_assert(V == 4 || V == 8 || V == 16, "V");
  halving_stage<V / 2, 32, V>(x, lane);
  double t = x[0];
  constexpr int LOG = V == 4 ? 2 : V == 8 ? 3 : 4;
  if constexpr (LOG <= 2) t += dpp_partner<8>(t);
  if constexpr (LOG <= 3) t += dpp_partner<4>(t);
  t += dpp_partner<2>(t);
  t += dpp_partner<1>(t);
  return t;
}

template <int V>
__device__ __forceinline__ int wave_multi_index(int lane) {
  constexpr int LOG = V == 4 ? 2 : V == 8 ? 3 : 4;
  int idx = 0;
#pragma unroll
  for (int k = 0; k < LOG; ++k) idx |= ((lane >> (5 - k)) & 1) << (LOG - 1 - k);
  return idx;
}

// One output's prefactor (4 pi, times sqrt 2 off the n = n' diagonal) and its three sums for pair b of a
// batch, without it: d/dr, d/dtheta, d/dphi parts from the P/Q rows at (n', l), (n, l) and the radial rows
// at n, n' (src/soap_turbo.f90:2026-2045).
__device__ __forceinline__ double der_pre(const int cm) {
  const int n = cm & 1023, np = (cm >> 10) & 1023;
  return (4.0 * 3.141592653589793) * (n == np ? 1.0 : 1.4142135623730951);
}

template <int PB>
__device__ __forceinline__ void der_sums(const int cm, int b, const double* __restrict__ PQ,
                                         const double* __restrict__ R, int L1, double& tr, double& ta, double& to) {
#pragma clang fp contract(fast)
  constexpr int PQS = 4 * PB + 1, RS = 2 * PB + 1;
  const int n = cm & 1023, np = (cm >> 10) & 1023, l = cm >> 20;
  const double* A = PQ + (np * L1 + l) * PQS + 4 * b;
  const double* C = PQ + (n * L1 + l) * PQS + 4 * b;
  const double* Rn = R + n * RS + 2 * b;
  const double* Rp = R + np * RS + 2 * b;
  tr = Rn[1] * A[0] + Rn[0] * A[1] + Rp[1] * C[0] + Rp[0] * C[1];
  ta = Rn[0] * A[2] + Rp[0] * C[2];
  to = Rn[0] * A[3] + Rp[0] * C[3];
}

template <int PB>
__device__ __forceinline__ void der_component(const int cm, int b, const double* __restrict__ PQ,
                                              const double* __restrict__ R, int L1, double& tr, double& ta,
                                              double& to) {
  constexpr int PQS = 4 * PB + 1, RS = 2 * PB + 1;
  const int n = cm & 1023, np = (cm >> 10) & 1023, l = cm >> 20;
  const double* A = PQ + (np * L1 + l) * PQS + 4 * b;   // P, Q, Qa, Qo at (np, l)
  const double* C = PQ + (n * L1 + l) * PQS + 4 * b;    // ... at (n, l)
  const double* Rn = R + n * RS + 2 * b;                // r, r' at n
  const double* Rp = R + np * RS + 2 * b;               // ... at np
  const double pre = (4.0 * 3.141592653589793) * (n == np ? 1.0 : 1.4142135623730951);
  tr = pre * (Rn[1] * A[0] + Rn[0] * A[1] + Rp[1] * C[0] + Rp[0] * C[1]);
  ta = pre * (Rn[0] * A[2] + Rp[0] * C[2]);
  to = pre * (Rn[0] * A[3] + Rp[0] * C[3]);
}

// Rotation of the spherical derivatives (d/dr, d/dtheta / r, d/dphi / (r sin theta))
// to Cartesian components for every pair (src/soap_turbo.f90:864-876):
// rot = {sin t cos f, sin t sin f, cos t, -cos t cos f / r, -cos t sin f / r,
// sin t / r, -sin f / r, cos f / r}.
__global__ void rotation_k(int n_pairs, const double* __restrict__ rjs, const double* __restrict__ thetas,
                           const double* __restrict__ phis, double* __restrict__ rotp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const double sin_t = sin(thetas[p]), cos_t = cos(thetas[p]);
  const double sin_f = sin(phis[p]), cos_f = cos(phis[p]);
  const double inv_r = 1.0 / rjs[p];
  double* rr = rotp + (size_t)p * 8;
  rr[0] = sin_t * cos_f; rr[1] = sin_t * sin_f; rr[2] = cos_t;
  rr[3] = -(cos_t * cos_f * inv_r); rr[4] = -(cos_t * sin_f * inv_r); rr[5] = sin_t * inv_r;
  rr[6] = -(sin_f * inv_r); rr[7] = cos_f * inv_r;
}

// DIRECT: every output has exactly one (component, P_el) entry (ocomp/oel);
// otherwise outputs sum their CSR entries (out_ptr/out_c/out_el), or, with
// no compression (out_ptr null), output s is component s.
//
// The batch inputs (angular rows, radial values, rotation) are double
// buffered in LDS: the loads of batch k + 1 are issued to registers before the
// P/Q products of batch k and written to the other buffer during its pass 1,
// so that their latency is hidden behind a batch of work. The FK/FN registers
// per thread cover k_max <= 48 (l_max <= 8) and n_max <= 32; larger shapes
// load the remainder when the buffer is written.
template <int NSPT, int NT, bool DIRECT>
__global__ __launch_bounds__(NT, DerShape<NSPT>::waves(NT)) void derivatives_k(
    int n_sites, int n_max, int l_max, int k_max, int n_soap, const int* __restrict__ site_ptr,
    const int4* __restrict__ comp, const int* __restrict__ out_ptr, const int* __restrict__ out_c,
    const double* __restrict__ out_el, const int4* __restrict__ ocomp, const double* __restrict__ oel,
    const double2* __restrict__ cnk, const double* __restrict__ rad,
    const double* __restrict__ rad_der, const double2* __restrict__ ang, const double2* __restrict__ ang_r,
    const double2* __restrict__ ang_a, const double2* __restrict__ ang_p, const double* __restrict__ soap,
    const double* __restrict__ sqrt_dot_p, const double* __restrict__ rotp, double* __restrict__ cart) {
  // fused multiply-adds here: the Jacobian's short sums change only in their last bits (the file's other
  // kernels keep the reference's separate roundings)
#pragma clang fp contract(fast)
  using Sh = DerShape<NSPT>;
  constexpr int PB = Sh::PB, PQS = Sh::PQS, RS = Sh::RS, AS = Sh::AS, NRED = Sh::NRED;
  extern __shared__ double2 smem2[];
  constexpr int NW = NT / 64;
  __shared__ double red[NW][NRED];
  __shared__ double rot[2][PB][8];
  const int i = blockIdx.x;
  const int L1 = l_max + 1, nl = n_max * L1;
  double2* cn = smem2;                                   // n_max * k_max
  double2* sAbuf = cn + n_max * k_max;                   // 2 x k_max * AS
  double* sRbuf = (double*)(sAbuf + 2 * k_max * AS);     // 2 x n_max * RS
  double* PQ = sRbuf + 2 * n_max * RS;                   // nl * PQS
  const int p0 = site_ptr[i], p1 = site_ptr[i + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;

  // batch inputs: global -> registers (issue), registers -> LDS buffer (complete). A batch's rows
  // of each input array are one contiguous range (pairs pb .. pb + nb - 1), so thread t loads
  // element t (+ NT, ...) of each array from a workgroup-uniform base: the address is the base plus
  // the thread's fixed offset. Its LDS position (row entry k or n, pair b) is fixed as well.
  constexpr int FK = (PB * 48 + NT - 1) / NT, FN = (PB * 32 + NT - 1) / NT;
  double2 fa[4][FK];
  double fr[2][FN], fo = 0.0;
  const gptr<double2> angs[4] = {(gptr<double2>)ang, (gptr<double2>)ang_r, (gptr<double2>)ang_a,
                                 (gptr<double2>)ang_p};
  const gptr<double> rads[2] = {(gptr<double>)rad, (gptr<double>)rad_der};
  auto fetch = [&](int pb) {
    const int na = min(PB, p1 - pb) * k_max, nr = min(PB, p1 - pb) * n_max;
    // the thread's offset is made opaque here, so that base + offset is formed in the loop (as a
    // uniform base and a 32-bit offset) and not kept across it per lane
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int kind = 0; kind < 4; ++kind) {
      const gptr<double2> src = angs[kind] + (size_t)pb * k_max;
#pragma unroll
      for (int it = 0; it < FK; ++it) {
        const int t = tid + it * NT;
        fa[kind][it] = t < na ? ld(src, t) : make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const gptr<double> src = rads[d] + (size_t)pb * n_max;
#pragma unroll
      for (int it = 0; it < FN; ++it) {
        const int t = tid + it * NT;
        fr[d][it] = t < nr ? ld(src, t) : 0.0;
      }
    }
    fo = 0.0;
    if (tid < PB * 8 && tid < 8 * min(PB, p1 - pb)) fo = ld((gptr<double>)rotp + (size_t)pb * 8, tid);
  };
  auto commit = [&](int pb, int buf) {
    double2* sA = sAbuf + buf * k_max * AS;
    double* sR = sRbuf + buf * n_max * RS;
#pragma unroll
    for (int it = 0; it < FK; ++it) {
      const int t = threadIdx.x + it * NT;
      if (t < PB * k_max) {
        const int b = t / k_max, k = t - b * k_max;
#pragma unroll
        for (int kind = 0; kind < 4; ++kind) sA[k * AS + 4 * b + kind] = fa[kind][it];
      }
    }
    const int na = min(PB, p1 - pb) * k_max, nr = min(PB, p1 - pb) * n_max;
    for (int t = threadIdx.x + FK * NT; t < PB * k_max; t += NT) {
      const int b = t / k_max, k = t - b * k_max;
#pragma unroll
      for (int kind = 0; kind < 4; ++kind)
        sA[k * AS + 4 * b + kind] = t < na ? ld(angs[kind] + (size_t)pb * k_max, t) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int it = 0; it < FN; ++it) {
      const int t = threadIdx.x + it * NT;
      if (t < PB * n_max) {
        const int b = t / n_max, n = t - b * n_max;
        sR[n * RS + 2 * b] = fr[0][it];
        sR[n * RS + 2 * b + 1] = fr[1][it];
      }
    }
    for (int t = threadIdx.x + FN * NT; t < PB * n_max; t += NT) {
      const int b = t / n_max, n = t - b * n_max;
#pragma unroll
      for (int d = 0; d < 2; ++d) sR[n * RS + 2 * b + d] = t < nr ? ld(rads[d] + (size_t)pb * n_max, t) : 0.0;
    }
    if (threadIdx.x < PB * 8) rot[buf][threadIdx.x / 8][threadIdx.x % 8] = fo;
  };

  // the site's c_nk with the m-weight of the P/Q sums (1 for m = 0, else 2: exact) applied once
  for (int t = threadIdx.x; t < n_max * k_max; t += NT) {
    const int k = t % k_max;
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= k) ++l;
    const double2 c = cnk[(size_t)i * n_max * k_max + t];
    cn[t] = k == l * (l + 1) / 2 ? c : make_double2(2.0 * c.x, 2.0 * c.y);
  }
  const double sdp = sqrt_dot_p[i];
  const double inv_p = 1.0 / sdp, inv_p2 = 1.0 / (sdp * sdp);
  const bool scaled = out_ptr != nullptr;
  double selfacc[NSPT][3];
  double qv[NSPT];
  int cmu[NSPT];
  double elu[NSPT];
#pragma unroll
  for (int u = 0; u < NSPT; ++u) {
    selfacc[u][0] = selfacc[u][1] = selfacc[u][2] = 0.0;
    const int s = threadIdx.x + u * NT;
    const bool live = s < n_soap;
    qv[u] = live ? soap[(size_t)i * n_soap + s] : 0.0;
    if (DIRECT) {
      cmu[u] = live ? pack_component(ocomp[s]) : 0;
      elu[u] = live ? der_pre(cmu[u]) * (scaled ? oel[s] : 1.0) * inv_p : 0.0;
    } else {
      cmu[u] = (live && !scaled) ? pack_component(comp[s]) : 0;
      elu[u] = live && !scaled ? der_pre(cmu[u]) * inv_p : 0.0;
    }
  }
  if (p0 + 1 < p1) {
    fetch(p0 + 1);
    commit(p0 + 1, 0);
  }
  int buf = 0;
  for (int pb = p0 + 1; pb < p1; pb += PB, buf ^= 1) {
    const int nb = min(PB, p1 - pb);
    const bool more = pb + PB < p1;
    const double2* sA = sAbuf + buf * k_max * AS;
    const double* sR = sRbuf + buf * n_max * RS;
    __syncthreads();
    if (more) fetch(pb + PB);
    // P, Q, Qa, Qo(n, l) for every pair of the batch: sum over m of w(m) Re(conj(A) c), as one
    // block-diagonal product per l on the FP64 matrix cores: rows n (tiles of 16), columns
    // j = 4 pair + kind (16), K = the 2 (l + 1) real terms (re, im of each m), padded to 4.
    {
      const int ntile = (n_max + 15) >> 4, npair = (ntile + 1) >> 1, njob = npair * L1;
      const int r16 = lane & 15, g = lane >> 4;
      // A job is one l and two row tiles (rows 32 tp .. 32 tp + 31): the tiles share the sA operand
      // and their two product chains are independent, so they interleave. Jobs in order of falling
      // l, dealt to the waves in a snake, so that their K steps even out.
      for (int q = 0; q < njob; ++q) {
        const int round = q / NW, pos = q - round * NW;
        if ((round & 1 ? NW - 1 - pos : pos) != w) continue;
        const int l = l_max - q / npair, tp = q - (q / npair) * npair;
        const int kb = l * (l + 1) / 2, nK = 2 * (l + 1), nsteps = (nK + 3) >> 2;
        const int n0 = 32 * tp + r16, n1 = n0 + 16;
        // The job's products with the step count NS fixed at compile time: every operand load (one
        // double: the real or imaginary part of the entry) is issued before any is used, so the job
        // waits for the LDS once. Rows past n_max read row n_max - 1 and columns past the batch's
        // pairs read column 0 (their results are not kept).
        const int nc0 = n0 < n_max ? n0 : n_max - 1, nc1 = n1 < n_max ? n1 : n_max - 1;
        const int rc = r16 < 4 * PB ? r16 : 0;
        // K entry kappa = 4 st + g is part (kappa & 1) of m = kappa / 2: the operands of step st sit a
        // fixed distance past step 0's (two entries of cn's row, two rows of sA)
        const double* ca0 = reinterpret_cast<const double*>(cn + nc0 * k_max + kb + (g >> 1)) + (g & 1);
        const double* ca1 = reinterpret_cast<const double*>(cn + nc1 * k_max + kb + (g >> 1)) + (g & 1);
        const double* cb = reinterpret_cast<const double*>(sA + (kb + (g >> 1)) * AS + rc) + (g & 1);
        f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
        auto product = [&](auto ns_) {
          constexpr int NS = decltype(ns_)::value;
          double av0[NS], av1[NS], bv[NS];
#pragma unroll
          for (int st = 0; st < NS; ++st) {
            av0[st] = ca0[4 * st];
            av1[st] = ca1[4 * st];
            bv[st] = cb[4 * AS * st];
          }
          // only the last step can reach past the 2 (l + 1) entries (m up to l + 1: inside the LDS)
          if (4 * (NS - 1) + g >= nK) bv[NS - 1] = av0[NS - 1] = av1[NS - 1] = 0.0;
          // every load is issued before the first product: left to itself the scheduler alternates
          // load, wait, product, and the job waits for the LDS once per step
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int st = 0; st < NS; ++st) {
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av0[st], bv[st], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av1[st], bv[st], acc1, 0, 0, 0);
          }
        };
        switch (nsteps) {  // l_max <= 12: at most 7 steps
          case 1: product(std::integral_constant<int, 1>{}); break;
          case 2: product(std::integral_constant<int, 2>{}); break;
          case 3: product(std::integral_constant<int, 3>{}); break;
          case 4: product(std::integral_constant<int, 4>{}); break;
          case 5: product(std::integral_constant<int, 5>{}); break;
          case 6: product(std::integral_constant<int, 6>{}); break;
          default: product(std::integral_constant<int, 7>{}); break;
        }
        // D(row, col) sits in lane col + 16 (row % 4), element row / 4
        if (r16 < 4 * PB) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int nn = 32 * tp + g + 4 * v;
            if (nn < n_max) PQ[(nn * L1 + l) * PQS + r16] = acc0[v];
            if (nn + 16 < n_max) PQ[((nn + 16) * L1 + l) * PQS + r16] = acc1[v];
          }
        }
      }
    }
    __syncthreads();
    // pass 1: the components of every pair of the batch, and q . d
    double v[PB][NSPT][3];
    double dots[NRED];
#pragma unroll
    for (int j = 0; j < NRED; ++j) dots[j] = 0.0;
#pragma unroll
    for (int u = 0; u < NSPT; ++u) {
      const int s = threadIdx.x + u * NT;
#pragma unroll
      for (int b = 0; b < PB; ++b) {
        double vr = 0.0, va = 0.0, vo = 0.0;
        if (DIRECT || !scaled) {
          // on every lane, with no branch, so that the LDS reads of all outputs and pairs can be issued
          // together: an output past n_soap has component (0, 0, 0) and factor 0
          double tr, ta, to;
          der_sums<PB>(cmu[u], b, PQ, sR, L1, tr, ta, to);
          vr = elu[u] * tr; va = elu[u] * ta; vo = elu[u] * to;
        } else if (s < n_soap) {
          {
            for (int e = out_ptr[s]; e < out_ptr[s + 1]; ++e) {
              double tr, ta, to;
              der_component<PB>(pack_component(comp[out_c[e]]), b, PQ, sR, L1, tr, ta, to);
              const double el = out_el[e];
              vr = vr + el * tr; va = va + el * ta; vo = vo + el * to;
            }
            vr *= inv_p; va *= inv_p; vo *= inv_p;
          }
        }
        v[b][u][0] = vr; v[b][u][1] = va; v[b][u][2] = vo;
        dots[3 * b] += qv[u] * vr; dots[3 * b + 1] += qv[u] * va; dots[3 * b + 2] += qv[u] * vo;
      }
    }
    {
      const double t = wave_multi_sum<NRED>(dots, lane);
      if ((lane & (64 / NRED - 1)) == 0) red[w][wave_multi_index<NRED>(lane)] = t;
    }
    // the next batch's inputs into the other buffer (last read by the batch before, which every
    // thread finished at this batch's first barrier). Here, before this batch's stores: vector memory
    // operations complete in order, so a wait for these loads placed after the stores would also wait
    // for the stores to be written. The explicit wait (vmcnt(0)) is on every path, so that no load is
    // pending past it in the compiler's view: it would otherwise wait again at the loop top, for the
    // stores too, before the fetch registers are reused
    __builtin_amdgcn_s_waitcnt(0x0f70);
    if (more) commit(pb + PB, buf ^ 1);
    __syncthreads();
    // pass 2: normalise, rotate, store, accumulate the self column
#pragma unroll
    for (int b = 0; b < PB; ++b) {
      if (b < nb) {
        double dot_r = red[0][3 * b], dot_a = red[0][3 * b + 1], dot_o = red[0][3 * b + 2];
#pragma unroll
        for (int q = 1; q < NW; ++q) {
          dot_r += red[q][3 * b]; dot_a += red[q][3 * b + 1]; dot_o += red[q][3 * b + 2];
        }
        const double* rr = rot[buf][b];
        const size_t p = pb + b;
#pragma unroll
        for (int u = 0; u < NSPT; ++u) {
          const int s = threadIdx.x + u * NT;
          if (s < n_soap) {
            const double sv = qv[u] * inv_p2;
            const double vr = v[b][u][0] - sv * dot_r;
            const double va = v[b][u][1] - sv * dot_a;
            const double vo = v[b][u][2] - sv * dot_o;
            const double c0 = rr[0] * vr + rr[3] * vo + rr[6] * va;
            const double c1 = rr[1] * vr + rr[4] * vo + rr[7] * va;
            const double c2 = rr[2] * vr + rr[5] * vo;
            // soap_cart_der is written once and read by the caller's contraction only: stores that do
            // not keep its lines in L2 over the batches' angular and radial rows
            double* C = cart + (p * n_soap + s) * 3;
            __builtin_nontemporal_store(c0, C);
            __builtin_nontemporal_store(c1, C + 1);
            __builtin_nontemporal_store(c2, C + 2);
            selfacc[u][0] = selfacc[u][0] - c0;
            selfacc[u][1] = selfacc[u][1] - c1;
            selfacc[u][2] = selfacc[u][2] - c2;
          }
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NSPT; ++u) {
    const int s = threadIdx.x + u * NT;
    if (s < n_soap) {
      double* C = cart + ((size_t)p0 * n_soap + s) * 3;
      C[0] = selfacc[u][0]; C[1] = selfacc[u][1]; C[2] = selfacc[u][2];
    }
  }
}

// pair(:,k) = sum_s G(i,s) soap_cart_der(:,s,k): what a caller of the
// full-Jacobian get_soap does to turn dE/dq into pair gradients. One
// workgroup per site holds G(i,:) in LDS; one wavefront per pair.
__global__ __launch_bounds__(256) void contract_cart_k(int n_sites, int n_pairs, int n_soap, int ldg,
                                                       const int* __restrict__ site_ptr,
                                                       const double* __restrict__ G, const double* __restrict__ cart,
                                                       double* __restrict__ pair) {
  extern __shared__ double gs[];
  const int i = blockIdx.x;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x) gs[s] = G[(size_t)s * ldg + i];
  __syncthreads();
  const int p0 = site_ptr[i], p1 = site_ptr[i + 1];
  const int len = 3 * n_soap;  // the pair's row of soap_cart_der, read as one contiguous stream
  // lane l reads elements e = l + 64 j; as 64 = 1 (mod 3), element e is component (l + j) % 3 of
  // output e / 3, so with j taken three at a time acc[u] collects component (l + u) % 3
  int s_u[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) s_u[u] = (lane + 64 * u) / 3;
  for (int p = p0 + w; p < p1; p += 4) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double* C = cart + (size_t)p * len;
    // 30 loads per lane in flight before they are used (a 621-output row is one chunk)
    for (int e0 = lane, sb = 0; e0 < len; e0 += 1920, sb += 640) {
      double x[30];
#pragma unroll
      for (int j = 0; j < 30; ++j) x[j] = e0 + 64 * j < len ? C[e0 + 64 * j] : 0.0;
#pragma unroll
      for (int j = 0; j < 30; ++j)
        if (e0 + 64 * j < len) acc[j % 3] += gs[sb + 64 * (j / 3) + s_u[j % 3]] * x[j];
    }
    // back to component order: component c is acc[(c - l) mod 3]
    const int r = lane % 3;
    const double c0 = r == 0 ? acc[0] : r == 1 ? acc[2] : acc[1];
    const double c1 = r == 0 ? acc[1] : r == 1 ? acc[0] : acc[2];
    const double c2 = r == 0 ? acc[2] : r == 1 ? acc[1] : acc[0];
    acc[0] = c0; acc[1] = c1; acc[2] = c2;
    const double t = wave_multi_sum<4>(acc, lane);
    const int q = wave_multi_index<4>(lane);
    if ((lane & 15) == 0 && q < 3) pair[3 * (size_t)p + q] = t;
  }
}

// ---------------------------------------------------------------- soap_contract
// One workgroup per site (src/soap_turbo.f90:1780-1877): back through the
// normalisation and compression, H(k,n) = w(m) sum_np Y(n,np,l) cnk(k,np),
// then one thread per neighbour pair against H. unc_ptr/unc_e list, per
// uncompressed component, the P entries that feed it, in P order.
__global__ __launch_bounds__(256) void contract_k(
    int n_sites, int n_max, int l_max, int k_max, int n_soap, int n_soap_unc, const int* __restrict__ site_ptr,
    const uint8_t* __restrict__ skip, const int* __restrict__ unc_ptr, const int* __restrict__ unc_e,
    const int* __restrict__ umap, const double* __restrict__ uel,
    const int32_t* __restrict__ P_i, const double* __restrict__ P_el, const double2* __restrict__ cnk,
    const double* __restrict__ rad, const double* __restrict__ rad_der, const double2* __restrict__ ang,
    const double2* __restrict__ ang_r, const double2* __restrict__ ang_a, const double2* __restrict__ ang_p,
    const double* __restrict__ soap, const double* __restrict__ G, double sgn, const double* __restrict__ sqrt_dot_p,
    const double* __restrict__ rjs, const double* __restrict__ thetas, const double* __restrict__ phis,
    double* __restrict__ pair) {
  extern __shared__ double2 smem2[];
  __shared__ double red[16];
  const int i = blockIdx.x;
  double2* cn = smem2;                       // n_max * k_max
  double2* H = cn + n_max * k_max;           // n_max * k_max
  double* Gp = (double*)(H + n_max * k_max); // n_soap
  double* Gu = Gp + n_soap;                  // n_soap_unc
  const int p0 = site_ptr[i], p1 = site_ptr[i + 1];
  for (int t = threadIdx.x; t < n_max * k_max; t += blockDim.x) cn[t] = cnk[(size_t)i * n_max * k_max + t];
  double qg_part = 0.0;
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x) qg_part += soap[(size_t)i * n_soap + s] * G[(size_t)s * n_sites + i];
  const double qg = block_sum(qg_part, red);
  const double inv = 1.0 / sqrt_dot_p[i];
  for (int s = threadIdx.x; s < n_soap; s += blockDim.x)
    Gp[s] = (G[(size_t)s * n_sites + i] - soap[(size_t)i * n_soap + s] * qg) * inv;
  __syncthreads();
  for (int c = threadIdx.x; c < n_soap_unc; c += blockDim.x) {
    double v;
    if (umap) {
      const int q = umap[c];
      v = unc_ptr ? (q >= 0 ? 0.0 + uel[c] * Gp[q] : 0.0) : Gp[c];
    } else if (unc_ptr) {
      v = 0.0;
      for (int e = unc_ptr[c]; e < unc_ptr[c + 1]; ++e) v = v + P_el[unc_e[e]] * Gp[P_i[unc_e[e]] - 1];
    } else {
      v = Gp[c];
    }
    Gu[c] = v;
  }
  __syncthreads();
  const int L1 = l_max + 1;
  for (int t = threadIdx.x; t < n_max * k_max; t += blockDim.x) {
    const int n = t / k_max, k = t - n * k_max;
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= k) ++l;
    const int m = k - l * (l + 1) / 2;
    double hr = 0.0, hi = 0.0;
    for (int np = 0; np < n_max; ++np) {
      const int a = n < np ? n : np, b = n < np ? np : n;
      const int c = (a * n_max - a * (a - 1) / 2 + (b - a)) * L1 + l;
      double x = 0.0;
      if (!skip[c]) x = Gu[c] * (a == b ? 1.0 : 1.4142135623730951);
      const double y = a == b ? 2.0 * x : x;
      const double2 cv = cn[np * k_max + k];
      hr = hr + y * cv.x;
      hi = hi + y * cv.y;
    }
    const double w = m == 0 ? 1.0 : 2.0;
    H[n * k_max + k] = make_double2(w * hr, w * hi);
  }
  __syncthreads();
  const double fourpi = 4.0 * 3.141592653589793;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int p = p0 + 1 + threadIdx.x; p < p1; p += blockDim.x) {
    const double2* A = ang + (size_t)p * k_max;
    const double2* Ar = ang_r + (size_t)p * k_max;
    const double2* Aa = ang_a + (size_t)p * k_max;
    const double2* Ap = ang_p + (size_t)p * k_max;
    double f_rad = 0.0, f_azi = 0.0, f_pol = 0.0;
    for (int n = 0; n < n_max; ++n) {
      double sd = 0.0, sr = 0.0, sa = 0.0, so = 0.0;
      for (int k = 0; k < k_max; ++k) {
        const double2 h = H[n * k_max + k];
        const double2 a = A[k], ar = Ar[k], aa = Aa[k], ap = Ap[k];
        sd = sd + a.x * h.x + a.y * h.y;
        sr = sr + ar.x * h.x + ar.y * h.y;
        sa = sa + aa.x * h.x + aa.y * h.y;
        so = so + ap.x * h.x + ap.y * h.y;
      }
      f_rad = f_rad + rad_der[(size_t)p * n_max + n] * sd + rad[(size_t)p * n_max + n] * sr;
      f_azi = f_azi + rad[(size_t)p * n_max + n] * sa;
      f_pol = f_pol + rad[(size_t)p * n_max + n] * so;
    }
    f_rad = fourpi * f_rad; f_azi = fourpi * f_azi; f_pol = fourpi * f_pol;
    const double sin_t = sin(thetas[p]), cos_t = cos(thetas[p]);
    const double sin_f = sin(phis[p]), cos_f = cos(phis[p]);
    const double inv_r = 1.0 / rjs[p];
    const double x0 = sgn * (sin_t * cos_f * f_rad - cos_t * cos_f * inv_r * f_pol - sin_f * inv_r * f_azi);
    const double x1 = sgn * (sin_t * sin_f * f_rad - cos_t * sin_f * inv_r * f_pol + cos_f * inv_r * f_azi);
    const double x2 = sgn * (cos_t * f_rad + sin_t * inv_r * f_pol);
    pair[3 * (size_t)p] = x0; pair[3 * (size_t)p + 1] = x1; pair[3 * (size_t)p + 2] = x2;
    s0 -= x0; s1 -= x1; s2 -= x2;
  }
  s0 = block_sum(s0, red);
  s1 = block_sum(s1, red + 4);
  s2 = block_sum(s2, red + 8);
  if (threadIdx.x == 0) {
    pair[3 * (size_t)p0] = s0; pair[3 * (size_t)p0 + 1] = s1; pair[3 * (size_t)p0 + 2] = s2;
  }
}

// ---------------------------------------------------------------- launchers
void launch_pack_pairs(hipStream_t st, int n_pairs, int n_species, const int32_t* mask, const int* site_ptr,
                       int n_sites, uint32_t* spmask, uint8_t* first) {
  if (n_pairs == 0) return;
  pack_pairs_k<<<(n_pairs + 255) / 256, 256, 0, st>>>(n_pairs, n_species, mask, site_ptr, n_sites, spmask, first);
}

void launch_fill(hipStream_t st, double* x, size_t n, double v) {
  if (n == 0) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  fill_k<<<(unsigned)blocks, 256, 0, st>>>(x, n, v);
}

template <int A>
static void radial_one(hipStream_t st, const RadialConsts& c, int n_pairs, const double* rjs, const uint32_t* spmask,
                       const uint8_t* first, double* rad, double* rad_der, double* rad_der2) {
  if (c.op)
    radial_op_k<A><<<(n_pairs + 127) / 128, 128, 0, st>>>(c, n_pairs, rjs, spmask, first, rad, rad_der);
  else if (rad_der2)
    radial_k<A, true><<<(n_pairs + 127) / 128, 128, 0, st>>>(c, n_pairs, rjs, spmask, first, rad, rad_der, rad_der2);
  else
    radial_k<A, false><<<(n_pairs + 127) / 128, 128, 0, st>>>(c, n_pairs, rjs, spmask, first, rad, rad_der, nullptr);
}

bool launch_radial(hipStream_t st, const RadialConsts& c, int n_pairs, const double* rjs, const uint32_t* spmask,
                   const uint8_t* first, double* rad, double* rad_der, double* rad_der2) {
  if (n_pairs == 0) return true;
  switch (c.alpha_max) {
#define RC(a) case a: radial_one<a>(st, c, n_pairs, rjs, spmask, first, rad, rad_der, rad_der2); return true;
    RC(1) RC(2) RC(3) RC(4) RC(5) RC(6) RC(7) RC(8) RC(9) RC(10) RC(11)
#undef RC
    default: return false;
  }
}

template <int L>
static void angular_one(hipStream_t st, const AngularConsts& c, int n_pairs, const double* rjs, const double* th,
                        const double* ph, const uint32_t* spmask, const double* preflm, double2* a, double2* ar,
                        double2* aa, double2* ap) {
  constexpr int KMAX = 1 + L * (L + 1) / 2 + L;
  constexpr size_t sh = 64 * KMAX * sizeof(double2);
  if (sh > 65536) (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&angular_k<L>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
  angular_k<L><<<(n_pairs + 63) / 64, 64, sh, st>>>(c, n_pairs, rjs, th, ph, spmask, preflm, a, ar, aa, ap);
}

bool launch_angular(hipStream_t st, const AngularConsts& c, int n_pairs, const double* rjs, const double* th,
                    const double* ph, const uint32_t* spmask, const double* preflm, double2* a, double2* ar,
                    double2* aa, double2* ap) {
  if (n_pairs == 0) return true;
  switch (c.l_max) {
#define AC(l) case l: angular_one<l>(st, c, n_pairs, rjs, th, ph, spmask, preflm, a, ar, aa, ap); return true;
    AC(0) AC(1) AC(2) AC(3) AC(4) AC(5) AC(6) AC(7) AC(8) AC(9) AC(10) AC(11) AC(12)
#undef AC
    default: return false;
  }
}

void launch_cnk(hipStream_t st, int n_sites, int n_max, int k_max, const int* site_ptr, const double* rad,
                const double2* ang, const int32_t* site_species, int max_mult, const int32_t* species_mult,
                const double* central, double2* cnk) {
  if (n_sites == 0) return;
  cnk_k<<<n_sites, 256, 0, st>>>(n_sites, n_max, k_max, site_ptr, rad, ang, site_species, max_mult, species_mult,
                                 central, cnk);
}

void launch_power(hipStream_t st, int n_sites, int n_max, int k_max, int n_soap, const int4* comp, const int* out_ptr,
                  const int* out_c, const double* out_el, const int4* ocomp, const double* oel, const double2* cnk,
                  double* soap, double* sqrt_dot_p) {
  if (n_sites == 0) return;
  size_t sh = sizeof(double2) * n_max * k_max;
  power_k<<<n_sites, 256, sh, st>>>(n_sites, n_max, k_max, n_soap, comp, out_ptr, out_c, out_el, ocomp, oel, cnk, soap,
                                    sqrt_dot_p);
}

void launch_normalise(hipStream_t st, int n_sites, int n_soap, const double* sqrt_dot_p, double* soap) {
  size_t n = (size_t)n_sites * n_soap;
  if (n == 0) return;
  normalise_k<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n_sites, n_soap, sqrt_dot_p, soap);
}

// The full Jacobian with one wavefront per pair (soap_derivatives_from_expansion,
// src/soap_turbo.f90:1927-2075, and the normalisation and rotation of :837-895): a workgroup of
// four wavefronts per site holds the site's cnk in LDS; each wavefront takes every fourth pair,
// stages its angular rows, radial values and rotation in its own LDS slice, forms the P, Q, Qa,
// Qo(n, l) table of the pair (sum over m of w(m) Re(conj(A) c)), then its lanes compute the
// pair's n_soap outputs (NSW per lane), reduce q . d across the wavefront with no workgroup
// barrier, and store the normalised, rotated column. The site's own column (minus the sum of the
// others) is combined from the four wavefronts' partials at the end.
template <int NSW, bool DIRECT>
__global__ __launch_bounds__(256, 2) void jacobian_wave_k(
    int n_sites, int n_max, int l_max, int k_max, int n_soap, const int* __restrict__ site_ptr,
    const int4* __restrict__ comp, const int* __restrict__ out_ptr, const int* __restrict__ out_c,
    const double* __restrict__ out_el, const int4* __restrict__ ocomp, const double* __restrict__ oel,
    const double2* __restrict__ cnk, const double* __restrict__ rad, const double* __restrict__ rad_der,
    const double2* __restrict__ ang, const double2* __restrict__ ang_r, const double2* __restrict__ ang_a,
    const double2* __restrict__ ang_p, const double* __restrict__ soap, const double* __restrict__ sqrt_dot_p,
    const double* __restrict__ rotp, double* __restrict__ cart) {
  extern __shared__ double2 smw[];
  const int i = blockIdx.x;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L1 = l_max + 1, nl = n_max * L1;
  double2* cn = smw;                                                  // n_max * k_max (site)
  double2* sA = cn + n_max * k_max + w * 4 * k_max;                   // [kind][k] of this wave's pair
  double* dbase = reinterpret_cast<double*>(cn + n_max * k_max + 4 * 4 * k_max);
  double* sR = dbase + w * 2 * n_max;                                 // r[n], r'[n]
  double* PQ = dbase + 4 * 2 * n_max + w * 4 * nl;                    // [(n L1 + l) 4 + kind]
  double* srot = dbase + 4 * 2 * n_max + 4 * 4 * nl + w * 8;          // rotation of the pair
  double* sq = dbase + 4 * 2 * n_max + 4 * 4 * nl + 32;               // per output: q (unnormalised)
  double* sel = sq + n_soap;                                          //   its compression element
  int* scm = reinterpret_cast<int*>(sel + n_soap);                    //   its packed component
  double* selfsum = reinterpret_cast<double*>(smw);                   // 3 n_soap, over all at the end
  const int p0 = site_ptr[i], p1 = site_ptr[i + 1];
  for (int t = threadIdx.x; t < n_max * k_max; t += 256) cn[t] = cnk[(size_t)i * n_max * k_max + t];
  const double sdp = sqrt_dot_p[i];
  const double inv_p = 1.0 / sdp, inv_p3 = 1.0 / (sdp * sdp * sdp);
  const bool scaled = out_ptr != nullptr;
  for (int s = threadIdx.x; s < n_soap; s += 256) {
    sq[s] = soap[(size_t)i * n_soap + s];
    if (DIRECT) {
      scm[s] = pack_component(ocomp[s]);
      sel[s] = oel[s];
    } else {
      scm[s] = scaled ? 0 : pack_component(comp[s]);
      sel[s] = 0.0;
    }
  }
  double self[NSW][3];
#pragma unroll
  for (int j = 0; j < NSW; ++j) self[j][0] = self[j][1] = self[j][2] = 0.0;
  __syncthreads();
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // one output's (d_rad, d_azi, d_pol) from the pair's table, as der_component
  auto component = [&](int cm, double& tr, double& ta, double& to) {
    const int n = cm & 1023, np = (cm >> 10) & 1023, l = cm >> 20;
    const double* A = PQ + (np * L1 + l) * 4;
    const double* C = PQ + (n * L1 + l) * 4;
    const double pre = (4.0 * 3.141592653589793) * (n == np ? 1.0 : 1.4142135623730951);
    tr = pre * (sR[n_max + n] * A[0] + sR[n] * A[1] + sR[n_max + np] * C[0] + sR[np] * C[1]);
    ta = pre * (sR[n] * A[2] + sR[np] * C[2]);
    to = pre * (sR[n] * A[3] + sR[np] * C[3]);
  };
  for (int p = p0 + 1 + w; p < p1; p += 4) {
    for (int t = lane; t < 4 * k_max; t += 64) {
      const int kind = t / k_max, k = t - kind * k_max;
      const auto src = pick4(kind, ang, ang_r, ang_a, ang_p);
      sA[t] = ld(src, (size_t)p * k_max + k);
    }
    for (int t = lane; t < 2 * n_max; t += 64)
      sR[t] = ld(pick2(t < n_max ? 0 : 1, rad, rad_der), (size_t)p * n_max + t % n_max);
    if (lane < 8) srot[lane] = rotp[(size_t)p * 8 + lane];
    wave_sync();
    // P, Q, Qa, Qo(n, l) = sum over m <= l of w(m) Re(conj(A_kind(l, m)) c(n, l, m))
    for (int e = lane; e < nl; e += 64) {
      const int n = e / L1, l = e - n * L1, kb = l * (l + 1) / 2;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (int m = 0; m <= l; ++m) {
        const double2 c = cn[n * k_max + kb + m];
        const double wm = m == 0 ? 1.0 : 2.0;
#pragma unroll
        for (int kind = 0; kind < 4; ++kind) {
          const double2 a = sA[kind * k_max + kb + m];
          acc[kind] = acc[kind] + wm * (a.x * c.x + a.y * c.y);
        }
      }
#pragma unroll
      for (int kind = 0; kind < 4; ++kind) PQ[e * 4 + kind] = acc[kind];
    }
    wave_sync();
    // output j's (d_rad, d_azi, d_pol) for the pair, through its compression entries
    auto output = [&](int j, double& vr, double& va, double& vo) {
      const int s = lane + 64 * j;
      vr = va = vo = 0.0;
      if (s >= n_soap) return;
      if (DIRECT || !scaled) {
        double tr, ta, to;
        component(scm[s], tr, ta, to);
        if (DIRECT && scaled) {
          const double el = sel[s];
          vr = 0.0 + el * tr; va = 0.0 + el * ta; vo = 0.0 + el * to;
        } else {
          vr = tr; va = ta; vo = to;
        }
      } else {
        for (int e = out_ptr[s]; e < out_ptr[s + 1]; ++e) {
          double tr, ta, to;
          component(pack_component(comp[out_c[e]]), tr, ta, to);
          const double el = out_el[e];
          vr = vr + el * tr; va = va + el * ta; vo = vo + el * to;
        }
      }
    };
    // q . d (the components are formed again below rather than held: registers for occupancy)
    double dots[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NSW; ++j) {
      double vr, va, vo;
      output(j, vr, va, vo);
      const double q = lane + 64 * j < n_soap ? sq[lane + 64 * j] : 0.0;
      dots[0] += q * vr; dots[1] += q * va; dots[2] += q * vo;
      if (j & 1) __builtin_amdgcn_sched_barrier(0);  // two outputs' loads in flight at a time
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      dots[0] += __shfl_xor(dots[0], o, 64);
      dots[1] += __shfl_xor(dots[1], o, 64);
      dots[2] += __shfl_xor(dots[2], o, 64);
    }
    const double rr[8] = {srot[0], srot[1], srot[2], srot[3], srot[4], srot[5], srot[6], srot[7]};
#pragma unroll
    for (int j = 0; j < NSW; ++j) {
      const int s = lane + 64 * j;
      if (s < n_soap) {
        double dr, da, dd;
        output(j, dr, da, dd);
        const double sv = sq[s];
        const double vr = dr * inv_p - sv * inv_p3 * dots[0];
        const double va = da * inv_p - sv * inv_p3 * dots[1];
        const double vo = dd * inv_p - sv * inv_p3 * dots[2];
        const double c0 = rr[0] * vr + rr[3] * vo + rr[6] * va;
        const double c1 = rr[1] * vr + rr[4] * vo + rr[7] * va;
        const double c2 = rr[2] * vr + rr[5] * vo;
        double* Cc = cart + ((size_t)p * n_soap + s) * 3;
        Cc[0] = c0; Cc[1] = c1; Cc[2] = c2;
        self[j][0] = self[j][0] - c0;
        self[j][1] = self[j][1] - c1;
        self[j][2] = self[j][2] - c2;
      }
      if (j & 1) __builtin_amdgcn_sched_barrier(0);
    }
    wave_sync();  // the slice is restaged for the next pair
  }
  // the site's own column: the wavefronts' partial sums, added in wavefront order
  __syncthreads();
  for (int q = 0; q < 4; ++q) {
    if (w == q) {
#pragma unroll
      for (int j = 0; j < NSW; ++j) {
        const int s = lane + 64 * j;
        if (s < n_soap)
#pragma unroll
          for (int c = 0; c < 3; ++c) selfsum[3 * s + c] = q == 0 ? self[j][c] : selfsum[3 * s + c] + self[j][c];
      }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < 3 * n_soap; t += 256) cart[(size_t)p0 * n_soap * 3 + t] = selfsum[t];
}

// LDS of jacobian_wave_k: the site's cnk, four wavefront slices (angular rows, radial values,
// the P/Q table, the rotation), the outputs' q, elements and components, and at the end the site
// column's 3 n_soap partial sums over all of it
static size_t jacobian_wave_lds(int n_max, int l_max, int k_max, int n_soap) {
  const size_t nl = (size_t)n_max * (l_max + 1);
  const size_t main = sizeof(double2) * ((size_t)n_max * k_max + 16 * (size_t)k_max) +
                      sizeof(double) * (8 * (size_t)n_max + 16 * nl + 32) + 20 * (size_t)n_soap;
  return std::max(main, sizeof(double) * 3 * (size_t)n_soap);
}

// outputs per lane of the compiled jacobian_wave_k shapes (0: none covers n_soap)
static int jacobian_wave_nsw(int n_soap) { return n_soap <= 384 ? 6 : n_soap <= 640 ? 10 : n_soap <= 832 ? 13 : 0; }

// (outputs per thread, threads per block) shapes of the Jacobian kernel for n_soap outputs: the
// default first (256 threads up to 1024 outputs, 512 threads up to 3072), then the other thread
// count when a compiled shape covers n_soap with it.
static int der_shapes(int n_soap, int (*shape)[2]) {
  int n = 0;
  const int first = n_soap <= ST_DER_NT256_MAX ? 256 : 512;
  for (int nt : {first, 768 - first}) {
    int nspt = (n_soap + nt - 1) / nt;
    if (nspt == 5) nspt = 6;
    const bool compiled = nt == 256 ? nspt <= 4 : (nspt >= 2 && nspt <= 6);
    if (compiled) {
      shape[n][0] = nspt;
      shape[n][1] = nt;
      ++n;
    }
  }
  return n;
}

static void der_config(int n_soap, int variant, int& nspt, int& nt) {
  int sh[2][2] = {{1, 256}, {1, 256}};
  const int n = der_shapes(n_soap, sh);
  const int v = variant >= 0 && variant < n ? variant : 0;
  nspt = sh[v][0];
  nt = sh[v][1];
}

static int der_pb(int n_soap, int variant) {
  int nspt, nt;
  der_config(n_soap, variant, nspt, nt);
  return nspt == 1 ? DerShape<1>::PB : nspt == 2 ? DerShape<2>::PB : nspt == 3 ? DerShape<3>::PB
       : nspt == 4 ? DerShape<4>::PB : DerShape<6>::PB;
}

int derivatives_variants(int n_soap) {
  int sh[2][2];
  return n_soap > 3072 ? 0 : der_shapes(n_soap, sh) + (jacobian_wave_nsw(n_soap) > 0);
}

// variants: the batched kernel's shapes (der_shapes; the first is the default), then the
// wave-per-pair kernel when a compiled shape covers n_soap
static bool is_wave_variant(int n_soap, int variant) {
  int sh[2][2];
  return variant > 0 && variant == der_shapes(n_soap, sh) && jacobian_wave_nsw(n_soap) > 0;
}
static int batched_variant(int, int variant) { return variant < 0 ? 0 : variant; }

size_t derivatives_lds(int n_max, int l_max, int k_max, int n_soap, int variant) {
  if (is_wave_variant(n_soap, variant)) return jacobian_wave_lds(n_max, l_max, k_max, n_soap);
  const int pb = der_pb(n_soap, batched_variant(n_soap, variant));
  // (one more angular row: the P/Q product's last step may read one row past a buffer, into this pad)
  return sizeof(double2) * (n_max * k_max + 2 * k_max * (4 * pb + 1)) +
         sizeof(double) * (2 * n_max * (2 * pb + 1) + n_max * (l_max + 1) * (4 * pb + 1)) +
         sizeof(double2) * (4 * pb + 1);
}

bool launch_derivatives(hipStream_t st, int n_sites, int n_pairs, int n_max, int l_max, int k_max, int n_soap,
                        const int* site_ptr, const int4* comp, const int* out_ptr, const int* out_c,
                        const double* out_el, const int4* ocomp, const double* oel,
                        const double2* cnk, const double* rad, const double* rad_der, const double2* ang,
                        const double2* ang_r, const double2* ang_a, const double2* ang_p, const double* soap,
                        const double* sqrt_dot_p, const double* rjs, const double* thetas, const double* phis,
                        double* rotp, double* cart, int variant) {
  if (n_sites == 0) return true;
  if (n_soap > 3072) return false;
  if (n_pairs > 0) rotation_k<<<(n_pairs + 255) / 256, 256, 0, st>>>(n_pairs, rjs, thetas, phis, rotp);
  if (is_wave_variant(n_soap, variant)) {
    const size_t shw = jacobian_wave_lds(n_max, l_max, k_max, n_soap);
    const int nsw = jacobian_wave_nsw(n_soap);
#define JW(N)                                                                                                     \
  if (nsw == N) {                                                                                                 \
    const void* f1 = reinterpret_cast<const void*>(&jacobian_wave_k<N, true>);                                    \
    const void* f0 = reinterpret_cast<const void*>(&jacobian_wave_k<N, false>);                                   \
    if (shw > 64 * 1024) {                                                                                        \
      (void)hipFuncSetAttribute(f1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shw);                       \
      (void)hipFuncSetAttribute(f0, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shw);                       \
    }                                                                                                             \
    if (ocomp)                                                                                                    \
      jacobian_wave_k<N, true><<<n_sites, 256, shw, st>>>(n_sites, n_max, l_max, k_max, n_soap, site_ptr, comp,   \
                                                          out_ptr, out_c, out_el, ocomp, oel, cnk, rad, rad_der,  \
                                                          ang, ang_r, ang_a, ang_p, soap, sqrt_dot_p, rotp, cart); \
    else                                                                                                          \
      jacobian_wave_k<N, false><<<n_sites, 256, shw, st>>>(n_sites, n_max, l_max, k_max, n_soap, site_ptr, comp,  \
                                                           out_ptr, out_c, out_el, ocomp, oel, cnk, rad, rad_der, \
                                                           ang, ang_r, ang_a, ang_p, soap, sqrt_dot_p, rotp,      \
                                                           cart);                                                 \
    return true;                                                                                                  \
  }
    JW(6) JW(10) JW(13)
#undef JW
    return false;
  }
  int nspt, nt;
  der_config(n_soap, batched_variant(n_soap, variant), nspt, nt);
  size_t sh = derivatives_lds(n_max, l_max, k_max, n_soap, variant);
#define DL(N, T)                                                                                                  \
  if (nspt == N && nt == T) {                                                                                     \
    if (sh > 64 * 1024) {                                                                                         \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&derivatives_k<N, T, true>),                       \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);                             \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&derivatives_k<N, T, false>),                      \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);                             \
    }                                                                                                             \
    if (ocomp)                                                                                                    \
      derivatives_k<N, T, true><<<n_sites, T, sh, st>>>(n_sites, n_max, l_max, k_max, n_soap, site_ptr, comp,    \
                                                        out_ptr, out_c, out_el, ocomp, oel, cnk, rad, rad_der,    \
                                                        ang, ang_r, ang_a, ang_p, soap, sqrt_dot_p, rotp, cart);  \
    else                                                                                                          \
      derivatives_k<N, T, false><<<n_sites, T, sh, st>>>(n_sites, n_max, l_max, k_max, n_soap, site_ptr, comp,   \
                                                         out_ptr, out_c, out_el, ocomp, oel, cnk, rad, rad_der,   \
                                                         ang, ang_r, ang_a, ang_p, soap, sqrt_dot_p, rotp, cart); \
    return true;                                                                                                  \
  }
  DL(1, 256) DL(2, 256) DL(3, 256) DL(4, 256) DL(2, 512) DL(3, 512) DL(4, 512) DL(6, 512)
#undef DL
  return false;
}

void launch_contract_cart(hipStream_t st, int n_sites, int n_pairs, int n_soap, const int* site_ptr, const double* G,
                          int ldg, const double* cart, double* pair) {
  if (n_pairs == 0 || n_sites == 0) return;
  contract_cart_k<<<n_sites, 256, sizeof(double) * n_soap, st>>>(n_sites, n_pairs, n_soap, ldg, site_ptr, G, cart,
                                                                  pair);
}

size_t contract_lds(int n_max, int k_max, int n_soap, int n_soap_unc) {
  return sizeof(double2) * 2 * n_max * k_max + sizeof(double) * (n_soap + n_soap_unc);
}

void launch_contract(hipStream_t st, int n_sites, int n_max, int l_max, int k_max, int n_soap, int n_soap_unc,
                     const int* site_ptr, const uint8_t* skip, const int* unc_ptr, const int* unc_e,
                     const int* umap, const double* uel,
                     const int32_t* P_i, const double* P_el, const double2* cnk, const double* rad,
                     const double* rad_der, const double2* ang, const double2* ang_r, const double2* ang_a,
                     const double2* ang_p, const double* soap, const double* G, double sgn, const double* sqrt_dot_p,
                     const double* rjs, const double* thetas, const double* phis, double* pair) {
  if (n_sites == 0) return;
  size_t sh = contract_lds(n_max, k_max, n_soap, n_soap_unc);
  contract_k<<<n_sites, 256, sh, st>>>(n_sites, n_max, l_max, k_max, n_soap, n_soap_unc, site_ptr, skip, unc_ptr, unc_e,
                                       umap, uel, P_i, P_el, cnk, rad, rad_der, ang, ang_r, ang_a, ang_p, soap, G, sgn,
                                       sqrt_dot_p, rjs, thetas, phis, pair);
}

}  // namespace st
