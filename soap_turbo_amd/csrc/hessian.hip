// Central-atom Hessian of the descriptor, contracted with weight vectors, in FP64:
// get_soap_central_hessian, src/soap_turbo.f90:1151-1700. For pair k2 = (site i, neighbour j),
//   hess(a, b, v, k2) = sum_d vecs(d, v, i) d^2 q_d / dr_ia dr_jb,
// and the self pair's block is minus the sum of the others. q = p / |p| is the normalised
// (compressed) descriptor.
//
// Kernels, on the resident FP64 state of the descriptor pass (cnk, the radial expansion with first
// and second derivatives, the normalised descriptor and its norm):
//   hess_geom_k      one thread per pair: the pair's geometry head (F_l and its radial derivatives,
//                    the unit vector, the derivatives of t = z/r and w = (x - iy)/r, w^m, Q_l^m(t));
//   hess_gacc_k      one workgroup per site: G(n, k, a) = sum_j dc^j_(n,k) / du_ja (pass A);
//   hess_d1_k        one workgroup per site: the central-atom derivative D1 = dp/dr_i from G and c
//                    (an accumulated form of minus the sum of the per-pair derivatives), and the
//                    per-site dot products of the normalisation;
//   hess_weights_k   one workgroup per (site, weight vector): the caller's vectors, p itself and
//                    the three columns of D1, each spread to its uncompressed S matrix, Z = S c and,
//                    for the caller's vectors and p, ZG = S G;
//   hess_contract_k  one workgroup per (site, weight vector), Z and ZG in LDS: per pair, Omega =
//                    C_k Q_l^m(t) w^m with its Cartesian gradient and Hessian, contracted (pass B;
//                    one launch for the columns of D1, one for the caller's vectors and p);
//   hess_norm_k      one workgroup per site: the quotient rule of the normalisation, and the self
//                    pair's blocks.
// Dot products of the per-pair derivative with a vector (p.dp2, w.dp2, D1.dp2) come out of the same
// contraction as first derivatives, so the per-pair derivative itself is never formed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "../../include/soap_turbo_hip.h"
#include "kernels.h"
#include "st_internal.h"

#pragma clang fp contract(off)

namespace st {

namespace {

constexpr int HT = 256;  // threads per workgroup
constexpr int MAXL = ST_MAX_LMAX;
constexpr int MAXK = 1 + MAXL * (MAXL + 1) / 2 + MAXL;

struct HessParams {
  int n_sites, n_pairs, n_max, k_max, l_max, n_soap, n_unc, n_vec, n_species, compressed;
  double rcut_max;
  double sigma_t[ST_MAX_SPECIES], sigma_t_scaling[ST_MAX_SPECIES];
  // per-site scratch: offsets (in doubles) of each array in one site's slot, and the slot size
  size_t o_gacc, o_z, o_zg, o_d1u, o_d1c, o_pvec, o_scal, stride;
};

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
// Re(a b), and a b, with fused multiply-adds
__device__ inline double re_mul(double2 a, double2 b) { return fma(a.x, b.x, -(a.y * b.y)); }
__device__ inline double2 cmul_f(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
// lane l's value of v, l uniform over the wavefront (v_readlane)
__device__ inline double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// (site, weight vector) of a workgroup of a 1-D grid of n_sites * nv. Workgroups are dealt to the
// eight XCDs (each with its own L2) in turn; this bijection gives each XCD a contiguous run of the
// site-major list, so the vectors of one site run together on one XCD and share its L2 copy of the
// site's data (c and G, the pairs' heads and radial values).
__device__ inline void site_and_vector(int nv, int& i, int& v) {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b & 7, j = b >> 3, q = nb >> 3, rmd = nb & 7;
  const int w = x * q + (x < rmd ? x : rmd) + j;
  i = w / nv;
  v = w - i * nv;
}

// dst[0, n) = src[0, n) by the workgroup, eight loads in flight per thread
__device__ inline void stage_lds(double2* __restrict__ dst, const double2* __restrict__ src, int n, int tx) {
  constexpr int SU = 8;
  for (int e0 = tx; e0 < n; e0 += HT * SU) {
    double2 t[SU];  // every element loaded (index clamped into range), so t stays in registers
#pragma unroll
    for (int u = 0; u < SU; ++u) t[u] = src[e0 + u * HT < n ? e0 + u * HT : n - 1];
#pragma unroll
    for (int u = 0; u < SU; ++u)
      if (e0 + u * HT < n) dst[e0 + u * HT] = t[u];
  }
}

// get_ilexp, src/soap_turbo_angular.f90:361-422, to lmax (<= MAXL + 2)
__device__ void ilexp_dev(int lmax, double xx, double* il) {
  const double xcut = 1e-7;
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
  if (lmax >= 1) {
    fact = fact * 3.0;
    il[1] = x2 / 1000.0 < xcut ? (x2 - x4) / fact : flm1;
  }
  for (int l = 2; l <= lmax; ++l) {
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

// One pair's geometry: pair_geometry, src/soap_turbo.f90:1585-1697. The head (up to Om) is
// per pair and is written by hess_geom_k once per pair; pass A builds Omega and its gradient, per
// k, in LDS after it (the contraction forms its terms from the head directly).
#define ST_PAIR_HEAD                                                  \
  double ok, invr, nh[3];                                             \
  double Fl[MAXL + 1], Fl1[MAXL + 1], Fl2[MAXL + 1];                  \
  double Qt[MAXK + 3];                                                \
  double tvec[3], thes[9];                                            \
  double2 wp[MAXL + 3]; /* w^m at m + 2 (w^-2 = w^-1 = 0) */          \
  double2 wvec[3], whes[9];
struct PairHead {
  ST_PAIR_HEAD
};
struct PairGeom {
  ST_PAIR_HEAD
  double2 Om[MAXK], dOm[3 * MAXK];
};
#undef ST_PAIR_HEAD
static_assert(offsetof(PairGeom, Om) == sizeof(PairHead), "the head leads PairGeom");
static_assert(sizeof(PairHead) % 16 == 0, "heads stay 16-byte aligned");
constexpr int GEOM_HEAD = (int)(sizeof(PairHead) / sizeof(double));  // doubles per pair

// pair_geometry in three steps, each run by one wavefront for its own pair; the workgroup
// synchronises between them, so every wavefront calls each step (with k2 < 0: no pair this round).
// Step 1, one lane: F_l and its two radial derivatives, the unit vector, and the Cartesian
// derivatives of t = z/r and w = (x - iy)/r.
template <class G>
__device__ void geom_scalars(G& g, const HessParams& P, int k2, const double* __restrict__ rjs,
                             const double* __restrict__ thetas, const double* __restrict__ phis,
                             const uint32_t* __restrict__ spmask, bool want2) {
  const int L = P.l_max;
  if (k2 < 0) { g.ok = 0.0; return; }
  const double rr = rjs[k2];
  const uint32_t msk = spmask[k2];
  g.ok = (!(rr >= P.rcut_max || rr < 1e-10) && msk != 0) ? 1.0 : 0.0;
  if (!g.ok) return;
  const int isp = __ffs(msk) - 1;
  const double invr = 1.0 / rr;
  g.invr = invr;
  const double sscal = P.sigma_t_scaling[isp];
  const double sig = P.sigma_t[isp] + sscal * rr;
  const double amp = P.rcut_max * P.rcut_max / (sig * sig);
  const double ampd = -2.0 * sscal / sig * amp;
  const double ampdd = 6.0 * (sscal * sscal) / (sig * sig) * amp;
  const double xarg = rr / sig, zz = xarg * xarg, invz = 1.0 / zz;
  double il[MAXL + 3];
  ilexp_dev(L + 2, xarg, il);
  const double zp = 2.0 * rr / (sig * sig) - 2.0 * (rr * rr) * sscal / (sig * sig * sig);
  const double zp2 = 2.0 / (sig * sig) - 8.0 * rr * sscal / (sig * sig * sig) +
                     6.0 * (rr * rr) * (sscal * sscal) / (sig * sig * sig * sig);
  for (int l = 0; l <= L; ++l) {
    const double rl = (double)l;
    const double dilz = il[l + 1] + (rl * invz - 1.0) * il[l];
    const double d2ilz = il[l + 2] + ((double)(2 * l + 1) * invz - 2.0) * il[l + 1] +
                         (rl * (rl - 1.0) * invz * invz - 2.0 * rl * invz + 1.0) * il[l];
    const double f0 = il[l], f1 = dilz * zp, f2 = d2ilz * zp * zp + dilz * zp2;
    g.Fl[l] = amp * f0;
    g.Fl1[l] = ampd * f0 + amp * f1;
    g.Fl2[l] = ampdd * f0 + 2.0 * ampd * f1 + amp * f2;
  }
  const double sin_t = sin(thetas[k2]), cos_t = cos(thetas[k2]);
  const double sin_f = sin(phis[k2]), cos_f = cos(phis[k2]);
  const double nh[3] = {sin_t * cos_f, sin_t * sin_f, cos_t};
  g.nh[0] = nh[0]; g.nh[1] = nh[1]; g.nh[2] = nh[2];
  const double tt = nh[2];
  const double2 wc = make_double2(nh[0], -nh[1]);
  const double2 ec[3] = {make_double2(1.0, 0.0), make_double2(0.0, -1.0), make_double2(0.0, 0.0)};
  for (int a = 0; a < 3; ++a) {
    g.tvec[a] = -tt * nh[a] * invr;
    g.wvec[a] = cadd(cscale(-nh[a] * invr, wc), cscale(invr, ec[a]));
  }
  g.tvec[2] += invr;
  if (want2) {
    const double invr2 = invr * invr;
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a) {
        g.thes[3 * a + b] = 3.0 * tt * nh[a] * nh[b] * invr2;
        double2 w3 = cscale(3.0 * nh[a] * nh[b], wc);
        w3.x -= ec[a].x * nh[b] + ec[b].x * nh[a];
        w3.y -= ec[a].y * nh[b] + ec[b].y * nh[a];
        g.whes[3 * a + b] = cscale(invr2, w3);
      }
    for (int a = 0; a < 3; ++a) {
      g.thes[3 * 2 + a] -= nh[a] * invr2;
      g.thes[3 * a + 2] -= nh[a] * invr2;
      g.thes[3 * a + a] -= tt * invr2;
      g.whes[3 * a + a] = make_double2(g.whes[3 * a + a].x - wc.x * invr2, g.whes[3 * a + a].y - wc.y * invr2);
    }
  }
  g.wp[0] = g.wp[1] = make_double2(0.0, 0.0);
  g.wp[2] = make_double2(1.0, 0.0);
  for (int m = 1; m <= L; ++m) g.wp[m + 2] = cmul(g.wp[m + 1], wc);
}

// Step 1, lanes m <= l_max: Q_l^m(t) = d^m P_l / dt^m, column m, the reference's recursion in l
template <class G>
__device__ void geom_qt(G& g, const HessParams& P, int k2, int m, const double* __restrict__ thetas) {
  const int L = P.l_max;
  if (k2 < 0 || m > L) return;
  const double tt = cos(thetas[k2]);
  double qmm = 1.0;
  for (int l = 1; l <= m; ++l) qmm = (double)(2 * l - 1) * qmm;
  g.Qt[m * (m + 1) / 2 + m] = qmm;
  if (m + 1 <= L) g.Qt[(m + 1) * (m + 2) / 2 + m] = (double)(2 * m + 1) * tt * qmm;
  for (int l = m + 2; l <= L; ++l) {
    const int kq = l * (l + 1) / 2 + m;
    g.Qt[kq] = ((double)(2 * l - 1) * tt * g.Qt[kq - l] - (double)(l + m - 1) * g.Qt[kq - 2 * l + 1]) / (double)(l - m);
  }
}

// Step 2, lanes over k: Omega_k = C_k Q_l^m(t) w^m and its gradient
__device__ void geom_omega(PairGeom& g, const HessParams& P, int lane, const double* __restrict__ Ck) {
  if (g.ok == 0.0) return;
  for (int k = lane; k < P.k_max; k += 64) {
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= k) ++l;
    const int mm = k - l * (l + 1) / 2;
    const double q0 = g.Qt[k], q1 = mm + 1 <= l ? g.Qt[k + 1] : 0.0;
    const double C = Ck[k], dm = (double)mm;
    const double2 w0 = g.wp[mm + 2], w1 = g.wp[mm + 1];
    g.Om[k] = cscale(C * q0, w0);
    for (int a = 0; a < 3; ++a)
      g.dOm[3 * k + a] = cscale(C, cadd(cscale(q1 * g.tvec[a], w0), cscale(q0 * dm, cmul(w1, g.wvec[a]))));
  }
}

// The per-pair head of the geometry, one thread per pair (every pair that is not a site's own)
__global__ __launch_bounds__(128) void hess_geom_k(HessParams P, const double* __restrict__ rjs,
                                                  const double* __restrict__ thetas, const double* __restrict__ phis,
                                                  const uint32_t* __restrict__ spmask, const uint8_t* __restrict__ first,
                                                  double* __restrict__ heads) {
  const int k2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (k2 >= P.n_pairs) return;
  // written in place (each head is 16-byte aligned: sizeof(PairHead) is a multiple of 16)
  PairHead& g = *reinterpret_cast<PairHead*>(heads + (size_t)k2 * GEOM_HEAD);
  geom_scalars(g, P, first[k2] ? -1 : k2, rjs, thetas, phis, spmask, true);
  if (g.ok != 0.0)
    for (int m = 0; m <= P.l_max; ++m) geom_qt(g, P, k2, m, thetas);
}

// The geometry of pair k2 (< 0: none) for this wavefront, in g: the head copied from hess_geom_k's
// output, then Omega by the lanes. Called by every thread of the workgroup, which synchronises
// around each step.
__device__ void wave_geometry(PairGeom& g, const HessParams& P, int k2, const double* __restrict__ heads,
                              const double* __restrict__ Ck) {
  const int lane = threadIdx.x & 63;
  __syncthreads();  // every wavefront is done with the previous round's geometry
  double* dst = reinterpret_cast<double*>(&g);
  if (k2 < 0) {
    if (lane == 0) g.ok = 0.0;
  } else {
    for (int q = lane; q < GEOM_HEAD; q += 64) dst[q] = heads[(size_t)k2 * GEOM_HEAD + q];
  }
  __syncthreads();
  geom_omega(g, P, lane, Ck);
  __syncthreads();
}

// block sum of NV values (every thread passes its partials; thread 0 gets the totals)
template <int NV>
__device__ void block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) red[w * NV + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = red[q] + red[NV + q] + red[2 * NV + q] + red[3 * NV + q];
  __syncthreads();
}

// pass A: G(a, n, k) = sum_j [g'(r) n_a Omega + g(r) dOmega_a] with g = 4 pi rad F_l; the four
// wavefronts build the geometry of four pairs per round, then every thread adds them to its
// elements in pair order
// EPT > 0: each thread keeps its (at most EPT) elements' sums in registers across the rounds;
// EPT = 0: they are read and written back in the scratch slot every round
template <int EPT>
__global__ __launch_bounds__(HT) void hess_gacc_k(HessParams P, const int* __restrict__ site_ptr,
                                                 const double* __restrict__ heads,
                                                 const double* __restrict__ Ck, const double* __restrict__ rad,
                                                 const double* __restrict__ rad_der, double* __restrict__ scratch) {
  __shared__ PairGeom gs[HT / 64];
  const int i = blockIdx.x, tx = threadIdx.x, wv = tx >> 6;
  const int n_max = P.n_max, k_max = P.k_max, E = n_max * k_max;
  const double fourpi = 4.0 * 3.141592653589793;
  double2* G = reinterpret_cast<double2*>(scratch + (size_t)i * P.stride + P.o_gacc);
  double2 reg[EPT > 0 ? EPT : 1][3];
  if constexpr (EPT > 0) {
#pragma unroll
    for (int u = 0; u < EPT; ++u)
#pragma unroll
      for (int a = 0; a < 3; ++a) reg[u][a] = make_double2(0.0, 0.0);
  } else {
    for (int e = tx; e < 3 * E; e += HT) G[e] = make_double2(0.0, 0.0);
  }
  const int k3 = site_ptr[i], k4 = site_ptr[i + 1];
  for (int r0 = k3 + 1; r0 < k4; r0 += HT / 64) {
    const int k2 = r0 + wv < k4 ? r0 + wv : -1;
    wave_geometry(gs[wv], P, k2, heads, Ck);
#pragma unroll
    for (int u = 0; u < (EPT > 0 ? EPT : 1 << 20); ++u) {
      const int e = tx + HT * u;
      if (e >= E) break;
      const int n = e / k_max, k = e - n * k_max;
      int l = 0;
      while ((l + 1) * (l + 2) / 2 <= k) ++l;
      double2 acc[3];
      if constexpr (EPT > 0) {
        acc[0] = reg[u][0]; acc[1] = reg[u][1]; acc[2] = reg[u][2];
      } else {
        acc[0] = G[e]; acc[1] = G[E + e]; acc[2] = G[2 * E + e];
      }
      for (int w = 0; w < HT / 64 && r0 + w < k4; ++w) {
        const PairGeom& g = gs[w];
        if (g.ok == 0.0) continue;
        const int kp = r0 + w;
        const double r_0 = fourpi * rad[(size_t)kp * n_max + n], r_1 = fourpi * rad_der[(size_t)kp * n_max + n];
        const double grad0 = r_0 * g.Fl[l], grad1 = r_1 * g.Fl[l] + r_0 * g.Fl1[l];
        for (int a = 0; a < 3; ++a)
          acc[a] = cadd(acc[a], cadd(cscale(grad1 * g.nh[a], g.Om[k]), cscale(grad0, g.dOm[3 * k + a])));
      }
      if constexpr (EPT > 0) {
        reg[u][0] = acc[0]; reg[u][1] = acc[1]; reg[u][2] = acc[2];
      } else {
        G[e] = acc[0]; G[E + e] = acc[1]; G[2 * E + e] = acc[2];
      }
    }
  }
  if constexpr (EPT > 0) {
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = tx + HT * u;
      if (e < E) { G[e] = reg[u][0]; G[E + e] = reg[u][1]; G[2 * E + e] = reg[u][2]; }
    }
  }
}

// D1 = dp/dr_i, the unnormalised p and the per-site dot products, one workgroup per site
__global__ __launch_bounds__(HT) void hess_d1_k(HessParams P, const double2* __restrict__ cnk,
                                                    const double* __restrict__ soap, const double* __restrict__ sdp,
                                                    const double* __restrict__ vecs, const int4* __restrict__ comp,
                                                    const uint8_t* __restrict__ skip, const int* __restrict__ out_ptr,
                                                    const int* __restrict__ out_c, const double* __restrict__ out_el,
                                                    const int* __restrict__ unc_ptr, const int* __restrict__ unc_e,
                                                    const int32_t* __restrict__ P_i, const double* __restrict__ P_el,
                                                    double* __restrict__ scratch) {
  __shared__ double red[4 * 8];
  extern __shared__ double2 cg[];  // the site's c [n][k], then G [a][n][k]
  const int i = blockIdx.x, tx = threadIdx.x;
  const int n_max = P.n_max, k_max = P.k_max, E = n_max * k_max;
  const int n_unc = P.n_unc, n_soap = P.n_soap, n_vec = P.n_vec;
  const double sq2 = 1.4142135623730951;
  double* base = scratch + (size_t)i * P.stride;
  double* D1u = base + P.o_d1u;
  double* D1c = base + P.o_d1c;
  double* pvec = base + P.o_pvec;
  double* scal = base + P.o_scal;
  // every element of c and G is read for each of the n_max components it pairs with: staged once
  stage_lds(cg, cnk + (size_t)i * E, E, tx);
  stage_lds(cg + E, reinterpret_cast<const double2*>(base + P.o_gacc), 3 * E, tx);
  __syncthreads();
  const double2* cn = cg;
  const double2* G = cg + E;
  // D1(n, n', l, a) = -base sum_m w(m) Re(G_n,a conj(c_n') + c_n conj(G_n',a)), skipped components 0
  for (int u = tx; u < n_unc; u += HT) {
    const int4 cm = comp[u];
    const int n = cm.x, np = cm.y, l = cm.z;
    double d[3] = {0.0, 0.0, 0.0};
    if (!skip[u]) {
      const double bs = n == np ? 1.0 : sq2;
      for (int m = 0; m <= l; ++m) {
        const int k = l * (l + 1) / 2 + m;
        const double w = m == 0 ? 1.0 : 2.0;
        const double2 c1 = cn[n * k_max + k], c2 = cn[np * k_max + k];
        for (int a = 0; a < 3; ++a) {
          const double2 g1 = G[a * E + n * k_max + k], g2 = G[a * E + np * k_max + k];
          d[a] += w * (g1.x * c2.x + g1.y * c2.y + c1.x * g2.x + c1.y * g2.y);
        }
      }
      for (int a = 0; a < 3; ++a) d[a] = -bs * d[a];
    }
    for (int a = 0; a < 3; ++a) D1u[a * n_unc + u] = d[a];
  }
  __syncthreads();
  // compressed D1 and the unnormalised descriptor p
  const double s = sdp[i];
  for (int q = tx; q < n_soap; q += HT) {
    for (int a = 0; a < 3; ++a) {
      double v = 0.0;
      if (P.compressed) {
        for (int e = out_ptr[q]; e < out_ptr[q + 1]; ++e) v += out_el[e] * D1u[a * n_unc + out_c[e]];
      } else {
        v = D1u[a * n_unc + q];
      }
      D1c[a * n_soap + q] = v;
    }
    pvec[q] = soap[(size_t)i * n_soap + q] * s;
  }
  __syncthreads();
  // per-site dot products: t1(a) = p.D1_a; for each caller vector, w.p and w.D1_a
  {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = tx; q < n_soap; q += HT)
      for (int a = 0; a < 3; ++a) v[a] += pvec[q] * D1c[a * n_soap + q];
    block_sum<4>(v, red);
    if (tx == 0)
      for (int a = 0; a < 3; ++a) scal[a] = v[a];
    for (int vv = 0; vv < n_vec; ++vv) {
      const double* w = vecs + ((size_t)i * n_vec + vv) * n_soap;
      double x[4] = {0.0, 0.0, 0.0, 0.0};
      for (int q = tx; q < n_soap; q += HT) {
        x[0] += w[q] * pvec[q];
        for (int a = 0; a < 3; ++a) x[a + 1] += w[q] * D1c[a * n_soap + q];
      }
      block_sum<4>(x, red);
      if (tx == 0)
        for (int a = 0; a < 4; ++a) scal[3 + 4 * vv + a] = x[a];
    }
  }
}

// Z(n, k) = sum_n' S(n, n', l) c(n', k) and (WITH_G) ZG(a, n, k) = sum_n' S(n, n', l) G(a, n', k),
// S(n, n', l) = 2 vu(n, n', l) on the diagonal and sqrt 2 vu off it, vu the spread weight vector in
// LDS. Each thread takes NB consecutive n at one k, so a column of c (and G) is read from memory
// once per NB outputs; consecutive threads take consecutive k.
template <bool WITH_G>
__device__ inline void spread_rows(int n_max, int k_max, int L1, int E, const double2* __restrict__ cn,
                                   const double2* __restrict__ G, const double* __restrict__ vu,
                                   double2* __restrict__ Z, double2* __restrict__ ZG, int tx) {
  constexpr int NB = 4;
  const double sq2 = 1.4142135623730951;
  const int nblk = (n_max + NB - 1) / NB;
  for (int t = tx; t < nblk * k_max; t += HT) {
    const int blk = t / k_max, k = t - blk * k_max, n0 = blk * NB;
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= k) ++l;
    double2 z[NB], zg[WITH_G ? NB : 1][3];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      z[u] = make_double2(0.0, 0.0);
      if constexpr (WITH_G)
#pragma unroll
        for (int a = 0; a < 3; ++a) zg[u][a] = make_double2(0.0, 0.0);
    }
    for (int np = 0; np < n_max; ++np) {
      const double2 cv = cn[np * k_max + k];
      double2 gv[3];
      if constexpr (WITH_G)
#pragma unroll
        for (int a = 0; a < 3; ++a) gv[a] = G[a * E + np * k_max + k];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int n = n0 + u;
        if (n >= n_max) break;
        const int a_ = n < np ? n : np, b_ = n < np ? np : n;
        const double sv = (n == np ? 2.0 : sq2) * vu[((a_ * (2 * n_max - a_ + 1)) / 2 + (b_ - a_)) * L1 + l];
        z[u] = cadd(z[u], cscale(sv, cv));
        if constexpr (WITH_G)
#pragma unroll
          for (int a = 0; a < 3; ++a) zg[u][a] = cadd(zg[u][a], cscale(sv, gv[a]));
      }
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int n = n0 + u;
      if (n >= n_max) break;
      Z[n * k_max + k] = z[u];
      if constexpr (WITH_G)
#pragma unroll
        for (int a = 0; a < 3; ++a) ZG[a * E + n * k_max + k] = zg[u][a];
    }
  }
}

// one weight vector per workgroup: its spread P^T v, Z = S c and (caller's vectors and p) ZG = S G
__global__ __launch_bounds__(HT) void hess_weights_k(HessParams P, const double2* __restrict__ cnk,
                                                    const double* __restrict__ vecs, const uint8_t* __restrict__ skip,
                                                    const int* __restrict__ unc_ptr, const int* __restrict__ unc_e,
                                                    const int32_t* __restrict__ P_i, const double* __restrict__ P_el,
                                                    double* __restrict__ scratch) {
  extern __shared__ double vunc[];  // n_unc
  const int tx = threadIdx.x;
  int i, v;
  site_and_vector(P.n_vec + 4, i, v);
  const int n_max = P.n_max, k_max = P.k_max, E = n_max * k_max, L1 = P.l_max + 1;
  const int n_unc = P.n_unc, n_soap = P.n_soap, n_vec = P.n_vec;
  double* base = scratch + (size_t)i * P.stride;
  const double2* G = reinterpret_cast<const double2*>(base + P.o_gacc);
  double2* Z = reinterpret_cast<double2*>(base + P.o_z);
  double2* ZG = reinterpret_cast<double2*>(base + P.o_zg);
  const double* D1c = base + P.o_d1c;
  const double* pvec = base + P.o_pvec;
  const double2* cn = cnk + (size_t)i * E;
  // weight vector v: the caller's (v < n_vec), p (v = n_vec), then the columns of D1
  {
    const double* vin = v < n_vec ? vecs + ((size_t)i * n_vec + v) * n_soap : v == n_vec ? pvec : D1c + (v - n_vec - 1) * n_soap;
    __syncthreads();
    // P^T v, skipped components cleared (spread_weights, src/soap_turbo.f90:1510-1525)
    for (int u = tx; u < n_unc; u += HT) {
      double x = 0.0;
      if (!skip[u]) {
        if (P.compressed) {
          for (int e = unc_ptr[u]; e < unc_ptr[u + 1]; ++e) x += P_el[unc_e[e]] * vin[P_i[unc_e[e]] - 1];
        } else {
          x = vin[u];
        }
      }
      vunc[u] = x;
    }
    __syncthreads();
    // Z = S c and, for the caller's vectors and p, ZG = S G
    if (v <= n_vec)
      spread_rows<true>(n_max, k_max, L1, E, cn, G, vunc, Z + (size_t)v * E, ZG + (size_t)v * 3 * E, tx);
    else
      spread_rows<false>(n_max, k_max, L1, E, cn, G, vunc, Z + (size_t)v * E, ZG, tx);
  }
}

// Sums over n of one pair's radial values against the staged Z(v) (and ZG(v)) at one k:
//   a_t = sum_n R_t(n) conj Z(n, k), t = 0, 1 (and 2 with SECOND), R_t = rad, rad', rad'';
//   b_t(a) = sum_n R_t(n) conj ZG(a, n, k), t = 0, 1 (SECOND only).
// zp points at Z(0, k) in LDS. The sums run unconjugated with fused multiply-adds and are
// conjugated at the end (a negation, exact). Lane n holds R_t(n) (64 n at a time), read out with
// v_readlane, so every lane of the wavefront must have loaded its values: the caller runs this
// with the whole wavefront, each lane's k clamped into range. U LDS reads per operand are issued
// before their sums, so their latencies overlap.
struct RadialSums {
  double2 a0, a1, a2, b0[3], b1[3];
};

template <bool SECOND, int U>
__device__ inline void radial_steps(const double2* __restrict__ zp, int E, int k_max, int j, double rv0, double rv1,
                                    double rv2, RadialSums& s) {
  double2 zc[U], gc[SECOND ? U : 1][3];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    zc[u] = zp[(j + u) * k_max];
    if constexpr (SECOND)
#pragma unroll
      for (int a = 0; a < 3; ++a) gc[u][a] = zp[(a + 1) * E + (j + u) * k_max];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const double r0 = readlane_d(rv0, j + u), r1 = readlane_d(rv1, j + u);
    s.a0.x = fma(r0, zc[u].x, s.a0.x); s.a0.y = fma(r0, zc[u].y, s.a0.y);
    s.a1.x = fma(r1, zc[u].x, s.a1.x); s.a1.y = fma(r1, zc[u].y, s.a1.y);
    if constexpr (SECOND) {
      const double r2 = readlane_d(rv2, j + u);
      s.a2.x = fma(r2, zc[u].x, s.a2.x); s.a2.y = fma(r2, zc[u].y, s.a2.y);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        s.b0[a].x = fma(r0, gc[u][a].x, s.b0[a].x); s.b0[a].y = fma(r0, gc[u][a].y, s.b0[a].y);
        s.b1[a].x = fma(r1, gc[u][a].x, s.b1[a].x); s.b1[a].y = fma(r1, gc[u][a].y, s.b1[a].y);
      }
    }
  }
}

template <bool SECOND>
__device__ inline void radial_sums(const double2* __restrict__ zp, int E, int k_max, int n_max, int lane,
                                   const double* __restrict__ r, const double* __restrict__ rd,
                                   const double* __restrict__ rh, RadialSums& s) {
  const double2 zero = make_double2(0.0, 0.0);
  s.a0 = s.a1 = s.a2 = zero;
#pragma unroll
  for (int a = 0; a < 3; ++a) s.b0[a] = s.b1[a] = zero;
  for (int c0 = 0; c0 < n_max; c0 += 64) {
    const int nl = c0 + lane;
    const double rv0 = nl < n_max ? r[nl] : 0.0, rv1 = nl < n_max ? rd[nl] : 0.0;
    const double rv2 = SECOND && nl < n_max ? rh[nl] : 0.0;
    const int c1 = n_max - c0 < 64 ? n_max - c0 : 64;
    const double2* zc = zp + (size_t)c0 * k_max;
    int j = 0;
    for (; j + 4 <= c1; j += 4) radial_steps<SECOND, 4>(zc, E, k_max, j, rv0, rv1, rv2, s);
    for (; j < c1; ++j) radial_steps<SECOND, 1>(zc, E, k_max, j, rv0, rv1, rv2, s);
  }
  s.a0.y = -s.a0.y; s.a1.y = -s.a1.y; s.a2.y = -s.a2.y;
#pragma unroll
  for (int a = 0; a < 3; ++a) { s.b0[a].y = -s.b0[a].y; s.b1[a].y = -s.b1[a].y; }
}

// One step of a transposing wavefront sum: lanes with bit d set keep the upper H of their 2H
// values and send the lower H to lane ^ d, which keeps the lower H and sends the upper.
template <int H>
__device__ inline void halve_sum(double (&u)[16], int lane, int d) {
  const bool hi = (lane & d) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double keep = hi ? u[i + H] : u[i], send = hi ? u[i] : u[i + H];
    u[i] = keep + __shfl_xor(send, d, 64);
  }
}

// Sums of acc[0..11] over the wavefront in 17 exchanges (72 for twelve separate butterflies): lane
// L ends with the sum of acc[(L >> 2) & 15] (zero for 12..15).
__device__ inline double wave_sum12(const double (&acc)[12], int lane) {
  double u[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) u[q] = q < 12 ? acc[q] : 0.0;
  halve_sum<8>(u, lane, 32);
  halve_sum<4>(u, lane, 16);
  halve_sum<2>(u, lane, 8);
  halve_sum<1>(u, lane, 4);
  double x = u[0];
  x += __shfl_xor(x, 2, 64);
  x += __shfl_xor(x, 1, 64);
  return x;
}

// pass B, contraction: one workgroup per (site, weight vector v), with Z(v) (and, for the caller's
// vectors and p, ZG(v)) staged in LDS once; its wavefronts walk the site's pairs, lanes over k.
// Per pair and v it writes 12 sums: v.dp/du_jb (3) and, with second derivatives, v.d2p/dr_ia
// du_jb before the normalisation (9).
// SECOND: the caller's vectors and p (second derivatives, Z and ZG staged); otherwise the three
// columns of D1 (first derivatives only, Z staged), a launch of its own with a quarter of the LDS
// and fewer registers
template <bool SECOND>
__global__ __launch_bounds__(HT) void hess_contract_k(HessParams P, const int* __restrict__ site_ptr,
                                                     const double* __restrict__ heads, const double* __restrict__ Ck,
                                                     const double* __restrict__ rad, const double* __restrict__ rad_der,
                                                     const double* __restrict__ rad_der2,
                                                     const double* __restrict__ scratch, double* __restrict__ res) {
  extern __shared__ double2 zs[];  // Z(v) [n][k], then ZG(v) [a][n][k]
  const int tx = threadIdx.x, lane = tx & 63, wv = tx >> 6;
  int i, v;
  site_and_vector(SECOND ? P.n_vec + 1 : 3, i, v);
  if (!SECOND) v += P.n_vec + 1;
  const int n_max = P.n_max, k_max = P.k_max, E = n_max * k_max, n_vec = P.n_vec, nvz = n_vec + 4;
  constexpr bool second = SECOND;
  const double fourpi = 4.0 * 3.141592653589793;
  const double* base = scratch + (size_t)i * P.stride;
  const double2* Z = reinterpret_cast<const double2*>(base + P.o_z) + (size_t)v * E;
  const double2* ZG = reinterpret_cast<const double2*>(base + P.o_zg) + (size_t)v * 3 * E;
  stage_lds(zs, Z, E, tx);
  if (second) stage_lds(zs + E, ZG, 3 * E, tx);
  __syncthreads();
  const int k3 = site_ptr[i], k4 = site_ptr[i + 1];
  for (int k2 = k3 + 1 + wv; k2 < k4; k2 += HT / 64) {
    const PairHead& h = *reinterpret_cast<const PairHead*>(heads + (size_t)k2 * GEOM_HEAD);
    if (h.ok == 0.0) continue;  // uniform over the wavefront
    const double* r = rad + (size_t)k2 * n_max;
    const double* rd = rad_der + (size_t)k2 * n_max;
    const double* rh = rad_der2 + (size_t)k2 * n_max;
    const double invr = h.invr;
    const double nh[3] = {h.nh[0], h.nh[1], h.nh[2]};
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) acc[q] = 0.0;
    // lanes over k, 64 at a time; the whole wavefront runs the radial sums (k clamped)
    for (int kb = 0; kb < k_max; kb += 64) {
      const int k = kb + lane, kc = k < k_max ? k : k_max - 1;
      int l = 0;
      while ((l + 1) * (l + 2) / 2 <= kc) ++l;
      const int m = kc - l * (l + 1) / 2;
      // Omega_k = C Q_l^m(t) w^m enters only through real parts of products with it, so each term
      // Re(y d^n Omega) is formed from y C w^m, y C w^(m-1) and y C w^(m-2) and the head's
      // derivatives of t and w (the chain rule of pair_geometry, src/soap_turbo.f90:1648-1697).
      // The head's per-k values are loaded before the radial sums, so their latency overlaps them.
      const double f0 = h.Fl[l], f1 = h.Fl1[l], f2 = h.Fl2[l], wmk = m > 0 ? 2.0 : 1.0;
      const double q0 = h.Qt[kc], q1 = m + 1 <= l ? h.Qt[kc + 1] : 0.0, q2 = m + 2 <= l ? h.Qt[kc + 2] : 0.0;
      const double2 w0 = h.wp[m + 2], w1 = h.wp[m + 1], w2 = h.wp[m];
      RadialSums s;
      if (second)
        radial_sums<true>(zs + kc, E, k_max, n_max, lane, r, rd, rh, s);
      else
        radial_sums<false>(zs + kc, E, k_max, n_max, lane, r, rd, rh, s);
      if (k >= k_max) continue;
      const double q0m = q0 * (double)m, q1m = q1 * (double)m, q0mm = q0 * (double)(m * (m - 1));
      double tv[3];
      double2 wv[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) { tv[a] = h.tvec[a]; wv[a] = h.wvec[a]; }
      const double c4 = fourpi * Ck[kc];
      // y0 = C z0, y1 = C z1 with z0 = 4 pi f0 a0, z1 = 4 pi (f0 a1 + f1 a0)
      const double2 y0 = cscale(c4 * f0, s.a0);
      const double2 y1 = cscale(c4, make_double2(fma(f0, s.a1.x, f1 * s.a0.x), fma(f0, s.a1.y, f1 * s.a0.y)));
      const double P0 = re_mul(y0, w0), P1 = re_mul(y1, w0);  // Re(z0 Omega) / q0, Re(z1 Omega) / q0
      const double2 U1 = cmul_f(y0, w1), V1 = cmul_f(y1, w1);
      double R1[3], S1[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) { R1[a] = re_mul(U1, wv[a]); S1[a] = re_mul(V1, wv[a]); }
      const double zo1 = q0 * P1;  // Re(z1 Omega)
      // first derivatives: Re(n_b z1 Omega + z0 dOmega_b), Re(z0 dOmega_b) = q1 t_b P0 + q0 m R1_b
#pragma unroll
      for (int b = 0; b < 3; ++b) acc[b] = fma(wmk, fma(nh[b], zo1, fma(q1 * tv[b], P0, q0m * R1[b])), acc[b]);
      if (second) {
        // yr = C (z2 - z1 / r), z2 = 4 pi (f0 a2 + 2 f1 a1 + f2 a0)
        const double2 z2 = make_double2(fma(f0, s.a2.x, fma(2.0 * f1, s.a1.x, f2 * s.a0.x)),
                                        fma(f0, s.a2.y, fma(2.0 * f1, s.a1.y, f2 * s.a0.y)));
        const double2 yr = make_double2(fma(c4, z2.x, -invr * y1.x), fma(c4, z2.y, -invr * y1.y));
        const double zor = q0 * re_mul(yr, w0);  // Re(zr Omega)
        const double2 U2 = cmul_f(y0, w2);
        double E1[3], H0[3], H1[3];
        double2 G1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          E1[a] = fma(q1 * tv[a], P1, q0m * S1[a]);  // Re(z1 dOmega_a)
          // C gg0_a and C gg1_a with gg0 = 4 pi f0 b0, gg1 = 4 pi (f0 b1 + f1 b0)
          const double2 g0 = cscale(c4 * f0, s.b0[a]);
          const double2 g1 = cscale(c4, make_double2(fma(f0, s.b1[a].x, f1 * s.b0[a].x),
                                                     fma(f0, s.b1[a].y, f1 * s.b0[a].y)));
          H1[a] = q0 * re_mul(g1, w0);  // Re(gg1_a Omega)
          H0[a] = re_mul(g0, w0);
          G1[a] = cmul_f(g0, w1);
        }
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            // Re(z0 d2Omega_ab)
            double d2 = fma(fma(q2 * tv[a], tv[b], q1 * h.thes[3 * a + b]), P0,
                            q1m * fma(tv[a], R1[b], tv[b] * R1[a]));
            d2 = fma(q0mm, re_mul(U2, cmul_f(wv[a], wv[b])), d2);
            d2 = fma(q0m, re_mul(U1, h.whes[3 * a + b]), d2);
            // Re(n_a n_b zr Omega + z1 (n_a dOmega_b + n_b dOmega_a) + z0 d2Omega_ab + [a = b] z1 Omega / r)
            double tE = fma(nh[a] * nh[b], zor, fma(nh[a], E1[b], fma(nh[b], E1[a], d2)));
            if (a == b) tE = fma(invr, zo1, tE);
            // Re(n_b gg1_a Omega + gg0_a dOmega_b)
            const double tD = fma(nh[b], H1[a], fma(q1 * tv[b], H0[a], q0m * re_mul(G1[a], wv[b])));
            acc[3 + 3 * a + b] = fma(-wmk, tE + tD, acc[3 + 3 * a + b]);
          }
      }
    }
    const double x = wave_sum12(acc, lane);
    if ((lane & 3) == 0 && (lane >> 2) < 12) res[((size_t)k2 * nvz + v) * 12 + (lane >> 2)] = x;
  }
}

// pass B, normalisation, q = p / s (src/soap_turbo.f90:1453-1485): with d1 = d/dr_ia, d2 = d/dr_jb,
//   w.d1d2 q = (w.d1d2 p)/s - [(w.d1 p)(p.d2 p) + (w.d2 p)(p.d1 p)]/s^3
//              - (w.p)[(d1 p.d2 p) + (p.d1d2 p)]/s^3 + 3 (w.p)(p.d1 p)(p.d2 p)/s^5
// then the self pair's block, minus the sum of the others in pair order. One workgroup per site.
__global__ __launch_bounds__(HT) void hess_norm_k(HessParams P, const int* __restrict__ site_ptr,
                                                 const double* __restrict__ heads, const double* __restrict__ sdp,
                                                 const double* __restrict__ scratch, const double* __restrict__ res,
                                                 double* __restrict__ hess) {
  const int i = blockIdx.x, tx = threadIdx.x;
  const int n_vec = P.n_vec, nvz = n_vec + 4, nt = 9 * n_vec;
  const double* scal = scratch + (size_t)i * P.stride + P.o_scal;
  const double s = sdp[i], invs = 1.0 / s, invs3 = invs / (s * s), invs5 = invs3 / (s * s);
  const int k3 = site_ptr[i], k4 = site_ptr[i + 1];
  const int np = k4 - k3 - 1;
  for (int idx = tx; idx < np * nt; idx += HT) {
    const int k2 = k3 + 1 + idx / nt, t = idx % nt;
    const int v = t / 9, a = (t % 9) / 3, b = t % 3;
    double h = 0.0;
    if (heads[(size_t)k2 * GEOM_HEAD] != 0.0) {
      const double* rw = res + (size_t)k2 * nvz * 12;
      const double* rp = rw + n_vec * 12;  // p's row: u2 = p.dp2, hv = p.d1d2 p
      const double t1a = scal[a], t2b = rp[b];
      const double wp = scal[3 + 4 * v], u1a = scal[3 + 4 * v + 1 + a];
      const double u2b = rw[v * 12 + b];
      const double c12 = rw[(n_vec + 1 + a) * 12 + b];  // D1_a . dp2_b
      h = rw[v * 12 + 3 + 3 * a + b] * invs - (u1a * t2b + u2b * t1a) * invs3 -
          wp * (c12 + rp[3 + 3 * a + b]) * invs3 + 3.0 * wp * t1a * t2b * invs5;
    }
    hess[(((size_t)k2 * n_vec + v) * 3 + b) * 3 + a] = h;
  }
  __syncthreads();  // the blocks of this site are written (global memory, same workgroup)
  for (int t = tx; t < nt; t += HT) {
    const int v = t / 9, a = (t % 9) / 3, b = t % 3;
    double sum = 0.0;
    for (int k2 = k3 + 1; k2 < k4; ++k2) sum += hess[(((size_t)k2 * n_vec + v) * 3 + b) * 3 + a];
    hess[(((size_t)k3 * n_vec + v) * 3 + b) * 3 + a] = -sum;
  }
}

}  // namespace

size_t hessian_scratch_per_site(int n_max, int k_max, int n_soap, int n_unc, int n_vec) {
  const size_t E = (size_t)n_max * k_max;
  const size_t n = 2 * E * 3 + 2 * E * (n_vec + 4) + 2 * E * 3 * (n_vec + 1) + 3 * (size_t)n_unc +
                   3 * (size_t)n_soap + (size_t)n_soap + 3 + 4 * (size_t)n_vec + 8;
  return (n + 31) / 32 * 32;  // each site's slot starts on a 256-byte boundary
}

size_t hessian_heads_per_pair() { return GEOM_HEAD; }
size_t hessian_geom_head() { return GEOM_HEAD; }

size_t hessian_lds(int n_max, int k_max, int n_unc) {
  const size_t geom = (HT / 64) * sizeof(PairGeom);        // pass A
  const size_t w = (size_t)n_unc * 8 + 32 * 8;              // weight vectors
  // contraction: Z and ZG of one vector; central-atom derivative: c and G, and its reduction array
  const size_t c = (size_t)n_max * k_max * 4 * sizeof(double2) + 32 * 8;
  return std::max(geom, std::max(w, c));
}

size_t hessian_res_per_pair(int n_vec) { return (size_t)(n_vec + 4) * 12; }

size_t hessian_scal_offset(int n_max, int k_max, int n_soap, int n_unc, int n_vec) {
  const size_t E = (size_t)n_max * k_max;
  return 2 * E * 3 + 2 * E * (n_vec + 4) + 2 * E * 3 * (n_vec + 1) + 3 * (size_t)n_unc + 3 * (size_t)n_soap +
         (size_t)n_soap;
}

bool launch_hessian(hipStream_t st, int n_sites, int n_pairs, int n_max, int l_max, int n_soap, int n_unc, int n_vec,
                    int compressed, int n_species, const double* sigma_t, const double* sigma_t_scaling,
                    double rcut_max, const int* site_ptr, const double* rjs, const double* thetas, const double* phis,
                    const uint32_t* spmask, const uint8_t* first, const double* Ck, const double* rad,
                    const double* rad_der, const double* rad_der2, const double2* cnk, const double* soap,
                    const double* sdp, const double* vecs, const int4* comp, const uint8_t* skip, const int* out_ptr,
                    const int* out_c, const double* out_el, const int* unc_ptr, const int* unc_e, const int32_t* P_i,
                    const double* P_el, double* scratch, double* heads, double* res, double* hess) {
  if (l_max > MAXL || n_species > ST_MAX_SPECIES || n_vec < 1 || (long long)n_sites * (n_vec + 4) > 0x7fffffffll)
    return false;
  HessParams P;
  P.n_sites = n_sites; P.n_pairs = n_pairs; P.n_max = n_max; P.k_max = 1 + l_max * (l_max + 1) / 2 + l_max;
  P.l_max = l_max; P.n_soap = n_soap; P.n_unc = n_unc; P.n_vec = n_vec; P.n_species = n_species;
  P.compressed = compressed; P.rcut_max = rcut_max;
  for (int s = 0; s < ST_MAX_SPECIES; ++s) {
    P.sigma_t[s] = s < n_species ? sigma_t[s] : 0.0;
    P.sigma_t_scaling[s] = s < n_species ? sigma_t_scaling[s] : 0.0;
  }
  const size_t E = (size_t)n_max * P.k_max;
  P.o_gacc = 0;
  P.o_z = P.o_gacc + 2 * E * 3;
  P.o_zg = P.o_z + 2 * E * (n_vec + 4);
  P.o_d1u = P.o_zg + 2 * E * 3 * (n_vec + 1);
  P.o_d1c = P.o_d1u + 3 * (size_t)n_unc;
  P.o_pvec = P.o_d1c + 3 * (size_t)n_soap;
  P.o_scal = P.o_pvec + n_soap;  // == hessian_scal_offset(...)
  P.stride = hessian_scratch_per_site(n_max, P.k_max, n_soap, n_unc, n_vec);
  if (n_sites == 0) return true;
  if (n_pairs > 0)
    hess_geom_k<<<(n_pairs + 127) / 128, 128, 0, st>>>(P, rjs, thetas, phis, spmask, first, heads);
  switch ((int)((E + HT - 1) / HT)) {
#define GA(n) case n: hess_gacc_k<n><<<n_sites, HT, 0, st>>>(P, site_ptr, heads, Ck, rad, rad_der, scratch); break;
    GA(1) GA(2) GA(3) GA(4) GA(5) GA(6) GA(7) GA(8) GA(9) GA(10) GA(11) GA(12)
#undef GA
    default: hess_gacc_k<0><<<n_sites, HT, 0, st>>>(P, site_ptr, heads, Ck, rad, rad_der, scratch);
  }
  const size_t cl = E * 4 * sizeof(double2);  // c and G (hess_d1_k); Z and ZG of one vector (hess_contract_k)
  if (cl > 64 * 1024) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&hess_d1_k), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)cl);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&hess_contract_k<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)cl);
  }
  hess_d1_k<<<n_sites, HT, cl, st>>>(P, cnk, soap, sdp, vecs, comp, skip, out_ptr, out_c, out_el, unc_ptr, unc_e, P_i,
                                      P_el, scratch);
  const size_t wl = (size_t)n_unc * 8;
  if (wl > 64 * 1024) (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&hess_weights_k),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)wl);
  hess_weights_k<<<n_sites * (n_vec + 4), HT, wl, st>>>(P, cnk, vecs, skip, unc_ptr, unc_e, P_i, P_el, scratch);
  hess_contract_k<false><<<n_sites * 3, HT, E * sizeof(double2), st>>>(P, site_ptr, heads, Ck, rad, rad_der, rad_der2,
                                                                       scratch, res);
  hess_contract_k<true><<<n_sites * (n_vec + 1), HT, cl, st>>>(P, site_ptr, heads, Ck, rad, rad_der, rad_der2, scratch,
                                                           res);
  hess_norm_k<<<n_sites, HT, 0, st>>>(P, site_ptr, heads, sdp, scratch, res, hess);
  return true;
}

}  // namespace st
