// GAP-form energy and its gradient with respect to the descriptors, on the FP32 descriptors the
// site kernel leaves resident: E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta over sparse
// points q_t (the SOAP kernel of a Gaussian-approximation potential; the energy model that
// supplies get_soap's callers with dE/dq).
//
// Two FP32 matrix-core products (v_mfma_f32_16x16x4f32), 64 x 64 output tiles per workgroup of four
// wavefronts (32 x 32 each), the K dimension in steps of 32 staged through LDS with the next step's
// global loads issued before the current step's products, and split over workgroups (summed in a
// fixed order) when the output tiles alone are too few to fill the chip:
//   K = Q T^T (sites x sparse points, dot products of descriptor rows), then per site (one wavefront)
//   W = delta^2 zeta alpha_t K^(zeta-1) over K and e_i = delta^2 sum_t alpha_t K^zeta;
//   dE/dq = W T (sites x descriptor components), the site-major FP32 dE/dq the adjoint site kernel
//   reads.
// Every sum runs in a fixed order, so repeated runs agree bitwise.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "adjoint.h"

namespace st {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32;
constexpr int AS = TK + 1;  // LDS row stride of the [row][k] tiles (odd: the 16 rows of an operand read differ in bank)
constexpr int BS = TN + 4;  // ... of the [k][col] tile
static_assert(TK == 32 || TK == 64, "the [k][col] tile is loaded in blocks of 32 rows");

// The 64 x 64 tile of C = A B^T (BT, B given by its rows of length K) or C = A B (B row-major K x N)
// at (m0, n0). Wavefront w computes rows 32 (w >> 1) .., columns 32 (w & 1) ..: acc[i][j][v] is
// C(m0 + 32 (w >> 1) + 16 i + 4 (lane >> 4) + v, n0 + 32 (w & 1) + 16 j + (lane & 15)).
template <bool BT>
__device__ __forceinline__ void gemm_tile(const float* __restrict__ A, int lda, const float* __restrict__ B, int ldb,
                                          int M, int N, int K, int m0, int n0, floatx4 (&acc)[2][2], float* As,
                                          float* Bs) {
  const int tx = threadIdx.x, lane = tx & 63, w = tx >> 6, wm = w >> 1, wn = w & 1;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // this thread's share of a step's tiles: A row ra, k ca .. ca + TK / 4 - 1; B the same ([n][k]), or k
  // rows rb and rb + 32, columns cb .. cb + 7 ([k][n])
  constexpr int NF = TK / 4;
  const int ra = tx >> 2, ca = (tx & 3) * NF;
  const int rb = BT ? ra : tx >> 3, cb = BT ? ca : (tx & 7) * 8;
  float fa[TK / 4], fb[TK / 4];
  auto fetch = [&](int k0) {
    const int gm = m0 + ra;
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      const int gk = k0 + ca + j;
      fa[j] = gm < M && gk < K ? A[(size_t)gm * lda + gk] : 0.f;
    }
    if (BT) {
      const int gn = n0 + rb;
#pragma unroll
      for (int j = 0; j < NF; ++j) {
        const int gk = k0 + cb + j;
        fb[j] = gn < N && gk < K ? B[(size_t)gn * ldb + gk] : 0.f;
      }
    } else {
#pragma unroll
      for (int h = 0; h < NF / 8; ++h) {
        const int gk = k0 + rb + 32 * h;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int gn = n0 + cb + j;
          fb[8 * h + j] = gk < K && gn < N ? B[(size_t)gk * ldb + gn] : 0.f;
        }
      }
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += TK) {
#pragma unroll
    for (int j = 0; j < NF; ++j) As[ra * AS + ca + j] = fa[j];
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      if (BT) Bs[rb * AS + cb + j] = fb[j];
      else Bs[(rb + 32 * (j / 8)) * BS + cb + (j % 8)] = fb[j];
    }
    __syncthreads();
    if (k0 + TK < K) fetch(k0 + TK);  // the next step's loads, in flight under this step's products
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[(32 * wm + 16 * i + li) * AS + kk + lk];
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = BT ? Bs[(32 * wn + 16 * j + li) * AS + kk + lk] : Bs[(kk + lk) * BS + 32 * wn + 16 * j + li];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
}

// One 64 x 64 tile of C = A B^T (BT) or A B over the K range [z kc, z kc + kc) of split z = blockIdx.z,
// written to C + z cz (row-major, leading dimension N); the splits are summed in order afterwards.
template <bool BT>
__global__ __launch_bounds__(256) void gemm_part_k(const float* __restrict__ A, int lda, const float* __restrict__ B,
                                                   int ldb, int M, int N, int K, int kc, float* __restrict__ C,
                                                   size_t cz) {
  __shared__ float As[TM * AS], Bs[BT ? TN * AS : TK * BS];
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int kb = blockIdx.z * kc, kn = min(K, kb + kc) - kb;
  floatx4 acc[2][2];
  gemm_tile<BT>(A + kb, lda, BT ? B + kb : B + (size_t)kb * ldb, ldb, M, N, kn, m0, n0, acc, As, Bs);
  float* Cz = C + blockIdx.z * cz;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + 32 * wm + 16 * i + 4 * lk + v, n = n0 + 32 * wn + 16 * j + li;
        if (m < M && n < N) Cz[(size_t)m * N + n] = acc[i][j][v];
      }
}

// One wavefront per site: K(i, t) = the sum of the nz splits (in order), then W(i, t) = delta^2 zeta
// alpha_t K^(zeta-1) written over split 0, and e_i = delta^2 sum_t alpha_t K^zeta.
__global__ __launch_bounds__(256) void gap_weights_k(int n_sites, int n_sparse, int nz, float* __restrict__ Kp,
                                                     size_t cz, const float* __restrict__ alpha, int zeta, float wscale,
                                                     double d2, double* __restrict__ e_site) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= n_sites) return;
  double e = 0.0;
  float* row = Kp + (size_t)i * n_sparse;
  for (int t = lane; t < n_sparse; t += 64) {
    float k = row[t];
    for (int z = 1; z < nz; ++z) k += row[z * cz + t];
    float km1 = 1.f;  // k^(zeta - 1)
    for (int z = 1; z < zeta; ++z) km1 *= k;
    e += (double)alpha[t] * (double)(km1 * k);
    row[t] = wscale * alpha[t] * km1;
  }
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
  if (lane == 0) e_site[i] = d2 * e;
}

// G = the sum of the nz splits of dE/dq (in order)
__global__ void split_sum_k(size_t n, int nz, const float* __restrict__ Gp, size_t cz, float* __restrict__ G) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    float g = Gp[e];
    for (int z = 1; z < nz; ++z) g += Gp[z * cz + e];
    G[e] = g;
  }
}

// total of the site energies (one workgroup, fixed order)
__global__ __launch_bounds__(256) void sum_k(int n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double part[4];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) t += x[i];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) *out = part[0] + part[1] + part[2] + part[3];
}

// splits of the K dimension that bring a launch of `tiles` output tiles to about 1024 workgroups,
// each split at least two steps of K
int k_splits(int tiles, int K) {
  int nz = std::max(1, std::min(8, 1024 / std::max(tiles, 1)));
  while (nz > 1 && (K + nz - 1) / nz < 2 * TK) --nz;
  return nz;
}

int ensure_buf(AdjBuf& b, size_t bytes, std::string& msg) {
  if (b.bytes >= bytes) return ST_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    msg = "hipMalloc failed";
    return ST_ERR_HIP;
  }
  b.bytes = bytes;
  return ST_OK;
}

}  // namespace

int gap_energy_grad(AdjointState& s, hipStream_t st, int n_sites, int n_soap, int n_sparse, const float* T,
                    const float* alpha, int zeta, double delta, double* e_site, double* energy_dev,
                    std::string& msg) {
  if (n_sites < 1) return ST_OK;
  const int mt = (n_sites + TM - 1) / TM;
  // K range per split (whole steps of K), and the splits that leaves
  int kc1 = k_splits(mt * ((n_sparse + TN - 1) / TN), n_soap);
  kc1 = ((n_soap + kc1 - 1) / kc1 + TK - 1) / TK * TK;
  const int nz1 = (n_soap + kc1 - 1) / kc1;
  int kc2 = k_splits(mt * ((n_soap + TN - 1) / TN), n_sparse);
  kc2 = ((n_sparse + kc2 - 1) / kc2 + TK - 1) / TK * TK;
  const int nz2 = (n_sparse + kc2 - 1) / kc2;
  // the splits of K (n_sites x n_sparse, split 0 then holds W), then those of dE/dq when it is split
  const size_t kz = (size_t)n_sites * n_sparse, gz = (size_t)n_sites * n_soap;
  const size_t kbytes = (kz * nz1 * 4 + 255) / 256 * 256;
  int rc;
  if ((rc = ensure_buf(s.gapK, kbytes + (nz2 > 1 ? gz * nz2 * 4 : 0), msg))) return rc;
  float* Kp = s.gapK.as<float>();
  float* Gp = reinterpret_cast<float*>(static_cast<char*>(s.gapK.p) + kbytes);
  gemm_part_k<true><<<dim3((n_sparse + TN - 1) / TN, mt, nz1), 256, 0, st>>>(s.soap.as<float>(), n_soap, T, n_soap,
                                                                            n_sites, n_sparse, n_soap, kc1, Kp, kz);
  gap_weights_k<<<(n_sites + 3) / 4, 256, 0, st>>>(n_sites, n_sparse, nz1, Kp, kz, alpha, zeta,
                                                   (float)(delta * delta * zeta), delta * delta, e_site);
  float* G = s.G32.as<float>();
  gemm_part_k<false><<<dim3((n_soap + TN - 1) / TN, mt, nz2), 256, 0, st>>>(Kp, n_sparse, T, n_soap, n_sites, n_soap,
                                                                           n_sparse, kc2, nz2 > 1 ? Gp : G, gz);
  if (nz2 > 1) split_sum_k<<<(int)std::min<size_t>((gz + 255) / 256, 4096), 256, 0, st>>>(gz, nz2, Gp, gz, G);
  sum_k<<<1, 256, 0, st>>>(n_sites, e_site, energy_dev);
  if (hipGetLastError() != hipSuccess) {
    msg = "GAP model launch failed";
    return ST_ERR_HIP;
  }
  return ST_OK;
}

}  // namespace st
