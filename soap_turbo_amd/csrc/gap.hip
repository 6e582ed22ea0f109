// GAP-form energy and its gradient with respect to the descriptors, on the FP32 descriptors the
// site kernel leaves resident: E = sum_i delta^2 sum_t alpha_t (q_i . q_t)^zeta over sparse
// points q_t (the SOAP kernel of a Gaussian-approximation potential; the energy model that
// supplies get_soap's callers with dE/dq). K = Q T^T and dE/dq = W T are plain FP32 GEMMs
// (rocBLAS); the elementwise step between them forms W = delta^2 zeta alpha_t K^(zeta-1) and the
// site energies.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "adjoint.h"

namespace st {

namespace {

// one wavefront per site: W(i, t) over K(i, t) in place, and e_i = delta^2 sum_t alpha_t K^zeta
__global__ __launch_bounds__(256) void gap_weights_k(int n_sites, int n_sparse, float* __restrict__ K,
                                                     const float* __restrict__ alpha, int zeta, double d2,
                                                     double* __restrict__ e_site) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= n_sites) return;
  double e = 0.0;
  float* row = K + (size_t)i * n_sparse;
  for (int t = lane; t < n_sparse; t += 64) {
    const float k = row[t];
    float km1 = 1.f;  // k^(zeta - 1)
    for (int z = 1; z < zeta; ++z) km1 *= k;
    e += (double)alpha[t] * (double)(km1 * k);
    row[t] = (float)(d2 * zeta) * alpha[t] * km1;
  }
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
  if (lane == 0) e_site[i] = d2 * e;
}

// total of the site energies (one workgroup)
__global__ __launch_bounds__(256) void sum_k(int n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double part[4];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) t += x[i];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) *out = part[0] + part[1] + part[2] + part[3];
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

int gap_energy_grad(AdjointState& s, hipStream_t st, void* blas, int n_sites, int n_soap, int n_sparse, const float* T,
                    const float* alpha, int zeta, double delta, double* e_site, double* energy_dev,
                    std::string& msg) {
  rocblas_handle h = static_cast<rocblas_handle>(blas);
  int rc;
  if ((rc = ensure_buf(s.gapK, (size_t)n_sites * n_sparse * 4, msg))) return rc;
  float* K = s.gapK.as<float>();
  const float one = 1.f, zero = 0.f;
  if (rocblas_set_stream(h, st) != rocblas_status_success) {
    msg = "rocblas_set_stream failed";
    return ST_ERR_HIP;
  }
  // K(n_sparse, n_sites) column-major = T Q^T with T (n_soap, n_sparse) and Q (n_soap, n_sites)
  // column-major: the row-major site-by-sparse matrix of dot products
  if (rocblas_sgemm(h, rocblas_operation_transpose, rocblas_operation_none, n_sparse, n_sites, n_soap, &one, T, n_soap,
                    s.soap.as<float>(), n_soap, &zero, K, n_sparse) != rocblas_status_success) {
    msg = "rocblas_sgemm (kernel values) failed";
    return ST_ERR_HIP;
  }
  gap_weights_k<<<(n_sites + 3) / 4, 256, 0, st>>>(n_sites, n_sparse, K, alpha, zeta, delta * delta, e_site);
  // dE/dq (n_soap, n_sites) column-major = T (n_soap, n_sparse) W (n_sparse, n_sites): site-major rows
  if (rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, n_soap, n_sites, n_sparse, &one, T, n_soap, K,
                    n_sparse, &zero, s.G32.as<float>(), n_soap) != rocblas_status_success) {
    msg = "rocblas_sgemm (descriptor gradient) failed";
    return ST_ERR_HIP;
  }
  sum_k<<<1, 256, 0, st>>>(n_sites, e_site, energy_dev);
  return ST_OK;
}

}  // namespace st
