// Forces and virial from pair gradients, and the FP32 adjoint path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "adjoint.h"

namespace st {

namespace {

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

__global__ void zero_k(double* x, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0.0;
}

// forces(:, a) -= pair(:, k) for every pair k whose neighbour is atom a;
// per-block partial virials -sum_k vec_k (x) pair_k.
template <class T>
__global__ __launch_bounds__(256) void scatter_k(int n_pairs, const T* __restrict__ pair,
                                                 const int32_t* __restrict__ pair_atom, const double* __restrict__ vec,
                                                 const double* __restrict__ rjs, const double* __restrict__ thetas,
                                                 const double* __restrict__ phis, double* __restrict__ forces,
                                                 double* __restrict__ vpart) {
  __shared__ double red[9][4];
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n_pairs; k += gridDim.x * blockDim.x) {
    const double g0 = pair[3 * (size_t)k], g1 = pair[3 * (size_t)k + 1], g2 = pair[3 * (size_t)k + 2];
    const int a = pair_atom[k];
    atomicAdd(forces + 3 * (size_t)a, -g0);
    atomicAdd(forces + 3 * (size_t)a + 1, -g1);
    atomicAdd(forces + 3 * (size_t)a + 2, -g2);
    double x, y, z;
    if (vec) {
      x = vec[3 * (size_t)k]; y = vec[3 * (size_t)k + 1]; z = vec[3 * (size_t)k + 2];
    } else {
      const double r = rjs[k], st_ = sin(thetas[k]);
      x = r * st_ * cos(phis[k]); y = r * st_ * sin(phis[k]); z = r * cos(thetas[k]);
    }
    const double r3[3] = {x, y, z}, g[3] = {g0, g1, g2};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) v[3 * i + j] -= r3[i] * g[j];
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    double t = v[q];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) red[q][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double t = 0.0;
    for (int j = 0; j < 4; ++j) t += red[threadIdx.x][j];
    vpart[(size_t)blockIdx.x * 9 + threadIdx.x] = t;
  }
}

// one wavefront per virial component, lanes striding over the block partials
__global__ __launch_bounds__(576) void virial_sum_k(int nb, const double* __restrict__ vpart,
                                                    double* __restrict__ virial) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double t = 0.0;
  for (int b = lane; b < nb; b += 64) t += vpart[(size_t)b * 9 + q];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if (lane == 0) virial[q] = t;
}

template <class T>
void scatter(hipStream_t st, int n_pairs, int n_atoms, const T* pair, const int32_t* pair_atom, const double* vec,
             const double* rjs, const double* thetas, const double* phis, double* forces, double* vpart,
             double* virial) {
  zero_k<<<256, 256, 0, st>>>(forces, (size_t)n_atoms * 3);
  int nb = (n_pairs + 255) / 256;
  if (nb > SCATTER_BLOCKS) nb = SCATTER_BLOCKS;
  if (nb < 1) nb = 1;
  scatter_k<T><<<nb, 256, 0, st>>>(n_pairs, pair, pair_atom, vec, rjs, thetas, phis, forces, vpart);
  virial_sum_k<<<1, 576, 0, st>>>(nb, vpart, virial);
}

int download_forces(AdjointState& s, hipStream_t st, int n_atoms, double* forces, double* virial, std::string& msg) {
  if (forces && hipMemcpyAsync(forces, s.forces.p, (size_t)n_atoms * 24, hipMemcpyDeviceToHost, st) != hipSuccess) {
    msg = "copy of forces failed";
    return ST_ERR_HIP;
  }
  if (virial && hipMemcpyAsync(virial, s.virial.p, 72, hipMemcpyDeviceToHost, st) != hipSuccess) {
    msg = "copy of virial failed";
    return ST_ERR_HIP;
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    msg = "stream synchronisation failed";
    return ST_ERR_HIP;
  }
  return ST_OK;
}

}  // namespace

void launch_scatter_f64(hipStream_t st, int n_pairs, int n_atoms, const double* pair, const int32_t* pair_atom,
                        const double* vec, const double* rjs, const double* thetas, const double* phis,
                        double* forces, double* vpart, double* virial) {
  scatter<double>(st, n_pairs, n_atoms, pair, pair_atom, vec, rjs, thetas, phis, forces, vpart, virial);
}

void launch_scatter_f32(hipStream_t st, int n_pairs, int n_atoms, const float* pair, const int32_t* pair_atom,
                        const double* vec, const double* rjs, const double* thetas, const double* phis,
                        double* forces, double* vpart, double* virial) {
  scatter<float>(st, n_pairs, n_atoms, pair, pair_atom, vec, rjs, thetas, phis, forces, vpart, virial);
}

int adjoint_download_forces(AdjointState& s, hipStream_t st, int n_atoms, double* forces, double* virial,
                            std::string& msg) {
  if (!s.have_forces) {
    msg = "no FP32 forces computed";
    return ST_ERR_STATE;
  }
  return download_forces(s, st, n_atoms, forces, virial, msg);
}

void adjoint_invalidate(AdjointState& s) {
  s.have_out = false;
  s.have_forces = false;
}

void adjoint_release(AdjointState& s) {
  AdjBuf* bufs[] = {&s.rtab, &s.atab, &s.tabinfo, &s.cnk, &s.soap, &s.sdp, &s.H, &s.pair, &s.forces,
                    &s.virial_part, &s.virial, &s.G32, &s.order, &s.atom_ptr, &s.atom_pairs, &s.vecbuf, &s.bins, &s.el32};
  for (AdjBuf* b : bufs)
    if (b->p) {
      (void)hipFree(b->p);
      b->p = nullptr;
      b->bytes = 0;
    }
}

}  // namespace st
