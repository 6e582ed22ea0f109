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

// forces(:, a) = -sum of pair(:, k) over the pairs k whose neighbour is atom a
// (atom_ptr/atom_pairs: those pairs per atom, in increasing k), gathered by 16
// lanes per atom, so no atomics and no clearing pass; each workgroup also
// leaves its partial virial -sum_k vec_k (x) pair_k over the pairs of its atoms.
// (Summing the partials in the last workgroup to finish needs a device-scope
// release per workgroup, which on this GPU costs more than a second launch.)
template <class T>
__global__ __launch_bounds__(256) void gather_k(int n_atoms, const T* __restrict__ pair,
                                                const int* __restrict__ atom_ptr, const int* __restrict__ atom_pairs,
                                                const double* __restrict__ vec, const double* __restrict__ rjs,
                                                const double* __restrict__ thetas, const double* __restrict__ phis,
                                                double* __restrict__ forces, double* __restrict__ vpart) {
  __shared__ double red[9][16];
  const int g = threadIdx.x >> 4, li = threadIdx.x & 15;   // 16 atom groups per workgroup
  const int a = blockIdx.x * 16 + g;
  double f[3] = {0.0, 0.0, 0.0}, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (a < n_atoms) {
    for (int e = atom_ptr[a] + li; e < atom_ptr[a + 1]; e += 16) {
      const int k = atom_pairs[e];
      const double g0 = pair[3 * (size_t)k], g1 = pair[3 * (size_t)k + 1], g2 = pair[3 * (size_t)k + 2];
      f[0] -= g0; f[1] -= g1; f[2] -= g2;
      double x, y, z;
      if (vec) {
        x = vec[3 * (size_t)k]; y = vec[3 * (size_t)k + 1]; z = vec[3 * (size_t)k + 2];
      } else {
        const double r = rjs[k], st_ = sin(thetas[k]);
        x = r * st_ * cos(phis[k]); y = r * st_ * sin(phis[k]); z = r * cos(thetas[k]);
      }
      const double r3[3] = {x, y, z}, gg[3] = {g0, g1, g2};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) v[3 * i + j] -= r3[i] * gg[j];
    }
  }
  // sums over the 16 lanes of each atom group
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] += __shfl_xor(f[c], o, 16);
#pragma unroll
    for (int q = 0; q < 9; ++q) v[q] += __shfl_xor(v[q], o, 16);
  }
  if (li == 0) {
    if (a < n_atoms) {
      forces[3 * (size_t)a] = f[0];
      forces[3 * (size_t)a + 1] = f[1];
      forces[3 * (size_t)a + 2] = f[2];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) red[q][g] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double t = 0.0;
    for (int j = 0; j < 16; ++j) t += red[threadIdx.x][j];
    vpart[(size_t)blockIdx.x * 9 + threadIdx.x] = t;
  }
}

// one wavefront per virial component, lanes striding over the workgroup partials; also sets the
// step's work counter (if any) back to zero
__global__ __launch_bounds__(576) void virial_sum_k(int nb, const double* __restrict__ vpart,
                                                    double* __restrict__ virial, unsigned long long* reset) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (reset && threadIdx.x == 0) *reset = 0ull;
  double t = 0.0;
  for (int b = lane; b < nb; b += 64) t += vpart[(size_t)b * 9 + q];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if (lane == 0) virial[q] = t;
}

// forces(:, a) = -sum of pair(:, k) over the pairs k whose neighbour is atom a, 16 lanes per atom (the
// FP32 path: its virial is summed over the pairs in their own order by virial_pairs_k, so the only reads
// at scattered addresses here are those of the pair gradients)
__global__ __launch_bounds__(256) void forces_k(int n_atoms, const float* __restrict__ pair,
                                                const int* __restrict__ atom_ptr, const int* __restrict__ atom_pairs,
                                                double* __restrict__ forces) {
  const int li = threadIdx.x & 15;
  const int a = blockIdx.x * 16 + (threadIdx.x >> 4);
  double f[3] = {0.0, 0.0, 0.0};
  if (a < n_atoms) {
    const int e1 = atom_ptr[a + 1];
    for (int e = atom_ptr[a] + li; e < e1; e += 16) {
      const int k = atom_pairs[e];
      f[0] -= pair[3 * (size_t)k];
      f[1] -= pair[3 * (size_t)k + 1];
      f[2] -= pair[3 * (size_t)k + 2];
    }
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] += __shfl_xor(f[c], o, 16);
  if (li == 0 && a < n_atoms) {
    forces[3 * (size_t)a] = f[0];
    forces[3 * (size_t)a + 1] = f[1];
    forces[3 * (size_t)a + 2] = f[2];
  }
}

// the partial virials -sum_k vec_k (x) pair_k of a fixed split of the pairs: thread t of workgroup b
// takes pairs b 256 + t, then every nb 256-th (reads in pair order), FP64 sums; one partial per workgroup
__global__ __launch_bounds__(256) void virial_pairs_k(int n_pairs, const float* __restrict__ pair,
                                                      const double* __restrict__ vec, const double* __restrict__ rjs,
                                                      const double* __restrict__ thetas,
                                                      const double* __restrict__ phis, double* __restrict__ vpart) {
  __shared__ double red[9][4];
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < (size_t)n_pairs; k += (size_t)gridDim.x * 256) {
    const double gg[3] = {pair[3 * k], pair[3 * k + 1], pair[3 * k + 2]};
    double r3[3];
    if (vec) {
      r3[0] = vec[3 * k]; r3[1] = vec[3 * k + 1]; r3[2] = vec[3 * k + 2];
    } else {
      const double r = rjs[k], st_ = sin(thetas[k]);
      r3[0] = r * st_ * cos(phis[k]); r3[1] = r * st_ * sin(phis[k]); r3[2] = r * cos(thetas[k]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) v[3 * i + j] -= r3[i] * gg[j];
  }
#pragma unroll
  for (int q = 0; q < 9; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 9; ++q) red[q][threadIdx.x >> 6] = v[q];
  __syncthreads();
  if (threadIdx.x < 9)
    vpart[(size_t)blockIdx.x * 9 + threadIdx.x] =
        red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

template <class T>
void gather(hipStream_t st, int n_atoms, const T* pair, const int* atom_ptr, const int* atom_pairs,
            const double* vec, const double* rjs, const double* thetas, const double* phis, double* forces,
            double* vpart, double* virial, unsigned long long* reset) {
  const int nb = gather_blocks(n_atoms);
  gather_k<T><<<nb, 256, 0, st>>>(n_atoms, pair, atom_ptr, atom_pairs, vec, rjs, thetas, phis, forces, vpart);
  virial_sum_k<<<1, 576, 0, st>>>(nb, vpart, virial, reset);
}

__global__ void counter_reset_k(unsigned long long* c) { *c = 0ull; }

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

void launch_forces_f64(hipStream_t st, int n_atoms, const double* pair, const int* atom_ptr, const int* atom_pairs,
                       const double* vec, const double* rjs, const double* thetas, const double* phis,
                       double* forces, double* vpart, double* virial) {
  gather<double>(st, n_atoms, pair, atom_ptr, atom_pairs, vec, rjs, thetas, phis, forces, vpart, virial, nullptr);
}

void launch_forces_f32(hipStream_t st, int n_atoms, int n_pairs, const float* pair, const int* atom_ptr,
                       const int* atom_pairs, const double* vec, const double* rjs, const double* thetas,
                       const double* phis, double* forces, double* vpart, double* virial, unsigned long long* reset) {
  forces_k<<<gather_blocks(n_atoms), 256, 0, st>>>(n_atoms, pair, atom_ptr, atom_pairs, forces);
  const int nb = virial_blocks(n_atoms, n_pairs);
  virial_pairs_k<<<nb, 256, 0, st>>>(n_pairs, pair, vec, rjs, thetas, phis, vpart);
  virial_sum_k<<<1, 576, 0, st>>>(nb, vpart, virial, reset);
}

void launch_counter_reset(hipStream_t st, unsigned long long* counter) { counter_reset_k<<<1, 1, 0, st>>>(counter); }

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
  s.g32_sites = s.g32_soap = 0;
  s.order_ok = false;
  s.cen_ok = false;
}

void adjoint_release(AdjointState& s) {
  AdjBuf* bufs[] = {&s.rtab, &s.atab, &s.tabinfo, &s.soap, &s.sdp, &s.pair, &s.forces,
                    &s.virial_part, &s.virial, &s.G32, &s.el32, &s.order, &s.tickets, &s.odir, &s.cenv, &s.gapK};
  for (AdjBuf* b : bufs)
    if (b->p) {
      (void)hipFree(b->p);
      b->p = nullptr;
      b->bytes = 0;
    }
}

}  // namespace st
