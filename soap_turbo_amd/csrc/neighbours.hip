// Device neighbour-list builder: periodic cubic box, cell list, and the
// get_soap inputs in the layout the reference's callers pass
// (src/soap_turbo.f90:139-169): per site its own atom first, then every other
// atom j with |r_j - r_i| < rcut under the minimum image, in increasing j;
// rjs, thetas = acos(z/r), phis = atan2(y, x); mask(k, s) = (species of the
// neighbour == s), as assign_species_multiplicity reads it
// (src/soap_turbo.f90:2188-2214). Also the pairs of each atom (CSR, increasing
// pair index) that the force gather reads.
//
// Pair vectors are formed with the same operations as the host generator, in
// FP64 without contraction, so the integer outputs and r_j - r_i are bit-exact.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"

#pragma clang fp contract(off)

namespace st {

namespace {

constexpr int NB_CAP = 1024;  // neighbours one wavefront collects for a site (or pairs for an atom)

__device__ __forceinline__ double wrap_d(double d, double L) { return d - L * rint(d / L); }

__device__ __forceinline__ int cell_coord(double x, double L, int nc) {
  double u = x - L * floor(x / L);  // into [0, L)
  int c = (int)(u * nc / L);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void nb_cells_k(int n_atoms, const double* __restrict__ pos, double L, int nc, int* __restrict__ cell_of,
                           int* __restrict__ cell_count) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_atoms) return;
  const int cx = cell_coord(pos[3 * a], L, nc), cy = cell_coord(pos[3 * a + 1], L, nc),
            cz = cell_coord(pos[3 * a + 2], L, nc);
  const int c = (cz * nc + cy) * nc + cx;
  cell_of[a] = c;
  atomicAdd(cell_count + c, 1);
}

// exclusive prefix sum of n counts into out[0..n] (out[n] = total), one workgroup of 1024
__global__ __launch_bounds__(1024) void scan_k(int n, const int* __restrict__ in, int* __restrict__ out) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024, b = t * chunk, e = min(n, b + chunk);
  int s = 0;
  for (int i = b; i < e; ++i) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = b; i < e; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (t == 1023) out[n] = part[1023];
}

__global__ void nb_bin_k(int n_atoms, const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                         int* __restrict__ cursor, int* __restrict__ cell_atoms) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_atoms) return;
  const int c = cell_of[a];
  cell_atoms[cell_start[c] + atomicAdd(cursor + c, 1)] = a;
}

// Visits the candidate atoms of site atom i (the 27 cells around it, or every atom when nc = 1),
// 64 at a time: every lane calls f(ok, j, dx, dy, dz) in step, ok when its candidate j != i has
// |d| < rcut, d = minimum-image r_j - r_i. One wavefront.
template <class F>
__device__ __forceinline__ void for_neighbours(int i, const double* __restrict__ pos, double L, double rc2, int nc,
                                               const int* __restrict__ cell_start,
                                               const int* __restrict__ cell_atoms, F&& f) {
  const int lane = threadIdx.x & 63;
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const int ncell = nc == 1 ? 1 : 27;
  const int cx = cell_coord(xi, L, nc), cy = cell_coord(yi, L, nc), cz = cell_coord(zi, L, nc);
  for (int q = 0; q < ncell; ++q) {
    int c = 0;
    if (nc > 1) {
      const int ex = (cx + q % 3 - 1 + nc) % nc, ey = (cy + (q / 3) % 3 - 1 + nc) % nc,
                ez = (cz + q / 9 - 1 + nc) % nc;
      c = (ez * nc + ey) * nc + ex;
    }
    const int b = cell_start[c], e = cell_start[c + 1];
    for (int m0 = b; m0 < e; m0 += 64) {
      const int m = m0 + lane;
      bool ok = false;
      int j = -1;
      double dx = 0.0, dy = 0.0, dz = 0.0;
      if (m < e) {
        j = cell_atoms[m];
        if (j != i) {
          dx = wrap_d(pos[3 * j] - xi, L);
          dy = wrap_d(pos[3 * j + 1] - yi, L);
          dz = wrap_d(pos[3 * j + 2] - zi, L);
          ok = dx * dx + dy * dy + dz * dz < rc2;
        }
      }
      f(ok, j, dx, dy, dz);
    }
  }
}

// n_neigh[s] = 1 + neighbours of site s; one wavefront per site
__global__ __launch_bounds__(256) void nb_count_k(int n_sites, const int* __restrict__ site_atom,
                                                  const double* __restrict__ pos, double L, double rc2, int nc,
                                                  const int* __restrict__ cell_start,
                                                  const int* __restrict__ cell_atoms, int* __restrict__ n_neigh) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= n_sites) return;
  int cnt = 0;
  for_neighbours(site_atom[s], pos, L, rc2, nc, cell_start, cell_atoms,
                 [&](bool ok, int, double, double, double) { cnt += ok; });
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) n_neigh[s] = 1 + cnt;
}

// The pairs of site s: collected in LDS, ranked by atom index, written in that order after the
// site's own pair. Sites with more than NB_CAP neighbours set *overflow.
__global__ __launch_bounds__(256) void nb_fill_k(int n_sites, int n_pairs, int n_species,
                                                 const int* __restrict__ site_atom, const int* __restrict__ species,
                                                 const double* __restrict__ pos, double L, double rc2, int nc,
                                                 const int* __restrict__ cell_start,
                                                 const int* __restrict__ cell_atoms, const int* __restrict__ site_ptr,
                                                 int* __restrict__ pair_atom, double* __restrict__ vec,
                                                 double* __restrict__ rjs, double* __restrict__ thetas,
                                                 double* __restrict__ phis, int32_t* __restrict__ mask,
                                                 int* __restrict__ overflow) {
  __shared__ int js[4][NB_CAP];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + w;
  if (s >= n_sites) return;
  const int i = site_atom[s];
  int* list = js[w];
  int n = 0;  // count of collected neighbours, the same in every lane
  for_neighbours(i, pos, L, rc2, nc, cell_start, cell_atoms, [&](bool ok, int j, double, double, double) {
    // appended in lane order: each passing lane after the passing lanes below it
    const unsigned long long act = __ballot(ok);
    const int below = __popcll(act & ((1ull << lane) - 1ull));
    if (ok && n + below < NB_CAP) list[n + below] = j;
    n += __popcll(act);
  });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int k0 = site_ptr[s];
  if (n > NB_CAP || k0 + 1 + n != site_ptr[s + 1]) {
    if (lane == 0) atomicOr(overflow, 1);
    return;
  }
  const int si = species[i] - 1;
  if (lane == 0) {
    pair_atom[k0] = i;
    vec[3 * (size_t)k0] = vec[3 * (size_t)k0 + 1] = vec[3 * (size_t)k0 + 2] = 0.0;
    rjs[k0] = thetas[k0] = phis[k0] = 0.0;
    for (int sp = 0; sp < n_species; ++sp) mask[(size_t)sp * n_pairs + k0] = si == sp;
  }
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  for (int m = lane; m < n; m += 64) {
    const int j = list[m];
    int rank = 0;
    for (int q = 0; q < n; ++q) rank += list[q] < j;
    const size_t k = (size_t)k0 + 1 + rank;
    const double x = wrap_d(pos[3 * j] - xi, L), y = wrap_d(pos[3 * j + 1] - yi, L), z = wrap_d(pos[3 * j + 2] - zi, L);
    const double r = sqrt(x * x + y * y + z * z);
    double c = z / r;
    c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
    pair_atom[k] = j;
    vec[3 * k] = x;
    vec[3 * k + 1] = y;
    vec[3 * k + 2] = z;
    rjs[k] = r;
    thetas[k] = acos(c);
    phis[k] = atan2(y, x);
    const int sj = species[j] - 1;
    for (int sp = 0; sp < n_species; ++sp) mask[(size_t)sp * n_pairs + k] = sj == sp;
  }
}

__global__ void site_species_k(int n_sites, const int* __restrict__ site_atom, const int* __restrict__ species,
                               int32_t* __restrict__ site_species, int32_t* __restrict__ mult) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sites) return;
  site_species[s] = species[site_atom[s]];
  mult[s] = 1;
}

__global__ void atom_count_k(int n_pairs, const int* __restrict__ pair_atom, int* __restrict__ count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_pairs) atomicAdd(count + pair_atom[k], 1);
}

__global__ void atom_bin_k(int n_pairs, const int* __restrict__ pair_atom, const int* __restrict__ atom_ptr,
                           int* __restrict__ cursor, int* __restrict__ atom_pairs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_pairs) {
    const int a = pair_atom[k];
    atom_pairs[atom_ptr[a] + atomicAdd(cursor + a, 1)] = k;
  }
}

// each atom's pairs into increasing pair index (rank by comparison; one wavefront per atom)
__global__ __launch_bounds__(256) void atom_sort_k(int n_atoms, const int* __restrict__ atom_ptr,
                                                   int* __restrict__ atom_pairs, int* __restrict__ overflow) {
  __shared__ int ks[4][NB_CAP];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int a = blockIdx.x * 4 + w;
  if (a >= n_atoms) return;
  const int b = atom_ptr[a], n = atom_ptr[a + 1] - b;
  if (n > NB_CAP) {
    if (lane == 0) atomicOr(overflow, 2);
    return;
  }
  int* list = ks[w];
  for (int m = lane; m < n; m += 64) list[m] = atom_pairs[b + m];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int m = lane; m < n; m += 64) {
    const int k = list[m];
    int rank = 0;
    for (int q = 0; q < n; ++q) rank += list[q] < k;
    atom_pairs[b + rank] = k;
  }
}

}  // namespace

int neighbour_cells(double L, double rcut) {
  int nc = (int)floor(L / rcut);
  if (nc < 3) nc = 1;  // a 27-cell stencil needs three cells a side; otherwise one cell of every atom
  if (nc > 128) nc = 128;
  return nc;
}

void launch_neighbour_bins(hipStream_t st, int n_atoms, const double* pos, double L, int nc, int* cell_of,
                           int* cell_count, int* cell_start, int* cursor, int* cell_atoms) {
  const int n_cells = nc * nc * nc;
  (void)hipMemsetAsync(cell_count, 0, (size_t)n_cells * 4, st);
  (void)hipMemsetAsync(cursor, 0, (size_t)n_cells * 4, st);
  const int nb = (n_atoms + 255) / 256;
  nb_cells_k<<<nb, 256, 0, st>>>(n_atoms, pos, L, nc, cell_of, cell_count);
  scan_k<<<1, 1024, 0, st>>>(n_cells, cell_count, cell_start);
  nb_bin_k<<<nb, 256, 0, st>>>(n_atoms, cell_of, cell_start, cursor, cell_atoms);
}

void launch_neighbour_count(hipStream_t st, int n_sites, const int* site_atom, const double* pos, double L,
                            double rcut, int nc, const int* cell_start, const int* cell_atoms, int* n_neigh,
                            int* site_ptr) {
  if (n_sites < 1) return;
  nb_count_k<<<(n_sites + 3) / 4, 256, 0, st>>>(n_sites, site_atom, pos, L, rcut * rcut, nc, cell_start, cell_atoms,
                                                 n_neigh);
  scan_k<<<1, 1024, 0, st>>>(n_sites, n_neigh, site_ptr);
}

void launch_neighbour_fill(hipStream_t st, int n_sites, int n_pairs, int n_species, const int* site_atom,
                           const int* species, const double* pos, double L, double rcut, int nc, const int* cell_start,
                           const int* cell_atoms, const int* site_ptr, int* pair_atom, double* vec, double* rjs,
                           double* thetas, double* phis, int32_t* mask, int32_t* site_species, int32_t* mult,
                           int* overflow) {
  if (n_sites < 1) return;
  nb_fill_k<<<(n_sites + 3) / 4, 256, 0, st>>>(n_sites, n_pairs, n_species, site_atom, species, pos, L, rcut * rcut,
                                                nc, cell_start, cell_atoms, site_ptr, pair_atom, vec, rjs, thetas,
                                                phis, mask, overflow);
  site_species_k<<<(n_sites + 255) / 256, 256, 0, st>>>(n_sites, site_atom, species, site_species, mult);
}

void launch_atom_pairs(hipStream_t st, int n_atoms, int n_pairs, const int* pair_atom, int* count, int* cursor,
                       int* atom_ptr, int* atom_pairs, int* overflow) {
  (void)hipMemsetAsync(count, 0, (size_t)n_atoms * 4, st);
  (void)hipMemsetAsync(cursor, 0, (size_t)n_atoms * 4, st);
  if (n_pairs > 0) atom_count_k<<<(n_pairs + 255) / 256, 256, 0, st>>>(n_pairs, pair_atom, count);
  scan_k<<<1, 1024, 0, st>>>(n_atoms, count, atom_ptr);
  if (n_pairs > 0) atom_bin_k<<<(n_pairs + 255) / 256, 256, 0, st>>>(n_pairs, pair_atom, atom_ptr, cursor, atom_pairs);
  atom_sort_k<<<(n_atoms + 3) / 4, 256, 0, st>>>(n_atoms, atom_ptr, atom_pairs, overflow);
}

}  // namespace st
