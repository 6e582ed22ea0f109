// Device neighbour-list builder: periodic cell (cubic, or general lattice
// vectors), cell list, and the get_soap inputs in the layout the reference's callers pass
// (src/soap_turbo.f90:139-169): per site its own atom first, then every other
// atom j with |r_j - r_i| < rcut under the minimum image, in increasing j;
// rjs, thetas = acos(z/r), phis = atan2(y, x); mask(k, s) = (species of the
// neighbour == s), as assign_species_multiplicity reads it
// (src/soap_turbo.f90:2188-2214). Also the pairs of each atom (CSR, increasing
// pair index) that the force gather reads.
//
// Pair vectors are formed with the same operations as the host generator, in
// FP64 without contraction, so the integer outputs and r_j - r_i are bit-exact.
// A general cell takes the image d - H rint(H^-1 d) of d = r_j - r_i (H: the
// lattice vectors as columns); with every width of the cell (the distance
// between opposite faces) above 2 rcut, that is the one image inside the
// cutoff. Atoms are binned on their fractional coordinates, each cell at least
// rcut wide across its faces, so the 27 cells around an atom hold every
// neighbour.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// the minimum image of d = r_j - r_i, in place
__device__ __forceinline__ void min_image(const Lattice& g, double& x, double& y, double& z) {
  if (!g.general) {
    x = wrap_d(x, g.L);
    y = wrap_d(y, g.L);
    z = wrap_d(z, g.L);
    return;
  }
  const double n0 = rint(g.hi[0] * x + g.hi[1] * y + g.hi[2] * z);
  const double n1 = rint(g.hi[3] * x + g.hi[4] * y + g.hi[5] * z);
  const double n2 = rint(g.hi[6] * x + g.hi[7] * y + g.hi[8] * z);
  x = x - (g.h[0] * n0 + g.h[1] * n1 + g.h[2] * n2);
  y = y - (g.h[3] * n0 + g.h[4] * n1 + g.h[5] * n2);
  z = z - (g.h[6] * n0 + g.h[7] * n1 + g.h[8] * n2);
}

// the cell of a position, per lattice direction
__device__ __forceinline__ void cell_coords(const Lattice& g, double x, double y, double z, int& cx, int& cy, int& cz) {
  if (!g.general) {
    cx = cell_coord(x, g.L, g.n[0]);
    cy = cell_coord(y, g.L, g.n[1]);
    cz = cell_coord(z, g.L, g.n[2]);
    return;
  }
  int c[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const double f = g.hi[3 * v] * x + g.hi[3 * v + 1] * y + g.hi[3 * v + 2] * z;
    const int k = (int)((f - floor(f)) * g.n[v]);
    c[v] = k < 0 ? 0 : (k >= g.n[v] ? g.n[v] - 1 : k);
  }
  cx = c[0];
  cy = c[1];
  cz = c[2];
}

__global__ void nb_cells_k(int n_atoms, const double* __restrict__ pos, const Lattice g, int* __restrict__ cell_of,
                           int* __restrict__ cell_count) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_atoms) return;
  int cx, cy, cz;
  cell_coords(g, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], cx, cy, cz);
  const int c = (cz * g.n[1] + cy) * g.n[0] + cx;
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

// Visits the candidate atoms of site atom i (the 27 cells around it, or every atom with a single
// cell), 64 at a time: every lane calls f(ok, j, dx, dy, dz) in step, ok when its candidate j != i
// has |d| < rcut, d = minimum-image r_j - r_i. One wavefront.
template <class F>
__device__ __forceinline__ void for_neighbours(int i, const double* __restrict__ pos, const Lattice& g, double rc2,
                                               const int* __restrict__ cell_start,
                                               const int* __restrict__ cell_atoms, F&& f) {
  const int lane = threadIdx.x & 63;
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const bool one = nx == 1;  // a single cell holds every atom (then ny = nz = 1)
  const int ncell = one ? 1 : 27;
  int cx, cy, cz;
  cell_coords(g, xi, yi, zi, cx, cy, cz);
  for (int q = 0; q < ncell; ++q) {
    int c = 0;
    if (!one) {
      const int ex = (cx + q % 3 - 1 + nx) % nx, ey = (cy + (q / 3) % 3 - 1 + ny) % ny,
                ez = (cz + q / 9 - 1 + nz) % nz;
      c = (ez * ny + ey) * nx + ex;
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
          dx = pos[3 * j] - xi;
          dy = pos[3 * j + 1] - yi;
          dz = pos[3 * j + 2] - zi;
          min_image(g, dx, dy, dz);
          ok = dx * dx + dy * dy + dz * dz < rc2;
        }
      }
      f(ok, j, dx, dy, dz);
    }
  }
}

// n_neigh[s] = 1 + neighbours of site s; one wavefront per site
__global__ __launch_bounds__(256) void nb_count_k(int n_sites, const int* __restrict__ site_atom,
                                                  const double* __restrict__ pos, const Lattice g, double rc2,
                                                  const int* __restrict__ cell_start,
                                                  const int* __restrict__ cell_atoms, int* __restrict__ n_neigh) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= n_sites) return;
  int cnt = 0;
  for_neighbours(site_atom[s], pos, g, rc2, cell_start, cell_atoms,
                 [&](bool ok, int, double, double, double) { cnt += ok; });
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) n_neigh[s] = 1 + cnt;
}

// The pairs of site s: collected in LDS, ranked by atom index, written in that order after the
// site's own pair. Sites with more than NB_CAP neighbours set *overflow.
__global__ __launch_bounds__(256) void nb_fill_k(int n_sites, int n_pairs, int n_species,
                                                 const int* __restrict__ site_atom, const int* __restrict__ species,
                                                 const double* __restrict__ pos, const Lattice g, double rc2,
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
  for_neighbours(i, pos, g, rc2, cell_start, cell_atoms, [&](bool ok, int j, double, double, double) {
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
    double x = pos[3 * j] - xi, y = pos[3 * j + 1] - yi, z = pos[3 * j + 2] - zi;
    min_image(g, x, y, z);
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

Lattice cubic_lattice(double L, double rcut) {
  Lattice g{};
  g.L = L;
  int nc = (int)floor(L / rcut);
  if (nc < 3) nc = 1;  // a 27-cell stencil needs three cells a side; otherwise one cell of every atom
  if (nc > 128) nc = 128;
  g.n[0] = g.n[1] = g.n[2] = nc;
  return g;
}

bool general_lattice(const double* cell, double rcut, Lattice& g, double& min_width) {
  g = Lattice{};
  g.general = 1;
  // h[3c + v]: component c of lattice vector v (cell[c + 3v], the vectors as columns)
  for (int c = 0; c < 3; ++c)
    for (int v = 0; v < 3; ++v) g.h[3 * c + v] = cell[c + 3 * v];
  const double* a = cell;
  const double* b = cell + 3;
  const double* d = cell + 6;
  // rows of H^-1 = (b x c, c x a, a x b) / det
  const double bc[3] = {b[1] * d[2] - b[2] * d[1], b[2] * d[0] - b[0] * d[2], b[0] * d[1] - b[1] * d[0]};
  const double ca[3] = {d[1] * a[2] - d[2] * a[1], d[2] * a[0] - d[0] * a[2], d[0] * a[1] - d[1] * a[0]};
  const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
  if (!(det > 0.0)) return false;  // a right-handed, non-degenerate cell
  const double* rows[3] = {bc, ca, ab};
  min_width = 1e300;
  for (int v = 0; v < 3; ++v) {
    const double* r = rows[v];
    for (int c = 0; c < 3; ++c) g.hi[3 * v + c] = r[c] / det;
    // the distance between the faces spanned by the other two vectors: det / |their cross product|
    const double w = det / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    min_width = std::min(min_width, w);
    int nv = (int)floor(w / rcut);
    if (nv > 128) nv = 128;
    g.n[v] = nv;
  }
  if (g.n[0] < 3 || g.n[1] < 3 || g.n[2] < 3) g.n[0] = g.n[1] = g.n[2] = 1;
  return true;
}

void launch_neighbour_bins(hipStream_t st, int n_atoms, const double* pos, const Lattice& g, int* cell_of,
                           int* cell_count, int* cell_start, int* cursor, int* cell_atoms) {
  const int n_cells = g.n[0] * g.n[1] * g.n[2];
  (void)hipMemsetAsync(cell_count, 0, (size_t)n_cells * 4, st);
  (void)hipMemsetAsync(cursor, 0, (size_t)n_cells * 4, st);
  const int nb = (n_atoms + 255) / 256;
  nb_cells_k<<<nb, 256, 0, st>>>(n_atoms, pos, g, cell_of, cell_count);
  scan_k<<<1, 1024, 0, st>>>(n_cells, cell_count, cell_start);
  nb_bin_k<<<nb, 256, 0, st>>>(n_atoms, cell_of, cell_start, cursor, cell_atoms);
}

void launch_neighbour_count(hipStream_t st, int n_sites, const int* site_atom, const double* pos, const Lattice& g,
                            double rcut, const int* cell_start, const int* cell_atoms, int* n_neigh, int* site_ptr) {
  if (n_sites < 1) return;
  nb_count_k<<<(n_sites + 3) / 4, 256, 0, st>>>(n_sites, site_atom, pos, g, rcut * rcut, cell_start, cell_atoms,
                                                 n_neigh);
  scan_k<<<1, 1024, 0, st>>>(n_sites, n_neigh, site_ptr);
}

void launch_neighbour_fill(hipStream_t st, int n_sites, int n_pairs, int n_species, const int* site_atom,
                           const int* species, const double* pos, const Lattice& g, double rcut,
                           const int* cell_start, const int* cell_atoms, const int* site_ptr, int* pair_atom,
                           double* vec, double* rjs, double* thetas, double* phis, int32_t* mask,
                           int32_t* site_species, int32_t* mult, int* overflow) {
  if (n_sites < 1) return;
  nb_fill_k<<<(n_sites + 3) / 4, 256, 0, st>>>(n_sites, n_pairs, n_species, site_atom, species, pos, g, rcut * rcut,
                                                cell_start, cell_atoms, site_ptr, pair_atom, vec, rjs, thetas, phis,
                                                mask, overflow);
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
