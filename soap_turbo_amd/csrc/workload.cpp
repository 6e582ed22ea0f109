// Synthetic atomic configurations and neighbour lists in the layout that
// get_soap consumes (src/soap_turbo.f90:139-169): per site, the site itself
// first (r = 0), then every other atom within the cutoff; per pair the
// distance, polar angle and azimuth of the neighbour relative to the site.
//
// Positions come from random sequential addition in a cubic periodic box,
// with slabs in z of independent densities (one slab for a uniform box).
// Neighbour lists are built for a cubic box or for general lattice vectors.
// The random stream is SplitMix64, so the same seed gives the same atoms on
// every host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline double wrap(double d, double L) { return d - L * std::nearbyint(d / L); }

struct CellGrid {
  int nc;
  double cs, L;
  std::vector<int> head, next;
  CellGrid(double L_, double min_cell, int n_atoms) : L(L_) {
    nc = (int)std::floor(L / min_cell);
    if (nc < 1) nc = 1;
    if (nc > 256) nc = 256;
    cs = L / nc;
    head.assign((size_t)nc * nc * nc, -1);
    next.assign(n_atoms, -1);
  }
  int coord(double x) const {
    int c = (int)std::floor(x / cs);
    c %= nc;
    if (c < 0) c += nc;
    return c;
  }
  size_t cell(int cx, int cy, int cz) const { return ((size_t)cz * nc + cy) * nc + cx; }
  void insert(int a, const double* p) {
    size_t c = cell(coord(p[0]), coord(p[1]), coord(p[2]));
    next[a] = head[c];
    head[c] = (int)a;
  }
};

}  // namespace

extern "C" {

// Random sequential addition. n_slabs slabs [zlo, zhi) each receive count[s]
// atoms; a trial point closer than dmin (minimum image) to an accepted atom is
// rejected. Returns the number of atoms placed (< total if max_trials ran out).
int64_t stw_place_atoms(uint64_t seed, double L, double dmin, int n_slabs, const double* zlo,
                        const double* zhi, const int64_t* count, int64_t max_trials_per_atom,
                        double* pos) {
  int64_t total = 0;
  for (int s = 0; s < n_slabs; ++s) total += count[s];
  SplitMix64 rng(seed);
  CellGrid grid(L, dmin > 0 ? dmin : L, (int)total);
  const double d2min = dmin * dmin;
  int64_t placed = 0;
  for (int s = 0; s < n_slabs; ++s) {
    for (int64_t a = 0; a < count[s]; ++a) {
      bool ok = false;
      double p[3];
      for (int64_t t = 0; t < max_trials_per_atom && !ok; ++t) {
        p[0] = rng.uniform() * L;
        p[1] = rng.uniform() * L;
        p[2] = zlo[s] + rng.uniform() * (zhi[s] - zlo[s]);
        ok = true;
        if (dmin > 0) {
          int cx = grid.coord(p[0]), cy = grid.coord(p[1]), cz = grid.coord(p[2]);
          if (grid.nc < 3) {
            // too few cells for a 27-cell stencil: test every placed atom
            for (int64_t b = 0; b < placed && ok; ++b) {
              double rx = wrap(pos[3 * b] - p[0], L), ry = wrap(pos[3 * b + 1] - p[1], L),
                     rz = wrap(pos[3 * b + 2] - p[2], L);
              if (rx * rx + ry * ry + rz * rz < d2min) ok = false;
            }
          } else
          for (int dz = -1; dz <= 1 && ok; ++dz)
            for (int dy = -1; dy <= 1 && ok; ++dy)
              for (int dx = -1; dx <= 1 && ok; ++dx) {
                int ex = ((cx + dx) % grid.nc + grid.nc) % grid.nc;
                int ey = ((cy + dy) % grid.nc + grid.nc) % grid.nc;
                int ez = ((cz + dz) % grid.nc + grid.nc) % grid.nc;
                for (int b = grid.head[grid.cell(ex, ey, ez)]; b >= 0; b = grid.next[b]) {
                  double rx = wrap(pos[3 * b] - p[0], L), ry = wrap(pos[3 * b + 1] - p[1], L),
                         rz = wrap(pos[3 * b + 2] - p[2], L);
                  if (rx * rx + ry * ry + rz * rz < d2min) { ok = false; break; }
                }
              }
        }
      }
      if (!ok) return placed;
      std::memcpy(pos + 3 * placed, p, sizeof p);
      grid.insert((int)placed, p);
      ++placed;
    }
  }
  return placed;
}

// Species drawn independently per atom with cumulative probabilities cum[0..ns-1].
void stw_draw_species(uint64_t seed, int64_t n_atoms, int ns, const double* cum, int32_t* species) {
  SplitMix64 rng(seed);
  for (int64_t a = 0; a < n_atoms; ++a) {
    double u = rng.uniform();
    int s = 0;
    while (s < ns - 1 && u >= cum[s]) ++s;
    species[a] = s;
  }
}

// Standard-normal deviates (Box-Muller on the same stream).
void stw_normal(uint64_t seed, int64_t n, double* out) {
  SplitMix64 rng(seed);
  const double two_pi = 6.283185307179586;
  for (int64_t i = 0; i < n; i += 2) {
    double u1 = rng.uniform(), u2 = rng.uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    double rad = std::sqrt(-2.0 * std::log(u1));
    out[i] = rad * std::cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = rad * std::sin(two_pi * u2);
  }
}

// Neighbour list of the sites (site_atom[0..n_sites-1] are atom indices):
// for each site its own atom first, then every other atom j with
// |r_j - r_i| < rcut under the minimum image, in increasing j. Requires
// L > 2 rcut. Two passes: with nbr == nullptr only n_neigh is filled and the
// total pair count returned; then nbr (pair -> atom) and vec (3 per pair,
// r_j - r_i) are filled.
int64_t stw_neighbours(int64_t n_atoms, const double* pos, double L, double rcut, int64_t n_sites,
                       const int64_t* site_atom, int32_t* n_neigh, int32_t* nbr, double* vec) {
  CellGrid grid(L, rcut, (int)n_atoms);
  for (int64_t a = n_atoms - 1; a >= 0; --a) grid.insert((int)a, pos + 3 * a);
  const double rc2 = rcut * rcut;
  const bool stencil = grid.nc >= 3;
  // Visits the neighbours of site s in increasing atom index (its own atom excluded): emit(b, rx, ry, rz).
  auto visit = [&](int64_t s, std::vector<int>& uniq, auto&& emit) {
    const int64_t i = site_atom[s];
    const double* pi = pos + 3 * i;
    uniq.clear();
    if (stencil) {
      int cx = grid.coord(pi[0]), cy = grid.coord(pi[1]), cz = grid.coord(pi[2]);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int ex = ((cx + dx) % grid.nc + grid.nc) % grid.nc;
            int ey = ((cy + dy) % grid.nc + grid.nc) % grid.nc;
            int ez = ((cz + dz) % grid.nc + grid.nc) % grid.nc;
            for (int b = grid.head[grid.cell(ex, ey, ez)]; b >= 0; b = grid.next[b]) uniq.push_back(b);
          }
    } else {
      for (int b = 0; b < (int)n_atoms; ++b) uniq.push_back(b);
    }
    // increasing j, no duplicates (small grids can revisit a cell)
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    for (int b : uniq) {
      if (b == i) continue;
      double rx = wrap(pos[3 * b] - pi[0], L), ry = wrap(pos[3 * b + 1] - pi[1], L),
             rz = wrap(pos[3 * b + 2] - pi[2], L);
      double r2 = rx * rx + ry * ry + rz * rz;
      if (r2 >= rc2) continue;
      emit(b, rx, ry, rz);
    }
  };
  // counts (sites in parallel; each site's list depends on nothing else)
#pragma omp parallel
  {
    std::vector<int> uniq;
#pragma omp for schedule(dynamic, 256)
    for (int64_t s = 0; s < n_sites; ++s) {
      int32_t cnt = 1;
      visit(s, uniq, [&](int, double, double, double) { ++cnt; });
      n_neigh[s] = cnt;
    }
  }
  int64_t total = 0;
  std::vector<int64_t> start((size_t)n_sites + 1, 0);
  for (int64_t s = 0; s < n_sites; ++s) start[s + 1] = start[s] + n_neigh[s];
  total = start[n_sites];
  if (!nbr) return total;
#pragma omp parallel
  {
    std::vector<int> uniq;
#pragma omp for schedule(dynamic, 256)
    for (int64_t s = 0; s < n_sites; ++s) {
      int64_t k = start[s];
      nbr[k] = (int32_t)site_atom[s];
      vec[3 * k] = vec[3 * k + 1] = vec[3 * k + 2] = 0.0;
      ++k;
      visit(s, uniq, [&](int b, double rx, double ry, double rz) {
        nbr[k] = b;
        vec[3 * k] = rx;
        vec[3 * k + 1] = ry;
        vec[3 * k + 2] = rz;
        ++k;
      });
    }
  }
  return total;
}

// The same list for a general periodic cell: cell[c + 3v] is component c of lattice vector v (the
// vectors as columns of H). A pair vector is d - H rint(H^-1 d) for d = r_j - r_i, formed with the
// operations the device builder uses; every width of the cell (the distance between opposite faces)
// must exceed 2 rcut (returns -1 otherwise, or for a left-handed or degenerate cell). Atoms are binned
// on their fractional coordinates into cells at least rcut wide across their faces.
int64_t stw_neighbours_cell(int64_t n_atoms, const double* pos, const double* cell, double rcut, int64_t n_sites,
                            const int64_t* site_atom, int32_t* n_neigh, int32_t* nbr, double* vec) {
  double h[9], hi[9];
  for (int c = 0; c < 3; ++c)
    for (int v = 0; v < 3; ++v) h[3 * c + v] = cell[c + 3 * v];
  const double* a = cell;
  const double* b = cell + 3;
  const double* d = cell + 6;
  // rows of H^-1 = (b x c, c x a, a x b) / det
  const double bc[3] = {b[1] * d[2] - b[2] * d[1], b[2] * d[0] - b[0] * d[2], b[0] * d[1] - b[1] * d[0]};
  const double ca[3] = {d[1] * a[2] - d[2] * a[1], d[2] * a[0] - d[0] * a[2], d[0] * a[1] - d[1] * a[0]};
  const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
  if (!(det > 0.0)) return -1;
  const double* rows[3] = {bc, ca, ab};
  int n[3];
  for (int v = 0; v < 3; ++v) {
    const double* r = rows[v];
    for (int c = 0; c < 3; ++c) hi[3 * v + c] = r[c] / det;
    const double w = det / std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(w > 2.0 * rcut)) return -1;
    n[v] = std::min(128, (int)std::floor(w / rcut));
  }
  const bool stencil = n[0] >= 3 && n[1] >= 3 && n[2] >= 3;
  if (!stencil) n[0] = n[1] = n[2] = 1;
  auto cell_of = [&](const double* p, int* c) {
    for (int v = 0; v < 3; ++v) {
      const double f = hi[3 * v] * p[0] + hi[3 * v + 1] * p[1] + hi[3 * v + 2] * p[2];
      int k = (int)((f - std::floor(f)) * n[v]);
      c[v] = k < 0 ? 0 : (k >= n[v] ? n[v] - 1 : k);
    }
  };
  const size_t n_cells = (size_t)n[0] * n[1] * n[2];
  std::vector<int> head(n_cells, -1), next((size_t)n_atoms, -1);
  for (int64_t at = n_atoms - 1; at >= 0; --at) {
    int c[3];
    cell_of(pos + 3 * at, c);
    const size_t k = ((size_t)c[2] * n[1] + c[1]) * n[0] + c[0];
    next[at] = head[k];
    head[k] = (int)at;
  }
  const double rc2 = rcut * rcut;
  auto visit = [&](int64_t s, std::vector<int>& uniq, auto&& emit) {
    const int64_t i = site_atom[s];
    const double* pi = pos + 3 * i;
    uniq.clear();
    int c[3];
    cell_of(pi, c);
    const int span = stencil ? 1 : 0;
    for (int dz = -span; dz <= span; ++dz)
      for (int dy = -span; dy <= span; ++dy)
        for (int dx = -span; dx <= span; ++dx) {
          const int ex = ((c[0] + dx) % n[0] + n[0]) % n[0];
          const int ey = ((c[1] + dy) % n[1] + n[1]) % n[1];
          const int ez = ((c[2] + dz) % n[2] + n[2]) % n[2];
          for (int q = head[((size_t)ez * n[1] + ey) * n[0] + ex]; q >= 0; q = next[q]) uniq.push_back(q);
        }
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    for (int q : uniq) {
      if (q == i) continue;
      double x = pos[3 * q] - pi[0], y = pos[3 * q + 1] - pi[1], z = pos[3 * q + 2] - pi[2];
      const double n0 = std::nearbyint(hi[0] * x + hi[1] * y + hi[2] * z);
      const double n1 = std::nearbyint(hi[3] * x + hi[4] * y + hi[5] * z);
      const double n2 = std::nearbyint(hi[6] * x + hi[7] * y + hi[8] * z);
      x = x - (h[0] * n0 + h[1] * n1 + h[2] * n2);
      y = y - (h[3] * n0 + h[4] * n1 + h[5] * n2);
      z = z - (h[6] * n0 + h[7] * n1 + h[8] * n2);
      if (x * x + y * y + z * z >= rc2) continue;
      emit(q, x, y, z);
    }
  };
#pragma omp parallel
  {
    std::vector<int> uniq;
#pragma omp for schedule(dynamic, 256)
    for (int64_t s = 0; s < n_sites; ++s) {
      int32_t cnt = 1;
      visit(s, uniq, [&](int, double, double, double) { ++cnt; });
      n_neigh[s] = cnt;
    }
  }
  std::vector<int64_t> start((size_t)n_sites + 1, 0);
  for (int64_t s = 0; s < n_sites; ++s) start[s + 1] = start[s] + n_neigh[s];
  if (!nbr) return start[n_sites];
#pragma omp parallel
  {
    std::vector<int> uniq;
#pragma omp for schedule(dynamic, 256)
    for (int64_t s = 0; s < n_sites; ++s) {
      int64_t k = start[s];
      nbr[k] = (int32_t)site_atom[s];
      vec[3 * k] = vec[3 * k + 1] = vec[3 * k + 2] = 0.0;
      ++k;
      visit(s, uniq, [&](int q, double x, double y, double z) {
        nbr[k] = q;
        vec[3 * k] = x;
        vec[3 * k + 1] = y;
        vec[3 * k + 2] = z;
        ++k;
      });
    }
  }
  return start[n_sites];
}

// Spherical coordinates of the pair vectors: r, theta = acos(z/r),
// phi = atan2(y, x); the site's own pair (r = 0) gets zeros.
void stw_spherical(int64_t n_pairs, const double* vec, double* rjs, double* thetas, double* phis) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n_pairs; ++k) {
    double x = vec[3 * k], y = vec[3 * k + 1], z = vec[3 * k + 2];
    double r = std::sqrt(x * x + y * y + z * z);
    rjs[k] = r;
    if (r > 0) {
      double c = z / r;
      if (c > 1) c = 1;
      if (c < -1) c = -1;
      thetas[k] = std::acos(c);
      phis[k] = std::atan2(y, x);
    } else {
      thetas[k] = 0;
      phis[k] = 0;
    }
  }
}

}  // extern "C"
