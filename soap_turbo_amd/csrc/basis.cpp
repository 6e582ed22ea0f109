// Host FP64 construction of the radial basis transformation W = S^{-1/2}.
//
// The reference (get_orthonormalization_matrix_poly3 / _poly3gauss,
// src/soap_turbo_radial.f90:1023-1219) builds S in closed form (poly3) or by
// the Gaussian-polynomial overlap recursion (poly3gauss), then takes an SVD of
// S, forms S^{1/2} = U sqrt(Sigma) V^T, and inverts it through a Cholesky
// factorisation (dpotrf/dpotri). The same steps are taken here with a cyclic
// Jacobi eigensolver (S is symmetric positive definite, so its SVD is its
// eigendecomposition) and a hand-written Cholesky inverse, in FP64, on the
// host: the matrices are at most 12 x 12 and are built once per basis. The
// decomposition runs in x87 extended precision (64-bit mantissa): at poly3
// alpha_max = 8, cond(S) = 1.3e13 and a double-precision route loses ~1e-4 of
// W's accuracy; in extended precision W matches S^{-1/2} to ~1e-9, and only
// the final rounding to double remains.
#include <dlfcn.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "st_internal.h"

namespace st {

// Overlap matrix of the primitive functions, column-major a x a.
void overlap_matrix(int basis, int a, double atom_sigma_r, double rcut_hard, double* S) {
  std::memset(S, 0, sizeof(double) * a * a);
  const int np = basis == 1 ? a - 1 : a;
  for (int i = 1; i <= np; ++i) {
    S[(i - 1) + (i - 1) * a] = 1.0;
    for (int j = i + 1; j <= np; ++j) {
      double v = std::sqrt((double)(5 + 2 * i) * (double)(5 + 2 * j)) / (double)(5 + i + j);
      S[(i - 1) + (j - 1) * a] = v;
      S[(j - 1) + (i - 1) * a] = v;
    }
  }
  if (basis != 1) return;
  // Gaussian at the origin against the polynomials: the recursion of
  // src/soap_turbo_radial.f90:1133-1172, in units of rcut_hard.
  const double sg = atom_sigma_r / rcut_hard, rc = 1.0;
  const double s2 = sg * sg, sq2 = std::sqrt(2.0), pi = std::acos(-1.0);
  double I_n = 0.0, N_n = 1.0;
  double N_np1 = std::sqrt(rc / 1.0);  // N_a(rc, -2) = sqrt(rc / (2*(-2)+5))
  double I_np1 = std::sqrt(pi / 2.0) * sg * std::erf(rc / sq2 / sg) / N_np1;
  double C2 = s2 / rc;
  for (int n = -1; n <= a - 1; ++n) {
    C2 = C2 * rc;
    double N_np2 = std::sqrt(rc / (double)(2 * n + 5));
    double I_np2 = s2 * (double)(n + 1) * N_n / N_np2 * I_n + N_np1 * rc / N_np2 * I_np1 - C2 / N_np2;
    if (n > 0) {
      double v = I_np2 * sq2 / std::sqrt(sg) / std::pow(pi, 0.25);
      S[(a - 1) + (n - 1) * a] = v;
      S[(n - 1) + (a - 1) * a] = v;
    }
    N_n = N_np1;
    N_np1 = N_np2;
    I_n = I_np1;
    I_np1 = I_np2;
  }
  S[(a - 1) + (a - 1) * a] = 1.0;
}

using real_x = long double;

// Cyclic Jacobi: A (column-major, symmetric) -> eigenvalues w, eigenvectors V.
static void jacobi_eigen(int n, std::vector<real_x> A, std::vector<real_x>& w, std::vector<real_x>& V) {
  V.assign((size_t)n * n, 0.0L);
  for (int i = 0; i < n; ++i) V[i + i * n] = 1.0L;
  for (int sweep = 0; sweep < 100; ++sweep) {
    real_x off = 0.0L, diag = 0.0L;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) (i == j ? diag : off) += A[i + j * n] * A[i + j * n];
    if (off <= 1e-40L * diag) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        real_x apq = A[p + q * n];
        if (apq == 0.0L) continue;
        real_x app = A[p + p * n], aqq = A[q + q * n];
        real_x theta = (aqq - app) / (2.0L * apq);
        real_x t = (theta >= 0 ? 1.0L : -1.0L) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0L));
        real_x c = 1.0L / std::sqrt(t * t + 1.0L), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J (columns p, q)
          real_x akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, q)
          real_x apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          real_x vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i + i * n];
}

// W = (S^{1/2})^{-1}, S^{1/2} from the eigendecomposition, inverse through
// S^{1/2} = U^T U (Cholesky, upper) as dpotrf/dpotri do. Returns false if S
// is not positive definite.
bool orthonormalization_matrix(int a, const double* S, double* W) {
  std::vector<real_x> A(S, S + (size_t)a * a), w, V;
  jacobi_eigen(a, A, w, V);
  std::vector<real_x> H((size_t)a * a, 0.0L);
  for (int i = 0; i < a; ++i)
    if (!(w[i] > 0.0L)) return false;
  for (int j = 0; j < a; ++j)
    for (int i = 0; i < a; ++i) {
      real_x s = 0.0L;
      for (int k = 0; k < a; ++k) s += V[i + k * a] * std::sqrt(w[k]) * V[j + k * a];
      H[i + j * a] = s;
    }
  std::vector<real_x> U((size_t)a * a, 0.0L);
  for (int j = 0; j < a; ++j) {
    real_x d = H[j + j * a];
    for (int k = 0; k < j; ++k) d -= U[k + j * a] * U[k + j * a];
    if (!(d > 0.0L)) return false;
    d = std::sqrt(d);
    U[j + j * a] = d;
    for (int i = j + 1; i < a; ++i) {
      real_x s = H[j + i * a];
      for (int k = 0; k < j; ++k) s -= U[k + j * a] * U[k + i * a];
      U[j + i * a] = s / d;
    }
  }
  std::vector<real_x> Ui((size_t)a * a, 0.0L);
  for (int j = 0; j < a; ++j) {
    Ui[j + j * a] = 1.0L / U[j + j * a];
    for (int i = j - 1; i >= 0; --i) {
      real_x s = 0.0L;
      for (int k = i + 1; k <= j; ++k) s += U[i + k * a] * Ui[k + j * a];
      Ui[i + j * a] = -s / U[i + i * a];
    }
  }
  for (int j = 0; j < a; ++j)
    for (int i = 0; i < a; ++i) {
      real_x s = 0.0L;
      for (int k = (i > j ? i : j); k < a; ++k) s += Ui[i + k * a] * Ui[j + k * a];
      W[i + j * a] = (double)s;
    }
  return true;
}

namespace {

// LAPACK/BLAS entry points with the Fortran calling convention (LP64 integers, the hidden lengths
// of the character arguments last).
using dgesvd_fn = void (*)(const char*, const char*, const int*, const int*, double*, const int*, double*, double*,
                           const int*, double*, const int*, double*, const int*, int*, size_t, size_t);
using dgemm_fn = void (*)(const char*, const char*, const int*, const int*, const int*, const double*, const double*,
                          const int*, const double*, const int*, const double*, double*, const int*, size_t, size_t);
using dpotr_fn = void (*)(const char*, const int*, double*, const int*, int*, size_t);

struct Lapack {
  dgesvd_fn gesvd = nullptr;
  dgemm_fn gemm = nullptr;
  dpotr_fn potrf = nullptr, potri = nullptr;
  std::string err;
};

// The host's LAPACK, loaded at first use: ST_LAPACK names the library, else the image's MKL
// runtime or a system liblapack.
const Lapack& lapack() {
  static Lapack L;
  static std::once_flag once;
  std::call_once(once, [] {
    std::vector<std::string> names;
    if (const char* e = std::getenv("ST_LAPACK")) names.push_back(e);
    for (const char* n : {"libmkl_rt.so.1", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", "liblapack.so.3",
                          "libopenblas.so.0"})
      names.push_back(n);
    for (const std::string& n : names) {
      void* h = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
      if (!h) continue;
      L.gesvd = reinterpret_cast<dgesvd_fn>(dlsym(h, "dgesvd_"));
      L.gemm = reinterpret_cast<dgemm_fn>(dlsym(h, "dgemm_"));
      L.potrf = reinterpret_cast<dpotr_fn>(dlsym(h, "dpotrf_"));
      L.potri = reinterpret_cast<dpotr_fn>(dlsym(h, "dpotri_"));
      if (L.gesvd && L.gemm && L.potrf && L.potri) return;
      L.gesvd = nullptr;
      dlclose(h);
    }
    L.err = "no LAPACK library found (set ST_LAPACK to one that exports dgesvd_, dgemm_, dpotrf_, dpotri_)";
  });
  return L;
}

}  // namespace

// W = S^{-1/2} by the reference's own LAPACK sequence (get_orthonormalization_matrix_poly3[gauss],
// src/soap_turbo_radial.f90:1052-1080, :1177-1205): the SVD of S, S^{1/2} = U sqrt(Sigma) V^T by two
// GEMMs, then its inverse from the Cholesky factor (dpotrf, dpotri), the upper triangle mirrored.
// With the LAPACK the reference was linked with, W is the reference's W.
bool orthonormalization_matrix_lapack(int a, const double* S, double* W, std::string& msg) {
  const Lapack& L = lapack();
  if (!L.gesvd) {
    msg = L.err;
    return false;
  }
  std::vector<double> Sw(S, S + (size_t)a * a), U((size_t)a * a), VT((size_t)a * a), Sc((size_t)a * a), svd(a),
      work((size_t)6 * a);
  const int lwork = 6 * a;
  int info = 0;
  L.gesvd("A", "A", &a, &a, Sw.data(), &a, svd.data(), U.data(), &a, VT.data(), &a, work.data(), &lwork, &info, 1, 1);
  if (info != 0) {
    msg = "dgesvd failed on the overlap matrix (info " + std::to_string(info) + ")";
    return false;
  }
  std::fill(Sw.begin(), Sw.end(), 0.0);
  for (int i = 0; i < a; ++i) Sw[i + (size_t)i * a] = std::sqrt(svd[i]);
  const double one = 1.0, zero = 0.0;
  L.gemm("N", "N", &a, &a, &a, &one, U.data(), &a, Sw.data(), &a, &zero, Sc.data(), &a, 1, 1);
  L.gemm("N", "N", &a, &a, &a, &one, Sc.data(), &a, VT.data(), &a, &zero, Sw.data(), &a, 1, 1);
  L.potrf("U", &a, Sw.data(), &a, &info, 1);
  if (info != 0) {
    msg = "the overlap matrix of the radial basis is not positive definite (dpotrf info " + std::to_string(info) + ")";
    return false;
  }
  L.potri("U", &a, Sw.data(), &a, &info, 1);
  for (int i = 0; i < a; ++i) {
    W[i + (size_t)i * a] = Sw[i + (size_t)i * a];
    for (int j = i + 1; j < a; ++j) W[i + (size_t)j * a] = W[j + (size_t)i * a] = Sw[i + (size_t)j * a];
  }
  return true;
}

}  // namespace st
