/* FP64 CPU restatement of the soap_turbo get_soap hot path -- TEST ORACLE ONLY.
 *
 * This file is test infrastructure: only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it, as the checker. The product path never
 * links or calls it.
 *
 * Every function restates one reference routine, cited by file:line, keeping
 * its operation order (so that, built without FMA contraction, it rounds as
 * the reference's x86-64 build does). Parity pins, by function:
 *   so_preflm          closed form sqrt((2l+1)/4pi (l-m)!/(l+m)!) (tests), reference output
 *   so_overlap_S       closed form / mpmath quadrature (tests), reference S
 *   so_radial          reference radial expansion, mpmath quadrature of the overlap integral,
 *                      finite differences of the value (unified seed)
 *   so_angular         scipy lpmv / spherical_in closed forms (tests), reference output
 *   so_get_soap        reference soap and soap_cart_der; invariants (|q| = 1, self = -sum)
 *   so_contract        reference soap_contract; contraction of the reference soap_cart_der
 */
#include "soap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXA 32           /* alpha_max + 8 slots for the recursions */
#define MAXK 1024         /* k_max(l_max + 1) for l_max <= 40 */

int so_k_max(int l_max) { return 1 + l_max * (l_max + 1) / 2 + l_max; }

static int n_max_of(const so_params* p) {
  int n = 0;
  for (int s = 0; s < p->n_species; ++s) n += p->alpha_max[s];
  return n;
}

int so_n_soap(const so_params* p, const so_compress* c) {
  int n_max = n_max_of(p);
  if (c && c->compress) {
    int m = 0;
    for (int i = 0; i < c->P_nonzero; ++i)
      if (c->P_i[i] > m) m = c->P_i[i];
    return m;
  }
  return n_max * (n_max + 1) / 2 * (p->l_max + 1);
}

/* get_preflm, src/soap_turbo_functions.f90:215-241 */
void so_preflm(int l_max, double* preflm) {
  int k = 0;
  for (int l = 0; l <= l_max; ++l) {
    double ppi = sqrt((double)(2 * l + 1) / 4.0 / acos(-1.0));
    double fact2 = 1.0;
    for (int i = 1; i <= l; ++i) fact2 = fact2 * (double)i;
    double fact1 = fact2;
    for (int m = 0; m <= l; ++m) {
      if (m > 0) {
        fact1 = fact1 / (double)(l + 1 - m);
        fact2 = fact2 * (double)(l + m);
      }
      preflm[k++] = ppi * sqrt(fact1 / fact2);
    }
  }
}

/* Overlap matrix of the primitive basis (column-major alpha x alpha):
 * poly3 src/soap_turbo_radial.f90:1041-1047; poly3gauss :1133-1172 (the
 * Gaussian-polynomial overlaps by their recursion). */
void so_overlap_S(int basis, int alpha_max, double atom_sigma_r, double rcut_hard, double* S) {
  const int a = alpha_max;
  memset(S, 0, sizeof(double) * a * a);
  int np = basis == 1 ? a - 1 : a;
  for (int i = 1; i <= np; ++i) {
    S[(i - 1) + (i - 1) * a] = 1.0;
    for (int j = i + 1; j <= np; ++j) {
      double v = sqrt((double)(5 + 2 * i) * (double)(5 + 2 * j)) / (double)(5 + i + j);
      S[(i - 1) + (j - 1) * a] = v;
      S[(j - 1) + (i - 1) * a] = v;
    }
  }
  if (basis != 1) return;
  double atom_sigma = atom_sigma_r / rcut_hard, rc = 1.0;
  double s2 = atom_sigma * atom_sigma, sq2 = sqrt(2.0), pi = acos(-1.0);
  double I_n = 0.0, N_n = 1.0;
  double N_np1 = sqrt(rc / (double)(2 * (-2) + 5));
  double I_np1 = sqrt(pi / 2.0) * atom_sigma * erf(rc / sq2 / atom_sigma) / N_np1;
  double C2 = s2 / rc;
  for (int n = -1; n <= a - 1; ++n) {
    C2 = C2 * rc;
    double N_np2 = sqrt(rc / (double)(2 * n + 5));
    double I_np2 = s2 * (double)(n + 1) * N_n / N_np2 * I_n + N_np1 * rc / N_np2 * I_np1 - C2 / N_np2;
    if (n > 0) {
      double v = I_np2 * sq2 / sqrt(atom_sigma) / pow(pi, 0.25);
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

/* One pair of one species: process_block, src/soap_turbo_radial.f90:550-1002,
 * for the per-call constants of radial_expansion_range (:394-547). Output
 * before global_scaling. */
typedef struct {
  int alpha_max, n_poly, n_hi, do_poly, with_gauss, do_der, unified, has_filter, scaling_poly, radial_enh;
  int do_central;
  double rcut_soft, atom_sigma, dr, inv_rcut, sqrt_rcut, inv_sqrt_rcut, inv_Nm2, sigma_sq0, N_gauss, q, drnf;
  double scaling, amp_scaling, central_weight;
  double ra[MAXA], rb[MAXA], rc[MAXA], R1[MAXA], R2[MAXA]; /* index n + 1 (ra(-1) at 0) */
  const double* W; int ldw;                                /* species block of W */
} radial_consts;

static void radial_setup(radial_consts* c, const so_params* p, int s, const double* W, int ldw, int off, int do_der) {
  memset(c, 0, sizeof *c);
  c->alpha_max = p->alpha_max[s];
  c->with_gauss = p->basis == 1;
  c->do_der = do_der;
  c->unified = !p->legacy_seed;
  c->scaling_poly = p->scaling_poly;
  c->radial_enh = p->radial_enhancement;
  c->scaling = p->atom_sigma_r_scaling[s];
  c->amp_scaling = p->amplitude_scaling[s];
  c->central_weight = p->central_weight[s];
  c->do_central = p->basis == 1 ? 0 : (p->central_weight[s] != 0.0);
  c->inv_rcut = 1.0 / p->rcut_hard[s];
  c->rcut_soft = p->rcut_soft[s] * c->inv_rcut;
  c->atom_sigma = p->atom_sigma_r[s] * c->inv_rcut;
  c->dr = 1.0 - c->rcut_soft;
  c->sqrt_rcut = sqrt(p->rcut_hard[s]);
  c->inv_sqrt_rcut = 1.0 / c->sqrt_rcut;
  c->sigma_sq0 = c->atom_sigma * c->atom_sigma;
  c->N_gauss = sqrt(2.0 / c->atom_sigma) / pow(acos(-1.0), 0.25);
  c->has_filter = c->dr > 0.0;
  if (c->has_filter) {
    c->drnf = c->dr / p->nf[s];
    c->q = c->drnf * c->drnf;
  }
  c->n_poly = c->with_gauss ? c->alpha_max - 1 : c->alpha_max;
  c->do_poly = c->n_poly >= 1;
  c->n_hi = c->n_poly + (do_der ? 2 : 0);
  double Nv[MAXA + 4]; /* Nv(n) at n + 3 */
  Nv[0] = 1.0;
  for (int n = -2; n <= c->alpha_max + 4; ++n) Nv[n + 3] = sqrt(1.0 / (double)(2 * n + 5));
  for (int n = -1; n <= c->n_hi; ++n) {
    c->ra[n + 1] = (double)(n + 1) * Nv[n - 2 + 3] / Nv[n + 3];
    c->rb[n + 1] = Nv[n - 1 + 3] / Nv[n + 3];
    c->rc[n + 1] = 1.0 / Nv[n + 3];
  }
  c->ra[0] = 0.0;
  c->inv_Nm2 = 1.0 / Nv[-2 + 3];
  for (int n = 1; n <= c->n_poly; ++n) {
    c->R1[n] = Nv[n + 1 + 3] / Nv[n + 3];
    c->R2[n] = Nv[n + 2 + 3] / Nv[n + 3];
  }
  c->W = W + off + (size_t)off * ldw;
  c->ldw = ldw;
}

static void radial_pair(const radial_consts* c, double rjs, int is_central, double* outv, double* outd) {
  const double pi = acos(-1.0), sq2 = sqrt(2.0), isq2 = 1.0 / sq2, sqpi2 = sqrt(pi / 2.0),
               sq2pi = sqrt(2.0 / pi), sq8pi = sqrt(8.0 / pi);
  const double sc = c->scaling, rs = c->rcut_soft;
  double T1[MAXA], T2[MAXA], pol[MAXA], DT[MAXA]; /* T(n) at n + 3; pol(n), DT(n) at n */
  memset(T2, 0, sizeof T2);
  memset(DT, 0, sizeof DT);
  double rj = rjs * c->inv_rcut;
  double sig = c->atom_sigma + sc * rj, s2 = sig * sig, isig = 1.0 / sig;
  double is2 = isig * isig, is3 = is2 * isig;
  double amp, ampd;
  /* amplitude, :589-635 */
  if (!c->scaling_poly) {
    amp = 1.0; ampd = 0.0;
  } else if (c->amp_scaling == 0.0) {
    amp = isig; ampd = -sc * is2;
  } else if (c->amp_scaling == 1.0) {
    double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
    double hp = 6.0 * (rj * rj) - 6.0 * rj;
    double g = isig * h;
    amp = h <= 1e-10 ? 0.0 : g;
    ampd = h <= 1e-10 ? 0.0 : hp * isig - sc * isig * g;
  } else {
    double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
    if (h <= 1e-10) {
      amp = 0.0; ampd = 0.0;
    } else {
      double hp = 6.0 * (rj * rj) - 6.0 * rj, aa = c->amp_scaling;
      amp = isig * pow(h, aa);
      ampd = aa * pow(h, aa - 1.0) * hp * isig - sc * isig * amp;
    }
  }
  /* central weight, :637-646 */
  if (c->do_central && c->central_weight != 1.0 && is_central) {
    amp = c->central_weight * amp;
    ampd = c->central_weight * ampd;
  }
  /* radial enhancement, :648-670 */
  if (c->radial_enh == 1) {
    double g = rj + sq2pi * sig, f = 1.0 + sq2pi * sc;
    ampd = amp * f + ampd * g;
    amp = amp * g;
  } else if (c->radial_enh == 2) {
    double g = (rj * rj) + s2 + sq8pi * sig * rj;
    double f = 2.0 * rj + 2.0 * sig * sc + sq8pi * sig + sq8pi * rj * sc;
    ampd = amp * f + ampd * g;
    amp = amp * g;
  }
  /* overlap on [0, rcut_soft], :672-698 */
  double es = exp(-0.5 * ((rs - rj) * (rs - rj)) * is2);
  double e0 = exp(-0.5 * (rj * rj) * is2);
  double C1 = s2 * es, C2 = s2 * e0;
  T1[0] = 0.0;
  T1[1] = sqpi2 * sig * (erf((rs - rj) * isq2 * isig) + erf(rj * isq2 * isig)) * c->inv_Nm2;
  for (int n = -1; n <= c->n_hi; ++n) {
    if (n >= 0) C1 = C1 * c->dr;
    T1[n + 3] = s2 * c->ra[n + 1] * T1[n + 1] - c->rb[n + 1] * (rj - 1.0) * T1[n + 2] + (C1 - C2) * c->rc[n + 1];
  }
  /* legacy derivative of that piece, :700-714 */
  if (c->do_der && c->do_poly && !c->unified) {
    double f = rj - 1.0, g = sc * f * isig;
    double ca = f * is2 * (g - 1.0), cb = is2 * (2.0 * g - 1.0), cc = sc * is3;
    for (int n = 1; n <= c->n_poly; ++n)
      DT[n] = ca * T1[n + 3] + c->R1[n] * cb * T1[n + 4] + c->R2[n] * cc * T1[n + 5];
  }
  /* Gaussian filter on [rcut_soft, rcut_hard], :716-805 */
  double pref = 0.0;
  if (c->has_filter) {
    double den = s2 + c->q, iD = 1.0 / den, sqiD = sqrt(iD);
    double sgf = sig * c->drnf * sqiD;
    double rjf = (s2 * rs + c->q * rj) * iD;
    double sf2 = sgf * sgf;
    int keep = (rs - rj) < 4.0 * sig;
    double isgf = 1.0 / sgf, isf2 = isgf * isgf;
    pref = keep ? exp(-0.5 * ((rs - rj) * (rs - rj)) * iD) : 0.0;
    double C2f = sf2 * exp(-0.5 * ((rs - rjf) * (rs - rjf)) * isf2);
    T2[0] = 0.0;
    T2[1] = sqpi2 * sgf * (erf((1.0 - rjf) * isq2 * isgf) - erf((rs - rjf) * isq2 * isgf)) * c->inv_Nm2;
    for (int n = -1; n <= c->n_hi; ++n) {
      if (n >= 0) C2f = C2f * c->dr;
      T2[n + 3] = sf2 * c->ra[n + 1] * T2[n + 1] - c->rb[n + 1] * (rjf - 1.0) * T2[n + 2] - C2f * c->rc[n + 1];
      if (n == -1 && c->unified)
        T2[2] = T2[2] + sf2 * exp(-0.5 * ((1.0 - rjf) * (1.0 - rjf)) * isf2) * c->rc[0];
    }
    if (c->do_der && c->do_poly && !c->unified) {
      double f = rs - rj;
      double dpref = pref * (f * iD + f * f * iD * iD * sig * sc);
      double drjf = (2.0 * sig * rs * sc + c->q) * iD - (s2 * rs + c->q * rj) * 2.0 * sig * sc * iD * iD;
      double dsgf = sc * c->drnf * sqiD * (1.0 - s2 * iD);
      double ff = rjf - 1.0, g = dsgf * ff * isgf;
      double ca = pref * ff * isf2 * (g - drjf);
      double cb = pref * isf2 * (2.0 * g - drjf);
      double cc = pref * dsgf * isgf * isf2;
      for (int n = 1; n <= c->n_poly; ++n)
        DT[n] = DT[n] + ca * T2[n + 3] + c->R1[n] * cb * T2[n + 4] + c->R2[n] * cc * T2[n + 5] + dpref * T2[n + 3];
    }
  }
  /* :807-812 */
  for (int n = 1; n <= c->n_hi; ++n) pol[n] = T1[n + 3] + pref * T2[n + 3];
  /* unified derivatives, :814-862 */
  if (c->do_der && c->do_poly && c->unified) {
    double f = 1.0 - rj;
    double g0 = f * is2 + sc * f * f * is3;
    double g1 = -(is2 + 2.0 * sc * f * is3);
    double g2 = sc * is3;
    for (int n = 1; n <= c->n_poly; ++n)
      DT[n] = g0 * pol[n] + c->R1[n] * g1 * pol[n + 1] + c->R2[n] * g2 * pol[n + 2];
  }
  /* Gaussian at the origin, :864-908 */
  if (c->with_gauss) {
    double ss2 = c->sigma_sq0 + s2, iss = 1.0 / ss2, siss = sqrt(iss);
    int gm = rj < 4.0 * (c->atom_sigma + sig);
    double ge = gm ? c->N_gauss * e0 : 0.0;
    double M0 = gm ? exp(-0.5 * (rj * rj) * iss) * sqpi2 * sig * c->atom_sigma * siss *
                         (1.0 + erf(c->atom_sigma * isig * rj * isq2 * siss)) * c->N_gauss
                   : 0.0;
    pol[c->alpha_max] = M0;
    if (c->do_der) {
      double sc0 = c->sigma_sq0 * s2 * iss, pp = s2 * rj * iss;
      double M1 = sc0 * ge - pp * M0;
      double M2 = sc0 * ge * (-rj) - pp * M1 + sc0 * M0;
      DT[c->alpha_max] = M1 * is2 + sc * M2 * is3;
    }
  }
  /* transform and store, :910-1000 */
  const int a = c->alpha_max;
  if (c->do_der) {
    double dtot[MAXA];
    for (int n = 1; n <= a; ++n) dtot[n] = amp * DT[n] + ampd * pol[n];
    for (int n = 1; n <= a; ++n) {
      double acc = c->W[(n - 1)] * dtot[1];
      for (int m = 2; m <= a; ++m) acc = acc + c->W[(n - 1) + (size_t)(m - 1) * c->ldw] * dtot[m];
      outd[n - 1] = acc * c->inv_sqrt_rcut;
    }
  }
  for (int n = 1; n <= a; ++n) {
    double acc = c->W[(n - 1)] * pol[1];
    for (int m = 2; m <= a; ++m) acc = acc + c->W[(n - 1) + (size_t)(m - 1) * c->ldw] * pol[m];
    outv[n - 1] = amp * acc * c->sqrt_rcut;
  }
}

/* get_radial_expansion_coefficients_poly3[gauss] for every species, then
 * global_scaling (src/soap_turbo.f90:491-544; src/soap_turbo_radial.f90:116-384). */
void so_radial(const so_params* p, const so_inputs* in, const double* W, int do_der, double* rad, double* rad_der) {
  const int n_max = n_max_of(p);
  memset(rad, 0, sizeof(double) * (size_t)n_max * in->n_pairs);
  if (do_der) memset(rad_der, 0, sizeof(double) * (size_t)n_max * in->n_pairs);
  int* first = (int*)malloc(sizeof(int) * (in->n_pairs + 1));
  {
    int k = 0;
    for (int i = 0; i < in->n_sites; ++i)
      for (int j = 0; j < in->n_neigh[i]; ++j) first[k++] = (j == 0);
  }
  int off = 0;
  for (int s = 0; s < p->n_species; ++s) {
    radial_consts c;
    radial_setup(&c, p, s, W, n_max, off, do_der);
    const double gs = p->global_scaling[s];
#pragma omp parallel for schedule(dynamic, 256)
    for (int k = 0; k < in->n_pairs; ++k) {
      if (first[k] && !c.do_central) continue;
      if (in->rjs[k] >= p->rcut_hard[s]) continue;
      if (!in->mask[k + (size_t)s * in->n_pairs]) continue;
      double v[MAXA], d[MAXA];
      radial_pair(&c, in->rjs[k], first[k], v, d);
      for (int n = 0; n < c.alpha_max; ++n) {
        rad[(size_t)k * n_max + off + n] = v[n] * gs;
        if (do_der) rad_der[(size_t)k * n_max + off + n] = d[n] * gs;
      }
    }
    off += p->alpha_max[s];
  }
  free(first);
}

/* get_plm_array, src/soap_turbo_angular.f90:108-160 (tables :55-70) */
static void plm_array(int lmax, double x, double* plm) {
  double x2 = x * x, sq = sqrt(1.0 - x2);
  plm[0] = 1.0;
  if (lmax < 1) return;
  plm[1] = x;
  plm[2] = -sq;
  if (lmax < 2) return;
  plm[3] = 1.5 * x2 - 0.5;
  plm[4] = -3.0 * x * sq;
  plm[5] = 3.0 - 3.0 * x2;
  for (int l = 3; l <= lmax; ++l) {
    int k = 0;
    for (int m = 0; m <= l - 2; ++m) {
      k = 1 + l * (l + 1) / 2 + m;
      double a = (double)(2 * l - 1) / (double)(l - m), b = (double)(l - 1 + m) / (double)(l - m);
      plm[k - 1] = a * x * plm[k - l - 1] - b * plm[k - 2 * l + 1 - 1];
    }
    k = k + 1;
    plm[k - 1] = x * (double)(2 * l - 1) * plm[k - l - 1];
    k = k + 1;
    plm[k - 1] = -((double)(2 * l - 1) * sq * plm[k - l - 1 - 1]);
  }
}

/* get_plm_array_der, src/soap_turbo_angular.f90:181-248 (tables :75-99) */
static void plm_array_der(int lmax, const double* plm, double* div_sin, double* der_mul_sin) {
  for (int l = 0; l <= lmax; ++l)
    for (int m = 0; m <= l; ++m) {
      int k = 1 + l * (l + 1) / 2 + m;
      double part1, part2;
      if (m == 0) {
        part1 = l == 0 ? 0.0 : -0.5 * plm[1 + l * (l + 1) / 2 + 1 - 1];
      } else {
        double der_a = 0.5 * (double)(l + m) * (double)(l - m + 1);
        part1 = der_a * plm[k - 1 - 1];
      }
      part2 = m == l ? 0.0 : -0.5 * plm[k + 1 - 1];
      der_mul_sin[k - 1] = part1 + part2;
    }
  for (int l = 0; l <= lmax; ++l) div_sin[l * (l + 1) / 2] = 0.0;
  for (int l = 0; l <= lmax; ++l)
    for (int m = 1; m <= l; ++m) {
      int k = 1 + l * (l + 1) / 2 + m;
      int kp = 1 + (l + 1) * (l + 2) / 2 + m + 1, km = 1 + (l + 1) * (l + 2) / 2 + m - 1;
      double div_a = 0.5 * (double)(l - m + 1) * (double)(l - m + 2);
      div_sin[k - 1] = div_a * plm[km - 1] + 0.5 * plm[kp - 1];
    }
}

/* get_ilexp, src/soap_turbo_angular.f90:361-422; fact_array from
 * get_semifactorial_array :344-356 */
static void ilexp(int lmax, double x, double* out) {
  const double xcut = 1e-7;
  double fact[64];
  double f = 1.0;
  for (int i = 1; i <= lmax; ++i) {
    f = f * (2.0 * (double)i + 1.0);
    fact[i] = f;
  }
  double x2 = x * x, x4 = x2 * x2;
  double inv_x2 = x2 > 0.0 ? 1.0 / x2 : 0.0;
  double x2pow = x2 * x2;
  double flm2, flm1;
  if (x > 0.0) {
    flm2 = fabs((1.0 - exp(-2.0 * x2)) / 2.0 / x2);
    flm1 = fabs((x2 - 1.0 + exp(-2.0 * x2) * (x2 + 1.0)) / 2.0 / x4);
  } else {
    flm2 = 1.0;
    flm1 = 0.0;
  }
  for (int l = 0; l <= lmax; ++l) {
    if (l == 0) {
      out[0] = x < xcut ? 1.0 - x2 : flm2;
    } else if (l == 1) {
      out[1] = x2 / 1000.0 < xcut ? (x2 - x4) / fact[1] : flm1;
    } else {
      double ratio = x2pow / fact[l], fl;
      if (ratio * (double)l < xcut)
        fl = ratio;
      else
        fl = fabs(flm2 - (2.0 * (double)l - 1.0) * inv_x2 * flm1);
      x2pow = x2pow * x2;
      flm2 = flm1;
      flm1 = fl;
      out[l] = fl;
    }
  }
}

/* get_ilexp_der, src/soap_turbo_angular.f90:436-463 */
static void ilexp_der(int lmax, const double* il, double rj, double sigma, double scaling, double* der) {
  double coeff1 = 2.0 * rj / (sigma * sigma);
  double coeff2 = 1.0 - scaling * rj / sigma;
  for (int l = 0; l <= lmax; ++l) der[l] = 0.0;
  if (rj < 1e-5) return;
  der[0] = coeff1 * (il[1] - il[0]);
  for (int l = 1; l <= lmax; ++l) der[l] = (-coeff1 - (double)(2 * l + 2) / rj) * il[l] + coeff1 * il[l - 1];
  for (int l = 0; l <= lmax; ++l) der[l] = der[l] * coeff2;
}

/* get_angular_expansion_coefficients, src/soap_turbo_angular.f90:485-617,
 * with get_eimphi_conjg :267-329 */
void so_angular(const so_params* p, const so_inputs* in, int do_der, double* ang, double* ang_r, double* ang_a,
                double* ang_p) {
  const int lmax = p->l_max, kmax = so_k_max(lmax);
  double rcut = 0.0;
  for (int s = 0; s < p->n_species; ++s)
    if (p->rcut_hard[s] > rcut) rcut = p->rcut_hard[s];
  const double rcut2 = rcut * rcut;
  double preflm[MAXK];
  so_preflm(lmax, preflm);
#pragma omp parallel for schedule(dynamic, 64)
  for (int k = 0; k < in->n_pairs; ++k) {
    double rj = in->rjs[k];
    int i_sp = -1;
    if (rj < rcut)
      for (int s = 0; s < p->n_species; ++s)
        if (in->mask[k + (size_t)s * in->n_pairs]) { i_sp = s; break; }
    double* A = ang + 2 * (size_t)k * kmax;
    if (i_sp < 0) {
      memset(A, 0, sizeof(double) * 2 * kmax);
      if (do_der) {
        memset(ang_r + 2 * (size_t)k * kmax, 0, sizeof(double) * 2 * kmax);
        memset(ang_a + 2 * (size_t)k * kmax, 0, sizeof(double) * 2 * kmax);
        memset(ang_p + 2 * (size_t)k * kmax, 0, sizeof(double) * 2 * kmax);
      }
      continue;
    }
    double phi = in->phis[k], x = cos(in->thetas[k]);
    double sigma = p->atom_sigma_t[i_sp] + p->atom_sigma_t_scaling[i_sp] * rj;
    double amplitude = rcut2 / (sigma * sigma);
    /* get_eimphi_conjg */
    double prefl[64], prefl_d[64], pm_re[64], pm_im[64];
    ilexp(lmax, rj / sigma, prefl);
    double cosm2 = cos(phi), cosphi2 = 2.0 * cosm2, sinm2 = -sin(phi), cosm1 = 1.0, sinm1 = 0.0;
    pm_re[0] = 1.0; pm_im[0] = 0.0;
    for (int l = 1; l <= lmax; ++l) {
      double cos0 = cosphi2 * cosm1 - cosm2, sin0 = cosphi2 * sinm1 - sinm2;
      cosm2 = cosm1; sinm2 = sinm1; cosm1 = cos0; sinm1 = sin0;
      pm_re[l] = cos0; pm_im[l] = -sin0;
    }
    if (do_der) ilexp_der(lmax, prefl, rj, sigma, p->atom_sigma_t_scaling[i_sp], prefl_d);
    double plm[MAXK];
    if (do_der) {
      double div_sin[MAXK], dms[MAXK];
      plm_array(lmax + 1, x, plm);
      plm_array_der(lmax, plm, div_sin, dms);
      double sig_fac = 2.0 * p->atom_sigma_t_scaling[i_sp] / sigma;
      double* Ar = ang_r + 2 * (size_t)k * kmax;
      double* Aa = ang_a + 2 * (size_t)k * kmax;
      double* Ap = ang_p + 2 * (size_t)k * kmax;
      int kk = 0;
      for (int l = 0; l <= lmax; ++l)
        for (int m = 0; m <= l; ++m, ++kk) {
          double e_re = prefl[l] * pm_re[m], e_im = prefl[l] * pm_im[m];
          double er_re = prefl_d[l] * pm_re[m], er_im = prefl_d[l] * pm_im[m];
          double u = amplitude * preflm[kk], v = u * plm[kk];
          double ec_re = v * e_re, ec_im = v * e_im;
          A[2 * kk] = ec_re; A[2 * kk + 1] = ec_im;
          Ar[2 * kk] = v * er_re - sig_fac * ec_re;
          Ar[2 * kk + 1] = v * er_im - sig_fac * ec_im;
          double t = u * div_sin[kk];
          double ez_re = t * e_re, ez_im = t * e_im;
          Aa[2 * kk] = -ez_im; Aa[2 * kk + 1] = ez_re;
          double t2 = u * dms[kk];
          Ap[2 * kk] = t2 * e_re; Ap[2 * kk + 1] = t2 * e_im;
        }
    } else {
      plm_array(lmax, x, plm);
      int kk = 0;
      for (int l = 0; l <= lmax; ++l)
        for (int m = 0; m <= l; ++m, ++kk) {
          double e_re = prefl[l] * pm_re[m], e_im = prefl[l] * pm_im[m];
          double t = amplitude * preflm[kk] * plm[kk];
          A[2 * kk] = t * e_re; A[2 * kk + 1] = t * e_im;
        }
    }
  }
}

typedef struct {
  int n_max, k_max, n_soap, n_soap_unc;
  int* k_first;
  char* skip;
  double* mult;      /* multiplicity per non-skipped (n, np, l, m) */
  double *rad, *rad_der, *ang, *ang_r, *ang_a, *ang_p, *cnk;
} so_state;

static void state_free(so_state* st) {
  free(st->k_first); free(st->skip); free(st->mult);
  free(st->rad); free(st->rad_der); free(st->ang); free(st->ang_r); free(st->ang_a); free(st->ang_p); free(st->cnk);
}

/* The expansion and cnk (src/soap_turbo.f90:394-664) */
static void build_state(so_state* st, const so_params* p, const so_inputs* in, const double* W, const double* S,
                        const so_compress* c, int do_der) {
  memset(st, 0, sizeof *st);
  const int n_max = n_max_of(p), k_max = so_k_max(p->l_max), L = p->l_max;
  st->n_max = n_max; st->k_max = k_max;
  st->n_soap_unc = n_max * (n_max + 1) / 2 * (L + 1);
  st->n_soap = so_n_soap(p, c);
  st->k_first = (int*)malloc(sizeof(int) * (in->n_sites + 1));
  int k2 = 0;
  for (int i = 0; i < in->n_sites; ++i) { st->k_first[i] = k2; k2 += in->n_neigh[i]; }
  st->skip = (char*)calloc(st->n_soap_unc, 1);
  if (c && c->compress) {
    memset(st->skip, 1, st->n_soap_unc);
    for (int i = 0; i < c->P_nonzero; ++i) st->skip[c->P_j[i] - 1] = 0;
  }
  /* multiplicity array, :707-745 */
  int cnt2 = 0, cnt = 0;
  st->mult = (double*)malloc(sizeof(double) * (size_t)(st->n_soap_unc) * (L + 1));
  for (int n = 1; n <= n_max; ++n)
    for (int np = n; np <= n_max; ++np)
      for (int l = 0; l <= L; ++l) {
        if (st->skip[cnt++]) continue;
        for (int m = 0; m <= l; ++m) {
          double mu = 1.0;
          if (n != np) mu = mu * sqrt(2.0);
          if (m > 0) mu = mu * 2.0;
          st->mult[cnt2++] = mu;
        }
      }
  size_t np_ = in->n_pairs;
  st->rad = (double*)malloc(sizeof(double) * n_max * np_);
  st->rad_der = (double*)malloc(sizeof(double) * n_max * np_);
  st->ang = (double*)malloc(sizeof(double) * 2 * k_max * np_);
  if (do_der) {
    st->ang_r = (double*)malloc(sizeof(double) * 2 * k_max * np_);
    st->ang_a = (double*)malloc(sizeof(double) * 2 * k_max * np_);
    st->ang_p = (double*)malloc(sizeof(double) * 2 * k_max * np_);
  }
  so_radial(p, in, W, do_der, st->rad, st->rad_der);
  so_angular(p, in, do_der, st->ang, st->ang_r, st->ang_a, st->ang_p);
  /* cnk, :619-664 */
  st->cnk = (double*)calloc(2 * (size_t)k_max * n_max * in->n_sites, sizeof(double));
  const double pi = acos(-1.0);
  int* i_beg = (int*)malloc(sizeof(int) * p->n_species);
  int* i_end = (int*)malloc(sizeof(int) * p->n_species);
  i_beg[0] = 1; i_end[0] = p->alpha_max[0];
  for (int s = 1; s < p->n_species; ++s) { i_beg[s] = i_end[s - 1] + 1; i_end[s] = i_beg[s] + p->alpha_max[s] - 1; }
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < in->n_sites; ++i) {
    double* cn = st->cnk + 2 * (size_t)i * n_max * k_max;
    int kk2 = st->k_first[i];
    for (int j = 0; j < in->n_neigh[i]; ++j, ++kk2) {
      for (int n = 0; n < n_max; ++n) {
        double fr = 4.0 * pi * st->rad[(size_t)kk2 * n_max + n];
        const double* A = st->ang + 2 * (size_t)kk2 * k_max;
        for (int k = 0; k < k_max; ++k) {
          cn[2 * (n * k_max + k)] += fr * A[2 * k];
          cn[2 * (n * k_max + k) + 1] += fr * A[2 * k + 1];
        }
      }
    }
    for (int m = 0; m < in->species_mult[i]; ++m) {
      int j = in->species[m + i * 1] - 1;  /* species(k, i), leading dim 1 */
      if (p->basis == 1 && p->central_weight[j] != 0.0) {
        double amplitude;
        if (p->radial_enhancement == 1) amplitude = sqrt(2.0 / pi) * p->atom_sigma_r[j] / p->rcut_hard[j];
        else if (p->radial_enhancement == 2)
          amplitude = (p->atom_sigma_r[j] * p->atom_sigma_r[j]) / (p->rcut_hard[j] * p->rcut_hard[j]);
        else amplitude = 1.0;
        double rc = p->rcut_hard[j];
        double fac = amplitude * p->central_weight[j] * sqrt(4.0 * pi) * pow(pi, 0.25) *
                     sqrt(p->atom_sigma_r[j] / 2.0) * (rc * rc * rc) / (p->atom_sigma_t[j] * p->atom_sigma_t[j]) /
                     p->atom_sigma_r[j];
        int a = p->alpha_max[j], off = i_beg[j] - 1, col = i_end[j] - 1;
        for (int n = 0; n < a; ++n) {
          double v = 0.0;
          for (int m2 = 0; m2 < a; ++m2)
            v = v + W[(off + n) + (size_t)(off + m2) * n_max] * S[(off + m2) + (size_t)col * n_max];
          cn[2 * ((off + n) * k_max + 0)] += fac * v;
        }
      }
    }
  }
  free(i_beg); free(i_end);
}

/* Power spectrum, compression and norm: src/soap_turbo.f90:750-813 */
static void power_spectrum(const so_state* st, const so_params* p, const so_compress* c, int n_sites, double* soap,
                           double* sqrt_dot_p) {
  const int n_max = st->n_max, k_max = st->k_max, L = p->l_max, n_soap = st->n_soap;
#pragma omp parallel
  {
    double* ts = (double*)calloc(st->n_soap_unc, sizeof(double));
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < n_sites; ++i) {
      const double* cn = st->cnk + 2 * (size_t)i * n_max * k_max;
      int cnt = 0, cnt2 = 0;
      for (int n = 0; n < n_max; ++n)
        for (int np = n; np < n_max; ++np)
          for (int l = 0; l <= L; ++l) {
            if (st->skip[cnt]) { cnt++; continue; }
            int kb = l * (l + 1) / 2;
            double acc = 0.0;
            for (int m = 0; m <= l; ++m) {
              int k = kb + m;
              double ur = cn[2 * (n * k_max + k)], ui = cn[2 * (n * k_max + k) + 1];
              double pr = cn[2 * (np * k_max + k)], pi_ = cn[2 * (np * k_max + k) + 1];
              acc = acc + st->mult[cnt2++] * (ur * pr + ui * pi_);
            }
            ts[cnt++] = acc;
          }
      double* q = soap + (size_t)i * n_soap;
      if (c && c->compress) {
        for (int k = 0; k < c->P_nonzero; ++k) q[c->P_i[k] - 1] = q[c->P_i[k] - 1] + c->P_el[k] * ts[c->P_j[k] - 1];
      } else {
        for (int s = 0; s < n_soap; ++s) q[s] = ts[s];
      }
      double d = 0.0;
      for (int s = 0; s < n_soap; ++s) d = d + q[s] * q[s];
      d = sqrt(d);
      if (d < 1e-5) d = 1.0;
      sqrt_dot_p[i] = d;
    }
    free(ts);
  }
}

int so_get_soap(const so_params* p, const so_inputs* in, const double* W, const double* S, const so_compress* c,
                int do_der, double* soap, double* cart, double* sqrt_dot_p_out) {
  so_state st;
  build_state(&st, p, in, W, S, c, do_der && cart);
  const int n_max = st.n_max, k_max = st.k_max, L = p->l_max, n_soap = st.n_soap;
  double* sdp = (double*)malloc(sizeof(double) * in->n_sites);
  power_spectrum(&st, p, c, in->n_sites, soap, sdp);
  if (do_der && cart) {
    const double pi = acos(-1.0), fourpi = 4.0 * pi, sq2 = sqrt(2.0);
    /* soap_derivatives_from_expansion (:1927-2075) fused with the rotation (:837-895) */
#pragma omp parallel
    {
      double* tr = (double*)calloc(st.n_soap_unc, sizeof(double));
      double* ta = (double*)calloc(st.n_soap_unc, sizeof(double));
      double* to = (double*)calloc(st.n_soap_unc, sizeof(double));
      double* dr_ = (double*)malloc(sizeof(double) * n_soap);
      double* da = (double*)malloc(sizeof(double) * n_soap);
      double* dp = (double*)malloc(sizeof(double) * n_soap);
      double* P = (double*)malloc(sizeof(double) * n_max * (L + 1) * 4);
      double *Q = P + n_max * (L + 1), *Qa = Q + n_max * (L + 1), *Qo = Qa + n_max * (L + 1);
#pragma omp for schedule(dynamic, 16)
      for (int i = 0; i < in->n_sites; ++i) {
        const double* cn = st.cnk + 2 * (size_t)i * n_max * k_max;
        const double* q = soap + (size_t)i * n_soap;
        double inv_p = 1.0 / sdp[i];
        double inv_p3 = 1.0 / (sdp[i] * sdp[i] * sdp[i]);
        int k3 = st.k_first[i];
        for (int j = 1; j < in->n_neigh[i]; ++j) {
          int k2 = st.k_first[i] + j;
          const double* A = st.ang + 2 * (size_t)k2 * k_max;
          const double* Ar = st.ang_r + 2 * (size_t)k2 * k_max;
          const double* Aa = st.ang_a + 2 * (size_t)k2 * k_max;
          const double* Ap = st.ang_p + 2 * (size_t)k2 * k_max;
          for (int n = 0; n < n_max; ++n)
            for (int l = 0; l <= L; ++l) {
              int kb = l * (l + 1) / 2;
              double ap = 0, aq = 0, aa = 0, ao = 0;
              for (int m = 0; m <= l; ++m) {
                int k = kb + m;
                double w = m == 0 ? 1.0 : 2.0;
                double cr = cn[2 * (n * k_max + k)], ci = cn[2 * (n * k_max + k) + 1];
                ap = ap + w * (A[2 * k] * cr + A[2 * k + 1] * ci);
                aq = aq + w * (Ar[2 * k] * cr + Ar[2 * k + 1] * ci);
                aa = aa + w * (Aa[2 * k] * cr + Aa[2 * k + 1] * ci);
                ao = ao + w * (Ap[2 * k] * cr + Ap[2 * k + 1] * ci);
              }
              P[n + l * n_max] = ap; Q[n + l * n_max] = aq; Qa[n + l * n_max] = aa; Qo[n + l * n_max] = ao;
            }
          int cnt = 0;
          const double* r = st.rad + (size_t)k2 * n_max;
          const double* rd = st.rad_der + (size_t)k2 * n_max;
          for (int n = 0; n < n_max; ++n)
            for (int np = n; np < n_max; ++np) {
              double base = n == np ? 1.0 : sq2, pre = fourpi * base;
              for (int l = 0; l <= L; ++l, ++cnt) {
                if (st.skip[cnt]) continue;
                tr[cnt] = pre * (rd[n] * P[np + l * n_max] + r[n] * Q[np + l * n_max] + rd[np] * P[n + l * n_max] +
                                 r[np] * Q[n + l * n_max]);
                ta[cnt] = pre * (r[n] * Qa[np + l * n_max] + r[np] * Qa[n + l * n_max]);
                to[cnt] = pre * (r[n] * Qo[np + l * n_max] + r[np] * Qo[n + l * n_max]);
              }
            }
          if (c && c->compress) {
            for (int s = 0; s < n_soap; ++s) dr_[s] = da[s] = dp[s] = 0.0;
            for (int k = 0; k < c->P_nonzero; ++k) {
              int pi_ = c->P_i[k] - 1, pj = c->P_j[k] - 1;
              dr_[pi_] = dr_[pi_] + c->P_el[k] * tr[pj];
              da[pi_] = da[pi_] + c->P_el[k] * ta[pj];
              dp[pi_] = dp[pi_] + c->P_el[k] * to[pj];
            }
          } else {
            for (int s = 0; s < n_soap; ++s) { dr_[s] = tr[s]; da[s] = ta[s]; dp[s] = to[s]; }
          }
          double dot_r = 0, dot_a = 0, dot_o = 0;
          for (int s = 0; s < n_soap; ++s) {
            dot_r = dot_r + q[s] * dr_[s];
            dot_a = dot_a + q[s] * da[s];
            dot_o = dot_o + q[s] * dp[s];
          }
          double sin_t = sin(in->thetas[k2]), cos_t = cos(in->thetas[k2]);
          double sin_f = sin(in->phis[k2]), cos_f = cos(in->phis[k2]);
          double inv_r = 1.0 / in->rjs[k2];
          double cx = sin_t * cos_f, cy = sin_t * sin_f, cz = cos_t;
          double px = -(cos_t * cos_f * inv_r), py = -(cos_t * sin_f * inv_r), pz = sin_t * inv_r;
          double ax = -(sin_f * inv_r), ay = cos_f * inv_r;
          double* C = cart + (size_t)k2 * n_soap * 3;
          double* C3 = cart + (size_t)k3 * n_soap * 3;
          for (int s = 0; s < n_soap; ++s) {
            double sv = q[s];
            double vr = dr_[s] * inv_p - sv * inv_p3 * dot_r;
            double va = da[s] * inv_p - sv * inv_p3 * dot_a;
            double vo = dp[s] * inv_p - sv * inv_p3 * dot_o;
            C[3 * s] = cx * vr + px * vo + ax * va;
            C[3 * s + 1] = cy * vr + py * vo + ay * va;
            C[3 * s + 2] = cz * vr + pz * vo;
          }
          for (int s = 0; s < n_soap; ++s) {
            C3[3 * s] = C3[3 * s] - C[3 * s];
            C3[3 * s + 1] = C3[3 * s + 1] - C[3 * s + 1];
            C3[3 * s + 2] = C3[3 * s + 2] - C[3 * s + 2];
          }
        }
      }
      free(tr); free(ta); free(to); free(dr_); free(da); free(dp); free(P);
    }
  }
  /* normalise, :904-909 */
  for (int i = 0; i < in->n_sites; ++i)
    for (int s = 0; s < n_soap; ++s) soap[(size_t)i * n_soap + s] = soap[(size_t)i * n_soap + s] / sdp[i];
  if (sqrt_dot_p_out) memcpy(sqrt_dot_p_out, sdp, sizeof(double) * in->n_sites);
  free(sdp);
  state_free(&st);
  return 0;
}

/* get_soap(contract = .true.) then soap_contract: src/soap_turbo.f90:1723-1884.
 * soap is written (normalised) as get_soap would; pair receives the gradients. */
int so_contract(const so_params* p, const so_inputs* in, const double* W, const double* S, const so_compress* c,
                const double* G, double sgn, double* soap, double* pair) {
  so_state st;
  build_state(&st, p, in, W, S, c, 1);
  const int n_max = st.n_max, k_max = st.k_max, L = p->l_max, n_soap = st.n_soap;
  double* sdp = (double*)malloc(sizeof(double) * in->n_sites);
  power_spectrum(&st, p, c, in->n_sites, soap, sdp);
  for (int i = 0; i < in->n_sites; ++i)
    for (int s = 0; s < n_soap; ++s) soap[(size_t)i * n_soap + s] = soap[(size_t)i * n_soap + s] / sdp[i];
  const double fourpi = 4.0 * acos(-1.0), sq2 = sqrt(2.0);
#pragma omp parallel
  {
    double* Gp = (double*)malloc(sizeof(double) * n_soap);
    double* Gu = (double*)malloc(sizeof(double) * st.n_soap_unc);
    double* Y = (double*)malloc(sizeof(double) * n_max * n_max * (L + 1));
    double* H = (double*)malloc(sizeof(double) * 2 * k_max * n_max);
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < in->n_sites; ++i) {
      const double* q = soap + (size_t)i * n_soap;
      double qg = 0.0;
      for (int k = 0; k < n_soap; ++k) qg = qg + q[k] * G[(size_t)k * in->n_sites + i];
      double inv = 1.0 / sdp[i];
      for (int k = 0; k < n_soap; ++k) Gp[k] = (G[(size_t)k * in->n_sites + i] - q[k] * qg) * inv;
      if (c && c->compress) {
        for (int k = 0; k < st.n_soap_unc; ++k) Gu[k] = 0.0;
        for (int k = 0; k < c->P_nonzero; ++k) Gu[c->P_j[k] - 1] = Gu[c->P_j[k] - 1] + c->P_el[k] * Gp[c->P_i[k] - 1];
      } else {
        for (int k = 0; k < n_soap; ++k) Gu[k] = Gp[k];
      }
      int cnt = 0;
      for (int n = 0; n < n_max; ++n)
        for (int np = n; np < n_max; ++np) {
          double base = n == np ? 1.0 : sq2;
          for (int l = 0; l <= L; ++l, ++cnt) {
            double x = 0.0;
            if (!st.skip[cnt]) x = Gu[cnt] * base;
            if (n == np) Y[n + n * n_max + (size_t)l * n_max * n_max] = 2.0 * x;
            else {
              Y[n + np * n_max + (size_t)l * n_max * n_max] = x;
              Y[np + n * n_max + (size_t)l * n_max * n_max] = x;
            }
          }
        }
      const double* cn = st.cnk + 2 * (size_t)i * n_max * k_max;
      for (int n = 0; n < n_max; ++n)
        for (int l = 0; l <= L; ++l) {
          int kb = l * (l + 1) / 2;
          for (int m = 0; m <= l; ++m) {
            int k = kb + m;
            double hr = 0.0, hi = 0.0;
            for (int np = 0; np < n_max; ++np) {
              double y = Y[n + np * n_max + (size_t)l * n_max * n_max];
              hr = hr + y * cn[2 * (np * k_max + k)];
              hi = hi + y * cn[2 * (np * k_max + k) + 1];
            }
            double w = m == 0 ? 1.0 : 2.0;
            H[2 * (n * k_max + k)] = w * hr;
            H[2 * (n * k_max + k) + 1] = w * hi;
          }
        }
      int k3 = st.k_first[i];
      pair[3 * (size_t)k3] = pair[3 * (size_t)k3 + 1] = pair[3 * (size_t)k3 + 2] = 0.0;
      for (int j = 1; j < in->n_neigh[i]; ++j) {
        int k2 = st.k_first[i] + j;
        const double* A = st.ang + 2 * (size_t)k2 * k_max;
        const double* Ar = st.ang_r + 2 * (size_t)k2 * k_max;
        const double* Aa = st.ang_a + 2 * (size_t)k2 * k_max;
        const double* Ap = st.ang_p + 2 * (size_t)k2 * k_max;
        const double* r = st.rad + (size_t)k2 * n_max;
        const double* rd = st.rad_der + (size_t)k2 * n_max;
        double f_rad = 0, f_azi = 0, f_pol = 0;
        for (int n = 0; n < n_max; ++n) {
          double sd = 0, sr = 0, sa = 0, so = 0;
          for (int k = 0; k < k_max; ++k) {
            double hr = H[2 * (n * k_max + k)], hi = H[2 * (n * k_max + k) + 1];
            sd = sd + A[2 * k] * hr + A[2 * k + 1] * hi;
            sr = sr + Ar[2 * k] * hr + Ar[2 * k + 1] * hi;
            sa = sa + Aa[2 * k] * hr + Aa[2 * k + 1] * hi;
            so = so + Ap[2 * k] * hr + Ap[2 * k + 1] * hi;
          }
          f_rad = f_rad + rd[n] * sd + r[n] * sr;
          f_azi = f_azi + r[n] * sa;
          f_pol = f_pol + r[n] * so;
        }
        f_rad = fourpi * f_rad; f_azi = fourpi * f_azi; f_pol = fourpi * f_pol;
        double sin_t = sin(in->thetas[k2]), cos_t = cos(in->thetas[k2]);
        double sin_f = sin(in->phis[k2]), cos_f = cos(in->phis[k2]);
        double inv_r = 1.0 / in->rjs[k2];
        double* pr = pair + 3 * (size_t)k2;
        pr[0] = sgn * (sin_t * cos_f * f_rad - cos_t * cos_f * inv_r * f_pol - sin_f * inv_r * f_azi);
        pr[1] = sgn * (sin_t * sin_f * f_rad - cos_t * sin_f * inv_r * f_pol + cos_f * inv_r * f_azi);
        pr[2] = sgn * (cos_t * f_rad + sin_t * inv_r * f_pol);
        pair[3 * (size_t)k3] -= pr[0];
        pair[3 * (size_t)k3 + 1] -= pr[1];
        pair[3 * (size_t)k3 + 2] -= pr[2];
      }
    }
    free(Gp); free(Gu); free(Y); free(H);
  }
  free(sdp);
  state_free(&st);
  return 0;
}
