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
  int alpha_max, n_poly, n_hi, do_poly, with_gauss, do_der, do_der2, unified, has_filter, scaling_poly, radial_enh;
  int do_central;
  double rcut_soft, atom_sigma, dr, inv_rcut, sqrt_rcut, inv_sqrt_rcut, inv_Nm2, sigma_sq0, N_gauss, q, drnf;
  double scaling, amp_scaling, central_weight;
  double ra[MAXA], rb[MAXA], rc[MAXA], R1[MAXA], R2[MAXA], R3[MAXA], R4[MAXA]; /* ra.. at n + 1 (ra(-1) at 0) */
  const double* W; int ldw;                                /* species block of W */
} radial_consts;

/* do_der: 0 values, 1 first derivatives, 2 also second derivatives (unified form only,
 * src/soap_turbo_radial.f90:308-320) */
static void radial_setup(radial_consts* c, const so_params* p, int s, const double* W, int ldw, int off, int do_der) {
  memset(c, 0, sizeof *c);
  c->alpha_max = p->alpha_max[s];
  c->with_gauss = p->basis == 1;
  c->do_der = do_der >= 1;
  c->do_der2 = do_der >= 2;
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
  c->n_hi = c->n_poly + (c->do_der2 ? 4 : c->do_der ? 2 : 0);
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
    c->R3[n] = Nv[n + 3 + 3] / Nv[n + 3];
    c->R4[n] = Nv[n + 4 + 3] / Nv[n + 3];
  }
  c->W = W + off + (size_t)off * ldw;
  c->ldw = ldw;
}

static void radial_pair(const radial_consts* c, double rjs, int is_central, double* outv, double* outd, double* outh) {
  const double pi = acos(-1.0), sq2 = sqrt(2.0), isq2 = 1.0 / sq2, sqpi2 = sqrt(pi / 2.0),
               sq2pi = sqrt(2.0 / pi), sq8pi = sqrt(8.0 / pi);
  const double sc = c->scaling, rs = c->rcut_soft;
  double T1[MAXA], T2[MAXA], pol[MAXA], DT[MAXA], D2T[MAXA]; /* T(n) at n + 3; pol(n), DT(n) at n */
  memset(T2, 0, sizeof T2);
  memset(DT, 0, sizeof DT);
  memset(D2T, 0, sizeof D2T);
  double rj = rjs * c->inv_rcut;
  double sig = c->atom_sigma + sc * rj, s2 = sig * sig, isig = 1.0 / sig;
  double is2 = isig * isig, is3 = is2 * isig;
  double amp, ampd, ampd2;
  /* amplitude and its first two derivatives, :572-635 */
  if (!c->scaling_poly) {
    amp = 1.0; ampd = 0.0; ampd2 = 0.0;
  } else if (c->amp_scaling == 0.0) {
    amp = isig; ampd = -sc * is2; ampd2 = 2.0 * sc * sc * is3;
  } else if (c->amp_scaling == 1.0) {
    double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
    double hp = 6.0 * (rj * rj) - 6.0 * rj;
    double hpp = 12.0 * rj - 6.0;
    double g = isig * h;
    amp = h <= 1e-10 ? 0.0 : g;
    ampd = h <= 1e-10 ? 0.0 : hp * isig - sc * isig * g;
    ampd2 = h <= 1e-10 ? 0.0 : hpp * isig - 2.0 * sc * hp * is2 + 2.0 * sc * sc * h * is3;
  } else {
    double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
    if (h <= 1e-10) {
      amp = 0.0; ampd = 0.0; ampd2 = 0.0;
    } else {
      double hp = 6.0 * (rj * rj) - 6.0 * rj, hpp = 12.0 * rj - 6.0, aa = c->amp_scaling;
      amp = isig * pow(h, aa);
      ampd = aa * pow(h, aa - 1.0) * hp * isig - sc * isig * amp;
      ampd2 = (aa * (aa - 1.0) * pow(h, aa - 2.0) * hp * hp + aa * pow(h, aa - 1.0) * hpp) * isig -
              2.0 * sc * aa * pow(h, aa - 1.0) * hp * is2 + 2.0 * sc * sc * pow(h, aa) * is3;
    }
  }
  /* central weight, :637-646 */
  if (c->do_central && c->central_weight != 1.0 && is_central) {
    amp = c->central_weight * amp;
    ampd = c->central_weight * ampd;
    ampd2 = c->central_weight * ampd2;
  }
  /* radial enhancement, :648-670: (a m)'' = a'' m + 2 a' m' + a m'' */
  if (c->radial_enh == 1) {
    double g = rj + sq2pi * sig, f = 1.0 + sq2pi * sc;
    ampd2 = ampd2 * g + 2.0 * ampd * f;
    ampd = amp * f + ampd * g;
    amp = amp * g;
  } else if (c->radial_enh == 2) {
    double g = (rj * rj) + s2 + sq8pi * sig * rj;
    double f = 2.0 * rj + 2.0 * sig * sc + sq8pi * sig + sq8pi * rj * sc;
    double h = 2.0 + 2.0 * sc * sc + 2.0 * sq8pi * sc;
    ampd2 = ampd2 * g + 2.0 * ampd * f + amp * h;
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
    /* second derivative, :838-861 */
    if (c->do_der2) {
      double is4 = is2 * is2, is5 = is4 * isig, is6 = is4 * is2;
      double k0 = -is2, k1 = -4.0 * sc * is3, k2 = (1.0 - 3.0 * sc * sc) * is4, k3 = 2.0 * sc * is5, k4 = sc * sc * is6;
      double B0 = k0 + k1 * f + k2 * f * f + k3 * (f * f * f) + k4 * (f * f * f * f);
      double B1 = -(k1 + 2.0 * k2 * f + 3.0 * k3 * f * f + 4.0 * k4 * (f * f * f));
      double B2 = k2 + 3.0 * k3 * f + 6.0 * k4 * f * f;
      double B3 = -(k3 + 4.0 * k4 * f);
      double B4 = k4;
      for (int n = 1; n <= c->n_poly; ++n)
        D2T[n] = B0 * pol[n] + c->R1[n] * B1 * pol[n + 1] + c->R2[n] * B2 * pol[n + 2] + c->R3[n] * B3 * pol[n + 3] +
                 c->R4[n] * B4 * pol[n + 4];
    }
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
      if (c->do_der2) {
        double M3 = sc0 * ge * (rj * rj) - pp * M2 + 2.0 * sc0 * M1;
        double M4 = sc0 * ge * (-(rj * rj * rj)) - pp * M3 + 3.0 * sc0 * M2;
        double is4 = is2 * is2, is5 = is4 * isig, is6 = is4 * is2;
        D2T[c->alpha_max] = -M0 * is2 - 4.0 * sc * M1 * is3 + (1.0 - 3.0 * sc * sc) * M2 * is4 + 2.0 * sc * M3 * is5 +
                            sc * sc * M4 * is6;
      }
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
    if (c->do_der2) {
      double d2tot[MAXA];
      for (int n = 1; n <= a; ++n) d2tot[n] = amp * D2T[n] + 2.0 * ampd * DT[n] + ampd2 * pol[n];
      for (int n = 1; n <= a; ++n) {
        double acc = c->W[(n - 1)] * d2tot[1];
        for (int m = 2; m <= a; ++m) acc = acc + c->W[(n - 1) + (size_t)(m - 1) * c->ldw] * d2tot[m];
        outh[n - 1] = acc * c->inv_sqrt_rcut * c->inv_rcut;
      }
    }
  }
  for (int n = 1; n <= a; ++n) {
    double acc = c->W[(n - 1)] * pol[1];
    for (int m = 2; m <= a; ++m) acc = acc + c->W[(n - 1) + (size_t)(m - 1) * c->ldw] * pol[m];
    outv[n - 1] = amp * acc * c->sqrt_rcut;
  }
}

/* ---- poly3operator: get_radial_expansion_coefficients_poly3operator,
 * src/soap_turbo_radial.f90:1549-2141, for one pair of one species (scaled units, rcut_hard = 1).
 * The atom is a piecewise cubic of half-width w = 2 sqrt(2 ln 2) sigma(r_j) about r_j
 * (g_aux, :2154-2171); inside rcut_soft it is integrated against the polynomial basis directly
 * (the "soft" region, :1598-1842), between rcut_soft and rcut_hard multiplied by the cubic filter
 * (the "buffer" region, :1845-2087). Every integral is a polynomial moment between two limits:
 * the A matrix (get_constant_poly_coeff, :2276-2297) holds the basis side, and the atom (and
 * filter) side enters through its derivatives at the limits. */
static void op_g(double r, double r0, double winv, int right, double* g) {
  /* g_aux_array, :2179-2217: the cubic and its scaled derivatives */
  double x = (r - r0) * winv, x2 = x * x, x3 = x2 * x, w2 = winv * winv, w3 = w2 * winv;
  if (!right) {
    g[0] = 1.0 - 3.0 * x2 - 2.0 * x3; g[1] = -6.0 * (x2 + x) * winv; g[2] = -3.0 * (2.0 * x + 1.0) * w2; g[3] = -2.0 * w3;
  } else {
    g[0] = 1.0 - 3.0 * x2 + 2.0 * x3; g[1] = 6.0 * (x2 - x) * winv; g[2] = 3.0 * (2.0 * x - 1.0) * w2; g[3] = 2.0 * w3;
  }
}

static void op_gd(double r, double r0, double winv, double ws, int right, double* g) {
  /* g_aux_der_array, :2227-2267: derivatives with respect to r0, width = width0 + ws r0 */
  double x = (r - r0) * winv, x2 = x * x, x3 = x2 * x, w2 = winv * winv, w3 = w2 * winv, w4 = w3 * winv;
  if (!right) {
    g[0] = 6.0 * (x + (ws + 1.0) * x2 + ws * x3) * winv;
    g[1] = 6.0 * (1.0 + 2.0 * (ws + 1.0) * x + 3.0 * ws * x2) * w2;
    g[2] = 6.0 * (ws + 1.0 + 3.0 * ws * x) * w3;
    g[3] = 6.0 * ws * w4;
  } else {
    g[0] = 6.0 * (x + (ws - 1.0) * x2 - ws * x3) * winv;
    g[1] = 6.0 * (1.0 + 2.0 * (ws - 1.0) * x - 3.0 * ws * x2) * w2;
    g[2] = 6.0 * (ws - 1.0 - 3.0 * ws * x) * w3;
    g[3] = -6.0 * ws * w4;
  }
}

static void op_powers(double r, int n, double* t) { /* M_radial_poly_array, :2399-2418: (1 - r)^(i-1) */
  t[0] = 1.0;
  for (int i = 1; i < n; ++i) t[i] = t[i - 1] * (1.0 - r);
}

/* -B . M_radial_monomial(r, 6), :2481-2500: minus the derivatives of the degree-6 polynomial with
 * coefficients B at r, v(q) = -sum_p B(p) (p-1)!/(p-q)! r^(p-q) */
static void op_bm(const double* B, double r, double* v) {
  double rt[7];
  rt[0] = 1.0;
  for (int p = 1; p < 7; ++p) rt[p] = r * rt[p - 1];
  for (int q = 0; q < 7; ++q) {
    double acc = 0.0;
    for (int pp = q; pp < 7; ++pp) {
      double f = 1.0;
      for (int t = pp - q + 1; t <= pp; ++t) f *= (double)t;
      acc += -B[pp] * (rt[pp - q] * f);
    }
    v[q] = acc;
  }
}

/* get_constant_poly_filter_coeff_array (and _der), :2307-2390: the degree-6 product of the atom's
 * cubic about 0 at -r_j and the filter's cubic at 0 */
static void op_filter_B(double rj, double winv, double ws, double rs, double fw, int right, int der, double* B) {
  double cf[4], col[4];
  { /* g_aux(0, r_filter, filter_width, "right"), :2154-2171 */
    double x = (0.0 - rs) / fw;
    cf[0] = 1.0 - 3.0 * (x * x) + 2.0 * (x * x * x);
    cf[1] = 6.0 * (x * x - x) / fw;
    cf[2] = 3.0 * (2.0 * x - 1.0) / (fw * fw);
    cf[3] = 2.0 / (fw * fw * fw);
  }
  if (der) op_gd(-rj, 0.0, winv, ws, right, col);
  else op_g(-rj, 0.0, winv, right, col);
  for (int r = 0; r < 7; ++r) {
    double acc = 0.0;
    for (int i = 0; i < 4; ++i) {
      int t = r - i;
      acc += (t >= 0 && t < 4 ? col[t] : 0.0) * cf[i];
    }
    B[r] = acc;
  }
}

void so_op_pair(const so_params* p, int s, double rjs_in, int is_central, double* outv, double* outd) {
  const int a = p->alpha_max[s];
  const double pi = acos(-1.0), K = 2.0 * sqrt(2.0 * log(2.0));
  const double rc_in = p->rcut_hard[s], rs_in = p->rcut_soft[s], sig_in = p->atom_sigma_r[s];
  const double sc = p->atom_sigma_r_scaling[s], aa = p->amplitude_scaling[s], cw = p->central_weight[s];
  const double rs = rs_in / rc_in, sig0 = sig_in / rc_in, fw = K * (1.0 - rs);
  double A[MAXA][7];
  for (int i = 1; i <= a; ++i) { /* get_constant_poly_coeff(alpha_max, 1, A) */
    A[i - 1][0] = 1.0 / sqrt(1.0 / (double)(2 * i + 5)) / (double)(i + 3);
    for (int j = 2; j <= 7; ++j) A[i - 1][j - 1] = A[i - 1][j - 2] * 1.0 / (double)(i + 3 + j - 1);
  }
  double ex[MAXA], exd[MAXA];
  for (int n = 0; n < a; ++n) ex[n] = exd[n] = 0.0;
  const double win = K * (sig_in + sc * rjs_in);  /* unscaled width, for the region tests */
  const int soft = rjs_in - win < rs_in;
  const int buffer = rs_in < rc_in && rjs_in + win > rs_in;
  const double rj = rjs_in / rc_in, sg = sig0 + sc * rj, s2 = sg * sg, w = K * sg, winv = 1.0 / w, ws = K * sc;
  double amp, ampd;
  /* amplitude, polynomial scaling only (:1701-1727) */
  double h = 1.0 + 2.0 * (rj * rj * rj) - 3.0 * (rj * rj);
  if (aa == 0.0) {
    amp = 1.0 / sg; ampd = -sc / s2;
  } else if (aa == 1.0) {
    amp = 1.0 / sg * h;
    ampd = 6.0 / sg * (rj * rj - rj) - sc / sg * amp;
  } else {
    amp = 1.0 / sg * pow(h, aa);
    ampd = 6.0 * aa / sg * (rj * rj - rj) * pow(h, aa - 1.0) - sc / sg * amp;
  }
  if (is_central) { amp = cw * amp; ampd = cw * ampd; }
  if (p->radial_enhancement == 1) {
    ampd = amp * (1.0 + sqrt(2.0 / pi) * sc) + ampd * (rj + sqrt(2.0 / pi) * sg);
    amp = amp * (rj + sqrt(2.0 / pi) * sg);
  } else if (p->radial_enhancement == 2) {
    ampd = amp * (2.0 * rj + 2.0 * sg * sc + sqrt(8.0 / pi) * sg + sqrt(8.0 / pi) * rj * sc) +
           ampd * (rj * rj + s2 + sqrt(8.0 / pi) * sg * rj);
    amp = amp * (rj * rj + s2 + sqrt(8.0 / pi) * sg * rj);
  }
  static const double vect[4] = {-1.0, -1.0, -2.0, -6.0};
  const int NT = (a + 4 > 7 ? a + 4 : 7);
  if (soft) {
    double sv[MAXA], sd[MAXA];
    double L[3] = {rj - w > 0.0 ? rj - w : 0.0, rj < rs ? rj : rs, rs < rj + w ? rs : rj + w};
    double I0[3][MAXA + 8];
    for (int t = 0; t < 3; ++t) op_powers(L[t], a + 4, I0[t]);
    for (int d = 0; d < 2; ++d) {
      double M[3][2][4]; /* [limit][piece 0 left 1 right][j] */
      for (int t = 0; t < 3; ++t)
        for (int pc = 0; pc < 2; ++pc) {
          double g[4];
          if (d == 0) op_g(L[t], rj, winv, pc, g);
          else op_gd(L[t], rj, winv, ws, pc, g);
          for (int j = 0; j < 4; ++j) M[t][pc][j] = I0[t][j] * g[j] * vect[j];
        }
      for (int n = 0; n < a; ++n) {
        double s12 = 0, s11 = 0, s23 = 0, s22 = 0;
        for (int j = 0; j < 4; ++j) {
          s12 += M[1][0][j] * A[n][j]; s11 += M[0][0][j] * A[n][j];
          s23 += M[2][1][j] * A[n][j]; s22 += M[1][1][j] * A[n][j];
        }
        double il = s12 * I0[1][n + 4] - s11 * I0[0][n + 4];
        double ir = s23 * I0[2][n + 4] - s22 * I0[1][n + 4];
        if (d == 0) sv[n] = il + ir;
        else sd[n] = il + ir;
      }
    }
    for (int n = 0; n < a; ++n) {
      ex[n] = amp * sv[n];
      exd[n] = amp * sd[n] + ampd * sv[n];
    }
  }
  if (buffer) {
    double L[3] = {rs > rj - w ? rs : rj - w, rj > rs ? rj : rs, 1.0 < rj + w ? 1.0 : rj + w};
    double I0[3][MAXA + 8];
    for (int t = 0; t < 3; ++t) op_powers(L[t], NT, I0[t]);
    double bd[MAXA], bb[MAXA];
    for (int d = 0; d < 2; ++d) {
      double B[2][7], M[3][2][7];
      for (int pc = 0; pc < 2; ++pc) op_filter_B(rj, winv, ws, rs, fw, pc, d, B[pc]);
      for (int t = 0; t < 3; ++t)
        for (int pc = 0; pc < 2; ++pc) {
          double v[7];
          op_bm(B[pc], L[t], v);
          for (int q = 0; q < 7; ++q) M[t][pc][q] = v[q] * I0[t][q];
        }
      for (int n = 0; n < a; ++n) {
        double s12 = 0, s11 = 0, s23 = 0, s22 = 0;
        for (int q = 0; q < 7; ++q) {
          s12 += M[1][0][q] * A[n][q]; s11 += M[0][0][q] * A[n][q];
          s23 += M[2][1][q] * A[n][q]; s22 += M[1][1][q] * A[n][q];
        }
        double v = (s12 * I0[1][n + 4] - s11 * I0[0][n + 4]) + (s23 * I0[2][n + 4] - s22 * I0[1][n + 4]);
        if (d == 0) bb[n] = v;
        else bd[n] = v;
      }
    }
    for (int n = 0; n < a; ++n) {
      ex[n] += amp * bb[n];
      exd[n] += amp * bd[n] + ampd * bb[n];
    }
  }
  for (int n = 0; n < a; ++n) { outv[n] = ex[n]; outd[n] = exd[n]; }
}

/* get_radial_expansion_coefficients_poly3[gauss] for every species, then
 * global_scaling (src/soap_turbo.f90:491-544; src/soap_turbo_radial.f90:116-384). */
void so_radial(const so_params* p, const so_inputs* in, const double* W, int do_der, double* rad, double* rad_der) {
  so_radial2(p, in, W, do_der, rad, rad_der, NULL);
}

/* The same with the second derivatives (rad_der2 non-NULL; do_der must be set and the
 * unified form selected, as for the reference, src/soap_turbo_radial.f90:308-320). */
void so_radial2(const so_params* p, const so_inputs* in, const double* W, int do_der, double* rad, double* rad_der,
                double* rad_der2) {
  const int n_max = n_max_of(p);
  memset(rad, 0, sizeof(double) * (size_t)n_max * in->n_pairs);
  if (do_der) memset(rad_der, 0, sizeof(double) * (size_t)n_max * in->n_pairs);
  if (rad_der2) memset(rad_der2, 0, sizeof(double) * (size_t)n_max * in->n_pairs);
  int* first = (int*)malloc(sizeof(int) * (in->n_pairs + 1));
  {
    int k = 0;
    for (int i = 0; i < in->n_sites; ++i)
      for (int j = 0; j < in->n_neigh[i]; ++j) first[k++] = (j == 0);
  }
  int off = 0;
  if (p->basis == 2) {
    /* poly3operator: every masked pair of the species (no cutoff test, the self pair weighted by
     * central_weight), then W, the sqrt(rcut_hard) of the change of variable and global_scaling
     * (src/soap_turbo_radial.f90:2089-2107, src/soap_turbo.f90:534-541) */
    for (int s = 0; s < p->n_species; ++s) {
      const int a = p->alpha_max[s];
      const double gs = p->global_scaling[s], sq = sqrt(p->rcut_hard[s]);
#pragma omp parallel for schedule(dynamic, 256)
      for (int k = 0; k < in->n_pairs; ++k) {
        if (!in->mask[k + (size_t)s * in->n_pairs]) continue;
        double v[MAXA], d[MAXA];
        so_op_pair(p, s, in->rjs[k], first[k], v, d);
        for (int n = 0; n < a; ++n) {
          double av = 0.0, ad = 0.0;
          for (int m = 0; m < a; ++m) {
            av += W[(off + n) + (size_t)(off + m) * n_max] * v[m];
            ad += W[(off + n) + (size_t)(off + m) * n_max] * d[m];
          }
          rad[(size_t)k * n_max + off + n] = av * sq * gs;
          if (do_der) rad_der[(size_t)k * n_max + off + n] = ad / sq * gs;
        }
      }
      off += a;
    }
    free(first);
    return;
  }
  for (int s = 0; s < p->n_species; ++s) {
    radial_consts c;
    radial_setup(&c, p, s, W, n_max, off, rad_der2 ? 2 : do_der);
    const double gs = p->global_scaling[s];
#pragma omp parallel for schedule(dynamic, 256)
    for (int k = 0; k < in->n_pairs; ++k) {
      if (first[k] && !c.do_central) continue;
      if (in->rjs[k] >= p->rcut_hard[s]) continue;
      if (!in->mask[k + (size_t)s * in->n_pairs]) continue;
      double v[MAXA], d[MAXA], h[MAXA];
      radial_pair(&c, in->rjs[k], first[k], v, d, h);
      for (int n = 0; n < c.alpha_max; ++n) {
        rad[(size_t)k * n_max + off + n] = v[n] * gs;
        if (do_der) rad_der[(size_t)k * n_max + off + n] = d[n] * gs;
        if (rad_der2) rad_der2[(size_t)k * n_max + off + n] = h[n] * gs;
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

/* ------------------------------------------------------------------------------------------
 * Central-atom Hessian, get_soap_central_hessian, src/soap_turbo.f90:1151-1700: for every pair
 * k2 = (site i, neighbour j), hess(k2, v, a, b) = sum_d vecs(i, v, d) d^2 q_d / dr_ia dr_jb, with
 * q the normalised (compressed) descriptor; the self pair holds minus the sum of the others.
 * Needs the unified radial form (legacy_seed = 0): the second radial derivatives exist only there.
 * ------------------------------------------------------------------------------------------ */
#include <complex.h>

/* the unnormalised Cartesian derivative of the compressed descriptor for one pair (the terms of
 * soap_derivatives_from_expansion, :1927-2075, compressed, then rotated as :837-876 without the
 * normalisation): d[3 s + a] */
static void pair_unnorm_der(const so_state* st, const so_params* p, const so_compress* c, const so_inputs* in,
                            const double* cn, int k2, double* tr, double* ta, double* to, double* P, double* d) {
  const int n_max = st->n_max, k_max = st->k_max, L = p->l_max, n_soap = st->n_soap;
  const double fourpi = 4.0 * acos(-1.0), sq2 = sqrt(2.0);
  double *Q = P + n_max * (L + 1), *Qa = Q + n_max * (L + 1), *Qo = Qa + n_max * (L + 1);
  const double* A = st->ang + 2 * (size_t)k2 * k_max;
  const double* Ar = st->ang_r + 2 * (size_t)k2 * k_max;
  const double* Aa = st->ang_a + 2 * (size_t)k2 * k_max;
  const double* Ap = st->ang_p + 2 * (size_t)k2 * k_max;
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
  const double* r = st->rad + (size_t)k2 * n_max;
  const double* rd = st->rad_der + (size_t)k2 * n_max;
  for (int n = 0; n < n_max; ++n)
    for (int np = n; np < n_max; ++np) {
      double base = n == np ? 1.0 : sq2, pre = fourpi * base;
      for (int l = 0; l <= L; ++l, ++cnt) {
        tr[cnt] = ta[cnt] = to[cnt] = 0.0;
        if (st->skip[cnt]) continue;
        tr[cnt] = pre * (rd[n] * P[np + l * n_max] + r[n] * Q[np + l * n_max] + rd[np] * P[n + l * n_max] +
                         r[np] * Q[n + l * n_max]);
        ta[cnt] = pre * (r[n] * Qa[np + l * n_max] + r[np] * Qa[n + l * n_max]);
        to[cnt] = pre * (r[n] * Qo[np + l * n_max] + r[np] * Qo[n + l * n_max]);
      }
    }
  double sin_t = sin(in->thetas[k2]), cos_t = cos(in->thetas[k2]);
  double sin_f = sin(in->phis[k2]), cos_f = cos(in->phis[k2]);
  double inv_r = 1.0 / in->rjs[k2];
  double cx = sin_t * cos_f, cy = sin_t * sin_f, cz = cos_t;
  double px = -cos_t * cos_f * inv_r, py = -cos_t * sin_f * inv_r, pz = sin_t * inv_r;
  double ax = -sin_f * inv_r, ay = cos_f * inv_r;
  for (int s = 0; s < 3 * n_soap; ++s) d[s] = 0.0;
  for (int e = 0; e < (c && c->compress ? c->P_nonzero : n_soap); ++e) {
    int s = c && c->compress ? c->P_i[e] - 1 : e, u = c && c->compress ? c->P_j[e] - 1 : e;
    double w = c && c->compress ? c->P_el[e] : 1.0;
    d[3 * s] += w * (cx * tr[u] + px * to[u] + ax * ta[u]);
    d[3 * s + 1] += w * (cy * tr[u] + py * to[u] + ay * ta[u]);
    d[3 * s + 2] += w * (cz * tr[u] + pz * to[u]);
  }
}

typedef struct {
  int k_max, l_max;
  const int *lofk, *mofk;
  const double* Ck;
  double nh[3], invr;
  double Fl[16], Fl1[16], Fl2[16];
  double complex *Om, *dOm, *d2Om; /* dOm[3 k + a], d2Om[9 k + 3 a + b] */
} hess_geom;

/* pair_geometry, src/soap_turbo.f90:1585-1697: F_l and its two radial derivatives, and
 * Omega_k = C_k Q_l^m(t) w^m with its Cartesian gradient and Hessian */
static int pair_geometry(hess_geom* g, const so_params* p, const so_inputs* in, int k2, double rcut_max, int want2) {
  const int L = g->l_max, k_max = g->k_max;
  double rr = in->rjs[k2];
  if (rr >= rcut_max || rr < 1e-10) return 0;
  int isp = -1;
  for (int s = 0; s < p->n_species; ++s)
    if (in->mask[k2 + (size_t)s * in->n_pairs]) { isp = s; break; }
  if (isp < 0) return 0;
  double invr = 1.0 / rr, invr2 = invr * invr;
  g->invr = invr;
  double sscal = p->atom_sigma_t_scaling[isp];
  double sig = p->atom_sigma_t[isp] + sscal * rr;
  double amp = rcut_max * rcut_max / (sig * sig);
  double ampd = -2.0 * sscal / sig * amp;
  double ampdd = 6.0 * (sscal * sscal) / (sig * sig) * amp;
  double xarg = rr / sig, zz = xarg * xarg, invz = 1.0 / zz;
  double il[20], dilz[16], d2ilz[16];
  ilexp(L + 2, xarg, il);
  for (int l = 0; l <= L; ++l) {
    double rl = (double)l;
    dilz[l] = il[l + 1] + (rl * invz - 1.0) * il[l];
    d2ilz[l] = il[l + 2] + ((double)(2 * l + 1) * invz - 2.0) * il[l + 1] +
               (rl * (rl - 1.0) * invz * invz - 2.0 * rl * invz + 1.0) * il[l];
  }
  double zp = 2.0 * rr / (sig * sig) - 2.0 * (rr * rr) * sscal / (sig * sig * sig);
  double zp2 = 2.0 / (sig * sig) - 8.0 * rr * sscal / (sig * sig * sig) +
               6.0 * (rr * rr) * (sscal * sscal) / (sig * sig * sig * sig);
  for (int l = 0; l <= L; ++l) {
    double f0 = il[l], f1 = dilz[l] * zp, f2 = d2ilz[l] * zp * zp + dilz[l] * zp2;
    g->Fl[l] = amp * f0;
    g->Fl1[l] = ampd * f0 + amp * f1;
    g->Fl2[l] = ampdd * f0 + 2.0 * ampd * f1 + amp * f2;
  }
  double sin_t = sin(in->thetas[k2]), cos_t = cos(in->thetas[k2]);
  double sin_f = sin(in->phis[k2]), cos_f = cos(in->phis[k2]);
  double* nh = g->nh;
  nh[0] = sin_t * cos_f; nh[1] = sin_t * sin_f; nh[2] = cos_t;
  double tt = nh[2];
  double complex wc = nh[0] - I * nh[1];
  const double complex ec[3] = {1.0, -I, 0.0};
  double tvec[3], thes[3][3];
  double complex wvec[3], whes[3][3];
  for (int a = 0; a < 3; ++a) {
    tvec[a] = -tt * nh[a] * invr;
    wvec[a] = -wc * nh[a] * invr + ec[a] * invr;
  }
  tvec[2] += invr;
  if (want2) {
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a) {
        thes[a][b] = 3.0 * tt * nh[a] * nh[b] * invr2;
        whes[a][b] = (3.0 * wc * nh[a] * nh[b] - ec[a] * nh[b] - ec[b] * nh[a]) * invr2;
      }
    for (int a = 0; a < 3; ++a) {
      thes[2][a] -= nh[a] * invr2;
      thes[a][2] -= nh[a] * invr2;
      thes[a][a] -= tt * invr2;
      whes[a][a] -= wc * invr2;
    }
  }
  double Qt[200];
  Qt[0] = 1.0;
  if (L >= 1) { Qt[1] = tt; Qt[2] = 1.0; }
  for (int ll = 2; ll <= L; ++ll) {
    int kb = ll * (ll + 1) / 2; /* 0-based k of (ll, 0) */
    for (int mm = 0; mm <= ll - 2; ++mm) {
      int kq = kb + mm;
      Qt[kq] = ((double)(2 * ll - 1) * tt * Qt[kq - ll] - (double)(ll + mm - 1) * Qt[kq - 2 * ll + 1]) / (double)(ll - mm);
    }
    Qt[kb + ll - 1] = (double)(2 * ll - 1) * tt * Qt[kb + ll - 1 - ll];
    Qt[kb + ll] = (double)(2 * ll - 1) * Qt[kb + ll - ll - 1];
  }
  double complex wp[20]; /* wpow(m) at m + 2 */
  wp[0] = wp[1] = 0.0;
  wp[2] = 1.0;
  for (int m = 1; m <= L; ++m) wp[m + 2] = wp[m + 1] * wc;
  for (int k = 0; k < k_max; ++k) {
    int ll = g->lofk[k], mm = g->mofk[k];
    double q0 = Qt[k], q1 = mm + 1 <= ll ? Qt[k + 1] : 0.0, q2 = mm + 2 <= ll ? Qt[k + 2] : 0.0;
    double C = g->Ck[k], dm = (double)mm;
    g->Om[k] = C * q0 * wp[mm + 2];
    for (int a = 0; a < 3; ++a)
      g->dOm[3 * k + a] = C * (q1 * tvec[a] * wp[mm + 2] + q0 * dm * wp[mm + 1] * wvec[a]);
    if (want2)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          g->d2Om[9 * k + 3 * a + b] =
              C * (q2 * tvec[a] * tvec[b] * wp[mm + 2] + q1 * thes[a][b] * wp[mm + 2] +
                   q1 * dm * wp[mm + 1] * (tvec[a] * wvec[b] + tvec[b] * wvec[a]) +
                   q0 * (double)(mm * (mm - 1)) * wp[mm] * wvec[a] * wvec[b] + q0 * dm * wp[mm + 1] * whes[a][b]);
  }
  return 1;
}

int so_central_hessian(const so_params* p, const so_inputs* in, const double* W, const double* S,
                       const so_compress* c, int n_vec, const double* vecs, double* soap, double* hess,
                       double* rad_der2_out, double* selftest_err) {
  if (p->legacy_seed) return 1; /* the second radial derivatives need the unified form */
  so_state st;
  build_state(&st, p, in, W, S, c, 1);
  const int n_max = st.n_max, k_max = st.k_max, L = p->l_max, n_soap = st.n_soap, n_unc = st.n_soap_unc;
  const size_t NP = in->n_pairs;
  double* rad2 = (double*)malloc(sizeof(double) * n_max * NP);
  {
    double* r0 = (double*)malloc(sizeof(double) * n_max * NP);
    double* r1 = (double*)malloc(sizeof(double) * n_max * NP);
    so_radial2(p, in, W, 1, r0, r1, rad2);
    free(r0); free(r1);
  }
  if (rad_der2_out) memcpy(rad_der2_out, rad2, sizeof(double) * n_max * NP);
  double* sdp = (double*)malloc(sizeof(double) * in->n_sites);
  memset(soap, 0, sizeof(double) * (size_t)in->n_sites * n_soap);
  power_spectrum(&st, p, c, in->n_sites, soap, sdp); /* unnormalised here */
  const double pi = acos(-1.0), fourpi = 4.0 * pi, sqrt2 = sqrt(2.0);
  double rcut_max = 0.0;
  for (int s = 0; s < p->n_species; ++s) rcut_max = p->rcut_hard[s] > rcut_max ? p->rcut_hard[s] : rcut_max;
  const int nv = n_vec + 1;
  int* lofk = (int*)malloc(sizeof(int) * k_max);
  int* mofk = (int*)malloc(sizeof(int) * k_max);
  double *wm = (double*)malloc(sizeof(double) * k_max), *Ck = (double*)malloc(sizeof(double) * k_max);
  double* preflm = (double*)malloc(sizeof(double) * k_max);
  so_preflm(L, preflm);
  for (int l = 0, k = 0; l <= L; ++l)
    for (int m = 0; m <= l; ++m, ++k) {
      lofk[k] = l; mofk[k] = m;
      wm[k] = m > 0 ? 2.0 : 1.0;
      Ck[k] = m % 2 == 0 ? preflm[k] : -preflm[k];
    }
  memset(hess, 0, sizeof(double) * 9 * (size_t)n_vec * NP);
  double chk_err = 0.0, chk_ref = 0.0;
#pragma omp parallel reduction(max : chk_err, chk_ref)
  {
    hess_geom g;
    g.k_max = k_max; g.l_max = L; g.lofk = lofk; g.mofk = mofk; g.Ck = Ck;
    g.Om = (double complex*)malloc(sizeof(double complex) * k_max);
    g.dOm = (double complex*)malloc(sizeof(double complex) * 3 * k_max);
    g.d2Om = (double complex*)malloc(sizeof(double complex) * 9 * k_max);
    double complex* Gacc = (double complex*)malloc(sizeof(double complex) * 3 * n_max * k_max); /* [a][n][k] */
    double complex* Zv = (double complex*)malloc(sizeof(double complex) * (size_t)nv * n_max * k_max);
    double complex* ZGv = (double complex*)malloc(sizeof(double complex) * (size_t)nv * 3 * n_max * k_max);
    double* Smat = (double*)malloc(sizeof(double) * n_max * n_max * (L + 1));
    double* vunc = (double*)malloc(sizeof(double) * n_unc);
    double* pvec = (double*)malloc(sizeof(double) * n_soap);
    double* dp1 = (double*)malloc(sizeof(double) * 3 * n_soap);
    double* dp2 = (double*)malloc(sizeof(double) * 3 * n_soap);
    double *tr = (double*)malloc(sizeof(double) * n_unc), *ta = (double*)malloc(sizeof(double) * n_unc),
           *to = (double*)malloc(sizeof(double) * n_unc);
    double* Pw = (double*)malloc(sizeof(double) * 4 * n_max * (L + 1));
    double* hv = (double*)malloc(sizeof(double) * 9 * nv);
    double* u2rec = (double*)malloc(sizeof(double) * 3 * nv);
    double* hsum = (double*)malloc(sizeof(double) * 9 * (n_vec > 0 ? n_vec : 1));
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < in->n_sites; ++i) {
      const double s = sdp[i], invs = 1.0 / s, invs3 = invs / (s * s), invs5 = invs3 / (s * s);
      const double* cn = st.cnk + 2 * (size_t)i * n_max * k_max;
      for (int d = 0; d < n_soap; ++d) pvec[d] = soap[(size_t)i * n_soap + d];
      const int k3 = st.k_first[i];
      /* pass A: G and the unnormalised central-atom derivative */
      memset(Gacc, 0, sizeof(double complex) * 3 * n_max * k_max);
      memset(dp1, 0, sizeof(double) * 3 * n_soap);
      for (int j = 1; j < in->n_neigh[i]; ++j) {
        int k2 = k3 + j;
        if (!pair_geometry(&g, p, in, k2, rcut_max, 0)) continue;
        pair_unnorm_der(&st, p, c, in, cn, k2, tr, ta, to, Pw, dp2);
        for (int d = 0; d < 3 * n_soap; ++d) dp1[d] -= dp2[d];
        for (int k = 0; k < k_max; ++k) {
          double f0 = g.Fl[lofk[k]], f1 = g.Fl1[lofk[k]];
          for (int n = 0; n < n_max; ++n) {
            double r0 = fourpi * st.rad[(size_t)k2 * n_max + n], r1 = fourpi * st.rad_der[(size_t)k2 * n_max + n];
            double grad0 = r0 * f0, grad1 = r1 * f0 + r0 * f1;
            for (int a = 0; a < 3; ++a)
              Gacc[(a * n_max + n) * k_max + k] += grad1 * g.nh[a] * g.Om[k] + grad0 * g.dOm[3 * k + a];
          }
        }
      }
      /* the weight vectors, uncompressed (P^T, skipped components cleared), their S matrices */
      for (int v = 0; v < nv; ++v) {
        const double* vin = v < n_vec ? vecs + ((size_t)i * n_vec + v) * n_soap : pvec;
        memset(vunc, 0, sizeof(double) * n_unc);
        if (c && c->compress) {
          for (int e = 0; e < c->P_nonzero; ++e) vunc[c->P_j[e] - 1] += c->P_el[e] * vin[c->P_i[e] - 1];
        } else {
          for (int d = 0; d < n_soap; ++d) vunc[d] = vin[d];
        }
        for (int u = 0; u < n_unc; ++u)
          if (st.skip[u]) vunc[u] = 0.0;
        memset(Smat, 0, sizeof(double) * n_max * n_max * (L + 1));
        int cnt = 0;
        for (int n = 0; n < n_max; ++n)
          for (int np = n; np < n_max; ++np)
            for (int l = 0; l <= L; ++l, ++cnt) {
              if (n == np) {
                Smat[(l * n_max + n) * n_max + n] = 2.0 * vunc[cnt];
              } else {
                Smat[(l * n_max + n) * n_max + np] = sqrt2 * vunc[cnt];
                Smat[(l * n_max + np) * n_max + n] = sqrt2 * vunc[cnt];
              }
            }
        for (int k = 0; k < k_max; ++k) {
          int l = lofk[k];
          for (int n = 0; n < n_max; ++n) {
            double complex zc = 0.0, b0[3] = {0.0, 0.0, 0.0};
            for (int np = 0; np < n_max; ++np) {
              double sv = Smat[(l * n_max + n) * n_max + np];
              zc += sv * (cn[2 * (np * k_max + k)] + I * cn[2 * (np * k_max + k) + 1]);
              for (int a = 0; a < 3; ++a) b0[a] += sv * Gacc[(a * n_max + np) * k_max + k];
            }
            Zv[((size_t)v * n_max + n) * k_max + k] = zc;
            for (int a = 0; a < 3; ++a) ZGv[(((size_t)v * 3 + a) * n_max + n) * k_max + k] = b0[a];
          }
        }
      }
      /* pass B: the per-pair contraction */
      memset(hsum, 0, sizeof(double) * 9 * (n_vec > 0 ? n_vec : 1));
      for (int j = 1; j < in->n_neigh[i]; ++j) {
        int k2 = k3 + j;
        if (!pair_geometry(&g, p, in, k2, rcut_max, 1)) continue;
        pair_unnorm_der(&st, p, c, in, cn, k2, tr, ta, to, Pw, dp2);
        memset(hv, 0, sizeof(double) * 9 * nv);
        memset(u2rec, 0, sizeof(double) * 3 * nv);
        const double* r = st.rad + (size_t)k2 * n_max;
        const double* rd = st.rad_der + (size_t)k2 * n_max;
        const double* rh = rad2 + (size_t)k2 * n_max;
        const double invr = g.invr, *nh = g.nh;
        for (int k = 0; k < k_max; ++k) {
          int l = lofk[k];
          double f0 = g.Fl[l], f1 = g.Fl1[l], f2 = g.Fl2[l], wmk = wm[k];
          double complex om = g.Om[k];
          const double complex* dom = g.dOm + 3 * k;
          for (int v = 0; v < nv; ++v) {
            double complex a0 = 0, a1 = 0, a2 = 0, b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
            for (int n = 0; n < n_max; ++n) {
              double complex zc = conj(Zv[((size_t)v * n_max + n) * k_max + k]);
              a0 += r[n] * zc;
              a1 += rd[n] * zc;
              a2 += rh[n] * zc;
              for (int a = 0; a < 3; ++a) {
                double complex gc = conj(ZGv[(((size_t)v * 3 + a) * n_max + n) * k_max + k]);
                b0[a] += r[n] * gc;
                b1[a] += rd[n] * gc;
              }
            }
            double complex z0 = fourpi * f0 * a0;
            double complex z1 = fourpi * (f0 * a1 + f1 * a0);
            double complex z2 = fourpi * (f0 * a2 + 2.0 * f1 * a1 + f2 * a0);
            double complex gg0[3], gg1[3];
            for (int a = 0; a < 3; ++a) {
              gg0[a] = fourpi * f0 * b0[a];
              gg1[a] = fourpi * (f0 * b1[a] + f1 * b0[a]);
            }
            for (int b = 0; b < 3; ++b) u2rec[3 * v + b] += wmk * creal(z1 * nh[b] * om + z0 * dom[b]);
            for (int b = 0; b < 3; ++b)
              for (int a = 0; a < 3; ++a) {
                double complex tE = (z2 - z1 * invr) * nh[a] * nh[b] * om + z1 * (nh[a] * dom[b] + nh[b] * dom[a]) +
                                    z0 * g.d2Om[9 * k + 3 * a + b];
                if (a == b) tE += z1 * invr * om;
                double complex tD = gg1[a] * nh[b] * om + gg0[a] * dom[b];
                hv[9 * v + 3 * a + b] -= wmk * (creal(tE) + creal(tD));
              }
          }
        }
        /* the self-test: the first derivative from the same z0, z1 against pair_dp */
        for (int v = 0; v < nv; ++v)
          for (int b = 0; b < 3; ++b) {
            const double* w_ = v < n_vec ? vecs + ((size_t)i * n_vec + v) * n_soap : pvec;
            double refd = 0.0;
            for (int d = 0; d < n_soap; ++d) refd += w_[d] * dp2[3 * d + b];
            double e = fabs(u2rec[3 * v + b] - refd);
            chk_err = e > chk_err ? e : chk_err;
            chk_ref = fabs(refd) > chk_ref ? fabs(refd) : chk_ref;
          }
        /* normalisation, q = p / s */
        double t1[3], t2[3], c12[3][3];
        for (int a = 0; a < 3; ++a) {
          t1[a] = t2[a] = 0.0;
          for (int d = 0; d < n_soap; ++d) { t1[a] += pvec[d] * dp1[3 * d + a]; t2[a] += pvec[d] * dp2[3 * d + a]; }
        }
        for (int b = 0; b < 3; ++b)
          for (int a = 0; a < 3; ++a) {
            c12[a][b] = 0.0;
            for (int d = 0; d < n_soap; ++d) c12[a][b] += dp1[3 * d + a] * dp2[3 * d + b];
          }
        for (int v = 0; v < n_vec; ++v) {
          const double* w_ = vecs + ((size_t)i * n_vec + v) * n_soap;
          double wp = 0.0, u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
          for (int d = 0; d < n_soap; ++d) wp += w_[d] * pvec[d];
          for (int a = 0; a < 3; ++a)
            for (int d = 0; d < n_soap; ++d) { u1[a] += w_[d] * dp1[3 * d + a]; u2[a] += w_[d] * dp2[3 * d + a]; }
          double* h = hess + ((size_t)k2 * n_vec + v) * 9;
          for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
              h[3 * a + b] = hv[9 * v + 3 * a + b] * invs - (u1[a] * t2[b] + u2[b] * t1[a]) * invs3 -
                             wp * (c12[a][b] + hv[9 * n_vec + 3 * a + b]) * invs3 + 3.0 * wp * t1[a] * t2[b] * invs5;
              hsum[9 * v + 3 * a + b] += h[3 * a + b];
            }
        }
      }
      for (int v = 0; v < n_vec; ++v)
        for (int e = 0; e < 9; ++e) hess[((size_t)k3 * n_vec + v) * 9 + e] = -hsum[9 * v + e];
    }
    free(g.Om); free(g.dOm); free(g.d2Om); free(Gacc); free(Zv); free(ZGv); free(Smat); free(vunc); free(pvec);
    free(dp1); free(dp2); free(tr); free(ta); free(to); free(Pw); free(hv); free(u2rec); free(hsum);
  }
  for (int i = 0; i < in->n_sites; ++i)
    for (int d = 0; d < n_soap; ++d) soap[(size_t)i * n_soap + d] /= sdp[i];
  if (selftest_err) *selftest_err = chk_err / (chk_ref > 1e-300 ? chk_ref : 1e-300);
  free(lofk); free(mofk); free(wm); free(Ck); free(preflm); free(rad2); free(sdp);
  state_free(&st);
  return 0;
}
