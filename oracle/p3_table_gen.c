/* p3_table_gen.c -- CPU restatement of the P3 ice lookup-table generator.
 *
 * TEST INFRASTRUCTURE ONLY: this file checks the device generator
 * (scream_amd/csrc/p3_table_gen.hip); nothing in the product calls it.
 *
 * Follows components/eam/tools/create_p3_lookupTable/create_p3_lookupTable_1.f90-v4.1
 * ("GEN" below, compiled by the reference in double precision): the same
 * constants, loops, branch structure and order of every sum and product, with
 * Fortran `x**r` (real r) as pow, `x**0.5` as sqrt, integer constant powers as
 * repeated products and `x**i` (integer variable i) as binary exponentiation.
 * Elementary functions: glibc, or (P3O_PORTABLE_MATH) the portable functions
 * the device code uses, which round identically on the CPU and the GPU.
 *
 * Output (per cell; cells are (i_rhor, i_Fr, i_Qnorm), 5 x 4 x 50):
 *   ice[cell][15]        qsave nsave uns ums nagg nrwat vdep eff nlarge nsmall
 *                        lsave refl vdep1 dmm rhomm          (GEN:1377-1404)
 *   coll[cell][30][6]    qsave nsave lamrs nrrain qrrain qsrain (GEN:1406-1424)
 * with GEN's clamp of values below 1e-99 to 0 applied.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "p3_oracle.h"

#ifdef P3O_PORTABLE_MATH
#include "../scream_amd/csrc/p3_math.h"
#define POW p3m_pow
#define EXP p3m_exp
#define LOG p3m_log
#else
#define POW pow
#define EXP exp
#define LOG log
#endif

#define N_RHOR 5
#define N_FR 4
#define N_QNORM 50
#define N_DRSCALE 30

/* GEN:56-115 constants (GEN writes pi = 3.14159) */
typedef struct {
  double pi, g, p, t, rho, mu, del0, c0, c1, c2, ds, cs, dg, bas, aas, bag, aag, dcrit, thrd, sxth;
} Consts;

static void consts_init(Consts* k) {
  k->thrd = 1. / 3.;
  k->sxth = 1. / 6.;
  k->pi = 3.14159;
  k->g = 9.861;
  k->p = 60000.;
  k->t = 253.15;
  k->rho = k->p / (287.15 * k->t);
  k->mu = 1.496E-6 * POW(k->t, 1.5) / (k->t + 120.) / k->rho;
  k->del0 = 5.83;
  k->c0 = 0.6;
  k->c1 = 4. / (k->del0 * k->del0 * sqrt(k->c0));
  k->c2 = k->del0 * k->del0 / 4.;
  k->ds = 1.9;
  k->cs = 0.0121;
  k->dg = 3.;
  k->bas = 1.88;
  k->aas = 0.2285 * POW(100., k->bas) / (100. * 100.);
  k->bag = 2.;
  k->aag = k->pi * 0.25;
  k->dcrit = POW(k->pi / (6. * k->cs) * 900., 1. / (k->ds - 3.));
}

/* Fortran real**integer: binary exponentiation from a product of 1 */
static double powi(double b, int e) {
  double r = 1.;
  for (;;) {
    if (e & 1) r *= b;
    e >>= 1;
    if (e == 0) break;
    b *= b;
  }
  return r;
}

/* GEN:1560-1812, the gamma function of W. J. Cody's SPECFUN (rational
 * approximation on (1,2), asymptotic series from 12), for the positive
 * arguments GEN uses. */
static double cody_gamma(double x) {
  static const double P[8] = {-1.71618513886549492533811E+0, 2.47656508055759199108314E+1,
                              -3.79804256470945635097577E+2, 6.29331155312818442661052E+2,
                              8.66966202790413211295064E+2, -3.14512729688483675254357E+4,
                              -3.61444134186911729807069E+4, 6.64561438202405440627855E+4};
  static const double Q[8] = {-3.08402300119738975254353E+1, 3.15350626979604161529144E+2,
                              -1.01515636749021914166146E+3, -3.10777167157231109440444E+3,
                              2.25381184209801510330112E+4, 4.75584627752788110767815E+3,
                              -1.34659959864969306392456E+5, -1.15132259675553483497211E+5};
  static const double C[7] = {-1.910444077728E-03, 8.4171387781295E-04, -5.952379913043012E-04,
                              7.93650793500350248E-04, -2.777777777777681622553E-03,
                              8.333333333333333331554247E-02, 5.7083835261E-03};
  const double sqrtpi = 0.9189385332046727417803297E0, xbig = 35.040E0, xminin = 1.18E-38,
               eps = 1.19E-7, xinf = 3.4E38;
  double y = x, res;
  int n = 0;
  if (y <= 0.) return xinf; /* not reached by GEN's arguments */
  if (y < eps) {
    res = y >= xminin ? 1. / y : xinf;
  } else if (y < 12.) {
    const double y1 = y;
    double z;
    if (y < 1.) {
      z = y;
      y = y + 1.;
    } else {
      n = (int)y - 1;
      y = y - (double)n;
      z = y - 1.;
    }
    double xnum = 0., xden = 1.;
    for (int i = 0; i < 8; ++i) {
      xnum = (xnum + P[i]) * z;
      xden = xden * z + Q[i];
    }
    res = xnum / xden + 1.;
    if (y1 < y) {
      res = res / y1;
    } else if (y1 > y) {
      for (int i = 0; i < n; ++i) {
        res = res * y;
        y = y + 1.;
      }
    }
  } else if (y <= xbig) {
    const double ysq = y * y;
    double sum = C[6];
    for (int i = 0; i < 6; ++i) sum = sum / ysq + C[i];
    sum = sum / y - y + sqrtpi;
    sum = sum + (y - 0.5) * LOG(y);
    res = EXP(sum);
  } else {
    res = xinf;
  }
  return res;
}

/* GEN:1438-1558, Q(a, x) = 1 - P(a, x) by the series (gser) or the continued
 * fraction (gcf) of Numerical Recipes, eps = 3e-7, at most 100 terms. */
static double gammq(double a, double x) {
  const double eps = 3.e-7, fpmin = 1.e-30;
  if (x < a + 1.) {
    const double gln = LOG(cody_gamma(a));
    if (x <= 0.) return 1.;
    double ap = a, sum = 1. / a, del = sum;
    for (int n = 1; n <= 100; ++n) {
      ap = ap + 1.;
      del = del * x / ap;
      sum = sum + del;
      if (fabs(del) < fabs(sum) * eps) break;
    }
    const double gamser = sum * EXP(-x + a * LOG(x) - gln);
    return 1. - gamser;
  }
  const double gln = LOG(cody_gamma(a));
  double b = x + 1. - a, c = 1. / fpmin, d = 1. / b, h = d;
  for (int i = 1; i <= 100; ++i) {
    const double an = (double)(-i) * ((double)i - a);
    b = b + 2.;
    d = an * d + b;
    if (fabs(d) < fpmin) d = fpmin;
    c = b + an / c;
    if (fabs(c) < fpmin) c = fpmin;
    d = 1. / d;
    const double del = d * c;
    h = h * del;
    if (fabs(del - 1.) < eps) break;
  }
  return EXP(-x + a * LOG(x) - gln) * h;
}

/* The mass- and area-size regime of one (rime density, rime fraction) pair. */
typedef struct {
  int i_Fr;
  double Fr, cgp, dcrits, dcritr, csr, dsr;
} Regime;

/* GEN:250-302: critical sizes and the partially-rimed m-D coefficients */
static void regime_init(const Consts* k, int i_rhor, int i_Fr, Regime* r) {
  static const double rr[N_RHOR] = {50., 250., 450., 650., 900.};
  const double crp = rr[i_rhor] * k->pi * k->sxth;
  static const double frs[N_FR] = {0., 0.333, 0.667, 1.};
  r->i_Fr = i_Fr;
  r->Fr = frs[i_Fr];
  r->cgp = crp;
  if (i_Fr == 0) {
    r->dcrits = 1.e+6;
    r->dcritr = r->dcrits;
    r->csr = k->cs;
    r->dsr = k->ds;
  } else if (i_Fr == 1 || i_Fr == 2) {
    const double Fr = r->Fr;
    for (;;) {
      r->dcrits = POW(k->cs / r->cgp, 1. / (k->dg - k->ds));
      r->dcritr = POW((1. + Fr / (1. - Fr)) * k->cs / r->cgp, 1. / (k->dg - k->ds));
      r->csr = k->cs * (1. + Fr / (1. - Fr));
      r->dsr = k->ds;
      const double rhodep = 1. / (r->dcritr - r->dcrits) * 6. * k->cs / (k->pi * (k->ds - 2.)) *
                            (POW(r->dcritr, k->ds - 2.) - POW(r->dcrits, k->ds - 2.));
      const double cgpold = r->cgp;
      r->cgp = crp * Fr + rhodep * (1. - Fr) * k->pi * k->sxth;
      if (fabs((r->cgp - cgpold) / r->cgp) < 0.01) break;
    }
  } else {
    r->dcrits = POW(k->cs / r->cgp, 1. / (k->dg - k->ds));
    r->dcritr = 1.e+6;
    r->csr = r->cgp;
    r->dsr = k->dg;
  }
}

/* GEN:333-369 (and its copies): m = cs1 D^ds1 and A = aas1 D^bas1 at size d.
 * small_aas is pi/4 as GEN spells it in each loop (the same double). */
static void mass_area(const Consts* k, const Regime* r, double d, double* cs1, double* ds1, double* bas1,
                      double* aas1) {
  if (d <= k->dcrit) {
    *cs1 = k->pi * k->sxth * 900.;
    *ds1 = 3.;
    *bas1 = 2.;
    *aas1 = k->pi / 4.;
  } else if (d <= r->dcrits) {
    *cs1 = k->cs;
    *ds1 = k->ds;
    *bas1 = k->bas;
    *aas1 = k->aas;
  } else if (d <= r->dcritr) {
    *cs1 = r->cgp;
    *ds1 = k->dg;
    *bas1 = k->bag;
    *aas1 = k->aag;
  } else {
    *cs1 = r->csr;
    *ds1 = r->dsr;
    if (r->i_Fr == 0) {
      *aas1 = k->aas;
      *bas1 = k->bas;
    } else {
      *bas1 = k->bas;
      const double dum1 = k->aas * POW(d, k->bas);
      const double dum2 = k->aag * POW(d, k->bag);
      const double m1 = *cs1 * POW(d, *ds1);
      const double m2 = k->cs * POW(d, k->ds);
      const double m3 = r->cgp * POW(d, k->dg);
      const double dum3 = dum1 + (m1 - m2) * (dum2 - dum1) / (m3 - m2);
      *aas1 = dum3 / POW(d, k->bas);
    }
  }
}

/* GEN:377-394: Mitchell and Heymsfield (2005) fall speed at size d1 */
static double ice_fallspeed(const Consts* k, double cs1, double ds1, double bas1, double aas1, double d1) {
  const double mr = k->mu * k->rho;
  const double xx = 2. * cs1 * k->g * k->rho * POW(d1, ds1 + 2. - bas1) / (aas1 * (mr * mr));
  const double sx = sqrt(xx);
  const double s1 = sqrt(1. + k->c1 * sx);
  /* a0 = b0 = 0: the turbulence terms are exact zeros */
  const double b1 = k->c1 * sx / (2. * (s1 - 1.) * s1) - 0. * 0. * POW(xx, 0.) / (k->c2 * ((s1 - 1.) * (s1 - 1.)));
  const double a1 = (k->c2 * ((s1 - 1.) * (s1 - 1.)) - 0. * POW(xx, 0.)) / POW(xx, b1);
  return a1 * POW(k->mu, 1. - 2. * b1) * POW(2. * cs1 * k->g / (k->rho * aas1), b1) *
         POW(d1, b1 * (ds1 - bas1 + 2.) - 1.);
}

/* GEN:480-495: rain fall speed (m/s) at size d1 */
static double rain_fallspeed(const Consts* k, double d1) {
  const double dia = d1;
  double amg = k->pi * k->sxth * 997. * (dia * dia * dia);
  amg = amg * 1000.;
  double dum2;
  if (dia <= 134.43e-6) dum2 = 4.5795e5 * POW(amg, 2. * k->thrd);
  else if (dia < 1511.64e-6) dum2 = 4.962e3 * POW(amg, k->thrd);
  else if (dia < 3477.84e-6) dum2 = 1.732e3 * POW(amg, k->sxth);
  else dum2 = 917.;
  return dum2 * 1.e-2;
}

/* The four partial moments of GEN:565-580 (q for a given lam and mu_i); clamp
 * applies GEN's max(.,0) of the search loop (absent after it). */
static double qsum(const Consts* k, const Regime* r, double lam, double mu_i, int clamp) {
  const double cs1 = k->pi * k->sxth * 900., ds1 = 3.;
  const double dum1 =
      POW(lam, -ds1 - mu_i - 1.) * cody_gamma(mu_i + ds1 + 1.) * (1. - gammq(mu_i + ds1 + 1., k->dcrit * lam));
  double dum2 = POW(lam, -k->ds - mu_i - 1.) * cody_gamma(mu_i + k->ds + 1.) * (gammq(mu_i + k->ds + 1., k->dcrit * lam));
  double dum = POW(lam, -k->ds - mu_i - 1.) * cody_gamma(mu_i + k->ds + 1.) * (gammq(mu_i + k->ds + 1., r->dcrits * lam));
  dum2 = dum2 - dum;
  if (clamp) dum2 = fmax(dum2, 0.);
  double dum3 = POW(lam, -k->dg - mu_i - 1.) * cody_gamma(mu_i + k->dg + 1.) * (gammq(mu_i + k->dg + 1., r->dcrits * lam));
  dum = POW(lam, -k->dg - mu_i - 1.) * cody_gamma(mu_i + k->dg + 1.) * (gammq(mu_i + k->dg + 1., r->dcritr * lam));
  dum3 = dum3 - dum;
  if (clamp) dum3 = fmax(dum3, 0.);
  const double dum4 =
      POW(lam, -r->dsr - mu_i - 1.) * cody_gamma(mu_i + r->dsr + 1.) * (gammq(mu_i + r->dsr + 1., r->dcritr * lam));
  return cs1 * dum1 + k->cs * dum2 + r->cgp * dum3 + r->csr * dum4;
}

static double mu_of_lam(double lam) {
  double mu_i = 0.076 * POW(lam / 100., 0.8) - 2.;
  mu_i = fmax(mu_i, 0.);
  mu_i = fmin(mu_i, 6.);
  return mu_i;
}

static double clamp99(double v) { return v < 1.e-99 ? 0. : v; }

/* One cell: GEN:502-1367 for (i_rhor, i_Fr, i_Qnorm). */
static void cell(const Consts* k, const Regime* r, const double* fall1, const double* fall2,
                 const double* fallr, int i_Qnorm, double* ice, double* coll) {
  const double Dm_max = 2000.e-6, Dm_min = 2.e-6;
  const double pi = k->pi, sxth = k->sxth;
  const double q = POW(261.7, (double)(i_Qnorm + 1 + 10) * 0.1) * 1.e-18;

  /* GEN:540-590: lambda with the smallest |q - q(lambda)| over 9000 candidates */
  double qerror = 1.e+20, lamf = 0.;
  for (int ii = 1; ii <= 9000; ++ii) {
    double lam = powi(1.0013, ii) * 100.;
    const double mu_i = mu_of_lam(lam);
    lam = fmax(lam, (mu_i + 1.) / Dm_max);
    lam = fmin(lam, (mu_i + 1.) / Dm_min);
    const double n0 = POW(lam, mu_i + 1.) / (cody_gamma(mu_i + 1.));
    const double qdum = n0 * qsum(k, r, lam, mu_i, 1);
    if (ii == 1) {
      qerror = fabs(q - qdum);
      lamf = lam;
    }
    if (fabs(q - qdum) < qerror) {
      lamf = lam;
      qerror = fabs(q - qdum);
    }
  }
  const double lam = lamf;
  const double mu_i = mu_of_lam(lam);
  const double n0 = q / qsum(k, r, lam, mu_i, 0);

  /* GEN:655-700: lambda limiters */
  double duml = (mu_i + 1.) / Dm_min;
  double n0dum = q / qsum(k, r, duml, mu_i, 0);
  const double nlarge = n0dum / (POW(duml, mu_i + 1.) / (cody_gamma(mu_i + 1.)));
  duml = (mu_i + 1.) / (Dm_max);
  n0dum = q / qsum(k, r, duml, mu_i, 0);
  const double nsmall = n0dum / (POW(duml, mu_i + 1.) / (cody_gamma(mu_i + 1.)));

  /* GEN:713-804: fall speeds, reflectivity, mean size and density */
  double sum1 = 0., sum2 = 0., sum3 = 0., sum4 = 0., sum5 = 0., sum6 = 0., sum7 = 0.;
  double dd = 2.e-6;
  for (int ii = 1; ii <= 10000; ++ii) {
    const double dum = (double)ii * dd - 1.e-6;
    double cs1, ds1, bas1, aas1;
    mass_area(k, r, dum, &cs1, &ds1, &bas1, &aas1);
    const double e = EXP(-lam * dum);
    const double pm = POW(dum, mu_i);
    const double f1 = fall1[ii - 1];
    sum1 = sum1 + f1 * pm * e * dd;
    sum2 = sum2 + f1 * cs1 * POW(dum, ds1 + mu_i) * e * dd;
    sum3 = sum3 + pm * e * dd;
    sum4 = sum4 + cs1 * POW(dum, ds1 + mu_i) * e * dd;
    const double a = cs1 / 917., b = 6. / pi;
    sum5 = sum5 + n0 * (a * a) * (b * b) * POW(dum, 2. * ds1 + mu_i) * e * dd;
    sum6 = sum6 + cs1 * POW(dum, ds1 + mu_i + 1.) * e * dd;
    const double m = cs1 * POW(dum, ds1);
    sum7 = sum7 + (m * m) / (pi * sxth * (dum * dum * dum)) * pm * e * dd;
  }
  const double uns = sum1 / sum3, ums = sum2 / sum4, refl = sum5, dmm = sum6 / sum4, rhomm = sum7 / sum4;

  /* GEN:810-928: self-aggregation over pairs of 1000 bins */
  double num[1000];
  dd = 20.e-6;
  for (int jj = 1000; jj >= 1; --jj) {
    const double d1 = (double)jj * dd - 10.e-6;
    num[jj - 1] = n0 * POW(d1, mu_i) * EXP(-lam * d1) * dd;
  }
  /* the per-bin factors of the pair term (the same values GEN forms in the loop) */
  double A1[1000], M1[1000], M2[1000], M3[1000], DA1[1000], DA2[1000], PB[1000], BAS2[1000];
  int big[1000];
  for (int jj = 1; jj <= 1000; ++jj) {
    const double d = (double)jj * 20. * 1.e-6 - 10.e-6;
    double cs1, ds1, bas1, aas1;
    mass_area(k, r, d, &cs1, &ds1, &bas1, &aas1);
    A1[jj - 1] = aas1 * POW(d, bas1);
    /* particle 2's interpolation at size d (GEN:870-885) uses particle 1's masses */
    M1[jj - 1] = r->csr * POW(d, r->dsr);
    M2[jj - 1] = k->cs * POW(d, k->ds);
    M3[jj - 1] = r->cgp * POW(d, k->dg);
    DA1[jj - 1] = k->aas * POW(d, k->bas);
    DA2[jj - 1] = k->aag * POW(d, k->bag);
    PB[jj - 1] = POW(d, k->bas);
    big[jj - 1] = d > r->dcritr;
    /* area term of particle 2 when it is not interpolated */
    if (d <= k->dcrit) BAS2[jj - 1] = (k->pi / 4.) * POW(d, 2.);
    else if (d <= r->dcrits) BAS2[jj - 1] = k->aas * POW(d, k->bas);
    else BAS2[jj - 1] = k->aag * POW(d, k->bag);
  }
  sum1 = 0.;
  for (int jj = 1000; jj >= 1; --jj) {
    for (int kk = 1; kk <= jj - 1; ++kk) {
      double a2;
      if (big[kk - 1] && r->i_Fr != 0) {
        const double dum3 = DA1[kk - 1] + (M1[jj - 1] - M2[jj - 1]) * (DA2[kk - 1] - DA1[kk - 1]) /
                                              (M3[jj - 1] - M2[jj - 1]);
        a2 = dum3 / PB[kk - 1] * PB[kk - 1];
      } else {
        a2 = BAS2[kk - 1];
      }
      const double delu = fabs(fall2[jj - 1] - fall2[kk - 1]);
      sum1 = sum1 + (A1[jj - 1] + a2) * delu * num[jj - 1] * num[kk - 1];
    }
  }
  const double nagg = sum1;

  /* GEN:940-1004: collection of cloud droplets */
  sum1 = 0.;
  dd = 2.e-6;
  for (int jj = 1; jj <= 10000; ++jj) {
    const double d1 = (double)jj * 2. * 1.e-6 - 1.e-6;
    double cs1, ds1, bas1, aas1;
    mass_area(k, r, d1, &cs1, &ds1, &bas1, &aas1);
    if (d1 >= 100.e-6)
      sum1 = sum1 + aas1 * POW(d1, bas1) * fall1[jj - 1] * n0 * POW(d1, mu_i) * EXP(-lam * d1) * dd;
  }
  const double nrwat = sum1;

  /* GEN:1023-1206: ice-rain collection for 30 rain sizes */
  double Ai[1000], Pi_[1000], Ei[1000], CSi[1000], Di[1000], numi[1000];
  dd = 20.e-6;
  for (int jj = 1; jj <= 1000; ++jj) {
    const double d1 = (double)jj * 20. * 1.e-6 - 10.e-6;
    double cs1, ds1, bas1, aas1;
    mass_area(k, r, d1, &cs1, &ds1, &bas1, &aas1);
    Ai[jj - 1] = aas1 * POW(d1, bas1);
    Pi_[jj - 1] = POW(d1, mu_i);
    Ei[jj - 1] = EXP(-lam * d1);
    CSi[jj - 1] = cs1;
    Di[jj - 1] = POW(d1, ds1);
    numi[jj - 1] = n0 * POW(d1, mu_i) * EXP(-lam * d1) * dd;
  }
  for (int i_Drscale = 1; i_Drscale <= N_DRSCALE; ++i_Drscale) {
    double dum = powi(1.24, i_Drscale) * 10.e-6;
    const double lamv = 1. / dum;
    const double lamrs = lamv;
    const double mu_r = 0.;
    double lamr = lamv * POW(cody_gamma(mu_r + 4.) / (6. * cody_gamma(mu_r + 1.)), k->thrd);
    const double lammax = (mu_r + 1.) * 1.e+5, lammin = (mu_r + 1.) * 200.;
    lamr = fmin(lamr, lammax);
    lamr = fmax(lamr, lammin);
    double Bk[1000], Ck[1000], Rk[1000], Fk[1000], numr[1000];
    for (int kk = 1; kk <= 1000; ++kk) {
      const double d2 = (double)kk * 20. * 1.e-6 - 10.e-6;
      Bk[kk - 1] = pi * 0.25 * (d2 * d2);
      Ck[kk - 1] = d2 * d2 * d2;
      Rk[kk - 1] = POW(d2, mu_r);
      Fk[kk - 1] = EXP(-lamr * d2);
      numr[kk - 1] = POW(d2, mu_r) * EXP(-lamr * d2) * dd;
    }
    double s1 = 0., s2 = 0., s6 = 0.;
    for (int jj = 0; jj < 1000; ++jj) {
      for (int kk = 0; kk < 1000; ++kk) {
        const double delu = fabs(fall2[jj] - fallr[kk]);
        const double area = Ai[jj] + Bk[kk];
        const double t = area * delu * n0 * Pi_[jj] * Ei[jj] * dd * numr[kk];
        s1 = s1 + t;
        s2 = s2 + t * pi * sxth * 997. * Ck[kk];
        s6 = s6 + area * delu * Rk[kk] * Fk[kk] * dd * numi[jj] * CSi[jj] * Di[jj];
      }
    }
    double* c = coll + (size_t)(i_Drscale - 1) * 6;
    c[0] = clamp99(q);
    c[1] = 1.;
    c[2] = clamp99(lamrs);
    c[3] = clamp99(s1);
    c[4] = clamp99(s2);
    c[5] = clamp99(s6);
  }

  /* GEN:1214-1277: vapour deposition with ventilation */
  sum1 = 0.;
  sum2 = 0.;
  dd = 2.e-6;
  for (int jj = 1; jj <= 10000; ++jj) {
    const double d1 = (double)jj * 2. * 1.e-6 - 1.e-6;
    double cap;
    if (d1 <= k->dcrit) cap = 1.;
    else if (d1 <= r->dcrits) cap = 0.48;
    else if (d1 <= r->dcritr) cap = 1.;
    else if (r->i_Fr == 0) cap = 0.48;
    else {
      const double m1 = r->csr * POW(d1, r->dsr);
      const double m2 = k->cs * POW(d1, k->ds);
      const double m3 = r->cgp * POW(d1, k->dg);
      cap = 0.48 + (m1 - m2) * (1. - 0.48) / (m3 - m2);
    }
    const double dum = sqrt(fall1[jj - 1] * d1);
    if (d1 < 100.e-6) {
      sum1 = sum1 + cap * n0 * POW(d1, mu_i + 1.) * EXP(-lam * d1) * dd;
    } else {
      sum1 = sum1 + cap * n0 * 0.65 * POW(d1, mu_i + 1.) * EXP(-lam * d1) * dd;
      sum2 = sum2 + cap * n0 * 0.44 * dum * POW(d1, mu_i + 1.) * EXP(-lam * d1) * dd;
    }
  }
  const double vdep = sum1, vdep1 = sum2;

  /* GEN:1284-1349: effective radius */
  sum1 = 0.;
  sum2 = 0.;
  for (int jj = 1; jj <= 10000; ++jj) {
    const double d1 = (double)jj * 2. * 1.e-6 - 1.e-6;
    double cs1, ds1, bas1, aas1;
    mass_area(k, r, d1, &cs1, &ds1, &bas1, &aas1);
    sum1 = sum1 + cs1 * POW(d1, ds1) * POW(d1, mu_i) * EXP(-lam * d1) * dd;
    sum2 = sum2 + aas1 * POW(d1, bas1) * POW(d1, mu_i) * EXP(-lam * d1) * dd;
  }
  const double eff = 3. * sum1 / (4. * sum2 * 916.7);

  const double v[15] = {q, 1., uns, ums, nagg, nrwat, vdep, eff, nlarge, nsmall, lam, refl, vdep1, dmm, rhomm};
  for (int j = 0; j < 15; ++j) ice[j] = clamp99(v[j]);
}

void p3o_table_gen(const unsigned char* mask, double* ice, double* coll, int nthreads) {
  Consts k;
  consts_init(&k);
  for (int rf = 0; rf < N_RHOR * N_FR; ++rf) {
    const int i_rhor = rf / N_FR, i_Fr = rf % N_FR;
    int any = 0;
    for (int c = 0; c < N_QNORM; ++c) any |= !mask || mask[rf * N_QNORM + c];
    if (!any) continue;
    Regime r;
    regime_init(&k, i_rhor, i_Fr, &r);
    static double fall1[10000], fall2[1000], fallr[1000];
    for (int jj = 1; jj <= 10000; ++jj) {
      const double d1 = (double)jj * 2. * 1.e-6 - 1.e-6;
      double cs1, ds1, bas1, aas1;
      mass_area(&k, &r, d1, &cs1, &ds1, &bas1, &aas1);
      fall1[jj - 1] = ice_fallspeed(&k, cs1, ds1, bas1, aas1, d1);
    }
    for (int jj = 1; jj <= 1000; ++jj) {
      const double d1 = (double)jj * 20. * 1.e-6 - 10.e-6;
      double cs1, ds1, bas1, aas1;
      mass_area(&k, &r, d1, &cs1, &ds1, &bas1, &aas1);
      fall2[jj - 1] = ice_fallspeed(&k, cs1, ds1, bas1, aas1, d1);
      fallr[jj - 1] = rain_fallspeed(&k, d1);
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads != 1)
    for (int c = 0; c < N_QNORM; ++c) {
      const size_t id = (size_t)rf * N_QNORM + c;
      if (mask && !mask[id]) continue;
      cell(&k, &r, fall1, fall2, fallr, c, ice + id * 15, coll + id * N_DRSCALE * 6);
    }
  }
}
