// p3_table_gen.hip -- the P3 ice lookup-table generator on the GPU.
//
// Computes the table that components/eam/tools/create_p3_lookupTable/
// create_p3_lookupTable_1.f90-v4.1 ("GEN") writes: for 5 rime densities x 4
// rime fractions x 50 normalized ice masses ("cells"), the size distribution
// whose mass moment matches the cell (a search over 9000 slopes), its fall
// speeds, aggregation, riming, ice-rain collection (30 rain sizes),
// deposition and effective-radius integrals.  GEN runs about 3e10 inner
// iterations in sequence (16 minutes on one core); here every cell, every
// candidate slope and every (cell, rain size) pair is a thread of its own.
//
// Every value is computed with GEN's formula and operation order, and every
// sum is accumulated in GEN's order by one thread, so the result does not
// depend on the launch shape.  Elementary functions are the portable ones of
// p3_math.h (the same bits on the CPU and the GPU).
//
// Layout: per-cell quantities are stored bin-major ([bin][cell]), so that the
// lanes of a wavefront, which hold consecutive cells, read consecutive words.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/p3x.h"
#include "p3_math.h"

namespace {

constexpr int kRhor = 5, kFr = 4, kQnorm = 50, kDrscale = 30;
constexpr int kRF = kRhor * kFr;        // (rime density, rime fraction) pairs
constexpr int kCells = kRF * kQnorm;    // 1000
constexpr int kSearch = 9000;           // candidate slopes (GEN:540)
constexpr int kFine = 10000;            // 2-micron bins (GEN:324)
constexpr int kCoarse = 1000;           // 20-micron bins (GEN:401)
constexpr int kIceCols = 15, kCollCols = 6;

__device__ __forceinline__ double d_pow(double x, double y) { return p3m_pow(x, y); }
__device__ __forceinline__ double d_exp(double x) { return p3m_exp(x); }
__device__ __forceinline__ double d_log(double x) { return p3m_log(x); }

// GEN:56-115 (GEN writes pi = 3.14159)
struct Consts {
  double pi, g, rho, mu, c1, c2, ds, cs, dg, bas, aas, bag, aag, dcrit, thrd, sxth;
};

__device__ Consts make_consts() {
  Consts k;
  k.thrd = 1. / 3.;
  k.sxth = 1. / 6.;
  k.pi = 3.14159;
  k.g = 9.861;
  const double p = 60000., t = 253.15;
  k.rho = p / (287.15 * t);
  k.mu = 1.496E-6 * d_pow(t, 1.5) / (t + 120.) / k.rho;
  const double del0 = 5.83, c0 = 0.6;
  k.c1 = 4. / (del0 * del0 * sqrt(c0));
  k.c2 = del0 * del0 / 4.;
  k.ds = 1.9;
  k.cs = 0.0121;
  k.dg = 3.;
  k.bas = 1.88;
  k.aas = 0.2285 * d_pow(100., k.bas) / (100. * 100.);
  k.bag = 2.;
  k.aag = k.pi * 0.25;
  k.dcrit = d_pow(k.pi / (6. * k.cs) * 900., 1. / (k.ds - 3.));
  return k;
}

// Fortran real**integer: binary exponentiation starting from 1
__device__ __forceinline__ double powi(double b, int e) {
  double r = 1.;
  for (;;) {
    if (e & 1) r *= b;
    e >>= 1;
    if (e == 0) break;
    b *= b;
  }
  return r;
}

// W. J. Cody's SPECFUN gamma (GEN:1560-1812) for the positive arguments GEN uses
__device__ double cody_gamma(double x) {
  const double P[8] = {-1.71618513886549492533811E+0, 2.47656508055759199108314E+1,
                       -3.79804256470945635097577E+2, 6.29331155312818442661052E+2,
                       8.66966202790413211295064E+2, -3.14512729688483675254357E+4,
                       -3.61444134186911729807069E+4, 6.64561438202405440627855E+4};
  const double Q[8] = {-3.08402300119738975254353E+1, 3.15350626979604161529144E+2,
                       -1.01515636749021914166146E+3, -3.10777167157231109440444E+3,
                       2.25381184209801510330112E+4, 4.75584627752788110767815E+3,
                       -1.34659959864969306392456E+5, -1.15132259675553483497211E+5};
  const double C[7] = {-1.910444077728E-03, 8.4171387781295E-04, -5.952379913043012E-04,
                       7.93650793500350248E-04, -2.777777777777681622553E-03,
                       8.333333333333333331554247E-02, 5.7083835261E-03};
  const double sqrtpi = 0.9189385332046727417803297E0, xbig = 35.040E0, xminin = 1.18E-38,
               eps = 1.19E-7, xinf = 3.4E38;
  double y = x;
  if (y <= 0.) return xinf;
  if (y < eps) return y >= xminin ? 1. / y : xinf;
  if (y < 12.) {
    const double y1 = y;
    double z;
    int n = 0;
    if (y < 1.) {
      z = y;
      y = y + 1.;
    } else {
      n = (int)y - 1;
      y = y - (double)n;
      z = y - 1.;
    }
    double xnum = 0., xden = 1.;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      xnum = (xnum + P[i]) * z;
      xden = xden * z + Q[i];
    }
    double res = xnum / xden + 1.;
    if (y1 < y) {
      res = res / y1;
    } else if (y1 > y) {
      for (int i = 0; i < n; ++i) {
        res = res * y;
        y = y + 1.;
      }
    }
    return res;
  }
  if (y > xbig) return xinf;
  const double ysq = y * y;
  double sum = C[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) sum = sum / ysq + C[i];
  sum = sum / y - y + sqrtpi;
  sum = sum + (y - 0.5) * d_log(y);
  return d_exp(sum);
}

// Q(a, x) = 1 - P(a, x): Numerical Recipes' gser / gcf as GEN has them
// (GEN:1438-1558), eps 3e-7, at most 100 terms
__device__ double gammq(double a, double x) {
  const double eps = 3.e-7, fpmin = 1.e-30;
  const double gln = d_log(cody_gamma(a));
  if (x < a + 1.) {
    if (x <= 0.) return 1.;
    double ap = a, sum = 1. / a, del = sum;
    for (int n = 1; n <= 100; ++n) {
      ap = ap + 1.;
      del = del * x / ap;
      sum = sum + del;
      if (fabs(del) < fabs(sum) * eps) break;
    }
    return 1. - sum * d_exp(-x + a * d_log(x) - gln);
  }
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
  return d_exp(-x + a * d_log(x) - gln) * h;
}

// The m-D / A-D regime of one (rime density, rime fraction) pair (GEN:250-302)
struct Regime {
  int i_Fr;
  double cgp, dcrits, dcritr, csr, dsr;
};

__device__ Regime make_regime(const Consts& k, int rf) {
  const double rr[kRhor] = {50., 250., 450., 650., 900.};
  const double frs[kFr] = {0., 0.333, 0.667, 1.};
  const int i_rhor = rf / kFr, i_Fr = rf % kFr;
  const double crp = rr[i_rhor] * k.pi * k.sxth, Fr = frs[i_Fr];
  Regime r;
  r.i_Fr = i_Fr;
  r.cgp = crp;
  if (i_Fr == 0) {
    r.dcrits = 1.e+6;
    r.dcritr = r.dcrits;
    r.csr = k.cs;
    r.dsr = k.ds;
  } else if (i_Fr < kFr - 1) {
    // fixed point for the density of the partially rimed part; GEN iterates to
    // 1 % without a bound, the bound here only guarantees termination
    for (int it = 0; it < 10000; ++it) {
      r.dcrits = d_pow(k.cs / r.cgp, 1. / (k.dg - k.ds));
      r.dcritr = d_pow((1. + Fr / (1. - Fr)) * k.cs / r.cgp, 1. / (k.dg - k.ds));
      r.csr = k.cs * (1. + Fr / (1. - Fr));
      r.dsr = k.ds;
      const double rhodep = 1. / (r.dcritr - r.dcrits) * 6. * k.cs / (k.pi * (k.ds - 2.)) *
                            (d_pow(r.dcritr, k.ds - 2.) - d_pow(r.dcrits, k.ds - 2.));
      const double cgpold = r.cgp;
      r.cgp = crp * Fr + rhodep * (1. - Fr) * k.pi * k.sxth;
      if (fabs((r.cgp - cgpold) / r.cgp) < 0.01) break;
    }
  } else {
    r.dcrits = d_pow(k.cs / r.cgp, 1. / (k.dg - k.ds));
    r.dcritr = 1.e+6;
    r.csr = r.cgp;
    r.dsr = k.dg;
  }
  return r;
}

// mass = cs1 D^ds1, projected area = aas1 D^bas1 at size d (GEN:333-369)
struct MassArea {
  double cs1, ds1, bas1, aas1;
};

__device__ MassArea mass_area(const Consts& k, const Regime& r, double d) {
  MassArea m;
  if (d <= k.dcrit) {
    m.cs1 = k.pi * k.sxth * 900.;
    m.ds1 = 3.;
    m.bas1 = 2.;
    m.aas1 = k.pi / 4.;
  } else if (d <= r.dcrits) {
    m.cs1 = k.cs;
    m.ds1 = k.ds;
    m.bas1 = k.bas;
    m.aas1 = k.aas;
  } else if (d <= r.dcritr) {
    m.cs1 = r.cgp;
    m.ds1 = k.dg;
    m.bas1 = k.bag;
    m.aas1 = k.aag;
  } else {
    m.cs1 = r.csr;
    m.ds1 = r.dsr;
    m.bas1 = k.bas;
    if (r.i_Fr == 0) {
      m.aas1 = k.aas;
    } else {
      // area interpolated by mass between unrimed and fully rimed ice
      const double dum1 = k.aas * d_pow(d, k.bas);
      const double dum2 = k.aag * d_pow(d, k.bag);
      const double m1 = m.cs1 * d_pow(d, m.ds1);
      const double m2 = k.cs * d_pow(d, k.ds);
      const double m3 = r.cgp * d_pow(d, k.dg);
      const double dum3 = dum1 + (m1 - m2) * (dum2 - dum1) / (m3 - m2);
      m.aas1 = dum3 / d_pow(d, k.bas);
    }
  }
  return m;
}

// Mitchell and Heymsfield (2005) fall speed (GEN:377-394); GEN's turbulence
// terms have a0 = b0 = 0 and contribute exact zeros
__device__ double ice_fallspeed(const Consts& k, const MassArea& m, double d1) {
  const double mr = k.mu * k.rho;
  const double xx = 2. * m.cs1 * k.g * k.rho * d_pow(d1, m.ds1 + 2. - m.bas1) / (m.aas1 * (mr * mr));
  const double sx = sqrt(xx);
  const double s1 = sqrt(1. + k.c1 * sx);
  const double b1 = k.c1 * sx / (2. * (s1 - 1.) * s1) - 0. * 0. * d_pow(xx, 0.) / (k.c2 * ((s1 - 1.) * (s1 - 1.)));
  const double a1 = (k.c2 * ((s1 - 1.) * (s1 - 1.)) - 0. * d_pow(xx, 0.)) / d_pow(xx, b1);
  return a1 * d_pow(k.mu, 1. - 2. * b1) * d_pow(2. * m.cs1 * k.g / (k.rho * m.aas1), b1) *
         d_pow(d1, b1 * (m.ds1 - m.bas1 + 2.) - 1.);
}

// rain fall speed (GEN:480-495)
__device__ double rain_fallspeed(const Consts& k, double dia) {
  double amg = k.pi * k.sxth * 997. * (dia * dia * dia);
  amg = amg * 1000.;
  double dum2;
  if (dia <= 134.43e-6) dum2 = 4.5795e5 * d_pow(amg, 2. * k.thrd);
  else if (dia < 1511.64e-6) dum2 = 4.962e3 * d_pow(amg, k.thrd);
  else if (dia < 3477.84e-6) dum2 = 1.732e3 * d_pow(amg, k.sxth);
  else dum2 = 917.;
  return dum2 * 1.e-2;
}

// The mass moment's four pieces over the regimes (GEN:565-580, 615-627);
// clamp = the max(., 0) GEN applies inside the search only.
__device__ double mass_moment(const Consts& k, const Regime& r, double lam, double mu_i, bool clamp) {
  const double cs1 = k.pi * k.sxth * 900., ds1 = 3.;
  const double dum1 =
      d_pow(lam, -ds1 - mu_i - 1.) * cody_gamma(mu_i + ds1 + 1.) * (1. - gammq(mu_i + ds1 + 1., k.dcrit * lam));
  double dum2 = d_pow(lam, -k.ds - mu_i - 1.) * cody_gamma(mu_i + k.ds + 1.) * (gammq(mu_i + k.ds + 1., k.dcrit * lam));
  double dum = d_pow(lam, -k.ds - mu_i - 1.) * cody_gamma(mu_i + k.ds + 1.) * (gammq(mu_i + k.ds + 1., r.dcrits * lam));
  dum2 = dum2 - dum;
  if (clamp) dum2 = fmax(dum2, 0.);
  double dum3 = d_pow(lam, -k.dg - mu_i - 1.) * cody_gamma(mu_i + k.dg + 1.) * (gammq(mu_i + k.dg + 1., r.dcrits * lam));
  dum = d_pow(lam, -k.dg - mu_i - 1.) * cody_gamma(mu_i + k.dg + 1.) * (gammq(mu_i + k.dg + 1., r.dcritr * lam));
  dum3 = dum3 - dum;
  if (clamp) dum3 = fmax(dum3, 0.);
  const double dum4 =
      d_pow(lam, -r.dsr - mu_i - 1.) * cody_gamma(mu_i + r.dsr + 1.) * (gammq(mu_i + r.dsr + 1., r.dcritr * lam));
  return cs1 * dum1 + k.cs * dum2 + r.cgp * dum3 + r.csr * dum4;
}

__device__ __forceinline__ double mu_of_lam(double lam) {
  double mu_i = 0.076 * d_pow(lam / 100., 0.8) - 2.;
  mu_i = fmax(mu_i, 0.);
  return fmin(mu_i, 6.);
}

__device__ __forceinline__ double q_of_cell(int c) { return d_pow(261.7, (double)(c + 1 + 10) * 0.1) * 1.e-18; }

__device__ __forceinline__ double d_fine(int jj) { return (double)jj * 2. * 1.e-6 - 1.e-6; }   // jj 1-based
__device__ __forceinline__ double d_coarse(int jj) { return (double)jj * 20. * 1.e-6 - 10.e-6; }

struct Cell {
  double q, lam, mu_i, n0, nlarge, nsmall;
};

// Device buffers of one generation.
struct Work {
  Regime* regime;     // [kRF]
  double* fall1;      // [kRF][kFine]
  double* fall2;      // [kRF][kCoarse]
  double* fallr;      // [kCoarse]
  double* err;        // [kCells][kSearch] |q - q(lambda)|
  Cell* cell;         // [kCells]
  double* bin_rf;     // [kBinRF][kRF][kCoarse]: per-regime bin factors
  double* bin_cell;   // [kBinCell][kCoarse][kCells]: per-cell bin factors, bin-major
  double* bin_rain;   // [kBinRain][kDrscale][kCoarse]: per-rain-size bin factors
  double* lamrs;      // [kDrscale]
  double* ice;        // [kCells][kIceCols]
  double* coll;       // [kCells][kDrscale][kCollCols]
};

// per-regime bin factors (20-micron bins)
enum { BR_A1, BR_M1, BR_M2, BR_M3, BR_DA1, BR_DA2, BR_PB, BR_A2, BR_BIG, BR_CS, BR_D, kBinRF };
// per-cell bin factors
enum { BC_NUM, BC_P, BC_E, kBinCell };
// per-rain-size bin factors
enum { RN_B, RN_C, RN_R, RN_F, RN_NUM, kBinRain };

__device__ __forceinline__ double& brf(const Work& w, int f, int rf, int j) {
  return w.bin_rf[((size_t)f * kRF + rf) * kCoarse + j];
}
__device__ __forceinline__ double& bcell(const Work& w, int f, int j, int c) {
  return w.bin_cell[((size_t)f * kCoarse + j) * kCells + c];
}
__device__ __forceinline__ double& brain(const Work& w, int f, int d, int k) {
  return w.bin_rain[((size_t)f * kDrscale + d) * kCoarse + k];
}

// ---- kernels

__global__ void k_regime(Work w) {
  const int rf = threadIdx.x;
  if (rf >= kRF) return;
  const Consts k = make_consts();
  w.regime[rf] = make_regime(k, rf);
}

// fall speeds on both size grids (GEN:320-497)
__global__ void k_fall(Work w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const Consts k = make_consts();
  if (t < kRF * kFine) {
    const int rf = t / kFine, jj = t % kFine + 1;
    const double d1 = d_fine(jj);
    w.fall1[t] = ice_fallspeed(k, mass_area(k, w.regime[rf], d1), d1);
  } else if (t < kRF * kFine + kRF * kCoarse) {
    const int u = t - kRF * kFine, rf = u / kCoarse, jj = u % kCoarse + 1;
    const double d1 = d_coarse(jj);
    w.fall2[u] = ice_fallspeed(k, mass_area(k, w.regime[rf], d1), d1);
    if (rf == 0) w.fallr[jj - 1] = rain_fallspeed(k, d1);
  }
}

// |q - q(lambda)| for every cell and candidate slope (GEN:540-590)
__global__ void k_search(Work w) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)kCells * kSearch) return;
  const int c = (int)(t / kSearch), ii = (int)(t % kSearch) + 1;
  const Consts k = make_consts();
  const Regime r = w.regime[c / kQnorm];
  const double Dm_max = 2000.e-6, Dm_min = 2.e-6;
  double lam = powi(1.0013, ii) * 100.;
  const double mu_i = mu_of_lam(lam);
  lam = fmax(lam, (mu_i + 1.) / Dm_max);
  lam = fmin(lam, (mu_i + 1.) / Dm_min);
  const double n0 = d_pow(lam, mu_i + 1.) / (cody_gamma(mu_i + 1.));
  const double qdum = n0 * mass_moment(k, r, lam, mu_i, true);
  w.err[t] = fabs(q_of_cell(c % kQnorm) - qdum);
}

// GEN's selection (first strict minimum, from candidate 1), then n0 and the
// lambda-limiter numbers (GEN:584-700)
__global__ void k_select(Work w) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kCells) return;
  const Consts k = make_consts();
  const Regime r = w.regime[c / kQnorm];
  const double Dm_max = 2000.e-6, Dm_min = 2.e-6;
  const double* e = w.err + (size_t)c * kSearch;
  double qerror = e[0];
  int best = 1;
  for (int ii = 2; ii <= kSearch; ++ii) {
    if (e[ii - 1] < qerror) {
      qerror = e[ii - 1];
      best = ii;
    }
  }
  double lam = powi(1.0013, best) * 100.;
  double mu_i = mu_of_lam(lam);
  lam = fmax(lam, (mu_i + 1.) / Dm_max);
  lam = fmin(lam, (mu_i + 1.) / Dm_min);
  mu_i = mu_of_lam(lam);
  Cell out;
  out.q = q_of_cell(c % kQnorm);
  out.lam = lam;
  out.mu_i = mu_i;
  out.n0 = out.q / mass_moment(k, r, lam, mu_i, false);
  double duml = (mu_i + 1.) / Dm_min;
  double n0dum = out.q / mass_moment(k, r, duml, mu_i, false);
  out.nlarge = n0dum / (d_pow(duml, mu_i + 1.) / (cody_gamma(mu_i + 1.)));
  duml = (mu_i + 1.) / (Dm_max);
  n0dum = out.q / mass_moment(k, r, duml, mu_i, false);
  out.nsmall = n0dum / (d_pow(duml, mu_i + 1.) / (cody_gamma(mu_i + 1.)));
  w.cell[c] = out;
}

// bin factors: per regime (t < kRF*kCoarse), per cell, per rain size
__global__ void k_bins(Work w) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const Consts k = make_consts();
  const size_t n_rf = (size_t)kRF * kCoarse, n_cell = (size_t)kCells * kCoarse;
  if (t < n_rf) {
    const int rf = (int)(t / kCoarse), j = (int)(t % kCoarse);
    const Regime& r = w.regime[rf];
    const double d = d_coarse(j + 1);
    const MassArea m = mass_area(k, r, d);
    brf(w, BR_A1, rf, j) = m.aas1 * d_pow(d, m.bas1);
    // particle 2 of the aggregation pair (GEN:863-889): GEN interpolates its
    // area with particle 1's masses, so keep the pieces separately
    brf(w, BR_M1, rf, j) = r.csr * d_pow(d, r.dsr);
    brf(w, BR_M2, rf, j) = k.cs * d_pow(d, k.ds);
    brf(w, BR_M3, rf, j) = r.cgp * d_pow(d, k.dg);
    brf(w, BR_DA1, rf, j) = k.aas * d_pow(d, k.bas);
    brf(w, BR_DA2, rf, j) = k.aag * d_pow(d, k.bag);
    brf(w, BR_PB, rf, j) = d_pow(d, k.bas);
    brf(w, BR_BIG, rf, j) = (d > r.dcritr && r.i_Fr != 0) ? 1. : 0.;
    if (d <= k.dcrit) brf(w, BR_A2, rf, j) = (k.pi / 4.) * d_pow(d, 2.);
    else if (d <= r.dcrits) brf(w, BR_A2, rf, j) = k.aas * d_pow(d, k.bas);
    else brf(w, BR_A2, rf, j) = k.aag * d_pow(d, k.bag);
    brf(w, BR_CS, rf, j) = m.cs1;
    brf(w, BR_D, rf, j) = d_pow(d, m.ds1);
  } else if (t < n_rf + n_cell) {
    const size_t u = t - n_rf;
    const int j = (int)(u / kCells), c = (int)(u % kCells);
    const Cell cl = w.cell[c];
    const double dd = 20.e-6;
    const double d1 = (double)(j + 1) * dd - 10.e-6;  // GEN:818 spells the bin this way
    bcell(w, BC_NUM, j, c) = cl.n0 * d_pow(d1, cl.mu_i) * d_exp(-cl.lam * d1) * dd;
    const double d = d_coarse(j + 1);
    bcell(w, BC_P, j, c) = d_pow(d, cl.mu_i);
    bcell(w, BC_E, j, c) = d_exp(-cl.lam * d);
  } else if (t < n_rf + n_cell + (size_t)kDrscale * kCoarse) {
    const size_t u = t - n_rf - n_cell;
    const int dr = (int)(u / kCoarse), kk = (int)(u % kCoarse);
    // GEN:1024-1062: the rain slope of this rain size (mu_r = 0)
    const double lamv = 1. / (powi(1.24, dr + 1) * 10.e-6);
    const double mu_r = 0.;
    double lamr = lamv * d_pow(cody_gamma(mu_r + 4.) / (6. * cody_gamma(mu_r + 1.)), k.thrd);
    lamr = fmin(lamr, (mu_r + 1.) * 1.e+5);
    lamr = fmax(lamr, (mu_r + 1.) * 200.);
    if (kk == 0) w.lamrs[dr] = lamv;
    const double d2 = d_coarse(kk + 1), dd = 20.e-6;
    brain(w, RN_B, dr, kk) = k.pi * 0.25 * (d2 * d2);
    brain(w, RN_C, dr, kk) = d2 * d2 * d2;
    brain(w, RN_R, dr, kk) = d_pow(d2, mu_r);
    brain(w, RN_F, dr, kk) = d_exp(-lamr * d2);
    brain(w, RN_NUM, dr, kk) = d_pow(d2, mu_r) * d_exp(-lamr * d2) * dd;
  }
}

// The four 2-micron integrals of a cell (GEN:713-804, 940-1004, 1214-1277,
// 1284-1349), one thread per (cell, integral), each summed in GEN's order.
__global__ void k_sums(Work w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * kCells) return;
  const int which = t / kCells, c = t % kCells;
  const Consts k = make_consts();
  const Regime r = w.regime[c / kQnorm];
  const Cell cl = w.cell[c];
  const double* fall1 = w.fall1 + (size_t)(c / kQnorm) * kFine;
  const double lam = cl.lam, mu_i = cl.mu_i, n0 = cl.n0, pi = k.pi, sxth = k.sxth;
  double* ice = w.ice + (size_t)c * kIceCols;
  const double dd = 2.e-6;
  if (which == 0) {  // fall speeds, reflectivity, mean size and density
    double s1 = 0., s2 = 0., s3 = 0., s4 = 0., s5 = 0., s6 = 0., s7 = 0.;
    for (int ii = 1; ii <= kFine; ++ii) {
      const double dum = (double)ii * dd - 1.e-6;
      const MassArea m = mass_area(k, r, dum);
      const double e = d_exp(-lam * dum), pm = d_pow(dum, mu_i), f1 = fall1[ii - 1];
      s1 = s1 + f1 * pm * e * dd;
      s2 = s2 + f1 * m.cs1 * d_pow(dum, m.ds1 + mu_i) * e * dd;
      s3 = s3 + pm * e * dd;
      s4 = s4 + m.cs1 * d_pow(dum, m.ds1 + mu_i) * e * dd;
      const double a = m.cs1 / 917., b = 6. / pi;
      s5 = s5 + n0 * (a * a) * (b * b) * d_pow(dum, 2. * m.ds1 + mu_i) * e * dd;
      s6 = s6 + m.cs1 * d_pow(dum, m.ds1 + mu_i + 1.) * e * dd;
      const double ms = m.cs1 * d_pow(dum, m.ds1);
      s7 = s7 + (ms * ms) / (pi * sxth * (dum * dum * dum)) * pm * e * dd;
    }
    ice[2] = s1 / s3;   // uns
    ice[3] = s2 / s4;   // ums
    ice[11] = s5;       // refl
    ice[13] = s6 / s4;  // dmm
    ice[14] = s7 / s4;  // rhomm
  } else if (which == 1) {  // collection of cloud droplets
    double s1 = 0.;
    for (int jj = 1; jj <= kFine; ++jj) {
      const double d1 = d_fine(jj);
      const MassArea m = mass_area(k, r, d1);
      if (d1 >= 100.e-6) s1 = s1 + m.aas1 * d_pow(d1, m.bas1) * fall1[jj - 1] * n0 * d_pow(d1, mu_i) * d_exp(-lam * d1) * dd;
    }
    ice[5] = s1;  // nrwat
  } else if (which == 2) {  // vapour deposition with ventilation
    double s1 = 0., s2 = 0.;
    for (int jj = 1; jj <= kFine; ++jj) {
      const double d1 = d_fine(jj);
      double cap;
      if (d1 <= k.dcrit) cap = 1.;
      else if (d1 <= r.dcrits) cap = 0.48;
      else if (d1 <= r.dcritr) cap = 1.;
      else if (r.i_Fr == 0) cap = 0.48;
      else {
        const double m1 = r.csr * d_pow(d1, r.dsr), m2 = k.cs * d_pow(d1, k.ds), m3 = r.cgp * d_pow(d1, k.dg);
        cap = 0.48 + (m1 - m2) * (1. - 0.48) / (m3 - m2);
      }
      const double dum = sqrt(fall1[jj - 1] * d1);
      if (d1 < 100.e-6) {
        s1 = s1 + cap * n0 * d_pow(d1, mu_i + 1.) * d_exp(-lam * d1) * dd;
      } else {
        s1 = s1 + cap * n0 * 0.65 * d_pow(d1, mu_i + 1.) * d_exp(-lam * d1) * dd;
        s2 = s2 + cap * n0 * 0.44 * dum * d_pow(d1, mu_i + 1.) * d_exp(-lam * d1) * dd;
      }
    }
    ice[6] = s1;   // vdep
    ice[12] = s2;  // vdep1
  } else {  // effective radius
    double s1 = 0., s2 = 0.;
    for (int jj = 1; jj <= kFine; ++jj) {
      const double d1 = d_fine(jj);
      const MassArea m = mass_area(k, r, d1);
      s1 = s1 + m.cs1 * d_pow(d1, m.ds1) * d_pow(d1, mu_i) * d_exp(-lam * d1) * dd;
      s2 = s2 + m.aas1 * d_pow(d1, m.bas1) * d_pow(d1, mu_i) * d_exp(-lam * d1) * dd;
    }
    ice[7] = 3. * s1 / (4. * s2 * 916.7);  // eff
  }
}

// self-aggregation over pairs of 20-micron bins (GEN:810-928), one thread per cell
__global__ void k_agg(Work w) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kCells) return;
  const int rf = c / kQnorm;
  const double* f2 = w.fall2 + (size_t)rf * kCoarse;
  double sum = 0.;
  for (int jj = kCoarse - 1; jj >= 0; --jj) {
    const double a1 = brf(w, BR_A1, rf, jj), nj = bcell(w, BC_NUM, jj, c), fj = f2[jj];
    const double m1 = brf(w, BR_M1, rf, jj), m2 = brf(w, BR_M2, rf, jj), m3 = brf(w, BR_M3, rf, jj);
    for (int kk = 0; kk < jj; ++kk) {
      double a2;
      if (brf(w, BR_BIG, rf, kk) != 0.) {
        const double da1 = brf(w, BR_DA1, rf, kk), pb = brf(w, BR_PB, rf, kk);
        const double dum3 = da1 + (m1 - m2) * (brf(w, BR_DA2, rf, kk) - da1) / (m3 - m2);
        a2 = dum3 / pb * pb;
      } else {
        a2 = brf(w, BR_A2, rf, kk);
      }
      const double delu = fabs(fj - f2[kk]);
      sum = sum + (a1 + a2) * delu * nj * bcell(w, BC_NUM, kk, c);
    }
  }
  w.ice[(size_t)c * kIceCols + 4] = sum;  // nagg
}

// ice-rain collection (GEN:1066-1200): one thread per (cell, rain size); the
// block's rain size is uniform, so its rain-bin factors are staged in LDS.
constexpr int kRainThreads = 64;
__global__ __launch_bounds__(kRainThreads) void k_rain(Work w) {
  __shared__ double s_fr[kCoarse], s_b[kCoarse], s_c[kCoarse], s_r[kCoarse], s_f[kCoarse], s_n[kCoarse];
  const int groups = (kCells + kRainThreads - 1) / kRainThreads;
  const int dr = blockIdx.x / groups;
  const int c = (blockIdx.x % groups) * kRainThreads + threadIdx.x;
  for (int kk = threadIdx.x; kk < kCoarse; kk += kRainThreads) {
    s_fr[kk] = w.fallr[kk];
    s_b[kk] = brain(w, RN_B, dr, kk);
    s_c[kk] = brain(w, RN_C, dr, kk);
    s_r[kk] = brain(w, RN_R, dr, kk);
    s_f[kk] = brain(w, RN_F, dr, kk);
    s_n[kk] = brain(w, RN_NUM, dr, kk);
  }
  __syncthreads();
  if (c >= kCells) return;
  const int rf = c / kQnorm;
  const Consts k = make_consts();
  const double pi = k.pi, sxth = k.sxth, dd = 20.e-6, n0 = w.cell[c].n0;
  const double* f2 = w.fall2 + (size_t)rf * kCoarse;
  double s1 = 0., s2 = 0., s6 = 0.;
  for (int jj = 0; jj < kCoarse; ++jj) {
    const double aj = brf(w, BR_A1, rf, jj), fj = f2[jj];
    const double pj = bcell(w, BC_P, jj, c), ej = bcell(w, BC_E, jj, c);
    const double numi = n0 * pj * ej * dd;
    const double csj = brf(w, BR_CS, rf, jj), dj = brf(w, BR_D, rf, jj);
    for (int kk = 0; kk < kCoarse; ++kk) {
      const double delu = fabs(fj - s_fr[kk]);
      const double area = aj + s_b[kk];
      const double t = area * delu * n0 * pj * ej * dd * s_n[kk];
      s1 = s1 + t;
      s2 = s2 + t * pi * sxth * 997. * s_c[kk];
      s6 = s6 + area * delu * s_r[kk] * s_f[kk] * dd * numi * csj * dj;
    }
  }
  double* o = w.coll + ((size_t)c * kDrscale + dr) * kCollCols;
  o[3] = s1;
  o[4] = s2;
  o[5] = s6;
}

__device__ __forceinline__ double clamp99(double v) { return v < 1.e-99 ? 0. : v; }

// the output rows with GEN's clamp of values below 1e-99 (GEN:1384-1420)
__global__ void k_rows(Work w) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kCells) return;
  const Cell cl = w.cell[c];
  double* ice = w.ice + (size_t)c * kIceCols;
  ice[0] = cl.q;
  ice[1] = 1.;
  ice[8] = cl.nlarge;
  ice[9] = cl.nsmall;
  ice[10] = cl.lam;
  for (int j = 0; j < kIceCols; ++j) ice[j] = clamp99(ice[j]);
  for (int d = 0; d < kDrscale; ++d) {
    double* o = w.coll + ((size_t)c * kDrscale + d) * kCollCols;
    o[0] = clamp99(cl.q);
    o[1] = 1.;
    o[2] = clamp99(w.lamrs[d]);
    for (int j = 3; j < kCollCols; ++j) o[j] = clamp99(o[j]);
  }
}

// GEN's e15.5 field (GEN:1374): 0.ddddd, five correctly rounded significant
// digits, then E+xx, right-aligned in 15 characters
void fortran_e15_5(char* out, double v) {
  if (v == 0.) {
    std::snprintf(out, 16, "%15s", "0.00000E+00");
    return;
  }
  char m[32];
  std::snprintf(m, sizeof m, "%.4e", v < 0 ? -v : v);  // d.dddde+xx, correctly rounded
  const int e = std::atoi(m + 7) + 1;
  char body[32];
  if (e >= -99 && e <= 99)
    std::snprintf(body, sizeof body, "%s0.%c%c%c%c%cE%c%02d", v < 0 ? "-" : "", m[0], m[2], m[3], m[4], m[5],
                  e < 0 ? '-' : '+', e < 0 ? -e : e);
  else  // three exponent digits drop the E, as Fortran does
    std::snprintf(body, sizeof body, "%s0.%c%c%c%c%c%c%03d", v < 0 ? "-" : "", m[0], m[2], m[3], m[4], m[5],
                  e < 0 ? '-' : '+', e < 0 ? -e : e);
  std::snprintf(out, 16, "%15s", body);
}

int dev_err(hipError_t e) { return e == hipErrorOutOfMemory ? P3X_ERR_NO_MEMORY : P3X_ERR_DEVICE; }

}  // namespace

extern "C" int p3x_ice_table_generate(double* ice_rows, double* coll_rows, void* stream, double* elapsed_ms) {
  if (!ice_rows || !coll_rows) return P3X_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const auto t0 = std::chrono::steady_clock::now();
  Work w{};
  std::vector<void*> allocs;
  hipError_t e = hipSuccess;
  auto alloc = [&](auto** p, size_t n) {
    if (e != hipSuccess) return;
    e = hipMalloc((void**)p, n * sizeof(**p));
    if (e == hipSuccess) allocs.push_back((void*)*p);
  };
  alloc(&w.regime, kRF);
  alloc(&w.fall1, (size_t)kRF * kFine);
  alloc(&w.fall2, (size_t)kRF * kCoarse);
  alloc(&w.fallr, kCoarse);
  alloc(&w.err, (size_t)kCells * kSearch);
  alloc(&w.cell, kCells);
  alloc(&w.bin_rf, (size_t)kBinRF * kRF * kCoarse);
  alloc(&w.bin_cell, (size_t)kBinCell * kCoarse * kCells);
  alloc(&w.bin_rain, (size_t)kBinRain * kDrscale * kCoarse);
  alloc(&w.lamrs, kDrscale);
  alloc(&w.ice, (size_t)kCells * kIceCols);
  alloc(&w.coll, (size_t)kCells * kDrscale * kCollCols);
  auto launch = [&](auto kern, size_t threads, int block) {
    if (e != hipSuccess) return;
    const unsigned grid = (unsigned)((threads + block - 1) / block);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, w);
    e = hipGetLastError();
  };
  launch(k_regime, kRF, 32);
  launch(k_fall, (size_t)kRF * kFine + (size_t)kRF * kCoarse, 256);
  launch(k_search, (size_t)kCells * kSearch, 256);
  launch(k_select, kCells, 64);
  launch(k_bins, (size_t)kRF * kCoarse + (size_t)kCells * kCoarse + (size_t)kDrscale * kCoarse, 256);
  launch(k_sums, (size_t)4 * kCells, 64);
  launch(k_agg, kCells, 64);
  if (e == hipSuccess) {
    const int groups = (kCells + kRainThreads - 1) / kRainThreads;
    hipLaunchKernelGGL(k_rain, dim3(groups * kDrscale), dim3(kRainThreads), 0, s, w);
    e = hipGetLastError();
  }
  launch(k_rows, kCells, 64);
  if (e == hipSuccess) e = hipMemcpyAsync(ice_rows, w.ice, sizeof(double) * kCells * kIceCols, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(coll_rows, w.coll, sizeof(double) * kCells * kDrscale * kCollCols, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  const hipError_t e_sync = hipStreamSynchronize(s);  // drain before freeing, whatever happened above
  for (void* p : allocs) (void)hipFree(p);
  if (e == hipSuccess) e = e_sync;
  if (e != hipSuccess) return dev_err(e);
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return P3X_OK;
}

extern "C" int p3x_ice_table_write(const char* path, const double* ice_rows, const double* coll_rows) {
  if (!path || !ice_rows || !coll_rows) return P3X_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return P3X_ERR_TABLE_FILE;
  std::fprintf(f, "VERSION 4.1.1\n");
  char buf[16];
  for (int rf = 0; rf < kRF; ++rf) {
    const int i_rhor = rf / kFr + 1, i_Fr = rf % kFr + 1;
    for (int q = 0; q < kQnorm; ++q) {
      std::fprintf(f, "%5d%5d", i_rhor, i_Fr);
      const double* v = ice_rows + ((size_t)rf * kQnorm + q) * kIceCols;
      for (int j = 0; j < kIceCols; ++j) {
        fortran_e15_5(buf, v[j]);
        std::fputs(buf, f);
      }
      std::fputc('\n', f);
    }
    for (int q = 0; q < kQnorm; ++q) {
      for (int d = 0; d < kDrscale; ++d) {
        std::fprintf(f, "%5d%5d", i_rhor, i_Fr);
        const double* v = coll_rows + (((size_t)rf * kQnorm + q) * kDrscale + d) * kCollCols;
        for (int j = 0; j < kCollCols; ++j) {
          fortran_e15_5(buf, v[j]);
          std::fputs(buf, f);
        }
        std::fputc('\n', f);
      }
    }
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? P3X_OK : P3X_ERR_TABLE_FILE;
}
