/*
 * p3_math.h -- FP64 elementary functions that give the same bits on the CPU
 * (gcc/clang, x86-64) and on CDNA4 (hipcc, gfx950).
 *
 * Why: P3 contains quantities that are roundoff residuals of near-complete
 * cancellations (e.g. an ice number that its sinks almost exactly remove) and
 * later decisions (table cells, which level's ice falls how fast) depend on
 * them.  Two correct libms that differ by one ulp therefore give columns whose
 * sedimentation differs at the 1e-5 level.  Built only from IEEE-754
 * operations that both targets round identically (+, -, *, /, sqrt, fma,
 * integer conversion and bit manipulation, with no floating-point
 * contraction), these functions make the HIP kernel and a CPU build of the
 * oracle produce identical results.  Accuracy is a few ulps; the published
 * methods used are named per function.
 *
 * Usable from C (gcc) and from HIP C++ host and device code.
 */
#ifndef P3_MATH_H
#define P3_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define P3M_FN static __host__ __device__ __forceinline__
#else
#include <math.h>
#define P3M_FN static inline
#endif

#define P3M_LN2_HI 6.93147180369123816490e-01    /* high 32 bits of ln 2 */
#define P3M_LN2_LO 1.9082149292705877e-10        /* ln 2 - P3M_LN2_HI     */
#define P3M_INVLN2 1.4426950408889634
#define P3M_INVLN10_HI 0.4342944819032518
#define P3M_INVLN10_LO 1.098319650216765e-17
#define P3M_SQRT2 1.4142135623730951
#define P3M_SQRT2PI 2.5066282746310007

P3M_FN uint64_t p3m_bits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
P3M_FN double p3m_from_bits(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }
P3M_FN double p3m_inf(void) { return p3m_from_bits(0x7ff0000000000000ULL); }
P3M_FN double p3m_nan(void) { return p3m_from_bits(0x7ff8000000000000ULL); }

/* x * 2^k for x in [0.5, 2) and any k that keeps the result finite or subnormal */
P3M_FN double p3m_scale2(double x, int k) {
  if (k > 1023) { x *= p3m_from_bits((uint64_t)(1023 + 1023) << 52); k -= 1023; }
  if (k < -1022) { x *= p3m_from_bits((uint64_t)(-1000 + 1023) << 52); k += 1000;
                   if (k < -1022) { x *= p3m_from_bits((uint64_t)(-1000 + 1023) << 52); k += 1000; } }
  return x * p3m_from_bits((uint64_t)(k + 1023) << 52);
}

/* exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, then the Taylor
 * series of e^r to r^13 (truncation < 1e-17), evaluated with fma. */
P3M_FN double p3m_exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return p3m_inf();
  if (x < -745.1332191019412) return 0.0;
  const double t = x * P3M_INVLN2;
  const int k = (int)(t + (t >= 0 ? 0.5 : -0.5));
  const double kd = (double)k;
  double r = fma(-kd, P3M_LN2_HI, x);
  r = fma(-kd, P3M_LN2_LO, r);
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return p3m_scale2(p, k);
}

/* log as an unevaluated sum hi + lo.  Reduction x = 2^e m, m in (sqrt(1/2), sqrt(2)],
 * f = m - 1, s = f/(2+f), log m = 2 atanh s = f - f^2/2 + s (f^2/2 + R) with
 * R = sum_{k>=1} 2 s^(2k)/(2k+1) (the formulation of Sun's fdlibm __ieee754_log,
 * here with the exact series coefficients up to s^22). */
P3M_FN void p3m_log_dd(double x, double* hi, double* lo) {
  uint64_t u = p3m_bits(x);
  int e = 0;
  /* One unsigned compare separates the positive normal finite arguments (the
   * only ones P3 passes) from zero, subnormals, negatives, infinities and NaN. */
  if (u - 0x0010000000000000ULL >= 0x7fe0000000000000ULL) {
    if (!(x == x) || x < 0) { *hi = p3m_nan(); *lo = 0; return; }
    if (x == 0) { *hi = -p3m_inf(); *lo = 0; return; }
    if (u >= 0x7ff0000000000000ULL) { *hi = x; *lo = 0; return; }
    x *= 18014398509481984.0; u = p3m_bits(x); e = -54;  /* subnormal */
  }
  e += (int)(u >> 52) - 1023;
  double m = p3m_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > P3M_SQRT2) { m *= 0.5; e += 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 2.0 / 23.0;
  R = fma(R, z, 2.0 / 21.0);
  R = fma(R, z, 2.0 / 19.0);
  R = fma(R, z, 2.0 / 17.0);
  R = fma(R, z, 2.0 / 15.0);
  R = fma(R, z, 2.0 / 13.0);
  R = fma(R, z, 2.0 / 11.0);
  R = fma(R, z, 2.0 / 9.0);
  R = fma(R, z, 2.0 / 7.0);
  R = fma(R, z, 2.0 / 5.0);
  R = fma(R, z, 2.0 / 3.0);
  R = R * z;
  const double hfsq = 0.5 * f * f;
  const double ed = (double)e;
  const double B = f - (hfsq - (s * (hfsq + R) + ed * P3M_LN2_LO));
  const double A = ed * P3M_LN2_HI;  /* exact */
  const double h = A + B;
  *hi = h;
  *lo = (A - h) + B;  /* |A| >= |B| whenever A != 0 */
}

P3M_FN double p3m_log(double x) {
  double h, l;
  p3m_log_dd(x, &h, &l);
  return h + l;
}

P3M_FN double p3m_log10(double x) {
  double h, l;
  p3m_log_dd(x, &h, &l);
  if (!(h == h) || h == p3m_inf() || h == -p3m_inf()) return h;
  return fma(h, P3M_INVLN10_HI, fma(h, P3M_INVLN10_LO, l * P3M_INVLN10_HI));
}

/* pow for the arguments P3 uses (x >= 0): exp(y log x) with log x in
 * double-double and the product's rounding error carried into exp.  A negative
 * base with an integral exponent is |x|^y with the sign of (-1)^y.  (No
 * recursion, so that the device compiler can inline it everywhere.) */
P3M_FN double p3m_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (!(x == x) || !(y == y)) return x + y;
  if (x == 1.0) return 1.0;
  if (y == 1.0) return x;
  if (y == 2.0) return x * x;
  double sgn = 1.0;
  if (x < 0) {
    const double yi = (double)(int64_t)y;
    if (yi != y || y > 9007199254740992.0 || y < -9007199254740992.0) return p3m_nan();
    sgn = ((int64_t)y & 1) ? -1.0 : 1.0;
    x = -x;
    if (x == 1.0) return sgn;
  }
  if (x == 0.0) return y > 0 ? 0.0 : p3m_inf();
  if (p3m_bits(x) == 0x7ff0000000000000ULL) return sgn * (y > 0 ? x : 0.0);
  double lh, ll;
  p3m_log_dd(x, &lh, &ll);
  const double z = y * lh;
  const double zl = fma(y, lh, -z) + y * ll;
  if (z > 709.8) return sgn * p3m_inf();
  if (z < -745.2) return sgn * 0.0;
  const double e = p3m_exp(z);
  return sgn * fma(e, zl, e);
}

/* cbrt: inverse cube root y ~ a^(-1/3) from the exponent bits, refined by the
 * division-free Newton iteration y <- y + y (1 - a y^3)/3, then one Newton step
 * on c^3 = a for c = a y^2 with the residual c^3 - a formed exactly with fma. */
P3M_FN double p3m_cbrt(double x) {
  if (x == 0.0 || !(x == x) || p3m_bits(x < 0 ? -x : x) >= 0x7ff0000000000000ULL) return x;
  const double sgn = x < 0 ? -1.0 : 1.0;
  double a = x < 0 ? -x : x;
  double post = 1.0;
  if (a < 2.2250738585072014e-308) { a *= 18014398509481984.0; post = 1.0 / 262144.0; }
  double y = p3m_from_bits(0x553ef0ff289dd796ULL - p3m_bits(a) / 3);
  for (int it = 0; it < 5; ++it) {
    const double y3 = y * y * y;
    y = fma(y * (1.0 / 3.0), fma(-a, y3, 1.0), y);
  }
  double c = a * y * y;
  const double c2 = c * c;
  const double e2 = fma(c, c, -c2);
  const double c3 = c2 * c;
  const double e3 = fma(c2, c, -c3) + e2 * c;
  const double r = (c3 - a) + e3;
  c = c - r / (3.0 * c2);
  return sgn * c * post;
}

/* expm1: Taylor series to x^16 for |x| < 1/2, exp(x) - 1 otherwise. */
P3M_FN double p3m_expm1(double x) {
  if (!(x == x)) return x;
  const double ax = x < 0 ? -x : x;
  if (ax < 0.5) {
    double p = 1.0 / 20922789888000.0;     /* 1/16! */
    p = fma(p, x, 1.0 / 1307674368000.0);  /* 1/15! */
    p = fma(p, x, 1.0 / 87178291200.0);
    p = fma(p, x, 1.0 / 6227020800.0);
    p = fma(p, x, 1.0 / 479001600.0);
    p = fma(p, x, 1.0 / 39916800.0);
    p = fma(p, x, 1.0 / 3628800.0);
    p = fma(p, x, 1.0 / 362880.0);
    p = fma(p, x, 1.0 / 40320.0);
    p = fma(p, x, 1.0 / 5040.0);
    p = fma(p, x, 1.0 / 720.0);
    p = fma(p, x, 1.0 / 120.0);
    p = fma(p, x, 1.0 / 24.0);
    p = fma(p, x, 1.0 / 6.0);
    p = fma(p, x, 0.5);
    return fma(p * x, x, x);
  }
  if (x < -40.0) return -1.0;
  return p3m_exp(x) - 1.0;
}

/* tanh(x) = e/(e+2) with e = expm1(2|x|). */
P3M_FN double p3m_tanh(double x) {
  if (!(x == x)) return x;
  const double ax = x < 0 ? -x : x;
  if (ax > 22.0) return x < 0 ? -1.0 : 1.0;
  if (ax < 3.725290298461914e-09) return x;
  const double e = p3m_expm1(2.0 * ax);
  const double t = e / (e + 2.0);
  return x < 0 ? -t : t;
}

/* tgamma for x > 0: recurrence Gamma(x) = (x-1) Gamma(x-1) down to [1, 2),
 * Gamma(x) = Gamma(x+1)/x below 1, and on [1, 2) the Lanczos approximation
 * with g = 7, n = 9 (coefficients as published with the method, e.g. in
 * Press et al. / Godfrey), Gamma(z+1) = sqrt(2 pi) t^(z+1/2) e^(-t) A(z),
 * t = z + 7.5, with the partial-fraction sum A taken over one common
 * denominator so that a single division is needed. */
P3M_FN double p3m_tgamma(double x) {
  if (!(x == x)) return x;
  if (x <= 0.0) return p3m_nan();
  if (x > 171.6) return p3m_inf();
  double prod = 1.0, div = 1.0;
  if (x < 1.0) { div = x; x = x + 1.0; }
  while (x >= 2.0) { x = x - 1.0; prod *= x; }
  const double z = x - 1.0;
  const double p0 = 0.99999999999980993, p1 = 676.5203681218851, p2 = -1259.1392167224028,
               p3 = 771.32342877765313, p4 = -176.61502916214059, p5 = 12.507343278686905,
               p6 = -0.13857109526572012, p7 = 9.9843695780195716e-6, p8 = 1.5056327351493116e-7;
  const double d1 = z + 1.0, d2 = z + 2.0, d3 = z + 3.0, d4 = z + 4.0, d5 = z + 5.0, d6 = z + 6.0,
               d7 = z + 7.0, d8 = z + 8.0;
  /* prefix / suffix products: 1/(z+i) = pre_i * suf_i / (d1...d8) */
  const double q12 = d1 * d2, q34 = d3 * d4, q56 = d5 * d6, q78 = d7 * d8;
  const double q1234 = q12 * q34, q5678 = q56 * q78;
  double num = p1 * (d2 * q34 * q5678);
  num = fma(p2, d1 * q34 * q5678, num);
  num = fma(p3, q12 * d4 * q5678, num);
  num = fma(p4, q12 * d3 * q5678, num);
  num = fma(p5, q1234 * d6 * q78, num);
  num = fma(p6, q1234 * d5 * q78, num);
  num = fma(p7, q1234 * q56 * d8, num);
  num = fma(p8, q1234 * q56 * d7, num);
  const double A = p0 + num / (q1234 * q5678);
  const double t = z + 7.5;
  double lh, ll;
  p3m_log_dd(t, &lh, &ll);
  const double zh = z + 0.5;
  /* w + wl = zh*(lh+ll) - t: exact product error by fma, exact sum error by
   * Fast2Sum (|t| > |zh*lh| on [1, 2)) */
  const double ph = zh * lh;
  const double pl = fma(zh, lh, -ph);
  const double w = ph - t;
  const double wl = (ph - (w + t)) + pl + zh * ll;
  const double e = p3m_exp(w);
  const double g = P3M_SQRT2PI * fma(e, wl, e) * A;
  return g * prod / div;
}

#endif /* P3_MATH_H */
