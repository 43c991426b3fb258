// p3_physics.hpp -- P3 per-level physics as CDNA4 device functions.
//
// Everything here is a pointwise function of one (column, level) point; the
// kernel in p3_kernel.hip arranges the points.  The rate functions are
// templated on the arithmetic type R of the bulk process rates:
//   R = double : the FP64 drop-in; expressions keep the reference's operand
//                order so results follow the reference's rounding.
//   R = float  : the mixed form; saturation, the conservation limiters
//                (water, number, supersaturation), the prognostic updates and
//                every small difference of large FP64 quantities (qv - qsat,
//                T - Tmelt, qv - qv_prev) stay FP64.
// Function-by-function citations are to
//   components/eamxx/src/physics/p3/impl/*_impl.hpp             (P3I)
//   components/eamxx/src/physics/share/physics_saturation_impl.hpp (PSH)
#pragma once

#include <hip/hip_runtime.h>

#include "p3_constants.hpp"
#include "p3_math.h"

namespace p3d {

using namespace p3c;

// ---------------------------------------------------------------- math shims
// FP64: the portable functions of p3_math.h, which round identically on the
// CPU (the oracle's portable build) and here, so the FP64 form reproduces the
// oracle bit for bit.  FP32 (mixed form): the hardware-backed float routines.
__device__ __forceinline__ double m_exp(double x) { return p3m_exp(x); }
__device__ __forceinline__ float m_exp(float x) { return ::expf(x); }
__device__ __forceinline__ double m_log(double x) { return p3m_log(x); }
__device__ __forceinline__ float m_log(float x) { return ::logf(x); }
__device__ __forceinline__ double m_log10(double x) { return p3m_log10(x); }
__device__ __forceinline__ float m_log10(float x) { return ::log10f(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return p3m_pow(x, y); }
__device__ __forceinline__ float m_pow(float x, float y) { return ::powf(x, y); }
__device__ __forceinline__ double m_cbrt(double x) { return p3m_cbrt(x); }
__device__ __forceinline__ float m_cbrt(float x) { return ::cbrtf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return ::sqrt(x); }  // IEEE correctly rounded
__device__ __forceinline__ float m_sqrt(float x) { return ::sqrtf(x); }
__device__ __forceinline__ double m_tgamma(double x) { return p3m_tgamma(x); }
__device__ __forceinline__ float m_tgamma(float x) { return ::tgammaf(x); }
__device__ __forceinline__ double m_expm1(double x) { return p3m_expm1(x); }
__device__ __forceinline__ float m_expm1(float x) { return ::expm1f(x); }
__device__ __forceinline__ double m_fabs(double x) { return ::fabs(x); }
__device__ __forceinline__ float m_fabs(float x) { return ::fabsf(x); }

// ekat::impl::min/max on scalars (ternary form, argument order preserved)
template <class T> __device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : b; }

// Error bits reported through the device error word.
enum : unsigned { ERR_BAD_TEMPERATURE = 1u, ERR_SUBSTEP_LIMIT = 2u, ERR_CHECK_VALUES = 4u };

// ---------------------------------------------------------------- saturation (FP64 always)
// MurphyKoop_svp, PSH:38-77 (with check_temperature, PSH:14-36, as a flag)
__host__ __device__ __forceinline__ double mk_svp(double t, bool ice, unsigned& err) {
  if (t <= 0 || t != t) err |= ERR_BAD_TEMPERATURE;  // t != t: NaN
  if (ice && t < Tmelt) {
    return p3m_exp(9.550426 - (5723.265 / t) + (3.53068 * p3m_log(t)) - (0.00728332 * t));
  }
  const double logt = p3m_log(t);
  return p3m_exp(54.842763 - (6763.22 / t) - (4.210 * logt) + (0.000367 * t) +
               (p3m_tanh(0.0415 * (t - 218.8)) *
                (53.878 - (1331.22 / t) - (9.44523 * logt) + 0.014025 * t)));
}
// qv_sat_dry, PSH:127-157
__device__ __forceinline__ double qv_sat_dry(double t, double p, bool ice, unsigned& err) {
  const double e = mk_svp(t, ice, err);
  return ep_2 * e / vmax(p, 1.e-3);
}

// ---------------------------------------------------------------- runtime options
struct Runtime {  // P3Runtime, components/eamxx/src/physics/p3/p3_functions.hpp:112-158
  double max_total_ni, autoconversion_prefactor, autoconversion_qc_exponent,
      autoconversion_nc_exponent, autoconversion_radius, accretion_prefactor,
      accretion_qc_exponent, accretion_qr_exponent, rain_selfcollection_prefactor,
      rain_selfcollection_breakup_diameter, constant_mu_rain, spa_ccn_to_nc_factor,
      cldliq_to_ice_collection_factor, rain_to_ice_collection_factor, min_rime_rho,
      max_rime_rho, immersion_freezing_exponent, deposition_nucleation_exponent,
      ice_sedimentation_factor;
  int do_ice_production;
};

// ---------------------------------------------------------------- tables
// Device copies of the lookup tables, per arithmetic type.  Besides the
// reference layouts (coll, revap) the ice and rain-velocity tables are
// repacked into records holding just the quantities one kernel interpolates,
// so that a table corner is one or two 16-byte loads instead of one load per
// quantity.  The values and the interpolation arithmetic are those of the
// reference tables; only the addresses differ.
template <class T> struct TableSet {
  const T* coll;     // [5][4][50][30][2]  (log10 values), reference layout
  const T* revap;    // [300][10]
  const T* ice_p2;   // [5][4][50][8]: ice quantities 1,2,3,4,6,7,9 and a pad (part 2)
  const T* ice_sed;  // [5][4][50][4]: quantities 0,1,6,7 (ice sedimentation)
  const T* ice_sed_d;  // [5][4][50][8]: the same 4 quantities, then each one's difference
                       // to the next size cell (0 in the last; ice sedimentation)
  const T* ice_p3;   // [5][4][50][4]: quantities 5,6,7,11 (part 3)
  const T* rain_v;   // [300][10][2]: vm, vn (rain sedimentation)
};
struct Tables {
  TableSet<double> d;  // the drop-in FP64 form
  TableSet<float> f;   // the mixed-precision form (half the gather bytes, no conversions)
};
constexpr int kIceP2Quant[7] = {1, 2, 3, 4, 6, 7, 9};
constexpr int kIceSedQuant[4] = {0, 1, 6, 7};
constexpr int kIceP3Quant[4] = {5, 6, 7, 11};

// The table set a precision form interpolates in.
template <class R> struct TabOf;
template <> struct TabOf<double> {
  static __device__ __forceinline__ const TableSet<double>& get(const Tables& t) { return t.d; }
};
template <> struct TabOf<float> {
  static __device__ __forceinline__ const TableSet<float>& get(const Tables& t) { return t.f; }
};

template <class R> struct TabIce { int dumi, dumii, dumjj; R dum1, dum4, dum5; };
template <class R> struct TabRain { int dumj; R dum3; };
template <class R> struct Tab3 { int dumii, dumjj; R rdumii, rdumjj; };

// lookup_ice, P3I/p3_table_ice_impl.hpp:80-138
template <class R>
__device__ __forceinline__ TabIce<R> lookup_ice(R qi, R ni, R qm, R rhop) {
  TabIce<R> t;
  t.dum1 = (m_log10(qi / ni) + R(18)) * R(lookup_table_1a_dum1_c) - R(10);
  t.dumi = (int)t.dum1;
  t.dum1 = vmin(t.dum1, R(isize));
  t.dum1 = vmax(t.dum1, R(1.));
  t.dumi = vmax(1, t.dumi);
  t.dumi = vmin(isize - 1, t.dumi);
  t.dum4 = (qm / qi) * R(3) + R(1);
  t.dumii = (int)t.dum4;
  t.dum4 = vmin(t.dum4, R(rimsize));
  t.dum4 = vmax(t.dum4, R(1.));
  t.dumii = vmax(1, t.dumii);
  t.dumii = vmin(rimsize - 1, t.dumii);
  if (rhop <= R(650)) t.dum5 = (rhop - R(50)) * R(0.005) + R(1);
  else t.dum5 = (rhop - R(650)) * R(0.004) + R(4);
  t.dumjj = (int)t.dum5;
  t.dum5 = vmin(t.dum5, R(densize));
  t.dum5 = vmax(t.dum5, R(1.));
  t.dumjj = vmax(1, t.dumjj);
  t.dumjj = vmin(densize - 1, t.dumjj);
  t.dumi -= 1;
  t.dumjj -= 1;
  t.dumii -= 1;
  return t;
}

// lookup_rain, P3I/p3_table_ice_impl.hpp:140-173
template <class R>
__device__ __forceinline__ TabRain<R> lookup_rain(R qr, R nr) {
  TabRain<R> t;
  const bool gt_small = qr > R(QSMALL) && nr > R(0.0);
  R dumlr = R(1);
  if (gt_small) dumlr = m_cbrt(qr / (R(Pi) * R(RHO_H2O) * nr));
  t.dum3 = (m_log10(R(1) * dumlr) + R(5)) * R(10.70415);
  t.dumj = (int)t.dum3;
  t.dum3 = vmin(t.dum3, R(rcollsize));
  t.dum3 = vmax(t.dum3, R(1));
  t.dumj = vmax(1, t.dumj);
  t.dumj = vmin(rcollsize - 1, t.dumj);
  if (!gt_small) { t.dumj = 1; t.dum3 = R(1); }
  t.dumj -= 1;
  return t;
}

__device__ __forceinline__ int ice_off(int jj, int ii, int i, int j) {
  return (((jj * rimsize + ii) * isize + i) * ice_table_size) + j;
}
__device__ __forceinline__ int coll_off(int jj, int ii, int i, int j, int k) {
  return ((((jj * rimsize + ii) * isize + i) * rcollsize + j) * collect_table_size) + k;
}

// Loads the first N values of a 16-byte-aligned record with 16-byte loads.
template <int N>
__device__ __forceinline__ void load_rec(const double* __restrict__ r, double (&v)[N]) {
  static_assert(N % 2 == 0, "records hold pairs");
#pragma unroll
  for (int q = 0; q < N; q += 2) {
    const double2 x = *reinterpret_cast<const double2*>(r + q);
    v[q] = x.x; v[q + 1] = x.y;
  }
}
template <int N>
__device__ __forceinline__ void load_rec(const float* __restrict__ r, float (&v)[N]) {
  static_assert(N % 4 == 0 || N == 2, "records hold quads or pairs");
  if constexpr (N == 2) {
    const float2 x = *reinterpret_cast<const float2*>(r);
    v[0] = x.x; v[1] = x.y;
  } else {
#pragma unroll
    for (int q = 0; q < N; q += 4) {
      const float4 x = *reinterpret_cast<const float4*>(r + q);
      v[q] = x.x; v[q + 1] = x.y; v[q + 2] = x.z; v[q + 3] = x.w;
    }
  }
}

// apply_table_ice, P3I/p3_table_ice_impl.hpp:175-217, for the NQ quantities of
// a repacked record of NR values per (density, rime, size) cell; each value is
// the reference's trilinear interpolation with its operation order.
template <int NQ, int NR, class R>
__device__ __forceinline__ void apply_table_ice_rec(const R* __restrict__ tab, const TabIce<R>& t, R (&out)[NQ],
                                                    int q0 = 0) {
  const int jj = t.dumjj, ii = t.dumii, i = t.dumi;
  const R w1 = (t.dum1 - R(i) - R(1));
  const R w4 = (t.dum4 - R(ii) - R(1));
  const R w5 = (t.dum5 - R(jj) - R(1));
  // values q0 .. q0+NQ-1 of the record of cell (a, b, i)
  auto rec = [&](int a, int b) { return tab + ((a * rimsize + b) * isize + i) * NR + q0; };
  R c[8][NQ];
  load_rec<NQ>(rec(jj, ii), c[0]);
  load_rec<NQ>(rec(jj, ii) + NR, c[1]);
  load_rec<NQ>(rec(jj, ii + 1), c[2]);
  load_rec<NQ>(rec(jj, ii + 1) + NR, c[3]);
  load_rec<NQ>(rec(jj + 1, ii), c[4]);
  load_rec<NQ>(rec(jj + 1, ii) + NR, c[5]);
  load_rec<NQ>(rec(jj + 1, ii + 1), c[6]);
  load_rec<NQ>(rec(jj + 1, ii + 1) + NR, c[7]);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    R iproc1 = c[0][q] + w1 * (c[1][q] - c[0][q]);
    R gproc1 = c[2][q] + w1 * (c[3][q] - c[2][q]);
    const R tmp1 = iproc1 + w4 * (gproc1 - iproc1);
    iproc1 = c[4][q] + w1 * (c[5][q] - c[4][q]);
    gproc1 = c[6][q] + w1 * (c[7][q] - c[6][q]);
    const R tmp2 = iproc1 + w4 * (gproc1 - iproc1);
    out[q] = tmp1 + w5 * (tmp2 - tmp1);
  }
}

// apply_table_ice_rec for records that also hold, DOFF values after each
// quantity, its difference to the same quantity of the next size cell,
// computed on the host in R.  c + w1 * d then has the bits of
// c + w1 * (c_next - c), so the four size-direction interpolations need no
// subtraction, and each corner pair is still two 16-byte reads (the value and
// its difference instead of the two cells).
template <int NQ, int NR, int DOFF, class R>
__device__ __forceinline__ void apply_table_ice_dif(const R* __restrict__ tab, const TabIce<R>& t, R (&out)[NQ],
                                                    int q0 = 0) {
  const int jj = t.dumjj, ii = t.dumii, i = t.dumi;
  const R w1 = (t.dum1 - R(i) - R(1));
  const R w4 = (t.dum4 - R(ii) - R(1));
  const R w5 = (t.dum5 - R(jj) - R(1));
  auto rec = [&](int a, int b) { return tab + ((a * rimsize + b) * isize + i) * NR + q0; };
  R c[4][NQ], d[4][NQ];
  load_rec<NQ>(rec(jj, ii), c[0]);
  load_rec<NQ>(rec(jj, ii) + DOFF, d[0]);
  load_rec<NQ>(rec(jj, ii + 1), c[1]);
  load_rec<NQ>(rec(jj, ii + 1) + DOFF, d[1]);
  load_rec<NQ>(rec(jj + 1, ii), c[2]);
  load_rec<NQ>(rec(jj + 1, ii) + DOFF, d[2]);
  load_rec<NQ>(rec(jj + 1, ii + 1), c[3]);
  load_rec<NQ>(rec(jj + 1, ii + 1) + DOFF, d[3]);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    R iproc1 = c[0][q] + w1 * d[0][q];
    R gproc1 = c[1][q] + w1 * d[1][q];
    const R tmp1 = iproc1 + w4 * (gproc1 - iproc1);
    iproc1 = c[2][q] + w1 * d[2][q];
    gproc1 = c[3][q] + w1 * d[3][q];
    const R tmp2 = iproc1 + w4 * (gproc1 - iproc1);
    out[q] = tmp1 + w5 * (tmp2 - tmp1);
  }
}

// apply_table_coll, P3I/p3_table_ice_impl.hpp:219-297, for both quantities of
// the collection table at once (they are adjacent in the reference layout).
template <class R>
__device__ __forceinline__ void apply_table_coll2(const R* __restrict__ tab, const TabIce<R>& ti,
                                                  const TabRain<R>& tr, R (&out)[2]) {
  const int jj = ti.dumjj, ii = ti.dumii, i = ti.dumi, j = tr.dumj;
  const R w1 = (ti.dum1 - R(i) - R(1));
  const R w3 = (tr.dum3 - R(j) - R(1));
  const R w4 = (ti.dum4 - R(ii) - R(1));
  const R w5 = (ti.dum5 - R(jj) - R(1));
  // one rime-density plane (8 corners) at a time: the same arithmetic as all 16
  // corners at once, with half the values live
  auto plane = [&](int a, R (&tmp)[2]) {
    R c[8][2];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      // corner bits: 4 = rime, 2 = size, 1 = rain size
      load_rec<2>(tab + coll_off(jj + a, ii + (n >> 2), i + ((n >> 1) & 1), j + (n & 1), 0), c[n]);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      auto T = [&](int b, int x, int y) { return c[b * 4 + x * 2 + y][q]; };
      R dproc1 = T(0, 0, 0) + w1 * (T(0, 1, 0) - T(0, 0, 0));
      R dproc2 = T(0, 0, 1) + w1 * (T(0, 1, 1) - T(0, 0, 1));
      const R iproc1 = dproc1 + w3 * (dproc2 - dproc1);
      dproc1 = T(1, 0, 0) + w1 * (T(1, 1, 0) - T(1, 0, 0));
      dproc2 = T(1, 0, 1) + w1 * (T(1, 1, 1) - T(1, 0, 1));
      const R gproc1 = dproc1 + w3 * (dproc2 - dproc1);
      tmp[q] = iproc1 + w4 * (gproc1 - iproc1);
    }
  };
  R tmp1[2], tmp2[2];
  plane(0, tmp1);
  plane(1, tmp2);
#pragma unroll
  for (int q = 0; q < 2; ++q) out[q] = tmp1[q] + w5 * (tmp2[q] - tmp1[q]);
}

// lookup (Table3), P3I/p3_table3_impl.hpp:14-62
template <class R>
__device__ __forceinline__ Tab3<R> lookup3(R mu_r, R lamr) {
  Tab3<R> t;
  const R dum1 = (mu_r + R(1)) / lamr;
  if (dum1 <= R(195.e-6)) {
    const R inv_dum3 = R(0.1);
    R rdumii = (dum1 * R(1.e6) + R(5)) * inv_dum3;
    rdumii = vmax(rdumii, R(1.0));
    rdumii = vmin(rdumii, R(20.0));
    int dumii = (int)rdumii;
    dumii = vmax(dumii, 1);
    dumii = vmin(dumii, 20);
    t.rdumii = rdumii;
    t.dumii = dumii;
  } else {
    const R inv_dum3 = R(THIRD) * R(0.1);
    R rdumii = (dum1 * R(1.e+6) - R(195)) * inv_dum3 + R(20);
    rdumii = vmax(rdumii, R(20.0));
    rdumii = vmin(rdumii, R(300.0));
    int dumii = (int)rdumii;
    dumii = vmax(dumii, 20);
    dumii = vmin(dumii, 299);
    t.rdumii = rdumii;
    t.dumii = dumii;
  }
  R rdumjj = mu_r + R(1);
  rdumjj = vmax(rdumjj, R(1));
  rdumjj = vmin(rdumjj, R(10));
  int dumjj = (int)rdumjj;
  dumjj = vmax(dumjj, 1);
  dumjj = vmin(dumjj, 9);
  t.rdumjj = rdumjj;
  t.dumjj = dumjj;
  return t;
}

// apply_table, P3I/p3_table3_impl.hpp:64-83 ([300][10] table)
template <class R>
__device__ __forceinline__ R apply_table3(const R* __restrict__ table, const Tab3<R>& t) {
  const R rdumii_m_dumii = t.rdumii - R(t.dumii);
  const R t_im1_jm1 = table[(t.dumii - 1) * VTABLE_DIM1 + (t.dumjj - 1)];
  const R dum1 = (t_im1_jm1 + rdumii_m_dumii * (table[t.dumii * VTABLE_DIM1 + (t.dumjj - 1)] - t_im1_jm1));
  const R t_im1_j = table[(t.dumii - 1) * VTABLE_DIM1 + t.dumjj];
  const R dum2 = (t_im1_j + rdumii_m_dumii * (table[t.dumii * VTABLE_DIM1 + t.dumjj] - t_im1_j));
  return dum1 + (t.rdumjj - R(t.dumjj)) * (dum2 - dum1);
}

// apply_table for the mass- and number-weighted rain fall speeds together,
// from the [300][10][2] (vm, vn) record table: out = {vm, vn}.
template <class R>
__device__ __forceinline__ void apply_table3_v(const R* __restrict__ table, const Tab3<R>& t, R (&out)[2]) {
  const R rdumii_m_dumii = t.rdumii - R(t.dumii);
  R a[2], b[2], c[2], d[2];
  load_rec<2>(table + ((t.dumii - 1) * VTABLE_DIM1 + (t.dumjj - 1)) * 2, a);
  load_rec<2>(table + (t.dumii * VTABLE_DIM1 + (t.dumjj - 1)) * 2, b);
  load_rec<2>(table + ((t.dumii - 1) * VTABLE_DIM1 + t.dumjj) * 2, c);
  load_rec<2>(table + (t.dumii * VTABLE_DIM1 + t.dumjj) * 2, d);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const R dum1 = (a[q] + rdumii_m_dumii * (b[q] - a[q]));
    const R dum2 = (c[q] + rdumii_m_dumii * (d[q] - c[q]));
    out[q] = dum1 + (t.rdumjj - R(t.dumjj)) * (dum2 - dum1);
  }
}

// ---------------------------------------------------------------- size distributions
// get_cloud_dsd2, P3I/p3_dsd2_impl.hpp:14-71 (iparam = 3, nu stays 0)
template <class R>
__device__ __forceinline__ void get_cloud_dsd2(R qc, R& nc, R& mu_c, R rho, R& lamc, R& cdist, R& cdist1) {
  lamc = R(0); cdist = R(0); cdist1 = R(0); mu_c = R(0);
  if (qc >= R(QSMALL)) {
    nc = vmax(nc, R(NSMALL));
    R m = R(0.0005714) * (nc * R(1.e-6) * rho) + R(0.2714);
    m = R(1) / (m * m) - R(1);
    m = vmax(m, R(2));
    m = vmin(m, R(15));
    mu_c = m;
    lamc = m_cbrt(R(CONS1) * nc * (m + R(3)) * (m + R(2)) * (m + R(1)) / qc);
    const R lammin = (m + R(1)) * R(2.5e+4);
    const R lammax = (m + R(1)) * R(1.e+6);
    const bool lt = lamc < lammin, gt = lamc > lammax;
    if (lt) lamc = lammin;
    if (gt) lamc = lammax;
    if (lt || gt)
      nc = R(6) * (lamc * lamc * lamc) * qc / (R(Pi) * R(RHO_H2O) * (m + R(3)) * (m + R(2)) * (m + R(1)));
    cdist = nc * (m + R(1)) / lamc;
    cdist1 = nc / m_tgamma(m + R(1));
  }
}

// get_rain_dsd2, P3I/p3_dsd2_impl.hpp:73-126
template <class R>
__device__ __forceinline__ void get_rain_dsd2(const Runtime& rt, R qr, R& nr, R& mu_r, R& lamr) {
  lamr = R(0);
  if (qr >= R(QSMALL)) {
    const R nr_lim = vmax(nr, R(NSMALL));
    mu_r = R(rt.constant_mu_rain);
    const R f = R(CONS1) * (mu_r + R(3)) * (mu_r + R(2)) * (mu_r + R(1));
    lamr = m_cbrt(f * nr_lim / qr);
    const R lammax = (mu_r + R(1.)) * R(1.e+5);
    const R lammin = (mu_r + R(1.)) * R(500);
    const bool lt = lamr < lammin, gt = lamr > lammax;
    nr = nr_lim;
    if (lt || gt) {
      if (lt) lamr = lammin;
      if (gt) lamr = lammax;
      nr = lamr * lamr * lamr * qr / f;
    }
  }
}

// calc_bulk_rho_rime, P3I/p3_ice_sed_impl.hpp:14-64
template <class R>
__device__ __forceinline__ R calc_bulk_rho_rime(const Runtime& rt, R qi_tot, R& qi_rim, R& bi_rim) {
  R rho_rime = R(0);
  const bool gt = bi_rim >= R(BSMALL), lt = bi_rim < R(BSMALL);
  if (gt) rho_rime = qi_rim / bi_rim;
  const bool lt_min = rho_rime < R(rt.min_rime_rho), gt_max = rho_rime > R(rt.max_rime_rho);
  if (gt && lt_min) rho_rime = R(rt.min_rime_rho);
  if (gt && gt_max) rho_rime = R(rt.max_rime_rho);
  const bool clamped = gt && (gt_max || lt_min);
  if (lt) { qi_rim = R(0); bi_rim = R(0); rho_rime = R(0); }
  // The reference sets bi_rim = qi_rim / rho_rime for a clamped density, then
  // again with qi_rim = qi_tot when the rime exceeds the total; the second
  // overrides the first, and `lt` excludes both, so one division serves.
  const bool over = qi_rim > qi_tot && rho_rime > R(0);
  if (over || clamped) bi_rim = (over ? qi_tot : qi_rim) / rho_rime;
  if (over) qi_rim = qi_tot;
  if (qi_rim < R(QSMALL)) { qi_rim = R(0); bi_rim = R(0); }
  return rho_rime;
}

// calculate_incloud_mixingratios, P3I/p3_incloud_mixingratios_impl.hpp:9-67
struct Incld { double qc, qr, qi, qm, nc, nr, ni, bm; };
__device__ __forceinline__ Incld incloud(double qc, double qr, double qi, double qm, double nc, double nr,
                                         double ni, double bm, double icfl, double icfi, double icfr) {
  Incld o;
  if (qc >= QSMALL) { o.qc = qc * icfl; o.nc = vmax(nc * icfl, 0.0); } else { o.qc = 0; o.nc = 0; }
  if (qi >= QSMALL) { o.qi = qi * icfi; o.ni = vmax(ni * icfi, 0.0); } else { o.qi = 0; o.ni = 0; }
  if (qi >= QSMALL && qm >= QSMALL) { o.qm = qm * icfi; o.bm = vmax(bm * icfl, 0.0); } else { o.qm = 0; o.bm = 0; }
  if (qr >= QSMALL) { o.qr = qr * icfr; o.nr = vmax(nr * icfr, 0.0); } else { o.qr = 0; o.nr = 0; }
  if (o.qc > incloud_limit || o.qi > incloud_limit || o.qr > precip_limit || o.bm > incloud_limit) {
    o.qc = vmin(o.qc, incloud_limit);
    o.qi = vmin(o.qi, incloud_limit);
    o.bm = vmin(o.bm, incloud_limit);
    o.qr = vmin(o.qr, precip_limit);
  }
  return o;
}

// impose_max_total_ni, P3I/p3_impose_max_total_ni_impl.hpp:9-26
template <class R>
__device__ __forceinline__ void impose_max_total_ni(R& ni, R max_total_ni, R inv_rho) {
  if (ni >= R(1e-20)) ni = ni * vmin(max_total_ni * inv_rho / ni, R(1));
}

}  // namespace p3d
