// p3_tables.cpp -- host-side P3 lookup tables (P3LookupTables contents).
//
// * ice / ice-rain collection tables: the text reader of
//   components/eamxx/src/physics/p3/impl/p3_table_ice_impl.hpp:16-78
//   (same header check, same columns kept, log10 of the collection values);
// * rain fall-speed / ventilation tables: the quadrature the reference runs at
//   init in Fortran, components/eam/src/physics/p3/scream/micro_p3.F90:218-341
//   (p3_init_b), whose results init_kokkos_tables copies
//   (p3_table3_impl.hpp:85-148);
// * dnu: the constants of p3_table3_impl.hpp:120-135.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/p3x.h"
#include "p3_constants.hpp"

namespace {

using namespace p3c;

int read_ice_table(const char* path, p3x_table_data* out) {
  std::ifstream in(path);
  if (!in) return P3X_ERR_TABLE_FILE;
  std::string version, version_val;
  in >> version >> version_val;
  if (version != "VERSION") return P3X_ERR_TABLE_VERSION;
  if (version_val != "4.1.1") return P3X_ERR_TABLE_VERSION;
  double dum_s;
  int dum_i;
  for (int jj = 0; jj < densize; ++jj) {
    for (int ii = 0; ii < rimsize; ++ii) {
      for (int i = 0; i < isize; ++i) {
        in >> dum_i >> dum_i;
        int j_idx = 0;
        for (int j = 0; j < 15; ++j) {
          in >> dum_s;
          if (j > 1 && j != 10)
            out->ice_table[((jj * rimsize + ii) * isize + i) * ice_table_size + j_idx++] = dum_s;
        }
      }
      for (int i = 0; i < isize; ++i) {
        for (int j = 0; j < rcollsize; ++j) {
          in >> dum_i >> dum_i;
          int k_idx = 0;
          for (int k = 0; k < 6; ++k) {
            in >> dum_s;
            if (k == 3 || k == 4)
              out->collect_table[(((jj * rimsize + ii) * isize + i) * rcollsize + j) * collect_table_size +
                                 k_idx++] = std::log10(dum_s);
          }
        }
      }
    }
  }
  if (!in) return P3X_ERR_TABLE_FILE;
  return P3X_OK;
}

// micro_p3.F90:218-341.  mu_r is the constant mu_r_constant = 1
// (micro_p3_utils.F90:33) for every one of the 10 mu columns, so the columns
// are identical and one is computed and replicated.
void build_rain_tables(p3x_table_data* out) {
  const double mu_r_constant = 1.0;
  const double pi = Pi;
  const double thrd = 1.0 / 3.0, sxth = 1.0 / 6.0;
  const double piov6 = pi * sxth;
  for (int i = 0; i < MU_R_TABLE_DIM; ++i) out->mu_r_table[i] = mu_r_constant;
  const double mu_r = mu_r_constant;
  for (int jj = 1; jj <= VTABLE_DIM0; ++jj) {
    double dm;
    if (jj <= 20) dm = ((double)jj * 10.0 - 5.0) * 1.e-6;
    else dm = ((double)(jj - 20) * 30.0 + 195.0) * 1.e-6;
    const double lamr = (mu_r + 1.0) / dm;
    double dum1 = 0, dum2 = 0, dum3 = 0, dum4 = 0, dum5 = 0;
    const double dd = 2.0;
    for (int kk = 1; kk <= 10000; ++kk) {
      const double dia = ((double)kk * dd - dd * 0.5) * 1.e-6;
      double amg = piov6 * 997.0 * (dia * dia * dia);
      amg = amg * 1000.0;
      double vt;
      if (dia * 1.e+6 <= 134.43) vt = 4.5795e+3 * std::pow(amg, 2.0 * thrd);
      else if (dia * 1.e+6 < 1511.64) vt = 4.962e+1 * std::pow(amg, thrd);
      else if (dia * 1.e+6 < 3477.84) vt = 1.732e+1 * std::pow(amg, sxth);
      else vt = 9.17;
      const double ld = std::log10(dia);
      const double e = std::exp(-lamr * dia);
      dum1 = dum1 + vt * std::pow(10.0, mu_r * ld + 4.0 * mu_r) * e * dd * 1.e-6;
      dum2 = dum2 + std::pow(10.0, mu_r * ld + 4.0 * mu_r) * e * dd * 1.e-6;
      dum3 = dum3 + vt * std::pow(10.0, (mu_r + 3.0) * ld + 4.0 * mu_r) * e * dd * 1.e-6;
      dum4 = dum4 + std::pow(10.0, (mu_r + 3.0) * ld + 4.0 * mu_r) * e * dd * 1.e-6;
      dum5 = dum5 + std::pow(vt * dia, 0.5) * std::pow(10.0, (mu_r + 1.0) * ld + 3.0 * mu_r) * e * dd * 1.e-6;
    }
    dum2 = std::max(dum2, 1.e-30);
    dum4 = std::max(dum4, 1.e-30);
    dum5 = std::max(dum5, 1.e-30);
    const double vn = dum1 / dum2;
    const double vm = dum3 / dum4;
    const double revap = std::pow(10.0, std::log10(dum5) + (mu_r + 1.0) * std::log10(lamr) - (3.0 * mu_r));
    for (int ii = 0; ii < VTABLE_DIM1; ++ii) {
      out->vn_table[(jj - 1) * VTABLE_DIM1 + ii] = vn;
      out->vm_table[(jj - 1) * VTABLE_DIM1 + ii] = vm;
      out->revap_table[(jj - 1) * VTABLE_DIM1 + ii] = revap;
    }
  }
}

}  // namespace

extern "C" int p3x_tables_load(const char* ice_table_file, p3x_table_data* out) {
  if (!ice_table_file || !out) return P3X_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  const int st = read_ice_table(ice_table_file, out);
  if (st != P3X_OK) return st;
  build_rain_tables(out);
  static const double dnu[16] = {0.000,  -0.557, -0.430, -0.307, -0.186, -0.067, -0.050, -0.167,
                                 -0.282, -0.397, -0.512, -0.626, -0.739, -0.853, -0.966, -0.966};
  std::memcpy(out->dnu, dnu, sizeof(dnu));
  return P3X_OK;
}
