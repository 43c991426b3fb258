// p3_constants.hpp -- physical constants and table dimensions used by P3.
//
// Values and defining expressions follow
//   components/eamxx/src/physics/share/physics_constants.hpp:27-113
//   components/eamxx/src/physics/p3/p3_functions.hpp:30-52
// The expressions are kept (not pre-evaluated) so every constant rounds the
// way the reference's constexpr does.
#pragma once

namespace p3c {

constexpr double Cpair = 1004.64;
constexpr double Rair = 287.042;
constexpr double RH2O = 461.505;
constexpr double RV = RH2O;
constexpr double RHO_H2O = 1000.0;
constexpr double INV_RHO_H2O = 1.0 / RHO_H2O;
constexpr double MWH2O = 18.016;
constexpr double MWdry = 28.966;
constexpr double ep_2 = MWH2O / MWdry;
constexpr double gravit = 9.80616;
constexpr double LatVap = 2501000.0;
constexpr double LatIce = 333700.0;
constexpr double CpLiq = 4188.0;
constexpr double Tmelt = 273.15;
constexpr double T_zerodegc = Tmelt;
constexpr double T_homogfrz = Tmelt - 40;
constexpr double T_rainfrz = Tmelt - 4;
constexpr double Pi = 3.14159265358979323;
constexpr double RHOW = RHO_H2O;
constexpr double RHO_RIMEMAX = 900.0;
constexpr double INV_RHO_RIMEMAX = 1.0 / RHO_RIMEMAX;
constexpr double THIRD = 1.0 / 3.0;
constexpr double SXTH = 1.0 / 6.0;
constexpr double PIOV3 = Pi * THIRD;
constexpr double PIOV6 = Pi * SXTH;
constexpr double BIMM = 2.0;
constexpr double CONS1 = PIOV6 * RHOW;
constexpr double CONS2 = 4. * PIOV3 * RHOW;
constexpr double CONS3 = 1.0 / (CONS2 * 1.562500000000000e-14);
constexpr double CONS5 = PIOV6 * BIMM;
constexpr double CONS6 = PIOV6 * PIOV6 * RHOW * BIMM;
constexpr double QSMALL = 1.e-14;
constexpr double QTENDSMALL = 1e-20;
constexpr double BSMALL = 1.e-15;
constexpr double NSMALL = 1.e-16;
constexpr double P0 = 100000.0;
constexpr double RD = Rair;
constexpr double RHO_1000MB = P0 / (RD * Tmelt);
constexpr double RHO_600MB = 60000 / (RD * 253.15);
constexpr double CP = Cpair;
constexpr double INV_CP = 1.0 / CP;
constexpr double dt_left_tol = 1.e-4;
constexpr double bcn = 2.;
constexpr double dropmass = 5.2e-7;
constexpr double NCCNST = 200.0e+6;
constexpr double incloud_limit = 5.1e-3;
constexpr double precip_limit = 1.0e-2;
constexpr double f1r = 0.78;
constexpr double f2r = 0.32;
constexpr double nmltratio = 1.0;

// ice lookup table, p3_functions.hpp:33-40
constexpr int densize = 5;
constexpr int rimsize = 4;
constexpr int isize = 50;
constexpr int ice_table_size = 12;
constexpr int rcollsize = 30;
constexpr int collect_table_size = 2;
constexpr int dnusize = 16;
constexpr double lookup_table_1a_dum1_c = 4.135985029041767e+00;
// rain tables, physics_constants.hpp:92-95
constexpr int VTABLE_DIM0 = 300;
constexpr int VTABLE_DIM1 = 10;
constexpr int MU_R_TABLE_DIM = 150;

constexpr int ICE_TABLE_LEN = densize * rimsize * isize * ice_table_size;                     // 12000
constexpr int COLL_TABLE_LEN = densize * rimsize * isize * rcollsize * collect_table_size;    // 60000
constexpr int RAIN_TABLE_LEN = VTABLE_DIM0 * VTABLE_DIM1;                                     // 3000

}  // namespace p3c
