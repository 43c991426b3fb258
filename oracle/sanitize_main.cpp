// sanitize_main.cpp -- runs the CPU code of this repository under AddressSanitizer
// and UndefinedBehaviorSanitizer (test infrastructure; built by `make -C oracle
// sanitize`, run by tests/test_sanitizers.py).
//
// Covered: the oracle's p3_main (p3o_main) and the P3Microphysics pre/post-amble,
// CldFraction and column-budget restatements (p3_oracle.c), the ice-table
// generator restatement (p3_table_gen.c) on a few cells, and the product's host
// table reader and rain-table quadrature (scream_amd/csrc/p3_tables.cpp,
// p3x_tables_load).  The GPU code cannot run here; its host side is the C ABI
// in scream_amd/csrc, whose only CPU-side computation is that table loader.
//
// usage: p3sanitize <table file> <input file> <output file>
//   input:  int32 ncol, nlev, predictNc, prescribedCCN; float64 dt; then the 23
//           p3_main input fields and the 3 history fields, each nlev*ncol
//           doubles, level-major, in the order of p3o_fields
//   output: the 10 state fields, then qv2qi_depos_tend, the 3 effective radii,
//           rho_qi, precip_total_tend, nevapr, the 3 history fields (nlev*ncol
//           each), the 2 flux fields ((nlev+1)*ncol each), the 2 surface fields
//           (ncol each), then int32 col_status[ncol] and sed_substeps[3*ncol].
// Prints one line per part and returns 0; a sanitizer finding aborts the process.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/p3x.h"
#include "p3_oracle.h"

namespace {

std::vector<double> read_doubles(FILE* f, size_t n) {
  std::vector<double> v(n);
  if (fread(v.data(), sizeof(double), n, f) != n) {
    fprintf(stderr, "p3sanitize: short input\n");
    exit(2);
  }
  return v;
}

void write_doubles(FILE* f, const std::vector<double>& v) { fwrite(v.data(), sizeof(double), v.size(), f); }

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s <table file> <input file> <output file>\n", argv[0]);
    return 2;
  }
  // host table reader + rain-table quadrature of the product library
  std::vector<p3x_table_data> tabv(1);  // 650 KB: on the heap
  p3x_table_data& tab = tabv[0];
  int st = p3x_tables_load(argv[1], &tab);
  if (st != 0) {
    fprintf(stderr, "p3x_tables_load: %d\n", st);
    return 3;
  }
  double csum = 0;
  for (double x : tab.ice_table) csum += x;
  for (double x : tab.collect_table) csum += x;
  for (double x : tab.vm_table) csum += x;
  printf("tables ok (checksum %.17g)\n", csum);

  FILE* in = fopen(argv[2], "rb");
  if (!in) return 2;
  int hdr[4];
  double dt;
  if (fread(hdr, sizeof(int), 4, in) != 4 || fread(&dt, sizeof(double), 1, in) != 1) return 2;
  const int ncol = hdr[0], nlev = hdr[1];
  const size_t n = (size_t)ncol * nlev, nf = (size_t)ncol * (nlev + 1);
  std::vector<std::vector<double>> fin(26);
  for (auto& v : fin) v = read_doubles(in, n);
  fclose(in);

  // p3_main on exactly sized buffers, so that any out-of-bounds access is caught
  std::vector<std::vector<double>> outs(9, std::vector<double>(n));
  std::vector<double> fl(nf), fi(nf), sl(ncol), si(ncol);
  p3o_fields f;
  double** state[] = {&f.qc, &f.nc, &f.qr, &f.nr, &f.qi, &f.qm, &f.ni, &f.bm, &f.qv, &f.th};
  for (int j = 0; j < 10; ++j) *state[j] = fin[j].data();
  const double** din[] = {&f.nc_nuceat_tend, &f.nccn, &f.ni_activated, &f.inv_qc_relvar, &f.cld_frac_i,
                          &f.cld_frac_l, &f.cld_frac_r, &f.pres, &f.dz, &f.dpres, &f.inv_exner,
                          &f.qv_prev, &f.t_prev};
  for (int j = 0; j < 13; ++j) *din[j] = fin[10 + j].data();
  double** dout[] = {&f.qv2qi_depos_tend, &f.diag_eff_radius_qc, &f.diag_eff_radius_qi,
                     &f.diag_eff_radius_qr, &f.rho_qi, &f.precip_total_tend, &f.nevapr};
  for (int j = 0; j < 7; ++j) *dout[j] = outs[j].data();
  f.liq_ice_exchange = fin[23].data();
  f.vap_liq_exchange = fin[24].data();
  f.vap_ice_exchange = fin[25].data();
  f.precip_liq_flux = fl.data();
  f.precip_ice_flux = fi.data();
  f.precip_liq_surf = sl.data();
  f.precip_ice_surf = si.data();
  std::vector<int> status(ncol), nsub(3 * (size_t)ncol), ice_idx(n);
  p3o_debug dbg{status.data(), nsub.data(), ice_idx.data()};
  p3o_runtime rt;
  p3o_runtime_defaults(&rt);
  p3o_tables ot{tab.ice_table, tab.collect_table, tab.vn_table, tab.vm_table, tab.revap_table,
                tab.mu_r_table, tab.dnu};
  const int nerr = p3o_main(&rt, &ot, &f, &dbg, ncol, nlev, dt, hdr[2], hdr[3], 2);
  printf("p3o_main ok (%d columns with a temperature error)\n", nerr);

  FILE* out = fopen(argv[3], "wb");
  if (!out) return 2;
  for (int j = 0; j < 10; ++j) write_doubles(out, fin[j]);
  for (int j = 0; j < 7; ++j) write_doubles(out, outs[j]);
  for (int j = 23; j < 26; ++j) write_doubles(out, fin[j]);
  write_doubles(out, fl);
  write_doubles(out, fi);
  write_doubles(out, sl);
  write_doubles(out, si);
  fwrite(status.data(), sizeof(int), status.size(), out);
  fwrite(nsub.data(), sizeof(int), nsub.size(), out);
  fclose(out);

  // pre-amble and post-amble on field-manager-like fields derived from the inputs
  {
    std::vector<std::vector<double>> a(34, std::vector<double>(n));
    for (size_t x = 0; x < n; ++x) {
      a[0][x] = fin[17][x];                 // p_mid = pres
      a[1][x] = fin[19][x] * 1.01;          // pseudo_density (wet)
      a[2][x] = fin[19][x];                 // pseudo_density_dry
      a[3][x] = fin[15][x];                 // cldfrac_tot
      a[4][x] = fin[9][x] / fin[20][x];     // T_mid = th / inv_exner
      for (int j = 0; j < 9; ++j) a[5 + j][x] = fin[j][x];  // qv .. bm (order below)
      a[14][x] = fin[21][x];                // qv_prev
      a[15][x] = fin[22][x];                // T_prev
    }
    std::vector<double> surf(6 * (size_t)ncol, 0.0);
    p3o_atm at;
    at.p_mid = a[0].data(); at.pseudo_density = a[1].data(); at.pseudo_density_dry = a[2].data();
    at.cldfrac_tot = a[3].data(); at.T_mid = a[4].data();
    at.qv = a[13].data(); at.qc = a[5].data(); at.nc = a[6].data(); at.qr = a[7].data(); at.nr = a[8].data();
    at.qi = a[9].data(); at.qm = a[10].data(); at.ni = a[11].data(); at.bm = a[12].data();
    at.qv_prev = a[14].data(); at.T_prev = a[15].data();
    at.th_atm = a[16].data(); at.inv_exner = a[17].data(); at.dz = a[18].data();
    at.cld_frac_l = a[19].data(); at.cld_frac_i = a[20].data(); at.cld_frac_r = a[21].data();
    at.eff_radius_qc = a[22].data(); at.eff_radius_qi = a[23].data(); at.eff_radius_qr = a[24].data();
    at.precip_liq_surf = sl.data(); at.precip_ice_surf = si.data();
    at.precip_liq_surf_mass = surf.data(); at.precip_ice_surf_mass = surf.data() + ncol;
    at.vapor_flux = surf.data() + 2 * (size_t)ncol; at.water_flux = surf.data() + 3 * (size_t)ncol;
    at.ice_flux = surf.data() + 4 * (size_t)ncol; at.heat_flux = surf.data() + 5 * (size_t)ncol;
    p3o_preamble(&at, ncol, nlev);
    p3o_postamble(&at, ncol, nlev, dt);
    std::vector<double> mass(ncol), energy(ncol), mass1(ncol), energy1(ncol), me(ncol), ee(ncol);
    std::vector<double> zero(n, 0.0), ps(ncol, 1.0e5), phis(ncol, 0.0);
    p3o_column_budget(a[1].data(), a[4].data(), zero.data(), zero.data(), at.qv, at.qc, at.qi, at.qr,
                      ps.data(), phis.data(), ncol, nlev, mass.data(), energy.data());
    p3o_column_budget(a[1].data(), a[4].data(), zero.data(), zero.data(), at.qv, at.qc, at.qi, at.qr,
                      ps.data(), phis.data(), ncol, nlev, mass1.data(), energy1.data());
    p3o_conservation_errors(mass.data(), energy.data(), mass1.data(), energy1.data(), at.vapor_flux,
                            at.water_flux, at.ice_flux, at.heat_flux, ncol, dt, me.data(), ee.data());
    std::vector<double> ci(n), ct(n), ci4(n), ct4(n);
    p3o_cld_fraction(at.qi, a[19].data(), ncol, nlev, 1e-12, 1e-5, ci.data(), ct.data(), ci4.data(), ct4.data());
    printf("pre/post-amble, column budget, cld fraction ok\n");
  }

  // the ice-table generator restatement on two cells
  {
    std::vector<unsigned char> mask(1000, 0);
    mask[0] = 1;
    mask[777] = 1;
    std::vector<double> ice(1000 * 15), coll(1000 * 30 * 6);
    p3o_table_gen(mask.data(), ice.data(), coll.data(), 2);
    printf("table generator ok (cell 777 first value %.6g)\n", ice[777 * 15]);
  }
  return 0;
}
