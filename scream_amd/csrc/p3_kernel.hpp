// p3_kernel.hpp -- the fused P3 step kernel (p3_main over a block of columns).
//
// Work decomposition (one launch = one p3_main call, or one column set of a
// batch):
//   * a workgroup owns CB consecutive columns; its NT = 64*waves threads split
//     every column's levels into NCH = NT/CB contiguous chunks;
//   * thread t handles column c = t % CB (so CB consecutive lanes read CB
//     consecutive doubles of a level-major row: one or more full cache lines)
//     and chunk t / CB;
//   * pointwise phases (part 1, part 2, homogeneous freezing, part 3) run on the
//     thread's own levels; per-column reductions (the early-exit flags,
//     find_top/find_bottom, the CFL maximum) and the upwind flux crossing a
//     chunk boundary go through LDS.
// Between phases a level is only ever touched by its owning thread, so the
// level-major scratch in HBM needs no synchronisation beyond program order.
//
// The step follows components/eamxx/src/physics/p3/impl/p3_main_impl.hpp:74-331.
#pragma once

#include "p3_part2.hpp"

namespace p3d {

enum Scr { SC_T = 0, SC_QVSL, SC_QVSI, SC_SUPI, SC_RHOFACR, SC_RHOFACI, SC_ACN, SC_V1, SC_V2, SC_COUNT };

struct KParams {
  // P3PrognosticState (in/out)
  double *qc, *nc, *qr, *nr, *qi, *qm, *ni, *bm, *qv, *th;
  // P3DiagnosticInputs
  const double *nc_nuceat_tend, *nccn, *ni_activated, *inv_qc_relvar, *cld_frac_i, *cld_frac_l,
      *cld_frac_r, *pres, *dz, *dpres, *inv_exner, *qv_prev, *t_prev;
  // P3DiagnosticOutputs
  double *qv2qi_depos_tend, *precip_liq_surf, *precip_ice_surf, *diag_eff_radius_qc,
      *diag_eff_radius_qi, *diag_eff_radius_qr, *rho_qi, *precip_liq_flux, *precip_ice_flux,
      *precip_total_tend, *nevapr;
  // P3HistoryOnly (in/out unless exch_zero)
  double *liq_ice_exchange, *vap_liq_exchange, *vap_ice_exchange;
  double* scratch;      // SC_COUNT x nlev x ld
  unsigned* err;        // device error word
  int *dbg_status, *dbg_substeps, *dbg_ice_idx;
  Tables tab;
  Runtime rt;
  int ncol, nlev, ld;   // ld: row stride (elements) of every level-major array
  double dt;
  int predictNc, prescribedCCN, exch_zero;
};

constexpr int kMaxSubsteps = 100000;

template <class R, int CB, int NT>
__global__ __launch_bounds__(NT) void p3_main_kernel(const KParams p) {
  constexpr int NCH = NT / CB;
  static_assert(NT % CB == 0, "block must hold whole chunk rows");
  __shared__ int s_i[NT];
  __shared__ double s_d[NT];
  __shared__ double s_f[4 * NT];

  const int tid = threadIdx.x;
  const int c = tid % CB;
  const int ch = tid / CB;
  const int i = blockIdx.x * CB + c;
  const bool col_ok = i < p.ncol;
  const int nlev = p.nlev;
  const int L = (nlev + NCH - 1) / NCH;
  const int k0 = ch * L < nlev ? ch * L : nlev;
  const int k1 = k0 + L < nlev ? k0 + L : nlev;
  const int kbot = nlev - 1;
  const bool owns_kbot = (k0 <= kbot) && (kbot < k1);
  const size_t ld = (size_t)p.ld;
  const double dt = p.dt, inv_dt = 1 / dt;
  const double latvap = LatVap, latice = LatIce, inv_cp = INV_CP;
  const Runtime& rt = p.rt;
  unsigned err = 0;

  auto scr = [&](int f, int k) -> double& { return p.scratch[((size_t)f * nlev + k) * ld + i]; };
  auto at = [&](int k) { return (size_t)k * ld + i; };

  auto red_or = [&](int v) {
    s_i[tid] = v;
    __syncthreads();
    int r = 0;
    for (int j = 0; j < NCH; ++j) r |= s_i[j * CB + c];
    __syncthreads();
    return r;
  };
  auto red_min = [&](int v) {
    s_i[tid] = v;
    __syncthreads();
    int r = v;
    for (int j = 0; j < NCH; ++j) r = min(r, s_i[j * CB + c]);
    __syncthreads();
    return r;
  };
  auto red_max = [&](int v) {
    s_i[tid] = v;
    __syncthreads();
    int r = v;
    for (int j = 0; j < NCH; ++j) r = max(r, s_i[j * CB + c]);
    __syncthreads();
    return r;
  };
  auto red_maxd = [&](double v) {
    s_d[tid] = v;
    __syncthreads();
    double r = v;
    for (int j = 0; j < NCH; ++j) { const double o = s_d[j * CB + c]; if (o > r) r = o; }
    __syncthreads();
    return r;
  };

  // ================= p3_main_init + p3_main_part1 (P3I/p3_main_impl_part1.hpp:18-185)
  int flags = 0;  // bit0 nucleationPossible, bit1 hydrometeorsPresent
  if (col_ok) {
    for (int k = k0; k < k1; ++k) {
      const size_t x = at(k);
      const double inv_exner = p.inv_exner[x];
      const double exner = 1 / inv_exner;
      double th = p.th[x];
      double qv = vmax(p.qv[x], 0.0);
      double qc = p.qc[x], nc = p.nc[x], qr = p.qr[x], nr = p.nr[x];
      double qi = p.qi[x], ni = p.ni[x], qm = p.qm[x], bm = p.bm[x];
      const double pres = p.pres[x];
      double T_atm = th * exner;
      const double rho = p.dpres[x] / p.dz[x] / gravit;
      const double inv_rho = 1 / rho;
      const double qv_sat_l = qv_sat_dry(T_atm, pres, false, err);
      const double qv_sat_i = qv_sat_dry(T_atm, pres, true, err);
      const double qv_supersat_i = qv / qv_sat_i - 1;
      double rhofacr, rhofaci, acn;
      {
        const R ir = R(inv_rho), Tr = R(T_atm);
        rhofacr = double(m_pow(R(RHO_1000MB) * ir, R(.54)));
        rhofaci = double(m_pow(R(RHO_600MB) * ir, R(.54)));
        const R dum = R(1.496e-6) * m_pow(Tr, R(1.5)) / (Tr + R(120));
        acn = double(R(gravit) * R(RHO_H2O) / (R(18) * dum));
      }
      if (T_atm < T_zerodegc && qv_supersat_i >= -0.05) flags |= 1;

      if (qc < QSMALL) {
        qv = qv + qc;
        th = th - inv_exner * qc * latvap * inv_cp;
        qc = 0; nc = 0;
      } else {
        flags |= 2;
        if (p.prescribedCCN) nc = vmax(nc, rt.spa_ccn_to_nc_factor * p.nccn[x] / (1 / p.cld_frac_l[x]));
        else if (p.predictNc) nc = vmax(nc + p.nc_nuceat_tend[x] * dt, 0.0);
        else nc = NCCNST * inv_rho;
      }
      if (qr < QSMALL) {
        qv = qv + qr;
        th = th - inv_exner * qr * latvap * inv_cp;
        qr = 0; nr = 0;
      } else {
        flags |= 2;
      }
      if (qi < QSMALL || (qi < 1.e-8 && qv_supersat_i < -0.1)) {
        qv = qv + qi;
        th = th - inv_exner * qi * (latvap + latice) * inv_cp;
        qi = 0; ni = 0; qm = 0; bm = 0;
      } else {
        flags |= 2;
      }
      if (qi >= QSMALL && qi < 1.e-8 && T_atm >= T_zerodegc) {
        qr = qr + qi;
        th = th - inv_exner * qi * latice * inv_cp;
        qi = 0; ni = 0; qm = 0; bm = 0;
      }
      T_atm = th * exner;

      p.qc[x] = qc; p.nc[x] = nc; p.qr[x] = qr; p.nr[x] = nr; p.qi[x] = qi;
      p.ni[x] = ni; p.qm[x] = qm; p.bm[x] = bm; p.qv[x] = qv; p.th[x] = th;
      scr(SC_T, k) = T_atm;
      scr(SC_QVSL, k) = qv_sat_l;
      scr(SC_QVSI, k) = qv_sat_i;
      scr(SC_SUPI, k) = qv_supersat_i;
      scr(SC_RHOFACR, k) = rhofacr;
      scr(SC_RHOFACI, k) = rhofaci;
      scr(SC_ACN, k) = acn;
    }
  }
  const int cflags = red_or(flags);
  // status: 0 skipped after part 1, 1 skipped after part 2, 2 full step
  int status = (col_ok && (cflags & 3)) ? 1 : 0;

  // ================= p3_main_part2 (P3I/p3_main_impl_part2.hpp:113-524)
  int hydro = 0;
  if (col_ok) {
    for (int k = k0; k < k1; ++k) {
      const size_t x = at(k);
      p.precip_liq_flux[x] = 0;
      bool done = false;
      if (status == 1) {
        Lvl s;
        s.qc = p.qc[x]; s.qr = p.qr[x]; s.qi = p.qi[x];
        s.T_atm = scr(SC_T, k);
        s.qv_supersat_i = scr(SC_SUPI, k);
        const bool skip_all = s.qc < QSMALL && s.qr < QSMALL && s.qi < QSMALL && s.T_atm < T_zerodegc &&
                              s.qv_supersat_i < -0.05;
        if (!skip_all) {
          s.nc = p.nc[x]; s.nr = p.nr[x]; s.qm = p.qm[x]; s.ni = p.ni[x]; s.bm = p.bm[x];
          s.qv = p.qv[x]; s.th = p.th[x];
          s.pres = p.pres[x];
          s.rho = p.dpres[x] / p.dz[x] / gravit;
          s.inv_rho = 1 / s.rho;
          s.qv_sat_l = scr(SC_QVSL, k);
          s.qv_sat_i = scr(SC_QVSI, k);
          s.rhofaci = scr(SC_RHOFACI, k);
          s.acn = scr(SC_ACN, k);
          s.cfl = p.cld_frac_l[x]; s.cfr = p.cld_frac_r[x]; s.cfi = p.cld_frac_i[x];
          s.icfl = 1 / s.cfl; s.icfr = 1 / s.cfr; s.icfi = 1 / s.cfi;
          s.inv_exner = p.inv_exner[x];
          s.qv_prev = p.qv_prev[x]; s.t_prev = p.t_prev[x];
          s.ni_activated = p.ni_activated[x];
          if (part2_level<R>(rt, p.tab, dt, p.predictNc, p.prescribedCCN, s, err)) hydro = 1;
          p.qc[x] = s.qc; p.nc[x] = s.nc; p.qr[x] = s.qr; p.nr[x] = s.nr; p.qi[x] = s.qi;
          p.ni[x] = s.ni; p.qm[x] = s.qm; p.bm[x] = s.bm; p.qv[x] = s.qv; p.th[x] = s.th;
          p.qv2qi_depos_tend[x] = s.qv2qi_depos_tend;
          p.precip_total_tend[x] = s.precip_total_tend;
          p.nevapr[x] = s.nevapr;
          p.vap_ice_exchange[x] = s.vap_ice_exchange;
          p.vap_liq_exchange[x] = s.vap_liq_exchange;
          p.liq_ice_exchange[x] = s.liq_ice_exchange;
          done = true;
        }
      }
      if (!done) {
        p.qv2qi_depos_tend[x] = 0;
        p.precip_total_tend[x] = 0;
        p.nevapr[x] = 0;
        if (p.exch_zero) {
          p.vap_ice_exchange[x] = 0;
          p.vap_liq_exchange[x] = 0;
          p.liq_ice_exchange[x] = 0;
        }
      }
    }
    if (owns_kbot) p.precip_liq_flux[at(nlev)] = 0;
  }
  if (red_or(hydro) && status == 1) status = 2;
  const bool active = status == 2;

  // ================= sedimentation (cloud, rain, ice), adaptive substepping
  double liq_prt_cloud = 0, liq_prt_rain = 0, ice_prt = 0;
  bool cloud_present = false, rain_present = false, ice_present = false;
  int nsub_species[3] = {0, 0, 0};

#pragma unroll 1
  for (int sp = 0; sp < 3; ++sp) {
    // find_top / find_bottom, P3I/p3_find_impl.hpp:14-92
    double* qx = sp == 0 ? p.qc : (sp == 1 ? p.qr : p.qi);
    int my_top = nlev, my_bot = -1;
    if (active) {
      for (int k = k0; k < k1; ++k) {
        if (qx[at(k)] >= QSMALL) {
          if (k < my_top) my_top = k;
          my_bot = k;
        }
      }
    }
    const int k_top = red_min(my_top);
    int k_bot = red_max(my_bot);
    const bool present = active && k_top < nlev;
    if (sp == 0) cloud_present = present;
    if (sp == 1) rain_present = present;
    if (sp == 2) ice_present = present;
    double dt_left = dt, prt_accum = 0;
    bool first = true;
    int nsub = 0;
    const int nfield = sp == 2 ? 4 : (sp == 1 ? 2 : (p.predictNc ? 2 : 1));

#pragma unroll 1
    while (true) {
      // The substep count is bounded by the CFL number (a few hundred at most for
      // physical inputs); the cap only guarantees termination for corrupt inputs.
      if (nsub >= kMaxSubsteps && dt_left > dt_left_tol) { err |= ERR_SUBSTEP_LIMIT; dt_left = 0; }
      const bool run = present && dt_left > dt_left_tol;
      if (!__syncthreads_or(run)) break;
      // ---- velocities and CFL number on the thread's part of [k_top, k_bot].
      // The upwind step below also reads the level under the window, whose
      // velocity the reference zeroes with the whole column (P3I/p3_ice_sed_impl.hpp:120-125);
      // it is written as 0 here so no velocity of an earlier species survives there.
      double co = 0;
      if (run) {
        const int k_under = (k_bot == kbot) ? k_bot : k_bot + 1;
        const int ka = max(k0, k_top), kb = min(k1 - 1, k_under);
        for (int k = ka; k <= kb; ++k) {
          const size_t x = at(k);
          const double rho = p.dpres[x] / p.dz[x] / gravit;
          double V1 = 0, V2 = 0;
          if (k > k_bot) {
            // below the window: no velocity
          } else if (sp == 0) {
            // cloud_sedimentation, P3I/p3_cloud_sed_impl.hpp:85-112
            const double cfl = p.cld_frac_l[x];
            const double qc = p.qc[x], nc = p.nc[x];
            double qc_incld, nc_incld;
            if (first) {
              const double icfl = 1 / cfl;
              if (qc >= QSMALL) { qc_incld = vmin(qc * icfl, incloud_limit); nc_incld = vmax(nc * icfl, 0.0); }
              else { qc_incld = 0; nc_incld = 0; }
            } else {
              qc_incld = qc / cfl;
              nc_incld = nc / cfl;
            }
            if (qc_incld > QSMALL) {
              R ncr = R(nc_incld), mu_c, lamc, cdist, cdist1;
              get_cloud_dsd2<R>(R(qc_incld), ncr, mu_c, R(rho), lamc, cdist, cdist1);
              p.nc[x] = double(ncr) * cfl;
              const R dum = R(1) / m_pow(lamc, R(bcn));
              const R acn = R(scr(SC_ACN, k));
              V1 = double(acn * m_tgamma(R(4 + bcn) + mu_c) * dum / m_tgamma(mu_c + R(4)));
              if (p.predictNc) V2 = double(acn * m_tgamma(R(1 + bcn) + mu_c) * dum / m_tgamma(mu_c + R(1)));
            }
          } else if (sp == 1) {
            // rain_sedimentation + compute_rain_fall_velocity, P3I/p3_rain_sed_impl.hpp:14-34,104-126
            const double cfr = p.cld_frac_r[x];
            const double qr = p.qr[x], nr = p.nr[x];
            double qr_incld, nr_incld;
            if (first) {
              const double icfr = 1 / cfr;
              if (qr >= QSMALL) { qr_incld = vmin(qr * icfr, precip_limit); nr_incld = vmax(nr * icfr, 0.0); }
              else { qr_incld = 0; nr_incld = 0; }
            } else {
              qr_incld = qr / cfr;
              nr_incld = nr / cfr;
            }
            if (qr_incld > QSMALL) {
              R nrr = R(nr_incld), mu_r = 0, lamr;
              get_rain_dsd2<R>(rt, R(qr_incld), nrr, mu_r, lamr);
              const Tab3<R> t3 = lookup3<R>(mu_r, lamr);
              const R rhofacr = R(scr(SC_RHOFACR, k));
              V1 = double(apply_table3<R>(p.tab.vm, t3) * rhofacr);
              V2 = double(apply_table3<R>(p.tab.vn, t3) * rhofacr);
              p.nr[x] = double(nrr) * cfr;
            }
          } else {
            // ice_sedimentation, P3I/p3_ice_sed_impl.hpp:134-166
            const double cfi = p.cld_frac_i[x];
            const double qi = p.qi[x], ni = p.ni[x], qm = p.qm[x], bm = p.bm[x];
            double qi_incld, ni_incld, qm_incld, bm_incld;
            if (first) {
              const double icfi = 1 / cfi, icfl = 1 / p.cld_frac_l[x];
              if (qi >= QSMALL) { qi_incld = vmin(qi * icfi, incloud_limit); ni_incld = vmax(ni * icfi, 0.0); }
              else { qi_incld = 0; ni_incld = 0; }
              if (qi >= QSMALL && qm >= QSMALL) { qm_incld = qm * icfi; bm_incld = vmin(vmax(bm * icfl, 0.0), incloud_limit); }
              else { qm_incld = 0; bm_incld = 0; }
            } else {
              qi_incld = qi / cfi;
              ni_incld = ni / cfi;
              qm_incld = qm / cfi;
              bm_incld = bm / cfi;
            }
            if (qi_incld > QSMALL) {
              R nir = vmax(R(ni_incld), R(NSMALL));
              R qmr = R(qm_incld), bmr = R(bm_incld);
              const R qir = R(qi_incld);
              const R rhop = calc_bulk_rho_rime<R>(rt, qir, qmr, bmr);
              p.qm[x] = double(qmr) * cfi;
              p.bm[x] = double(bmr) * cfi;
              const TabIce<R> ti = lookup_ice<R>(qir, nir, qmr, rhop);
              const R t_ni_fall = apply_table_ice<R>(p.tab.ice, 0, ti);
              const R t_qi_fall = apply_table_ice<R>(p.tab.ice, 1, ti);
              const R t_lammax = apply_table_ice<R>(p.tab.ice, 6, ti);
              const R t_lammin = apply_table_ice<R>(p.tab.ice, 7, ti);
              nir = vmin(nir, t_lammax * nir);
              nir = vmax(nir, t_lammin * nir);
              p.ni[x] = double(nir) * cfi;
              const R rhofaci = R(scr(SC_RHOFACI, k));
              V1 = double(R(rt.ice_sedimentation_factor) * t_qi_fall * rhofaci);
              V2 = double(R(rt.ice_sedimentation_factor) * t_ni_fall * rhofaci);
            }
          }
          scr(SC_V1, k) = V1;
          scr(SC_V2, k) = V2;
          const double Co = V1 * dt_left * (1 / p.dz[x]);
          if (Co > co) co = Co;
        }
      }
      const double Co_max = red_maxd(co);

      // ---- generalized_sedimentation + calc_first_order_upwind_step, P3I/p3_upwind_impl.hpp:14-129
      int k_temp = k_bot;
      double dt_sub = 0;
      double* rf[4];
      int vsel[4];
      if (sp == 0) { rf[0] = p.qc; rf[1] = p.nc; vsel[0] = SC_V1; vsel[1] = SC_V2; }
      else if (sp == 1) { rf[0] = p.qr; rf[1] = p.nr; vsel[0] = SC_V1; vsel[1] = SC_V2; }
      else {
        rf[0] = p.qi; rf[1] = p.ni; rf[2] = p.qm; rf[3] = p.bm;
        vsel[0] = SC_V1; vsel[1] = SC_V2; vsel[2] = SC_V1; vsel[3] = SC_V1;
      }
      if (run) {
        const int tmpint1 = (int)(Co_max + 1);
        dt_sub = dt_left / tmpint1;
        k_temp = (k_bot == kbot) ? k_bot : k_bot + 1;
        // flux leaving the last level of this chunk, for the chunk below
        const int kl = k1 - 1;
        if (kl >= k_top && kl <= k_temp && k0 < k1) {
          const size_t x = at(kl);
          const double rho = p.dpres[x] / p.dz[x] / gravit;
          for (int f = 0; f < nfield; ++f) s_f[f * NT + tid] = scr(vsel[f], kl) * rf[f][x] * rho;
        }
      }
      __syncthreads();
      if (run) {
        const int ka = max(k0, k_top), kb = min(k1 - 1, k_temp);
        double Fa[4] = {0, 0, 0, 0};
        if (ka > k_top && ka <= kb)
          for (int f = 0; f < nfield; ++f) Fa[f] = s_f[f * NT + tid - CB];
        for (int k = ka; k <= kb; ++k) {
          const size_t x = at(k);
          const double rho = p.dpres[x] / p.dz[x] / gravit;
          const double inv_rho = 1 / rho;
          const double inv_dz = 1 / p.dz[x];
          double F0 = 0;
          for (int f = 0; f < nfield; ++f) {
            const double r = rf[f][x];
            const double F = scr(vsel[f], k) * r * rho;
            const double fluxdiv = (Fa[f] - F) * inv_dz;
            rf[f][x] = r + fluxdiv * dt_sub * inv_rho;
            Fa[f] = F;
            if (f == 0) F0 = F;
          }
          if (sp == 1) p.precip_liq_flux[at(k + 1)] += F0;  // rain_sed :140-153 (k <= new k_qxbot)
          if (k == kbot && k_bot == kbot) prt_accum += F0 * dt_sub;
        }
        if (k_bot != kbot) k_bot += 1;
        dt_left -= dt_sub;
        nsub++;
        first = false;
      }
    }
    if (sp == 0) liq_prt_cloud = prt_accum;
    if (sp == 1) liq_prt_rain = prt_accum;
    if (sp == 2) ice_prt = prt_accum;
    nsub_species[sp] = nsub;
  }

  // ================= homogeneous_freezing (P3I/p3_ice_sed_impl.hpp:207-270) + p3_main_part3
  if (col_ok) {
    for (int k = k0; k < k1; ++k) {
      const size_t x = at(k);
      p.precip_ice_flux[x] = 0;
      double reff_c = 10.e-6, reff_i = 25.e-6, reff_r = 500.e-6, rho_qi = 0;
      if (active) {
        double qc = p.qc[x], nc = p.nc[x], qr = p.qr[x], nr = p.nr[x], qi = p.qi[x], ni = p.ni[x];
        double qm = p.qm[x], bm = p.bm[x], qv = p.qv[x], th = p.th[x];
        const double inv_exner = p.inv_exner[x];
        if (rt.do_ice_production) {
          const bool t_lt = scr(SC_T, k) < T_homogfrz;
          const bool qc_ge = t_lt && qc >= QSMALL;
          const bool qr_ge = t_lt && qr >= QSMALL;
          const double Qc_nuc = qc, Qr_nuc = qr;
          const double Nc_nuc = vmax(nc, NSMALL), Nr_nuc = vmax(nr, NSMALL);
          if (qc_ge) {
            qm = qm + Qc_nuc; qi = qi + Qc_nuc; bm = bm + Qc_nuc * INV_RHO_RIMEMAX; ni = ni + Nc_nuc;
            th = th + inv_exner * Qc_nuc * latice * inv_cp;
          }
          if (qr_ge) {
            qm = qm + Qr_nuc; qi = qi + Qr_nuc; bm = bm + Qr_nuc * INV_RHO_RIMEMAX; ni = ni + Nr_nuc;
            th = th + inv_exner * Qr_nuc * latice * inv_cp;
          }
          if (qc_ge) { qc = 0; nc = 0; }
          if (qr_ge) { qr = 0; nr = 0; }
        }
        // p3_main_part3, P3I/p3_main_impl_part3.hpp:60-199 (only the fields p3_main outputs)
        const double rho = p.dpres[x] / p.dz[x] / gravit;
        const double inv_rho = 1 / rho;
        const double cfl = p.cld_frac_l[x], cfr = p.cld_frac_r[x], cfi = p.cld_frac_i[x];
        double vap_liq = p.vap_liq_exchange[x];
        // cloud
        if (qc >= QSMALL) {
          R ncr = R(nc / cfl), mu_c, lamc, cd, cd1;
          get_cloud_dsd2<R>(R(qc / cfl), ncr, mu_c, R(rho), lamc, cd, cd1);
          nc = double(ncr) * cfl;
          reff_c = double(R(0.5) * (mu_c + R(3)) / lamc);
        } else {
          qv = qv + qc;
          th = th - inv_exner * qc * latvap * inv_cp;
          vap_liq = vap_liq - qc;
          qc = 0; nc = 0;
        }
        // rain
        if (qr >= QSMALL) {
          R nrr = R(nr / cfr), mu_r = 0, lamr;
          get_rain_dsd2<R>(rt, R(qr / cfr), nrr, mu_r, lamr);
          nr = double(nrr) * cfr;
          reff_r = double(R(1.5) / lamr);
        } else {
          qv = qv + qr;
          th = th - inv_exner * qr * latvap * inv_cp;
          vap_liq = vap_liq - qr;
          qr = 0; nr = 0;
        }
        // ice
        {
          const bool qi_gt = qi >= QSMALL;
          ni = vmax(ni, NSMALL);
          const double qi_incld = qi / cfi;
          double ni_incld = ni / cfi;
          R qmr = R(qm / cfi), bmr = R(bm / cfi);
          double qm_incld = qm / cfi, bm_incld = bm / cfi;
          R rhop = 0;
          if (qi_gt) {
            rhop = calc_bulk_rho_rime<R>(rt, R(qi_incld), qmr, bmr);
            qm_incld = double(qmr); bm_incld = double(bmr);
            qm = qm_incld * cfi;
            bm = bm_incld * cfi;
          }
          impose_max_total_ni<double>(ni_incld, rt.max_total_ni, inv_rho);
          if (qi_gt) {
            const TabIce<R> ti = lookup_ice<R>(R(qi_incld), R(ni_incld), R(qm_incld), rhop);
            if (p.dbg_ice_idx) p.dbg_ice_idx[at(k)] = ti.dumjj * 10000 + ti.dumii * 100 + ti.dumi;
            const R t_eff = apply_table_ice<R>(p.tab.ice, 5, ti);
            const R t_lammax = apply_table_ice<R>(p.tab.ice, 6, ti);
            const R t_lammin = apply_table_ice<R>(p.tab.ice, 7, ti);
            const R t_rhomm = apply_table_ice<R>(p.tab.ice, 11, ti);
            ni_incld = vmin(ni_incld, double(t_lammax) * ni_incld);
            ni_incld = vmax(ni_incld, double(t_lammin) * ni_incld);
            reff_i = double(t_eff);
            rho_qi = double(t_rhomm);
          } else if (p.dbg_ice_idx) {
            p.dbg_ice_idx[at(k)] = -1;
          }
          ni = ni_incld * cfi;
          if (qm < QSMALL && qi_gt) { qm = 0; bm = 0; }
          if (!qi_gt) {
            qv = qv + qi;
            th = th - inv_exner * qi * (latvap + latice) * inv_cp;
            qi = 0; ni = 0; qm = 0; bm = 0;
          }
        }
        if (qr < QSMALL) nr = 0;
        p.qc[x] = qc; p.nc[x] = nc; p.qr[x] = qr; p.nr[x] = nr; p.qi[x] = qi;
        p.ni[x] = ni; p.qm[x] = qm; p.bm[x] = bm; p.qv[x] = qv; p.th[x] = th;
        p.vap_liq_exchange[x] = vap_liq;
      } else if (p.dbg_ice_idx) {
        p.dbg_ice_idx[at(k)] = -1;
      }
      p.diag_eff_radius_qc[x] = reff_c;
      p.diag_eff_radius_qi[x] = reff_i;
      p.diag_eff_radius_qr[x] = reff_r;
      p.rho_qi[x] = rho_qi;
    }
    if (owns_kbot) {
      p.precip_ice_flux[at(nlev)] = 0;
      double liq = 0, ice = 0;
      if (cloud_present) liq = liq_prt_cloud * INV_RHO_H2O * inv_dt;
      if (rain_present) liq += liq_prt_rain * INV_RHO_H2O * inv_dt;
      if (ice_present) ice += ice_prt * INV_RHO_H2O * inv_dt;
      p.precip_liq_surf[i] = liq;
      p.precip_ice_surf[i] = ice;
    }
    if (ch == 0) {
      if (p.dbg_status) p.dbg_status[i] = status;
      if (p.dbg_substeps)
        for (int s = 0; s < 3; ++s) p.dbg_substeps[(size_t)s * p.ncol + i] = nsub_species[s];
    }
  }
  if (err) atomicOr(p.err, err);
}

}  // namespace p3d
