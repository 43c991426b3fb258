// p3_part2.hpp -- per-level process-rate chain of p3_main_part2 as a device function.
//
// Follows components/eamxx/src/physics/p3/impl/p3_main_impl_part2.hpp:113-524
// for one (column, level) point that is not skip_all.  R is the arithmetic of
// the bulk process rates (see p3_physics.hpp); cell averaging, the
// conservation limiters and the prognostic updates are FP64 in both forms.
#pragma once

#include "p3_physics.hpp"

namespace p3d {

// Per-level values p3_main_part2 reads and updates.
struct Lvl {
  // prognostic state (in/out)
  double qc, nc, qr, nr, qi, qm, ni, bm, qv, th;
  // part-1 quantities and inputs
  double T_atm, pres, rho, inv_rho, qv_sat_l, qv_sat_i, qv_supersat_i, rhofaci, acn;
  double cfl, cfr, cfi, icfl, icfr, icfi, inv_exner, qv_prev, t_prev, ni_activated;
  // outputs
  double qv2qi_depos_tend, precip_total_tend, nevapr;
  double vap_ice_exchange, vap_liq_exchange, liq_ice_exchange;
};

// Returns true when a hydrometeor survives the end-of-part-2 clipping at this
// level (the per-level contribution to hydrometeorsPresent, :465-508).
template <class R>
__device__ bool part2_level(const Runtime& rt, const Tables& tb, double dt, bool predictNc,
                            bool prescribedCCN, double e_melt, Lvl& s, unsigned& err) {
  const double inv_dt = 1 / dt;
  const double latvap = LatVap, latice = LatIce, inv_cp = INV_CP;
  const double T_atm = s.T_atm, rho = s.rho, inv_rho = s.inv_rho, pres = s.pres;
  const double cld_frac_l = s.cfl, cld_frac_r = s.cfr, cld_frac_i = s.cfi;

  // in-cloud values on entry: calculate_incloud_mixingratios of the current state
  const Incld in0 = incloud(s.qc, s.qr, s.qi, s.qm, s.nc, s.nr, s.ni, s.bm, s.icfl, s.icfi, s.icfr);
  R qc_incld = R(in0.qc), qr_incld = R(in0.qr), qi_incld = R(in0.qi), qm_incld = R(in0.qm);
  R nc_incld = R(in0.nc), nr_incld = R(in0.nr), ni_incld = R(in0.ni), bm_incld = R(in0.bm);

  R qc2qr_accret_tend = 0, qc2qr_autoconv_tend = 0, nc_accret_tend = 0,
    nc_selfcollect_tend = 0, nc2nr_autoconv_tend = 0, nr_selfcollect_tend = 0, ncautr = 0;
  R qi2qv_sublim_tend = 0, nr_ice_shed_tend = 0, qc2qi_hetero_freeze_tend = 0, qr2qi_collect_tend = 0,
    qc2qr_ice_shed_tend = 0, qi2qr_melt_tend = 0, qc2qi_collect_tend = 0, qr2qi_immers_freeze_tend = 0,
    qv2qi_nucleat_tend = 0, ni2nr_melt_tend = 0, nc_collect_tend = 0, ncshdc = 0,
    nc2ni_immers_freeze_tend = 0, nr_collect_tend = 0, ni_selfcollect_tend = 0, ni_nucleat_tend = 0,
    qv2qi_vapdep_tend = 0, qc2qi_berg_tend = 0, nr2ni_immers_freeze_tend = 0, ni_sublim_tend = 0;
  double rho_qm_cloud = 400;
  double evap_d = 0, nr_evap_d = 0;  // rain evaporation rates (finished in FP64)
  bool wetgrowth = false;
  R mu_c = 0, lamc = 0, cdist = 0, cdist1 = 0, mu_r = 0, lamr = 0, cdistr = 0, logn0r = 0;

  const bool skip_micro = !(qc_incld >= R(QSMALL) || qr_incld >= R(QSMALL) || qi_incld >= R(QSMALL));
  if (!skip_micro) {
    const R Tr = R(T_atm), rhor = R(rho), rhofaci = R(s.rhofaci);
    const R dT_melt = R(T_atm - Tmelt);  // FP64 difference (exact in the FP64 form)
    // get_time_space_phys_variables, P3I/p3_get_time_space_phys_variables_impl.hpp:10-46
    const R mu = R(1.496e-6) * m_pow(Tr, R(1.5)) / (Tr + R(120));
    const R dv = R(8.794e-5) * m_pow(Tr, R(1.81)) / R(pres);
    const R sc = mu / (rhor * dv);
    const R dum_t = R(1) / (R(RV) * (Tr * Tr));
    const R dqsdt = R(latvap) * R(s.qv_sat_l) * dum_t;
    const R dqsidt = R(latvap + latice) * R(s.qv_sat_i) * dum_t;
    const R ab = R(1) + dqsdt * R(latvap) * R(INV_CP);
    const R abi = R(1) + dqsidt * R(latvap + latice) * R(INV_CP);
    const R kap = R(1.414e+3) * mu;
    R eii;
    if (T_atm < 253.15) eii = R(0.001);
    else if (T_atm < 273.15) eii = R(0.001) + R(T_atm - 253.15) * (R(0.3) - R(0.001)) / R(20);
    else eii = R(0.3);

    get_cloud_dsd2<R>(qc_incld, nc_incld, mu_c, rhor, lamc, cdist, cdist1);
    s.nc = double(nc_incld) * cld_frac_l;

    get_rain_dsd2<R>(rt, qr_incld, nr_incld, mu_r, lamr);
    // get_cdistr_logn0r, P3I/p3_dsd2_impl.hpp:128-148
    if (qr_incld >= R(QSMALL)) {
      cdistr = nr_incld / m_tgamma(mu_r + R(1));
      logn0r = m_log10(cdistr) + (mu_r + R(1)) * m_log10(lamr);
    }
    s.nr = double(nr_incld) * cld_frac_r;

    impose_max_total_ni<R>(ni_incld, R(rt.max_total_ni), R(inv_rho));

    R tv_qi_fallspd = 0, tv_ni_self_collect = 0, tv_qc2qi_collect = 0, tv_qi2qr_melting = 0,
      tv_nr_collect = 0, tv_qr2qi_collect = 0, tv_lammax = 0, tv_lammin = 0, tv_vent_melt = 0;
    if (qi_incld >= R(QSMALL)) {
      ni_incld = vmax(ni_incld, R(NSMALL));
      nr_incld = vmax(nr_incld, R(NSMALL));
      const R rhop = calc_bulk_rho_rime<R>(rt, qi_incld, qm_incld, bm_incld);
      s.qm = double(qm_incld) * cld_frac_i;
      s.bm = double(bm_incld) * cld_frac_i;
      const TabIce<R> ti = lookup_ice<R>(qi_incld, ni_incld, qm_incld, rhop);
      const TabRain<R> tr = lookup_rain<R>(qr_incld, nr_incld);
      // ice quantities 1,2,3,4 then 6,7,9 of the part 2 records (in FP64 two at a
      // time: 8 corners of 2 values live at once instead of 4, fewer spills)
      R v[4];
      if constexpr (sizeof(R) == 8) {
        R a[2];
        apply_table_ice_rec<2, 8, R>(TabOf<R>::get(tb).ice_p2, ti, a, 0);
        v[0] = a[0]; v[1] = a[1];
        apply_table_ice_rec<2, 8, R>(TabOf<R>::get(tb).ice_p2, ti, a, 2);
        v[2] = a[0]; v[3] = a[1];
      } else {
        apply_table_ice_rec<4, 8, R>(TabOf<R>::get(tb).ice_p2, ti, v, 0);
      }
      tv_qi_fallspd = v[0];
      tv_ni_self_collect = v[1];
      tv_qc2qi_collect = v[2];
      tv_qi2qr_melting = v[3];
      if constexpr (sizeof(R) == 8) {
        R a[2];
        apply_table_ice_rec<2, 8, R>(TabOf<R>::get(tb).ice_p2, ti, a, 4);
        v[0] = a[0]; v[1] = a[1];
        apply_table_ice_rec<2, 8, R>(TabOf<R>::get(tb).ice_p2, ti, a, 6);
        v[2] = a[0];
      } else {
        apply_table_ice_rec<4, 8, R>(TabOf<R>::get(tb).ice_p2, ti, v, 4);
      }
      tv_lammax = v[0];
      tv_lammin = v[1];
      tv_vent_melt = v[2];
      if (qr_incld >= R(QSMALL)) {
        R c[2];
        apply_table_coll2<R>(TabOf<R>::get(tb).coll, ti, tr, c);
        tv_nr_collect = c[0];
        tv_qr2qi_collect = c[1];
      }
      ni_incld = vmin(ni_incld, tv_lammax * ni_incld);
      ni_incld = vmax(ni_incld, tv_lammin * ni_incld);
    }
    // ventilation factor shared by melting, wet growth and the relaxation timescale
    // (every use below requires qi_incld >= QSMALL)
    const R vent = (qi_incld >= R(QSMALL))
                       ? tv_qi2qr_melting + tv_vent_melt * m_cbrt(sc) * m_sqrt(rhofaci * rhor / mu)
                       : R(0);

    if (rt.do_ice_production) {
      // ice_cldliq_collection, P3I/p3_ice_collection_impl.hpp:9-62
      {
        const bool t_is_negative = T_atm <= Tmelt;
        const bool both = qi_incld >= R(QSMALL) && qc_incld >= R(QSMALL);
        const R f = R(rt.cldliq_to_ice_collection_factor);
        if (both && t_is_negative)
          qc2qi_collect_tend = rhofaci * tv_qc2qi_collect * qc_incld * f * rhor * ni_incld;
        if (both) nc_collect_tend = rhofaci * tv_qc2qi_collect * nc_incld * f * rhor * ni_incld;
        if (both && !t_is_negative) {
          qc2qr_ice_shed_tend = rhofaci * tv_qc2qi_collect * qc_incld * f * rhor * ni_incld;
          nc_collect_tend = rhofaci * tv_qc2qi_collect * nc_incld * f * rhor * ni_incld;
          ncshdc = qc2qr_ice_shed_tend * R(1.0 / dropmass);
        }
      }
      // ice_rain_collection, P3I/p3_ice_collection_impl.hpp:64-106
      {
        const bool t_is_negative = T_atm <= Tmelt;
        const bool both = qi_incld >= R(QSMALL) && qr_incld >= R(QSMALL);
        const R f = R(rt.rain_to_ice_collection_factor);
        if (both && t_is_negative) {
          qr2qi_collect_tend = m_pow(R(10.0), tv_qr2qi_collect + logn0r) * rhor * rhofaci * f * ni_incld;
          nr_collect_tend = m_pow(R(10.0), tv_nr_collect + logn0r) * rhor * rhofaci * f * ni_incld;
        }
        if (both && !t_is_negative)
          nr_collect_tend = m_pow(R(10.0), tv_nr_collect + logn0r) * rhor * rhofaci * f * ni_incld;
      }
      // ice_self_collection, P3I/p3_ice_collection_impl.hpp:108-151
      if (qi_incld >= R(QSMALL)) {
        const bool qm_pos = qm_incld > R(0.0);
        R tmp1 = R(0.0), Eii_fact = R(0.0);
        if (qm_pos) tmp1 = qm_incld / qi_incld;
        if (tmp1 < R(0.6) && qm_pos) Eii_fact = R(1);
        if (tmp1 >= R(0.6) && tmp1 < R(0.9) && qm_pos) Eii_fact = R(1) - (tmp1 - R(0.6)) / R(0.3);
        if (tmp1 >= R(0.9) && qm_pos) Eii_fact = R(0);
        if (!qm_pos) Eii_fact = R(1);
        ni_selfcollect_tend = tv_ni_self_collect * rhor * eii * Eii_fact * rhofaci * ni_incld * ni_incld;
      }
    }

    // ice_melting, P3I/p3_ice_melting_impl.hpp:11-52
    if (qi_incld >= R(QSMALL) && T_atm > Tmelt) {
      const double qsat0 = ep_2 * e_melt / vmax(pres, 1.e-3);  // qv_sat_dry(Tmelt, pres, liquid)
      const R dqs = R(qsat0 - s.qv);
      qi2qr_melt_tend = (vent * (dT_melt * kap - rhor * R(latvap) * dv * dqs) * R(2) * R(Pi) / R(latice)) * ni_incld;
      qi2qr_melt_tend = vmax(qi2qr_melt_tend, R(0));
      ni2nr_melt_tend = qi2qr_melt_tend * (ni_incld / qi_incld);
    }

    if (rt.do_ice_production) {
      // ice_cldliq_wet_growth, P3I/p3_ice_cldliq_wet_growth_impl.hpp:11-79
      const bool t_is_negative = T_atm < Tmelt;
      const bool any_if = qi_incld >= R(QSMALL) && (qc_incld + qr_incld) >= R(1.0e-6) && t_is_negative;
      const bool any_if_col = any_if && (qc2qi_collect_tend + qr2qi_collect_tend) >= R(1.0e-10);
      if (any_if) {
        const double qsat0 = ep_2 * e_melt / vmax(pres, 1.e-3);  // qv_sat_dry(Tmelt, pres, liquid)
        const R dqs = R(qsat0 - s.qv);
        R qc_growth_rate = (vent * R(Pi * 2) * (rhor * R(latvap) * dv * dqs - dT_melt * kap) /
                            (R(latice) + R(CpLiq) * dT_melt)) * ni_incld;
        qc_growth_rate = vmax(qc_growth_rate, R(0.0));
        const R dum = vmax(R(0.0), (qc2qi_collect_tend + qr2qi_collect_tend) - qc_growth_rate);
        const bool dum_ge_small = dum >= R(1.0e-10);
        if (dum_ge_small) {
          nr_ice_shed_tend = nr_ice_shed_tend + dum * R(1.923e+6);
          const R dum1 = R(1.0) / (qc2qi_collect_tend + qr2qi_collect_tend);
          if (any_if_col) {
            qc2qr_ice_shed_tend = qc2qr_ice_shed_tend + dum * qc2qi_collect_tend * dum1;
            qc2qi_collect_tend = vmax(R(0), qc2qi_collect_tend - dum * qc2qi_collect_tend * dum1);
            qr2qi_collect_tend = vmax(R(0), qr2qi_collect_tend - dum * qr2qi_collect_tend * dum1);
          }
        }
        wetgrowth = dum_ge_small;
      }
    }

    // ice_relaxation_timescale, P3I/p3_ice_relaxation_timescale_impl.hpp:9-40
    R epsi = R(0), epsi_tot = R(0);
    if (qi_incld >= R(QSMALL) && T_atm < Tmelt) {
      epsi = (vent * R(2.0) * R(Pi) * rhor * dv) * ni_incld;
      epsi_tot = epsi_tot + epsi;
    }

    // calc_rime_density, P3I/p3_calc_rime_density_impl.hpp:14-73
    if (qc2qi_collect_tend >= R(QSMALL) && T_atm < T_zerodegc) {
      const R vtrmi1 = tv_qi_fallspd * rhofaci;
      if (qc_incld >= R(QSMALL)) {
        const R Vt_qc = R(s.acn) * m_tgamma(R(4 + bcn) + mu_c) / (m_pow(lamc, R(bcn)) * m_tgamma(R(4) + mu_c));
        const R D_c = (R(4) + mu_c) / lamc;
        const R V_impact = m_fabs(vtrmi1 - Vt_qc);
        const R inv_Tc = R(1) / vmin(R(-0.001), R(T_atm - T_zerodegc));
        const R Ri = vmax(R(1), vmin(R(-0.5e+6) * D_c * V_impact * inv_Tc, R(12)));
        if (Ri <= R(8.0)) rho_qm_cloud = double((R(0.051) + R(0.114) * Ri - R(0.0055) * (Ri * Ri)) * R(1000));
        else rho_qm_cloud = double(R(611.) + R(72.25) * (Ri - R(8)));
      }
    }

    if (rt.do_ice_production) {
      const R aimm = R(rt.immersion_freezing_exponent);
      const R dT_frz = R(T_zerodegc - T_atm);
      // cldliq_immersion_freezing, P3I/p3_cldliq_imm_freezing_impl.hpp:15-52
      if (qc_incld >= R(QSMALL) && T_atm <= T_rainfrz) {
        const R expAimmDt = m_exp(aimm * dT_frz);
        const R x = R(1) / lamc;
        const R inv_lamc3 = x * x * x;
        if constexpr (sizeof(R) == sizeof(double)) {
          qc2qi_hetero_freeze_tend = R(1.0) * R(CONS6) * cdist1 * m_tgamma(R(7) + mu_c) * expAimmDt * (inv_lamc3 * inv_lamc3);
        } else {
          // lamc^-6 is below the FP32 normal range; apply the two lamc^-3 factors apart
          qc2qi_hetero_freeze_tend = R(CONS6) * (cdist1 * inv_lamc3) * m_tgamma(R(7) + mu_c) * expAimmDt * inv_lamc3;
        }
        nc2ni_immers_freeze_tend = R(CONS5) * cdist1 * m_tgamma(R(4.0) + mu_c) * expAimmDt * inv_lamc3;
      }
      // rain_immersion_freezing, P3I/p3_rain_imm_freezing_impl.hpp:14-46
      if (qr_incld >= R(QSMALL) && T_atm <= T_rainfrz) {
        qr2qi_immers_freeze_tend = R(CONS6) * m_exp(m_log(cdistr) + m_log(m_tgamma(R(7.) + mu_r)) - R(6.) * m_log(lamr)) *
                                   m_exp(aimm * dT_frz);
        nr2ni_immers_freeze_tend = R(CONS5) * m_exp(m_log(cdistr) + m_log(m_tgamma(R(4.) + mu_r)) - R(3.) * m_log(lamr)) *
                                   m_exp(aimm * dT_frz);
      }
    }

    // calc_liq_relaxation_timescale, P3I/p3_calc_liq_relaxation_timescale_impl.hpp:14-44
    R epsr = R(0);
    if (qr_incld >= R(QSMALL)) {
      const Tab3<R> t3 = lookup3<R>(mu_r, lamr);
      epsr = R(2) * R(Pi) * cdistr * rhor * dv *
             (R(f1r) * m_tgamma(mu_r + R(2)) / lamr + R(f2r) * m_sqrt(rhor / mu) * m_cbrt(sc) * apply_table3<R>(TabOf<R>::get(tb).revap, t3));
    }

    // evaporate_rain, P3I/p3_evaporate_rain_impl.hpp:83-201
    {
      const double ssat_r_d = s.qv - s.qv_sat_l;  // FP64 difference of near-equal values
      const R ssat_r = R(ssat_r_d);
      const R cld_frac = ((qc_incld + qi_incld) < R(1.e-6)) ? R(0) : R(cld_frac_l);
      const R cfr = R(cld_frac_r);
      if (qr_incld >= R(QSMALL) && ssat_r_d < 0 && cfr > cld_frac) {
        const R inv_dt_r = R(1 / dt);
        const R tau_r = R(1) / epsr;
        const R dqv_dt = R((s.qv - s.qv_prev) * (1 / dt));
        const R dT = R(T_atm - s.t_prev);
        R eps_eff, A_c;
        if (T_atm < Tmelt) {
          const R fac = R(1) + R(latvap + latice) * R(1 / Cpair) * dqsdt;
          eps_eff = epsr + epsi_tot * fac / abi;
          A_c = dqv_dt - dqsdt * dT * inv_dt_r - R(s.qv_sat_l - s.qv_sat_i) * fac / abi * epsi_tot;
        } else {
          eps_eff = epsr;
          A_c = dqv_dt - dqsdt * dT * inv_dt_r;
        }
        if (eps_eff < R(1e-20)) eps_eff = R(1e-20);
        const R tau_eff = R(1) / eps_eff;
        // The limits below can remove all rain; they run in FP64 on the FP64
        // in-cloud mass (qr_incld is unchanged since entry) so that complete
        // evaporation leaves the same clipped residual in both forms.
        const double qr_incld_d = in0.qr;
        const double inv_dt_d = 1 / dt;
        double ev;
        if (qr_incld < R(1e-12) && s.qv / s.qv_sat_l < 0.999) {
          ev = qr_incld_d * inv_dt_d;
        } else {
          const R dt_over_tau = R(dt) / tau_eff;
          const R tscale_weight = -m_expm1(-dt_over_tau) / dt_over_tau;
          const R equilib_tend = -A_c / ab * tau_eff / tau_r;
          const R instant_tend = -ssat_r / (ab * tau_r);
          ev = double(instant_tend * tscale_weight + equilib_tend * (R(1) - tscale_weight));
        }
        const double abd = ab, cfd = cld_frac;
        if (ev > -ssat_r_d * inv_dt_d / abd) ev = -ssat_r_d * inv_dt_d / abd;
        if (ev < 0) ev = 0;
        if (ev > qr_incld_d * inv_dt_d) ev = qr_incld_d * inv_dt_d;
        ev = ev * (cld_frac_r - cfd) / cld_frac_r;
        evap_d = ev;
        nr_evap_d = ev * (double(nr_incld) / qr_incld_d);
      }
    }

    if (rt.do_ice_production) {
      // ice_deposition_sublimation, P3I/p3_ice_deposition_sublimation_impl.hpp:10-65
      if (qi_incld > R(QSMALL)) {
        const R qi_tend = vmin(epsi / abi, R(inv_dt)) * R(s.qv - s.qv_sat_i);
        const bool neg = qi_tend < R(0);
        if (neg) {
          qi2qv_sublim_tend = -qi_tend;
          ni_sublim_tend = qi2qv_sublim_tend * (ni_incld / qi_incld);
        }
        if (T_atm < T_zerodegc && !neg) qv2qi_vapdep_tend = qi_tend;
        if (T_atm < T_zerodegc) qc2qi_berg_tend = vmax(epsi / abi * R(s.qv_sat_l - s.qv_sat_i), R(0));
      }
    }
  }  // !skip_micro

  // ice_nucleation, P3I/p3_ice_nucleation_impl.hpp:9-60
  if (rt.do_ice_production) {
    const R mi0 = R(4.0 * PIOV3 * 900.0 * 1.e-18);
    const bool lt = T_atm < (Tmelt - 15.0);
    const bool ge = s.qv_supersat_i >= 0.05;
    const bool do_log = (!predictNc || prescribedCCN);
    const bool any_if_log = lt && ge && !do_log;
    const bool any_if_not_log = lt && ge && do_log;
    if (any_if_not_log) {
      const R ir = R(inv_rho);
      R dum = R(0.005) * m_exp(R(rt.deposition_nucleation_exponent) * R(Tmelt - T_atm)) * R(1.0e3) * ir;
      dum = vmin(dum, R(1.0e5) * ir);
      const R N_nuc = vmax(R(0.0), (dum - R(s.ni)) * R(inv_dt));
      if (N_nuc >= R(NSMALL)) {
        const R Q_nuc = vmax(R(0.0), (dum - R(s.ni)) * mi0 * R(inv_dt));
        qv2qi_nucleat_tend = Q_nuc;
        ni_nucleat_tend = N_nuc;
      }
    } else if (any_if_log) {
      ni_nucleat_tend = vmax(R(0.0), R((s.ni_activated - s.ni) * inv_dt));
      qv2qi_nucleat_tend = ni_nucleat_tend * mi0;
    }
  }

  // cloud_water_autoconversion, P3I/p3_autoconversion_impl.hpp:10-57
  {
    const double radius = rt.autoconversion_radius;
    const R cons3 = R((radius == 25.0e-6) ? CONS3 : 1.0 / (CONS2 * p3m_pow(radius, 3.0)));
    if (qc_incld >= R(1e-8)) {
      qc2qr_autoconv_tend = R(1.0) * R(rt.autoconversion_prefactor) * m_pow(qc_incld, R(rt.autoconversion_qc_exponent)) *
                            m_pow(nc_incld * R(1.e-6) * R(rho), R(-rt.autoconversion_nc_exponent));
      ncautr = qc2qr_autoconv_tend * cons3;
      nc2nr_autoconv_tend = qc2qr_autoconv_tend * nc_incld / qc_incld;
    }
    if (qc2qr_autoconv_tend == R(0)) nc2nr_autoconv_tend = R(0);
    if (nc2nr_autoconv_tend == R(0)) qc2qr_autoconv_tend = R(0);
  }
  // droplet_self_collection, P3I/p3_droplet_self_coll_impl.hpp:14-30: stays 0
  // cloud_rain_accretion, P3I/p3_cloud_rain_acc_impl.hpp:15-56
  if (qr_incld >= R(QSMALL) && qc_incld >= R(QSMALL)) {
    if (rt.accretion_qc_exponent == rt.accretion_qr_exponent)
      qc2qr_accret_tend = R(1.0) * R(rt.accretion_prefactor) * m_pow(qc_incld * qr_incld, R(rt.accretion_qr_exponent));
    else
      qc2qr_accret_tend = R(1.0) * R(rt.accretion_prefactor) * m_pow(qc_incld, R(rt.accretion_qc_exponent)) *
                          m_pow(qr_incld, R(rt.accretion_qr_exponent));
    nc_accret_tend = qc2qr_accret_tend * nc_incld / qc_incld;
    if (nc_accret_tend == R(0)) qc2qr_accret_tend = R(0);
    if (qc2qr_accret_tend == R(0)) nc_accret_tend = R(0);
  }
  // rain_self_collection, P3I/p3_rain_self_collection_impl.hpp:9-55
  if (qr_incld >= R(QSMALL)) {
    const R bd = R(rt.rain_selfcollection_breakup_diameter);
    const R dum2 = m_cbrt((qr_incld) / (R(Pi) * R(RHO_H2O) * nr_incld));
    R dum;
    if (dum2 < bd) dum = R(1);
    else dum = R(2) - m_exp(R(2300) * (dum2 - bd));
    nr_selfcollect_tend = dum * R(rt.rain_selfcollection_prefactor) * nr_incld * qr_incld * R(rho);
  }

  // ---------------- FP64 from here: cell averages, limiters, updates ----------------
  double t_qc2qr_accret = qc2qr_accret_tend, t_qr2qv_evap = evap_d,
         t_qc2qr_autoconv = qc2qr_autoconv_tend, t_nc_accret = nc_accret_tend,
         t_nc_selfcollect = nc_selfcollect_tend, t_nc2nr_autoconv = nc2nr_autoconv_tend,
         t_nr_selfcollect = nr_selfcollect_tend, t_nr_evap = nr_evap_d, t_ncautr = ncautr,
         t_qi2qv_sublim = qi2qv_sublim_tend, t_nr_ice_shed = nr_ice_shed_tend,
         t_qc2qi_hetero = qc2qi_hetero_freeze_tend, t_qr2qi_collect = qr2qi_collect_tend,
         t_qc2qr_ice_shed = qc2qr_ice_shed_tend, t_qi2qr_melt = qi2qr_melt_tend,
         t_qc2qi_collect = qc2qi_collect_tend, t_qr2qi_immers = qr2qi_immers_freeze_tend,
         t_ni2nr_melt = ni2nr_melt_tend, t_nc_collect = nc_collect_tend, t_ncshdc = ncshdc,
         t_nc2ni_immers = nc2ni_immers_freeze_tend, t_nr_collect = nr_collect_tend,
         t_ni_selfcollect = ni_selfcollect_tend, t_qv2qi_vapdep = qv2qi_vapdep_tend,
         t_nr2ni_immers = nr2ni_immers_freeze_tend, t_ni_sublim = ni_sublim_tend,
         t_qv2qi_nucleat = qv2qi_nucleat_tend, t_ni_nucleat = ni_nucleat_tend,
         t_qc2qi_berg = qc2qi_berg_tend;

  // back_to_cell_average, P3I/p3_back_to_cell_average_impl.hpp:14-76
  {
    const double ir_cldm = vmin(cld_frac_i, cld_frac_r);
    const double il_cldm = vmin(cld_frac_i, cld_frac_l);
    const double lr_cldm = vmin(cld_frac_l, cld_frac_r);
    t_qc2qr_accret *= lr_cldm;
    t_qr2qv_evap *= cld_frac_r;
    t_qc2qr_autoconv *= cld_frac_l;
    t_nc_accret *= lr_cldm;
    t_nc_selfcollect *= cld_frac_l;
    t_nc2nr_autoconv *= cld_frac_l;
    t_nr_selfcollect *= cld_frac_r;
    t_nr_evap *= cld_frac_r;
    t_ncautr *= lr_cldm;
    t_qi2qv_sublim *= cld_frac_i;
    t_nr_ice_shed *= il_cldm;
    t_qc2qi_hetero *= il_cldm;
    t_qr2qi_collect *= ir_cldm;
    t_qc2qr_ice_shed *= il_cldm;
    t_qi2qr_melt *= cld_frac_i;
    t_qc2qi_collect *= il_cldm;
    t_qr2qi_immers *= cld_frac_r;
    t_ni2nr_melt *= cld_frac_i;
    t_nc_collect *= il_cldm;
    t_ncshdc *= il_cldm;
    t_nc2ni_immers *= cld_frac_l;
    t_nr_collect *= ir_cldm;
    t_ni_selfcollect *= cld_frac_i;
    t_qv2qi_vapdep *= cld_frac_i;
    t_nr2ni_immers *= cld_frac_r;
    t_ni_sublim *= cld_frac_i;
    t_qc2qi_berg *= il_cldm;
  }

  // cloud_water_conservation, P3I/p3_conservation_impl.hpp:10-40
  {
    const double sinks = (t_qc2qr_autoconv + t_qc2qr_accret + t_qc2qi_collect + t_qc2qi_hetero + t_qc2qr_ice_shed +
                          t_qc2qi_berg) * dt;
    const double sources = s.qc;
    double ratio;
    if (sinks > sources && sinks >= QTENDSMALL) {
      ratio = sources / sinks;
      t_qc2qr_autoconv *= ratio;
      t_qc2qr_accret *= ratio;
      t_qc2qi_collect *= ratio;
      t_qc2qi_hetero *= ratio;
      t_qc2qr_ice_shed *= ratio;
      t_qc2qi_berg *= ratio;
    } else {
      ratio = 1;
    }
    if (sources > QTENDSMALL) {
      t_qv2qi_vapdep = t_qv2qi_vapdep * (1 - ratio);
      t_qi2qv_sublim = t_qi2qv_sublim * (1 - ratio);
    }
  }
  // rain_water_conservation, P3I/p3_conservation_impl.hpp:42-65
  {
    const double sinks = (t_qr2qv_evap + t_qr2qi_collect + t_qr2qi_immers) * dt;
    const double sources = s.qr + (t_qc2qr_autoconv + t_qc2qr_accret + t_qi2qr_melt + t_qc2qr_ice_shed) * dt;
    if (sinks > sources && sinks >= QTENDSMALL) {
      const double ratio = sources / sinks;
      t_qr2qv_evap *= ratio;
      t_qr2qi_collect *= ratio;
      t_qr2qi_immers *= ratio;
    }
  }
  // ice_water_conservation, P3I/p3_conservation_impl.hpp:67-96
  {
    const double sinks = (t_qi2qv_sublim + t_qi2qr_melt) * dt;
    const double sources = s.qi + (t_qv2qi_vapdep + t_qv2qi_nucleat + t_qr2qi_collect + t_qc2qi_collect + t_qr2qi_immers +
                                   t_qc2qi_hetero + t_qc2qi_berg) * dt;
    if (sinks > sources && sinks >= QTENDSMALL) {
      const double ratio = sources / sinks;
      t_qi2qv_sublim *= ratio;
      t_qi2qr_melt *= ratio;
    }
  }
  // nc_conservation, P3I/p3_nc_conservation_impl.hpp:16-28
  {
    const double sink_nc = (t_nc_collect + t_nc2ni_immers + t_nc_accret + t_nc2nr_autoconv) * dt;
    const double source_nc = s.nc + t_nc_selfcollect * dt;
    if (sink_nc > source_nc) {
      const double ratio = source_nc / sink_nc;
      t_nc_collect *= ratio;
      t_nc2ni_immers *= ratio;
      t_nc_accret *= ratio;
      t_nc2nr_autoconv *= ratio;
    }
  }
  // nr_conservation, P3I/p3_nr_conservation_impl.hpp:16-28
  {
    const double sink_nr = (t_nr_collect + t_nr2ni_immers + t_nr_selfcollect + t_nr_evap) * dt;
    const double source_nr = s.nr + (t_ni2nr_melt * nmltratio + t_nr_ice_shed + t_ncshdc + t_nc2nr_autoconv) * dt;
    if (sink_nr > source_nr) {
      const double ratio = source_nr / sink_nr;
      t_nr_collect *= ratio;
      t_nr2ni_immers *= ratio;
      t_nr_selfcollect *= ratio;
      t_nr_evap *= ratio;
    }
  }
  // ni_conservation, P3I/p3_ni_conservation_impl.hpp:16-27
  {
    const double sink_ni = (t_ni2nr_melt + t_ni_sublim + t_ni_selfcollect) * dt;
    const double source_ni = s.ni + (t_ni_nucleat + t_nr2ni_immers + t_nc2ni_immers) * dt;
    if (sink_ni > source_ni) {
      const double ratio = source_ni / sink_ni;
      t_ni2nr_melt *= ratio;
      t_ni_sublim *= ratio;
      t_ni_selfcollect *= ratio;
    }
  }
  // ice_supersat_conservation, P3I/p3_ice_supersat_conservation_impl.hpp:14-43
  {
    const double t_atm = s.th / s.inv_exner;
    const double latsublim2 = (latvap + latice) * (latvap + latice);
    const double qv_sink = t_qv2qi_vapdep + t_qv2qi_nucleat;
    if (qv_sink > QSMALL && cld_frac_i > 1e-20) {
      double qv_avail = (s.qv + (t_qi2qv_sublim + t_qr2qv_evap) * dt - s.qv_sat_i) /
                        (1 + latsublim2 * s.qv_sat_i / (CP * RH2O * (t_atm * t_atm))) / dt;
      qv_avail = vmax(qv_avail, 0.0);
      if (qv_sink > qv_avail) {
        const double fract = qv_avail / qv_sink;
        t_qv2qi_nucleat = t_qv2qi_nucleat * fract;
        t_qv2qi_vapdep = t_qv2qi_vapdep * fract;
      }
    }
  }
  // prevent_liq_supersaturation, P3I/p3_prevent_liq_supersaturation_impl.hpp:15-78
  {
    const double qv_sources = t_qi2qv_sublim + t_qr2qv_evap;
    if (qv_sources >= QSMALL) {
      const double qv_sinks = t_qv2qi_vapdep + t_qv2qi_nucleat;
      const double T_endstep = T_atm + ((qv_sinks - t_qi2qv_sublim) * (latvap + latice) * inv_cp -
                                        t_qr2qv_evap * latvap * inv_cp) * dt;
      const double qsl = qv_sat_dry(T_endstep, pres, false, err);
      const double A = latvap * qsl * dt * inv_cp / (RV * T_endstep * T_endstep) *
                       ((latvap + latice) * t_qi2qv_sublim + latvap * t_qr2qv_evap);
      double frac = (qsl - s.qv + qv_sinks * dt + A) / (qv_sources * dt + A);
      frac = vmax(0.0, frac);
      frac = vmin(1.0, frac);
      t_qi2qv_sublim = frac * t_qi2qv_sublim;
      t_qr2qv_evap = frac * t_qr2qv_evap;
    }
  }

  // update_prognostic_ice, P3I/p3_update_prognostics_impl.hpp:9-87
  {
    const double inv_exner = s.inv_exner;
    s.qc = s.qc + (-t_qc2qi_hetero - t_qc2qi_collect - t_qc2qr_ice_shed - t_qc2qi_berg) * dt;
    if (predictNc) s.nc = s.nc + (-t_nc_collect - t_nc2ni_immers) * dt;
    s.qr = s.qr + (-t_qr2qi_collect + t_qi2qr_melt - t_qr2qi_immers + t_qc2qr_ice_shed) * dt;
    s.nr = s.nr + (-t_nr_collect - t_nr2ni_immers + nmltratio * t_ni2nr_melt + t_nr_ice_shed + t_ncshdc) * dt;
    if (s.qi >= QSMALL) {
      s.bm = s.bm - ((t_qi2qv_sublim + t_qi2qr_melt) / s.qi) * dt * s.bm;
      s.qm = s.qm - ((t_qi2qv_sublim + t_qi2qr_melt) * s.qm / s.qi) * dt;
      s.qi = s.qi - (t_qi2qv_sublim + t_qi2qr_melt) * dt;
    }
    const double dum = (t_qr2qi_collect + t_qc2qi_collect + t_qr2qi_immers + t_qc2qi_hetero) * dt;
    s.qi = s.qi + (t_qv2qi_vapdep + t_qv2qi_nucleat + t_qc2qi_berg) * dt + dum;
    s.qm = s.qm + dum;
    s.bm = s.bm + (t_qr2qi_collect * INV_RHO_RIMEMAX + t_qc2qi_collect / rho_qm_cloud +
                   (t_qr2qi_immers + t_qc2qi_hetero) * INV_RHO_RIMEMAX) * dt;
    s.ni = s.ni + (t_ni_nucleat - t_ni2nr_melt - t_ni_sublim - t_ni_selfcollect + t_nr2ni_immers + t_nc2ni_immers) * dt;
    if (s.qm < 0) { s.qm = 0; s.bm = 0; }
    if (wetgrowth) { s.qm = s.qi; s.bm = s.qm * INV_RHO_RIMEMAX; }
    s.qv = s.qv + (-t_qv2qi_vapdep + t_qi2qv_sublim - t_qv2qi_nucleat) * dt;
    s.th = s.th + inv_exner * ((t_qv2qi_vapdep - t_qi2qv_sublim + t_qv2qi_nucleat) * (latvap + latice) * INV_CP +
                               (t_qr2qi_collect + t_qc2qi_collect + t_qc2qi_hetero + t_qr2qi_immers - t_qi2qr_melt +
                                t_qc2qi_berg) * latice * INV_CP) * dt;

    // update_prognostic_liquid, P3I/p3_update_prognostics_impl.hpp:89-122
    s.qc = s.qc + (-t_qc2qr_accret - t_qc2qr_autoconv) * dt;
    s.qr = s.qr + (t_qc2qr_accret + t_qc2qr_autoconv - t_qr2qv_evap) * dt;
    if (predictNc || prescribedCCN) s.nc = s.nc + (-t_nc_accret - t_nc2nr_autoconv + t_nc_selfcollect) * dt;
    else s.nc = NCCNST * inv_rho;
    s.nr = s.nr + (t_ncautr - t_nr_selfcollect - t_nr_evap) * dt;
    s.qv = s.qv + t_qr2qv_evap * dt;
    s.th = s.th + inv_exner * (-t_qr2qv_evap * latvap * INV_CP) * dt;

    // outputs, P3I/p3_main_impl_part2.hpp:456-463
    s.qv2qi_depos_tend = t_qv2qi_vapdep - t_qi2qv_sublim + t_qv2qi_nucleat;
    s.precip_total_tend = t_qc2qr_accret + t_qc2qr_autoconv + t_qc2qr_ice_shed + t_qc2qi_collect;
    s.nevapr = t_qi2qv_sublim + t_qr2qv_evap;
    s.vap_ice_exchange = t_qv2qi_vapdep - t_qi2qv_sublim + t_qv2qi_nucleat;
    s.vap_liq_exchange = -t_qr2qv_evap;
    s.liq_ice_exchange = t_qc2qi_hetero + t_qr2qi_immers - t_qi2qr_melt + t_qc2qi_berg + t_qc2qi_collect + t_qr2qi_collect;

    // clipping, P3I/p3_main_impl_part2.hpp:465-508
    bool present = false;
    if (s.qc < QSMALL) {
      s.qv = s.qv + s.qc;
      s.th = s.th - inv_exner * s.qc * latvap * inv_cp;
      s.qc = 0; s.nc = 0;
    } else present = true;
    if (s.qr < QSMALL) {
      s.qv = s.qv + s.qr;
      s.th = s.th - inv_exner * s.qr * latvap * inv_cp;
      s.qr = 0; s.nr = 0;
    } else present = true;
    if (s.qi < QSMALL) {
      s.qv = s.qv + s.qi;
      s.th = s.th - inv_exner * s.qi * (latvap + latice) * inv_cp;
      s.qi = 0; s.ni = 0; s.qm = 0; s.bm = 0;
    } else present = true;

    // P3I/p3_main_impl_part2.hpp:514-517
    double ni_incld_d = s.ni / cld_frac_i;
    impose_max_total_ni<double>(ni_incld_d, rt.max_total_ni, inv_rho);
    s.ni = ni_incld_d * cld_frac_i;
    return present;
  }
}

}  // namespace p3d
