/*
 * p3_oracle.h -- CPU restatement of EAMxx P3 microphysics (P3Functions::p3_main).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP product is compared
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links it and never falls back to it.
 *
 * Algorithm followed: components/eamxx/src/physics/p3/impl/p3_main_impl.hpp
 * (monolithic path, :74-331) with pack size 1 and kdir = -1, ktop = 0,
 * kbot = nlev-1 (p3_main_impl.hpp:97-99).  Every function in p3_oracle.c
 * cites the reference file:line it restates.
 *
 * Layout: every per-level array is level-major, element (k, i) at k*ncol + i
 * (k = 0 is the model top).  precip_liq_flux / precip_ice_flux have nlev+1
 * rows.  precip_liq_surf / precip_ice_surf have ncol entries.
 */
#ifndef P3_ORACLE_H
#define P3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* P3Runtime defaults: components/eamxx/src/physics/p3/p3_functions.hpp:112-158 */
typedef struct p3o_runtime {
  double max_total_ni;
  double autoconversion_prefactor;
  double autoconversion_qc_exponent;
  double autoconversion_nc_exponent;
  double autoconversion_radius;
  double accretion_prefactor;
  double accretion_qc_exponent;
  double accretion_qr_exponent;
  double rain_selfcollection_prefactor;
  double rain_selfcollection_breakup_diameter;
  double constant_mu_rain;
  double spa_ccn_to_nc_factor;
  double cldliq_to_ice_collection_factor;
  double rain_to_ice_collection_factor;
  double min_rime_rho;
  double max_rime_rho;
  double immersion_freezing_exponent;
  double deposition_nucleation_exponent;
  double ice_sedimentation_factor;
  int do_ice_production;
} p3o_runtime;

/* P3LookupTables (p3_functions.hpp:279-290), row-major:
 *   ice_table     [5][4][50][12]
 *   collect_table [5][4][50][30][2]   (log10 of the table text values)
 *   vn, vm, revap [300][10]
 *   mu_r          [150]
 *   dnu           [16]                                                      */
typedef struct p3o_tables {
  const double* ice_table;
  const double* collect_table;
  const double* vn_table;
  const double* vm_table;
  const double* revap_table;
  const double* mu_r_table;
  const double* dnu;
} p3o_tables;

/* All arrays of P3PrognosticState, P3DiagnosticInputs, P3DiagnosticOutputs
 * and P3HistoryOnly (p3_functions.hpp:161-276), level-major. */
typedef struct p3o_fields {
  /* prognostic state, in/out */
  double *qc, *nc, *qr, *nr, *qi, *qm, *ni, *bm, *qv, *th;
  /* diagnostic inputs */
  const double *nc_nuceat_tend, *nccn, *ni_activated, *inv_qc_relvar;
  const double *cld_frac_i, *cld_frac_l, *cld_frac_r;
  const double *pres, *dz, *dpres, *inv_exner, *qv_prev, *t_prev;
  /* diagnostic outputs */
  double *qv2qi_depos_tend, *precip_liq_surf, *precip_ice_surf;
  double *diag_eff_radius_qc, *diag_eff_radius_qi, *diag_eff_radius_qr;
  double *rho_qi, *precip_liq_flux, *precip_ice_flux;
  double *precip_total_tend, *nevapr;
  /* history (in/out: p3_main only overwrites/updates them, the caller zeroes) */
  double *liq_ice_exchange, *vap_liq_exchange, *vap_ice_exchange;
} p3o_fields;

/* Optional integer diagnostics (any pointer may be NULL).
 *   col_status [ncol]    : 0 skipped after part 1, 1 skipped after part 2, 2 full
 *   sed_substeps [3*ncol]: number of CFL substeps (cloud, rain, ice) of column i
 *                          at [0*ncol+i], [1*ncol+i], [2*ncol+i]
 *   ice_table_idx [nlev*ncol]: part-3 ice table cell packed as
 *                          dumjj*10000 + dumii*100 + dumi (0-based) or -1       */
typedef struct p3o_debug {
  int* col_status;
  int* sed_substeps;
  int* ice_table_idx;
} p3o_debug;

/* The P3Microphysics fields its pre-amble and post-amble touch
 * (components/eamxx/src/physics/p3/eamxx_p3_process_interface.hpp:63-353):
 * field-manager inputs, the state it converts in place, and the locals it
 * builds for p3_main.  Level-major; water_flux == NULL turns the optional
 * column-conservation boundary fluxes off. */
typedef struct p3o_atm {
  /* inputs */
  const double *p_mid, *pseudo_density, *pseudo_density_dry, *cldfrac_tot;
  /* state: wet <-> dry in place */
  double *T_mid, *qv, *qc, *nc, *qr, *nr, *qi, *qm, *ni, *bm, *qv_prev, *T_prev;
  /* pre-amble outputs (p3_main inputs) */
  double *th_atm, *inv_exner, *dz, *cld_frac_l, *cld_frac_i, *cld_frac_r;
  /* post-amble: p3_main outputs and the accumulated surface masses */
  double *eff_radius_qc, *eff_radius_qi, *eff_radius_qr;
  const double *precip_liq_surf, *precip_ice_surf;
  double *precip_liq_surf_mass, *precip_ice_surf_mass;
  double *vapor_flux, *water_flux, *ice_flux, *heat_flux;
} p3o_atm;

void p3o_preamble(const p3o_atm* a, int ncol, int nlev);
void p3o_postamble(const p3o_atm* a, int ncol, int nlev, double dt);

/* Column total mass and energy, and the relative errors of a step against
 * the boundary fluxes, of the mass-and-energy column conservation check. */
void p3o_column_budget(const double* pseudo_density, const double* T_mid, const double* u, const double* v,
                       const double* qv, const double* qc, const double* qi, const double* qr,
                       const double* ps, const double* phis, int ncol, int nlev,
                       double* mass, double* energy);
void p3o_conservation_errors(const double* mass0, const double* energy0, const double* mass1,
                             const double* energy1, const double* vapor_flux, const double* water_flux,
                             const double* ice_flux, const double* heat_flux, int ncol, double dt,
                             double* mass_err, double* energy_err);

/* The CldFraction process before P3: ice cloud fraction 1 where qi > threshold,
 * total = max(ice, liquid), for both thresholds.  Level-major, ld = ncol. */
void p3o_cld_fraction(const double* qi, const double* cldfrac_liq, int ncol, int nlev, double ice_threshold,
                      double ice_4out_threshold, double* cldfrac_ice, double* cldfrac_tot,
                      double* cldfrac_ice_4out, double* cldfrac_tot_4out);

/* SHOC's post-process that produces P3's inv_qc_relvar (and clips cldfrac_liq,
 * which may be NULL).  Dense arrays of ncol*nlev points. */
void p3o_shoc_postprocess(const double* qc, const double* qc2, double* cldfrac_liq, double* inv_qc_relvar,
                          int ncol, int nlev);

/* Defaults of P3Runtime. */
void p3o_runtime_defaults(p3o_runtime* rt);

/* One p3_main call over ncol columns.  Returns 0, or a positive count of
 * columns in which the saturation temperature check failed (T <= 0 or NaN,
 * physics_saturation_impl.hpp:14-36). */
int p3o_main(const p3o_runtime* rt, const p3o_tables* tab, const p3o_fields* f,
             const p3o_debug* dbg, int ncol, int nlev, double dt,
             int predictNc, int prescribedCCN, int nthreads);
/* p3o_main with the early exits of the reference's small-kernel dispatch
 * (small_kernel_exits = 1: a column whose hydrometeors part 2 removed still runs
 * sedimentation, homogeneous freezing and part 3) or the monolithic ones (0). */
int p3o_main_ex(const p3o_runtime* rt, const p3o_tables* tab, const p3o_fields* f,
                const p3o_debug* dbg, int ncol, int nlev, double dt,
                int predictNc, int prescribedCCN, int nthreads, int small_kernel_exits);

/* Building blocks exported for unit tests. */
double p3o_qv_sat_dry(double t, double p, int ice);
double p3o_elementary(int fn, double x, double y);
void p3o_lookup_ice(double qi, double ni, double qm, double rhop,
                    int* dumi, int* dumii, int* dumjj,
                    double* dum1, double* dum4, double* dum5);
void p3o_lookup_rain(double qr, double nr, int* dumj, double* dum3);
double p3o_apply_table_ice(const p3o_tables* tab, int idx, double qi, double ni,
                           double qm, double rhop);
double p3o_apply_table_coll(const p3o_tables* tab, int idx, double qi, double ni,
                            double qm, double rhop, double qr, double nr);
void p3o_upwind_step(int nk, int k_bot, int k_top, double dt_sub, int nfield,
                     const double* rho, const double* inv_rho, const double* inv_dz,
                     double* const* flux, const double* const* V, double* const* r);

/* The ice lookup-table generator (oracle/p3_table_gen.c, its own library):
 * cells (i_rhor, i_Fr, i_Qnorm) with mask[cell] set (mask NULL = all 1000);
 * ice[cell][15] and coll[cell][30][6] as the generator's output rows. */
void p3o_table_gen(const unsigned char* mask, double* ice, double* coll, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
