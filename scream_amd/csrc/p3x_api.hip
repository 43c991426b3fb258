// p3x_api.hip -- the C ABI of include/p3x.h: context, table upload, launch.
//
// A context owns the device copies of the lookup tables (P3LookupTables,
// components/eamxx/src/physics/p3/p3_functions.hpp:279-290), the level-major
// scratch the fused kernel keeps between phases (the counterpart of the
// reference's WorkspaceManager, eamxx_p3_process_interface.cpp:423-424), an
// error word and timing events.  Nothing is allocated per call unless a
// larger problem than any before needs more scratch.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "../../include/p3x.h"
#include "p3_launch.hpp"

struct p3x_context {
  int device = 0;
  double* d_ice = nullptr;
  double* d_coll = nullptr;
  double* d_vn = nullptr;
  double* d_vm = nullptr;
  double* d_revap = nullptr;
  double* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  unsigned* d_err = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipStream_t last_stream = nullptr;
  int pending_timing = 0;
};

namespace {

int dev_fail(hipError_t e) { return e == hipErrorOutOfMemory ? P3X_ERR_NO_MEMORY : P3X_ERR_DEVICE; }

// Tuned default block shape per (precision, nlev); see the autotuner in
// scream_amd/tune.py for how these were chosen.
void default_shape(int precision, int nk, int* cb, int* nt) {
  (void)precision;
  (void)nk;
  *cb = 32;
  *nt = 256;
}

}  // namespace

extern "C" void p3x_runtime_defaults(p3x_runtime* rt) {
  if (!rt) return;
  rt->max_total_ni = 740.0e3;
  rt->autoconversion_prefactor = 1350.0;
  rt->autoconversion_qc_exponent = 2.47;
  rt->autoconversion_nc_exponent = 1.79;
  rt->autoconversion_radius = 25.0e-6;
  rt->accretion_prefactor = 67.0;
  rt->accretion_qc_exponent = 1.15;
  rt->accretion_qr_exponent = 1.15;
  rt->rain_selfcollection_prefactor = 5.78;
  rt->rain_selfcollection_breakup_diameter = 0.00028;
  rt->constant_mu_rain = 1.0;
  rt->spa_ccn_to_nc_factor = 1.0;
  rt->cldliq_to_ice_collection_factor = 0.5;
  rt->rain_to_ice_collection_factor = 1.0;
  rt->min_rime_rho = 50.0;
  rt->max_rime_rho = 900.0;
  rt->immersion_freezing_exponent = 0.65;
  rt->deposition_nucleation_exponent = 0.304;
  rt->ice_sedimentation_factor = 1.0;
  rt->do_ice_production = 1;
}

extern "C" void p3x_launch_defaults(p3x_launch* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->precision = P3X_FP64;
}

extern "C" const char* p3x_strerror(int st) {
  switch (st) {
    case P3X_OK: return "ok";
    case P3X_ERR_ARG: return "invalid argument";
    case P3X_ERR_TABLE_FILE: return "cannot read the P3 ice lookup table";
    case P3X_ERR_TABLE_VERSION: return "P3 ice lookup table: expected header VERSION 4.1.1";
    case P3X_ERR_DEVICE: return "HIP runtime error";
    case P3X_ERR_TEMPERATURE: return "temperature <= 0 or NaN reached the saturation functions";
    case P3X_ERR_NO_MEMORY: return "device memory allocation failed";
    case P3X_ERR_SUBSTEPS: return "sedimentation needed more than 100000 CFL substeps (corrupt input)";
    default: return "unknown P3 status";
  }
}

extern "C" int p3x_init(const p3x_table_data* t, p3x_context** out) {
  if (!t || !out) return P3X_ERR_ARG;
  *out = nullptr;
  p3x_context* c = new p3x_context();
  hipError_t e = hipGetDevice(&c->device);
  auto up = [&](double** dst, const double* src, size_t n) {
    if (e != hipSuccess) return;
    e = hipMalloc((void**)dst, n * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(*dst, src, n * sizeof(double), hipMemcpyHostToDevice);
  };
  up(&c->d_ice, t->ice_table, p3c::ICE_TABLE_LEN);
  up(&c->d_coll, t->collect_table, p3c::COLL_TABLE_LEN);
  up(&c->d_vn, t->vn_table, p3c::RAIN_TABLE_LEN);
  up(&c->d_vm, t->vm_table, p3c::RAIN_TABLE_LEN);
  up(&c->d_revap, t->revap_table, p3c::RAIN_TABLE_LEN);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_err, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(c->d_err, 0, sizeof(unsigned));
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    p3x_finalize(c);
    return dev_fail(e);
  }
  *out = c;
  return P3X_OK;
}

extern "C" void p3x_finalize(p3x_context* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  // teardown: nothing useful can be done with a failure here
  (void)hipFree(c->d_ice);
  (void)hipFree(c->d_coll);
  (void)hipFree(c->d_vn);
  (void)hipFree(c->d_vm);
  (void)hipFree(c->d_revap);
  (void)hipFree(c->d_scratch);
  (void)hipFree(c->d_err);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  delete c;
}

extern "C" int p3x_sync(p3x_context* c) {
  if (!c) return P3X_ERR_ARG;
  hipError_t e = hipStreamSynchronize(c->last_stream);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  unsigned err = 0;
  e = hipMemcpy(&err, c->d_err, sizeof(unsigned), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  if (err) {
    (void)hipMemset(c->d_err, 0, sizeof(unsigned));
    return (err & p3d::ERR_BAD_TEMPERATURE) ? P3X_ERR_TEMPERATURE : P3X_ERR_SUBSTEPS;
  }
  return P3X_OK;
}

extern "C" int p3x_main(p3x_context* c, const p3x_runtime* runtime, const p3x_prognostic_state* ps,
                        const p3x_diagnostic_inputs* di, const p3x_diagnostic_outputs* dout,
                        const p3x_infrastructure* infra, const p3x_history_only* hist, int nj, int nk,
                        const p3x_launch* opts, double* elapsed_us) {
  if (!c || !runtime || !ps || !di || !dout || !infra || !hist) return P3X_ERR_ARG;
  if (nj < 0 || nk <= 0 || !(infra->dt > 0)) return P3X_ERR_ARG;
  p3x_launch o;
  if (opts) o = *opts;
  else p3x_launch_defaults(&o);
  const int ld = o.ld > 0 ? o.ld : nj;
  if (ld < nj) return P3X_ERR_ARG;
  if (o.precision != P3X_FP64 && o.precision != P3X_MIXED) return P3X_ERR_ARG;
  int cb = o.columns_per_block, nt = o.threads_per_block;
  if (cb == 0 || nt == 0) default_shape(o.precision, nk, &cb, &nt);
  hipStream_t s = (hipStream_t)o.stream;
  if (nj == 0) return P3X_OK;

  const void* must[] = {ps->qc, ps->nc, ps->qr, ps->nr, ps->qi, ps->qm, ps->ni, ps->bm, ps->qv, ps->th,
                        di->nc_nuceat_tend, di->nccn, di->ni_activated, di->inv_qc_relvar, di->cld_frac_i,
                        di->cld_frac_l, di->cld_frac_r, di->pres, di->dz, di->dpres, di->inv_exner,
                        di->qv_prev, di->t_prev, dout->qv2qi_depos_tend, dout->precip_liq_surf,
                        dout->precip_ice_surf, dout->diag_eff_radius_qc, dout->diag_eff_radius_qi,
                        dout->diag_eff_radius_qr, dout->rho_qi, dout->precip_liq_flux, dout->precip_ice_flux,
                        dout->precip_total_tend, dout->nevapr, hist->liq_ice_exchange,
                        hist->vap_liq_exchange, hist->vap_ice_exchange};
  for (const void* p : must)
    if (!p) return P3X_ERR_ARG;

  // scratch: SC_COUNT level-major fields of nk x ld doubles
  const size_t need = (size_t)p3d::SC_COUNT * nk * ld * sizeof(double);
  if (need > c->scratch_bytes) {
    hipError_t e = hipStreamSynchronize(c->last_stream);
    if (e == hipSuccess && c->d_scratch) e = hipFree(c->d_scratch);
    c->d_scratch = nullptr;
    c->scratch_bytes = 0;
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_scratch, need);
    if (e != hipSuccess) return dev_fail(e);
    c->scratch_bytes = need;
  }

  p3d::KParams k;
  k.qc = ps->qc; k.nc = ps->nc; k.qr = ps->qr; k.nr = ps->nr; k.qi = ps->qi;
  k.qm = ps->qm; k.ni = ps->ni; k.bm = ps->bm; k.qv = ps->qv; k.th = ps->th;
  k.nc_nuceat_tend = di->nc_nuceat_tend; k.nccn = di->nccn; k.ni_activated = di->ni_activated;
  k.inv_qc_relvar = di->inv_qc_relvar; k.cld_frac_i = di->cld_frac_i; k.cld_frac_l = di->cld_frac_l;
  k.cld_frac_r = di->cld_frac_r; k.pres = di->pres; k.dz = di->dz; k.dpres = di->dpres;
  k.inv_exner = di->inv_exner; k.qv_prev = di->qv_prev; k.t_prev = di->t_prev;
  k.qv2qi_depos_tend = dout->qv2qi_depos_tend; k.precip_liq_surf = dout->precip_liq_surf;
  k.precip_ice_surf = dout->precip_ice_surf; k.diag_eff_radius_qc = dout->diag_eff_radius_qc;
  k.diag_eff_radius_qi = dout->diag_eff_radius_qi; k.diag_eff_radius_qr = dout->diag_eff_radius_qr;
  k.rho_qi = dout->rho_qi; k.precip_liq_flux = dout->precip_liq_flux; k.precip_ice_flux = dout->precip_ice_flux;
  k.precip_total_tend = dout->precip_total_tend; k.nevapr = dout->nevapr;
  k.liq_ice_exchange = hist->liq_ice_exchange; k.vap_liq_exchange = hist->vap_liq_exchange;
  k.vap_ice_exchange = hist->vap_ice_exchange;
  k.scratch = c->d_scratch;
  k.err = c->d_err;
  k.dbg_status = o.col_status;
  k.dbg_substeps = o.sed_substeps;
  k.dbg_ice_idx = o.ice_table_idx;
  k.tab.ice = c->d_ice; k.tab.coll = c->d_coll; k.tab.vn = c->d_vn; k.tab.vm = c->d_vm; k.tab.revap = c->d_revap;
  static_assert(sizeof(p3d::Runtime) == sizeof(p3x_runtime), "runtime layouts differ");
  std::memcpy(&k.rt, runtime, sizeof(k.rt));
  k.ncol = nj; k.nlev = nk; k.ld = ld;
  k.dt = infra->dt;
  k.predictNc = infra->predictNc ? 1 : 0;
  k.prescribedCCN = infra->prescribedCCN ? 1 : 0;
  k.exch_zero = infra->exchange_zeroed_on_entry ? 1 : 0;

  hipError_t e = hipEventRecord(c->ev0, s);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  e = o.precision == P3X_FP64 ? p3d::launch_p3_fp64(k, cb, nt, s) : p3d::launch_p3_mixed(k, cb, nt, s);
  if (e == hipErrorInvalidValue) return P3X_ERR_ARG;
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  e = hipEventRecord(c->ev1, s);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  c->last_stream = s;
  if (o.async) return P3X_OK;
  const int st = p3x_sync(c);
  if (elapsed_us) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) *elapsed_us = 1000.0 * ms;
  }
  return st;
}
