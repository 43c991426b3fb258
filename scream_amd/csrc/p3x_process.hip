// p3x_process.hip -- the P3Microphysics atmosphere process around p3_main:
// pre-amble (wet -> dry mixing ratios, dz, exner, theta, cloud fractions with
// maximum-overlap rain fraction), p3_main, and post-amble (the rescaled
// temperature change, dry -> wet, effective radii in microns, the accumulated
// surface precipitation masses and the optional boundary fluxes).  The pre-amble
// runs inside the part 1 kernel and the post-amble inside the sedimentation +
// part 3 kernel (p3d::Amble), so a process step launches the p3_main kernels only.
//
// Follows components/eamxx/src/physics/p3/eamxx_p3_process_interface.hpp:63-261,
// eamxx_p3_process_interface.cpp:249-340 (which fields feed p3_main) and
// eamxx_p3_run.cpp:5-45 (the order of run_impl); PhysicsFunctions from
// components/eamxx/src/share/util/scream_common_physics_functions_impl.hpp.
// The pre/post-amble arithmetic is FP64 in every precision form; pow uses the
// same portable function as the kernels, so the process is reproducible
// against the oracle bit for bit.
#include <hip/hip_runtime.h>

#include <cstring>

#include "p3x_internal.hpp"

using namespace p3x_detail;

namespace {

struct CldFracParams {
  const double *qi, *liq;
  double *ice, *tot, *ice4, *tot4;
  double thr, thr4;
  size_t n;
};

// calc_icefrac (strict threshold) and calc_totalfrac for both thresholds,
// cld_fraction_main_impl.hpp:52-95; one thread per column-level.
__global__ __launch_bounds__(256) void p3_cld_fraction_kernel(const CldFracParams a) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n) return;
  const double qi = a.qi[x], liq = a.liq[x];
  const double ice = qi > a.thr ? 1.0 : 0.0, ice4 = qi > a.thr4 ? 1.0 : 0.0;
  a.ice[x] = ice;
  a.ice4[x] = ice4;
  // max(ice, liq) as a < b ? b : a (std::max's order: a NaN liquid fraction gives ice)
  a.tot[x] = ice < liq ? liq : ice;
  a.tot4[x] = ice4 < liq ? liq : ice4;
}

struct ShocPostParams {
  const double *qc, *qc2;
  double *cldfrac_liq, *inv_qc_relvar;
  size_t n;
};

// SHOCPostprocess's P3 inputs, eamxx_shoc_process_interface.hpp:288-300: the
// relative-variance ratio clamped to [0.001, 10] (1 where qc or its variance is
// 0) and the liquid cloud fraction clipped to at most 1; one thread per point.
// ekat's scalar min/max are taken as std::min/std::max (they differ only for a
// NaN argument, which no test feeds).
__global__ __launch_bounds__(256) void p3_shoc_post_kernel(const ShocPostParams a) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n) return;
  const double qc = a.qc[x], qc2 = a.qc2[x];
  double r = 1;
  if (qc != 0 && qc2 != 0) {
    const double q = qc * qc / qc2;
    const double lo = 0.001 < q ? q : 0.001;  // max(0.001, q)
    r = lo < 10.0 ? lo : 10.0;                 // min(10, .)
  }
  a.inv_qc_relvar[x] = r;
  if (a.cldfrac_liq) {
    const double c = a.cldfrac_liq[x];
    a.cldfrac_liq[x] = 1.0 < c ? 1.0 : c;       // min(c, 1)
  }
}

struct BudgetParams {
  p3x_budget_fields f;
  double *mass, *energy;
  int ncol, nlev;
  int cs, ks;  // element (i, k) at i*cs + k*ks
};

// One thread per column, levels in order: the sums round like the oracle's.
__global__ __launch_bounds__(256) void p3_column_budget_kernel(const BudgetParams b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.ncol) return;
  const p3x_budget_fields& f = b.f;
  const double g = p3c::gravit, cp = p3c::CP, lv = p3c::LatVap, lf = p3c::LatIce;
  double m = 0, e = 0;
  for (int k = 0; k < b.nlev; ++k) {
    const size_t x = (size_t)i * b.cs + (size_t)k * b.ks;
    const double dp = f.pseudo_density[x];
    const double qv = f.qv[x], qc = f.qc[x], qi = f.qi[x], qr = f.qr[x];
    m += (qv + qc + qi + qr) * dp / g;
    const double u2 = f.u ? f.u[x] * f.u[x] : 0.0;
    const double v2 = f.v ? f.v[x] * f.v[x] : 0.0;
    e += (f.T_mid[x] * cp + 0.5 * (u2 + v2) + (lv + lf) * qv + lf * (qc + qr)) * dp / g;
  }
  e += (f.phis ? f.phis[i] : 0.0) * (f.ps ? f.ps[i] : 0.0) / g;
  b.mass[i] = m;
  b.energy[i] = e;
}

struct ErrParams {
  const double *m0, *e0, *m1, *e1, *vapor, *water, *ice, *heat;
  double *merr, *eerr;
  int ncol;
  double dt;
};

__global__ __launch_bounds__(256) void p3_conservation_errors_kernel(const ErrParams a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncol) return;
  const double rho_w = p3c::RHO_H2O, lv = p3c::LatVap, lf = p3c::LatIce;
  const double tm_exp = a.m0[i] + (a.vapor[i] - a.water[i] * rho_w) * a.dt;
  a.merr[i] = fabs(a.m1[i] - tm_exp) / a.m0[i];
  const double te_exp =
      a.e0[i] + (a.vapor[i] * (lv + lf) - (a.water[i] - a.ice[i]) * rho_w * lf + a.heat[i]) * a.dt;
  a.eerr[i] = fabs(a.e1[i] - te_exp) / a.e0[i];
}

}  // namespace

extern "C" int p3x_column_budget(p3x_context* c, const p3x_budget_fields* f, int ncol, int nlev, double* mass,
                                 double* energy, void* stream) {
  if (!c || !f || !mass || !energy || ncol < 0 || nlev <= 0) return P3X_ERR_ARG;
  if (!f->pseudo_density || !f->T_mid || !f->qv || !f->qc || !f->qi || !f->qr) return P3X_ERR_ARG;
  p3d::KParams g;
  if (set_geometry(ncol, nlev, f->layout, f->ld, 0, &g) != P3X_OK) return P3X_ERR_ARG;
  if (ncol == 0) return P3X_OK;
  const BudgetParams b{*f, mass, energy, ncol, nlev, g.cs, g.ks};
  hipLaunchKernelGGL(p3_column_budget_kernel, dim3((ncol + 255) / 256), dim3(256), 0, (hipStream_t)stream, b);
  return hipGetLastError() == hipSuccess ? P3X_OK : P3X_ERR_DEVICE;
}

extern "C" int p3x_shoc_postprocess(p3x_context* c, const double* qc, const double* qc2, double* cldfrac_liq,
                                    double* inv_qc_relvar, int ncol, int nlev, void* stream) {
  if (!c || !qc || !qc2 || !inv_qc_relvar || ncol <= 0 || nlev <= 0) return P3X_ERR_ARG;
  const ShocPostParams a{qc, qc2, cldfrac_liq, inv_qc_relvar, (size_t)ncol * nlev};
  const size_t blocks = (a.n + 255) / 256;
  hipLaunchKernelGGL(p3_shoc_post_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? P3X_OK : P3X_ERR_DEVICE;
}

extern "C" int p3x_cld_fraction(p3x_context* c, const double* qi, const double* cldfrac_liq, int ncol,
                                int nlev, double ice_threshold, double ice_4out_threshold,
                                double* cldfrac_ice, double* cldfrac_tot, double* cldfrac_ice_4out,
                                double* cldfrac_tot_4out, void* stream) {
  if (!c || !qi || !cldfrac_liq || !cldfrac_ice || !cldfrac_tot || !cldfrac_ice_4out || !cldfrac_tot_4out ||
      ncol <= 0 || nlev <= 0)
    return P3X_ERR_ARG;
  CldFracParams a{qi, cldfrac_liq, cldfrac_ice, cldfrac_tot, cldfrac_ice_4out, cldfrac_tot_4out,
                  ice_threshold, ice_4out_threshold, (size_t)ncol * nlev};
  const size_t blocks = (a.n + 255) / 256;
  hipLaunchKernelGGL(p3_cld_fraction_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? P3X_OK : P3X_ERR_DEVICE;
}

extern "C" int p3x_conservation_errors(p3x_context* c, const double* m0, const double* e0, const double* m1,
                                       const double* e1, const double* vapor_flux, const double* water_flux,
                                       const double* ice_flux, const double* heat_flux, int ncol, double dt,
                                       double* mass_err, double* energy_err, void* stream) {
  if (!c || !m0 || !e0 || !m1 || !e1 || !vapor_flux || !water_flux || !ice_flux || !heat_flux || !mass_err ||
      !energy_err || ncol < 0)
    return P3X_ERR_ARG;
  if (ncol == 0) return P3X_OK;
  const ErrParams a{m0, e0, m1, e1, vapor_flux, water_flux, ice_flux, heat_flux, mass_err, energy_err, ncol, dt};
  hipLaunchKernelGGL(p3_conservation_errors_kernel, dim3((ncol + 255) / 256), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? P3X_OK : P3X_ERR_DEVICE;
}

struct p3x_process {
  p3x_context* ctx = nullptr;
  p3x_runtime rt{};
  int ncol = 0, nlev = 0, predictNc = 1, prescribedCCN = 0, it = 0;
  int layout = P3X_LEVEL_MAJOR, ld = 0, ld_iface = 0;  // of the fields and of the local arrays
  double* mem = nullptr;  // every process-local array below
  double *inv_exner, *th_atm, *cld_frac_l, *cld_frac_i, *dz, *qv2qi_depos_tend, *rho_qi;
  double *precip_liq_flux, *precip_ice_flux, *precip_liq_surf, *precip_ice_surf, *unused;
  size_t local_bytes = 0;  // of mem
};

extern "C" int p3x_process_create(p3x_context* c, const p3x_runtime* rt, int ncol, int nlev, int predictNc,
                                  int prescribedCCN, p3x_process** out) {
  return p3x_process_create_layout(c, rt, ncol, nlev, predictNc, prescribedCCN, P3X_LEVEL_MAJOR, 0, 0, out);
}

extern "C" int p3x_process_create_layout(p3x_context* c, const p3x_runtime* rt, int ncol, int nlev, int predictNc,
                                         int prescribedCCN, int layout, int ld, int ld_iface, p3x_process** out) {
  if (!c || !rt || !out || ncol <= 0 || nlev <= 0 || nlev > 64 * p3d::kMaxLevelsPerLane) return P3X_ERR_ARG;
  *out = nullptr;
  p3d::KParams g;
  if (set_geometry(ncol, nlev, layout, ld, ld_iface, &g) != P3X_OK) return P3X_ERR_ARG;
  p3x_process* p = new p3x_process();
  p->ctx = c;
  p->rt = *rt;
  p->ncol = ncol;
  p->nlev = nlev;
  p->predictNc = predictNc ? 1 : 0;
  p->prescribedCCN = prescribedCCN ? 1 : 0;
  // the local arrays have the fields' strides, so that one set of strides indexes both
  p->layout = layout;
  p->ld = layout == P3X_LEVEL_MAJOR ? g.ks : g.cs;
  p->ld_iface = layout == P3X_LEVEL_MAJOR ? 0 : g.fcs;
  const size_t lev = layout == P3X_LEVEL_MAJOR ? (size_t)nlev * g.ks : (size_t)ncol * g.cs;
  const size_t iface = layout == P3X_LEVEL_MAJOR ? (size_t)(nlev + 1) * g.fks : (size_t)ncol * g.fcs;
  const size_t total = 8 * lev + 2 * iface + 2 * (size_t)ncol;
  p->local_bytes = total * sizeof(double);
  const hipError_t e = hipMalloc((void**)&p->mem, total * sizeof(double));
  if (e != hipSuccess) {
    delete p;
    return dev_fail(e);
  }
  double* m = p->mem;
  p->inv_exner = m; m += lev;
  p->th_atm = m; m += lev;
  p->cld_frac_l = m; m += lev;
  p->cld_frac_i = m; m += lev;
  p->dz = m; m += lev;
  p->qv2qi_depos_tend = m; m += lev;
  p->rho_qi = m; m += lev;
  p->unused = m; m += lev;  // nccn when it is not prescribed (never read by p3_main then)
  p->precip_liq_flux = m; m += iface;
  p->precip_ice_flux = m; m += iface;
  p->precip_liq_surf = m; m += ncol;
  p->precip_ice_surf = m;
  // the context's scratch for this shape, once, as initialize_impl sizes its buffers
  const int st = p3x_reserve(c, ncol, nlev);
  if (st != P3X_OK) {
    (void)hipFree(p->mem);
    delete p;
    return st;
  }
  *out = p;
  return P3X_OK;
}

// The p3_main arguments of one process step on the field-manager fields f (the
// process's views, eamxx_p3_process_interface.cpp:284-337) and the fused
// pre/post-amble; the launch options must name the process's own layout (or
// none).  P3X_ERR_ARG for a missing field or another layout.
int p3x_detail::process_args(p3x_process* p, const p3x_atm_fields* f, double dt, const p3x_launch* opts,
                             ProcessArgs* out) {
  if (!p || !f || !(dt > 0)) return P3X_ERR_ARG;
  const void* must[] = {f->p_mid, f->p_dry_mid, f->pseudo_density, f->pseudo_density_dry, f->cldfrac_tot,
                        f->nc_nuceat_tend, f->ni_activated, f->inv_qc_relvar, f->T_mid, f->qv, f->qc, f->nc,
                        f->qr, f->nr, f->qi, f->qm, f->ni, f->bm, f->qv_prev_micro_step, f->T_prev_micro_step,
                        f->precip_liq_surf_mass, f->precip_ice_surf_mass, f->eff_radius_qc, f->eff_radius_qi,
                        f->eff_radius_qr, f->precip_total_tend, f->nevapr, f->rainfrac,
                        f->micro_liq_ice_exchange, f->micro_vap_liq_exchange, f->micro_vap_ice_exchange};
  for (const void* q : must)
    if (!q) return P3X_ERR_ARG;
  if (p->prescribedCCN && !f->nccn) return P3X_ERR_ARG;
  const bool some_flux = f->vapor_flux || f->water_flux || f->ice_flux || f->heat_flux;
  const bool all_flux = f->vapor_flux && f->water_flux && f->ice_flux && f->heat_flux;
  if (some_flux && !all_flux) return P3X_ERR_ARG;
  ProcessArgs& r = *out;
  if (opts) r.o = *opts;
  else p3x_launch_defaults(&r.o);
  // the layout is the process's own (its local arrays were allocated in it)
  if ((r.o.layout != P3X_LEVEL_MAJOR && r.o.layout != p->layout) || (r.o.ld != 0 && r.o.ld != p->ld) ||
      (r.o.ld_iface != 0 && r.o.ld_iface != p->ld_iface))
    return P3X_ERR_ARG;
  r.o.layout = p->layout;
  r.o.ld = p->ld;
  r.o.ld_iface = p->ld_iface;
  // p3_main on the process's views (eamxx_p3_process_interface.cpp:284-337); the
  // pre-amble runs inside part 1 and the post-amble inside part 3.
  r.ps = {f->qc, f->nc, f->qr, f->nr, f->qi, f->qm, f->ni, f->bm, f->qv, p->th_atm};
  r.di = {f->nc_nuceat_tend, p->prescribedCCN ? f->nccn : p->unused, f->ni_activated,
          f->inv_qc_relvar, p->cld_frac_i, p->cld_frac_l, f->rainfrac, f->p_dry_mid,
          p->dz, f->pseudo_density_dry, p->inv_exner, f->qv_prev_micro_step,
          f->T_prev_micro_step};
  r.dout = {p->qv2qi_depos_tend, p->precip_liq_surf, p->precip_ice_surf,
            f->eff_radius_qc, f->eff_radius_qi, f->eff_radius_qr, p->rho_qi,
            p->precip_liq_flux, p->precip_ice_flux, f->precip_total_tend, f->nevapr};
  r.hist = {f->micro_liq_ice_exchange, f->micro_vap_liq_exchange, f->micro_vap_ice_exchange};
  // run_impl zeroes the three exchange fields before p3_main (eamxx_p3_run.cpp:27-29)
  r.inf = {dt, p->it + 1, 0, p->ncol - 1, 0, p->nlev - 1, p->predictNc, p->prescribedCCN, nullptr, 1};
  p3d::Amble& a = r.amble;
  a = p3d::Amble{};
  a.on = 1;
  a.p_mid = f->p_mid;
  a.dp_wet = f->pseudo_density;
  a.cldfrac_tot = f->cldfrac_tot;
  a.T_mid = f->T_mid;
  a.T_prev = f->T_prev_micro_step;
  a.qv_prev = f->qv_prev_micro_step;
  a.dz = p->dz;
  a.inv_exner = p->inv_exner;
  a.cld_frac_l = p->cld_frac_l;
  a.cld_frac_i = p->cld_frac_i;
  a.rainfrac = f->rainfrac;
  a.precip_liq_mass = f->precip_liq_surf_mass;
  a.precip_ice_mass = f->precip_ice_surf_mass;
  a.vapor_flux = f->vapor_flux;
  a.water_flux = f->water_flux;
  a.ice_flux = f->ice_flux;
  a.heat_flux = f->heat_flux;
  return P3X_OK;
}

int p3x_detail::process_kparams(p3x_process* p, const p3x_atm_fields* f, double dt, p3d::KParams* k) {
  ProcessArgs r;
  int st = process_args(p, f, dt, nullptr, &r);
  if (st != P3X_OK) return st;
  st = fill_params(p->ctx, &p->rt, &r.ps, &r.di, &r.dout, &r.inf, &r.hist, p->ncol, p->nlev, r.o.layout, r.o.ld,
                   r.o.ld_iface, nullptr, nullptr, nullptr, nullptr, k);
  if (st != P3X_OK) return st;
  k->amble = r.amble;
  return P3X_OK;
}

size_t p3x_detail::process_local_bytes(const p3x_process* p) { return p->local_bytes; }
p3x_context* p3x_detail::process_context(const p3x_process* p) { return p->ctx; }

extern "C" int p3x_process_run(p3x_process* p, const p3x_atm_fields* f, double dt, const p3x_launch* opts) {
  ProcessArgs r;
  const int st0 = process_args(p, f, dt, opts, &r);
  if (st0 != P3X_OK) return st0;
  ++p->it;
  p3x_launch mo = r.o;
  mo.async = 1;
  const int st = main_impl(p->ctx, &p->rt, &r.ps, &r.di, &r.dout, &r.inf, &r.hist, p->ncol, p->nlev, &mo, nullptr,
                           &r.amble);
  if (st != P3X_OK) return st;
  if (r.o.async) return P3X_OK;
  return p3x_sync(p->ctx);
}

// The neighbour stages of a served process step (include/p3x.h, p3x_step_chain),
// enqueued on s around the step's kernels in the order of MicrophysicsStep:
// SHOC's post-process, CldFraction and the first column budget before; the
// second budget and the conservation errors after.
int p3x_detail::chain_check(const p3x_process* p, const p3x_atm_fields* f, const p3x_step_chain* c) {
  if (!c) return P3X_OK;
  const bool dense = p->layout == P3X_LEVEL_MAJOR ? p->ld == p->ncol : p->ld == p->nlev;
  if ((c->qc2 || c->cldfrac_liq) && !dense) return P3X_ERR_ARG;  // the pointwise stages take dense arrays
  if (c->qc2 && !f->inv_qc_relvar) return P3X_ERR_ARG;
  if (c->cldfrac_liq && (!c->cldfrac_ice || !c->cldfrac_ice_4out || !c->cldfrac_tot_4out)) return P3X_ERR_ARG;
  if (c->mass_err && (!c->energy_err || !c->budget || !f->vapor_flux)) return P3X_ERR_ARG;
  return P3X_OK;
}

int p3x_detail::chain_before(const p3x_process* p, const p3x_atm_fields* f, const p3x_step_chain* c, hipStream_t s) {
  if (!c) return P3X_OK;
  int st = P3X_OK;
  if (c->qc2)
    st = p3x_shoc_postprocess(p->ctx, f->qc, c->qc2, c->cldfrac_liq, const_cast<double*>(f->inv_qc_relvar),
                              p->ncol, p->nlev, s);
  if (st == P3X_OK && c->cldfrac_liq)
    st = p3x_cld_fraction(p->ctx, f->qi, c->cldfrac_liq, p->ncol, p->nlev, c->ice_threshold, c->ice_4out_threshold,
                          c->cldfrac_ice, const_cast<double*>(f->cldfrac_tot), c->cldfrac_ice_4out,
                          c->cldfrac_tot_4out, s);
  if (st == P3X_OK && c->mass_err) {
    const p3x_budget_fields b{f->pseudo_density, f->T_mid, c->u, c->v, f->qv, f->qc, f->qi, f->qr, c->ps, c->phis,
                              p->layout, p->ld};
    st = p3x_column_budget(p->ctx, &b, p->ncol, p->nlev, c->budget, c->budget + p->ncol, s);
  }
  return st;
}

int p3x_detail::chain_after(const p3x_process* p, const p3x_atm_fields* f, const p3x_step_chain* c, double dt,
                            hipStream_t s) {
  if (!c || !c->mass_err) return P3X_OK;
  const int n = p->ncol;
  const p3x_budget_fields b{f->pseudo_density, f->T_mid, c->u, c->v, f->qv, f->qc, f->qi, f->qr, c->ps, c->phis,
                            p->layout, p->ld};
  int st = p3x_column_budget(p->ctx, &b, n, p->nlev, c->budget + 2 * n, c->budget + 3 * n, s);
  if (st == P3X_OK)
    st = p3x_conservation_errors(p->ctx, c->budget, c->budget + n, c->budget + 2 * n, c->budget + 3 * n,
                                 f->vapor_flux, f->water_flux, f->ice_flux, f->heat_flux, n, dt, c->mass_err,
                                 c->energy_err, s);
  return st;
}

extern "C" void p3x_process_destroy(p3x_process* p) {
  if (!p) return;
  (void)hipDeviceSynchronize();
  (void)hipFree(p->mem);
  delete p;
}
