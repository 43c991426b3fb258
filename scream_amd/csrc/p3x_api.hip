// p3x_api.hip -- the C ABI of include/p3x.h: context, table upload, launch.
//
// A context owns the device copies of the lookup tables (P3LookupTables,
// components/eamxx/src/physics/p3/p3_functions.hpp:279-290), the scratch the
// kernels of a step share (the counterpart of the reference's
// WorkspaceManager, eamxx_p3_process_interface.cpp:423-424), an error word and
// timing events.  Nothing is allocated per call unless a larger problem than
// any before needs more scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/p3x.h"
#include "p3x_internal.hpp"

// The scratch of p3x_main_noalloc: one fixed array per GPU, allocated with the
// code object and held until the process ends.  In host code this name is only
// the runtime's handle for the symbol, not a device address: p3x_init resolves
// the array's address on its device (hipGetSymbolAddress) and the calls pass
// that address to the kernels.
static __device__ __align__(256) unsigned char scratch[P3X_NOALLOC_SCRATCH_BYTES];

namespace {
// Device address of `scratch` per device ordinal, resolved once per device.
std::mutex g_noalloc_mu;
std::vector<double*> g_noalloc_ptr;

// With `dev` current: the address of `scratch` on it (resolved on first use).
hipError_t noalloc_scratch_on(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (dev < 0) return hipErrorInvalidDevice;
  if ((size_t)dev < g_noalloc_ptr.size() && g_noalloc_ptr[dev]) {
    *out = g_noalloc_ptr[dev];
    return hipSuccess;
  }
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  if (g_noalloc_ptr.size() <= (size_t)dev) g_noalloc_ptr.resize(dev + 1, nullptr);
  g_noalloc_ptr[dev] = (double*)ptr;
  *out = (double*)ptr;
  return hipSuccess;
}
}  // namespace


namespace p3x_detail {

int dev_fail(hipError_t e) { return e == hipErrorOutOfMemory ? P3X_ERR_NO_MEMORY : P3X_ERR_DEVICE; }

int status_of_err(unsigned err) {
  if (err & p3d::ERR_BAD_TEMPERATURE) return P3X_ERR_TEMPERATURE;
  if (err & p3d::ERR_SUBSTEP_LIMIT) return P3X_ERR_SUBSTEPS;
  if (err & p3d::ERR_CHECK_VALUES) return P3X_ERR_CHECK_VALUES;
  return err ? P3X_ERR_DEVICE : P3X_OK;
}

// Default launch shape per (precision, nlev), from sweeps on MI355X at 21600x72
// and 16384x128 in both precisions.  Part 1: one thread per (level, column).
// Part 2: compiled for two wavefronts per SIMD (in FP64 that spills a little,
// which costs less than the latency a single wavefront leaves exposed).
void default_shape(int precision, int nk, p3d::Shape* sh) {
  sh->cb = 32;
  sh->nt = 256;
  sh->wpe = 2;
  // Part 1 one thread per (level, column): part 1 + 2 take 267 instead of 317 us
  // (FP64) and 180 instead of 221 us (mixed) at 21600 x 72.  Every form uses it
  // unless an explicit cb x nt shape asks for the block kernel (the batch tuner
  // tries those too).  Part 3 runs in a kernel of its own in every form.
  sh->p1lvl = 1;
  // Sedimentation lanes per column, per species group (kernel trace, us):
  //                     liquid 16 / 32 / 64     ice 16 / 32 / 64
  //   FP64  21600 x 72     115 / 184 / 247     766 / 767 / 891
  //   FP64  16384 x 128    111 /  77 / 105     284 / 184 / 171
  //   mixed 21600 x 72      63 /  84 / 121     412 / 363 / 427
  //   mixed 16384 x 128     51 /  49 /  60     131 / 122 / 133
  // (at most 5 levels per lane keep two wavefronts per SIMD; the batched form
  // runs both groups on lpc, which its tuner also searches).  Since the group
  // reductions run on DPP inside a 16-lane row, 16 lanes per column pay off for
  // the FP64 ice group at up to 80 levels: whole step at 21600 x 72, four runs
  // of each build alternating, 1143 us against 1159 at 32 lanes (mixed: no
  // difference; at 128 levels 32 or 64 lanes for both groups were slower).
  sh->lpc = nk <= 80 ? 16 : (nk <= 32 * 5 ? 32 : 64);
  if (precision == P3X_MIXED) sh->lpc_ice = nk <= 32 * 5 ? 32 : 64;
  else sh->lpc_ice = nk <= 80 ? 16 : (nk <= 96 ? 32 : 64);
}

// Offsets (in elements) of the tables in one device allocation; each is a
// multiple of 4 elements, so every record starts 16-byte aligned.
namespace TableLayout {
constexpr size_t kCells = (size_t)p3c::densize * p3c::rimsize * p3c::isize;
constexpr size_t coll = 0;
constexpr size_t revap = coll + p3c::COLL_TABLE_LEN;
constexpr size_t ice_p2 = revap + p3c::RAIN_TABLE_LEN;
constexpr size_t ice_sed = ice_p2 + kCells * 8;
constexpr size_t ice_p3 = ice_sed + kCells * 4;
constexpr size_t rain_v = ice_p3 + kCells * 4;
constexpr size_t ice_sed_d = rain_v + (size_t)p3c::RAIN_TABLE_LEN * 2;
constexpr size_t total = ice_sed_d + kCells * 8;
static_assert(revap % 4 == 0 && ice_p2 % 4 == 0 && ice_sed % 4 == 0 && ice_p3 % 4 == 0 && rain_v % 4 == 0 &&
              ice_sed_d % 4 == 0, "records must stay 16-byte aligned");
}  // namespace TableLayout

template <class T>
p3d::TableSet<T> table_set(const T* base) {
  p3d::TableSet<T> s;
  s.coll = base + TableLayout::coll;
  s.revap = base + TableLayout::revap;
  s.ice_p2 = base + TableLayout::ice_p2;
  s.ice_sed = base + TableLayout::ice_sed;
  s.ice_sed_d = base + TableLayout::ice_sed_d;
  s.ice_p3 = base + TableLayout::ice_p3;
  s.rain_v = base + TableLayout::rain_v;
  return s;
}

// The device table set in type T from the reference tables: collection and
// rain-evaporation tables as they are, the ice table repacked into per-kernel
// records of the quantities each kernel interpolates, and the two rain
// fall-speed tables interleaved.
template <class T>
std::vector<T> pack_tables(const p3x_table_data* t) {
  using namespace TableLayout;
  std::vector<T> v(total, T(0));
  for (size_t j = 0; j < (size_t)p3c::COLL_TABLE_LEN; ++j) v[coll + j] = (T)t->collect_table[j];
  for (size_t j = 0; j < (size_t)p3c::RAIN_TABLE_LEN; ++j) v[revap + j] = (T)t->revap_table[j];
  for (size_t cell = 0; cell < kCells; ++cell) {
    const double* r = t->ice_table + cell * p3c::ice_table_size;
    for (int q = 0; q < 7; ++q) v[ice_p2 + cell * 8 + q] = (T)r[p3d::kIceP2Quant[q]];
    for (int q = 0; q < 4; ++q) v[ice_sed + cell * 4 + q] = (T)r[p3d::kIceSedQuant[q]];
    // the differences to the next size cell, in T as the kernels would form them
    const bool last = cell % p3c::isize == (size_t)p3c::isize - 1;
    for (int q = 0; q < 4; ++q) {
      const T a = (T)r[p3d::kIceSedQuant[q]];
      const T b = last ? a : (T)r[p3c::ice_table_size + p3d::kIceSedQuant[q]];
      v[ice_sed_d + cell * 8 + q] = a;
      v[ice_sed_d + cell * 8 + 4 + q] = (T)(b - a);
    }
    for (int q = 0; q < 4; ++q) v[ice_p3 + cell * 4 + q] = (T)r[p3d::kIceP3Quant[q]];
  }
  for (size_t j = 0; j < (size_t)p3c::RAIN_TABLE_LEN; ++j) {
    v[rain_v + 2 * j] = (T)t->vm_table[j];
    v[rain_v + 2 * j + 1] = (T)t->vn_table[j];
  }
  return v;
}

int set_geometry(int nj, int nk, int layout, int ld, int ld_iface, p3d::KParams* k) {
  if (nj < 0 || nk <= 0) return P3X_ERR_ARG;
  k->ncol = nj;
  k->nlev = nk;
  k->dbg_stride = nj;
  k->sfs = (size_t)nj * nk;
  if (layout == P3X_LEVEL_MAJOR) {
    const int l = ld > 0 ? ld : nj;
    if (l < nj || (ld_iface != 0 && ld_iface != l)) return P3X_ERR_ARG;
    k->cs = 1; k->ks = l;
    k->fcs = 1; k->fks = l;
    k->scs = 1; k->sks = nj;
    k->colmajor = 0;
    return P3X_OK;
  }
  if (layout == P3X_COLUMN_MAJOR) {
    const int l = ld > 0 ? ld : nk;
    const int lf = ld_iface > 0 ? ld_iface : std::max(l, nk + 1);
    if (l < nk || lf < nk + 1) return P3X_ERR_ARG;
    // column-major point counts stay below 2^31 (the limit the pointwise launches were built for)
    if ((long long)nj * nk >= (1ll << 31)) return P3X_ERR_ARG;
    k->cs = l; k->ks = 1;
    k->fcs = lf; k->fks = 1;
    k->scs = nk; k->sks = 1;
    k->colmajor = 1;
    return P3X_OK;
  }
  return P3X_ERR_ARG;
}

size_t level_span(const p3d::KParams& k) {
  if (k.ncol <= 0) return 0;
  return (size_t)(k.ncol - 1) * k.cs + (size_t)(k.nlev - 1) * k.ks + 1;
}

// KParams of one column set; scratch points at SC_COUNT*nj*nk doubles followed
// by nj ints.
int fill_params(const p3x_context* c, const p3x_runtime* runtime, const p3x_prognostic_state* ps,
                const p3x_diagnostic_inputs* di, const p3x_diagnostic_outputs* dout,
                const p3x_infrastructure* infra, const p3x_history_only* hist, int nj, int nk, int layout,
                int ld, int ld_iface, int* col_status, int* sed_substeps, int* ice_table_idx, double* scratch,
                p3d::KParams* out) {
  p3d::KParams& k = *out;
  k = p3d::KParams{};
  const int st = set_geometry(nj, nk, layout, ld, ld_iface, &k);
  if (st != P3X_OK) return st;
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
  k.scratch = scratch;
  k.colstat = scratch ? (int*)(scratch + (size_t)p3d::SC_COUNT * k.sfs) : nullptr;
  k.err = c->d_err;
  k.dbg_status = col_status;
  k.dbg_substeps = sed_substeps;
  k.dbg_ice_idx = ice_table_idx;
  k.tab.d = table_set(c->d_tab);
  k.tab.f = table_set(c->d_tab_f);
  static_assert(sizeof(p3d::Runtime) == sizeof(p3x_runtime), "runtime layouts differ");
  std::memcpy(&k.rt, runtime, sizeof(k.rt));
  k.dt = infra->dt;
  k.predictNc = infra->predictNc ? 1 : 0;
  k.prescribedCCN = infra->prescribedCCN ? 1 : 0;
  k.exch_zero = infra->exchange_zeroed_on_entry ? 1 : 0;
  // qv_sat_dry(Tmelt, p) = ep_2 e_melt / max(p, 1e-3) for melting and wet growth: the same
  // portable arithmetic on the host as on the device
  unsigned no_err = 0;
  k.e_melt = p3d::mk_svp(p3c::Tmelt, false, no_err);
  k.amble = p3d::Amble{};  // no fused pre/post-amble
  return P3X_OK;
}

// Every field pointer of a column set is non-null.
bool fields_present(const p3x_prognostic_state* ps, const p3x_diagnostic_inputs* di,
                    const p3x_diagnostic_outputs* dout, const p3x_history_only* hist) {
  const void* must[] = {ps->qc, ps->nc, ps->qr, ps->nr, ps->qi, ps->qm, ps->ni, ps->bm, ps->qv, ps->th,
                        di->nc_nuceat_tend, di->nccn, di->ni_activated, di->inv_qc_relvar, di->cld_frac_i,
                        di->cld_frac_l, di->cld_frac_r, di->pres, di->dz, di->dpres, di->inv_exner,
                        di->qv_prev, di->t_prev, dout->qv2qi_depos_tend, dout->precip_liq_surf,
                        dout->precip_ice_surf, dout->diag_eff_radius_qc, dout->diag_eff_radius_qi,
                        dout->diag_eff_radius_qr, dout->rho_qi, dout->precip_liq_flux, dout->precip_ice_flux,
                        dout->precip_total_tend, dout->nevapr, hist->liq_ice_exchange,
                        hist->vap_liq_exchange, hist->vap_ice_exchange};
  for (const void* p : must)
    if (!p) return false;
  return true;
}

// Bytes of scratch one column set needs (SC_COUNT dense fields of nj x nk, then
// nj column statuses), rounded up to 256.
size_t scratch_bytes_for(int nj, int nk) {
  const size_t b = (size_t)p3d::SC_COUNT * nk * nj * sizeof(double) + (size_t)nj * sizeof(int);
  return (b + 255) & ~(size_t)255;
}

}  // namespace p3x_detail

using namespace p3x_detail;

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
    case P3X_ERR_BUDGET: return "the member does not fit the server's HBM budget";
    case P3X_ERR_CHECK_VALUES: return "check_values: a temperature outside (173, 323) K after the step";
    case P3X_ERR_SCRATCH_LIMIT:
      return "p3x_main_noalloc: the column set needs more scratch than the fixed array holds "
             "(P3X_NOALLOC_SCRATCH_BYTES)";
    default: return "unknown P3 status";
  }
}

extern "C" int p3x_init(const p3x_table_data* t, p3x_context** out) {
  if (!t || !out) return P3X_ERR_ARG;
  *out = nullptr;
  p3x_context* c = new p3x_context();
  hipError_t e = hipGetDevice(&c->device);
  auto up = [&](auto** dst, const auto& host) {
    const size_t bytes = host.size() * sizeof(host[0]);
    if (e == hipSuccess) e = hipMalloc((void**)dst, bytes);
    if (e == hipSuccess) e = hipMemcpy(*dst, host.data(), bytes, hipMemcpyHostToDevice);
  };
  up(&c->d_tab, pack_tables<double>(t));
  up(&c->d_tab_f, pack_tables<float>(t));
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_err, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(c->d_err, 0, sizeof(unsigned));
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipEventCreate(&c->ev_mid);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device);
  if (e != hipSuccess) {
    p3x_finalize(c);
    return dev_fail(e);
  }
  // p3x_main_noalloc's array on this device (the device is current here); a
  // failure only disables that entry point (it then returns P3X_ERR_DEVICE)
  if (noalloc_scratch_on(c->device, &c->noalloc_scratch) != hipSuccess) {
    (void)hipGetLastError();
    c->noalloc_scratch = nullptr;
  }
  *out = c;
  return P3X_OK;
}

extern "C" void p3x_finalize(p3x_context* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  // teardown: nothing useful can be done with a failure here
  (void)hipFree(c->d_tab);
  (void)hipFree(c->d_tab_f);
  (void)hipFree(c->d_scratch);
  (void)hipFree(c->d_err);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev_mid) (void)hipEventDestroy(c->ev_mid);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  delete c;
}

// Scratch of at least need bytes; a larger allocation replaces the old one
// once the last call that used it has ended.
int p3x_detail::grow_scratch(p3x_context* c, size_t need) {
  if (need <= c->scratch_bytes) return P3X_OK;
  hipError_t e = hipStreamSynchronize(c->main_stream);
  if (e == hipSuccess && c->d_scratch) e = hipFree(c->d_scratch);
  c->d_scratch = nullptr;
  c->scratch_bytes = 0;
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_scratch, need);
  if (e != hipSuccess) return dev_fail(e);
  c->scratch_bytes = need;
  return P3X_OK;
}

extern "C" int p3x_reserve(p3x_context* c, int ld, int nk) {
  if (!c || ld <= 0 || nk <= 0 || nk > 64 * p3d::kMaxLevelsPerLane) return P3X_ERR_ARG;
  return grow_scratch(c, scratch_bytes_for(ld, nk));
}

extern "C" int p3x_sync(p3x_context* c) {
  if (!c) return P3X_ERR_ARG;
  hipError_t e = hipStreamSynchronize(c->last_stream);
  if (e == hipSuccess && c->main_launched && c->main_stream != c->last_stream)
    e = hipStreamSynchronize(c->main_stream);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  unsigned err = 0;
  e = hipMemcpy(&err, c->d_err, sizeof(unsigned), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  if (err) {
    (void)hipMemset(c->d_err, 0, sizeof(unsigned));
    return status_of_err(err);
  }
  return P3X_OK;
}

extern "C" int p3x_main(p3x_context* c, const p3x_runtime* runtime, const p3x_prognostic_state* ps,
                        const p3x_diagnostic_inputs* di, const p3x_diagnostic_outputs* dout,
                        const p3x_infrastructure* infra, const p3x_history_only* hist, int nj, int nk,
                        const p3x_launch* opts, double* elapsed_us) {
  return main_impl(c, runtime, ps, di, dout, infra, hist, nj, nk, opts, elapsed_us, nullptr);
}

extern "C" int p3x_main_noalloc(p3x_context* c, const p3x_runtime* runtime, const p3x_prognostic_state* ps,
                                const p3x_diagnostic_inputs* di, const p3x_diagnostic_outputs* dout,
                                const p3x_infrastructure* infra, const p3x_history_only* hist, int nj, int nk,
                                const p3x_launch* opts, double* elapsed_us) {
  if (!c) return P3X_ERR_ARG;
  if (!c->noalloc_scratch) return P3X_ERR_DEVICE;
  return main_impl(c, runtime, ps, di, dout, infra, hist, nj, nk, opts, elapsed_us, nullptr, c->noalloc_scratch,
                   (size_t)P3X_NOALLOC_SCRATCH_BYTES);
}

int p3x_detail::main_impl(p3x_context* c, const p3x_runtime* runtime, const p3x_prognostic_state* ps,
                          const p3x_diagnostic_inputs* di, const p3x_diagnostic_outputs* dout,
                          const p3x_infrastructure* infra, const p3x_history_only* hist, int nj, int nk,
                          const p3x_launch* opts, double* elapsed_us, const p3d::Amble* amble,
                          double* fixed_scratch, size_t fixed_bytes) {
  if (!c || !runtime || !ps || !di || !dout || !infra || !hist) return P3X_ERR_ARG;
  if (nj < 0 || nk <= 0 || !(infra->dt > 0)) return P3X_ERR_ARG;
  p3x_launch o;
  if (opts) o = *opts;
  else p3x_launch_defaults(&o);
  {
    p3d::KParams g;
    if (set_geometry(nj, nk, o.layout, o.ld, o.ld_iface, &g) != P3X_OK) return P3X_ERR_ARG;
  }
  if (o.precision != P3X_FP64 && o.precision != P3X_MIXED) return P3X_ERR_ARG;
  if (nk > 64 * p3d::kMaxLevelsPerLane) return P3X_ERR_ARG;
  p3d::Shape sh;
  {
    const int st = resolve_shape(o.precision, nk, o.columns_per_block, o.threads_per_block,
                                 o.sed_lanes_per_column, o.part12_waves_per_simd, &sh);
    if (st != P3X_OK) return st;
  }
  hipStream_t s = (hipStream_t)o.stream;
  if (nj == 0) return P3X_OK;
  if (!fields_present(ps, di, dout, hist)) return P3X_ERR_ARG;

  double* scr = fixed_scratch;
  if (scr) {
    if (scratch_bytes_for(nj, nk) > fixed_bytes) return P3X_ERR_SCRATCH_LIMIT;
  } else {
    const int st = grow_scratch(c, scratch_bytes_for(nj, nk));
    if (st != P3X_OK) return st;
    scr = c->d_scratch;
  }

  p3d::KParams k;
  {
    const int st = fill_params(c, runtime, ps, di, dout, infra, hist, nj, nk, o.layout, o.ld, o.ld_iface,
                               o.col_status, o.sed_substeps, o.ice_table_idx, scr, &k);
    if (st != P3X_OK) return st;
  }
  k.dbg_check = o.check_values;
  if (amble) k.amble = *amble;

  // Every call on this context shares its scratch: a call on another stream than
  // the previous one waits on the device for the previous one to end.
  hipError_t e = hipSuccess;
  if (c->main_launched && c->main_stream != s) {
    e = hipStreamWaitEvent(s, c->ev1, 0);
    if (e != hipSuccess) return P3X_ERR_DEVICE;
  }
  e = hipEventRecord(c->ev0, s);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  // the liquid sedimentation kernel beside the ice kernel (21600 x 72 FP64: 1255 -> 1216 us per step)
  const p3d::SedFork fork{c->side, c->ev_fork, c->ev_join};
  e = o.precision == P3X_FP64 ? p3d::launch_p3_fp64(k, sh, s, c->ev_mid, &fork)
                              : p3d::launch_p3_mixed(k, sh, s, c->ev_mid, &fork);
  if (e == hipErrorInvalidValue) return P3X_ERR_ARG;
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  e = hipEventRecord(c->ev1, s);
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  c->last_stream = s;
  c->main_stream = s;
  c->main_launched = 1;
  if (o.async) return P3X_OK;
  const int st = p3x_sync(c);
  if (elapsed_us) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) *elapsed_us = 1000.0 * ms;
  }
  return st;
}

extern "C" int p3x_last_kernel_times(p3x_context* c, double* part12_us, double* sed_us) {
  if (!c) return P3X_ERR_ARG;
  if (hipEventSynchronize(c->ev1) != hipSuccess) return P3X_ERR_DEVICE;
  float a = 0, b = 0;
  if (hipEventElapsedTime(&a, c->ev0, c->ev_mid) != hipSuccess) return P3X_ERR_DEVICE;
  if (hipEventElapsedTime(&b, c->ev_mid, c->ev1) != hipSuccess) return P3X_ERR_DEVICE;
  if (part12_us) *part12_us = 1000.0 * a;
  if (sed_us) *sed_us = 1000.0 * b;
  return P3X_OK;
}

// ============================================================ batched form

namespace p3x_detail {

int resolve_shape(int precision, int nk_max, int cb, int nt, int lpc, int wpe, p3d::Shape* sh) {
  default_shape(precision, nk_max, sh);
  if (cb != 0 || nt != 0) {  // an explicit block shape runs at one wavefront per SIMD unless told otherwise
    sh->cb = cb;
    sh->nt = nt;
    sh->wpe = 1;
    sh->p1lvl = 0;
  }
  if (lpc != 0) sh->lpc = sh->lpc_ice = lpc;
  if (wpe != 0) sh->wpe = wpe;
  if (sh->wpe != 1 && sh->wpe != 2) return P3X_ERR_ARG;
  for (int l : {sh->lpc, sh->lpc_ice}) {
    if (l != 16 && l != 32 && l != 64) return P3X_ERR_ARG;
    if (nk_max > l * p3d::kMaxLevelsPerLane) return P3X_ERR_ARG;
  }
  return P3X_OK;
}

int layout_batch(const p3d::KParams* sets, int n, const p3d::Shape& sh, Layout* out) {
  Layout& L = *out;
  L.first.assign(n + 1, 0);
  L.first2.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    L.first[i + 1] = L.first[i] + (sets[i].ncol + sh.cb - 1) / sh.cb;
    L.first2[i + 1] = L.first2[i] + p3d::pw_tiles(sets[i]);
  }
  const int cols_sed = p3d::kSedWaves * (64 / sh.lpc);
  std::vector<int> cls_of(n);
  for (int i = 0; i < n; ++i) {
    cls_of[i] = p3d::sed_levels_per_lane(sets[i].nlev, sh.lpc);
    if (cls_of[i] == 0) return P3X_ERR_ARG;
  }
  L.sed_order.clear();
  L.sed_first.clear();
  L.cls.clear();
  for (int l = 1; l <= p3d::kMaxLevelsPerLane; ++l) {
    Layout::Cls c{l, (int)L.sed_order.size(), (int)L.sed_first.size(), 0, 0};
    for (int i = 0; i < n; ++i) {
      if (cls_of[i] != l) continue;
      if (c.nsets == 0) L.sed_first.push_back(0);
      L.sed_order.push_back(i);
      c.total += (sets[i].ncol + cols_sed - 1) / cols_sed;
      L.sed_first.push_back(c.total);
      c.nsets++;
    }
    if (c.nsets > 0) L.cls.push_back(c);
  }
  return P3X_OK;
}

int grid_for(int ncu, int blocks_per_cu, int total, int resident) {
  const long g = blocks_per_cu < 0 ? total
                 : (long)ncu * (blocks_per_cu > 0 ? blocks_per_cu : std::max(1, resident));
  return (int)std::max(1L, std::min<long>(g, std::max(total, 1)));
}

hipError_t occupancy(int precision, const p3d::Shape& sh, int L, int* r1, int* r2, int* rsed) {
  return precision == P3X_FP64 ? p3d::occupancy_p3_fp64(sh, L, r1, r2, rsed)
                               : p3d::occupancy_p3_mixed(sh, L, r1, r2, rsed);
}

hipError_t launch_plan(int precision, const p3d::BatchPlan& plan, hipStream_t s) {
  return precision == P3X_FP64 ? p3d::launch_p3_batch_fp64(plan, s) : p3d::launch_p3_batch_mixed(plan, s);
}

}  // namespace p3x_detail

struct p3x_batch {
  p3x_context* ctx = nullptr;
  p3x_batch_options opt{};
  std::vector<p3d::KParams> sets;  // in the caller's order
  double* d_scratch = nullptr;
  p3d::KParams* d_sets = nullptr;  // caller's order (part 1 and part 2 launches)
  int* d_first = nullptr;          // part 1 blocks, then part 2 tiles (2 x (nsets + 1))
  p3d::KParams* d_sets_sed = nullptr;  // grouped by sedimentation class
  int* d_first_sed = nullptr;
  p3d::BatchPlan plan{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void free_plan(p3x_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  (void)hipFree(b->d_first);
  (void)hipFree(b->d_sets_sed);
  (void)hipFree(b->d_first_sed);
  b->d_first = nullptr;
  b->d_sets_sed = nullptr;
  b->d_first_sed = nullptr;
}

hipError_t launch_batch(const p3x_batch* b, hipStream_t s) { return launch_plan(b->opt.precision, b->plan, s); }

// Lay the sets out for b->opt: virtual-block offsets of both launches, the
// sedimentation classes, grid sizes and (use_graph) the graph.
int build_plan(p3x_batch* b) {
  p3x_context* c = b->ctx;
  const int n = (int)b->sets.size();
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  free_plan(b);
  int nk_max = 1;
  for (const auto& k : b->sets) nk_max = std::max(nk_max, k.nlev);
  p3d::Shape sh;
  int st = resolve_shape(b->opt.precision, nk_max, b->opt.columns_per_block, b->opt.threads_per_block,
                         b->opt.sed_lanes_per_column, b->opt.part12_waves_per_simd, &sh);
  if (st != P3X_OK) return st;
  Layout lay;
  st = layout_batch(b->sets.data(), n, sh, &lay);
  if (st != P3X_OK) return st;
  std::vector<p3d::KParams> sed_sets;
  for (int i : lay.sed_order) sed_sets.push_back(b->sets[i]);

  std::vector<int> firsts(lay.first);
  firsts.insert(firsts.end(), lay.first2.begin(), lay.first2.end());
  e = hipMalloc((void**)&b->d_first, firsts.size() * sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(b->d_first, firsts.data(), firsts.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_sets_sed, sed_sets.size() * sizeof(p3d::KParams));
  if (e == hipSuccess)
    e = hipMemcpy(b->d_sets_sed, sed_sets.data(), sed_sets.size() * sizeof(p3d::KParams), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_first_sed, lay.sed_first.size() * sizeof(int));
  if (e == hipSuccess)
    e = hipMemcpy(b->d_first_sed, lay.sed_first.data(), lay.sed_first.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return dev_fail(e);

  b->plan = p3d::BatchPlan{};
  b->plan.shape = sh;
  b->plan.sets = b->d_sets;
  b->plan.first = b->d_first;
  b->plan.first2 = b->d_first + (n + 1);
  b->plan.nsets = n;
  b->plan.nclass = (int)lay.cls.size();
  for (int ci = 0; ci < b->plan.nclass; ++ci) {
    const Layout::Cls& lc = lay.cls[ci];
    p3d::SedClass& sc = b->plan.cls[ci];
    sc.L = lc.L;
    sc.sets = b->d_sets_sed + lc.set_off;
    sc.first = b->d_first_sed + lc.first_off;
    sc.nsets = lc.nsets;
    int r1 = 1, r2 = 1, rsed = 1;
    e = occupancy(b->opt.precision, sh, sc.L, &r1, &r2, &rsed);
    if (e == hipErrorInvalidValue) return P3X_ERR_ARG;
    if (e != hipSuccess) return P3X_ERR_DEVICE;
    sc.grid = grid_for(c->ncu, b->opt.blocks_per_cu, lc.total, rsed);
    if (ci == 0) {
      b->plan.grid = grid_for(c->ncu, b->opt.blocks_per_cu, lay.first[n], r1);
      b->plan.grid2 = grid_for(c->ncu, b->opt.blocks_per_cu, lay.first2[n], r2);
      b->plan.total2 = lay.first2[n];
    }
  }

  if (b->opt.use_graph) {
    hipStream_t cs = nullptr;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      const hipError_t le = launch_batch(b, cs);
      e = hipStreamEndCapture(cs, &b->graph);
      if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
    if (e == hipErrorInvalidValue) return P3X_ERR_ARG;
    if (e != hipSuccess) return dev_fail(e);
  }
  return P3X_OK;
}

// The level-major arrays p3_main updates in place (state, and the exchange
// fields unless the set declares them zeroed on entry), for the autotuner's
// save/restore.
void inout_arrays(const p3d::KParams& k, std::vector<double*>* out) {
  double* f[] = {k.qc, k.nc, k.qr, k.nr, k.qi, k.qm, k.ni, k.bm, k.qv, k.th};
  for (double* p : f) out->push_back(p);
  if (!k.exch_zero) {
    out->push_back(k.liq_ice_exchange);
    out->push_back(k.vap_liq_exchange);
    out->push_back(k.vap_ice_exchange);
  }
}

}  // namespace

extern "C" void p3x_batch_defaults(p3x_batch_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->precision = P3X_MIXED;
  o->blocks_per_cu = -1;
  o->use_graph = 1;
}

extern "C" int p3x_batch_create(p3x_context* c, const p3x_runtime* runtime, const p3x_column_set* sets,
                                int nsets, const p3x_batch_options* opts, p3x_batch** out) {
  if (!c || !runtime || !out || nsets <= 0 || !sets) return P3X_ERR_ARG;
  *out = nullptr;
  p3x_batch_options o;
  if (opts) o = *opts;
  else p3x_batch_defaults(&o);
  if (o.precision != P3X_FP64 && o.precision != P3X_MIXED) return P3X_ERR_ARG;
  std::vector<size_t> offs(nsets);
  size_t total = 0;
  for (int i = 0; i < nsets; ++i) {
    const p3x_column_set& s = sets[i];
    p3d::KParams g;
    if (s.nj <= 0 || s.nk > 64 * p3d::kMaxLevelsPerLane) return P3X_ERR_ARG;
    if (set_geometry(s.nj, s.nk, s.layout, s.ld, s.ld_iface, &g) != P3X_OK) return P3X_ERR_ARG;
    if (!(s.infrastructure.dt > 0)) return P3X_ERR_ARG;
    if (!fields_present(&s.prognostic_state, &s.diagnostic_inputs, &s.diagnostic_outputs, &s.history_only))
      return P3X_ERR_ARG;
    offs[i] = total;
    total += scratch_bytes_for(s.nj, s.nk);
  }
  p3x_batch* b = new p3x_batch();
  b->ctx = c;
  b->opt = o;
  hipError_t e = hipMalloc((void**)&b->d_scratch, total);
  if (e == hipSuccess) {
    b->sets.resize(nsets);
    for (int i = 0; i < nsets; ++i) {
      const p3x_column_set& s = sets[i];
      fill_params(c, runtime, &s.prognostic_state, &s.diagnostic_inputs, &s.diagnostic_outputs,
                  &s.infrastructure, &s.history_only, s.nj, s.nk, s.layout, s.ld, s.ld_iface, s.col_status,
                  s.sed_substeps, s.ice_table_idx, (double*)((char*)b->d_scratch + offs[i]), &b->sets[i]);
      b->sets[i].dbg_check = s.check_values;
    }
    e = hipMalloc((void**)&b->d_sets, nsets * sizeof(p3d::KParams));
  }
  if (e == hipSuccess) e = hipMemcpy(b->d_sets, b->sets.data(), nsets * sizeof(p3d::KParams), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    p3x_batch_destroy(b);
    return dev_fail(e);
  }
  const int st = build_plan(b);
  if (st != P3X_OK) {
    p3x_batch_destroy(b);
    return st;
  }
  *out = b;
  return P3X_OK;
}

extern "C" int p3x_batch_run(p3x_batch* b, void* stream) {
  if (!b) return P3X_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const hipError_t e = b->exec ? hipGraphLaunch(b->exec, s) : launch_batch(b, s);
  if (e == hipErrorInvalidValue) return P3X_ERR_ARG;
  if (e != hipSuccess) return P3X_ERR_DEVICE;
  b->ctx->last_stream = s;
  return P3X_OK;
}

extern "C" int p3x_batch_tune(p3x_batch* b, void* stream, int reps, p3x_batch_options* chosen,
                              double* best_us) {
  if (!b || reps <= 0) return P3X_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  // save the in/out arrays so that every trial starts from the same state
  std::vector<double*> arrs;
  std::vector<size_t> bytes;
  for (const auto& k : b->sets) {
    const size_t n0 = arrs.size();
    inout_arrays(k, &arrs);
    for (size_t a = n0; a < arrs.size(); ++a) bytes.push_back(level_span(k) * sizeof(double));
  }
  size_t total = 0;
  for (size_t x : bytes) total += x;
  char* backup = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&backup, total);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  auto copy_all = [&](bool save) {
    size_t off = 0;
    for (size_t a = 0; a < arrs.size() && e == hipSuccess; ++a) {
      e = save ? hipMemcpyAsync(backup + off, arrs[a], bytes[a], hipMemcpyDeviceToDevice, s)
               : hipMemcpyAsync(arrs[a], backup + off, bytes[a], hipMemcpyDeviceToDevice, s);
      off += bytes[a];
    }
  };
  copy_all(true);
  int st = P3X_OK;
  // mean device microseconds of one step with b->opt (state restored around it)
  auto trial = [&](double* us) {
    st = build_plan(b);
    if (st != P3X_OK) return;
    copy_all(false);
    if (e == hipSuccess) st = p3x_batch_run(b, s);  // warm-up
    double sum = 0;
    for (int r = 0; r < reps && st == P3X_OK && e == hipSuccess; ++r) {
      copy_all(false);
      if (e == hipSuccess) e = hipEventRecord(e0, s);
      if (e == hipSuccess) st = p3x_batch_run(b, s);
      if (e == hipSuccess && st == P3X_OK) e = hipEventRecord(e1, s);
      if (e == hipSuccess && st == P3X_OK) e = hipEventSynchronize(e1);
      float ms = 0;
      if (e == hipSuccess && st == P3X_OK) e = hipEventElapsedTime(&ms, e0, e1);
      sum += 1000.0 * ms;
    }
    *us = sum / reps;
  };
  p3x_batch_options best = b->opt;
  double best_t = 0;
  if (e == hipSuccess) trial(&best_t);
  static const int shapes[][3] = {{16, 128, 1}, {16, 256, 1}, {32, 256, 1}, {64, 256, 1}, {32, 512, 1},
                                  {64, 512, 1},  {16, 256, 2}, {32, 256, 2}, {64, 256, 2}};
  static const int lanes[] = {16, 32, 64};
  static const int bpcs[] = {0, -1, 1, 2, 4};
  auto consider = [&](const p3x_batch_options& cand) {
    if (e != hipSuccess || st != P3X_OK) return;
    b->opt = cand;
    double t = 0;
    trial(&t);
    if (st == P3X_ERR_ARG) {  // shape not usable for these sets: skip it
      st = P3X_OK;
      return;
    }
    if (e == hipSuccess && st == P3X_OK && t < best_t) {
      best_t = t;
      best = cand;
    }
  };
  for (const auto& sh : shapes) {
    p3x_batch_options cand = best;
    cand.columns_per_block = sh[0];
    cand.threads_per_block = sh[1];
    cand.part12_waves_per_simd = sh[2];
    consider(cand);
  }
  for (int l : lanes) {
    p3x_batch_options cand = best;
    cand.sed_lanes_per_column = l;
    consider(cand);
  }
  for (int g : bpcs) {
    p3x_batch_options cand = best;
    cand.blocks_per_cu = g;
    consider(cand);
  }
  b->opt = best;
  if (e == hipSuccess && st == P3X_OK) st = build_plan(b);
  copy_all(false);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(backup);
  b->ctx->last_stream = s;
  if (e != hipSuccess) return dev_fail(e);
  if (st != P3X_OK) return st;
  if (chosen) *chosen = best;
  if (best_us) *best_us = best_t;
  return p3x_sync(b->ctx);
}

extern "C" void p3x_batch_destroy(p3x_batch* b) {
  if (!b) return;
  (void)hipDeviceSynchronize();
  free_plan(b);
  (void)hipFree(b->d_sets);
  (void)hipFree(b->d_scratch);
  delete b;
}
