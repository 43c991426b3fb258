// p3_kernel.hpp -- p3_main_init, part 1 and part 2 of a P3 step (the first
// two of its four kernels; sedimentation and part 3 are in p3_sed.hpp).
//
// Part 1 (part1_level) runs in one of two kernels:
//  * p3_part1_lvl_kernel, one thread per (level, column) on the part 2 tiles
//    below; the column's early-exit flag (nucleationPossible ||
//    hydrometeorsPresent) is the OR of its levels' flags, set with atomicOr
//    into status words p3_colstat_clear_kernel zeroed first;
//  * p3_part1_kernel, a workgroup of CB consecutive columns whose NT threads
//    split every column's levels into NCH = NT/CB contiguous chunks (thread t:
//    column t % CB, chunk t / CB), the flag an OR-reduction through LDS.
// Either way the column status goes to KParams::colstat.
//
// Part 2 (p3_part2_kernel): pointwise in the level once the column's part-1
// status is known, so one thread per (level, column): a workgroup is 64
// consecutive columns x 4 levels, one level per wavefront.  Every wavefront
// then carries the same work whatever the column count, and a grid of ncol/64
// x nlev/4 blocks keeps every CU busy to the end (one block of whole columns
// per CU left most of the chip idle in a second, partial round).  A level
// with hydrometeors after part 2 ORs bit 2 into its column's status, so the
// sedimentation kernel sees 0 (skipped after part 1), 1 (after part 2) or 3
// (full step).
//
// The step follows components/eamxx/src/physics/p3/impl/p3_main_impl.hpp:74-260.
#pragma once

#include "p3_part2.hpp"

namespace p3d {

// Scratch fields (ncol x nlev doubles each, in the layout of the column set)
// that carry per-level results of part 1 to part 2 and to the sedimentation
// kernel.
enum Scr { SC_T = 0, SC_QVSL, SC_QVSI, SC_SUPI, SC_RHOFACR, SC_RHOFACI, SC_ACN, SC_COUNT };

// The P3Microphysics pre- and post-amble (eamxx_p3_process_interface.hpp:63-261),
// fused into part 1 and part 3 when `on`.  The prognostic pointers then hold the
// field manager's wet mixing ratios: part 1 reads them wet and writes them dry,
// part 3 writes them wet again; th is the process's own array.
struct Amble {
  int on;
  const double *p_mid, *dp_wet, *cldfrac_tot;  // p_mid, pseudo_density, cldfrac_tot
  double *T_mid, *T_prev;                      // T_mid (in, out), T_prev_micro_step (out)
  double* qv_prev;                             // qv_prev_micro_step: wet in, dry in p3_main, wet out
  double *dz, *inv_exner, *cld_frac_l, *cld_frac_i, *rainfrac;  // P3 inputs part 1 writes
  double *precip_liq_mass, *precip_ice_mass;   // per column, accumulated
  double *vapor_flux, *water_flux, *ice_flux, *heat_flux;  // per column; NULL = off
};

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
  double* scratch;      // SC_COUNT fields of sfs doubles
  int* colstat;         // per column: 0 skipped after part 1, 1 after part 2, 3 full step
  unsigned* err;        // device error word
  int *dbg_status, *dbg_substeps, *dbg_ice_idx;
  int* dbg_check;       // per column: check_values flags (NULL = the check is off)
  Tables tab;
  Runtime rt;
  int ncol, nlev;
  // Layout.  Element (column i, level k) of a level field is at i*cs + k*ks, of
  // an interface field (nlev + 1 levels) at i*fcs + k*fks: level-major
  // (cs = fcs = 1, ks = fks = the row stride) or column-major, the reference's
  // (column, level) LayoutRight views (ks = fks = 1, cs and fcs the column
  // strides).  Scratch field f holds (i, k) at f*sfs + i*scs + k*sks, dense, in
  // the same order.
  int cs, ks, fcs, fks;
  int scs, sks;
  size_t sfs;
  int colmajor;         // 1: column-major; the pointwise kernels then map lanes to levels
  int dbg_stride;       // species stride of dbg_substeps (>= ncol)
  double dt;
  int predictNc, prescribedCCN, exch_zero;
  double e_melt;        // saturation vapour pressure over liquid at Tmelt (qv_sat_dry(Tmelt, p) numerator)
  Amble amble;

  __host__ __device__ __forceinline__ size_t at(int i, int k) const { return (size_t)i * cs + (size_t)k * ks; }
  __host__ __device__ __forceinline__ size_t atf(int i, int k) const { return (size_t)i * fcs + (size_t)k * fks; }
  __host__ __device__ __forceinline__ double& sc(int f, int i, int k) const {
    return scratch[(size_t)f * sfs + (size_t)i * scs + (size_t)k * sks];
  }
};

__host__ __device__ __forceinline__ double exner_function(double pres) {  // PhysicsFunctions, p^(Rd/cp)
  return p3m_pow(pres / P0, Rair * INV_CP);
}

constexpr int kMaxSubsteps = 100000;
// Column status bits: part 1 found work, part 2 left hydrometeors.
constexpr int kStatusPart1 = 1, kStatusPart2 = 2, kStatusFull = 3;

// p3_main_init + p3_main_part1 (P3I/p3_main_impl_part1.hpp:18-185) for level k of
// column i; returns its flags: bit 0 nucleationPossible, bit 1 hydrometeorsPresent.
template <class R>
__device__ __forceinline__ int part1_level(const KParams& p, int k, int i, unsigned& err) {
  const double dt = p.dt;
  const double latvap = LatVap, latice = LatIce, inv_cp = INV_CP;
  const Runtime& rt = p.rt;
  auto scr = [&](int f, int kk) -> double& { return p.sc(f, i, kk); };
  const size_t x = p.at(i, k);
  int flags = 0;
  double inv_exner, th, qv, qc, nc, qr, nr, qi, ni, qm, bm, dz, cfl;
  if (p.amble.on) {
    // p3_preamble, eamxx_p3_process_interface.hpp:63-124
    const Amble& a = p.amble;
    const double pmid = a.p_mid[x], T = a.T_mid[x], cft = a.cldfrac_tot[x];
    const double dp = a.dp_wet[x], dp_dry = p.dpres[x];
    const double c1 = -1.0 + 1.0 / ep_2;  // calculate_dz with the wet qv's virtual temperature
    const double T_virtual = T * (1.0 + c1 * p.qv[x]);
    dz = (Rair / gravit) * dp * T_virtual / pmid;
    a.dz[x] = dz;
    // wet -> dry mixing ratios (calculate_drymmr_from_wetmmr_dp_based)
    qc = p.qc[x] * dp / dp_dry; nc = p.nc[x] * dp / dp_dry; qr = p.qr[x] * dp / dp_dry;
    nr = p.nr[x] * dp / dp_dry; qi = p.qi[x] * dp / dp_dry; ni = p.ni[x] * dp / dp_dry;
    qm = p.qm[x] * dp / dp_dry; bm = p.bm[x] * dp / dp_dry;
    qv = p.qv[x] * dp / dp_dry;
    a.qv_prev[x] = a.qv_prev[x] * dp / dp_dry;
    inv_exner = 1.0 / exner_function(pmid);
    a.inv_exner[x] = inv_exner;
    th = T / exner_function(pmid);
    // cloud fractions with the 1e-4 floor and the maximum-overlap rain fraction
    cfl = cft > 0.0001 ? cft : 0.0001;
    a.cld_frac_l[x] = cfl;
    a.cld_frac_i[x] = cfl;
    double cfr = cfl;
    if (k != 0) {
      const double above = a.cldfrac_tot[x - p.ks];
      cfr = above > cfr ? above : cfr;
    }
    a.rainfrac[x] = cfr;
  } else {
    inv_exner = p.inv_exner[x];
    th = p.th[x];
    qv = p.qv[x];
    qc = p.qc[x]; nc = p.nc[x]; qr = p.qr[x]; nr = p.nr[x];
    qi = p.qi[x]; ni = p.ni[x]; qm = p.qm[x]; bm = p.bm[x];
    dz = p.dz[x];
    cfl = p.prescribedCCN ? p.cld_frac_l[x] : 1.0;
  }
  // the state as it is in memory: a field part 1 leaves bit for bit unchanged
  // is not stored again (with the pre-amble every field changes to dry)
  const double in[10] = {qc, nc, qr, nr, qi, ni, qm, bm, qv, th};
  qv = vmax(qv, 0.0);
  const double exner = 1 / inv_exner;
  const double pres = p.pres[x];
  double T_atm = th * exner;
  const double rho = p.dpres[x] / dz / gravit;
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
    if (p.prescribedCCN) nc = vmax(nc, rt.spa_ccn_to_nc_factor * p.nccn[x] / (1 / cfl));
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

  {
    double* const f[10] = {p.qc, p.nc, p.qr, p.nr, p.qi, p.ni, p.qm, p.bm, p.qv, p.th};
    const double out[10] = {qc, nc, qr, nr, qi, ni, qm, bm, qv, th};
#pragma unroll
    for (int q = 0; q < 10; ++q)
      if (p.amble.on || __double_as_longlong(out[q]) != __double_as_longlong(in[q])) f[q][x] = out[q];
  }
  scr(SC_T, k) = T_atm;
  scr(SC_QVSL, k) = qv_sat_l;
  scr(SC_QVSI, k) = qv_sat_i;
  scr(SC_SUPI, k) = qv_supersat_i;
  scr(SC_RHOFACR, k) = rhofacr;
  scr(SC_RHOFACI, k) = rhofaci;
  scr(SC_ACN, k) = acn;
  return flags;
}

// One block's work: columns bid*CB .. bid*CB+CB-1 of the column set p.
// s_i is an LDS array of NT ints.
template <class R, int CB, int NT>
__device__ __forceinline__ void part1_block(const KParams& p, int bid, int* s_i) {
  constexpr int NCH = NT / CB;
  static_assert(NT % CB == 0, "block must hold whole chunk rows");

  const int tid = threadIdx.x;
  const int c = tid % CB;
  const int ch = tid / CB;
  const int i = bid * CB + c;
  const bool col_ok = i < p.ncol;
  const int nlev = p.nlev;
  const int L = (nlev + NCH - 1) / NCH;
  const int k0 = ch * L < nlev ? ch * L : nlev;
  const int k1 = k0 + L < nlev ? k0 + L : nlev;
  unsigned err = 0;

  auto red_or = [&](int v) {
    s_i[tid] = v;
    __syncthreads();
    int r = 0;
    for (int j = 0; j < NCH; ++j) r |= s_i[j * CB + c];
    __syncthreads();
    return r;
  };

  int flags = 0;  // bit0 nucleationPossible, bit1 hydrometeorsPresent
  if (col_ok) {
    for (int k = k0; k < k1; ++k) flags |= part1_level<R>(p, k, i, err);
  }
  const int cflags = red_or(flags);
  if (col_ok && ch == 0) p.colstat[i] = (cflags & 3) ? kStatusPart1 : 0;
  if (col_ok && ch == 0 && p.dbg_check) p.dbg_check[i] = 0;
  if (err) atomicOr(p.err, err);
}

template <class R, int CB, int NT>
__global__ __launch_bounds__(NT) void p3_part1_kernel(const KParams p) {
  __shared__ int s_i[NT];
  part1_block<R, CB, NT>(p, blockIdx.x, s_i);
}

// Tiles of the pointwise kernels (part 1, part 2, part 3): 256 threads, one
// (level, column) point each, arranged so that a wavefront's loads coalesce:
//  * level-major: 64 consecutive columns x 4 levels, one level per wavefront,
//    columns fastest, so a wavefront reads 64 consecutive doubles of a row
//    (and holds points of one level, whose work is alike);
//  * column-major, two arrangements:
//    - PW_RUNS (part 1, part 3, memory-bound): 256 consecutive points of the
//      column-by-column order, so a wavefront reads 512 contiguous bytes;
//    - PW_BLOCKS (part 2): a wavefront holds 8 columns x 8 consecutive levels
//      (levels fastest, 8 runs of 64 bytes per load), a block 32 columns x 8
//      levels, so that a wavefront's points are of nearby levels: the upper
//      levels are mostly dry and skip part 2's processes, and a wavefront
//      spanning a whole column would run them for its few wet lanes (part 2
//      330 -> 270 us at 21600 x 72 against PW_RUNS).
//    Both take the same grid (the larger of their tile counts); tiles past an
//    arrangement's end hold no points.
constexpr int kPart2Cols = 64, kPart2Levels = 4, kTileThreads = kPart2Cols * kPart2Levels;
constexpr int kBlkCols = 32, kBlkLevels = 8;
enum PwMap { PW_RUNS = 0, PW_BLOCKS = 1 };
__host__ __device__ __forceinline__ int part2_tiles(int ncol, int nlev) {
  return ((ncol + kPart2Cols - 1) / kPart2Cols) * ((nlev + kPart2Levels - 1) / kPart2Levels);
}
__host__ __device__ __forceinline__ int pw_tiles(const KParams& p) {
  if (p.colmajor) {
    const int runs = (int)(((long long)p.ncol * p.nlev + kTileThreads - 1) / kTileThreads);
    const int blocks = ((p.ncol + kBlkCols - 1) / kBlkCols) * ((p.nlev + kBlkLevels - 1) / kBlkLevels);
    return runs > blocks ? runs : blocks;
  }
  return part2_tiles(p.ncol, p.nlev);
}
// The point of this thread in tile `tile`; false past the end of the set.
// (Column-major point counts are below 2^31: the host refuses larger sets.)
template <int MAP>
__device__ __forceinline__ bool pw_point(const KParams& p, int tile, int& i, int& k) {
  if (p.colmajor) {
    if constexpr (MAP == PW_RUNS) {
      const unsigned e = (unsigned)tile * kTileThreads + threadIdx.x;
      const unsigned nl = (unsigned)p.nlev;
      i = (int)(e / nl);
      k = (int)(e - (unsigned)i * nl);
      return i < p.ncol;
    } else {
      const int ntc = (p.ncol + kBlkCols - 1) / kBlkCols;
      const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
      i = (tile % ntc) * kBlkCols + w * 8 + (l >> 3);
      k = (tile / ntc) * kBlkLevels + (l & 7);
      return i < p.ncol && k < p.nlev;
    }
  }
  const int ntc = (p.ncol + kPart2Cols - 1) / kPart2Cols;
  i = (tile % ntc) * kPart2Cols + (threadIdx.x & 63);
  k = (tile / ntc) * kPart2Levels + (threadIdx.x >> 6);
  return i < p.ncol && k < p.nlev;
}

// ORs `bit` into the status word of column i for every lane whose `set` is
// true, one atomic per column and wavefront.  Level-major lanes hold distinct
// columns.  Column-major lanes hold runs of one column's consecutive levels:
// under PW_RUNS lane l's run starts min(l, k) lanes below it, under PW_BLOCKS
// the runs are the aligned groups of 8 lanes; the lowest lane of a run with
// `set` makes the atomic.  Every lane of the wavefront that is still running
// must call this.
template <int MAP>
__device__ __forceinline__ void colstat_or(const KParams& p, int i, int k, bool set, int bit) {
  if (!p.colmajor) {
    if (set) atomicOr(&p.colstat[i], bit);
    return;
  }
  const unsigned long long m = __ballot(set);
  const int l = threadIdx.x & 63;
  unsigned long long run_below;
  if constexpr (MAP == PW_RUNS) {
    const int below = k < l ? k : l;
    run_below = ((1ull << l) - 1) ^ ((1ull << (l - below)) - 1);
  } else {
    run_below = (0xFFull << (l & ~7)) & ((1ull << l) - 1);
  }
  if (set && !(m & run_below)) atomicOr(&p.colstat[i], bit);
}

// Part 1 with one thread per (level, column), on the pointwise tiles: a
// column's status is the OR of its levels' flags, set atomically in a word that
// p3_colstat_clear_kernel zeroed before.  Same per-level arithmetic as
// part1_block; many more wavefronts, so the chip stays busy.
template <class R>
__device__ __forceinline__ void part1_tile(const KParams& p, int tile) {
  int i, k;
  if (!pw_point<PW_RUNS>(p, tile, i, k)) return;
  unsigned err = 0;
  const int flags = part1_level<R>(p, k, i, err);
  colstat_or<PW_RUNS>(p, i, k, (flags & 3) != 0, kStatusPart1);
  if (k == 0 && p.dbg_check) p.dbg_check[i] = 0;  // part 3 ORs the column's check_values flags
  if (err) atomicOr(p.err, err);
}

template <class R>
__global__ __launch_bounds__(kPart2Cols * kPart2Levels) void p3_part1_lvl_kernel(const KParams p) {
  part1_tile<R>(p, blockIdx.x);
}

// Zeroes the column status words before p3_part1_lvl_kernel (templated only so
// that each precision's translation unit has its own copy).
template <class R>
__global__ __launch_bounds__(256) void p3_colstat_clear_kernel(int* colstat, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) colstat[i] = 0;
}

// ================= p3_main_part2 (P3I/p3_main_impl_part2.hpp:113-524) for the
// level of this wavefront and the column of this lane in tile `tile`.
template <class R>
__device__ __forceinline__ void part2_tile(const KParams& p, int tile) {
  int i, k;
  if (!pw_point<PW_BLOCKS>(p, tile, i, k)) return;
  const size_t x = p.at(i, k);
  auto scr = [&](int f) -> double { return p.sc(f, i, k); };
  unsigned err = 0;
  bool done = false, hydro = false;
  // The column status and the level's skip-test inputs are loaded together (one
  // memory round trip for the decision); bit 1 only: other levels of the column
  // may already have set bit 2.
  {
    Lvl s;
    const int cstat = p.colstat[i];
    s.qc = p.qc[x]; s.qr = p.qr[x]; s.qi = p.qi[x];
    s.T_atm = scr(SC_T);
    s.qv_supersat_i = scr(SC_SUPI);
    const bool skip_all = s.qc < QSMALL && s.qr < QSMALL && s.qi < QSMALL && s.T_atm < T_zerodegc &&
                          s.qv_supersat_i < -0.05;
    if ((cstat & kStatusPart1) && !skip_all) {
      s.nc = p.nc[x]; s.nr = p.nr[x]; s.qm = p.qm[x]; s.ni = p.ni[x]; s.bm = p.bm[x];
      s.qv = p.qv[x]; s.th = p.th[x];
      s.pres = p.pres[x];
      s.rho = p.dpres[x] / p.dz[x] / gravit;
      s.inv_rho = 1 / s.rho;
      s.qv_sat_l = scr(SC_QVSL);
      s.qv_sat_i = scr(SC_QVSI);
      s.rhofaci = scr(SC_RHOFACI);
      s.acn = scr(SC_ACN);
      s.cfl = p.cld_frac_l[x]; s.cfr = p.cld_frac_r[x]; s.cfi = p.cld_frac_i[x];
      s.icfl = 1 / s.cfl; s.icfr = 1 / s.cfr; s.icfi = 1 / s.cfi;
      s.inv_exner = p.inv_exner[x];
      s.qv_prev = p.qv_prev[x]; s.t_prev = p.t_prev[x];
      s.ni_activated = p.ni_activated[x];
      const double in[10] = {s.qc, s.nc, s.qr, s.nr, s.qi, s.ni, s.qm, s.bm, s.qv, s.th};
      // hydrometeorsPresent after part 2 (P3I/p3_main_impl.hpp:255-260): a column flag
      hydro = part2_level<R>(p.rt, p.tab, p.dt, p.predictNc, p.prescribedCCN, p.e_melt, s, err);
      {
        // mixed form: a state field whose bits part 2 left unchanged is not stored
        // again (in FP64 the ten held originals cost more registers than the
        // stores they save)
        double* const f[10] = {p.qc, p.nc, p.qr, p.nr, p.qi, p.ni, p.qm, p.bm, p.qv, p.th};
        const double out[10] = {s.qc, s.nc, s.qr, s.nr, s.qi, s.ni, s.qm, s.bm, s.qv, s.th};
#pragma unroll
        for (int q = 0; q < 10; ++q)
          if (sizeof(R) == 8 || __double_as_longlong(out[q]) != __double_as_longlong(in[q])) f[q][x] = out[q];
      }
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
  // hydrometeorsPresent after part 2 (P3I/p3_main_impl.hpp:255-260): a column flag
  colstat_or<PW_BLOCKS>(p, i, k, hydro, kStatusPart2);
  if (err) atomicOr(p.err, err);
}

// WPE: minimum wavefronts per SIMD the register allocation must allow (1 or 2).
template <class R, int WPE>
__global__ __launch_bounds__(kPart2Cols * kPart2Levels, WPE) void p3_part2_kernel(const KParams p) {
  part2_tile<R>(p, blockIdx.x);
}

}  // namespace p3d
