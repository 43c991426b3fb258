// p3_sed.hpp -- second kernel of a P3 step: cloud, rain and ice sedimentation,
// homogeneous freezing and p3_main_part3, with the surface precipitation.
//
// Work decomposition: one column per group of LPC lanes (LPC = 16, 32 or 64,
// aligned inside the 64-lane wavefront); lane j of a group owns the L
// contiguous levels j*L .. j*L+L-1 of its column and keeps their
// sedimentation state in registers for the whole substep loop.  Everything a
// substep needs from other levels goes through cross-lane operations inside
// the group: the CFL maximum and find_top/find_bottom are butterfly
// reductions, and the upwind flux entering a lane's first level is the last
// level's flux of the lane above (one shuffle).  A column therefore runs
// exactly its own number of substeps, without block barriers, and reads its
// state from HBM once per species.
//
// Follows components/eamxx/src/physics/p3/impl/p3_main_impl.hpp:262-312
// (sedimentation, homogeneous freezing, part 3), with
// p3_cloud_sed_impl.hpp:14-147, p3_rain_sed_impl.hpp:14-170,
// p3_ice_sed_impl.hpp:66-270, p3_upwind_impl.hpp:14-151 and
// p3_main_impl_part3.hpp:18-204.
#pragma once

#include "p3_kernel.hpp"

namespace p3d {

// Reductions over aligned groups of LPC lanes, all lanes getting the result.
// Step o combines each lane with a lane of the other half of its 2o-lane
// block; before the step both halves are uniform, so any pairing of the halves
// does: inside a 16-lane row by DPP (o = 1, 2: quad permutes; 4, 8: the
// half-row and row mirrors), across rows by ds_swizzle (o = 16, xor inside 32
// lanes) and across the wavefront's halves by ds_bpermute (o = 32).  A lane
// reads only lanes of its own group, and every lane of a group must be active
// (a DPP read from an inactive lane returns the reader's own value): the
// sedimentation code calls these with whole groups, never with part of one.
template <int O>
__device__ __forceinline__ int partner(int v) {
  if constexpr (O == 1) return __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
  else if constexpr (O == 2) return __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
  else if constexpr (O == 4) return __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
  else if constexpr (O == 8) return __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false);  // row_mirror
  else if constexpr (O == 16) return __builtin_amdgcn_ds_swizzle(v, 0x401F);  // lane ^ 16 within 32
  else return __shfl_xor(v, 32, 64);
}
template <int O>
__device__ __forceinline__ double partner(double v) {
  return __hiloint2double(partner<O>(__double2hiint(v)), partner<O>(__double2loint(v)));
}

template <int LPC>
__device__ __forceinline__ int grp_min(int v) {
  if constexpr (LPC > 1) {
    v = grp_min<LPC / 2>(v);
    v = min(v, partner<LPC / 2>(v));
  }
  return v;
}
template <int LPC>
__device__ __forceinline__ int grp_max(int v) {
  if constexpr (LPC > 1) {
    v = grp_max<LPC / 2>(v);
    v = max(v, partner<LPC / 2>(v));
  }
  return v;
}
template <int LPC>
__device__ __forceinline__ double grp_maxd(double v) {
  if constexpr (LPC > 1) {
    v = grp_maxd<LPC / 2>(v);
    const double w = partner<LPC / 2>(v);
    v = w > v ? w : v;
  }
  return v;
}
// (Used where one lane of the group holds the value and the others 0, so the
// order of the additions does not matter.)
template <int LPC>
__device__ __forceinline__ double grp_sumd(double v) {
  if constexpr (LPC > 1) {
    v = grp_sumd<LPC / 2>(v);
    v += partner<LPC / 2>(v);
  }
  return v;
}

// Sedimentation runs as two kernels over the same column blocks, one per
// species group: the liquid group (cloud, then rain, with the rain flux and
// the liquid surface precipitation) and the ice group.  The groups touch
// disjoint state fields, so their order does not change a bit; apart, each
// kernel holds only its own group's registers and field pointers.
enum SedGroup { SED_LIQ = 0, SED_ICE = 1 };

// Wavefronts per SIMD the sedimentation kernels are compiled for: up to 5
// levels per lane two fit (256 registers); the FP64 ice group at up to 3
// levels per lane runs at three (168 registers and a 76-byte spill at 3 levels:
// 736 -> 714 us at 21600 x 72; the FP32 ice group holds 127 registers anyway).
template <class R, int G, int L> constexpr int sed_waves_per_simd() {
  return (G == SED_ICE && sizeof(R) == 8 && L <= 3) ? 3 : (L <= 5 ? 2 : 1);
}

// The ice group interpolates its 4 table quantities from an LDS copy of the
// ice-sedimentation records.  Where the blocks the kernel's occupancy puts on
// one CU (blocks of 4 wavefronts) can each hold it in the CU's 160 KB of LDS,
// the copy is of the records that also carry each quantity's difference to
// the next size cell (64 KB in FP64, 32 KB in FP32), which saves the four
// size-direction subtractions per quantity; elsewhere (the FP64 kernels at
// three wavefronts per SIMD) it is of the plain records (32 KB).
constexpr int kIceCells = densize * rimsize * isize;
template <class R, int L> constexpr bool ice_rec_dif() {
  return sed_waves_per_simd<R, SED_ICE, L>() * kIceCells * 8 * (int)sizeof(R) <= 160 * 1024;
}
template <class R, int L> constexpr int ice_lds_len() { return kIceCells * (ice_rec_dif<R, L>() ? 8 : 4); }

// Per-lane geometry of the owned levels.
template <int L>
struct Geo {
  double rho[L], inv_rho[L], inv_dz[L];
};

// One species' sedimentation over the group's column (SP = 0 cloud, 1 rain,
// 2 ice), with lane j of the group owning the L levels k_top + j*L ..
// k_top + j*L + L-1 from the species' top level k_top down: sedimentation only
// moves mass down, so the levels above k_top take no part, and spreading only
// [k_top, nlev) over the lanes keeps them busy (rain, below the freezing
// level, fills a quarter of a column).  The state of one column's run, held
// in registers: load() reads the species' fields in that mapping, substep()
// is one CFL substep, store() writes the fields back (the rain run also
// stores the rain flux at the interface below each owned level).  Every lane
// of the group calls each of them together (the CFL maximum, find_bottom and
// the upwind flux are cross-lane operations inside the group).  DIF: the ice
// records in ice_tab carry their size-direction differences (the layout the
// kernel staged, chosen by its own levels per lane, not by this run's L).
template <int SP, class R, int LPC, int L, bool DIF>
struct SedRun {
  static constexpr int NF = SP == 2 ? 4 : 2;  // fields held (cloud always holds nc: its DSD rewrites it)
  Geo<L> g;
  double q[NF][L], cf[L], coef[L], cfl_ice[L], pflux[L];
  int kbase, k_top, k_bot;
  double dt_left, prt_accum;  // prt_accum: mass through the surface, on the lane owning the lowest level
  int nsub;

  __device__ __forceinline__ static void fields(const KParams& p, double* (&fld)[NF]) {
    if constexpr (SP == 0) { fld[0] = p.qc; fld[1] = p.nc; }
    if constexpr (SP == 1) { fld[0] = p.qr; fld[1] = p.nr; }
    if constexpr (SP == 2) { fld[0] = p.qi; fld[1] = p.ni; fld[2] = p.qm; fld[3] = p.bm; }
  }

  __device__ __forceinline__ void load(const KParams& p, int i, int j, int top) {
    const int nlev = p.nlev;
    double* fld[NF];
    fields(p, fld);
    const double* cfp = SP == 0 ? p.cld_frac_l : (SP == 1 ? p.cld_frac_r : p.cld_frac_i);
    const int coef_f = SP == 0 ? SC_ACN : (SP == 1 ? SC_RHOFACR : SC_RHOFACI);
    k_top = top;
    kbase = k_top + j * L;
    int my_bot = -1;
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = kbase + t;
      const bool ok = k < nlev;
      const size_t x = p.at(i, ok ? k : 0);
      const double dz = ok ? p.dz[x] : 1.0;
      const double dp = ok ? p.dpres[x] : 1.0;
      g.rho[t] = dp / dz / gravit;
      g.inv_rho[t] = 1 / g.rho[t];
      g.inv_dz[t] = 1 / dz;
#pragma unroll
      for (int f = 0; f < NF; ++f) q[f][t] = ok ? fld[f][x] : 0.0;
      cf[t] = ok ? cfp[x] : 1.0;
      coef[t] = ok ? p.sc(coef_f, i, ok ? k : 0) : 0.0;
      cfl_ice[t] = (SP == 2 && ok) ? p.cld_frac_l[x] : 1.0;
      pflux[t] = 0;
      // find_bottom, P3I/p3_find_impl.hpp:52-92 (find_top gave k_top)
      if (ok && q[0][t] >= QSMALL) my_bot = k;
    }
    k_bot = grp_max<LPC>(my_bot);
    dt_left = p.dt;
    prt_accum = 0;
    nsub = 0;
  }

  // One substep; `first` (the first substep: in-cloud values limited) is a
  // constant after inlining wherever the caller knows it.
  __device__ __forceinline__ void substep(const KParams& p, int j, bool first, const R* __restrict__ ice_tab) {
    const int nlev = p.nlev, kbot = nlev - 1;
    const Runtime& rt = p.rt;
    const int nfield = SP == 2 ? 4 : (SP == 1 ? 2 : (p.predictNc ? 2 : 1));
    double V1[L], V2[L];
    double co = 0;
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = kbase + t;
      V1[t] = 0;
      V2[t] = 0;
      if (k >= k_top && k <= k_bot) {  // k_bot < nlev, so k is a real level
        const double rho = g.rho[t];
        if constexpr (SP == 0) {
          // cloud_sedimentation, P3I/p3_cloud_sed_impl.hpp:85-112
          // in-cloud values in the arithmetic type of the form (the state stays FP64)
          const R cfl = R(cf[t]);
          const R qc = R(q[0][t]), nc = R(q[1][t]);
          R qc_incld, nc_incld;
          if (first) {
            const R icfl = R(1) / cfl;
            if (q[0][t] >= QSMALL) { qc_incld = vmin(qc * icfl, R(incloud_limit)); nc_incld = vmax(nc * icfl, R(0)); }
            else { qc_incld = 0; nc_incld = 0; }
          } else {
            qc_incld = qc / cfl;
            nc_incld = nc / cfl;
          }
          if (qc_incld > R(QSMALL)) {
            R ncr = nc_incld, mu_c, lamc, cdist, cdist1;
            get_cloud_dsd2<R>(qc_incld, ncr, mu_c, R(rho), lamc, cdist, cdist1);
            q[1][t] = double(ncr) * cf[t];
            const R dum = R(1) / m_pow(lamc, R(bcn));
            const R acn = R(coef[t]);
            V1[t] = double(acn * m_tgamma(R(4 + bcn) + mu_c) * dum / m_tgamma(mu_c + R(4)));
            if (p.predictNc) V2[t] = double(acn * m_tgamma(R(1 + bcn) + mu_c) * dum / m_tgamma(mu_c + R(1)));
          }
        } else if constexpr (SP == 1) {
          // rain_sedimentation + compute_rain_fall_velocity, P3I/p3_rain_sed_impl.hpp:14-34,104-126
          const R cfr = R(cf[t]);
          const R qr = R(q[0][t]), nr = R(q[1][t]);
          R qr_incld, nr_incld;
          if (first) {
            const R icfr = R(1) / cfr;
            if (q[0][t] >= QSMALL) { qr_incld = vmin(qr * icfr, R(precip_limit)); nr_incld = vmax(nr * icfr, R(0)); }
            else { qr_incld = 0; nr_incld = 0; }
          } else {
            qr_incld = qr / cfr;
            nr_incld = nr / cfr;
          }
          if (qr_incld > R(QSMALL)) {
            R nrr = nr_incld, mu_r = 0, lamr;
            get_rain_dsd2<R>(rt, qr_incld, nrr, mu_r, lamr);
            const Tab3<R> t3 = lookup3<R>(mu_r, lamr);
            const R rhofacr = R(coef[t]);
            R v[2];  // vm, vn
            apply_table3_v<R>(TabOf<R>::get(p.tab).rain_v, t3, v);
            V1[t] = double(v[0] * rhofacr);
            V2[t] = double(v[1] * rhofacr);
            q[1][t] = double(nrr) * cf[t];
          }
        } else {
          // ice_sedimentation, P3I/p3_ice_sed_impl.hpp:134-166
          const R cfi = R(cf[t]);
          const R qi = R(q[0][t]), ni = R(q[1][t]), qm = R(q[2][t]), bm = R(q[3][t]);
          R qi_incld, ni_incld, qm_incld, bm_incld;
          if (first) {
            const R icfi = R(1) / cfi, icfl = R(1) / R(cfl_ice[t]);
            if (q[0][t] >= QSMALL) { qi_incld = vmin(qi * icfi, R(incloud_limit)); ni_incld = vmax(ni * icfi, R(0)); }
            else { qi_incld = 0; ni_incld = 0; }
            if (q[0][t] >= QSMALL && q[2][t] >= QSMALL) { qm_incld = qm * icfi; bm_incld = vmin(vmax(bm * icfl, R(0)), R(incloud_limit)); }
            else { qm_incld = 0; bm_incld = 0; }
          } else {
            qi_incld = qi / cfi;
            ni_incld = ni / cfi;
            qm_incld = qm / cfi;
            bm_incld = bm / cfi;
          }
          if (qi_incld > R(QSMALL)) {
            R nir = vmax(ni_incld, R(NSMALL));
            R qmr = qm_incld, bmr = bm_incld;
            const R qir = qi_incld;
            const R rhop = calc_bulk_rho_rime<R>(rt, qir, qmr, bmr);
            q[2][t] = double(qmr) * cf[t];
            q[3][t] = double(bmr) * cf[t];
            const TabIce<R> ti = lookup_ice<R>(qir, nir, qmr, rhop);
            // quantities 0, 1, 6, 7 of the records (in FP64 as two halves, which bounds the registers)
            R v[4];
            if constexpr (sizeof(R) == 4) {
              if constexpr (DIF) apply_table_ice_dif<4, 8, 4, R>(ice_tab, ti, v, 0);
              else apply_table_ice_rec<4, 4, R>(ice_tab, ti, v, 0);
            } else {
              R a[2], b[2];
              if constexpr (DIF) {
                apply_table_ice_dif<2, 8, 4, R>(ice_tab, ti, a, 0);
                apply_table_ice_dif<2, 8, 4, R>(ice_tab, ti, b, 2);
              } else {
                apply_table_ice_rec<2, 4, R>(ice_tab, ti, a, 0);
                apply_table_ice_rec<2, 4, R>(ice_tab, ti, b, 2);
              }
              v[0] = a[0]; v[1] = a[1]; v[2] = b[0]; v[3] = b[1];
            }
            const R t_ni_fall = v[0], t_qi_fall = v[1], t_lammax = v[2], t_lammin = v[3];
            nir = vmin(nir, t_lammax * nir);
            nir = vmax(nir, t_lammin * nir);
            q[1][t] = double(nir) * cf[t];
            const R rhofaci = R(coef[t]);
            V1[t] = double(R(rt.ice_sedimentation_factor) * t_qi_fall * rhofaci);
            V2[t] = double(R(rt.ice_sedimentation_factor) * t_ni_fall * rhofaci);
          }
        }
        const double Co = V1[t] * dt_left * g.inv_dz[t];
        if (Co > co) co = Co;
      }
    }
    const double Co_max = grp_maxd<LPC>(co);

    // generalized_sedimentation + calc_first_order_upwind_step, P3I/p3_upwind_impl.hpp:14-129
    const int tmpint1 = (int)(Co_max + 1);
    const double dt_sub = dt_left / tmpint1;
    const int k_temp = (k_bot == kbot) ? k_bot : k_bot + 1;
    // fluxes F = V r rho.  Only the levels in [k_top, k_temp] update, and
    // they read F of their own level and of the level above, which is in
    // that range too (or is k_top's upper boundary, flux 0): F elsewhere is
    // never read, so it is not masked.
    double F[NF][L];
#pragma unroll
    for (int t = 0; t < L; ++t) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double V = (SP == 2) ? (f == 1 ? V2[t] : V1[t]) : (f == 0 ? V1[t] : V2[t]);
        F[f][t] = V * q[f][t] * g.rho[t];
      }
    }
    // flux out of the last level of the lane above (lane 0's first level is
    // k_top, whose incoming flux is 0)
    double Fin[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) Fin[f] = __shfl_up(F[f][L - 1], 1, LPC);
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = kbase + t;
      if (k >= k_top && k <= k_temp) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          if (f < nfield) {
            // level kbase + t is k_top exactly when t = 0 and j = 0
            const double Fa = t == 0 ? (j == 0 ? 0.0 : Fin[f]) : F[f][t > 0 ? t - 1 : 0];
            const double fluxdiv = (Fa - F[f][t]) * g.inv_dz[t];
            q[f][t] = q[f][t] + fluxdiv * dt_sub * g.inv_rho[t];
          }
        }
        if (SP == 1) pflux[t] += F[0][t];  // rain_sed :140-153, interface k+1
        if (k == kbot && k_bot == kbot) prt_accum += F[0][t] * dt_sub;
      }
    }
    if (k_bot != kbot) k_bot += 1;
    dt_left -= dt_sub;
    nsub++;
  }

  // The substeps changed levels k_top .. k_bot only (k_bot has moved down one
  // level after each substep, past the last one's lowest updated level): the
  // levels below keep their loaded values and are not stored again.
  __device__ __forceinline__ void store(const KParams& p, int i) const {
    const int nlev = p.nlev;
    double* fld[NF];
    fields(p, fld);
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = kbase + t;
      if (k < nlev) {
        const size_t x = p.at(i, k);
        if (k <= k_bot) {
#pragma unroll
          for (int f = 0; f < NF; ++f) fld[f][x] = q[f][t];
        }
        if (SP == 1) p.precip_liq_flux[p.atf(i, k + 1)] = pflux[t];  // rain_sed :140-153, interface k+1
      }
    }
  }
};

// One species' run over the group's column, whole: load, the substeps (the
// first peeled off the loop, so that neither carries the other's branch),
// store.  prt (every lane) is the mass that left through the surface.
template <int SP, class R, int LPC, int L, bool DIF>
__device__ __forceinline__ void sediment_run(const KParams& p, int i, int j, int k_top, double& prt,
                                          int& nsub_out, unsigned& err, const R* __restrict__ ice_tab) {
  SedRun<SP, R, LPC, L, DIF> r;
  r.load(p, i, j, k_top);
  if (r.dt_left > dt_left_tol) r.substep(p, j, true, ice_tab);
#pragma unroll 1
  while (r.dt_left > dt_left_tol) {
    // The substep count is bounded by the CFL number (a few hundred at most for
    // physical inputs); the cap only guarantees termination for corrupt inputs.
    if (r.nsub >= kMaxSubsteps) { err |= ERR_SUBSTEP_LIMIT; break; }
    r.substep(p, j, false, ice_tab);
  }
  r.store(p, i);
  prt = grp_sumd<LPC>(r.prt_accum);  // only the lane owning kbot accumulated
  nsub_out = r.nsub;
}

// find_top (P3I/p3_find_impl.hpp:14-50) of NS species (mixing ratios f[s]) over
// the group's column in the standard mapping (lane j: levels j*L .. j*L+L-1),
// all loads in one pass; top[s] = nlev when species s is absent (or the column
// does not run).
template <int LPC, int L, int NS>
__device__ __forceinline__ void find_tops(const KParams& p, int i, int j, bool run,
                                          const double* const (&f)[NS], int (&top)[NS]) {
  const int nlev = p.nlev;
  int my[NS];
#pragma unroll
  for (int sp = 0; sp < NS; ++sp) my[sp] = nlev;
  if (run) {
    double q[NS][L];
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = j * L + t;
      const size_t x = p.at(i, k < nlev ? k : 0);
#pragma unroll
      for (int sp = 0; sp < NS; ++sp) q[sp][t] = f[sp][x];
    }
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
      const int k = j * L + t;
#pragma unroll
      for (int sp = 0; sp < NS; ++sp)
        if (k < nlev && q[sp][t] >= QSMALL) my[sp] = k;
    }
  }
#pragma unroll
  for (int sp = 0; sp < NS; ++sp) top[sp] = grp_min<LPC>(my[sp]);
}

// One species' substeps on the fewest levels per lane that cover [k_top, nlev).
// Every lane of the wavefront calls this (k_top = nlev for a lane whose column
// does not run or has none of the species).
template <int SP, class R, int LPC, int L>
__device__ __forceinline__ void sediment(const KParams& p, int i, int j, int k_top, double& prt, int& nsub_out,
                                         unsigned& err, const R* __restrict__ ice_tab = nullptr) {
  const int nlev = p.nlev;
  prt = 0;
  nsub_out = 0;
  // levels per lane for [k_top, nlev), the largest over the wavefront's
  // columns (all its lanes are here), so that the wavefront takes one path
  int lw = k_top < nlev ? (nlev - k_top + LPC - 1) / LPC : 0;
#pragma unroll
  for (int o = 32; o >= LPC; o >>= 1) lw = max(lw, __shfl_xor(lw, o, 64));
  if (k_top >= nlev) return;
  // the ice records' layout is the one the kernel staged for its own L
  constexpr bool dif = ice_rec_dif<R, L>();
  if (lw <= 1) sediment_run<SP, R, LPC, 1, dif>(p, i, j, k_top, prt, nsub_out, err, ice_tab);
  else if (L > 2 && lw <= 2) sediment_run<SP, R, LPC, (L > 2 ? 2 : L), dif>(p, i, j, k_top, prt, nsub_out, err, ice_tab);
  else sediment_run<SP, R, LPC, L, dif>(p, i, j, k_top, prt, nsub_out, err, ice_tab);
}

// homogeneous_freezing (P3I/p3_ice_sed_impl.hpp:207-270) + p3_main_part3
// (P3I/p3_main_impl_part3.hpp:60-199, only the fields p3_main outputs) for
// level k of column i, after sedimentation.
template <class R>
__device__ __forceinline__ void part3_level(const KParams& p, int k, int i, bool active, unsigned& err) {
  const int nlev = p.nlev;
  const size_t x = p.at(i, k);
  const double latvap = LatVap, latice = LatIce, inv_cp = INV_CP;
  const Runtime& rt = p.rt;
  p.precip_ice_flux[p.atf(i, k)] = 0;
  if (k == nlev - 1) p.precip_ice_flux[p.atf(i, nlev)] = 0;
  double reff_c = 10.e-6, reff_i = 25.e-6, reff_r = 500.e-6, rho_qi = 0;
  double qc = 0, nc = 0, qr = 0, nr = 0, qi = 0, ni = 0, qm = 0, bm = 0, qv = 0, th = 0;
  if (active || p.amble.on) {
    qc = p.qc[x]; nc = p.nc[x]; qr = p.qr[x]; nr = p.nr[x]; qi = p.qi[x]; ni = p.ni[x];
    qm = p.qm[x]; bm = p.bm[x]; qv = p.qv[x]; th = p.th[x];
  }
  // the state as it is in memory: without the post-amble, a field that part 3
  // leaves bit for bit unchanged is not stored again
  const double in[10] = {qc, nc, qr, nr, qi, ni, qm, bm, qv, th};
  if (active) {
    const double inv_exner = p.inv_exner[x];
    if (rt.do_ice_production) {
      const bool t_lt = p.sc(SC_T, i, k) < T_homogfrz;
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
        if (p.dbg_ice_idx) p.dbg_ice_idx[x] = ti.dumjj * 10000 + ti.dumii * 100 + ti.dumi;
        R v[4];  // quantities 5, 6, 7, 11
        apply_table_ice_rec<4, 4, R>(TabOf<R>::get(p.tab).ice_p3, ti, v);
        const R t_eff = v[0], t_lammax = v[1], t_lammin = v[2], t_rhomm = v[3];
        ni_incld = vmin(ni_incld, double(t_lammax) * ni_incld);
        ni_incld = vmax(ni_incld, double(t_lammin) * ni_incld);
        reff_i = double(t_eff);
        rho_qi = double(t_rhomm);
      } else if (p.dbg_ice_idx) {
        p.dbg_ice_idx[x] = -1;
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
    p.vap_liq_exchange[x] = vap_liq;
    if (p.dbg_check) {
      // check_values (P3I/p3_check_values_impl.hpp:25-79, called after part 3 at
      // P3I/p3_main_impl.hpp:316-324 in debug builds) on T = th exner and qv: T outside
      // (173, 323) K is the trap (ERR_CHECK_VALUES), qv outside (0, 40e-3) only a flag
      const double T = th * (1 / inv_exner);
      const int f = (!(T > 173. && T < 323.) ? 1 : 0) | (!(qv > 0. && qv < 40.e-3) ? 2 : 0);
      if (f) atomicOr(&p.dbg_check[i], f);
      if (f & 1) err |= ERR_CHECK_VALUES;
    }
  } else if (p.dbg_ice_idx) {
    p.dbg_ice_idx[x] = -1;
  }
  if (p.amble.on) {
    // p3_postamble, eamxx_p3_process_interface.hpp:200-241: the temperature change
    // rescaled by dp_dry/dp, dry -> wet mixing ratios, radii in microns
    const Amble& a = p.amble;
    const double dp = a.dp_wet[x], dp_dry = p.dpres[x];
    const double T_before = a.T_mid[x];
    double T = (th * exner_function(a.p_mid[x]) - T_before) * dp_dry / dp;
    T += T_before;
    a.T_mid[x] = T;
    a.T_prev[x] = T;
    qv = qv * dp_dry / dp;
    p.qc[x] = qc * dp_dry / dp; p.nc[x] = nc * dp_dry / dp; p.qr[x] = qr * dp_dry / dp;
    p.nr[x] = nr * dp_dry / dp; p.qi[x] = qi * dp_dry / dp; p.ni[x] = ni * dp_dry / dp;
    p.qm[x] = qm * dp_dry / dp; p.bm[x] = bm * dp_dry / dp; p.qv[x] = qv;
    a.qv_prev[x] = qv;
    reff_c *= 1e6;
    reff_i *= 1e6;
    reff_r *= 1e6;
  } else if (active) {
    double* const f[10] = {p.qc, p.nc, p.qr, p.nr, p.qi, p.ni, p.qm, p.bm, p.qv, p.th};
    const double out[10] = {qc, nc, qr, nr, qi, ni, qm, bm, qv, th};
#pragma unroll
    for (int q = 0; q < 10; ++q)
      if (__double_as_longlong(out[q]) != __double_as_longlong(in[q])) f[q][x] = out[q];
  }
  p.diag_eff_radius_qc[x] = reff_c;
  p.diag_eff_radius_qi[x] = reff_i;
  p.diag_eff_radius_qr[x] = reff_r;
  p.rho_qi[x] = rho_qi;
}

// The ice group's per-column outputs once column i's run has ended (any one
// lane of the group writes them): the surface precipitation, its post-amble
// accumulation (eamxx_p3_process_interface.hpp:245-258) and the substep count.
__device__ __forceinline__ void ice_column_out(const KParams& p, int i, bool ran, double prt, int nsub) {
  double ice = 0;
  if (ran) ice += prt * INV_RHO_H2O * (1 / p.dt);
  p.precip_ice_surf[i] = ice;
  if (p.amble.on) p.amble.precip_ice_mass[i] += ice * RHO_H2O * p.dt;
  if (p.dbg_substeps) p.dbg_substeps[(size_t)2 * p.dbg_stride + i] = nsub;
}

// One wavefront's work for group G: the 64/LPC columns starting at column
// col0, with LPC lanes per column.
template <class R, int LPC, int L, int G>
__device__ __forceinline__ void sed_wave(const KParams& p, int col0, const R* __restrict__ ice_tab) {
  static_assert(64 % LPC == 0, "a column group must divide the wavefront");
  const int lane = threadIdx.x & 63;
  const int j = lane % LPC;
  const int i = col0 + lane / LPC;
  const int nlev = p.nlev, kbot = nlev - 1;
  const int kbase = j * L;
  unsigned err = 0;

  // Every lane of the wavefront takes part in sediment() (its wavefront-wide
  // choice of levels per lane); columns past ncol or without a full step only
  // skip the arithmetic.
  const bool live = i < p.ncol;
  const int status = live ? p.colstat[i] : 0;
  const bool run = live && status == kStatusFull;
  const double inv_dt = 1 / p.dt;
  if constexpr (G == SED_LIQ) {
    double prt_c, prt_r;
    int ns_c, ns_r, top[2];
    find_tops<LPC, L, 2>(p, i, j, run, {p.qc, p.qr}, top);
    const int top_c = top[0], top_r = top[1];
    sediment<0, R, LPC, L>(p, i, j, top_c, prt_c, ns_c, err);
    sediment<1, R, LPC, L>(p, i, j, top_r, prt_r, ns_r, err);
    if (live) {
#pragma unroll
      for (int t = 0; t < L; ++t) {
        const int k = kbase + t;
        if (k >= nlev) continue;
        // no rain flux above the rain's top level (the rain run stored the rest)
        if (k < top_r) p.precip_liq_flux[p.atf(i, k + 1)] = 0;
        if (k == kbot) {
          double liq = 0;
          if (top_c < nlev) liq = prt_c * INV_RHO_H2O * inv_dt;
          if (top_r < nlev) liq += prt_r * INV_RHO_H2O * inv_dt;
          p.precip_liq_surf[i] = liq;
          // the per-column post-amble, eamxx_p3_process_interface.hpp:245-258 (the
          // boundary fluxes, which need both groups, are written by part 3)
          if (p.amble.on) p.amble.precip_liq_mass[i] += liq * RHO_H2O * p.dt;
        }
      }
      if (j == 0) {
        p.precip_liq_flux[p.atf(i, 0)] = 0;  // the top interface
        // 0 skipped after part 1, 1 skipped after part 2, 2 full step
        if (p.dbg_status) p.dbg_status[i] = status == kStatusFull ? 2 : status;
        if (p.dbg_substeps) {
          p.dbg_substeps[i] = ns_c;
          p.dbg_substeps[(size_t)p.dbg_stride + i] = ns_r;
        }
      }
    }
  } else {
    double prt_i;
    int ns_i, top[1];
    find_tops<LPC, L, 1>(p, i, j, run, {p.qi}, top);
    const int top_i = top[0];
    sediment<2, R, LPC, L>(p, i, j, top_i, prt_i, ns_i, err, ice_tab);
    if (live && j == 0) ice_column_out(p, i, top_i < nlev, prt_i, ns_i);
  }
  if (err) atomicOr(p.err, err);
}

// One block's work for group G: WPB wavefronts of 64/LPC columns each, the
// CB = WPB * 64/LPC columns starting at column bid * CB of the column set p,
// with LPC lanes per column (part 3 follows in p3_part3_lvl_kernel).
template <class R, int LPC, int L, int WPB, int G>
__device__ __forceinline__ void sed_block(const KParams& p, int bid, const R* __restrict__ ice_tab) {
  constexpr int CPW = 64 / LPC;  // columns per wavefront
  constexpr int CB = WPB * CPW;  // columns per block
  static_assert((CB & (CB - 1)) == 0, "columns per block must be a power of two");
  const int wave = threadIdx.x >> 6;
  sed_wave<R, LPC, L, G>(p, bid * CB + wave * CPW, ice_tab);
}

// Part 3 as a kernel of its own, one thread per (level, column) on the
// pointwise tiles, after the sedimentation kernel.
template <class R>
__device__ __forceinline__ void part3_tile(const KParams& p, int tile) {
  int i, k;
  if (!pw_point<PW_RUNS>(p, tile, i, k)) return;
  unsigned err = 0;
  part3_level<R>(p, k, i, p.colstat[i] == kStatusFull, err);
  if (k == 0 && p.amble.on && p.amble.water_flux) {
    // the post-amble's boundary fluxes (eamxx_p3_process_interface.hpp:245-258),
    // from the surface precipitation both sedimentation kernels wrote
    const Amble& a = p.amble;
    const double liq = p.precip_liq_surf[i], ice = p.precip_ice_surf[i];
    a.vapor_flux[i] = 0.0;
    a.water_flux[i] = liq + ice;
    a.ice_flux[i] = ice;
    a.heat_flux[i] = 0.0;
  }
  if (err) atomicOr(p.err, err);
}

template <class R>
__global__ __launch_bounds__(kTileThreads) void p3_part3_lvl_kernel(const KParams p) {
  part3_tile<R>(p, blockIdx.x);
}

// Workgroups go to the 8 XCDs round-robin, and each XCD has its own L2.  A
// block's lanes read a few columns of many levels, so neighbouring blocks share
// every cache line they touch: give each XCD a contiguous range of blocks, so
// that the blocks sharing lines run on one XCD at about the same time.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds, n = b / kXcds;
  return x * q + (x < r ? x : r) + n;
}

// Copies the ice-sedimentation records into the block's LDS copy (16-byte
// chunks), then waits for the whole block.
template <class R, int LEN>
__device__ __forceinline__ void stage_ice_table(const R* __restrict__ g, R* s) {
  constexpr int n = LEN * (int)sizeof(R) / 16;
  static_assert(LEN * sizeof(R) % 16 == 0, "16-byte chunks");
  for (int c = threadIdx.x; c < n; c += blockDim.x)
    reinterpret_cast<uint4*>(s)[c] = reinterpret_cast<const uint4*>(g)[c];
  __syncthreads();
}

// The ice table's records in the form the ice group of L levels per lane reads.
template <class R, int L>
__device__ __forceinline__ const R* ice_lds_source(const KParams& p) {
  return ice_rec_dif<R, L>() ? TabOf<R>::get(p.tab).ice_sed_d : TabOf<R>::get(p.tab).ice_sed;
}

// The ice group interpolates the ice table from an LDS copy of its records,
// staged once per block, rather than through the L1/L2 caches: 16 gathers per
// level and substep become LDS reads (ice sedimentation 3-13 % faster at
// 21600 x 72 and 16384 x 128 with the plain records).
template <class R, int LPC, int L, int WPB, int G>
__global__ __launch_bounds__(64 * WPB, (sed_waves_per_simd<R, G, L>())) void p3_sed_kernel(const KParams p) {
  if constexpr (G == SED_ICE) {
    static_assert(WPB == 4 || !ice_rec_dif<R, L>(), "the LDS budget assumes blocks of 4 wavefronts");
    __shared__ __align__(16) R s_tab[ice_lds_len<R, L>()];
    stage_ice_table<R, ice_lds_len<R, L>()>(ice_lds_source<R, L>(p), s_tab);
    sed_block<R, LPC, L, WPB, G>(p, xcd_block(blockIdx.x, gridDim.x), s_tab);
  } else {
    sed_block<R, LPC, L, WPB, G>(p, xcd_block(blockIdx.x, gridDim.x), nullptr);
  }
}

// Both species groups in one launch, for shapes at which they run on the same
// lanes and levels per lane: blocks
// [0, nb_ice) run the ice group and the blocks after them the liquid group,
// over the same column blocks.  Workgroups are dispatched in order, so the
// short liquid blocks take the CUs that the long ice blocks free at the end of
// the ice work, the longest work first.  Two kernels side by side on two
// streams interleave their blocks from the start instead, and the liquid
// blocks then delay the ice blocks; nor is there a fork or a join between
// streams.  Each block runs the code of the two-kernel form, so the results
// are the same bits.  The kernel is compiled for the ice group's wavefronts
// per SIMD; the shapes it is launched at are those where the liquid group's
// code also fits that register budget without a spill (in FP64 not at 3
// levels per lane, where the ice group runs three wavefronts and the liquid
// group needs 173 VGPRs).
template <class R, int LPC, int L, int WPB>
__global__ __launch_bounds__(64 * WPB, (sed_waves_per_simd<R, SED_ICE, L>())) void p3_sed_fused_kernel(
    const KParams p, int nb_ice) {
  static_assert(WPB == 4 || !ice_rec_dif<R, L>(), "the LDS budget assumes blocks of 4 wavefronts");
  __shared__ __align__(16) R s_tab[ice_lds_len<R, L>()];
  const int b = blockIdx.x;  // uniform over the block: the whole block takes one branch
  if (b < nb_ice) {
    stage_ice_table<R, ice_lds_len<R, L>()>(ice_lds_source<R, L>(p), s_tab);
    sed_block<R, LPC, L, WPB, SED_ICE>(p, xcd_block(b, nb_ice), s_tab);
  } else {
    // block b runs on XCD b % 8; the liquid blocks' offset nb_ice only rotates
    // which XCD receives which contiguous range of column blocks
    sed_block<R, LPC, L, WPB, SED_LIQ>(p, xcd_block(b - nb_ice, (int)gridDim.x - nb_ice), nullptr);
  }
}

}  // namespace p3d
