"""Seeded synthetic P3 columns (level-major SoA, element (k, i) at k*ncol + i).

The distributions are the ones the benchmark configurations name:

* pressure: interfaces log-spaced from 10 hPa (top, k = 0) to 1000 hPa (surface);
  midpoints are the log-midpoints; ``dpres`` is the interface difference;
* temperature: 300 K at the surface, 6.5 K/km lapse rate, floored at 200 K,
  plus N(0, 1 K) noise; heights from hydrostatic integration of the noise-free
  profile;
* RH ~ U(0, 1.05); ``qv = RH * qv_sat_dry(T, p)`` (Murphy-Koop over liquid,
  components/eamxx/src/physics/share/physics_saturation_impl.hpp:38-157);
* qc ~ LogNormal(median 2e-4, sigma 1) where RH > 0.98 and T > 253 K, nc = 1e8/kg
  there; qr ~ LogNormal(median 1e-5) below the freezing level (noise-free
  T > 273.15 K), nr = 1e5/kg there; qi ~ LogNormal(median 5e-5) where T < 268 K,
  ni = 1e5/kg, qm = 0.3 qi, bm = qm / 500;
* cloud fraction ``clip((RH - 0.8) / 0.2, 0, 1)``; cld_frac_l = cld_frac_i =
  max(cf, 1e-4); cld_frac_r from the maximum-overlap rule of the P3 pre-amble
  (components/eamxx/src/physics/p3/eamxx_p3_process_interface.hpp:108-126);
* inv_exner = (1e5 / p)^(Rd/cp); th = T inv_exner; qv_prev = qv; t_prev = T;
  inv_qc_relvar = 1; nc_nuceat_tend = 0; ni_activated = 0; nccn = 1e8/kg.

Everything is FP64 numpy; the generator is deterministic for a given seed.
"""
from __future__ import annotations

import numpy as np

RAIR = 287.042
CPAIR = 1004.64
GRAVIT = 9.80616
TMELT = 273.15
EP_2 = 18.016 / 28.966

STATE_FIELDS = ("qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm", "qv", "th")
INPUT_FIELDS = ("nc_nuceat_tend", "nccn", "ni_activated", "inv_qc_relvar", "cld_frac_i",
                "cld_frac_l", "cld_frac_r", "pres", "dz", "dpres", "inv_exner", "qv_prev",
                "t_prev")
OUTPUT_FIELDS = ("qv2qi_depos_tend", "diag_eff_radius_qc", "diag_eff_radius_qi",
                 "diag_eff_radius_qr", "rho_qi", "precip_total_tend", "nevapr")
FLUX_FIELDS = ("precip_liq_flux", "precip_ice_flux")
SURF_FIELDS = ("precip_liq_surf", "precip_ice_surf")
HIST_FIELDS = ("liq_ice_exchange", "vap_liq_exchange", "vap_ice_exchange")


def murphy_koop_liquid(t: np.ndarray) -> np.ndarray:
    """Saturation vapour pressure over liquid [Pa], Murphy & Koop (2005) eq. 10."""
    logt = np.log(t)
    return np.exp(54.842763 - (6763.22 / t) - (4.210 * logt) + (0.000367 * t) +
                  (np.tanh(0.0415 * (t - 218.8)) *
                   (53.878 - (1331.22 / t) - (9.44523 * logt) + 0.014025 * t)))


def pressure_levels(nlev: int):
    """Interface and midpoint pressures [Pa], top (10 hPa) to surface (1000 hPa)."""
    p_int = np.exp(np.linspace(np.log(1000.0), np.log(100000.0), nlev + 1))
    p_mid = np.sqrt(p_int[:-1] * p_int[1:])
    return p_int, p_mid


def base_temperature(nlev: int):
    """Noise-free temperature and height at the midpoints (hydrostatic, lapse 6.5 K/km)."""
    p_int, p_mid = pressure_levels(nlev)
    z_int = np.zeros(nlev + 1)
    t_mid = np.zeros(nlev)
    z_mid = np.zeros(nlev)
    for k in range(nlev - 1, -1, -1):  # integrate upward from the surface interface
        z_lo = z_int[k + 1]
        # half-layer to the midpoint, then the other half to the upper interface
        t_lo = max(300.0 - 6.5e-3 * z_lo, 200.0)
        z_m = z_lo + RAIR / GRAVIT * t_lo * np.log(p_int[k + 1] / p_mid[k])
        t_m = max(300.0 - 6.5e-3 * z_m, 200.0)
        z_int[k] = z_m + RAIR / GRAVIT * t_m * np.log(p_mid[k] / p_int[k])
        t_mid[k], z_mid[k] = t_m, z_m
    return t_mid, z_mid


def murphy_koop_ice(t: np.ndarray) -> np.ndarray:
    """Saturation vapour pressure over ice [Pa], Murphy & Koop (2005) eq. 7."""
    return np.exp(9.550426 - (5723.265 / t) + (3.53068 * np.log(t)) - (0.00728332 * t))


def make_columns(ncol: int, nlev: int, seed: int = 0, sigma: float = 1.0,
                 dry_every: int = 0, evap_every: int = 0) -> dict:
    """All p3_main input arrays for ``ncol`` columns, each shaped (nlev, ncol).

    ``dry_every = n`` makes every n-th column (0, n, 2n, ...) free of
    hydrometeors at half of ice/liquid saturation, so p3_main skips it after
    part 1.  ``evap_every = m`` makes every m-th column (offset 1) the same but
    with a trace of rain (2e-14 kg/kg) at its lowest level, which evaporates
    completely in part 2, so p3_main skips the column after part 2.
    """
    rng = np.random.default_rng(seed)
    shape = (nlev, ncol)
    p_int, p_mid = pressure_levels(nlev)
    t_base, _ = base_temperature(nlev)
    pres = np.broadcast_to(p_mid[:, None], shape).copy()
    dpres = np.broadcast_to(np.diff(p_int)[:, None], shape).copy()
    T = t_base[:, None] + rng.standard_normal(shape)
    rh = rng.uniform(0.0, 1.05, shape)
    qv = rh * (EP_2 * murphy_koop_liquid(T) / np.maximum(pres, 1e-3))

    def lognormal(median):
        return median * np.exp(sigma * rng.standard_normal(shape))

    qc = np.where((rh > 0.98) & (T > 253.0), lognormal(2e-4), 0.0)
    nc = np.where(qc > 0, 1e8, 0.0)
    below_fl = np.broadcast_to((t_base > TMELT)[:, None], shape)
    qr = np.where(below_fl, lognormal(1e-5), 0.0)
    nr = np.where(qr > 0, 1e5, 0.0)
    qi = np.where(T < 268.0, lognormal(5e-5), 0.0)
    ni = np.where(qi > 0, 1e5, 0.0)
    qm = 0.3 * qi
    bm = qm / 500.0

    special = np.zeros(ncol, bool)
    if dry_every:
        special[::dry_every] = True
    evap = np.zeros(ncol, bool)
    if evap_every:
        evap[1::evap_every] = True
        special |= evap
    if special.any():
        es = np.where(T < TMELT, murphy_koop_ice(T), murphy_koop_liquid(T))
        qv_dry = 0.5 * EP_2 * es / np.maximum(pres, 1e-3)
        for a in (qc, nc, qr, nr, qi, ni, qm, bm):
            a[:, special] = 0.0
        qv[:, special] = qv_dry[:, special]
        rh[:, special] = 0.5
        qr[nlev - 1, evap] = 2e-14
        nr[nlev - 1, evap] = 1e3

    cf = np.clip((rh - 0.8) / 0.2, 0.0, 1.0)
    cld_l = np.maximum(cf, 1e-4)
    cld_i = cld_l.copy()
    cld_r = cld_l.copy()
    cld_r[1:] = np.where(cf[:-1] > cld_r[1:], cf[:-1], cld_r[1:])

    inv_exner = (1e5 / pres) ** (RAIR / CPAIR)
    th = T * inv_exner
    dz = RAIR / GRAVIT * dpres * T / pres

    f = dict(qc=qc, nc=nc, qr=qr, nr=nr, qi=qi, qm=qm, ni=ni, bm=bm, qv=qv, th=th,
             nc_nuceat_tend=np.zeros(shape), nccn=np.full(shape, 1e8),
             ni_activated=np.zeros(shape), inv_qc_relvar=np.ones(shape),
             cld_frac_i=cld_i, cld_frac_l=cld_l, cld_frac_r=cld_r, pres=pres, dz=dz,
             dpres=dpres, inv_exner=inv_exner, qv_prev=qv.copy(), t_prev=T.copy())
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in f.items()}


def empty_outputs(ncol: int, nlev: int) -> dict:
    """Zeroed output and history arrays (the caller zeroes the exchange fields)."""
    out = {k: np.zeros((nlev, ncol)) for k in OUTPUT_FIELDS + HIST_FIELDS}
    out.update({k: np.zeros((nlev + 1, ncol)) for k in FLUX_FIELDS})
    out.update({k: np.zeros(ncol) for k in SURF_FIELDS})
    return out


ATM_IN_FIELDS = ("p_mid", "p_dry_mid", "pseudo_density", "pseudo_density_dry", "cldfrac_tot",
                 "nc_nuceat_tend", "nccn", "ni_activated", "inv_qc_relvar")
ATM_UPDATED_FIELDS = ("T_mid", "qv", "qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm",
                      "qv_prev_micro_step", "T_prev_micro_step", "precip_liq_surf_mass",
                      "precip_ice_surf_mass")
ATM_COMPUTED_FIELDS = ("eff_radius_qc", "eff_radius_qi", "eff_radius_qr", "precip_total_tend",
                       "nevapr", "rainfrac", "micro_liq_ice_exchange", "micro_vap_liq_exchange",
                       "micro_vap_ice_exchange")
ATM_FLUX_FIELDS = ("vapor_flux", "water_flux", "ice_flux", "heat_flux")


def make_atm_fields(ncol: int, nlev: int, seed: int = 0, **kw) -> dict:
    """Field-manager fields of the P3Microphysics process (wet mixing ratios) from
    the same generator as make_columns: its mixing ratios are taken as wet;
    pseudo_density is its dpres and pseudo_density_dry = pseudo_density (1 - qv);
    p_mid is its pres and p_dry_mid = p_mid (1 - qv); cldfrac_tot is its cloud
    fraction before the 1e-4 floor; T_mid = th / inv_exner; the accumulated
    precipitation masses start at 0."""
    c = make_columns(ncol, nlev, seed=seed, **kw)
    T = c["th"] / c["inv_exner"]
    qv = c["qv"]
    cf = np.where(c["cld_frac_l"] <= 1e-4, 0.0, c["cld_frac_l"])
    f = dict(p_mid=c["pres"], p_dry_mid=c["pres"] * (1 - qv), pseudo_density=c["dpres"],
             pseudo_density_dry=c["dpres"] * (1 - qv), cldfrac_tot=cf,
             nc_nuceat_tend=c["nc_nuceat_tend"], nccn=c["nccn"], ni_activated=c["ni_activated"],
             inv_qc_relvar=c["inv_qc_relvar"], T_mid=T, qv=qv, qc=c["qc"], nc=c["nc"], qr=c["qr"],
             nr=c["nr"], qi=c["qi"], qm=c["qm"], ni=c["ni"], bm=c["bm"],
             qv_prev_micro_step=qv.copy(), T_prev_micro_step=T.copy(),
             precip_liq_surf_mass=np.zeros(ncol), precip_ice_surf_mass=np.zeros(ncol))
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in f.items()}
