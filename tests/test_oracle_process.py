"""The oracle's P3Microphysics pre-amble and post-amble (CPU).

Pinned by what the mathematics fixes:
* the pre-amble followed by the post-amble with no microphysics in between
  returns the wet state (round trip), T unchanged, qv_prev = qv;
* exner(P0) = 1, so at p = P0 inv_exner = 1 and theta = T;
* dz with qv = 0 is the dry hydrostatic thickness Rd/g dp T / p;
* the rain fraction is the maximum-overlap rule, checked by brute force;
* a whole process step closes the wet column water budget: the water stepped
  by p3_main on the dry grid is the same water in wet terms;
* the accumulated surface masses and the boundary fluxes follow from the
  surface fluxes.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from scream_amd.inputs import make_atm_fields

RD, G, CP, EP_2 = 287.042, 9.80616, 1004.64, 18.016 / 28.966


def _atm(f):
    nlev, ncol = f["qc"].shape
    a = {k: f[k].copy() for k in ("p_mid", "pseudo_density", "pseudo_density_dry", "cldfrac_tot", "T_mid",
                                   "qv", "qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm")}
    a["qv_prev"] = f["qv_prev_micro_step"].copy()
    a["T_prev"] = f["T_prev_micro_step"].copy()
    for k in oracle.ATM_LOCAL + ("eff_radius_qc", "eff_radius_qi", "eff_radius_qr"):
        a[k] = np.zeros((nlev, ncol))
    for k in ("precip_liq_surf", "precip_ice_surf", "precip_liq_surf_mass", "precip_ice_surf_mass"):
        a[k] = np.zeros(ncol)
    return a


def test_round_trip_without_microphysics():
    lib = oracle.load_oracle()
    f = make_atm_fields(40, 72, seed=3)
    a = _atm(f)
    lib.p3o_preamble(C.byref(oracle._atm_struct(a)), 40, 72)
    lib.p3o_postamble(C.byref(oracle._atm_struct(a)), 40, 72, 300.0)
    for k in ("qv", "qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm"):
        np.testing.assert_allclose(a[k], f[k], rtol=4e-16, atol=0, err_msg=k)
    np.testing.assert_allclose(a["T_mid"], f["T_mid"], rtol=1e-14)
    np.testing.assert_array_equal(a["T_prev"], a["T_mid"])
    np.testing.assert_array_equal(a["qv_prev"], a["qv"])


def test_exner_reference_pressure_and_dry_dz():
    lib = oracle.load_oracle()
    f = make_atm_fields(8, 72, seed=4)
    f["p_mid"][:] = 1e5
    f["qv"][:] = 0.0
    a = _atm(f)
    lib.p3o_preamble(C.byref(oracle._atm_struct(a)), 8, 72)
    assert np.all(a["inv_exner"] == 1.0)
    np.testing.assert_array_equal(a["th_atm"], f["T_mid"])
    np.testing.assert_allclose(a["dz"], RD / G * f["pseudo_density"] * f["T_mid"] / 1e5, rtol=1e-15)


def test_wet_dz_uses_virtual_temperature():
    """With moisture, dz is the hypsometric thickness of the virtual temperature
    T_v = T (1 + (Mdry/Mw - 1) qv), the textbook form with the constants' molecular
    weights (18.016 and 28.966): the ratio dz(qv) / dz(0) - 1 is exactly that
    coefficient times qv, so a wrong coefficient or the (1 + qv/eps)/(1 + qv) form
    (2e-4 away at qv = 0.02) fails."""
    lib = oracle.load_oracle()
    f = make_atm_fields(16, 72, seed=7)
    rng = np.random.default_rng(7)
    f["qv"] = rng.uniform(1e-4, 3e-2, f["qv"].shape)
    a = _atm(f)
    lib.p3o_preamble(C.byref(oracle._atm_struct(a)), 16, 72)
    c1 = 28.966 / 18.016 - 1.0
    want = RD / G * f["pseudo_density"] * (f["T_mid"] * (1.0 + c1 * f["qv"])) / f["p_mid"]
    np.testing.assert_allclose(a["dz"], want, rtol=2e-15, atol=0)
    f0 = dict(f)
    f0["qv"] = np.zeros_like(f["qv"])
    a0 = _atm(f0)
    lib.p3o_preamble(C.byref(oracle._atm_struct(a0)), 16, 72)
    np.testing.assert_allclose(a["dz"] / a0["dz"] - 1.0, c1 * f["qv"], rtol=1e-11, atol=0)


def test_rain_fraction_maximum_overlap_brute_force():
    lib = oracle.load_oracle()
    f = make_atm_fields(16, 72, seed=5)
    a = _atm(f)
    lib.p3o_preamble(C.byref(oracle._atm_struct(a)), 16, 72)
    cft = f["cldfrac_tot"]
    for i in range(16):
        for k in range(72):
            want = max(cft[k, i], 1e-4)
            if k > 0:
                want = max(want, cft[k - 1, i])
            assert a["cld_frac_r"][k, i] == want
            assert a["cld_frac_l"][k, i] == max(cft[k, i], 1e-4) == a["cld_frac_i"][k, i]


@pytest.mark.parametrize("dt", [300.0, 1800.0])
def test_process_step_wet_water_budget(oracle_tables, dt):
    f = make_atm_fields(64, 72, seed=6, dry_every=8, evap_every=9)
    out, dbg, nerr = oracle.run_process(f, oracle_tables, dt, fluxes=True)
    assert nerr == 0
    wet = lambda s: ((s["qv"] + s["qc"] + s["qr"] + s["qi"]) * f["pseudo_density"] / G).sum(axis=0)
    before = wet(f)
    after = wet(out) + out["precip_liq_surf_mass"] + out["precip_ice_surf_mass"]
    assert np.all(np.abs(after - before) <= 1e-12 * before)
    # boundary fluxes (eamxx_p3_process_interface.hpp:253-258)
    np.testing.assert_allclose(out["water_flux"] * 1000.0 * dt,
                               out["precip_liq_surf_mass"] + out["precip_ice_surf_mass"], rtol=1e-15)
    assert np.all(out["vapor_flux"] == 0) and np.all(out["heat_flux"] == 0)
    assert np.all(out["ice_flux"] <= out["water_flux"])
    # radii in microns: the part-3 defaults 10, 25 and 500 um where nothing is present
    assert out["eff_radius_qc"].max() >= 10.0 and out["eff_radius_qr"].max() >= 500.0


@pytest.mark.parametrize("dt", [100.0, 300.0, 1800.0])
def test_process_step_passes_column_conservation_check(oracle_tables, dt):
    """The reference's mass-and-energy column conservation check, run around a
    process step with P3's boundary fluxes: both close to roundoff.  A 1e-9
    relative change of the post-step temperature (a plausible slip in the
    post-amble's rescaled dT) is caught."""
    f = make_atm_fields(64, 72, seed=7, dry_every=8, evap_every=9)
    m0, e0 = oracle.column_budget(f)
    out, _, nerr = oracle.run_process(f, oracle_tables, dt, fluxes=True)
    assert nerr == 0
    after = dict(out, pseudo_density=f["pseudo_density"])
    m1, e1 = oracle.column_budget(after)
    fl = [out[k] for k in ("vapor_flux", "water_flux", "ice_flux", "heat_flux")]
    me, ee = oracle.conservation_errors(m0, e0, m1, e1, *fl, dt)
    assert me.max() < 1e-14 and ee.max() < 1e-14, (me.max(), ee.max())
    bad = dict(after, T_mid=after["T_mid"] * (1 + 1e-9))
    _, e_bad = oracle.column_budget(bad)
    _, ee_bad = oracle.conservation_errors(m0, e0, m1, e_bad, *fl, dt)
    assert ee_bad.max() > 1e-10


def test_column_budget_closed_forms():
    """Mass of a column of uniform mixing ratios is q * sum(dp) / g; the energy of
    an isothermal dry calm column is cp T sum(dp) / g plus phis ps / g."""
    nlev, ncol = 72, 5
    dp = np.full((nlev, ncol), 1000.0)
    z = np.zeros((nlev, ncol))
    q = np.full((nlev, ncol), 1e-3)
    f = dict(pseudo_density=dp, T_mid=np.full((nlev, ncol), 250.0), qv=q, qc=z, qi=z, qr=z,
             ps=np.full(ncol, 1e5), phis=np.full(ncol, 100.0))
    m, e = oracle.column_budget(f)
    np.testing.assert_allclose(m, 1e-3 * 72000.0 / 9.80616, rtol=1e-14)
    f["qv"] = z
    m, e = oracle.column_budget(f)
    np.testing.assert_allclose(e, (1004.64 * 250.0 * 72000.0 + 100.0 * 1e5) / 9.80616, rtol=1e-14)


def test_shoc_postprocess_closed_form():
    """SHOC's post-process (eamxx_shoc_process_interface.hpp:288-300): 1 where qc or
    its variance is zero; qc^2/qc2 inside [0.001, 10]; the clamps outside it; the
    liquid cloud fraction clipped to at most 1 and otherwise untouched."""
    qc = np.array([[0.0, 1e-4, 1e-4, 2e-4, 1e-6, 3e-5]])
    qc2 = np.array([[1e-9, 0.0, 1e-8, 4e-9, 1e-8, 9e-10]])
    liq = np.array([[0.2, 1.0, 1.5, 0.0, 0.999, 7.0]])
    r, cl = oracle.shoc_postprocess(qc, qc2, liq)
    want = np.array([[1.0, 1.0, 1.0, 10.0, 0.001, 1.0]])
    # 1e-4^2/1e-8 = 1, 2e-4^2/4e-9 = 10 (clamped at 10), 1e-6^2/1e-8 = 1e-4 -> 0.001,
    # 3e-5^2/9e-10 = 1
    np.testing.assert_allclose(r, want, rtol=1e-15)
    mid = oracle.shoc_postprocess(np.array([[3e-4]]), np.array([[2e-8]]))[0]
    assert mid[0, 0] == 3e-4 * 3e-4 / 2e-8  # inside the bounds: the ratio itself, 4.5
    np.testing.assert_array_equal(cl, np.array([[0.2, 1.0, 1.0, 0.0, 0.999, 1.0]]))
    assert oracle.shoc_postprocess(qc, qc2)[1] is None
