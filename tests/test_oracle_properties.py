"""Properties that pin the oracle where the golden vectors leave room (CPU).

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong index or sign, a swapped operand) fails it:
* saturation vapour pressure at the Murphy & Koop (2005) Table C1 check values
  (temperatures listed in components/eamxx/src/physics/share/tests/
  physics_saturation_run_and_cmp.cpp:96-104);
* trilinear / 4-D table interpolation reproduces a table that is linear in
  its indices exactly (the idea of p3_ice_tables_unit_tests.cpp:63-97);
* ice/rain table indices equal a brute-force floor-and-clamp;
* the first-order upwind step conserves mass including the outflow and
  creates no negative values (p3_upwind_unit_tests.cpp:41-200);
* a whole p3_main step closes the column total-water budget to roundoff and
  leaves no negative mass (mass_and_energy_column_conservation_check.cpp).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from scream_amd.inputs import make_columns

EP_2 = 18.016 / 28.966

# Murphy & Koop (2005) Table C1, "values recommended for checking computer codes" [Pa]
MK_C1 = [(150.0, 6.1061e-6, 1.5621e-5), (180.0, 0.0053975, 0.011239), (210.0, 0.70202, 1.2335),
         (240.0, 27.272, 37.667), (273.16, 611.657, 611.657), (300.0, None, 3536.8)]


@pytest.mark.parametrize("T,e_ice,e_liq", MK_C1)
def test_murphy_koop_check_values(T, e_ice, e_liq):
    lib = oracle.load_oracle()
    p = 1e5
    liq = lib.p3o_qv_sat_dry(T, p, 0) * p / EP_2
    assert abs(liq / e_liq - 1) < 5e-5
    if e_ice is not None and T < 273.15:
        ice = lib.p3o_qv_sat_dry(T, p, 1) * p / EP_2
        assert abs(ice / e_ice - 1) < 5e-5


def test_qv_sat_pressure_floor():
    lib = oracle.load_oracle()
    # qv_sat_dry divides by max(p, 1e-3) (physics_saturation_impl.hpp:155)
    assert lib.p3o_qv_sat_dry(250.0, 0.0, 1) == lib.p3o_qv_sat_dry(250.0, 1e-3, 1)


def _linear_tables(oracle_tables):
    """Tables that are exactly linear in their (0-based) indices."""
    t = dict(oracle_tables)
    jj, ii, i, j = np.meshgrid(np.arange(5), np.arange(4), np.arange(50), np.arange(12), indexing="ij")
    t["ice"] = 1.0 + 0.5 * jj + 0.25 * ii + 0.125 * i + 2.0 * j
    jj, ii, i, r, k = np.meshgrid(np.arange(5), np.arange(4), np.arange(50), np.arange(30), np.arange(2),
                                  indexing="ij")
    t["coll"] = -3.0 + 0.5 * jj + 0.25 * ii + 0.125 * i + 0.0625 * r + 4.0 * k
    return t


def _tabs(t):
    return oracle._tables(t)


def test_ice_interpolation_reproduces_linear_table(oracle_tables):
    lib = oracle.load_oracle()
    t = _linear_tables(oracle_tables)
    tb, keep = _tabs(t)
    rng = np.random.default_rng(7)
    for _ in range(300):
        qi = 10 ** rng.uniform(-9, -3)
        ni = 10 ** rng.uniform(2, 7)
        qm = qi * rng.uniform(0, 0.99)
        rhop = rng.uniform(50, 900)
        idx = int(rng.integers(0, 12))
        dumi, dumii, dumjj = C.c_int(), C.c_int(), C.c_int()
        d1, d4, d5 = C.c_double(), C.c_double(), C.c_double()
        lib.p3o_lookup_ice(qi, ni, qm, rhop, C.byref(dumi), C.byref(dumii), C.byref(dumjj),
                           C.byref(d1), C.byref(d4), C.byref(d5))
        got = lib.p3o_apply_table_ice(C.byref(tb), idx, qi, ni, qm, rhop)
        # linear in the 1-based continuous coordinates dum1/dum4/dum5
        want = 1.0 + 0.5 * (d5.value - 1) + 0.25 * (d4.value - 1) + 0.125 * (d1.value - 1) + 2.0 * idx
        assert abs(got - want) <= 1e-12 * abs(want)


def test_collection_interpolation_reproduces_linear_table(oracle_tables):
    lib = oracle.load_oracle()
    t = _linear_tables(oracle_tables)
    tb, keep = _tabs(t)
    rng = np.random.default_rng(8)
    for _ in range(300):
        qi = 10 ** rng.uniform(-9, -3)
        ni = 10 ** rng.uniform(2, 7)
        qm = qi * rng.uniform(0, 0.99)
        rhop = rng.uniform(50, 900)
        qr = 10 ** rng.uniform(-8, -3)
        nr = 10 ** rng.uniform(2, 6)
        idx = int(rng.integers(0, 2))
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        d1, d4, d5 = C.c_double(), C.c_double(), C.c_double()
        lib.p3o_lookup_ice(qi, ni, qm, rhop, C.byref(a), C.byref(b), C.byref(c),
                           C.byref(d1), C.byref(d4), C.byref(d5))
        j, d3 = C.c_int(), C.c_double()
        lib.p3o_lookup_rain(qr, nr, C.byref(j), C.byref(d3))
        got = lib.p3o_apply_table_coll(C.byref(tb), idx, qi, ni, qm, rhop, qr, nr)
        want = (-3.0 + 0.5 * (d5.value - 1) + 0.25 * (d4.value - 1) + 0.125 * (d1.value - 1) +
                0.0625 * (d3.value - 1) + 4.0 * idx)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_ice_indices_brute_force():
    lib = oracle.load_oracle()
    rng = np.random.default_rng(9)
    for _ in range(2000):
        qi = 10 ** rng.uniform(-14, -2)
        ni = 10 ** rng.uniform(-16, 8)
        qm = qi * rng.uniform(0, 1.5)
        rhop = rng.uniform(0, 1200)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        d1, d4, d5 = C.c_double(), C.c_double(), C.c_double()
        lib.p3o_lookup_ice(qi, ni, qm, rhop, C.byref(a), C.byref(b), C.byref(c),
                           C.byref(d1), C.byref(d4), C.byref(d5))
        x1 = (math.log10(qi / ni) + 18) * 4.135985029041767 - 10  # C libm, as the oracle
        x4 = (qm / qi) * 3 + 1
        x5 = (rhop - 50) * 0.005 + 1 if rhop <= 650 else (rhop - 650) * 0.004 + 4
        assert a.value == min(max(int(x1), 1), 49) - 1
        assert b.value == min(max(int(x4), 1), 3) - 1
        assert c.value == min(max(int(x5), 1), 4) - 1
        assert d1.value == min(max(x1, 1.0), 50.0)


def test_lookup_golden(oracle_tables):
    """Fortran find_lookupTable_indices + access_lookup_table on the reference
    unit-test literals (tests/golden/lookup.npz) -- bit for bit."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "lookup.npz"))
    lib = oracle.load_oracle(math="fortran")
    tb, keep = _tabs(oracle_tables)
    for s, (qi, ni, qm, rhop) in enumerate(z["ice_in"]):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        d1, d4, d5 = C.c_double(), C.c_double(), C.c_double()
        lib.p3o_lookup_ice(qi, ni, qm, rhop, C.byref(a), C.byref(b), C.byref(c),
                           C.byref(d1), C.byref(d4), C.byref(d5))
        assert (a.value, b.value, c.value) == tuple(z["ice_idx"][s])
        assert (d1.value, d4.value, d5.value) == tuple(z["ice_w"][s])
        for idx in range(12):
            assert lib.p3o_apply_table_ice(C.byref(tb), idx, qi, ni, qm, rhop) == z["ice_proc"][s, idx]
        qr, nr = z["rain_in"][s]
        if qr > 1e-14:  # the Fortran tests qr >= qsmall, the C++ qr > qsmall: equal off the boundary
            j, d3 = C.c_int(), C.c_double()
            lib.p3o_lookup_rain(qr, nr, C.byref(j), C.byref(d3))
            assert (j.value, d3.value) == (z["coll_j"][s], z["coll_d3"][s])
            for idx in range(2):
                got = lib.p3o_apply_table_coll(C.byref(tb), idx, qi, ni, qm, rhop, qr, nr)
                assert got == z["coll_proc"][s, idx]


def test_upwind_conserves_mass_and_positivity():
    lib = oracle.load_oracle()
    rng = np.random.default_rng(10)
    dp = C.POINTER(C.c_double)
    for trial in range(50):
        nk = 40
        rho = rng.uniform(0.05, 1.2, nk)
        dz = rng.uniform(100, 600, nk)
        inv_dz, inv_rho = 1 / dz, 1 / rho
        V = rng.uniform(0, 10, nk)
        r = rng.uniform(0, 1e-3, nk)
        k_top, k_bot = int(rng.integers(0, 10)), nk - 1
        co = np.max(V * inv_dz)
        dt_sub = 0.999 / co  # CFL-limited substep
        flux = np.zeros(nk)
        r0 = r.copy()
        ptr = lambda a: a.ctypes.data_as(dp)
        fl = (dp * 1)(ptr(flux))
        vv = (dp * 1)(ptr(V))
        rr = (dp * 1)(ptr(r))
        lib.p3o_upwind_step(nk, k_bot, k_top, dt_sub, 1, ptr(rho), ptr(inv_rho), ptr(inv_dz), fl, vv, rr)
        mass0 = (r0 * rho * dz).sum()
        mass1 = (r * rho * dz).sum() + flux[k_bot] * dt_sub
        assert abs(mass1 - mass0) <= 1e-13 * mass0
        assert (r >= -1e-18).all()
        # levels above the window are untouched
        np.testing.assert_array_equal(r[:k_top], r0[:k_top])


@pytest.mark.parametrize("dt,nlev", [(1800.0, 72), (100.0, 128)])
def test_step_budget_and_positivity(oracle_tables, dt, nlev):
    inp = make_columns(64, nlev, seed=33, dry_every=8, evap_every=9)
    out, dbg, nerr = oracle.run_oracle(inp, oracle_tables, dt, True, False)
    g = 9.80616
    before = ((inp["qv"] + inp["qc"] + inp["qr"] + inp["qi"]) * inp["dpres"] / g).sum(axis=0)
    after = ((out["qv"] + out["qc"] + out["qr"] + out["qi"]) * inp["dpres"] / g).sum(axis=0)
    after += (out["precip_liq_surf"] + out["precip_ice_surf"]) * 1000.0 * dt
    assert np.all(np.abs(after - before) <= 1e-12 * before)
    for k in ("qc", "qr", "qi", "qm", "qv"):
        assert out[k].min() >= 0, k
    # precipitation only leaves through the surface and is non-negative
    assert out["precip_liq_surf"].min() >= 0 and out["precip_ice_surf"].min() >= 0
