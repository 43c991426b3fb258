"""The CldFraction oracle (oracle.cld_fraction; physics/cld_fraction/
cld_fraction_main_impl.hpp:12-95, thresholds from
eamxx_cld_fraction_process_interface.cpp:60-62) against hand-worked cases and
the invariants its definition fixes.  The reference has no unit test of its own
for this process; its multi-process tests run it between shoc and p3
(tests/multi-process/dynamics_physics/homme_shoc_cld_p3_rrtmgp/input.yaml:33)."""
import numpy as np

import oracle


def test_known_answers():
    # one level, columns chosen at and around both default thresholds
    qi = np.array([[0.0, 1e-12, np.nextafter(1e-12, 1.0), 1e-8, 1e-5, 2e-5, 1e-3, 1e-13]])
    liq = np.array([[0.0, 0.25, 0.5, 0.25, 0.75, 0.0, 1.0, 0.9]])
    out = oracle.cld_fraction(qi, liq)
    # the threshold is strict: qi equal to it is not ice cloud
    np.testing.assert_array_equal(out["cldfrac_ice"], [[0, 0, 1, 1, 1, 1, 1, 0]])
    np.testing.assert_array_equal(out["cldfrac_ice_for_analysis"], [[0, 0, 0, 0, 0, 1, 1, 0]])
    np.testing.assert_array_equal(out["cldfrac_tot"], [[0.0, 0.25, 1, 1, 1, 1, 1, 0.9]])
    np.testing.assert_array_equal(out["cldfrac_tot_for_analysis"], [[0.0, 0.25, 0.5, 0.25, 0.75, 1, 1, 0.9]])


def test_thresholds_are_parameters():
    qi = np.array([[1e-9, 1e-7], [1e-6, 0.0]])
    liq = np.full((2, 2), 0.125)
    out = oracle.cld_fraction(qi, liq, ice_threshold=1e-8, ice_4out_threshold=1e-6)
    np.testing.assert_array_equal(out["cldfrac_ice"], [[0, 1], [1, 0]])
    np.testing.assert_array_equal(out["cldfrac_ice_for_analysis"], [[0, 0], [0, 0]])
    np.testing.assert_array_equal(out["cldfrac_tot"], [[0.125, 1], [1, 0.125]])


def test_invariants_on_random_columns():
    rng = np.random.default_rng(3)
    nlev, ncol = 72, 500
    qi = np.where(rng.random((nlev, ncol)) < 0.4, 0.0, 10.0 ** rng.uniform(-16, -2, (nlev, ncol)))
    liq = np.where(rng.random((nlev, ncol)) < 0.3, 0.0, rng.random((nlev, ncol)))
    out = oracle.cld_fraction(qi, liq)
    ice, tot = out["cldfrac_ice"], out["cldfrac_tot"]
    ice4, tot4 = out["cldfrac_ice_for_analysis"], out["cldfrac_tot_for_analysis"]
    # fractions are 0 or 1 for ice; the larger threshold marks a subset of the cells
    assert set(np.unique(ice)) <= {0.0, 1.0} and set(np.unique(ice4)) <= {0.0, 1.0}
    assert np.all(ice4 <= ice)
    # the total is the liquid fraction or a full ice cloud, never below either
    assert np.all((tot == liq) | (tot == 1.0)) and np.all(tot >= liq) and np.all(tot >= ice)
    assert np.all(tot4 <= tot)
    # where there is no ice cloud the total is the liquid fraction, bit for bit
    np.testing.assert_array_equal(tot[ice == 0], liq[ice == 0])
    assert np.all(tot[ice == 1] == 1.0)
