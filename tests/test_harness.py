"""The regression harness (scream_amd.harness) on the CPU: baseline file format,
the comparison rule, the mixed initial condition, and the harness driving the
reference's own Fortran P3 against the oracle (bit for bit with the Fortran
spelling of cbrt/expm1), plus the committed Fortran baseline fixture."""
import io
import os

import numpy as np
import pytest

import oracle
from scream_amd import harness

from harness_runners import FortranRunner, OracleRunner

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TABLE_DIR = os.path.join(os.path.dirname(GOLDEN), "..", "scream_amd", "data")
FIXTURE = os.path.join(GOLDEN, "harness_mixed_i3_k72_s3.bin")


def test_record_round_trip_and_layout():
    rng = np.random.default_rng(0)
    d = {k: rng.standard_normal(v.shape) for k, v in harness.new_data(3, 40).items()}
    buf = io.BytesIO()
    harness.write_record(buf, d)
    raw = buf.getvalue()
    # first field qv: dim 2, extents (3, 40), then 120 doubles
    assert np.frombuffer(raw[:12], dtype=np.int32).tolist() == [2, 3, 40]
    np.testing.assert_array_equal(np.frombuffer(raw[12:12 + 960]), d["qv"].ravel())
    per_2d, per_1d = 12 + 3 * 40 * 8, 12 + 3 * 8
    per_if = 12 + 3 * 41 * 8
    assert len(raw) == 31 * per_2d + 2 * per_1d + 2 * per_if
    buf.seek(0)
    back = harness.read_record(buf, 3, 40)
    for k in harness.FIELD_ORDER:
        np.testing.assert_array_equal(back[k], d[k])
    buf.seek(0)
    with pytest.raises(ValueError):
        harness.read_record(buf, 3, 41)


def test_compare_rule():
    d = {k: np.ones(v.shape) for k, v in harness.new_data(2, 36).items()}
    e = {k: v.copy() for k, v in d.items()}
    assert harness.compare(0.0, d, e, io.StringIO()) == 0
    e["qc"][1, 3] = 1.0 + 1e-9
    assert harness.compare(0.0, d, e, io.StringIO()) == 1
    assert harness.compare(1e-8, d, e, io.StringIO()) == 0  # |diff| <= tol * max|ref|
    e["qc"][0, 0] = np.nan
    assert harness.compare(1e-8, d, e, io.StringIO()) == 1


def test_mixed_initial_condition():
    d = harness.make_mixed(3, 72)
    nk = 72
    assert np.all(d["nccn_prescribed"] == 1e8)
    assert np.all(d["qc"][:, nk - 20] == 1e-4) and np.all(d["qc"][:, nk - 6] == 0.0)
    assert np.all(d["qi"][:, nk - 1] == 1e-9) and np.all(d["qv"][:, nk - 1] == 5e-2)
    assert np.all(d["qc"][:, 35] == 1e-7) and np.all(d["qi"][:, nk - 25] == 5e-8)
    # T = th / (P0/p)^(Rd/cp): 150 K at the top, +150 K over the column, column offsets for i > 0
    T = d["th_atm"] / (1e5 / d["pres"]) ** (287.042 / 1004.64)
    np.testing.assert_allclose(T[0], 150 + 150 / nk * np.arange(nk), rtol=1e-14)
    np.testing.assert_array_equal(d["t_prev"][0], 150 + 150 / nk * np.arange(nk))
    assert np.all(d["qv_prev"] == d["qv"])
    assert np.all(d["dz"] > 0)
    with pytest.raises(ValueError):
        harness.make_mixed(3, 30)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
def test_harness_fortran_vs_oracle_bit_exact(oracle_tables, tmp_path):
    base = str(tmp_path / "ref.bin")
    harness.generate_baseline(base, FortranRunner(TABLE_DIR), nsteps=3)
    out = io.StringIO()
    assert harness.run_and_cmp(base, OracleRunner(oracle_tables, math="fortran"), tol=0.0, nsteps=3,
                               out=out) == 0, out.getvalue()


def test_committed_fortran_baseline_vs_oracle(oracle_tables):
    """The committed baseline (written by the reference Fortran through this harness)
    against the oracle, with the reference harness's own rule at tol 1e-12."""
    out = io.StringIO()
    nerr = harness.run_and_cmp(FIXTURE, OracleRunner(oracle_tables), tol=1e-12, nsteps=3, out=out)
    assert nerr == 0, out.getvalue()
    out = io.StringIO()
    assert harness.run_and_cmp(FIXTURE, OracleRunner(oracle_tables, math="fortran"), tol=0.0, nsteps=3,
                               out=out) == 0, out.getvalue()
