"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds oracle/_san/p3sanitize: the oracle's p3_main,
pre/post-amble, CldFraction and column budget (p3_oracle.c), the table
generator restatement (p3_table_gen.c) on two cells, and the product's host
table loader (scream_amd/csrc/p3_tables.cpp), all with -fsanitize=address,undefined
and no recovery, so that any finding aborts the run.  The sanitized p3_main must
also reproduce the ordinary oracle build bit for bit (same source, same
floating-point flags).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from scream_amd import p3x
from scream_amd.inputs import make_columns

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_san", "p3sanitize")


@pytest.fixture(scope="module")
def sanitized():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitized build failed:\n" + r.stderr[-3000:])
    return EXE


@pytest.mark.parametrize("ncol,nlev,dt,pnc,pccn", [(40, 72, 1800.0, 1, 0), (24, 128, 100.0, 0, 1)])
def test_cpu_code_clean_under_asan_ubsan(sanitized, oracle_tables, tmp_path, ncol, nlev, dt, pnc, pccn):
    inp = make_columns(ncol, nlev, seed=17, dry_every=7, evap_every=9)
    hist = {k: np.zeros((nlev, ncol)) for k in oracle.HIST}
    src = tmp_path / "in.bin"
    with open(src, "wb") as fh:
        np.array([ncol, nlev, pnc, pccn], np.int32).tofile(fh)
        np.array([dt], np.float64).tofile(fh)
        for k in oracle.STATE + oracle.INPUTS:
            np.ascontiguousarray(inp[k], np.float64).tofile(fh)
        for k in oracle.HIST:
            hist[k].tofile(fh)
    dst = tmp_path / "out.bin"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([sanitized, p3x.ICE_TABLE, str(src), str(dst)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, f"sanitizer run failed ({r.returncode}):\n{r.stdout}\n{r.stderr[-4000:]}"
    assert "table generator ok" in r.stdout
    want, wdbg, _ = oracle.run_oracle({**inp, **hist}, oracle_tables, dt, bool(pnc), bool(pccn))
    raw = np.fromfile(dst, np.uint8)
    n, nf = ncol * nlev, ncol * (nlev + 1)
    names = list(oracle.STATE) + ["qv2qi_depos_tend", "diag_eff_radius_qc", "diag_eff_radius_qi",
                                  "diag_eff_radius_qr", "rho_qi", "precip_total_tend", "nevapr"] + \
        list(oracle.HIST)
    sizes = [n] * len(names) + [nf, nf, ncol, ncol]
    names += ["precip_liq_flux", "precip_ice_flux", "precip_liq_surf", "precip_ice_surf"]
    off = 0
    for k, m in zip(names, sizes):
        got = raw[off:off + 8 * m].view(np.float64)
        off += 8 * m
        assert np.array_equal(got.view(np.int64), np.ascontiguousarray(want[k], np.float64).ravel().view(np.int64)), k
    ints = raw[off:].view(np.int32)
    assert np.array_equal(ints[:ncol], wdbg["col_status"])
    assert np.array_equal(ints[ncol:].reshape(3, ncol), wdbg["sed_substeps"])
