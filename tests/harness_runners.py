"""Runners of the regression harness (scream_amd.harness) backed by the checkers:
the oracle (C restatement) and the reference Fortran (oracle/_ref).  Test code only."""
import numpy as np

import oracle
from scream_amd import harness

# FortranData name -> oracle field name
_TO_ORACLE = {"th_atm": "th", "nccn_prescribed": "nccn"}
_IN = oracle.STATE + oracle.INPUTS + oracle.HIST


def _to_levels(d):
    """FortranData (ncol, nlev) arrays -> the oracle's level-major inputs."""
    out = {}
    for k in _IN:
        src = {v: k2 for k2, v in _TO_ORACLE.items()}.get(k, k)
        if src in d:
            out[k] = np.ascontiguousarray(d[src].T)
    return out


def _from_levels(d, res):
    for k, v in res.items():
        dst = {"th": "th_atm", "nccn": "nccn_prescribed"}.get(k, k)
        if dst in d:
            d[dst][...] = v if v.ndim == 1 else v.T


class OracleRunner:
    def __init__(self, tables, math="glibc"):
        self.tables, self.math = tables, math

    def __call__(self, d, dt, it, pnc, pccn):
        harness.zero_outputs(d)
        res, _, nerr = oracle.run_oracle(_to_levels(d), self.tables, dt, pnc, pccn, math=self.math)
        assert nerr == 0
        _from_levels(d, {k: res[k] for k in oracle.STATE + oracle.OUTS + oracle.HIST})
        return 0.0


class FortranRunner:
    def __init__(self, table_dir):
        self.table_dir = table_dir

    def __call__(self, d, dt, it, pnc, pccn):
        res = oracle.run_ref(_to_levels(d), self.table_dir, dt, pnc, pccn, it=it)
        keep = oracle.STATE + oracle.HIST + ("precip_liq_surf", "precip_ice_surf", "diag_eff_radius_qc",
                                             "diag_eff_radius_qi", "diag_eff_radius_qr", "rho_qi",
                                             "qv2qi_depos_tend", "precip_liq_flux", "precip_ice_flux")
        _from_levels(d, {k: res[k] for k in keep})
        return 0.0
