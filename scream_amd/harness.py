"""Reference-style regression harness for the P3 step.

Mirrors components/eamxx/src/physics/p3/tests/p3_run_and_cmp.cpp:
  * the "mixed" initial condition of components/eamxx/src/physics/p3/p3_ic_cases.cpp:11-118;
  * the FortranData arrays and their order (components/eamxx/src/physics/p3/p3_f90.cpp:19-88),
    each (ncol, nlev) row-major like the reference's LayoutRight views;
  * the baseline file: for every field in that order, int32 dim (always 2),
    int32 extents[dim] (a 1-D field is (ncol, 1)), then the float64 data
    (p3_run_and_cmp.cpp:175-183);
  * the cases: predict_nc x prescribed_CCN in {no, yes} (or one of them),
    nsteps steps each, one baseline record per step (p3_run_and_cmp.cpp:53-130);
  * the comparison: per field, |ref - new| > tol * max|ref| counts an error,
    and so does a NaN or Inf in either (share/util/scream_utils.hpp:306-340).

The implementation under test is the HIP step (scream_amd.p3x) by default; any
callable with the signature of ``GpuRunner.__call__`` can stand in for it.

Command line (the reference's options)::

    python -m scream_amd.harness -g -b base.bin     # generate a baseline
    python -m scream_amd.harness -b base.bin -t 0   # run and compare
      [-s steps=6] [-dt 300] [-i ncol=3] [-k nlev=72] [-p both] [-c both] [-r repeat]
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

# p3_f90.cpp:77-86, the order fields are written and read
FIELD_ORDER = ("qv", "th_atm", "pres", "dz", "nc_nuceat_tend", "nccn_prescribed", "ni_activated",
               "inv_qc_relvar", "qc", "nc", "qr", "nr", "qi", "ni", "qm", "bm", "precip_liq_surf",
               "precip_ice_surf", "diag_eff_radius_qc", "diag_eff_radius_qi", "diag_eff_radius_qr",
               "rho_qi", "dpres", "inv_exner", "qv2qi_depos_tend", "precip_liq_flux",
               "precip_ice_flux", "cld_frac_r", "cld_frac_l", "cld_frac_i", "liq_ice_exchange",
               "vap_liq_exchange", "vap_ice_exchange", "qv_prev", "t_prev")
ONE_D = ("precip_liq_surf", "precip_ice_surf")
INTERFACE = ("precip_liq_flux", "precip_ice_flux")

P0, RD, CP = 100000.0, 287.042, 1004.64


def _shape(name, ncol, nlev):
    if name in ONE_D:
        return (ncol,)
    return (ncol, nlev + 1) if name in INTERFACE else (ncol, nlev)


def new_data(ncol: int, nlev: int) -> dict:
    """Zeroed FortranData arrays (the reference's views start at zero)."""
    return {k: np.zeros(_shape(k, ncol, nlev)) for k in FIELD_ORDER}


# What the reference's wrapper zeroes before every step: the arrays after the first
# NUM_INPUT_ARRAYS = 24 of its list (p3_functions_f90.cpp:1926-1945,
# p3_functions_f90.hpp:753); the Fortran zeroes the exchange fields itself
# (micro_p3.F90:1321-1323).
OUTPUTS_ZEROED = ("diag_eff_radius_qi", "diag_eff_radius_qr", "rho_qi",
                  "qv2qi_depos_tend", "liq_ice_exchange", "vap_liq_exchange", "vap_ice_exchange",
                  "precip_liq_flux", "precip_ice_flux", "precip_liq_surf", "precip_ice_surf")


def zero_outputs(d: dict) -> None:
    for k in OUTPUTS_ZEROED:
        d[k][...] = 0.0


def make_mixed(ncol: int, nlev: int) -> dict:
    """The "mixed" initial condition, p3_ic_cases.cpp:11-118 (needs nlev >= 36)."""
    if nlev < 36:
        raise ValueError("the mixed initial condition needs at least 36 levels")
    d = new_data(ncol, nlev)
    nk = nlev
    g = 9.8  # p3_ic_cases.cpp:102, this generator's own gravity
    for i in range(ncol):
        for k in range(nk):
            # (100.0 + i/ncol - k/nk) with the reference's integer divisions
            d["nccn_prescribed"][i, k] = (100.0 + (i // ncol) - (k // nk)) * 1e6
        for k in range(15):
            d["qc"][i, nk - 20 + k] = 1e-4 * (1 - k / 14.0)
        d["nc"][i, :] = 1e6
        for k in range(20):
            d["qr"][i, nk - 20 + k] = 1e-5 * (1 - k / 19.0)
        d["nr"][i, :] = 1e6
        for k in range(15):
            d["qi"][i, nk - 20 + k] = 1e-4
        d["ni"][i, :] = 1e6
        for k in range(15):
            d["qm"][i, nk - 20 + k] = 1e-4 * (1 - k / 14.0)
            d["bm"][i, nk - 20 + k] = 1e-2
        for k in range(nk):
            tmp = -5e-4 + 1e-3 / nk * k
            d["qv"][i, k] = tmp if tmp > 0 else 0.0
        for k in range(15):
            d["qv"][i, nk - 20 + k] = 5e-3
        for k in range(nk):
            d["pres"][i, k] = 100 + 1e5 / nk * k
            d["dpres"][i, k] = 1e5 / nk
            d["inv_exner"][i, k] = (1e5 / d["pres"][i, k]) ** (287.15 / 1005.0)
            d["cld_frac_i"][i, k] = 1.0
            d["cld_frac_l"][i, k] = 1.0
            d["cld_frac_r"][i, k] = 1.0
            d["inv_qc_relvar"][i, k] = 1.0
        T = np.empty(nk)
        for k in range(nk):
            T[k] = 150 + 150 / nk * k
            if i > 0:
                T[k] += ((i % 3) - 0.5) / nk * k
            d["th_atm"][i, k] = T[k] * (P0 / d["pres"][i, k]) ** (RD / CP)
        d["qi"][i, nk - 1] = 1e-9
        d["qv"][i, nk - 1] = 5e-2
        d["qr"][i, nk - 1] = 1e-6
        d["qc"][i, nk - 1] = 1e-6
        d["qi"][i, nk - 25] = 5e-8
        d["qc"][i, 35] = 1e-7
        d["qv"][i, 35] = 1e-6
        d["qv"][i, 33] = 1e-4
        for k in range(nk):
            d["qv_prev"][i, k] = d["qv"][i, k]
            d["t_prev"][i, k] = T[k]
        p = d["pres"][i]
        for k in range(nk):
            plo = max(float(i), p[0] - 0.5 * (p[1] - p[0]) / (1 - 0)) if k == 0 else 0.5 * (p[k - 1] + p[k])
            phi = p[nk - 1] + 0.5 * (p[nk - 1] - p[nk - 2]) / (1 - 0) if k == nk - 1 else 0.5 * (p[k] + p[k + 1])
            d["dz"][i, k] = RD * T[k] / (g * p[k]) * (phi - plo)
    return d


def write_record(fh, d: dict):
    """One baseline record (p3_run_and_cmp.cpp:175-183)."""
    for k in FIELD_ORDER:
        a = np.ascontiguousarray(d[k], dtype=np.float64)
        ext = (a.shape[0], 1) if a.ndim == 1 else a.shape
        fh.write(np.array([2], dtype=np.int32).tobytes())
        fh.write(np.array(ext, dtype=np.int32).tobytes())
        fh.write(a.tobytes())


def read_record(fh, ncol: int, nlev: int) -> dict:
    """One baseline record, checked against the expected dims and extents
    (p3_run_and_cmp.cpp:185-209)."""
    d = {}
    for k in FIELD_ORDER:
        shape = _shape(k, ncol, nlev)
        ext = (shape[0], 1) if len(shape) == 1 else shape
        dim = np.frombuffer(fh.read(4), dtype=np.int32)
        if dim.size != 1 or dim[0] != 2:
            raise ValueError(f"field {k}: expected dim 2, got {dim}")
        got = tuple(np.frombuffer(fh.read(8), dtype=np.int32))
        if got != tuple(ext):
            raise ValueError(f"field {k}: expected extents {ext}, got {got}")
        n = int(np.prod(shape))
        a = np.frombuffer(fh.read(8 * n), dtype=np.float64)
        if a.size != n:
            raise ValueError(f"field {k}: baseline file ends early")
        d[k] = a.reshape(shape).copy()
    return d


def compare(tol: float, ref: dict, new: dict, out=sys.stdout) -> int:
    """Errors per the reference's rule (share/util/scream_utils.hpp:306-340)."""
    nerr = 0
    for k in FIELD_ORDER:
        a, b = np.ravel(ref[k]), np.ravel(new[k])
        den = np.abs(a).max() if a.size else 0.0
        bad1 = ~np.isfinite(a) | ~np.isfinite(b)
        num = np.abs(a - b)
        bad2 = ~bad1 & (num > tol * den)
        n1, n2 = int(bad1.sum()), int(bad2.sum())
        if n1:
            print(f"{k} has {n1} infs + nans.", file=out)
        if n2:
            print(f"{k} > tol {n2} times. Max rel diff= {num[bad2].max() / den} normalized by ref impl "
                  f"val={den}.", file=out)
        nerr += n1 + n2
    return nerr


class GpuRunner:
    """Advances FortranData arrays one P3 step with the HIP library (p3x_main)."""

    def __init__(self, precision: int | None = None):
        import torch
        from scream_amd import p3x
        self.p3x, self.torch = p3x, torch
        self.ctx = p3x.Context()
        self.precision = p3x.P3X_FP64 if precision is None else precision

    def __call__(self, d: dict, dt: float, it: int, predict_nc: bool, prescribed_ccn: bool) -> float:
        p3x, torch = self.p3x, self.torch
        zero_outputs(d)
        ncol, nlev = d["qc"].shape
        dev = torch.device("cuda:0")
        f = {}
        for k in p3x.ALL_FIELDS:
            src = {"th": "th_atm", "nccn": "nccn_prescribed"}.get(k, k)
            if src in d:
                f[k] = torch.from_numpy(np.ascontiguousarray(d[src].T)).to(dev)
            elif k in ("precip_total_tend", "nevapr"):
                f[k] = torch.zeros((nlev, ncol), dtype=torch.float64, device=dev)
        f["precip_liq_surf"] = torch.from_numpy(d["precip_liq_surf"].copy()).to(dev)
        f["precip_ice_surf"] = torch.from_numpy(d["precip_ice_surf"].copy()).to(dev)
        torch.cuda.synchronize()
        us = self.ctx.main(f, ncol, nlev, dt, predict_nc, prescribed_ccn, precision=self.precision)
        for k in p3x.ALL_FIELDS:
            dst = {"th": "th_atm", "nccn": "nccn_prescribed"}.get(k, k)
            if dst in d:
                v = f[k].cpu().numpy()
                d[dst][...] = v if v.ndim == 1 else v.T
        return us


def cases(predict_nc: str = "both", prescribed_ccn: str = "both"):
    """(predict_nc, prescribed_CCN) pairs in the reference's order (p3_run_and_cmp.cpp:57-76)."""
    pi = {"yes": [True], "no": [False], "both": [False, True]}
    return [(p, c) for p in pi[predict_nc] for c in pi[prescribed_ccn]]


def generate_baseline(filename: str, runner, nsteps=6, dt=300.0, ncol=3, nlev=72, predict_nc="both",
                      prescribed_ccn="both") -> None:
    with open(filename, "wb") as fh:
        for pnc, pccn in cases(predict_nc, prescribed_ccn):
            d = make_mixed(ncol, nlev)
            for _ in range(nsteps):
                runner(d, dt, nsteps, pnc, pccn)  # d.it = nsteps, as set_params does
                write_record(fh, d)


def run_and_cmp(filename: str, runner, tol=0.0, nsteps=6, dt=300.0, ncol=3, nlev=72, predict_nc="both",
                prescribed_ccn="both", out=sys.stdout) -> int:
    nerr = 0
    with open(filename, "rb") as fh:
        for case_num, (pnc, pccn) in enumerate(cases(predict_nc, prescribed_ccn), 1):
            d = make_mixed(ncol, nlev)
            for it in range(nsteps):
                print(f"--- checking case # {case_num}, timestep # {it + 1} of {nsteps} ---", file=out)
                ref = read_record(fh, ncol, nlev)
                runner(d, dt, nsteps, pnc, pccn)
                ne = compare(tol, ref, d, out)
                if ne:
                    print("Ref impl failed.", file=out)
                nerr += ne
    return nerr


def timing(runner, repeat, nsteps=6, dt=300.0, ncol=3, nlev=72, predict_nc="both", prescribed_ccn="both",
           out=sys.stdout) -> None:
    """The reference's -r mode: per case, mean over ``repeat`` runs of the summed step times,
    after one cold run (p3_run_and_cmp.cpp:87-121)."""
    for pnc, pccn in cases(predict_nc, prescribed_ccn):
        total = 0.0
        for r in range(-1, repeat):
            d = make_mixed(ncol, nlev)
            for _ in range(nsteps):
                us = runner(d, dt, nsteps, pnc, pccn)
                if r >= 0:
                    total += us
        print(f"Running P3 with ni={ncol}, nk={nlev}, dt={dt}, ts={nsteps}, predict_nc={int(pnc)}, "
              f"prescribed_CCN={int(pccn)}", file=out)
        print(f"Time = {1e-6 * total / repeat:1.3e} seconds", file=out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("-g", "--generate", action="store_true")
    ap.add_argument("-b", "--baseline-file", default="p3_baseline.bin")
    ap.add_argument("-t", "--tol", type=float, default=0.0)
    ap.add_argument("-s", "--steps", type=int, default=6)
    ap.add_argument("-dt", "--dt", type=float, default=300.0)
    ap.add_argument("-i", "--ncol", type=int, default=3)
    ap.add_argument("-k", "--nlev", type=int, default=72)
    ap.add_argument("-r", "--repeat", type=int, default=0)
    ap.add_argument("-p", "--predict-nc", default="both", choices=["yes", "no", "both"])
    ap.add_argument("-c", "--prescribed-ccn", default="both", choices=["yes", "no", "both"])
    a = ap.parse_args(argv)
    runner = GpuRunner()
    kw = dict(nsteps=a.steps, dt=a.dt, ncol=a.ncol, nlev=a.nlev, predict_nc=a.predict_nc,
              prescribed_ccn=a.prescribed_ccn)
    if a.repeat > 0:
        timing(runner, a.repeat, **kw)
        return 0
    if a.generate:
        generate_baseline(a.baseline_file, runner, **kw)
        return 0
    t0 = time.perf_counter()
    nerr = run_and_cmp(a.baseline_file, runner, tol=a.tol, **kw)
    print(f"{nerr} errors ({time.perf_counter() - t0:.1f} s)")
    return 1 if nerr else 0


if __name__ == "__main__":
    sys.exit(main())
