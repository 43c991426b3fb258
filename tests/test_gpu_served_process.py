"""The served neighbour step: process members (p3x_server_add_process) whose
requests each run SHOC's post-process (inv_qc_relvar), CldFraction, the column
budget, the whole P3Microphysics step and the conservation errors.  Every
member's fields after interleaved, batched requests must equal, bit for bit, the
same stages run alone on one stream (p3x.shoc_postprocess and
p3x.MicrophysicsStep); one member's first step is also checked against the
oracle's chain in FP64."""
import numpy as np
import pytest

import oracle
from scream_amd import p3x
from scream_amd.inputs import ATM_COMPUTED_FIELDS, ATM_FLUX_FIELDS, ATM_UPDATED_FIELDS, make_atm_fields

pytestmark = pytest.mark.gpu

# ncol, nlev, layout, seed
MEMBERS = [(200, 72, p3x.P3X_LEVEL_MAJOR, 31), (77, 128, p3x.P3X_LEVEL_MAJOR, 32),
           (150, 72, p3x.P3X_COLUMN_MAJOR, 33)]
CHAIN_OUT = ("cldfrac_ice", "cldfrac_ice_for_analysis", "cldfrac_tot_for_analysis")


@pytest.fixture(scope="module")
def ctx(host_tables):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = p3x.Context(host_tables)
    yield c
    c.close()


def _inputs(ncol, nlev, seed):
    """Field-manager fields plus SHOC's outputs for the chain: cldfrac_liq (with a few
    values above 1 that SHOC's post-process clips) and qc2, the qc variance, chosen so
    that the ratio qc^2/qc2 lands below, inside and above SHOC's [0.001, 10] clamp."""
    f = make_atm_fields(ncol, nlev, seed=seed, dry_every=7, evap_every=9)
    rng = np.random.default_rng(seed + 100)
    f["cldfrac_liq"] = np.where(f["qc"] > 1e-7, rng.uniform(0.2, 1.1, f["qc"].shape), 0.0)
    ratio = 10.0 ** rng.uniform(-4, 2, f["qc"].shape)
    f["qc2"] = np.where(f["qc"] != 0, f["qc"] ** 2 / ratio, rng.uniform(0, 1e-9, f["qc"].shape))
    f["inv_qc_relvar"] = np.full(f["qc"].shape, np.nan)  # produced by the chain
    return f


def _device(f, layout):
    import torch
    dev = torch.device("cuda:0")
    t = (lambda a: np.ascontiguousarray(a.T)) if layout == p3x.P3X_COLUMN_MAJOR else (lambda a: a)
    d = {k: torch.from_numpy(t(v) if v.ndim == 2 else v).to(dev) for k, v in f.items()}
    shape = d["qc"].shape
    ncol = f["qc"].shape[1]
    for k in ATM_COMPUTED_FIELDS + CHAIN_OUT:
        d[k] = torch.full(shape, np.nan, dtype=torch.float64, device=dev)
    for k in ATM_FLUX_FIELDS + ("mass_err", "energy_err"):
        d[k] = torch.full((ncol,), np.nan, dtype=torch.float64, device=dev)
    d["budget"] = torch.zeros((4, ncol), dtype=torch.float64, device=dev)
    return d


def _chain(d):
    return {"qc2": d["qc2"], "cldfrac_liq": d["cldfrac_liq"], "cldfrac_ice": d["cldfrac_ice"],
            "cldfrac_ice_4out": d["cldfrac_ice_for_analysis"], "cldfrac_tot_4out": d["cldfrac_tot_for_analysis"],
            "mass_err": d["mass_err"], "energy_err": d["energy_err"], "budget": d["budget"]}


FIELDS_OUT = ATM_UPDATED_FIELDS + ATM_COMPUTED_FIELDS + ATM_FLUX_FIELDS + CHAIN_OUT + (
    "cldfrac_tot", "inv_qc_relvar", "cldfrac_liq", "mass_err", "energy_err")


@pytest.mark.parametrize("precision", [p3x.P3X_FP64, p3x.P3X_MIXED])
def test_served_chain_equals_alone(ctx, oracle_tables, precision):
    import torch
    sp = torch.cuda.current_stream().cuda_stream
    dts = [(0, 300.0), (1, 1800.0), (2, 100.0), (0, 300.0), (1, 1800.0), (2, 100.0), (0, 1800.0)]
    inputs = [_inputs(n, k, s) for n, k, _, s in MEMBERS]
    served = [_device(f, lay) for f, (_, _, lay, _) in zip(inputs, MEMBERS)]
    alone = [_device(f, lay) for f, (_, _, lay, _) in zip(inputs, MEMBERS)]

    # served: every request queued at once, batched by the scheduler
    opts = p3x.server_defaults(precision=precision, slots=2, max_batch=3)
    srv = p3x.Server(ctx, opts)
    procs = [p3x.P3Microphysics(ctx, n, k, layout=lay) for n, k, lay, _ in MEMBERS]
    mids = [srv.add_process(pr, d, _chain(d)) for pr, d in zip(procs, served)]
    print(f"served chain: precision {precision}, {len(dts)} requests", flush=True)
    for tag, (m, dt) in enumerate(dts):
        srv.submit(mids[m], dt, tag)
    srv.drain()
    done = srv.poll(64)
    assert len(done) == len(dts) and all(c.status == p3x.P3X_OK for c in done), [c.status for c in done]
    srv.close()
    torch.cuda.synchronize()

    # alone: the same stages, one member step after another on one stream
    steps = [p3x.MicrophysicsStep(ctx, n, k, layout=lay) for n, k, lay, _ in MEMBERS]
    for m, dt in dts:
        n, k, lay, _ = MEMBERS[m]
        d = alone[m]
        p3x.shoc_postprocess(ctx, d["qc"], d["qc2"], d["inv_qc_relvar"], n, k, cldfrac_liq=d["cldfrac_liq"],
                             stream=sp)
        me, ee = steps[m].run(d, dt, precision=precision, stream=sp)
        d["mass_err"].copy_(me)
        d["energy_err"].copy_(ee)
    torch.cuda.synchronize()
    for m, (a, b) in enumerate(zip(served, alone)):
        for key in FIELDS_OUT:
            np.testing.assert_array_equal(a[key].cpu().numpy(), b[key].cpu().numpy(),
                                          err_msg=f"member {m} {key}")
    for s in steps:
        s.close()
    for p in procs:
        p.close()


def test_served_chain_first_step_matches_oracle(ctx, oracle_tables):
    """Member 0's first request in FP64 against the oracle: SHOC's post-process, then
    CldFraction, the budgets, the process and the errors."""
    import torch
    n, k, lay, seed = MEMBERS[0]
    f = _inputs(n, k, seed)
    d = _device(f, lay)
    srv = p3x.Server(ctx, p3x.server_defaults(precision=p3x.P3X_FP64))
    proc = p3x.P3Microphysics(ctx, n, k)
    mid = srv.add_process(proc, d, _chain(d))
    srv.submit(mid, 300.0, 1)
    srv.drain()
    assert srv.poll(4)[0].status == p3x.P3X_OK
    srv.close()
    torch.cuda.synchronize()
    want = dict(f)
    r, liq = oracle.shoc_postprocess(f["qc"], f["qc2"], f["cldfrac_liq"])
    want["inv_qc_relvar"], want["cldfrac_liq"] = r, liq
    cf = oracle.cld_fraction(want["qi"], want["cldfrac_liq"])
    want["cldfrac_tot"] = cf["cldfrac_tot"]
    m0, e0 = oracle.column_budget(want)
    res, _, nerr = oracle.run_process(want, oracle_tables, 300.0, True, False, math="portable", fluxes=True)
    assert nerr == 0
    for key in ATM_UPDATED_FIELDS + ATM_COMPUTED_FIELDS + ATM_FLUX_FIELDS:
        np.testing.assert_array_equal(d[key].cpu().numpy(), res[key], err_msg=key)
    np.testing.assert_array_equal(d["inv_qc_relvar"].cpu().numpy(), r)
    for key in p3x.CLD_FRACTION_OUT:
        np.testing.assert_array_equal(d[key].cpu().numpy(), cf[key], err_msg=key)
    after = {**want, **{key: res[key] for key in ATM_UPDATED_FIELDS}}
    m1, e1 = oracle.column_budget(after)
    wme, wee = oracle.conservation_errors(m0, e0, m1, e1, res["vapor_flux"], res["water_flux"],
                                          res["ice_flux"], res["heat_flux"], 300.0)
    np.testing.assert_array_equal(d["mass_err"].cpu().numpy(), wme)
    np.testing.assert_array_equal(d["energy_err"].cpu().numpy(), wee)
    proc.close()
