"""CldFraction on the GPU (p3x_cld_fraction) against the oracle bit for bit, and
the microphysics step around P3 (CldFraction, then P3Microphysics with the
column conservation check, p3x.MicrophysicsStep) against the oracle's chain:
oracle.cld_fraction, oracle.run_process, oracle.column_budget and
oracle.conservation_errors, over consecutive steps."""
import numpy as np
import pytest

import oracle
from scream_amd import p3x
from scream_amd.inputs import ATM_COMPUTED_FIELDS, ATM_FLUX_FIELDS, ATM_UPDATED_FIELDS, make_atm_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(host_tables):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = p3x.Context(host_tables)
    yield c
    c.close()


def _qi_liq(nlev, ncol, seed):
    rng = np.random.default_rng(seed)
    qi = np.where(rng.random((nlev, ncol)) < 0.4, 0.0, 10.0 ** rng.uniform(-16, -2, (nlev, ncol)))
    qi[rng.random((nlev, ncol)) < 0.05] = 1e-12  # exactly at the model threshold
    qi[rng.random((nlev, ncol)) < 0.05] = 1e-5   # exactly at the analysis threshold
    liq = np.where(rng.random((nlev, ncol)) < 0.3, 0.0, rng.random((nlev, ncol)))
    return qi, liq


@pytest.mark.parametrize("ncol,nlev", [(1, 1), (300, 72), (21600, 72)])
def test_cld_fraction_matches_oracle(ctx, ncol, nlev):
    import torch
    dev = torch.device("cuda:0")
    qi, liq = _qi_liq(nlev, ncol, seed=ncol + nlev)
    d = {"qi": torch.from_numpy(qi).to(dev), "cldfrac_liq": torch.from_numpy(liq).to(dev)}
    for k in p3x.CLD_FRACTION_OUT:
        d[k] = torch.full((nlev, ncol), np.nan, dtype=torch.float64, device=dev)
    print(f"cld_fraction {ncol}x{nlev}", flush=True)
    p3x.CldFraction(ctx, ncol, nlev).run(d, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    want = oracle.cld_fraction(qi, liq)
    for k in p3x.CLD_FRACTION_OUT:
        np.testing.assert_array_equal(d[k].cpu().numpy(), want[k], err_msg=k)


def test_microphysics_step_matches_oracle_chain(ctx, oracle_tables):
    import torch
    dev = torch.device("cuda:0")
    ncol, nlev = 333, 72
    f = make_atm_fields(ncol, nlev, seed=17, dry_every=7, evap_every=9)
    # CldFraction's inputs: the liquid fraction SHOC would give, and the state's qi
    rng = np.random.default_rng(5)
    f["cldfrac_liq"] = np.where(f["qc"] > 1e-7, rng.uniform(0.2, 1.0, f["qc"].shape), 0.0)
    d = {k: torch.from_numpy(v).to(dev) for k, v in f.items()}
    for k in ATM_COMPUTED_FIELDS + p3x.CLD_FRACTION_OUT:
        d[k] = torch.full((nlev, ncol), np.nan, dtype=torch.float64, device=dev)
    for k in ATM_FLUX_FIELDS:
        d[k] = torch.full((ncol,), np.nan, dtype=torch.float64, device=dev)
    step = p3x.MicrophysicsStep(ctx, ncol, nlev)
    want = dict(f)
    for n, dt in enumerate((300.0, 1800.0)):
        print(f"microphysics step {n} dt={dt}", flush=True)
        me, ee = step.run(d, dt, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        cf = oracle.cld_fraction(want["qi"], want["cldfrac_liq"])
        for k in p3x.CLD_FRACTION_OUT:
            np.testing.assert_array_equal(d[k].cpu().numpy(), cf[k], err_msg=f"step {n} {k}")
        want["cldfrac_tot"] = cf["cldfrac_tot"]
        m0, e0 = oracle.column_budget(want)
        res, _, nerr = oracle.run_process(want, oracle_tables, dt, True, False, math="portable", fluxes=True)
        assert nerr == 0
        for k in ATM_UPDATED_FIELDS + ATM_COMPUTED_FIELDS + ATM_FLUX_FIELDS:
            np.testing.assert_array_equal(d[k].cpu().numpy(), res[k], err_msg=f"step {n} {k}")
        after = {**want, **{k: res[k] for k in ATM_UPDATED_FIELDS}}
        m1, e1 = oracle.column_budget(after)
        wme, wee = oracle.conservation_errors(m0, e0, m1, e1, res["vapor_flux"], res["water_flux"],
                                              res["ice_flux"], res["heat_flux"], dt)
        np.testing.assert_array_equal(me.cpu().numpy(), wme, err_msg=f"step {n} mass_err")
        np.testing.assert_array_equal(ee.cpu().numpy(), wee, err_msg=f"step {n} energy_err")
        # the step closes the column budgets to roundoff
        assert wme.max() < 1e-14 and wee.max() < 1e-14
        want = after
    step.close()
