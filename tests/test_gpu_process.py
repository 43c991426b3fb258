"""The P3Microphysics process on the GPU (p3x_process_*): pre-amble, p3_main
and post-amble on the field-manager fields must reproduce the oracle's
run_impl chain (oracle.run_process, the kernels' elementary functions) bit for
bit, over consecutive steps."""
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


@pytest.mark.parametrize("ncol,nlev,pnc,pccn,fluxes", [(300, 72, True, False, True),
                                                      (257, 128, False, True, False),
                                                      (64, 72, True, True, True)])
def test_process_matches_oracle(ctx, oracle_tables, ncol, nlev, pnc, pccn, fluxes):
    import torch
    dev = torch.device("cuda:0")
    f = make_atm_fields(ncol, nlev, seed=ncol + nlev, dry_every=7, evap_every=9)
    d = {k: torch.from_numpy(v).to(dev) for k, v in f.items()}
    for k in ATM_COMPUTED_FIELDS:
        d[k] = torch.full((nlev, ncol), np.nan, dtype=torch.float64, device=dev)
    if fluxes:
        for k in ATM_FLUX_FIELDS:
            d[k] = torch.full((ncol,), np.nan, dtype=torch.float64, device=dev)
    proc = p3x.P3Microphysics(ctx, ncol, nlev, predict_nc=pnc, prescribed_ccn=pccn)
    want = dict(f)
    for step, dt in enumerate((300.0, 1800.0)):
        proc.run(d, dt, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        res, _, nerr = oracle.run_process(want, oracle_tables, dt, pnc, pccn, math="portable", fluxes=fluxes)
        assert nerr == 0
        for k in ATM_UPDATED_FIELDS + ATM_COMPUTED_FIELDS + (ATM_FLUX_FIELDS if fluxes else ()):
            np.testing.assert_array_equal(d[k].cpu().numpy(), res[k], err_msg=f"step {step} {k}")
        want = {**want, **{k: res[k] for k in ATM_UPDATED_FIELDS}}
    proc.close()


def test_conservation_check_matches_oracle_and_closes(ctx, oracle_tables):
    """p3x_column_budget / p3x_conservation_errors reproduce the oracle bit for bit,
    and a GPU process step passes the check to roundoff."""
    import torch
    dev = torch.device("cuda:0")
    ncol, nlev, dt = 400, 72, 1800.0
    f = make_atm_fields(ncol, nlev, seed=91, dry_every=7, evap_every=9)
    d = {k: torch.from_numpy(v).to(dev) for k, v in f.items()}
    for k in ATM_COMPUTED_FIELDS:
        d[k] = torch.zeros((nlev, ncol), dtype=torch.float64, device=dev)
    for k in ATM_FLUX_FIELDS:
        d[k] = torch.zeros((ncol,), dtype=torch.float64, device=dev)
    vec = lambda: torch.zeros(ncol, dtype=torch.float64, device=dev)
    m0, e0, m1, e1, me, ee = vec(), vec(), vec(), vec(), vec(), vec()
    sp = torch.cuda.current_stream().cuda_stream
    p3x.column_budget(ctx, d, ncol, nlev, m0, e0, sp)
    proc = p3x.P3Microphysics(ctx, ncol, nlev)
    proc.run(d, dt, stream=sp)
    p3x.column_budget(ctx, d, ncol, nlev, m1, e1, sp)
    p3x.conservation_errors(ctx, m0, e0, m1, e1, d["vapor_flux"], d["water_flux"], d["ice_flux"],
                            d["heat_flux"], ncol, dt, me, ee, sp)
    torch.cuda.synchronize()
    proc.close()
    wm0, we0 = oracle.column_budget(f)
    np.testing.assert_array_equal(m0.cpu().numpy(), wm0)
    np.testing.assert_array_equal(e0.cpu().numpy(), we0)
    after = {k: d[k].cpu().numpy() for k in ("T_mid", "qv", "qc", "qi", "qr")}
    after["pseudo_density"] = f["pseudo_density"]
    wm1, we1 = oracle.column_budget(after)
    np.testing.assert_array_equal(m1.cpu().numpy(), wm1)
    np.testing.assert_array_equal(e1.cpu().numpy(), we1)
    assert me.max().item() < 1e-14 and ee.max().item() < 1e-14
