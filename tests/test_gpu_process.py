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
