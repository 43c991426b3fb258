"""Helpers shared by the parity tests: run the HIP path, compare with the oracle."""
import numpy as np

from scream_amd import p3x

FLOAT_FIELDS = p3x.STATE + p3x.OUTS + p3x.HIST


def device_fields(inp, exch=None):
    """Device copies of ``inp`` plus NaN-filled outputs, zeroed (or ``exch``) exchange
    fields, and integer debug arrays filled with -7; returns (fields, debug)."""
    import torch
    dev = torch.device("cuda:0")
    nlev, ncol = inp["qc"].shape
    d = {k: torch.from_numpy(np.ascontiguousarray(inp[k])).to(dev) for k in p3x.STATE + p3x.INPUTS}
    for k in p3x.OUTS + p3x.HIST:
        if k in ("precip_liq_surf", "precip_ice_surf"):
            shape = (ncol,)
        elif k in ("precip_liq_flux", "precip_ice_flux"):
            shape = (nlev + 1, ncol)
        else:
            shape = (nlev, ncol)
        d[k] = torch.full(shape, np.nan, dtype=torch.float64, device=dev)
    for k in p3x.HIST:
        d[k] = (torch.from_numpy(np.ascontiguousarray(exch[k])).to(dev) if exch is not None
                else torch.zeros((nlev, ncol), dtype=torch.float64, device=dev))
    dbg = dict(col_status=torch.full((ncol,), -7, dtype=torch.int32, device=dev),
               sed_substeps=torch.full((3, ncol), -7, dtype=torch.int32, device=dev),
               ice_table_idx=torch.full((nlev, ncol), -7, dtype=torch.int32, device=dev))
    return d, dbg


def to_host(d, dbg):
    return ({k: v.cpu().numpy() for k, v in d.items()}, {k: v.cpu().numpy() for k, v in dbg.items()})


def run_gpu(ctx, inp, dt, predict_nc=True, prescribed_ccn=False, exch=None, **kw):
    """One p3x_main on device copies of ``inp``; returns (fields, debug, microseconds)."""
    import torch
    nlev, ncol = inp["qc"].shape
    d, dbg = device_fields(inp, exch)
    torch.cuda.synchronize()
    us = ctx.main(d, ncol, nlev, dt, predict_nc, prescribed_ccn, debug=dbg,
                  stream=torch.cuda.current_stream().cuda_stream, **kw)
    torch.cuda.synchronize()
    return ({k: v.cpu().numpy() for k, v in d.items()},
            {k: v.cpu().numpy() for k, v in dbg.items()}, us)


def rel_err(a, b):
    """Pointwise |a-b| / max(|a|,|b|) (0 where both are 0); NaN positions must agree."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    if (nan_a != nan_b).any():
        return np.inf
    a = np.where(nan_a, 0, a)
    b = np.where(nan_b, 0, b)
    s = np.maximum(np.abs(a), np.abs(b))
    d = np.abs(a - b)
    r = np.where(s > 0, d / np.where(s > 0, s, 1), 0.0)
    return float(r.max()) if r.size else 0.0


def field_errors(got, want, fields=FLOAT_FIELDS):
    return {k: rel_err(got[k], want[k]) for k in fields}


def water_budget(inp, out, dt):
    """Column total water before and after (kg/m2), the reference's definition
    (share/property_checks/mass_and_energy_column_conservation_check.cpp:155-175,283-313):
    sum_k (qv+qc+qr+qi) dp/g  +  (precip_liq_surf+precip_ice_surf) rho_w dt."""
    g = 9.80616
    dp = inp["dpres"]
    before = ((inp["qv"] + inp["qc"] + inp["qr"] + inp["qi"]) * dp / g).sum(axis=0)
    after = ((out["qv"] + out["qc"] + out["qr"] + out["qi"]) * dp / g).sum(axis=0)
    after = after + (out["precip_liq_surf"] + out["precip_ice_surf"]) * 1000.0 * dt
    return before, after
