"""The reference's own field layout at the boundary: EAMxx hands p3_main its
field-manager views as (column, level) LayoutRight arrays, pack-padded per
column (eamxx_p3_process_interface.cpp:254-265, eamxx_p3_run.cpp:31-36).
p3x_main, the batched form and the P3Microphysics process take them as
P3X_COLUMN_MAJOR with the views' column strides, and must give the oracle's
results bit for bit -- the same bits as the level-major layout -- while the
padding of every array stays untouched."""
import numpy as np
import pytest

import oracle
from scream_amd import p3x
from scream_amd.inputs import ATM_COMPUTED_FIELDS, ATM_FLUX_FIELDS, ATM_UPDATED_FIELDS, make_atm_fields, make_columns

from p3_compare import FLOAT_FIELDS, device_fields, run_gpu, to_host

pytestmark = pytest.mark.gpu

SENTINEL = -12345.5
IFACE = ("precip_liq_flux", "precip_ice_flux")
SURF = ("precip_liq_surf", "precip_ice_surf")


@pytest.fixture(scope="module")
def ctx(host_tables):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = p3x.Context(host_tables)
    yield c
    c.close()


def _cm(a, ld, fill=SENTINEL):
    """Level-major (levels, ncol) -> column-major (ncol, ld) with padding `fill`."""
    lev, ncol = a.shape
    out = np.full((ncol, ld), fill)
    out[:, :lev] = a.T
    return out


def cm_device_fields(inp, ld, ld_iface, exch=None):
    """Column-major device copies of the level-major ``inp`` (padded to ld / ld_iface
    with a sentinel), NaN outputs, zero (or ``exch``) exchange fields and integer
    debug arrays; returns (fields, debug)."""
    import torch
    dev = torch.device("cuda:0")
    nlev, ncol = inp["qc"].shape
    d = {k: torch.from_numpy(_cm(inp[k], ld)).to(dev) for k in p3x.STATE + p3x.INPUTS}
    for k in p3x.OUTS + p3x.HIST:
        if k in SURF:
            d[k] = torch.full((ncol,), np.nan, dtype=torch.float64, device=dev)
        elif k in IFACE:
            d[k] = torch.from_numpy(_cm(np.full((nlev + 1, ncol), np.nan), ld_iface)).to(dev)
        elif k in p3x.HIST:
            src = exch[k] if exch is not None else np.zeros((nlev, ncol))
            d[k] = torch.from_numpy(_cm(src, ld)).to(dev)
        else:
            d[k] = torch.from_numpy(_cm(np.full((nlev, ncol), np.nan), ld)).to(dev)
    idx = np.full((ncol, ld), -9, np.int32)
    idx[:, :nlev] = -7
    dbg = dict(col_status=torch.full((ncol,), -7, dtype=torch.int32, device=dev),
               sed_substeps=torch.full((3, ncol), -7, dtype=torch.int32, device=dev),
               ice_table_idx=torch.from_numpy(idx).to(dev))
    return d, dbg


def cm_to_host(d, dbg, nlev):
    """Back to level-major numpy, asserting the padding kept its sentinel."""
    out = {}
    for k, v in d.items():
        a = v.cpu().numpy()
        if a.ndim == 2:
            lev = nlev + 1 if k in IFACE else nlev
            pad = a[:, lev:]
            assert np.all(pad == SENTINEL), f"padding of {k} was written"
            a = np.ascontiguousarray(a[:, :lev].T)
        out[k] = a
    g = {k: v.cpu().numpy() for k, v in dbg.items()}
    assert np.all(g["ice_table_idx"][:, nlev:] == -9), "padding of ice_table_idx was written"
    g["ice_table_idx"] = np.ascontiguousarray(g["ice_table_idx"][:, :nlev].T)
    return out, g


def run_cm(ctx, inp, dt, pnc, pccn, ld, ld_iface, exch=None, **kw):
    import torch
    nlev, ncol = inp["qc"].shape
    d, dbg = cm_device_fields(inp, ld, ld_iface, exch)
    torch.cuda.synchronize()
    ctx.main(d, ncol, nlev, dt, pnc, pccn, debug=dbg, stream=torch.cuda.current_stream().cuda_stream,
             layout=p3x.P3X_COLUMN_MAJOR, ld=ld if ld != nlev else 0, ld_iface=ld_iface, **kw)
    torch.cuda.synchronize()
    return cm_to_host(d, dbg, nlev)


def _assert_equal(got, gdbg, want, wdbg, label):
    for k in FLOAT_FIELDS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{label} {k}")
    for k in ("col_status", "sed_substeps", "ice_table_idx"):
        np.testing.assert_array_equal(gdbg[k], wdbg[k], err_msg=f"{label} {k}")


# ncol, nlev, dt, predictNc, prescribedCCN, seed, ld, ld_iface: dense views (pack size 1,
# as on the GPU) and views padded to a pack of 16 (npack(nlev) * 16, npack(nlev + 1) * 16)
CASES = [(256, 72, 300.0, True, False, 11, 72, 73),
         (300, 72, 1800.0, True, False, 12, 80, 80),
         (257, 128, 100.0, False, True, 13, 128, 144),
         (200, 72, 1800.0, False, False, 14, 80, 80),
         (97, 3, 1800.0, True, True, 15, 16, 16)]


@pytest.mark.parametrize("ncol,nlev,dt,pnc,pccn,seed,ld,ldi", CASES)
def test_column_major_matches_oracle(ctx, oracle_tables, ncol, nlev, dt, pnc, pccn, seed, ld, ldi):
    if nlev < 72:
        inp = {k: np.ascontiguousarray(v[62:62 + nlev]) for k, v in
               make_columns(ncol, 72, seed=seed, dry_every=7, evap_every=9).items()}
    else:
        inp = make_columns(ncol, nlev, seed=seed, dry_every=7, evap_every=9)
    print(f"column-major {ncol}x{nlev} ld {ld} ld_iface {ldi} dt {dt}", flush=True)
    want, wdbg, nerr = oracle.run_oracle(inp, oracle_tables, dt, pnc, pccn, math="portable")
    assert nerr == 0
    got, gdbg = run_cm(ctx, inp, dt, pnc, pccn, ld, ldi)
    _assert_equal(got, gdbg, want, wdbg, f"cm {ncol}x{nlev}")


def test_column_major_history_and_shapes(ctx, oracle_tables):
    """Nonzero exchange fields in, an explicit block shape for part 1 and 64
    sedimentation lanes: still the oracle's bits."""
    inp = make_columns(130, 72, seed=17, dry_every=5, evap_every=6)
    rng = np.random.default_rng(17)
    exch = {k: rng.uniform(-1e-7, 1e-7, inp["qc"].shape) for k in p3x.HIST}
    want, wdbg, _ = oracle.run_oracle({**inp, **{k: v.copy() for k, v in exch.items()}},
                                      oracle_tables, 900.0, True, False, math="portable")
    for cb, nt, lpc in ((0, 0, 0), (32, 256, 64), (16, 128, 16)):
        print(f"column-major history, block {cb}x{nt} lanes {lpc}", flush=True)
        got, gdbg = run_cm(ctx, inp, 900.0, True, False, 80, 80, exch=exch, columns_per_block=cb,
                           threads_per_block=nt, sed_lanes_per_column=lpc)
        _assert_equal(got, gdbg, want, wdbg, f"history {cb}x{nt} {lpc}")


def test_column_major_mixed_equals_level_major(ctx):
    """The mixed form does the same arithmetic in either layout."""
    inp = make_columns(700, 128, seed=31, dry_every=7, evap_every=9)
    lm, lmdbg, _ = run_gpu(ctx, inp, 100.0, True, False, precision=p3x.P3X_MIXED)
    got, gdbg = run_cm(ctx, inp, 100.0, True, False, 128, 129, precision=p3x.P3X_MIXED)
    _assert_equal(got, gdbg, lm, lmdbg, "mixed cm vs lm")


def test_column_major_full_size_g64(ctx, oracle_tables):
    """21600 x 72, dt 1800 s, dense column-major views: the oracle's bits."""
    inp = make_columns(21600, 72, seed=1)
    print("column-major G64", flush=True)
    want, wdbg, _ = oracle.run_oracle(inp, oracle_tables, 1800.0, True, False, math="portable")
    got, gdbg = run_cm(ctx, inp, 1800.0, True, False, 72, 73)
    _assert_equal(got, gdbg, want, wdbg, "cm G64")


def test_column_major_bad_strides_refused(ctx):
    import torch
    dev = torch.device("cuda:0")
    f = {k: torch.zeros((4, 80), dtype=torch.float64, device=dev)
         for k in p3x.STATE + p3x.INPUTS + p3x.OUTS + p3x.HIST}
    for ld, ldi in ((71, 0), (72, 72), (80, 40)):
        with pytest.raises(p3x.P3Error) as e:
            ctx.main(f, 4, 72, 300.0, layout=p3x.P3X_COLUMN_MAJOR, ld=ld, ld_iface=ldi)
        assert e.value.status == p3x.P3X_ERR_ARG
    with pytest.raises(p3x.P3Error):
        ctx.main(f, 4, 72, 300.0, layout=7)
    with pytest.raises(p3x.P3Error):  # level-major interface rows share the row stride
        ctx.main(f, 4, 72, 300.0, ld_iface=9)


def test_batch_mixes_layouts(ctx):
    """A batch whose sets come in both layouts: each set equals its own p3x_main."""
    import torch
    specs = [(300, 72, 1800.0, p3x.P3X_COLUMN_MAJOR, 80, 80), (64, 128, 100.0, p3x.P3X_LEVEL_MAJOR, 0, 0),
             (777, 128, 300.0, p3x.P3X_COLUMN_MAJOR, 128, 129), (1, 72, 300.0, p3x.P3X_COLUMN_MAJOR, 72, 73)]
    inps = [make_columns(n, k, seed=200 + j, dry_every=7, evap_every=9) for j, (n, k, *_) in enumerate(specs)]
    for prec in (p3x.P3X_FP64, p3x.P3X_MIXED):
        devs, sets = [], []
        for inp, (n, k, dt, lay, ld, ldi) in zip(inps, specs):
            if lay == p3x.P3X_COLUMN_MAJOR:
                d, dbg = cm_device_fields(inp, ld, ldi)
            else:
                d, dbg = device_fields(inp)
            devs.append((d, dbg))
            sets.append(p3x.column_set(d, n, k, dt, debug=dbg, layout=lay, ld=ld, ld_iface=ldi))
        b = p3x.Batch(ctx, sets, p3x.batch_defaults(precision=prec))
        torch.cuda.synchronize()
        b.run(torch.cuda.current_stream().cuda_stream)
        assert ctx.sync() == 0
        b.close()
        for j, ((d, dbg), inp, (n, k, dt, lay, ld, ldi)) in enumerate(zip(devs, inps, specs)):
            want, wdbg, _ = run_gpu(ctx, inp, dt, True, False, precision=prec)
            got, gdbg = cm_to_host(d, dbg, k) if lay == p3x.P3X_COLUMN_MAJOR else to_host(d, dbg)
            _assert_equal(got, gdbg, want, wdbg, f"batch set {j} precision {prec}")


@pytest.mark.parametrize("ncol,nlev,ld,ldi,pnc,pccn", [(300, 72, 80, 80, True, False), (257, 128, 128, 129, False, True)])
def test_process_column_major_matches_oracle(ctx, oracle_tables, ncol, nlev, ld, ldi, pnc, pccn):
    """P3Microphysics on column-major field-manager views (pack-padded): the
    pre-amble, p3_main and post-amble reproduce the oracle's run_impl bit for bit
    over two steps, and the budget check reads the same views."""
    import torch
    dev = torch.device("cuda:0")
    f = make_atm_fields(ncol, nlev, seed=ncol + nlev, dry_every=7, evap_every=9)
    d = {}
    for k, v in f.items():
        d[k] = torch.from_numpy(_cm(v, ld) if v.ndim == 2 else v.copy()).to(dev)
    for k in ATM_COMPUTED_FIELDS:
        d[k] = torch.from_numpy(_cm(np.full((nlev, ncol), np.nan), ld)).to(dev)
    for k in ATM_FLUX_FIELDS:
        d[k] = torch.full((ncol,), np.nan, dtype=torch.float64, device=dev)
    proc = p3x.P3Microphysics(ctx, ncol, nlev, predict_nc=pnc, prescribed_ccn=pccn,
                              layout=p3x.P3X_COLUMN_MAJOR, ld=ld, ld_iface=ldi)
    sp = torch.cuda.current_stream().cuda_stream
    vec = lambda: torch.zeros(ncol, dtype=torch.float64, device=dev)
    m0, e0 = vec(), vec()
    p3x.column_budget(ctx, d, ncol, nlev, m0, e0, sp, layout=p3x.P3X_COLUMN_MAJOR, ld=ld)
    want = dict(f)
    wm0, we0 = oracle.column_budget(f)
    for step, dt in enumerate((300.0, 1800.0)):
        print(f"process column-major {ncol}x{nlev} ld {ld} step {step}", flush=True)
        proc.run(d, dt, stream=sp)
        torch.cuda.synchronize()
        res, _, nerr = oracle.run_process(want, oracle_tables, dt, pnc, pccn, math="portable", fluxes=True)
        assert nerr == 0
        for k in ATM_UPDATED_FIELDS + ATM_COMPUTED_FIELDS + ATM_FLUX_FIELDS:
            a = d[k].cpu().numpy()
            if a.ndim == 2:
                assert np.all(a[:, nlev:] == SENTINEL), f"padding of {k}"
                a = a[:, :nlev].T
            np.testing.assert_array_equal(a, res[k], err_msg=f"step {step} {k}")
        want = {**want, **{k: res[k] for k in ATM_UPDATED_FIELDS}}
    proc.close()
    np.testing.assert_array_equal(m0.cpu().numpy(), wm0)
    np.testing.assert_array_equal(e0.cpu().numpy(), we0)
