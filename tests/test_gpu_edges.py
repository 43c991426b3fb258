"""Edge shapes of p3x_main against the oracle: one column, one to three levels
(cut from the lower troposphere of a 72-level column so that cloud, rain and
ice are all present), the largest supported level count (512, 64 lanes of 8
levels in sedimentation), a padded leading dimension (ld > ncol; the padding
must stay untouched), and an empty column set (nj = 0, nothing launched)."""
import numpy as np
import pytest

import oracle
from scream_amd import p3x
from scream_amd.inputs import make_columns

from p3_compare import FLOAT_FIELDS, field_errors, run_gpu

pytestmark = pytest.mark.gpu

RTOL_FP64 = 1e-10
MATH = "portable"


@pytest.fixture(scope="module")
def ctx(host_tables):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = p3x.Context(host_tables)
    yield c
    c.close()


def _levels(inp, k0, k1):
    return {k: np.ascontiguousarray(v[k0:k1]) for k, v in inp.items()}


def _check(got, gdbg, want, wdbg, label):
    errs = field_errors(got, want)
    bad = {k: v for k, v in errs.items() if not v <= RTOL_FP64}
    assert not bad, f"[{label}] fields beyond {RTOL_FP64}: {bad}"
    for k in ("col_status", "sed_substeps", "ice_table_idx"):
        np.testing.assert_array_equal(gdbg[k], wdbg[k], err_msg=f"{label} {k}")


@pytest.mark.parametrize("ncol,k0,k1,dt", [(1, 66, 67, 1800.0), (1, 0, 1, 300.0), (2, 63, 65, 300.0),
                                           (7, 62, 65, 1800.0), (33, 56, 72, 1800.0)])
def test_few_levels(ctx, oracle_tables, ncol, k0, k1, dt):
    inp = _levels(make_columns(ncol, 72, seed=50 + ncol, dry_every=7, evap_every=9), k0, k1)
    print(f"few levels: {ncol} columns, levels {k0}..{k1 - 1}, dt {dt}", flush=True)
    want, wdbg, _ = oracle.run_oracle(inp, oracle_tables, dt, True, False, math=MATH)
    got, gdbg, _ = run_gpu(ctx, inp, dt, True, False)
    _check(got, gdbg, want, wdbg, f"{ncol}x{k1 - k0}")


def test_most_levels(ctx, oracle_tables):
    ncol, nlev, dt = 65, 512, 100.0
    inp = make_columns(ncol, nlev, seed=61, dry_every=7, evap_every=9)
    print(f"most levels: {ncol}x{nlev}", flush=True)
    want, wdbg, _ = oracle.run_oracle(inp, oracle_tables, dt, True, False, math=MATH)
    got, gdbg, _ = run_gpu(ctx, inp, dt, True, False)
    _check(got, gdbg, want, wdbg, f"{ncol}x{nlev}")


def test_too_many_levels_is_an_argument_error(ctx):
    import torch
    dev = torch.device("cuda:0")
    nlev, ncol = 513, 4
    f = {k: torch.zeros((nlev + 1, ncol), dtype=torch.float64, device=dev)
         for k in p3x.STATE + p3x.INPUTS + p3x.OUTS + p3x.HIST}
    with pytest.raises(p3x.P3Error):
        ctx.main(f, ncol, nlev, 300.0)


def test_padded_leading_dimension(ctx, oracle_tables):
    """Level-major arrays with ld = ncol + 13: the first ncol entries of each row
    are the columns; the 13 padding entries hold a sentinel that must survive."""
    import torch
    dev = torch.device("cuda:0")
    ncol, nlev, ld, dt = 300, 72, 313, 1800.0
    inp = make_columns(ncol, nlev, seed=71, dry_every=7, evap_every=9)
    sentinel = -12345.5
    d = {}
    for k in p3x.STATE + p3x.INPUTS:
        a = np.full((nlev, ld), sentinel)
        a[:, :ncol] = inp[k]
        d[k] = torch.from_numpy(a).to(dev)
    for k in p3x.OUTS + p3x.HIST:
        if k in ("precip_liq_surf", "precip_ice_surf"):
            d[k] = torch.full((ncol,), np.nan, dtype=torch.float64, device=dev)
            continue
        rows = nlev + 1 if k in ("precip_liq_flux", "precip_ice_flux") else nlev
        a = np.full((rows, ld), sentinel)
        a[:, :ncol] = 0.0 if k in p3x.HIST else np.nan
        d[k] = torch.from_numpy(a).to(dev)
    print(f"padded: {ncol}x{nlev} ld {ld}", flush=True)
    ctx.main(d, ncol, nlev, dt, True, False, ld=ld, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    want, _, _ = oracle.run_oracle(inp, oracle_tables, dt, True, False, math=MATH)
    got = {}
    for k, v in d.items():
        a = v.cpu().numpy()
        if a.ndim == 2:
            np.testing.assert_array_equal(a[:, ncol:], sentinel, err_msg=f"padding of {k}")
            a = np.ascontiguousarray(a[:, :ncol])
        got[k] = a
    errs = field_errors(got, want)
    bad = {k: v for k, v in errs.items() if not v <= RTOL_FP64}
    assert not bad, bad


def test_no_columns(ctx):
    import torch
    dev = torch.device("cuda:0")
    nlev = 72
    f = {k: torch.full((nlev + 1, 1), 7.0, dtype=torch.float64, device=dev)
         for k in p3x.STATE + p3x.INPUTS + p3x.OUTS + p3x.HIST}
    ctx.main(f, 0, nlev, 300.0)
    torch.cuda.synchronize()
    for k, v in f.items():
        assert bool((v == 7.0).all()), k


def test_async_calls_on_two_streams_share_scratch_safely(ctx, oracle_tables):
    """Two asynchronous p3x_main calls of one context on two streams (different
    column sets, same scratch) give the results each gives alone."""
    import torch
    from p3_compare import device_fields
    inp_a = make_columns(700, 72, seed=81, dry_every=7, evap_every=9)
    inp_b = make_columns(500, 128, seed=82, dry_every=7, evap_every=9)
    want_a, _, _ = oracle.run_oracle(inp_a, oracle_tables, 1800.0, True, False, math=MATH)
    want_b, _, _ = oracle.run_oracle(inp_b, oracle_tables, 300.0, True, False, math=MATH)
    da, dbg_a = device_fields(inp_a)
    db, dbg_b = device_fields(inp_b)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    print("two streams: 700x72 then 500x128, async", flush=True)
    ctx.main(da, 700, 72, 1800.0, True, False, stream=sa.cuda_stream, async_=True, debug=dbg_a)
    ctx.main(db, 500, 128, 300.0, True, False, stream=sb.cuda_stream, async_=True, debug=dbg_b)
    assert ctx.sync() == 0
    torch.cuda.synchronize()
    for got, want in ((da, want_a), (db, want_b)):
        errs = field_errors({k: v.cpu().numpy() for k, v in got.items()}, want)
        bad = {k: v for k, v in errs.items() if not v <= RTOL_FP64}
        assert not bad, bad


@pytest.mark.parametrize("bad", [-5.0, float("nan")])
def test_bad_temperature_is_reported_not_fatal(ctx, oracle_tables, bad):
    """A level with T <= 0 or NaN (the reference's EKAT_KERNEL_REQUIRE abort in
    check_temperature, share/physics_saturation_impl.hpp:30-35) makes the call
    return P3X_ERR_TEMPERATURE; the context stays usable and the next call on
    good columns is exact."""
    import torch
    from p3_compare import device_fields
    inp = make_columns(64, 72, seed=91)
    inp["th"] = inp["th"].copy()
    inp["th"][40, 17] = bad
    d, dbg = device_fields(inp)
    print(f"bad temperature {bad}", flush=True)
    with pytest.raises(p3x.P3Error) as e:
        ctx.main(d, 64, 72, 300.0, True, False, stream=torch.cuda.current_stream().cuda_stream, debug=dbg)
    assert e.value.status == p3x.P3X_ERR_TEMPERATURE
    good = make_columns(64, 72, seed=92, dry_every=7, evap_every=9)
    want, _, _ = oracle.run_oracle(good, oracle_tables, 300.0, True, False, math=MATH)
    got, _, _ = run_gpu(ctx, good, 300.0, True, False)
    errs = field_errors(got, want)
    assert max(errs.values()) <= RTOL_FP64, errs


def test_reserve_then_smaller_sets(ctx, oracle_tables):
    """p3x_reserve sizes the scratch once; later calls within it give exact results,
    and shapes beyond the supported levels are refused."""
    with pytest.raises(p3x.P3Error):
        ctx.reserve(100, 513)
    with pytest.raises(p3x.P3Error):
        ctx.reserve(0, 72)
    ctx.reserve(4096, 128)
    for ncol, nlev, seed in ((300, 72, 101), (1000, 128, 102)):
        inp = make_columns(ncol, nlev, seed=seed, dry_every=7, evap_every=9)
        want, _, _ = oracle.run_oracle(inp, oracle_tables, 300.0, True, False, math=MATH)
        got, _, _ = run_gpu(ctx, inp, 300.0, True, False)
        errs = field_errors(got, want)
        assert max(errs.values()) <= RTOL_FP64, errs


def _check_values_expected(inp, out, dbg):
    """check_values' flags per column from a step's outputs: bit 0 T = th exner outside
    (173, 323) K, bit 1 qv outside (0, 40e-3), on the fully stepped columns only
    (components/eamxx/src/physics/p3/impl/p3_check_values_impl.hpp:35-60)."""
    T = out["th"] * (1 / inp["inv_exner"])
    qv = out["qv"]
    f = (~((T > 173.0) & (T < 323.0))).astype(np.int32) | 2 * (~((qv > 0.0) & (qv < 40.0e-3))).astype(np.int32)
    return np.where(dbg["col_status"] == 2, np.bitwise_or.reduce(f, axis=0), 0).astype(np.int32)


def test_check_values_flags_and_trap(ctx, oracle_tables):
    """The opt-in check_values: qv out of bounds is a flag only (status OK); a
    temperature out of bounds is the trap (P3X_ERR_CHECK_VALUES); the flags equal the
    bounds applied to the oracle's outputs, and without the option nothing is checked."""
    import torch
    from p3_compare import device_fields
    # column 12 runs the full step (columns 0, 7, 14, ... are dry, 1, 10, 19, ... skip
    # after part 2, and check_values looks at fully stepped columns only)
    inp = make_columns(128, 72, seed=93, dry_every=7, evap_every=9)
    inp["qv"] = inp["qv"].copy()
    inp["qv"][60, 12] = 0.05
    want, wdbg, _ = oracle.run_oracle(inp, oracle_tables, 300.0, True, False, math=MATH)
    d, dbg = device_fields(inp)
    dbg["check_values"] = torch.full((128,), -7, dtype=torch.int32, device="cuda:0")
    print("check_values: qv flag", flush=True)
    ctx.main(d, 128, 72, 300.0, True, False, debug=dbg, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    exp = _check_values_expected(inp, want, wdbg)
    assert wdbg["col_status"][12] == 2
    assert exp[12] & 2 and not (exp & 1).any()
    np.testing.assert_array_equal(dbg["check_values"].cpu().numpy(), exp)

    hot = make_columns(128, 72, seed=94, dry_every=7, evap_every=9)
    hot["th"] = hot["th"].copy()
    hot["th"][70, 3] = 330.0 * hot["inv_exner"][70, 3]
    want, wdbg, _ = oracle.run_oracle(hot, oracle_tables, 300.0, True, False, math=MATH)
    d, dbg = device_fields(hot)
    dbg["check_values"] = torch.full((128,), -7, dtype=torch.int32, device="cuda:0")
    print("check_values: T trap", flush=True)
    with pytest.raises(p3x.P3Error) as e:
        ctx.main(d, 128, 72, 300.0, True, False, debug=dbg, stream=torch.cuda.current_stream().cuda_stream)
    assert e.value.status == p3x.P3X_ERR_CHECK_VALUES
    exp = _check_values_expected(hot, want, wdbg)
    assert exp[3] & 1
    np.testing.assert_array_equal(dbg["check_values"].cpu().numpy(), exp)
    # the step itself is the plain one, and without the option the same step is OK
    errs = field_errors({k: v.cpu().numpy() for k, v in d.items()}, want)
    assert max(errs.values()) <= RTOL_FP64, errs
    got, _, _ = run_gpu(ctx, hot, 300.0, True, False)
    assert max(field_errors(got, want).values()) <= RTOL_FP64
