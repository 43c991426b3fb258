"""The batched form (p3x_batch_*): many column sets with different ncol, nlev,
dt and flags advanced in one launch per kernel.  Every set must reproduce, bit
for bit, a p3x_main call on that set alone -- with and without graph capture,
FP64 and mixed, and after the autotuner has run on the batch's data."""
import numpy as np
import pytest

import oracle
from scream_amd import p3x
from scream_amd.inputs import make_columns

from p3_compare import FLOAT_FIELDS, device_fields, field_errors, run_gpu, to_host

pytestmark = pytest.mark.gpu

# ncol, nlev, dt, predictNc, prescribedCCN, seed -- uneven sizes, both level counts
SETS = [(300, 72, 1800.0, True, False, 61), (1, 128, 100.0, True, False, 62),
        (777, 128, 300.0, False, True, 63), (64, 72, 100.0, True, True, 64),
        (1500, 72, 300.0, False, False, 65), (33, 128, 1800.0, True, False, 66)]


@pytest.fixture(scope="module")
def ctx(host_tables):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = p3x.Context(host_tables)
    yield c
    c.close()


def _inputs():
    return [make_columns(n, k, seed=s, dry_every=7, evap_every=9) for n, k, _, _, _, s in SETS]


def _run_batch(ctx, inps, options, tune=False):
    import torch
    devs = [device_fields(inp) for inp in inps]
    sets = [p3x.column_set(d, inp["qc"].shape[1], inp["qc"].shape[0], dt, pnc, pccn, debug=dbg)
            for (d, dbg), inp, (_, _, dt, pnc, pccn, _) in zip(devs, inps, SETS)]
    b = p3x.Batch(ctx, sets, options)
    torch.cuda.synchronize()
    if tune:
        chosen, us = b.tune(reps=1)
        print(f"tuned: {chosen.as_dict()} {us:.1f} us")
    b.run(torch.cuda.current_stream().cuda_stream)
    assert ctx.sync() == 0
    b.close()
    return [to_host(d, dbg) for d, dbg in devs]


def _singles(ctx, inps, precision):
    return [run_gpu(ctx, inp, dt, pnc, pccn, precision=precision)[:2]
            for inp, (_, _, dt, pnc, pccn, _) in zip(inps, SETS)]


def _assert_identical(got, want):
    for s, ((g, gd), (w, wd)) in enumerate(zip(got, want)):
        for k in FLOAT_FIELDS:
            np.testing.assert_array_equal(g[k], w[k], err_msg=f"set {s} {k}")
        for k in wd:
            np.testing.assert_array_equal(gd[k], wd[k], err_msg=f"set {s} {k}")


@pytest.mark.parametrize("use_graph", [0, 1])
@pytest.mark.parametrize("blocks_per_cu", [0, -1, 1])
def test_batch_fp64_equals_single_sets(ctx, use_graph, blocks_per_cu):
    inps = _inputs()
    got = _run_batch(ctx, inps, p3x.batch_defaults(precision=p3x.P3X_FP64, use_graph=use_graph,
                                                   blocks_per_cu=blocks_per_cu))
    _assert_identical(got, _singles(ctx, inps, p3x.P3X_FP64))


def test_batch_mixed_equals_single_sets(ctx):
    inps = _inputs()
    got = _run_batch(ctx, inps, p3x.batch_defaults(precision=p3x.P3X_MIXED))
    _assert_identical(got, _singles(ctx, inps, p3x.P3X_MIXED))


def test_batch_fp64_matches_oracle(ctx, oracle_tables):
    inps = _inputs()
    got = _run_batch(ctx, inps, p3x.batch_defaults(precision=p3x.P3X_FP64, sed_lanes_per_column=16))
    for s, (inp, (_, _, dt, pnc, pccn, _)) in enumerate(zip(inps, SETS)):
        want, wdbg, _ = oracle.run_oracle(inp, oracle_tables, dt, pnc, pccn, math="portable")
        errs = field_errors(got[s][0], want)
        assert max(errs.values()) <= 1e-10, (s, errs)
        for k in wdbg:
            np.testing.assert_array_equal(got[s][1][k], wdbg[k], err_msg=f"set {s} {k}")


def test_tuned_batch_unchanged_results(ctx):
    """The autotuner restores the state it perturbs: the step after tuning equals a
    single-set step on the original inputs."""
    inps = _inputs()
    got = _run_batch(ctx, inps, p3x.batch_defaults(precision=p3x.P3X_MIXED), tune=True)
    _assert_identical(got, _singles(ctx, inps, p3x.P3X_MIXED))


def test_bench_config_mixed_64_sets(ctx):
    """The bench's batched configuration itself: 64 sets, ncol log-uniform in 256-8192,
    nlev from {72, 128}, dt from {100, 300, 1800} s (the same draws, seed 4, inputs seeded
    2000 + set), mixed precision, graph-captured after the autotuner; every set bit for
    bit its own p3x_main call."""
    import torch
    rng = np.random.default_rng(4)
    nsets = 64
    ncols = np.round(np.exp(rng.uniform(np.log(256), np.log(8192), nsets))).astype(int)
    nlevs = rng.choice([72, 128], nsets)
    dts = rng.choice([100.0, 300.0, 1800.0], nsets)
    specs = [(int(n), int(k), float(dt)) for n, k, dt in zip(ncols, nlevs, dts)]
    print(f"bench batch: {nsets} sets, {sum(n * k for n, k, _ in specs)} column-levels", flush=True)
    inps = [make_columns(n, k, seed=2000 + s) for s, (n, k, _) in enumerate(specs)]
    devs = [device_fields(inp) for inp in inps]
    sets = [p3x.column_set(d, n, k, dt, debug=dbg) for (d, dbg), (n, k, dt) in zip(devs, specs)]
    b = p3x.Batch(ctx, sets, p3x.batch_defaults(precision=p3x.P3X_MIXED, use_graph=1))
    chosen, us = b.tune(reps=1)
    print(f"tuned: {chosen.as_dict()} {us:.1f} us", flush=True)
    b.run(torch.cuda.current_stream().cuda_stream)
    assert ctx.sync() == 0
    b.close()
    got = [to_host(d, dbg) for d, dbg in devs]
    del devs, sets
    for s, (inp, (n, k, dt)) in enumerate(zip(inps, specs)):
        g, gd = got[s]
        w, wd = run_gpu(ctx, inp, dt, True, False, precision=p3x.P3X_MIXED)[:2]
        for key in FLOAT_FIELDS:
            np.testing.assert_array_equal(g[key], w[key], err_msg=f"set {s} ({n}x{k}) {key}")
        for key in wd:
            np.testing.assert_array_equal(gd[key], wd[key], err_msg=f"set {s} ({n}x{k}) {key}")
