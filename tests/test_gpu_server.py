"""The served form (p3x_server_*): step requests from several ensemble members,
interleaved and continuously batched over concurrent slots, must leave every
member exactly where running its own steps alone (p3x_main) leaves it."""
import numpy as np
import pytest

from scream_amd import p3x
from scream_amd.inputs import make_columns

from p3_compare import FLOAT_FIELDS, device_fields, run_gpu, to_host

pytestmark = pytest.mark.gpu

# ncol, nlev, predictNc, prescribedCCN, seed
MEMBERS = [(300, 128, True, False, 71), (64, 72, True, True, 72), (1500, 128, False, False, 73),
           (1, 72, True, False, 74), (777, 128, True, False, 75), (129, 72, False, True, 76)]
DTS = [300.0, 100.0, 1800.0]  # three steps per member


@pytest.fixture(scope="module")
def ctx(host_tables):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = p3x.Context(host_tables)
    yield c
    c.close()


def _alone(ctx, inp, pnc, pccn, precision):
    """Member stepped alone through p3x_main, DTS in order."""
    import torch
    d, dbg = device_fields(inp)
    nlev, ncol = inp["qc"].shape
    for dt in DTS:
        for k in p3x.HIST:
            d[k].zero_()
        ctx.main(d, ncol, nlev, dt, pnc, pccn, precision=precision, debug=dbg,
                 stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return to_host(d, dbg)


def _served(ctx, inps, precision, **opts):
    import torch
    srv = p3x.Server(ctx, p3x.server_defaults(precision=precision, **opts))
    devs = []
    for inp, (ncol, nlev, pnc, pccn, _) in zip(inps, MEMBERS):
        d, dbg = device_fields(inp)
        devs.append((d, dbg))
        srv.add_member(p3x.column_set(d, ncol, nlev, 1.0, pnc, pccn, exchange_zeroed=True, debug=dbg))
    torch.cuda.synchronize()
    tag = 0
    for dt in DTS:
        for m in range(len(MEMBERS)):
            srv.submit(m, dt, tag)
            tag += 1
    srv.drain()
    done = srv.poll(max_n=1024)
    stats = srv.stats()
    srv.close()
    assert len(done) == tag and all(c.status == 0 for c in done)
    assert stats.completed == tag and stats.p99_us >= stats.p50_us > 0
    # per member, completions come back in submission order
    for m in range(len(MEMBERS)):
        tags = [c.tag for c in done if c.member == m]
        assert tags == sorted(tags)
    print(f"served {tag} requests: p50 {stats.p50_us:.0f} us p99 {stats.p99_us:.0f} us")
    return [to_host(d, dbg) for d, dbg in devs]


def _check(got, want):
    for m, ((g, gd), (w, wd)) in enumerate(zip(got, want)):
        for k in FLOAT_FIELDS:  # exchange fields too: zeroed-on-entry writes every entry
            np.testing.assert_array_equal(g[k], w[k], err_msg=f"member {m} {k}")
        for k in wd:
            np.testing.assert_array_equal(gd[k], wd[k], err_msg=f"member {m} {k}")


@pytest.mark.parametrize("precision", [p3x.P3X_FP64, p3x.P3X_MIXED])
@pytest.mark.parametrize("slots,max_batch", [(4, 32), (2, 2), (3, 1)])
def test_served_equals_alone(ctx, precision, slots, max_batch):
    inps = [make_columns(n, k, seed=s, dry_every=7, evap_every=9) for n, k, _, _, s in MEMBERS]
    want = [_alone(ctx, inp, pnc, pccn, precision) for inp, (_, _, pnc, pccn, _) in zip(inps, MEMBERS)]
    got = _served(ctx, inps, precision, slots=slots, max_batch=max_batch)
    _check(got, want)


def test_scratch_budget_admission(ctx):
    """A budget that holds only a couple of members per slot still serves every request,
    with the same results; a member that can never fit is refused at registration."""
    inps = [make_columns(n, k, seed=s, dry_every=7, evap_every=9) for n, k, _, _, s in MEMBERS]
    want = [_alone(ctx, inp, pnc, pccn, p3x.P3X_FP64) for inp, (_, _, pnc, pccn, _) in zip(inps, MEMBERS)]
    big = 7 * 128 * 1500 * 8 + 1500 * 4  # scratch of the largest member
    got = _served(ctx, inps, p3x.P3X_FP64, slots=2, scratch_budget_bytes=2 * 1.2 * big)
    _check(got, want)
    srv = p3x.Server(ctx, p3x.server_defaults(slots=2, scratch_budget_bytes=2 * 0.5 * big))
    d, dbg = device_fields(inps[2])
    with pytest.raises(p3x.P3Error) as e:
        srv.add_member(p3x.column_set(d, 1500, 128, 1.0))
    assert e.value.status == p3x.P3X_ERR_ARG
    srv.close()
