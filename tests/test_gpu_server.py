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
    assert e.value.status == p3x.P3X_ERR_BUDGET
    srv.close()


def test_hbm_budget_counts_resident_members(ctx):
    """Members are admitted while their resident arrays plus the slots' scratch fit the
    HBM budget; the next one is refused with P3X_ERR_BUDGET and nothing changes."""
    inp = make_columns(1000, 128, seed=77)
    d, _ = device_fields(inp)
    scratch = 64 << 20
    # a member spans 33 per-level fields, 2 interface fields and 2 surface fields
    one = (33 * 128 * 1000 + 2 * 129 * 1000 + 2 * 1000) * 8
    srv = p3x.Server(ctx, p3x.server_defaults(slots=2, scratch_budget_bytes=scratch,
                                              hbm_budget_bytes=scratch + 2.5 * one))
    srv.add_member(p3x.column_set(d, 1000, 128, 1.0, exchange_zeroed=True))
    srv.add_member(p3x.column_set(d, 1000, 128, 1.0, exchange_zeroed=True))
    with pytest.raises(p3x.P3Error) as e:
        srv.add_member(p3x.column_set(d, 1000, 128, 1.0, exchange_zeroed=True))
    assert e.value.status == p3x.P3X_ERR_BUDGET
    st = srv.stats()
    assert st.members == 2 and st.resident_bytes == 2 * one and st.scratch_bytes == scratch
    srv.close()


def test_poisson_arrivals_and_members_joining(ctx):
    """Requests arrive as a Poisson stream (exponential gaps) from members of uneven
    size and depth with random dt; two members register only after serving has started.
    Every member ends exactly where its own steps, run alone in submission order, leave it."""
    import time

    import torch
    rng = np.random.default_rng(123)
    spec = [(300, 128, 81), (64, 72, 82), (2048, 128, 83), (1, 72, 84), (777, 128, 85), (129, 72, 86),
            (4096, 72, 87), (513, 128, 88)]
    inps = [make_columns(n, k, seed=s, dry_every=7, evap_every=9) for n, k, s in spec]
    srv = p3x.Server(ctx, p3x.server_defaults(precision=p3x.P3X_MIXED, slots=3, max_batch=8))
    devs, ids, dts = [], [], {}

    def join(m):
        d, dbg = device_fields(inps[m])
        devs.append((m, d, dbg))
        ids.append(srv.add_member(p3x.column_set(d, spec[m][0], spec[m][1], 1.0, exchange_zeroed=True,
                                                 debug=dbg)))
        dts[ids[-1]] = []

    for m in range(6):
        join(m)
    torch.cuda.synchronize()
    nreq, tag = 90, 0
    gaps = rng.exponential(1.0 / 1500.0, nreq)  # about 1500 requests/s
    t = time.perf_counter()
    for r in range(nreq):
        if r == 30:
            join(6)  # while requests are in flight
        if r == 55:
            join(7)
        t += gaps[r]
        while time.perf_counter() < t:
            pass
        member = int(rng.choice(ids))
        dt = float(rng.choice([100.0, 300.0, 1800.0]))
        srv.submit(member, dt, tag)
        dts[member].append(dt)
        tag += 1
    srv.drain()
    done = srv.poll(max_n=1024)
    st = srv.stats()
    srv.close()
    assert len(done) == nreq and all(c.status == 0 for c in done)
    assert st.members == 8 and st.completed == nreq
    print(f"poisson: {nreq} requests, p50 {st.p50_us:.0f} us p99 {st.p99_us:.0f} us", flush=True)
    for (m, d, dbg), mid in zip(devs, ids):
        inp = inps[m]
        wd, wdbg = device_fields(inp)
        nlev, ncol = inp["qc"].shape
        for dt in dts[mid]:
            ctx.main(wd, ncol, nlev, dt, True, False, precision=p3x.P3X_MIXED, debug=wdbg,
                     exchange_zeroed=True, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        got, want = to_host(d, dbg), to_host(wd, wdbg)
        for k in FLOAT_FIELDS:
            np.testing.assert_array_equal(got[0][k], want[0][k], err_msg=f"member {m} {k}")


def test_failing_request_fails_alone(ctx):
    """One member carries a level with T <= 0 (the reference aborts there,
    share/physics_saturation_impl.hpp:30-35).  Its requests complete with
    P3X_ERR_TEMPERATURE; every other request of the same batches completes OK and
    leaves its member exactly as running alone does."""
    import torch
    inps = [make_columns(n, k, seed=s, dry_every=7, evap_every=9) for n, k, _, _, s in MEMBERS]
    bad = 2
    inps[bad]["th"] = inps[bad]["th"].copy()
    inps[bad]["th"][40, 17] = -5.0
    srv = p3x.Server(ctx, p3x.server_defaults(precision=p3x.P3X_FP64, slots=1, max_batch=32))
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
    srv.close()
    assert len(done) == tag
    for c in done:
        assert c.status == (p3x.P3X_ERR_TEMPERATURE if c.member == bad else 0), (c.member, c.status)
    for m, ((d, dbg), inp, (_, _, pnc, pccn, _)) in enumerate(zip(devs, inps, MEMBERS)):
        if m == bad:
            continue
        want = _alone(ctx, inp, pnc, pccn, p3x.P3X_FP64)
        g, gd = to_host(d, dbg)
        for k in FLOAT_FIELDS:
            np.testing.assert_array_equal(g[k], want[0][k], err_msg=f"member {m} {k}")
