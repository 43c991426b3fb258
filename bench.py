#!/usr/bin/env python3
"""Benchmark of the MI355X P3 microphysics step (p3_main drop-in).

A "step" is one p3_main call over one batch of synthetic columns, with the
inputs already resident in HBM.  Headline workload: the FP64 drop-in on the
ne30pg2-sized physics grid, 21600 columns x 72 levels, dt = 1800 s.  Before
every step the prognostic state is restored from a pristine device copy (so
every step sees the same work); that copy is inside the timed region.

Also reported (keys "forms"): the FP64 and mixed-precision forms on one GPU's
share of a high-resolution grid (16384 x 128, dt = 100 s) with the mixed
form's per-field error against the FP64 result, the column-level roofline of
the dominant kernel, and the oracle (CPU restatement) timed on this host.

Prints ONE JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "P3 column-levels/s, FP64 vs mixed at stated error; served p50/p99 step latency"
BYTES_PER_COLUMN_LEVEL = 360.0   # 23 FP64 fields read + 22 written per column-level
# The dominant kernel, sedimentation + part 3: of the step's interface fields it must read
# the 10 state fields, the 3 cloud fractions, dz, dpres, inv_exner and vap_liq_exchange, and
# write the 10 state fields, 3 effective radii, rho_qi, the 2 precipitation fluxes and
# vap_liq_exchange: 34 FP64 fields per column-level.
SED_BYTES_PER_COLUMN_LEVEL = 272.0
HBM_PEAK_GBS = 8000.0            # MI355X published HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-forms", action="store_true", help="skip the 16384x128 forms")
    ap.add_argument("--no-forms-batch", action="store_true",
                    help="skip the mixed batched leg of the 16384x128 forms")
    ap.add_argument("--no-column-major", action="store_true",
                    help="skip the headline step on column-major (column, level) arrays")
    ap.add_argument("--no-process", action="store_true",
                    help="skip the P3Microphysics process (run_impl) leg")
    ap.add_argument("--batch-sets", type=int, default=64,
                    help="column sets of the batched-form leg; 0 skips it")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="length of the served-form run (Poisson arrivals); 0 skips it")
    ap.add_argument("--no-table-gen", action="store_true",
                    help="skip timing the ice lookup-table generator")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="turn two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of "
                         "'bench.py --no-cpu --no-forms' and of 'bench.py --copy-probe' "
                         "into profiles/traffic.json")
    return ap.parse_args()


class Workload:
    """Device-resident fields for one column set, plus a pristine copy of the state.
    layout: level-major (nlev, ncol) arrays, or column-major (ncol, nlev), the order of
    the reference's field-manager views."""

    def __init__(self, ncol, nlev, seed, layout=0):
        import numpy as np
        import torch
        from scream_amd import p3x
        from scream_amd.inputs import make_columns
        self.ncol, self.nlev, self.layout = ncol, nlev, layout
        dev = torch.device("cuda:0")
        inp = make_columns(ncol, nlev, seed=seed)
        if layout == p3x.P3X_COLUMN_MAJOR:
            inp = {k: np.ascontiguousarray(v.T) for k, v in inp.items()}
        # the 10 state fields are views of one allocation, so that restoring the
        # pristine state before a step is a single device copy
        ns = len(p3x.STATE)
        self.state = torch.from_numpy(np.stack([inp[k] for k in p3x.STATE])).to(dev)
        self.f = {k: self.state[j] for j, k in enumerate(p3x.STATE)}
        self.f.update({k: torch.from_numpy(inp[k]).to(dev) for k in p3x.INPUTS})
        for k in p3x.OUTS + p3x.HIST:
            rows = nlev + 1 if "flux" in k else nlev
            shape = (ncol,) if "surf" in k else ((ncol, rows) if layout else (rows, ncol))
            self.f[k] = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.pristine = self.state.clone()
        assert self.state.shape[0] == ns

    def restore(self):
        self.state.copy_(self.pristine)


def time_steps(ctx, w, dt, precision, steps, warmup):
    """Wall ms/step (sync-bracketed) and mean kernel us/step (HIP events on the kernel's stream)."""
    import torch
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    for _ in range(warmup):
        w.restore()
        ctx.main(w.f, w.ncol, w.nlev, dt, True, False, precision=precision, stream=sp,
                 async_=True, exchange_zeroed=True, layout=w.layout)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        w.restore()
        e0.record(stream)
        ctx.main(w.f, w.ncol, w.nlev, dt, True, False, precision=precision, stream=sp,
                 async_=True, exchange_zeroed=True, layout=w.layout)
        e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = ctx.sync()
    if st != 0:
        raise RuntimeError(f"p3x status {st}")
    kern_us = [1000.0 * e0.elapsed_time(e1) for e0, e1 in evs]
    return 1000.0 * (t1 - t0) / steps, statistics.mean(kern_us)


def kernel_split(ctx, w, dt, precision, steps):
    """Mean device microseconds of a step in two intervals, (the part 1 and part 2 kernels, the
    sedimentation and part 3 kernels), from HIP events the library records on the kernels'
    stream around and between them."""
    import torch
    sp = torch.cuda.current_stream().cuda_stream
    a, b = [], []
    for _ in range(steps):
        w.restore()
        ctx.main(w.f, w.ncol, w.nlev, dt, True, False, precision=precision, stream=sp,
                 async_=True, exchange_zeroed=True, layout=w.layout)
        x, y = ctx.last_kernel_times()
        a.append(x)
        b.append(y)
    return statistics.mean(a), statistics.mean(b)


def process_step(ctx, ncol, nlev, dt, steps, warmup, seed=1, layout=0):
    """The P3Microphysics process step (run_impl: pre-amble, p3_main, post-amble, the
    pre/post-amble fused into the p3_main kernels) on wet field-manager fields of the
    same generator: wall ms/step and HIP-event kernel us/step, the updated fields
    restored from one pristine copy before every step (outside the events).  layout:
    level-major, or column-major (the field manager's (column, level) views, dense)."""
    import numpy as np
    import torch
    from scream_amd import p3x
    from scream_amd.inputs import (ATM_COMPUTED_FIELDS, ATM_FLUX_FIELDS, ATM_UPDATED_FIELDS,
                                   make_atm_fields)
    dev = torch.device("cuda:0")
    f = make_atm_fields(ncol, nlev, seed=seed)
    if layout == p3x.P3X_COLUMN_MAJOR:
        f = {k: np.ascontiguousarray(v.T) for k, v in f.items()}
    flat = torch.from_numpy(np.concatenate([f[k].ravel() for k in ATM_UPDATED_FIELDS])).to(dev)
    d, off = {}, 0
    for k in ATM_UPDATED_FIELDS:
        n = f[k].size
        d[k] = flat[off:off + n].view(f[k].shape)
        off += n
    d.update({k: torch.from_numpy(v).to(dev) for k, v in f.items() if k not in d})
    for k in ATM_COMPUTED_FIELDS:
        d[k] = torch.zeros((ncol, nlev) if layout else (nlev, ncol), dtype=torch.float64, device=dev)
    for k in ATM_FLUX_FIELDS:
        d[k] = torch.zeros((ncol,), dtype=torch.float64, device=dev)
    pristine = flat.clone()
    proc = p3x.P3Microphysics(ctx, ncol, nlev, predict_nc=True, prescribed_ccn=False, layout=layout)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    for _ in range(warmup):
        flat.copy_(pristine)
        proc.run(d, dt, stream=sp, async_=True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        flat.copy_(pristine)
        e0.record(stream)
        proc.run(d, dt, stream=sp, async_=True)
        e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = ctx.sync()
    proc.close()
    if st != 0:
        raise RuntimeError(f"p3x status {st}")
    ms = 1000.0 * (t1 - t0) / steps
    kern_us = statistics.mean(1000.0 * e0.elapsed_time(e1) for e0, e1 in evs)
    lay = "column-major" if layout == p3x.P3X_COLUMN_MAJOR else "level-major"
    return {"workload": f"P3Microphysics run_impl FP64, {ncol}x{nlev}, dt={dt:g} s, boundary fluxes on, "
                        f"{lay} fields",
            "column_levels_per_s": ncol * nlev / (ms / 1000.0), "ms_per_step": ms,
            "kernel_us_per_step": kern_us}


def mixed_error(ctx, w, dt):
    """Per-field max |mixed - fp64| / max|fp64| and RMS ratio on one step (outputs of the FP64 form
    match the FP64 oracle to 1e-10, so this is the mixed form's error against the oracle)."""
    import torch
    from scream_amd import p3x
    sp = torch.cuda.current_stream().cuda_stream
    fields = p3x.STATE + p3x.OUTS + p3x.HIST
    w.restore()
    ctx.main(w.f, w.ncol, w.nlev, dt, True, False, precision=p3x.P3X_FP64, stream=sp, exchange_zeroed=True)
    ref = {k: w.f[k].clone() for k in fields}
    w.restore()
    ctx.main(w.f, w.ncol, w.nlev, dt, True, False, precision=p3x.P3X_MIXED, stream=sp, exchange_zeroed=True)
    out = {}
    for k in fields:
        d = (w.f[k] - ref[k]).abs()
        scale = ref[k].abs().max().item()
        rms_ref = ref[k].pow(2).mean().sqrt().item()
        out[k] = {"max_rel_to_field_max": (d.max().item() / scale) if scale > 0 else 0.0,
                  "rms_rel": ((d.pow(2).mean().sqrt().item() / rms_ref) if rms_ref > 0 else 0.0)}
    return out


class SetBatch:
    """Device-resident column sets for the batched form: every set's 10 state fields are
    views of one allocation, so that restoring all of them is one device copy."""

    def __init__(self, specs, seed0, inputs=None):
        import numpy as np
        import torch
        from scream_amd import p3x
        from scream_amd.inputs import make_columns
        dev = torch.device("cuda:0")
        self.specs = specs  # (ncol, nlev, dt)
        inps = inputs or [make_columns(n, k, seed=seed0 + s) for s, (n, k, _) in enumerate(specs)]
        total = sum(10 * n * k for n, k, _ in specs)
        self.state = torch.empty(total, dtype=torch.float64, device=dev)
        self.fields, off = [], 0
        for inp, (n, k, _) in zip(inps, specs):
            f = {}
            for key in p3x.STATE:
                f[key] = self.state[off:off + n * k].view(k, n)
                f[key].copy_(torch.from_numpy(inp[key]))
                off += n * k
            f.update({key: torch.from_numpy(inp[key]).to(dev) for key in p3x.INPUTS})
            for key in p3x.OUTS + p3x.HIST:
                shape = (n,) if "surf" in key else ((k + 1, n) if "flux" in key else (k, n))
                f[key] = torch.zeros(shape, dtype=torch.float64, device=dev)
            self.fields.append(f)
        self.pristine = self.state.clone()
        self.column_levels = float(sum(n * k for n, k, _ in specs))

    def restore(self):
        self.state.copy_(self.pristine)

    def column_sets(self):
        from scream_amd import p3x
        return [p3x.column_set(f, n, k, dt, exchange_zeroed=True) for f, (n, k, dt) in zip(self.fields, self.specs)]


def run_batched(ctx, sb, precision, steps, tune_reps=2):
    """One graph-captured batch of the sets of ``sb``, launch shape from the autotuner; the
    state of every set is restored before each step (inside the wall clock, outside the
    events around the batch).  Beside it, the same sets stepped one p3x_main call after
    another (``singles``, same restore)."""
    import torch
    from scream_amd import p3x
    sp = torch.cuda.current_stream().cuda_stream
    b = p3x.Batch(ctx, sb.column_sets(), p3x.batch_defaults(precision=precision, use_graph=1))
    chosen, tuned_us = b.tune(sp, reps=tune_reps)
    for _ in range(2):
        sb.restore()
        b.run(sp)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        sb.restore()
        e0.record()
        b.run(sp)
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = ctx.sync()
    b.close()
    if st != 0:
        raise RuntimeError(f"batched form: p3x status {st}")
    kern_ms = statistics.mean(e0.elapsed_time(e1) for e0, e1 in evs)
    # the same sets through p3x_main, one call per set
    def singles():
        for f, (n, k, dt) in zip(sb.fields, sb.specs):
            ctx.main(f, n, k, dt, True, False, precision=precision, stream=sp, async_=True,
                     exchange_zeroed=True)
    sb.restore()
    singles()
    torch.cuda.synchronize()
    sev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for e0, e1 in sev:
        sb.restore()
        e0.record()
        singles()
        e1.record()
    torch.cuda.synchronize()
    if ctx.sync() != 0:
        raise RuntimeError("batched form: singles failed")
    singles_ms = statistics.mean(e0.elapsed_time(e1) for e0, e1 in sev)
    return {"sets": len(sb.specs), "column_levels_per_step": sb.column_levels, "steps": steps,
            "ms_per_step": 1000.0 * wall / steps, "kernel_ms_per_step": kern_ms,
            "column_levels_per_s": sb.column_levels / (kern_ms / 1000.0),
            "wall_column_levels_per_s": sb.column_levels * steps / wall,
            "singles_ms_per_step": singles_ms, "batched_over_singles": singles_ms / kern_ms,
            "precision": "mixed" if precision == p3x.P3X_MIXED else "fp64", "graph": True,
            "tuned": chosen.as_dict(), "tuned_us_per_step": tuned_us,
            "state": "restored before every step (one device copy, outside the events)"}


def forms_batched(ctx, ncol, nlev, dt, forms, steps):
    """The mixed form batched on the 16384 x 128 workload: the same columns as 8 sets of
    2048 (one graph replay per step, autotuned), against the plain FP64 drop-in on the
    whole set (forms["fp64"])."""
    import numpy as np
    from scream_amd import p3x
    from scream_amd.inputs import make_columns
    whole = make_columns(ncol, nlev, seed=2)
    parts = [{k: np.ascontiguousarray(v[..., j * (ncol // 8):(j + 1) * (ncol // 8)])
              for k, v in whole.items()} for j in range(8)]
    sb = SetBatch([(ncol // 8, nlev, dt)] * 8, 0, inputs=parts)
    del whole, parts
    mb = run_batched(ctx, sb, p3x.P3X_MIXED, steps)
    del sb
    mb["workload"] = f"mixed batched form, 8 column sets of {ncol // 8} x {nlev}, dt={dt:g} s"
    mb["over_fp64_drop_in"] = forms["fp64"]["kernel_us_per_step"] / (1000.0 * mb["kernel_ms_per_step"])
    return mb


def batched(ctx, nsets, steps, seed=4):
    """Batched form: nsets independent column sets, ncol log-uniform in [256, 8192],
    nlev from {72, 128}, dt from {100, 300, 1800} s, mixed precision, one graph replay
    per step; launch shape from the autotuner."""
    import numpy as np
    from scream_amd import p3x
    rng = np.random.default_rng(seed)
    ncols = np.round(np.exp(rng.uniform(np.log(256), np.log(8192), nsets))).astype(int)
    nlevs = rng.choice([72, 128], nsets)
    dts = rng.choice([100.0, 300.0, 1800.0], nsets)
    specs = [(int(n), int(k), float(dt)) for n, k, dt in zip(ncols, nlevs, dts)]
    sb = SetBatch(specs, 2000)
    out = run_batched(ctx, sb, p3x.P3X_MIXED, steps)
    out["workload"] = (f"{nsets} column sets, ncol log-uniform 256-8192, nlev 72/128, dt 100/300/1800 s "
                       f"(seed {seed})")
    del sb
    return out


def served(ctx, seconds, rate=200.0, members=32, nlev=128, dt=300.0, seed=5):
    """Served form: Poisson arrivals at ``rate`` requests/s over ``seconds``, each a step of
    one of ``members`` ensemble members drawn uniformly; member ncol ~ LogNormal(median 2048,
    sigma 1 in ln-space) clipped to [128, 16384], nlev 128, mixed precision, state resident."""
    import numpy as np
    import torch
    from scream_amd import p3x
    from scream_amd.inputs import make_columns
    rng = np.random.default_rng(seed)
    ncols = np.clip(np.round(2048 * np.exp(rng.normal(0.0, 1.0, members))), 128, 16384).astype(int)
    srv = p3x.Server(ctx, p3x.server_defaults(precision=p3x.P3X_MIXED))
    keep = []
    dev = torch.device("cuda:0")
    for m, n in enumerate(ncols):
        inp = make_columns(int(n), nlev, seed=1000 + m)
        f = {k: torch.from_numpy(inp[k]).to(dev) for k in p3x.STATE + p3x.INPUTS}
        for k in p3x.OUTS + p3x.HIST:
            shape = (n,) if "surf" in k else ((nlev + 1, n) if "flux" in k else (nlev, n))
            f[k] = torch.zeros(shape, dtype=torch.float64, device=dev)
        keep.append(f)
        srv.add_member(p3x.column_set(f, int(n), nlev, dt, exchange_zeroed=True))
    torch.cuda.synchronize()
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5)))
    arrivals = arrivals[arrivals < seconds]
    who = rng.integers(0, members, len(arrivals))
    t0 = time.perf_counter()
    for i, (t, m) in enumerate(zip(arrivals, who)):
        wait = t0 + t - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        srv.submit(int(m), dt, i)
    srv.drain()
    wall = time.perf_counter() - t0
    bad = [c for c in srv.poll(max_n=len(arrivals) + 1) if c.status != 0]
    st = srv.stats()
    srv.close()
    if bad:
        raise RuntimeError(f"served form: {len(bad)} requests failed (status {bad[0].status})")
    return {"requests": int(st.completed), "offered_rate_per_s": rate, "seconds": seconds,
            "achieved_rate_per_s": st.completed / wall, "p50_ms": st.p50_us / 1000.0,
            "p99_ms": st.p99_us / 1000.0, "max_ms": st.max_us / 1000.0,
            "column_levels_per_s": st.column_levels / wall,
            "members": members, "nlev": nlev, "dt": dt,
            "member_ncol": {"min": int(ncols.min()), "median": float(np.median(ncols)),
                            "max": int(ncols.max())},
            "precision": "mixed", "slots": 4, "max_batch": 32}


# One GPU's share of the GPU box's host CPUs: worker pools on the box are sized to it,
# whatever the host's own CPU count is.
CPU_SHARE = 16


def host_facts():
    """The host's CPU count, the CPUs this process may run on, and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"host_cpus": os.cpu_count(), "affinity_cpus": aff, "cpu_model": model, "cpu_share": CPU_SHARE}


def cpu_baseline():
    """The CPU path timed on this host, in the same run, on (1) the headline workload itself
    (21600 columns x 72 levels, dt 1800 s, same generator) and (2) the reference's CPU
    configuration (256 x 72, dt 300 s): the wall time of the step alone (no init, no table
    load), median of repeats after a discarded cold run.  With oracle/_ref built (the
    reference's own Fortran P3, micro_p3.F90, compiled from its sources) that is the
    baseline ("reference"), called on column chunks from a pool of threads, one per CPU of
    this GPU's share of the host; the C restatement of the oracle (OpenMP over columns, same
    threads) is timed beside it ("port_value")."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import oracle
    from scream_amd import p3x
    from scream_amd.inputs import make_columns
    t = p3x.tables_load()
    a = lambda x, sh: np.ctypeslib.as_array(x).reshape(sh).copy()
    tabs = dict(ice=a(t.ice_table, (5, 4, 50, 12)), coll=a(t.collect_table, (5, 4, 50, 30, 2)),
                vn=a(t.vn_table, (300, 10)), vm=a(t.vm_table, (300, 10)),
                revap=a(t.revap_table, (300, 10)), mu_r=a(t.mu_r_table, (150,)), dnu=a(t.dnu, (16,)))
    facts = host_facts()
    nthreads = max(1, min(CPU_SHARE, facts["affinity_cpus"]))

    def median_time(fn, reps):
        fn()
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        return statistics.median(times)

    have_ref = oracle.ref_available()
    table_dir = os.path.join(ROOT, "scream_amd", "data")
    if have_ref:
        oracle.load_ref(table_dir)  # p3_init (table file I/O) once, before any thread calls p3_main
        threading.stack_size(256 << 20)  # the Fortran keeps its column arrays on the stack

    def leg(ncol, nlev, dt, seed, chunk, reps):
        inp = make_columns(ncol, nlev, seed=seed)
        port = median_time(lambda: oracle.run_oracle(inp, tabs, dt, True, False, nthreads=nthreads), reps)
        r = {"workload": f"{ncol}x{nlev}, dt {dt:g} s", "port_value": ncol * nlev / port, "port_s": port}
        if have_ref:
            chunks = [{k: np.ascontiguousarray(v[:, c:c + chunk]) for k, v in inp.items()}
                      for c in range(0, ncol, chunk)]
            with ThreadPoolExecutor(nthreads) as ex:
                run = lambda: list(ex.map(lambda ch: oracle.run_ref(ch, table_dir, dt, True, False), chunks))
                ref = median_time(run, reps)
            r.update({"value": ncol * nlev / ref, "ref_s": ref, "chunk_columns": chunk})
        return r

    g64 = leg(21600, 72, 1800.0, 1, 64, 5)
    cpu_cfg = leg(256, 72, 300.0, 3, 16, 51)
    kind = "reference" if have_ref else "port"
    what = ("reference Fortran p3_main (micro_p3.F90 via oracle/_ref)" if have_ref
            else "oracle/p3_oracle.c FP64 (OpenMP)")
    out = {"value": g64["value"] if have_ref else g64["port_value"], "unit": "column-levels/s",
           "cores": nthreads, "kind": kind,
           "sample": f"{what}, the headline workload 21600x72 dt 1800 s (same generator), whole, on "
                     f"{nthreads} threads (one GPU's share of the host), median of 5 runs after a cold run",
           "port_value": g64["port_value"], "g64": g64, "cpu_config": cpu_cfg, **facts}
    return out


COPY_PROBE_BYTES = 1 << 30


def table_generator(reps=3):
    """Wall milliseconds of the device ice lookup-table generator (the whole
    5 x 4 x 50-cell table of create_p3_lookupTable_1, about 3e10 inner iterations
    when run in sequence), median of reps after one untimed run."""
    from scream_amd import p3x
    p3x.ice_table_generate()
    ms = sorted(p3x.ice_table_generate()[2] for _ in range(reps))
    return {"ms": ms[len(ms) // 2], "cells": 1000, "reps": reps}


def copy_probe():
    """A plain streaming copy of 1 GiB as one elementwise kernel (dst = src * 1), three
    times: under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per dispatch calibrate the
    counters (1 GiB read, 1 GiB written)."""
    import torch
    n = COPY_PROBE_BYTES // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES}))


def copy_bandwidth(reps=10):
    """Measured streaming bandwidth of this GPU: a 1 GiB device-to-device copy (read
    + write = 2 GiB moved per copy), timed with HIP events, best of reps."""
    import torch
    n = COPY_PROBE_BYTES // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) / 1000.0)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * COPY_PROBE_BYTES / best / 1e9


def _pmc_means(directory, counter, prefix="*"):
    """Mean counter value per dispatch, by kernel name, from the rocprofv3 csv output
    files named <prefix>_counter_collection.csv under directory."""
    import csv
    import glob
    sums = {}
    for path in glob.glob(os.path.join(directory, "**", f"{prefix}_counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                k = row["Kernel_Name"]
                s = sums.setdefault(k, [0.0, 0])
                s[0] += float(row["Counter_Value"])
                s[1] += 1
    return {k: v[0] / v[1] for k, v in sums.items()}


def summarize_pmc(fetch_dir, write_dir, out_path=None):
    """HBM bytes per step of the headline workload from FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled by the bytes-per-KiB the same counter reports for the copy probe;
    written to out_path (default profiles/traffic.json)."""
    fetch, write = _pmc_means(fetch_dir, "FETCH_SIZE", "bench"), _pmc_means(write_dir, "WRITE_SIZE", "bench")
    pick = lambda d, key: {k: v for k, v in d.items() if key in k}
    copy_f = max(_pmc_means(fetch_dir, "FETCH_SIZE", "copy").values())
    copy_w = max(_pmc_means(write_dir, "WRITE_SIZE", "copy").values())
    scale_f = COPY_PROBE_BYTES / copy_f  # bytes per counter KiB, reads
    scale_w = COPY_PROBE_BYTES / copy_w  # bytes per counter KiB, writes
    kern = {}
    # part 1 is p3_part1_lvl_kernel (one thread per level and column) or p3_part1_kernel;
    # sedimentation is one launch of both species groups (p3_sed_fused_kernel) or two
    # kernels (the ice group and the liquid group), summed
    for key, name in (("p3_part1", "p3_part1_kernel"), ("p3_part2", "p3_part2_kernel"),
                      ("p3_sed", "p3_sed_kernel"), ("p3_part3", "p3_part3_kernel")):
        f = pick(fetch, key)
        w = pick(write, key)
        if f and w:
            kern[name] = {"FETCH_SIZE_KiB": sum(f.values()), "WRITE_SIZE_KiB": sum(w.values()),
                          "bytes": sum(f.values()) * scale_f + sum(w.values()) * scale_w,
                          "kernels": sorted(f)}
    out = {"g64_fp64": {
        "workload": "p3_main FP64 drop-in, 21600x72, dt=1800 s (bench.py --steps 5 --warmup 2 --no-cpu "
                    "--no-forms --no-process --batch-sets 0 --serve-seconds 0 --no-table-gen "
                    "--no-column-major)",
        "copy_probe": {"bytes_each_way": COPY_PROBE_BYTES, "FETCH_SIZE_KiB": copy_f,
                       "WRITE_SIZE_KiB": copy_w, "read_bytes_per_KiB": scale_f,
                       "write_bytes_per_KiB": scale_w},
        "kernels": kern,
        "bytes_per_step": sum(v["bytes"] for v in kern.values())}}
    path = out_path or os.path.join(ROOT, "profiles", "traffic.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def traffic_for(workload, kernel=None, path=None):
    """HBM bytes per step (or per launch of `kernel`) measured by rocprofv3 counters
    (profiles/traffic.json unless path is given), or None."""
    path = path or os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh).get(workload)
    if rec is None:
        return None
    if kernel is None:
        return rec.get("bytes_per_step")
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes")


def main():
    args = parse()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "P3 columns are independent, so N GPUs would each run this workload on their "
              "own column range", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from scream_amd import p3x
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.copy_probe:
        copy_probe()
        return
    ctx = p3x.Context()
    copy_gbs = copy_bandwidth()

    ncol, nlev, dt = 21600, 72, 1800.0
    w = Workload(ncol, nlev, seed=1)
    ms, kern_us = time_steps(ctx, w, dt, p3x.P3X_FP64, args.steps, args.warmup)
    k12_us, sed_us = kernel_split(ctx, w, dt, p3x.P3X_FP64, args.steps)
    colev = ncol * nlev
    value = colev / (ms / 1000.0)
    achieved = SED_BYTES_PER_COLUMN_LEVEL * colev / (sed_us * 1e-6) / 1e9
    # the dominant stage: sedimentation, then part 3 in its own kernel
    t_sed, t_p3 = traffic_for("g64_fp64", "p3_sed_kernel"), traffic_for("g64_fp64", "p3_part3_kernel")
    traffic = t_sed + t_p3 if (t_sed is not None and t_p3 is not None) else None
    workload = f"p3_main FP64 drop-in, {ncol}x{nlev}, dt={dt:g} s"
    line = {
        "metric": METRIC, "value": value, "unit": "column-levels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "ncol": ncol, "nlev": nlev, "dt": dt,
                   "predictNc": True, "prescribedCCN": False, "kernel_us_per_step": kern_us,
                   "part12_kernel_us": k12_us, "sed_kernel_us": sed_us,
                   "step_hbm_frac": BYTES_PER_COLUMN_LEVEL * colev / (kern_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                   "copy_GBps": copy_gbs,
                   "step_frac_of_copy": BYTES_PER_COLUMN_LEVEL * colev / (kern_us * 1e-6) / 1e9 / copy_gbs,
                   "step_traffic_bytes": traffic_for("g64_fp64")},
        # the dominant stage (the sedimentation kernel and the part 3 kernel after it):
        # algorithmic bytes per step over its event-timed duration, and its measured HBM bytes
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "p3_sed_fused_kernel (ice and liquid groups in one launch) + p3_part3_lvl_kernel"},
    }
    del w
    if not args.no_column_major:
        # the same step on the reference's own (column, level) layout of the fields
        wc = Workload(ncol, nlev, seed=1, layout=p3x.P3X_COLUMN_MAJOR)
        ms_c, kern_c = time_steps(ctx, wc, dt, p3x.P3X_FP64, args.steps, args.warmup)
        line["column_major"] = {"workload": workload + ", column-major (column, level) arrays",
                                "column_levels_per_s": colev / (ms_c / 1000.0), "ms_per_step": ms_c,
                                "kernel_us_per_step": kern_c, "kernel_time_over_level_major": kern_c / kern_us}
        del wc
    if not args.no_process:
        line["process"] = process_step(ctx, ncol, nlev, dt, args.steps, args.warmup)
        pc = process_step(ctx, ncol, nlev, dt, args.steps, args.warmup, layout=p3x.P3X_COLUMN_MAJOR)
        pc["kernel_time_over_level_major"] = pc["kernel_us_per_step"] / line["process"]["kernel_us_per_step"]
        line["process_column_major"] = pc
    if not args.no_forms:
        ncol2, nlev2, dt2 = 16384, 128, 100.0
        w2 = Workload(ncol2, nlev2, seed=2)
        forms = {}
        for name, prec in (("fp64", p3x.P3X_FP64), ("mixed", p3x.P3X_MIXED)):
            ms2, k2 = time_steps(ctx, w2, dt2, prec, max(5, args.steps // 2), args.warmup)
            forms[name] = {"column_levels_per_s": ncol2 * nlev2 / (ms2 / 1000.0), "ms_per_step": ms2,
                           "kernel_us_per_step": k2,
                           "hbm_frac": BYTES_PER_COLUMN_LEVEL * ncol2 * nlev2 / (k2 * 1e-6) / 8e12}
        forms["mixed_over_fp64"] = forms["mixed"]["column_levels_per_s"] / forms["fp64"]["column_levels_per_s"]
        forms["mixed_over_fp64_kernel"] = forms["fp64"]["kernel_us_per_step"] / forms["mixed"]["kernel_us_per_step"]
        # the mixed form batched: the same 16384 x 128 columns as 8 sets of 2048 (one graph
        # replay per step, autotuned), against the plain FP64 drop-in on the whole set
        if not args.no_forms_batch:
            forms["mixed_batched"] = forms_batched(ctx, ncol2, nlev2, dt2, forms, max(5, args.steps // 2))
        errs = mixed_error(ctx, w2, dt2)
        forms["mixed_error_vs_fp64"] = {
            "worst_field": max(errs, key=lambda k: errs[k]["max_rel_to_field_max"]),
            "max_rel_to_field_max": max(e["max_rel_to_field_max"] for e in errs.values()),
            "per_field": errs}
        line["forms"] = {f"{ncol2}x{nlev2}_dt{dt2:g}": forms}
        del w2
    if args.batch_sets > 0:
        line["batched"] = batched(ctx, args.batch_sets, max(5, args.steps // 2))
    if args.serve_seconds > 0:
        line["served"] = served(ctx, args.serve_seconds)
    if not args.no_table_gen:
        line["table_generator"] = table_generator()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
