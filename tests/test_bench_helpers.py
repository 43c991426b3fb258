"""bench.py's host-side helpers: the HBM-traffic summary from rocprofv3 counter
files (calibrated by the copy probe) and the multi-GPU refusal; no GPU needed."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _write(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_summarize_pmc_scales_by_the_copy_probe(tmp_path):
    gib = bench.COPY_PROBE_BYTES
    fdir, wdir = tmp_path / "f", tmp_path / "w"
    # the copy probe reads 1 GiB and writes 1 GiB: 2048 B per FETCH_SIZE KiB, 1024 B per WRITE_SIZE KiB
    _write(str(fdir / "x" / "copy_counter_collection.csv"),
           [{"Kernel_Name": "copy", "Counter_Name": "FETCH_SIZE", "Counter_Value": gib / 2048}])
    _write(str(wdir / "x" / "copy_counter_collection.csv"),
           [{"Kernel_Name": "copy", "Counter_Name": "WRITE_SIZE", "Counter_Value": gib / 1024}])
    rows_f, rows_w = [], []
    for name, f, w in (("p3_part1_kernel", 100.0, 50.0), ("p3_part2_kernel", 200.0, 80.0),
                       ("p3_sed_kernel", 300.0, 90.0)):
        for rep in range(3):  # three dispatches each, means taken per kernel
            rows_f.append({"Kernel_Name": f"void p3d::{name}<double>(p3d::KParams)", "Counter_Name": "FETCH_SIZE",
                           "Counter_Value": f + rep - 1})
            rows_w.append({"Kernel_Name": f"void p3d::{name}<double>(p3d::KParams)", "Counter_Name": "WRITE_SIZE",
                           "Counter_Value": w})
    rows_f.append({"Kernel_Name": "void p3d::p3_sed_kernel<double>(p3d::KParams)", "Counter_Name": "SQ_WAVES",
                   "Counter_Value": 1e9})  # other counters are ignored
    _write(str(fdir / "x" / "bench_counter_collection.csv"), rows_f)
    _write(str(wdir / "x" / "bench_counter_collection.csv"), rows_w)
    out = tmp_path / "traffic.json"
    bench.summarize_pmc(str(fdir), str(wdir), out_path=str(out))
    k = bench.traffic_for("g64_fp64", "p3_sed_kernel", path=str(out))
    assert k == pytest.approx(300.0 * 2048 + 90.0 * 1024)
    step = bench.traffic_for("g64_fp64", path=str(out))
    assert step == pytest.approx((100 + 200 + 300) * 2048 + (50 + 80 + 90) * 1024)
    assert bench.traffic_for("no_such_workload", path=str(out)) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
