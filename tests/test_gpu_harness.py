"""The regression harness (scream_amd.harness) with the HIP step as the
implementation under test:
* against a baseline the oracle writes through the same harness with the
  kernels' elementary functions: bit for bit (tol 0), FP64;
* against the committed baseline the reference's own Fortran P3 wrote (mixed
  initial condition, 4 flag cases x 3 steps): within the reference harness's
  rule at tol 1e-10 (the Fortran's libm and the kernels' portable functions
  round differently in the last bit)."""
import io
import os

import pytest

from scream_amd import harness

from harness_runners import OracleRunner

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def gpu_runner():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return harness.GpuRunner()


def test_gpu_vs_oracle_baseline_bit_exact(gpu_runner, oracle_tables, tmp_path):
    base = str(tmp_path / "oracle.bin")
    harness.generate_baseline(base, OracleRunner(oracle_tables, math="portable"), nsteps=6, ncol=5, nlev=72)
    out = io.StringIO()
    nerr = harness.run_and_cmp(base, gpu_runner, tol=0.0, nsteps=6, ncol=5, nlev=72, out=out)
    assert nerr == 0, out.getvalue()


def test_gpu_vs_fortran_baseline(gpu_runner):
    out = io.StringIO()
    nerr = harness.run_and_cmp(os.path.join(GOLDEN, "harness_mixed_i3_k72_s3.bin"), gpu_runner, tol=1e-10,
                               nsteps=3, out=out)
    assert nerr == 0, out.getvalue()
