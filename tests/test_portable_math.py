"""The portable FP64 elementary functions (scream_amd/csrc/p3_math.h) (CPU).

The HIP FP64 form uses them so that it reproduces the oracle's portable
flavour bit for bit.  Pinned here three ways:
* accuracy against mpmath at 120 bits on random arguments over the ranges P3
  uses (ulp bounds per function);
* exact special values;
* the portable-flavour oracle agrees with the glibc flavour (the one pinned bit
  for bit to the Fortran reference) within 1e-10 on the golden column sets,
  with identical integer outputs.
"""
import glob
import math
import os

import numpy as np
import pytest

import oracle

mp = pytest.importorskip("mpmath")
mp.mp.prec = 120

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# name, id, argument sampler, mpmath reference, max ulp
FUNCS = [
    ("exp", 0, lambda r, n: (r.uniform(-740, 709, n), None), lambda x, y: mp.exp(x), 1.0),
    ("log", 1, lambda r, n: (10 ** r.uniform(-300, 300, n), None), lambda x, y: mp.log(x), 1.0),
    ("log10", 2, lambda r, n: (10 ** r.uniform(-30, 30, n), None), lambda x, y: mp.log10(x), 1.5),
    ("pow", 3, lambda r, n: (10 ** r.uniform(-8, 8, n), r.uniform(-3, 3, n)), lambda x, y: mp.power(x, y), 2.0),
    ("cbrt", 4, lambda r, n: (10 ** r.uniform(-300, 300, n) * r.choice([-1, 1], n), None),
     lambda x, y: mp.sign(x) * mp.cbrt(abs(x)), 1.0),
    ("tgamma", 5, lambda r, n: (r.uniform(0.05, 30, n), None), lambda x, y: mp.gamma(x), 32.0),
    ("expm1", 6, lambda r, n: (np.concatenate([r.uniform(-0.5, 0.5, n // 2), r.uniform(-40, 40, n - n // 2)]), None),
     lambda x, y: mp.expm1(x), 1.5),
    ("tanh", 7, lambda r, n: (r.uniform(-25, 25, n), None), lambda x, y: mp.tanh(x), 3.0),
]


def _ulps(got, ref):
    ref = mp.mpf(ref)
    if ref == 0:
        return 0.0 if got == 0 else math.inf
    e = math.frexp(float(ref))[1]
    ulp = max(mp.mpf(2) ** (e - 53), mp.mpf(2) ** -1074)
    return float(abs(mp.mpf(got) - ref) / ulp)


@pytest.mark.parametrize("name,fn,sample,ref,bound", FUNCS, ids=[f[0] for f in FUNCS])
def test_accuracy_against_mpmath(name, fn, sample, ref, bound):
    lib = oracle.load_oracle("portable")
    rng = np.random.default_rng(100 + fn)
    xs, ys = sample(rng, 600)
    ys = np.zeros_like(xs) if ys is None else ys
    worst = 0.0
    for x, y in zip(xs, ys):
        got = lib.p3o_elementary(fn, float(x), float(y))
        worst = max(worst, _ulps(got, ref(mp.mpf(float(x)), mp.mpf(float(y)))))
    assert worst <= bound, f"{name}: {worst:.2f} ulp"


def test_exact_special_values():
    lib = oracle.load_oracle("portable")
    f = lambda fn, x, y=0.0: lib.p3o_elementary(fn, x, y)
    assert f(0, 0.0) == 1.0 and f(1, 1.0) == 0.0 and f(2, 1.0) == 0.0
    assert f(0, -800.0) == 0.0 and f(0, 800.0) == math.inf
    assert f(1, 0.0) == -math.inf and math.isnan(f(1, -1.0))
    for x in (1.7, 3.25e-5, 812.0):
        assert f(3, x, 2.0) == x * x and f(3, x, 1.0) == x and f(3, x, 0.0) == 1.0
    assert f(3, 0.0, 2.5) == 0.0 and f(3, 0.0, -1.5) == math.inf
    assert f(4, 27.0) == 3.0 and f(4, -8.0) == -2.0 and f(4, 0.0) == 0.0
    for n, g in ((1.0, 1.0), (2.0, 1.0), (3.0, 2.0), (5.0, 24.0), (7.0, 720.0)):
        assert abs(f(5, n) / g - 1) < 4e-15
    assert f(6, 0.0) == 0.0 and f(6, -60.0) == -1.0
    assert f(7, 0.0) == 0.0 and f(7, 30.0) == 1.0 and f(7, -30.0) == -1.0


CASES = sorted(glob.glob(os.path.join(GOLDEN, "p3_main_*.npz")))


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[8:-4] for p in CASES])
def test_portable_flavour_close_to_glibc(path, oracle_tables):
    z = np.load(path)
    inp = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    ncol, nlev, dt, pnc, pccn, seed = z["meta"]
    a, da, _ = oracle.run_oracle(inp, oracle_tables, float(dt), bool(pnc), bool(pccn), math="glibc")
    b, db, _ = oracle.run_oracle(inp, oracle_tables, float(dt), bool(pnc), bool(pccn), math="portable")
    for k in oracle.STATE + oracle.OUTS + oracle.HIST:
        s = np.maximum(np.abs(a[k]), np.abs(b[k]))
        r = np.where(s > 0, np.abs(a[k] - b[k]) / np.where(s > 0, s, 1), 0)
        assert r.max() <= 1e-10, k
    for k in da:
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)
