"""The portable elementary functions (ocotillo_amd/csrc/oco_math.h), host
build: accuracy against correctly rounded values (decimal, 60 digits),
special values, and the committed tables against a fresh generation."""
import ctypes
import math
import os
import subprocess
import sys
from decimal import Decimal, getcontext

import numpy as np
import pytest

from conftest import REPO

getcontext().prec = 60
EXP, LOG, LOG10, EXP10, POW, SQRT = range(6)


@pytest.fixture(scope="module")
def probe():
    from oracle.oracle import LIBDIR, build
    path = os.path.join(LIBDIR, "libmathprobe.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.mp_eval.argtypes = [ctypes.c_int, ctypes.c_long, DP, DP, DP]

    def f(fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(x)
        L.mp_eval(fn, x.size, x, y, out)
        return out
    return f


def ulp_err(got, exact):
    b = float(exact)
    if got == b:
        return 0.0
    return float(abs(Decimal(got) - exact) / Decimal(math.ulp(b)))


CASES = {
    EXP: (lambda r: r.uniform(-708, 709, 3000), lambda x, y: Decimal(x).exp()),
    LOG: (lambda r: np.exp(r.uniform(-700, 700, 3000)), lambda x, y: Decimal(x).ln()),
    LOG10: (lambda r: np.exp(r.uniform(-50, 50, 3000)), lambda x, y: Decimal(x).log10()),
    EXP10: (lambda r: r.uniform(-300, 300, 3000), lambda x, y: (Decimal(x) * Decimal(10).ln()).exp()),
}


@pytest.mark.parametrize("fn", sorted(CASES))
def test_accuracy_against_correct_rounding(probe, fn):
    rng = np.random.default_rng(11 + fn)
    gen, exact = CASES[fn]
    x = gen(rng)
    got = probe(fn, x)
    errs = [ulp_err(g, exact(v, None)) for g, v in zip(got, x)]
    assert max(errs) < 0.51
    assert np.mean(np.array(errs) > 0) < 0.01


def test_log_near_one_is_relative_accurate(probe):
    x = 1.0 + np.concatenate([np.linspace(-2 ** -7, 2 ** -7, 801), [1e-12, -1e-12, 2 ** -52, -2 ** -53]])
    x = x[x != 1.0]
    got = probe(LOG, x)
    assert max(ulp_err(g, Decimal(v).ln()) for g, v in zip(got, x)) < 0.51


def test_pow_accuracy(probe):
    rng = np.random.default_rng(3)
    x = rng.uniform(1e-4, 1e6, 2000)
    y = np.where(rng.uniform(size=2000) < 0.5, 1.5, rng.uniform(-3.5, 3.5, 2000))
    got = probe(POW, x, y)
    assert max(ulp_err(g, (Decimal(b) * Decimal(a).ln()).exp()) for g, a, b in zip(got, x, y)) < 0.51


def test_subnormal_and_range_edges(probe):
    x = np.array([-745.0, -744.44, -730.0, -708.5, 709.78, 709.7827128933840])
    got = probe(EXP, x)
    for g, v in zip(got, x):
        assert ulp_err(g, Decimal(v).exp()) <= 1.0 or Decimal(v).exp() < Decimal("4.9e-324")
    assert probe(LOG, np.array([5e-324]))[0] == pytest.approx(math.log(5e-324), rel=1e-15)


def test_special_values(probe):
    inf, nan = float("inf"), float("nan")
    e = probe(EXP, np.array([0.0, -0.0, inf, -inf, nan, 710.0, -746.0]))
    assert e[0] == 1.0 and e[1] == 1.0 and e[2] == inf and e[3] == 0.0 and math.isnan(e[4])
    assert e[5] == inf and e[6] == 0.0
    l = probe(LOG, np.array([1.0, 0.0, -1.0, inf, nan]))
    assert l[0] == 0.0 and l[1] == -inf and math.isnan(l[2]) and l[3] == inf and math.isnan(l[4])
    assert probe(LOG10, np.array([1.0, 10.0, 1000.0])).tolist() == [0.0, 1.0, 3.0]
    p = probe(POW, np.array([0.0, 0.0, 1.0, 2.0, inf, 4.0]), np.array([2.0, -1.0, nan, 0.0, 2.0, 0.5]))
    assert p.tolist()[:5] == [0.0, inf, 1.0, 1.0, inf] and p[5] == 2.0
    assert probe(EXP10, np.array([0.0, 1.0, 2.0]))[0] == 1.0
    assert probe(SQRT, np.array([2.0]))[0] == math.sqrt(2.0)


def test_tables_regenerate_identically():
    r = subprocess.run([sys.executable, os.path.join(REPO, "ocotillo_amd", "csrc", "gen_math_tables.py"), "--check"])
    assert r.returncode == 0
