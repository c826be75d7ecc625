"""The C-ABI library loads and exports every entry point include/mallard_amd.h
declares; host-side helpers behave (no GPU calls here)."""
import math
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mallard_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mlr_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported(mlr):
    names = _declared()
    assert len(names) >= 30
    L = mlr.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert set(names) <= set(mlr._SIGS)


def test_null_arguments_return_errors(mlr):
    L = mlr.lib()
    assert L.mlr_ctx_create_generated(None, None, None) != 0
    assert L.mlr_get_sizes(None, None) != 0
    assert L.mlr_calc_rhs(None, None) != 0


def test_host_pow6_is_correctly_rounded(mlr):
    from fractions import Fraction
    rng = np.random.default_rng(4)
    x = np.exp(rng.uniform(np.log(1e-12), np.log(1e3), 3000))
    h = mlr.host_pow6(x)
    exact = np.array([float(Fraction(v) ** 6) for v in x])
    assert np.array_equal(h, exact)


def test_host_pow_tracks_glibc(mlr):
    rng = np.random.default_rng(5)
    b = rng.uniform(0.01, 10.0, 50000)
    e = rng.uniform(0.1, 8.0, 50000)
    h = mlr.host_pow(b, e)
    g = np.array([math.pow(u, v) for u, v in zip(b, e)])
    ulps = np.abs(h.view(np.int64) - g.view(np.int64))
    assert ulps.max() <= 1 and np.mean(ulps > 0) < 0.005
    # IEEE special operands follow the platform pow
    sp = np.array([0.0, -2.0, np.inf, np.nan, 1.0])
    assert np.array_equal(mlr.host_pow(sp, np.full(5, 2.5)), np.array([math.pow(v, 2.5) if v >= 0 or np.isnan(v) else np.nan
                                                                      for v in sp]), equal_nan=True)
