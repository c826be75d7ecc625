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


def test_ensemble_refuses_bad_dt_mode_before_allocating(mlr):
    """An ensemble checks its time-step mode when it is created, before any
    case allocates device memory (so no GPU is touched here): a fixed dt of
    zero or below (cfl = 0) and a negative CFL number are refused with the
    reference's messages (solver.cpp:580-590)."""
    import pytest
    gens = [mlr.mesh_gen("cartesian_tri", 8, 8, 1.0, 1.0, 0.0, 0)]
    for cfl, dt, msg in ((0.0, 0.0, "fixed dt must be positive"), (0.0, -1e-3, "dt negative"),
                         (-0.5, 1e-3, "dt negative")):
        with pytest.raises(mlr.MallardError, match=msg):
            mlr.Ensemble(gens, mlr.numerics(recon="TENO", order=2, cfl=cfl, dt=dt))


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


def _glibc_pow(x, y):
    """C pow from the C library itself (Python's math.pow raises where C99
    returns infinities and NaNs)."""
    import ctypes
    import ctypes.util
    libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
    f = libm.pow
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double, ctypes.c_double]
    return np.array([f(float(a), float(b)) for a, b in zip(x, y)])


SPECIAL_X = np.array([0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, np.inf, -np.inf, np.nan, 5e-324, -5e-324,
                      1e-310, 1e300, -1e300, 3.0, -3.0])
SPECIAL_Y = np.array([0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 3.0, -3.0, 2.5, -2.5, np.inf, -np.inf, np.nan, 1e300,
                      -1e300, 2.0 ** 53, 2.0 ** 53 + 2, 7.000000000000001, 0.5, 1e-320])


def special_pow_operands():
    """Every pairing of zeros of both signs, +-1, infinities, NaN,
    subnormals, huge values, odd, even and non-integer exponents."""
    X, Y = np.meshgrid(SPECIAL_X, SPECIAL_Y)
    return X.ravel(), Y.ravel()


def test_host_pow_special_operands_follow_c99(mlr):
    """Special operands are handled case by case (C99 Annex F.9.4.4) and give
    glibc's results bit for bit, signed zeros, infinities and NaNs included;
    subnormal bases, negative bases with integer exponents and results near
    overflow stay within 1 ulp of glibc, results in the subnormal range
    within 1 unit of 2^-1074."""
    x, y = special_pow_operands()
    h, g = mlr.host_pow(x, y), _glibc_pow(x, y)
    same = (h.view(np.int64) == g.view(np.int64)) | (np.isnan(h) & np.isnan(g))
    assert same.all(), list(zip(x[~same], y[~same], h[~same], g[~same]))
    rng = np.random.default_rng(11)
    cases = [(-np.exp(rng.uniform(-5, 5, 4000)), rng.integers(-40, 40, 4000).astype(float)),
             (rng.uniform(1e-320, 2.2e-308, 4000), rng.uniform(0.01, 1.0, 4000)),
             (np.exp(rng.uniform(700, 712, 4000) / 7), np.full(4000, 7.0)),
             (np.exp(-rng.uniform(700, 750, 4000) / 7), np.full(4000, 7.0)),
             (np.exp(rng.uniform(-700, 700, 4000)), rng.uniform(-3, 3, 4000))]
    for bx, by in cases:
        h, g = mlr.host_pow(bx, by), _glibc_pow(bx, by)
        assert (np.isinf(h) == np.isinf(g)).all()
        fin = np.isfinite(g)
        assert np.abs(h[fin].view(np.int64) - g[fin].view(np.int64)).max() <= 1


def test_operator_entry_points_refuse_bad_arguments_on_the_host(mlr):
    """The operator entry points check their arguments before touching the
    GPU: unknown Riemann solver or integrator codes are refused with the
    reference's messages, NULL arrays with an error status."""
    import ctypes as C
    L = mlr.lib()
    one = np.zeros(8)
    p = one.ctypes.data_as(C.c_void_p)
    assert L.mlr_riemann_flux(7, 0, 1, p, p, p, p) != 0
    assert "Unknown Riemann solver type" in L.mlr_last_error(None).decode()
    assert L.mlr_riemann_flux(2, 0, -1, p, p, p, p) != 0
    assert L.mlr_riemann_flux(2, 0, 1, None, p, p, p) != 0
    assert L.mlr_riemann_flux(2, 0, 0, None, None, None, None) == 0  # nothing to do
    cb = mlr.RHS_FN(lambda u, r, user: 0)
    vecs = (C.c_void_p * 4)(p, p, p, p)
    assert L.mlr_integrator_take_step(None, 5, 0.1, 8, vecs, vecs, cb, None) != 0
    assert "Unknown time integrator type" in L.mlr_last_error(None).decode()
    assert L.mlr_integrator_take_step(None, 2, 0.1, 8, None, vecs, cb, None) != 0
    assert L.mlr_calc_face_values(None, p, p) != 0
    assert L.mlr_calc_rhs_of(None, p, 3, p) != 0
    assert L.mlr_calc_rhs_device(None, p, 3, p) != 0
