"""The CPU oracle against the values the reference's own unit tests hold
(tests/golden/reference_unit_vectors.json), at the reference's 1e-6 and, where
the value is exact in FP64, bitwise."""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_unit_vectors.json")))


def _h(g, rho, p):
    e = p / ((g - 1.0) * rho)
    return e + p / rho


def _normal(name):
    if name == "x":
        return np.array([1.0, 0.0])
    if name == "y":
        return np.array([0.0, 1.0])
    s = 1.0 / math.sqrt(2.0)
    return np.array([s, s])


@pytest.mark.parametrize("case", GOLDEN["riemann"], ids=lambda c: f"{c['solver']}-{c['normal']}-{c['states']}")
def test_riemann_fluxes(oracle, case):
    st = GOLDEN[case["states"]]
    g = st["gamma"]
    rl, ul0, ul1, pl = st["left"]
    rr, ur0, ur1, pr = st["right"]
    F = oracle.riemann_flux(case["solver"], _normal(case["normal"]), rl, [ul0, ul1], pl, g, _h(g, rl, pl),
                            rr, [ur0, ur1], pr, g, _h(g, rr, pr))
    assert np.allclose(F, case["flux"], atol=1e-6, rtol=0)


def test_hll_and_roe_consistency(oracle):
    """HLL has no reference test; both it and Roe must reduce to the physical
    flux for equal states and agree with HLLC's mass flux sign on Sod."""
    g = 1.4
    for s in ("HLL", "Roe"):
        F = oracle.riemann_flux(s, [1.0, 0.0], 1.0, [0.3, -0.2], 1.0, g, _h(g, 1.0, 1.0), 1.0, [0.3, -0.2], 1.0,
                                g, _h(g, 1.0, 1.0))
        E = 1.0 / 0.4 + 0.5 * (0.09 + 0.04)
        assert np.allclose(F, [0.3, 0.3 * 0.3 + 1.0, -0.2 * 0.3, (E + 1.0) * 0.3], rtol=1e-14)
        F = oracle.riemann_flux(s, [1.0, 0.0], 1.0, [0.0, 0.0], 1.0, g, _h(g, 1, 1), 0.125, [0.0, 0.0], 0.1, g,
                                _h(g, 0.125, 0.1))
        assert F[0] > 0.3 and F[3] > 0.9


def test_physics_constants(oracle):
    P = GOLDEN["physics"]
    out = np.zeros(3)
    oracle.lib().orc_physics_constants(P["gamma"], P["p_ref"], P["T_ref"], P["rho_ref"], oracle._p(out))
    assert np.allclose(out, [P["R"], P["cp"], P["cv"]], atol=1e-6, rtol=0)
    # primitive round trip (physics_test.cpp:60-111)
    R, cp, cv = out
    rt = P["roundtrip"]
    u, p, T = rt["u"], rt["p"], rt["T"]
    rho = p / (R * T)
    e = cv * T
    h = e + p / rho
    E = e + 0.5 * (u[0] * u[0] + u[1] * u[1])
    cons = np.array([rho, rho * u[0], rho * u[1], rho * E])
    prim = np.zeros(5)
    k = np.array([P["gamma"], R, cp, cv, P["p_min"], P["p_max"]])
    oracle.lib().orc_primitives(oracle._p(k), oracle._p(cons), oracle._p(prim))
    assert np.allclose(prim, [u[0], u[1], p, T, h], atol=1e-6, rtol=0)


def test_small_dense_algebra(oracle):
    L = oracle.lib()
    p = oracle._p
    for key, fn in (("invert2", L.orc_invert2), ("invert3", L.orc_invert3)):
        A = np.array(GOLDEN[key]["A"])
        out = np.zeros_like(A)
        fn(p(A), p(out))
        assert np.allclose(out, GOLDEN[key]["Ainv"], atol=1e-6)
    for c in GOLDEN["gemm"]:
        A, B = np.array(c["A"], float), np.array(c["B"], float)
        m, n, pp, q = c["dims"]
        C = np.zeros(len(c["C"]))
        L.orc_gemm(p(A), p(B), p(C), m, n, pp, q, c["tA"], c["tB"])
        assert np.array_equal(C, c["C"])
    for c in GOLDEN["qr"]:
        A = np.array(c["A"])
        R = np.zeros_like(A)
        L.orc_qr_noq(p(A), p(R), c["m"], c["n"])
        assert np.allclose(R, c["R"], atol=1e-6)
    for c in GOLDEN["forward_substitution"]:
        Lm, B = np.array(c["L"], float), np.array(c["B"], float)
        X = np.zeros(len(c["X"]))
        L.orc_forward_substitution(p(Lm), p(B), p(X), *c["mnp"], c["tL"], c["tB"])
        assert np.allclose(X, c["X"], atol=1e-6)
    for c in GOLDEN["back_substitution"]:
        U, B = np.array(c["U"], float), np.array(c["B"], float)
        X = np.zeros(len(c["X"]))
        L.orc_back_substitution(p(U), p(B), p(X), *c["mnp"], c["tU"], c["tB"])
        assert np.allclose(X, c["X"], atol=1e-6)
    for c in GOLDEN["triangle_area"]:
        v = [np.array(x, float) for x in c["v"]]
        assert L.orc_triangle_area(p(v[0]), p(v[1]), p(v[2])) == c["area"]
    for c in GOLDEN["triangle_J"]:
        v = [np.array(x, float) for x in c["v"]]
        J, Ji = np.zeros(4), np.zeros(4)
        L.orc_triangle_J(p(v[0]), p(v[1]), p(v[2]), p(J))
        L.orc_invert2(p(J), p(Ji))
        assert np.array_equal(J, c["J"]) and np.array_equal(Ji, c["Jinv"])


@pytest.mark.parametrize("integrator", ["FE", "RK4", "SSPRK3"])
def test_integrators_constant_rhs(oracle, integrator):
    """time_integrator_test.cpp: rhs = U0 = i_cell * 4 + i, dt = 0.1 -> 1.1 U0."""
    U = np.arange(8, dtype=float)
    rhs = np.arange(8, dtype=float)
    oracle.lib().orc_integrator_constant_rhs(oracle.INTEGRATOR[integrator], 0.1, oracle._p(U), 8, oracle._p(rhs))
    assert np.allclose(U, 1.1 * np.arange(8), atol=1e-6, rtol=0)


def _phys_flux(n, rho, u, p, g):
    un = u[0] * n[0] + u[1] * n[1]
    E = p / (g - 1.0) / rho + 0.5 * (u[0] ** 2 + u[1] ** 2)
    return np.array([rho * un, rho * u[0] * un + p * n[0], rho * u[1] * un + p * n[1], (rho * E + p) * un])


def test_roe_textbook_properties(oracle):
    """Roe's flux (no reference counterpart) pinned by properties of the
    exact Roe linearisation: a stationary contact passes no mass or energy
    and only its pressure as momentum; a supersonic pair whose waves all run
    one way is the upwind side's physical flux; the flux is rotation
    invariant (normal and tangential components follow the normal)."""
    g = 1.4
    # stationary contact: density jump, equal pressure, no velocity
    F = oracle.riemann_flux("Roe", [1.0, 0.0], 1.0, [0.0, 0.0], 1.0, g, _h(g, 1.0, 1.0), 0.3, [0.0, 0.0], 1.0, g,
                            _h(g, 0.3, 1.0))
    assert abs(F[0]) < 1e-15 and abs(F[3]) < 1e-15 and abs(F[1] - 1.0) < 1e-15 and abs(F[2]) < 1e-15
    # supersonic to the right (all eigenvalues > 0): the left state's flux
    uL, uR = [3.5, 0.2], [3.2, -0.1]
    F = oracle.riemann_flux("Roe", [1.0, 0.0], 1.0, uL, 1.0, g, _h(g, 1.0, 1.0), 0.8, uR, 0.7, g, _h(g, 0.8, 0.7))
    assert np.allclose(F, _phys_flux([1.0, 0.0], 1.0, uL, 1.0, g), rtol=1e-12, atol=1e-13)
    # ... and to the left: the right state's flux
    uL, uR = [-3.5, 0.2], [-3.2, -0.1]
    F = oracle.riemann_flux("Roe", [1.0, 0.0], 1.0, uL, 1.0, g, _h(g, 1.0, 1.0), 0.8, uR, 0.7, g, _h(g, 0.8, 0.7))
    assert np.allclose(F, _phys_flux([1.0, 0.0], 0.8, uR, 0.7, g), rtol=1e-12, atol=1e-13)
    # rotation invariance: rotate states and normal together
    th = 0.7
    c, s = np.cos(th), np.sin(th)
    rot = lambda v: [c * v[0] - s * v[1], s * v[0] + c * v[1]]
    uL, uR = [0.4, 0.1], [-0.2, 0.3]
    F0 = oracle.riemann_flux("Roe", [1.0, 0.0], 1.0, uL, 1.0, g, _h(g, 1.0, 1.0), 0.5, uR, 0.4, g, _h(g, 0.5, 0.4))
    F1 = oracle.riemann_flux("Roe", rot([1.0, 0.0]), 1.0, rot(uL), 1.0, g, _h(g, 1.0, 1.0), 0.5, rot(uR), 0.4, g,
                             _h(g, 0.5, 0.4))
    assert abs(F1[0] - F0[0]) < 1e-13 and abs(F1[3] - F0[3]) < 1e-13
    assert np.allclose([F1[1], F1[2]], rot([F0[1], F0[2]]), atol=1e-13)
