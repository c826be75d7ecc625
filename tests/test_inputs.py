"""Input compatibility on the CPU: the reference's three example inputs
(tests/golden/examples/, copied from examples/*/input.toml) parse to the
effective settings the reference's Solver::init reads, including its quirks:
n_steps defaults to UINT32_MAX (solver.cpp:270), the examples'
quadrature_type / quadrature_order keys are ignored (face_reconstruction.cpp:
113-116 reads quadrature_*_cell / _face), and the basis defaults apply. The
exprtk subset is pinned by the closed forms of the examples' expressions and
by operator semantics."""
import os

import numpy as np
import pytest
import tomli

HERE = os.path.join(os.path.dirname(__file__), "golden", "examples")
UINT32_MAX = 4294967295.0


def _bcs(num):
    return [(num.bc[i].zone.decode(), num.bc[i].type) for i in range(num.n_bc)]


def test_sod_example(mlr):
    g, num, run, init = mlr.parse_config(os.path.join(HERE, "sod.toml"))
    assert (g.type, g.nx, g.ny, g.Lx, g.Ly, g.jitter) == (mlr.MESH["cartesian"], 1000, 1, 1.0, 0.001, 0.0)
    assert num.recon == mlr.RECON["FO"]
    assert num.riemann == mlr.RIEMANN["HLLC"] and num.integrator == mlr.INTEGRATOR["SSPRK3"]
    assert (num.gamma, num.p_ref, num.T_ref, num.rho_ref) == (1.4, 101325.0, 298.15, 1.225)
    assert (num.p_min, num.p_max) == (-1e20, 1e20)
    assert run["cfl"] == 1.0 and run["t_stop"] == 0.2 and run["n_steps"] == UINT32_MAX
    assert run["check_nan"] == 1.0 and run["check_interval"] == 100
    assert init == "analytical"
    assert _bcs(num) == [(z, mlr.BC["symmetry"]) for z in ("left", "right", "top", "bottom")]


def test_riemann_2d_example(mlr):
    g, num, run, init = mlr.parse_config(os.path.join(HERE, "riemann_2d.toml"))
    assert (g.type, g.nx, g.ny) == (mlr.MESH["cartesian_tri"], 1024, 1024)
    assert num.recon == mlr.RECON["TENO"] and num.basis == mlr.BASIS["legendre"] and num.order == 3
    assert num.stencil_factor == 2.0
    # quadrature_type / quadrature_order are not the keys the reference reads:
    # the defaults (Dunavant p+1, Gauss-Legendre (p+1)/2) stay in force
    assert (num.quad_order_cell, num.quad_order_face) == (4, 2)  # p + 1, (p + 1) // 2 at p = 3
    assert num.teno_variant == mlr.TENO_VARIANT["reference"]
    assert run["cfl"] == 0.1 and run["t_stop"] == 0.3 and run["n_steps"] == UINT32_MAX
    assert init == "analytical"


def test_wedge_example(mlr):
    g, num, run, init = mlr.parse_config(os.path.join(HERE, "wedge.toml"))
    assert (g.type, g.nx, g.ny, g.Lx, g.Ly) == (mlr.MESH["wedge"], 150, 50, 4.0, 1.5)
    assert num.p_min == 1e-6 and run["check_nan"] == 0.0
    assert init == "constant"
    b = {num.bc[i].zone.decode(): num.bc[i] for i in range(num.n_bc)}
    assert b["left"].type == mlr.BC["upt"] and tuple(b["left"].u) == (600.0, 0.0)
    assert (b["left"].p, b["left"].T) == (101325.0, 300.0)
    assert b["right"].type == mlr.BC["p_out"] and b["right"].p == 101325.0
    assert b["top"].type == b["bottom"].type == mlr.BC["symmetry"]


def test_example_expressions_closed_form(mlr):
    rng = np.random.default_rng(0)
    x, y = rng.uniform(0, 1, 4000), rng.uniform(0, 1, 4000)
    x[:4], y[:4] = [0.8, 0.7999999, 0.5, 0.0], [0.8, 0.9, 0.79999999, 1.0]  # on and beside the split
    r2 = tomli.load(open(os.path.join(HERE, "riemann_2d.toml"), "rb"))["initialize"]
    l, r, b, t = x < 0.8, x >= 0.8, y < 0.8, y >= 0.8
    rho = 1.5 * r * t + 0.532258064516129 * l * t + 0.137992831541219 * l * b + 0.532258064516129 * r * b
    u = 0.0 * r * t + 1.206045378311055 * l * t + 1.206045378311055 * l * b + 0.0 * r * b
    v = 0.0 * r * t + 0.0 * l * t + 1.206045378311055 * l * b + 1.206045378311055 * r * b
    p = 1.5 * r * t + 0.3 * l * t + 0.029032258064516 * l * b + 0.3 * r * b
    assert np.array_equal(mlr.eval_expr(r2["rho"], x, y), rho)
    assert np.array_equal(mlr.eval_expr(r2["u"][0], x, y), u)
    assert np.array_equal(mlr.eval_expr(r2["u"][1], x, y), v)
    assert np.array_equal(mlr.eval_expr(r2["p"], x, y), p)
    sod = tomli.load(open(os.path.join(HERE, "sod.toml"), "rb"))["initialize"]
    assert np.array_equal(mlr.eval_expr(sod["rho"], x, y), np.where(x < 0.5, 1.0, 0.125))
    assert np.array_equal(mlr.eval_expr(sod["u"][0], x, y), np.zeros_like(x))


@pytest.mark.parametrize("expr,ref", [
    ("1 + 2 * 3", lambda x, y: 7.0),
    ("2 ^ 3 ^ 2", lambda x, y: 2.0 ** 9),            # right associative
    ("-x ^ 2", lambda x, y: -(x ** 2)),
    ("(x + y) / 2 - x * y", lambda x, y: (x + y) / 2 - x * y),
    ("var a := x > 0.5; var b := y <= 0.25; a and b", lambda x, y: ((x > 0.5) & (y <= 0.25)) * 1.0),
    ("x < 0.3 or not (y > 0.6)", lambda x, y: ((x < 0.3) | ~(y > 0.6)) * 1.0),
    ("x == y", lambda x, y: (x == y) * 1.0),
    ("sin(pi * x) * exp(y)", lambda x, y: np.sin(np.pi * x) * np.exp(y)),
    ("max(x, y) - min(x, 2 * y) + sqrt(abs(x - y))", lambda x, y: np.maximum(x, y) - np.minimum(x, 2 * y) +
     np.sqrt(np.abs(x - y))),
    ("1 + 0.2 * cos(2 * pi * x) * tanh(y / 0.1)", lambda x, y: 1 + 0.2 * np.cos(2 * np.pi * x) * np.tanh(y / 0.1)),
])
def test_expression_semantics(mlr, expr, ref):
    rng = np.random.default_rng(1)
    x, y = rng.uniform(0, 1, 500), rng.uniform(0, 1, 500)
    x[:3] = y[:3]
    got = mlr.eval_expr(expr, x, y)
    assert np.allclose(got, ref(x, y), rtol=1e-15, atol=0)


def test_expression_errors(mlr):
    with pytest.raises(mlr.MallardError):
        mlr.eval_expr("1 + ", [0.0], [0.0])
    with pytest.raises(mlr.MallardError):
        mlr.eval_expr("var q := 1; z + q", [0.0], [0.0])
    with pytest.raises(mlr.MallardError):
        mlr.eval_expr("erfc(x)", [0.0], [0.0])


def test_navier_stokes_keys(mlr, tmp_path):
    """[physics] type = "navier-stokes" and the no-slip wall (extension keys)."""
    base = open(os.path.join(HERE, "wedge.toml")).read()
    ns = base.replace('type = "euler"', 'type = "navier-stokes"\nmu_ref = 1.8e-5\nsutherland_S = 110.4\nPr = 0.71')
    ns = ns.replace('name = "bottom"\ntype = "symmetry"', 'name = "bottom"\ntype = "wall_noslip"')
    assert ns != base
    p = tmp_path / "ns.toml"
    p.write_text(ns)
    g, num, run, init = mlr.parse_config(p)
    assert num.viscous == 1 and num.mu_ref == 1.8e-5 and num.sutherland_S == 110.4 and num.prandtl == 0.71
    assert num.T_mu_ref == 298.15  # defaults to T_ref
    b = {num.bc[i].zone.decode(): num.bc[i] for i in range(num.n_bc)}
    assert b["bottom"].type == mlr.BC["wall_noslip"]
    g, num, run, init = mlr.parse_config(os.path.join(HERE, "wedge.toml"))
    assert num.viscous == 0
    p.write_text(ns.replace("mu_ref = 1.8e-5\n", ""))
    with pytest.raises(mlr.MallardError, match="Missing mu_ref for physics: navier-stokes."):
        mlr.parse_config(p)
    p.write_text(base.replace('type = "euler"', 'type = "stokes"'))
    with pytest.raises(mlr.MallardError, match="Unknown physics type: stokes."):
        mlr.parse_config(p)


@pytest.mark.parametrize("run,msg", [
    ("n_steps = 5", "Either dt or cfl must be specified."),
    ("cfl = 0.5\ndt = 1e-3\nn_steps = 5", "Only one of dt or cfl can be specified."),
    ("cfl = 0.5", "Either n_steps, t_stop, or t_wall_stop must be specified."),
])
def test_run_section_errors(mlr, tmp_path, run, msg):
    """The reference's [run] checks and messages (solver.cpp:246-258)."""
    src = open(os.path.join(HERE, "sod.toml")).read()
    head, tail = src.split("[run]", 1)
    rest = tail.split("\n[", 1)
    p = tmp_path / "in.toml"
    p.write_text(head + "[run]\n" + run + "\n" + ("\n[" + rest[1] if len(rest) > 1 else ""))
    with pytest.raises(mlr.MallardError) as e:
        mlr.parse_config(p)
    assert str(e.value) == msg


def test_run_section_fixed_dt(mlr, tmp_path):
    """dt instead of cfl selects the fixed step (solver.cpp:260-264); cfl
    then reads 0 and n_steps keeps its UINT32_MAX default."""
    src = open(os.path.join(HERE, "sod.toml")).read()
    head, tail = src.split("[run]", 1)
    rest = tail.split("\n[", 1)
    p = tmp_path / "in.toml"
    p.write_text(head + "[run]\ndt = 2.5e-4\nt_stop = 0.01\n" + ("\n[" + rest[1] if len(rest) > 1 else ""))
    run = mlr.parse_config(p)[2]
    assert run["cfl"] == 0.0 and run["dt"] == 2.5e-4
    assert run["t_stop"] == 0.01 and run["n_steps"] == UINT32_MAX


def test_positivity_limiter_key(mlr, tmp_path):
    """[numerics] positivity_limiter (extension key, default off)."""
    base = open(os.path.join(HERE, "sod.toml")).read()
    assert "[numerics]" in base
    p = tmp_path / "pp.toml"
    p.write_text(base.replace("[numerics]", "[numerics]\npositivity_limiter = true", 1))
    assert mlr.parse_config(p)[1].positivity == 1
    assert mlr.parse_config(os.path.join(HERE, "sod.toml"))[1].positivity == 0
