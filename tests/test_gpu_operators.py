"""The reference's strategy interfaces on their own, through the C ABI on an
MI355X: RiemannSolver::calc_flux (mlr_riemann_flux),
FaceReconstruction::calc_face_values on a caller's state
(mlr_calc_face_values), the interior and boundary face sums of calc_rhs
(mlr_calc_rhs_of / mlr_calc_rhs_device; the boundary sum is every
Boundary::apply) and TimeIntegrator::take_step with a caller's rhs function
(mlr_integrator_take_step).

Bar: the reference's own golden fluxes (riemann_solver_test.cpp, 1e-6 as
there, and the oracle's value bit for bit); bit-exact against the oracle
wherever both run the same IEEE operations; the two-rarefaction estimate's
pow differs from glibc's by at most 1 ulp, so random states are compared to
1e-13 relative there.
"""
import json
import math
import os

import numpy as np
import pytest

from hipmem import DeviceArray

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_unit_vectors.json")))
RIEMANN2D = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                      [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])


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


def _side(g, rho, u, v, p):
    return [rho, u, v, p, g, _h(g, rho, p)]


def _oracle_fluxes(oracle, kind, n, L, R, guard=0):
    return np.array([oracle.riemann_flux(kind, n[i], L[i, 0], L[i, 1:3], L[i, 3], L[i, 4], L[i, 5],
                                         R[i, 0], R[i, 1:3], R[i, 3], R[i, 4], R[i, 5], guard=guard)
                     for i in range(len(n))])


def test_riemann_flux_reference_golden(mlr, oracle):
    """The reference's Sod and equal-state fluxes for Rusanov and HLLC
    (riemann_solver_test.cpp:43-265), all cases in one batched call."""
    cases = GOLDEN["riemann"]
    for kind in ("Rusanov", "HLLC"):
        sel = [c for c in cases if c["solver"] == kind]
        n = np.array([_normal(c["normal"]) for c in sel])
        L = np.array([_side(GOLDEN[c["states"]]["gamma"], *GOLDEN[c["states"]]["left"]) for c in sel])
        R = np.array([_side(GOLDEN[c["states"]]["gamma"], *GOLDEN[c["states"]]["right"]) for c in sel])
        F = mlr.riemann_flux(kind, n, L, R)
        want = np.array([c["flux"] for c in sel])
        assert np.allclose(F, want, atol=1e-6, rtol=0), kind
        # the 15-digit HLLC values of the x and y normals agree to rounding
        for i, c in enumerate(sel):
            if kind == "HLLC" and c["states"] == "sod_states" and c["normal"] != "diag":
                assert np.allclose(F[i], c["flux"], rtol=2e-15, atol=2e-16)
        assert np.array_equal(F, _oracle_fluxes(oracle, kind, n, L, R)), kind


def test_riemann_flux_matches_oracle_on_random_states(mlr, oracle):
    """Every solver on random states and normals (both subsonic and
    supersonic pairs, all branches of the adaptive pressure estimate)."""
    rng = np.random.default_rng(17)
    n_pts = 3000
    th = rng.uniform(0.0, 2.0 * np.pi, n_pts)
    n = np.stack([np.cos(th), np.sin(th)], axis=1)
    g = 1.4

    def sides():
        rho = rng.uniform(0.05, 3.0, n_pts)
        p = np.exp(rng.uniform(np.log(0.01), np.log(5.0), n_pts))
        u = rng.uniform(-3.0, 3.0, (n_pts, 2))
        return np.array([_side(g, rho[i], u[i, 0], u[i, 1], p[i]) for i in range(n_pts)])

    L, R = sides(), sides()
    for kind in ("Rusanov", "HLL", "HLLC", "Roe"):
        F = mlr.riemann_flux(kind, n, L, R)
        Fo = _oracle_fluxes(oracle, kind, n, L, R)
        # strong two-sided expansions drive the reference's two-rarefaction
        # estimate to a negative base of pow (NaN) on both sides alike
        assert np.array_equal(np.isnan(F), np.isnan(Fo)), kind
        ok = np.isfinite(Fo).all(axis=1)
        assert ok.mean() > 0.8, kind
        scale = np.maximum(np.abs(Fo[ok]), 1e-12)
        assert (np.abs(F[ok] - Fo[ok]) / scale).max() < 1e-13, kind
        exact = np.mean((F[ok] == Fo[ok]).all(axis=1))
        assert exact > (0.999 if kind in ("Rusanov", "Roe") else 0.95), (kind, exact)
    # Rusanov and Roe take no pow: bit for bit
    for kind in ("Rusanov", "Roe"):
        assert np.array_equal(mlr.riemann_flux(kind, n, L, R), _oracle_fluxes(oracle, kind, n, L, R))
    # the vacuum guard only changes points whose star-region terms go
    # negative, and leaves no NaN
    Fg = mlr.riemann_flux("HLLC", n, L, R, guard=1)
    F0 = mlr.riemann_flux("HLLC", n, L, R, guard=0)
    ok = np.isfinite(F0).all(axis=1)
    assert np.array_equal(Fg[ok], F0[ok]) and np.isfinite(Fg).all()
    Fgo = _oracle_fluxes(oracle, "HLLC", n, L, R, guard=1)
    assert (np.abs(Fg - Fgo) / np.maximum(np.abs(Fgo), 1e-12)).max() < 1e-13


def test_riemann_flux_refuses_bad_input(mlr):
    n = np.array([[1.0, 0.0]])
    s = np.array([_side(1.4, 1.0, 0.0, 0.0, 1.0)])
    with pytest.raises(mlr.MallardError, match="Unknown Riemann solver type"):
        mlr.riemann_flux(9, n, s, s)
    assert mlr.riemann_flux("HLLC", np.zeros((0, 2)), np.zeros((0, 6)), np.zeros((0, 6))).shape == (0, 4)


def _pair(mlr, oracle, kind, nx, ny, jitter, seed, bcs=None, **kw):
    num = mlr.numerics(bcs=bcs, **kw)
    dev = mlr.Solver(gen=mlr.mesh_gen(kind, nx, ny, 1.0, 1.0, jitter, seed), num=num)
    m = oracle.Mesh(kind, nx, ny, 1.0, 1.0, jitter, seed)
    okw = {k: v for k, v in kw.items() if k not in ("cfl", "dt", "stencil_factor")}
    orc = oracle.Solver(m, oracle.make_config(bcs=bcs, **okw), 1)
    return dev, orc, m


def _smooth_state(m, seed):
    """A smooth, non-uniform state at the cell centroids."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.5, 6)
    x, y = m.cell_coords[:, 0], m.cell_coords[:, 1]
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * a[0] * x) * np.cos(2 * np.pi * a[1] * y)
    u = 0.3 * np.cos(2 * np.pi * a[2] * y)
    v = -0.2 * np.sin(2 * np.pi * a[3] * x)
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (a[4] * x + a[5] * y))
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=1)


@pytest.mark.parametrize("recon,variant,order", [("TENO", "corrected", 2), ("TENO", "reference", 2),
                                                  ("TENO", "corrected", 3), ("FO", "corrected", 1)])
def test_calc_face_values_of_caller_state(mlr, oracle, recon, variant, order):
    """Face values of a state the solver does not hold equal the oracle's
    (bitwise for the corrected weighting and first order; 1e-12 for the
    reference weighting's x^6), and the solver's own state, time and dt are
    left as they were."""
    dev, orc, m = _pair(mlr, oracle, "cartesian_tri", 18, 14, 0.2, 5, recon=recon, order=order,
                        teno_variant=variant, cfl=0.4)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    dev.run_steps(2)
    U0, W0 = dev.state()
    t0 = dev.time()
    U = _smooth_state(m, 3)
    fv = dev.calc_face_values(U)
    orc.calc_rhs(U)
    fo = orc.face_values(dev.Q)
    written = ~np.isnan(fv)
    assert written.sum() > 0
    if variant == "reference" and recon == "TENO":
        assert np.allclose(fv[written], fo[written], rtol=1e-12, atol=0)
    else:
        assert np.array_equal(fv[written], fo[written])
    # (the reference weighting's own state is already non-finite after two
    # steps: compare bit patterns)
    U1, W1 = dev.state()
    assert np.array_equal(U0.view(np.int64), U1.view(np.int64))
    assert np.array_equal(W0.view(np.int64), W1.view(np.int64)) and dev.time() == t0


@pytest.mark.parametrize("recon", ["FO", "TENO"])
def test_rhs_interior_and_boundary_parts(mlr, oracle, recon):
    """calc_rhs_interior and calc_rhs_boundaries (every Boundary::apply) of a
    caller's state, each divided by the volume, equal the oracle's parts bit
    for bit; they add up to the full residual to rounding. Walls, fixed state,
    p_out and extrapolation zones."""
    bcs = [("left", "upt", {"u": (0.4, 0.0), "p": 1.0, "T": 1.0 / 287.0}),
           ("right", "p_out", {"p": 0.9}),
           ("top", "wall_adiabatic", {}),
           ("bottom", "extrapolation", {})]
    dev, orc, m = _pair(mlr, oracle, "cartesian_tri", 16, 12, 0.15, 9, bcs=bcs, recon=recon, order=2,
                        teno_variant="corrected", cfl=0.4, p_ref=1.0, T_ref=1.0 / 287.0, rho_ref=1.0)
    U = _smooth_state(m, 8)
    parts = {}
    for k in (1, 2, 3):
        parts[k] = dev.calc_rhs_of(U, k)
        assert np.array_equal(parts[k], orc.calc_rhs_parts(U, k)), k
    assert np.array_equal(parts[3], dev.calc_rhs_of(U))
    scale = np.abs(parts[3]).max(axis=0)
    assert (np.abs(parts[1] + parts[2] - parts[3]).max(axis=0) <= 1e-12 * scale).all()
    # only cells with a boundary face carry a boundary part
    cof = m.cells_of_face[:m.n_faces_real]
    bcells = np.unique(cof[cof[:, 1] < 0, 0])
    inner = np.setdiff1d(np.arange(m.n_cells), bcells)
    assert not parts[2][inner].any() and parts[2][bcells].any()


def test_wall_boundary_part_closed_form(mlr, oracle):
    """First order with adiabatic walls on every side: the boundary part of
    each wall cell is -(1/V) sum over its wall faces of area (0, p n, 0)
    (boundary_wall_adiabatic.cpp:38-70), with p the cell's pressure."""
    bcs = [(z, "wall_adiabatic", {}) for z in ("left", "right", "top", "bottom")]
    dev, orc, m = _pair(mlr, oracle, "cartesian_tri", 10, 7, 0.2, 4, bcs=bcs, recon="FO", order=1)
    U = _smooth_state(m, 2)
    r = dev.calc_rhs_of(U, mlr.RHS_BOUNDARY)
    p = 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / U[:, 0])
    want = np.zeros_like(r)
    cof = m.cells_of_face
    for f in range(m.n_faces_real):
        if cof[f, 1] >= 0:
            continue
        c = cof[f, 0]
        nv = m.face_normals[f]
        a = m.face_area[f]
        nu = nv / math.hypot(nv[0], nv[1])
        want[c, 1] -= a * p[c] * nu[0]
        want[c, 2] -= a * p[c] * nu[1]
    want /= m.cell_volume[:, None]
    assert np.allclose(r, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_calc_rhs_device_equals_host_entry(mlr):
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 20, 20, 1.0, 1.0, 0.2, 3),
                     num=mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4))
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    U, _ = dev.state()
    perm = dev.permutation().astype(np.int64)
    Ud = DeviceArray(mlr, U.shape, U[perm])
    rd = DeviceArray(mlr, U.shape)
    for parts in (mlr.RHS_INTERIOR, mlr.RHS_BOUNDARY, mlr.RHS_ALL):
        dev.calc_rhs_device(Ud.ptr, rd.ptr, parts)
        r = np.empty_like(U)
        r[perm] = rd.get()
        assert np.array_equal(r, dev.calc_rhs_of(U, parts)), parts
    assert np.array_equal(r, dev.calc_rhs())


def test_integrator_constant_rhs_golden(mlr, oracle):
    """time_integrator_test.cpp:22-181: with a constant rhs equal to U0 and
    dt = 0.1, every integrator ends at 1.1 U0 (1e-6 as there), and equals the
    oracle's integrator sequence bit for bit."""
    G = GOLDEN["integrator_constant_rhs"]
    nc = G["n_cells"]
    U0 = np.arange(4 * nc, dtype=np.float64)
    const = DeviceArray(mlr, 4 * nc, U0)
    copy = mlr.lib().hipMemcpy
    for kind, nsol, nrhs in (("FE", 1, 1), ("RK4", 2, 4), ("SSPRK3", 2, 3)):
        sol = [DeviceArray(mlr, 4 * nc, U0)] + [DeviceArray(mlr, 4 * nc) for _ in range(nsol - 1)]
        rhs = [DeviceArray(mlr, 4 * nc) for _ in range(nrhs)]

        def fn(_U, r):
            if copy(r, const.ptr, 32 * nc, 3) != 0:  # device to device, synchronous
                raise RuntimeError("copy failed")

        mlr.integrator_take_step(kind, G["dt"], 4 * nc, [a.ptr for a in sol], [a.ptr for a in rhs], fn)
        got = sol[0].get()
        assert np.allclose(got, 1.1 * U0, atol=1e-6, rtol=0), kind
        want = U0.copy()
        oracle.lib().orc_integrator_constant_rhs(oracle.INTEGRATOR[kind], G["dt"], oracle._p(want), 4 * nc,
                                                 oracle._p(U0.copy()))
        assert np.array_equal(got, want), kind


def test_integrator_linear_rhs_amplification(mlr, oracle):
    """dU/dt = lam U: one step multiplies U by the integrator's stability
    polynomial (FE 1+z, SSPRK3 1+z+z^2/2+z^3/6, RK4 + z^4/24), and the device
    sequence equals the oracle's bit for bit. The rhs function reads the
    stage state back to the host, scales it and writes the residual."""
    lam, dt = -1.7, 0.13
    z = lam * dt
    poly = {"FE": 1 + z, "SSPRK3": 1 + z + z * z / 2 + z ** 3 / 6,
            "RK4": 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24}
    U0 = np.linspace(0.5, 2.0, 40)
    for kind, nsol, nrhs in (("FE", 1, 1), ("RK4", 2, 4), ("SSPRK3", 2, 3)):
        sol = [DeviceArray(mlr, 40, U0)] + [DeviceArray(mlr, 40) for _ in range(nsol - 1)]
        rhs = [DeviceArray(mlr, 40) for _ in range(nrhs)]
        arrays = {a.ptr: a for a in sol + rhs}

        def fn(U, r):
            arrays[r].upload(lam * arrays[U].get())

        mlr.integrator_take_step(kind, dt, 40, [a.ptr for a in sol], [a.ptr for a in rhs], fn)
        got = sol[0].get()
        assert np.allclose(got, poly[kind] * U0, rtol=1e-14, atol=0), kind
        want = U0.copy()
        oracle.lib().orc_integrator_linear_rhs(oracle.INTEGRATOR[kind], dt, lam, oracle._p(want), 40)
        assert np.array_equal(got, want), kind


@pytest.mark.parametrize("integrator", ["SSPRK3", "RK4", "FE"])
def test_integrator_with_device_residual_equals_solver_step(mlr, integrator):
    """TimeIntegrator::take_step driven by mlr_calc_rhs_device on the
    solver's stream reproduces the fused step (mlr_step) bit for bit: the
    fused stage updates are the reference's copy / axpy / axpby sequence."""
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 24, 20, 1.0, 1.0, 0.2, 21),
                     num=mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4,
                                      integrator=integrator))
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    U, _ = dev.state()
    perm = dev.permutation().astype(np.int64)
    dt = dev.calc_dt()
    nsol, nrhs = {"FE": (1, 1), "RK4": (2, 4), "SSPRK3": (2, 3)}[integrator]
    sol = [DeviceArray(mlr, U.shape, U[perm])] + [DeviceArray(mlr, U.shape) for _ in range(nsol - 1)]
    rhs = [DeviceArray(mlr, U.shape) for _ in range(nrhs)]
    mlr.integrator_take_step(integrator, dt, U.size, [a.ptr for a in sol], [a.ptr for a in rhs],
                             lambda u, r: dev.calc_rhs_device(u, r), solver=dev)
    got = np.empty_like(U)
    got[perm] = sol[0].get()
    dev.take_step()
    assert np.array_equal(got, dev.state()[0])


def test_roe_flux_resolves_isolated_shocks_exactly(mlr):
    """The HIP Roe flux at lone Rankine-Hugoniot shocks (both directions, at
    rest and moving, three normals) equals the upwind side's physical flux."""
    from shocks import cases, enthalpy
    g = 1.4
    cs = cases(g)
    n = np.array([c[0] for c in cs])
    side = lambda S: [S[0], S[1][0], S[1][1], S[2], g, enthalpy(g, S[0], S[2])]
    L = np.array([side(c[1]) for c in cs])
    R = np.array([side(c[2]) for c in cs])
    want = np.array([c[3] for c in cs])
    F = mlr.riemann_flux("Roe", n, L, R)
    for i in range(len(cs)):
        assert np.allclose(F[i], want[i], rtol=1e-12, atol=1e-12 * np.abs(want[i]).max()), i
