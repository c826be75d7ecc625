"""Pins of the CPU oracle where the reference holds no output values: closed
forms, invariants and independent brute-force restatements, each chosen so a
plausible mistake (dropped term, wrong index or sign, transposed operand)
fails it."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle_pins import central_stencil

RIEMANN2D = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                      [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])


# ------------------------------------------------------------------ mesh
@pytest.mark.parametrize("kind,nx,ny,jit", [("cartesian", 7, 5, 0.0), ("cartesian", 6, 4, 0.2),
                                           ("cartesian_tri", 7, 5, 0.0), ("cartesian_tri", 9, 6, 0.2),
                                           ("wedge", 8, 4, 0.0)])
def test_mesh_invariants(oracle, kind, nx, ny, jit):
    Lx, Ly = 2.0, 1.5
    m = oracle.Mesh(kind, nx, ny, Lx, Ly, jit, 3)
    tri = kind == "cartesian_tri"
    assert m.n_cells == (2 if tri else 1) * nx * ny
    assert m.n_faces_real == (3 if tri else 2) * nx * ny + nx + ny
    assert m.n_faces == (6 * nx * ny + nx + ny if tri else m.n_faces_real)
    if kind != "wedge":
        assert math.isclose(m.cell_volume.sum(), Lx * Ly, rel_tol=1e-13)
    assert (m.cell_volume > 0).all()
    # every real face: normal magnitude = area, pointing out of cells_of_face[f][0]
    nf = m.n_faces_real
    n = m.face_normals[:nf]
    assert np.allclose(np.hypot(n[:, 0], n[:, 1]), m.face_area[:nf], rtol=1e-14)
    nof = m.nodes_of_face.reshape(-1, 2)[:nf]
    mid = 0.5 * (m.node_coords[nof[:, 0]] + m.node_coords[nof[:, 1]])
    c0 = m.cells_of_face[:nf, 0]
    assert (np.einsum("ij,ij->i", mid - m.cell_coords[c0], n) > 0).all()
    # closed cells: sum of outward area-normals vanishes
    k = 3 if tri else 4
    foc = m.faces_of_cell.reshape(-1, k)
    for c in range(m.n_cells):
        s = np.zeros(2)
        for f in foc[c]:
            s += n[f] if m.cells_of_face[f, 0] == c else -n[f]
        assert np.abs(s).max() < 1e-14 * max(Lx, Ly)
    # zones partition the real faces: interior ones have two cells
    allz = np.concatenate(list(m.zones.values()))
    assert np.array_equal(np.sort(allz), np.arange(nf))
    assert (m.cells_of_face[m.zones["interior"], 1] >= 0).all()
    for z in ("left", "right", "top", "bottom"):
        assert (m.cells_of_face[m.zones[z], 1] == -1).all()
        assert len(m.zones[z]) == (ny if z in ("left", "right") else nx)
    if tri:
        # over-allocated phantom faces: nodes (0, 0), cells (0, 0), NaN normals
        assert (m.cells_of_face[nf:] == 0).all()
        assert np.isnan(m.face_normals[nf:]).all()


# ------------------------------------------------------------------ stencils
def test_central_stencils_brute_force(oracle):
    """Independent ring search (Python sets + stable sort) on small meshes."""
    for jit in (0.0, 0.25):
        m = oracle.Mesh("cartesian_tri", 6, 5, 1.0, 1.0, jit, 17)
        s = oracle.Solver(m, oracle.make_config(order=2))
        T = s.teno()
        og, os_, st = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"]
        for c in range(m.n_cells):
            a, b = os_[og[c]], os_[og[c] + 1]
            assert list(st[a:b]) == central_stencil(m, c, 12)


def test_directional_stencil_properties(oracle):
    """Sizes and membership rules of the directional stencils (the stencils
    themselves are pinned by the brute-force restatement in oracle_pins)."""
    m = oracle.Mesh("cartesian_tri", 8, 8, 1.0, 1.0, 0.2, 5)
    s = oracle.Solver(m, oracle.make_config(order=2))
    T = s.teno()
    og, os_, st = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"]
    for c in range(m.n_cells):
        for sl in range(4):
            a, b = os_[og[c] + sl], os_[og[c] + sl + 1]
            ids = st[a:b]
            assert b - a in (0, 12)
            if sl == 0:
                assert b - a == 12
            if b - a:
                assert ids[0] == c and len(set(ids)) == 12
            if sl > 0:
                f = m.faces_of_cell[3 * c + sl - 1]
                assert (b - a == 0) == (m.cells_of_face[f, 1] == -1)


# ------------------------------------------------------------------ OI, ipt
def _leg(p):
    return np.polynomial.legendre.leg2poly([0] * p + [1])


def _tri_avg(px, py):
    """Average of x^a y^b over the unit right triangle: 2 a! b! / (a+b+2)!."""
    tot = Fraction(0)
    for a, ca in enumerate(px):
        for b, cb in enumerate(py):
            if ca == 0 or cb == 0:
                continue
            tot += Fraction(ca) * Fraction(cb) * 2 * Fraction(math.factorial(a) * math.factorial(b),
                                                               math.factorial(a + b + 2))
    return float(tot)


@pytest.mark.parametrize("order", [1, 2])
def test_oscillation_indicator_and_means_closed_form(oracle, order):
    m = oracle.Mesh("cartesian_tri", 4, 4)
    s = oracle.Solver(m, oracle.make_config(order=order))
    T = s.teno()
    K = T["n_dof"]
    poly = T["poly_indices"].reshape(K, 2)
    # lexicographic index pairs
    assert [tuple(x) for x in poly] == [(p - j, j) for p in range(order + 1) for j in range(p + 1)]
    # means of psi_k over the target triangle (cell 0 maps to the reference triangle)
    ipt = [_tri_avg(_leg(px), _leg(py)) for px, py in poly]
    assert np.allclose(T["integral_psi_target"], ipt, atol=1e-15)
    # OI[i][j] = sum_{k>=1} avg(D^k psi_i D^k psi_j), exact rational integration
    OI = np.zeros((K, K))
    for i in range(K):
        for j in range(K):
            acc = 0.0
            for kx, ky in poly[1:]:
                di = np.polynomial.polynomial.polymul(np.polynomial.polynomial.polyder(_leg(poly[i][0]), kx) if kx <= poly[i][0] else [0],
                                                      [1])
                dj = np.polynomial.polynomial.polyder(_leg(poly[j][0]), kx) if kx <= poly[j][0] else [0]
                ei = np.polynomial.polynomial.polyder(_leg(poly[i][1]), ky) if ky <= poly[i][1] else [0]
                ej = np.polynomial.polynomial.polyder(_leg(poly[j][1]), ky) if ky <= poly[j][1] else [0]
                acc += _tri_avg(np.polynomial.polynomial.polymul(di, dj), np.polynomial.polynomial.polymul(ei, ej))
            OI[i, j] = acc
    assert np.allclose(T["oscillation_indicator"].reshape(K, K), OI, atol=1e-13)


# ------------------------------------------------------------------ TENO kernel
def _tri_vertices(m, c):
    return m.node_coords[m.nodes_of_cell[3 * c:3 * c + 3]]


def _cell_avg_quadratic(v, coef):
    """Exact triangle average of a + b x + c y + d x^2 + e x y + f y^2."""
    a, b, c, d, e, f = coef
    x, y = v[:, 0], v[:, 1]
    mx, my = x.mean(), y.mean()
    xx = (np.sum(x * x) + x[0] * x[1] + x[0] * x[2] + x[1] * x[2]) / 6.0
    yy = (np.sum(y * y) + y[0] * y[1] + y[0] * y[2] + y[1] * y[2]) / 6.0
    xy = (2 * np.sum(x * y) + x[0] * (y[1] + y[2]) + x[1] * (y[0] + y[2]) + x[2] * (y[0] + y[1])) / 12.0
    return a + b * mx + c * my + d * xx + e * xy + f * yy


def test_corrected_reconstruction_exact_for_quadratics(oracle):
    """Zero-mean face evaluation reproduces any quadratic at the face
    midpoints: pins A+, transformed areas, quadrature, basis and ipt."""
    m = oracle.Mesh("cartesian_tri", 10, 10, 1.0, 1.0, 0.2, 9)
    s = oracle.Solver(m, oracle.make_config(order=2, teno_variant="corrected",
                                            bcs=[(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]))
    # a physical state (positive density and pressure), so the positivity
    # scaling of the corrected variant stays inactive
    coefs = [(1.0, 0.3, -0.2, 0.5, -0.7, 0.25), (0.5, 0.1, 0.0, 0.0, 0.0, 0.0), (0.2, 0.0, 0.0, -0.3, 0.2, 0.4),
             (3.0, -0.1, 0.2, 0.1, 0.1, -0.2)]
    U = np.zeros((m.n_cells, 4))
    for c in range(m.n_cells):
        v = _tri_vertices(m, c)
        U[c] = [_cell_avg_quadratic(v, k) for k in coefs]
    s.calc_rhs(U)
    fv = s.face_values(1)
    nof = m.nodes_of_face.reshape(-1, 2)
    for f in m.zones["interior"]:
        mid = 0.5 * (m.node_coords[nof[f, 0]] + m.node_coords[nof[f, 1]])
        x, y = mid
        ex = [k[0] + k[1] * x + k[2] * y + k[3] * x * x + k[4] * x * y + k[5] * y * y for k in coefs]
        assert np.allclose(fv[f, 0, 0], ex, rtol=0, atol=5e-12)
        assert np.allclose(fv[f, 0, 1], ex, rtol=0, atol=5e-12)
    assert s.decisions().all()


def test_reference_face_offset_is_cell_constant(oracle):
    """The reference evaluates psi_k + ipt_k/ta_0 instead of psi_k - ipt_k,
    so in central-stencil cells its face values exceed the zero-mean ones by
    sum_k a_k (ipt_k/ta_0 + ipt_k), the same on every face of the cell."""
    m = oracle.Mesh("cartesian_tri", 8, 8, 1.0, 1.0, 0.15, 2)
    bcs = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
    cc = m.cell_coords
    U = np.stack([1 + 0.2 * cc[:, 0], 0.1 + 0 * cc[:, 0], 0.3 - 0.1 * cc[:, 1], 2.5 + 0.1 * cc[:, 0] * cc[:, 1]], 1)
    out = {}
    for var in ("reference", "corrected"):
        s = oracle.Solver(m, oracle.make_config(order=2, teno_variant=var, bcs=bcs))
        s.calc_rhs(U)
        out[var] = (s.face_values(1), s.decisions())
    d = out["reference"][0] - out["corrected"][0]
    central = out["reference"][1].all(axis=1)
    checked = 0
    for c in np.where(central)[0]:
        vals = []
        for f in m.faces_of_cell[3 * c:3 * c + 3]:
            side = 0 if m.cells_of_face[f, 0] == c else 1
            vals.append(d[f, 0, side])
        vals = np.array(vals)
        assert np.allclose(vals, vals[0], rtol=1e-9, atol=1e-13)
        checked += 1
    assert checked > 50 and np.abs(d[~np.isnan(d)]).max() > 1e-3


# ------------------------------------------------------------------ residual
@pytest.mark.parametrize("recon,variant", [("FO", "reference"), ("TENO", "reference"), ("TENO", "corrected")])
def test_free_stream_preserved(oracle, recon, variant):
    m = oracle.Mesh("cartesian_tri", 9, 7, 1.0, 1.0, 0.2, 4)
    bcs = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
    s = oracle.Solver(m, oracle.make_config(recon=recon, order=2, teno_variant=variant, bcs=bcs))
    s.init_constant([0.7, -0.4], 1.3, 0.01)
    U = s.state()[0]
    r = s.calc_rhs(U)
    assert np.abs(r).max() < 1e-12 * np.abs(U).max() / m.cell_volume.min()


def test_mass_conserved_with_symmetry_walls(oracle):
    m = oracle.Mesh("cartesian_tri", 12, 10, 1.0, 1.0, 0.2, 8)
    s = oracle.Solver(m, oracle.make_config(recon="TENO", order=2, teno_variant="corrected"))
    s.init_quadrant(0.5, 0.5, RIEMANN2D)
    U = s.state()[0]
    r = s.calc_rhs(U)
    tot = (r * m.cell_volume[:, None]).sum(axis=0)
    scale = (np.abs(r) * m.cell_volume[:, None]).sum(axis=0)
    # symmetry walls carry no mass or energy; momentum gets the wall pressure
    assert abs(tot[0]) < 1e-13 * scale[0]
    assert abs(tot[3]) < 1e-13 * scale[3]


# ------------------------------------------------------------------ Sod, exact solution
def _sod_exact(x, t, L=(1.0, 0.0, 1.0), R=(0.125, 0.0, 0.1), g=1.4, x0=0.5):
    """Exact Riemann solution (Toro, ch. 4): Newton on the pressure function."""
    rl, ul, pl = L
    rr, ur, pr = R
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)

    def f(p, rk, pk, ak):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B)), math.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
        return 2 * ak / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1), 1 / (rk * ak) * (p / pk) ** (-(g + 1) / (2 * g))

    p = 0.5 * (pl + pr)
    for _ in range(100):
        fl, dl = f(p, rl, pl, al)
        fr, dr = f(p, rr, pr, ar)
        p -= (fl + fr + ur - ul) / (dl + dr)
    fl, _ = f(p, rl, pl, al)
    fr, _ = f(p, rr, pr, ar)
    us = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    out = np.zeros_like(x)
    for i, xi in enumerate(x):
        s = (xi - x0) / t
        if s < us:  # left of contact: rarefaction (Sod)
            shl, stl = ul - al, us - al * (p / pl) ** ((g - 1) / (2 * g))
            if s < shl:
                out[i] = rl
            elif s > stl:
                out[i] = rl * (p / pl) ** (1 / g)
            else:
                out[i] = rl * (2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)) ** (2 / (g - 1))
        else:  # right: shock
            rs = rr * (p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1)
            S = ur + ar * math.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
            out[i] = rs if s < S else rr
    return out


def test_fo_sod_converges_to_exact_solution(oracle):
    m = oracle.Mesh("cartesian", 400, 1, 1.0, 0.01)
    s = oracle.Solver(m, oracle.make_config(recon="FO", riemann="HLLC"), 2)
    sod = np.array([[0.125, 0, 0, 0.1], [1.0, 0, 0, 1.0], [1.0, 0, 0, 1.0], [0.125, 0, 0, 0.1]])
    s.init_quadrant(0.5, -1.0, sod)
    while s.time() < 0.2:
        dt, _ = s.calc_dt(0.5)
        s.take_step(dt)
    U = s.state()[0]
    ex = _sod_exact(m.cell_coords[:, 0], s.time())
    l1 = np.mean(np.abs(U[:, 0] - ex))
    assert l1 < 0.01, l1


def _smooth_state(m):
    x, y = m.cell_coords[:, 0], m.cell_coords[:, 1]
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    u, v, p = 0.3 + 0 * x, -0.2 + 0 * x, 1.0 + 0.1 * np.cos(2 * np.pi * x)
    R = 101325.0 / (298.15 * 1.225)
    e = R / 0.4 * (p / (rho * R))
    return np.stack([rho, rho * u, rho * v, rho * (e + 0.5 * (u * u + v * v))], axis=1)


def test_reference_weighting_diverges_corrected_does_not(oracle):
    """The reference TENO weighting on a smooth jittered-triangle field goes
    non-finite within one SSP-RK3 step; the corrected weighting (ENO among
    the passing directional stencils, zero-mean face evaluation, positivity
    scaling) stays finite and bounded for 40 steps."""
    m = oracle.Mesh("cartesian_tri", 24, 24, 1.0, 1.0, 0.1, 5)
    bcs = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
    U0 = _smooth_state(m)
    ref = oracle.Solver(m, oracle.make_config(teno_variant="reference", bcs=bcs), 1)
    ref.set_state(U0)
    assert np.isfinite(ref.calc_rhs(U0)).all()
    ref.run(1, 0.4)
    assert not np.isfinite(ref.state()[0]).all()
    cor = oracle.Solver(m, oracle.make_config(teno_variant="corrected", bcs=bcs), 1)
    cor.set_state(U0)
    cor.run(40, 0.4)
    U = cor.state()[0]
    assert np.isfinite(U).all()
    assert 0.75 < U[:, 0].min() and U[:, 0].max() < 1.25


@pytest.mark.parametrize("integrator,coefs", [("FE", [1, 1]), ("SSPRK3", [1, 1, 1 / 2, 1 / 6]),
                                              ("RK4", [1, 1, 1 / 2, 1 / 6, 1 / 24])])
def test_integrator_amplification_closed_form(oracle, integrator, coefs):
    """The solver's stage sequence on dU/dt = lam U multiplies U by the
    truncated exponential of z = lam dt (pins every stage coefficient, not only
    their sum as the constant-RHS test does)."""
    for z in (-0.9, -0.37, 0.25, 1.3):
        U = np.array([1.0, -2.5, 3.0])
        assert oracle.lib().orc_integrator_linear_rhs(oracle.INTEGRATOR[integrator], 0.1, z / 0.1,
                                                      oracle._p(U), 3) == 0
        amp = sum(c * z ** k for k, c in enumerate(coefs))
        assert np.allclose(U, amp * np.array([1.0, -2.5, 3.0]), rtol=1e-14, atol=0)


def _gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w


def _poly2(px, py, x, y):
    return np.polynomial.polynomial.polyval(x, _leg(px)) * np.polynomial.polynomial.polyval(y, _leg(py))


def _tri_mean(f, t0, t1, t2, n=8):
    """Mean of f over a triangle by the collapsed Gauss rule (exact far beyond p=3)."""
    u, wu = _gl01(n)
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu) * (1 - U)
    xi, eta = U, V * (1 - U)
    x = t0[0] + (t1[0] - t0[0]) * xi + (t2[0] - t0[0]) * eta
    y = t0[1] + (t1[1] - t0[1]) * xi + (t2[1] - t0[1]) * eta
    return 2.0 * np.sum(W * f(x, y))


@pytest.mark.parametrize("order,jit", [(1, 0.2), (2, 0.0), (2, 0.2), (3, 0.2)])
def test_pseudo_inverse_matches_numpy_svd(oracle, order, jit):
    """The stored (M-1)x(K-1) block of A+ equals numpy's SVD pseudo-inverse of
    the least-squares matrix rebuilt here independently: rows are stencil
    members, columns the Legendre basis means over each member (target-cell
    frame) minus the target-cell means, each row scaled by the member's area
    in that frame (face_reconstruction.cpp:477-741)."""
    m = oracle.Mesh("cartesian_tri", 10, 9, 1.0, 1.0, jit, 3)
    s = oracle.Solver(m, oracle.make_config(order=order))
    T = s.teno()
    K, M = T["n_dof"], T["max_cells"]
    poly = T["poly_indices"].reshape(K, 2)
    ipt = T["integral_psi_target"]
    og, os_, st, orm = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"], T["offsets_recon"]
    L = oracle.lib()
    worst = 0.0
    for c in range(0, m.n_cells, 5):
        nodes = m.nodes_of_cell[3 * c:3 * c + 3]
        w = [m.node_coords[n] for n in nodes]
        J = np.zeros(4)
        L.orc_triangle_J(oracle._p(np.ascontiguousarray(w[0])), oracle._p(np.ascontiguousarray(w[1])),
                         oracle._p(np.ascontiguousarray(w[2])), oracle._p(J))
        Ji = np.linalg.inv(J.reshape(2, 2))
        for sl in range(og[c + 1] - og[c]):
            si = og[c] + sl
            ids = st[os_[si]:os_[si + 1]]
            if len(ids) == 0:
                continue
            B = np.zeros((len(ids) - 1, K - 1))
            for r, nb in enumerate(ids[1:]):
                t = [Ji @ (m.node_coords[n] - w[0]) for n in m.nodes_of_cell[3 * nb:3 * nb + 3]]
                area = 0.5 * abs((t[1][0] - t[0][0]) * (t[2][1] - t[0][1]) - (t[2][0] - t[0][0]) * (t[1][1] - t[0][1]))
                for k in range(1, K):
                    mean = _tri_mean(lambda x, y: _poly2(poly[k][0], poly[k][1], x, y), *t)
                    B[r, k - 1] = area * (mean - ipt[k])
            ref = np.linalg.pinv(B)
            got = T["recon"][orm[si]:orm[si] + K * M].reshape(K, M)
            assert np.all(got[0] == 0) and np.all(got[:, 0] == 0)
            err = np.abs(got[1:, 1:] - ref).max() / np.abs(ref).max()
            worst = max(worst, err)
    assert worst < 1e-10, worst


@pytest.mark.parametrize("order", [2, 3])
def test_precompute_independent_of_thread_count(oracle, order):
    """The oracle searches stencils and builds the pseudo-inverses cell by
    cell in parallel; every array equals the one-thread build bit for bit."""
    m = oracle.Mesh("cartesian_tri", 30, 26, 1.0, 1.0, 0.2, 41)
    cfg = oracle.make_config(order=order, teno_variant="corrected")
    one = oracle.Solver(m, cfg, 1).teno()
    many = oracle.Solver(m, cfg, 8).teno()
    for k, v in one.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, many[k]), k


@pytest.mark.parametrize("recon,bc", [("TENO", "extrapolation"), ("TENO", "symmetry"), ("FO", "wall_adiabatic")])
def test_canonical_mode_equals_one_thread(oracle, recon, bc):
    """The canonical gather at 8 threads reproduces the one-thread run bit
    for bit (residual parts included); the atomic mode agrees to rounding."""
    m = oracle.Mesh("cartesian_tri", 40, 34, 1.0, 1.0, 0.2, 8)
    bcs = [(z, bc, {}) for z in ("left", "right", "top", "bottom")]
    cfg = oracle.make_config(recon=recon, order=2, teno_variant="corrected", bcs=bcs)
    one = oracle.Solver(m, cfg, 1)
    can = oracle.Solver(m, cfg, 8, canonical=True)
    for s in (one, can):
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
    U = one.state()[0]
    for parts in (1, 2, 3):
        assert np.array_equal(one.calc_rhs_parts(U, parts), can.calc_rhs_parts(U, parts)), parts
    assert np.array_equal(one.run(15, 0.4), can.run(15, 0.4))
    assert np.array_equal(one.state()[0], can.state()[0])
    at = oracle.Solver(m, cfg, 8)
    at.init_quadrant(0.5, 0.5, RIEMANN2D)
    r = at.calc_rhs(U)
    r1 = oracle.Solver(m, cfg, 1).calc_rhs(U)
    assert np.allclose(r, r1, rtol=1e-13, atol=1e-13 * np.abs(r1).max())


def _as_arrays(m):
    names = list(m.zones.keys())
    faces = [m.zones[z] for z in names]
    return dict(node_coords=m.node_coords, offsets_nodes_of_cell=m.offsets_nodes_of_cell,
                nodes_of_cell=m.nodes_of_cell, faces_of_cell=m.faces_of_cell, nodes_of_face=m.nodes_of_face,
                cells_of_face=m.cells_of_face.reshape(-1), zone_names=names,
                zone_offsets=np.concatenate([[0], np.cumsum([len(f) for f in faces])]).astype(np.uint32),
                zone_faces=np.concatenate(faces).astype(np.uint32))


def test_mesh_from_arrays_equals_generator(oracle):
    """A generated mesh handed back as arrays rebuilds the same geometry, and
    a TENO residual on it is bit-identical (zones keep their names and
    order)."""
    g = oracle.Mesh("cartesian_tri", 12, 9, 1.0, 1.0, 0.2, 3)
    a = oracle.Mesh("arrays", arrays=_as_arrays(g))
    for k in ("cell_coords", "cell_volume", "face_area", "face_normals", "cells_of_face", "nodes_of_face"):
        assert np.array_equal(getattr(g, k), getattr(a, k), equal_nan=True), k
    assert list(a.zones) == list(g.zones)
    bcs = [("left", "extrapolation", {}), ("right", "symmetry", {}), ("top", "wall_adiabatic", {}),
           ("bottom", "extrapolation", {})]
    cg = oracle.make_config(order=2, teno_variant="corrected", bcs=bcs)
    ca = oracle.make_config(order=2, teno_variant="corrected", bcs=bcs, zones=list(a.zones))
    sg, sa = oracle.Solver(g, cg, 1), oracle.Solver(a, ca, 1)
    sg.init_quadrant(0.5, 0.5, RIEMANN2D)
    U = sg.state()[0]
    assert np.array_equal(sg.calc_rhs(U), sa.calc_rhs(U))
