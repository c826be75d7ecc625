"""Closed-form and brute-force pins of the oracle's functions that the
reference holds no output values for. Each check takes the oracle binding
(oracle_py, or a copy of it bound to another build) and raises AssertionError
when the oracle disagrees. test_oracle_pins.py runs them on the oracle;
test_oracle_mutations.py runs them on builds of the oracle with a planted
mistake and requires each to fail there.

Every expected value is derived here from the geometry, the textbook formula
or a restatement that shares no code with the oracle:
  - dt / spectral radius (solver.cpp:592-742): per-face terms of a uniform or
    cell-wise state on unjittered meshes, where the centroid offsets and
    normals have closed forms;
  - the TENO cut-off's ENO branch (face_reconstruction.cpp:948-1031) on a
    piecewise-constant state with a straight jump;
  - directional (Type-4) stencils (face_reconstruction.cpp:288-404): a
    brute-force restatement with the cone test by Cramer's rule;
  - the wall, fixed (u, p, T) and p_out boundary fluxes
    (boundary_wall_adiabatic.cpp:38-70, boundary_upt.cpp:83-100,
    boundary_p_out.cpp:57-100) through residuals with closed forms;
  - the HLL flux (riemann_solver.h:377-439) against Toro's HLL formula with
    the adaptive pressure estimate (Toro, Riemann Solvers and Numerical
    Methods for Fluid Dynamics, 3rd ed., sections 9.5 and 10.5);
  - the positivity-preserving flux limiter (no reference counterpart; Hu,
    Adams & Shu 2013): its blend factor against the roots of the density
    (linear) and pressure (quadratic) constraints, and its low-order flux
    against the Euler flux and the mirror-state Lax-Friedrichs flux.
"""
import math

import numpy as np

GAMMA = 1.4
R_GAS = 101325.0 / (298.15 * 1.225)  # p_ref / (T_ref rho_ref) of the reference's defaults
CV = R_GAS / (GAMMA - 1.0)


# ------------------------------------------------------------------ rings
def neighbors1(m, c):
    """The cell and its face neighbours, ascending (mesh.cpp:127-165)."""
    out = {c}
    for f in m.faces_of_cell[m.offsets_faces_of_cell[c]:m.offsets_faces_of_cell[c + 1]]:
        a, b = m.cells_of_face[f]
        if b >= 0:
            out.add(int(b) if a == c else int(a))
    return sorted(out)


def next_ring(m, rings, tgt):
    """Cells adjacent to the last ring that no ring holds yet, in discovery
    order, then stably sorted by squared distance to the target (equal to
    libstdc++'s introsort for rings of up to 16 cells)."""
    old = set().union(*[set(r) for r in rings])
    nr = []
    for c in rings[-1]:
        for x in neighbors1(m, c):
            if x not in old and x not in nr:
                nr.append(x)
    d = [((m.cell_coords[x, 0] - m.cell_coords[tgt, 0]) ** 2 + (m.cell_coords[x, 1] - m.cell_coords[tgt, 1]) ** 2, i, x)
         for i, x in enumerate(nr)]
    d.sort(key=lambda t: t[0])
    assert len(nr) <= 16
    return [x for _, _, x in d]


def central_stencil(m, c, M):
    st, rings = [c], [[c]]
    while len(st) < M:
        rings.append(next_ring(m, rings, c))
        for x in rings[-1]:
            if len(st) == M:
                break
            st.append(x)
    return st


# ------------------------------------------------------------------ dt
def _cart_tri_classes(m, nx, ny, h, k):
    """(kind, i, j) of every cell of an unjittered cartesian_tri mesh from its
    centroid: 'r' is the lower-right triangle of quad (i, j), centroid at
    (i + 2/3, j + 1/3) cells, 'l' the upper-left one at (i + 1/3, j + 2/3)."""
    out = []
    for c in range(m.n_cells):
        x, y = m.cell_coords[c] / (h, k)
        i, j = int(math.floor(x)), int(math.floor(y))
        fx = x - i
        kind = "r" if abs(fx - 2.0 / 3.0) < 1e-6 else "l"
        assert abs(fx - (2.0 / 3.0 if kind == "r" else 1.0 / 3.0)) < 1e-6
        out.append((kind, i, j))
    return out


def _prims(U):
    rho = U[:, 0]
    u, v = U[:, 1] / rho, U[:, 2] / rho
    e = U[:, 3] / rho - 0.5 * (u * u + v * v)
    p = (GAMMA - 1.0) * rho * e
    return rho, u, v, p


def check_dt_closed_form(orc):
    """Spectral radius per cell and dt on unjittered cartesian_tri and
    cartesian meshes with a cell-wise random state. Per face, with n the unit
    normal: u_f, c_f the means of the two cells' velocity and sound speed
    (the cell itself across a boundary face), dx_n = |s . n| where s joins
    the centroids, or is (2 sqrt V, 2 sqrt V) on a boundary face (the
    reference's hack, solver.cpp:662-667); then
    lambda = 1.37 (3/n_f) sum |u_f . n| / dx_n + 1.37 sqrt((3/n_f) sum (c_f/dx_n)^2)
    and dt = cfl / max lambda. For triangles of an h x k quad split along its
    rising diagonal dx_n is 2h/3 (vertical faces), 2k/3 (horizontal) and
    2hk / (3 sqrt(h^2 + k^2)) (diagonal); for quads h and k."""
    rng = np.random.default_rng(7)
    for kind, nx, ny, Lx, Ly in (("cartesian_tri", 6, 4, 1.2, 0.5), ("cartesian", 5, 4, 1.0, 0.6)):
        m = orc.Mesh(kind, nx, ny, Lx, Ly, 0.0, 0)
        s = orc.Solver(m, orc.make_config(recon="FO"))
        nc = m.n_cells
        rho = rng.uniform(0.5, 1.5, nc)
        u, v = rng.uniform(-0.6, 0.6, nc), rng.uniform(-0.6, 0.6, nc)
        p = rng.uniform(0.4, 1.6, nc)
        U = np.stack([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)], 1)
        s.set_state(U)
        rho, u, v, p = _prims(s.state()[0])
        cs = np.sqrt(GAMMA * p / rho)
        h, k = Lx / nx, Ly / ny
        lam = np.zeros(nc)
        if kind == "cartesian_tri":
            cls = _cart_tri_classes(m, nx, ny, h, k)
            idx = {cl: c for c, cl in enumerate(cls)}
            L = math.hypot(h, k)
            bnd = 2.0 * math.sqrt(h * k / 2.0)
            for c, (t, i, j) in enumerate(cls):
                # (neighbour or None for a boundary face, unit normal, interior dx_n)
                if t == "r":
                    faces = [(idx.get(("l", i, j - 1)), (0.0, 1.0), 2 * k / 3),
                             (idx.get(("l", i + 1, j)), (1.0, 0.0), 2 * h / 3),
                             (idx[("l", i, j)], (k / L, -h / L), 2 * h * k / (3 * L))]
                else:
                    faces = [(idx.get(("r", i - 1, j)), (1.0, 0.0), 2 * h / 3),
                             (idx.get(("r", i, j + 1)), (0.0, 1.0), 2 * k / 3),
                             (idx[("r", i, j)], (k / L, -h / L), 2 * h * k / (3 * L))]
                conv = acou = 0.0
                for nb, n, dxn in faces:
                    if nb is None:
                        nb, dxn = c, bnd * abs(n[0] + n[1])
                    uf, vf = 0.5 * (u[c] + u[nb]), 0.5 * (v[c] + v[nb])
                    cf = 0.5 * (cs[c] + cs[nb])
                    conv += abs(uf * n[0] + vf * n[1]) / dxn
                    acou += (cf / dxn) ** 2
                lam[c] = 1.37 * conv + 1.37 * math.sqrt(acou)
        else:
            ij = {}
            for c in range(nc):
                i, j = (int(math.floor(t)) for t in m.cell_coords[c] / (h, k))
                ij[(i, j)] = c
            bnd = 2.0 * math.sqrt(h * k)
            for (i, j), c in ij.items():
                faces = [(ij.get((i - 1, j)), (1.0, 0.0), h), (ij.get((i + 1, j)), (1.0, 0.0), h),
                         (ij.get((i, j - 1)), (0.0, 1.0), k), (ij.get((i, j + 1)), (0.0, 1.0), k)]
                conv = acou = 0.0
                for nb, n, dxn in faces:
                    if nb is None:
                        nb, dxn = c, bnd
                    uf, vf = 0.5 * (u[c] + u[nb]), 0.5 * (v[c] + v[nb])
                    cf = 0.5 * (cs[c] + cs[nb])
                    conv += abs(uf * n[0] + vf * n[1]) / dxn
                    acou += (cf / dxn) ** 2
                lam[c] = 1.37 * 0.75 * conv + 1.37 * math.sqrt(0.75 * acou)
        cfl = 0.37
        dt, cfl_local = s.calc_dt(cfl)
        assert math.isclose(dt, cfl / lam.max(), rel_tol=1e-13), (kind, dt, cfl / lam.max())
        np.testing.assert_allclose(cfl_local, dt * lam, rtol=1e-13, atol=0, err_msg=kind)


# ------------------------------------------------------------------ TENO ENO branch
def _leg(p):
    return np.polynomial.legendre.leg2poly([0] * p + [1])


def _psi(px, py, x, y):
    return np.polynomial.polynomial.polyval(x, _leg(px)) * np.polynomial.polynomial.polyval(y, _leg(py))


def check_eno_branch(orc):
    """A piecewise-constant state with a straight jump across the domain
    (all four conserved variables jump). A stencil entirely on the cell's
    side of the jump has b = 0, so a = 0, SI = 0 and raw weight 1e72; a
    stencil across it has a finite smoothness indicator.
      - Central stencil on one side: the central stencil is selected and
        every face value is the cell average, in both weightings.
      - Central stencil across the jump, some directional stencil on one
        side: the ENO branch runs (decision 0). Corrected weighting: the
        passing (one-sided) stencils reconstruct nothing, the failing ones
        and the central one are dropped, so every face value is exactly the
        cell average. Reference weighting (face_reconstruction.cpp:964-1031):
        the central stencil keeps its raw weight w_0 = (SI_0 + 1e-12)^-6 and is
        not normalised; the directional weights, 1/K for the passing ones and
        raw for the failing ones, are normalised over the directional
        stencils; each stencil s adds w_s sum_k a_sk (psi_k(x_q) + ipt_k / ta_s0)
        with a_s = A+_s b_s, b_s,i = ta_s,i (u_i - u_c)."""
    m = orc.Mesh("cartesian_tri", 10, 8, 1.0, 1.0, 0.2, 11)
    bcs = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
    UL = np.array([1.0, 0.2, -0.1, 1.0])
    UR = np.array([0.4, -0.3, 0.25, 0.3])

    def cons(w):
        r, uu, vv, pp = w
        return np.array([r, r * uu, r * vv, pp / (GAMMA - 1.0) + 0.5 * r * (uu * uu + vv * vv)])

    cc = m.cell_coords
    left = cc[:, 0] < 0.43 + 0.1 * cc[:, 1]
    U = np.where(left[:, None], cons(UL)[None, :], cons(UR)[None, :])
    out = {}
    for var in ("corrected", "reference"):
        s = orc.Solver(m, orc.make_config(order=2, teno_variant=var, bcs=bcs))
        s.calc_rhs(U)
        out[var] = (s.face_values(1), s.decisions(), s.teno() if var == "reference" else None)
    T = out["reference"][2]
    K, M = T["n_dof"], T["max_cells"]
    poly = T["poly_indices"].reshape(K, 2)
    ipt = T["integral_psi_target"]
    OI = T["oscillation_indicator"].reshape(K, K)
    og, os_, st, orm = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"], T["offsets_recon"]
    ta = T["transformed_areas"]
    nof = m.nodes_of_face.reshape(-1, 2)
    n_central = n_eno = 0
    worst = 0.0
    for c in range(m.n_cells):
        groups = []
        for sl in range(og[c + 1] - og[c]):
            a, b = os_[og[c] + sl], os_[og[c] + sl + 1]
            groups.append((og[c] + sl, st[a:b], a))
        one_sided = [len(ids) > 0 and bool(np.all(left[ids] == left[c])) for _, ids, _ in groups]
        faces = m.faces_of_cell[3 * c:3 * c + 3]
        sides = [0 if m.cells_of_face[f, 0] == c else 1 for f in faces]
        if one_sided[0]:
            n_central += 1
            for key in ("corrected", "reference"):
                fv, dec = out[key][0], out[key][1]
                assert dec[c].all(), (key, c)
                for f, sd in zip(faces, sides):
                    assert np.array_equal(fv[f, 0, sd], U[c]), (key, c)
            continue
        if not any(one_sided[1:]):
            continue
        n_eno += 1
        # corrected: exactly the cell average, ENO decision
        fv, dec = out["corrected"][0], out["corrected"][1]
        assert not dec[c].any(), c
        for f, sd in zip(faces, sides):
            assert np.array_equal(fv[f, 0, sd], U[c]), ("corrected", c, fv[f, 0, sd], U[c])
        # reference: raw central weight plus the normalised directional ones
        w0 = m.node_coords[m.nodes_of_cell[3 * c:3 * c + 3]]
        J = np.array([[w0[1][0] - w0[0][0], w0[2][0] - w0[0][0]], [w0[1][1] - w0[0][1], w0[2][1] - w0[0][1]]])
        Ji = np.linalg.inv(J)
        coef, wraw = [], []
        for si, ids, a in groups:
            if len(ids) == 0:
                coef.append(None)
                wraw.append(np.zeros(4))
                continue
            A = T["recon"][orm[si]:orm[si] + K * M].reshape(K, M)
            bm = ta[a:a + M, None] * (U[ids] - U[c][None, :])
            av = A @ bm  # K x 4
            si_ = np.einsum("kv,kj,jv->v", av, OI, av)
            coef.append(av)
            wraw.append(1.0 / (si_ + 1e-12) ** 6)
        wraw = np.array(wraw)  # G x 4
        dec = out["reference"][1]
        assert not dec[c].any(), c
        wt = wraw.copy()
        for vv in range(4):
            sd = wraw[1:, vv].sum()
            for g in range(1, len(groups)):
                if wraw[g, vv] / sd > 1e-5:
                    wt[g, vv] = 1.0 / K
            wt[1:, vv] /= wt[1:, vv].sum()
        fv = out["reference"][0]
        for f, sd in zip(faces, sides):
            mid = 0.5 * (m.node_coords[nof[f, 0]] + m.node_coords[nof[f, 1]])
            xq = Ji @ (mid - w0[0])
            psi = np.array([_psi(px, py, xq[0], xq[1]) for px, py in poly])
            exp = U[c].copy()
            mag = np.zeros(4)  # sum of the terms' magnitudes
            for g, (si, ids, a) in enumerate(groups):
                if coef[g] is None:
                    continue
                exp += wt[g] * ((psi + ipt / ta[a]) @ coef[g])
                mag += wt[g] * (np.abs(psi + ipt / ta[a]) @ np.abs(coef[g]))
            got = fv[f, 0, sd]
            # (SI + 1e-12)^-6 turns rounding in the smoothness indicators into
            # relative errors up to ~1e-9 of the terms
            worst = max(worst, float((np.abs(got - exp) / mag).max()))
            # the kept raw weights move the face value away from the average
            assert np.all(np.abs(got - U[c]) > 1e-5 * mag), (c, got, U[c])
    assert n_central > 20 and n_eno > 10, (n_central, n_eno)
    assert worst < 1e-8, worst


# ------------------------------------------------------------------ directional stencils
def directional_stencils(m, c, M):
    """Type-4 directional stencils restated from face_reconstruction.cpp:288-404:
    per face f of c, members are added ring by ring while the centroid offset
    d lies in the cone of the sub-triangle (centroid, face nodes), i.e.
    d = alpha e1 + beta e2 with alpha, beta >= 0 for the edges e1, e2 from
    the centroid to the face's nodes (Cramer's rule here); a stencil of a
    boundary face is empty; one that stops short of M members is filled
    outward, ring by ring from its own members. Returns the stencils and the
    smallest |alpha|, |beta| met, so near-ties can be recognised."""
    f0 = m.offsets_faces_of_cell[c]
    faces = m.faces_of_cell[f0:m.offsets_faces_of_cell[c + 1]]
    nof = m.nodes_of_face.reshape(-1, 2)
    cones = []
    for f in faces:
        e1 = m.node_coords[nof[f, 0]] - m.cell_coords[c]
        e2 = m.node_coords[nof[f, 1]] - m.cell_coords[c]
        cones.append((e1, e2, e1[0] * e2[1] - e2[0] * e1[1]))
    interior = [m.cells_of_face[f, 1] != -1 for f in faces]
    S = [[c] for _ in faces]
    rings = [[c]]
    tie = math.inf
    while True:
        rings.append(next_ring(m, rings, c))
        grew = [False] * len(faces)
        for s, (e1, e2, det) in enumerate(cones):
            for x in rings[-1]:
                if len(S[s]) == M:
                    break
                d = m.cell_coords[x] - m.cell_coords[c]
                al = (d[0] * e2[1] - e2[0] * d[1]) / det
                be = (e1[0] * d[1] - d[0] * e1[1]) / det
                tie = min(tie, abs(al), abs(be))
                if al >= 0.0 and be >= 0.0:
                    S[s].append(x)
                    grew[s] = True
        if not any(interior[s] and len(S[s]) < M and grew[s] for s in range(len(faces))):
            break
    for s in range(len(faces)):
        if not interior[s]:
            S[s] = []
            continue
        r2 = [list(S[s])]
        while len(S[s]) < M:
            r2.append(next_ring(m, r2, c))
            for x in r2[-1]:
                if len(S[s]) == M:
                    break
                S[s].append(x)
    return S, tie


def check_directional_stencils(orc):
    """The oracle's directional stencils equal the restatement on jittered
    6 x 5 and 8 x 8 cartesian_tri meshes (p = 2, 12 members), and every
    member added before the outward fill lies in its face's cone. Cells where
    a cone coordinate is within 1e-12 of zero (a tie the two evaluation
    orders might split differently) are skipped; there must be few."""
    checked = skipped = filled = 0
    for nx, ny, seed in ((6, 5, 17), (8, 8, 5)):
        m = orc.Mesh("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, seed)
        s = orc.Solver(m, orc.make_config(order=2))
        T = s.teno()
        M = T["max_cells"]
        og, os_, st = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"]
        for c in range(m.n_cells):
            exp, tie = directional_stencils(m, c, M)
            if tie < 1e-12:
                skipped += 1
                continue
            for sl in range(1, og[c + 1] - og[c]):
                a, b = os_[og[c] + sl], os_[og[c] + sl + 1]
                assert list(st[a:b]) == exp[sl - 1], (nx, ny, c, sl, list(st[a:b]), exp[sl - 1])
            checked += 1
            # stencils that needed the outward fill (near the boundary)
            nof = m.nodes_of_face.reshape(-1, 2)
            for sl, f in enumerate(m.faces_of_cell[3 * c:3 * c + 3]):
                ids = exp[sl]
                if not ids:
                    continue
                e1 = m.node_coords[nof[f, 0]] - m.cell_coords[c]
                e2 = m.node_coords[nof[f, 1]] - m.cell_coords[c]
                inside = [bool(np.all(np.linalg.solve(np.column_stack([e1, e2]), m.cell_coords[x] - m.cell_coords[c])
                                      >= 0)) for x in ids[1:]]
                filled += not all(inside)
    assert checked > 0.9 * (checked + skipped) and filled > 0, (checked, skipped, filled)


# ------------------------------------------------------------------ boundaries
def _uniform(rho, u, v, p, n):
    return np.tile([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)], (n, 1))


def _boundary_faces(m, zones):
    out = {}
    for z in zones:
        for f in m.zones[z]:
            out.setdefault(int(m.cells_of_face[f, 0]), []).append(int(f))
    return out


def _unit(m, f):
    n = m.face_normals[f]
    return n / math.hypot(n[0], n[1])


def check_boundary_fluxes(orc):
    """Residuals with closed forms on a jittered triangle mesh, first order,
    uniform state (interior faces then carry the physical flux, which sums
    to zero around a closed cell):
      - adiabatic slip walls on all four sides pass only the pressure,
        (0, p n, 0): a cell's residual is (1/V) sum over its wall faces of
        u_n (rho, rho u, rho v, rho E + p) A;
      - fixed (u, p, T) equal to the interior state: the ghost state is the
        cell's, so the residual vanishes;
      - p_out with a back pressure unequal to the cell's: a supersonic cell
        extrapolates its pressure (residual zero); a subsonic one takes the
        back pressure with extrapolated T and velocity, rho = p_out / (R T),
        so its residual is (1/V) sum over its p_out faces of
        (F(U) - F_HLLC(U, ghost)) A."""
    zones = ("left", "right", "top", "bottom")
    m = orc.Mesh("cartesian_tri", 7, 6, 1.0, 0.8, 0.2, 9)
    nc = m.n_cells
    V = m.cell_volume
    bf = _boundary_faces(m, zones)
    rho, u, v, p = 1.1, 0.35, -0.2, 0.8

    # walls
    s = orc.Solver(m, orc.make_config(recon="FO", bcs=[(z, "wall_adiabatic", {}) for z in zones]))
    U = _uniform(rho, u, v, p, nc)
    r = s.calc_rhs(U)
    exp = np.zeros((nc, 4))
    E = U[0, 3]
    for c, fl in bf.items():
        for f in fl:
            n = _unit(m, f)
            un = u * n[0] + v * n[1]
            exp[c] += un * np.array([rho, rho * u, rho * v, E + p]) * m.face_area[f] / V[c]
    scale = np.abs(exp).max()
    assert scale > 0.1
    np.testing.assert_allclose(r, exp, rtol=0, atol=1e-12 * scale)

    # fixed (u, p, T) at the interior state
    T = p / (rho * R_GAS)
    s = orc.Solver(m, orc.make_config(recon="FO", bcs=[(z, "upt", {"u": (u, v), "p": p, "T": T}) for z in zones]))
    r = s.calc_rhs(U)
    assert np.abs(r).max() <= 1e-11 * np.abs(U).max() / V.min(), np.abs(r).max()

    # p_out: supersonic (|u| > c) and subsonic cells
    c0 = math.sqrt(GAMMA * p / rho)
    pb = 0.55
    for uu, vv, sup in ((1.3 * c0, 0.4 * c0, True), (0.3 * c0, -0.2 * c0, False)):
        s = orc.Solver(m, orc.make_config(recon="FO", bcs=[(z, "p_out", {"p": pb}) for z in zones]))
        U = _uniform(rho, uu, vv, p, nc)
        r = s.calc_rhs(U)
        if sup:
            assert np.abs(r).max() <= 1e-11 * np.abs(U).max() / V.min(), np.abs(r).max()
            continue
        Tl = p / (rho * R_GAS)
        rg = pb / (R_GAS * Tl)
        hg = CV * Tl + pb / rg
        hl = CV * Tl + p / rho
        exp = np.zeros((nc, 4))
        for c, fl in bf.items():
            for f in fl:
                n = _unit(m, f)
                un = uu * n[0] + vv * n[1]
                Fphys = np.array([rho * un, rho * uu * un + p * n[0], rho * vv * un + p * n[1], (U[0, 3] + p) * un])
                Frs = orc.riemann_flux("HLLC", n, rho, [uu, vv], p, GAMMA, hl, rg, [uu, vv], pb, GAMMA, hg)
                exp[c] += (Fphys - Frs) * m.face_area[f] / V[c]
        scale = np.abs(exp).max()
        assert scale > 0.01
        np.testing.assert_allclose(r, exp, rtol=0, atol=1e-11 * scale)


# ------------------------------------------------------------------ HLL
def toro_hll(WL, WR, n, g=GAMMA):
    """Toro's HLL flux (eq. 10.21) with the wave speeds of eqs. 10.59-10.60
    and the adaptive noniterative pressure estimate of section 9.5.3 (PVRS,
    switching to two rarefactions below p_min and two shocks otherwise,
    Q_user = 2), for primitive states (rho, u, v, p) and unit normal n."""
    (rl, ul, vl, pl), (rr, ur, vr, pr) = WL, WR
    unl, unr = ul * n[0] + vl * n[1], ur * n[0] + vr * n[1]
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    ppv = 0.5 * (pl + pr) - 0.125 * (unr - unl) * (rl + rr) * (al + ar)
    pmin, pmax = min(pl, pr), max(pl, pr)
    if pmax / pmin < 2.0 and pmin <= ppv <= pmax:
        ps = ppv
    elif ppv < pmin:
        z = (g - 1.0) / (2.0 * g)
        ps = ((al + ar - 0.5 * (g - 1.0) * (unr - unl)) / (al / pl ** z + ar / pr ** z)) ** (1.0 / z)
    else:
        p0 = max(0.0, ppv)
        gl = math.sqrt(2.0 / ((g + 1.0) * rl) / (p0 + (g - 1.0) / (g + 1.0) * pl))
        gr = math.sqrt(2.0 / ((g + 1.0) * rr) / (p0 + (g - 1.0) / (g + 1.0) * pr))
        ps = (gl * pl + gr * pr - (unr - unl)) / (gl + gr)

    def q(pk):
        return 1.0 if ps <= pk else math.sqrt(1.0 + (g + 1.0) / (2.0 * g) * (ps / pk - 1.0))

    SL, SR = unl - al * q(pl), unr + ar * q(pr)

    def UF(r, u, v, p, un):
        E = p / (g - 1.0) + 0.5 * r * (u * u + v * v)
        return (np.array([r, r * u, r * v, E]),
                np.array([r * un, r * u * un + p * n[0], r * v * un + p * n[1], (E + p) * un]))

    Ul, Fl = UF(rl, ul, vl, pl, unl)
    Ur, Fr = UF(rr, ur, vr, pr, unr)
    if SL >= 0.0:
        return Fl
    if SR <= 0.0:
        return Fr
    return (SR * Fl - SL * Fr + SL * SR * (Ur - Ul)) / (SR - SL)


def check_hll_toro(orc):
    """The oracle's HLL flux equals Toro's on Sod (two-shock estimate), a
    symmetric expansion (two-rarefaction estimate), a weak wave (PVRS) and a
    supersonic pair (upwind flux), along the axes and a diagonal."""
    cases = [((1.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1)),
             ((1.0, -1.2, 0.3, 0.4), (0.9, 1.1, -0.2, 0.45)),
             ((1.0, 0.1, -0.05, 1.0), (0.9, 0.0, 0.02, 0.8)),
             ((1.0, 3.0, 0.5, 1.0), (0.5, 2.9, 0.4, 0.7))]
    r2 = 1.0 / math.sqrt(2.0)
    for WL, WR in cases:
        for n in ((1.0, 0.0), (0.0, 1.0), (r2, r2), (-0.6, 0.8)):
            hl = WL[3] / WL[0] * GAMMA / (GAMMA - 1.0)
            hr = WR[3] / WR[0] * GAMMA / (GAMMA - 1.0)
            got = orc.riemann_flux("HLL", n, WL[0], WL[1:3], WL[3], GAMMA, hl, WR[0], WR[1:3], WR[3], GAMMA, hr)
            exp = toro_hll(WL, WR, n)
            np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13 * max(1.0, np.abs(exp).max()),
                                       err_msg=str((WL, WR, n)))


def _euler_flux(U, n):
    rho, m0, m1, E = U
    u, v = m0 / rho, m1 / rho
    p = (GAMMA - 1.0) * (E - 0.5 * rho * (u * u + v * v))
    un = u * n[0] + v * n[1]
    return np.array([rho * un, m0 * un + p * n[0], m1 * un + p * n[1], (E + p) * un]), u, v, p


def check_positivity_limiter(orc):
    """Blend factor theta of one bracket U + s (F_lo + theta (F - F_lo)):
    1 when the high-order bracket is admissible; at a density violation the
    root of the linear density, rho = 1e-13; at a pressure violation the
    largest root of p = 1e-13, a quadratic in theta, approached from below by
    the bisection; 0 when the low-order bracket is not admissible or the flux
    is not finite. Low-order flux: the Euler flux for equal states and for an
    extrapolation boundary, and at symmetry and slip walls the Lax-Friedrichs
    flux against the mirror state, (0, (rho u_n^2 + p + alpha rho u_n) n, 0)
    with alpha = |u_n| + c."""
    bcs = [("left", "symmetry", {}), ("right", "wall_adiabatic", {}), ("top", "extrapolation", {}),
           ("bottom", "extrapolation", {})]
    m = orc.Mesh("cartesian_tri", 3, 3, 1.0, 1.0, 0.0, 0)
    s = orc.Solver(m, orc.make_config(order=2, teno_variant="corrected", bcs=bcs, positivity=True), 1)
    U = np.array([1.0, 0.0, 0.0, 2.5])  # rho 1, at rest, p 1
    zero = np.zeros(4)
    assert s.pp_theta(U, -1.0, np.array([0.1, 0.02, -0.03, 0.2]), zero) == 1.0
    assert s.pp_theta(U, -1.0, np.array([0.1, 0.02, -0.03, 0.2]), zero, 0.3) == 0.3
    # density: B(theta) = (1 - 2 theta, 0, 0, 2.5)
    th = s.pp_theta(U, -1.0, np.array([2.0, 0.0, 0.0, 0.0]), zero)
    assert abs(th - (1.0 - 1e-13) / 2.0) <= 1e-16, th
    # pressure, linear: p(B(theta)) = 0.4 (2.5 - 5 theta)
    th = s.pp_theta(U, -1.0, np.array([0.0, 0.0, 0.0, 5.0]), zero)
    root = (2.5 - 1e-13 / 0.4) / 5.0
    assert 0.0 <= root - th <= 2e-12, (th, root)
    # pressure, quadratic: B(theta) = (1, -3 theta, 0, 2.5), p = 0.4 (2.5 - 4.5 theta^2)
    th = s.pp_theta(U, -1.0, np.array([0.0, 3.0, 0.0, 0.0]), zero)
    root = math.sqrt((2.5 - 1e-13 / 0.4) / 4.5)
    assert 0.0 <= root - th <= 2e-12, (th, root)
    # a mass flux that leaves rho > 0 moves the pressure root: p rho is quadratic in theta
    F = np.array([0.9, 3.0, 0.0, 0.0])
    th = s.pp_theta(U, -1.0, F, zero)
    r_rho = (1.0 - 1e-13) / 0.9
    # p(theta) rho(theta) / 0.4 = 2.5 (1 - 0.9 theta) - 4.5 theta^2 = 1e-13 (1 - 0.9 theta) / 0.4
    a, b, c = -4.5, -2.25 + 0.9e-13 / 0.4, 2.5 - 1e-13 / 0.4
    r_p = (-b - math.sqrt(b * b - 4 * a * c)) / (2 * a)
    assert r_p < r_rho and 0.0 <= r_p - th <= 2e-12, (th, r_p)
    assert s.pp_theta(U, -1.0, F, np.array([2.0, 0.0, 0.0, 0.0])) == 0.0  # low-order bracket rho < 0
    assert s.pp_theta(U, -1.0, np.array([0.0, np.nan, 0.0, 0.0]), zero) == 0.0
    assert s.pp_theta(U, -1.0, np.array([0.0, np.inf, 0.0, 0.0]), zero) == 0.0
    # low-order fluxes
    rng = np.random.default_rng(5)
    for _ in range(20):
        rho, u, v, p = rng.uniform(0.2, 2.0), rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(0.1, 2.0)
        Um = np.array([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)])
        ang = rng.uniform(0, 2 * math.pi)
        n = np.array([math.cos(ang), math.sin(ang)])
        Fe, _, _, _ = _euler_flux(Um, n)
        np.testing.assert_allclose(s.pp_low_flux(-1, n, Um, Um), Fe, rtol=0, atol=1e-13 * np.abs(Fe).max())
        r2, u2, v2, p2 = rng.uniform(0.2, 2.0), rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(0.1, 2.0)
        U2 = np.array([r2, r2 * u2, r2 * v2, p2 / (GAMMA - 1.0) + 0.5 * r2 * (u2 * u2 + v2 * v2)])
        F2, _, _, _ = _euler_flux(U2, n)
        a2 = max(abs(u * n[0] + v * n[1]) + math.sqrt(GAMMA * p / rho),
                 abs(u2 * n[0] + v2 * n[1]) + math.sqrt(GAMMA * p2 / r2))
        lf = 0.5 * (Fe + F2) - 0.5 * a2 * (U2 - Um)
        np.testing.assert_allclose(s.pp_low_flux(-1, n, Um, U2), lf, rtol=0, atol=1e-13 * np.abs(lf).max())
        np.testing.assert_allclose(s.pp_low_flux(2, n, Um), Fe, rtol=0, atol=1e-13 * np.abs(Fe).max())
        un = u * n[0] + v * n[1]
        alpha = abs(un) + math.sqrt(GAMMA * p / rho)
        mom = rho * un * un + p + alpha * rho * un
        exp = np.array([0.0, mom * n[0], mom * n[1], 0.0])
        for b in (0, 1):
            got = s.pp_low_flux(b, n, Um)
            np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13 * max(1.0, abs(mom)), err_msg=str((b, got, exp)))


def check_roe_shocks(orc):
    """Roe's flux (no reference counterpart; Roe 1981) at lone normal shocks
    built from the Rankine-Hugoniot relations: the jump is an eigenvector of
    the Roe matrix with the shock speed as eigenvalue, so the flux is the
    upwind side's physical flux. Shocks of Mach 1.5, 2 and 4, both
    directions, at rest and moving, three normals, tangential velocity."""
    from shocks import cases, enthalpy
    for n, L, R, want in cases(GAMMA):
        F = orc.riemann_flux("Roe", list(n), L[0], list(L[1]), L[2], GAMMA, enthalpy(GAMMA, L[0], L[2]), R[0],
                             list(R[1]), R[2], GAMMA, enthalpy(GAMMA, R[0], R[2]))
        np.testing.assert_allclose(F, want, rtol=1e-12, atol=1e-12 * np.abs(want).max(), err_msg=str((n, L, R)))


CHECKS = {
    "dt": check_dt_closed_form,
    "eno": check_eno_branch,
    "directional": check_directional_stencils,
    "boundary": check_boundary_fluxes,
    "hll": check_hll_toro,
    "positivity": check_positivity_limiter,
    "roe": check_roe_shocks,
}
