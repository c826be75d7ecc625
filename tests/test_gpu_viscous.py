"""Navier-Stokes extension on the GPU (no reference counterpart: the
reference solves Euler). The viscous part of the residual, R_NS - R_Euler on
the same state, is checked against the numpy restatement
(oracle/viscous_ref.py) and against closed forms the scheme reproduces exactly:
no viscous term for uniform flow (bit-exact), zero viscous force and heating
mu a^2 for the linear shear u = a y, zero heat-flux divergence for a linear
temperature, wall drag opposing the flow at a no-slip wall, and a smaller dt."""
import numpy as np
import pytest

from oracle import viscous_ref as vr

pytestmark = pytest.mark.gpu

GAMMA = 1.4
EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
# nondimensional physics: R = p_ref / (T_ref rho_ref) = 1
PHYS = dict(gamma=GAMMA, p_ref=1.0, T_ref=1.0, rho_ref=1.0, p_min=-1e20, p_max=1e20)


def _pair(mlr, n, mu=0.05, S=-1.0, bcs=EXTRAP, seed=4, recon="FO"):
    gen = mlr.mesh_gen("cartesian_tri", n, n, 1.0, 1.0, 0.2, seed)
    kw = dict(recon=recon, order=2, teno_variant="corrected", riemann="HLLC", cfl=0.4, bcs=bcs, **PHYS)
    ns = mlr.Solver(gen=gen, num=mlr.numerics(physics="navier-stokes", mu_ref=mu, T_mu_ref=1.0, sutherland_S=S,
                                              prandtl=0.72, **kw))
    eu = mlr.Solver(gen=gen, num=mlr.numerics(**kw))
    return ns, eu, mlr.Layout(gen, mlr.numerics(recon="FO", bcs=bcs)).mesh_arrays()


def _state(cc, rho, u, v, p):
    U = np.zeros((len(cc), 4))
    U[:, 0] = rho
    U[:, 1] = rho * u
    U[:, 2] = rho * v
    U[:, 3] = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def _dR(ns, eu, U):
    ns.set_state(U)
    eu.set_state(U)
    return ns.calc_rhs() - eu.calc_rhs(), eu.calc_rhs()


def _deep_interior(mesh, layers=2):
    """Cells whose faces, and their neighbours' faces, are all interior."""
    cof = mesh["cells_of_face"]
    foc = mesh["faces_of_cell"].reshape(-1, 3)
    N = len(foc)
    ok = (cof[foc, 1] >= 0).all(axis=1)
    for _ in range(layers - 1):
        nb = np.where(cof[foc, 0] == np.arange(N)[:, None], cof[foc, 1], cof[foc, 0])
        ok = ok & np.where(nb >= 0, ok[np.maximum(nb, 0)], False).all(axis=1)
    return ok


@pytest.mark.parametrize("S", [-1.0, 0.4])
def test_viscous_residual_matches_restatement(mlr, S):
    ns, eu, mesh = _pair(mlr, 24, mu=0.05, S=S)
    cc = mesh["cell_coords"]
    x, y = cc[:, 0], cc[:, 1]
    U = _state(cc, 1.0 + 0.2 * np.sin(3 * x) * np.cos(2 * y), 0.5 + 0.3 * np.sin(4 * y + x),
               -0.2 + 0.25 * np.cos(3 * x - 2 * y), 1.0 + 0.1 * np.cos(5 * x * y))
    dR, R = _dR(ns, eu, U)
    ref = vr.viscous_residual(mesh, U, GAMMA, 1.0, 0.05, 1.0, S, 0.72)
    tol = 1e-10 * np.abs(ref).max() + 1e-13 * np.abs(R).max()
    assert np.abs(dR - ref).max() <= tol, np.abs(dR - ref).max() / np.abs(ref).max()


def test_uniform_flow_has_no_viscous_terms(mlr):
    ns, eu, mesh = _pair(mlr, 20)
    n = len(mesh["cell_coords"])
    U = _state(mesh["cell_coords"], np.full(n, 1.2), np.full(n, 0.7), np.full(n, -0.3), np.full(n, 0.9))
    ns.set_state(U)
    eu.set_state(U)
    assert np.array_equal(ns.calc_rhs(), eu.calc_rhs())


def test_linear_shear_closed_form(mlr):
    mu, a = 0.05, 0.8
    ns, eu, mesh = _pair(mlr, 28, mu=mu)
    cc = mesh["cell_coords"]
    n = len(cc)
    U = _state(cc, np.full(n, 1.0), 0.3 + a * cc[:, 1], np.zeros(n), np.full(n, 1.0))
    dR, R = _dR(ns, eu, U)
    inner = _deep_interior(mesh)
    assert inner.sum() > 500
    h = 1.0 / 28
    scale = mu * a / h
    # no net viscous force: the stress is uniform
    assert np.abs(dR[inner, 1:3]).max() <= 1e-9 * scale + 1e-13 * np.abs(R).max()
    # viscous heating d/dy (u tau_xy) = mu a^2, exact for a linear u
    assert np.abs(dR[inner, 3] - mu * a * a).max() <= 1e-9 * scale + 1e-13 * np.abs(R).max()
    assert np.all(dR[:, 0] == 0.0)


def test_linear_temperature_has_no_heat_flux_divergence(mlr):
    mu = 0.05
    ns, eu, mesh = _pair(mlr, 28, mu=mu)
    cc = mesh["cell_coords"]
    n = len(cc)
    T = 1.0 + 0.5 * cc[:, 0] - 0.3 * cc[:, 1]
    p = np.full(n, 1.0)
    U = _state(cc, p / T, np.zeros(n), np.zeros(n), p)
    dR, R = _dR(ns, eu, U)
    inner = _deep_interior(mesh)
    k = mu * GAMMA / (GAMMA - 1.0) / 0.72
    scale = k * 0.6 * 28
    assert np.abs(dR[inner]).max() <= 1e-9 * scale + 1e-13 * np.abs(R).max()


def test_noslip_wall_drags_the_flow(mlr):
    bcs = [("left", "extrapolation", {}), ("right", "extrapolation", {}), ("top", "extrapolation", {}),
           ("bottom", "wall_noslip", {})]
    ns, eu, mesh = _pair(mlr, 24, mu=0.02, bcs=bcs)
    cc = mesh["cell_coords"]
    n = len(cc)
    U = _state(cc, np.full(n, 1.0), np.full(n, 0.8), np.zeros(n), np.full(n, 1.0))
    dR, _ = _dR(ns, eu, U)
    cof = mesh["cells_of_face"]
    foc = mesh["faces_of_cell"].reshape(-1, 3)
    fy = mesh["node_coords"][mesh["nodes_of_face"].reshape(-1, 2)][:, :, 1].max(axis=1)
    wall = ((cof[foc, 1] < 0) & (fy[foc] < 1e-12)).any(axis=1)
    assert wall.sum() >= 24
    N = len(foc)
    nb = np.where(cof[foc, 0] == np.arange(N)[:, None], cof[foc, 1], cof[foc, 0])
    near = wall | np.where(nb >= 0, wall[np.maximum(nb, 0)], False).any(axis=1)
    vol = mesh["cell_volume"]
    assert (dR[wall, 1] * vol[wall]).sum() < 0.0       # x-momentum lost to the wall
    assert (dR[wall, 1] < 0.0).mean() > 0.9
    assert np.all(dR[~near] == 0.0)                    # uniform flow away from it


def test_viscous_time_step_is_smaller(mlr):
    ns, eu, mesh = _pair(mlr, 20, mu=0.02)
    for s in (ns, eu):
        s.init_constant([0.5, 0.1], 1.0, 1.0)
    assert ns.calc_dt() < eu.calc_dt()


def test_viscous_runs_steps_with_teno(mlr):
    bcs = [("left", "upt", {"u": (2.5, 0.0), "p": 1.0, "T": 1.0 / GAMMA}), ("right", "extrapolation", {}),
           ("top", "extrapolation", {}), ("bottom", "wall_noslip", {})]
    ns, _, mesh = _pair(mlr, 32, mu=2.5 * 1.4 / 1000.0, S=110.4 / 288.15 / GAMMA, bcs=bcs, recon="TENO")
    ns.init_constant([2.5, 0.0], 1.0, 1.0 / GAMMA)
    ns.run_steps(20)
    U, W = ns.state()
    assert np.isfinite(U).all() and (U[:, 0] > 0).all() and (W[:, 2] > 0).all()
    # a boundary layer: the flow slows next to the wall
    cc = mesh["cell_coords"]
    u = U[:, 1] / U[:, 0]
    low = cc[:, 1] < 1.0 / 32
    mid = np.abs(cc[:, 1] - 0.5) < 0.05
    assert u[low].mean() < u[mid].mean()


def test_delaunay_wedge_viscous_roe(mlr):
    """The viscous wedge problem on a small graded Delaunay mesh: TENO p=2 +
    Roe + viscous terms, supersonic inflow, no-slip wedge surface."""
    from mallard_amd import meshgen
    mesh = meshgen.wedge_mesh(n_cells=30000, seed=2)
    Tinf = 1.0 / GAMMA
    bcs = [("left", "upt", {"u": (2.5, 0.0), "p": 1.0, "T": Tinf}), ("right", "extrapolation", {}),
           ("top", "extrapolation", {}), ("symmetry", "symmetry", {}), ("wall", "wall_noslip", {})]
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", riemann="Roe", cfl=0.4, bcs=bcs,
                       physics="navier-stokes", mu_ref=1.4 * 2.5 / 1e4, T_mu_ref=Tinf,
                       sutherland_S=110.4 / 288.15 * Tinf, prandtl=0.72, **PHYS)
    s = mlr.Solver(mesh=mesh, num=num)
    s.init_constant([2.5, 0.0], 1.0, Tinf)
    s.run_steps(40)
    U, W = s.state()
    assert np.isfinite(U).all() and (U[:, 0] > 0).all() and (W[:, 2] > 0).all()
    t, steps = s.time()
    assert steps == 40 and t > 0
    # the wedge compresses the flow: pressure rises in the cells on the wedge
    # surface, and the no-slip wall slows them down
    names, off, zf = mesh["zone_names"], mesh["zone_offsets"], mesh["zone_faces"]
    k = names.index("wall")
    wall_cells = np.unique(mesh["cells_of_face"].reshape(-1, 2)[zf[off[k]:off[k + 1]], 0])
    print("wall cells: p mean %.4f, u mean %.4f" % (W[wall_cells, 2].mean(), W[wall_cells, 0].mean()))
    assert W[wall_cells, 2].mean() > 1.05
    assert W[wall_cells, 0].mean() < 2.5
