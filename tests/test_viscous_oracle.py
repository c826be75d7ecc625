"""The numpy restatement of the viscous terms (oracle/viscous_ref.py), pinned
by closed forms on jittered triangle meshes (host only): no viscous terms for
uniform flow, zero viscous force and heating mu a^2 for the linear shear
u = a y, zero heat-flux divergence for a linear temperature with constant
viscosity, and the Sutherland law's reference value and growth."""
import numpy as np

from oracle import viscous_ref as vr

GAMMA = 1.4
EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]


def _mesh(mlr, n, seed=4):
    gen = mlr.mesh_gen("cartesian_tri", n, n, 1.0, 1.0, 0.2, seed)
    return mlr.Layout(gen, mlr.numerics(recon="FO", bcs=EXTRAP)).mesh_arrays()


def _state(rho, u, v, p):
    U = np.zeros((len(rho), 4))
    U[:, 0], U[:, 1], U[:, 2] = rho, rho * u, rho * v
    U[:, 3] = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def _deep_interior(mesh):
    cof = mesh["cells_of_face"]
    foc = mesh["faces_of_cell"].reshape(-1, 3)
    N = len(foc)
    ok = (cof[foc, 1] >= 0).all(axis=1)
    nb = np.where(cof[foc, 0] == np.arange(N)[:, None], cof[foc, 1], cof[foc, 0])
    return ok & np.where(nb >= 0, ok[np.maximum(nb, 0)], False).all(axis=1)


def test_uniform_flow(mlr):
    m = _mesh(mlr, 12)
    n = len(m["cell_coords"])
    U = _state(np.full(n, 1.1), np.full(n, 0.4), np.full(n, -0.2), np.full(n, 0.8))
    assert np.abs(vr.viscous_residual(m, U, GAMMA, 1.0, 0.1, 1.0, -1.0, 0.72)).max() == 0.0


def test_linear_shear(mlr):
    m = _mesh(mlr, 16)
    cc = m["cell_coords"]
    n = len(cc)
    mu, a = 0.1, 0.7
    U = _state(np.ones(n), 0.2 + a * cc[:, 1], np.zeros(n), np.ones(n))
    r = vr.viscous_residual(m, U, GAMMA, 1.0, mu, 1.0, -1.0, 0.72)
    inner = _deep_interior(m)
    assert inner.sum() > 100
    assert np.abs(r[inner, 1:3]).max() < 1e-12
    assert np.abs(r[inner, 3] - mu * a * a).max() < 1e-12
    assert np.all(r[:, 0] == 0.0)
    # a plausible mistake (stress without its transpose term) breaks it
    assert abs(r[~inner, 1]).max() > 1e-6  # boundary cells feel the one-sided data


def test_linear_temperature(mlr):
    m = _mesh(mlr, 16)
    cc = m["cell_coords"]
    n = len(cc)
    T = 1.0 + 0.4 * cc[:, 0] + 0.2 * cc[:, 1]
    U = _state(1.0 / T, np.zeros(n), np.zeros(n), np.ones(n))
    r = vr.viscous_residual(m, U, GAMMA, 1.0, 0.1, 1.0, -1.0, 0.72)
    assert np.abs(r[_deep_interior(m)]).max() < 1e-11


def test_sutherland():
    T = np.array([1.0, 2.0])
    mu = vr.sutherland(T, 0.3, 1.0, 0.4)
    assert mu[0] == 0.3
    assert abs(mu[1] - 0.3 * 2.0 ** 1.5 * 1.4 / 2.4) < 1e-15


def _zoned_mesh(oracle, n, seed=4):
    """A jittered triangle mesh with its zones (generator zones: right, top,
    left, bottom), as the dictionary viscous_residual takes."""
    m = oracle.Mesh("cartesian_tri", n, n, 1.0, 1.0, 0.2, seed)
    names = list(m.zones.keys())
    faces = [m.zones[z] for z in names]
    return dict(node_coords=m.node_coords, cell_coords=m.cell_coords, cell_volume=m.cell_volume,
                face_area=m.face_area, face_normals=m.face_normals, faces_of_cell=m.faces_of_cell,
                nodes_of_face=m.nodes_of_face, cells_of_face=m.cells_of_face, zone_names=names,
                zone_offsets=np.concatenate([[0], np.cumsum([len(f) for f in faces])]).astype(np.uint32),
                zone_faces=np.concatenate(faces).astype(np.uint32))


def test_noslip_wall_linear_shear(oracle):
    """u = a y over a no-slip bottom wall (u = 0 there): the wall holds the
    field's own value, so every cell away from the other sides, wall cells
    included, has zero viscous force and heating mu a^2."""
    m = _zoned_mesh(oracle, 20)
    cc = m["cell_coords"]
    n = len(cc)
    mu, a = 0.1, 0.7
    U = _state(np.ones(n), a * cc[:, 1], np.zeros(n), np.ones(n))
    bc = {"bottom": ("wall_noslip", {})}
    r = vr.viscous_residual(m, U, GAMMA, 1.0, mu, 1.0, -1.0, 0.72, bc=bc)
    sel = (cc[:, 0] > 0.2) & (cc[:, 0] < 0.8) & (cc[:, 1] < 0.6)
    wall = sel & (cc[:, 1] < 0.05)
    assert wall.sum() > 10
    assert np.abs(r[sel, 1:3]).max() < 1e-12
    assert np.abs(r[sel, 3] - mu * a * a).max() < 1e-12
    # with the wall taken as an extrapolation boundary the wall cells differ
    r2 = vr.viscous_residual(m, U, GAMMA, 1.0, mu, 1.0, -1.0, 0.72)
    assert np.abs(r2[wall, 3] - mu * a * a).max() > 1e-4


def test_fixed_state_boundary_linear_field(oracle):
    """u = b x with a fixed-state (u, p, T) left boundary holding u = 0 and
    the field's temperature: zero viscous force and heating (4/3) mu b^2 up
    to the boundary."""
    m = _zoned_mesh(oracle, 20)
    cc = m["cell_coords"]
    n = len(cc)
    mu, b, T0 = 0.1, 0.6, 1.0
    U = _state(np.ones(n), b * cc[:, 0], np.zeros(n), np.ones(n) * T0)
    bc = {"left": ("upt", {"u": (0.0, 0.0), "T": T0})}
    r = vr.viscous_residual(m, U, GAMMA, 1.0, mu, 1.0, -1.0, 0.72, bc=bc)
    sel = (cc[:, 1] > 0.2) & (cc[:, 1] < 0.8) & (cc[:, 0] < 0.6)
    edge = sel & (cc[:, 0] < 0.05)
    assert edge.sum() > 10
    assert np.abs(r[sel, 1:3]).max() < 1e-12
    assert np.abs(r[sel, 3] - 4.0 / 3.0 * mu * b * b).max() < 1e-12
    r2 = vr.viscous_residual(m, U, GAMMA, 1.0, mu, 1.0, -1.0, 0.72)
    assert np.abs(r2[edge, 3] - 4.0 / 3.0 * mu * b * b).max() > 1e-4


def test_slip_wall_carries_no_viscous_flux(oracle):
    """Symmetry and slip walls carry no viscous stress or heat flux: with
    them on all four sides, interior faces cancel in pairs and the
    volume-weighted residual of a sheared, heated field sums to zero; with
    extrapolation sides it does not."""
    m = _zoned_mesh(oracle, 12)
    cc, V = m["cell_coords"], m["cell_volume"]
    n = len(cc)
    U = _state(np.ones(n), 0.5 * cc[:, 1] ** 2, 0.3 * cc[:, 0], 1.0 + 0.2 * cc[:, 0])
    scale = np.abs(vr.viscous_residual(m, U, GAMMA, 1.0, 0.1, 1.0, -1.0, 0.72) * V[:, None]).sum()
    for kind in ("symmetry", "wall_adiabatic"):
        bc = {z: (kind, {}) for z in ("left", "right", "top", "bottom")}
        r = vr.viscous_residual(m, U, GAMMA, 1.0, 0.1, 1.0, -1.0, 0.72, bc=bc)
        assert np.abs((r * V[:, None]).sum(axis=0)).max() < 1e-14 * scale, kind
    r = vr.viscous_residual(m, U, GAMMA, 1.0, 0.1, 1.0, -1.0, 0.72)
    assert np.abs((r * V[:, None]).sum(axis=0)).max() > 1e-6 * scale
