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
