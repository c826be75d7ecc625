"""Derived TENO face basis on the GPU: the plain p = 2 pass recomputes the
face-point basis values in the target cell's frame from the cell's node
coordinates (tune[4] = 1) instead of reading the stored table. The
arithmetic is the precompute's, in its order, so face values, residuals and
trajectories equal the stored-table pass bit for bit, and (corrected
weighting) the oracle's, at every occupancy target, for both weightings, on
meshes whose cell count is not a multiple of the 64-cell block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RIEMANN2D = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                      [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])
EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
# (occupancy target, face-basis source: 0 stored, 1 derived)
SHAPES = [(0, 0), (0, 1), (4, 1), (5, 1)]


def _solver(mlr, nx, ny, seed, variant, bcs=None, positivity=False):
    num = mlr.numerics(recon="TENO", order=2, teno_variant=variant, cfl=0.4, bcs=bcs, positivity=positivity)
    return mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, seed), num=num)


def _random_state(n, seed):
    """Cell states mixing smooth regions, jumps and near-vacuum cells, so the
    cut-off takes the central, directional and ENO branches and the
    positivity scaling engages."""
    rng = np.random.default_rng(seed)
    rho = np.where(rng.random(n) < 0.5, 1.0, 0.125) * (1.0 + 0.05 * rng.standard_normal(n))
    rho[rng.random(n) < 0.02] = 1e-9
    u = rng.uniform(-1.0, 1.0, n)
    v = rng.uniform(-1.0, 1.0, n)
    p = np.where(rng.random(n) < 0.5, 1.0, 0.1) * (1.0 + 0.05 * rng.standard_normal(n))
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=1)


@pytest.mark.parametrize("variant", ["corrected", "reference"])
def test_face_values_equal_stored_basis(mlr, variant):
    out = []
    for occ, tables in SHAPES:
        s = _solver(mlr, 37, 29, 41, variant)
        s.set_tune(occ, 256, 256, 0, tables)
        assert s.tune()[0] == occ and s.tune()[4] == tables
        out.append(s.calc_face_values(_random_state(s.n_cells, 5)))
        s.close()
    assert np.isfinite(out[0]).any()
    for k in range(1, len(SHAPES)):
        assert np.array_equal(out[0], out[k], equal_nan=True), SHAPES[k]


def test_residual_equals_oracle(mlr, oracle):
    num_kw = dict(recon="TENO", order=2, teno_variant="corrected")
    m = oracle.Mesh("cartesian_tri", 21, 19, 1.0, 1.0, 0.2, 7)
    orc = oracle.Solver(m, oracle.make_config(**num_kw), 1)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    r_orc = orc.calc_rhs(orc.state()[0])
    for occ in (0, 4, 5):
        s = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 21, 19, 1.0, 1.0, 0.2, 7), num=mlr.numerics(**num_kw))
        s.set_tune(occ, 256, 256, 0, 1)
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        assert np.array_equal(s.calc_rhs(), r_orc), occ
        s.close()


@pytest.mark.parametrize("kind,positivity", [("extrapolation", False), ("symmetry", True)])
def test_trajectory_equals_stored_basis(mlr, kind, positivity):
    bcs = [(z, kind, {}) for z in ("left", "right", "top", "bottom")]
    out = []
    for occ, tables in SHAPES:
        s = _solver(mlr, 45, 33, 9, "corrected", bcs=bcs, positivity=positivity)
        s.set_tune(occ, 128, 128, 0, tables)
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        dts = s.run_steps(20)
        out.append((s.state()[0], dts))
        s.close()
    assert np.isfinite(out[0][0]).all()
    for k in range(1, len(SHAPES)):
        assert np.array_equal(out[0][0], out[k][0]), SHAPES[k]
        assert np.array_equal(out[0][1], out[k][1]), SHAPES[k]


def test_derived_basis_refused_where_not_derivable(mlr):
    """Monomial basis and p = 3 keep the stored face basis; asking for the
    derived one is an error, not a silent switch."""
    for kw in (dict(basis="monomial", order=2), dict(order=3)):
        num = mlr.numerics(recon="TENO", teno_variant="corrected", cfl=0.4, **kw)
        s = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 24, 24, 1.0, 1.0, 0.2, 3), num=num)
        with pytest.raises(mlr.MallardError, match="derived TENO face basis"):
            s.set_tune(0, 256, 256, 0, 1)
        s.close()
