"""TENO tables beyond 2^32 entries (runs on an MI355X).

The reference indexes its reconstruction matrices with uint32_t offsets
(face_reconstruction.cpp:478, :893), so at p = 3 (684 entries per cell) it
wraps beyond 6.28M triangles. Here the tables are indexed in 64 bits. A
jittered 1800 x 1800 triangle mesh (6.48M cells, 4.43G pseudo-inverse
entries) checks this through a closed form that does not depend on the mesh
size: every TENO candidate reproduces a linear field exactly. With linear
conserved variables and extrapolation boundaries, both sides of each face
then carry the field's value there, the HLLC flux is the physical one, and
the two-point Gauss rule integrates the linear mass flux exactly. So the
mass residual of every cell away from the boundary is -div(rho u), a
constant. (Cells on the boundary are left out: the corrected variant's
extrapolation ghost is the cell mean, not the face value.)

A table read from another cell's entries (wrapped offsets) reconstructs a
jittered cell with the wrong geometry; that moves its mass residual by
roughly jitter x gradient x area / volume, about 0.1 here, so the bound of
1e-8 tells the two apart. The oracle's closed form
on a 30 x 30 mesh holds to 1e-13.
"""
import numpy as np
import pytest

EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]


def linear_state(xy):
    x, y = xy[:, 0], xy[:, 1]
    return np.stack([1.0 + 0.2 * x + 0.1 * y, 0.3 + 0.2 * x - 0.1 * y,
                     -0.2 + 0.1 * x + 0.3 * y, 5.0 + 0.5 * x + 0.4 * y], axis=1)


DIV_MOMENTUM = 0.2 + 0.3


def inner_cells(m):
    """Cells with no boundary face."""
    cof = m.cells_of_face[:m.n_faces_real]
    inner = np.ones(m.n_cells, bool)
    inner[cof[cof[:, 1] < 0, 0]] = False
    return inner


def test_linear_field_closed_form_oracle(oracle):
    """The closed form itself, on the oracle (small mesh, p = 2 and 3)."""
    for order in (2, 3):
        m = oracle.Mesh("cartesian_tri", 30, 30, 1.0, 1.0, 0.2, 5)
        s = oracle.Solver(m, oracle.make_config(recon="TENO", order=order, teno_variant="corrected",
                                                bcs=EXTRAP), 1)
        r = s.calc_rhs(linear_state(m.cell_coords))
        inner = inner_cells(m)
        assert np.abs(r[inner, 0] + DIV_MOMENTUM).max() <= 1e-12
        # the boundary cells differ from it (the closed form is not vacuous there)
        assert np.abs(r[~inner, 0] + DIV_MOMENTUM).max() > 1e-6


@pytest.mark.gpu
def test_p3_tables_beyond_uint32(mlr, oracle):
    nx = ny = 1800
    num = mlr.numerics(recon="TENO", order=3, teno_variant="corrected", bcs=EXTRAP)
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, 5), num=num)
    n = dev.n_cells
    assert (dev.G, dev.M, dev.K) == (4, 20, 10)
    assert n * dev.G * (dev.M - 1) * (dev.K - 1) > 2 ** 32  # pseudo-inverse entries
    m = oracle.Mesh("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, 5)
    dev.set_state(linear_state(m.cell_coords))
    r = dev.calc_rhs()
    err = np.abs(r[inner_cells(m), 0] + DIV_MOMENTUM)
    print("p=3, %d cells: max |R_rho + div(rho u)| = %.3e" % (n, err.max()))
    assert np.isfinite(r).all()
    assert err.max() <= 1e-8
