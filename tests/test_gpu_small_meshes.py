"""The smallest meshes against the oracle (runs on an MI355X).

A single cell whose every face is a boundary face, strips one cell high, and
TENO meshes of one partial 64-cell tile, one full tile and a strip. The bars
are those of the other parity tests:
- first order: residual, dts and states bit-exact;
- corrected TENO: the residual bit-exact, and 20 SSP-RK3 steps within 1e-10
  relative L-inf (the device pow and glibc pow in the two-rarefaction
  estimate differ by 1 ulp on a few operands).
"""
import numpy as np
import pytest

from test_gpu_parity import RIEMANN2D, pair, rel_linf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,nx,ny", [("cartesian", 1, 1), ("cartesian_tri", 1, 1), ("cartesian", 3, 1),
                                        ("cartesian_tri", 5, 1), ("cartesian", 1, 4)])
def test_fo_smallest_meshes_bitexact(mlr, oracle, kind, nx, ny):
    dev, orc, m = pair(mlr, oracle, kind, nx, ny, recon="FO", integrator="SSPRK3", cfl=0.4)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    assert np.array_equal(dev.calc_rhs(), orc.calc_rhs(orc.state()[0]))
    assert np.array_equal(dev.run_steps(20), orc.run(20, 0.4))
    assert np.array_equal(dev.state()[0], orc.state()[0])


@pytest.mark.parametrize("nx,ny,order", [(3, 3, 2), (4, 4, 2), (8, 4, 2), (16, 1, 2), (2, 2, 1), (6, 5, 3)])
def test_teno_smallest_meshes(mlr, oracle, nx, ny, order):
    """18, 32 and 64 cells (one partial tile, one full tile), a strip one
    cell high, and the smallest meshes tried for p = 1 and p = 3."""
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", nx, ny, recon="TENO", order=order,
                       teno_variant="corrected", cfl=0.4)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    assert np.array_equal(dev.calc_rhs(), orc.calc_rhs(orc.state()[0]))
    dts = dev.run_steps(20)
    dto = orc.run(20, 0.4)
    U1, U2 = dev.state()[0], orc.state()[0]
    assert np.isfinite(U2).all()
    assert rel_linf(U1, U2) <= 1e-10
    assert np.allclose(dts, dto, rtol=1e-12)
