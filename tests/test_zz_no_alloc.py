"""The allocation-free Riemann entry point (mlr_riemann_flux_noalloc) against
the allocating one (mlr_riemann_flux), through the C ABI on an MI355X: equal
bit for bit on two calls in a row, the second with fewer points (the fixed
scratch keeps the first call's contents), and batches beyond the scratch are
refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _states(rng, n, g=1.4):
    rho = rng.uniform(0.05, 3.0, n)
    p = np.exp(rng.uniform(np.log(0.01), np.log(5.0), n))
    u = rng.uniform(-3.0, 3.0, (n, 2))
    e = p / ((g - 1.0) * rho)
    return np.stack([rho, u[:, 0], u[:, 1], p, np.full(n, g), e + p / rho], axis=1)


def test_noalloc_riemann_flux_equals_allocating_entry(mlr):
    cap = mlr.riemann_flux_noalloc_capacity()
    assert cap == 65536
    rng = np.random.default_rng(99)
    for kind in ("HLLC", "HLL", "Rusanov", "Roe"):
        for n in (cap, 1000):
            th = rng.uniform(0.0, 2.0 * np.pi, n)
            nrm = np.stack([np.cos(th), np.sin(th)], axis=1)
            L, R = _states(rng, n), _states(rng, n)
            got = mlr.riemann_flux_noalloc(kind, nrm, L, R, guard=1)
            want = mlr.riemann_flux(kind, nrm, L, R, guard=1)
            assert got.shape == (n, 4)
            assert np.array_equal(got, want), (kind, n)
    with pytest.raises(mlr.MallardError, match="exceed the fixed scratch"):
        z = np.zeros((cap + 1, 2))
        mlr.riemann_flux_noalloc("HLLC", z, np.zeros((cap + 1, 6)), np.zeros((cap + 1, 6)))
    # an empty batch is a no-op
    assert mlr.riemann_flux_noalloc("HLLC", np.zeros((0, 2)), np.zeros((0, 6)), np.zeros((0, 6))).shape == (0, 4)
