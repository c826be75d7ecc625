"""The persistent ensemble step (one cooperative launch per step over every
case, grid-wide barriers between the dt, TENO and flux phases) against the
graph of per-phase launches: every case's state, time and step count equal
bit for bit, for TENO and first order, both integrators, every Riemann
solver, fixed and CFL time steps, cases of one partial block beside larger
ones, and a switch between the two modes mid-run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
SIZES = [(16, 16), (3, 3), (40, 40), (33, 20), (16, 1), (64, 64)]  # (1, 1) too small for TENO stencils
SIZES_FO = [(16, 16), (1, 1), (40, 40), (33, 20), (5, 1), (64, 64)]


def _init(mlr, ens, seed):
    from mallard_amd import indicator as ind
    rng = np.random.default_rng(seed)
    for k in range(ens.n):
        ens.init_quadrant(k, rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), ind.random_quadrant_states(rng))


def _ensemble(mlr, num, seed, sizes=SIZES):
    gens = [mlr.mesh_gen("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, 300 + k) for k, (nx, ny) in enumerate(sizes)]
    ens = mlr.Ensemble(gens, num)
    _init(mlr, ens, seed)
    return ens


def _same(a, b):
    for k in range(a.n):
        Ua, Ub = a.state(k)[0], b.state(k)[0]
        assert np.array_equal(Ua.view(np.int64), Ub.view(np.int64)), k
        assert a.time(k) == b.time(k), k


@pytest.mark.parametrize("recon,integrator,riemann", [
    ("TENO", "SSPRK3", "HLLC"), ("TENO", "RK4", "HLLC"), ("TENO", "SSPRK3", "Rusanov"),
    ("TENO", "SSPRK3", "HLL"), ("TENO", "SSPRK3", "Roe"), ("FO", "SSPRK3", "HLLC"), ("FO", "RK4", "Roe")])
def test_persistent_step_equals_graph(mlr, recon, integrator, riemann):
    num = mlr.numerics(recon=recon, order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP,
                       integrator=integrator, riemann=riemann)
    sizes = SIZES if recon == "TENO" else SIZES_FO
    ref = _ensemble(mlr, num, 5, sizes)
    ref.run_steps(10)
    ens = _ensemble(mlr, num, 5, sizes)
    ens.set_persistent(True)
    ens.run_steps(10)
    print("persistent grid: %d blocks" % ens.persistent_grid)
    assert ens.persistent_grid >= 256
    _same(ens, ref)
    for k in range(ens.n):
        assert ens.time(k)[1] == 10


def test_persistent_fixed_dt_and_mode_switch(mlr):
    """Fixed dt, the reference variant, and steps alternating between the
    persistent launch and the graph."""
    num = mlr.numerics(recon="TENO", order=2, teno_variant="reference", cfl=0.0, dt=2e-4, bcs=EXTRAP)
    ref = _ensemble(mlr, num, 9)
    ref.set_tune(block=256)
    ref.run_steps(9)
    ens = _ensemble(mlr, num, 9)
    ens.set_persistent(True)
    ens.run_steps(3)
    ens.set_persistent(False)
    ens.set_tune(block=256)
    ens.run_steps(3)
    ens.set_persistent(True)
    ens.run_steps(3)
    _same(ens, ref)
    for k in range(ens.n):
        t, steps = ens.time(k)
        assert steps == 9 and abs(t - 9 * 2e-4) < 1e-15


def test_persistent_refuses_what_it_does_not_run(mlr):
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    ens = _ensemble(mlr, num, 2)
    theta = np.zeros(mlr.N_PARAMS, np.float32)
    ens.set_indicator(theta, 0.9)
    with pytest.raises(mlr.MallardError, match="plain TENO"):
        ens.set_persistent(True)
    ens.set_indicator(None)
    ens.set_persistent(True)
    with pytest.raises(mlr.MallardError, match="plain TENO"):
        ens.set_indicator(theta, 0.9)
    with pytest.raises(mlr.MallardError, match="256 cells"):
        ens.set_tune(block=128)
    ens.run_steps(2)
    num1 = mlr.numerics(recon="TENO", order=1, cfl=0.4, bcs=EXTRAP)
    e1 = _ensemble(mlr, num1, 2)
    with pytest.raises(mlr.MallardError, match="p = 2"):
        e1.set_persistent(True)


def test_persistent_watchdog_reports_and_recovers(mlr):
    """A zero wait limit makes the blocks that reach a grid barrier early give
    up at once (every block still ends its launch): run_steps raises. With
    the limit restored, the barrier state has been reset: the same ensemble,
    set back to its initial states, then runs exactly as the graph does."""
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    ref = _ensemble(mlr, num, 12)
    ref.run_steps(4)
    ens = _ensemble(mlr, num, 12)
    ens.set_persistent(True)
    ens.persistent_wait_limit(0.0)
    with pytest.raises(mlr.MallardError, match="waited longer"):
        ens.run_steps(1)
    with pytest.raises(mlr.MallardError, match="within"):
        ens.persistent_wait_limit(-1.0)
    ens.persistent_wait_limit(2000.0)
    _init(mlr, ens, 12)
    ens.run_steps(4)
    for k in range(ens.n):
        assert np.array_equal(ens.state(k)[0].view(np.int64), ref.state(k)[0].view(np.int64)), k
