"""Batched ensembles on the GPU: every case equals its standalone run bit for
bit (batch invariance), with uneven mesh sizes in one launch per phase."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]


def _cases(mlr, seed=3):
    from mallard_amd import indicator as ind
    rng = np.random.default_rng(seed)
    sizes = [(16, 16), (24, 24), (40, 40), (33, 20), (64, 64)]
    gens = [mlr.mesh_gen("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, 100 + k) for k, (nx, ny) in enumerate(sizes)]
    states = [ind.random_quadrant_states(rng) for _ in sizes]
    splits = [(rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7)) for _ in sizes]
    return gens, states, splits


@pytest.mark.parametrize("integrator", ["SSPRK3", "RK4"])
def test_ensemble_equals_standalone_runs(mlr, integrator):
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP,
                       integrator=integrator)
    gens, states, splits = _cases(mlr)
    ens = mlr.Ensemble(gens, num)
    for k in range(len(gens)):
        ens.init_quadrant(k, *splits[k], states[k])
    ens.run_steps(12)
    ens.run_steps_profiled(3)
    for k, g in enumerate(gens):
        s = mlr.Solver(gen=g, num=num)
        s.init_quadrant(*splits[k], states[k])
        s.run_steps(15)
        U, _ = ens.state(k)
        assert np.array_equal(U, s.state()[0]), k
        assert ens.time(k) == s.time(), k


@pytest.mark.parametrize("split", [True, False])
def test_ensemble_with_gate_equals_standalone(mlr, split):
    """With the gate on, the batched passes (split or one kernel) equal the
    standalone runs made the other way, states and gated counts."""
    from oracle import indicator_ref as ir
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    gens, states, splits = _cases(mlr, seed=8)
    rng = np.random.default_rng(1)
    th = (rng.normal(0, 0.3, mlr.N_PARAMS)).astype(np.float32)
    th[ir.OFF_B3] = 2.0  # a net that gates some cells and not others
    ens = mlr.Ensemble(gens, num)
    ens.set_indicator(th, 0.9, split=split)
    for k in range(len(gens)):
        ens.init_quadrant(k, *splits[k], states[k])
    ens.run_steps(10)
    gated = ens.gated()
    total = 0
    for k, g in enumerate(gens):
        s = mlr.Solver(gen=g, num=num)
        s.set_indicator(th, 0.9, split=not split)
        s.init_quadrant(*splits[k], states[k])
        s.run_steps(10)
        total += s.gated()
        assert np.array_equal(ens.state(k)[0], s.state()[0]), k
    assert 0 < gated == total


@pytest.mark.parametrize("recon,sizes", [("TENO", [(3, 3), (40, 40), (8, 4), (16, 1)]),
                                         ("FO", [(1, 1), (5, 1), (24, 24), (2, 3)])])
def test_ensemble_tiny_cases_equal_standalone(mlr, recon, sizes):
    """Cases of one partial tile or block (down to a single cell) beside a
    larger one: each still equals its standalone run."""
    from mallard_amd import indicator as ind
    num = mlr.numerics(recon=recon, order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    rng = np.random.default_rng(17)
    gens = [mlr.mesh_gen("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, 40 + k) for k, (nx, ny) in enumerate(sizes)]
    states = [ind.random_quadrant_states(rng) for _ in sizes]
    ens = mlr.Ensemble(gens, num)
    for k in range(len(gens)):
        ens.init_quadrant(k, 0.5, 0.5, states[k])
    ens.run_steps(8)
    for k, g in enumerate(gens):
        s = mlr.Solver(gen=g, num=num)
        s.init_quadrant(0.5, 0.5, states[k])
        s.run_steps(8)
        assert np.array_equal(ens.state(k)[0], s.state()[0]), k
        assert ens.time(k) == s.time(), k


@pytest.mark.parametrize("gate", [False, True])
def test_ensemble_cases_equal_their_oracle_runs(mlr, oracle, gate):
    """Six cases from 16^2 to 96^2 quads with random quadrant states of the
    learned indicator's training distribution (rho, p ~ U[0.1, 2], u, v ~
    U[-1.5, 1.5]), advanced 20 steps in one batch, each against its own
    oracle run: dt within 1e-12 and the state within 1e-10 relative L-inf
    (the two-rarefaction pow differs from glibc's by 1 ulp on 0.08% of
    inputs). With a gate that never fires the batch is unchanged. (The draw
    is one whose six cases all stay physical for 20 steps; about one random
    case in six reaches vacuum within 20 steps in the oracle.)"""
    from mallard_amd import indicator as ind
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    rng = np.random.default_rng(34)
    sizes = [16, 24, 40, 56, 72, 96]
    gens = [mlr.mesh_gen("cartesian_tri", n, n, 1.0, 1.0, 0.2, 500 + k) for k, n in enumerate(sizes)]
    states = [ind.random_quadrant_states(rng) for _ in sizes]
    splits = [(rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7)) for _ in sizes]
    ens = mlr.Ensemble(gens, num)
    if gate:
        theta = np.zeros(mlr.N_PARAMS, np.float32)
        theta[-1] = -50.0
        ens.set_indicator(theta, 0.9)
    for k in range(len(sizes)):
        ens.init_quadrant(k, *splits[k], states[k])
    ens.run_steps(20)
    if gate:
        assert ens.gated() == 0
    for k, n in enumerate(sizes):
        m = oracle.Mesh("cartesian_tri", n, n, 1.0, 1.0, 0.2, 500 + k)
        orc = oracle.Solver(m, oracle.make_config(order=2, teno_variant="corrected", bcs=EXTRAP), 8,
                            canonical=True)
        orc.init_quadrant(*splits[k], states[k])
        orc.run(20, 0.4)
        U, _ = ens.state(k)
        Uo = orc.state()[0]
        t, steps = ens.time(k)
        assert steps == 20
        assert abs(t / orc.time() - 1.0) <= 1e-12, k
        bad, bado = ~np.isfinite(U), ~np.isfinite(Uo)
        assert np.array_equal(bad, bado), k
        ok = ~bado.any(axis=1)
        scale = np.abs(Uo[ok]).max(axis=0)
        diff = (np.abs(U[ok] - Uo[ok]).max(axis=0) / scale).max()
        print("case %d (%d^2): relative L-inf %.3g, bitwise %s" % (k, n, diff, np.array_equal(U[ok], Uo[ok])))
        assert diff <= 1e-10, k
