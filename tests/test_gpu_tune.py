"""Launch-shape autotuner on the GPU: every shape it can pick (TENO occupancy
target, flux and spectral block sizes, stored or derived TENO face basis) gives
bit-identical results, with and without the learned gate, for single
solvers and ensembles; autotuning leaves the state and the gated-cell count
untouched."""
import numpy as np
import pytest

from oracle import indicator_ref as ir

pytestmark = pytest.mark.gpu

EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
RIEMANN2D = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                      [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])
SHAPES = [(0, 256, 256, 0, 0), (5, 64, 128, 6, 0), (6, 128, 64, 5, 0), (0, 256, 256, 0, 1), (4, 128, 64, 5, 1),
          (5, 64, 256, 6, 1)]


def _num(mlr):
    return mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)


def _half_gate():
    """A network whose logits spread around the threshold: some cells gated, some not."""
    rng = np.random.default_rng(4)
    th = rng.normal(0.0, 0.2, ir.N_PARAMS).astype(np.float32)
    th[ir.OFF_B3] = 2.2
    return th


@pytest.mark.parametrize("gate", [False, True])
def test_every_shape_gives_identical_results(mlr, gate):
    gen = mlr.mesh_gen("cartesian_tri", 48, 48, 1.0, 1.0, 0.2, 6)
    out, counts = [], []
    for shape in SHAPES:
        s = mlr.Solver(gen=gen, num=_num(mlr))
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        if gate:
            s.set_indicator(_half_gate(), 0.9)
        s.set_tune(*shape)
        assert s.tune() == shape
        s.run_steps(6)
        s.run_steps_profiled(2)
        out.append(s.state()[0])
        counts.append(s.gated() if gate else 0)
    for k in range(1, len(SHAPES)):
        assert np.array_equal(out[0], out[k]), SHAPES[k]
        assert counts[0] == counts[k]
    if gate:
        assert 0 < counts[0] < 48 * 48 * 2 * 3 * 8


def test_p3_occupancy_targets_give_identical_results(mlr):
    """Plain TENO at p = 3 in its three occupancy targets (the compiler's
    choice, 4 and 5 waves per SIMD) gives the same states bit for bit."""
    gen = mlr.mesh_gen("cartesian_tri", 40, 40, 1.0, 1.0, 0.2, 12)
    num = mlr.numerics(recon="TENO", order=3, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    out = []
    for occ in (0, 4, 5):
        s = mlr.Solver(gen=gen, num=num)
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        s.set_tune(occ, 128, 128, 0, 0)
        s.run_steps(5)
        out.append(s.state()[0])
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    s = mlr.Solver(gen=gen, num=num)
    s.init_quadrant(0.5, 0.5, RIEMANN2D)
    assert s.autotune()[0] in (0, 4, 5)


def test_autotune_leaves_state_and_counts(mlr):
    s = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 40, 40, 1.0, 1.0, 0.2, 9), num=_num(mlr))
    s.init_quadrant(0.5, 0.5, RIEMANN2D)
    s.set_indicator(_half_gate(), 0.9)
    s.set_tune(0, 256, 256, 0)
    s.run_steps(3)
    U0, g0 = s.state()[0], s.gated()
    occ, fb, sb, occ2, tables = s.autotune()
    assert occ in (0, 4, 5, 6) and occ2 in (0, 5, 6) and fb in (64, 128, 256) and sb in (64, 128, 256)
    assert tables in (0, 1)
    assert s.tune() == (occ, fb, sb, occ2, tables)
    assert np.array_equal(s.state()[0], U0)
    assert s.gated() == g0
    # and the steps after it equal the steps of an untuned twin
    t = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 40, 40, 1.0, 1.0, 0.2, 9), num=_num(mlr))
    t.init_quadrant(0.5, 0.5, RIEMANN2D)
    t.set_indicator(_half_gate(), 0.9)
    t.set_tune(0, 256, 256, 0)
    t.run_steps(3)
    s.run_steps(4)
    t.run_steps(4)
    assert np.array_equal(s.state()[0], t.state()[0])


def test_ensemble_shapes_give_identical_results(mlr):
    from mallard_amd import indicator as ind
    rng = np.random.default_rng(8)
    sizes = [20, 36, 52]
    gens = [mlr.mesh_gen("cartesian_tri", n, n, 1.0, 1.0, 0.2, 50 + n) for n in sizes]
    states = [ind.random_quadrant_states(rng) for _ in sizes]
    res = []
    for occ, bs, occ2 in [(0, 256, 0), (6, 64, 5), (5, 128, 6)]:
        ens = mlr.Ensemble(gens, _num(mlr))
        for k in range(len(gens)):
            ens.init_quadrant(k, 0.5, 0.5, states[k])
        ens.set_indicator(_half_gate(), 0.9)
        ens.set_tune(occ, bs, occ2)
        ens.run_steps(5)
        res.append([ens.state(k)[0] for k in range(len(gens))])
        res[-1].append(ens.gated())
        ens.close()
    for r in res[1:]:
        for k in range(len(gens)):
            assert np.array_equal(res[0][k], r[k]), k
        assert res[0][-1] == r[-1]
    ens = mlr.Ensemble(gens, _num(mlr))
    for k in range(len(gens)):
        ens.init_quadrant(k, 0.5, 0.5, states[k])
    occ, bs, bs2, occ2, fb = ens.autotune()
    assert occ in (0, 4, 5, 6) and occ2 in (0, 5, 6) and bs in (64, 128, 256) and bs2 == bs and fb in (0, 1)


def test_ensemble_derived_face_basis_gives_identical_results(mlr):
    """The plain batched TENO pass with each case's face basis derived from
    its node coordinates, at three occupancy targets, equals the stored-basis
    pass bit for bit, and each case equals its standalone solver."""
    sizes = [(20, 20), (33, 27), (48, 48)]
    gens = [mlr.mesh_gen("cartesian_tri", nx, ny, 1.0, 1.0, 0.2, 70 + nx) for nx, ny in sizes]
    res = []
    for occ, fb in [(0, 0), (0, 1), (4, 1), (5, 1)]:
        ens = mlr.Ensemble(gens, _num(mlr))
        for k in range(len(gens)):
            ens.init_quadrant(k, 0.5, 0.5, RIEMANN2D)
        ens.set_tune(occ, 128, 0, fb)
        ens.run_steps(8)
        res.append([ens.state(k)[0] for k in range(len(gens))])
        ens.close()
    for r in res[1:]:
        for k in range(len(gens)):
            assert np.array_equal(res[0][k], r[k]), k
    for k, gen in enumerate(gens):
        s = mlr.Solver(gen=gen, num=_num(mlr))
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        s.set_tune(0, 128, 128, 0, 1)
        s.run_steps(8)
        assert np.array_equal(s.state()[0], res[0][k]), k
        s.close()
