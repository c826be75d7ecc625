"""The positivity-preserving flux limiter on the GPU (Hu, Adams & Shu 2013;
no reference counterpart; the oracle's copy is pinned in oracle_pins):

- cases of the small benchmark ensemble that reach vacuum within 50 steps
  run physical with it and equal the oracle's limited runs (dt within 1e-12,
  state within 1e-10 relative L-inf: the two-rarefaction pow differs from
  glibc's by 1 ulp on 0.08% of inputs), with extrapolation and with
  symmetry walls;
- where the limiter never acts the states equal the unlimited run bit for
  bit;
- the full-size four-quadrant problem with symmetry walls (2048^2 quads,
  8,388,608 triangles), which goes non-finite by step 40 without it, runs
  100 steps physical with it;
- in an ensemble every case equals its standalone limited run bit for bit;
- RK4, viscous runs and the persistent ensemble step refuse it."""
import numpy as np
import pytest

from test_oracle_positivity import small_ensemble_case

pytestmark = pytest.mark.gpu

EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
SYM = [(z, "symmetry", {}) for z in ("left", "right", "top", "bottom")]
QUADRANT = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                     [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])


def _physical(U, W):
    return bool(np.isfinite(U).all() and (U[:, 0] > 0).all() and (W[:, 2] > 0).all())


@pytest.mark.parametrize("k,bcs", [(17, "extrapolation"), (20, "extrapolation"), (41, "extrapolation"),
                                   (69, "extrapolation"), (17, "symmetry"), (105, "symmetry")])
def test_limited_runs_equal_oracle(mlr, oracle, k, bcs):
    n, st = small_ensemble_case(k)
    bc = EXTRAP if bcs == "extrapolation" else SYM
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=bc, positivity=True)
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", n, n, 1.0, 1.0, 0.2, 2024 + k), num=num)
    dev.init_quadrant(0.5, 0.5, st)
    dts = dev.run_steps(50)
    U, W = dev.state()
    m = oracle.Mesh("cartesian_tri", n, n, 1.0, 1.0, 0.2, 2024 + k)
    orc = oracle.Solver(m, oracle.make_config(order=2, teno_variant="corrected", bcs=bc, positivity=True), 8,
                        canonical=True)
    orc.init_quadrant(0.5, 0.5, st)
    dto = orc.run(50, 0.4)
    Uo = orc.state()[0]
    assert orc.pp_limited() > 0
    assert _physical(U, W)
    assert np.abs(dts / dto - 1.0).max() <= 1e-12
    diff = (np.abs(U - Uo).max(axis=0) / np.abs(Uo).max(axis=0)).max()
    print("case %d (%d^2, %s): %d limited faces in the oracle, relative L-inf %.3g" %
          (k, n, bcs, orc.pp_limited(), diff))
    assert diff <= 1e-10


def test_untouched_run_is_bitwise_unlimited(mlr, oracle):
    st = np.array([[1.0, 0.1, 0.0, 1.0], [0.6, 0.1, 0.0, 0.7], [0.5, 0.0, 0.1, 0.6], [0.8, 0.0, 0.0, 0.9]])
    out = []
    for pp in (False, True):
        num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP, positivity=pp)
        s = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 24, 24, 1.0, 1.0, 0.2, 9), num=num)
        s.init_quadrant(0.5, 0.5, st)
        s.run_steps(20)
        out.append((s.state()[0], s.time()))
    assert np.array_equal(out[0][0].view(np.int64), out[1][0].view(np.int64)) and out[0][1] == out[1][1]


def test_full_size_symmetry_walls_stay_physical(mlr):
    res = {}
    for pp in (False, True):
        num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=SYM, positivity=pp)
        s = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 2048, 2048, 1.0, 1.0, 0.2, 2024), num=num)
        s.init_quadrant(0.5, 0.5, QUADRANT)
        s.run_steps(50 if not pp else 100)
        U, W = s.state()
        res[pp] = _physical(U, W)
        if pp:
            print("limited: min rho %.4g, min p %.4g after 100 steps" % (U[:, 0].min(), W[:, 2].min()))
        s.close()
    assert not res[False] and res[True]


def test_ensemble_limited_cases_equal_standalone(mlr):
    ks = [17, 20, 41, 69, 105, 116]
    cases = [small_ensemble_case(k) for k in ks]
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP, positivity=True)
    gens = [mlr.mesh_gen("cartesian_tri", n, n, 1.0, 1.0, 0.2, 2024 + k) for k, (n, _) in zip(ks, cases)]
    ens = mlr.Ensemble(gens, num)
    for i, (_, st) in enumerate(cases):
        ens.init_quadrant(i, 0.5, 0.5, st)
    ens.run_steps(50)
    with pytest.raises(mlr.MallardError, match="positivity"):
        ens.set_persistent(True)
    for i, g in enumerate(gens):
        s = mlr.Solver(gen=g, num=num)
        s.init_quadrant(0.5, 0.5, cases[i][1])
        s.run_steps(50)
        U, W = ens.state(i)
        assert _physical(U, W), ks[i]
        assert np.array_equal(U.view(np.int64), s.state()[0].view(np.int64)), ks[i]
        assert ens.time(i) == s.time(), ks[i]


def test_refusals(mlr):
    g = mlr.mesh_gen("cartesian_tri", 8, 8, 1.0, 1.0, 0.0, 0)
    with pytest.raises(mlr.MallardError, match="positivity"):
        mlr.Solver(gen=g, num=mlr.numerics(recon="TENO", integrator="RK4", bcs=EXTRAP, positivity=True))
    with pytest.raises(mlr.MallardError, match="positivity"):
        mlr.Solver(gen=g, num=mlr.numerics(recon="FO", bcs=EXTRAP, physics="navier-stokes", mu_ref=1e-3,
                                           positivity=True))
