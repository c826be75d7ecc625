"""The oracle's positivity-preserving flux limiter on whole runs (its pieces
are pinned in oracle_pins.check_positivity_limiter): random four-quadrant
cases of the small benchmark ensemble that reach vacuum within 50 steps stay
physical with it, and a run it never touches equals the unlimited run bit
for bit."""
import numpy as np
import pytest

EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]


def small_ensemble_case(k):
    """Case k of the 1024-case ensemble of 8..48 quads per side (sizes drawn
    log-uniformly with seed 2112, states with random_quadrant_states from
    seed 2024, mesh seed 2024 + k, jitter 0.2, split at (0.5, 0.5))."""
    from mallard_amd import indicator as ind
    sizes = np.round(np.exp(np.random.default_rng(2112).uniform(np.log(8), np.log(48), 1024))).astype(int)
    rng = np.random.default_rng(2024)
    states = [ind.random_quadrant_states(rng) for _ in range(k + 1)]
    return int(sizes[k]), states[k]


def physical(s):
    U, W = s.state()
    return bool(np.isfinite(U).all() and (U[:, 0] > 0).all() and (W[:, 2] > 0).all())


@pytest.mark.parametrize("k", [17, 20, 41, 69])
def test_vacuum_cases_stay_physical(oracle, k):
    n, st = small_ensemble_case(k)
    runs = {}
    for pp in (False, True):
        m = oracle.Mesh("cartesian_tri", n, n, 1.0, 1.0, 0.2, 2024 + k)
        s = oracle.Solver(m, oracle.make_config(order=2, teno_variant="corrected", bcs=EXTRAP, positivity=pp), 4,
                          canonical=True)
        s.init_quadrant(0.5, 0.5, st)
        try:
            s.run(50, 0.4)
            runs[pp] = (physical(s), s.pp_limited())
        except RuntimeError as e:  # a non-finite spectral radius: "dt negative: nan"
            runs[pp] = (False, str(e))
    assert not runs[False][0], runs
    assert runs[True][0] and runs[True][1] > 0, runs


def test_untouched_run_equals_unlimited(oracle):
    """Sod-like data the limiter never needs: identical states."""
    st = np.array([[1.0, 0.1, 0.0, 1.0], [0.6, 0.1, 0.0, 0.7], [0.5, 0.0, 0.1, 0.6], [0.8, 0.0, 0.0, 0.9]])
    out = []
    for pp in (False, True):
        m = oracle.Mesh("cartesian_tri", 24, 24, 1.0, 1.0, 0.2, 9)
        s = oracle.Solver(m, oracle.make_config(order=2, teno_variant="corrected", bcs=EXTRAP, positivity=pp), 4,
                          canonical=True)
        s.init_quadrant(0.5, 0.5, st)
        s.run(20, 0.4)
        out.append((s.state()[0], s.time(), s.pp_limited()))
    assert out[1][2] == 0
    assert np.array_equal(out[0][0].view(np.int64), out[1][0].view(np.int64)) and out[0][1] == out[1][1]


def test_rk4_refused(oracle):
    m = oracle.Mesh("cartesian_tri", 4, 4, 1.0, 1.0, 0.0, 0)
    with pytest.raises(RuntimeError, match="positivity"):
        oracle.Solver(m, oracle.make_config(order=2, integrator="RK4", bcs=EXTRAP, positivity=True), 1)
