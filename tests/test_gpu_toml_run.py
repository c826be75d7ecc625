"""The drop-in path end to end on the GPU: Solver::init(path) and
Solver::run (mlr_ctx_create_toml, mlr_run) on the reference's own Sod
example input, against the CPU oracle on the same mesh and state."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sod_example_runs_to_the_oracle_state(mlr, oracle, tmp_path):
    """examples/sod/input.toml (1000 x 1 quads, FO, HLLC, SSPRK3, CFL 1,
    t_stop 0.2, symmetry walls, analytical initial state, VTU output every
    100 steps) run by mlr_run from the state the input defines: the run
    stops at the first step that reaches t_stop (dt is not clipped), as the
    oracle's time does after the same number of steps, and the write_data
    table's files are written. The rarefaction's two-rarefaction estimates
    take the device's double-double pow where the oracle calls glibc's,
    which differ by one ulp on about 0.1% of arguments (here first at step
    55), so state and time are compared to 1e-10 relative L-inf."""
    src = open(os.path.join(HERE, "golden", "examples", "sod.toml")).read()
    prefix = str(tmp_path / "sod_all")
    assert 'prefix = "./solut/all/sod_all"' in src
    p = tmp_path / "sod.toml"
    p.write_text(src.replace('prefix = "./solut/all/sod_all"', 'prefix = "%s"' % prefix))
    s = mlr.Solver(toml_path=p)
    U0, W0 = s.state()  # the analytical initial state, as the input file defines it
    s.run()
    t, steps = s.time()
    U = s.state()[0]
    s.close()

    m = oracle.Mesh("cartesian", 1000, 1, 1.0, 0.001, 0.0, 0)
    orc = oracle.Solver(m, oracle.make_config(recon="FO", riemann="HLLC", integrator="SSPRK3"), 1)
    orc.set_state(U0, W0)
    dts = orc.run(steps, 1.0)
    assert orc.time() >= 0.2 > orc.time() - dts[-1]  # the step that crossed t_stop was the last
    assert t >= 0.2 and abs(t - orc.time()) <= 1e-12 * t
    Uo = orc.state()[0]
    assert np.isfinite(U).all()
    scale = np.maximum(np.abs(Uo).max(axis=0), 1e-300)  # rho v is zero throughout
    assert (np.abs(U - Uo).max(axis=0) / scale).max() <= 1e-10
    files = sorted(os.listdir(tmp_path))
    vtus = [f for f in files if f.startswith("sod_all_") and f.endswith(".vtu")]
    assert len(vtus) >= steps // 100
    head = open(tmp_path / vtus[0], "rb").read(4096)
    for name in (b"CFL", b"RHO", b"RHOU_X", b"P", b"T", b"H"):
        assert b'Name="' + name + b'"' in head


def _sod_with_run(tmp_path, run):
    """sod.toml with its [run] table replaced and its output redirected."""
    src = open(os.path.join(HERE, "golden", "examples", "sod.toml")).read()
    src = src.replace('prefix = "./solut/all/sod_all"', 'prefix = "%s"' % str(tmp_path / "sod_all"))
    head, tail = src.split("[run]", 1)
    rest = tail.split("\n[", 1)
    p = tmp_path / "sod.toml"
    p.write_text(head + "[run]\n" + run + "\n" + ("\n[" + rest[1] if len(rest) > 1 else ""))
    return p


def test_sod_example_n_steps_stop(mlr, oracle, tmp_path):
    """n_steps ends the run after exactly that many steps (solver.cpp:375-403),
    at the oracle's time and state (1e-10, as above)."""
    s = mlr.Solver(toml_path=_sod_with_run(tmp_path, "n_steps = 37\ncfl = 1.0"))
    U0, W0 = s.state()
    s.run()
    t, steps = s.time()
    U = s.state()[0]
    s.close()
    assert steps == 37
    m = oracle.Mesh("cartesian", 1000, 1, 1.0, 0.001, 0.0, 0)
    orc = oracle.Solver(m, oracle.make_config(recon="FO", riemann="HLLC", integrator="SSPRK3"), 1)
    orc.set_state(U0, W0)
    orc.run(37, 1.0)
    assert abs(t - orc.time()) <= 1e-12 * t
    Uo = orc.state()[0]
    scale = np.maximum(np.abs(Uo).max(axis=0), 1e-300)
    assert (np.abs(U - Uo).max(axis=0) / scale).max() <= 1e-10


def test_sod_example_wall_clock_stop(mlr, tmp_path):
    """t_wall_stop ends a run that no step or time limit would end
    (solver.cpp:394-397): the clock starts when the context is made, so the
    run ends no earlier than t_wall_stop after the creation call began, and
    within one step after the limit (here well under a few seconds). With
    an n_steps limit the run cannot reach, the wall clock still decides."""
    import time
    for extra in ("", "n_steps = 1000000000\n"):
        t0 = time.perf_counter()
        s = mlr.Solver(toml_path=_sod_with_run(tmp_path, extra + "t_wall_stop = 0.5\ndt = 1.0e-6"))
        t1 = time.perf_counter()
        s.run()
        t2 = time.perf_counter()
        t, steps = s.time()
        U = s.state()[0]
        s.close()
        assert t2 - t0 >= 0.5, extra
        assert t2 - t1 <= 0.5 + 3.0, extra
        assert 0 < steps < 1000000000 and abs(t - steps * 1.0e-6) <= 1e-9
        assert np.isfinite(U).all()
