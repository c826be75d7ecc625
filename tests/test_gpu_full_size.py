"""Parity at the benchmark's full size: the 8,388,608-cell four-quadrant
Riemann workload (2048^2 quads split into triangles, jitter 0.2, seed 2024,
TENO p = 2 corrected, HLLC, SSP-RK3, CFL 0.4) run 100 steps on the GPU
against the CPU oracle's run of the same workload, in two boundary sets:
extrapolation on the four sides (tests/golden/riemann2048_corrected_100.npz),
and symmetry walls on the four sides with the positivity-preserving flux
limiter on (tests/golden/riemann2048_symmetry_pp_100.npz), both written by
tests/golden/make_riemann2048_fixture.py.

Bar (the acceptance tolerance of the path): conserved variables within 1e-10
relative L-inf (per variable, relative to the field's largest magnitude) on
the sampled cells and in every whole-field norm, and each step's dt within
1e-12 relative. Both accumulate each cell's faces in the one-thread order;
the oracle's two-rarefaction estimate takes glibc's pow and the GPU its own
double-double pow (within 1 ulp of glibc, different on 0.08% of inputs), so
the two agree to rounding rather than bit for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
QUADRANT = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                     [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])


@pytest.mark.parametrize("fixture,kind,positivity", [
    ("riemann2048_corrected_100.npz", "extrapolation", False),
    ("riemann2048_symmetry_pp_100.npz", "symmetry", True),
], ids=["extrapolation", "symmetry_positivity"])
def test_riemann2048_100_steps_against_oracle(mlr, fixture, kind, positivity):
    ref = np.load(os.path.join(GOLDEN, fixture))
    nx, jitter, seed, steps, cfl = ref["meta"][:5]
    nx, seed, steps = int(nx), int(seed), int(steps)
    bcs = [(z, kind, {}) for z in ("left", "right", "top", "bottom")]
    num = mlr.numerics(recon="TENO", basis="legendre", order=2, stencil_factor=2.0, teno_variant="corrected",
                       riemann="HLLC", integrator="SSPRK3", cfl=cfl, bcs=bcs, positivity=positivity)
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", nx, nx, 1.0, 1.0, jitter, seed), num=num)
    assert dev.n_cells == 2 * nx * nx
    dev.init_quadrant(0.5, 0.5, QUADRANT)
    dts = dev.run_steps(steps)
    U, _ = dev.state()
    vol = dev.volumes()
    dev.close()
    print("max relative dt difference %.3g" % float(np.abs(dts / ref["dts"] - 1.0).max()))
    assert np.abs(dts / ref["dts"] - 1.0).max() <= 1e-12
    scale = ref["linf"]
    diff = np.abs(U[ref["ids"]] - ref["U"]).max(axis=0) / scale
    print("sample relative L-inf per variable", diff)
    assert (diff <= 1e-10).all()
    assert (np.abs(np.abs(U).max(axis=0) - ref["linf"]) <= 1e-10 * scale).all()
    assert (np.abs(U.min(axis=0) - ref["umin"]) <= 1e-10 * scale).all()
    l1 = (np.abs(U) * vol[:, None]).sum(axis=0)
    assert (np.abs(l1 / ref["l1"] - 1.0) <= 1e-10).all()
    total = (U * vol[:, None]).sum(axis=0)
    assert (np.abs(total - ref["total"]) <= 1e-10 * np.maximum(np.abs(ref["total"]), ref["l1"])).all()
