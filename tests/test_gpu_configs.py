"""The workloads beyond the bench's own, each against the CPU oracle on an
MI355X:

- the Sod shock tube on triangles with TENO (cartesian_tri 1000 x 10 on
  [0, 1] x [0, 0.01], HLLC, SSP-RK3, CFL 0.5, to t = 0.2), against the oracle
  run committed as tests/golden/sod_tri_teno_t02.npz
  (tests/golden/make_sod_tri_fixture.py) and against Toro's exact solution;
  with the reference weighting, against a live oracle run up to its first
  non-finite step;
- the Mach 2.5 wedge pipeline on a graded Delaunay mesh (TENO + Roe with
  fixed-state inflow, pressure outflow, slip wall and extrapolation), the
  inviscid residual and 20 steps against the oracle on the same mesh arrays,
  and the Navier-Stokes steps against the oracle's Euler residual plus the
  numpy viscous restatement, stepped by the reference's SSP-RK3 sequence.

Bar: 1e-10 relative L-inf per variable (relative to the field's largest
magnitude) and 1e-12 relative per dt; the two-rarefaction pow differs from
glibc's by 1 ulp on 0.08% of inputs, so trajectories are not bit for bit.
"""
import os

import numpy as np
import pytest

from test_oracle_invariants import _sod_exact

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SOD = np.array([[0.125, 0.0, 0.0, 0.1], [1.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]])


def rel_linf(a, b):
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return float((np.abs(a - b).max(axis=0) / scale).max())


def test_config1_sod_teno_triangles(mlr, oracle):
    ref = np.load(os.path.join(GOLDEN, "sod_tri_teno_t02.npz"))
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", riemann="HLLC", integrator="SSPRK3",
                       cfl=0.5)
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 1000, 10, 1.0, 0.01, 0.0, 0), num=num)
    dev.init_quadrant(0.5, -1.0, SOD)
    n = len(ref["dts"])
    dts = dev.run_steps(n)
    t, steps = dev.time()
    assert steps == n and t >= 0.2
    assert np.abs(dts / ref["dts"] - 1.0).max() <= 1e-12
    assert abs(t / float(ref["t"]) - 1.0) <= 1e-12
    U, _ = dev.state()
    scale = ref["linf"]
    diff = np.abs(U[ref["ids"]] - ref["U"]).max(axis=0) / scale
    print("Sod TENO triangles: %d steps to t = %.6f, relative L-inf per variable %s" % (n, t, diff))
    assert (diff <= 1e-10).all()
    assert (np.abs(np.abs(U).max(axis=0) - ref["linf"]) <= 1e-10 * scale).all()
    vol = dev.volumes()
    assert (np.abs((np.abs(U) * vol[:, None]).sum(axis=0) / ref["l1"] - 1.0) <= 1e-10).all()
    # against the exact solution (Toro): L1 of the density over the domain
    m = oracle.Mesh("cartesian_tri", 1000, 10, 1.0, 0.01, 0.0, 0)
    ex = _sod_exact(m.cell_coords[:, 0], t)
    l1 = float(np.sum(np.abs(U[:, 0] - ex) * m.cell_volume) / np.sum(m.cell_volume))
    print("L1 density error against the exact solution: %.4g" % l1)
    assert l1 < 4e-3


def test_config1_reference_weighting_fails_alike(mlr, oracle):
    """The reference weighting goes non-finite on this problem: the GPU and
    the oracle reach their first non-finite state at the same step, with the
    same cells."""
    num = mlr.numerics(recon="TENO", order=2, teno_variant="reference", riemann="HLLC", integrator="SSPRK3",
                       cfl=0.5)
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 1000, 10, 1.0, 0.01, 0.0, 0), num=num)
    m = oracle.Mesh("cartesian_tri", 1000, 10, 1.0, 0.01, 0.0, 0)
    orc = oracle.Solver(m, oracle.make_config(order=2, teno_variant="reference"), 16, canonical=True)
    dev.init_quadrant(0.5, -1.0, SOD)
    orc.init_quadrant(0.5, -1.0, SOD)
    first = None
    for step in range(1, 201):
        d = dev.run_steps(1)
        o = orc.run(1, 0.5)
        U, Uo = dev.state()[0], orc.state()[0]
        bad, bado = ~np.isfinite(U), ~np.isfinite(Uo)
        if bad.any() or bado.any():
            first = step
            assert np.array_equal(bad, bado), step
            break
        assert abs(d[0] / o[0] - 1.0) <= 1e-12, step
        assert rel_linf(U, Uo) <= 1e-10, step
    print("reference weighting: first non-finite state at step", first)
    assert first is not None


GAMMA = 1.4
# nondimensional physics: R = p_ref / (T_ref rho_ref) = 1
PHYS = dict(gamma=GAMMA, p_ref=1.0, T_ref=1.0, rho_ref=1.0, p_min=-1e20, p_max=1e20)
TINF = 1.0 / GAMMA  # Mach 2.5 inflow: a = 1
# the wedge problem's zones: inflow "left", outflow "right" and "top",
# "symmetry" ahead of the wedge, the wedge surface "wall"
WEDGE_BCS = [("left", "upt", {"u": (2.5, 0.0), "p": 1.0, "T": TINF}), ("right", "extrapolation", {}),
             ("top", "extrapolation", {}), ("symmetry", "symmetry", {}), ("wall", "wall_adiabatic", {})]


def _wedge_pair(mlr, oracle, bcs, **kw):
    from mallard_amd import meshgen
    mesh = meshgen.wedge_mesh(n_cells=30000, seed=2)
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", riemann="Roe", cfl=0.4, bcs=bcs,
                       **PHYS, **kw)
    dev = mlr.Solver(mesh=mesh, num=num)
    m = oracle.Mesh("arrays", arrays=mesh)
    obcs = [(z, "wall_adiabatic" if k == "wall_noslip" else k, p) for z, k, p in bcs]
    cfg = oracle.make_config(order=2, teno_variant="corrected", riemann="Roe", bcs=obcs, zones=mesh["zone_names"],
                             **PHYS)
    return dev, oracle.Solver(m, cfg, 16, canonical=True), m, mesh


def test_config3_wedge_inviscid_against_oracle(mlr, oracle):
    """TENO + Roe on the graded Delaunay wedge mesh (30k cells) with the
    problem's boundary set: the residual and 20 steps equal the oracle's on
    the same mesh arrays (Roe takes no pow, so bit for bit is expected)."""
    dev, orc, m, mesh = _wedge_pair(mlr, oracle, WEDGE_BCS)
    dev.init_constant([2.5, 0.0], 1.0, TINF)
    orc.init_constant([2.5, 0.0], 1.0, TINF)
    U0 = orc.state()[0]
    assert np.array_equal(dev.state()[0], U0)
    r, ro = dev.calc_rhs(), orc.calc_rhs(U0)
    print("wedge residual: %.1f%% of entries bitwise, relative L-inf %.3g" % (100 * np.mean(r == ro), rel_linf(r, ro)))
    assert rel_linf(r, ro) <= 1e-12
    dts, dto = dev.run_steps(20), orc.run(20, 0.4)
    U, Uo = dev.state()[0], orc.state()[0]
    print("wedge 20 steps: dt max rel %.3g, state relative L-inf %.3g, bitwise %s"
          % (np.abs(dts / dto - 1).max(), rel_linf(U, Uo), np.array_equal(U, Uo)))
    assert np.abs(dts / dto - 1.0).max() <= 1e-12
    assert rel_linf(U, Uo) <= 1e-10
    # the wedge compresses the flow: oblique-shock pressure ratio 2.47 for
    # Mach 2.5 and 15 degrees, reached on the wedge surface after the shock forms
    assert np.isfinite(U).all()


def test_config3_wedge_navier_stokes_against_restatement(mlr, oracle):
    """The viscous wedge (TENO + Roe, Sutherland viscosity, adiabatic no-slip
    wedge surface): the residual equals the oracle's Euler residual plus the
    numpy viscous restatement, and 20 steps equal the reference's SSP-RK3
    sequence driven by that sum, with the GPU's dt values."""
    from oracle import viscous_ref as vr
    bcs = [b if b[0] != "wall" else ("wall", "wall_noslip", {}) for b in WEDGE_BCS]
    mu_ref, S, Pr = 1.4 * 2.5 / 1e4, 110.4 / 288.15 * TINF, 0.72
    dev, orc, m, mesh = _wedge_pair(mlr, oracle, bcs, physics="navier-stokes", mu_ref=mu_ref, T_mu_ref=TINF,
                                    sutherland_S=S, prandtl=Pr)
    md = dict(mesh)
    md.update(cell_coords=m.cell_coords, cell_volume=m.cell_volume, face_area=m.face_area,
              face_normals=m.face_normals, cells_of_face=m.cells_of_face)
    vbc = {z: (k, p) for z, k, p in bcs}

    def R(U):
        return orc.calc_rhs(U) + vr.viscous_residual(md, U, GAMMA, 1.0, mu_ref, TINF, S, Pr, bc=vbc)

    dev.init_constant([2.5, 0.0], 1.0, TINF)
    U = dev.state()[0]
    # a state with shear and heat conduction everywhere: two steps in
    U[:, 1] *= 1.0 + 0.05 * np.sin(7.0 * m.cell_coords[:, 1])
    U[:, 3] += 0.02 * np.cos(5.0 * m.cell_coords[:, 0])
    dev.set_state(U)
    r = dev.calc_rhs()
    print("NS wedge residual: relative L-inf %.3g" % rel_linf(r, R(U)))
    assert rel_linf(r, R(U)) <= 1e-12
    dts = dev.run_steps(20)
    for dt in dts:  # time_integrator.cpp:135-163
        k1 = R(U)
        Ut = U + dt * k1
        k2 = R(Ut)
        Ut = 3.0 / 4.0 * U + 1.0 / 4.0 * Ut
        Ut = Ut + dt / 4.0 * k2
        k3 = R(Ut)
        U = U + dt * (1.0 / 6.0) * k1
        U = U + dt * (1.0 / 6.0) * k2
        U = U + dt * (2.0 / 3.0) * k3
    Ud = dev.state()[0]
    print("NS wedge 20 steps: relative L-inf %.3g" % rel_linf(Ud, U))
    assert np.isfinite(Ud).all()
    assert rel_linf(Ud, U) <= 1e-10
