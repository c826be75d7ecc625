"""HIP path vs the CPU oracle on the same meshes and states (runs on an MI355X).

Bar: bit-exact for first-order / corrected-TENO residuals and trajectories,
where the device and the oracle execute the same IEEE operations; within
1e-10 relative L-inf after 100 SSP-RK3 steps in every case. The reference
TENO variant's raw weights use x^6: the device computes it correctly
rounded, glibc's pow(x, 6.0) misrounds by 1 ulp in about 0.1% of arguments,
so those residuals are compared to 1e-12 relative instead of bitwise.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RIEMANN2D = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                      [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])
SOD = np.array([[0.125, 0.0, 0.0, 0.1], [1.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 1.0],
                [0.125, 0.0, 0.0, 0.1]])


def pair(mlr, oracle, kind, nx, ny, Lx=1.0, Ly=1.0, jitter=0.0, seed=0, bcs=None, threads=1, canonical=False,
         **kw):
    num = mlr.numerics(bcs=bcs, **kw)
    dev = mlr.Solver(gen=mlr.mesh_gen(kind, nx, ny, Lx, Ly, jitter, seed), num=num)
    m = oracle.Mesh(kind, nx, ny, Lx, Ly, jitter, seed)
    okw = {k: v for k, v in kw.items() if k not in ("cfl", "dt", "n_threads", "stencil_factor",
                                                        "quad_order_cell", "quad_order_face")}
    if "stencil_factor" in kw:
        okw["factor"] = kw["stencil_factor"]
    cfg = oracle.make_config(bcs=bcs, **okw)
    orc = oracle.Solver(m, cfg, threads, canonical=canonical)
    return dev, orc, m


def rel_linf(a, b):
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return float((np.abs(a - b).max(axis=0) / scale).max())


def test_device_math_matches_ieee(mlr):
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-30, 30, 200000))
    y = np.exp(rng.uniform(-30, 30, 200000)) * rng.choice([-1.0, 1.0], 200000)
    assert np.array_equal(mlr.probe_math(0, x), np.sqrt(x))
    assert np.array_equal(mlr.probe_math(1, x, y), x / y)
    # x^6: device == host correctly rounded routine, bitwise
    z = np.exp(rng.uniform(np.log(1e-12), np.log(1e4), 200000))
    assert np.array_equal(mlr.probe_math(3, z), mlr.host_pow6(z))
    # general pow (two-rarefaction estimate): the double-double pow is the
    # same bits on host and device and within 1 ulp of glibc, rarely off
    b = rng.uniform(0.01, 10.0, 20000)
    e = rng.uniform(0.1, 8.0, 20000)
    d = mlr.probe_math(2, b, e)
    assert np.array_equal(d, mlr.host_pow(b, e))
    g = np.array([math.pow(u, v) for u, v in zip(b, e)])
    ulps = np.abs(d.view(np.int64) - g.view(np.int64))
    assert ulps.max() <= 1
    assert np.mean(ulps > 0) < 0.005
    # special operands: the device takes the same case-by-case branches
    from test_capi import special_pow_operands
    sx, sy = special_pow_operands()
    sd, sh = mlr.probe_math(2, sx, sy), mlr.host_pow(sx, sy)
    assert np.array_equal(sd.view(np.int64), sh.view(np.int64)) or np.array_equal(sd, sh, equal_nan=True)
    assert (np.isnan(sd) == np.isnan(sh)).all() and (np.signbit(sd) == np.signbit(sh))[~np.isnan(sd)].all()
    native = mlr.probe_math(4, b, e)
    print("pow vs glibc: double-double %.4f%%, platform pow %.4f%% differ by 1 ulp"
          % (100.0 * np.mean(ulps > 0), 100.0 * np.mean(native != g)))


@pytest.mark.parametrize("riemann", ["HLLC", "HLL", "Rusanov", "Roe"])
def test_fo_quads_sod_bitexact(mlr, oracle, riemann):
    dev, orc, m = pair(mlr, oracle, "cartesian", 64, 3, 1.0, 0.01, recon="FO", riemann=riemann,
                       integrator="SSPRK3", cfl=0.5)
    dev.init_quadrant(0.5, -1.0, SOD)
    orc.init_quadrant(0.5, -1.0, SOD)
    U0, W0 = dev.state()
    Uo, Wo = orc.state()
    assert np.array_equal(U0, Uo) and np.array_equal(W0, Wo)
    assert np.array_equal(dev.calc_rhs(), orc.calc_rhs(Uo))
    dts = dev.run_steps(30)
    dto = orc.run(30, 0.5)
    assert np.array_equal(dts, dto)
    U1, W1 = dev.state()
    U2, W2 = orc.state()
    assert np.array_equal(U1, U2)
    assert np.array_equal(W1, W2)


@pytest.mark.parametrize("integrator", ["FE", "RK4", "SSPRK3"])
def test_fo_tri_integrators_bitexact(mlr, oracle, integrator):
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", 12, 10, jitter=0.2, seed=3, recon="FO",
                       integrator=integrator, cfl=0.4)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    dts = dev.run_steps(10)
    dto = orc.run(10, 0.4)
    assert np.array_equal(dts, dto)
    assert np.array_equal(dev.state()[0], orc.state()[0])


def test_fo_boundary_conditions_bitexact(mlr, oracle):
    bcs = [("left", "upt", {"u": (600.0, 0.0), "p": 101325.0, "T": 300.0}),
           ("right", "p_out", {"p": 101325.0}),
           ("top", "wall_adiabatic", {}),
           ("bottom", "extrapolation", {})]
    dev, orc, m = pair(mlr, oracle, "wedge", 30, 10, 4.0, 1.5, recon="FO", bcs=bcs, cfl=0.5,
                       p_min=1e-6)
    dev.init_constant([550.0, 10.0], 90000.0, 290.0)
    orc.init_constant([550.0, 10.0], 90000.0, 290.0)
    assert np.array_equal(dev.calc_rhs(), orc.calc_rhs(orc.state()[0]))
    assert np.array_equal(dev.run_steps(20), orc.run(20, 0.5))
    assert np.array_equal(dev.state()[0], orc.state()[0])


@pytest.mark.parametrize("order", [1, 2, 3])
def test_teno_reference_variant_residual(mlr, oracle, order):
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", 14, 12, jitter=0.2, seed=11, recon="TENO",
                       basis="legendre", order=order, teno_variant="reference")
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    Uo = orc.state()[0]
    r_dev = dev.calc_rhs()
    r_orc = orc.calc_rhs(Uo)
    nq = dev.Q
    fv_dev = dev.face_values()
    fv_orc = orc.face_values(nq)
    written = ~np.isnan(fv_dev)
    assert np.allclose(fv_dev[written], fv_orc[written], rtol=1e-12, atol=0)
    exact = np.mean(fv_dev[written] == fv_orc[written])
    assert exact > 0.99
    d_dev = dev.decisions()
    d_orc = orc.decisions()
    bits = (d_orc[:, 0] | (d_orc[:, 1] << 1) | (d_orc[:, 2] << 2) | (d_orc[:, 3] << 3)).astype(np.uint8)
    assert np.array_equal(d_dev, bits)
    # the reference weighting yields non-physical states (NaN residuals) at
    # the same cells on both sides
    assert np.array_equal(np.isnan(r_dev), np.isnan(r_orc))
    fin = np.isfinite(r_orc)
    scale = np.abs(r_orc[fin]).max()
    assert np.allclose(r_dev[fin], r_orc[fin], rtol=1e-12, atol=1e-12 * scale)


@pytest.mark.parametrize("order,factor", [(1, 2.0), (2, 2.0), (2, 3.0), (3, 2.0)])
def test_teno_corrected_residual_bitexact(mlr, oracle, order, factor):
    """p=2 with factor 2 runs the unrolled kernel with 8-bit tile positions;
    factor 3 and p=3 take the 16-bit tile form, p=3 also two face points."""
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", 20, 20, jitter=0.2, seed=7, recon="TENO",
                       order=order, teno_variant="corrected", stencil_factor=factor)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    r_dev = dev.calc_rhs()
    r_orc = orc.calc_rhs(orc.state()[0])
    assert np.array_equal(r_dev, r_orc)


def test_device_stencils_match_oracle(mlr, oracle):
    """Stencil ids read back from the uploaded tile layout (positions in each
    64-cell block's list) equal the oracle's stencils, in reference numbering."""
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", 24, 24, jitter=0.2, seed=11, recon="TENO",
                       order=2, teno_variant="corrected")
    T = orc.teno()
    og, os_, st = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"]
    for c in range(0, m.n_cells, 13):
        ids = dev.device_stencils(c)
        for sl in range(4):
            si = og[c] + sl
            ref = st[os_[si]:os_[si + 1]]
            got = ids[sl][ids[sl] != 0xFFFFFFFF]
            assert np.array_equal(got, ref), (c, sl)


def test_teno_corrected_100_steps(mlr, oracle):
    """Config-2 numerics (TENO p=2 + HLLC + SSPRK3, CFL 0.4) on a 32^2 jittered mesh."""
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", 32, 32, jitter=0.2, seed=2048, recon="TENO",
                       order=2, teno_variant="corrected", cfl=0.4)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    dts = dev.run_steps(100)
    dto = orc.run(100, 0.4)
    U1 = dev.state()[0]
    U2 = orc.state()[0]
    assert np.isfinite(U2).all()
    err = rel_linf(U1, U2)
    print("corrected TENO, 100 steps: rel L-inf %.3e, identical cells %.4f" % (err, np.mean(U1 == U2)))
    assert err <= 1e-10
    assert np.allclose(dts, dto, rtol=1e-12)


def test_teno_p3_corrected_50_steps(mlr, oracle):
    """The p = 3 secondary case (the reference example's basis order: K = 10,
    20-member stencils, two Gauss points per face, the 16-bit tile form) for
    50 steps of the four-quadrant problem on a 96^2 jittered mesh, against
    the oracle in its canonical order."""
    bcs = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", 96, 96, jitter=0.2, seed=3033, recon="TENO",
                       order=3, teno_variant="corrected", cfl=0.4, bcs=bcs, threads=8, canonical=True)
    assert (dev.K, dev.M, dev.Q) == (10, 20, 2)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    dts = dev.run_steps(50)
    dto = orc.run(50, 0.4)
    U1, U2 = dev.state()[0], orc.state()[0]
    assert np.isfinite(U2).all()
    err = rel_linf(U1, U2)
    print("corrected TENO p=3, 50 steps: rel L-inf %.3e, identical cells %.4f" % (err, np.mean(U1 == U2)))
    assert err <= 1e-10
    assert np.allclose(dts, dto, rtol=1e-12)


def test_teno_smooth_field_steps(mlr, oracle):
    """Smooth field: corrected variant for 40 steps. With the reference
    weighting the first residual is finite and must match, but the oracle's
    first SSP-RK3 step already ends non-finite, so no steps are compared."""
    bcs = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]
    for variant, n in (("corrected", 40), ("reference", 1)):
        dev, orc, m = pair(mlr, oracle, "cartesian_tri", 24, 24, jitter=0.1, seed=5, recon="TENO",
                           order=2, teno_variant=variant, cfl=0.4, bcs=bcs)
        x, y = m.cell_coords[:, 0], m.cell_coords[:, 1]
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
        u, v, p = 0.3 + 0 * x, -0.2 + 0 * x, 1.0 + 0.1 * np.cos(2 * np.pi * x)
        R = 101325.0 / (298.15 * 1.225)
        cv = R * 1.4 / 0.4 / 1.4
        e = cv * (p / (rho * R))
        U = np.stack([rho, rho * u, rho * v, rho * (e + 0.5 * (u * u + v * v))], axis=1)
        dev.set_state(U)
        orc.set_state(U)
        if variant == "reference":
            r1, r2 = dev.calc_rhs(), orc.calc_rhs(U)
            assert np.array_equal(np.isnan(r1), np.isnan(r2))
            fin = np.isfinite(r2)
            assert np.allclose(r1[fin], r2[fin], rtol=1e-12, atol=1e-12 * np.abs(r2[fin]).max())
            continue
        dev.run_steps(n)
        orc.run(n, 0.4)
        U1, U2 = dev.state()[0], orc.state()[0]
        assert np.isfinite(U2).all()
        err = rel_linf(U1, U2)
        print("%s TENO smooth, %d steps: rel L-inf %.3e" % (variant, n, err))
        assert err <= 1e-10


def test_run_steps_graph_equals_eager(mlr):
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4)
    gen = mlr.mesh_gen("cartesian_tri", 20, 20, jitter=0.2, seed=9)
    a = mlr.Solver(gen=gen, num=num)
    b = mlr.Solver(gen=gen, num=num)
    a.init_quadrant(0.5, 0.5, RIEMANN2D)
    b.init_quadrant(0.5, 0.5, RIEMANN2D)
    a.run_steps(12)
    for _ in range(12):
        b.calc_dt()
        b.take_step()
    assert np.array_equal(a.state()[0], b.state()[0])
    assert a.time() == b.time()


def test_determinism(mlr):
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4)
    gen = mlr.mesh_gen("cartesian_tri", 24, 24, jitter=0.2, seed=1)
    out = []
    for _ in range(2):
        s = mlr.Solver(gen=gen, num=num)
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        s.run_steps(20)
        out.append(s.state()[0])
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("order", [1, 2, 3])
def test_gpu_precompute_matches_host(mlr, order):
    """The pseudo-inverses and member areas the GPU builds at upload equal
    the host precompute's bit for bit (Legendre basis, factor 2)."""
    gen = mlr.mesh_gen("cartesian_tri", 18, 14, 1.0, 1.0, 0.2, 21)
    num = mlr.numerics(recon="TENO", order=order, teno_variant="corrected")
    dev = mlr.Solver(gen=gen, num=num)
    L = mlr.Layout(gen, num)
    for c in range(0, L.n_cells, 3):
        a_d, t_d = dev.device_matrices(c)
        a_h, t_h, _ = L.cell_matrices(c)
        assert np.array_equal(a_d, a_h), c
        assert np.array_equal(t_d[:, 1:], t_h[:, 1:]), c


def test_fixed_dt_step_equals_cfl_step(mlr, oracle):
    """A fixed dt (cfl = 0, dt given; solver.cpp:580-590 skips the spectral
    radius) takes exactly that dt every step, and its first step equals the
    CFL run's first step when given the dt that run chose."""
    dev, orc, m = pair(mlr, oracle, "cartesian_tri", 12, 10, jitter=0.2, seed=3, recon="TENO", order=2,
                       teno_variant="corrected", cfl=0.4)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    orc.init_quadrant(0.5, 0.5, RIEMANN2D)
    dt0 = dev.run_steps(1)[0]
    assert dt0 == orc.run(1, 0.4)[0]
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.0, dt=dt0)
    fx = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 12, 10, 1.0, 1.0, 0.2, 3), num=num)
    fx.init_quadrant(0.5, 0.5, RIEMANN2D)
    dts = fx.run_steps(1)
    assert dts[0] == dt0
    assert np.array_equal(fx.state()[0], dev.state()[0])
    dts = fx.run_steps(4)
    assert np.all(dts == dt0)
    t, steps = fx.time()
    assert steps == 5 and abs(t - 5 * dt0) <= 1e-15


def test_negative_or_zero_fixed_dt_refused(mlr):
    """A negative fixed dt, a zero one, and a negative CFL number (which the
    C ABI takes as CFL mode, so the reference's negative-dt error applies,
    solver.cpp:580-590) are refused before any step, also when a fixed dt is
    given beside the negative CFL number."""
    gen = mlr.mesh_gen("cartesian_tri", 6, 6, 1.0, 1.0, 0.0, 0)
    for cfl, dt, msg in ((0.0, -1e-3, "dt negative"), (0.0, 0.0, "fixed dt must be positive"),
                         (-0.5, 1e-3, "dt negative"), (-0.5, 0.0, "dt negative")):
        s = mlr.Solver(gen=gen, num=mlr.numerics(recon="FO", cfl=cfl, dt=dt))
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        U0 = s.state()[0]
        with pytest.raises(mlr.MallardError, match=msg):
            s.run_steps(1)
        assert s.time() == (0.0, 0) and np.array_equal(s.state()[0], U0)


def test_dt_history_ring_wraps(mlr):
    """Runs longer than the device's dt ring are issued in chunks that each
    read the ring back before the next overwrites it: with a 7-entry ring,
    20 steps in one call, then calls of 3 and 11 steps that start and end
    inside the ring, return the same dt values, time and state as a solver
    with the default ring."""
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4)
    gen = mlr.mesh_gen("cartesian_tri", 12, 10, 1.0, 1.0, 0.2, 5)
    a = mlr.Solver(gen=gen, num=num)
    b = mlr.Solver(gen=gen, num=num)
    b.set_dt_history(7)
    for s in (a, b):
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
    dts = []
    for n in (20, 3, 11):
        da = a.run_steps(n)
        db = b.run_steps(n)
        assert np.array_equal(da, db), n
        assert a.time() == b.time()
        dts.extend(db)
    assert b.time()[1] == 34 and len(set(dts)) > 1
    t = 0.0
    for d in dts:  # the host's running sum, step by step
        t += d
    assert b.time()[0] == t
    assert np.array_equal(a.state()[0], b.state()[0])
    with pytest.raises(mlr.MallardError, match="dt history"):
        b.set_dt_history(0)
