"""Learned troubled-cell indicator on the GPU against the CPU oracle
(oracle/indicator_ref.py, oracle.cpp's TENO decisions).

Bars: features and labels bit-exact; logits within 1e-3 (FP32 accumulation
order of the matrix cores and bf16 re-rounding of hidden activations differ
from the oracle's float64 sums); gradients within 2e-3 of the largest entry;
gating with a network that never fires is bit-identical to plain TENO."""
import numpy as np
import pytest

from oracle import indicator_ref as ir

pytestmark = pytest.mark.gpu

RIEMANN2D = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                      [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])
EXTRAP = [(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")]


def _pair(mlr, oracle, n, seed=5):
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", n, n, 1.0, 1.0, 0.2, seed), num=num)
    m = oracle.Mesh("cartesian_tri", n, n, 1.0, 1.0, 0.2, seed)
    orc = oracle.Solver(m, oracle.make_config(order=2, teno_variant="corrected", bcs=EXTRAP), 1)
    return dev, orc, m


def _central(orc, nc):
    T = orc.teno()
    og, os_, st = T["offsets_stencil_groups"], T["offsets_stencils"], T["stencils"]
    return np.stack([st[os_[og[c]]:os_[og[c]] + 12] for c in range(nc)]).astype(np.int64)


def _small_dataset(mlr, n_cases=24, nx=32, seed=3):
    from mallard_amd import indicator as ind
    return ind.build_dataset(n_cases=n_cases, nx=nx, snapshots=(0, 6, 15), seed=seed)


def test_features_and_labels_match_oracle(mlr, oracle):
    dev, orc, m = _pair(mlr, oracle, 24)
    dev.init_quadrant(0.5, 0.5, RIEMANN2D)
    dev.run_steps(5)
    U, _ = dev.state()
    X, y = dev.indicator_features()
    Xr = ir.features(U, _central(orc, m.n_cells), 1.4)
    assert np.array_equal(X, Xr)
    orc.set_state(U)
    orc.calc_rhs(U)
    d = orc.decisions()
    valid = ~np.all(Xr == 0, axis=1)
    yr = np.where(valid, (d.sum(axis=1) == 4).astype(np.float32), -1.0).astype(np.float32)
    assert np.array_equal(y, yr)
    assert (y == 1).any() and (y == 0).any()  # both classes present


def test_mlp_forward_matches_oracle(mlr):
    from mallard_amd import indicator as ind
    rng = np.random.default_rng(2)
    th = ind.init_params(4)
    th[ir.OFF_B1:ir.OFF_W2] = rng.normal(0, 0.1, 64)
    th[ir.OFF_B2:ir.OFF_W3] = rng.normal(0, 0.1, 64)
    th[ir.OFF_B3] = 0.3
    X = ir.bf16_bits(rng.normal(0, 1, (1000, 48)).astype(np.float32))  # not a multiple of 16 rows
    z = mlr.mlp_forward(th, X)
    zr, _ = ir.forward(th, X)
    err = np.abs(z - zr)
    assert err.max() <= 1e-3 * (1 + np.abs(zr).max())
    assert np.median(err) <= 1e-5 * (1 + np.abs(zr).max())


def test_gradients_match_oracle(mlr):
    from mallard_amd import indicator as ind
    ds, y = _small_dataset(mlr, n_cases=6, nx=24)
    X, yy = ds.get()
    rng = np.random.default_rng(9)
    idx = rng.choice(len(yy), 700, replace=False).astype(np.uint32)  # 700: a partial last block
    th = ind.init_params(11)
    th[ir.OFF_B3] = -0.2
    g, loss = mlr.mlp_grad(th, ds, idx, pos_w=1.0, neg_w=4.0)
    lr_, gr = ir.loss_grad(th, X[idx], yy[idx], pos_w=1.0, neg_w=4.0)
    assert abs(loss - lr_) <= 1e-4 * max(1.0, abs(lr_))
    scale = np.abs(gr).max()
    assert np.abs(g - gr).max() <= 2e-3 * scale, np.abs(g - gr).max() / scale


def test_training_is_deterministic_and_learns(mlr):
    from mallard_amd import indicator as ind
    ds, y = _small_dataset(mlr)
    pool = ind.make_pool(y, seed=1)
    p1, h1 = ind.train_indicator(ds, pool, steps=400, batch=1024, lr=3e-3, seed=5)
    p2, h2 = ind.train_indicator(ds, pool, steps=400, batch=1024, lr=3e-3, seed=5)
    assert np.array_equal(p1, p2) and np.array_equal(h1, h2)
    assert h1[-50:].mean() < 0.6 * h1[:50].mean()


@pytest.mark.parametrize("every_stage", [False, True])
def test_gate_that_never_fires_is_plain_teno(mlr, every_stage):
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4)
    gen = mlr.mesh_gen("cartesian_tri", 40, 40, 1.0, 1.0, 0.2, 8)
    a, b = mlr.Solver(gen=gen, num=num), mlr.Solver(gen=gen, num=num)
    for s in (a, b):
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
    th = np.zeros(mlr.N_PARAMS, np.float32)
    th[ir.OFF_B3] = -50.0  # p = sigmoid(-50) < 0.9 everywhere
    b.set_indicator(th, 0.9, every_stage=every_stage)
    a.run_steps(20)
    b.run_steps(20)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert b.gated() == 0


@pytest.mark.parametrize("order", [2, 3])
def test_gate_that_always_fires_on_smooth_field_is_plain_teno(mlr, order):
    """Where plain TENO selects the central stencil everywhere, forcing it
    changes nothing but rounding: the gated cells evaluate the central
    stencil through its precombined linear operator (face-point basis x A+ x
    areas) instead of coefficients and basis. Every cell of every pass is
    counted, and a uniform state stays bit-identical. p = 2 runs the split
    kernels, p = 3 (two face points per face, 20-member stencils) the
    one-kernel runtime-loop build."""
    num = mlr.numerics(recon="TENO", order=order, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    gen = mlr.mesh_gen("cartesian_tri", 32, 32, 1.0, 1.0, 0.2, 12)
    a, b = mlr.Solver(gen=gen, num=num), mlr.Solver(gen=gen, num=num)
    th = np.zeros(mlr.N_PARAMS, np.float32)
    th[ir.OFF_B3] = 50.0
    b.set_indicator(th, 0.9)
    # uniform: the first stage sees exactly equal members
    for s in (a, b):
        s.init_constant([0.4, -0.3], 1.0e5, 290.0)
    assert np.array_equal(a.calc_rhs(), b.calc_rhs())
    # smooth: plain TENO takes the central stencil in every cell
    lay = mlr.Layout(gen, mlr.numerics(recon="FO", bcs=EXTRAP))
    cc = lay.mesh_arrays()["cell_coords"]
    rho = 1.0 + 0.01 * np.sin(2 * np.pi * cc[:, 0]) * np.cos(2 * np.pi * cc[:, 1])
    U = np.stack([rho, 0.3 * rho, -0.2 * rho, 1.0 / 0.4 + 0.5 * rho * 0.13], axis=1)
    for s in (a, b):
        s.set_state(U)
    a.run_steps(5)
    b.gated(reset=True)
    b.run_steps(5)
    assert (a.decisions() == 0xF).all()
    Ua, Ub = a.state()[0], b.state()[0]
    assert np.abs(Ua - Ub).max() <= 1e-13 * np.abs(Ua).max()
    assert b.gated() == a.n_cells * 3 * 5


@pytest.mark.parametrize("every_stage", [False, True])
def test_trained_indicator_gains_at_matched_l1(mlr, every_stage):
    """The trained network skips the directional stencils in most cells of a
    Riemann problem it was not trained on, and the solution stays within
    1e-3 relative L1 of plain TENO per variable, whether it judges the cells
    once per step (the default) or on every stage."""
    from mallard_amd import indicator as ind
    ds, y = ind.build_dataset(n_cases=64, nx=48, snapshots=(0, 8, 20), seed=21)
    params, hist = ind.train_indicator(ds, ind.make_pool(y, seed=2), steps=1500, batch=2048, lr=2e-3, seed=3)
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    gen = mlr.mesh_gen("cartesian_tri", 128, 128, 1.0, 1.0, 0.2, 77)
    a, b = mlr.Solver(gen=gen, num=num), mlr.Solver(gen=gen, num=num)
    for s in (a, b):
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
    b.set_indicator(params, 0.9, every_stage=every_stage)
    a.run_steps(100)
    b.run_steps(100)
    Ua, Ub = a.state()[0], b.state()[0]
    l1 = ind.rel_l1(Ub, Ua, a.volumes())
    frac = b.gated() / (a.n_cells * 3 * 100)
    print("gated fraction %.3f, rel L1 per variable %s" % (frac, np.array2string(l1, precision=2)))
    assert np.isfinite(Ub).all()
    assert frac > 0.5
    assert l1.max() <= 1e-3


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("every_stage", [False, True])
def test_gate_never_firing_is_plain_at_each_order(mlr, order, every_stage):
    """The gated kernels of every stencil size (p = 2: 12 members, the
    unrolled kernel; p = 3: 20 members, the runtime-loop kernel) leave the
    solution bit-identical to plain TENO when the network never fires."""
    num = mlr.numerics(recon="TENO", order=order, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    gen = mlr.mesh_gen("cartesian_tri", 24, 24, 1.0, 1.0, 0.2, 19)
    a, b = mlr.Solver(gen=gen, num=num), mlr.Solver(gen=gen, num=num)
    for s in (a, b):
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
    th = np.zeros(mlr.N_PARAMS, np.float32)
    th[ir.OFF_B3] = -50.0
    b.set_indicator(th, 0.9, every_stage=every_stage)
    a.run_steps(6)
    b.run_steps(6)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert b.gated() == 0


@pytest.mark.parametrize("integrator,every_stage", [("SSPRK3", False), ("SSPRK3", True), ("RK4", False),
                                                    ("FE", False)])
def test_split_gated_passes_equal_one_kernel(mlr, integrator, every_stage):
    """Gated passes split in two kernels (the gated cells; then the ungated
    cells of the blocks holding any) give the one-kernel results bit for
    bit: time steps, states, TENO decision bits, face values and gated
    counts, on a mesh whose last 64-cell block is partial and a network that
    gates about half of the cells, so most waves hold both kinds."""
    rng = np.random.default_rng(4)
    th = rng.normal(0.0, 0.2, mlr.N_PARAMS).astype(np.float32)
    th[ir.OFF_B3] = 2.2
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP, integrator=integrator)
    gen = mlr.mesh_gen("cartesian_tri", 37, 37, 1.0, 1.0, 0.2, 23)
    stages = {"SSPRK3": 3, "RK4": 4, "FE": 1}[integrator]
    out = []
    for split in (False, True):
        s = mlr.Solver(gen=gen, num=num)
        s.init_quadrant(0.5, 0.5, RIEMANN2D)
        s.set_indicator(th, 0.9, every_stage=every_stage, split=split)
        dts = s.run_steps(7)
        g = s.gated()
        out.append((s.state()[0], s.decisions(), s.face_values(), g, dts, s.cfl_local()))
        s.close()
    assert 0 < out[0][3] < 37 * 37 * 2 * stages * 7
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2], equal_nan=True)  # NaN: the missing side of boundary faces
    assert out[0][3] == out[1][3]
    assert np.array_equal(out[0][4], out[1][4])
    assert np.array_equal(out[0][5], out[1][5])


@pytest.mark.parametrize("split", [False, True])
def test_fused_gate_decisions_equal_feature_and_mlp_kernels(mlr, split):
    """The network evaluated inside the TENO pass (features from the LDS
    tile, the network staged in LDS) decides exactly as the standalone
    feature kernel followed by the inference kernel: gated where the logit
    reaches logit(0.9) and the stencil is physical, for every cell of a
    developed Riemann state."""
    rng = np.random.default_rng(9)
    th = rng.normal(0.0, 0.2, mlr.N_PARAMS).astype(np.float32)
    th[ir.OFF_B3] = 0.0
    num = mlr.numerics(recon="TENO", order=2, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    s = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 45, 45, 1.0, 1.0, 0.2, 31), num=num)
    s.init_quadrant(0.5, 0.5, RIEMANN2D)
    s.run_steps(12)
    X, y = s.indicator_features()
    thr = np.float32(np.log(0.9 / 0.1))
    # the output bias that puts the threshold at the median logit of the physical cells
    z0 = mlr.mlp_forward(th, X)
    th[ir.OFF_B3] = thr - np.median(z0[y >= 0])
    z = mlr.mlp_forward(th, X)
    want = ((z >= thr) & (y >= 0)).astype(np.uint8)
    s.set_indicator(th, 0.9, split=split)
    s.calc_rhs()
    got = s.gate_decisions()
    assert 0.05 < want.mean() < 0.95
    assert np.array_equal(got, want)


def test_gate_needs_a_central_stencil_of_twelve(mlr):
    num = mlr.numerics(recon="TENO", order=1, teno_variant="corrected", cfl=0.4, bcs=EXTRAP)
    s = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", 8, 8, 1.0, 1.0, 0.2, 1), num=num)
    with pytest.raises(mlr.MallardError, match="central stencils of at least 12 cells"):
        s.set_indicator(np.zeros(mlr.N_PARAMS, np.float32), 0.9)
