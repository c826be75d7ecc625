"""CPU checks of the indicator oracle (oracle/indicator_ref.py): bf16
rounding against torch's, the backward formulas against central finite
differences of the exact (unrounded) network, and the features' closed forms."""
import numpy as np
import pytest

from oracle import indicator_ref as ir


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 20000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0, 3.3895e38], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ir.bf16_bits(x), ref)
    # ties go to even: 1 + 2^-8 -> 1.0, 1 + 3*2^-8 -> 1 + 2^-6... (bit patterns)
    assert ir.bf16_bits(np.float32(1.0 + 2 ** -8))[()] == 0x3F80


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = rng.standard_normal(ir.N_PARAMS) * 0.2
    X = rng.standard_normal((37, ir.N_FEAT))
    y = (rng.random(37) > 0.4).astype(np.float64)
    loss, g = ir.loss_grad(theta, X, y, pos_w=1.0, neg_w=3.0, exact=True)
    picks = np.concatenate([rng.choice(ir.OFF_B1, 12, replace=False),
                            ir.OFF_B1 + rng.choice(64, 4, replace=False),
                            ir.OFF_W2 + rng.choice(4096, 12, replace=False),
                            ir.OFF_B2 + rng.choice(64, 4, replace=False),
                            ir.OFF_W3 + rng.choice(64, 4, replace=False), [ir.OFF_B3]])
    h = 1e-6
    for p in picks:
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (ir.loss_grad(tp, X, y, 1.0, 3.0, exact=True)[0] - ir.loss_grad(tm, X, y, 1.0, 3.0, exact=True)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(g).max()), (p, fd, g[p])


def test_features_closed_forms():
    g = 1.4
    # cell 0 and its 11 members: members 1..5 equal, members 6..11 denser
    # (same velocity and pressure)
    U = np.zeros((12, 4))
    rho, u, v, p = 1.0, 0.5, -0.25, 1.0   # binary fractions: velocity ratios are exact
    U[:] = [rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)]
    U[6:, 0] *= 1.5
    U[6:, 1:3] *= 1.5
    U[6:, 3] = p / (g - 1) + 0.5 * 1.5 * rho * (u * u + v * v)
    central = np.tile(np.arange(12), (12, 1))
    X = ir.bf16_value(ir.features(U, central, g)).astype(np.float64)
    # scale slots: rho differences peak at 0.5 = 2^-1 -> ilogb -1 -> -1/32;
    # velocity differences vanish -> -2.5; pressure ones are rounding noise
    assert X[0, 0] == -1 / 32 and X[0, 1] == -2.5 and X[0, 2] == -2.5
    assert X[0, 3] <= -40 / 32
    assert np.all(X[0, 4:24] == 0)                                # identical members
    assert np.all(X[0, 24::4] == 1.0)                             # pattern = d / max|d|
    assert np.all(X[0, 25::4] == 0) and np.all(X[0, 26::4] == 0)
    # a uniform stencil has the zero pattern and the -2.5 scale everywhere
    Uu = np.tile(U[0], (12, 1))
    Xu = ir.bf16_value(ir.features(Uu, central, g)).astype(np.float64)
    assert np.all(Xu[:, :4] == -2.5) and np.all(Xu[:, 4:] == 0)
    # scale-free: the same pattern at amplitude 1e-6 differs only in the scale slots
    Us = U.copy()
    Us[6:, 0] = rho * (1 + 1e-6)
    Us[6:, 1:3] = Us[6:, 0:1] * np.array([u, v])
    Us[6:, 3] = p / (g - 1) + 0.5 * Us[6:, 0] * (u * u + v * v)
    Xs = ir.bf16_value(ir.features(Us, central, g)).astype(np.float64)
    assert np.array_equal(Xs[0, 24::4], X[0, 24::4]) and Xs[0, 0] == -20 / 32
    # a non-physical member empties the row
    U2 = U.copy()
    U2[7, 3] = -1.0
    assert np.all(ir.features(U2, central, g)[0] == 0)
