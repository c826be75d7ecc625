"""CPU oracle of the learned troubled-cell indicator.

TEST INFRASTRUCTURE ONLY: loaded by tests/ (and bench.py's cpu_baseline leg
never needs it). numpy restatement of the features, the 48-64-64-1 network
(bf16 weights and first-layer activations, FP32 accumulation, tanh GELU,
sigmoid) and the gradient of the weighted mean binary cross-entropy with the
same rounding points as the GPU kernels, so the GPU's forward and backward
passes can be checked against it. The indicator has no counterpart in the
reference; its labels are the reference's TENO cut-off decisions
(face_reconstruction.cpp:948-981: central stencil selected for all four
variables).
"""
import numpy as np

N_FEAT, N_HID, N_MEMBERS = 48, 64, 12
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_FEAT * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1
GELU_K = np.float32(0.7978845608028654)


def bf16_bits(x):
    """float32 -> bfloat16 bit patterns, round to nearest even (NaN stays NaN)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def bf16_value(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def bf16(x):
    return bf16_value(bf16_bits(x))


def _bits(x):
    return bf16_bits(np.asarray(x, dtype=np.float64).astype(np.float32))


def _ilogb(x):
    """Unbiased binary exponent of positive finite doubles (C ilogb)."""
    return np.frexp(x)[1] - 1


def features(U, central, gamma):
    """bf16 features [N][48] of conserved states U[N][4] (FP64) with central
    stencils central[N][12] (target first), reference numbering: per member
    1..11 the differences (rho_m ir_c - 1, (u_m - u_c) ia_c, (v_m - v_c) ia_c,
    p_m ip_c - 1), with ir = 1/rho, ia = 1/a, ip = 1/p and velocities and
    pressures formed with 1/rho, multiplied by the reciprocal of their
    stencil-wide maximum magnitude s_v per variable (0 for s_v = 0); member 0
    carries max(-2, ilogb(s_v)/32) (-2.5 for s_v = 0). Rows with a
    non-physical cell or member are zero (valid[i] False)."""
    U = np.asarray(U, dtype=np.float64)
    rho = U[:, 0]
    with np.errstate(all="ignore"):
        ir = 1.0 / rho
        u, v = U[:, 1] * ir, U[:, 2] * ir
        p = (gamma - 1.0) * (U[:, 3] - 0.5 * (U[:, 1] * U[:, 1] + U[:, 2] * U[:, 2]) * ir)
        ok = (rho > 0.0) & (p > 0.0)
        ia = np.where(ok, 1.0 / np.sqrt(gamma * p * ir), 1.0)
        ip = 1.0 / p
        D = np.zeros((len(U), N_MEMBERS, 4))
        for m in range(1, N_MEMBERS):
            j = central[:, m]
            D[:, m, 0] = rho[j] * ir - 1.0
            D[:, m, 1] = (u[j] - u) * ia
            D[:, m, 2] = (v[j] - v) * ia
            D[:, m, 3] = p[j] * ip - 1.0
            ok &= (rho[j] > 0.0) & (p[j] > 0.0) & np.isfinite(D[:, m]).all(axis=1)
        s = np.abs(D[:, 1:]).max(axis=1)
        pos = s > 0.0
        f0 = np.where(pos, np.maximum(-2.0, _ilogb(np.where(pos, s, 1.0)) / 32.0), -2.5)
        inv = np.where(pos, 1.0 / np.where(pos, s, 1.0), 0.0)
        pat = D * inv[:, None, :]
    X = np.zeros((len(U), N_FEAT), np.uint16)
    X[:, 0:4] = _bits(f0)
    X[:, 4:] = _bits(pat[:, 1:].reshape(len(U), -1))
    X[~ok] = 0
    return X


def valid_rows(U, central, gamma):
    """Rows features() fills (physical cell and members)."""
    return ~np.all(features(U, central, gamma) == 0, axis=1)


def unpack(theta, dtype=np.float32):
    t = np.asarray(theta, dtype=dtype)
    W1 = t[OFF_W1:OFF_B1].reshape(N_FEAT, N_HID)
    W2 = t[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    return W1, t[OFF_B1:OFF_W2], W2, t[OFF_B2:OFF_W3], t[OFF_W3:OFF_B3], t[OFF_B3]


def gelu(x):
    x = x.astype(np.float32)
    t = np.tanh(GELU_K * (x + np.float32(0.044715) * x * x * x))
    return np.float32(0.5) * x * (np.float32(1.0) + t)


def gelu_d(x):
    x = x.astype(np.float32)
    t = np.tanh(GELU_K * (x + np.float32(0.044715) * x * x * x))
    return (np.float32(0.5) * (1 + t) + np.float32(0.5) * x * (1 - t * t) * GELU_K *
            (1 + np.float32(3 * 0.044715) * x * x))


def forward(theta, Xbits, exact=False):
    """Logits z and the intermediates. Default: the GPU's rounding points
    (bf16 weights, inputs and first-layer activations; FP32 values; the
    accumulations in float64, which bounds the GPU's FP32 accumulation error).
    exact=True: everything in float64 with no rounding (X given as values),
    the smooth function whose derivatives loss_grad returns."""
    W1, b1, W2, b2, w3, b3 = unpack(theta, np.float64 if exact else np.float32)
    if exact:
        f64 = lambda a: np.asarray(a, dtype=np.float64)
        X = f64(Xbits)
        pre1 = X @ f64(W1) + f64(b1)
        h1 = gelu64(pre1)
        pre2 = h1 @ f64(W2) + f64(b2)
        h2 = gelu64(pre2)
        z = h2 @ f64(w3) + np.float64(b3)
        return z, dict(X=X, pre1=pre1, h1=h1, pre2=pre2, h2=h2, W2=f64(W2))
    X = bf16_value(Xbits).astype(np.float64)
    pre1 = (X @ bf16(W1).astype(np.float64) + b1).astype(np.float32)
    h1 = bf16(gelu(pre1))
    pre2 = (h1.astype(np.float64) @ bf16(W2).astype(np.float64) + b2).astype(np.float32)
    h2 = gelu(pre2)
    z = (h2.astype(np.float64) @ w3.astype(np.float64) + np.float64(b3)).astype(np.float32)
    return z, dict(X=X, pre1=pre1, h1=h1, pre2=pre2, h2=h2, W2=bf16(W2).astype(np.float64))


def gelu64(x):
    k = np.sqrt(2.0 / np.pi)
    return 0.5 * x * (1.0 + np.tanh(k * (x + 0.044715 * x ** 3)))


def gelu64_d(x):
    k = np.sqrt(2.0 / np.pi)
    t = np.tanh(k * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * k * (1 + 3 * 0.044715 * x * x)


def loss_grad(theta, Xbits, y, pos_w=1.0, neg_w=1.0, exact=False):
    """Weighted mean BCE (with logits) and its gradient w.r.t. the FP32
    master parameters (straight through the bf16 rounding of the weights)."""
    W1, b1, W2, b2, w3, b3 = unpack(theta, np.float64 if exact else np.float32)
    z, c = forward(theta, Xbits, exact)
    gd = gelu64_d if exact else (lambda a: gelu_d(a).astype(np.float64))
    z = z.astype(np.float64)
    y = np.asarray(y, dtype=np.float64)
    w = np.where(y > 0.5, pos_w, neg_w)
    B = len(y)
    loss = np.sum(w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))) / B
    p = 1.0 / (1.0 + np.exp(-z))
    dz = w * (p - y) / B
    g = np.zeros(N_PARAMS)
    h2 = c["h2"].astype(np.float64)
    g[OFF_W3:OFF_B3] = dz @ h2
    g[OFF_B3] = dz.sum()
    dpre2 = dz[:, None] * w3.astype(np.float64)[None, :] * gd(c["pre2"])
    g[OFF_W2:OFF_B2] = (c["h1"].astype(np.float64).T @ dpre2).ravel()
    g[OFF_B2:OFF_W3] = dpre2.sum(0)
    dh1 = dpre2 @ c["W2"].T
    dpre1 = dh1 * gd(c["pre1"])
    g[OFF_W1:OFF_B1] = (c["X"].T @ dpre1).ravel()
    g[OFF_B1:OFF_W2] = dpre1.sum(0)
    return loss, g
