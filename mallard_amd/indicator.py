"""Training and evaluation of the learned troubled-cell indicator.

Everything numeric runs on the GPU through the C ABI: the solver produces
states and plain-TENO labels, the feature kernel writes the training rows, and
the network is trained by the library's own forward/backward kernels (Adam,
FP32 master weights, BF16 MFMA forward). This module only orchestrates.

Data: seeded random four-quadrant Riemann problems (rho, p ~ U[0.1, 2],
u, v ~ U[-1.5, 1.5] per quadrant) on a jittered triangle mesh, sampled at a
few step counts; a row is labelled smooth when plain TENO selects the central
stencil for all four conserved variables.
"""
import numpy as np

from . import Dataset, Solver, mesh_gen, mlp_train, numerics, N_PARAMS


def random_quadrant_states(rng):
    """(4, 4) array of (rho, u, v, p) for the UR, UL, LL, LR quadrants."""
    s = np.empty((4, 4))
    s[:, 0] = rng.uniform(0.1, 2.0, 4)
    s[:, 1] = rng.uniform(-1.5, 1.5, 4)
    s[:, 2] = rng.uniform(-1.5, 1.5, 4)
    s[:, 3] = rng.uniform(0.1, 2.0, 4)
    return s


def init_params(seed=0):
    """He-uniform weights, zero biases (FP32 master vector)."""
    rng = np.random.default_rng(seed)
    th = np.zeros(N_PARAMS, np.float32)
    o = 0
    for fan_in, fan_out in ((48, 64), (64, 64)):
        lim = np.sqrt(6.0 / fan_in)
        th[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    th[o:o + 64] = rng.uniform(-np.sqrt(6.0 / 64), np.sqrt(6.0 / 64), 64)
    return th


def build_dataset(n_cases=512, nx=64, snapshots=(0, 8, 20), seed=1234, jitter=0.2, order=2,
                  variant="corrected", cfl=0.4):
    """Rows of every cell of every case at the given step counts (plain TENO
    decisions as labels). Returns the device Dataset and the host labels."""
    num = numerics(recon="TENO", order=order, teno_variant=variant, cfl=cfl,
                   bcs=[(z, "extrapolation", {}) for z in ("left", "right", "top", "bottom")])
    dev = Solver(gen=mesh_gen("cartesian_tri", nx, nx, 1.0, 1.0, jitter, seed), num=num)
    ds = Dataset(n_cases * len(snapshots) * dev.n_cells)
    rng = np.random.default_rng(seed)
    for _ in range(n_cases):
        dev.init_quadrant(rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), random_quadrant_states(rng))
        done = 0
        for k in snapshots:
            if k > done:
                dev.run_steps(k - done)
                done = k
            # a case whose expansion reaches vacuum stops contributing rows
            U, W = dev.state()
            if not (np.isfinite(U).all() and (U[:, 0] > 0).all() and (W[:, 2] > 0).all()):
                break
            dev.collect_features(ds)
    _, y = ds.get(0, len(ds), features=False)
    return ds, y


def make_pool(y, pos_per_neg=2.0, seed=0):
    """All non-smooth rows plus a seeded sample of smooth ones (rows without
    features, label -1, are left out)."""
    rng = np.random.default_rng(seed)
    neg = np.flatnonzero((y >= 0.0) & (y < 0.5))
    pos = np.flatnonzero(y >= 0.5)
    k = min(len(pos), int(pos_per_neg * max(len(neg), 1)))
    pos = rng.choice(pos, k, replace=False) if k < len(pos) else pos
    return np.sort(np.concatenate([neg, pos])).astype(np.uint32)


def train_indicator(ds, pool, steps=3000, batch=4096, lr=2e-3, seed=7, pos_w=1.0, neg_w=4.0):
    """Adam on the GPU from init_params(seed). Non-smooth rows weigh more:
    calling a rough cell smooth costs accuracy, the converse only speed."""
    return mlp_train(init_params(seed), ds, pool, steps, batch, lr, seed=seed, pos_w=pos_w, neg_w=neg_w)


def rel_l1(Ua, Ub, vol):
    """Per-variable volume-weighted L1 difference relative to the L1 norm of Ub."""
    w = vol[:, None]
    return (np.abs(Ua - Ub) * w).sum(0) / np.maximum((np.abs(Ub) * w).sum(0), 1e-300)
