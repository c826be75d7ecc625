"""Writes tests/golden/riemann2048_corrected_100.npz: the CPU oracle's run of
the full-size four-quadrant Riemann workload, for the GPU parity test at that
size (tests/test_gpu_full_size.py).

Workload: 2048 x 2048 quads split into 8,388,608 triangles, interior nodes
jittered by +/-0.2h (seed 2024); TENO p = 2 (Legendre, stencil factor 2,
corrected weighting), HLLC, SSP-RK3, CFL 0.4, extrapolation boundaries on the
four sides; 100 steps from the quadrant state (UR (1.5, 0, 0, 1.5), UL
(0.5323, 1.206, 0, 0.3), LL (0.138, 1.206, 1.206, 0.029), LR (0.5323, 0,
1.206, 0.3)).

The fixture holds the 100 dt values, the final conserved state of a seeded
sample of cells (10,000 drawn uniformly plus each variable's extreme cells),
and whole-field norms per variable (max |U|, sum |U| V, sum U V, min U).

With --symmetry the four sides are symmetry walls (the boundary set the
survey's configuration names) and the positivity-preserving flux limiter is
on, which that set needs to stay physical for 100 steps; the default output
is then tests/golden/riemann2048_symmetry_pp_100.npz.

The oracle runs with OpenMP in its canonical mode: each cell gathers its
face contributions in the order of a one-thread run, so the fixture does not
depend on the thread count (about 30 GB of host memory; minutes on 16
cores). Usage:
    python tests/golden/make_riemann2048_fixture.py [out.npz] [--threads N] [--nx N] [--symmetry]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

QUADRANT = np.array([[1.5, 0.0, 0.0, 1.5], [0.5323, 1.206, 0.0, 0.3],
                     [0.138, 1.206, 1.206, 0.029], [0.5323, 0.0, 1.206, 0.3]])
NX, JITTER, SEED, STEPS, CFL, N_SAMPLE = 2048, 0.2, 2024, 100, 0.4, 10000


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out", nargs="?", default=None)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--nx", type=int, default=NX)
    ap.add_argument("--symmetry", action="store_true",
                    help="symmetry walls on the four sides, with the positivity limiter")
    args = ap.parse_args()
    if args.out is None:
        args.out = os.path.join(HERE, "riemann2048_symmetry_pp_100.npz" if args.symmetry
                                else "riemann2048_corrected_100.npz")
    import oracle_py as orc
    threads = args.threads or int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1
    t0 = time.perf_counter()
    m = orc.Mesh("cartesian_tri", args.nx, args.nx, 1.0, 1.0, JITTER, SEED)
    kind = "symmetry" if args.symmetry else "extrapolation"
    bcs = [(z, kind, {}) for z in ("left", "right", "top", "bottom")]
    s = orc.Solver(m, orc.make_config(order=2, teno_variant="corrected", bcs=bcs, positivity=args.symmetry),
                   n_threads=threads, canonical=True)
    s.init_quadrant(0.5, 0.5, QUADRANT)
    print(f"setup {time.perf_counter() - t0:.1f} s, {m.n_cells} cells, {threads} threads", flush=True)
    dts = np.zeros(STEPS)
    t1 = time.perf_counter()
    for k in range(0, STEPS, 10):
        dts[k:k + 10] = s.run(10, CFL)
        print(f"step {k + 10}: {time.perf_counter() - t1:.1f} s", flush=True)
    U, _ = s.state()
    if not np.isfinite(U).all():
        raise RuntimeError("oracle state is not finite")
    V = m.cell_volume
    rng = np.random.default_rng(20260)
    ids = set(rng.choice(m.n_cells, min(N_SAMPLE, m.n_cells), replace=False).tolist())
    for v in range(4):
        ids.add(int(np.argmin(U[:, v])))
        ids.add(int(np.argmax(U[:, v])))
    ids = np.array(sorted(ids), dtype=np.uint32)
    np.savez_compressed(
        args.out, ids=ids, U=U[ids], dts=dts,
        linf=np.abs(U).max(axis=0), l1=(np.abs(U) * V[:, None]).sum(axis=0),
        total=(U * V[:, None]).sum(axis=0), umin=U.min(axis=0),
        meta=np.array([args.nx, JITTER, SEED, STEPS, CFL, threads], dtype=np.float64))
    print(f"wrote {args.out}: {len(ids)} sampled cells, t = {dts.sum():.6g}, "
          f"{s.pp_limited()} limited faces, {time.perf_counter() - t0:.1f} s in all", flush=True)


if __name__ == "__main__":
    main()
