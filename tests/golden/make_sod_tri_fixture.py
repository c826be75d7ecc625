"""Writes tests/golden/sod_tri_teno_t02.npz: the CPU oracle's run of the Sod
shock tube on triangles with TENO, for the GPU comparison in
tests/test_gpu_configs.py.

Workload: cartesian_tri 1000 x 10 on [0, 1] x [0, 0.01] (20,000 triangles,
no jitter), left state (rho, u, v, p) = (1, 0, 0, 1) for x < 0.5, right
(0.125, 0, 0, 0.1); TENO p = 2 (Legendre, corrected weighting), HLLC,
SSP-RK3, CFL 0.5, symmetry on the four sides; steps until t >= 0.2 (dt is
not clipped at t_stop, solver.cpp:375-403). The reference cannot run TENO on
the quads of its own Sod example (face_reconstruction.cpp:485-487), so the
triangle mesh of the same size is used.

The fixture holds every step's dt, the final state of the 2,000 cells of the
middle row (0.005 <= y < 0.006) and whole-field norms. The oracle runs in its
canonical mode (one-thread face order at any thread count). Usage:
    python tests/golden/make_sod_tri_fixture.py [out.npz] [--threads N]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SOD = np.array([[0.125, 0.0, 0.0, 0.1], [1.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]])
NX, NY, LX, LY, CFL, T_STOP = 1000, 10, 1.0, 0.01, 0.5, 0.2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out", nargs="?", default=os.path.join(HERE, "sod_tri_teno_t02.npz"))
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    import oracle_py as orc
    threads = args.threads or int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1
    m = orc.Mesh("cartesian_tri", NX, NY, LX, LY, 0.0, 0)
    s = orc.Solver(m, orc.make_config(order=2, teno_variant="corrected"), n_threads=threads, canonical=True)
    # the split line sits below the domain: every cell takes the upper
    # states, left (x < 0.5) and right
    s.init_quadrant(0.5, -1.0, SOD)
    t0 = time.perf_counter()
    dts = []
    while s.time() < T_STOP:
        dts.extend(s.run(50 if s.time() < 0.95 * T_STOP else 1, CFL))
        if len(dts) % 200 == 0:
            print(f"step {len(dts)}: t = {s.time():.5f}, {time.perf_counter() - t0:.1f} s", flush=True)
    U, _ = s.state()
    if not np.isfinite(U).all():
        raise RuntimeError("oracle state is not finite")
    V = m.cell_volume
    y = m.cell_coords[:, 1]
    ids = np.nonzero((y >= 0.005) & (y < 0.006))[0].astype(np.uint32)
    np.savez_compressed(args.out, ids=ids, U=U[ids], dts=np.array(dts), t=s.time(),
                        linf=np.abs(U).max(axis=0), l1=(np.abs(U) * V[:, None]).sum(axis=0),
                        umin=U.min(axis=0))
    print(f"wrote {args.out}: {len(dts)} steps to t = {s.time():.6f}, {len(ids)} cells, "
          f"{time.perf_counter() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
