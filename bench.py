#!/usr/bin/env python3
"""Throughput of the explicit Euler residual + SSP-RK3 step on one MI355X.

Workload: the 2D four-quadrant Riemann problem on a 2048x2048 quad mesh split
into 8,388,608 triangles (interior nodes jittered by +/-0.2h from a seed),
TENO reconstruction (p=2, Legendre basis, stencil factor 2), HLLC, SSP-RK3,
CFL 0.4, transmissive (extrapolation) boundaries on the four sides, FP64
throughout. A "step" is one full
SSP-RK3 step (3 stages, each a TENO face-state pass + a flux/residual/stage
update pass) plus its CFL time-step reduction; the metric counts
cell-updates per RK stage: n_cells * 3 * steps / seconds.

The timed region launches every kernel on the solver's stream with a HIP
event after each launch (mlr_run_steps_profiled), so the same run yields the
whole-step throughput and the average duration of each kernel. The same
number of steps replayed from the captured hipGraph is timed afterwards and
reported as ms_per_step_graph. The plain and learned runs use the exact
TENO tables (bit-identical to the oracle). Afterwards the same steps are
timed on compact tables (plain_compact_tables: member areas folded into the
pseudo-inverses, face basis per reference-triangle edge), with their largest
relative difference from the exact tables' run (--exact-tables skips it),
and on a second solver with symmetry walls on the four sides and the
positivity-preserving flux limiter (symmetry_walls; --no-walls skips it).

cpu_baseline: the CPU oracle (the restatement of the reference's OpenMP path:
AoS state, per-face atomics) timed on this host's cores on bounded samples of
the same numerics: a 1024x1024 (2,097,152-triangle) mesh, whose figure is
`value`, and a 512x512 mesh beside it, each for about --cpu-seconds (at most
the workload's 100 steps), with OMP_PROC_BIND=close OMP_PLACES=cores; and
the serial canonical form (one core, one-thread face order) on a 256x256
mesh. The report names the host CPU, nproc, the cores this process may use
and the OpenMP threads used.

Usage: python bench.py [--gpus 1] [--steps 100] [--warmup 3]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the cores this process may use, read before the OpenMP runtime binds the
# main thread to its first place
ALLOWED_CORES = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
# the OpenMP runtime reads its placement when the first library using it
# loads, so it is set before anything is imported
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

METRIC = "cell-updates/s per RK stage, learned indicator vs plain TENO at matched L1"
QUADRANT = np.array([[1.5, 0.0, 0.0, 1.5],        # upper right
                     [0.5323, 1.206, 0.0, 0.3],   # upper left
                     [0.138, 1.206, 1.206, 0.029],  # lower left
                     [0.5323, 0.0, 1.206, 0.3]])  # lower right
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "teno_kernel_riemann2048_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def numerics_kw(args):
    bcs = [(z, args.bc, {}) for z in ("left", "right", "top", "bottom")]
    return dict(recon="TENO", basis="legendre", order=args.order, stencil_factor=2.0,
                teno_variant=args.variant, riemann="HLLC", integrator=args.integrator, cfl=0.4, bcs=bcs,
                positivity=args.positivity)


def teno_bytes(n_cells, n_stencils, K, M, S, Q, tile_entries, pos_bytes, compact=False, derived=False):
    """Algorithmic HBM bytes of one TENO face-state launch.

    alg: each compulsory element counted once, per non-empty stencil the
    reference's K x M pseudo-inverse (8 B), M ids (4 B) and M transformed
    areas (8 B), plus the cell's own state (32 B).
    layout: what the packed device layout must move at least: the
    (K-1) x (M-1) non-zero block of the pseudo-inverse, M-1 areas and in-tile
    positions, the tile lists (4 B per entry), the cell state, its info word,
    the face-point basis values, the face states written (S*Q points x 4
    variables) and the decision byte. Compact tables read neither the areas
    (folded into the pseudo-inverse) nor the per-cell face basis; the
    derived face basis reads the cell's node coordinates (48 B) instead of
    the face basis.
    """
    alg = n_stencils * (8 * K * M + 12 * M) + 32 * n_cells
    area_b = 0 if compact else 8
    basis_b = 0 if (compact or derived) else 8 * S * Q * (K - 1)
    layout = (n_stencils * (8 * (K - 1) * (M - 1) + (area_b + pos_bytes) * (M - 1)) + 4 * tile_entries +
              n_cells * (32 + 4 + basis_b + 8 * S * Q * 4 + 1 + (48 if derived else 0)))
    return alg, layout


def step_roofline(teno_alg, N, n_stages, wall):
    """The plain path's whole step against HBM peak: algorithmic bytes of a
    stage are the TENO launch's plus 224 B per cell outside it (the stage
    input, step-start and output states, 3 x 32 B; the cell geometry, 68 B;
    1.5 faces' normals, areas, cells and nodes, 60 B), over the timed steps'
    wall time."""
    per_stage = teno_alg + 224 * N
    achieved = per_stage * n_stages / wall / 1e9
    return {"alg_bytes_per_stage": per_stage, "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
            "frac": achieved / 8000.0}


def learned_roofline(learned, N, dev):
    """The gated TENO kernel of the learned run against HBM peak. Algorithmic
    bytes per launch, in the learned form's accounting: a gated cell needs
    only its central stencil (the K x M pseudo-inverse, M ids and M areas),
    an ungated one all its non-empty stencils, every cell its state; averaged
    with the measured gated fraction. (The kernel reads less for a gated
    cell: its precombined linear operator, S*Q x (M-1) values, in place of
    the pseudo-inverse, areas and face basis.)"""
    if learned is None:
        return None
    per_st = 8 * dev.K * dev.M + 12 * dev.M
    g = learned["gated_fraction"]
    alg = 32 * N + per_st * (g * N + (1.0 - g) * dev.n_stencils)
    achieved = alg / (learned["teno_ms"] * 1e-3) / 1e9
    return {"kernel": "teno_kernel (learned gate; network evaluated once per step)", "bound": "hbm",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "alg_bytes_per_launch": alg, "avg_launch_ms": learned["teno_ms"]}


def check_state(dev, what):
    U, W = dev.state()
    bad = ~np.isfinite(U).all(axis=1) | (U[:, 0] <= 0) | (W[:, 2] <= 0)
    if bad.any():
        raise RuntimeError(f"non-physical state after the {what}: {int(bad.sum())} cells, "
                           f"min rho {np.nanmin(U[:, 0]):.3g}, min p {np.nanmin(W[:, 2]):.3g}")
    log(f"[bench] {what}: state physical, min rho {U[:, 0].min():.3g}, min p {W[:, 2].min():.3g}")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_sample(orc, args, n, threads, canonical, seconds):
    """Steps of the oracle on an n x n sample of the workload: cell-updates/s
    per RK stage over steps 2.. (the first step, with setup, is not timed)."""
    m = orc.Mesh("cartesian_tri", n, n, 1.0, 1.0, args.jitter, args.seed)
    bcs = [(z, args.bc, {}) for z in ("left", "right", "top", "bottom")]
    s = orc.Solver(m, orc.make_config(order=args.order, teno_variant=args.variant, bcs=bcs), n_threads=threads,
                   canonical=canonical)
    s.init_quadrant(0.5, 0.5, QUADRANT)
    t = time.perf_counter()
    s.run(1, 0.4)
    one = time.perf_counter() - t
    # at most the workload's 100 steps (1 + 99) from the initial state
    k = min(99, max(2, int(seconds / max(one, 1e-3))))
    t = time.perf_counter()
    s.run(k, 0.4)
    el = time.perf_counter() - t
    U, _ = s.state()
    if not np.isfinite(U).all():
        raise RuntimeError("oracle sample produced non-finite state")
    return m.n_cells, k, el, m.n_cells * 3 * k / el


def cpu_baseline(args):
    """Oracle (OpenMP, atomics) on bounded samples of the same workload, on
    every core this process may use (the GPU's share of the host where a
    scheduler grants one: OMP_NUM_THREADS caps it), plus one serial core."""
    from oracle import oracle_py as orc
    nproc = os.cpu_count() or 1
    allowed = ALLOWED_CORES
    omp_cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(allowed, omp_cap) if omp_cap > 0 else allowed
    sizes = []
    for n in (args.cpu_nx, args.cpu_nx // 2):
        cells, k, el, v = cpu_sample(orc, args, n, threads, False, args.cpu_seconds)
        sizes.append({"mesh": f"{n}x{n}", "cells": cells, "steps": k, "seconds": round(el, 3), "value": v})
        log(f"[bench] CPU oracle {n}x{n} on {threads} threads: {v / 1e6:.2f} M cell-updates/s ({k} steps, {el:.1f} s)")
    n1 = max(16, args.cpu_nx // 4)
    cells, k, el, v1 = cpu_sample(orc, args, n1, 1, True, args.cpu_seconds * 2.0 / 3.0)
    log(f"[bench] CPU oracle {n1}x{n1} serial canonical: {v1 / 1e6:.3f} M cell-updates/s ({k} steps, {el:.1f} s)")
    big = sizes[0]
    per_cell_spread = max(x["value"] for x in sizes) / min(x["value"] for x in sizes)
    return {"value": big["value"], "unit": "cell-updates/s per RK stage", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (OpenMP, per-face atomics, OMP_PROC_BIND=close OMP_PLACES=cores) on "
                      f"{big['mesh']} quads split into {big['cells']} jittered triangles, same numerics, "
                      f"{big['steps']} SSP-RK3 steps in {big['seconds']:.2f} s",
            "cpu_model": cpu_model(), "nproc": nproc, "cores_allowed": allowed,
            "omp_num_threads": omp_cap or None, "sizes": sizes, "size_spread": per_cell_spread,
            "serial": {"value": v1, "cores": 1, "mesh": f"{n1}x{n1}", "cells": cells, "steps": k,
                       "seconds": round(el, 3), "form": "one thread, canonical face order"}}


def learned_run(args, mlr, dev, U_plain, value_plain, wall_plain):
    """Train the indicator on the GPU (seeded random quadrant problems on a
    64x64 jittered mesh), then time the same steps with the gate on and
    compare the final state with the plain run's."""
    from mallard_amd import indicator as ind
    t = time.perf_counter()
    if args.indicator_file:
        # parameters saved by an earlier run (--save-indicator): no dataset, no training
        params = np.load(args.indicator_file, allow_pickle=False).astype(np.float32)
        y, hist = np.zeros(0), np.zeros(50)
        log(f"[bench] indicator parameters loaded from {args.indicator_file}")
    else:
        ds, y = ind.build_dataset(n_cases=args.train_cases, nx=64, snapshots=(0, 8, 20), seed=args.seed)
        params, hist = ind.train_indicator(ds, ind.make_pool(y, seed=args.seed), steps=3000, batch=4096, lr=2e-3,
                                           seed=args.seed)
        del ds
        log(f"[bench] indicator: {len(y)} rows ({int((y < 0.5).sum())} non-smooth), trained in "
            f"{time.perf_counter() - t:.1f} s, final loss {hist[-50:].mean():.4f}")
        if args.save_indicator:
            np.save(args.save_indicator, params)
    train_s = time.perf_counter() - t
    N = dev.n_cells

    def timed(every_stage):
        dev.set_indicator(params, args.threshold, every_stage=every_stage, split=bool(args.gate_split))
        dev.init_quadrant(0.5, 0.5, QUADRANT)
        tune = dev.autotune()  # the gated kernel's shapes, before the timed region
        dev.gated(reset=True)
        t = time.perf_counter()
        st = dev.run_steps_profiled(args.steps)
        wall = time.perf_counter() - t
        gated = dev.gated(reset=True)
        check_state(dev, "learned-indicator steps")
        l1 = ind.rel_l1(dev.state()[0], U_plain, dev.volumes())
        value = N * 3 * args.steps / wall
        log(f"[bench] learned ({'every stage' if every_stage else 'once per step'}): {wall:.3f} s wall, "
            f"x{value / value_plain:.2f}, gated {gated / (N * 3 * args.steps):.3f}, rel L1 {l1}")
        return {"value": value, "ms_per_step": wall / args.steps * 1e3, "speedup": value / value_plain,
                "l1_rel": [float(x) for x in l1], "matched_l1": bool(np.all(l1 <= 1e-3)),
                "gated_fraction": gated / (N * 3 * args.steps),
                "teno_ms": st["teno_ms"] / max(st["teno_n"], 1), "tune": list(tune)}

    out = timed(False)
    out["network_cadence"] = "once per step (first stage), decisions reused by stages 2 and 3"
    every = timed(True)
    dev.set_indicator(None)
    out["every_stage"] = {k: every[k] for k in ("value", "ms_per_step", "speedup", "l1_rel", "gated_fraction",
                                                "teno_ms")}
    out.update({"train_s": train_s, "train_rows": int(len(y)), "train_loss": float(hist[-50:].mean()),
                "threshold": args.threshold,
                "gate_kernels": "split: gated cells, then the ungated cells' blocks" if args.gate_split
                else "one kernel for gated and ungated cells"})
    return out


def ensemble_sizes(n_cases, seed, lo=64, hi=1024):
    """Quads per side of each case: log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_cases))).astype(int)


def ensemble_run(args, mlr):
    """Batched ensemble: independent four-quadrant problems with seeded random
    states on jittered triangle meshes of uneven sizes, advanced together (one
    launch per kernel phase over all cases, one hipGraph per step), plain
    and with the learned gate. Returns the JSON record."""
    from mallard_amd import indicator as ind
    sizes = ensemble_sizes(args.ens_cases, args.ens_size_seed, args.ens_min_side, args.ens_max_side)
    rng = np.random.default_rng(args.seed)
    states = [ind.random_quadrant_states(rng) for _ in sizes]
    gens = [mlr.mesh_gen("cartesian_tri", int(n), int(n), 1.0, 1.0, args.jitter, args.seed + k)
            for k, n in enumerate(sizes)]
    t0 = time.perf_counter()
    log(f"[bench] building {len(sizes)} cases ({sizes.min()}^2 .. {sizes.max()}^2 quads) ...")
    ens = mlr.Ensemble(gens, mlr.numerics(**numerics_kw(args)))
    N = int(ens.n_cells.sum())
    log(f"[bench] ensemble setup {time.perf_counter() - t0:.1f} s: {N} cells in {len(sizes)} cases")

    def init():
        for k in range(len(sizes)):
            ens.init_quadrant(k, 0.5, 0.5, states[k])

    def physical(k):
        U, W = ens.state(k)
        return bool(np.isfinite(U).all() and (U[:, 0] > 0).all() and (W[:, 2] > 0).all())

    init()
    tune = ens.autotune()
    log(f"[bench] autotuned (TENO occupancy, block, block, reuse occupancy, face basis): {tune}")
    if args.warmup > 0:
        ens.run_steps(args.warmup)
    init()
    t = time.perf_counter()
    ens.run_steps(args.ens_steps)
    wall = time.perf_counter() - t
    value_plain = N * 3 * args.ens_steps / wall
    ok = [physical(k) for k in range(len(sizes))]
    log(f"[bench] plain: {wall:.3f} s for {args.ens_steps} steps, {value_plain / 1e9:.3f} G cell-updates/s, "
        f"{sum(ok)}/{len(ok)} cases physical")
    U_plain = [ens.state(k)[0] for k in range(len(sizes))]
    plain = {"value": value_plain, "ms_per_step": wall / args.ens_steps * 1e3, "cases_physical": int(sum(ok)),
             "tune": list(tune)}

    persistent = None
    if not args.no_persistent and not args.positivity:
        # every step as one cooperative launch over all cases; states must equal the graph run's bit for bit
        ens.set_persistent(True)
        init()
        if args.warmup > 0:
            ens.run_steps(args.warmup)
        init()
        t = time.perf_counter()
        ens.run_steps(args.ens_steps)
        wall_p = time.perf_counter() - t
        same = all(np.array_equal(ens.state(k)[0].view(np.int64), U_plain[k].view(np.int64))
                   for k in range(len(sizes)))
        value_p = N * 3 * args.ens_steps / wall_p
        log(f"[bench] persistent: {wall_p:.3f} s, x{value_p / value_plain:.2f} vs graph, "
            f"{ens.persistent_grid} blocks, bitwise equal {same}")
        persistent = {"value": value_p, "ms_per_step": wall_p / args.ens_steps * 1e3,
                      "speedup_vs_graph": value_p / value_plain, "grid": ens.persistent_grid,
                      "bitwise_equal_graph": bool(same)}
        ens.set_persistent(False)

    learned = None
    if not args.no_learned:
        t = time.perf_counter()
        ds, y = ind.build_dataset(n_cases=args.train_cases, nx=64, snapshots=(0, 8, 20), seed=args.seed)
        params, hist = ind.train_indicator(ds, ind.make_pool(y, seed=args.seed), steps=3000, batch=4096, lr=2e-3,
                                           seed=args.seed)
        del ds
        train_s = time.perf_counter() - t
        ens.set_indicator(params, args.threshold)
        init()
        tune_g = ens.autotune()
        ens.gated(reset=True)
        t = time.perf_counter()
        ens.run_steps(args.ens_steps)
        wall_g = time.perf_counter() - t
        gated = ens.gated(reset=True)
        num = np.zeros(4)
        den = np.zeros(4)
        for k in range(len(sizes)):
            if not (ok[k] and physical(k)):
                continue
            v = ens.volumes(k)[:, None]
            num += (np.abs(ens.state(k)[0] - U_plain[k]) * v).sum(0)
            den += (np.abs(U_plain[k]) * v).sum(0)
        l1 = num / np.maximum(den, 1e-300)
        value_g = N * 3 * args.ens_steps / wall_g
        log(f"[bench] learned: {wall_g:.3f} s, x{value_g / value_plain:.2f}, gated "
            f"{gated / (N * 3 * args.ens_steps):.3f}, rel L1 {l1}")
        learned = {"value": value_g, "ms_per_step": wall_g / args.ens_steps * 1e3,
                   "speedup": value_g / value_plain, "l1_rel": [float(x) for x in l1],
                   "matched_l1": bool(np.all(l1 <= 1e-3)), "gated_fraction": gated / (N * 3 * args.ens_steps),
                   "train_s": train_s, "tune": list(tune_g), "threshold": args.threshold}
    matched = learned is not None and learned["matched_l1"]
    return {
        "metric": METRIC,
        "value": learned["value"] if matched else value_plain,
        "unit": "cell-updates/s per RK stage",
        "n_gpus": 1,
        "steps": args.ens_steps,
        "warmup": args.warmup,
        "ms_per_step": learned["ms_per_step"] if matched else plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched ensemble: {len(sizes)} four-quadrant Riemann problems, random states "
                        f"(rho, p ~ U[0.1, 2], u, v ~ U[-1.5, 1.5], seed {args.seed}) on jittered triangle "
                        f"meshes of {sizes.min()}^2 .. {sizes.max()}^2 quads (log-uniform draw, seed "
                        f"{args.ens_size_seed}), TENO p={args.order} + HLLC + SSPRK3, CFL 0.4"
                        f"{', positivity limiter' if args.positivity else ''}, one launch "
                        f"per kernel phase, hipGraph per step (the persistent record: one cooperative "
                        f"launch per step)",
            "n_cells": N,
            "n_cases": len(sizes),
            "indicator": "learned MLP gate, threshold %g" % args.threshold if matched else "none",
        },
        "plain": plain,
        "persistent": persistent,
        "learned": learned,
    }


def wedge_run(args, mlr):
    """Viscous wedge: Mach 2.5 Navier-Stokes flow over a 15-degree wedge of
    length 1 at Re = 1e5 (Sutherland viscosity, Pr 0.72), on a seeded graded
    Delaunay mesh; supersonic inflow, symmetry ahead of the wedge, adiabatic
    no-slip wedge surface, extrapolation outflow; TENO p=2 + Roe + viscous
    face gradients, SSP-RK3. Returns the JSON record (no reference
    counterpart: the reference solves Euler only)."""
    from mallard_amd import meshgen
    t0 = time.perf_counter()
    log(f"[bench] building the Delaunay wedge mesh ({args.wedge_cells} cells requested) ...")
    mesh = meshgen.wedge_mesh(n_cells=args.wedge_cells, seed=args.seed)
    t_mesh = time.perf_counter() - t0
    gamma, Tinf, uinf = 1.4, 1.0 / 1.4, 2.5
    bcs = [("left", "upt", {"u": (uinf, 0.0), "p": 1.0, "T": Tinf}), ("right", "extrapolation", {}),
           ("top", "extrapolation", {}), ("symmetry", "symmetry", {}), ("wall", "wall_noslip", {})]
    num = mlr.numerics(recon="TENO", basis="legendre", order=2, stencil_factor=2.0, teno_variant="corrected",
                       riemann="Roe", integrator="SSPRK3", cfl=0.4, bcs=bcs, gamma=gamma, p_ref=1.0, T_ref=1.0,
                       rho_ref=1.0, physics="navier-stokes", mu_ref=1.4 * uinf / 1e5, T_mu_ref=Tinf,
                       sutherland_S=110.4 / 288.15 * Tinf, prandtl=0.72)
    dev = mlr.Solver(mesh=mesh, num=num)
    N = dev.n_cells
    log(f"[bench] wedge setup {time.perf_counter() - t0:.1f} s (mesh {t_mesh:.1f} s): {N} cells")
    dev.init_constant([uinf, 0.0], 1.0, Tinf)
    tune = dev.autotune()
    if args.warmup > 0:
        dev.run_steps(args.warmup)
    dev.init_constant([uinf, 0.0], 1.0, Tinf)
    t = time.perf_counter()
    st = dev.run_steps_profiled(args.wedge_steps)
    wall = time.perf_counter() - t
    check_state(dev, "wedge steps")
    value = N * 3 * args.wedge_steps / wall
    log(f"[bench] wedge: {wall:.3f} s for {args.wedge_steps} steps, {value / 1e9:.3f} G cell-updates/s")
    return {
        "metric": METRIC, "value": value, "unit": "cell-updates/s per RK stage", "n_gpus": 1,
        "steps": args.wedge_steps, "warmup": args.warmup, "ms_per_step": wall / args.wedge_steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"viscous wedge: Mach 2.5 over a 15-degree wedge of length 1, Re 1e5, "
                               f"Sutherland, Pr 0.72, graded Delaunay mesh (seed {args.seed}), TENO p=2 + Roe + "
                               f"viscous face gradients, SSPRK3, CFL 0.4", "n_cells": N},
        "kernels_ms": {"teno": st["teno_ms"] / max(st["teno_n"], 1),
                       "gradients_flux_update": st["flux_ms"] / max(st["flux_n"], 1),
                       "spectral_dt": st["dt_ms"] / max(st["dt_n"], 1)},
        "tune": list(tune),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=2048, help="quads per side (2 triangles each)")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--jitter", type=float, default=0.2)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--variant", default="corrected", choices=["corrected", "reference"])
    ap.add_argument("--bc", default="extrapolation", choices=["extrapolation", "symmetry"],
                    help="boundary condition on the four sides (transmissive by default)")
    ap.add_argument("--integrator", default="SSPRK3", choices=["SSPRK3", "FE"],
                    help="FE adds one plain state copy per step (a known-size kernel for "
                         "calibrating HBM byte counters); the metric is defined for SSPRK3")
    ap.add_argument("--cpu-nx", type=int, default=1024,
                    help="quads per side of the CPU baseline's main sample (also timed at half that)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-indicator run")
    ap.add_argument("--exact-tables", action="store_true",
                    help="skip the extra timing on compact TENO tables")
    ap.add_argument("--positivity", action="store_true",
                    help="the positivity-preserving flux limiter in the main (or ensemble) run")
    ap.add_argument("--no-walls", action="store_true",
                    help="skip the extra timing with symmetry walls and the positivity limiter")
    ap.add_argument("--train-cases", type=int, default=512)
    ap.add_argument("--threshold", type=float, default=0.9)
    ap.add_argument("--gate-split", type=int, default=1, choices=[0, 1],
                    help="1: gated passes run as two kernels (gated cells, then ungated); 0: one kernel")
    ap.add_argument("--save-indicator", default=None, help="save the trained indicator parameters (.npy)")
    ap.add_argument("--indicator-file", default=None,
                    help="use saved indicator parameters instead of training (profiling runs)")
    ap.add_argument("--workload", default="riemann", choices=["riemann", "ensemble", "wedge"],
                    help="riemann: the 2048^2 problem (default); ensemble: the batched ensemble; "
                         "wedge: the viscous wedge")
    ap.add_argument("--wedge-cells", type=int, default=4_000_000)
    ap.add_argument("--wedge-steps", type=int, default=200)
    ap.add_argument("--ens-cases", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=50)
    ap.add_argument("--ens-size-seed", type=int, default=2112,
                    help="seed of the case-size draw (2112: 70.3M cells in 256 cases)")
    ap.add_argument("--ens-min-side", type=int, default=64, help="smallest case: quads per side")
    ap.add_argument("--ens-max-side", type=int, default=1024, help="largest case: quads per side")
    ap.add_argument("--no-persistent", action="store_true",
                    help="ensemble: skip the timing of the one-launch-per-step (persistent) mode")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported: "
              "the residual path runs on a single MI355X and has no multi-GPU mode.", file=sys.stderr)
        sys.exit(2)

    import mallard_amd as mlr

    if args.workload == "ensemble":
        print(json.dumps(ensemble_run(args, mlr)), flush=True)
        return
    if args.workload == "wedge":
        print(json.dumps(wedge_run(args, mlr)), flush=True)
        return

    t0 = time.perf_counter()
    log(f"[bench] building {args.nx}x{args.nx} jittered triangle mesh + TENO p={args.order} layout ...")
    dev = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", args.nx, args.nx, 1.0, 1.0, args.jitter, args.seed),
                     num=mlr.numerics(**numerics_kw(args)))
    dev.init_quadrant(0.5, 0.5, QUADRANT)
    N = dev.n_cells
    log(f"[bench] setup {time.perf_counter() - t0:.1f} s: {N} cells, {dev.n_stencils} stencils, "
        f"device layout {dev.device_bytes / 1e9:.2f} GB")

    # launch shapes picked on this mesh before anything is timed
    tune = dev.autotune()
    log(f"[bench] autotuned launch shapes (TENO occupancy, flux block, spectral block, reuse occupancy, "
        f"TENO face basis stored 0 / derived 1): {tune}")
    if args.warmup > 0:
        dev.run_steps(args.warmup)
    log(f"[bench] warmup {args.warmup} steps done; timing {args.steps} steps")

    def plain_run(what):
        # every timed block is the workload itself: steps from the initial state
        dev.init_quadrant(0.5, 0.5, QUADRANT)
        # timed region: kernel-by-kernel launches with HIP events on the solver's stream
        t = time.perf_counter()
        st = dev.run_steps_profiled(args.steps)   # returns after the stream has drained
        wall = time.perf_counter() - t
        log(f"[bench] timed ({what}): {wall:.3f} s wall, {st['total_ms']:.1f} ms by events")
        check_state(dev, "timed steps")
        dev.init_quadrant(0.5, 0.5, QUADRANT)
        t = time.perf_counter()
        dev.run_steps(args.steps)
        wall_graph = time.perf_counter() - t
        check_state(dev, "graph-replayed steps")
        return st, wall, wall_graph, dev.state()[0]

    st, wall, wall_graph, U_plain = plain_run("exact tables")
    stages = {"SSPRK3": 3, "FE": 1}[args.integrator]
    teno_ms = st["teno_ms"] / max(st["teno_n"], 1)
    flux_ms = st["flux_ms"] / max(st["flux_n"], 1)
    alg, lay = teno_bytes(N, dev.n_stencils, dev.K, dev.M, dev.S, dev.Q, dev.tile_entries, dev.tile_pos_bytes,
                          derived=tune[4] == 1)
    achieved = alg / (teno_ms * 1e-3) / 1e9
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    # summary, used only when it was measured on this workload and on the
    # kernel source being run
    traffic = traffic_compact = None
    if os.path.exists(TRAFFIC_FILE):
        tr = json.load(open(TRAFFIC_FILE))
        src = open(os.path.join(ROOT, "mallard_amd", "csrc", "kernels.hip"), "rb").read()
        if (tr.get("n_cells") == N and tr.get("order") == args.order and
                tr.get("kernels_sha256") == hashlib.sha256(src).hexdigest()[:16]):
            traffic = tr["bytes_per_launch"]
            traffic_compact = tr.get("compact_bytes_per_launch")

    # what HBM delivers to plain 16-byte streaming on this GPU (2 GiB buffers)
    nb = 2 << 30
    stream_copy = 2 * nb / (mlr.probe_stream(0, nb, 10) * 1e-3) / 1e9
    stream_read = nb / (mlr.probe_stream(1, nb, 10) * 1e-3) / 1e9
    # the same ceiling by other methods: the runtime's device-to-device copy,
    # 64 contiguous bytes per lane, non-temporal loads
    memcpy_d2d = 2 * nb / (mlr.probe_stream(2, nb, 10) * 1e-3) / 1e9
    read_wide = nb / (mlr.probe_stream(3, nb, 10) * 1e-3) / 1e9
    read_nt = nb / (mlr.probe_stream(4, nb, 10) * 1e-3) / 1e9
    log(f"[bench] HBM probes (GB/s): copy {stream_copy:.0f}, read {stream_read:.0f}, hipMemcpy D2D "
        f"{memcpy_d2d:.0f}, read 64 B/lane {read_wide:.0f}, non-temporal read {read_nt:.0f}")

    value_plain = N * stages * args.steps / wall
    learned = None
    if not args.no_learned and args.integrator == "SSPRK3" and args.order == 2:
        learned = learned_run(args, mlr, dev, U_plain, value_plain, wall)

    compact = None
    if not args.exact_tables and args.variant == "corrected" and args.order == 2:
        # the same steps on compact tables: member areas folded into A+ and the
        # face basis per reference-triangle edge (same quantities in exact
        # arithmetic, 472 fewer bytes per cell in the TENO pass, rounding
        # that the TENO cut-off can amplify); the tables change for good
        dev.set_indicator(None)
        dev.compact_tables()
        dev.autotune()
        if args.warmup > 0:
            dev.run_steps(args.warmup)
        stc, wallc, wall_graphc, Uc = plain_run("compact tables")
        sc = np.maximum(np.abs(U_plain).max(axis=0), 1e-300)
        dev_c = float((np.abs(Uc - U_plain).max(axis=0) / sc).max())
        teno_c = stc["teno_ms"] / max(stc["teno_n"], 1)
        alg_c, lay_c = teno_bytes(N, dev.n_stencils, dev.K, dev.M, dev.S, dev.Q, dev.tile_entries,
                                  dev.tile_pos_bytes, compact=True)
        ach_c = alg_c / (teno_c * 1e-3) / 1e9
        compact = {"value": N * stages * args.steps / wallc, "ms_per_step": wallc / args.steps * 1e3,
                   "ms_per_step_graph": wall_graphc / args.steps * 1e3, "teno_ms": teno_c,
                   "rel_linf_vs_exact_tables": dev_c,
                   "roofline": {"achieved": ach_c, "frac": ach_c / HBM_PEAK_GBS, "traffic": traffic_compact,
                                "layout_bytes_per_launch": lay_c}}
        log(f"[bench] compact tables: {compact['ms_per_step']:.2f} ms per step (TENO {teno_c:.3f} ms), relative "
            f"L-inf against the exact tables after {args.steps} steps {dev_c:.3g}")

    walls = None
    if not args.no_walls and args.bc == "extrapolation" and args.integrator == "SSPRK3":
        # the same mesh and data with symmetry walls on all four sides, which
        # reach non-finite states by step 40 without the positivity limiter
        kw = dict(numerics_kw(args), bcs=[(z, "symmetry", {}) for z in ("left", "right", "top", "bottom")],
                  positivity=True)
        ws = mlr.Solver(gen=mlr.mesh_gen("cartesian_tri", args.nx, args.nx, 1.0, 1.0, args.jitter, args.seed),
                        num=mlr.numerics(**kw))
        ws.init_quadrant(0.5, 0.5, QUADRANT)
        ws.autotune()
        if args.warmup > 0:
            ws.run_steps(args.warmup)
        ws.init_quadrant(0.5, 0.5, QUADRANT)
        t = time.perf_counter()
        stw = ws.run_steps_profiled(args.steps)
        wall_w = time.perf_counter() - t
        Uw, Ww = ws.state()
        ok = bool(np.isfinite(Uw).all() and (Uw[:, 0] > 0).all() and (Ww[:, 2] > 0).all())
        ws.init_quadrant(0.5, 0.5, QUADRANT)
        t = time.perf_counter()
        ws.run_steps(args.steps)
        wall_wg = time.perf_counter() - t
        walls = {"value": N * stages * args.steps / wall_w, "ms_per_step": wall_w / args.steps * 1e3,
                 "ms_per_step_graph": wall_wg / args.steps * 1e3,
                 "flux_ms": stw["flux_ms"] / max(stw["flux_n"], 1), "teno_ms": stw["teno_ms"] / max(stw["teno_n"], 1),
                 "physical_after_steps": ok, "min_rho": float(np.nanmin(Uw[:, 0])), "min_p": float(np.nanmin(Ww[:, 2])),
                 "boundaries": "symmetry on all four sides", "positivity_limiter": True}
        ws.close()
        log(f"[bench] symmetry walls + positivity limiter: {walls['ms_per_step']:.2f} ms per step (flux "
            f"{walls['flux_ms']:.3f} ms against {flux_ms:.3f} ms unlimited), physical after {args.steps} steps: {ok}")

    matched = learned is not None and learned["matched_l1"]
    value = learned["value"] if matched else value_plain
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s per RK stage",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": (learned["ms_per_step"] if matched else wall / args.steps * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"2D four-quadrant Riemann, {args.nx}x{args.nx} quads split into {N} triangles "
                        f"jittered +/-{args.jitter}h (seed {args.seed}), TENO p={args.order} legendre "
                        f"({args.variant} weighting) + HLLC + {args.integrator}, CFL 0.4, {args.bc} boundaries",
            "n_cells": N,
            "indicator": ("learned MLP gate (48-64-64-1 BF16 MFMA, threshold %g), L1 within 1e-3 of plain "
                          "FP64 TENO" % args.threshold) if matched else "none (plain FP64 TENO)",
        },
        "plain": {"value": value_plain, "ms_per_step": wall / args.steps * 1e3,
                  "ms_per_step_graph": wall_graph / args.steps * 1e3,
                  "tables": ("exact, face basis derived in the kernel from node coordinates"
                             if tune[4] == 1 else "exact, stored") + " (bit-identical to the oracle)",
                  "whole_step": step_roofline(alg, N, stages * args.steps, wall)},
        "plain_compact_tables": compact,
        "symmetry_walls": walls,
        "learned": learned,
        "roofline": {
            "kernel": "teno_kernel (plain FP64 TENO run)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "alg_bytes_per_launch": alg,
            "layout_bytes_per_launch": lay,
            "layout_GBs": lay / (teno_ms * 1e-3) / 1e9,
            "avg_launch_ms": teno_ms,
            "stream_copy_GBs": stream_copy,
            "stream_read_GBs": stream_read,
            "memcpy_d2d_GBs": memcpy_d2d,
            "stream_read_wide_GBs": read_wide,
            "stream_read_nontemporal_GBs": read_nt,
        },
        "roofline_learned": learned_roofline(learned, N, dev),
        "kernels_ms": {"teno": teno_ms, "flux_update": flux_ms,
                       "spectral_dt": st["dt_ms"] / max(st["dt_n"], 1),
                       "copy": st["copy_ms"] / max(st["copy_n"], 1) if st["copy_n"] else None},
        "tune": {"teno_occupancy": tune[0], "flux_block": tune[1], "spectral_block": tune[2],
                 "teno_occupancy_reuse": tune[3], "teno_face_basis": "derived" if tune[4] == 1 else "stored"},
    }
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        try:
            out["cpu_baseline"] = cpu_baseline(args)
        except Exception as e:  # the GPU figures stand on their own
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
