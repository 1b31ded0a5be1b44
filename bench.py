#!/usr/bin/env python3
"""Benchmark: implicit NR timesteps/s of the wing-disc ellipse (~40k prisms) on one MI355X.

One step = one implicit timestep: growth update (with the growth-rotation rule), nodal viscosity
surfaces, and the full Newton-Raphson loop until ||dU||_2 <= 1e-8 (assembly of forces and the
Jacobian, fixed DoFs, block-Jacobi preconditioned Krylov solve, update).  Inputs are resident in
HBM before the timed region.  Prints one JSON line.

The dominant kernel (k_cg1: one iteration of the single-kernel CG, i.e. the BSR SpMV with the
vector updates and dot products fused in) is timed with HIP events on the stream it runs on,
around each replay of a captured chunk of iterations; its algorithmic bytes per launch are
nb*(72+4) + (N+1)*4 (matrix) + 16*3N*8 (x, p, r, w, s, u, m, q read and written once) + N*72
(block inverses).
A second leg runs the batched 256-disc sweep (config 4: one tissue segment per disc in one
handle, autotuned tiling) and, for comparison, the same discs one handle at a time.
A third leg times the discrete adjoint on the inverse-growth workload (config 5: 30 steps of
the same ellipse, gradient of an apical-position loss with respect to 38,826 per-element rate
scales and 3 layer moduli) against its forward run, with a central-difference check.
The CPU baseline is the FP64 CPU oracle (same element math, OpenMP element loop, CSR assembly
and an MKL PARDISO direct transposed solve) timed on the host cores over one full timestep.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {
    -1: "k_cg1: one CG iteration (BSR3 FP64 SpMV, CSR-stream tiles, fused vector updates and dot products)",
    3: "k_cg1: one CG iteration (BSR3 FP64 SpMV, CSR-stream tiles, fused vector updates and dot products)",
    1: "k_cg_q: BSR3 FP64 SpMV with the CG direction update and one dot product",
    0: "k_bicg_spmv: BSR3 FP64 SpMV with fused dot products",
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=120, help="timed timesteps (config 2: 120 of dt = 60 s)")
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="config2", choices=["config1", "config2"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision runs")
    p.add_argument("--no-batch", action="store_true", help="skip the batched 256-disc sweep")
    p.add_argument("--batch-steps", type=int, default=3)
    p.add_argument("--no-adjoint", action="store_true", help="skip the config-5 adjoint leg")
    p.add_argument("--adjoint-fd", type=int, default=7, help="parameters checked by central differences")
    p.add_argument("--krylov", type=int, default=-1)
    p.add_argument("--matrix-fp32", action="store_true")
    p.add_argument("--no-graph", action="store_true")
    p.add_argument("--preconditioner", default="auto", choices=["auto", "block-jacobi", "two-level"])
    p.add_argument("--coarse-aggregates", type=int, default=0)
    p.add_argument("--coarse-refresh", type=int, default=100)
    p.add_argument("--no-kernel-timing", action="store_true", help="skip the HIP-event timing of the dominant kernel")
    p.add_argument("--tiling", type=int, nargs=2, metavar=("WG", "TILE_BLOCKS"),
                   help="Krylov tiling of the plain FP64 run instead of the tiling search (profiling runs)")
    return p.parse_args()


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch of the dominant Krylov kernel from the committed rocprofv3 PMC summary,
    scaled by the copy-kernel calibration recorded in the same file (None if absent)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+-]+)", txt, re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(tissue, dt):
    """One full timestep of the CPU oracle (growth, surfaces, node masses and the literal O(N^2)
    packing detection, NR loop with PARDISO)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU oracle, timed as the baseline only

    o = O.OracleTissue(tissue, solver="auto")
    cores = O.lib().tfo_num_threads()
    t0 = time.perf_counter()
    it, _ = o.timestep(dt)
    sec = time.perf_counter() - t0
    return {"value": 1.0 / sec, "unit": "timesteps/s", "cores": int(cores), "kind": "port",
            "sample": f"one full implicit timestep (growth + surfaces + packing detection ({len(o.pairs)} pairs) + "
                      f"{it} NR iterations, literal 729-term element tangent, CSR assembly, {o.solver_kind()} "
                      f"transposed solve) of the same workload",
            "seconds": sec, "nr_iterations": it}


def elastic_forces(s):
    """FP64 nodal forces at the current positions (un = uk, so no viscous term)."""
    s.constructUnMatrix()
    s.calculateForcesAndJacobianMatrixNR(60.0)
    return s.gSum


def run_timed(NewtonRaphsonSolver, tissue, opts, warmup, steps, timing, tune=True, tiling=None):
    """The workload's timesteps from its initial state on a fresh handle.  Outside the timed
    region, a second handle runs the warm-up steps (loading every kernel) and, from that state,
    the tiling search (or takes the given tiling); the timed handle takes its tiling before its
    first step."""
    dt = tissue.dt
    t = NewtonRaphsonSolver(tissue, opts)
    if tiling:
        t.set_tiling(*tiling)
    for _ in range(warmup):
        t.runOneStep(dt)
    wg = cap = None
    if tiling:
        wg, cap = tiling
    elif tune:
        t.autotune(dt, n_iters=64)
        wg, cap, _ = t.tiling()
    t.close()
    s = NewtonRaphsonSolver(tissue, opts)
    if wg is not None:
        s.set_tiling(wg, cap)
    for k in ("assemble_elements", "assemble_gather"):  # eager launches, events around each
        s.kernel_timing(k, enable=timing)
    s.positions  # synchronises the stream
    t0 = time.perf_counter()
    stats = [s.runOneStep(dt) for _ in range(steps)]
    x = s.positions
    sec = time.perf_counter() - t0
    return s, stats, sec, x


def trajectory(NewtonRaphsonSolver, tissue, opts, steps, tiling, ref=None):
    """Untimed run of the same timesteps on a fresh handle at the given tiling: the positions after
    every step, or (ref given) the largest relative displacement error against ref over all steps."""
    s = NewtonRaphsonSolver(tissue, opts)
    s.set_tiling(tiling[0], tiling[1])
    xs, err = [], 0.0
    for k in range(steps):
        s.runOneStep(tissue.dt)
        x = s.positions
        if ref is None:
            xs.append(x)
        else:
            err = max(err, float(np.max(np.abs(x - ref[k])) / np.max(np.abs(ref[k]))))
    s.close()
    return xs if ref is None else err


def batched_leg(NewtonRaphsonSolver, steps, warmup, cpu=True):
    """The 256-disc sweep (one tissue segment per disc in one handle) against the same discs run
    one handle at a time, plus a bounded CPU-oracle sample scaled by prism count."""
    from tissuefolding_paperversion_amd import configs

    ts = configs.config4()
    dt = 60.0
    n_prisms = sum(t.n_elements for t in ts)
    s = NewtonRaphsonSolver(ts)
    for _ in range(warmup):
        s.runOneStep(dt)
    s.autotune(dt, n_iters=64)
    s.positions
    t0 = time.perf_counter()
    stats = [s.runOneStep(dt) for _ in range(steps)]
    s.positions
    sec = time.perf_counter() - t0
    seg_it = s.segment_stats()[0]
    wg, cap, tiles = s.tiling()
    s.close()
    # the same discs, one handle each, one timed step after one untimed step
    seq = 0.0
    for t in ts:
        h = NewtonRaphsonSolver(t)
        h.runOneStep(dt)
        h.positions
        t1 = time.perf_counter()
        h.runOneStep(dt)
        h.positions
        seq += time.perf_counter() - t1
        h.close()
    out = {"workload": f"config 4: {len(ts)} discs, {n_prisms} prisms, {sum(t.n_nodes for t in ts)} nodes "
                       "(log-uniform 500-8000 prisms, 2-3 layers, per-disc E, nu, growth), dt = 60 s",
           "unit": "disc-timesteps/s", "value": len(ts) * steps / sec, "steps": steps,
           "ms_per_step": 1e3 * sec / steps, "prism_timesteps_per_s": n_prisms * steps / sec,
           "nr_iterations_per_step_max_over_discs": float(np.mean([st["nr_iterations"] for st in stats])),
           "nr_iterations_last_step_per_disc": {"min": int(seg_it.min()), "max": int(seg_it.max())},
           "krylov_iterations_per_nr_iteration": float(sum(st["krylov_iterations"] for st in stats)
                                                       / max(1, sum(st["nr_iterations"] for st in stats))),
           "converged_steps": int(sum(st["converged"] for st in stats)),
           "tiling": {"workgroup": wg, "tile_blocks": cap, "tiles": tiles},
           "one_handle_per_disc": {"value": len(ts) / seq, "ms_per_step": 1e3 * seq,
                                   "batched_speedup": (len(ts) * steps / sec) / (len(ts) / seq)}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the CPU oracle, timed as the baseline only

        done, prisms, t2 = 0, 0, time.perf_counter()
        for t in ts:
            O.OracleTissue(t, solver="auto").timestep(dt)
            done += 1
            prisms += t.n_elements
            if time.perf_counter() - t2 > 8.0:
                break
        csec = time.perf_counter() - t2
        out["cpu_baseline"] = {"value": len(ts) / (csec * n_prisms / prisms), "unit": "disc-timesteps/s",
                               "cores": int(O.lib().tfo_num_threads()), "kind": "port",
                               "sample": f"one timestep of the first {done} discs ({prisms} prisms) in sequence, "
                                         "scaled to the whole sweep by prism count"}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): the timestep of one tissue (or one "
              "batch of tissues) is measured on one MI355X; runs over several GPUs are not implemented.",
              file=sys.stderr)
        sys.exit(2)
    from tissuefolding_paperversion_amd import configs
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver

    tissue = configs.config2() if a.workload == "config2" else configs.config1()
    dt = tissue.dt
    opts = {"krylov": a.krylov}
    if a.matrix_fp32:
        opts["matrix_precision"] = 1
    if a.no_graph:
        opts["use_graph"] = 0
    opts["preconditioner"] = {"auto": -1, "block-jacobi": 0, "two-level": 1}[a.preconditioner]
    opts["coarse_aggregates"] = a.coarse_aggregates
    opts["coarse_refresh"] = a.coarse_refresh
    timing = 0 if a.no_kernel_timing else 1
    s, stats, sec, x_plain = run_timed(NewtonRaphsonSolver, tissue, opts, a.warmup, a.steps, timing, tiling=a.tiling)
    # the dominant Krylov kernel: 64 launches back to back on the final Jacobian, right after the
    # timed steps (HIP events around one graph replay; the timed steps replay their Krylov
    # chunks as graphs with nothing between the launches)
    spmv_n = 64 if timing else 0
    spmv_ms = s.kernel_probe(spmv_n) if timing else 0.0
    el_ms, el_n = s.kernel_timing("assemble_elements", enable=0)
    ga_ms, ga_n = s.kernel_timing("assemble_gather", enable=0)
    spmv_bytes = s.kernel_bytes("spmv")
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0
    nr = [st["nr_iterations"] for st in stats]
    two_level = bool(stats[-1]["two_level"])
    kr = [st["krylov_iterations"] for st in stats]
    out = {
        "metric": "NR timesteps/s at 40k prisms, plain FP64 vs mixed (disp err <=1e-8 rel)",
        "value": a.steps / sec,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * sec / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Delaunay ellipse mesh and seeded Gaussian-bump growth map)",
        "config": {"workload": f"{tissue.name}: {tissue.n_elements} prisms, {tissue.n_nodes} nodes, "
                               f"{3 * tissue.n_nodes} DoF, {a.steps} timesteps of dt=60 s from t = 0, plain FP64",
                   "prisms": tissue.n_elements, "nodes": tissue.n_nodes, "bsr_blocks": s.nb,
                   "krylov": "auto (single-kernel CG, BiCGSTAB fallback)" if a.krylov < 0 else str(a.krylov),
                   "matrix_values": "fp32" if a.matrix_fp32 else "fp64",
                   "preconditioner": ("two-level" if stats[-1]["two_level"] else "block-jacobi") +
                                     (" (auto)" if a.preconditioner == "auto" else ""),
                   "tiling": dict(zip(("workgroup", "tile_blocks", "tiles"), s.tiling()))},
        "nr_iterations_per_step": float(np.mean(nr)),
        "krylov_iterations_per_nr_iteration": float(np.sum(kr) / max(1, np.sum(nr))),
        "krylov_iterations_by_newton_iteration": [float(np.mean([st["krylov_by_newton"][k] for st in stats]))
                                                  for k in range(int(max(nr)) if len(nr) else 0) if k < 4],
        "converged_steps": int(sum(st["converged"] for st in stats)),
        "krylov_fallbacks": int(sum(st["krylov_fallbacks"] for st in stats)),
        "ms_per_step_by_phase": {k: float(np.mean([st[k] for st in stats]))
                                 for k in ("ms_grow", "ms_assemble", "ms_solve", "ms_total")},
        "packing": {"pairs_per_step_max": int(max(st["packing_pairs"] for st in stats)),
                    "ms_growth_surfaces_detection_per_step": float(np.mean([st["ms_grow"] for st in stats]))},
        "roofline": {"kernel": KERNEL_NAMES[1] if two_level else KERNEL_NAMES.get(a.krylov, KERNEL_NAMES[-1]),
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles("k_cg_q" if two_level else "k_cg1", a.workload),
                     "avg_launch_ms": spmv_ms, "launches": spmv_n,
                     "measured": "64 launches back to back on the final Jacobian after the timed steps, "
                                 "HIP events around one graph replay",
                     "algorithmic_bytes_per_launch": spmv_bytes},
        "kernels": {"assemble_elements_ms": el_ms, "assemble_gather_ms": ga_ms, "spmv_ms": spmv_ms},
    }
    if not a.no_mixed:
        # the same workload with MFMA element tangents and FP32 Jacobian values (FP64 residual,
        # update, Krylov vectors and dots): throughput, and errors against the plain FP64 run
        f_plain = elastic_forces(s)
        fscale = np.max(np.abs(f_plain))
        til = s.tiling()
        x_ref = trajectory(NewtonRaphsonSolver, tissue, opts, a.steps, til)  # FP64, every step
        mixed = {}
        for name, tp in (("fp64_tangent_fp32_matrix", 0), ("fp32_mfma", 1), ("bf16_mfma", 2)):
            mopts = dict(opts, tangent_precision=tp, matrix_precision=1)
            sm, mstats, msec, x_mixed = run_timed(NewtonRaphsonSolver, tissue, mopts, a.warmup, a.steps, 0)
            disp_err = float(np.max(np.abs(x_mixed - x_plain)) / np.max(np.abs(x_plain)))
            f_err = float(np.max(np.abs(elastic_forces(sm) - f_plain)) / fscale)
            step_err = trajectory(NewtonRaphsonSolver, tissue, mopts, a.steps, til, ref=x_ref)
            mixed[name] = {"value": a.steps / msec, "ms_per_step": 1e3 * msec / a.steps,
                           "disp_err_rel_max_over_steps": step_err,
                           "disp_err_rel_after_all_steps": disp_err, "force_err_rel": f_err,
                           "nr_iterations_per_step": float(np.mean([st["nr_iterations"] for st in mstats])),
                           "krylov_iterations_per_nr_iteration":
                               float(np.sum([st["krylov_iterations"] for st in mstats]) /
                                     max(1, np.sum([st["nr_iterations"] for st in mstats]))),
                           "krylov_iterations_by_newton_iteration":
                               [float(np.mean([st["krylov_by_newton"][k] for st in mstats]))
                                for k in range(min(4, max(st["nr_iterations"] for st in mstats)))],
                           "ms_per_step_by_phase": {k: float(np.mean([st[k] for st in mstats]))
                                                    for k in ("ms_grow", "ms_assemble", "ms_solve")},
                           "converged_steps": int(sum(st["converged"] for st in mstats)),
                           "krylov_fallbacks": int(sum(st["krylov_fallbacks"] for st in mstats)),
                           "speedup_vs_fp64": (a.steps / msec) / (a.steps / sec)}
            sm.close()
        # the same pair under the block-Jacobi preconditioner alone: the mixed form's effect where
        # the SpMV dominates the solve
        bj = dict(opts, preconditioner=0)
        sb, bstats, bsec, x_bj = run_timed(NewtonRaphsonSolver, tissue, bj, a.warmup, a.steps, 0)
        sb.close()
        sm, mstats, msec, x_mixed = run_timed(NewtonRaphsonSolver, tissue, dict(bj, tangent_precision=1,
                                                                                matrix_precision=1),
                                              a.warmup, a.steps, 0)
        sm.close()
        mixed["block_jacobi"] = {"fp64_value": a.steps / bsec, "fp32_mfma_value": a.steps / msec,
                                 "speedup_vs_fp64": bsec / msec,
                                 "disp_err_rel_after_all_steps": float(np.max(np.abs(x_mixed - x_bj)) /
                                                                       np.max(np.abs(x_bj))),
                                 "nr_iterations_per_step": [float(np.mean([st["nr_iterations"] for st in bstats])),
                                                            float(np.mean([st["nr_iterations"] for st in mstats]))]}
        out["mixed"] = mixed
    if not a.no_adjoint:
        from tissuefolding_paperversion_amd.adjoint import config5_check
        adj = config5_check(n_fd=a.adjoint_fd)
        adj.pop("rows")
        out["adjoint"] = adj
    if not a.no_batch:
        out["batched"] = batched_leg(NewtonRaphsonSolver, a.batch_steps, 1, cpu=not a.no_cpu_baseline)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(configs.config2() if a.workload == "config2" else configs.config1(), dt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
