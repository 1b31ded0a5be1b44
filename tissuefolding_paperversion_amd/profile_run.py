"""Short, profiler-friendly run of the NR hot path.

Runs a few implicit timesteps of a workload, optionally without graph capture so that every
kernel is an ordinary dispatch a counter-collecting profiler samples one by one.  The HBM byte
counters it yields are calibrated against ``lib/tf_copy_calib`` (a copy kernel of known size)::

    python -m tissuefolding_paperversion_amd.profile_run --workload config2 --steps 1 --no-graph
"""
from __future__ import annotations

import argparse
import json
import sys


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--workload", default="config2", choices=["config1", "config2"])
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-graph", action="store_true")
    p.add_argument("--tiling", type=int, nargs=2, metavar=("WG", "TILE_BLOCKS"),
                   help="Krylov tiling (workgroup size, blocks per tile) instead of the default")
    p.add_argument("--preconditioner", default="auto", choices=["auto", "block-jacobi", "two-level"])
    a = p.parse_args(argv)

    from . import configs
    from .solver import NewtonRaphsonSolver

    t = configs.config2() if a.workload == "config2" else configs.config1()
    pc = {"auto": -1, "block-jacobi": 0, "two-level": 1}[a.preconditioner]
    s = NewtonRaphsonSolver(t, {"use_graph": 0 if a.no_graph else 1, "preconditioner": pc})
    if a.tiling:
        s.set_tiling(*a.tiling)
    for _ in range(a.warmup):
        s.runOneStep(t.dt)
    stats = [s.runOneStep(t.dt) for _ in range(a.steps)]
    wg, cap, tiles = s.tiling()
    print(json.dumps({"workload": t.name, "steps": a.steps, "nr_iterations": [st["nr_iterations"] for st in stats],
                      "krylov_iterations": [st["krylov_iterations"] for st in stats],
                      "krylov_by_newton": [st["krylov_by_newton"][:st["nr_iterations"]] for st in stats],
                      "persistent_solves": [st["krylov_persistent"] for st in stats],
                      "spmv_algorithmic_bytes": s.kernel_bytes("spmv"), "tiling": [wg, cap, tiles]}), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
