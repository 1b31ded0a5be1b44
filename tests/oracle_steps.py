"""Runs the CPU oracle over the first timesteps of a named workload in a process of its own and
saves the positions and NR iteration counts after every step:

    python tests/oracle_steps.py <workload> <steps> <out.npz>

<workload> names a function of tissuefolding_paperversion_amd.configs (config1, config2, ...).  A
process of its own keeps torch (loaded by other tests) away from the oracle's MKL PARDISO solver,
which fails in a process that has loaded torch's bundled MKL.  Test infrastructure only."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402  (the checker)
from tissuefolding_paperversion_amd import configs  # noqa: E402


def main():
    name, steps, out = sys.argv[1], int(sys.argv[2]), sys.argv[3]
    t = getattr(configs, name)()
    o = O.OracleTissue(t, solver="auto")
    xs, its = [], []
    for k in range(steps):
        it, d = o.timestep(t.dt)
        print(f"oracle {name} step {k}: {it} NR iterations, last |dU| {d:.2e}, solver {o.solver_kind()}", flush=True)
        its.append(it)
        xs.append(o.pos.copy())
    np.savez(out, x=np.array(xs), its=np.array(its), solver=np.array(o.solver_kind()))


if __name__ == "__main__":
    main()
