"""The C++ side of the boundary, run: ``lib/tf_driver`` (csrc/tf_driver.cpp) drives the C ABI through
the header-only facade include/tf_nr.hpp with the reference's NewtonRaphsonSolver call sequence
(constructUnMatrix, then calculateForcesAndJacobianMatrixNR / solveForDeltaU /
checkConvergenceViaDeltaU until convergence, as Simulation::updateStepNR, SC/Simulation.cpp:6449-6590)
on the reference's own Wind48hr_02.mesh.  Its final positions must equal the Python facade's
bitwise (same library, same calls) and the CPU oracle's within 1e-8 with the same Newton counts."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err
from tissuefolding_paperversion_amd import configs
from tissuefolding_paperversion_amd.mesh import read_mesh

pytestmark = pytest.mark.gpu

MESH = os.path.join(ROOT, "tests", "golden", "Wind48hr_02.mesh")
EXE = os.path.join(ROOT, "tissuefolding_paperversion_amd", "lib", "tf_driver")
STEPS, E, NU, G_HR = 3, 1000.0, 0.3, 0.05


def _driver_tissue():
    """The tissue tf_driver builds from the mesh file: one material, basal z fixed, x and y of the
    first basal node and y of the second one fixed, uniform in-plane growth."""
    m = read_mesh(MESH)
    n, ne = m.n_nodes, m.n_elements
    lam, mu = configs.lame(np.full(ne, E), np.full(ne, NU))
    fixed = np.zeros(3 * n, np.uint8)
    basal = np.where(m.node_placement == 0)[0]
    fixed[3 * basal + 2] = 1
    fixed[3 * basal[0]] = fixed[3 * basal[0] + 1] = 1
    fixed[3 * basal[1] + 1] = 1
    rates = np.tile([G_HR / 3600.0, G_HR / 3600.0, 0.0], (ne, 1))
    return configs.Tissue(m, lam, mu, np.zeros(ne), np.zeros(3 * n), fixed, rates, np.zeros(ne, np.uint8), 60.0,
                          STEPS, "Wind48hr_02")


def test_cpp_driver_matches_python_and_oracle(tmp_path, oracle):
    # the driver reads the file as it is: the mesh's prisms must already be oriented the way
    # read_mesh leaves them, or the two would describe different prisms
    raw = open(MESH).read().split()
    n = int(raw[0])
    ne = int(raw[1 + 6 * n])
    conn_raw = np.array(raw[2 + 6 * n:], dtype=float).reshape(ne, 25)[:, 1:7].astype(np.int32)
    t = _driver_tissue()
    assert np.array_equal(conn_raw, t.mesh.conn)

    out = tmp_path / "positions.txt"
    print(f"running {EXE} on {MESH}", flush=True)
    r = subprocess.run([EXE, MESH, str(STEPS), str(E), str(NU), str(G_HR), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    x_cpp = np.loadtxt(out).ravel()
    its = [0] * STEPS
    for step, it in re.findall(r"^step (\d+) iteration (\d+)", r.stdout, re.M):
        its[int(step)] = max(its[int(step)], int(it) + 1)

    # the Python facade, the same calls
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    s = NewtonRaphsonSolver(t)
    py_its = []
    for _ in range(STEPS):
        s.grow(t.dt)
        s.updateNodeViscositySurfaces()
        s.constructUnMatrix()
        k, conv = 0, False
        while not conv:
            s.calculateForcesAndJacobianMatrixNR(t.dt)
            s.solveForDeltaU()
            _, conv = s.updateUkInIteration()
            k += 1
            if not conv and k > 20:
                break
        py_its.append(k)
    assert py_its == its
    assert np.array_equal(s.positions, x_cpp)

    # the oracle
    o = oracle.OracleTissue(t, packing=False)
    o_its = [o.timestep(t.dt)[0] for _ in range(STEPS)]
    assert o_its == its
    assert rel_err(x_cpp, o.pos) < 1e-8
    assert rel_err(x_cpp, t.mesh.pos.ravel()) > 1e-5  # the tissue did grow
