"""GPU parity of host-given growth increments (angled, ring and grid growth functions) against the
oracle: Fg after one growth update within 1e-13 relative with the same rotated-element count,
and a model-input run with grid and ring functions within 1e-8 relative of the oracle's."""
import math

import numpy as np
import pytest

from conftest import rel_err
from test_growth_functions import write_grid_input
from tissuefolding_paperversion_amd import configs, modelinput

pytestmark = pytest.mark.gpu


def test_growth_increments_match_oracle(oracle):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    t = configs.disc_tissue(radius=15.0, n_steps=1)
    s = NewtonRaphsonSolver(t)
    o = oracle.OracleTissue(t)
    rng = np.random.default_rng(5)
    inc = np.eye(3)[None] + 0.01 * rng.standard_normal((t.n_elements, 3, 3))
    # rotate the tissue by 5 degrees about z so that the growth-rotation rule rotates the increments
    c, sn = math.cos(math.radians(5)), math.sin(math.radians(5))
    R = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1.0]])
    x = (t.mesh.pos @ R.T).ravel()
    s.positions = x
    o.pos = x.copy()
    s.set_growth_increments(inc)
    n_gpu = s.grow(t.dt)
    o.growth_inc = inc
    n_orc = o.grow(t.dt)
    fg, _ = s.get_fg()
    print(f"rotated increments: {n_gpu} (oracle {n_orc}) of {t.n_elements}", flush=True)
    assert n_gpu == n_orc > 0
    assert rel_err(fg, o.Fg.reshape(-1, 3, 3)) < 1e-13
    s.set_growth_increments(None)  # back to the rates
    s.grow(t.dt)


@pytest.mark.parametrize("pinned", [1, 0])
def test_grid_model_input_run_matches_oracle(oracle, tmp_path, pinned):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    mi = modelinput.read(write_grid_input(tmp_path, pinned=pinned))
    t = modelinput.build_tissue(mi)
    s = NewtonRaphsonSolver(t)
    n = 5
    stats = modelinput.run(mi, t, s, n_steps=n)
    o = oracle.OracleTissue(t)
    tr = modelinput.GrowthTracker(mi, t.mesh)
    for k in range(n):
        o.growth_inc = tr.increments(k * mi.dt, lambda: o.pos)
        o.timestep(mi.dt)
        assert stats[k]["nr_iterations"] == o.stats[-1]["nr_iters"]
    print(f"grid growth pinned={pinned}: NR {[st['nr_iterations'] for st in stats]}", flush=True)
    assert all(st["converged"] == 1 for st in stats)
    assert rel_err(s.positions, o.pos) < 1e-8
