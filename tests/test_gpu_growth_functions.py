"""GPU parity of the growth increments against the oracle: host-given increments (Fg after one
growth update within 1e-13 relative, same rotated-element count); growth functions evaluated on
the device (uniform at an angle, ring, grid pinned or current, nearest-corner or bilinear, several
at once, activity windows, re-pinning) against the oracle's host restatement within 1e-13; and a
model-input run with grid and ring functions within 1e-8 relative of the oracle's."""
import math

import numpy as np
import pytest

import growth_ref as GR  # the oracle's host growth functions (test infrastructure)
from conftest import rel_err
from test_growth_functions import write_grid_input
from tissuefolding_paperversion_amd import configs, modelinput
from tissuefolding_paperversion_amd import growth as G

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
    tr = GR.HostTracker(mi, t.mesh)
    for k in range(n):
        o.growth_inc = tr.increments(k * mi.dt, lambda: o.pos)
        o.timestep(mi.dt)
        assert stats[k]["nr_iterations"] == o.stats[-1]["nr_iters"]
    print(f"grid growth pinned={pinned}: NR {[st['nr_iterations'] for st in stats]}", flush=True)
    assert all(st["converged"] == 1 for st in stats)
    assert rel_err(s.positions, o.pos) < 1e-8


def _functions(nx=6, ny=5):
    rng = np.random.default_rng(21)
    rates = np.zeros((nx, ny, 3))
    rates[:, :, 0] = rng.uniform(1e-6, 2e-5, (nx, ny))
    rates[:, :, 1] = rng.uniform(1e-6, 2e-5, (nx, ny))
    rates[:, :, 2] = rng.uniform(-2e-6, 2e-6, (nx, ny))
    rates[1, 2, 1] = rates[1, 2, 0] * 1.05  # a nearly isotropic corner: its angle is eliminated
    angles = rng.uniform(-math.pi, math.pi, (nx, ny))
    grid = G.GridGrowth(0.0, 1e9, rates, angles, 0.0, 0.8)
    uni = G.UniformGrowth(0.0, 500.0, (3e-5, 1e-5, 2e-6), math.radians(35))
    ring = G.RingGrowth(100.0, 1e9, (5.0, -3.0), 8.0, 30.0, (4e-5, 2e-5, 0.0), math.radians(-20))
    return grid, uni, ring


@pytest.mark.parametrize("pinned", [1, 0])
@pytest.mark.parametrize("interp", [0, 1])
def test_device_growth_functions_match_the_oracle(pinned, interp):
    """Increments the device evaluates from the functions against the oracle's, at three step
    start times (the uniform function active only in the first two, the ring from the second),
    after the positions moved (current grids follow them, pinned ones do not), and after a
    re-pin on the moved positions."""
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    t = configs.wing_disc_tissue(a=60.0, b=36.0, h=3.66, z_levels=(0.0, 2.0, 7.0, 12.0), seed=2, growth_seed=3,
                                 name="growth-function ellipse")
    grid, uni, ring = _functions()
    fns = [grid, uni, ring]
    geo = G.GrowthGeometry(t.mesh)
    s = NewtonRaphsonSolver(t)
    s.set_growth_functions(fns, geo.initial_z, geo.rel_src, pinned, interp)
    rel_pin = GR.relative_positions(geo, t.mesh.pos)
    pos = t.mesh.pos.copy()
    dt = 60.0
    rng = np.random.default_rng(4)
    worst = 0.0
    for k, time in enumerate((0.0, 120.0, 600.0)):
        if k == 2:  # move the nodes: a stretch and a shift
            pos = pos * np.array([1.07, 0.96, 1.0]) + np.array([2.0, -1.5, 0.3]) + 0.05 * rng.standard_normal(pos.shape)
            s.positions = pos.ravel()
        s.set_time(time)
        s.grow(dt)
        inc = s.growth_increments()
        ref = GR.increments(fns, geo, pos, time, dt, bool(pinned), interp, rel_pin)
        err = rel_err(inc, ref)
        worst = max(worst, err)
        assert err < 1e-13, (time, err)
        assert not np.allclose(inc, np.eye(3)[None])
    # re-pinned on the moved positions
    s.pin_growth_positions()
    s.set_time(600.0)
    s.grow(dt)
    rel_pin = GR.relative_positions(geo, pos)
    err = rel_err(s.growth_increments(), GR.increments(fns, geo, pos, 600.0, dt, bool(pinned), interp, rel_pin))
    assert err < 1e-13
    # the time advances with the timesteps; an empty list returns to the rates
    s.set_time(0.0)
    s.runOneStep(dt)
    assert s.get_time() == dt
    s.set_growth_functions([], geo.initial_z, geo.rel_src)
    with pytest.raises(Exception):
        s.growth_increments()
    print(f"device growth functions (pinned={pinned}, interpolation={interp}): max rel error {max(worst, err):.1e}",
          flush=True)
