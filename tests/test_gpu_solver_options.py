"""Inexact-Newton forcing, Krylov warm starts and the XCD tile map change how the linear
systems are solved, not the Newton iterates they produce: against exact solves from zero the NR
iteration counts must be equal and the positions within 1e-10 relative, with fewer Krylov
iterations."""
import numpy as np
import pytest

from conftest import rel_err
from tissuefolding_paperversion_amd import configs

pytestmark = pytest.mark.gpu


def _solver(tissue, **opts):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    return NewtonRaphsonSolver(tissue, opts or None)


def _ellipse():
    t = configs.wing_disc_tissue(a=60.0, b=36.0, h=3.66, z_levels=(0.0, 2.0, 7.0, 12.0), seed=2, growth_seed=3,
                                 name="options ellipse")
    return t


def test_forcing_and_warm_start_keep_the_newton_iterates():
    t = _ellipse()
    exact = _solver(t, krylov_forcing=0, krylov_warm_start=0)
    fast = _solver(t)
    assert fast.options()["krylov_forcing"] == 1 and fast.options()["krylov_warm_start"] == 1
    kr_e = kr_f = 0
    for step in range(6):
        a, b = exact.runOneStep(t.dt), fast.runOneStep(t.dt)
        assert a["converged"] == 1 and b["converged"] == 1
        assert a["nr_iterations"] == b["nr_iterations"]
        if step >= 2:
            kr_e += a["krylov_iterations"]
            kr_f += b["krylov_iterations"]
        assert rel_err(fast.positions, exact.positions) < 1e-10
    assert kr_f < 0.8 * kr_e, (kr_f, kr_e)


def test_warm_start_alone():
    t = _ellipse()
    cold = _solver(t, krylov_forcing=0, krylov_warm_start=0)
    warm = _solver(t, krylov_forcing=0, krylov_warm_start=1)
    kc = kw = 0
    for step in range(5):
        a, b = cold.runOneStep(t.dt), warm.runOneStep(t.dt)
        assert a["nr_iterations"] == b["nr_iterations"]
        if step >= 1:
            kc += a["krylov_iterations"]
            kw += b["krylov_iterations"]
    assert rel_err(warm.positions, cold.positions) < 1e-10
    assert kw < kc, (kw, kc)


def test_xcd_tile_map_is_bitwise_neutral():
    t = _ellipse()
    a, b = _solver(t, xcd_tile_map=0), _solver(t, xcd_tile_map=1)
    for _ in range(2):
        a.runOneStep(t.dt)
        b.runOneStep(t.dt)
    assert np.array_equal(a.positions, b.positions)


@pytest.mark.parametrize("prec", [1, 0])  # two-level and block Jacobi
def test_inverted_prism_stops_the_step(prec):
    """A prism turned inside out (det Fe <= 0 at its Gauss points) ends the NR loop with
    TF_NEG_JACOBIAN before any update, with or without a warm start."""
    from tissuefolding_paperversion_amd import _lib as L
    t = _ellipse()
    s = _solver(t, preconditioner=prec)
    s.runOneStep(t.dt)
    s.runOneStep(t.dt)  # the warm-start history is filled
    x = s.positions.reshape(-1, 3).copy()
    conn = t.mesh.conn
    e = 0
    top, bottom = conn[e, 3:], conn[e, :3]
    x[top, 2] = x[bottom, 2].min() - 5.0  # apical triangle pushed below the basal one
    s.positions = x.ravel()
    before = s.positions
    with pytest.raises(L.TfError) as err:
        s.runOneStep(t.dt)
    assert err.value.code == L.TF_NEG_JACOBIAN
    assert "det Fe" in str(err.value)
    assert np.array_equal(s.positions, before)


@pytest.mark.parametrize("prec", [1, 0], ids=["two_level", "block_jacobi"])
def test_one_read_per_newton_iteration_is_bitwise_neutral(prec):
    """With a prediction from the previous timestep, a Newton iteration runs its assembly, warm
    start, solve, update and convergence test on the device behind one gate, and the host reads one
    report per iteration (nr_graph = 1).  The iterates are those of the host-checked loop
    (nr_graph = 0), bit for bit, and from the third timestep on the host waits once per Newton
    iteration, or a little more where a prediction falls short."""
    t = _ellipse()
    host, dev = _solver(t, nr_graph=0, preconditioner=prec), _solver(t, nr_graph=1, preconditioner=prec)
    waits = nr = 0
    for step in range(6):
        a, b = host.runOneStep(t.dt), dev.runOneStep(t.dt)
        assert a["converged"] == 1 and b["converged"] == 1
        assert (a["nr_iterations"], a["krylov_iterations"]) == (b["nr_iterations"], b["krylov_iterations"])
        assert np.array_equal(host.positions, dev.positions), step
        assert a["host_waits"] > a["nr_iterations"]
        if step >= 2:
            waits += b["host_waits"]
            nr += b["nr_iterations"]
    assert waits <= 1.5 * nr, (waits, nr)


@pytest.mark.parametrize("case", ["two_level", "block_jacobi", "fp32_first_iteration"])
def test_newton_iteration_graphs_replay_the_one_read_iterates(case):
    """nr_graph = 2 replays a Newton iteration (assembly, warm start, two-level set-up, the solve's
    first burst, update, convergence test) from one captured hipGraph per shape of its launches.
    The iterates equal those of the eagerly enqueued one-read path bit for bit (graphs captured on
    first use, replayed afterwards), the host reads what the eager path reads, and the two-level
    solves run as the same persistent launches."""
    t = _ellipse()
    opts = {"preconditioner": 0} if case == "block_jacobi" else {"preconditioner": 1}
    if case == "fp32_first_iteration":
        opts.update(matrix_precision=1)
    eager, graph = _solver(t, nr_graph=1, **opts), _solver(t, nr_graph=2, **opts)
    builds = []
    waits = nr = 0
    for step in range(8):
        a, b = eager.runOneStep(t.dt), graph.runOneStep(t.dt)
        assert a["converged"] == 1 and b["converged"] == 1
        assert (a["nr_iterations"], a["krylov_iterations"]) == (b["nr_iterations"], b["krylov_iterations"])
        assert np.array_equal(eager.positions, graph.positions), step
        assert b["host_waits"] == a["host_waits"], step  # the same reads, from a graph or eager launches
        # the persistent two-level solve runs inside the captured iterations too
        assert b["krylov_persistent"] == a["krylov_persistent"], step
        builds.append(b["nr_graph_builds"])
        if step >= 2:
            waits += b["host_waits"]
            nr += b["nr_iterations"]
    assert builds[0] == 0 and sum(builds) > 0  # nothing to predict from in the first timestep
    assert waits <= 1.75 * nr, (waits, nr)  # one read per iteration, more where a prediction falls short


def test_newton_iteration_graphs_of_several_tissues():
    """A handle of several tissues under nr_graph = 2: each tissue's positions equal those of the
    eager one-read path bit for bit."""
    ts = [configs.batch_disc(n, layers, E, 0.3, rate, seed)
          for n, layers, E, rate, seed in [(600, 2, 800.0, 1.0, 1), (1500, 3, 2000.0, 2.5, 2), (900, 2, 4000.0, 0.7, 3)]]
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    eager = NewtonRaphsonSolver(ts, {"nr_graph": 1})
    graph = NewtonRaphsonSolver(ts, {"nr_graph": 2})
    for step in range(6):
        a, b = eager.runOneStep(ts[0].dt), graph.runOneStep(ts[0].dt)
        assert a["converged"] == 1 and b["converged"] == 1
        assert np.array_equal(eager.positions, graph.positions), step
