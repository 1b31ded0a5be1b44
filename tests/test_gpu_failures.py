"""Failure statuses of the NR step, each forced from the host so that every kernel still exits
normally: the iteration cap of the NR loop (SC/Simulation.cpp:6567-6571: after maxIteration
iterations the reference prints "did not converge" and continues), a linear solve that cannot
reach its tolerance (the reference's Pardiso failures exit the process,
SC/NewtonRaphsonSolver.cpp:579-600; here the step returns TF_KRYLOV_STALL), a singular Jacobian,
and a position that is not finite.  After each one the handle stays usable."""
import ctypes as C
import time

import numpy as np
import pytest

from conftest import rel_err
from tissuefolding_paperversion_amd import _lib as L
from tissuefolding_paperversion_amd import configs

pytestmark = pytest.mark.gpu

EXACT = dict(krylov_forcing=0, krylov_warm_start=0)


def _solver(tissue, **opts):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    return NewtonRaphsonSolver(tissue, opts or None)


def _ellipse():
    return configs.wing_disc_tissue(a=60.0, b=36.0, h=3.66, z_levels=(0.0, 2.0, 7.0, 12.0), seed=2, growth_seed=3,
                                    name="failure ellipse")


def _raw_timestep(s, dt):
    st = L.StepStats()
    rc = s._lib.tf_timestep(s.h, dt, C.byref(st))
    return rc, st.as_dict()


def test_nr_cap_returns_not_converged_after_the_reference_count(oracle):
    """nr_max_iter = 1: the loop stops when the count exceeds the cap, so two updates are applied
    (the reference's iteratorK > maxIteration test), the step reports TF_NOT_CONVERGED, and the
    positions are the oracle's after the same two Newton iterations."""
    t = _ellipse()
    t.rates = t.rates * 20.0  # a first step that needs more than two Newton iterations
    s = _solver(t, nr_max_iter=1, **EXACT)
    out = s.runOneStep(t.dt)
    assert out["status"] == L.TF_NOT_CONVERGED
    assert out["nr_iterations"] == 2 and out["converged"] == 0
    o = oracle.OracleTissue(t)
    o.grow(t.dt)
    o.update_viscosity_surfaces()
    o.detect_packing()
    it, d = o.nr_step(t.dt, max_iter=1)
    assert it == 2 and d > 1e-8
    assert rel_err(s.positions, o.pos) < 1e-8
    # the handle goes on: the next step with the reference's cap converges
    s.set_options(nr_max_iter=20)
    out = s.runOneStep(t.dt)
    assert out["status"] == L.TF_OK and out["converged"] == 1


@pytest.mark.parametrize("prec", [0, 1])  # block Jacobi, two-level
def test_krylov_stall_is_reported_and_the_update_is_not_applied(prec):
    """A Krylov cap of one iteration at tolerance 1e-12: CG stops at the cap with a residual far
    above 1e-2 of the right-hand side (unusable), its BiCGSTAB fallback too, so the step returns
    TF_KRYLOV_STALL without applying the update, and the handle goes on with the default cap."""
    t = _ellipse()
    s = _solver(t, preconditioner=prec, krylov_max_iter=1, **EXACT)
    before = s.positions
    rc, st = _raw_timestep(s, t.dt)
    assert rc == L.TF_KRYLOV_STALL, L.STATUS_NAMES.get(rc, rc)
    assert "did not reach its tolerance" in s._lib.tf_last_error(s.h).decode()
    assert st["krylov_fallbacks"] >= 1 and st["converged"] == 0
    assert np.array_equal(s.positions, before)
    s.set_options(krylov_max_iter=20000)
    out = s.runOneStep(t.dt)
    assert out["status"] == L.TF_OK and out["converged"] == 1 and out["krylov_fallbacks"] == 0


def test_capped_solve_with_a_small_residual_is_an_inexact_newton_step():
    """A Krylov cap of 250 block-Jacobi CG iterations at tolerance 1e-12: the solves stop short of
    the tolerance (about 650 iterations on this mesh) but well below 1e-2 of the right-hand side
    (2e-5 for the first one, by a CPU run of the same PCG), so they are kept as inexact Newton
    steps (no BiCGSTAB restart), counted in krylov_stalls, and Newton still converges to the
    fixed point of exact solves."""
    t = _ellipse()
    exact = _solver(t, preconditioner=0, **EXACT)
    capped = _solver(t, preconditioner=0, krylov_max_iter=250, **EXACT)
    for _ in range(2):
        a, b = exact.runOneStep(t.dt), capped.runOneStep(t.dt)
        assert a["status"] == L.TF_OK and b["status"] == L.TF_OK
        assert b["converged"] == 1 and b["krylov_stalls"] >= 1 and b["krylov_fallbacks"] == 0
        assert a["krylov_stalls"] == 0
    assert rel_err(capped.positions, exact.positions) < 1e-8


def test_singular_jacobian_exits_at_the_iteration_cap():
    """No fixed DoFs and no external viscosity: M is singular (rigid-body modes).  A right-hand
    side with a net force has no solution, so the solve must stop at its cap (CG, then BiCGSTAB)
    with TF_KRYLOV_STALL -- in bounded time, with no fault."""
    t = _ellipse()
    t.fixed = np.zeros_like(t.fixed)
    t.eta_ext = np.zeros_like(t.eta_ext)
    s = _solver(t, krylov_max_iter=400, **EXACT)
    s.constructUnMatrix()
    s.calculateForcesAndJacobianMatrixNR(t.dt)
    b = np.zeros(3 * t.n_nodes)
    b[0::3] = 1.0  # a net x force: orthogonal to nothing in the null space of M
    t0 = time.perf_counter()
    with pytest.raises(L.TfError) as err:
        s.solve_system(b)
    assert err.value.code == L.TF_KRYLOV_STALL
    assert time.perf_counter() - t0 < 60.0


def test_nan_position_returns_nan_and_the_handle_recovers():
    """A NaN node position makes the residual not finite: the step returns TF_NAN before any
    update; with finite positions restored the next step converges as on a fresh handle."""
    t = _ellipse()
    s = _solver(t)
    ref = _solver(t)
    s.runOneStep(t.dt)
    ref.runOneStep(t.dt)
    good = s.positions
    x = good.copy()
    x[3 * 7 + 1] = np.nan
    s.positions = x
    with pytest.raises(L.TfError) as err:
        s.runOneStep(t.dt)
    assert err.value.code == L.TF_NAN
    assert "NaN" in str(err.value)
    # the growth update of the failed step has been applied; restore the reference handle's state
    s.positions = good
    Fg, InvFg = ref.get_fg()
    s.set_growth_state(Fg, InvFg)
    a = s.runOneStep(t.dt)
    b = ref.runOneStep(t.dt)
    assert a["status"] == L.TF_OK and a["converged"] == 1
    assert a["nr_iterations"] == b["nr_iterations"]
    assert rel_err(s.positions, ref.positions) < 1e-8


def test_kernel_probe_leaves_the_next_solve_unchanged():
    """The timing probe runs the solver with its own iteration cap.  The captured Krylov chunks
    carry the cap, so a later solve must not replay a chunk captured under the probe's cap: same
    Krylov counts, no fallback, identical positions."""
    t = _ellipse()
    a, b = _solver(t, **EXACT), _solver(t, **EXACT)
    for s in (a, b):
        s.runOneStep(t.dt)
    assert b.kernel_probe(64) > 0.0
    sa, sb = a.runOneStep(t.dt), b.runOneStep(t.dt)
    assert sa["krylov_iterations"] > 64  # long enough for the probe's cap to matter
    assert sb["krylov_fallbacks"] == 0 and sa["krylov_fallbacks"] == 0
    assert sb["krylov_iterations"] == sa["krylov_iterations"]
    assert np.array_equal(a.positions, b.positions)
