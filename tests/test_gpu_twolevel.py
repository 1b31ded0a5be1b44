"""Two-level preconditioner (block Jacobi plus a rigid-body coarse space over node aggregates).

The preconditioner changes how each linear system is solved, not its solution.  Bars:
- single solves within 1e-8 relative of the block-Jacobi solve and of a direct sparse solve of
  the same system (krylov_rtol 1e-12), in fewer Krylov iterations;
- fixed DoFs and bound slaves of du exactly as block Jacobi leaves them (zero, or the master's);
- trajectories with the block-Jacobi NR iteration counts and positions within 1e-10 relative
  (1e-8 against the oracle);
- handles with several tissues fall back to block Jacobi (bitwise equal results).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import rel_err
from test_oracle_constraints import bound_pairs
from tissuefolding_paperversion_amd import configs
from tissuefolding_paperversion_amd import _lib as L

pytestmark = pytest.mark.gpu

TL = dict(preconditioner=L.PRECOND_TWO_LEVEL)


def _solver(tissue, **opts):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    return NewtonRaphsonSolver(tissue, opts or None)


def _ellipse():
    return configs.wing_disc_tissue(a=60.0, b=36.0, h=3.66, z_levels=(0.0, 2.0, 7.0, 12.0), seed=2, growth_seed=3,
                                    name="two-level ellipse")


def _assembled(t, perturb=0.05, bind=False, **opts):
    s = _solver(t, **opts)
    if bind:
        s.set_bound_dofs(bound_pairs(t))
    rng = np.random.default_rng(11)
    X = t.mesh.pos.ravel().copy()
    s.updateNodeViscositySurfaces()
    s.constructUnMatrix()
    s.positions = X + perturb * rng.standard_normal(X.shape)
    s.calculateForcesAndJacobianMatrixNR(t.dt)
    return s


@pytest.mark.parametrize("make,aggregates", [(_ellipse, 0), (_ellipse, 7), (lambda: configs.disc_tissue(radius=30.0), 0)])
def test_single_solve_matches_block_jacobi(make, aggregates):
    t = make()
    bj = _assembled(t)
    du_bj = bj.solveForDeltaU()
    it_bj = bj.last_krylov["iterations"]
    tl = _assembled(t, coarse_aggregates=aggregates, **TL)
    du_tl = tl.solveForDeltaU()
    it_tl = tl.last_krylov["iterations"]
    print(f"{t.name} aggregates={aggregates}: block Jacobi {it_bj} its, two-level {it_tl} its", flush=True)
    assert tl.last_krylov["rel_residual"] <= 1e-12
    assert rel_err(du_tl, du_bj) < 1e-8
    ref = spla.spsolve(tl.jacobian_csr().tocsc(), tl.gSum)  # direct sparse solve of the same system
    assert rel_err(du_tl, ref) < 1e-8
    fixed = np.flatnonzero(t.fixed)
    assert np.all(du_tl[fixed] == 0.0)
    assert it_tl < it_bj


def test_aggregate_option_rebuilds_coarse_space():
    """The coarse space is built when the handle is created; setting coarse_aggregates afterwards
    (as the constructor's options do) rebuilds it, which shows in the iteration count."""
    t = _ellipse()
    its, dus = {}, {}
    for a in (0, 7):
        s = _assembled(t, coarse_aggregates=a, **TL)
        dus[a] = s.solveForDeltaU()
        its[a] = s.last_krylov["iterations"]
    print(f"two-level iterations: default aggregates {its[0]}, 7 aggregates {its[7]}", flush=True)
    assert its[7] > its[0]
    assert rel_err(dus[7], dus[0]) < 1e-8


@pytest.mark.parametrize("make,thin", [(_ellipse, True), (lambda: configs.disc_tissue(radius=30.0), False)])
def test_column_blocks_and_nodal_blocks_agree(make, thin):
    """The smoother on node columns (4 nodes on the ellipse, 3 on the two-layer disc) and the
    3x3 nodal one (column_blocks = 0) solve the same system to the same answer.  The columns take
    fewer iterations where the layers are thin against the in-plane spacing (the ellipse: 2-5 um
    layers, 3.7 um edges) and about as many where they are not (the disc: 5 um layers, 5.85 um
    edges)."""
    t = make()
    du, its = {}, {}
    for cb in (1, 0):
        s = _assembled(t, column_blocks=cb, **TL)
        du[cb] = s.solveForDeltaU()
        its[cb] = s.last_krylov["iterations"]
        assert s.last_krylov["rel_residual"] <= 1e-12
    ref = spla.spsolve(s.jacobian_csr().tocsc(), s.gSum)
    print(f"{t.name}: column blocks {its[1]} its, nodal blocks {its[0]} its", flush=True)
    assert rel_err(du[1], ref) < 1e-8 and rel_err(du[0], ref) < 1e-8
    assert its[1] < its[0] if thin else its[1] <= its[0] + 2


def test_bound_slaves_follow_masters():
    t = configs.disc_tissue(radius=30.0)
    pairs = np.asarray(bound_pairs(t))
    bj = _assembled(t, bind=True)
    tl = _assembled(t, bind=True, **TL)
    du_bj, du_tl = bj.solveForDeltaU(), tl.solveForDeltaU()
    assert rel_err(du_tl, du_bj) < 1e-8
    assert np.array_equal(du_tl[pairs[:, 0]], du_tl[pairs[:, 1]])


@pytest.mark.parametrize("prec", [L.PREC_FP64, L.PREC_FP32])
def test_trajectory_keeps_the_newton_iterates(prec):
    t = _ellipse()
    bj = _solver(t, matrix_precision=prec)
    tl = _solver(t, matrix_precision=prec, **TL)
    kb = kt = 0
    for step in range(6):
        a, b = bj.runOneStep(t.dt), tl.runOneStep(t.dt)
        print(f"step {step}: NR {a['nr_iterations']} / {b['nr_iterations']}, "
              f"Krylov {a['krylov_iterations']} / {b['krylov_iterations']}", flush=True)
        assert a["converged"] == 1 and b["converged"] == 1
        assert a["nr_iterations"] == b["nr_iterations"]
        kb += a["krylov_iterations"]
        kt += b["krylov_iterations"]
        assert rel_err(tl.positions, bj.positions) < 1e-10
    assert kt < kb, (kt, kb)


def test_trajectory_matches_oracle_with_fixed_dofs(oracle):
    t = configs.disc_tissue(radius=20.0)
    s = _solver(t, **TL)
    o = oracle.OracleTissue(t)
    X0 = t.mesh.pos.ravel().copy()
    fixed = np.flatnonzero(t.fixed)
    for step in range(4):
        st = s.runOneStep(t.dt)
        it_o, _ = o.timestep(t.dt)
        assert st["converged"] == 1
        assert st["nr_iterations"] == it_o
        assert rel_err(s.positions, o.pos) < 1e-8
        assert np.array_equal(s.positions[fixed], X0[fixed])


def test_packing_pairs_with_two_level(oracle):
    t = configs.facing_discs_tissue()
    s = _solver(t, **TL)
    o = oracle.OracleTissue(t)
    for step in range(3):
        st = s.runOneStep(t.dt)
        it_o, _ = o.timestep(t.dt)
        assert st["packing_pairs"] == o.stats[-1]["packing_pairs"] > 0
        assert st["nr_iterations"] == it_o
        assert rel_err(s.positions, o.pos) < 1e-8


def test_several_tissues_match_their_own_handles(oracle):
    """Two-level on a handle of several tissues: every tissue has its own aggregates and its own
    block of the coarse operator, its partial sums and rebuild decisions are its own, so each
    tissue's Newton and Krylov iterations and positions equal a handle of its own bit for bit; the
    first tissue also matches the oracle."""
    ts = [_ellipse(), configs.wing_disc_tissue(a=50.0, b=40.0, h=3.66, seed=5, growth_seed=6, n_steps=3, name="e2"),
          configs.wing_disc_tissue(a=70.0, b=30.0, h=3.66, seed=7, growth_seed=8, n_steps=3, name="e3")]
    batch = _solver(ts, **TL)
    singles = [_solver(t, **TL) for t in ts]
    o = oracle.OracleTissue(ts[0])
    offs = np.cumsum([0] + [3 * t.n_nodes for t in ts])
    for step in range(3):
        st = batch.runOneStep(60.0)
        assert st["converged"] == 1 and st["two_level"] == 1 and st["krylov_fallbacks"] == 0
        it = batch.segment_stats()[0]
        for k, s in enumerate(singles):
            sk = s.runOneStep(60.0)
            assert sk["two_level"] == 1 and it[k] == sk["nr_iterations"], (step, k)
        it_o, _ = o.timestep(60.0)
        assert it[0] == it_o
    xb = batch.positions
    for k, s in enumerate(singles):
        assert np.array_equal(xb[offs[k]:offs[k + 1]], s.positions), k
    assert rel_err(xb[:offs[1]], o.pos) < 1e-8


def test_fp64_preconditioner_storage(oracle):
    """precond_fp64 holds the column-block inverses and the coarse inverse in FP64, so every array
    of the solve is FP64: the same Newton iterates as the default FP32/FP16 storage (both within
    1e-8 of the oracle, same Newton counts) at a similar Krylov count."""
    t = _ellipse()
    a, b = _solver(t, **TL), _solver(t, precond_fp64=1, **TL)
    o = oracle.OracleTissue(t)
    ka = kb = 0
    for _ in range(3):
        sa, sb = a.runOneStep(t.dt), b.runOneStep(t.dt)
        it_o, _ = o.timestep(t.dt)
        assert sa["two_level"] == 1 and sb["two_level"] == 1
        assert sa["nr_iterations"] == sb["nr_iterations"] == it_o
        ka += sa["krylov_iterations"]
        kb += sb["krylov_iterations"]
    print(f"Krylov iterations over 3 steps: FP32/FP16 storage {ka}, FP64 storage {kb}", flush=True)
    assert kb <= 1.2 * ka
    assert rel_err(a.positions, o.pos) < 1e-8 and rel_err(b.positions, o.pos) < 1e-8
    assert b.kernel_bytes("coarse") > 1.5 * a.kernel_bytes("coarse")  # 8-byte inverses instead of 2 and 4


def test_option_validation():
    t = configs.disc_tissue(radius=10.0, n_steps=1)
    s = _solver(t)
    with pytest.raises(L.TfError):
        s.set_options(preconditioner=2)
    with pytest.raises(L.TfError):
        s.set_options(coarse_aggregates=-1)
    with pytest.raises(L.TfError):
        s.set_options(column_blocks=2)
    with pytest.raises(L.TfError):
        s.set_options(precond_fp64=2)
    with pytest.raises(L.TfError):
        s.set_options(krylov=2)  # the unused solver kind


@pytest.mark.parametrize("make", [_ellipse, configs.config2], ids=["ellipse", "config2"])
def test_coarse_inverse_matches_a_dense_inverse(make):
    """The coarse operator's inverse from the blocked Gauss-Jordan sweep (FP64, 64-wide pivot
    blocks, no pivoting) against numpy's LU inverse of the same operator: within 1e-10 relative,
    symmetric to rounding; the operator itself symmetric positive definite."""
    s = _assembled(make(), **TL)
    s.solveForDeltaU()  # a solve outside the Newton loop rebuilds the coarse operator
    Ac, Ai = s.coarse_operator()
    nc = Ac.shape[0]
    print(f"coarse operator: {nc} unknowns", flush=True)
    assert nc > 64  # more than one pivot block
    assert np.max(np.abs(Ac - Ac.T)) <= 1e-12 * np.max(np.abs(Ac))
    np.linalg.cholesky(Ac)
    ref = np.linalg.inv(Ac)
    assert np.max(np.abs(Ai - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert np.array_equal(Ai, Ai.T)


@pytest.mark.parametrize("case", ["ellipse", "ellipse_fp64_storage", "ellipse_cooperative", "config2"])
def test_persistent_solve_equals_two_launches(case):
    """tl_persistent runs a two-level solve as one launch (grid barriers between the halves of an
    iteration, the coarse rows held in LDS) instead of two launches per iteration: an ordinary launch
    of a grid the device holds at once (1), or a cooperative launch (2).  The arithmetic is the same
    in the same order: positions, Newton and Krylov counts are equal bit for bit over several
    timesteps."""
    t = configs.config2() if case == "config2" else _ellipse()
    opts = dict(TL)
    if case == "ellipse_fp64_storage":
        opts["precond_fp64"] = 1
    steps = 3 if case == "config2" else 5
    mode = 2 if case == "ellipse_cooperative" else 1
    two, one = _solver(t, tl_persistent=0, **opts), _solver(t, tl_persistent=mode, **opts)
    for step in range(steps):
        a, b = two.runOneStep(t.dt), one.runOneStep(t.dt)
        assert a["converged"] == 1 and b["converged"] == 1
        assert a["two_level"] == 1 and b["two_level"] == 1
        assert (a["nr_iterations"], a["krylov_iterations"], a["krylov_by_newton"]) == \
               (b["nr_iterations"], b["krylov_iterations"], b["krylov_by_newton"]), step
        assert np.array_equal(two.positions, one.positions), step
        assert a["krylov_persistent"] == 0 and b["krylov_persistent"] == b["nr_iterations"]


def test_persistent_barrier_timeout_falls_back():
    """A grid-barrier wait of the persistent solve that gives up (here forced from the host with a
    limit of one poll, as a grid that is not co-resident would meet it) sets the abort word: every
    workgroup leaves the launch, the solve is reported stalled and the step falls back to BiCGSTAB
    from zero.  The steps still converge, to the positions of the persistent run."""
    t = _ellipse()
    ref, forced = _solver(t, **TL), _solver(t, tl_barrier_polls=1, **TL)
    fell_back = 0
    for step in range(3):
        a, b = ref.runOneStep(t.dt), forced.runOneStep(t.dt)
        assert a["converged"] == 1 and b["converged"] == 1, step
        assert a["krylov_fallbacks"] == 0
        fell_back += b["krylov_fallbacks"]
        scale = np.max(np.abs(ref.positions))
        assert np.max(np.abs(ref.positions - forced.positions)) <= 1e-8 * scale, step
    assert fell_back > 0
