"""GPU parity of implicit packing (through the C ABI) against the oracle.

Bars: node masses within 1e-13 relative; detection gives the oracle's pair list exactly (same
pairs, same order) with weights within 1e-14 and thresholds within 1e-13; assembled g and M with
active pairs within 1e-10 relative of the oracle's (fixed DoFs and binding included);
trajectories within 1e-8 relative with the same NR iteration and pair counts per step.
"""
import numpy as np
import pytest

from conftest import rel_err
from test_oracle_constraints import bound_pairs
from test_oracle_packing import max_first, permuted
from tissuefolding_paperversion_amd import configs
from tissuefolding_paperversion_amd import _lib as L

pytestmark = pytest.mark.gpu


def _solver(tissue, **opts):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    return NewtonRaphsonSolver(tissue, opts or None)


def _facing(perm=None, **kw):
    t = configs.facing_discs_tissue(**kw)
    if perm == "max_first":
        return max_first(t)
    return t if perm is None else permuted(t, perm)


def test_node_masses_match_oracle(oracle):
    t = _facing()
    s = _solver(t)
    rng = np.random.default_rng(8)
    Fg = np.tile(np.eye(3), (t.n_elements, 1, 1)) + 0.05 * rng.standard_normal((t.n_elements, 3, 3))
    s.set_fg(Fg)
    m = s.node_masses()
    ref = oracle.node_masses(t.n_nodes, t.mesh.conn, t.mesh.xref, Fg)
    assert rel_err(m, ref) < 1e-13
    s.set_ablated_nodes([3, 4])
    m2 = s.node_masses()
    assert m2[3] == 0.0 and m2[4] == 0.0


@pytest.mark.parametrize("perm", [None, 5, 9, "max_first"])
def test_detection_matches_oracle(oracle, perm):
    t = _facing(perm)
    s = _solver(t)
    n = s.detectPacingNodes()
    o = oracle.OracleTissue(t)
    n_o = o.detect_packing()
    print(f"facing discs perm={perm}: {n} pairs (oracle {n_o})", flush=True)
    assert n == n_o > 20
    pairs, w, T, grid = s.packing_pairs()
    np.testing.assert_array_equal(pairs, o.pairs)
    assert np.max(np.abs(w - o.pw)) < 1e-14
    assert abs(T[0] - o.pT) < 1e-13 * o.pT
    assert rel_err(grid[0], o.pgrid) < 1e-13


def test_detection_per_tissue_in_a_batch(oracle):
    ts = [_facing(), configs.disc_tissue(radius=15.0, n_steps=1), _facing(7), _facing(gap=3.5)]
    s = _solver(ts)
    n = s.detectPacingNodes()
    pairs, w, T, grid = s.packing_pairs()
    off = np.cumsum([0] + [t.n_nodes for t in ts])
    exp_p, exp_w = [], []
    for k, t in enumerate(ts):
        o = oracle.OracleTissue(t)
        o.detect_packing()
        exp_p.append(o.pairs + off[k])
        exp_w.append(o.pw)
        assert abs(T[k] - o.pT) < 1e-13 * o.pT
        assert rel_err(grid[k], o.pgrid) < 1e-13
    exp_p = np.concatenate(exp_p)
    print(f"batch of {len(ts)} tissues: {n} pairs", flush=True)
    assert n == len(exp_p)
    np.testing.assert_array_equal(pairs, exp_p)
    assert np.max(np.abs(w - np.concatenate(exp_w))) < 1e-14


def _assembly_case(oracle, t, perturb=0.05, bind=False, **opts):
    s = _solver(t, **opts)
    o = oracle.OracleTissue(t)
    if bind:
        pairs = bound_pairs(t)
        s.set_bound_dofs(pairs)
        o.bound = pairs
    rng = np.random.default_rng(4)
    X = t.mesh.pos.ravel().copy()
    x = X + perturb * rng.standard_normal(X.shape)
    s.updateNodeViscositySurfaces()
    n = s.detectPacingNodes()
    s.constructUnMatrix()
    s.positions = x
    s.calculateForcesAndJacobianMatrixNR(t.dt)
    o.update_viscosity_surfaces()
    assert o.detect_packing() == n > 0
    K, g = o.assemble(x, X, t.dt, dense=True)
    return s, o, K, g


@pytest.mark.parametrize("perm", [None, 9])
def test_assembly_with_pairs_matches_oracle(oracle, perm):
    t = _facing(perm, eta_ext=20.0)
    s, o, K, g = _assembly_case(oracle, t)
    assert rel_err(s.gSum, g) < 1e-10
    assert rel_err(s.jacobian_csr().toarray(), K.T) < 1e-10
    du = s.solveForDeltaU()
    assert rel_err(du, o.solve(K, g)) < 1e-8


def test_assembly_with_pairs_and_bound_dofs(oracle):
    t = _facing(eta_ext=20.0)
    s, o, K, g = _assembly_case(oracle, t, bind=True)
    assert rel_err(s.gSum, g) < 1e-10
    assert rel_err(s.jacobian_csr().toarray(), K.T) < 1e-10


def test_fp32_storage_with_pairs(oracle):
    t = _facing(eta_ext=20.0)
    s, o, K, g = _assembly_case(oracle, t, matrix_precision=L.PREC_FP32)
    assert rel_err(s.gSum, g) < 1e-10
    assert rel_err(s.jacobian_csr().toarray(), K.T) < 1e-10  # FP64 copy; the solve reads FP32


def test_pairs_from_the_host_match_detection(oracle):
    t = _facing(eta_ext=20.0)
    s, o, K, g = _assembly_case(oracle, t)
    g_det, M_det = s.gSum, s.jacobian_csr().toarray()
    s2 = _solver(t)
    s2.updateNodeViscositySurfaces()
    s2.set_packing_pairs(o.pairs, o.pw, [o.pT])
    s2.constructUnMatrix()
    rng = np.random.default_rng(4)
    X = t.mesh.pos.ravel().copy()
    s2.positions = X + 0.05 * rng.standard_normal(X.shape)
    s2.calculateForcesAndJacobianMatrixNR(t.dt)
    assert rel_err(s2.gSum, g_det) < 1e-13
    assert rel_err(s2.jacobian_csr().toarray(), M_det) < 1e-13
    with pytest.raises(L.TfError):
        s2.set_packing_pairs([[0, 1], [1, 0]], np.ones((2, 3)), [1.0])


@pytest.mark.parametrize("perm,krylov", [(None, L.KRYLOV_CG1), (None, L.KRYLOV_BICGSTAB), (9, L.KRYLOV_CG1)])
def test_trajectory_with_packing_matches_oracle(oracle, perm, krylov):
    t = _facing(perm)
    s = _solver(t, krylov=krylov)
    o = oracle.OracleTissue(t)
    for step in range(4):
        st = s.runOneStep(t.dt)
        it_o, _ = o.timestep(t.dt)
        print(f"step {step}: pairs {st['packing_pairs']} / {o.stats[-1]['packing_pairs']}, "
              f"NR {st['nr_iterations']} / {it_o}", flush=True)
        assert st["converged"] == 1
        assert st["packing_pairs"] == o.stats[-1]["packing_pairs"]
        assert st["nr_iterations"] == it_o
        assert rel_err(s.positions, o.pos) < 1e-8


def test_packing_switch_and_adjoint_guard():
    t = _facing()
    s = _solver(t, implicit_packing=0)
    st = s.runOneStep(t.dt)
    assert st["packing_pairs"] == 0
    s.set_options(implicit_packing=1)
    st = s.runOneStep(t.dt)
    assert st["packing_pairs"] > 0
    with pytest.raises(L.TfError, match="packing"):
        s.adjoint_step(t.dt, np.zeros(3 * t.n_nodes), np.tile(np.eye(3), (t.n_elements, 1, 1)), np.zeros(t.n_elements))
