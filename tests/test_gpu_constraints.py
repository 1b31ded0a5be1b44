"""GPU parity of slave/master DoF binding and ablated nodes (through the C ABI) against the oracle.

Bars as for the unconstrained path: assembled g and M within 1e-10 relative of the oracle's
g_bound and K_bound^T, trajectories within 1e-8 relative with the same NR iteration counts.
"""
import numpy as np
import pytest

from conftest import rel_err
from test_oracle_constraints import bound_pairs, remove_elements_around
from tissuefolding_paperversion_amd import configs
from tissuefolding_paperversion_amd import _lib as L
from tissuefolding_paperversion_amd import mesh as M

pytestmark = pytest.mark.gpu


def _solver(tissue, **opts):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    return NewtonRaphsonSolver(tissue, opts or None)


def _disc(eta_ext=30.0):
    t = configs.disc_tissue(radius=12.0, h=4.0, eta_ext=eta_ext, name="bound disc")
    t.rates[:, :2] = 0.2 / 3600.0
    return t


def test_bound_assembly_matches_oracle(oracle):
    t = _disc()
    pairs = bound_pairs(t)
    s = _solver(t)
    nb0 = s.nb
    s.set_bound_dofs(pairs)
    assert s.nb >= nb0
    rng = np.random.default_rng(4)
    X = t.mesh.pos.ravel().copy()
    x = X + 0.05 * rng.standard_normal(X.shape)
    s.positions = X
    s.updateNodeViscositySurfaces()
    s.constructUnMatrix()
    s.positions = x
    s.calculateForcesAndJacobianMatrixNR(60.0)
    o = oracle.OracleTissue(t)
    o.pos = X.copy()
    o.update_viscosity_surfaces()
    o.bound = pairs
    K, g = o.assemble(x, X, 60.0, dense=True)
    assert rel_err(s.gSum, g) < 1e-10
    assert rel_err(s.jacobian_csr().toarray(), K.T) < 1e-10
    du = s.solveForDeltaU()
    ref = o.equate_slaves(o.solve(K, g))
    assert rel_err(du, ref) < 1e-8
    for sd, md in pairs:
        assert du[sd] == du[md]


@pytest.mark.parametrize("krylov", [L.KRYLOV_CG1, L.KRYLOV_BICGSTAB])
def test_bound_trajectory_matches_oracle(oracle, krylov):
    t = _disc()
    pairs = bound_pairs(t)
    s = _solver(t, krylov=krylov)
    s.set_bound_dofs(pairs)
    o = oracle.OracleTissue(t)
    o.bound = pairs
    X = t.mesh.pos.ravel().copy()
    for _ in range(3):
        st = s.runOneStep(t.dt)
        it_o, _ = o.timestep(t.dt)
        assert st["converged"] == 1
        assert st["nr_iterations"] == it_o
        assert rel_err(s.positions, o.pos) < 1e-8
    u = s.positions - X
    for sd, md in pairs:
        assert abs(u[sd] - u[md]) < 1e-12 * np.max(np.abs(u))


def test_unbinding_restores_plain_path():
    t = _disc()
    a, b = _solver(t), _solver(t)
    b.set_bound_dofs(bound_pairs(t))
    b.set_bound_dofs([])
    assert a.nb == b.nb
    for _ in range(2):
        a.runOneStep(t.dt)
        b.runOneStep(t.dt)
    assert np.array_equal(a.positions, b.positions)


def test_invalid_binding_rejected():
    t = _disc()
    s = _solver(t)
    n = 3 * t.n_nodes
    for bad in ([(0, 0)], [(n, 1)], [(-1, 1)], [(3, 4), (3, 5)], [(3, 4), (6, 3)], [(3, 4), (4, 7)]):
        with pytest.raises(L.TfError) as ex:
            s.set_bound_dofs(bad)
        assert ex.value.code == L.TF_INVALID
    ts = [configs.disc_tissue(radius=8.0, h=4.0, seed=3, name="a"), configs.disc_tissue(radius=9.0, h=4.0, seed=4, name="b")]
    sb = _solver(ts)
    with pytest.raises(L.TfError):
        sb.set_bound_dofs([(0, 3 * ts[0].n_nodes + 1)])   # across tissues


def test_ablated_node_matches_oracle(oracle):
    t = configs.disc_tissue(radius=12.0, h=4.0, eta_ext=0.0, name="ablated disc")
    t.rates[:, :2] = 0.2 / 3600.0
    apical = np.where(t.mesh.node_placement == M.APICAL)[0]
    node = int(apical[len(apical) // 2])
    ta = remove_elements_around(t, node)
    s = _solver(ta)
    s.set_ablated_nodes([node])
    o = oracle.OracleTissue(ta)
    o.ablated = [node]
    X = ta.mesh.pos.ravel().copy()
    rng = np.random.default_rng(5)
    x = X + 0.05 * rng.standard_normal(X.shape)
    s.positions = X
    s.constructUnMatrix()
    s.positions = x
    s.calculateForcesAndJacobianMatrixNR(60.0)
    K, g = o.assemble(x, X, 60.0, dense=True)
    assert rel_err(s.jacobian_csr().toarray(), K.T) < 1e-10
    assert rel_err(s.gSum, g) < 1e-10
    s.positions = X
    o.pos = X.copy()
    for _ in range(2):
        st = s.runOneStep(ta.dt)
        it_o, _ = o.timestep(ta.dt)
        assert st["converged"] == 1 and st["nr_iterations"] == it_o
        assert rel_err(s.positions, o.pos) < 1e-8
    assert np.array_equal(s.positions[3 * node:3 * node + 3], X[3 * node:3 * node + 3])
