"""Parity at the benchmark sizes: the workloads the benchmark is quoted on, run whole on the GPU and
checked against the oracle (or, where the oracle has no counterpart, against the FP64 path or
finite differences).  The loop they exercise is Simulation::updateStepNR (SC/Simulation.cpp:6449-6590)
with runOneStep's growth and surface updates around it.

- config 2, the 38,826-prism wing-disc ellipse: two timesteps against the CPU oracle (literal element
  tangent, CSR assembly, direct transposed solve) within 1e-8 with the oracle's NR counts; all 120
  timesteps converge with no Krylov fallback and stay within 1e-8 of a run with exact linear solves
  (same NR counts at every step).
- config 3, the same mesh with FP32 and BF16 MFMA element tangents (FP32 Jacobian values): every one
  of 20 timesteps within 1e-8 of the FP64 run.
- config 4, 256 discs of 500-8000 prisms in one handle (two-level, one coarse block per disc): all
  60 timesteps converge for every disc; sampled discs equal their own single-handle runs bit for
  bit (NR counts too) and two small ones the oracle within 1e-8 over three steps.
- config 5, the adjoint of 30 timesteps of the ellipse: 64 parameters (3 layer moduli, 61 rate
  scales) against central differences, h = 1e-6, forward runs at nr_tol 1e-11 with exact solves.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_err
from tissuefolding_paperversion_amd import configs

pytestmark = pytest.mark.gpu


def _solver(tissue, **opts):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    return NewtonRaphsonSolver(tissue, opts or None)


# ---- config 2 ----------------------------------------------------------------------------------
def test_config2_two_timesteps_match_the_oracle(tmp_path):
    out = tmp_path / "oracle_config2.npz"
    print("config 2: oracle, 2 timesteps (subprocess)", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "oracle_steps.py"), "config2", "2", str(out)],
                       capture_output=True, text=True, timeout=110)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = np.load(out)
    t = configs.config2()
    assert t.n_elements == 38826
    s = _solver(t)
    for k in range(2):
        st = s.runOneStep(t.dt)
        assert st["converged"] == 1 and st["two_level"] == 1
        assert st["nr_iterations"] == int(ref["its"][k]), (k, st["nr_iterations"], ref["its"][k])
        err = rel_err(s.positions, ref["x"][k])
        print(f"config 2 step {k}: {st['nr_iterations']} NR iterations (oracle {ref['its'][k]}), "
              f"max rel position error {err:.2e}", flush=True)
        assert err <= 1e-8


def test_config2_all_120_timesteps():
    t = configs.config2()
    fast = _solver(t)
    exact = _solver(t, krylov_forcing=0, krylov_warm_start=0)
    nr = []
    for k in range(t.n_steps):
        a = fast.runOneStep(t.dt)
        b = exact.runOneStep(t.dt)
        assert a["converged"] == 1 and b["converged"] == 1, k
        assert a["krylov_fallbacks"] == 0 and a["krylov_stalls"] == 0, k
        assert a["nr_iterations"] == b["nr_iterations"], (k, a["nr_iterations"], b["nr_iterations"])
        nr.append(a["nr_iterations"])
    err = rel_err(fast.positions, exact.positions)
    print(f"config 2, 120 timesteps: NR iterations per step {np.mean(nr):.2f} (min {min(nr)}, max {max(nr)}); "
          f"forcing and warm starts vs exact solves: max rel position error {err:.2e}", flush=True)
    assert err <= 1e-8
    assert rel_err(fast.positions, t.mesh.pos.ravel()) > 1e-3  # the tissue grew


# ---- config 3 ----------------------------------------------------------------------------------
@pytest.mark.parametrize("tangent", [1, 2], ids=["fp32_mfma", "bf16_mfma"])
def test_config3_mfma_tangents_track_fp64(tangent):
    t = configs.config2()
    ref = _solver(t)
    mixed = _solver(t, tangent_precision=tangent, matrix_precision=1)
    worst = 0.0
    for k in range(20):
        a = ref.runOneStep(t.dt)
        b = mixed.runOneStep(t.dt)
        assert a["converged"] == 1 and b["converged"] == 1, k
        err = rel_err(mixed.positions, ref.positions)
        worst = max(worst, err)
        assert err <= 1e-8, (k, err)
    print(f"config 3, tangent precision {tangent}: max rel displacement error over 20 steps {worst:.2e}", flush=True)


# ---- config 4 ----------------------------------------------------------------------------------
def test_config4_full_sweep(oracle):
    ts = configs.config4()
    assert len(ts) == 256
    batch = _solver(ts)
    offs = np.cumsum([0] + [3 * t.n_nodes for t in ts])
    small = [int(k) for k in np.argsort([t.n_elements for t in ts])[:2]]
    ors = {k: oracle.OracleTissue(ts[k]) for k in small}
    iters, kry, nrs = [], 0, 0
    for step in range(60):
        st = batch.runOneStep(60.0)
        it, dn, cv = batch.segment_stats()
        assert st["converged"] == 1 and cv.all() and np.all(dn <= 1e-8), step
        assert st["krylov_fallbacks"] == 0 and st["two_level"] == 1, step
        kry += st["krylov_iterations"]
        nrs += st["nr_iterations"]
        iters.append(it.copy())
        if step < 3:
            xb = batch.positions
            for k, o in ors.items():
                it_o, _ = o.timestep(60.0)
                assert it[k] == it_o, (step, k, it[k], it_o)
                assert rel_err(xb[offs[k]:offs[k + 1]], o.pos) < 1e-8
    xb = batch.positions
    print(f"config 4: {sum(t.n_elements for t in ts)} prisms, NR iterations per disc-step "
          f"{np.mean(iters):.2f} (max {np.max(iters)}), Krylov iterations per NR iteration {kry / nrs:.1f} "
          "(two-level, one coarse block per disc)", flush=True)
    assert kry / nrs <= 40
    sizes = np.array([t.n_elements for t in ts])
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 128, 255})
    for k in sample:
        s = _solver(ts[k], preconditioner=1, column_blocks=0)  # two-level with nodal smoother blocks, as in the
        # handle of discs of 2 and 3 layers
        for step in range(60):
            st = s.runOneStep(60.0)
            assert st["nr_iterations"] == iters[step][k], (k, step)
        assert np.array_equal(s.positions, xb[offs[k]:offs[k + 1]]), k


# ---- config 5 ----------------------------------------------------------------------------------
ACCURATE = dict(nr_tol=1e-11, krylov_rtol=1e-13)


@pytest.fixture(scope="module")
def config5_gradient():
    from tissuefolding_paperversion_amd.adjoint import GrowthRollout, apical_loss
    from tissuefolding_paperversion_amd.mesh import APICAL

    t, s_target, T, K = configs.config5()
    assert (T, K, t.n_elements) == (30, 5, 38826)
    ro = GrowthRollout(t, T, checkpoint_every=K, opts=ACCURATE)
    E0 = ro.layer_moduli()
    target = ro.forward(s_target, E0)
    apical = np.where(np.repeat(t.mesh.node_placement == APICAL, 3))[0]
    s0 = np.ones(t.n_elements)
    L0, dLdx = apical_loss(ro.forward(s0, E0), target, apical)
    assert all(st["nr_iterations"] <= 6 and st["rotated_elements"] == 0 for st in ro.stats)
    s_bar, E_bar = ro.backward(dLdx)
    gmax = max(np.max(np.abs(s_bar)), np.max(np.abs(E_bar * E0)))
    idx = np.random.default_rng(6).choice(t.n_elements, 61, replace=False)
    params = [("E_layer", k) for k in range(3)] + [("rate_scale", int(e)) for e in idx]

    def loss(s, E):
        return apical_loss(ro.forward(s, E), target, apical)[0]

    return dict(ro=ro, E0=E0, s0=s0, s_bar=s_bar, E_bar=E_bar, gmax=gmax, params=params, loss=loss, L0=L0)


@pytest.mark.parametrize("part", range(4))
def test_config5_adjoint_matches_central_differences(config5_gradient, part):
    """|adjoint - FD| <= 1e-5 max(|FD|, 1e-2 max|gradient|) for each of the 64 parameters (16 per
    part).  The floor is the noise floor of the differences: with h = 1e-6 the loss changes by ~1e-6
    of a gradient entry, and the forward runs' rounding leaves ~1e-8 max|gradient| in the FD value
    (measured: the same FD values at nr_tol 1e-10 and 1e-11)."""
    g = config5_gradient
    h = 1e-6
    worst = 0.0
    for kind, i in g["params"][16 * part:16 * (part + 1)]:
        if kind == "E_layer":
            Ep, Em = g["E0"].copy(), g["E0"].copy()
            Ep[i] *= 1 + h
            Em[i] *= 1 - h
            fd = (g["loss"](g["s0"], Ep) - g["loss"](g["s0"], Em)) / (2 * h)
            adj = g["E_bar"][i] * g["E0"][i]  # d L / d ln E_i
        else:
            sp, sm = g["s0"].copy(), g["s0"].copy()
            sp[i] += h
            sm[i] -= h
            fd = (g["loss"](sp, g["E0"]) - g["loss"](sm, g["E0"])) / (2 * h)
            adj = g["s_bar"][i]
        worst = max(worst, abs(adj - fd) / g["gmax"])
        assert abs(adj - fd) <= 1e-5 * max(abs(fd), 1e-2 * g["gmax"]), (kind, i, adj, fd)
    print(f"config 5 part {part}: max |adjoint - FD| / max|gradient| = {worst:.2e}", flush=True)
