"""Gradients of a multi-step implicit growth simulation by the discrete adjoint.

The parameters are a per-element growth-rate scale ``s`` (rates_e = s_e * base_rates_e) and a
Young's modulus per prism layer ``E_layer`` (lambda, mu from E and the element's nu,
Prism.cpp:491-516).  :class:`GrowthRollout` runs T timesteps of one tissue on the GPU (growth with
the rotation rule, viscosity surfaces, the NR loop) and keeps a checkpoint of (x, Fg, InvFg)
every ``checkpoint_every`` steps.  Its backward pass walks the checkpoint windows in reverse:
it re-runs the window forward from its checkpoint, then takes one reverse step per timestep on the
device (``tf_adjoint_step``: the transposed solve M^T y = xbar and the vector-Jacobian products
of the residual with respect to InvFg, lambda, mu, the previous positions and the growth
increment).  :class:`TissueRolloutFn` exposes the rollout as a ``torch.autograd.Function`` so that
a loss written in torch back-propagates into ``s`` and ``E_layer``.

Converged NR steps are treated as exact roots of the residual, and the growth-rotation angle of
each step (a thresholded decision) is held fixed.  Supported: eta_int = 0, no bound or ablated
DoFs (the library rejects the rest).
"""
from __future__ import annotations

import numpy as np

from .configs import Tissue, lame
from .solver import NewtonRaphsonSolver


class GrowthRollout:
    def __init__(self, tissue: Tissue, n_steps: int, checkpoint_every: int = 5, opts: dict | None = None):
        if tissue.mesh.layer is None or tissue.nu is None:
            raise ValueError("the tissue needs per-element layer indices and Poisson ratios")
        if np.any(tissue.eta_int != 0):
            raise ValueError("the adjoint supports eta_int = 0 only")
        self.t = tissue
        self.T = int(n_steps)
        self.K = max(1, int(checkpoint_every))
        self.dt = float(tissue.dt)
        self.base = np.asarray(tissue.rates, float).copy()
        self.layer = np.asarray(tissue.mesh.layer, np.int64)
        self.n_layers = int(self.layer.max()) + 1
        self.nu = np.asarray(tissue.nu, float)
        # exact linear solves from zero: with inexact-Newton forcing or warm starts the solves
        # depend on the NR history, and a window re-run from its checkpoint would not reproduce the
        # forward states bit for bit
        self.solver = NewtonRaphsonSolver(tissue, dict({"krylov_forcing": 0, "krylov_warm_start": 0}, **(opts or {})))
        self.x0 = tissue.mesh.pos.ravel().copy()
        E = tissue.n_elements
        self.eye = np.tile(np.eye(3), (E, 1, 1))
        self.checkpoints = {}
        self.stats = []

    def layer_moduli(self) -> np.ndarray:
        """Per-layer Young's modulus of the tissue as built (the mean over each layer)."""
        Ey = np.asarray(self.t.E_young, float)
        return np.array([Ey[self.layer == k].mean() for k in range(self.n_layers)])

    # -- parameters and state -------------------------------------------------------------------
    def _set_params(self, s, E_layer):
        s = np.asarray(s, float).ravel()
        E_layer = np.asarray(E_layer, float).ravel()
        self.solver.set_growth_rates(s[:, None] * self.base)
        lam, mu = lame(E_layer[self.layer], self.nu)
        self.solver.set_material(lam, mu, None)
        self._E_layer = E_layer

    def _state(self):
        Fg, Fi = self.solver.get_fg()
        return self.solver.positions, Fg, Fi

    def _restore(self, x, Fg, Fi):
        self.solver.positions = x
        self.solver.set_growth_state(Fg, Fi)

    # -- forward ---------------------------------------------------------------------------------
    def forward(self, s, E_layer) -> np.ndarray:
        """Runs the T steps from the initial state; returns the final positions (3N,)."""
        self._set_params(s, E_layer)
        self._restore(self.x0, self.eye, self.eye)
        self.checkpoints = {0: (self.x0.copy(), self.eye.copy(), self.eye.copy())}
        self.stats = []
        for n in range(self.T):
            st = self.solver.runOneStep(self.dt)
            if st["converged"] != 1:
                raise RuntimeError(f"step {n}: Newton-Raphson did not converge")
            self.stats.append(st)
            if (n + 1) % self.K == 0 and n + 1 < self.T:
                self.checkpoints[n + 1] = self._state()
        return self.solver.positions

    # -- backward --------------------------------------------------------------------------------
    def backward(self, xbar_T) -> tuple[np.ndarray, np.ndarray]:
        """dL/ds and dL/dE_layer from dL/dx_T (after :meth:`forward` with the same parameters)."""
        sv = self.solver
        sv.adjoint_reset()
        xbar = np.asarray(xbar_T, float).ravel().copy()
        starts = sorted(self.checkpoints)
        self.adjoint_krylov = []
        for c0 in reversed(starts):
            c1 = min(c0 + self.K, self.T)
            # re-run the window from its checkpoint, keeping every state and rotation angle
            x, Fg, Fi = self.checkpoints[c0]
            self._restore(x, Fg, Fi)
            states = [(x, Fg, Fi)]
            angles = []
            for n in range(c0, c1):
                sv.runOneStep(self.dt)
                angles.append(sv.growth_rotation())
                states.append(self._state())
            for n in reversed(range(c0, c1)):
                xn, Fgn, _ = states[n - c0]
                xn1, Fgn1, Fin1 = states[n - c0 + 1]
                # the converged state of step n: surfaces at x_n, un = x_n, uk = x_{n+1}, Fg_{n+1}
                sv.positions = xn
                sv.updateNodeViscositySurfaces()
                sv.constructUnMatrix()
                sv.positions = xn1
                sv.set_growth_state(Fgn1, Fin1)
                xbar = sv.adjoint_step(self.dt, xbar, Fgn, angles[n - c0])
                self.adjoint_krylov.append(sv.last_krylov["iterations"])
        rate_bar, lam_bar, mu_bar, _ = sv.adjoint_get()
        s_bar = np.sum(rate_bar * self.base, axis=1)
        nu = self.nu
        e_bar = lam_bar * nu / ((1 + nu) * (1 - 2 * nu)) + mu_bar / (2 * (1 + nu))
        E_bar = np.bincount(self.layer, weights=e_bar, minlength=self.n_layers)
        self.x0_bar = xbar
        return s_bar, E_bar


_FN = None


def _rollout_fn():
    """The torch.autograd.Function, built on first use so that importing this module does not
    load torch."""
    global _FN
    if _FN is None:
        import torch

        class TissueRolloutFn(torch.autograd.Function):
            """x_T = rollout(s, E_layer), differentiable in both arguments."""

            @staticmethod
            def forward(ctx, s, E_layer, rollout: GrowthRollout):
                x = rollout.forward(s.detach().cpu().double().numpy(), E_layer.detach().cpu().double().numpy())
                ctx.rollout = rollout
                ctx.meta = (s.dtype, s.device, E_layer.dtype, E_layer.device)
                return torch.from_numpy(x.copy()).to(dtype=s.dtype, device=s.device)

            @staticmethod
            def backward(ctx, xbar):
                s_bar, E_bar = ctx.rollout.backward(xbar.detach().cpu().double().numpy())
                sd, sdev, ed, edev = ctx.meta
                return (torch.from_numpy(s_bar).to(dtype=sd, device=sdev),
                        torch.from_numpy(E_bar).to(dtype=ed, device=edev), None)

        _FN = TissueRolloutFn
    return _FN


def __getattr__(name):
    if name == "TissueRolloutFn":
        return _rollout_fn()
    raise AttributeError(name)


def apical_loss(x, target, apical_dofs):
    """L = 1/2 sum over apical nodes of |x - x*|^2 and its gradient dL/dx."""
    d = np.zeros_like(x)
    d[apical_dofs] = x[apical_dofs] - target[apical_dofs]
    return 0.5 * float(np.dot(d, d)), d


FD_OPTS = dict(nr_tol=1e-11, krylov_rtol=1e-13)  # forward runs of the difference check: no solver noise


def config5_check(n_fd: int = 64, seed: int = 6, h: float = 1e-6, n_steps: int | None = None, opts: dict | None = None,
                  log=None, fd_opts: dict | None = None) -> dict:
    """The inverse-growth workload (config 5): target run, forward and backward timings (with
    ``opts``), and central differences (h relative) of the loss on ``n_fd`` parameters -- the
    three layer moduli plus ``n_fd - 3`` seeded random rate scales -- against the adjoint
    gradient of a second rollout at ``fd_opts`` (default: Newton to 1e-11 with exact Krylov
    solves, so that the differences carry no solver noise).  Errors are reported per parameter,
    relative to the largest gradient entry, and against the bar |adjoint - FD| <= 1e-5 max(|FD|,
    1e-2 max|gradient|)."""
    import time

    from . import configs
    from .mesh import APICAL

    t, s_target, T, K = configs.config5()
    if n_steps:
        T = n_steps
    rt = GrowthRollout(t, T, checkpoint_every=K, opts=opts)  # the timed rollout
    E0 = rt.layer_moduli()
    target = rt.forward(s_target, E0)
    apical = np.where(np.repeat(t.mesh.node_placement == APICAL, 3))[0]
    s0 = np.ones(t.n_elements)
    t0 = time.perf_counter()
    x = rt.forward(s0, E0)
    t_fwd = time.perf_counter() - t0
    _, dLdx = apical_loss(x, target, apical)
    t0 = time.perf_counter()
    rt.backward(dLdx)
    t_bwd = time.perf_counter() - t0
    # the difference check on a rollout at tight tolerances (its own target run)
    fo = dict(opts or {}, **(FD_OPTS if fd_opts is None else fd_opts))
    ro = GrowthRollout(t, T, checkpoint_every=K, opts=fo)
    target = ro.forward(s_target, E0)
    L0, dLdx = apical_loss(ro.forward(s0, E0), target, apical)
    s_bar, E_bar = ro.backward(dLdx)
    gmax = max(np.max(np.abs(s_bar)), np.max(np.abs(E_bar * E0)))
    idx = np.random.default_rng(seed).choice(t.n_elements, max(0, n_fd - 3), replace=False)
    rows = []
    for k in range(3 if n_fd >= 3 else n_fd):
        Ep, Em = E0.copy(), E0.copy()
        Ep[k] *= 1 + h
        Em[k] *= 1 - h
        fd = (apical_loss(ro.forward(s0, Ep), target, apical)[0] - apical_loss(ro.forward(s0, Em), target, apical)[0]) / (2 * h)
        adj = E_bar[k] * E0[k]   # d L / d ln E_k, on the scale of the rate-scale gradients
        rows.append(("E_layer", k, adj, fd))
        if log:
            log(rows[-1])
    for e in idx:
        sp, sm = s0.copy(), s0.copy()
        sp[e] += h
        sm[e] -= h
        fd = (apical_loss(ro.forward(sp, E0), target, apical)[0] - apical_loss(ro.forward(sm, E0), target, apical)[0]) / (2 * h)
        rows.append(("rate_scale", int(e), s_bar[e], fd))
        if log:
            log(rows[-1])
    err_rel_max = max(abs(a - f) / gmax for _, _, a, f in rows) if rows else 0.0
    per = [abs(a - f) / max(abs(f), 1e-300) for _, _, a, f in rows]
    bar = all(abs(a - f) <= 1e-5 * max(abs(f), 1e-2 * gmax) for _, _, a, f in rows)
    return {"workload": f"config 5: {t.name}, {t.n_elements} prisms, {T} steps, checkpoint every {K}",
            "loss": L0, "forward_s": t_fwd, "backward_s": t_bwd, "backward_over_forward": t_bwd / t_fwd,
            "nr_iterations_per_step": float(np.mean([st["nr_iterations"] for st in rt.stats])),
            "adjoint_krylov_per_step": float(np.mean(rt.adjoint_krylov)),
            "rotated_elements": int(sum(st["rotated_elements"] for st in rt.stats)),
            "fd_parameters": len(rows), "fd_h": h, "fd_options": fo, "fd_within_bar": bool(bar),
            "fd_err_max_rel_to_max_grad": err_rel_max,
            "fd_err_per_param_median": float(np.median(per)) if per else 0.0,
            "fd_err_per_param_max": float(np.max(per)) if per else 0.0,
            "rows": [(k, int(i), float(a), float(f)) for k, i, a, f in rows]}


if __name__ == "__main__":
    import argparse
    import json

    p = argparse.ArgumentParser(description="config-5 gradient check (GPU)")
    p.add_argument("--n-fd", type=int, default=64)
    p.add_argument("--seed", type=int, default=6)
    a = p.parse_args()
    out = config5_check(a.n_fd, a.seed, log=lambda r: print(json.dumps(r), flush=True))
    print(json.dumps(out), flush=True)
