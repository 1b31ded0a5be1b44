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
        self.solver = NewtonRaphsonSolver(tissue, opts)
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


try:
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

except ImportError:  # torch is optional for the numpy interface
    TissueRolloutFn = None


def apical_loss(x, target, apical_dofs):
    """L = 1/2 sum over apical nodes of |x - x*|^2 and its gradient dL/dx."""
    d = np.zeros_like(x)
    d[apical_dofs] = x[apical_dofs] - target[apical_dofs]
    return 0.5 * float(np.dot(d, d)), d
