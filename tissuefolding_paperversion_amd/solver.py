"""Host-side facade over the C ABI whose methods mirror the reference's ``NewtonRaphsonSolver``.

The reference drives one NR step from ``Simulation::updateStepNR`` (Simulation.cpp:6449-6590)
through these calls, in this order::

    setMatricesToZeroAtTheBeginningOfIteration / constructUnMatrix / initialteUkMatrix
    loop: calculateDisplacementMatrix, calculateForcesAndJacobianMatrixNR,
          writeForcesTogeAndgvInternal, writeImplicitElementalKToJacobian,
          calculateExternalViscousForcesForNR, addImplicitKViscousExternalToJacobian,
          calculateSumOfInternalForces, addExernalForces, calcutateFixedK, solveForDeltaU,
          checkConvergenceViaDeltaU, updateUkInIteration

:class:`NewtonRaphsonSolver` keeps those names; the element-side calls are merged into
``calculateForcesAndJacobianMatrixNR`` (one device pass builds g and the Jacobian), and the whole
loop is also available as one call (:meth:`updateStepNR`).  All state lives on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .configs import Tissue


class NewtonRaphsonSolver:
    def __init__(self, tissues, opts: dict | None = None):
        """``tissues``: one :class:`Tissue` or a list of them (independent segments of one handle)."""
        if isinstance(tissues, Tissue):
            tissues = [tissues]
        self.tissues = list(tissues)
        self._lib = L.load()
        node_off = np.cumsum([0] + [t.n_nodes for t in self.tissues])
        elem_off = np.cumsum([0] + [t.n_elements for t in self.tissues])
        self.node_seg_ptr = L.i32(node_off)
        self.elem_seg_ptr = L.i32(elem_off)
        cat = lambda f: np.concatenate([f(t) for t in self.tissues])  # noqa: E731
        self._keep = dict(
            pos=L.f64(cat(lambda t: t.mesh.pos.ravel())),
            conn=L.i32(np.concatenate([t.mesh.conn.ravel() + node_off[k] for k, t in enumerate(self.tissues)])),
            xref=L.f64(cat(lambda t: t.mesh.xref.ravel())),
            lam=L.f64(cat(lambda t: t.lam)), mu=L.f64(cat(lambda t: t.mu)), eta=L.f64(cat(lambda t: t.eta_int)),
            eta_ext=L.f64(cat(lambda t: t.eta_ext)), fixed=L.u8(cat(lambda t: t.fixed)),
            exposed=L.u8(cat(lambda t: t.exposed)), rate=L.f64(cat(lambda t: t.rates.ravel())),
            placement=np.ascontiguousarray(cat(lambda t: t.mesh.node_placement), dtype=np.int8),
        )
        k = self._keep
        d = L.MeshDesc()
        d.n_nodes, d.n_elements = int(node_off[-1]), int(elem_off[-1])
        d.node_pos, d.conn, d.ref_pos = L.ptr(k["pos"]), L.ptr(k["conn"]), L.ptr(k["xref"])
        d.lam, d.mu, d.eta_int = L.ptr(k["lam"]), L.ptr(k["mu"]), L.ptr(k["eta"])
        d.eta_ext, d.fixed, d.exposed, d.growth_rate = L.ptr(k["eta_ext"]), L.ptr(k["fixed"]), L.ptr(k["exposed"]), L.ptr(k["rate"])
        d.node_placement = L.ptr(k["placement"])
        d.n_segments = len(self.tissues)
        d.node_seg_ptr = L.ptr(self.node_seg_ptr) if len(self.tissues) > 1 else None
        d.elem_seg_ptr = L.ptr(self.elem_seg_ptr) if len(self.tissues) > 1 else None
        h = C.c_void_p()
        rc = self._lib.tf_create(C.byref(d), C.byref(h))
        if rc != L.TF_OK:
            raise L.TfError(rc, self._lib.tf_last_error(None).decode())
        self.h = h
        self.N, self.E = d.n_nodes, d.n_elements
        nb = C.c_int32()
        self._lib.tf_get_sizes(self.h, None, None, C.byref(nb), None)
        self.nb = nb.value
        if opts:
            self.set_options(**opts)

    # ---- lifetime / options --------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self._lib.tf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, allow=()):
        if rc != L.TF_OK and rc not in allow:
            raise L.TfError(rc, self._lib.tf_last_error(self.h).decode())
        return rc

    def options(self) -> dict:
        o = L.SolverOpts()
        self._check(self._lib.tf_get_options(self.h, C.byref(o)))
        return {k: getattr(o, k) for k, _ in o._fields_}

    def set_options(self, **kw):
        o = L.SolverOpts()
        self._check(self._lib.tf_get_options(self.h, C.byref(o)))
        for k, v in kw.items():
            setattr(o, k, v)
        self._check(self._lib.tf_set_options(self.h, C.byref(o)))

    # ---- state ------------------------------------------------------------------------------
    @property
    def positions(self) -> np.ndarray:
        x = np.zeros(3 * self.N)
        self._check(self._lib.tf_get_positions(self.h, L.ptr(x)))
        return x

    @positions.setter
    def positions(self, x):
        x = L.f64(x).ravel()
        self._check(self._lib.tf_set_positions(self.h, L.ptr(x)))

    def get_fg(self):
        a, b = np.zeros(9 * self.E), np.zeros(9 * self.E)
        self._check(self._lib.tf_get_fg(self.h, L.ptr(a), L.ptr(b)))
        return a.reshape(-1, 3, 3), b.reshape(-1, 3, 3)

    def set_fg(self, Fg):
        Fg = L.f64(Fg).ravel()
        self._check(self._lib.tf_set_fg(self.h, L.ptr(Fg)))

    def set_growth_increments(self, dFg):
        """Per-element growth increments (E, 3, 3) for the following growth updates; None returns
        to the diagonal exp(rate dt) increments."""
        a = None if dFg is None else L.f64(dFg).ravel()
        self._check(self._lib.tf_set_growth_increments(self.h, L.ptr(a)))

    def set_growth_functions(self, functions, initial_z, rel_src, pinned=True, interpolation=1):
        """Growth functions (growth.UniformGrowth, RingGrowth, GridGrowth) evaluated on the device
        at every growth update from the handle's time; an empty list returns to the rates."""
        recs = (L.GrowthFunction * max(1, len(functions)))()
        keep = []
        for k, f in enumerate(functions):
            r = recs[k]
            r.kind, r.t0, r.t1 = f.kind, float(f.t0), float(f.t1)
            if f.kind in (1, 2):
                r.rate[:] = [float(v) for v in f.rate]
                r.angle = float(f.angle)
            if f.kind == 2:
                r.centre[:] = [float(f.centre[0]), float(f.centre[1])]
                r.inner, r.outer = float(f.inner), float(f.outer)
            if f.kind == 3:
                rates = L.f64(f.rates)
                angles = L.f64(f.angles)
                keep += [rates, angles]
                r.nx, r.ny = angles.shape
                r.zmin, r.zmax = float(f.zmin), float(f.zmax)
                r.grid_rates = rates.ctypes.data_as(C.POINTER(C.c_double))
                r.grid_angles = angles.ctypes.data_as(C.POINTER(C.c_double))
        z = L.f64(initial_z).ravel()
        src = L.i32(rel_src).ravel()
        self._check(self._lib.tf_set_growth_functions(self.h, len(functions), recs, L.ptr(z), L.ptr(src),
                                                      int(bool(pinned)), int(interpolation)))

    def pin_growth_positions(self):
        self._check(self._lib.tf_pin_growth_positions(self.h))

    def set_time(self, t):
        self._check(self._lib.tf_set_time(self.h, float(t)))

    def get_time(self):
        t = C.c_double()
        self._check(self._lib.tf_get_time(self.h, C.byref(t)))
        return t.value

    def growth_increments(self):
        """(E, 3, 3) growth increments of the last growth update (device functions or host-given)."""
        out = np.zeros((self.E, 3, 3))
        self._check(self._lib.tf_get_growth_increments(self.h, L.ptr(out)))
        return out

    def set_growth_rates(self, rates):
        r = L.f64(rates).ravel()
        self._check(self._lib.tf_set_growth_rates(self.h, L.ptr(r)))

    def set_material(self, lam=None, mu=None, eta=None):
        a = None if lam is None else L.f64(lam)
        b = None if mu is None else L.f64(mu)
        c = None if eta is None else L.f64(eta)
        self._check(self._lib.tf_set_material(self.h, L.ptr(a), L.ptr(b), L.ptr(c)))

    # ---- DoF constraints ------------------------------------------------------------------------
    def set_bound_dofs(self, pairs):
        """Bind slave DoFs to master DoFs: ``pairs`` is a sequence of (slave DoF, master DoF)
        (``NewtonRaphsonSolver::slaveMasterList``); an empty sequence removes all binding."""
        p = L.i32(np.asarray(pairs, dtype=np.int32).reshape(-1, 2)) if len(pairs) else None
        n = 0 if p is None else p.shape[0]
        self._check(self._lib.tf_set_bound_dofs(self.h, n, L.ptr(p)))
        nb = C.c_int32()
        self._lib.tf_get_sizes(self.h, None, None, C.byref(nb), None)
        self.nb = nb.value

    def set_ablated_nodes(self, nodes):
        a = L.i32(np.asarray(nodes, dtype=np.int32)) if len(nodes) else None
        self._check(self._lib.tf_set_ablated_nodes(self.h, 0 if a is None else len(a), L.ptr(a)))

    # ---- implicit packing -----------------------------------------------------------------------
    def node_masses(self) -> np.ndarray:
        """Node::mass from the current Fg (Simulation::updateNodeMasses)."""
        m = np.zeros(self.N)
        self._check(self._lib.tf_get_node_masses(self.h, L.ptr(m)))
        return m

    def detectPacingNodes(self) -> int:
        """Node masses and packing pairs on the current positions (Simulation::detectPacingNodes);
        returns the pair count."""
        n = C.c_int32()
        self._check(self._lib.tf_detect_packing(self.h, C.byref(n)))
        self._refresh_nb()
        return n.value

    def set_packing_pairs(self, pairs, weights, threshold):
        """Replace the pair list: ``pairs`` (P, 2) node ids, ``weights`` (P, 3), ``threshold``
        one packing threshold per tissue."""
        p = L.i32(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        w = L.f64(np.asarray(weights, float).reshape(-1, 3))
        t = L.f64(np.atleast_1d(np.asarray(threshold, float)))
        self._check(self._lib.tf_set_packing_pairs(self.h, p.shape[0], L.ptr(p) if len(p) else None,
                                                   L.ptr(w) if len(w) else None, L.ptr(t)))
        self._refresh_nb()

    def packing_pairs(self):
        """(pairs (P, 2), weights (P, 3), threshold per tissue (S,), detection grids (S, 10, 5))."""
        n = C.c_int32()
        self._check(self._lib.tf_get_packing_pairs(self.h, C.byref(n), None, None, None, None))
        P, S = n.value, len(self.tissues)
        p, w = np.zeros(2 * P, np.int32), np.zeros(3 * P)
        t, g = np.zeros(S), np.zeros(50 * S)
        self._check(self._lib.tf_get_packing_pairs(self.h, C.byref(n), L.ptr(p), L.ptr(w), L.ptr(t), L.ptr(g)))
        return p.reshape(-1, 2), w.reshape(-1, 3), t, g.reshape(S, 10, 5)

    def _refresh_nb(self):
        nb = C.c_int32()
        self._lib.tf_get_sizes(self.h, None, None, C.byref(nb), None)
        self.nb = nb.value

    # ---- per step -----------------------------------------------------------------------------
    def grow(self, dt) -> int:
        n = C.c_int32()
        self._check(self._lib.tf_grow(self.h, dt, C.byref(n)))
        return n.value

    def updateNodeViscositySurfaces(self):
        self._check(self._lib.tf_update_viscosity_surfaces(self.h))

    def viscosity_surfaces(self):
        s = np.zeros(self.N)
        self._check(self._lib.tf_get_viscosity_surfaces(self.h, L.ptr(s)))
        return s

    # ---- NewtonRaphsonSolver call sequence -----------------------------------------------------
    def constructUnMatrix(self):
        """constructUnMatrix + initialteUkMatrix: un <- x, uk <- un."""
        self._check(self._lib.tf_begin_step(self.h))

    initialteUkMatrix = constructUnMatrix

    def calculateForcesAndJacobianMatrixNR(self, dt):
        """Element forces/tangents, their scatter, external viscosity, fixed DoFs (device)."""
        return self._check(self._lib.tf_assemble(self.h, dt), allow=(L.TF_NEG_JACOBIAN,))

    @property
    def gSum(self) -> np.ndarray:
        g = np.zeros(3 * self.N)
        self._check(self._lib.tf_get_residual(self.h, L.ptr(g)))
        return g

    def jacobian_bsr(self):
        rp = np.zeros(self.N + 1, np.int32)
        col = np.zeros(self.nb, np.int32)
        val = np.zeros(9 * self.nb)
        self._check(self._lib.tf_get_jacobian_bsr(self.h, L.ptr(rp), L.ptr(col), L.ptr(val)))
        return rp, col, val.reshape(-1, 3, 3)

    def jacobian_csr(self):
        """Natural-orientation Jacobian M as a scipy CSR matrix (3N x 3N)."""
        import scipy.sparse as sp
        rp, col, val = self.jacobian_bsr()
        return sp.bsr_matrix((val, col, rp), shape=(3 * self.N, 3 * self.N)).tocsr()

    def element_dump(self, e, dt):
        fe, fv, Ke, Kv = np.zeros(18), np.zeros(18), np.zeros(324), np.zeros(324)
        self._check(self._lib.tf_element_dump(self.h, int(e), dt, L.ptr(fe), L.ptr(fv), L.ptr(Ke), L.ptr(Kv)))
        return dict(fe=fe, fv=fv, Ke=Ke.reshape(18, 18), Kv=Kv.reshape(18, 18))

    def solveForDeltaU(self):
        du = np.zeros(3 * self.N)
        it = C.c_int32()
        rr = C.c_double()
        rc = self._check(self._lib.tf_solve(self.h, L.ptr(du), C.byref(it), C.byref(rr)), allow=(L.TF_KRYLOV_STALL,))
        self.last_krylov = dict(iterations=it.value, rel_residual=rr.value, status=rc)
        return du

    def updateUkInIteration(self):
        """checkConvergenceViaDeltaU + updateUkInIteration: returns (||dU||, converged)."""
        d = C.c_double()
        c = C.c_int32()
        self._check(self._lib.tf_update_uk(self.h, C.byref(d), C.byref(c)))
        return d.value, bool(c.value)

    def updateStepNR(self, dt):
        st = L.StepStats()
        rc = self._check(self._lib.tf_nr_step(self.h, dt, C.byref(st)), allow=(L.TF_NOT_CONVERGED,))
        out = st.as_dict()
        out["status"] = rc
        return out

    def runOneStep(self, dt):
        """Growth, viscosity surfaces, node masses and packing detection (option implicit_packing)
        and the NR loop of one timestep."""
        st = L.StepStats()
        rc = self._check(self._lib.tf_timestep(self.h, dt, C.byref(st)), allow=(L.TF_NOT_CONVERGED,))
        self._refresh_nb()
        out = st.as_dict()
        out["status"] = rc
        return out

    def segment_stats(self):
        S = len(self.tissues)
        it, dn, cv = np.zeros(S, np.int32), np.zeros(S), np.zeros(S, np.int32)
        self._check(self._lib.tf_get_segment_stats(self.h, L.ptr(it), L.ptr(dn), L.ptr(cv)))
        return it, dn, cv

    # ---- linear algebra and the discrete adjoint -----------------------------------------------------
    def set_growth_state(self, Fg, InvFg):
        a, b = L.f64(Fg).ravel(), L.f64(InvFg).ravel()
        self._check(self._lib.tf_set_growth_state(self.h, L.ptr(a), L.ptr(b)))

    def growth_rotation(self):
        tz = np.zeros(self.E)
        self._check(self._lib.tf_get_growth_rotation(self.h, L.ptr(tz)))
        return tz

    def apply_jacobian(self, x, transpose=False):
        x = L.f64(x).ravel()
        y = np.zeros(3 * self.N)
        self._check(self._lib.tf_apply_jacobian(self.h, int(bool(transpose)), L.ptr(x), L.ptr(y)))
        return y

    def solve_system(self, b, transpose=False):
        b = L.f64(b).ravel()
        x = np.zeros(3 * self.N)
        it, rr = C.c_int32(), C.c_double()
        self._check(self._lib.tf_solve_system(self.h, int(bool(transpose)), L.ptr(b), L.ptr(x), C.byref(it), C.byref(rr)))
        self.last_krylov = dict(iterations=it.value, rel_residual=rr.value, status=L.TF_OK)
        return x

    def adjoint_elements(self, y):
        y = L.f64(y).ravel()
        fi, lb, mb = np.zeros(9 * self.E), np.zeros(self.E), np.zeros(self.E)
        self._check(self._lib.tf_adjoint_elements(self.h, L.ptr(y), L.ptr(fi), L.ptr(lb), L.ptr(mb)))
        return fi.reshape(-1, 3, 3), lb, mb

    def adjoint_reset(self):
        self._check(self._lib.tf_adjoint_reset(self.h))

    def adjoint_step(self, dt, xbar_next, fg_prev, theta_z):
        xb = L.f64(xbar_next).ravel()
        fp = L.f64(fg_prev).ravel()
        tz = L.f64(theta_z).ravel()
        out = np.zeros(3 * self.N)
        it = C.c_int32()
        self._check(self._lib.tf_adjoint_step(self.h, dt, L.ptr(xb), L.ptr(fp), L.ptr(tz), L.ptr(out), C.byref(it)))
        self.last_krylov = dict(iterations=it.value, status=L.TF_OK)
        return out

    def adjoint_get(self):
        rb, lb, mb, fb = np.zeros(3 * self.E), np.zeros(self.E), np.zeros(self.E), np.zeros(9 * self.E)
        self._check(self._lib.tf_adjoint_get(self.h, L.ptr(rb), L.ptr(lb), L.ptr(mb), L.ptr(fb)))
        return rb.reshape(-1, 3), lb, mb, fb.reshape(-1, 3, 3)

    def coarse_size(self):
        """Coarse unknowns of the two-level preconditioner (0 before its first build)."""
        n = C.c_int32(0)
        rc = self._lib.tf_get_coarse_operator(self.h, C.byref(n), None, None)
        return n.value if rc == L.TF_OK else 0

    def coarse_operator(self):
        """(Ac, Ac^-1): the two-level coarse operator of the current Jacobian and the FP64 inverse
        the last coarse rebuild computed (diagnostics)."""
        n = C.c_int32()
        self._check(self._lib.tf_get_coarse_operator(self.h, C.byref(n), None, None))
        Ac, Ai = np.zeros((n.value, n.value)), np.zeros((n.value, n.value))
        self._check(self._lib.tf_get_coarse_operator(self.h, C.byref(n), L.ptr(Ac), L.ptr(Ai)))
        return Ac, Ai

    # ---- tuning -----------------------------------------------------------------------------------
    def set_tiling(self, wg, tile_cap):
        self._check(self._lib.tf_set_tiling(self.h, int(wg), int(tile_cap)))

    def tiling(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self._lib.tf_get_tiling(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def autotune(self, dt=60.0, n_iters=64):
        """Time the single-kernel CG under every supported tiling and keep the fastest."""
        n = C.c_int32()
        wg, cap, ms = np.zeros(32, np.int32), np.zeros(32, np.int32), np.zeros(32)
        self._check(self._lib.tf_autotune(self.h, dt, int(n_iters), C.byref(n), L.ptr(wg), L.ptr(cap), L.ptr(ms)))
        k = n.value
        return [dict(wg=int(wg[i]), tile_cap=int(cap[i]), ms_per_iter=float(ms[i])) for i in range(k)]

    # ---- measurement ------------------------------------------------------------------------------
    def kernel_timing(self, name, enable=-1):
        avg, n = C.c_double(), C.c_int64()
        self._check(self._lib.tf_kernel_timing(self.h, name.encode(), enable, C.byref(avg), C.byref(n)))
        return avg.value, n.value

    def kernel_bytes(self, name):
        b = C.c_double()
        self._check(self._lib.tf_kernel_bytes(self.h, name.encode(), C.byref(b)))
        return b.value

    def kernel_probe(self, n_launches=64, name="spmv"):
        """Average device ms of one launch of a Krylov kernel ("spmv": the SpMV launch of the
        active solver, "coarse": the two-level coarse-correction launch), n_launches back to back
        on the current Jacobian (HIP events around one graph replay)."""
        ms = C.c_double()
        self._check(self._lib.tf_kernel_probe_named(self.h, name.encode(), n_launches, C.byref(ms)))
        return ms.value

    def krylov_timing(self, enable=-1):
        """(replay ms, iterations launched, iterations taken) of the Krylov chunk replays since the
        last reset (HIP events on the handle's stream); enable 1/0 resets and switches on/off."""
        ms, launched, its = C.c_double(), C.c_int64(), C.c_int64()
        self._check(self._lib.tf_krylov_timing(self.h, enable, C.byref(ms), C.byref(launched), C.byref(its)))
        return ms.value, launched.value, its.value

    def copy_bandwidth(self, nbytes=1 << 30, reps=5):
        """Best read + write GB/s of a device-to-device copy of nbytes (the achievable HBM rate)."""
        bw = C.c_double()
        self._check(self._lib.tf_copy_bandwidth(self.h, int(nbytes), reps, C.byref(bw)))
        return bw.value
