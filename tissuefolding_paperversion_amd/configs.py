"""Physical set-up of a tissue for the Newton-Raphson path, and the named workloads.

The set-up rules follow the reference:

* Young's modulus by element placement, Poisson ratio clamped to [0, 0.5], Lame parameters
  (``Prism::setElasticProperties``, Prism.cpp:491-516);
* internal viscosity by placement (``ShapeBase::setViscosity``, ShapeBase.cpp:893-902);
* external (nodal) viscosity by node placement, circumference nodes of the mid-layer take the
  minimum of the apical and basal values (``Node::setExternalViscosity``, Node.cpp:68-116);
* growth rates are given per hour and converted to per second
  (``ModelInputObject.cpp:412-420``); the per-step increment is ``exp(rate * dt)`` along x, y, z
  (``ShapeBase::calculateFgFromRates``, ShapeBase.cpp:396-421).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .mesh import APICAL, BASAL, MIDLINE, LATERAL, TissueMesh, disc_mesh, ellipse_mesh


def lame(E, nu):
    nu = np.clip(np.asarray(nu, float), 0.0, 0.5)
    E = np.asarray(E, float)
    lam = E * nu / (1 + nu) / (1 - 2.0 * nu)
    mu = E / 2.0 / (1 + nu)
    return lam, mu


@dataclasses.dataclass
class Tissue:
    """Everything the NR path needs about one tissue (arrays in the reference's DoF order)."""
    mesh: TissueMesh
    lam: np.ndarray        # (E,)
    mu: np.ndarray         # (E,)
    eta_int: np.ndarray    # (E,)
    eta_ext: np.ndarray    # (3N,)
    fixed: np.ndarray      # (3N,) uint8
    rates: np.ndarray      # (E, 3) growth rate per second
    exposed: np.ndarray    # (E,) uint8 exposed-face bits
    dt: float = 60.0
    n_steps: int = 1
    name: str = ""
    E_young: np.ndarray | None = None   # (E,) Young's modulus (for per-layer gradients)
    nu: np.ndarray | None = None

    @property
    def n_nodes(self):
        return self.mesh.n_nodes

    @property
    def n_elements(self):
        return self.mesh.n_elements


def assign_material(mesh: TissueMesh, E_apical, E_basal, E_mid, nu):
    pl, _ = mesh.element_placement()
    E = np.full(mesh.n_elements, float(E_mid))
    E[pl == BASAL] = E_basal
    E[pl == APICAL] = E_apical
    lam, mu = lame(E, np.full(mesh.n_elements, nu))
    return E, lam, mu


def assign_external_viscosity(mesh: TissueMesh, apical: float, basal: float, extend_to_inner: bool = False):
    """Node::setExternalViscosity (Node.cpp:68-116) for every node, same value on x, y, z."""
    v = np.zeros(mesh.n_nodes)
    p = mesh.node_placement
    v[p == BASAL] = basal
    v[p == APICAL] = apical
    mid = (p != BASAL) & (p != APICAL)
    v[mid & mesh.at_circumference] = min(apical, basal)
    if extend_to_inner:
        inner = mid & ~mesh.at_circumference & ((p == MIDLINE) | (p == LATERAL))
        v[inner] = min(apical, basal)
    return np.repeat(v, 3)


def per_hour(rate):
    return np.asarray(rate, float) * (1.0 / 3600.0)


def _nearest(pos, target):
    return int(np.argmin(np.linalg.norm(pos - np.asarray(target), axis=1)))


def disc_tissue(radius=50.0, h=5.85, z_levels=(0.0, 5.0, 10.0), seed=1, E=1000.0, nu=0.3,
                growth_per_hour=(0.01, 0.01, 0.0), eta_ext=0.0, eta_int=0.0, dt=60.0, n_steps=60,
                name="disc") -> Tissue:
    """Plain FP64 disc: basal z fixed, centre x/y pinned, rim node at (+r, 0) pinned in y.

    The extra rim y-pin removes the free in-plane rotation of an un-damped disc; with it the
    stress-free scaling field x(t) = exp(r t) X (in plane) satisfies every boundary condition,
    which gives this workload an exact answer.
    """
    mesh = disc_mesh(radius, h, z_levels, seed)
    Ey, lam, mu = assign_material(mesh, E, E, E, nu)
    fixed = np.zeros(3 * mesh.n_nodes, np.uint8)
    basal = np.where(mesh.node_placement == BASAL)[0]
    fixed[3 * basal + 2] = 1
    for z in z_levels[:1]:
        c = _nearest(mesh.pos, (0.0, 0.0, z))
        fixed[3 * c + 0] = 1
        fixed[3 * c + 1] = 1
        r = _nearest(mesh.pos, (radius, 0.0, z))
        fixed[3 * r + 1] = 1
    rates = np.tile(per_hour(growth_per_hour), (mesh.n_elements, 1))
    return Tissue(mesh, lam, mu, np.full(mesh.n_elements, float(eta_int)),
                  assign_external_viscosity(mesh, eta_ext, eta_ext), fixed, rates, mesh.exposed_faces(),
                  dt, n_steps, name, Ey, np.full(mesh.n_elements, nu))


def gaussian_bump_rates(mesh: TissueMesh, a: float, b: float, n_bumps: int = 8, seed: int = 3):
    """Per-element growth rates (per second) from a sum of seeded anisotropic Gaussian bumps,
    evaluated at each prism's reference centroid.  Each bump has its own centre inside the
    outline, widths in x and y, and peak rates in x and y drawn from 0.5-3 %/hr; z rate is 0."""
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(0, 1, n_bumps)) * 0.8
    th = rng.uniform(0, 2 * math.pi, n_bumps)
    cx, cy = a * r * np.cos(th), b * r * np.sin(th)
    sx = rng.uniform(0.15, 0.35, n_bumps) * a
    sy = rng.uniform(0.15, 0.35, n_bumps) * b
    px = rng.uniform(0.5, 3.0, n_bumps) * 0.01
    py = rng.uniform(0.5, 3.0, n_bumps) * 0.01
    cen = mesh.xref.reshape(-1, 6, 3).mean(axis=1)
    rates = np.zeros((mesh.n_elements, 3))
    for k in range(n_bumps):
        g = np.exp(-0.5 * ((cen[:, 0] - cx[k]) / sx[k]) ** 2 - 0.5 * ((cen[:, 1] - cy[k]) / sy[k]) ** 2)
        rates[:, 0] += px[k] * g
        rates[:, 1] += py[k] * g
    return per_hour(rates)


def wing_disc_tissue(a=200.0, b=120.0, h=3.66, z_levels=(0.0, 2.0, 7.0, 12.0), seed=2, growth_seed=3,
                     E_apical=1000.0, E_mid=1000.0, E_basal=5000.0, nu=0.3, eta_ext=100.0, dt=60.0,
                     n_steps=120, name="wing-disc ellipse") -> Tissue:
    """Ellipse of three prism layers (apical, columnar, basal ECM) with a Gaussian-bump growth
    map.  External viscosity on apical and basal surfaces (and mid-layer rim nodes) removes the
    rigid-body modes, so no DoF is fixed."""
    mesh = ellipse_mesh(a, b, h, z_levels, seed)
    Ey, lam, mu = assign_material(mesh, E_apical, E_basal, E_mid, nu)
    fixed = np.zeros(3 * mesh.n_nodes, np.uint8)
    rates = gaussian_bump_rates(mesh, a, b, 8, growth_seed)
    return Tissue(mesh, lam, mu, np.zeros(mesh.n_elements), assign_external_viscosity(mesh, eta_ext, eta_ext),
                  fixed, rates, mesh.exposed_faces(), dt, n_steps, name, Ey, np.full(mesh.n_elements, nu))


def config1() -> Tissue:
    """Plain FP64 baseline disc: r = 50 um, ~500 triangles x 2 layers of 5 um, E = 1000 Pa,
    nu = 0.3, 1 %/hr in-plane growth, no viscosity, dt = 60 s, 60 steps."""
    return disc_tissue(name="disc r=50um, 2 layers (~1000 prisms)")


def config2(h: float = 3.66) -> Tissue:
    """Wing-disc-like ellipse 200 x 120 um, ~13k triangles x 3 layers (~40k prisms)."""
    return wing_disc_tissue(h=h, name="wing-disc ellipse 200x120um, 3 layers (~40k prisms)")


def batch_disc(n_prisms_target: float, n_layers: int, E: float, nu: float, rate_per_hour: float, seed: int,
               h: float = 3.0, layer_h: float = 4.0, name: str = "disc") -> Tissue:
    """One disc of a batched sweep: radius chosen so that triangles x layers is close to
    ``n_prisms_target`` (hex lattice of spacing h: a triangle covers sqrt(3)/4 h^2), boundary
    conditions of :func:`disc_tissue`, and a radially graded in-plane growth rate
    r(rho) = rate * (1.5 - (rho/R)^2), whose mean over the disc is ``rate`` and whose gradient
    stresses the tissue (so that every step needs several Newton iterations)."""
    n_tri = n_prisms_target / n_layers
    radius = math.sqrt(n_tri * math.sqrt(3.0) / 4.0 * h * h / math.pi)
    z = tuple(layer_h * k for k in range(n_layers + 1))
    t = disc_tissue(radius=radius, h=h, z_levels=z, seed=seed, E=E, nu=nu, growth_per_hour=(0.0, 0.0, 0.0),
                    name=name)
    cen = t.mesh.xref.reshape(-1, 6, 3).mean(axis=1)
    rho2 = (cen[:, 0] ** 2 + cen[:, 1] ** 2) / radius ** 2
    r = rate_per_hour * 0.01 * (1.5 - rho2)
    t.rates = per_hour(np.stack([r, r, np.zeros_like(r)], axis=1))
    return t


def facing_discs_tissue(radius: float = 20.0, h: float = 5.85, z_levels=(0.0, 5.0, 10.0), gap: float = 1.5,
                        seed: int = 1, E: float = 1000.0, nu: float = 0.3, growth_per_hour=(0.01, 0.01, 0.0),
                        eta_ext: float = 0.0, dt: float = 60.0, n_steps: int = 5,
                        name: str = "facing discs") -> Tissue:
    """Packing workload: two discs in one tissue, the second the mirror image of the first in
    z, so that their apical surfaces face each other across ``gap`` (below the packing
    threshold, 0.4 x the mean side length).  Every apical node of the lower disc then packs
    against the node above it.  The mirror disc lists its prism nodes as (0, 2, 1, 3, 5, 4), so
    that its basal cap (on top) keeps a positive orientation, and its basal nodes are fixed in z
    like the lower disc's; each disc has the in-plane pins of :func:`disc_tissue`."""
    lower = disc_tissue(radius=radius, h=h, z_levels=z_levels, seed=seed, E=E, nu=nu,
                        growth_per_hour=growth_per_hour, eta_ext=eta_ext, dt=dt, n_steps=n_steps)
    m = lower.mesh
    n = m.n_nodes
    top = float(z_levels[-1])
    pos_b = m.pos.copy()
    pos_b[:, 2] = 2.0 * top + gap - pos_b[:, 2]
    conn_b = m.conn[:, [0, 2, 1, 3, 5, 4]] + n
    pos = np.concatenate([m.pos, pos_b])
    conn = np.concatenate([m.conn, conn_b]).astype(np.int32)
    mesh = TissueMesh(pos, conn, pos[conn].reshape(-1, 18).copy(), np.concatenate([m.node_placement] * 2),
                      np.concatenate([m.node_type] * 2), np.concatenate([m.at_circumference] * 2),
                      None if m.layer is None else np.concatenate([m.layer] * 2))
    _, lam, mu = assign_material(mesh, E, E, E, nu)
    return Tissue(mesh, lam, mu, np.zeros(mesh.n_elements), np.concatenate([lower.eta_ext] * 2),
                  np.concatenate([lower.fixed] * 2), np.concatenate([lower.rates] * 2), mesh.exposed_faces(),
                  dt, n_steps, name, np.full(mesh.n_elements, float(E)), np.full(mesh.n_elements, nu))


def config4(n_discs: int = 256, seed: int = 4, min_prisms: float = 500.0, max_prisms: float = 8000.0) -> list[Tissue]:
    """Batched sweep: ``n_discs`` independent discs with prism counts log-uniform in
    [min_prisms, max_prisms], 2-3 layers, E ~ U(500, 5000) Pa, nu ~ U(0.25, 0.45) and mean growth
    rate ~ U(0.5, 3) %/hr (all seeded), dt = 60 s, 60 steps."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_discs):
        target = math.exp(rng.uniform(math.log(min_prisms), math.log(max_prisms)))
        layers = int(rng.integers(2, 4))
        E = rng.uniform(500.0, 5000.0)
        nu = rng.uniform(0.25, 0.45)
        rate = rng.uniform(0.5, 3.0)
        out.append(batch_disc(target, layers, E, nu, rate, seed=1000 * seed + k,
                              name=f"disc {k}: {layers} layers, E={E:.0f} Pa, nu={nu:.3f}, {rate:.2f} %/hr"))
    return out


def config5(seed_target: int = 5, n_steps: int = 30, checkpoint_every: int = 5, h: float = 3.66):
    """Inverse-growth workload: the config-2 ellipse over ``n_steps`` steps; the target is a
    forward run with every element's growth rate scaled by (1 + 0.05 N(0,1)) (seeded); the loss
    is 1/2 sum over apical nodes of |x - x*|^2; parameters are the per-element rate scales
    (nominal 1) and the three per-layer Young's moduli.  Returns (tissue, target_scale,
    n_steps, checkpoint_every)."""
    t = config2(h=h)
    t.n_steps = n_steps
    rng = np.random.default_rng(seed_target)
    scale = 1.0 + 0.05 * rng.standard_normal(t.n_elements)
    return t, scale, n_steps, checkpoint_every
