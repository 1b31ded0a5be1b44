"""Growth functions evaluated on the host, as the reference's ``Simulation::calculateGrowth`` does
(Simulation.cpp:9651-9670): every step starts each element's growth increment at the identity and
left-multiplies the increment of every growth function active at the step's start time
(``ShapeBase::updateGrowthIncrement``, ShapeBase.cpp:567-586).  The product goes to the device
with ``tf_set_growth_increments``; the device applies the growth-rotation rule and Fg <- dFg Fg.

Function types (``GrowthFunctionTypes.h``), columnar layer only:

* 1, uniform (``calculateGrowthUniform``, Simulation.cpp:9757-9797): R diag(exp(r dt)) R^T with R
  the rotation by the function's xy angle (``calculateFgFromRates``, ShapeBase.cpp:396-421);
* 2, ring (``calculateGrowthRing``, Simulation.cpp:9801-9898): the same with the rate scaled by
  1 - (d - r_in) / (r_out - r_in) for elements whose centre lies between the radii;
* 3, grid (``calculateGrowthGridBased``, Simulation.cpp:10065-10117): rates and xy angles on an
  nx x ny grid over the tissue's bounding box, taken at the element's relative position (initial
  or current) from the cell's four corners, nearest corner or bilinear
  (``calculateFgFromGridCorners``, ShapeBase.cpp:432-565), with the corners' angles first brought
  next to each other (``preCalculateAnglesForCompatibleAveraging``, GrowthFunctionTypes.h:434-587)
  and the z-range test on the element's initial relative height (``isGrowthRateApplicable``,
  ShapeBase.cpp:277-300).

Values the reference reads into ``float`` variables are rounded to float here as well.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import APICAL, BASAL, TissueMesh

F32 = np.float32


def _rot(angle: float) -> np.ndarray:
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _rate_increment(rates: np.ndarray, dt: float, R: np.ndarray) -> np.ndarray:
    """R diag(exp(r dt)) R^T for rates (E, 3)."""
    d = np.exp(rates * dt)
    inc = np.zeros((len(rates), 3, 3))
    inc[:, 0, 0], inc[:, 1, 1], inc[:, 2, 2] = d[:, 0], d[:, 1], d[:, 2]
    return np.einsum("ij,ejk,lk->eil", R, inc, R)


@dataclass
class UniformGrowth:
    t0: float
    t1: float
    rate: tuple                      # per second, float-rounded as the reader stores it
    angle: float = 0.0               # radians
    kind: int = 1


@dataclass
class RingGrowth:
    t0: float
    t1: float
    centre: tuple                    # (x, y)
    inner: float
    outer: float
    rate: tuple                      # maximum rate per second at the outer radius
    angle: float = 0.0
    kind: int = 2


@dataclass
class GridGrowth:
    t0: float
    t1: float
    rates: np.ndarray                # (nx, ny, 3) per second
    angles: np.ndarray               # (nx, ny) radians
    zmin: float = 0.0
    zmax: float = 1.0
    kind: int = 3
    corners: dict = field(default_factory=dict, repr=False)

    def __post_init__(self):
        self.corners = compatible_corners(self.rates, self.angles)


def read_grid_file(path: str):
    """Growth-rate grid file (ModelInputObject::readGrowthType3, ModelInputObject.cpp:634-676):
    nx ny, then rates (per hour, 3 per point) and angles (degrees) with rows from the last y
    index down to 0 and x ascending within a row.  Returns (rates per second, angles in radians)."""
    tok = open(path).read().split()
    nx, ny = int(tok[0]), int(tok[1])
    vals = np.array([float(v) for v in tok[2:]])
    need = nx * ny * 4
    if vals.size < need:
        raise ValueError(f"{path}: {vals.size} values, {need} expected")
    r = vals[:3 * nx * ny].reshape(ny, nx, 3)[::-1].transpose(1, 0, 2)      # [i][j][k], j from the bottom
    a = vals[3 * nx * ny:need].reshape(ny, nx)[::-1].T
    rates = r.astype(F32).astype(np.float64) * (1.0 / 3600.0)
    angles = a * math.pi / 180.0
    return rates, angles


def compatible_corners(rates: np.ndarray, angles: np.ndarray):
    """GridBasedGrowthFunction::preCalculateAnglesForCompatibleAveraging: for every grid cell, the
    four corners' growths and angles, angles of corners whose xy aspect ratio is below 1.1
    eliminated (0), the rest brought into the [0, pi] interval and within 45 degrees of the first
    kept corner, rotating by 90 degrees with x/y growths swapped where needed."""
    nx, ny = angles.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        AR = rates[:, :, 0] / rates[:, :, 1]
    AR = np.where(AR < 0, -AR, AR)
    AR = np.where(AR < 1.0, 1.0 / AR, AR)
    over = ~(AR < 1.1)  # NaN stays over the threshold, as in the reference
    out = {}
    q = math.pi
    for i in range(nx - 1):
        for j in range(ny - 1):
            idx = [(i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1)]
            g = [rates[a, b].copy() for a, b in idx]
            elim = [not over[a, b] for a, b in idx]
            ang = [0.0 if elim[k] else float(angles[a, b]) for k, (a, b) in enumerate(idx)]
            for k in range(4):
                while ang[k] < 0:
                    ang[k] += q
                while ang[k] > q:
                    ang[k] -= q
            selected, sel, tet = False, 3, 0.0
            for k in range(4):
                if not elim[k] and not selected:
                    tet = ang[k]
                    selected = True
                if selected:
                    sel = k
                    break
            flip = [False] * 4
            for k in range(sel, 4):
                d = ang[k] - tet
                if d > 0:
                    if q / 4 <= d < 3 * q / 4:
                        ang[k] -= q / 2; flip[k] = True
                    elif 3 * q / 4 <= d < 5 * q / 4:
                        ang[k] -= q
                    elif 5 * q / 4 <= d < 7 * q / 4:
                        ang[k] -= 6 * q / 4; flip[k] = True
                    elif d >= 7 * q / 4:
                        ang[k] -= 2 * q
                elif d < 0:
                    if -3 * q / 4 < d <= -q / 4:
                        ang[k] += q / 2; flip[k] = True
                    elif -5 * q / 4 < d <= -3 * q / 4:
                        ang[k] += q
                    elif -7 * q / 4 < d <= -5 * q / 4:
                        ang[k] += 6 * q / 4; flip[k] = True
                    elif d <= -7 * q / 4:
                        ang[k] += 2 * q
            for k in range(4):
                if flip[k]:
                    g[k][0], g[k][1] = g[k][1], g[k][0]
            out[(i, j)] = (np.array(g), np.array(ang), np.array(elim))
    return out


# ---- element geometry the growth functions need --------------------------------------------------
def reference_bounding_box(pos: np.ndarray):
    """Simulation::calculateBoundingBox (Simulation.cpp:9148-9189): "if (x < min) ... else if
    (x > max)" from (1e5, -1e5), so a node that lowers the running minimum is never compared with
    the maximum.  Returns (low (3,), high (3,))."""
    lo = np.minimum(1e5, pos.min(axis=0))
    hi = np.full(3, -1e5)
    for d in range(3):
        x = pos[:, d]
        run = np.minimum.accumulate(np.concatenate([[1e5], x[:-1]]))
        cand = ~(x < run)
        if cand.any():
            hi[d] = max(-1e5, float(x[cand].max()))
    return lo, hi


def element_columns(mesh: TissueMesh):
    """Prism::constructElementStackList (Prism.cpp:2542-2583): for every basal element, the
    elements above it (each one's basal nodes are the previous one's apical nodes), basal first.
    Returns {apical element: column list}."""
    pl, spans = mesh.element_placement()
    conn = mesh.conn
    by_basal = {}
    for e in range(mesh.n_elements):
        by_basal.setdefault(frozenset(conn[e, :3].tolist()), e)
    cols = {}
    for e in np.where((pl == BASAL) & ~spans)[0]:
        col = [int(e)]
        while True:
            nxt = by_basal.get(frozenset(conn[col[-1], 3:].tolist()))
            if nxt is None or nxt == col[-1]:
                break
            col.append(int(nxt))
        if pl[col[-1]] == APICAL:
            cols[col[-1]] = col
    return cols


def tissue_height(mesh: TissueMesh, pos: np.ndarray):
    """Simulation::calculateTissueHeight (Simulation.cpp:3281-3330): from the first basal node,
    through the first element that holds the node as a basal node, adding |P0 - P3| of that
    element, up to an apical node.  Returns (height, layers)."""
    pl = mesh.node_placement
    basal = np.where(pl == BASAL)[0]
    if not len(basal):
        raise ValueError("no basal node")
    node = int(basal[0])
    h, layers = 0.0, 0
    conn = mesh.conn
    while pl[node] != APICAL:
        hit = np.where((conn[:, :3] == node).any(axis=1))[0]
        if not len(hit):
            raise ValueError("cannot walk from the basal to the apical surface")
        e = int(hit[0])
        d = pos[conn[e, 0]] - pos[conn[e, 3]]
        h += math.sqrt(float(d @ d))
        layers += 1
        node = int(conn[e, 3 + list(conn[e, :3]).index(node)])
    return h, layers


class GrowthGeometry:
    """Relative positions of the elements in the tissue's bounding box and their initial relative
    heights, kept as the reference keeps them (ShapeBase::calculateRelativePosInBoundingBox,
    Simulation::updateRelativePositionsToApicalPositioning, ShapeBase::setInitialZPosition)."""

    def __init__(self, mesh: TissueMesh, pos=None):
        self.mesh = mesh
        pos = mesh.pos if pos is None else np.asarray(pos, float).reshape(-1, 3)
        self.columns = element_columns(mesh)
        self.height, self.layers = tissue_height(mesh, pos)
        lo, _ = reference_bounding_box(pos)
        cz = pos[mesh.conn][:, :, 2].sum(axis=1) / 6
        rz = 1.0 - (cz - lo[2]) / self.height
        self.initial_z = np.clip(rz, 0.0, 1.0)
        self.initial_rel = self.relative_positions(pos)

    def relative_positions(self, pos):
        """(E, 2) relative x, y of every element's centre in the current bounding box, copied down
        the column from its apical element as the reference does (to the first layers - 2
        elements of the column)."""
        pos = np.asarray(pos, float).reshape(-1, 3)
        lo, hi = reference_bounding_box(pos)
        cen = pos[self.mesh.conn].sum(axis=1) / 6
        rel = np.stack([(cen[:, 0] - lo[0]) / (hi[0] - lo[0]), (cen[:, 1] - lo[1]) / (hi[1] - lo[1])], axis=1)
        for ap, col in self.columns.items():
            for k in range(max(0, self.layers - 2)):
                if k < len(col):
                    rel[col[k]] = rel[ap]
        return rel


def grid_index(rel: np.ndarray, nx: int, ny: int):
    """ShapeBase::convertRelativePosToGridIndex (ShapeBase.cpp:633-660)."""
    px = rel[:, 0] * float(F32(nx - 1))
    py = rel[:, 1] * float(F32(ny - 1))
    ix = np.floor(px).astype(int)
    iy = np.floor(py).astype(int)
    fx, fy = px - ix, py - iy
    hi = ix >= nx - 1
    ix[hi], fx[hi] = nx - 2, 1.0
    lo = ~hi & (ix < 0)
    ix[lo], fx[lo] = 0, 0.0
    hi = iy >= ny - 1
    iy[hi], fy[hi] = ny - 2, 1.0
    lo = ~hi & (iy < 0)
    iy[lo], fy[lo] = 0, 0.0
    return ix, iy, fx, fy


def grid_increments(gf: GridGrowth, rel: np.ndarray, initial_z: np.ndarray, dt: float, interpolation: int):
    """calculateFgFromGridCorners for every element (columnar tissue, tissue weight 1)."""
    nx, ny = gf.angles.shape
    ix, iy, fx, fy = grid_index(rel, nx, ny)
    E = len(rel)
    inc = np.tile(np.eye(3), (E, 1, 1))
    for e in range(E):
        if initial_z[e] < gf.zmin or initial_z[e] > gf.zmax:
            continue
        g, ang, elim = gf.corners[(int(ix[e]), int(iy[e]))]
        X, Y = fx[e], fy[e]
        angle = 0.0
        if interpolation == 0:
            k = (3 if Y > 0.5 else 1) if X > 0.5 else (2 if Y > 0.5 else 0)
            growth = g[k].copy()
        else:
            w = np.array([(1.0 - X) * (1.0 - Y), X * (1.0 - Y), (1.0 - X) * Y, X * Y])
            frac_elim = 0.0
            for k in range(4):
                if elim[k]:
                    frac_elim += w[k]
            angle = ang[0] * w[0] + ang[1] * w[1] + ang[2] * w[2] + ang[3] * w[3]
            if frac_elim > 0:
                angle = 0.0 if frac_elim >= 0.9999999 else angle / (1.0 - frac_elim)
            growth = g[0] * w[0] + g[1] * w[1] + g[2] * w[2] + g[3] * w[3]
        d = np.diag(np.exp(growth * dt))
        if angle != 0.0:
            R = _rot(angle)
            d = R @ d @ R.T
        inc[e] = d
    return inc


def increments(functions, geo: GrowthGeometry, pos, time_sec: float, dt: float, pinned: bool = True,
               interpolation: int = 1):
    """Per-element growth increments (E, 3, 3) of the step starting at time_sec, the product of
    the active functions' increments in their order (later functions multiply from the left)."""
    pos = np.asarray(pos, float).reshape(-1, 3)
    E = geo.mesh.n_elements
    total = np.tile(np.eye(3), (E, 1, 1))
    rel = None
    for f in functions:
        if not (f.t0 <= time_sec < f.t1):
            continue
        if f.kind == 1:
            inc = _rate_increment(np.tile(np.asarray(f.rate, float), (E, 1)), dt, _rot(f.angle))
        elif f.kind == 2:
            cen = pos[geo.mesh.conn].sum(axis=1) / 6
            cx, cy = float(F32(f.centre[0])), float(F32(f.centre[1]))
            r_in, r_out = F32(f.inner), F32(f.outer)
            d2 = (cx - cen[:, 0]) ** 2 + (cy - cen[:, 1]) ** 2
            inside = (d2 > float(r_in * r_in)) & (d2 < float(r_out * r_out))
            dist = np.sqrt(d2).astype(F32)
            sf = 1.0 - ((dist - r_in) / (r_out - r_in)).astype(F32).astype(np.float64)
            rates = np.where(inside[:, None], np.asarray(f.rate, float)[None, :] * sf[:, None], 0.0)
            inc = _rate_increment(rates, dt, _rot(f.angle))
            inc[~inside] = np.eye(3)
        elif f.kind == 3:
            if pinned:
                r = geo.initial_rel
            else:
                if rel is None:
                    rel = geo.relative_positions(pos)
                r = rel
            inc = grid_increments(f, r, geo.initial_z, dt, interpolation)
        else:
            raise NotImplementedError(f"growth function type {f.kind}")
        total = np.einsum("eij,ejk->eik", inc, total)
    return total
