"""Growth functions of a model input, as the device evaluates them every growth update
(``tf_set_growth_functions``; ``Simulation::calculateGrowth``, Simulation.cpp:9651-9670): every
step starts each element's growth increment at the identity and left-multiplies the increment of
every growth function active at the step's start time (``ShapeBase::updateGrowthIncrement``,
ShapeBase.cpp:567-586); the device then applies the growth-rotation rule and Fg <- dFg Fg.

Function types (``GrowthFunctionTypes.h``), columnar layer only:

* 1, uniform (``calculateGrowthUniform``, Simulation.cpp:9757-9797): R diag(exp(r dt)) R^T with R
  the rotation by the function's xy angle (``calculateFgFromRates``, ShapeBase.cpp:396-421);
* 2, ring (``calculateGrowthRing``, Simulation.cpp:9801-9898): the same with the rate scaled by
  1 - (d - r_in) / (r_out - r_in) for elements whose centre lies between the radii;
* 3, grid (``calculateGrowthGridBased``, Simulation.cpp:10065-10117): rates and xy angles on an
  nx x ny grid over the tissue's bounding box, taken at the element's relative position (pinned
  or current) from the cell's four corners, nearest corner or bilinear
  (``calculateFgFromGridCorners``, ShapeBase.cpp:432-565), with the corners' angles first brought
  next to each other (``preCalculateAnglesForCompatibleAveraging``, GrowthFunctionTypes.h:434-587)
  and the z-range test on the element's initial relative height (``isGrowthRateApplicable``,
  ShapeBase.cpp:277-300).

This module holds the functions' parameters, the grid-file reader, and the per-element geometry
set up once from the initial mesh (initial relative heights, the column each element copies its
relative position from).  Values the reference reads into ``float`` variables are rounded to
float here as well.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .mesh import APICAL, BASAL, TissueMesh

F32 = np.float32


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
    rate: tuple                      # rate per second at the inner radius (0 at the outer one)
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
    """The per-element geometry the growth functions read, set up from the initial mesh as the
    reference sets it up: the initial relative height of every element's centre
    (ShapeBase::setInitialZPosition) and, for the relative positions in the bounding box
    (ShapeBase::calculateRelativePosInBoundingBox), the element each element takes its position
    from -- its column's apical element for the first layers - 2 elements of the column
    (Simulation::updateRelativePositionsToApicalPositioning), else itself."""

    def __init__(self, mesh: TissueMesh, pos=None):
        self.mesh = mesh
        pos = mesh.pos if pos is None else np.asarray(pos, float).reshape(-1, 3)
        self.columns = element_columns(mesh)
        self.height, self.layers = tissue_height(mesh, pos)
        lo, _ = reference_bounding_box(pos)
        cz = pos[mesh.conn][:, :, 2].sum(axis=1) / 6
        rz = 1.0 - (cz - lo[2]) / self.height
        self.initial_z = np.clip(rz, 0.0, 1.0)
        self.rel_src = np.arange(mesh.n_elements, dtype=np.int32)
        for ap, col in self.columns.items():
            for k in range(max(0, self.layers - 2)):
                if k < len(col):
                    self.rel_src[col[k]] = ap
