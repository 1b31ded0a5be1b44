"""Tissue meshes: seeded 2-D triangulations extruded into layers of 6-node prisms.

Host-side data preparation for the Newton-Raphson path.  Covers:

* the reference's ``.mesh`` text format (reader ``Simulation::initiateNodesFromMeshInput`` /
  ``initiateElementsFromMeshInput`` / ``initiatePrismFromMeshInput``, Simulation.cpp:2074-2111,
  :2222-2239; writer ``ToolBox/MeshGeneration/2DEllipse/src/Ellipse.cpp:531-586``);
* the prism node order (nodes 0-2 basal, 3-5 apical, counter-clockwise seen from apical,
  ``Prism::checkRotationConsistency3D``, Prism.cpp:364-412);
* node and element placement (``ShapeBase::setTissuePlacement``, ShapeBase.cpp:705-772) and the
  exposed apical/basal faces that carry external viscosity
  (``Prism::assignExposedSurfaceAreaIndices``, Prism.cpp:2674-2711);
* a seeded generator for discs and ellipses (jittered hexagonal point set inside the outline,
  evenly spaced outline points, Delaunay triangulation) that replaces the reference's external
  ``triangle`` call.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Sequence

import numpy as np

# node placements (Node::tissuePlacement)
BASAL, APICAL, MIDLINE, LATERAL = 0, 1, 2, 3


@dataclasses.dataclass
class TissueMesh:
    pos: np.ndarray            # (N, 3) float64 node positions
    conn: np.ndarray           # (E, 6) int32 node ids, 0-2 basal, 3-5 apical
    xref: np.ndarray           # (E, 18) float64 element-local reference positions
    node_placement: np.ndarray  # (N,) int8
    node_type: np.ndarray      # (N,) int8, 0 = columnar
    at_circumference: np.ndarray  # (N,) bool
    layer: np.ndarray | None = None  # (E,) int32 prism layer index from the bottom, if known

    @property
    def n_nodes(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.conn.shape[0])

    def element_placement(self) -> tuple[np.ndarray, np.ndarray]:
        """(placement, spans_whole_tissue) per element, ShapeBase.cpp:705-772."""
        npl = self.node_placement[self.conn]
        has_ap = (npl == APICAL).any(axis=1)
        has_ba = (npl == BASAL).any(axis=1)
        has_lat = (npl == LATERAL).any(axis=1)
        pl = np.full(self.n_elements, MIDLINE, np.int8)
        spans = np.zeros(self.n_elements, bool)
        pl[has_ba & ~has_ap] = BASAL
        pl[has_ap & ~has_ba] = APICAL
        spans[has_ap & has_ba] = True
        pl[has_lat] = LATERAL
        spans[has_lat] = False
        return pl, spans

    def exposed_faces(self) -> np.ndarray:
        """bit0: basal face (0,1,2) exposed, bit1: apical face (3,4,5) exposed (columnar tissue)."""
        pl, spans = self.element_placement()
        ex = np.zeros(self.n_elements, np.uint8)
        ex[pl == APICAL] |= 2
        ex[pl == BASAL] |= 1
        both = (pl == MIDLINE) & spans
        ex[both] |= 3
        return ex


# --------------------------------------------------------------------------------------------
# .mesh format
# --------------------------------------------------------------------------------------------
def read_mesh(path: str) -> TissueMesh:
    """Read the reference's .mesh text format (whitespace separated)."""
    with open(path) as f:
        tok = f.read().split()
    it = iter(tok)
    n = int(next(it))
    pos = np.zeros((n, 3))
    plc = np.zeros(n, np.int8)
    typ = np.zeros(n, np.int8)
    circ = np.zeros(n, bool)
    for i in range(n):
        pos[i] = [float(next(it)) for _ in range(3)]
        plc[i] = int(next(it))
        typ[i] = int(next(it))
        circ[i] = int(next(it)) != 0
    ne = int(next(it))
    conn = np.zeros((ne, 6), np.int32)
    xref = np.zeros((ne, 18))
    for e in range(ne):
        shape = int(next(it))
        if shape != 1:
            raise ValueError(f"element {e}: shape type {shape} is not a prism")
        conn[e] = [int(next(it)) for _ in range(6)]
        xref[e] = [float(next(it)) for _ in range(18)]
    m = TissueMesh(pos, conn, xref, plc, typ, circ)
    fix_orientation(m)
    return m


def write_mesh(mesh: TissueMesh, path: str) -> None:
    with open(path, "w") as f:
        f.write(f"{mesh.n_nodes}\n")
        for i in range(mesh.n_nodes):
            p = mesh.pos[i]
            f.write(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r} {int(mesh.node_placement[i])} {int(mesh.node_type[i])} "
                    f"{int(mesh.at_circumference[i])}\n")
        f.write(f"{mesh.n_elements}\n")
        for e in range(mesh.n_elements):
            ids = " ".join(str(int(v)) for v in mesh.conn[e])
            ref = " ".join(repr(float(v)) for v in mesh.xref[e])
            f.write(f"1 {ids} {ref}\n")
        f.write("0\n")


def fix_orientation(mesh: TissueMesh) -> int:
    """Swap nodes 1<->2 and 4<->5 of prisms whose basal triangle is clockwise seen from apical
    (Prism::checkRotationConsistency3D, Prism.cpp:364-412).  Returns the number of swaps."""
    P = mesh.pos[mesh.conn]
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    view = P[:, 0] - P[:, 3]
    bad = np.einsum("ij,ij->i", view, n) > 0
    if bad.any():
        c = mesh.conn[bad].copy()
        c[:, [1, 2, 4, 5]] = c[:, [2, 1, 5, 4]]
        mesh.conn[bad] = c
        x = mesh.xref[bad].reshape(-1, 6, 3).copy()
        x[:, [1, 2, 4, 5]] = x[:, [2, 1, 5, 4]]
        mesh.xref[bad] = x.reshape(-1, 18)
    return int(bad.sum())


# --------------------------------------------------------------------------------------------
# Generator
# --------------------------------------------------------------------------------------------
def _hex_lattice(halfx: float, halfy: float, h: float) -> np.ndarray:
    dy = h * math.sqrt(3.0) / 2.0
    ny = int(math.ceil(halfy / dy)) + 1
    nx = int(math.ceil(halfx / h)) + 2
    pts = []
    for j in range(-ny, ny + 1):
        off = 0.5 * h if (j % 2) else 0.0
        xs = np.arange(-nx, nx + 1) * h + off
        pts.append(np.stack([xs, np.full_like(xs, j * dy)], 1))
    return np.concatenate(pts)


def _ellipse_outline(a: float, b: float, h: float) -> np.ndarray:
    """Even number of outline points, equally spaced in arc length, with points at (+-a, 0)."""
    t = np.linspace(0.0, 2 * math.pi, 20001)
    xy = np.stack([a * np.cos(t), b * np.sin(t)], 1)
    s = np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(xy, axis=0), axis=1))])
    total = s[-1]
    nb = max(8, int(round(total / h)))
    nb += nb % 2
    # first half by arc length from theta = 0 to pi; second half mirrored in y -> (+-a, 0) exact
    half = nb // 2
    targets = np.arange(half + 1) * (total / 2) / half
    th = np.interp(targets, s, t)
    upper = np.stack([a * np.cos(th), b * np.sin(th)], 1)
    upper[0] = [a, 0.0]
    upper[-1] = [-a, 0.0]
    lower = upper[1:-1][::-1].copy()
    lower[:, 1] *= -1.0
    return np.concatenate([upper[:-1], [upper[-1]], lower])


def planar_ellipse(a: float, b: float, h: float, seed: int, jitter: float = 0.18):
    """Seeded planar triangulation of an ellipse (a = b gives a disc).

    Returns (points (P,2), triangles (T,3) counter-clockwise, boundary mask (P,)).  The centre
    point (0, 0) is always a vertex, and the outline has vertices at (+-a, 0).
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    outline = _ellipse_outline(a, b, h)
    lat = _hex_lattice(a, b, h)
    centre = np.all(np.abs(lat) < 1e-12, axis=1)
    jit = rng.uniform(-jitter * h, jitter * h, size=lat.shape)
    jit[centre] = 0.0
    lat = lat + jit
    margin = 0.55 * h
    ai, bi = a - margin, b - margin
    inside = (lat[:, 0] / ai) ** 2 + (lat[:, 1] / bi) ** 2 < 1.0
    interior = lat[inside]
    pts = np.concatenate([outline, interior])
    bnd = np.zeros(len(pts), bool)
    bnd[: len(outline)] = True
    tri = Delaunay(pts).simplices.astype(np.int64)
    # drop degenerate triangles (collinear outline triples) and orient counter-clockwise
    p = pts[tri]
    area2 = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])
    keep = np.abs(area2) > 1e-9 * h * h
    tri, area2 = tri[keep], area2[keep]
    neg = area2 < 0
    tri[neg] = tri[neg][:, [0, 2, 1]]
    return pts, tri.astype(np.int32), bnd


def extrude(pts2: np.ndarray, tri: np.ndarray, boundary: np.ndarray, z_levels: Sequence[float]) -> TissueMesh:
    """Extrude a planar triangulation into len(z_levels)-1 layers of prisms (columnar tissue)."""
    z = np.asarray(z_levels, float)
    L = len(z)
    P = len(pts2)
    pos = np.zeros((P * L, 3))
    plc = np.full(P * L, MIDLINE, np.int8)
    for l in range(L):
        pos[l * P:(l + 1) * P, :2] = pts2
        pos[l * P:(l + 1) * P, 2] = z[l]
    plc[:P] = BASAL
    plc[(L - 1) * P:] = APICAL
    circ = np.tile(boundary, L)
    conns = []
    layers = []
    for l in range(L - 1):
        c = np.concatenate([tri + l * P, tri + (l + 1) * P], axis=1)
        conns.append(c)
        layers.append(np.full(len(tri), l, np.int32))
    conn = np.concatenate(conns).astype(np.int32)
    xref = pos[conn].reshape(-1, 18).copy()
    m = TissueMesh(pos, conn, xref, plc, np.zeros(P * L, np.int8), circ, np.concatenate(layers))
    fix_orientation(m)
    return m


def disc_mesh(radius: float = 50.0, h: float = 5.85, z_levels=(0.0, 5.0, 10.0), seed: int = 1) -> TissueMesh:
    pts, tri, bnd = planar_ellipse(radius, radius, h, seed)
    return extrude(pts, tri, bnd, z_levels)


def ellipse_mesh(a: float = 200.0, b: float = 120.0, h: float = 3.66, z_levels=(0.0, 2.0, 7.0, 12.0),
                 seed: int = 2) -> TissueMesh:
    pts, tri, bnd = planar_ellipse(a, b, h, seed)
    return extrude(pts, tri, bnd, z_levels)


def single_prism(scale: float = 1.0) -> TissueMesh:
    """One right prism with a unit-ish triangular base (basal z = 0, apical z = 1)."""
    pts = np.array([[0.0, 0.0], [1.0, 0.1], [0.2, 0.9]]) * scale
    tri = np.array([[0, 1, 2]], np.int32)
    return extrude(pts, tri, np.ones(3, bool), [0.0, scale])
