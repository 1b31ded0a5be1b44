"""Closed-form pins for two oracle branches that only GPU-vs-oracle comparisons touched before:
the growth-rotation rule and the nodal viscosity surfaces.

Growth rotation (SC/ShapeBase.cpp:1656-1811, :1215-1221).  For current positions x = Rz(t) X (a
rigid rotation of the reference shape about z) every Gauss point has F = Rz(t), so TriPointF =
Rz(t) and its SVD rotation U V^T = Rz(t).  disassembleRotationMatrixForZ takes
tethaZ = atan2(m01, m00) = -t and, when |tethaZ| > 0.017, rebuilds [[c, -s], [s, c]] from
c = cos(-t), s = sin(-t), i.e. Rz(-t); CalculateGrowthRotationByF transposes it into
GrowthStrainsRotMat = Rz(t), and growShapeByFg rotates the increment to GRM^T inc GRM.  From
Fg = I one growth step therefore gives Fg = Rz(t)^T D Rz(t) with D = diag(exp(r dt)), and D itself
when |t| <= 0.017.  With anisotropic D the xy entry of Rz(t)^T D Rz(t) changes sign under a
transposed product or a flipped atan2 sign, so either mistake fails here.

Viscosity surfaces (SC/ShapeBase.cpp:4041-4077, SC/Prism.cpp:2424-2503).  Each exposed face (the
apical face of an apical prism, the basal face of a basal prism) gives a third of its area to each
of its three nodes.  A third of a triangle's area at each vertex integrates every linear function
exactly over the triangle, so for a flat extruded mesh sum_n S_A(n) = 2 A and
sum_n S_A(n) (x_n, y_n) = 2 A (cx, cy), with A and (cx, cy) the area and centroid of the outline
polygon by the shoelace formula -- an independent route (the polygon, not the triangles).  A missing
/3, a face given to the wrong nodes or an interior face counted fails one of them.
"""
import numpy as np
import pytest

from tissuefolding_paperversion_amd import configs


def _rz(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _small_ellipse():
    return configs.wing_disc_tissue(a=30.0, b=18.0, h=3.66, z_levels=(0.0, 2.0, 7.0, 12.0), seed=4, growth_seed=5,
                                    name="pin ellipse")


@pytest.mark.parametrize("theta", [0.3, -0.5, 0.02, 0.015, -0.016, 2.0])
def test_growth_rotation_closed_form(oracle, theta):
    t = _small_ellipse()
    rx, ry, rz = 3.0e-5, 0.4e-5, 1.0e-5  # per second, anisotropic in x-y
    t.rates = np.tile([rx, ry, rz], (t.n_elements, 1))
    dt = 60.0
    o = oracle.OracleTissue(t)
    X = t.mesh.pos
    centre = np.array([7.0, -4.0, 1.0])  # F is independent of the rotation centre
    o.pos = ((X - centre) @ _rz(theta).T + centre).ravel().copy()
    n_rot = o.grow(dt)
    D = np.diag(np.exp(np.array([rx, ry, rz]) * dt))
    R = _rz(theta)
    rotated = abs(theta) > 0.017
    expect = R.T @ D @ R if rotated else D
    Fg = o.Fg.reshape(-1, 3, 3)
    assert np.max(np.abs(Fg - expect)) < 1e-12, (theta, Fg[0], expect)
    assert n_rot == (t.n_elements if rotated else 0)
    InvFg = o.InvFg.reshape(-1, 3, 3)
    assert np.max(np.abs(InvFg @ Fg - np.eye(3))) < 1e-13
    if rotated and abs(np.sin(2 * theta)) > 0.1:
        # the opposite orientation (R D R^T) is measurably different: the test can tell them apart
        assert np.max(np.abs(R @ D @ R.T - expect)) > 1e-6


def _shoelace(p):
    x, y = p[:, 0], p[:, 1]
    xn, yn = np.roll(x, -1), np.roll(y, -1)
    cr = x * yn - xn * y
    A = 0.5 * cr.sum()
    return A, (x + xn).dot(cr) / (6 * A), (y + yn).dot(cr) / (6 * A)


def test_viscosity_surfaces_integrate_the_outline(oracle):
    t = _small_ellipse()
    o = oracle.OracleTissue(t)
    shift = np.array([17.0, -9.0, 0.0])  # a centroid away from the origin
    pos = t.mesh.pos + shift
    o.pos = pos.ravel().copy()
    o.update_viscosity_surfaces()
    Sa = o.Sa
    # the outline: the first points of the basal layer, counter-clockwise
    nb = int(np.argmin(t.mesh.at_circumference))  # boundary points come first in every layer
    P = t.n_nodes // 4
    assert t.mesh.at_circumference[:nb].all() and not t.mesh.at_circumference[nb:P].any()
    A, cx, cy = _shoelace(pos[:nb, :2])
    assert A > 0
    assert abs(Sa.sum() - 2 * A) < 1e-10 * A
    assert abs(Sa.dot(pos[:, 0]) - 2 * A * cx) < 1e-10 * A * max(1.0, abs(cx))
    assert abs(Sa.dot(pos[:, 1]) - 2 * A * cy) < 1e-10 * A * max(1.0, abs(cy))
    # only the apical and basal surfaces carry area: the two inner node layers get none
    assert np.all(Sa[P:3 * P] == 0.0)
    assert abs(Sa[:P].sum() - A) < 1e-10 * A and abs(Sa[3 * P:].sum() - A) < 1e-10 * A
    assert np.all(Sa[:P] > 0) and np.all(Sa[3 * P:] > 0)
