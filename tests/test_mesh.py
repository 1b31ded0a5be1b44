"""Host-side mesh layer: .mesh format, prism orientation, placements, generator quality."""
import os

import numpy as np

from tissuefolding_paperversion_amd import configs, mesh as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_reference_fixture_meshes():
    m = M.read_mesh(os.path.join(GOLD, "OnePrism.mesh"))
    assert m.n_nodes == 6 and m.n_elements == 1
    assert np.allclose(m.xref[0].reshape(6, 3), m.pos[m.conn[0]])
    m2 = M.read_mesh(os.path.join(GOLD, "Wind48hr_02.mesh"))
    assert m2.n_nodes == 64 and m2.n_elements == 44
    # every prism is counter-clockwise seen from apical after reading
    P = m2.pos[m2.conn]
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    assert np.all(np.einsum("ij,ij->i", P[:, 0] - P[:, 3], n) <= 0)


def test_mesh_roundtrip(tmp_path):
    m = M.disc_mesh(radius=12.0, h=4.0, seed=3)
    p = tmp_path / "d.mesh"
    M.write_mesh(m, str(p))
    r = M.read_mesh(str(p))
    assert np.array_equal(r.conn, m.conn)
    assert np.array_equal(r.pos, m.pos)
    assert np.array_equal(r.xref, m.xref)
    assert np.array_equal(r.node_placement, m.node_placement)
    assert np.array_equal(r.at_circumference, m.at_circumference)


def test_orientation_fix_swaps_clockwise_prisms():
    m = M.single_prism()
    m.conn[0] = m.conn[0][[0, 2, 1, 3, 5, 4]]
    m.xref[0] = m.pos[m.conn[0]].ravel()
    assert M.fix_orientation(m) == 1
    assert M.fix_orientation(m) == 0


def test_placements_follow_reference_rules():
    m = M.disc_mesh(radius=15.0, h=4.0, z_levels=(0, 3, 6, 9), seed=2)
    pl, spans = m.element_placement()
    L = m.layer
    assert np.all(pl[L == 0] == M.BASAL) and np.all(pl[L == 2] == M.APICAL) and np.all(pl[L == 1] == M.MIDLINE)
    assert not spans.any()
    ex = m.exposed_faces()
    assert np.all(ex[L == 0] == 1) and np.all(ex[L == 2] == 2) and np.all(ex[L == 1] == 0)
    one = M.disc_mesh(radius=15.0, h=4.0, z_levels=(0, 3), seed=2)
    pl1, sp1 = one.element_placement()
    assert np.all(pl1 == M.MIDLINE) and sp1.all() and np.all(one.exposed_faces() == 3)


def test_external_viscosity_assignment():
    m = M.disc_mesh(radius=15.0, h=4.0, z_levels=(0, 3, 6, 9), seed=2)
    v = configs.assign_external_viscosity(m, 100.0, 40.0).reshape(-1, 3)[:, 0]
    p = m.node_placement
    assert np.all(v[p == M.BASAL] == 40.0) and np.all(v[p == M.APICAL] == 100.0)
    mid = p == M.MIDLINE
    assert np.all(v[mid & m.at_circumference] == 40.0) and np.all(v[mid & ~m.at_circumference] == 0.0)


def test_generator_sizes_and_quality():
    t1 = configs.config1()
    assert 900 <= t1.n_elements <= 1200
    t2 = configs.config2()
    assert 37000 <= t2.n_elements <= 41000
    for m in (t1.mesh, M.ellipse_mesh(a=40, b=25, h=3.0, seed=4)):
        P = m.pos[m.conn[:, :3]][:, :, :2]
        e = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 1], P[:, 0] - P[:, 2]], 1)
        ang = []
        for k in range(3):
            a, b = -e[:, k - 1], e[:, k]
            c = np.einsum("ij,ij->i", a, b) / np.linalg.norm(a, axis=1) / np.linalg.norm(b, axis=1)
            ang.append(np.degrees(np.arccos(np.clip(c, -1, 1))))
        assert np.min(ang) > 15.0
    # the pinned nodes of config 1 lie on the symmetry line y = 0 and at the centre
    fx = np.where(t1.fixed.reshape(-1, 3)[:, 1])[0]
    assert np.allclose(t1.mesh.pos[fx, 1], 0.0)


def test_growth_map_is_seeded():
    a = configs.gaussian_bump_rates(M.ellipse_mesh(a=40, b=25, h=4.0, seed=4), 40, 25, seed=3)
    b = configs.gaussian_bump_rates(M.ellipse_mesh(a=40, b=25, h=4.0, seed=4), 40, 25, seed=3)
    assert np.array_equal(a, b) and a[:, 2].max() == 0.0 and a[:, :2].max() * 3600 > 0.004
