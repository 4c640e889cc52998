"""The oracle's mesh colliders: rigid-particle sampling (src_mpm/sampling/sample_trimesh.rs)
and the mesh part of the contact distance field (src_mpm_shaders/solver/p2g_cdf.rs),
pinned by geometry: samples on their triangles and covering them, shared edges sampled
once, a closed box mesh giving the same node field as the equal cuboid (projected by
parry's rules), and a triangle's field against a barycentric projection."""
import math

import numpy as np
import pytest

from nexus_amd import scenes
import oracle as O


def _tri_points(v, n, rng):
    r = rng.random((n, 2))
    flip = r.sum(1) > 1
    r[flip] = 1 - r[flip]
    return v[0] + r[:, :1] * (v[1] - v[0]) + r[:, 1:] * (v[2] - v[0])


def _bary(p, a, b, c):
    v0, v1, v2 = b - a, c - a, p - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    l1 = (d11 * d20 - d01 * d21) / den
    l2 = (d00 * d21 - d01 * d20) / den
    return np.stack([1 - l1 - l2, l1, l2], -1)


@pytest.mark.parametrize("tri", [
    [[0, 0, 0], [1, 0, 0], [0.3, 0.7, 0.1]],
    [[0.2, -0.1, 0.4], [-0.3, 0.5, 0.0], [0.6, 0.2, -0.5]],
    [[0, 0, 0], [0.05, 0.9, 0], [0.0, 0.0, 0.04]],          # thin
])
def test_samples_lie_on_their_triangle_and_cover_it(tri):
    """Every sample is on its triangle, none at a vertex; every point of the triangle has
    a sample or a vertex within h/2 (rows spaced h / sqrt 2 make square cells whose
    half-diagonal is h/2: a spacing of h would leave gaps of 0.7 h)."""
    from scipy.spatial import cKDTree
    h = 1.0 / 64
    v = np.array(tri, np.float32)
    pts, ids = O.sample_trimesh(v, np.array([[0, 1, 2]], np.uint32), h)
    assert len(pts) > 10 and np.all(ids == 0)
    vd = v.astype(np.float64)
    n = np.cross(vd[1] - vd[0], vd[2] - vd[0])
    n /= np.linalg.norm(n)
    assert np.abs((pts - vd[0]) @ n).max() < 1e-6
    lam = _bary(pts.astype(np.float64), *vd)
    assert lam.min() > -1e-5
    assert np.min(np.linalg.norm(pts[:, None, :] - vd[None], axis=-1)) > 1e-6
    q = _tri_points(vd, 20000, np.random.default_rng(0))
    d, _ = cKDTree(np.concatenate([pts, vd])).query(q)
    assert d.max() <= 0.5 * h * 1.001


def test_shared_edges_are_sampled_once():
    """Two triangles sharing the diagonal of a quad: the quad's samples are those of the
    two triangles alone minus one copy of the diagonal's ceil(L / s) - 1 points."""
    h = 1.0 / 32
    v = np.array([[0, 0, 0], [0.5, 0, 0], [0.5, 0.3, 0.1], [0, 0.3, 0.1]], np.float32)
    t1, t2 = np.array([[0, 1, 2]], np.uint32), np.array([[0, 2, 3]], np.uint32)
    n1 = len(O.sample_trimesh(v, t1, h)[0])
    n2 = len(O.sample_trimesh(v, t2, h)[0])
    nq = len(O.sample_trimesh(v, np.concatenate([t1, t2]), h)[0])
    s = np.float32(h) / np.sqrt(np.float32(2))
    diag = v[2] - v[0]
    L = np.sqrt(np.float32(diag @ diag))
    assert nq == n1 + n2 - (int(math.ceil(np.float32(L / s))) - 1)
    # a closed box: 12 triangles, 18 distinct edges
    bv, bt = scenes.box_mesh((0.2, 0.1, 0.15))
    pts, ids = O.sample_trimesh(bv, bt, h)
    assert len(np.unique(np.round(pts / 1e-6).astype(np.int64), axis=0)) == len(pts)


def _shell(c, he, h, n, seed):
    rng = np.random.default_rng(seed)
    return c + (rng.random((n, 3)) - 0.5) * 2 * (np.asarray(he) + 3 * h)


@pytest.mark.parametrize("bits", [32, 64])
def test_box_mesh_field_equals_the_cuboids(bits):
    """A closed, turned box as a mesh and as a cuboid: on every node both reach, the same
    distance and side (the cuboid's by parry's projection, the mesh's from its faces);
    the cuboid's field also covers nodes near edges and corners that no face interior
    reaches, the mesh's nodes up to two cells off its faces."""
    h = 1.0 / 32
    c, he = np.array([0.5, 0.5, 0.5]), np.array([0.15, 0.1, 0.12])
    parts = scenes.particles(_shell(c, he, h, 6000, 1), h / 4, 1000.0, scenes.model_neo_hookean(1e5, 0.3))
    q = np.array([math.sin(0.2) * 0.6, math.sin(0.2) * 0.8, 0, math.cos(0.2)], np.float32)
    vv, tt = scenes.box_mesh(he)
    sa = scenes.make_scene(parts, h, 1e-4, colliders=scenes.cuboid(c, he, rotation=q), cpic=True, gravity=(0, 0, 0))
    sb = scenes.make_scene(parts, h, 1e-4, cpic=True, gravity=(0, 0, 0))
    sb.meshes = [scenes.trimesh(vv, tt, c, rotation=q)]
    oa, ob = O.Oracle(sa, bits), O.Oracle(sb, bits)
    oa.step(1)
    ob.step(1)
    ga, gb = oa.grid(), ob.grid()
    assert np.array_equal(oa.blocks()["vid"], ob.blocks()["vid"])
    ma, mb = (ga["cdf_aff"] & 1).astype(bool), (gb["cdf_aff"] & 1).astype(bool)
    both = ma & mb
    assert both.sum() > 500 and (mb & ~ma).sum() > 0 and (ma & ~mb).sum() > 0
    assert np.abs(ga["cdf_dist"][both] - gb["cdf_dist"][both]).max() < (1e-6 if bits == 32 else 1e-9)
    assert np.array_equal((ga["cdf_aff"][both] >> 16) & 1, (gb["cdf_aff"][both] >> 16) & 1)
    assert np.all(gb["cdf_closest"][mb] == 0) and np.all(gb["cdf_closest"][~mb] == 0xFFFFFFFF)


def test_triangle_field_against_barycentric_projection():
    """One posed triangle: a node carries the mesh's affinity bit only where its
    projection onto the plane falls inside the triangle, with the plane distance and the
    side of the normal ab x ac; every node of an active block within 0.4 h of the
    triangle's interior is reached (a sample lies within h/2 of its projection)."""
    h = 1.0 / 32
    tri = np.array([[-0.2, -0.02, -0.15], [0.22, 0.03, -0.1], [0.0, -0.01, 0.2]], np.float32)
    c = np.array([0.5, 0.4, 0.5])
    q = np.array([0.0, math.sin(0.15), 0.0, math.cos(0.15)], np.float32)
    rng = np.random.default_rng(3)
    x = c + (rng.random((8000, 3)) - 0.5) * np.array([0.5, 0.2, 0.5])
    parts = scenes.particles(x, h / 4, 1000.0, scenes.model_neo_hookean(1e5, 0.3))
    sc = scenes.make_scene(parts, h, 1e-4, cpic=True, gravity=(0, 0, 0))
    sc.meshes = [scenes.trimesh(tri, np.array([[0, 1, 2]], np.uint32), c, rotation=q)]
    o = O.Oracle(sc, 64)
    o.step(1)
    g, b = o.grid(), o.blocks()
    # world triangle (rotation about y by 0.3 rad)
    ang = 0.3
    R = np.array([[math.cos(ang), 0, math.sin(ang)], [0, 1, 0], [-math.sin(ang), 0, math.cos(ang)]])
    wt = tri.astype(np.float64) @ R.T + c
    loc = np.arange(64)
    cells = b["vid"][:, None, :] * 4 + np.stack([loc & 3, (loc >> 2) & 3, loc >> 4], -1)[None]
    nodes = cells.reshape(-1, 3) * h
    n = np.cross(wt[1] - wt[0], wt[2] - wt[0])
    n /= np.linalg.norm(n)
    sd = (nodes - wt[0]) @ n
    lam = _bary(nodes - np.outer(sd, n), *wt)
    aff = g["cdf_aff"].reshape(-1)
    hit = (aff & 1).astype(bool)
    assert hit.sum() > 100
    assert lam[hit].min() > -1e-6
    # (1e-6: the pose's f32 quaternion against the exact rotation)
    assert np.allclose(g["cdf_dist"].reshape(-1)[hit], np.abs(sd[hit]), rtol=0, atol=1e-6)
    assert np.array_equal(((aff[hit] >> 16) & 1).astype(bool), sd[hit] < 0)
    must = (lam.min(1) > 1e-3) & (np.abs(sd) <= 0.4 * h)
    assert must.sum() > 30 and np.all(hit[must])


def test_sand_rests_on_a_mesh_ramp():
    """Sand dropped on a mesh ramp (CPIC) lands on it and stays above its plane (within a
    tenth of a cell), held by the mesh's contact field alone."""
    sc = scenes.sand_on_mesh_ramp(4000, seed=21)
    ramp = sc.meshes[0]
    o = O.Oracle(sc, 32)
    o.step(160)
    q = ramp.body["rotation"][0].astype(np.float64)
    ang = 2 * math.atan2(q[1], q[3])
    R = np.array([[math.cos(ang), 0, math.sin(ang)], [0, 1, 0], [-math.sin(ang), 0, math.cos(ang)]])
    loc = (o.s.x.astype(np.float64) - ramp.body["translation"][0]) @ R        # body frame
    over = (np.abs(loc[:, 0]) < 2.8) & (np.abs(loc[:, 2]) < 2.8)
    d = (loc[:, 1] - 0.2 * loc[:, 0]) / math.sqrt(1.04)                       # plane y = 0.2 x
    assert over.sum() > 1000 and (o.s.affinity & 0xFFFF != 0).sum() > 1000
    assert d[over].min() > -0.1 * sc.cell_width
    assert np.all(o.s.enabled == 1)
