"""Pins of the oracle's projection onto triangle-mesh colliders in the grid collider pass
(CPIC off; grid_update_collide.rs:53-67, TriMesh::project_local_point,
src_rbd_shaders/shapes/trimesh.rs:155-391): the nearest point against a dense sampling
of the triangles, the inside test against closed and open meshes whose inside is known,
and a substep against a closed box mesh equal to the same substep against the cuboid of
that box."""
import numpy as np
import pytest

import oracle as O
from nexus_amd import scenes

from test_oracle import box_boundary_closest

RNG = np.random.default_rng(777)


def _tri_samples(a, b, c, k=400):
    u, v = np.meshgrid(np.linspace(0, 1, k), np.linspace(0, 1, k), indexing="ij")
    keep = u + v <= 1.0
    u, v = u[keep], v[keep]
    return a[None] + u[:, None] * (b - a)[None] + v[:, None] * (c - a)[None]


def test_triangle_projection_is_the_nearest_point():
    """One triangle, points all around it (vertex, edge and face regions on both sides):
    the projection's distance equals the least distance to a dense sampling of the
    triangle (within the sampling step) and the projection lies on the triangle."""
    body = scenes.trimesh(np.zeros((3, 3)), np.array([[0, 1, 2]]), (0.0, 0.0, 0.0)).body
    for _ in range(120):
        tri = RNG.normal(size=(3, 3)).astype(np.float32)
        a, b, c = (tri[i].astype(np.float64) for i in range(3))
        samples = _tri_samples(a, b, c)
        step = max(np.linalg.norm(b - a), np.linalg.norm(c - a)) / 399
        p = a + RNG.normal(size=3) * 1.5
        proj, _ = O.project_trimesh(tri, np.array([[0, 1, 2]]), body, p, 1e9, 64)
        d_sample = np.sqrt(((samples - p) ** 2).sum(1)).min()
        d = np.linalg.norm(proj - p)
        assert d <= d_sample + 1e-9 and d >= d_sample - step
        # on the triangle: barycentric coordinates in [0, 1]
        M = np.stack([b - a, c - a], 1)
        uv, *_ = np.linalg.lstsq(M, proj - a, rcond=None)
        assert np.allclose(M @ uv, proj - a, atol=1e-9)
        assert uv.min() >= -1e-9 and uv.sum() <= 1 + 1e-9


def test_max_distance_and_ties():
    """Nothing within max_dist: no projection.  Two copies of one triangle: the later
    one wins the tie (dist <= best), which changes nothing visible."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    body = scenes.trimesh(tri, np.array([[0, 1, 2]]), (0.0, 0.0, 0.0)).body
    proj, _ = O.project_trimesh(tri, np.array([[0, 1, 2]]), body, np.array([0.2, 0.2, 0.3]), 0.29, 64)
    assert proj is None
    proj, inside = O.project_trimesh(tri, np.array([[0, 1, 2]]), body, np.array([0.2, 0.2, 0.3]), 0.31, 64)
    assert np.allclose(proj, [0.2, 0.2, 0.0]) and not inside
    proj, inside = O.project_trimesh(tri, np.array([[0, 1, 2], [0, 1, 2]]), body, np.array([0.2, 0.2, -0.3]), 1.0, 64)
    assert np.allclose(proj, [0.2, 0.2, 0.0]) and inside


@pytest.mark.parametrize("bits", [32, 64])
def test_closed_box_mesh_matches_the_box(bits):
    """A posed box mesh (12 triangles, outward winding): for points within reach of its
    surface, inside and outside, near faces, edges and corners, the distance to the
    projection is the distance to the box's boundary and the inside flag (vertex, edge
    and face pseudo-normals) is the box's inside test."""
    he = np.array([0.3, 0.2, 0.25])
    v, t = scenes.box_mesh(he)
    q = RNG.normal(size=4)
    q /= np.linalg.norm(q)
    q = q.astype(np.float32)
    c = np.array([0.1, -0.2, 0.3])
    mesh = scenes.trimesh(v, t, c, rotation=q)
    x, y, z, w = q.astype(np.float64)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    he32 = np.abs(v).max(0).astype(np.float64)
    tol = 1e-5 if bits == 32 else 1e-6   # the f32 quaternion is unit to ~1e-7
    n_in = n_out = 0
    for _ in range(400):
        local = RNG.uniform(-1.3, 1.3, size=3) * he32
        p = R @ local + c
        proj, inside = O.project_trimesh(v, t, mesh.body, p, 10.0, bits)
        d, _ = box_boundary_closest(he32, local)
        assert np.isclose(np.linalg.norm(proj - p), d, atol=tol)
        solid = bool(np.all(np.abs(local) <= he32))
        if d > 1e-6:
            assert inside == solid, (local, d)
            n_in += solid
            n_out += not solid
    assert n_in > 50 and n_out > 50


def test_open_ramp_sides():
    """An open ramp mesh wound with upward normals: points above it are outside, below
    it inside, whichever feature (face, inner edge, vertex) is nearest."""
    v, t = scenes.ramp_mesh(0.5, 0.5, 0.2, nx=3, nz=3)
    mesh = scenes.trimesh(v, t, (0.0, 0.0, 0.0))
    for _ in range(300):
        x, z = RNG.uniform(-0.45, 0.45, size=2)
        above = RNG.uniform(0.01, 0.1) * (1 if RNG.random() < 0.5 else -1)
        p = np.array([x, 0.2 * x / 0.5 + above, z])
        proj, inside = O.project_trimesh(v, t, mesh.body, p, 1.0, 64)
        assert proj is not None and inside == (above < 0)


def _box_scenes(h=1 / 64):
    """Sand falling onto (and into) a turned slab: once as a cuboid collider, once as the
    slab's mesh (both separate with friction, the same body).  The slab is 2.56 h thick:
    every node inside it is within 1.5 h of a face, where the mesh query looks (deeper
    nodes of a thicker solid see no triangle, while a cuboid reports every inside node)."""
    rng = np.random.Generator(np.random.PCG64(5))
    n = 6000
    pos = np.array([0.5, 0.45, 0.5]) + rng.uniform(-0.12, 0.12, size=(n, 3))
    parts = scenes.particles(pos, h / 4, 2200.0, scenes.model_sand(3.5e5, 0.3), velocity=np.tile([0.3, -2.0, 0.1], (n, 1)))
    q = np.array([np.sin(0.15), 0.0, np.sin(0.1), np.cos(0.15)], np.float64)
    q = (q / np.linalg.norm(q)).astype(np.float32)
    he = (0.2, 0.02, 0.15)
    cub = scenes.cuboid((0.5, 0.36, 0.5), he, rotation=q, bc=scenes.BC_SEPARATE, friction=0.4)
    a = scenes.make_scene(parts, h, 1e-4, colliders=cub)
    b = scenes.make_scene(parts, h, 1e-4)
    v, t = scenes.box_mesh(he)
    b.meshes = [scenes.trimesh(v, t, (0.5, 0.36, 0.5), rotation=q, bc=scenes.BC_SEPARATE, friction=0.4)]
    return a, b


def test_box_mesh_node_pass_equals_the_cuboid():
    """Three substeps (FP64) of sand against the slab as a cuboid and as a mesh: the
    same particles, to rounding; the mesh reached the nodes (the result differs from the
    scene without any collider)."""
    a, b = _box_scenes()
    oa, ob = O.Oracle(a, 64), O.Oracle(b, 64)
    oa.step(3)
    ob.step(3)
    for f in ("x", "v", "C", "F"):
        ga, gb = getattr(oa.s, f), getattr(ob.s, f)
        scale = np.abs(ga).max()
        assert np.abs(ga - gb).max() <= 1e-9 * scale, f
    c = a.copy()
    c.colliders = np.zeros(0, dtype=scenes.CUBOID_DTYPE)
    oc = O.Oracle(c, 64)
    oc.step(3)
    assert np.abs(oc.s.v - ob.s.v).max() > 1e-3


def _wound(v, tris, up):
    """Triangles re-wound so that their normals have a positive component along up."""
    out = []
    for a, b, c in tris:
        n = np.cross(v[b] - v[a], v[c] - v[a])
        out.append((a, b, c) if np.dot(n, up) > 0 else (a, c, b))
    return np.array(out, np.uint32)


@pytest.mark.parametrize("bc_last", [False, True], ids=["valley_as_ca", "valley_as_bc"])
def test_valley_edge_pseudo_normal(bc_last):
    """A sharp V valley (walls at 3:1, faces wound toward the opening): points below the
    valley line are nearest to the valley edge, and only its edge pseudo-normal (the sum
    of both faces' normals) puts all of them inside; either face's normal alone would call
    those with |y|/3 < |x| outside.  The valley edge is the bc edge of one triangle and
    the ca edge of another; the later one wins the distance tie, so the two orders test
    both edge slots."""
    s = 3.0
    v = np.array([[0, 0, -1], [0, 0, 1], [-1, s, -1], [-1, s, 1], [1, s, -1], [1, s, 1]], np.float32)
    # A0 B1 L1=2 L2=3 R1=4 R2=5; the valley A-B is the bc edge of (2, 1, 0) and the ca
    # edge of (1, 4, 0), both wound toward the opening as given
    tris = [(2, 1, 0), (2, 3, 1), (1, 4, 0), (1, 5, 4)]
    wound = _wound(v.astype(np.float64), tris, np.array([0.0, 1.0, 0.0]))
    assert tuple(wound[0]) == (2, 1, 0) and tuple(wound[2]) == (1, 4, 0)
    if bc_last:
        wound = wound[[2, 3, 0, 1]]
    mesh = scenes.trimesh(v, wound, (0.0, 0.0, 0.0))
    checked = 0
    for _ in range(200):
        y = -RNG.uniform(0.02, 0.2)
        x = RNG.uniform(abs(y) / 3 * 1.1, 3 * abs(y) * 0.9) * (1 if RNG.random() < 0.5 else -1)
        p = np.array([x, y, RNG.uniform(-0.8, 0.8)])
        proj, inside = O.project_trimesh(v, wound, mesh.body, p, 1.0, 64)
        assert np.allclose(proj[:2], 0.0, atol=1e-9)        # on the valley line
        assert inside
        checked += 1
    assert checked == 200


def test_pit_vertex_pseudo_normal_is_angle_weighted():
    """A steep square pit (walls at 3:1) whose one wall is a fan of 12 thin triangles and
    the others single triangles: below the pit vertex the vertex's pseudo-normal decides,
    and only the angle-weighted one (each wall weighted by its angle at the vertex,
    however it is split) puts every point whose nearest feature is the vertex inside."""
    s = 3.0
    rim = [np.array(c, np.float64) for c in ([-1, s, -1], [1, s, -1], [1, s, 1], [-1, s, 1])]
    v = [np.zeros(3)]
    tris = []
    k = 12
    # wall 0 (rim 0 -> 1) as a fan of k triangles
    for i in range(k + 1):
        v.append(rim[0] + (rim[1] - rim[0]) * i / k)
    for i in range(k):
        tris.append((0, 1 + i, 2 + i))
    base = len(v)
    v += [rim[1], rim[2], rim[3], rim[0]]
    for i in range(3):
        tris.append((0, base + i, base + i + 1))
    v = np.array(v, np.float32)
    wound = _wound(v.astype(np.float64), tris, np.array([0.0, 1.0, 0.0]))
    mesh = scenes.trimesh(v, wound, (0.0, 0.0, 0.0))
    n_vertex = 0
    for _ in range(400):
        d = RNG.normal(size=3)
        d[1] = -abs(d[1]) * 0.5
        p = d / np.linalg.norm(d) * RNG.uniform(0.02, 0.2)
        proj, inside = O.project_trimesh(v, wound, mesh.body, p, 1.0, 64)
        if np.linalg.norm(proj) < 1e-9:                      # the pit vertex is nearest
            n_vertex += 1
            assert inside, p
    assert n_vertex > 50
