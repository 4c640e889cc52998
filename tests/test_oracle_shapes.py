"""Pins of the oracle's analytic collider shapes (ball, capsule, cone, cylinder) against
geometry that does not use its formulas: the Euclidean distance from a point to the
shape's surface by closed forms of the solid (sphere, finite cylinder) or by dense
sampling of the surface's meridian (capsule axis, cone), the surface equations the
projection must satisfy, and the inside test of the solid.  Shape::project_point_on_boundary,
src_rbd_shaders/shapes/shape.rs:378-399; grid_update_collide.rs:30-99 for the node
pass and grid_update_cdf.rs:18-60 for the contact field."""
import numpy as np
import pytest

import oracle as O
from nexus_amd import scenes

RNG = np.random.default_rng(4321)


def _pose():
    q = RNG.normal(size=4)
    q /= np.linalg.norm(q)
    t = RNG.normal(size=3)
    x, y, z, w = q.astype(np.float32).astype(np.float64)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return q, t, R


def _project(col, R, local):
    """World projection of the body-frame point local, returned in the body frame."""
    t = col["translation"][0].astype(np.float64)
    p = R @ local + t
    proj, inside = O.project_cuboid(col, p, 64)
    return R.T @ (proj - t), inside


def _segment_distance(a, b, p, samples=200001):
    """Distance from p to the segment a-b by dense sampling (2D or 3D)."""
    s = np.linspace(0.0, 1.0, samples)[:, None]
    pts = a[None, :] + s * (b - a)[None, :]
    return float(np.sqrt(((pts - p[None, :]) ** 2).sum(1)).min())


def test_ball_projection():
    """The projection is the sphere point on the ray from the centre: |proj| = r along
    local; inside iff |local| <= r; at the centre the direction is undefined (NaN), as
    parry3d's Ball::project_local_point divides by |local|."""
    for _ in range(200):
        q, t, R = _pose()
        r = RNG.uniform(0.1, 1.0)
        col = scenes.ball(t, r, rotation=q)
        local = RNG.uniform(-1.5, 1.5, size=3) * r
        lp, inside = _project(col, R, local)
        r32 = float(np.float32(r))
        assert inside == (np.linalg.norm(local) <= r32)
        assert np.allclose(lp, local * (r32 / np.linalg.norm(local)), atol=1e-6)
    col = scenes.ball((0.5, 0.5, 0.5), 0.2)
    proj, inside = O.project_cuboid(col, np.array([0.5, 0.5, 0.5]), 64)
    assert inside and np.isnan(proj).all()


def test_capsule_projection():
    """|proj - p| = |d - r| with d the sampled distance from p to the axis segment, the
    projection lies at distance r from the segment, inside iff d <= r; on the axis the
    projection leaves perpendicular to it at distance r."""
    for _ in range(150):
        q, t, R = _pose()
        a, b = RNG.uniform(-0.5, 0.5, size=3), RNG.uniform(-0.5, 0.5, size=3)
        r = RNG.uniform(0.05, 0.4)
        col = scenes.capsule(t, a, b, r, rotation=q)
        a32, b32 = col["dims"][0, 0:3].astype(np.float64), col["dims"][0, 3:6].astype(np.float64)
        r32 = float(col["dims"][0, 6])
        lo, hi = np.minimum(a32, b32) - 1.5 * r32, np.maximum(a32, b32) + 1.5 * r32
        local = RNG.uniform(lo, hi)
        lp, inside = _project(col, R, local)
        d = _segment_distance(a32, b32, local)
        tol = 2e-5 * max(1.0, np.linalg.norm(b32 - a32))
        assert np.isclose(np.linalg.norm(lp - local), abs(d - r32), atol=tol)
        assert np.isclose(_segment_distance(a32, b32, lp), r32, atol=tol)
        if abs(d - r32) > tol:
            assert inside == (d <= r32)
    # a point on the axis
    q, t, R = _pose()
    a, b = np.array([0.0, -0.3, 0.1]), np.array([0.2, 0.4, -0.1])
    col = scenes.capsule(t, a, b, 0.15, rotation=q)
    mid = 0.5 * (col["dims"][0, 0:3].astype(np.float64) + col["dims"][0, 3:6].astype(np.float64))
    lp, inside = _project(col, R, mid)
    ax = b - a
    assert inside
    assert np.isclose(np.linalg.norm(lp - mid), float(np.float32(0.15)), atol=1e-6)
    assert abs(np.dot(lp - mid, ax)) < 1e-6


def _cylinder_distance(hh, r, p):
    """Euclidean distance from p to the surface of the solid cylinder |y| <= hh, rho <= r."""
    rho = np.hypot(p[0], p[2])
    dx, dy = rho - r, abs(p[1]) - hh
    if dx <= 0 and dy <= 0:
        return -min(-dx, -dy)
    return float(np.hypot(max(dx, 0.0), max(dy, 0.0)))


def test_cylinder_projection():
    """|proj - p| equals the distance to the solid cylinder's surface (outside: the
    distance to the solid; inside: the nearest of the caps and the side), the projection
    lies on a cap or on the side, inside iff |y| <= hh and rho <= r."""
    for _ in range(300):
        q, t, R = _pose()
        hh, r = RNG.uniform(0.1, 1.0), RNG.uniform(0.1, 1.0)
        col = scenes.cylinder(t, hh, r, rotation=q)
        hh32, r32 = float(col["dims"][0, 0]), float(col["dims"][0, 1])
        local = RNG.uniform(-1.5, 1.5, size=3) * np.array([r32, hh32, r32])
        lp, inside = _project(col, R, local)
        sd = _cylinder_distance(hh32, r32, local)
        assert inside == (sd <= 0)
        assert np.isclose(np.linalg.norm(lp - local), abs(sd), atol=1e-6)
        rho = np.hypot(lp[0], lp[2])
        on_cap = np.isclose(abs(lp[1]), hh32, atol=1e-6) and rho <= r32 + 1e-6
        on_side = np.isclose(rho, r32, atol=1e-6) and abs(lp[1]) <= hh32 + 1e-6
        assert on_cap or on_side
    # equal distances to the top and to the side: the side wins (cylinder.rs:111-120)
    col = scenes.cylinder((0.0, 0.0, 0.0), 0.5, 0.5)
    proj, inside = O.project_cuboid(col, np.array([0.25, 0.25, 0.0]), 64)
    assert inside and np.allclose(proj, [0.5, 0.25, 0.0])


def test_cone_projection():
    """The cone is a solid of revolution: the distance from p to its surface is the 2D
    distance, in p's meridian half-plane (rho, y), to the base segment (0..r at y = -hh)
    or to the side segment (apex (0, hh) to rim (r, -hh)), both sampled densely.  The
    projection lies on the base disc or on the side; inside iff -hh <= y <= hh and
    rho <= r (hh - y) / (2 hh)."""
    for _ in range(150):
        q, t, R = _pose()
        hh, r = RNG.uniform(0.1, 1.0), RNG.uniform(0.1, 1.0)
        col = scenes.cone(t, hh, r, rotation=q)
        hh32, r32 = float(col["dims"][0, 0]), float(col["dims"][0, 1])
        local = RNG.uniform(-1.5, 1.5, size=3) * np.array([r32, hh32, r32])
        lp, inside = _project(col, R, local)
        m = np.array([np.hypot(local[0], local[2]), local[1]])
        d = min(_segment_distance(np.array([0.0, -hh32]), np.array([r32, -hh32]), m),
                _segment_distance(np.array([0.0, hh32]), np.array([r32, -hh32]), m))
        tol = 2e-5 * max(r32, hh32)
        assert np.isclose(np.linalg.norm(lp - local), d, atol=tol)
        solid = -hh32 <= local[1] <= hh32 and m[0] <= r32 * (hh32 - local[1]) / (2 * hh32)
        if d > tol:
            assert inside == solid
        rho = np.hypot(lp[0], lp[2])
        on_base = np.isclose(lp[1], -hh32, atol=1e-6) and rho <= r32 + 1e-6
        on_side = np.isclose(rho, r32 * (hh32 - lp[1]) / (2 * hh32), atol=1e-6) and -hh32 - 1e-6 <= lp[1] <= hh32 + 1e-6
        assert on_base or on_side


def test_node_pass_on_a_ball():
    """The node pass against a ball: within the margin h of the sphere a slip condition
    removes the radial velocity (n = the outward radial direction, boundary_condition.rs:
    77-114); inside it a stick condition takes the body's velocity, here its angular
    velocity at the node; beyond 1.5 h the node is untouched."""
    h, dt = 1.0 / 64, float(np.float32(1e-4))
    sim = np.array([0, -9.81, 0, dt, h], np.float32)
    c, r = np.array([0.5, 0.5, 0.5]), 0.2
    n = np.array([1.0, 2.0, -2.0]) / 3.0
    vel = np.array([0.3, -0.4, 0.5])
    slip = scenes.ball(c, r, bc=scenes.BC_SLIP)
    node = c + n * (float(np.float32(r)) + 0.5 * h)
    out = O.grid_node_bc(slip, sim, node, vel, 64)
    assert np.allclose(out, vel - n * np.dot(vel, n), atol=1e-6)
    w = np.array([0.0, 0.0, 3.0])
    stick = scenes.ball(c, r, bc=scenes.BC_STICK, angvel=w)
    node = c + n * 0.1
    out = O.grid_node_bc(stick, sim, node, vel, 64)
    assert np.allclose(out, np.cross(w, node - c), atol=1e-6)
    far = c + n * (float(np.float32(r)) + 2.0 * h)
    assert np.allclose(O.grid_node_bc(slip, sim, far, vel, 64), vel)


def test_contact_field_of_shapes():
    """The contact field's node record with every shape kind in one list: the distance to
    the nearest surface within reach (closed forms for the ball and the cylinder), the
    affinity bit of each shape in reach with its side, and the nearest shape's id."""
    h = 1.0 / 64
    cols = scenes.as_colliders(scenes.ball((0.3, 0.5, 0.5), 0.1),
                               scenes.cylinder((0.7, 0.5, 0.5), 0.1, 0.1),
                               scenes.cuboid((0.5, 0.1, 0.5), (0.4, 0.05, 0.4)),
                               scenes.capsule((0.5, 0.8, 0.5), (-0.1, 0.0, 0.0), (0.1, 0.0, 0.0), 0.05),
                               scenes.cone((0.5, 0.5, 0.8), 0.1, 0.1))
    r32 = float(np.float32(0.1))
    # outside the ball within reach: ball only
    p = np.array([0.3 + r32 + 0.7 * h, 0.5, 0.5])
    d, aff, closest = O.node_cdf(cols, h, p, 64)
    assert closest == 0 and np.isclose(d, 0.7 * h, atol=1e-7) and aff == 0b1
    # inside the cylinder, nearer its side than its caps: cylinder, inside bit set
    p = np.array([0.7 + r32 - 0.3 * h, 0.5, 0.5])
    d, aff, closest = O.node_cdf(cols, h, p, 64)
    assert closest == 1 and np.isclose(d, 0.3 * h, atol=1e-7) and aff == (0b10 | (0b10 << 16))
    # nowhere near any shape
    d, aff, closest = O.node_cdf(cols, h, np.array([0.05, 0.9, 0.05]), 64)
    assert aff == 0 and closest == 0xFFFFFFFF


@pytest.mark.parametrize("bits", [32, 64])
def test_shape_records_keep_the_cuboid_path(bits):
    """A cuboid given as a shape record projects exactly as the cuboid record."""
    for _ in range(50):
        q, t, R = _pose()
        he = RNG.uniform(0.1, 1.0, size=3)
        a = scenes.cuboid(t, he, rotation=q)
        b = scenes.as_colliders(a)
        p = t + RNG.uniform(-1.5, 1.5, size=3)
        pa, ia = O.project_cuboid(a, p, bits)
        pb, ib = O.project_cuboid(b, p, bits)
        assert ia == ib and np.array_equal(pa, pb)
