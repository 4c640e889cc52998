"""The oracle's CPIC transfers (use_cpic = true) against closed forms.

- the 4x4 solve of the moving-least-squares fit against numpy;
- node contact distance fields of posed cuboids against the distance to the box computed
  here in the box frame (inside: nearest face, outside: clamp), the 1.5 h reach, the
  closest-collider rule (strictly smaller wins) and the affinity / side bits;
- the particle fit reproduces a plane's signed distance exactly (the fit is linear and
  the stencil's weights reproduce linear fields), with the side of a newly reached
  collider from the weighted distances and a known side kept; below the determinant
  threshold the particle field is zero (g2p_cdf.rs:309-330);
- whole substeps: mass and momentum split between the compatible and incompatible
  fields add up to the particles' totals, and with every collider out of reach the CPIC
  substep equals the plain one bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from nexus_amd import scenes

RNG = np.random.Generator(np.random.PCG64(20260))
NONE = 0xFFFFFFFF


def test_m4_solve_matches_numpy():
    for _ in range(20):
        a = RNG.normal(size=(4, 4))
        m = a @ a.T + 0.1 * np.eye(4)
        b = RNG.normal(size=4)
        det, x = O.m4_solve(m, b)
        assert np.isclose(det, np.linalg.det(m), rtol=1e-10)
        assert np.allclose(x, np.linalg.solve(m, b), rtol=1e-9, atol=1e-12)


def _box_distance(c, p):
    """Distance to the boundary of a posed box and whether p is inside (box frame)."""
    q = np.array(c["rotation"][0], np.float64)
    t = np.array(c["translation"][0], np.float64)
    he = np.array(c["half_extents"][0], np.float64)
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    local = R.T @ (np.asarray(p, np.float64) - t)
    inside = np.all(np.abs(local) <= he)
    if inside:
        return float(np.min(he - np.abs(local))), True
    return float(np.linalg.norm(local - np.clip(local, -he, he))), False


def test_node_cdf_cuboids():
    h = 0.25
    box = scenes.cuboid((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    d, a, c = O.node_cdf(box, h, (0.0, 1.2, 0.0))
    assert np.isclose(d, 0.2) and a == 0x1 and c == 0
    d, a, c = O.node_cdf(box, h, (0.0, 0.9, 0.0))
    assert np.isclose(d, 0.1) and a == 0x10001 and c == 0
    d, a, c = O.node_cdf(box, h, (0.0, 1.5, 0.0))                 # beyond 1.5 h
    assert np.isclose(d, 1e18) and a == 0 and c == NONE
    d, a, c = O.node_cdf(box, h, (1.2, 1.3, 0.0))                 # edge, within reach on each axis
    assert np.isclose(d, np.hypot(0.2, 0.3)) and a == 0x1
    # two colliders: the strictly closer wins, both are recorded
    two = np.concatenate([box, scenes.cuboid((0.0, 2.5, 0.0), (1.0, 1.0, 1.0))])
    d, a, c = O.node_cdf(two, h, (0.0, 1.3, 0.0))
    assert c == 1 and np.isclose(d, 0.2) and a == 0x3
    d, a, c = O.node_cdf(two, h, (0.0, 1.25, 0.0))                # tie: the first keeps it
    assert c == 0 and np.isclose(d, 0.25) and a == 0x3
    # posed boxes against the distance computed in the box frame
    for _ in range(30):
        q = RNG.normal(size=4)
        q /= np.linalg.norm(q)
        col = scenes.cuboid(RNG.uniform(-0.2, 0.2, 3), RNG.uniform(0.3, 0.8, 3), rotation=q)
        p = RNG.uniform(-1.2, 1.2, 3)
        d, a, c = O.node_cdf(col, 0.5, p)
        ref, inside = _box_distance(col, p)
        if a:
            # the pose is stored in f32, so the quaternion is unit only to ~1e-7
            assert np.isclose(d, ref, rtol=1e-6, atol=1e-7)
            assert bool(a & 0x10000) == inside


def _stencil(p, h, cols):
    """Node CDFs of the particle's 27 stencil nodes (NBH order) and refp."""
    base = np.round(np.asarray(p) / h) - 1.0
    refp = base * h - np.asarray(p)
    dist, aff = [], []
    for s in O.NBH_SHIFTS:
        d, a, _ = O.node_cdf(cols, h, (base + s) * h)
        dist.append(d)
        aff.append(a)
    return refp, np.array(dist), np.array(aff, np.uint32)


@pytest.mark.parametrize("bits", [32, 64])
def test_particle_cdf_reconstructs_a_plane(bits):
    """The top face y = 0 of a large box; a particle 0.2 below it.  Nodes inside read
    their distance with the inside side bit, nodes outside without it."""
    h = 1.0
    floor = scenes.cuboid((0.0, -50.0, 0.0), (50.0, 50.0, 50.0))
    p = np.array([0.3, -0.2, 0.1])
    refp, dist, aff = _stencil(p, h, floor)
    assert np.all(aff & 1)
    tol = 1e-12 if bits == 64 else 2e-6
    # it was outside: the nodes inside disagree with its side, the fit goes negative
    a, n, sd = O.particle_cdf(refp, h, dist, aff, prev=0x1, bits=bits)
    assert a == 0x1
    assert abs(sd - (-0.2)) < tol and np.allclose(n, [0, 1, 0], atol=tol)
    # newly reached while inside: the weighted distances are negative, so the side bit
    # is set and the fit is the depth, measured positive into the box
    a, n, sd = O.particle_cdf(refp, h, dist, aff, prev=0x0, bits=bits)
    assert a == 0x10001
    assert abs(sd - 0.2) < tol and np.allclose(n, [0, -1, 0], atol=tol)
    # a known inside side is kept
    a, n, sd = O.particle_cdf(refp, h, dist, aff, prev=0x10001, bits=bits)
    assert a == 0x10001 and abs(sd - 0.2) < tol
    # newly reached from outside: no side bit, positive distance above the face
    q = np.array([0.3, 0.35, 0.1])
    refq, dq, aq = _stencil(q, h, floor)
    a, n, sd = O.particle_cdf(refq, h, dq, aq, prev=0x0, bits=bits)
    assert a == 0x1 and abs(sd - 0.35) < tol and np.allclose(n, [0, 1, 0], atol=tol)


def test_particle_cdf_below_the_determinant_threshold_is_zero():
    """With h = 1/64 the fit's matrix has det ~ (h^2/4)^3 < 1e-8: the reference stores a
    zero CDF, affinity included (g2p_cdf.rs:309-330)."""
    h = 1.0 / 64
    floor = scenes.cuboid((0.0, -50.0, 0.0), (50.0, 50.0, 50.0))
    refp, dist, aff = _stencil(np.array([0.3, -0.2 * h, 0.1]), h, floor)
    assert np.all(aff & 1)
    a, n, sd = O.particle_cdf(refp, h, dist, aff, prev=0x1)
    assert a == 0 and sd == 0.0 and np.all(n == 0)


def _block_scene(h, center, n_side, velocity, colliders, cpic, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    s = h / 2
    idx = np.stack(np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = np.asarray(center) - n_side * s / 2 + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, idx.shape)
    vel = np.tile(np.asarray(velocity, np.float64), (pos.shape[0], 1))
    parts = scenes.particles(pos, s / 2, 1000.0, scenes.model_neo_hookean(1e5, 0.3), velocity=vel)
    return scenes.make_scene(parts, h, 1e-3, colliders=colliders, cpic=cpic)


def test_cpic_split_conserves_mass_and_momentum():
    """Every particle contribution lands in exactly one of the two node fields."""
    h = 0.25
    floor = scenes.cuboid((2.0, 0.5, 2.0), (2.0, 0.5, 2.0))
    sc = _block_scene(h, (2.0, 1.3, 2.0), 8, (0.3, -2.0, 0.1), floor, cpic=True)
    o = O.Oracle(sc, 64)
    o.step(3)
    g = o.grid()
    assert g["m_i"].sum() > 0 and g["m"].sum() > 0              # both fields in use
    s = o.s
    # the grid of the last substep was built from the particles before it: replay it
    o2 = O.Oracle(sc, 64)
    o2.step(2)
    m_tot = o2.s.mass[o2.s.enabled != 0].sum()
    p_tot = (o2.s.mass[:, None] * o2.s.v)[o2.s.enabled != 0].sum(0)
    assert np.isclose(g["m"].sum() + g["m_i"].sum(), m_tot, rtol=1e-12)
    assert np.allclose(g["mv"].sum((0, 1)) + g["mv_i"].sum((0, 1)), p_tot, rtol=1e-10, atol=1e-12)
    assert np.all(np.isfinite(s.x))


def test_cpic_without_colliders_in_reach_equals_plain_substep():
    h = 0.25
    far = scenes.cuboid((40.0, 40.0, 40.0), (1.0, 1.0, 1.0))
    a = O.Oracle(_block_scene(h, (2.0, 2.0, 2.0), 6, (0.5, -1.0, 0.2), far, cpic=True), 32)
    b = O.Oracle(_block_scene(h, (2.0, 2.0, 2.0), 6, (0.5, -1.0, 0.2), far, cpic=False), 32)
    a.step(4)
    b.step(4)
    for f in ("x", "v", "C", "F"):
        assert np.array_equal(getattr(a.s, f), getattr(b.s, f)), f
    assert np.all(a.s.affinity == 0)


def test_collider_pose_integration():
    """Body integration of the massless cuboids at the end of each substep: the
    translation accumulates v dt in f32 (bit-exact against a float32 loop in the FP32
    mirror) and a constant angular velocity about z turns the quaternion by w dt per
    substep: after n substeps it is (0, 0, sin(n w dt / 2), cos(n w dt / 2))."""
    h, dt, n = 1.0 / 64, 1e-3, 50
    parts = scenes.particles(np.array([[0.5, 0.5, 0.5]]), h / 4, 1000.0, scenes.model_neo_hookean(1e5, 0.3))
    cols = np.concatenate([scenes.cuboid((2.0, 2.0, 2.0), (0.1, 0.2, 0.3), linvel=(0.7, -0.2, 0.05)),
                           scenes.cuboid((-2.0, 2.0, 2.0), (0.1, 0.2, 0.3), angvel=(0.0, 0.0, 3.0)),
                           scenes.cuboid((2.0, -2.0, 2.0), (0.1, 0.2, 0.3))])
    sc = scenes.make_scene(parts, h, dt, colliders=cols)
    o = O.Oracle(sc, 32)
    o.step(n)
    t = np.array(cols["translation"][0], np.float32)
    v = np.array(cols["linvel"][0], np.float32)
    for _ in range(n):
        t = (t + v * np.float32(sc.dt)).astype(np.float32)
    assert np.array_equal(o.s.colliders["translation"][0], t)
    ang = n * 3.0 * np.float64(np.float32(sc.dt))
    assert np.allclose(o.s.colliders["rotation"][1], [0, 0, np.sin(ang / 2), np.cos(ang / 2)], atol=2e-6)
    assert np.array_equal(o.s.colliders["translation"][2], cols["translation"][2])
    assert np.allclose(o.s.colliders["rotation"][2], [0, 0, 0, 1])


def test_dynamic_body_integration():
    """A dynamic body coupled two ways: the pose advances with the velocity from before
    the substep's gravity, then gravity is added along the axes with a nonzero inverse
    mass (rigid_impulses.rs:66-140, impulses compiled out).  Bit-exact against a float32
    loop in the FP32 mirror.  A body turning about an off-centre centre of mass keeps
    that point fixed and its translation on a circle around it."""
    h, dt, n = 1.0 / 64, 1e-3, 40
    g = (0.5, -9.81, 0.25)
    parts = scenes.particles(np.array([[0.5, 0.5, 0.5]]), h / 4, 1000.0, scenes.model_neo_hookean(1e5, 0.3))
    lc = (0.1, 0.0, 0.0)
    cols = np.concatenate([scenes.cuboid((2.0, 2.0, 2.0), (0.1, 0.2, 0.3), linvel=(0.7, 0.2, 0.0),
                                         inv_mass=(0.5, 0.5, 0.5)),
                           scenes.cuboid((-2.0, 2.0, 2.0), (0.1, 0.2, 0.3), inv_mass=(0.0, 2.0, 0.0)),
                           scenes.cuboid((2.0, -2.0, 2.0), (0.1, 0.2, 0.3), angvel=(0.0, 0.0, 3.0),
                                         local_com=lc)])
    sc = scenes.make_scene(parts, h, dt, colliders=cols, gravity=g)
    o = O.Oracle(sc, 32)
    o.step(n)
    f32 = np.float32
    dtf, gf = f32(sc.dt), np.array(g, np.float32)
    for i, mask in ((0, np.array([1, 1, 1], np.float32)), (1, np.array([0, 1, 0], np.float32))):
        t = np.array(cols["translation"][i], np.float32)
        v = np.array(cols["linvel"][i], np.float32)
        for _ in range(n):
            t = (t + v * dtf).astype(np.float32)
            v = (v + (gf * mask) * dtf).astype(np.float32)
        assert np.array_equal(o.s.colliders["translation"][i], t), i
        assert np.array_equal(o.s.colliders["linvel"][i], v), i
    # the one-way body keeps its velocity and turns about its centre of mass
    assert np.array_equal(o.s.colliders["linvel"][2], cols["linvel"][2])
    q = o.s.colliders["rotation"][2].astype(np.float64)
    ang = n * 3.0 * np.float64(dtf)
    assert np.allclose(q, [0, 0, np.sin(ang / 2), np.cos(ang / 2)], atol=2e-6)
    com0 = cols["translation"][2].astype(np.float64) + np.array(lc)
    t = o.s.colliders["translation"][2].astype(np.float64)
    # world com = t + R(q) lc stays where it started
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    assert np.allclose(t + R @ np.array(lc), com0, atol=1e-5)
    assert np.isclose(np.linalg.norm(t - com0), 0.1, atol=1e-5)
    assert not np.allclose(t, cols["translation"][2])
