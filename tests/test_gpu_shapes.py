"""GPU parity of the analytic collider shapes (ball, capsule, cone, cylinder, with a
cuboid floor) against the oracle: the node pass (grid velocity after the collider
boundary conditions, node by node), the particles after several substeps, and the
CPIC contact field (affinity words and closest colliders exactly).  Also the shape
records' argument checks."""
import numpy as np
import pytest

from nexus_amd import mpm, scenes

from parity import assert_field_close, oracle_pair

pytestmark = pytest.mark.gpu


def _rot(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    return np.array([*(a * np.sin(angle / 2)), np.cos(angle / 2)], np.float32)


def shapes(scale=1.0, cpic=False):
    """Four shapes in a row over a floor, one boundary condition each, two of them moving."""
    k = scale
    cols = scenes.as_colliders(
        scenes.ball((0.3 * k, 0.35 * k, 0.3 * k), 0.08 * k, bc=scenes.BC_SEPARATE, friction=0.3,
                    linvel=(0.1 * k, 0.0, 0.0)),
        scenes.capsule((0.7 * k, 0.36 * k, 0.3 * k), (-0.08 * k, 0.0, 0.0), (0.08 * k, 0.0, 0.0), 0.05 * k,
                       rotation=_rot((0, 0, 1), 0.4), bc=scenes.BC_SLIP, angvel=(0.0, 1.0, 0.0)),
        scenes.cone((0.3 * k, 0.34 * k, 0.7 * k), 0.08 * k, 0.08 * k, rotation=_rot((1, 0, 0), 0.3),
                    bc=scenes.BC_STICK),
        scenes.cylinder((0.7 * k, 0.35 * k, 0.7 * k), 0.06 * k, 0.07 * k, rotation=_rot((0, 0, 1), 0.7),
                        bc=scenes.BC_SEPARATE, friction=0.5),
        scenes.cuboid((0.5 * k, 0.2 * k, 0.5 * k), (0.45 * k, 0.03 * k, 0.45 * k), bc=scenes.BC_SLIP),
    )
    return cols


def slab_scene(n=20000, seed=31):
    """Sand at about one particle per cell filling the region around the shapes, falling."""
    h = 1 / 64
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = np.array([0.2, 0.245, 0.2]), np.array([0.8, 0.46, 0.8])
    pos = lo + rng.random((n, 3)) * (hi - lo)
    r = 0.5 * (np.prod(hi - lo) / n) ** (1 / 3)
    vel = np.tile([0.2, -1.5, 0.1], (n, 1))
    parts = scenes.particles(pos, r, 2200.0, scenes.model_sand(3.5e5, 0.3, friction_deg=30.0), velocity=vel)
    return scenes.make_scene(parts, h, 1e-4, colliders=shapes())


def test_shape_node_pass_matches_oracle(pipeline):
    """Grid velocity after the collider pass, node by node, one substep."""
    sc = slab_scene()
    st = mpm.MpmState.from_scene(pipeline, sc)
    st.set_debug_grid(True)
    pipeline.step(st)
    can = st.read_sort_canonical()
    grid = st.read_grid(momentum=True)
    o32, o64 = oracle_pair(sc, 1)
    b = o64.blocks()
    gidx = {tuple(map(int, can["vid"][i])): i for i in range(can["vid"].shape[0])}
    order = np.array([gidx[tuple(map(int, v))] for v in b["vid"]])
    g64, g32 = o64.grid(), o32.grid()
    gm = grid["m"][order]
    has = gm > 0
    assert np.all(g64["m"][~has] <= 1e-12 * g64["m"].max())
    assert_field_close("velocity", grid["vel"][order][has], g32["vel"][has], g64["vel"][has])
    # the shapes acted: nodes whose velocity left free fall (the initial velocity plus
    # g dt), around the shapes and the floor
    vy = -1.5 + float(np.float32(-9.81)) * float(np.float32(sc.dt))
    free = np.abs(g64["vel"][..., 1] - vy) < 1e-6
    assert (~free & has).sum() > 1000 and (free & has).sum() > 1000
    st.close()


def test_shape_contact_matches_oracle(pipeline):
    """15 substeps of the slab settling onto the shapes: x, v, C, F and the plastic state."""
    sc = slab_scene(12000, seed=32)
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(15):
        pipeline.step(st)
    p = st.read_particles()
    o32, o64 = oracle_pair(sc, 15)
    assert np.array_equal(p["enabled"], o32.s.enabled)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    words = p["model"].view(np.float32)
    assert_field_close("plastic state", words[:, :3], o32.s.pstate, o64.s.pstate)
    cols = st.read_colliders()
    for f in ("translation", "rotation"):
        assert np.allclose(cols[f], o64.s.colliders[f], rtol=0, atol=2e-6), f
    assert np.all(cols["half_extents"][:4] == 0) and np.array_equal(cols["half_extents"][4], sc.colliders["half_extents"][4])
    st.close()


def test_shape_contact_field_matches_oracle(pipeline):
    """CPIC with the shapes at h = 0.25 (the scene 16x larger): the node contact field
    after one substep (affinity words and closest colliders exactly, distances to the
    bar), then 8 substeps of particles, normals and signed distances."""
    H = 0.25
    k = 16.0
    rng = np.random.Generator(np.random.PCG64(33))
    lo, hi = np.array([0.2, 0.245, 0.2]) * k, np.array([0.8, 0.46, 0.8]) * k
    n = 6000
    pos = lo + rng.random((n, 3)) * (hi - lo)
    r = 0.5 * (np.prod(hi - lo) / n) ** (1 / 3)
    vel = np.tile([0.5, -3.0, 0.2], (n, 1))
    parts = scenes.particles(pos, r, 1000.0, scenes.model_neo_hookean(1e5, 0.3), velocity=vel)
    sc = scenes.make_scene(parts, H, 1e-3, colliders=shapes(k), cpic=True)
    st = mpm.MpmState.from_scene(pipeline, sc)
    st.set_debug_grid(True)
    pipeline.step(st)
    can = st.read_sort_canonical()
    gc = st.read_grid_cpic()
    o32, o64 = oracle_pair(sc, 1)
    b = o64.blocks()
    gidx = {tuple(map(int, can["vid"][i])): i for i in range(can["vid"].shape[0])}
    order = np.array([gidx[tuple(map(int, v))] for v in b["vid"]])
    G64, G32 = o64.grid(), o32.grid()
    assert np.array_equal(gc["cdf_aff"][order], G64["cdf_aff"])
    assert np.array_equal(gc["cdf_closest"][order], G64["cdf_closest"])
    reach = G64["cdf_aff"] != 0
    for i in range(5):
        assert ((G64["cdf_aff"] >> i) & 1).any(), i                 # every shape is in reach somewhere
    assert_field_close("cdf distance", gc["cdf_dist"][order][reach], G32["cdf_dist"][reach], G64["cdf_dist"][reach])
    st.close()
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(8):
        pipeline.step(st)
    p = st.read_particles()
    cdf = st.read_cdf()
    o32, o64 = oracle_pair(sc, 8)
    assert np.array_equal(p["enabled"], o32.s.enabled)
    assert np.array_equal(cdf["affinity"], o64.s.affinity)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    assert_field_close("cdf normal", cdf["normal"], o32.s.normal, o64.s.normal)
    assert_field_close("cdf signed distance", cdf["signed_distance"], o32.s.signed_distance, o64.s.signed_distance)
    st.close()


def test_shape_records_are_checked(pipeline):
    st = mpm.MpmState(pipeline, 1024, 16)
    bad = scenes.ball((0, 0, 0), 0.1)
    bad["shape"] = 5
    with pytest.raises(mpm.MpmError, match="unknown shape kind"):
        st.set_colliders(bad)
    with pytest.raises(mpm.MpmError, match="radius"):
        st.set_colliders(scenes.ball((0, 0, 0), -0.1))
    with pytest.raises(mpm.MpmError, match="finite"):
        st.set_colliders(scenes.capsule((0, 0, 0), (0, np.inf, 0), (0, 1, 0), 0.1))
    with pytest.raises(mpm.MpmError, match="16"):
        st.set_colliders(scenes.as_colliders(*[scenes.ball((0, 0, 0), 0.1)] * 17))
    st.set_colliders(shapes())
    assert st.read_colliders().shape == (5,)
    st.close()
