"""GPU parity of the CPIC path (use_cpic = true) against the oracle.

The contact distance field only exists where its least-squares matrix is regular
(det > 1e-8, g2p_cdf.rs:309), which for the quadratic stencil needs cell widths of a
few tenths (the reference's default cell width is 1, src/state.rs:193): the scenes
here use h = 0.25.  Integer outputs (node and particle affinity words, closest
colliders) are compared exactly; x, v, C, F, the normals and signed distances with
the parity bar of the plain path.
"""
import numpy as np
import pytest

from nexus_amd import mpm, scenes

from parity import assert_field_close, oracle_pair

pytestmark = pytest.mark.gpu

H = 0.25


def block(n_side, center, velocity, model=None, density=1000.0, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    s = H / 2
    idx = np.stack(np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = np.asarray(center) - n_side * s / 2 + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, idx.shape)
    vel = np.tile(np.asarray(velocity, np.float64), (pos.shape[0], 1))
    model = model if model is not None else scenes.model_neo_hookean(1e5, 0.3)
    return scenes.particles(pos, s / 2, density, model, velocity=vel)


def contact_scene(bc=scenes.BC_STICK, friction=0.0, steps_hint=12, seed=7, tilt=0.15, moving=True):
    """A block moving down and sideways onto a tilted, moving floor slab."""
    parts = block(10, (4.0, 2.05, 4.0), (0.6, -3.0, 0.2), seed=seed)
    q = np.array([np.sin(tilt / 2), 0.0, 0.0, np.cos(tilt / 2)], np.float32)
    floor = scenes.cuboid((4.0, 0.9, 4.0), (2.5, 0.5, 2.5), rotation=q, bc=bc, friction=friction,
                          linvel=(0.2, 0.0, 0.0) if moving else (0, 0, 0), angvel=(0.0, 0.3, 0.0) if moving else (0, 0, 0))
    side = scenes.cuboid((5.05, 2.0, 4.0), (0.2, 1.0, 1.0), bc=scenes.BC_SLIP)
    sc = scenes.make_scene(parts, H, 1e-3, colliders=np.concatenate([floor, side]), cpic=True)
    bf = np.ones(sc.n, np.float32)
    bf[::3] = 0.5                                 # per-particle friction factors
    sc.boundary_friction = bf
    return sc


def run(pipeline, sc, steps):
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(steps):
        pipeline.step(st)
    return st


def check_against_oracle(st, sc, steps):
    p = st.read_particles()
    cdf = st.read_cdf()
    o32, o64 = oracle_pair(sc, steps)
    assert np.array_equal(p["enabled"], o32.s.enabled)
    assert np.array_equal(cdf["affinity"], o64.s.affinity)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    assert_field_close("cdf normal", cdf["normal"], o32.s.normal, o64.s.normal)
    assert_field_close("cdf signed distance", cdf["signed_distance"], o32.s.signed_distance, o64.s.signed_distance)
    return p, cdf, o64


def test_cpic_grid_field_matches_oracle(pipeline):
    """After one substep: the node contact field (affinity words and closest colliders
    exactly, distances to the bar) and the incompatible mass and velocity."""
    sc = contact_scene()
    st = mpm.MpmState.from_scene(pipeline, sc)
    st.set_debug_grid(True)
    pipeline.step(st)
    can = st.read_sort_canonical()
    g = st.read_grid(momentum=True)
    gc = st.read_grid_cpic()
    o32, o64 = oracle_pair(sc, 1)
    b = o64.blocks()
    gidx = {tuple(map(int, can["vid"][i])): i for i in range(can["vid"].shape[0])}
    order = np.array([gidx[tuple(map(int, v))] for v in b["vid"]])
    G64, G32 = o64.grid(), o32.grid()
    assert np.array_equal(gc["cdf_aff"][order], G64["cdf_aff"])
    assert np.array_equal(gc["cdf_closest"][order], G64["cdf_closest"])
    reach = G64["cdf_aff"] != 0
    assert reach.any() and (G64["cdf_aff"] >> 16).any()
    assert_field_close("cdf distance", gc["cdf_dist"][order][reach], G32["cdf_dist"][reach], G64["cdf_dist"][reach])
    assert_field_close("mass", g["m"][order], G32["m"], G64["m"])
    assert_field_close("incompatible mass", gc["m_i"][order], G32["m_i"], G64["m_i"])
    assert G64["m_i"].sum() > 0
    has = gc["m_i"][order] > 0
    assert_field_close("incompatible velocity", gc["vel_i"][order][has], G32["vel_i"][has], G64["vel_i"][has])
    st.close()


@pytest.mark.parametrize("bc,friction", [(scenes.BC_STICK, 0.0), (scenes.BC_SLIP, 0.0),
                                         (scenes.BC_SEPARATE, 0.4), (scenes.BC_NON_REFLECTING, 0.0)],
                         ids=["stick", "slip", "separate", "non_reflecting"])
def test_cpic_contact_matches_oracle(pipeline, bc, friction):
    """12 substeps of contact under each boundary condition, moving tilted floor."""
    sc = contact_scene(bc, friction)
    st = run(pipeline, sc, 12)
    p, cdf, o64 = check_against_oracle(st, sc, 12)
    assert (cdf["affinity"] & 1).sum() > 0 and (cdf["affinity"] & 2).sum() > 0   # both colliders reached
    st.close()


def test_cpic_penetration_penalty_matches_oracle(pipeline):
    """A fast block driven into a fixed slab: particles get behind the surface, so the
    reprojection and the penalty impulse act (signed distance < -0.05 h)."""
    parts = block(8, (4.0, 1.6, 4.0), (0.0, -9.0, 0.0), seed=11)
    slab = scenes.cuboid((4.0, 0.9, 4.0), (2.0, 0.5, 2.0), bc=scenes.BC_SLIP)
    sc = scenes.make_scene(parts, H, 1e-3, colliders=slab, cpic=True)
    st = run(pipeline, sc, 30)
    p, cdf, o64 = check_against_oracle(st, sc, 30)
    assert (o64.s.signed_distance < -0.05 * H).any()
    st.close()


def test_cpic_batch_is_bitwise_alone(pipeline):
    """A CPIC scene in a batch with a plain scene gives the same bits as alone, and the
    plain scene inside the CPIC world gives the same bits as in a plain world."""
    sc_c = contact_scene(scenes.BC_SEPARATE, 0.3)
    sc_p = scenes.sand_column(6000, seed=5, lo=(0.4, 0.1, 0.4), hi=(0.5, 0.3, 0.5), n_colliders=8)
    ref_c = run(pipeline, sc_c, 6)
    ref_p = run(pipeline, sc_p, 6)
    a_c, a_p = ref_c.read_particles(), ref_p.read_particles()
    a_cdf = ref_c.read_cdf()
    batch = mpm.Batch(pipeline)
    s1 = mpm.MpmState.from_scene(pipeline, sc_p)
    s2 = mpm.MpmState.from_scene(pipeline, sc_c)
    batch.add(s1)
    batch.add(s2)
    batch.run(6)
    b_c, b_p = s2.read_particles(), s1.read_particles()
    b_cdf = s2.read_cdf()
    for f in ("x", "v", "C", "F", "enabled"):
        assert np.array_equal(a_c[f], b_c[f]), f
        assert np.array_equal(a_p[f], b_p[f]), f
    for f in ("affinity", "normal", "signed_distance"):
        assert np.array_equal(a_cdf[f], b_cdf[f]), f
    batch.close()


def test_cpic_switch_resets_the_contact_field(pipeline):
    """Turning CPIC on for a scene that has been stepped without it starts from a zero
    contact field, like a new scene built from the same state."""
    sc = contact_scene(scenes.BC_STICK, moving=False)
    plain = sc.copy()
    plain.cpic = False
    st = run(pipeline, plain, 3)
    st.set_cpic(True)
    p = st.read_particles()
    rebuilt = scenes.Scene(cell_width=sc.cell_width, dt=sc.dt, gravity=sc.gravity, colliders=sc.colliders, cpic=True,
                           boundary_friction=sc.boundary_friction,
                           **{k: p[k] for k in ("x", "v", "F", "C", "mass", "enabled", "fixed", "tag", "model",
                                                 "damping")},
                           vol0=p["vol0"], radius=p["radius"])
    ref = run(pipeline, rebuilt, 4)
    for _ in range(4):
        pipeline.step(st)
    a, b = st.read_particles(), ref.read_particles()
    for f in ("x", "v", "C", "F"):
        assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(st.read_cdf()["affinity"], ref.read_cdf()["affinity"])
    st.close()
    ref.close()


def test_cpic_sand_column_matches_oracle(pipeline):
    """The benchmark's CPIC workload (the sand column at 64x scale: 16 rotated cuboids,
    Drucker-Prager) at 20,000 particles, 4 substeps."""
    sc = scenes.sand_column_cpic(20000)
    st = run(pipeline, sc, 4)
    p, cdf, o64 = check_against_oracle(st, sc, 4)
    assert (cdf["affinity"] & 0xFFFF).any()
    st.close()
