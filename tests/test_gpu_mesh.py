"""GPU parity of the mesh colliders (CPIC): the library's rigid-particle sampler against
the oracle's, the node contact field a mesh adds (gpu_p2g_cdf, p2g_cdf.rs:60-259), sand
on a mesh ramp, a moving mesh, batching, and the boundary's error cases."""
import math

import numpy as np
import pytest

from nexus_amd import mpm, scenes
import oracle as O

from parity import assert_field_close, oracle_pair

pytestmark = pytest.mark.gpu

H = 0.25


@pytest.mark.parametrize("mesh", ["box", "ramp", "random"])
def test_library_sampler_equals_the_oracles(pipeline, mesh):
    """nxm_sample_trimesh and the oracle's sample_mesh give the same points, bit for
    bit, in the same order."""
    if mesh == "box":
        v, t = scenes.box_mesh((0.7, 0.3, 0.45))
    elif mesh == "ramp":
        v, t = scenes.ramp_mesh(3.0, 2.0, 0.6, nx=6, nz=4)
    else:
        rng = np.random.default_rng(5)
        v = rng.normal(size=(30, 3)).astype(np.float32)
        t = rng.integers(0, 30, size=(40, 3)).astype(np.uint32)
    for h in (H, 0.1):
        a, ta = mpm.sample_trimesh(pipeline, v, t, h)
        b, tb = O.sample_trimesh(v, t, h)
        assert len(a) > 0 and np.array_equal(a, b) and np.array_equal(ta, tb)


def _box_scene(seed=1):
    c, he = np.array([4.013, 3.007, 3.991]), np.array([0.9, 0.6, 0.75])
    rng = np.random.default_rng(seed)
    x = c + (rng.random((6000, 3)) - 0.5) * 2 * (he + 3 * H)
    parts = scenes.particles(x, H / 4, 1000.0, scenes.model_neo_hookean(1e5, 0.3),
                             velocity=np.tile([0.3, -1.0, 0.2], (x.shape[0], 1)))
    q = np.array([math.sin(0.2) * 0.6, math.sin(0.2) * 0.8, 0, math.cos(0.2)], np.float32)
    vv, tt = scenes.box_mesh(he)
    slab = scenes.cuboid((4.0, 1.0, 4.0), (3.0, 0.5, 3.0), bc=scenes.BC_SLIP)
    sc = scenes.make_scene(parts, H, 1e-3, colliders=slab, cpic=True)
    sc.meshes = [scenes.trimesh(vv, tt, c, rotation=q, bc=scenes.BC_SEPARATE, friction=0.3,
                                linvel=(0.2, 0.0, -0.1), angvel=(0.0, 0.4, 0.0))]
    return sc


def test_mesh_grid_field_matches_oracle(pipeline):
    """After one substep: the merged node field of a turned box mesh and a cuboid slab
    (affinity words and closest colliders exactly, distances to the bar)."""
    sc = _box_scene()
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
    mesh = (G64["cdf_aff"] & 2) != 0
    assert mesh.sum() > 500 and ((G64["cdf_aff"] >> 17) & 1)[mesh].any()     # inside nodes too
    assert (G64["cdf_closest"] == 1).sum() > 500
    reach = G64["cdf_aff"] != 0
    assert_field_close("cdf distance", gc["cdf_dist"][order][reach], G32["cdf_dist"][reach], G64["cdf_dist"][reach])
    st.close()


def _check(st, sc, steps):
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


def test_moving_box_mesh_matches_oracle(pipeline):
    """A moving, turning box mesh pushing through a falling block for 12 substeps:
    particles, contact fields and the mesh's pose against the oracle."""
    sc = _box_scene(seed=2)
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(12):
        pipeline.step(st)
    p, cdf, o64 = _check(st, sc, 12)
    assert (cdf["affinity"] & 2).sum() > 100
    cols = st.read_colliders()
    assert cols.shape[0] == 2 and np.all(cols["half_extents"][1] == 0)
    for f in ("translation", "rotation"):
        assert np.allclose(cols[f], o64.s.colliders[f], rtol=0, atol=2e-6), f
    st.close()


def test_sand_on_mesh_ramp_matches_oracle(pipeline):
    """Sand dropped on a ramp mesh standing on a cuboid floor: 100 substeps (the sand
    reaches the ramp near substep 80) against the oracle."""
    sc = scenes.sand_on_mesh_ramp(6000, seed=23)
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(100):
        pipeline.step(st)
    p, cdf, o64 = _check(st, sc, 100)
    assert (cdf["affinity"] & 2).sum() > 100
    st.close()


def test_mesh_scene_batch_is_bitwise_alone(pipeline):
    """A mesh scene batched with a cuboid-only CPIC scene and a plain scene gives the
    same bits as alone."""
    sm = scenes.sand_on_mesh_ramp(4000, seed=24)
    sp = scenes.sand_column(5000, seed=5, lo=(0.4, 0.1, 0.4), hi=(0.5, 0.3, 0.5), n_colliders=8)
    sb = _box_scene(seed=3)
    refs = []
    for sc in (sm, sp, sb):
        st = mpm.MpmState.from_scene(pipeline, sc)
        for _ in range(8):
            pipeline.step(st)
        refs.append((st.read_particles(), st.read_colliders()))
        st.close()
    batch = mpm.Batch(pipeline)
    states = [mpm.MpmState.from_scene(pipeline, sc) for sc in (sp, sm, sb)]
    for s in states:
        batch.add(s)
    batch.run(8)
    for (a, ca), s in zip([refs[1], refs[0], refs[2]], states):
        b, cb = s.read_particles(), s.read_colliders()
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(a[f], b[f]), f
        for f in ("translation", "rotation"):
            assert np.array_equal(ca[f], cb[f]), f
    batch.close()


def test_mesh_collider_boundary_errors(pipeline):
    """Meshes act with CPIC (the contact field) or without it (the node pass), and CPIC
    switches under them; indices and vertices are checked; cuboids and meshes together are
    limited to 16; setting cuboids keeps the meshes."""
    sc = scenes.sand_on_mesh_ramp(500, seed=25)
    mesh = sc.meshes[0]
    plain = sc.copy()
    plain.cpic = False
    plain.meshes = []
    st = mpm.MpmState.from_scene(pipeline, plain)
    st.set_trimesh_colliders([mesh])
    pipeline.step(st)
    st.set_cpic(True)
    pipeline.step(st)
    st.set_cpic(False)
    pipeline.step(st)
    st.set_cpic(True)
    bad = scenes.TriMesh(mesh.vertices, np.array([[0, 1, 999]], np.uint32), mesh.body)
    with pytest.raises(mpm.MpmError):
        st.set_trimesh_colliders([bad])
    nan = mesh.vertices.copy()
    nan[0, 0] = np.nan
    with pytest.raises(mpm.MpmError, match="non-finite"):
        st.set_trimesh_colliders([scenes.TriMesh(nan, mesh.triangles, mesh.body)])
    with pytest.raises(mpm.MpmError):
        st.set_colliders(np.concatenate([sc.colliders] * 16))
    pipeline.step(st)
    st.set_colliders(sc.colliders)              # the mesh stays, with its pose
    assert st.read_colliders().shape[0] == 2
    pipeline.step(st)
    st.set_trimesh_colliders([])
    st.set_cpic(False)
    pipeline.step(st)
    st.close()
