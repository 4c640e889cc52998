"""GPU render readback (WgPrepReadback, src_mpm/solver/prep_readback.rs) against the
oracle's restatement (oracle/readback.py) computed from the particles read back: every
render mode on a stepped CPIC contact scene with groups, phases and a failed particle,
and the mesh colliders' rigid particles."""
import numpy as np
import pytest

from nexus_amd import mpm, scenes
import oracle as O
import readback as RB

pytestmark = pytest.mark.gpu

H = 0.25


def _scene():
    rng = np.random.default_rng(4)
    s = H / 2
    idx = np.stack(np.meshgrid(*(np.arange(8),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = np.array([4.0, 1.85, 4.0]) - 4 * s + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, idx.shape)
    parts = scenes.particles(pos, s / 2, 1000.0, scenes.model_neo_hookean(1e5, 0.3),
                             velocity=np.tile([0.4, -3.0, 0.1], (pos.shape[0], 1)))
    parts["F"][7] = np.array([1, 0, 0, 0, 1, 0, 0, 0, -1], np.float32)      # inverted: fails in the first substep
    floor = scenes.cuboid((4.0, 0.9, 4.0), (2.5, 0.5, 2.5), rotation=(0.0743, 0.0, 0.0, 0.99724), bc=scenes.BC_SLIP)
    sc = scenes.make_scene(parts, H, 1e-3, colliders=floor, cpic=True)
    sc.group_id = (np.arange(sc.n) % 9).astype(np.uint32)
    sc.phase = rng.random(sc.n).astype(np.float32)
    return sc


@pytest.mark.parametrize("palette", [None, [[0.1, 0.2, 0.3, 1.0], [0.9, 0.8, 0.7, 0.5]]], ids=["default", "given"])
def test_readback_modes_match_the_oracle(pipeline, palette):
    sc = _scene()
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(8):
        pipeline.step(st)
    p = st.read_particles()
    cdf = st.read_cdf()
    assert p["enabled"][7] == 0 and (cdf["affinity"] != 0).any()
    for mode in range(7):
        r = st.prep_readback(mode, palette)
        expect = RB.colors(mode, p["v"], p["F"], p["phase"], p["group"], p["enabled"], cdf["affinity"], cdf["normal"],
                           cdf["signed_distance"], palette, sc.cell_width, sc.dt)
        atol = 5e-3 if mode == 1 else 1e-6                  # volume: 200 x the singular values' rounding
        assert np.allclose(r["color"], expect, rtol=1e-6, atol=atol), mode
        assert np.array_equal(r["position"], p["x"]), mode
        d = RB.deformation(p["F"], p["radius"]).reshape(-1, 4, 4).transpose(0, 2, 1)
        assert np.allclose(r["deformation"], d, rtol=1e-6, atol=0), mode
    st.close()


def test_rigid_readback_matches_the_oracle(pipeline):
    sc = scenes.sand_on_mesh_ramp(2000, seed=31)
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(3):
        pipeline.step(st)
    r = st.prep_readback_rigid()
    mesh = sc.meshes[0]
    pts, _ = O.sample_trimesh(mesh.vertices, mesh.triangles, sc.cell_width)
    body = st.read_colliders()[-1]
    world = RB.quat_rotate(body["rotation"], pts) + body["translation"].astype(np.float64)
    assert r["position"].shape == (len(pts), 3)
    assert np.allclose(r["position"], world, rtol=0, atol=2e-5)
    pal = RB.DEFAULT_GROUP_COLORS
    assert np.array_equal(r["color"], pal[np.arange(len(pts)) % len(pal)])
    assert np.allclose(r["deformation"][:, :3, :3], np.eye(3) * np.float32(sc.cell_width) * np.float32(0.4))
    st.close()
