"""Edge cases of the GPU path: empty scenes, diverged (inverted) particles, every
boundary-condition type in contact, and the error paths (packable domain, grid
capacity, unsupported model, collider count).  Oracle comparisons use the same bar as
the one-substep parity tests."""
import numpy as np
import pytest

from nexus_amd import mpm, scenes

from parity import assert_field_close, oracle_pair

pytestmark = pytest.mark.gpu


def _block(n_side, center, h, model, density, velocity=(0.0, 0.0, 0.0), seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    s = h / 2
    idx = np.stack(np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = np.asarray(center) - n_side * s / 2 + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, idx.shape)
    vel = np.tile(np.asarray(velocity, np.float64), (pos.shape[0], 1))
    return scenes.particles(pos, s / 2, density, model, velocity=vel)


def test_empty_scene_steps(pipeline):
    st = mpm.MpmState(pipeline, 1024, 0)
    st.set_cell_width(1 / 64, 1024)
    st.set_simulation_params((0.0, -9.81, 0.0), 1e-4)
    pipeline.step(st)
    st.synchronize()
    assert len(st) == 0 and st.read_positions().shape == (0, 3)
    st.close()


def test_inverted_particle_is_disabled(pipeline):
    """A particle whose deformation gradient is a reflection (det F < 0) fails the
    particle update's det F > 0 test: it is disabled with zero mass (particle_update.rs:
    184-200), as in the oracle, and contributes nothing afterwards."""
    h = 1 / 64
    parts = _block(8, (0.5, 0.5, 0.5), h, scenes.model_neo_hookean(1e5, 0.3), 1000.0)
    parts["F"][3] = np.array([-1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    sc = scenes.make_scene(parts, h, 1e-4)
    st = mpm.MpmState.from_scene(pipeline, sc)
    pipeline.step(st)
    pipeline.step(st)
    p = st.read_particles()
    o32, o64 = oracle_pair(sc, 2)
    assert p["enabled"][3] == 0 and p["mass"][3] == 0.0
    assert np.array_equal(p["enabled"], o32.s.enabled)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    st.close()


@pytest.mark.parametrize("bc,friction", [(scenes.BC_STICK, 0.0), (scenes.BC_SLIP, 0.0),
                                         (scenes.BC_SEPARATE, 0.4), (scenes.BC_NON_REFLECTING, 0.0)],
                         ids=["stick", "slip", "separate", "non_reflecting"])
def test_boundary_conditions_in_contact(pipeline, bc, friction):
    """A block moving down and sideways onto a tilted floor cuboid (with a moving body
    velocity), for each boundary condition: 6 substeps against the oracle."""
    h = 1 / 64
    parts = _block(10, (0.5, 0.345, 0.5), h, scenes.model_sand(3.5e5, 0.3, friction_deg=30.0), 2200.0,
                   velocity=(0.8, -1.5, 0.3))
    q = np.array([np.sin(0.1), 0.0, 0.0, np.cos(0.1)], np.float32)     # 0.2 rad about x
    floor = scenes.cuboid((0.5, 0.25, 0.5), (0.3, 0.05, 0.3), rotation=q, bc=bc, friction=friction,
                          linvel=(0.1, 0.0, 0.0), angvel=(0.0, 0.5, 0.0))
    sc = scenes.make_scene(parts, h, 1e-4, colliders=floor)
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(6):
        pipeline.step(st)
    p = st.read_particles()
    o32, o64 = oracle_pair(sc, 6)
    assert np.array_equal(p["enabled"], o32.s.enabled)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    st.close()


def test_snow_compaction_matches_oracle(pipeline):
    """Snow (models/snow.rs, tag 5) thrown onto a stick floor: within a few substeps the
    singular values of F leave the yield box [1 - 2.5e-2, 1 + 7.5e-3], so the return map
    clamps them and the plastic determinant (model word 0) and the hardening change.
    10 substeps against the oracle: x, v, C, F and the plastic determinant."""
    h = 1 / 64
    parts = _block(10, (0.5, 0.3, 0.5), h, scenes.model_snow(1.4e5, 0.2), 400.0, velocity=(0.5, -2.0, 0.2))
    floor = scenes.cuboid((0.5, 0.2, 0.5), (0.4, 0.05, 0.4), bc=scenes.BC_STICK)
    sc = scenes.make_scene(parts, h, 1e-4, colliders=floor)
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(10):
        pipeline.step(st)
    p = st.read_particles()
    o32, o64 = oracle_pair(sc, 10)
    assert np.array_equal(p["enabled"], o32.s.enabled)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    pdet = p["model"][:, 0].view(np.float32)
    assert np.abs(pdet - 1.0).max() > 1e-3                      # plasticity acted
    assert_field_close("plastic det", pdet, o32.s.pstate[:, 0], o64.s.pstate[:, 0])
    assert np.array_equal(p["model"][:, 1:], sc.model[:, 1:])    # parameters come back unchanged
    st.close()


def test_all_models_in_one_scene(pipeline):
    """One block of each model tag (corotated linear, neo-Hookean, Drucker-Prager over the
    linear and the neo-Hookean backbone with the reference's default hardening law, fluid,
    snow) thrown onto a floor and onto each other: 40 substeps against the oracle, the
    per-particle plastic state included."""
    h = 1 / 64
    # lower row on the floor: corotated, sand (linear backbone), snow; upper row falling onto
    # it: neo-Hookean, sand (neo-Hookean backbone), fluid
    models = [(scenes.model_linear(2e5, 0.3), 1000.0), (scenes.model_sand(3.5e5, 0.3), 2200.0),
              (scenes.model_snow(1.4e5, 0.2), 400.0), (scenes.model_neo_hookean(2e5, 0.3), 1000.0),
              (scenes.model_sand(3.5e5, 0.3, neo_hookean=True), 2200.0), (scenes.model_fluid(1e4), 1000.0)]
    parts = []
    for i, (model, rho) in enumerate(models):
        c = (0.35 + 0.1 * (i % 3), 0.28 + 0.06 * (i // 3), 0.5)
        vel = (0.8 * (1 - 2 * (i % 2)), -4.0, 0.2)
        parts.append(_block(6, c, h, model, rho, velocity=vel, seed=20 + i))
    # The floor's top face is kept off the grid planes: a node exactly on a face has a zero
    # distance and no normal (grid_update_collide.rs:84-87), where FP32 and FP64 geometry
    # disagree for the separate condition.
    floor = scenes.cuboid((0.5, 0.2013, 0.5), (0.4, 0.05, 0.4), bc=scenes.BC_SEPARATE, friction=0.5)
    sc = scenes.make_scene(scenes.concat(parts), h, 1e-4, colliders=floor)
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(40):
        pipeline.step(st)
    p = st.read_particles()
    o32, o64 = oracle_pair(sc, 40)
    assert np.array_equal(p["enabled"], o32.s.enabled)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    words = p["model"].view(np.float32)
    sand = np.isin(p["tag"], (2, 3))
    snow = p["tag"] == 5
    assert_field_close("sand plastic state", words[sand, :3], o32.s.pstate[sand], o64.s.pstate[sand])
    assert_field_close("snow plastic det", words[snow, 0], o32.s.pstate[snow, 0], o64.s.pstate[snow, 0])
    assert (np.abs(o64.s.pstate[sand, 0] - 1) > 1e-4).any() and (np.abs(o64.s.pstate[snow, 0] - 1) > 1e-4).any()
    st.close()


def test_domain_error_is_reported(pipeline):
    """A particle outside the packable block range (|block x| > 1023) is reported as
    NXM_ERR_DOMAIN at the next synchronizing call; the step still runs for the others."""
    h = 1 / 64
    parts = _block(6, (0.5, 0.5, 0.5), h, scenes.model_neo_hookean(1e5, 0.3), 1000.0)
    parts["x"][0] = (1.0e3, 0.5, 0.5)
    sc = scenes.make_scene(parts, h, 1e-4)
    st = mpm.MpmState.from_scene(pipeline, sc, grid_size=4096)
    pipeline.step(st)
    with pytest.raises(mpm.MpmError) as e:
        st.synchronize()
    assert e.value.code == 4
    st.close()


def test_grid_capacity_error_is_reported(pipeline):
    """More active blocks than the grid capacity: NXM_ERR_GRID_CAPACITY, no fault."""
    h = 1 / 64
    rng = np.random.Generator(np.random.PCG64(3))
    pos = rng.uniform(0.1, 0.9, size=(4000, 3))
    sc = scenes.make_scene(scenes.particles(pos, h / 4, 1000.0, scenes.model_fluid(1e4)), h, 1e-4)
    st = mpm.MpmState.from_scene(pipeline, sc, grid_size=64)
    pipeline.step(st)
    with pytest.raises(mpm.MpmError) as e:
        st.synchronize()
    assert e.value.code == 3
    st.close()


def test_rejected_inputs(pipeline):
    h = 1 / 64
    st = mpm.MpmState(pipeline, 1024, 16)
    st.set_cell_width(h, 1024)
    unknown = _block(2, (0.5, 0.5, 0.5), h, (6, np.zeros(24, np.uint32)), 400.0)
    with pytest.raises(mpm.MpmError) as e:
        st.add_particles(scenes.make_scene(unknown, h, 1e-4))
    assert e.value.code == 6
    cols = np.concatenate([scenes.cuboid((0.1 * i, 0, 0), (0.01, 0.01, 0.01)) for i in range(17)])
    with pytest.raises(mpm.MpmError) as e:
        st.set_colliders(cols)
    assert e.value.code == 5
    st.close()
