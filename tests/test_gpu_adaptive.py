"""Adaptive substeps (nxm_scene_set_adaptive_dt): each substep's length is min(dt, the
CFL bound of the state at its start), chosen on the device.

- An elastic block thrown at 3 m/s onto a floor with dt = 2e-3, three times its CFL
  bound: without adaptive substeps it blows up; with them every substep takes the bound,
  the lengths match the oracle's adaptive schedule (FP32 mirror) and the state matches
  the oracle stepped with the same lengths (x, v, C, F to the parity bar); the clock holds
  their sum.
- Adaptive scenes of different stiffness stepped in one batch with graph replay take
  their own lengths, bit for bit as when stepped alone (no host round trip in a batch).
"""
import numpy as np
import pytest

from nexus_amd import mpm, scenes

import oracle as O
from parity import assert_field_close

pytestmark = pytest.mark.gpu


def _thrown_block(n_side=16, young=1e5, dt=2e-3, seed=5, center=(0.5, 0.2, 0.5)):
    h = 1 / 64
    rng = np.random.Generator(np.random.PCG64(seed))
    s = h / 2
    idx = np.stack(np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = np.asarray(center) - n_side * s / 2 + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, idx.shape)
    vel = np.tile([0.3, -3.0, 0.0], (pos.shape[0], 1))
    parts = scenes.particles(pos, s / 2, 1000.0, scenes.model_neo_hookean(young, 0.3), velocity=vel)
    floor = scenes.cuboid((0.5, 0.05, 0.5), (0.45, 0.08, 0.45), bc=scenes.BC_SLIP)
    return scenes.make_scene(parts, h, dt, colliders=floor)


def test_without_adaptive_substeps_the_block_blows_up(pipeline):
    sc = _thrown_block()
    st = mpm.MpmState.from_scene(pipeline, sc)
    for _ in range(40):
        pipeline.step(st)
    p = st.read_particles()
    st.close()
    on = p["enabled"] != 0
    vmax = np.abs(p["v"][on]).max() if on.any() else np.inf
    assert (~on).any() or not np.isfinite(vmax) or vmax > 30.0, vmax


def test_adaptive_substeps_follow_the_bound_and_match_the_oracle(pipeline):
    sc = _thrown_block()
    steps = 40
    st = mpm.MpmState.from_scene(pipeline, sc)
    st.set_adaptive_dt(True)
    o32, o64 = O.Oracle(sc, 32), O.Oracle(sc, 64)
    lengths = []
    for k in range(steps):
        pipeline.step(st)
        t, dt, n = st.read_clock()
        assert n == k + 1
        # the mirror's bound of its own state (the oracle's adaptive schedule)
        _, secs = O.timestep_bound(o32.s)
        expect = float(min(np.float32(sc.dt), np.float32(secs)))
        assert abs(dt - expect) <= 1e-4 * expect, (k, dt, expect)
        assert dt < sc.dt
        for o in (o32, o64):                    # both oracles take the GPU's length
            o.s.sim[3] = np.float32(dt)
            o.step(1)
        lengths.append(dt)
    t, dt, n = st.read_clock()
    assert t == sum(float(np.float32(x)) for x in lengths)
    p = st.read_particles()
    assert np.array_equal(p["enabled"], o32.s.enabled) and p["enabled"].all()
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    st.close()


def test_adaptive_scenes_in_a_batch_take_their_own_lengths(pipeline):
    mix = [_thrown_block(10, 1e5, seed=1), _thrown_block(12, 4e5, seed=2, center=(0.3, 0.25, 0.5)),
           scenes.elastic_cube(10)]
    alone = []
    for sc in mix:
        st = mpm.MpmState.from_scene(pipeline, sc)
        st.set_adaptive_dt(True)
        for _ in range(20):
            pipeline.step(st)
        alone.append((st.read_particles(), st.read_clock()))
        st.close()
    batch = mpm.Batch(pipeline)
    states = [mpm.MpmState.from_scene(pipeline, sc) for sc in mix]
    for s in states:
        s.set_adaptive_dt(True)
        batch.add(s)
    batch.run(4)
    batch.set_graph(True)
    batch.run(16)
    batch.synchronize()
    clocks = [s.read_clock() for s in states]
    assert clocks[0][1] != clocks[1][1]       # the stiffer block takes shorter substeps
    for s, (ref, ref_clock), clock in zip(states, alone, clocks):
        assert clock == ref_clock
        got = s.read_particles()
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(got[f], ref[f]), f
    batch.close()
    for s in states:
        s.close()
