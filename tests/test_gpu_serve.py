"""Served scenes evolve exactly as they do alone: requests join and leave one batch at
different substeps (reproducible step clock), each is read back after its horizon, and
its particle state equals the same scene stepped alone for the same number of substeps,
bit for bit."""
import numpy as np
import pytest

from nexus_amd import mpm, scenes, serve

pytestmark = pytest.mark.gpu


def _mix():
    rng = lambda i: np.random.Generator(np.random.PCG64([0x5E4E, i]))
    return [
        (0, scenes.material_box(rng(0), 3000, 0, 1 / 64), 6),
        (0, scenes.material_box(rng(1), 5000, 1, 1 / 96), 9),
        (2, scenes.material_box(rng(2), 4000, 2, 1 / 128), 4),
        (3, scenes.elastic_cube(10), 7),
        (12, scenes.material_box(rng(4), 2000, 1, 1 / 80), 3),
    ]


def test_served_equals_alone(pipeline):
    mix = _mix()
    alone = []
    for _, sc, horizon in mix:
        st = mpm.MpmState.from_scene(pipeline, sc)
        for _ in range(horizon):
            pipeline.step(st)
        alone.append(st.read_particles())
        st.close()
    reqs = [serve.Request(a, sc, h, rid=i, keep="particles") for i, (a, sc, h) in enumerate(mix)]
    srv = serve.Server(pipeline)
    rep = srv.run(reqs, clock="steps")
    srv.close()
    assert rep.completed == len(mix) and rep.rejected == 0
    assert rep.peak_active >= 3
    for r, ref in zip(reqs, alone):
        assert r.steps == r.horizon
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(r.result[f], ref[f]), (r.rid, f)


def test_served_under_tight_budget(pipeline):
    mix = _mix()[:3]
    reqs = [serve.Request(0, sc, h, rid=i) for i, (_, sc, h) in enumerate(mix)]
    probe = mpm.MpmState.from_scene(pipeline, mix[1][1])
    budget = probe.device_bytes() + 1
    probe.close()
    srv = serve.Server(pipeline, hbm_budget_bytes=budget)
    rep = srv.run(reqs, clock="steps")
    srv.close()
    assert rep.completed == 3 and rep.peak_active <= 2
    for r, (_, sc, h) in zip(reqs, mix):
        st = mpm.MpmState.from_scene(pipeline, sc)
        for _ in range(h):
            pipeline.step(st)
        assert np.array_equal(r.result, st.read_positions()), r.rid
        st.close()


def test_served_on_the_wall_clock_equals_alone(pipeline):
    """Wall-clock arrivals: four ingest threads build and pack the scenes in parallel
    (nxm_scene_stage_upload) while the batch steps; each result still equals the scene
    stepped alone, bit for bit, and the requests are admitted in arrival order."""
    mix = _mix()
    alone = []
    for _, sc, horizon in mix:
        st = mpm.MpmState.from_scene(pipeline, sc)
        for _ in range(horizon):
            pipeline.step(st)
        alone.append(st.read_particles())
        st.close()
    reqs = [serve.Request(0.02 * a, sc, h, rid=i, keep="particles") for i, (a, sc, h) in enumerate(mix)]
    srv = serve.Server(pipeline, ingest_workers=4)
    rep = srv.run(reqs, clock="wall")
    srv.close()
    assert rep.completed == len(mix) and rep.rejected == 0
    admitted = [r.admitted for r in sorted(reqs, key=lambda r: r.arrival)]
    assert admitted == sorted(admitted)
    for r, ref in zip(reqs, alone):
        assert r.steps == r.horizon
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(r.result[f], ref[f]), (r.rid, f)


def test_served_adaptive_scenes_equal_alone(pipeline):
    """With adaptive substeps every served scene takes its own substep lengths on the
    device; each result equals the scene stepped alone with adaptive
    substeps, bit for bit."""
    mix = _mix()
    alone = []
    for _, sc, horizon in mix:
        st = mpm.MpmState.from_scene(pipeline, sc)
        st.set_adaptive_dt(True)
        for _ in range(horizon):
            pipeline.step(st)
        alone.append(st.read_particles())
        st.close()
    reqs = [serve.Request(a, sc, h, rid=i, keep="particles") for i, (a, sc, h) in enumerate(mix)]
    srv = serve.Server(pipeline, adaptive_dt=True)
    rep = srv.run(reqs, clock="steps")
    srv.close()
    assert rep.completed == len(mix)
    for r, ref in zip(reqs, alone):
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(r.result[f], ref[f]), (r.rid, f)
