"""The batched and served workloads at their full scale, checked against scenes stepped
alone and against the oracle.

- Batched: the 256 scenes of the batched workload (N log-uniform in [2e3, 2e5], 9.67M
  particles, neo-Hookean / Drucker-Prager / fluid, h = 1/64..1/128) stepped as one batch
  with the launch-shape tuner and then graph replay.  The smallest and largest scene and
  one scene of each material equal the same scenes stepped alone, bit for bit; the
  smallest also matches the oracle (FP64, with the FP32 mirror's error as the bar).
- Served: one second of the served workload's Poisson stream (200 requests/s, N
  log-uniform in [1e3, 5e5], horizons 50-500 substeps) on the step clock (arrival second
  t -> batch substep 1000 t).  Every request completes; the largest, the longest-lived
  and three more equal the scene stepped alone for its horizon, bit for bit, and the
  smallest short-horizon request matches the oracle.
"""
import numpy as np
import pytest

from nexus_amd import mpm, scenes, serve

from parity import assert_field_close, oracle_pair

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "C", "F", "enabled")


def _alone(pipeline, scene, substeps):
    st = mpm.MpmState.from_scene(pipeline, scene)
    pipeline.run(st, substeps)
    out = st.read_particles()
    st.close()
    return out


def _check_oracle(got, scene, substeps):
    o32, o64 = oracle_pair(scene, substeps)
    assert np.array_equal(got["enabled"], o32.s.enabled)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, got[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)


def test_batch_of_256_scenes_at_workload_scale(pipeline):
    scs = scenes.batch_scenes(256)
    n = np.array([sc.n for sc in scs])
    assert n.sum() > 9_000_000
    mats = np.array([int(sc.tag[0]) for sc in scs])
    picks = {int(np.argmin(n)), int(np.argmax(n))}
    for tag in np.unique(mats):
        picks.add(int(np.flatnonzero(mats == tag)[0]))
    assert len(np.unique(mats)) == 3
    substeps = 48

    batch = mpm.Batch(pipeline)
    states = [mpm.MpmState.from_scene(pipeline, sc) for sc in scs]
    for st in states:
        batch.add(st)
    batch.set_autotune(True)
    done = 0
    while batch.autotune_state()[0] and done < substeps - 8:
        batch.run(1)
        done += 1
    assert not batch.autotune_state()[0], "the tuner did not finish"
    batch.set_graph(True)
    batch.run(substeps - done)
    batch.synchronize()
    got = {i: states[i].read_particles() for i in sorted(picks)}
    batch.close()
    for st in states:
        st.close()

    for i in sorted(picks):
        ref = _alone(pipeline, scs[i], substeps)
        for f in FIELDS:
            assert np.array_equal(got[i][f], ref[f]), (i, scs[i].name, f)
    small = int(np.argmin(n))
    _check_oracle(got[small], scs[small], substeps)


def test_served_stream_at_workload_scale(pipeline):
    specs = scenes.served_requests(1.0)
    assert len(specs) > 150
    n = np.array([sp.n for sp in specs])
    hz = np.array([sp.horizon for sp in specs])
    picks = {int(np.argmax(n)), int(np.argmax(hz))}
    rng = np.random.Generator(np.random.PCG64(0x5E57))
    for i in rng.permutation(len(specs)):
        if len(picks) >= 5:
            break
        picks.add(int(i))
    short = np.flatnonzero(hz <= 120)
    small = int(short[np.argmin(n[short])])
    picks.add(small)

    reqs = [serve.Request(1000.0 * sp.arrival, sp, sp.horizon, rid=i,
                          keep="particles" if i in picks else "none") for i, sp in enumerate(specs)]
    srv = serve.Server(pipeline, hbm_budget_bytes=200e9)
    rep = srv.run(reqs, clock="steps")
    srv.close()
    assert rep.completed == len(specs) and rep.rejected == 0
    assert rep.peak_active >= 20
    for i in sorted(picks):
        r = reqs[i]
        assert r.steps == r.horizon
        ref = _alone(pipeline, specs[i].build(), specs[i].horizon)
        for f in FIELDS:
            assert np.array_equal(r.result[f], ref[f]), (i, specs[i].n, specs[i].horizon, f)
    _check_oracle(reqs[small].result, specs[small].build(), specs[small].horizon)
