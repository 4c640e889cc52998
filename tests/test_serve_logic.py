"""Host logic of the continuous-batching server (admission, FIFO order, horizons,
budget), with the device batch replaced by a recording stand-in."""
import types

import numpy as np
import pytest

from nexus_amd import serve


class FakeState:
    def __init__(self, n):
        self.n = n
        self.steps = 0
        self.closed = False

    @classmethod
    def from_scene(cls, pipeline, sc):
        return cls(sc.n)

    def device_bytes(self):
        return 1000 * self.n

    def stage_upload(self):
        pass

    def read_positions(self):
        return np.full((self.n, 3), float(self.steps), np.float32)

    def read_particles(self):
        return {"steps": self.steps}

    def close(self):
        self.closed = True


class FakeBatch:
    log = []

    def __init__(self, pipeline):
        self.members = []

    def add(self, st):
        assert st not in self.members
        self.members.append(st)

    def remove(self, st):
        self.members.remove(st)

    def discard(self, st):
        self.members.remove(st)
        st.close()

    def step(self):
        FakeBatch.log.append(len(self.members))
        for m in self.members:
            m.steps += 1

    def synchronize(self):
        pass

    def close(self):
        pass


@pytest.fixture
def fake(monkeypatch):
    FakeBatch.log = []
    monkeypatch.setattr(serve.mpm, "Batch", FakeBatch)
    monkeypatch.setattr(serve.mpm, "MpmState", FakeState)


def spec(n):
    return types.SimpleNamespace(build=lambda: types.SimpleNamespace(n=n))


def test_each_request_gets_its_horizon(fake):
    reqs = [serve.Request(a, spec(n), h, rid=i)
            for i, (a, n, h) in enumerate([(0, 10, 3), (0, 20, 5), (2, 30, 2), (10, 40, 1)])]
    srv = serve.Server(None, hbm_budget_bytes=1e12)
    rep = srv.run(reqs, clock="steps")
    assert rep.completed == 4 and rep.rejected == 0
    for r in reqs:
        assert r.steps == r.horizon
        assert np.all(r.result == r.horizon)          # read back after exactly horizon substeps
        assert r.finished - r.admitted == r.horizon
    assert reqs[2].admitted == 2 and reqs[3].admitted == 10   # idle gap skipped to the arrival
    assert rep.particle_substeps == sum(r.n * r.horizon for r in reqs)
    assert rep.substeps == 6                          # steps 0-4 and step 10
    assert FakeBatch.log == [2, 2, 3, 2, 1, 1]       # r0 leaves after step 2, r2 after 3, r1 after 4


def test_budget_admits_fifo(fake):
    reqs = [serve.Request(0, spec(100), 2, rid=i) for i in range(4)]
    srv = serve.Server(None, hbm_budget_bytes=150_000)      # one scene (100 kB) at a time
    rep = srv.run(reqs, clock="steps")
    assert rep.completed == 4 and rep.peak_active == 1
    fin = [r.finished for r in reqs]
    assert fin == sorted(fin) and len(set(fin)) == 4


def test_particle_cap(fake):
    reqs = [serve.Request(0, spec(60), 1, rid=i) for i in range(3)]
    rep = serve.Server(None, hbm_budget_bytes=1e12, max_active_particles=100).run(reqs, clock="steps")
    assert rep.completed == 3 and rep.peak_active == 1


def test_oversize_rejected(fake):
    reqs = [serve.Request(0, spec(10), 1), serve.Request(0, spec(10_000), 1), serve.Request(1, spec(10), 1)]
    rep = serve.Server(None, hbm_budget_bytes=1_000_000).run(reqs, clock="steps")
    assert rep.completed == 2 and rep.rejected == 1
    assert reqs[1].error and reqs[1].state is None


def test_wall_clock_ingest(fake):
    reqs = [serve.Request(0.0, spec(10), 3), serve.Request(0.02, spec(20), 2), serve.Request(0.05, spec(5), 4)]
    rep = serve.Server(None, hbm_budget_bytes=1e12).run(reqs, clock="wall")
    assert rep.completed == 3
    for r in reqs:
        assert r.steps == r.horizon and r.admitted >= r.arrival
    s = rep.summary()
    assert s["end_to_end_s_p99"] >= s["end_to_end_s_p50"] >= 0
    assert s["particle_substeps"] == 10 * 3 + 20 * 2 + 5 * 4


def test_weighted_percentiles():
    rep = serve.ServeReport(0, 0, 0, 0, 1.0, 1.0, step_latency_ms=[1.0, 10.0], step_weights=[99, 1])
    assert rep._weighted_pct(0.5) == 1.0
    assert rep._weighted_pct(0.995) == 10.0


def test_error_releases_states(fake, monkeypatch):
    """A failing substep propagates, and every prepared or admitted scene is released."""
    def boom(self):
        raise RuntimeError("device error")
    monkeypatch.setattr(FakeBatch, "step", boom)
    reqs = [serve.Request(0, spec(10), 3), serve.Request(0, spec(20), 3)]
    with pytest.raises(RuntimeError):
        serve.Server(None, hbm_budget_bytes=1e12).run(reqs, clock="steps")
    assert all(r.state is None for r in reqs)
