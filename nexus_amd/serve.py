"""Continuous-batching server for MPM scenes on one GPU.

Requests (a scene and a horizon in substeps) arrive over time.  Admitted scenes join one
``Batch`` (one launch per stage for all of them); after every batch substep the scenes
that reached their horizon are read back (their particle positions are the response) and
leave the batch, and waiting scenes are admitted in arrival order while their device
memory fits the HBM budget.  A scene's results do not depend on the company it keeps:
P2G sums in exact fixed point and every other stage works per particle or per node, so a
served scene evolves bit for bit as it does alone.

The reference has no serving layer; its MPM state is one scene stepped by one caller
(``MpmPipeline::step``, src_mpm/pipeline.rs:385).
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import mpm
from . import scenes as _scenes


@dataclass(eq=False)
class Request:
    """One request.  ``arrival`` is in seconds (wall clock) or in batch substeps (step
    clock).  ``scene`` is a Scene or any object with ``build() -> Scene``.  Requests compare
    by identity."""

    arrival: float
    scene: object
    horizon: int
    rid: int = 0
    keep: str = "positions"            # "positions", "particles" or "none"
    # filled in by the server
    n: int = 0
    state: mpm.MpmState | None = None
    bytes: int = 0
    admitted: float = float("nan")
    finished: float = float("nan")
    steps: int = 0
    result: object = None
    error: str | None = None


@dataclass
class ServeReport:
    completed: int
    rejected: int
    substeps: int                       # batch substeps run
    particle_substeps: int              # sum over scenes of particles x substeps
    makespan_s: float                   # first arrival to last completion (wall)
    busy_s: float                       # wall time spent in batch substeps
    step_latency_ms: list = field(default_factory=list)      # one per batch substep
    step_weights: list = field(default_factory=list)         # active scenes in that substep
    end_to_end_s: list = field(default_factory=list)         # arrival -> response, per request
    wait_s: list = field(default_factory=list)               # arrival -> admission, per request
    peak_active: int = 0
    peak_bytes: int = 0
    step_end_s: list = field(default_factory=list)          # end of each substep (wall clock)
    step_particles: list = field(default_factory=list)      # resident particles in it
    arrival_window_s: tuple = (0.0, 0.0)                    # first and last arrival

    def _weighted_pct(self, q: float) -> float:
        if not self.step_latency_ms:
            return float("nan")
        lat = np.asarray(self.step_latency_ms)
        w = np.asarray(self.step_weights, dtype=np.float64)
        order = np.argsort(lat)
        cw = np.cumsum(w[order])
        k = int(np.searchsorted(cw, q * cw[-1]))
        return float(lat[order][min(k, len(lat) - 1)])

    def summary(self) -> dict:
        """Per-substep latency percentiles over (scene, substep) pairs, end-to-end latency
        percentiles over requests, sustained particle-substeps/s over the makespan."""
        # Throughput while requests keep arriving (the second half of the arrival window,
        # once the batch has filled), beside the whole run's.
        a0, a1 = self.arrival_window_s
        mid = 0.5 * (a0 + a1)
        ends = np.asarray(self.step_end_s)
        parts = np.asarray(self.step_particles, np.float64)
        sel = (ends > mid) & (ends <= a1)
        loaded = float(parts[sel].sum() / (a1 - mid)) if a1 > mid and sel.any() else float("nan")
        e2e = np.asarray(self.end_to_end_s) if self.end_to_end_s else np.asarray([np.nan])
        wait = np.asarray(self.wait_s) if self.wait_s else np.asarray([np.nan])
        return {
            "completed": self.completed,
            "rejected": self.rejected,
            "batch_substeps": self.substeps,
            "particle_substeps": self.particle_substeps,
            "makespan_s": self.makespan_s,
            "sustained_particle_substeps_per_s": self.particle_substeps / self.makespan_s if self.makespan_s > 0 else 0.0,
            "busy_particle_substeps_per_s": self.particle_substeps / self.busy_s if self.busy_s > 0 else 0.0,
            "loaded_particle_substeps_per_s": loaded,
            "substep_latency_ms_p50": self._weighted_pct(0.50),
            "substep_latency_ms_p99": self._weighted_pct(0.99),
            "end_to_end_s_p50": float(np.percentile(e2e, 50)),
            "end_to_end_s_p99": float(np.percentile(e2e, 99)),
            "queue_wait_s_p50": float(np.percentile(wait, 50)),
            "queue_wait_s_p99": float(np.percentile(wait, 99)),
            "peak_active_scenes": self.peak_active,
            "peak_device_bytes": self.peak_bytes,
        }


class Server:
    """Admission by HBM budget (and optionally by resident particles), FIFO order,
    continuous batching at substep granularity."""

    def __init__(self, pipeline: mpm.MpmPipeline, hbm_budget_bytes: float = 200e9,
                 max_active_particles: int | None = None, ingest_workers: int = 4, adaptive_dt: bool = False):
        """adaptive_dt: every admitted scene takes adaptive substeps (each substep's length
        min(its dt, the CFL bound of its state), chosen on the device), so that a request
        whose dt is too long for its material or speed still steps stably."""
        self.pipeline = pipeline
        self.adaptive_dt = bool(adaptive_dt)
        self.budget = float(hbm_budget_bytes)
        self.max_particles = max_active_particles
        self.ingest_workers = max(1, int(ingest_workers))
        self.batch = mpm.Batch(pipeline)

    def close(self):
        self.batch.close()

    def _prepare(self, r: Request) -> None:
        sc = r.scene if isinstance(r.scene, _scenes.Scene) else r.scene.build()
        r.n = sc.n
        r.state = mpm.MpmState.from_scene(self.pipeline, sc)   # host-side until admitted
        if self.adaptive_dt:
            r.state.set_adaptive_dt(True)
        r.state.stage_upload()          # packed here, so that admission only queues copies
        r.bytes = r.state.device_bytes()

    def run(self, requests: list[Request], clock: str = "wall") -> ServeReport:
        """Serves every request to completion.  clock="wall": arrivals in seconds from
        the start of the call; ingest threads build each arrived request's host-side scene
        (and pack it for upload) while the GPU steps the batch.  clock="steps": arrivals in batch
        substeps, requests prepared in the serving loop (reproducible)."""
        if clock not in ("wall", "steps"):
            raise ValueError("clock must be 'wall' or 'steps'")
        wall = clock == "wall"
        order = sorted(requests, key=lambda r: r.arrival)
        pending = deque(order)
        queue: deque[Request] = deque()
        active: list[Request] = []
        rep = ServeReport(0, 0, 0, 0, 0.0, 0.0)
        used = 0
        resident = 0
        step_idx = 0
        t0 = time.perf_counter()
        first_arrival = order[0].arrival if order else 0.0

        def now() -> float:
            return (time.perf_counter() - t0) if wall else float(step_idx)

        # Ingest: requests move to ``ready`` once arrived and prepared.
        ready: deque[Request] = deque()
        cv = threading.Condition()
        ingest_error: list[BaseException] = []
        stop = threading.Event()

        # Ingest workers prepare arrived requests in parallel (the native calls release the
        # interpreter lock); requests become ready in arrival order, so admission stays FIFO.
        prepared: set[int] = set()
        next_ready = [0]

        def ingest(w: int):
            try:
                for k in range(w, len(order), self.ingest_workers):
                    r = order[k]
                    while not stop.is_set():
                        delay = r.arrival - (time.perf_counter() - t0)
                        if delay <= 0:
                            break
                        stop.wait(min(delay, 0.05))
                    if stop.is_set():
                        return
                    self._prepare(r)
                    with cv:
                        prepared.add(k)
                        while next_ready[0] in prepared:
                            prepared.discard(next_ready[0])
                            ready.append(order[next_ready[0]])
                            next_ready[0] += 1
                        cv.notify()
            except BaseException as e:  # surfaced by the serving loop
                with cv:
                    ingest_error.append(e)
                    cv.notify()

        workers = []
        if wall:
            pending.clear()
            for w in range(min(self.ingest_workers, max(len(order), 1))):
                t = threading.Thread(target=ingest, args=(w,), name=f"nxm-ingest-{w}", daemon=True)
                t.start()
                workers.append(t)
        remaining = len(order)
        try:
            while remaining:
                if wall:
                    with cv:
                        if ingest_error:
                            raise ingest_error[0]
                        if not (active or queue or ready):
                            cv.wait(timeout=0.05)
                        while ready:
                            queue.append(ready.popleft())
                else:
                    t = now()
                    while pending and pending[0].arrival <= t:
                        queue.append(pending.popleft())
                # admission, FIFO: the head of the queue blocks the rest while it does not fit
                while queue:
                    r = queue[0]
                    if r.state is None:
                        self._prepare(r)
                    if r.bytes > self.budget:
                        queue.popleft()
                        r.error = f"needs {r.bytes} bytes, over the HBM budget"
                        r.state.close()
                        r.state = None
                        rep.rejected += 1
                        remaining -= 1
                        continue
                    if active and (used + r.bytes > self.budget or
                                   (self.max_particles is not None and resident + r.n > self.max_particles)):
                        break
                    queue.popleft()
                    self.batch.add(r.state)
                    used += r.bytes
                    resident += r.n
                    r.admitted = now()
                    active.append(r)
                rep.peak_active = max(rep.peak_active, len(active))
                rep.peak_bytes = max(rep.peak_bytes, used)
                if not active:
                    if not wall and pending:
                        step_idx = max(step_idx, int(np.ceil(pending[0].arrival)))
                    continue
                ts = time.perf_counter()
                self.batch.step()
                self.batch.synchronize()
                te = time.perf_counter()
                step_idx += 1
                rep.substeps += 1
                rep.busy_s += te - ts
                rep.step_latency_ms.append(1e3 * (te - ts))
                rep.step_weights.append(len(active))
                rep.particle_substeps += resident
                if wall:
                    rep.step_end_s.append(te - t0)
                    rep.step_particles.append(resident)
                done = []
                for r in active:
                    r.steps += 1
                    if r.steps >= r.horizon:
                        done.append(r)
                for r in done:
                    if r.keep == "positions":
                        r.result = r.state.read_positions()
                    elif r.keep == "particles":
                        r.result = r.state.read_particles()
                    self.batch.discard(r.state)      # results read: the state is not needed
                    r.state = None
                    used -= r.bytes
                    resident -= r.n
                    r.finished = now()
                    active.remove(r)
                    rep.completed += 1
                    remaining -= 1
                    if wall:
                        rep.end_to_end_s.append(r.finished - r.arrival)
                        rep.wait_s.append(r.admitted - r.arrival)
        finally:
            stop.set()
            for t in workers:
                t.join()
            # On an error, release what was prepared or admitted but not finished.
            for r in order:
                if r.state is not None:
                    if r in active:
                        self.batch.discard(r.state)
                    else:
                        r.state.close()
                    r.state = None
        rep.makespan_s = (time.perf_counter() - t0) - (first_arrival if wall else 0.0)
        if order:
            rep.arrival_window_s = (order[0].arrival, order[-1].arrival)
        return rep
