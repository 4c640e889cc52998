"""Particle chunks over one scene: the bookkeeping NexusState keeps on top of
GpuParticles::append / swap_remove (src/state.rs:665-774), for emitters and serving.

A chunk is a group of particles added together (``add_particles``) and possibly extended
(``extend_chunk``); it can be removed whole (``remove_chunk``) or in part
(``remove_particles_from_chunk``, the highest particle slots first).  Removal is the
scene's swap-remove (``nxm_scene_swap_remove``): the last particle moves into each freed
slot, and the slot-to-chunk table follows the moves the library reports.
"""
from __future__ import annotations

import numpy as np


class ParticleChunks:
    """Chunk handles over one state with ``add_particles(scene)``, ``swap_remove(ids) ->
    [(from, to)]`` and ``len()`` (an ``mpm.MpmState``)."""

    def __init__(self, state):
        self.state = state
        self.slot2chunk: list[int] = [-1] * len(state)   # particles present before: no chunk
        self.sizes: dict[int, int] = {}
        self._next = 0

    def add_particles(self, scene) -> int:
        """Appends the scene's particles as a new chunk; returns its handle."""
        chunk = self._next
        self._next += 1
        self.sizes[chunk] = 0
        self.extend_chunk(chunk, scene)
        return chunk

    def extend_chunk(self, chunk: int, scene) -> None:
        if chunk not in self.sizes:
            raise KeyError(f"unknown chunk {chunk}")
        self.state.add_particles(scene)
        self.slot2chunk.extend([chunk] * scene.n)
        self.sizes[chunk] += scene.n

    def slots(self, chunk: int) -> np.ndarray:
        """The chunk's particle slots (ids in the scene), ascending."""
        return np.array([i for i, c in enumerate(self.slot2chunk) if c == chunk], np.int64)

    def remove_chunk(self, chunk: int) -> None:
        """Removes every particle of the chunk and drops the handle."""
        self._swap_remove(self.slots(chunk))
        del self.sizes[chunk]

    def remove_particles_from_chunk(self, chunk: int, count: int) -> int:
        """Removes up to count particles of the chunk, highest slots first; the chunk is
        kept even when emptied.  Returns how many were removed."""
        slots = self.slots(chunk)[::-1][:max(count, 0)]
        self._swap_remove(slots)
        self.sizes[chunk] = max(self.sizes[chunk] - len(slots), 0)
        return int(len(slots))

    def _swap_remove(self, slots) -> None:
        if len(slots) == 0:
            return
        moves = self.state.swap_remove(np.asarray(slots, np.uint32))
        for src, dst in moves:
            self.slot2chunk[dst] = self.slot2chunk[src]
        del self.slot2chunk[len(self.state):]
