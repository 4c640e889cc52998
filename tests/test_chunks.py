"""Particle chunks (nexus_amd/chunks.py, NexusState's chunk bookkeeping) over a host
stand-in of the scene with the library's swap-remove rule: after any sequence of adds,
extends and removals every slot's chunk is the chunk its particle was added to."""
import numpy as np

from nexus_amd.chunks import ParticleChunks


class FakeState:
    """Particle labels in slot order; swap_remove as GpuParticles::swap_remove (distinct ids
    in descending order, the last particle moved into each freed slot)."""

    def __init__(self):
        self.labels = []

    def __len__(self):
        return len(self.labels)

    def add_particles(self, scene):
        self.labels.extend(scene.labels)

    def swap_remove(self, ids):
        moves = []
        for i in sorted({int(k) for k in ids}, reverse=True):
            if i >= len(self.labels):
                continue
            last = len(self.labels) - 1
            if i != last:
                self.labels[i] = self.labels[last]
                moves.append((last, i))
            self.labels.pop()
        return moves


class Block:
    def __init__(self, chunk_label, n):
        self.labels = [chunk_label] * n
        self.n = n


def test_chunk_bookkeeping_follows_swap_remove():
    rng = np.random.default_rng(0)
    st = FakeState()
    ch = ParticleChunks(st)
    handles = []
    for step in range(60):
        op = rng.integers(0, 4)
        if op == 0 or not handles:
            h = ch._next
            handles.append(ch.add_particles(Block(h, int(rng.integers(1, 40)))))
        elif op == 1:
            h = handles[rng.integers(len(handles))]
            ch.extend_chunk(h, Block(h, int(rng.integers(1, 20))))
        elif op == 2:
            h = handles.pop(rng.integers(len(handles)))
            ch.remove_chunk(h)
            assert h not in st.labels
        else:
            h = handles[rng.integers(len(handles))]
            before = st.labels.count(h)
            k = int(rng.integers(0, 30))
            assert ch.remove_particles_from_chunk(h, k) == min(k, before)
            assert st.labels.count(h) == before - min(k, before) == ch.sizes[h]
        assert ch.slot2chunk == st.labels
        for h in handles:
            assert ch.sizes[h] == st.labels.count(h)
