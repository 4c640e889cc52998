"""The block sort of steady-state substeps, compared in canonical form.

Only the first substep after an upload (or any host change of the arena) builds its
block keys and hash insertions in a separate key pass.  Every later substep takes them
from the previous substep's G2P kernels, which key each particle at its new position (and,
in CPIC worlds, from the contact G2P kernel on the side stream for the contact particles).
An extra or missing block would change neither x, v, C nor F, so the field tests cannot
see it; this test steps k substeps, reads the positions the next substep starts from,
steps once more, and compares that substep's sort with the oracle's sort of the same
positions: the active block set P U (P + {0,1}^3), per block the primary and total
particle counts and the ten bucket counts, per particle its primary block and local cell
(src_mpm_shaders/grid/sort.rs:104-163, 287-491)."""
import numpy as np
import pytest

import oracle as O
from nexus_amd import mpm, scenes

from parity import compare_sort
from test_gpu_cpic import contact_scene

pytestmark = pytest.mark.gpu


def _steady_sort(pipeline, sc, k, batch_with=None):
    st = mpm.MpmState.from_scene(pipeline, sc)
    batch = None
    if batch_with is not None:
        other = mpm.MpmState.from_scene(pipeline, batch_with)
        batch = mpm.Batch(pipeline)
        batch.add(other)
        batch.add(st)
        for _ in range(k):
            batch.step()
    else:
        for _ in range(k):
            pipeline.step(st)
    x = st.read_positions()
    if batch is not None:
        batch.step()
    else:
        pipeline.step(st)
    can = st.read_sort_canonical()
    # the oracle's sort of the same positions (everything else about the state is
    # irrelevant to the sort: it keys every particle, enabled or not)
    sc2 = sc.copy()
    sc2.x = np.ascontiguousarray(x, np.float32)
    o32 = O.Oracle(sc2, 32)
    o32.step(1)
    compare_sort(can, o32, sc2.x, sc.cell_width)
    if batch is not None:
        batch.close()
    st.close()
    return can


@pytest.mark.parametrize("k", [1, 7])
def test_steady_state_sort_plain(pipeline, k):
    sc = scenes.sand_column(20000, seed=21, lo=(0.4, 0.1, 0.4), hi=(0.5, 0.3, 0.5), n_colliders=16)
    can = _steady_sort(pipeline, sc, k)
    assert can["vid"].shape[0] > 100


def test_steady_state_sort_fluid_in_a_batch(pipeline):
    """Clumped fluid (many particles per cell) stepped inside a batch beside another scene."""
    sc = scenes.fluid_clumps(30000, seed=22, n_clumps=4, cell_width=1 / 256, sigma=(0.01, 0.03))
    _steady_sort(pipeline, sc, 5, batch_with=scenes.elastic_cube(8))


def test_steady_state_sort_cpic(pipeline):
    """CPIC: the contact particles are keyed by the contact G2P kernel on the side stream."""
    sc = contact_scene()
    _steady_sort(pipeline, sc, 8)
