"""Error paths of the worlds' bookkeeping, reached from the host.

A failed upload into a world (forced with the library's test hook nxm_debug_fail_uploads,
which fails before any copy is queued, as a failed device copy would) must leave the
scene on the host and detached: its slot freed, no range in the arena, and every object
still usable and destroyable in any order.  Before this was so, a failed
nxm_batch_add_scene left the batch's slot pointing at the scene; destroying the scene
then left that pointer dangling, and the batch's next table rebuild, step or destruction
read freed memory (a segmentation fault at interpreter teardown).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from nexus_amd import mpm, scenes

from parity import assert_field_close, oracle_pair

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cube(n_side=6, center=(0.5, 0.5, 0.5), seed=3):
    return scenes.elastic_cube(n_side, seed=seed, side=0.08, center=center)


def test_failed_batch_add_leaves_scene_detached(pipeline):
    """The failed scene can be destroyed while the batch keeps stepping; the member that
    was there all along matches the oracle."""
    a_sc = _cube()
    a = mpm.MpmState.from_scene(pipeline, a_sc)
    b = mpm.MpmState.from_scene(pipeline, _cube(4, (0.3, 0.3, 0.3), seed=4))
    batch = mpm.Batch(pipeline)
    batch.add(a)
    batch.step()
    mpm.debug_fail_uploads(1)
    try:
        with pytest.raises(mpm.MpmError) as e:
            batch.add(b)
        assert e.value.code == 7 and "injected" in str(e.value)
    finally:
        mpm.debug_fail_uploads(0)
    assert len(batch) == 1
    b.close()
    batch.step()
    batch.step()
    batch.synchronize()
    p = a.read_particles()
    o32, o64 = oracle_pair(a_sc, 3)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    batch.remove(a)
    batch.close()
    a.close()


def test_failed_reentry_after_swap_remove(pipeline):
    """swap_remove takes a batch member to the host and back; when the way back fails the
    member is on the host with its particles removed, the batch still steps, and removing
    it from the batch does not touch the slot it no longer holds."""
    sc = _cube()
    st = mpm.MpmState.from_scene(pipeline, sc)
    other_sc = _cube(4, (0.3, 0.3, 0.3), seed=5)
    other = mpm.MpmState.from_scene(pipeline, other_sc)
    batch = mpm.Batch(pipeline)
    batch.add(st)
    batch.add(other)
    batch.step()
    mpm.debug_fail_uploads(1)
    try:
        with pytest.raises(mpm.MpmError):
            st.swap_remove([0, 5])
    finally:
        mpm.debug_fail_uploads(0)
    assert len(st) == sc.n - 2
    batch.step()
    batch.remove(st)
    # stepped alone, the scene carries on from its host state
    pipeline.step(st)
    st.synchronize()
    assert np.all(np.isfinite(st.read_positions()))
    p = other.read_particles()
    o32, o64 = oracle_pair(other_sc, 2)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    batch.close()
    st.close()
    other.close()


def test_failed_append_keeps_particle_count(pipeline):
    """Particles appended to a resident scene whose upload fails are not added on either
    side: the scene keeps stepping as before and matches the oracle."""
    sc = _cube()
    st = mpm.MpmState.from_scene(pipeline, sc)
    pipeline.step(st)
    mpm.debug_fail_uploads(1)
    try:
        with pytest.raises(mpm.MpmError):
            st.add_particles(_cube(3, (0.3, 0.3, 0.3), seed=6))
    finally:
        mpm.debug_fail_uploads(0)
    assert len(st) == sc.n
    pipeline.step(st)
    p = st.read_particles()
    o32, o64 = oracle_pair(sc, 2)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)
    st.close()


def test_launch_config_limits(pipeline):
    """The largest P2G workgroup count the tuner tries (512 per CU) is accepted, one step
    above it is refused on the host, and the batch steps on with the accepted shape."""
    st = mpm.MpmState.from_scene(pipeline, scenes.elastic_cube(5, side=0.08, center=(0.5, 0.5, 0.5)))
    batch = mpm.Batch(pipeline)
    batch.add(st)
    cu = batch.launch_config()[0] // 32            # the default is 32 per CU
    batch.set_launch_config(cu * 512, 4, cu * 64)
    with pytest.raises(mpm.MpmError) as e:
        batch.set_launch_config(cu * 512 + 8, 4, cu * 64)
    assert e.value.code == 1
    assert tuple(batch.launch_config()) == (cu * 512, 4, cu * 64)
    batch.step()
    batch.synchronize()
    assert np.all(np.isfinite(st.read_particles()["x"]))
    batch.close()
    st.close()


_TEARDOWN = r"""
import sys
sys.path.insert(0, sys.argv[1])
from nexus_amd import mpm, scenes
pipe = mpm.MpmPipeline(0)
a = mpm.MpmState.from_scene(pipe, scenes.elastic_cube(5, side=0.08, center=(0.5, 0.5, 0.5)))
b = mpm.MpmState.from_scene(pipe, scenes.elastic_cube(4, side=0.08, center=(0.3, 0.3, 0.3)))
batch = mpm.Batch(pipe)
batch.add(a)
batch.step()
try:
    batch.set_launch_config(7, 16, 2048)          # rejected on the host (not a multiple of 8)
except mpm.MpmError as e:
    assert e.code == 1
mpm.debug_fail_uploads(1)
try:
    batch.add(b)
except mpm.MpmError as e:
    assert e.code == 7
mpm.debug_fail_uploads(0)
b.close()
batch.step()
batch.synchronize()
pipe.close()        # the context first: its deletion waits for the batch and the scene
batch.step()
batch.synchronize()
batch.close()
a.close()
print("teardown ok", flush=True)
"""


def test_error_path_teardown_in_a_fresh_process():
    """The whole sequence in a child process, destroying the context before its batch and
    scene: the process must exit normally (exit status 0, no fault at teardown)."""
    r = subprocess.run([sys.executable, "-c", _TEARDOWN, ROOT], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "teardown ok" in r.stdout
