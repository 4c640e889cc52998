"""The workloads at their full sizes.

- Neo-Hookean cube, 32,768 particles, 100 substeps: the scene stays in free fall for the
  whole run (the walls are more than 0.27 m away), so the closed form of symplectic Euler
  holds: v = -g n dt, COM_y(n) = COM_y(0) - g dt^2 n (n + 1) / 2, F = I.
- Drucker-Prager sand column, 1,048,576 particles, and clumped fluid, 8,388,608 particles:
  one substep against the oracle, the sort bit-exact in canonical form and x, v, C, F
  within the parity bar.
- The same two workloads over many substeps (the block keys then come from the previous
  substep's G2P): two runs are bit-identical, and so is the run of the scene inside a batch
  beside another scene (P2G sums in exact fixed point, every other stage works per particle
  or per node, so neither the atomics' order nor the batch changes a bit).
"""
import numpy as np
import pytest

from nexus_amd import mpm, scenes

from parity import assert_field_close, compare_sort, oracle_pair

pytestmark = pytest.mark.gpu


def test_cube_100_substeps_free_fall(pipeline):
    sc = scenes.elastic_cube()
    assert sc.n == 32768
    st = mpm.MpmState.from_scene(pipeline, sc)
    n = 100
    for _ in range(n):
        pipeline.step(st)
    p = st.read_particles()
    st.close()
    dt = float(np.float32(sc.dt))
    g = 9.81
    assert np.allclose(p["v"][:, 1].astype(np.float64), -g * n * dt, rtol=2e-5, atol=0)
    assert np.abs(p["v"][:, [0, 2]]).max() < 1e-6
    com0 = sc.x.astype(np.float64).mean(0)
    com = p["x"].astype(np.float64).mean(0)
    assert abs(com[1] - (com0[1] - g * dt * dt * n * (n + 1) / 2)) < 2e-7
    eye = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    assert np.abs(p["F"] - eye).max() < 1e-5
    assert p["enabled"].all()


@pytest.mark.parametrize("make", [scenes.sand_column, scenes.fluid_clumps], ids=["sand_1M", "fluid_8M"])
def test_one_substep_full_size(pipeline, make):
    sc = make()
    x0 = sc.x.copy()
    st = mpm.MpmState.from_scene(pipeline, sc)
    pipeline.step(st)
    can = st.read_sort_canonical()
    p = st.read_particles()
    st.close()
    o32, o64 = oracle_pair(sc, 1)
    compare_sort(can, o32, x0, sc.cell_width)
    assert np.array_equal(p["enabled"], o32.s.enabled)
    for f in ("x", "v", "C", "F"):
        assert_field_close(f, p[f], getattr(o32.s, f), getattr(o64.s, f), o64.s)


@pytest.mark.parametrize("make,substeps", [(scenes.sand_column, 20), (scenes.fluid_clumps, 4)],
                         ids=["sand_1M", "fluid_8M"])
def test_full_size_runs_are_bit_identical(pipeline, make, substeps):
    sc = make()
    runs = []
    for _ in range(2):
        st = mpm.MpmState.from_scene(pipeline, sc)
        for _ in range(substeps):
            pipeline.step(st)
        runs.append(st.read_particles())
        st.close()
    other = scenes.elastic_cube(12)
    batch = mpm.Batch(pipeline)
    states = [mpm.MpmState.from_scene(pipeline, s) for s in (other, sc)]
    for s in states:
        batch.add(s)
    batch.run(substeps)
    batch.synchronize()
    runs.append(states[1].read_particles())
    batch.close()
    for s in states:
        s.close()
    assert runs[0]["enabled"].sum() > 0.99 * sc.n
    for got in runs[1:]:
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(got[f], runs[0][f]), f
