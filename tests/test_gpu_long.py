"""Long horizon: 1000 substeps on the GPU against the FP64 oracle's fixtures
(tests/golden/long_*.npz, written by tests/golden/make_golden.py).

Centre of mass and kinetic energy must agree within 1e-3 (relative; the centre of
mass is compared through its displacement from the start, relative to the
displacement's size plus 1e-3 of a cell).  Four of the scenes are dominated by the
material and the colliders (sand landing on a tilted floor, a bouncing elastic cube,
fluid splashing in a box, the full 1,048,576-particle sand column); three by gravity.
"""
import os
import sys

import numpy as np
import pytest

from nexus_amd import mpm

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import EVERY, STEPS, long_horizon_scenes, observables  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = 1e-3


def gpu_observables(state):
    p = state.read_particles()

    class S:
        pass

    s = S()
    s.mass, s.x, s.v, s.enabled = p["mass"], p["x"], p["v"], p["enabled"]
    return observables(s)


@pytest.mark.parametrize("name", ["sand_column_65k", "fluid_clumps_65k", "neo_hookean_cube_32k", "sand_landing",
                                  "cube_bounce", "fluid_box", "sand_column_1m"])
def test_long_horizon_com_and_kinetic_energy(pipeline, name):
    path = os.path.join(HERE, "golden", f"long_{name}.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    gold = np.load(path)
    scene = long_horizon_scenes()[name]
    st = mpm.MpmState.from_scene(pipeline, scene)
    com0 = gold["com64"][0]
    for i, k in enumerate(range(0, STEPS + 1, EVERY)):
        if k:
            pipeline.run(st, EVERY)
        com, ke, en = gpu_observables(st)
        ref_com, ref_ke = gold["com64"][i], gold["ke64"][i]
        disp = np.linalg.norm(ref_com - com0)
        assert np.linalg.norm(com - ref_com) <= TOL * disp + 1e-3 * scene.cell_width, (k, com, ref_com)
        assert abs(ke - ref_ke) <= TOL * ref_ke + 1e-30, (k, ke, ref_ke)
        assert en == gold["enabled64"][i]
    st.close()
