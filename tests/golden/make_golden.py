"""Writes the long-horizon fixtures: centre of mass and kinetic energy of the FP64
oracle (oracle/, the CPU restatement of the reference substep) every 50 substeps
over 1000 substeps, for three seeded scenes.  The reference itself holds no
vectors for this path and cannot be built here; these fixtures pin the GPU path to
the oracle at horizons the tests cannot afford to recompute on every run.

    python tests/golden/make_golden.py [name ...]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

from nexus_amd import scenes  # noqa: E402
import oracle as O  # noqa: E402

STEPS = 1000
EVERY = 50


def long_horizon_scenes():
    return {
        "sand_column_65k": scenes.sand_column(65536),
        "fluid_clumps_65k": scenes.fluid_clumps(65536, n_clumps=8),
        "neo_hookean_cube_32k": scenes.elastic_cube(32),
    }


def observables(s):
    m = s.mass.astype(np.float64)
    x = s.x.astype(np.float64)
    v = s.v.astype(np.float64)
    M = m.sum()
    com = (m[:, None] * x).sum(0) / M
    ke = 0.5 * (m * (v * v).sum(1)).sum()
    return com, ke, int(s.enabled.sum())


def run(name, scene, bits):
    o = O.Oracle(scene, bits)
    coms, kes, ens = [], [], []
    for k in range(0, STEPS + 1, EVERY):
        if k:
            o.step(EVERY, keep_debug=False)
        c, e, en = observables(o.s)
        coms.append(c)
        kes.append(e)
        ens.append(en)
    return np.array(coms), np.array(kes), np.array(ens)


def main(names):
    for name, sc in long_horizon_scenes().items():
        if names and name not in names:
            continue
        t = time.time()
        c64, k64, e64 = run(name, sc, 64)
        c32, k32, e32 = run(name, sc, 32)
        np.savez(os.path.join(HERE, f"long_{name}.npz"), steps=np.arange(0, STEPS + 1, EVERY), com64=c64, ke64=k64,
                 enabled64=e64, com32=c32, ke32=k32, enabled32=e32)
        print(f"{name}: {time.time() - t:.0f} s; KE fp64 {k64[-1]:.6e} fp32 {k32[-1]:.6e}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
