"""Writes the long-horizon fixtures: centre of mass and kinetic energy of the FP64
oracle (oracle/, the CPU restatement of the reference substep) every 50 substeps
over 1000 substeps, for seeded scenes (the FP32 mirror's beside them).  The reference
itself holds no vectors for this path and cannot be built here; these fixtures pin the
GPU path to the oracle at horizons the tests cannot afford to recompute on every run.

Three scenes are dominated by gravity (free fall of a cube, sparse sand and fluid
clumps); the other four by the material and the colliders: Drucker-Prager sand landing
on a tilted floor, an elastic cube bouncing on a floor, fluid clumps splashing into a
box, and the full 1,048,576-particle sand column among its 16 cuboids.

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


def _lattice(lo, hi, spacing, seed):
    """Jittered lattice (+-spacing/4) filling the box [lo, hi]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    k = np.maximum(np.round((hi - lo) / spacing).astype(int), 1)
    idx = np.stack(np.meshgrid(*(np.arange(c) for c in k), indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    return lo + (idx + 0.5) * spacing + rng.uniform(-0.25 * spacing, 0.25 * spacing, idx.shape)


def _quat_z(deg):
    a = np.deg2rad(deg) / 2
    return (0.0, 0.0, float(np.sin(a)), float(np.cos(a)))


def sand_landing():
    """Drucker-Prager sand (the sand column's material) in a 0.16 x 0.12 x 0.16 m block,
    8 particles per cell at h = 1/128, its bottom 5-27 mm above a floor tilted by 8 degrees
    (separate, friction 0.5), beside two turned cuboids standing on the floor (stick and
    slip); dt = 1e-4."""
    h = 1.0 / 128
    pos = _lattice((0.42, 0.117, 0.42), (0.58, 0.237, 0.58), h / 2, 0x4E584D10)
    parts = scenes.particles(pos, h / 4, 2200.0, scenes.model_sand(3.5e5, 0.3, friction_deg=30.0))
    cols = np.concatenate([
        scenes.cuboid((0.5, 0.05, 0.5), (0.45, 0.05, 0.45), rotation=_quat_z(8.0), bc=scenes.BC_SEPARATE,
                      friction=0.5),
        scenes.cuboid((0.28, 0.12, 0.5), (0.03, 0.03, 0.12), rotation=(0.0, 0.3826834, 0.0, 0.9238795),
                      bc=scenes.BC_STICK),
        scenes.cuboid((0.67, 0.14, 0.47), (0.02, 0.04, 0.1), rotation=(0.0, 0.0, 0.2588190, 0.9659258),
                      bc=scenes.BC_SLIP),
    ])
    return scenes.make_scene(parts, h, 1e-4, colliders=cols, substeps=STEPS, name="sand landing on a tilted floor")


def cube_bounce():
    """A neo-Hookean cube (E = 1e5, nu = 0.3, density 1000) of side 0.1 m, 8 particles per
    cell at h = 1/128, dropped from 5 mm onto a floor (separate, no friction): it lands at
    about 32 ms, rebounds and flies again; dt = 1e-4."""
    h = 1.0 / 128
    pos = _lattice((0.45, 0.105, 0.45), (0.55, 0.205, 0.55), h / 2, 0x4E584D11)
    parts = scenes.particles(pos, h / 4, 1000.0, scenes.model_neo_hookean(1e5, 0.3))
    floor = scenes.cuboid((0.5, 0.05, 0.5), (0.45, 0.05, 0.45), bc=scenes.BC_SEPARATE, friction=0.0)
    return scenes.make_scene(parts, h, 1e-4, colliders=floor, substeps=STEPS, name="neo-Hookean cube bounce")


def fluid_box():
    """Weakly-compressible fluid (K = 1e4, gamma = 7, density 1000) in 6 Gaussian clumps
    (sigma 8-15 mm) 1-4 cm above the floor of a stick-walled box [0.3, 0.7]^3, 8 particles
    per cell at h = 1/128; dt = 1e-4."""
    h = 1.0 / 128
    rng = np.random.Generator(np.random.PCG64(0x4E584D12))
    out = []
    for _ in range(6):
        c = np.array([rng.uniform(0.38, 0.62), rng.uniform(0.34, 0.37), rng.uniform(0.38, 0.62)])
        sig = rng.uniform(0.008, 0.015)
        pts = c + rng.normal(scale=sig, size=(8000, 3))
        out.append(pts[np.all((pts > 0.31) & (pts < 0.69), axis=1)])
    parts = scenes.particles(np.concatenate(out), h / 4, 1000.0, scenes.model_fluid(1e4))
    walls = scenes.box_walls(h, lo=0.3, hi=0.7)
    return scenes.make_scene(parts, h, 1e-4, colliders=walls, substeps=STEPS, name="fluid clumps in a box")


def long_horizon_scenes():
    return {
        "sand_column_65k": scenes.sand_column(65536),
        "fluid_clumps_65k": scenes.fluid_clumps(65536, n_clumps=8),
        "neo_hookean_cube_32k": scenes.elastic_cube(32),
        "sand_landing": sand_landing(),
        "cube_bounce": cube_bounce(),
        "fluid_box": fluid_box(),
        "sand_column_1m": scenes.sand_column(),
    }


def free_fall_ke(scene, steps=STEPS):
    """Kinetic energy of the scene's particles in free fall after ``steps`` substeps
    (symplectic Euler from its initial velocities)."""
    m = scene.mass.astype(np.float64)
    v = scene.v.astype(np.float64) + np.asarray(scene.gravity) * scene.dt * steps
    return 0.5 * (m * (v * v).sum(1)).sum()


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
        print(f"{name}: {time.time() - t:.0f} s; KE fp64 {k64[-1]:.6e} fp32 {k32[-1]:.6e} "
              f"free fall {free_fall_ke(sc):.6e}; COM fp32-fp64 {np.abs(c32 - c64).max():.3e}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
