"""Shared comparison helpers for the GPU-vs-oracle parity tests."""
from __future__ import annotations

import numpy as np

import oracle as O

# One FP32 substep must match within 1e-5 relative (norm-wise over the field), or be
# no further from the FP64 oracle than the reference algorithm run in FP32 is (times
# a small factor): the GPU result is then exactly as good as the reference's own FP32
# arithmetic allows.
REL_TOL = 1e-5
FP32_FACTOR = 8.0


def field_errors(gpu: np.ndarray, o32: np.ndarray, o64: np.ndarray):
    g = np.asarray(gpu, np.float64)
    a = np.asarray(o32, np.float64)
    b = np.asarray(o64, np.float64)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    e_gpu = float(np.max(np.abs(g - b))) if b.size else 0.0
    e_32 = float(np.max(np.abs(a - b))) if b.size else 0.0
    return e_gpu, e_32, scale


def assert_field_close(name: str, gpu, o32, o64):
    e_gpu, e_32, scale = field_errors(gpu, o32, o64)
    bound = max(REL_TOL * scale, FP32_FACTOR * e_32, 1e-30)
    assert e_gpu <= bound, f"{name}: |gpu - fp64| = {e_gpu:.3e} > {bound:.3e} (fp32 mirror {e_32:.3e}, scale {scale:.3e})"
    return e_gpu, e_32, scale


def oracle_pair(scene, steps: int = 1):
    o32 = O.Oracle(scene, 32)
    o64 = O.Oracle(scene, 64)
    o32.step(steps)
    o64.step(steps)
    return o32, o64


def oracle_canonical_per_particle(o: O.Oracle, n: int):
    """Per particle (primary block vid, local cell) from the oracle's last sort."""
    b = o.blocks()
    vid_of = np.zeros((n, 3), np.int32)
    for blk in range(b["vid"].shape[0]):
        start = int(b["first"][blk])
        cnt = int(b["np"][blk])
        ids = b["sorted"][start:start + cnt]
        vid_of[ids] = b["vid"][blk]
    return vid_of, b


def compare_sort(gpu_can: dict, o32: O.Oracle, x_before: np.ndarray, cell_width: float):
    """Bit-exact comparison of the canonical sort: active block set, per-block
    num_particles, num_particles_with_extras and bucket counts, and per particle its
    primary block and local cell."""
    n = x_before.shape[0]
    vid_of, b = oracle_canonical_per_particle(o32, n)
    ovid, oloc = O.block_keys(np.asarray(x_before, np.float32), cell_width, 32)
    assert np.array_equal(vid_of, ovid)
    o_map = {tuple(map(int, b["vid"][i])): i for i in range(b["vid"].shape[0])}
    g_map = {tuple(map(int, gpu_can["vid"][i])): i for i in range(gpu_can["vid"].shape[0])}
    assert set(o_map) == set(g_map), f"block sets differ: {len(set(o_map) ^ set(g_map))} blocks"
    for key, oi in o_map.items():
        gi = g_map[key]
        assert int(b["np"][oi]) == int(gpu_can["np"][gi]), key
        assert int(b["npe"][oi]) == int(gpu_can["npe"][gi]), key
        assert np.array_equal(b["buckets"][oi], gpu_can["buckets"][gi]), key
    gvid = gpu_can["vid"][gpu_can["particle_block"]]
    assert np.array_equal(gvid, ovid)
    cell = oloc[:, 0] + 4 * oloc[:, 1] + 16 * oloc[:, 2]
    assert np.array_equal(gpu_can["particle_cell"], cell)
