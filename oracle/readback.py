"""TEST INFRASTRUCTURE ONLY: numpy restatement of the reference's render readback
(src_mpm_shaders/solver/prep_readback.rs), the checker of nxm_scene_prep_readback and
nxm_scene_prep_readback_rigid.  Only tests may import it.

Inputs are the particle state as read back (id order); outputs are the 24 floats of
ReadbackData per particle: colour (4), deformation (column-major 4x4), position (3), pad.
"""
from __future__ import annotations

import numpy as np

# GpuReadbackData::DEFAULT_GROUP_COLORS (src_mpm/solver/prep_readback.rs:58-65).
DEFAULT_GROUP_COLORS = np.array([
    [124 / 255, 144 / 255, 1, 1], [8 / 255, 144 / 255, 1, 1], [124 / 255, 7 / 255, 1, 1],
    [124 / 255, 144 / 255, 7 / 255, 1], [200 / 255, 37 / 255, 1, 1], [124 / 255, 230 / 255, 25 / 255, 1],
], np.float32)

f32 = np.float32


def deformation(F: np.ndarray, radius: np.ndarray) -> np.ndarray:
    """compute_deformation (prep_readback.rs:108-118): diag(2 r) times F clamped entrywise
    to [-4, 4], padded to 4x4 (Mat4::from_mat3).  F: (n, 9) column-major.  Returns (n, 16)."""
    n = F.shape[0]
    out = np.zeros((n, 16), np.float32)
    d = (radius.astype(np.float32) * f32(2.0))[:, None]
    cl = np.clip(F.astype(np.float32), f32(-4.0), f32(4.0))
    for c in range(3):
        out[:, 4 * c:4 * c + 3] = d * cl[:, 3 * c:3 * c + 3]
    out[:, 15] = 1.0
    return out


def colors(mode: int, v, F, phase, group, enabled, affinity, normal, signed_distance, palette, h, dt):
    """compute_color (prep_readback.rs:170-243), singular values largest first (volume)."""
    n = v.shape[0]
    pal = DEFAULT_GROUP_COLORS if palette is None or len(palette) == 0 else np.asarray(palette, np.float32)
    base = pal[np.asarray(group, np.int64) % len(pal)].astype(np.float32)
    col = base.copy()
    w = base[:, 3]
    if mode == 2:
        col[:, :3] = np.abs(v.astype(np.float32)) * f32(dt) * f32(100.0) + f32(0.2)
    elif mode == 1:
        s = np.linalg.svd(F.reshape(n, 3, 3).transpose(0, 2, 1).astype(np.float64), compute_uv=False)
        s = np.sort(s, axis=1)[:, ::-1]
        col[:, :3] = ((1.0 - s) / 0.005 + 0.2).astype(np.float32)
    elif mode == 3:
        ph = phase.astype(np.float32)
        col[:, 0] = 0.0
        col[:, 1] = f32(0.4) * ph
        col[:, 2] = f32(0.4) * (f32(1.0) - ph)
    elif mode == 4:
        nz = np.any(normal != 0, axis=1)
        col[:, :3] = np.where(nz[:, None], (normal.astype(np.float32) + f32(1.0)) * f32(0.5), f32(0.0))
    elif mode == 5:
        d = signed_distance.astype(np.float32) / (f32(h) * f32(1.5))
        col[:, :3] = 0.0
        col[:, 0] = np.where(d > 0, 0.0, np.abs(d))
        col[:, 1] = np.where(d > 0, np.abs(d), 0.0)
    elif mode == 6:
        aff = np.asarray(affinity, np.uint32)
        inside = (aff >> 16) & (aff & 0xFFFF)
        col[:, :3] = 0.0
        col[:, 0] = np.where((aff != 0) & (inside != 0), 1.0, 0.0)
        col[:, 1] = np.where((aff != 0) & (inside == 0), 1.0, 0.0)
    col[:, 3] = w
    col[np.asarray(enabled) == 0] = np.array([1, 0, 0, 1], np.float32)
    return col


def quat_rotate(q: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Rotation of points p (k, 3) by the unit quaternion q = (x, y, z, w)."""
    b, w = q[:3].astype(np.float64), float(q[3])
    p = p.astype(np.float64)
    return p * (w * w - b @ b) + np.outer(p @ b, b) * 2.0 + np.cross(b, p) * (2.0 * w)
