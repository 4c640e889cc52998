"""Pins of the oracle's render readback restatement (oracle/readback.py) by closed forms:
the deformation of F = I, the clamp, the singular values of a rotated diagonal stretch
against the stretch itself, the sign colours' truth table, failed particles in red and
the palette's wrap-around."""
import math

import numpy as np

import readback as RB


def test_deformation_closed_forms():
    F = np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1], [9, 0, 0, 0, -7, 0, 0, 0.5, 1]], np.float32)
    r = np.array([0.25, 0.1], np.float32)
    d = RB.deformation(F, r).reshape(2, 4, 4)          # [column][row]
    assert np.allclose(d[0, :3, :3], np.eye(3) * 0.5) and d[0, 3, 3] == 1 and not d[0, 3, :3].any()
    assert np.allclose(d[1, 0, 0], 0.8) and np.allclose(d[1, 1, 1], -0.8) and np.allclose(d[1, 2, 1], 0.1)


def test_volume_colours_are_the_stretches():
    a, b, c = 1.004, 0.997, 0.999
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Fm = R @ np.diag([a, b, c]) @ R.T
    F = Fm.T.reshape(1, 9).astype(np.float32)             # column-major
    col = RB.colors(1, np.zeros((1, 3)), F, np.zeros(1), np.zeros(1), np.ones(1), np.zeros(1), np.zeros((1, 3)),
                    np.zeros(1), None, 0.1, 1e-3)
    expect = [(1 - s) / 0.005 + 0.2 for s in sorted([a, b, c], reverse=True)]
    assert np.allclose(col[0, :3], expect, atol=2e-3)


def test_sign_colours_failed_red_and_palette_wrap():
    aff = np.array([0, 0x00000001, 0x00010001, 0x00020001], np.uint32)   # none / outside / inside / other side bit only
    n = aff.size
    col = RB.colors(6, np.zeros((n, 3)), np.tile(np.eye(3).reshape(9), (n, 1)), np.zeros(n), np.arange(n) + 5,
                    np.array([1, 1, 1, 0]), aff, np.zeros((n, 3)), np.zeros(n), None, 0.1, 1e-3)
    assert np.allclose(col[0, :3], 0) and np.allclose(col[1, :3], [0, 1, 0]) and np.allclose(col[2, :3], [1, 0, 0])
    assert np.allclose(col[3], [1, 0, 0, 1])                                  # disabled
    base = RB.colors(0, np.zeros((n, 3)), np.tile(np.eye(3).reshape(9), (n, 1)), np.zeros(n), np.arange(n) + 5,
                     np.ones(n), aff, np.zeros((n, 3)), np.zeros(n), None, 0.1, 1e-3)
    assert np.array_equal(base[1], RB.DEFAULT_GROUP_COLORS[0]) and np.array_equal(base[0], RB.DEFAULT_GROUP_COLORS[5])
