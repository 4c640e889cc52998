"""Shared comparison helpers for the GPU-vs-oracle parity tests."""
from __future__ import annotations

import os

import numpy as np

import oracle as O

# One FP32 substep must match the FP64 oracle within 1e-5 of the field's largest value
# (max-norm), or be no further from it than the reference algorithm itself run in FP32
# (the op-for-op FP32 mirror) times FP32_FACTOR: the GPU result is then as good as the
# reference's own FP32 arithmetic allows.  The second clause matters for fields whose
# FP32 floor lies above 1e-5: C on stiff materials, where the stress term is a difference
# of large nearly equal values.
#
# C has a third clause, its FP32 resolution.  C = m grad v - (V0 (4/h^2) dt) tau(F).  One
# ulp of the grid velocities moves grad v by about eps32 |v| 6/h (|dpt| <= 1.5 h, sum of
# weights 1, 4/h^2), and where F is not exactly I in FP32, tau's own cancellation
# (mu F F^T against mu I, for instance) leaves about eps32 M |F| of rounding in it (M =
# lambda + 2 mu, the material's P-wave modulus), times V0 (4/h^2) dt in C.  No FP32
# evaluation of C resolves it more finely, so where C's FP64 value is below that (a body
# in free fall, where C is rounding noise around zero) the relative and mirror clauses
# compare noise with noise; there C passes when |gpu - fp64| <= ULP_FACTOR x that
# resolution.
REL_TOL = 1e-5
FP32_FACTOR = 2.0
ULP_FACTOR = 4.0
EPS32 = float(np.finfo(np.float32).eps)

# Every comparison made in this process: (test, field, |gpu - fp64|, |fp32 mirror - fp64|,
# scale, clause, C's FP32 resolution or 0) -- summarized at the end of the session
# (tests/conftest.py).
RECORDS: list = []


def field_errors(gpu: np.ndarray, o32: np.ndarray, o64: np.ndarray):
    g = np.asarray(gpu, np.float64)
    a = np.asarray(o32, np.float64)
    b = np.asarray(o64, np.float64)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    e_gpu = float(np.max(np.abs(g - b))) if b.size else 0.0
    e_32 = float(np.max(np.abs(a - b))) if b.size else 0.0
    return e_gpu, e_32, scale


def p_wave_modulus(state) -> np.ndarray:
    """Per particle the modulus that sets its stress's response to strain: lambda + 2 mu
    of the elastic backbone (snow with its hardening), K gamma for the fluid."""
    w = np.asarray(state.params, np.float64)
    tag = np.asarray(state.tag)
    m = np.zeros(tag.shape[0])
    for t, (il, im) in {0: (0, 1), 1: (0, 1), 2: (7, 8), 3: (7, 8), 5: (4, 5)}.items():
        sel = tag == t
        m[sel] = np.abs(w[sel, il]) + 2 * np.abs(w[sel, im])
    snow = tag == 5
    if snow.any():
        jp = np.asarray(state.pstate, np.float64)[snow, 0]
        m[snow] *= np.exp(np.clip(w[snow, 3] * (1.0 - jp), -50, 50))
    fl = tag == 4
    m[fl] = np.abs(w[fl, 0] * w[fl, 1])
    return m


def c_resolution(state) -> float:
    """FP32 resolution of C over the enabled particles of an (FP64) oracle state:
    eps32 (m |v| 6/h + [F != I in FP32] M |F| V0 (4/h^2) dt), the largest over the
    particles."""
    s = state
    dt, h = float(s.sim[3]), float(s.sim[4])
    on = np.asarray(s.enabled) != 0
    if not on.any():
        return 0.0
    F = np.asarray(s.F, np.float64)
    eye = np.eye(3).reshape(9)
    not_identity = np.abs(F - eye).max(1) > EPS32 / 2
    v = np.abs(np.asarray(s.v, np.float64)).max(1)
    res = EPS32 * (np.asarray(s.mass, np.float64) * v * 6.0 / h
                   + not_identity * p_wave_modulus(s) * np.abs(F).max(1) * np.asarray(s.vol0, np.float64)
                   * 4.0 / h ** 2 * dt)
    return float(res[on].max())


def assert_field_close(name: str, gpu, o32, o64, state=None):
    """Asserts the one-substep bar and returns (|gpu - fp64|, |fp32 - fp64|, scale, clause):
    clause "rel" when the 1e-5 relative bound holds, "fp32" when only the mirror bound
    does, "ulp" when only C's FP32 resolution does (``state``: the FP64 oracle state the
    field belongs to; only C uses it)."""
    e_gpu, e_32, scale = field_errors(gpu, o32, o64)
    rel_bound = REL_TOL * scale
    ulp_bound = ULP_FACTOR * c_resolution(state) if (name == "C" and state is not None) else 0.0
    if e_gpu <= rel_bound:
        clause = "rel"
    elif e_gpu <= FP32_FACTOR * e_32:
        clause = "fp32"
    elif e_gpu <= ulp_bound:
        clause = "ulp"
    else:
        clause = "FAIL"
    test = os.environ.get("PYTEST_CURRENT_TEST", "").split(" ")[0]
    RECORDS.append((test, name, e_gpu, e_32, scale, clause, ulp_bound / ULP_FACTOR if ulp_bound else 0.0))
    bound = max(rel_bound, FP32_FACTOR * e_32, ulp_bound, 1e-30)
    assert e_gpu <= bound, (f"{name}: |gpu - fp64| = {e_gpu:.3e} > {bound:.3e} (fp32 mirror {e_32:.3e}, "
                            f"scale {scale:.3e}, relative {e_gpu / max(scale, 1e-30):.2e}, "
                            f"C resolution bound {ulp_bound:.3e})")
    return e_gpu, e_32, scale, clause


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
    first = np.asarray(b["first"], np.int64)
    cnt = np.asarray(b["np"], np.int64)
    # the primaries of block k are sorted[first[k] : first[k] + np[k]]
    blk_of = np.repeat(np.arange(cnt.size), cnt)
    pos = np.repeat(first, cnt) + (np.arange(blk_of.size) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    vid_of[np.asarray(b["sorted"])[pos]] = np.asarray(b["vid"])[blk_of]
    return vid_of, b


def compare_sort(gpu_can: dict, o32: O.Oracle, x_before: np.ndarray, cell_width: float):
    """Bit-exact comparison of the canonical sort: active block set, per-block
    num_particles, num_particles_with_extras and bucket counts, and per particle its
    primary block and local cell."""
    n = x_before.shape[0]
    vid_of, b = oracle_canonical_per_particle(o32, n)
    ovid, oloc = O.block_keys(np.asarray(x_before, np.float32), cell_width, 32)
    assert np.array_equal(vid_of, ovid)
    # blocks matched by virtual id (vectorized: both lists ordered by vid)
    ov, gv = np.asarray(b["vid"], np.int64), np.asarray(gpu_can["vid"], np.int64)
    assert ov.shape == gv.shape, f"block counts differ: oracle {ov.shape[0]}, gpu {gv.shape[0]}"
    oo = np.lexsort((ov[:, 2], ov[:, 1], ov[:, 0]))
    go = np.lexsort((gv[:, 2], gv[:, 1], gv[:, 0]))
    assert np.array_equal(ov[oo], gv[go]), "block sets differ"
    assert np.array_equal(np.asarray(b["np"])[oo], np.asarray(gpu_can["np"])[go])
    assert np.array_equal(np.asarray(b["npe"])[oo], np.asarray(gpu_can["npe"])[go])
    assert np.array_equal(np.asarray(b["buckets"])[oo], np.asarray(gpu_can["buckets"])[go])
    gvid = gpu_can["vid"][gpu_can["particle_block"]]
    assert np.array_equal(gvid, ovid)
    cell = oloc[:, 0] + 4 * oloc[:, 1] + 16 * oloc[:, 2]
    assert np.array_equal(gpu_can["particle_cell"], cell)
