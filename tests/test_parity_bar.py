"""The one-substep parity bar's third clause for C, its FP32 resolution (tests/parity.py),
checked against the FP32 noise the reference algorithm itself shows: the op-for-op FP32
mirror's C error against the FP64 oracle must lie within the clause's bound, and the bound
must not exceed that noise by more than a small factor (else the clause would let through
errors far above what FP32 arithmetic explains)."""
import numpy as np

from nexus_amd import scenes

from parity import ULP_FACTOR, c_resolution, field_errors, oracle_pair


def _block(n_side, center, h, model, density, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    s = h / 2
    idx = np.stack(np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = np.asarray(center) - n_side * s / 2 + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, idx.shape)
    return scenes.particles(pos, s / 2, density, model)


def _ratio(scene, steps):
    o32, o64 = oracle_pair(scene, steps)
    e32, _, scale = field_errors(o32.s.C, o32.s.C, o64.s.C)
    return e32 / c_resolution(o64.s), scale


def test_free_fall_noise_matches_resolution():
    """Free fall: C's FP64 value is zero up to 1e-19, so its FP32 value is pure rounding
    noise of grad v; the mirror's noise is between 1/8 and 1 of the resolution."""
    h = 1 / 64
    sc = scenes.make_scene(_block(8, (0.5, 0.5, 0.5), h, scenes.model_neo_hookean(1e5, 0.3), 1000.0), h, 1e-4)
    for steps in (2, 5):
        r, scale = _ratio(sc, steps)
        assert scale < 1e-15
        assert 1 / 8 <= r <= 1.0, r


def test_strained_noise_within_resolution():
    """A neo-Hookean block squeezed by 2 % (F != I, stress term nonzero) and a sand
    block after several substeps: the mirror's C error stays within the clause's bound."""
    h = 1 / 64
    parts = _block(8, (0.5, 0.5, 0.5), h, scenes.model_neo_hookean(1e5, 0.3), 1000.0)
    parts["F"][:] = np.array([0.98, 0, 0, 0, 1.01, 0, 0, 0, 1.0], np.float32)
    r, scale = _ratio(scenes.make_scene(parts, h, 1e-4), 3)
    assert scale > 0 and r <= ULP_FACTOR, r
    sand = scenes.sand_column(4000, seed=11, lo=(0.4, 0.1, 0.4), hi=(0.45, 0.2, 0.45), n_colliders=16)
    r, scale = _ratio(sand, 10)
    assert r <= ULP_FACTOR, r
