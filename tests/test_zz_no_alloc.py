"""nxm_pipeline_step_noalloc (one substep of particles held in caller-owned device arrays,
every other buffer in the library's static scratch) against nxm_pipeline_step on the same
scenes: bit for bit on x, v, C, F, mass, plastic state and the enabled flags, on three
calls in a row -- a Drucker-Prager sand column among 16 cuboids, then a smaller mixed
scene (neo-Hookean, fluid and snow particles in a walled box), then the sand column's
second substep from the first call's output.  The caller's buffers here are torch tensors;
the entry point itself allocates nothing."""
import ctypes as ct

import numpy as np
import pytest

from nexus_amd import mpm, scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _materials(sc):
    """Material table and per-particle flags: one material per distinct (tag, parameter
    words, damping, boundary friction); the sand and snow tags' leading plastic-state
    words are per particle, not part of the material."""
    n = sc.n
    words = np.ascontiguousarray(sc.model, np.uint32).reshape(n, 24).copy()
    tag = np.asarray(sc.tag, np.uint32)
    pw = np.where(np.isin(tag, (2, 3)), 3, np.where(tag == 5, 1, 0))
    keyw = words.copy()
    for k in range(3):
        keyw[pw > k, k] = 0
    damping = np.asarray(sc.damping, np.float32)
    bf = (np.ones(n, np.float32) if getattr(sc, "boundary_friction", None) is None
          else np.asarray(sc.boundary_friction, np.float32))
    key = np.concatenate([tag[:, None], keyw, damping.view(np.uint32)[:, None], bf.view(np.uint32)[:, None]], 1)
    uniq, first, cls = np.unique(key, axis=0, return_index=True, return_inverse=True)
    cls = cls.reshape(-1)
    mats = []
    for i in first:
        m = mpm.Material()
        m.model_tag = int(tag[i])
        for k in range(24):
            m.model_data[k] = int(words[i, k])
        m.damping = float(damping[i])
        m.boundary_friction = float(bf[i])
        mats.append(m)
    flags = (np.asarray(sc.enabled, np.uint32) & 1) | ((np.asarray(sc.fixed, np.uint32) & 1) << 1) | \
        (cls.astype(np.uint32) << 2)
    plastic = np.zeros((n, 3), np.float32)
    wf = words.view(np.float32)
    for k in range(3):
        plastic[pw > k, k] = wf[pw > k, k]
    return mats, flags, plastic


class _Device:
    """The caller's device arrays for one scene (torch tensors)."""

    def __init__(self, sc):
        mats, flags, plastic = _materials(sc)
        self.materials = mats
        dev = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
        self.t = dict(x=dev(sc.x), v=dev(sc.v), F=dev(sc.F), C=dev(sc.C), mass=dev(sc.mass), vol0=dev(sc.vol0),
                      plastic=dev(plastic), flags=dev(flags.view(np.int32), torch.int32))
        self.errors = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        self.n = sc.n

    def particles(self):
        t = self.t
        return mpm.DeviceParticles(t["x"].data_ptr(), t["v"].data_ptr(), t["F"].data_ptr(), t["C"].data_ptr(),
                                   t["mass"].data_ptr(), t["vol0"].data_ptr(), t["plastic"].data_ptr(),
                                   t["flags"].data_ptr())

    def read(self):
        out = {k: v.cpu().numpy() for k, v in self.t.items()}
        out["flags"] = out["flags"].view(np.uint32)
        return out


def _step_noalloc(pipeline, sc, d, cap):
    mpm.step_noalloc(pipeline, cell_width=sc.cell_width, grid_capacity=cap, gravity=sc.gravity, dt=sc.dt,
                     materials=d.materials, colliders=sc.colliders, particles=d.particles(), n=d.n,
                     errors=d.errors.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert int(d.errors.item()) == 0


def _compare(got, ref, sc):
    for f in ("x", "v", "C", "F", "mass"):
        assert np.array_equal(got[f], ref[f]), f
    assert np.array_equal(got["flags"] & 1, ref["enabled"])
    tag = np.asarray(sc.tag)
    sand, snow = np.isin(tag, (2, 3)), tag == 5
    ps = ref["model"].view(np.float32)
    assert np.array_equal(got["plastic"][sand], ps[sand, :3])
    assert np.array_equal(got["plastic"][snow, 0], ps[snow, 0])


def _mixed_scene():
    h = 1 / 64
    rng = np.random.Generator(np.random.PCG64(0x0A11))
    parts = []
    for k, (model, rho) in enumerate(((scenes.model_neo_hookean(1e5, 0.3), 1000.0), (scenes.model_fluid(1e4), 1000.0),
                                      (scenes.model_snow(1.4e5, 0.2), 400.0))):
        c = np.array([0.3 + 0.2 * k, 0.4, 0.5])
        pos = c + rng.uniform(-0.04, 0.04, size=(1500, 3))
        parts.append(scenes.particles(pos, h / 4, rho, model, velocity=np.tile([0.0, -0.5, 0.1 * k], (1500, 1))))
    return scenes.make_scene(scenes.concat(parts), h, 1e-4, colliders=scenes.box_walls(h, lo=0.1, hi=0.9))


def test_noalloc_matches_pipeline_step(pipeline):
    sand = scenes.sand_column(20000, seed=11, lo=(0.4, 0.1, 0.4), hi=(0.5, 0.3, 0.5), n_colliders=16)
    mixed = _mixed_scene()
    assert mixed.n < sand.n

    caps = {id(sand): 4096, id(mixed): 1024}

    def reference(sc, steps):
        st = mpm.MpmState.from_scene(pipeline, sc, grid_size=caps[id(sc)])
        for _ in range(steps):
            pipeline.step(st)
        out = st.read_particles()
        st.close()
        return out

    d_sand = _Device(sand)
    _step_noalloc(pipeline, sand, d_sand, caps[id(sand)])
    _compare(d_sand.read(), reference(sand, 1), sand)

    d_mixed = _Device(mixed)
    _step_noalloc(pipeline, mixed, d_mixed, caps[id(mixed)])
    _compare(d_mixed.read(), reference(mixed, 1), mixed)

    _step_noalloc(pipeline, sand, d_sand, caps[id(sand)])
    _compare(d_sand.read(), reference(sand, 2), sand)


def test_noalloc_rejects_what_does_not_fit(pipeline):
    sand = scenes.sand_column(2000, seed=3, lo=(0.4, 0.1, 0.4), hi=(0.5, 0.3, 0.5), n_colliders=0)
    d = _Device(sand)
    with pytest.raises(mpm.MpmError, match="scratch"):
        mpm.step_noalloc(pipeline, cell_width=sand.cell_width, grid_capacity=1 << 16, gravity=sand.gravity,
                         dt=sand.dt, materials=d.materials, colliders=sand.colliders, particles=d.particles(),
                         n=d.n, errors=d.errors.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    with pytest.raises(mpm.MpmError, match="materials"):
        mpm.step_noalloc(pipeline, cell_width=sand.cell_width, grid_capacity=1024, gravity=sand.gravity,
                         dt=sand.dt, materials=[], colliders=sand.colliders, particles=d.particles(), n=d.n)
