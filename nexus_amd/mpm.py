"""Python mirror of the nexus_mpm pipeline API over the C ABI (include/nexus_mpm.h).

Names follow the reference's Rust API so that a caller of nexus_mpm finds the same
objects and calls:

  MpmPipeline(device)                      MpmPipeline::new (src_mpm/pipeline.rs:366)
  MpmState.empty(pipeline, capacities)     MpmState::empty (:107)
  state.set_simulation_params(g, dt)       MpmState::set_simulation_params (:169)
  state.write_substep_params(n)            MpmState::write_substep_params (:183)
  state.set_cell_width(h, capacity)        MpmState::set_cell_width (:200)
  state.add_particles(...)                 GpuParticles::append (src_mpm/solver/particle.rs:533)
  state.set_colliders(records)             MpmState::set_coupling (:220), one-way analytic colliders
  pipeline.step(state)                     MpmPipeline::step (:385)
  state.read_positions()                   GpuParticles::read_positions (particle.rs:634)

The library is the HIP build in nexus_amd/libnexus_mpm.so; there is no CPU
fallback: every call fails loudly if the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

import numpy as np

from . import scenes as _scenes

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnexus_mpm.so")

NXM_OK = 0
STATUS_NAMES = {
    1: "NXM_ERR_INVALID_ARGUMENT", 2: "NXM_ERR_OUT_OF_MEMORY", 3: "NXM_ERR_GRID_CAPACITY", 4: "NXM_ERR_DOMAIN",
    5: "NXM_ERR_TOO_MANY_COLLIDERS", 6: "NXM_ERR_UNSUPPORTED_MODEL", 7: "NXM_ERR_DEVICE", 8: "NXM_ERR_BUSY",
}

# Every symbol include/nexus_mpm.h declares.
EXPORTS = [
    "nxm_ctx_create", "nxm_ctx_destroy", "nxm_last_error", "nxm_version",
    "nxm_scene_create", "nxm_scene_destroy", "nxm_scene_set_simulation_params", "nxm_scene_write_substep_params",
    "nxm_scene_set_cell_width", "nxm_scene_add_particles", "nxm_scene_add_particles_soa", "nxm_scene_set_colliders",
    "nxm_scene_set_shape_colliders",
    "nxm_scene_len", "nxm_scene_device_bytes", "nxm_scene_set_graph", "nxm_scene_timestep_bound",
    "nxm_scene_set_adaptive_dt", "nxm_scene_read_clock", "nxm_scene_swap_remove", "nxm_batch_set_autotune", "nxm_scene_set_autotune",
    "nxm_batch_get_launch_config", "nxm_batch_set_launch_config", "nxm_scene_get_launch_config",
    "nxm_batch_get_autotune", "nxm_scene_get_autotune", "nxm_scene_read_positions", "nxm_scene_read_particles_soa", "nxm_scene_sort_size",
    "nxm_scene_read_sort_canonical", "nxm_scene_read_grid", "nxm_scene_synchronize", "nxm_scene_set_debug_grid",
    "nxm_debug_fail_uploads",
    "nxm_scene_set_cpic", "nxm_scene_read_cdf", "nxm_scene_read_grid_cpic", "nxm_scene_read_colliders",
    "nxm_scene_stage_upload", "nxm_scene_set_trimesh_colliders", "nxm_sample_trimesh",
    "nxm_scene_prep_readback", "nxm_scene_prep_readback_rigid",
    "nxm_pipeline_step", "nxm_pipeline_run", "nxm_pipeline_step_noalloc", "nxm_scene_profile_begin", "nxm_scene_profile_end", "nxm_scene_profile_sampled",
    "nxm_batch_create", "nxm_batch_destroy", "nxm_batch_add_scene", "nxm_batch_remove_scene", "nxm_batch_step",
    "nxm_batch_run", "nxm_batch_synchronize", "nxm_batch_len", "nxm_batch_set_graph",
    "nxm_batch_profile_begin", "nxm_batch_profile_end",
]


class _TriMesh(ct.Structure):
    """nxm_trimesh."""
    _fields_ = [("vertices", ct.c_void_p), ("triangles", ct.c_void_p), ("num_vertices", ct.c_uint32),
                ("num_triangles", ct.c_uint32), ("body", ct.c_ubyte * 96)]


def sample_trimesh(pipeline_or_lib, vertices, triangles, spacing: float):
    """nxm_sample_trimesh: the library's rigid particles of a mesh (points, triangle ids)."""
    L = getattr(pipeline_or_lib, "lib", pipeline_or_lib)
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    n = ct.c_size_t(0)
    _check(L.nxm_sample_trimesh(_ptr(v), v.shape[0], _ptr(t), t.shape[0], ct.c_float(spacing), None, None, 0,
                                ct.byref(n)))
    pts = np.zeros((max(n.value, 1), 3), np.float32)
    tri = np.zeros(max(n.value, 1), np.uint32)
    _check(L.nxm_sample_trimesh(_ptr(v), v.shape[0], _ptr(t), t.shape[0], ct.c_float(spacing), _ptr(pts),
                                _ptr(tri), n.value, ct.byref(n)))
    return pts[:n.value], tri[:n.value]


class MpmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


class _Caps(ct.Structure):
    _fields_ = [("grid_size", ct.c_uint32), ("particles_capacity", ct.c_uint32)]


class _SimParams(ct.Structure):
    _fields_ = [("gravity", ct.c_float * 3), ("dt", ct.c_float)]


class _SoA(ct.Structure):
    _fields_ = [(name, ct.c_void_p) for name in (
        "position", "velocity", "def_grad", "affine", "init_volume", "init_radius", "mass", "damping", "phase",
        "boundary_friction", "group_id", "enabled", "fixed", "model_tag", "model_data")]


class LaunchConfig(ct.Structure):
    """Launch shapes of a substep (nxm_launch_config)."""
    _fields_ = [("p2g_groups", ct.c_uint32), ("p2g_block_cost", ct.c_uint32), ("grid_groups", ct.c_uint32)]

    def as_tuple(self):
        return (int(self.p2g_groups), int(self.p2g_block_cost), int(self.grid_groups))


class _StageTimes(ct.Structure):
    _fields_ = [("sort_ms", ct.c_float), ("p2g_ms", ct.c_float), ("grid_update_ms", ct.c_float),
                ("g2p_update_ms", ct.c_float), ("total_ms", ct.c_float), ("p2g_kernel_ms", ct.c_float),
                ("g2p_update_kernel_ms", ct.c_float)]


class _Canonical(ct.Structure):
    _fields_ = [(name, ct.c_void_p) for name in (
        "block_vid", "block_np", "block_npe", "block_buckets", "particle_block", "particle_cell")]


_LIB = None


def library() -> ct.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m nexus_amd.build`")
    L = ct.CDLL(LIB_PATH)
    P, I = ct.c_void_p, ct.c_int
    L.nxm_last_error.restype = ct.c_char_p
    L.nxm_version.restype = ct.c_char_p
    L.nxm_ctx_create.argtypes = [I, ct.POINTER(P)]
    L.nxm_ctx_destroy.argtypes = [P]
    L.nxm_scene_create.argtypes = [P, P, ct.POINTER(P)]
    L.nxm_scene_destroy.argtypes = [P]
    L.nxm_scene_set_simulation_params.argtypes = [P, P]
    L.nxm_scene_write_substep_params.argtypes = [P, ct.c_uint32]
    L.nxm_scene_set_cell_width.argtypes = [P, ct.c_float, ct.c_uint32]
    L.nxm_scene_add_particles.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_add_particles_soa.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_set_colliders.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_set_shape_colliders.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_set_trimesh_colliders.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_prep_readback.argtypes = [P, ct.c_uint32, P, ct.c_uint32, P, ct.c_size_t]
    L.nxm_scene_prep_readback_rigid.argtypes = [P, P, ct.c_uint32, P, ct.c_size_t, ct.POINTER(ct.c_size_t)]
    L.nxm_sample_trimesh.argtypes = [P, ct.c_uint32, P, ct.c_uint32, ct.c_float, P, P, ct.c_size_t,
                                     ct.POINTER(ct.c_size_t)]
    L.nxm_scene_len.argtypes = [P]
    L.nxm_scene_len.restype = ct.c_size_t
    L.nxm_scene_device_bytes.argtypes = [P]
    L.nxm_scene_device_bytes.restype = ct.c_size_t
    L.nxm_scene_set_graph.argtypes = [P, ct.c_int]
    L.nxm_scene_set_graph.restype = ct.c_int
    L.nxm_scene_timestep_bound.argtypes = [P, ct.POINTER(ct.c_float)]
    L.nxm_scene_timestep_bound.restype = ct.c_int
    L.nxm_scene_swap_remove.argtypes = [P, P, ct.c_size_t, P, P, ct.POINTER(ct.c_size_t)]
    L.nxm_scene_swap_remove.restype = ct.c_int
    for f in ("nxm_batch_set_autotune", "nxm_scene_set_autotune"):
        getattr(L, f).argtypes = [P, ct.c_int]
        getattr(L, f).restype = ct.c_int
    for f in ("nxm_batch_get_autotune", "nxm_scene_get_autotune"):
        getattr(L, f).argtypes = [P, ct.POINTER(ct.c_int), ct.POINTER(ct.c_uint32)]
        getattr(L, f).restype = ct.c_int
    for f in ("nxm_batch_get_launch_config", "nxm_scene_get_launch_config", "nxm_batch_set_launch_config"):
        getattr(L, f).argtypes = [P, ct.POINTER(LaunchConfig)]
        getattr(L, f).restype = ct.c_int
    L.nxm_scene_read_positions.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_read_particles_soa.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_sort_size.argtypes = [P, ct.POINTER(ct.c_size_t)]
    L.nxm_scene_read_sort_canonical.argtypes = [P, P]
    L.nxm_scene_read_grid.argtypes = [P, P, P, P]
    L.nxm_scene_synchronize.argtypes = [P]
    L.nxm_scene_set_debug_grid.argtypes = [P, I]
    L.nxm_debug_fail_uploads.argtypes = [ct.c_uint32]
    L.nxm_scene_set_cpic.argtypes = [P, I]
    L.nxm_pipeline_step_noalloc.argtypes = [P, P, P, ct.c_uint32, P, P]
    L.nxm_scene_set_adaptive_dt.argtypes = [P, ct.c_int]
    L.nxm_scene_set_adaptive_dt.restype = ct.c_int
    L.nxm_scene_read_clock.argtypes = [P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_float), ct.POINTER(ct.c_uint64)]
    L.nxm_scene_read_clock.restype = ct.c_int
    L.nxm_pipeline_step_noalloc.restype = ct.c_int
    L.nxm_scene_stage_upload.argtypes = [P]
    L.nxm_scene_read_colliders.argtypes = [P, P, ct.c_size_t]
    L.nxm_scene_read_cdf.argtypes = [P, P, P, P, ct.c_size_t]
    L.nxm_scene_read_grid_cpic.argtypes = [P, P, P, P, P, P]
    L.nxm_pipeline_step.argtypes = [P, P, P]
    L.nxm_pipeline_run.argtypes = [P, P, ct.c_uint32]
    L.nxm_batch_create.argtypes = [P, ct.POINTER(P)]
    L.nxm_batch_destroy.argtypes = [P]
    L.nxm_batch_add_scene.argtypes = [P, P]
    L.nxm_batch_remove_scene.argtypes = [P, P]
    L.nxm_batch_step.argtypes = [P, P]
    L.nxm_batch_run.argtypes = [P, ct.c_uint32]
    L.nxm_batch_synchronize.argtypes = [P]
    L.nxm_batch_len.argtypes = [P]
    L.nxm_batch_len.restype = ct.c_size_t
    L.nxm_batch_set_graph.argtypes = [P, I]
    for f in ("nxm_scene_profile_begin", "nxm_batch_profile_begin"):
        getattr(L, f).argtypes = [P]
    L.nxm_scene_profile_sampled.argtypes = [P, ct.c_uint32]
    for f in ("nxm_scene_profile_end", "nxm_batch_profile_end"):
        getattr(L, f).argtypes = [P, P, ct.POINTER(ct.c_uint32)]
    _LIB = L
    return L


def debug_fail_uploads(count: int) -> None:
    """Test hook: the next `count` particle uploads into a world fail with NXM_ERR_DEVICE
    before any copy (nxm_debug_fail_uploads); 0 ends it."""
    _check(library().nxm_debug_fail_uploads(int(count)))


def _check(code: int) -> None:
    if code != NXM_OK:
        raise MpmError(code, library().nxm_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


@dataclass
class StageTimes:
    sort_ms: float
    p2g_ms: float
    grid_update_ms: float
    g2p_update_ms: float
    total_ms: float
    p2g_kernel_ms: float = 0.0          # the kernels alone (their own start/end timestamps)
    g2p_update_kernel_ms: float = 0.0

    @staticmethod
    def of(t: "_StageTimes") -> "StageTimes":
        return StageTimes(t.sort_ms, t.p2g_ms, t.grid_update_ms, t.g2p_update_ms, t.total_ms, t.p2g_kernel_ms,
                          t.g2p_update_kernel_ms)


class MpmPipeline:
    """The compiled pipeline on one device (MpmPipeline::new, src_mpm/pipeline.rs:366)."""

    def __init__(self, device: int = 0):
        self.lib = library()
        h = ct.c_void_p()
        _check(self.lib.nxm_ctx_create(device, ct.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.nxm_ctx_destroy(self.handle)
            self.handle = None

    def step(self, state: "MpmState", timings: bool = False):
        """MpmPipeline::step (src_mpm/pipeline.rs:385): enqueues one substep."""
        if timings:
            t = _StageTimes()
            _check(self.lib.nxm_pipeline_step(self.handle, state.handle, ct.byref(t)))
            return StageTimes.of(t)
        _check(self.lib.nxm_pipeline_step(self.handle, state.handle, None))
        return None

    def run(self, state: "MpmState", substeps: int) -> None:
        _check(self.lib.nxm_pipeline_run(self.handle, state.handle, substeps))


class MpmState:
    """Device-resident MPM scene (MpmState, src_mpm/pipeline.rs:65)."""

    def __init__(self, pipeline: MpmPipeline, grid_size: int = 32768, particles_capacity: int = 65536):
        self.lib = pipeline.lib
        self.pipeline = pipeline
        caps = _Caps(grid_size, particles_capacity)
        h = ct.c_void_p()
        _check(self.lib.nxm_scene_create(pipeline.handle, ct.byref(caps), ct.byref(h)))
        self.handle = h

    @classmethod
    def empty(cls, pipeline: MpmPipeline, grid_size: int = 32768, particles_capacity: int = 65536) -> "MpmState":
        return cls(pipeline, grid_size, particles_capacity)

    @classmethod
    def from_scene(cls, pipeline: MpmPipeline, scene: "_scenes.Scene", grid_size: int | None = None) -> "MpmState":
        """Builds a device scene from a host Scene (positions, models, colliders, params)."""
        cap = grid_size if grid_size is not None else default_grid_capacity(scene)
        st = cls(pipeline, cap, scene.n)
        st.set_cell_width(scene.cell_width, cap)
        st.set_simulation_params(scene.gravity, scene.dt)
        if scene.colliders.size:
            st.set_colliders(scene.colliders)
        if getattr(scene, "cpic", False):
            st.set_cpic(True)
        if getattr(scene, "meshes", None):
            st.set_trimesh_colliders(scene.meshes)
        st.add_particles(scene)
        return st

    def read_colliders(self) -> np.ndarray:
        """The colliders' current poses (integrated with their velocities every substep)."""
        n = int(getattr(self, "num_colliders", 0))
        out = np.zeros(n, dtype=_scenes.CUBOID_DTYPE)
        _check(self.lib.nxm_scene_read_colliders(self.handle, _ptr(out) if n else None, n))
        return out

    def stage_upload(self) -> None:
        """Packs the (not yet resident) particles for the next upload, on this thread."""
        _check(self.lib.nxm_scene_stage_upload(self.handle))

    def set_cpic(self, enabled: bool) -> None:
        """MpmState::use_cpic: compatible particle-in-cell transfers against the colliders."""
        _check(self.lib.nxm_scene_set_cpic(self.handle, int(bool(enabled))))

    def read_cdf(self) -> dict:
        """Per-particle contact field of the last step: affinity word, normal, signed distance."""
        n = len(self)
        out = dict(affinity=np.zeros(n, np.uint32), normal=np.zeros((n, 3), np.float32),
                   signed_distance=np.zeros(n, np.float32))
        _check(self.lib.nxm_scene_read_cdf(self.handle, _ptr(out["affinity"]), _ptr(out["normal"]),
                                           _ptr(out["signed_distance"]), n))
        return out

    def read_grid_cpic(self) -> dict:
        """CPIC grid of the last step (read_grid's node order): node contact field and the
        incompatible mass and velocity."""
        nb = self.sort_size()
        out = dict(cdf_dist=np.zeros((nb, 64), np.float32), cdf_aff=np.zeros((nb, 64), np.uint32),
                   cdf_closest=np.zeros((nb, 64), np.uint32), m_i=np.zeros((nb, 64), np.float32),
                   vel_i=np.zeros((nb, 64, 3), np.float32))
        _check(self.lib.nxm_scene_read_grid_cpic(self.handle, _ptr(out["cdf_dist"]), _ptr(out["cdf_aff"]),
                                                 _ptr(out["cdf_closest"]), _ptr(out["m_i"]), _ptr(out["vel_i"])))
        return out

    def close(self):
        if self.handle:
            self.lib.nxm_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return int(self.lib.nxm_scene_len(self.handle))

    def device_bytes(self) -> int:
        """Device memory the scene occupies while resident in a world or batch."""
        return int(self.lib.nxm_scene_device_bytes(self.handle))

    def set_simulation_params(self, gravity, dt: float) -> None:
        p = _SimParams((ct.c_float * 3)(*gravity), dt)
        _check(self.lib.nxm_scene_set_simulation_params(self.handle, ct.byref(p)))

    def write_substep_params(self, num_substeps: int) -> None:
        _check(self.lib.nxm_scene_write_substep_params(self.handle, num_substeps))

    def set_cell_width(self, cell_width: float, grid_capacity: int) -> None:
        _check(self.lib.nxm_scene_set_cell_width(self.handle, cell_width, grid_capacity))

    def set_colliders(self, colliders: np.ndarray) -> None:
        """Analytic colliders: cuboid records (nxm_scene_set_colliders) or shape records of
        any kind (scenes.COLLIDER_DTYPE, nxm_scene_set_shape_colliders)."""
        if np.asarray(colliders).dtype == _scenes.COLLIDER_DTYPE:
            c = np.ascontiguousarray(colliders, dtype=_scenes.COLLIDER_DTYPE).reshape(-1)
            _check(self.lib.nxm_scene_set_shape_colliders(self.handle, _ptr(c) if c.size else None, c.size))
        else:
            c = np.ascontiguousarray(colliders, dtype=_scenes.CUBOID_DTYPE).reshape(-1)
            _check(self.lib.nxm_scene_set_colliders(self.handle, _ptr(c) if c.size else None, c.size))
        self.num_cuboids = int(c.size)
        self.num_colliders = self.num_cuboids + int(getattr(self, "num_meshes", 0))

    def prep_readback(self, mode: int = 0, group_colors=None) -> dict:
        """WgPrepReadback: per particle (id order) render colour (n, 4), deformation (n, 4, 4)
        column-major, and position (n, 3)."""
        n = len(self)
        out = np.zeros((max(n, 1), 24), np.float32)
        pal = None if group_colors is None else np.ascontiguousarray(group_colors, np.float32).reshape(-1, 4)
        _check(self.lib.nxm_scene_prep_readback(self.handle, int(mode), _ptr(pal) if pal is not None else None,
                                                0 if pal is None else pal.shape[0], _ptr(out), n))
        out = out[:n]
        return {"color": out[:, 0:4], "deformation": out[:, 4:20].reshape(n, 4, 4).transpose(0, 2, 1),
                "position": out[:, 20:23]}

    def prep_readback_rigid(self, group_colors=None) -> dict:
        """The mesh colliders' rigid particles as render data (see prep_readback)."""
        pal = None if group_colors is None else np.ascontiguousarray(group_colors, np.float32).reshape(-1, 4)
        cnt = ct.c_size_t(0)
        _check(self.lib.nxm_scene_prep_readback_rigid(self.handle, _ptr(pal) if pal is not None else None,
                                                      0 if pal is None else pal.shape[0], None, 0, ct.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 24), np.float32)
        _check(self.lib.nxm_scene_prep_readback_rigid(self.handle, _ptr(pal) if pal is not None else None,
                                                      0 if pal is None else pal.shape[0], _ptr(out), cnt.value,
                                                      ct.byref(cnt)))
        out = out[:cnt.value]
        k = cnt.value
        return {"color": out[:, 0:4], "deformation": out[:, 4:20].reshape(k, 4, 4).transpose(0, 2, 1),
                "position": out[:, 20:23]}

    def set_trimesh_colliders(self, meshes) -> None:
        """Mesh colliders (scenes.TriMesh) of a CPIC scene; their ids follow the cuboids'."""
        keep = []
        arr = (_TriMesh * max(len(meshes), 1))()
        for k, m in enumerate(meshes):
            v = np.ascontiguousarray(m.vertices, np.float32).reshape(-1, 3)
            t = np.ascontiguousarray(m.triangles, np.uint32).reshape(-1, 3)
            b = np.ascontiguousarray(m.body, dtype=_scenes.CUBOID_DTYPE).reshape(1)
            keep += [v, t]
            arr[k].vertices = v.ctypes.data
            arr[k].triangles = t.ctypes.data
            arr[k].num_vertices = v.shape[0]
            arr[k].num_triangles = t.shape[0]
            ct.memmove(ct.addressof(arr[k].body), b.ctypes.data, _scenes.CUBOID_DTYPE.itemsize)
        _check(self.lib.nxm_scene_set_trimesh_colliders(self.handle, ct.cast(arr, ct.c_void_p) if meshes else None,
                                                           len(meshes)))
        self.num_meshes = len(meshes)
        self.num_colliders = int(getattr(self, "num_cuboids", 0)) + self.num_meshes

    def set_debug_grid(self, enabled: bool = True) -> None:
        _check(self.lib.nxm_scene_set_debug_grid(self.handle, int(enabled)))

    def set_graph(self, enabled: bool = True) -> None:
        """Standalone steps replay one captured HIP graph per substep."""
        _check(self.lib.nxm_scene_set_graph(self.handle, int(enabled)))

    def swap_remove(self, ids) -> list[tuple[int, int]]:
        """GpuParticles::swap_remove: removes particles by id; the last particle moves
        into each freed id.  Returns the (from, to) moves."""
        a = np.ascontiguousarray(ids, np.uint32)
        f = np.zeros(max(a.size, 1), np.uint32)
        t = np.zeros(max(a.size, 1), np.uint32)
        k = ct.c_size_t(0)
        _check(self.lib.nxm_scene_swap_remove(self.handle, _ptr(a) if a.size else None, a.size, _ptr(f), _ptr(t),
                                              ct.byref(k)))
        return [(int(f[i]), int(t[i])) for i in range(k.value)]

    def set_autotune(self, enabled: bool = True) -> None:
        """The next standalone substeps try launch shapes and keep the fastest."""
        _check(self.lib.nxm_scene_set_autotune(self.handle, int(enabled)))

    def launch_config(self) -> tuple:
        c = LaunchConfig()
        _check(self.lib.nxm_scene_get_launch_config(self.handle, ct.byref(c)))
        return c.as_tuple()

    def autotune_state(self) -> tuple[bool, int]:
        """(tuning, substeps): whether the tuner is still trying launch shapes, and the
        substeps its current or last run has used."""
        t, k = ct.c_int(0), ct.c_uint32(0)
        _check(self.lib.nxm_scene_get_autotune(self.handle, ct.byref(t), ct.byref(k)))
        return bool(t.value), int(k.value)

    def set_adaptive_dt(self, enabled: bool = True) -> None:
        """Each substep's length becomes min(dt, the CFL bound of the state at its start)."""
        _check(self.lib.nxm_scene_set_adaptive_dt(self.handle, int(bool(enabled))))

    def read_clock(self) -> tuple[float, float, int]:
        """(simulated time, last substep's length, substeps) of this scene."""
        t, dt, k = ct.c_double(0), ct.c_float(0), ct.c_uint64(0)
        _check(self.lib.nxm_scene_read_clock(self.handle, ct.byref(t), ct.byref(dt), ct.byref(k)))
        return float(t.value), float(dt.value), int(k.value)

    def timestep_bound(self) -> float:
        """CFL bound on the substep length of the current state (WgTimestepBounds)."""
        dt = ct.c_float(0.0)
        _check(self.lib.nxm_scene_timestep_bound(self.handle, ct.byref(dt)))
        return float(dt.value)

    def add_particles(self, scene: "_scenes.Scene") -> None:
        arrs = dict(
            position=np.ascontiguousarray(scene.x, np.float32), velocity=np.ascontiguousarray(scene.v, np.float32),
            def_grad=np.ascontiguousarray(scene.F, np.float32), affine=np.ascontiguousarray(scene.C, np.float32),
            init_volume=np.ascontiguousarray(scene.vol0, np.float32),
            init_radius=np.ascontiguousarray(scene.radius, np.float32),
            mass=np.ascontiguousarray(scene.mass, np.float32), damping=np.ascontiguousarray(scene.damping, np.float32),
            enabled=np.ascontiguousarray(scene.enabled, np.uint32), fixed=np.ascontiguousarray(scene.fixed, np.uint32),
            model_tag=np.ascontiguousarray(scene.tag, np.uint32),
            model_data=np.ascontiguousarray(scene.model, np.uint32),
        )
        bf = getattr(scene, "boundary_friction", None)
        if bf is not None:
            arrs["boundary_friction"] = np.ascontiguousarray(bf, np.float32)
        if getattr(scene, "group_id", None) is not None:
            arrs["group_id"] = np.ascontiguousarray(scene.group_id, np.uint32)
        if getattr(scene, "phase", None) is not None:
            arrs["phase"] = np.ascontiguousarray(scene.phase, np.float32)
        soa = _SoA(**{k: _ptr(v) for k, v in arrs.items()})
        _check(self.lib.nxm_scene_add_particles_soa(self.handle, ct.byref(soa), scene.n))

    def synchronize(self) -> None:
        _check(self.lib.nxm_scene_synchronize(self.handle))

    def profile_begin(self, every: int = 1) -> None:
        """Start recording per-stage timestamps (no synchronization), on one step in `every`."""
        if every == 1:
            _check(self.lib.nxm_scene_profile_begin(self.handle))
        else:
            _check(self.lib.nxm_scene_profile_sampled(self.handle, int(every)))

    def profile_end(self):
        """Synchronize; return (per-step average StageTimes, number of steps)."""
        t = _StageTimes()
        n = ct.c_uint32(0)
        _check(self.lib.nxm_scene_profile_end(self.handle, ct.byref(t), ct.byref(n)))
        return StageTimes.of(t), int(n.value)

    def read_positions(self) -> np.ndarray:
        n = len(self)
        out = np.zeros((n, 3), np.float32)
        _check(self.lib.nxm_scene_read_positions(self.handle, _ptr(out), n))
        return out

    def read_particles(self) -> dict:
        n = len(self)
        out = dict(
            x=np.zeros((n, 3), np.float32), v=np.zeros((n, 3), np.float32), F=np.zeros((n, 9), np.float32),
            C=np.zeros((n, 9), np.float32), vol0=np.zeros(n, np.float32), radius=np.zeros(n, np.float32),
            mass=np.zeros(n, np.float32), damping=np.zeros(n, np.float32), phase=np.zeros(n, np.float32),
            boundary_friction=np.zeros(n, np.float32), group=np.zeros(n, np.uint32),
            enabled=np.zeros(n, np.uint32), fixed=np.zeros(n, np.uint32), tag=np.zeros(n, np.uint32),
            model=np.zeros((n, 24), np.uint32),
        )
        soa = _SoA(position=_ptr(out["x"]), velocity=_ptr(out["v"]), def_grad=_ptr(out["F"]), affine=_ptr(out["C"]),
                   init_volume=_ptr(out["vol0"]), init_radius=_ptr(out["radius"]), mass=_ptr(out["mass"]),
                   damping=_ptr(out["damping"]), phase=_ptr(out["phase"]),
                   boundary_friction=_ptr(out["boundary_friction"]), group_id=_ptr(out["group"]),
                   enabled=_ptr(out["enabled"]), fixed=_ptr(out["fixed"]), model_tag=_ptr(out["tag"]),
                   model_data=_ptr(out["model"]))
        _check(self.lib.nxm_scene_read_particles_soa(self.handle, ct.byref(soa), n))
        return out

    def sort_size(self) -> int:
        nb = ct.c_size_t(0)
        _check(self.lib.nxm_scene_sort_size(self.handle, ct.byref(nb)))
        return int(nb.value)

    def read_sort_canonical(self) -> dict:
        nb = self.sort_size()
        n = len(self)
        out = dict(vid=np.zeros((nb, 3), np.int32), np=np.zeros(nb, np.uint32), npe=np.zeros(nb, np.uint32),
                   buckets=np.zeros((nb, 10), np.uint32), particle_block=np.zeros(n, np.uint32),
                   particle_cell=np.zeros(n, np.uint32))
        c = _Canonical(block_vid=_ptr(out["vid"]), block_np=_ptr(out["np"]), block_npe=_ptr(out["npe"]),
                       block_buckets=_ptr(out["buckets"]), particle_block=_ptr(out["particle_block"]),
                       particle_cell=_ptr(out["particle_cell"]))
        _check(self.lib.nxm_scene_read_sort_canonical(self.handle, ct.byref(c)))
        return out

    def read_grid(self, momentum: bool = False) -> dict:
        nb = self.sort_size()
        out = dict(m=np.zeros((nb, 64), np.float32), vel=np.zeros((nb, 64, 3), np.float32))
        mom = np.zeros((nb, 64, 3), np.float32) if momentum else None
        _check(self.lib.nxm_scene_read_grid(self.handle, _ptr(out["m"]), _ptr(mom), _ptr(out["vel"])))
        if momentum:
            out["mv"] = mom
        return out


class Batch:
    """Many independent scenes stepped together, one launch per stage."""

    def __init__(self, pipeline: MpmPipeline):
        self.lib = pipeline.lib
        self.pipeline = pipeline
        h = ct.c_void_p()
        _check(self.lib.nxm_batch_create(pipeline.handle, ct.byref(h)))
        self.handle = h
        self.members: list[MpmState] = []

    def close(self):
        if self.handle:
            self.lib.nxm_batch_destroy(self.handle)
            self.handle = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return int(self.lib.nxm_batch_len(self.handle))

    def add(self, state: MpmState) -> None:
        _check(self.lib.nxm_batch_add_scene(self.handle, state.handle))
        self.members.append(state)

    def remove(self, state: MpmState) -> None:
        """Takes the scene out of the batch; its state comes back to the host."""
        _check(self.lib.nxm_batch_remove_scene(self.handle, state.handle))
        self.members.remove(state)

    def discard(self, state: MpmState) -> None:
        """Destroys a member scene without reading its state back."""
        self.members.remove(state)
        state.close()

    def step(self, timings: bool = False):
        if timings:
            t = _StageTimes()
            _check(self.lib.nxm_batch_step(self.handle, ct.byref(t)))
            return StageTimes.of(t)
        _check(self.lib.nxm_batch_step(self.handle, None))
        return None

    def run(self, substeps: int) -> None:
        _check(self.lib.nxm_batch_run(self.handle, substeps))

    def synchronize(self) -> None:
        _check(self.lib.nxm_batch_synchronize(self.handle))

    def set_graph(self, enabled: bool) -> None:
        _check(self.lib.nxm_batch_set_graph(self.handle, int(enabled)))

    def set_autotune(self, enabled: bool = True) -> None:
        """The next substeps try launch shapes (coordinate descent) and keep the fastest."""
        _check(self.lib.nxm_batch_set_autotune(self.handle, int(enabled)))

    def launch_config(self) -> tuple:
        c = LaunchConfig()
        _check(self.lib.nxm_batch_get_launch_config(self.handle, ct.byref(c)))
        return c.as_tuple()

    def autotune_state(self) -> tuple[bool, int]:
        """(tuning, substeps) of the batch's launch-shape tuner."""
        t, k = ct.c_int(0), ct.c_uint32(0)
        _check(self.lib.nxm_batch_get_autotune(self.handle, ct.byref(t), ct.byref(k)))
        return bool(t.value), int(k.value)

    def set_launch_config(self, p2g_groups: int, p2g_block_cost: int, grid_groups: int) -> None:
        c = LaunchConfig(p2g_groups, p2g_block_cost, grid_groups)
        _check(self.lib.nxm_batch_set_launch_config(self.handle, ct.byref(c)))

    def profile_begin(self) -> None:
        _check(self.lib.nxm_batch_profile_begin(self.handle))

    def profile_end(self):
        t = _StageTimes()
        n = ct.c_uint32(0)
        _check(self.lib.nxm_batch_profile_end(self.handle, ct.byref(t), ct.byref(n)))
        return StageTimes.of(t), int(n.value)


def default_grid_capacity(scene: "_scenes.Scene") -> int:
    """A block capacity covering the scene's occupied blocks with room for motion: the
    blocks of the particle bounding box (plus a one-block halo), times two."""
    if scene.n == 0:
        return 1024
    lo = np.floor(scene.x.min(0) / scene.cell_width / 4) - 2
    hi = np.floor(scene.x.max(0) / scene.cell_width / 4) + 2
    vol = float(np.prod(hi - lo + 1))
    est = min(vol, 8.0 * scene.n + 64)
    cap = 1024
    while cap < 2 * est:
        cap *= 2
    return min(cap, 1 << 26)


# ---- one substep without allocation (nxm_pipeline_step_noalloc) ----

class Material(ct.Structure):
    """nxm_material: model tag and words, damping and boundary friction factor."""
    _fields_ = [("model_tag", ct.c_uint32), ("model_data", ct.c_uint32 * 24), ("damping", ct.c_float),
                ("boundary_friction", ct.c_float)]


class DeviceParticles(ct.Structure):
    """nxm_device_particles: device addresses (integers) of the caller's arrays."""
    _fields_ = [(name, ct.c_void_p) for name in (
        "position", "velocity", "def_grad", "affine", "mass", "init_volume", "plastic", "flags")]


class _NoallocScene(ct.Structure):
    _fields_ = [("cell_width", ct.c_float), ("grid_capacity", ct.c_uint32), ("params", _SimParams),
                ("materials", ct.c_void_p), ("num_materials", ct.c_uint32), ("colliders", ct.c_void_p),
                ("num_colliders", ct.c_uint32)]


def step_noalloc(pipeline: MpmPipeline, *, cell_width: float, grid_capacity: int, gravity, dt: float,
                 materials, colliders, particles: DeviceParticles, n: int, errors: int | None = None,
                 stream: int | None = None) -> None:
    """One substep of particles held in caller-owned device arrays, with no allocation
    (nxm_pipeline_step_noalloc).  ``materials``: a sequence of Material; ``colliders``: a
    CUBOID_DTYPE array; ``errors`` / ``stream``: a device address / a hipStream_t (ints)."""
    lib = pipeline.lib
    mats = (Material * max(len(materials), 1))(*materials)
    cols = np.ascontiguousarray(colliders, dtype=_scenes.CUBOID_DTYPE).reshape(-1)
    sc = _NoallocScene(ct.c_float(cell_width), grid_capacity, _SimParams((ct.c_float * 3)(*gravity), dt),
                       ct.cast(mats, ct.c_void_p), len(materials), _ptr(cols) if cols.size else None, cols.size)
    _check(lib.nxm_pipeline_step_noalloc(pipeline.handle, ct.byref(sc), ct.byref(particles), n, errors, stream))
