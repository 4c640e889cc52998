"""ctypes front end of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  It loads oracle/build/liboracle32.so (FP32 mirror of the reference
arithmetic) or liboracle64.so (FP64) built by oracle/Makefile.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_BUILD selects another build of the same sources (oracle/Makefile's sanitizer
# build, build-asan/).
BUILD = os.path.abspath(os.environ.get("ORACLE_BUILD", os.path.join(HERE, "build")))

_LIBS: dict[int, ct.CDLL] = {}


# Collider records as the C oracle takes them (orc_cuboid = include/nexus_mpm.h's
# nxm_collider): the cuboid record's body fields, then the shape kind and 7 dimension
# words.  Cuboid-only records (the body fields alone) get the cuboid kind.
COLLIDER_DTYPE = np.dtype([
    ("half_extents", "<f4", 3), ("rotation", "<f4", 4), ("translation", "<f4", 3), ("linvel", "<f4", 3),
    ("angvel", "<f4", 3), ("bc_type", "<u4"), ("bc_friction", "<f4"), ("inv_mass", "<f4", 3),
    ("local_com", "<f4", 3), ("shape", "<u4"), ("dims", "<f4", 7)])
BODY_BYTES = 96
SHAPE_CUBOID, SHAPE_TRIMESH = 1, 6


def collider_records(cols, shape: int = SHAPE_CUBOID) -> np.ndarray:
    a = np.ascontiguousarray(cols).reshape(-1)
    if a.dtype.itemsize == COLLIDER_DTYPE.itemsize:
        return a.view(COLLIDER_DTYPE).copy()
    assert a.dtype.itemsize == BODY_BYTES, a.dtype
    out = np.zeros(a.shape[0], COLLIDER_DTYPE)
    out.view(np.uint8).reshape(-1, COLLIDER_DTYPE.itemsize)[:, :BODY_BYTES] = a.view(np.uint8).reshape(-1, BODY_BYTES)
    out["shape"] = shape
    return out


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(bits: int = 32) -> ct.CDLL:
    if bits in _LIBS:
        return _LIBS[bits]
    path = os.path.join(BUILD, f"liboracle{bits}.so")
    if not os.path.exists(path):
        build()
    L = ct.CDLL(path)
    P = ct.c_void_p
    L.orc_ws_create.restype = P
    L.orc_ws_destroy.argtypes = [P]
    L.orc_step.restype = ct.c_int
    L.orc_step.argtypes = [P, ct.c_int64] + [P] * 14 + [ct.c_int, ct.c_int]
    L.orc_step_ex.restype = ct.c_int
    L.orc_step_ex.argtypes = [P, ct.c_int64] + [P] * 14 + [ct.c_int, ct.c_int, P]
    L.orc_ws_grid_cpic.restype = ct.c_int
    L.orc_ws_grid_cpic.argtypes = [P] * 7
    L.orc_node_cdf.restype = ct.c_int
    L.orc_node_cdf.argtypes = [P, ct.c_int, ct.c_float, P, P, P, P]
    L.orc_particle_cdf.restype = ct.c_int
    L.orc_particle_cdf.argtypes = [P, ct.c_float, P, P, ct.c_uint32, P, P, P]
    L.orc_m4_solve.restype = ct.c_double if bits == 64 else ct.c_float
    L.orc_m4_solve.argtypes = [P, P, P]
    L.orc_ws_set_meshes.argtypes = [P, P, ct.c_int]
    L.orc_ws_set_meshes.restype = ct.c_int
    L.orc_project_trimesh.argtypes = [P, P, ct.c_double if bits == 64 else ct.c_float, P, P, P]
    L.orc_project_trimesh.restype = ct.c_int
    L.orc_sample_trimesh.restype = ct.c_int64
    L.orc_sample_trimesh.argtypes = [P, ct.c_int64, P, ct.c_int64, ct.c_float, P, P, ct.c_int64]
    for f in ("orc_ws_nblocks", "orc_ws_nprimary_blocks", "orc_ws_nsorted"):
        getattr(L, f).restype = ct.c_int64
        getattr(L, f).argtypes = [P]
    L.orc_ws_blocks.argtypes = [P] * 6
    L.orc_ws_sorted_ids.argtypes = [P, P]
    L.orc_ws_grid.argtypes = [P] * 4
    L.orc_svd3.argtypes = [P] * 4
    L.orc_model_update.argtypes = [ct.c_uint32] + [P] * 5
    L.orc_model_update.restype = ct.c_int
    L.orc_project_cuboid.argtypes = [P] * 4
    L.orc_grid_node_bc.argtypes = [P, ct.c_int, P, P, P]
    L.orc_block_keys.argtypes = [ct.c_int64, P, ct.c_float, P, P]
    L.orc_timestep_bound.argtypes = [ct.c_int64] + [P] * 8 + [ct.c_float]
    L.orc_timestep_bound.restype = ct.c_uint32
    R = ct.c_double if bits == 64 else ct.c_float
    L.orc_particle_update.restype = ct.c_int
    L.orc_particle_update.argtypes = [ct.c_uint32, P, ct.c_float, ct.c_float, R, ct.c_uint32, R, P, P, R, R, P, P,
                                      P, P, P, P, P, P, P]
    _LIBS[bits] = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def real(bits: int):
    return np.float32 if bits == 32 else np.float64


class _TriMesh(ct.Structure):
    _fields_ = [("vertices", ct.c_void_p), ("triangles", ct.c_void_p), ("samples", ct.c_void_p),
                ("sample_triangle", ct.c_void_p), ("nv", ct.c_int64), ("nt", ct.c_int64), ("ns", ct.c_int64)]


class _Cpic(ct.Structure):
    _fields_ = [("affinity", ct.c_void_p), ("normal", ct.c_void_p), ("signed_distance", ct.c_void_p),
                ("rigid_vel", ct.c_void_p), ("boundary_friction", ct.c_void_p), ("meshes", ct.c_void_p),
                ("nmesh", ct.c_int)]


def sample_trimesh(vertices, triangles, spacing: float):
    """The oracle's sample_mesh (src_mpm/sampling/sample_trimesh.rs:60-231): (points (k, 3)
    f32, triangle of each point (k,) u32)."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    L = lib(32)
    n = L.orc_sample_trimesh(_p(v), v.shape[0], _p(t), t.shape[0], ct.c_float(spacing), None, None, 0)
    pts = np.zeros((max(n, 1), 3), np.float32)
    tri = np.zeros(max(n, 1), np.uint32)
    L.orc_sample_trimesh(_p(v), v.shape[0], _p(t), t.shape[0], ct.c_float(spacing), _p(pts), _p(tri), n)
    return pts[:n], tri[:n]


@dataclass
class OracleState:
    """Particle state held by the oracle in its own precision."""

    bits: int
    x: np.ndarray
    v: np.ndarray
    C: np.ndarray
    F: np.ndarray
    mass: np.ndarray
    vol0: np.ndarray
    damping: np.ndarray
    pstate: np.ndarray
    enabled: np.ndarray
    fixed: np.ndarray
    tag: np.ndarray
    params: np.ndarray  # (n, 24) f32 view of the model words
    sim: np.ndarray     # gx, gy, gz, dt, h (f32)
    colliders: np.ndarray
    cpic: bool = False
    affinity: np.ndarray | None = None          # CPIC per-particle state and outputs
    normal: np.ndarray | None = None
    signed_distance: np.ndarray | None = None
    rigid_vel: np.ndarray | None = None
    boundary_friction: np.ndarray | None = None
    meshes: list | None = None                   # per mesh collider: its arrays (vertices, ...)

    @staticmethod
    def from_scene(scene, bits: int = 32) -> "OracleState":
        R = real(bits)
        # Per-particle plastic state: the sand tags' first three words, snow's first word.
        sand = np.isin(scene.tag, (2, 3))
        snow = np.asarray(scene.tag) == 5
        words = scene.model.view(np.float32).reshape(-1, 24)
        pstate = np.zeros((scene.n, 3), R)
        pstate[sand] = words[sand, 0:3]
        pstate[snow, 0] = words[snow, 0]
        sim = np.array([*scene.gravity, scene.dt, scene.cell_width], np.float32)
        # Mesh colliders (CPIC only): their bodies follow the cuboids in the collider
        # array; their rigid particles are sampled at the cell width
        # (src_mpm/pipeline.rs:310-316).
        # Without CPIC the node pass projects the grid nodes onto the meshes instead
        # (grid_update_collide.rs:53-67), and no rigid particles are sampled.
        meshes = list(getattr(scene, "meshes", None) or [])
        cpic = bool(getattr(scene, "cpic", False))
        cols = collider_records(scene.colliders)
        mesh_arrays = []
        if meshes:
            cols = np.concatenate([cols, collider_records(np.concatenate([m.body for m in meshes]), SHAPE_TRIMESH)])
            for m in meshes:
                v = np.ascontiguousarray(m.vertices, np.float32).reshape(-1, 3)
                t = np.ascontiguousarray(m.triangles, np.uint32).reshape(-1, 3)
                if cpic:
                    pts, tri = sample_trimesh(v, t, float(np.float32(scene.cell_width)))
                else:
                    pts, tri = np.zeros((0, 3), np.float32), np.zeros(0, np.uint32)
                mesh_arrays.append(dict(vertices=v, triangles=t, samples=np.ascontiguousarray(pts),
                                        sample_triangle=np.ascontiguousarray(tri)))
        return OracleState(
            bits=bits,
            x=np.array(scene.x, R), v=np.array(scene.v, R),
            C=np.array(scene.C, R), F=np.array(scene.F, R),
            mass=np.array(scene.mass, R), vol0=np.array(scene.vol0, R),
            damping=np.array(scene.damping, R), pstate=pstate,
            enabled=np.ascontiguousarray(scene.enabled, np.uint32).copy(),
            fixed=np.array(scene.fixed, np.uint32), tag=np.array(scene.tag, np.uint32),
            params=np.array(words, np.float32), sim=sim,
            colliders=cols,
            cpic=bool(getattr(scene, "cpic", False)),
            meshes=mesh_arrays,
            affinity=np.zeros(scene.n, np.uint32), normal=np.zeros((scene.n, 3), R),
            signed_distance=np.zeros(scene.n, R), rigid_vel=np.zeros((scene.n, 3), R),
            boundary_friction=(np.ones(scene.n, R) if getattr(scene, "boundary_friction", None) is None
                               else np.array(scene.boundary_friction, R)),
        )

    @property
    def n(self):
        return self.x.shape[0]


class Oracle:
    """Runs substeps of the reference algorithm on the CPU."""

    def __init__(self, scene, bits: int = 32):
        self.L = lib(bits)
        self.bits = bits
        self.ws = self.L.orc_ws_create()
        self.s = OracleState.from_scene(scene, bits)

    def __del__(self):
        try:
            self.L.orc_ws_destroy(self.ws)
        except Exception:
            pass

    def step(self, count: int = 1, keep_debug: bool = True) -> None:
        s = self.s
        cp = None
        self._meshes = (_TriMesh * max(len(s.meshes), 1))(*[
            _TriMesh(_p(m["vertices"]), _p(m["triangles"]), _p(m["samples"]), _p(m["sample_triangle"]),
                     m["vertices"].shape[0], m["triangles"].shape[0], m["samples"].shape[0]) for m in s.meshes])
        if not s.cpic:
            self.L.orc_ws_set_meshes(self.ws, ct.cast(self._meshes, ct.c_void_p), len(s.meshes))
        if s.cpic:
            cp = _Cpic(_p(s.affinity), _p(s.normal), _p(s.signed_distance), _p(s.rigid_vel),
                       _p(s.boundary_friction), ct.cast(self._meshes, ct.c_void_p), len(s.meshes))
        for _ in range(count):
            rc = self.L.orc_step_ex(
                self.ws, s.n, _p(s.x), _p(s.v), _p(s.C), _p(s.F), _p(s.mass), _p(s.vol0), _p(s.damping),
                _p(s.pstate), _p(s.enabled), _p(s.fixed), _p(s.tag), _p(s.params), _p(s.sim),
                _p(s.colliders) if s.colliders.size else None, int(s.colliders.size), int(keep_debug),
                ct.byref(cp) if cp is not None else None)
            if rc != 0:
                raise RuntimeError(f"oracle step failed: {rc}")

    def step_adaptive(self, dt_max: float) -> float:
        """One substep whose length is min(dt_max, the CFL bound of the current state),
        the bound converted to seconds as the reference's host does (units / 1e12 in f32,
        src_mpm/solver/timestep_bound.rs:45); returns that length."""
        units, secs = timestep_bound(self.s)
        dt = np.float32(dt_max)
        if units != 0xFFFFFFFF and units > 0 and np.float32(secs) < dt:
            dt = np.float32(secs)
        self.s.sim[3] = dt
        self.step(1)
        return float(dt)

    # ---- inspection of the last substep ----
    def blocks(self):
        L, ws = self.L, self.ws
        nb = L.orc_ws_nblocks(ws)
        vid = np.zeros((nb, 3), np.int32)
        np_ = np.zeros(nb, np.uint32)
        npe = np.zeros(nb, np.uint32)
        buckets = np.zeros((nb, 10), np.uint32)
        first = np.zeros(nb, np.uint32)
        L.orc_ws_blocks(ws, _p(vid), _p(np_), _p(npe), _p(buckets), _p(first))
        ids = np.zeros(L.orc_ws_nsorted(ws), np.uint32)
        L.orc_ws_sorted_ids(ws, _p(ids))
        return dict(vid=vid, np=np_, npe=npe, buckets=buckets, first=first, sorted=ids,
                    nprimary=int(L.orc_ws_nprimary_blocks(ws)))

    def grid(self):
        L, ws = self.L, self.ws
        nb = L.orc_ws_nblocks(ws)
        R = real(self.bits)
        mv = np.zeros((nb, 64, 3), R)
        m = np.zeros((nb, 64), R)
        vel = np.zeros((nb, 64, 3), R)
        L.orc_ws_grid(ws, _p(mv), _p(m), _p(vel))
        out = dict(mv=mv, m=m, vel=vel)
        if self.s.cpic:
            dist = np.zeros((nb, 64), R)
            aff = np.zeros((nb, 64), np.uint32)
            closest = np.zeros((nb, 64), np.uint32)
            mv_i = np.zeros((nb, 64, 3), R)
            m_i = np.zeros((nb, 64), R)
            vel_i = np.zeros((nb, 64, 3), R)
            L.orc_ws_grid_cpic(ws, _p(dist), _p(aff), _p(closest), _p(mv_i), _p(m_i), _p(vel_i))
            out.update(cdf_dist=dist, cdf_aff=aff, cdf_closest=closest, mv_i=mv_i, m_i=m_i, vel_i=vel_i)
        return out


def node_cdf(cols: np.ndarray, cell_width: float, pt, bits: int = 64):
    """(distance, affinity word, closest collider or 0xFFFFFFFF) of one grid node."""
    R = real(bits)
    p = np.ascontiguousarray(pt, R)
    d = np.zeros(1, R)
    a = np.zeros(1, np.uint32)
    c = np.zeros(1, np.uint32)
    cols = collider_records(cols)
    rc = lib(bits).orc_node_cdf(_p(cols) if cols.size else None, int(cols.size), ct.c_float(cell_width), _p(p),
                                _p(d), _p(a), _p(c))
    assert rc == 0
    return float(d[0]), int(a[0]), int(c[0])


def particle_cdf(refp, cell_width: float, dist27, aff27, prev: int, bits: int = 64):
    """(affinity word, normal, signed distance) of a particle from its 27 stencil nodes
    (NBH_SHIFTS order; refp = base node - particle)."""
    R = real(bits)
    r = np.ascontiguousarray(refp, R)
    d = np.ascontiguousarray(dist27, R)
    a = np.ascontiguousarray(aff27, np.uint32)
    aff = np.zeros(1, np.uint32)
    n = np.zeros(3, R)
    sd = np.zeros(1, R)
    lib(bits).orc_particle_cdf(_p(r), ct.c_float(cell_width), _p(d), _p(a), ct.c_uint32(prev), _p(aff), _p(n), _p(sd))
    return int(aff[0]), n, float(sd[0])


def m4_solve(m: np.ndarray, b: np.ndarray, bits: int = 64):
    R = real(bits)
    mm = np.ascontiguousarray(m, R).reshape(16)
    bb = np.ascontiguousarray(b, R)
    out = np.zeros(4, R)
    det = lib(bits).orc_m4_solve(_p(mm), _p(bb), _p(out))
    return float(det), out


NBH_SHIFTS = np.array([
    (2, 2, 2), (2, 0, 2), (2, 1, 2), (0, 2, 2), (0, 0, 2), (0, 1, 2), (1, 2, 2), (1, 0, 2), (1, 1, 2),
    (2, 2, 0), (2, 0, 0), (2, 1, 0), (0, 2, 0), (0, 0, 0), (0, 1, 0), (1, 2, 0), (1, 0, 0), (1, 1, 0),
    (2, 2, 1), (2, 0, 1), (2, 1, 1), (0, 2, 1), (0, 0, 1), (0, 1, 1), (1, 2, 1), (1, 0, 1), (1, 1, 1)])


def canonical_sort(blocks: dict) -> dict:
    """Canonical form of a block sort: {vid tuple: (np, npe, bucket counts, per-bucket
    sorted id tuples)}.  Independent of header order and of the order inside buckets."""
    out = {}
    ids = blocks["sorted"]
    for b in range(blocks["vid"].shape[0]):
        cnt = blocks["buckets"][b]
        start = int(blocks["first"][b])
        lists = []
        for k in range(10):
            c = int(cnt[k])
            lists.append(tuple(sorted(int(i) for i in ids[start:start + c])))
            start += c
        out[tuple(int(t) for t in blocks["vid"][b])] = (int(blocks["np"][b]), int(blocks["npe"][b]),
                                                       tuple(int(c) for c in cnt), tuple(lists))
    return out


def svd3(F: np.ndarray, bits: int = 64):
    R = real(bits)
    Fc = np.ascontiguousarray(F.reshape(9), R)
    U = np.zeros(9, R)
    s = np.zeros(3, R)
    V = np.zeros(9, R)
    lib(bits).orc_svd3(_p(Fc), _p(U), _p(s), _p(V))
    return U, s, V


def model_update(tag: int, words: np.ndarray, F: np.ndarray, gradv: np.ndarray, pstate=None, bits: int = 64):
    R = real(bits)
    params = np.ascontiguousarray(np.asarray(words, np.uint32).view(np.float32), np.float32)
    st = np.zeros(3, R) if pstate is None else np.ascontiguousarray(pstate, R).copy()
    Fc = np.ascontiguousarray(F.reshape(9), R).copy()
    g = np.ascontiguousarray(gradv.reshape(9), R)
    tau = np.zeros(9, R)
    rc = lib(bits).orc_model_update(tag, _p(params), _p(st), _p(Fc), _p(g), _p(tau))
    return rc, Fc, st, tau


def project_cuboid(col: np.ndarray, pt, bits: int = 64):
    R = real(bits)
    p = np.ascontiguousarray(pt, R)
    out = np.zeros(3, R)
    inside = ct.c_int(0)
    lib(bits).orc_project_cuboid(_p(collider_records(col)), _p(p), _p(out), ct.byref(inside))
    return out, bool(inside.value)


def project_trimesh(vertices, triangles, body, pt, max_dist: float, bits: int = 64):
    """(world projection or None, inside) of a point onto a posed mesh collider within
    max_dist (body: the mesh's collider record)."""
    R = real(bits)
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    m = _TriMesh(_p(v), _p(t), None, None, v.shape[0], t.shape[0], 0)
    b = collider_records(body, SHAPE_TRIMESH)
    p = np.ascontiguousarray(pt, R)
    out = np.zeros(3, R)
    inside = ct.c_int(0)
    ok = lib(bits).orc_project_trimesh(ct.byref(m), _p(b), max_dist, _p(p), _p(out), ct.byref(inside))
    return (out if ok else None), bool(inside.value)


def grid_node_bc(cols: np.ndarray, sim: np.ndarray, node_pos, vel, bits: int = 64):
    R = real(bits)
    p = np.ascontiguousarray(node_pos, R)
    v = np.ascontiguousarray(vel, R).copy()
    cols = collider_records(cols)
    lib(bits).orc_grid_node_bc(_p(cols) if cols.size else None, int(cols.size),
                               _p(np.ascontiguousarray(sim, np.float32)), _p(p), _p(v))
    return v


def block_keys(x: np.ndarray, cell_width: float, bits: int = 32):
    R = real(bits)
    xc = np.ascontiguousarray(x, R)
    n = xc.shape[0]
    vid = np.zeros((n, 3), np.int32)
    loc = np.zeros((n, 3), np.uint32)
    lib(bits).orc_block_keys(n, _p(xc), ct.c_float(cell_width), _p(vid), _p(loc))
    return vid, loc


def timestep_bound(state: "OracleState") -> tuple[int, float]:
    """CFL timestep bound of the oracle's current particle state: the reference's
    fixed-point value (1e-12 s units) and the seconds it stands for."""
    s = state
    params = np.array(s.params, np.float32)
    snow = s.tag == 5
    params[snow, 0] = s.pstate[snow, 0]          # current plastic determinant
    u = lib(s.bits).orc_timestep_bound(s.n, _p(s.v), _p(s.C), _p(s.F), _p(s.mass), _p(s.vol0), _p(s.enabled),
                                       _p(s.tag), _p(params), ct.c_float(float(s.sim[4])))
    return int(u), float(np.float32(u) / np.float32(1.0e12))


def particle_update(tag: int, words, cell_width: float, dt: float, *, x, gv, gradv=None, div=0.0, F=None, C=None,
                    mass=1.0, vol0=1.0, damping=0.0, fixed=0, pstate=None, psd=0.0, normal=(0, 0, 0),
                    rigid_vel=(0, 0, 0), enabled=1, bits: int = 64):
    """One particle through the oracle's particle update (particle_update.rs:60-201) from
    given G2P outputs (velocity gv, gradient gradv as a 3x3 array, divergence div) and
    contact field.  Returns a dict of the updated x, v, C (3x3), F (3x3), pstate, mass,
    enabled and the model rc."""
    R = real(bits)
    params = np.ascontiguousarray(np.asarray(words, np.uint32).view(np.float32), np.float32)
    cm = lambda M: np.ascontiguousarray(np.asarray(M, R).T.reshape(9))   # column-major
    xx = np.array(x, R)
    vv = np.zeros(3, R)
    CC = cm(np.zeros((3, 3)) if C is None else C)
    FF = cm(np.eye(3) if F is None else F)
    g = cm(np.zeros((3, 3)) if gradv is None else gradv)
    st = np.zeros(3, R) if pstate is None else np.array(pstate, R)
    m = np.array([mass], R)
    en = np.array([enabled], np.uint32)
    rc = lib(bits).orc_particle_update(tag, _p(params), ct.c_float(cell_width), ct.c_float(dt), damping, fixed, vol0,
                                       _p(np.array(gv, R)), _p(g), div, psd, _p(np.array(normal, R)),
                                       _p(np.array(rigid_vel, R)), _p(xx), _p(vv), _p(CC), _p(FF), _p(st), _p(m), _p(en))
    return dict(x=xx, v=vv, C=CC.reshape(3, 3).T, F=FF.reshape(3, 3).T, pstate=st, mass=float(m[0]),
                enabled=int(en[0]), rc=rc)
