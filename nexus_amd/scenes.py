"""Seeded scene generators for the MLS-MPM workloads.

A scene is a ``Scene``: structure-of-arrays particle data in the layout the C ABI
uploads (``include/nexus_mpm.h``), the simulation parameters and the colliders.

Particle construction follows the reference's ``Particle::new`` /
``ParticleDynamics::new`` (src_mpm/solver/particle.rs:37-43, 102-122): initial
volume (2r)^3, mass = density * volume, F = I, affine = 0, enabled, not fixed.
Model words follow ``From<ParticleModel> for GpuParticleModel``
(src_mpm/solver/particle_model.rs:282-325): a tag plus 24 u32 words.

Workloads (seeds are 0x4E584D00 + workload number; per-scene streams derive from
the workload seed and the scene index):

  1  neo-Hookean cube, 32^3 jittered lattice, h = 1/64, six stick walls, dt = 1e-4
  2  Drucker-Prager sand column, 1,048,576 particles, h = 1/256, 16 cuboids, dt = 5e-5
  3  weakly-compressible fluid, 8,388,608 particles in 64 Gaussian clumps, h = 1/512, dt = 2e-5
  4  256 independent scenes, 2k-200k particles each, mixed materials, h in [1/128, 1/64]
  5  arrivals of scenes as in 4 with 1k-500k particles and 50-500 substep horizons
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MODEL_ELASTIC_LINEAR = 0
MODEL_ELASTIC_NEO_HOOKEAN = 1
MODEL_SAND_LINEAR = 2
MODEL_SAND_NEO_HOOKEAN = 3
MODEL_FLUID = 4
MODEL_SNOW = 5

BC_STICK = 0
BC_SLIP = 1
BC_SEPARATE = 2
BC_NON_REFLECTING = 3

SEED_BASE = 0x4E584D00

# nxm_cuboid layout: half_extents[3], rotation[4] (x, y, z, w), translation[3],
# linvel[3], angvel[3], bc_type (u32), bc_friction (f32), inv_mass[3] (zero unless the
# body is dynamic and coupled two ways), local_com[3] (centre of mass, body frame).
CUBOID_DTYPE = np.dtype(
    [
        ("half_extents", "<f4", 3),
        ("rotation", "<f4", 4),
        ("translation", "<f4", 3),
        ("linvel", "<f4", 3),
        ("angvel", "<f4", 3),
        ("bc_type", "<u4"),
        ("bc_friction", "<f4"),
        ("inv_mass", "<f4", 3),
        ("local_com", "<f4", 3),
    ]
)


def f32(x):
    return np.float32(x)


def lame_lambda_mu(young: float, poisson: float):
    """Lame parameters in f32 arithmetic (src_mpm/models/mod.rs:19-30)."""
    e, nu = f32(young), f32(poisson)
    lam = e * nu / ((f32(1) + nu) * (f32(1) - f32(2) * nu))
    mu = e / (f32(2) * (f32(1) + nu))
    return f32(lam), f32(mu)


def _words(values) -> np.ndarray:
    w = np.zeros(24, dtype=np.uint32)
    arr = np.asarray(values, dtype=np.float32)
    w[: arr.size] = arr.view(np.uint32)
    return w


def model_neo_hookean(young: float, poisson: float):
    """ParticleModel::elastic_neo_hookean (src_mpm/solver/particle_model.rs:59-64)."""
    lam, mu = lame_lambda_mu(young, poisson)
    return MODEL_ELASTIC_NEO_HOOKEAN, _words([lam, mu, 0.5])


def model_linear(young: float, poisson: float):
    """ParticleModel::elastic (src_mpm/solver/particle_model.rs:52-57)."""
    lam, mu = lame_lambda_mu(young, poisson)
    return MODEL_ELASTIC_LINEAR, _words([lam, mu, 0.5])


def model_sand(young: float, poisson: float, friction_deg: float | None = None, *, neo_hookean=False):
    """Drucker-Prager sand (src_mpm/solver/particle_model.rs:67-77, 143-153).

    With ``friction_deg`` the hardening law is flattened to a constant angle
    (ha = angle, hb = hd = 0), which the reference allows because the
    DruckerPragerPlasticity fields are public; without it the reference's default
    hardening (35, 9, 0.2, 10 degrees; src_mpm/models/drucker_prager.rs:46-56) is used.
    """
    lam, mu = lame_lambda_mu(young, poisson)
    if friction_deg is None:
        ha, hb, hc, hd = (np.float32(np.deg2rad(35.0)), np.float32(np.deg2rad(9.0)),
                          f32(0.2), np.float32(np.deg2rad(10.0)))
    else:
        ha, hb, hc, hd = np.float32(np.deg2rad(friction_deg)), f32(0), f32(0.2), f32(0)
    tag = MODEL_SAND_NEO_HOOKEAN if neo_hookean else MODEL_SAND_LINEAR
    # plastic state (det Fp, hardening q, log volume gain) = (1, 1, 0)
    return tag, _words([1.0, 1.0, 0.0, ha, hb, hc, hd, lam, mu, 0.0, lam, mu, 0.5])


def model_snow(young: float, poisson: float, critical_compression: float = 2.5e-2,
               critical_stretch: float = 7.5e-3, hardening: float = 10.0):
    """ParticleModel::snow / snow_with_params (src_mpm/solver/particle_model.rs:218-251;
    defaults :45-49): plastic determinant 1, the yield box, the hardening coefficient and
    the linear elastic backbone."""
    lam, mu = lame_lambda_mu(young, poisson)
    return MODEL_SNOW, _words([1.0, critical_compression, critical_stretch, hardening, lam, mu, 0.5])


def model_fluid(bulk: float, gamma: float = 7.0, viscosity: float = 0.001, tensile: float = 0.25):
    """ParticleModel::fluid / water (src_mpm/solver/particle_model.rs:161-178)."""
    return MODEL_FLUID, _words([bulk, gamma, viscosity, 0.5, tensile])


@dataclass
class Scene:
    """Host copy of one MPM scene (structure of arrays, float32/uint32)."""

    x: np.ndarray            # (n, 3) f32 positions
    v: np.ndarray            # (n, 3) f32 velocities
    F: np.ndarray            # (n, 9) f32 deformation gradient, column-major
    C: np.ndarray            # (n, 9) f32 APIC affine matrix, column-major
    mass: np.ndarray         # (n,) f32
    vol0: np.ndarray         # (n,) f32 initial volume
    radius: np.ndarray       # (n,) f32 initial radius
    damping: np.ndarray      # (n,) f32
    enabled: np.ndarray      # (n,) u32
    fixed: np.ndarray        # (n,) u32
    tag: np.ndarray          # (n,) u32 model tag
    model: np.ndarray        # (n, 24) u32 model words
    cell_width: float
    dt: float
    gravity: tuple = (0.0, -9.81, 0.0)
    colliders: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=CUBOID_DTYPE))
    substeps: int = 1
    name: str = ""
    cpic: bool = False                       # MpmState::use_cpic (src_mpm/pipeline.rs:69)
    boundary_friction: np.ndarray | None = None   # (n,) f32, None = 1 (Particle default)
    meshes: list = field(default_factory=list)    # TriMesh colliders (CPIC), ids after the cuboids

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "Scene":
        out = Scene(**{k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                       for k, v in self.__dict__.items()})
        out.meshes = list(self.meshes)
        return out


def particles(positions, radius, density, model, *, velocity=None) -> dict:
    """Arrays for particles built like Particle::new (src_mpm/solver/particle.rs:37-43)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    n = pos.shape[0]
    r = np.broadcast_to(np.asarray(radius, dtype=np.float32), (n,)).copy()
    vol = (r * f32(2.0)) ** 3
    vol = vol.astype(np.float32)
    mass = (vol * np.float32(density)).astype(np.float32)
    eye = np.zeros((n, 9), dtype=np.float32)
    eye[:, 0] = eye[:, 4] = eye[:, 8] = 1.0
    tag, words = model
    vel = np.zeros((n, 3), np.float32) if velocity is None else np.ascontiguousarray(velocity, np.float32).reshape(n, 3)
    return dict(
        x=pos, v=vel, F=eye, C=np.zeros((n, 9), np.float32), mass=mass, vol0=vol, radius=r,
        damping=np.zeros(n, np.float32), enabled=np.ones(n, np.uint32), fixed=np.zeros(n, np.uint32),
        tag=np.full(n, tag, np.uint32), model=np.tile(words, (n, 1)),
    )


def concat(parts: list[dict]) -> dict:
    return {k: np.concatenate([p[k] for p in parts], axis=0) for k in parts[0]}

# Analytic collider records (nxm_collider): the cuboid record's body fields, then the
# shape kind (the reference's Shape tags, src_rbd_shaders/shapes/shape.rs:29-33) and its
# dimensions in the body frame.
SHAPE_BALL, SHAPE_CUBOID, SHAPE_CAPSULE, SHAPE_CONE, SHAPE_CYLINDER = 0, 1, 2, 3, 4
COLLIDER_DTYPE = np.dtype(CUBOID_DTYPE.descr + [("shape", "<u4"), ("dims", "<f4", 7)])


def as_colliders(*records) -> np.ndarray:
    """Collider records of any kind (cuboid() or shape records) as one COLLIDER_DTYPE
    array, in the given order (the collider ids)."""
    out = []
    for r in records:
        r = np.asarray(r).reshape(-1)
        if r.dtype == COLLIDER_DTYPE:
            out.append(r)
            continue
        c = np.zeros(r.shape[0], COLLIDER_DTYPE)
        for f in CUBOID_DTYPE.names:
            c[f] = r[f]
        c["shape"] = SHAPE_CUBOID
        out.append(c)
    return np.concatenate(out) if out else np.zeros(0, COLLIDER_DTYPE)


def _shape(kind, dims, center, rotation, bc, friction, linvel, angvel, inv_mass, local_com):
    c = as_colliders(cuboid(center, (0.0, 0.0, 0.0), rotation, bc, friction, linvel, angvel, inv_mass, local_com))
    c["shape"] = kind
    d = np.zeros(7, np.float32)
    d[:len(dims)] = dims
    c["dims"] = d
    return c


def ball(center, radius, rotation=(0.0, 0.0, 0.0, 1.0), bc=BC_STICK, friction=0.0, linvel=(0, 0, 0),
         angvel=(0, 0, 0), inv_mass=(0, 0, 0), local_com=(0, 0, 0)):
    """A ball collider (rapier Ball)."""
    return _shape(SHAPE_BALL, (radius,), center, rotation, bc, friction, linvel, angvel, inv_mass, local_com)


def capsule(center, a, b, radius, rotation=(0.0, 0.0, 0.0, 1.0), bc=BC_STICK, friction=0.0, linvel=(0, 0, 0),
            angvel=(0, 0, 0), inv_mass=(0, 0, 0), local_com=(0, 0, 0)):
    """A capsule collider: segment a-b (body frame) swept by radius (parry Capsule)."""
    return _shape(SHAPE_CAPSULE, (*a, *b, radius), center, rotation, bc, friction, linvel, angvel, inv_mass,
                  local_com)


def cone(center, half_height, radius, rotation=(0.0, 0.0, 0.0, 1.0), bc=BC_STICK, friction=0.0, linvel=(0, 0, 0),
         angvel=(0, 0, 0), inv_mass=(0, 0, 0), local_com=(0, 0, 0)):
    """A cone collider along the body's y axis, apex up (parry Cone)."""
    return _shape(SHAPE_CONE, (half_height, radius), center, rotation, bc, friction, linvel, angvel, inv_mass,
                  local_com)


def cylinder(center, half_height, radius, rotation=(0.0, 0.0, 0.0, 1.0), bc=BC_STICK, friction=0.0,
             linvel=(0, 0, 0), angvel=(0, 0, 0), inv_mass=(0, 0, 0), local_com=(0, 0, 0)):
    """A cylinder collider along the body's y axis (parry Cylinder)."""
    return _shape(SHAPE_CYLINDER, (half_height, radius), center, rotation, bc, friction, linvel, angvel, inv_mass,
                  local_com)


def cuboid(center, half_extents, rotation=(0.0, 0.0, 0.0, 1.0), bc=BC_STICK, friction=0.0,
           linvel=(0, 0, 0), angvel=(0, 0, 0), inv_mass=(0, 0, 0), local_com=(0, 0, 0)):
    """One collider record.  inv_mass / local_com describe a dynamic body coupled two
    ways (gravity acts on it along the axes with a nonzero inverse mass, and it turns
    about its centre of mass); the defaults are a body coupled one way."""
    c = np.zeros(1, dtype=CUBOID_DTYPE)
    c["half_extents"] = half_extents
    c["rotation"] = rotation
    c["translation"] = center
    c["linvel"] = linvel
    c["angvel"] = angvel
    c["bc_type"] = bc
    c["bc_friction"] = friction
    c["inv_mass"] = inv_mass
    c["local_com"] = local_com
    return c


@dataclass
class TriMesh:
    """A triangle-mesh collider: vertices (nv, 3) and triangles (nt, 3) in the body frame,
    and its body as one CUBOID_DTYPE record (pose, velocities, boundary condition, mass
    properties; half extents unused)."""

    vertices: np.ndarray
    triangles: np.ndarray
    body: np.ndarray


def trimesh(vertices, triangles, center, rotation=(0.0, 0.0, 0.0, 1.0), bc=BC_STICK, friction=0.0,
            linvel=(0, 0, 0), angvel=(0, 0, 0), inv_mass=(0, 0, 0), local_com=(0, 0, 0)) -> TriMesh:
    body = cuboid(center, (0.0, 0.0, 0.0), rotation, bc, friction, linvel, angvel, inv_mass, local_com)
    return TriMesh(np.ascontiguousarray(vertices, np.float32).reshape(-1, 3),
                   np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3), body)


def box_mesh(half_extents):
    """A closed box as 8 vertices and 12 triangles wound counter-clockwise seen from
    outside (so ab x ac points out of the box)."""
    hx, hy, hz = (float(e) for e in half_extents)
    v = np.array([[sx * hx, sy * hy, sz * hz] for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)], np.float32)
    # vertex index = (sx > 0) + 2 (sy > 0) + 4 (sz > 0)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = []
    for a, b, c, d in quads:
        t += [(a, b, c), (a, c, d)]
    return v, np.array(t, np.uint32)


def ramp_mesh(half_x, half_z, rise, nx=4, nz=4):
    """An open, tessellated ramp y = rise (x / half_x) over [-half_x, half_x] x [-half_z,
    half_z], triangles wound so that their normals point up (+y side)."""
    xs = np.linspace(-half_x, half_x, nx + 1, dtype=np.float64)
    zs = np.linspace(-half_z, half_z, nz + 1, dtype=np.float64)
    v = np.array([[x, rise * x / half_x, z] for z in zs for x in xs], np.float32)
    t = []
    for j in range(nz):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c, d = a + 1, a + nx + 2, a + nx + 1
            t += [(a, d, c), (a, c, b)]
    return v, np.array(t, np.uint32)


def box_walls(h: float, bc=BC_STICK, friction=0.0, lo=0.0, hi=1.0) -> np.ndarray:
    """Six cuboid slabs of thickness 4h enclosing [lo, hi]^3 (the reference has no
    built-in walls: its examples build them from fixed cuboids, e.g.
    crates/examples3d/mpm_sand3.rs:56-72)."""
    t = 2.0 * h
    mid = 0.5 * (lo + hi)
    ext = 0.5 * (hi - lo) + 2.0 * t
    walls = []
    for axis in range(3):
        for side in (lo - t, hi + t):
            c = [mid, mid, mid]
            c[axis] = side
            he = [ext, ext, ext]
            he[axis] = t
            walls.append(cuboid(c, he, bc=bc, friction=friction))
    return np.concatenate(walls)


def random_unit_quaternion(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return q.astype(np.float32)


def make_scene(parts: dict, cell_width, dt, *, colliders=None, gravity=(0.0, -9.81, 0.0), substeps=1, name="",
               cpic=False):
    cols = np.zeros(0, dtype=CUBOID_DTYPE) if colliders is None else colliders
    return Scene(cell_width=float(np.float32(cell_width)), dt=float(np.float32(dt)), gravity=gravity,
                 colliders=cols, substeps=substeps, name=name, cpic=cpic, **parts)


# --------------------------------------------------------------------------- #
# Workloads


def elastic_cube(n_side: int = 32, seed: int = SEED_BASE + 1, *, cell_width=1.0 / 64, side=0.25,
                 center=(0.5, 0.6, 0.5)) -> Scene:
    """Workload 1: neo-Hookean cube (E = 1e5, nu = 0.3, density 1000) on an n_side^3
    lattice with +-spacing/4 uniform jitter, six stick walls, dt = 1e-4, 100 substeps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = side / n_side
    idx = np.stack(np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij"), -1)
    idx = idx.reshape(-1, 3).astype(np.float64)
    jitter = rng.uniform(-0.25 * s, 0.25 * s, size=idx.shape)
    pos = np.asarray(center) - side / 2 + (idx + 0.5) * s + jitter
    parts = particles(pos, s / 2, 1000.0, model_neo_hookean(1e5, 0.3))
    return make_scene(parts, cell_width, 1e-4, colliders=box_walls(cell_width), substeps=100,
                      name=f"neo-Hookean cube {n_side ** 3}")


def sand_column(n: int = 1_048_576, seed: int = SEED_BASE + 2, *, cell_width=1.0 / 256,
                n_colliders: int = 16, lo=(0.35, 0.05, 0.35), hi=(0.65, 0.65, 0.65)) -> Scene:
    """Workload 2: Drucker-Prager sand (friction 30 degrees, E = 3.5e5, nu = 0.3,
    density 2200), uniform random positions in a 0.3 x 0.6 x 0.3 m box, 16 fixed
    cuboids with random poses and half-extents 0.02-0.1 m (separate, friction 1),
    h = 1/256, dt = 5e-5, 1000 substeps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = np.asarray(lo), np.asarray(hi)
    pos = lo + rng.random((n, 3)) * (hi - lo)
    vol = float(np.prod(hi - lo))
    r = 0.5 * (vol / n) ** (1.0 / 3.0)
    parts = particles(pos, r, 2200.0, model_sand(3.5e5, 0.3, friction_deg=30.0))
    # colliders from their own stream: the same 16 boxes for any particle count
    rng = np.random.Generator(np.random.PCG64([seed, 1]))
    cols = []
    for _ in range(n_colliders):
        q = random_unit_quaternion(rng)
        c = rng.random(3)
        he = rng.uniform(0.02, 0.1, size=3)
        cols.append(cuboid(c, he, rotation=q, bc=BC_SEPARATE, friction=1.0))
    colliders = np.concatenate(cols) if cols else None
    return make_scene(parts, cell_width, 5e-5, colliders=colliders, substeps=1000,
                      name=f"Drucker-Prager sand column {n}")


def scaled(scene: Scene, k: float, *, cpic: bool | None = None, name: str | None = None) -> Scene:
    """The same scene k times larger: positions, radii, colliders and the cell width scale
    by k, volumes and masses by k^3, dt by k (the same Courant number)."""
    out = scene.copy()
    f = np.float32(k)
    out.x = (out.x * f).astype(np.float32)
    out.radius = (out.radius * f).astype(np.float32)
    out.vol0 = (out.vol0 * f ** 3).astype(np.float32)
    out.mass = (out.mass * f ** 3).astype(np.float32)
    cols = out.colliders.copy()
    if cols.size:
        cols["translation"] = cols["translation"] * f
        cols["half_extents"] = cols["half_extents"] * f
        if cols.dtype == COLLIDER_DTYPE:
            cols["dims"] = cols["dims"] * f
    out.colliders = cols
    out.cell_width = float(np.float32(scene.cell_width * k))
    out.dt = float(np.float32(scene.dt * k))
    if cpic is not None:
        out.cpic = cpic
    if name is not None:
        out.name = name
    return out


def sand_column_cpic(n: int = 1_048_576, seed: int = SEED_BASE + 2) -> Scene:
    """The sand column 64 times larger (h = 0.25, dt = 3.2e-3, the same Courant number)
    with CPIC: the contact distance field needs cell widths of a few tenths (its
    least-squares matrix is singular below, g2p_cdf.rs:309)."""
    return scaled(sand_column(n, seed), 64.0, cpic=True, name=f"Drucker-Prager sand column {n}, x64, CPIC")


def sand_on_mesh_ramp(n: int = 6000, seed: int = SEED_BASE + 4, *, angle=0.2) -> Scene:
    """Drucker-Prager sand (h = 0.25, dt = 3.2e-3, CPIC) dropped onto a tessellated ramp
    mesh (60 triangles, turned about y by `angle`, boundary condition separate with
    friction 0.4) that stands on a slip floor cuboid: a mesh collider and a cuboid in one
    contact field.  The ramp's edges are kept off the grid's node lines."""
    h = 0.25
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.array([4.03, 3.31, 4.02]) + rng.random((n, 3)) * np.array([2.0, 1.2, 2.0])
    parts = particles(x, h / 4, 2200.0, model_sand(3.5e5, 0.3, friction_deg=30.0))
    q = np.array([0.0, math.sin(angle / 2), 0.0, math.cos(angle / 2)], np.float32)
    rv, rt = ramp_mesh(3.0, 3.0, 0.6, nx=6, nz=5)
    ramp = trimesh(rv, rt, (5.013, 2.4137, 4.987), rotation=q, bc=BC_SEPARATE, friction=0.4)
    floor = cuboid((5.0, 0.5013, 5.0), (10.0, 0.5, 10.0), bc=BC_SLIP)
    sc = make_scene(parts, h, 3.2e-3, colliders=floor, cpic=True, name=f"sand on a mesh ramp {n}")
    sc.meshes = [ramp]
    return sc


def fluid_clumps(n: int = 8_388_608, seed: int = SEED_BASE + 3, *, n_clumps: int = 64,
                 cell_width=1.0 / 512, sigma=(0.01, 0.05)) -> Scene:
    """Workload 3: weakly-compressible fluid (K = 1e4, gamma = 7, density 1000),
    positions from 64 isotropic Gaussian clumps (sigma 0.01-0.05 m, centres in
    [0.1, 0.9]^3, samples outside the unit box redrawn), h = 1/512, r = h/4,
    dt = 2e-5, 200 substeps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per = n // n_clumps
    out = []
    for k in range(n_clumps):
        m = per if k < n_clumps - 1 else n - per * (n_clumps - 1)
        c = rng.uniform(0.1, 0.9, size=3)
        s = rng.uniform(sigma[0], sigma[1])
        pts = np.empty((0, 3))
        while pts.shape[0] < m:
            cand = c + rng.normal(scale=s, size=(m - pts.shape[0], 3))
            ok = np.all((cand > 0.0) & (cand < 1.0), axis=1)
            pts = np.concatenate([pts, cand[ok]])
        out.append(pts[:m])
    pos = np.concatenate(out)
    parts = particles(pos, cell_width / 4, 1000.0, model_fluid(1e4))
    return make_scene(parts, cell_width, 2e-5, substeps=200, name=f"fluid clumps {n}")


def material_box(rng, n: int, material: int, cell_width: float, dt: float = 1e-4) -> Scene:
    """A jittered box of ``n`` particles (2 per cell per axis) at a random position
    in the unit domain, for the batched and served workloads.  material: 0 neo-Hookean
    (C1 parameters), 1 Drucker-Prager (C2 parameters), 2 fluid (C3 parameters)."""
    s = cell_width / 2
    side = int(math.ceil(n ** (1.0 / 3.0)))
    idx = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"), -1)
    idx = idx.reshape(-1, 3)[:n].astype(np.float64)
    extent = side * s
    origin = rng.uniform(0.05, max(0.05, 0.95 - extent), size=3)
    pos = origin + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, size=idx.shape)
    if material == 0:
        parts = particles(pos, s / 2, 1000.0, model_neo_hookean(1e5, 0.3))
    elif material == 1:
        parts = particles(pos, s / 2, 2200.0, model_sand(3.5e5, 0.3, friction_deg=30.0))
    else:
        parts = particles(pos, s / 2, 1000.0, model_fluid(1e4))
    return make_scene(parts, cell_width, dt, name=f"box {n} mat {material}")


def batch_scenes(count: int = 256, seed: int = SEED_BASE + 4, *, n_range=(2_000, 200_000),
                 substeps: int = 500) -> list[Scene]:
    """Workload 4: ``count`` independent scenes; N log-uniform in n_range, material
    uniform over {neo-Hookean, Drucker-Prager, fluid}, h = 1/k with k uniform in
    [64, 128], dt = 1e-4."""
    out = []
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64([seed, i]))
        n = int(math.floor(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        mat = int(rng.integers(0, 3))
        h = 1.0 / int(rng.integers(64, 129))
        sc = material_box(rng, n, mat, h)
        sc.substeps = substeps
        sc.name = f"scene {i}: {n} particles, material {mat}, h=1/{round(1 / h)}"
        out.append(sc)
    return out


def served_requests(duration_s: float, seed: int = SEED_BASE + 5, *, rate=200.0, n_range=(1_000, 500_000),
                    horizon=(50, 500)) -> list["RequestSpec"]:
    """Workload 5: Poisson arrivals (rate per second) over duration_s of scenes as in
    workload 4, with N log-uniform in n_range and step horizons uniform in horizon.
    Returns lightweight specs in arrival order; ``spec.build()`` makes the Scene."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        i = len(out)
        srng = np.random.Generator(np.random.PCG64([seed, i]))
        n = int(math.floor(math.exp(srng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        mat = int(srng.integers(0, 3))
        k = int(srng.integers(64, 129))
        steps = int(srng.integers(horizon[0], horizon[1] + 1))
        out.append(RequestSpec(t, i, seed, n, mat, k, steps))
    return out


class RequestSpec:
    """One served request: arrival time, scene parameters and horizon (substeps)."""

    __slots__ = ("arrival", "index", "seed", "n", "material", "inv_h", "horizon")

    def __init__(self, arrival, index, seed, n, material, inv_h, horizon):
        self.arrival, self.index, self.seed = float(arrival), int(index), int(seed)
        self.n, self.material, self.inv_h, self.horizon = int(n), int(material), int(inv_h), int(horizon)

    def build(self) -> Scene:
        rng = np.random.Generator(np.random.PCG64([self.seed, self.index, 1]))
        sc = material_box(rng, self.n, self.material, 1.0 / self.inv_h)
        sc.substeps = self.horizon
        sc.name = f"request {self.index}: {self.n} particles, material {self.material}, h=1/{self.inv_h}"
        return sc
