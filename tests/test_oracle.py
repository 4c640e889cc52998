"""Pins of the CPU oracle (oracle/) against facts independent of its own formulas.

The reference holds no vectors for this path (no MPM tests exist in dimforge/nexus),
so each function of the oracle is checked against integer brute force, a library
routine (LAPACK SVD), an energy derivative, a geometric brute force, or a closed-form
property of the whole transfer (APIC affine exactness, conservation, free fall).
"""
import math

import numpy as np
import pytest

import oracle as O
from nexus_amd import scenes

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------------------- keys

def round_half_away(q: float) -> int:
    return int(math.floor(abs(q) + 0.5)) * (1 if q >= 0 else -1)


@pytest.mark.parametrize("bits", [32, 64])
def test_block_keys_brute_force_with_ties(bits):
    """assoc_cell = round(x/h) - 1 (half away from zero), block = floor(assoc/4),
    local = assoc - 4 block (src_mpm_shaders/solver/particle.rs:208-225).  h is a
    power of two so x/h is exact; ties at k + 0.5 and negative coordinates included."""
    h = 1.0 / 64
    qs = [k / 2 for k in range(-40, 41)] + [k / 4 + 0.0625 for k in range(-40, 41)]
    x = np.array([[q * h, -q * h, (q + 0.5) * h] for q in qs], np.float64)
    vid, loc = O.block_keys(x, h, bits)
    for i, p in enumerate(x):
        for d in range(3):
            cell = round_half_away(p[d] / h) - 1
            blk = cell // 4
            assert vid[i, d] == blk, (p[d] / h, vid[i, d], blk)
            assert loc[i, d] == cell - 4 * blk


def test_sort_invariants_on_random_scene():
    """Active blocks = P U (P + {0,1}^3); each particle sits in 1 + #spill blocks; bucket
    counts add up; primaries precede extras (src_mpm_shaders/grid/sort.rs:287-491)."""
    sc = scenes.sand_column(3000, seed=7, lo=(0.3, 0.3, 0.3), hi=(0.36, 0.4, 0.35), n_colliders=0)
    o = O.Oracle(sc, 32)
    o.step(1)
    b = o.blocks()
    vid, loc = O.block_keys(sc.x, sc.cell_width, 32)
    P = {tuple(v) for v in vid}
    expect = set()
    for p in P:
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    expect.add((p[0] + dx, p[1] + dy, p[2] + dz))
    got = {tuple(map(int, v)) for v in b["vid"]}
    assert got == expect
    assert b["nprimary"] == len(P)
    counts = np.zeros(sc.n, int)
    for blk in range(b["vid"].shape[0]):
        ids = b["sorted"][b["first"][blk]: b["first"][blk] + b["npe"][blk]]
        np.add.at(counts, ids, 1)
        assert b["buckets"][blk].sum() == b["npe"][blk]
        assert b["buckets"][blk][:4].sum() == b["np"][blk]
        prim = b["sorted"][b["first"][blk]: b["first"][blk] + b["np"][blk]]
        assert all(tuple(vid[i]) == tuple(b["vid"][blk]) for i in prim)
    spills = np.prod(1 + (loc >= 2), axis=1)
    assert np.array_equal(counts, spills)


# ---------------------------------------------------------------------------- SVD

@pytest.mark.parametrize("kind", ["random", "near_identity", "repeated", "rank_deficient"])
def test_svd_against_lapack(kind):
    for _ in range(50):
        if kind == "random":
            F = RNG.normal(size=(3, 3))
        elif kind == "near_identity":
            F = np.eye(3) + 1e-4 * RNG.normal(size=(3, 3))
        elif kind == "repeated":
            Q, _ = np.linalg.qr(RNG.normal(size=(3, 3)))
            F = Q @ np.diag([1.5, 1.5, 0.7]) @ np.linalg.qr(RNG.normal(size=(3, 3)))[0]
        else:
            F = np.outer(RNG.normal(size=3), RNG.normal(size=3))
        U, s, V = O.svd3(F.T.reshape(9), 64)  # column-major
        U = U.reshape(3, 3).T
        V = V.reshape(3, 3).T
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-12 * max(1, np.abs(F).max()))
        assert np.allclose(U.T @ U, np.eye(3), atol=1e-12)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.sort(np.abs(s)), np.sort(ref), atol=1e-12 * max(1, ref.max()))
        if np.linalg.det(F) > 1e-12:
            assert np.all(s > 0)


# ---------------------------------------------------------------------------- models

def colmajor(M):
    return np.asarray(M, np.float64).T.reshape(9)


def from_colmajor(a):
    return np.asarray(a).reshape(3, 3).T


def numeric_kirchhoff(energy, F, eps=1e-6):
    """tau = dPsi/dF F^T by central differences of the strain energy."""
    P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            dF = np.zeros((3, 3))
            dF[i, j] = eps
            P[i, j] = (energy(F + dF) - energy(F - dF)) / (2 * eps)
    return P @ F.T


def random_F(scale=0.1):
    F = np.eye(3) + scale * RNG.normal(size=(3, 3))
    if np.linalg.det(F) < 0:
        F[:, 0] *= -1
    return F


def test_neo_hookean_stress_is_energy_derivative():
    """NeoHookeanModel (models/neo_hookean_elasticity.rs:26-39) against
    Psi = mu/2 (tr F^T F - 3) - mu ln J + lam/2 (ln J)^2."""
    tag, words = scenes.model_neo_hookean(1e5, 0.3)
    lam, mu = words[:2].view(np.float32).astype(np.float64)
    for _ in range(10):
        F = random_F()
        rc, Fo, _, tau = O.model_update(tag, words, colmajor(F), np.zeros(9))
        assert rc == 0
        J = np.linalg.det(F)
        psi = lambda G: mu / 2 * (np.trace(G.T @ G) - 3) - mu * np.log(np.linalg.det(G)) + lam / 2 * np.log(np.linalg.det(G)) ** 2
        assert np.allclose(from_colmajor(tau), numeric_kirchhoff(psi, F), rtol=1e-6, atol=1e-3)
        assert np.allclose(from_colmajor(Fo), F)
        assert J > 0


def test_corotated_stress_is_energy_derivative():
    """LinearElasticModel (models/linear_elasticity.rs:30-54) against
    Psi = mu |F - R|^2 + lam/2 (J - 1)^2 with R the polar rotation of F."""
    tag, words = scenes.model_linear(1e5, 0.3)
    lam, mu = words[:2].view(np.float32).astype(np.float64)

    def psi(G):
        U, s, Vt = np.linalg.svd(G)
        R = U @ Vt
        return mu * np.sum((G - R) ** 2) + lam / 2 * (np.linalg.det(G) - 1) ** 2

    for _ in range(10):
        F = random_F()
        rc, _, _, tau = O.model_update(tag, words, colmajor(F), np.zeros(9))
        assert rc == 0
        assert np.allclose(from_colmajor(tau), numeric_kirchhoff(psi, F), rtol=1e-5, atol=1e-2)


def test_fluid_pressure_tait_and_viscosity():
    """FluidModel (models/fluid.rs:54-84): p = K ((1/J)^gamma - 1) for J <= 1 (Tait,
    evaluated here with a power), -K ts (J - 1) above; tau = -p J I + 2 eta J dev(sym grad v)."""
    tag, words = scenes.model_fluid(1e4)
    K, gamma, eta, _, ts = words[:5].view(np.float32).astype(np.float64)
    for J in (0.8, 0.95, 1.0, 1.05, 1.3):
        d = J ** (1 / 3)
        g = 0.5 * RNG.normal(size=(3, 3))
        rc, _, _, tau = O.model_update(tag, words, colmajor(np.eye(3) * d), colmajor(g))
        p = K * ((1 / J) ** gamma - 1) if J <= 1 else -K * ts * (J - 1)
        sr = 0.5 * (g + g.T)
        dev = sr - np.trace(sr) / 3 * np.eye(3)
        assert np.allclose(from_colmajor(tau), -p * J * np.eye(3) + 2 * eta * J * dev, rtol=1e-9, atol=1e-9)


def test_snow_return_map():
    """SnowPlasticity::project (models/snow.rs:83-113): singular values inside the box
    [1 - theta_c, 1 + theta_s] leave F and det Fp alone; outside, they are clamped to the
    box (F = U diag(clamp s) V^T), det Fp takes the removed volume change (clamped to
    [0.1, 4]), and the stress is the corotated stress with lambda and mu scaled by
    exp(xi (1 - det Fp))."""
    tag, words = scenes.model_snow(1.4e5, 0.2)
    p = words.view(np.float32).astype(np.float64)
    tc, ts, xi, lam, mu = p[1], p[2], p[3], p[4], p[5]
    Q1 = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
    Q2 = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
    Q1 *= np.sign(np.linalg.det(Q1))
    Q2 *= np.sign(np.linalg.det(Q2))

    def corotated(F, l, m):
        U, s, Vt = np.linalg.svd(F)
        J = np.prod(s)
        return 2 * m * (F - U @ Vt) @ F.T + l * (J - 1) * J * np.eye(3)

    # inside the box: elastic
    s = np.array([1.004, 0.99, 0.98])
    F = Q1 @ np.diag(s) @ Q2.T
    rc, Fo, st, tau = O.model_update(tag, words, colmajor(F), np.zeros(9), pstate=[0.9, 0, 0])
    assert rc == 0 and np.allclose(from_colmajor(Fo), F, atol=1e-13) and np.isclose(st[0], 0.9, rtol=1e-14)
    hard = np.exp(xi * (1 - 0.9))
    assert np.allclose(from_colmajor(tau), corotated(F, lam * hard, mu * hard), rtol=1e-9, atol=1e-6)
    # stretched and compressed beyond the box
    s = np.array([1.05, 0.9, 0.99])
    F = Q1 @ np.diag(s) @ Q2.T
    rc, Fo, st, tau = O.model_update(tag, words, colmajor(F), np.zeros(9), pstate=[1.0, 0, 0])
    c = np.clip(s, 1 - tc, 1 + ts)
    assert np.allclose(from_colmajor(Fo), Q1 @ np.diag(c) @ Q2.T, atol=1e-12)
    pdet = np.prod(s) / np.prod(c)
    assert np.isclose(st[0], pdet, rtol=1e-12)
    hard = np.exp(xi * (1 - pdet))
    assert np.allclose(from_colmajor(tau), corotated(Q1 @ np.diag(c) @ Q2.T, lam * hard, mu * hard),
                       rtol=1e-9, atol=1e-6)
    # the plastic determinant is kept in [0.1, 4]
    s = np.array([0.5, 0.5, 0.5])
    rc, Fo, st, _ = O.model_update(tag, words, colmajor(Q1 @ np.diag(s) @ Q2.T), np.zeros(9), pstate=[0.2, 0, 0])
    assert st[0] == 0.1
    s = np.array([1.6, 1.6, 1.6])
    rc, Fo, st, _ = O.model_update(tag, words, colmajor(Q1 @ np.diag(s) @ Q2.T), np.zeros(9), pstate=[3.0, 0, 0])
    assert st[0] == 4.0
    # no yield box: purely elastic, hardening from the current state
    t2, w2 = scenes.model_snow(1.4e5, 0.2, critical_compression=0.0, critical_stretch=0.0)
    F = Q1 @ np.diag([1.2, 0.7, 1.0]) @ Q2.T
    rc, Fo, st, tau = O.model_update(t2, w2, colmajor(F), np.zeros(9), pstate=[0.5, 0, 0])
    assert np.allclose(from_colmajor(Fo), F, atol=1e-14) and st[0] == 0.5
    hard = np.exp(xi * 0.5)
    assert np.allclose(from_colmajor(tau), corotated(F, lam * hard, mu * hard), rtol=1e-9, atol=1e-5)


def dp_setup(friction=30.0):
    tag, words = scenes.model_sand(3.5e5, 0.3, friction_deg=friction)
    p = words.view(np.float32).astype(np.float64)
    return tag, words, dict(lam=p[7], mu=p[8], coh=p[9], ha=p[3])


def dp_alpha(ha):
    s = math.sin(ha)
    return math.sqrt(2 / 3) * 2 * s / (3 - s)


def test_drucker_prager_return_map():
    """DruckerPragerPlasticity::project (models/drucker_prager.rs:224-267): a state
    inside the cone is left alone; tension returns to the apex (F = R, singular values
    exp(cohesion/3) = 1); a state outside is projected onto the yield surface
    |dev e'| + (3 lam + 2 mu)/(2 mu) (tr e' - c) alpha = 0 with tr e' = tr e and
    q increased by the plastic multiplier."""
    tag, words, P = dp_setup()
    alpha = dp_alpha(P["ha"])
    k = (3 * P["lam"] + 2 * P["mu"]) / (2 * P["mu"])
    Q1 = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
    Q2 = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
    Q1 *= np.sign(np.linalg.det(Q1))
    Q2 *= np.sign(np.linalg.det(Q2))
    # elastic: compressive with small deviator
    s = np.exp(np.array([-0.010, -0.011, -0.012]))
    F = Q1 @ np.diag(s) @ Q2.T
    rc, Fo, st, _ = O.model_update(tag, words, colmajor(F), np.zeros(9), pstate=[1, 1, 0])
    assert np.allclose(from_colmajor(Fo), F, atol=1e-14) and np.allclose(st, [1, 1, 0])
    # tension
    s = np.exp(np.array([0.01, 0.02, -0.005]))
    F = Q1 @ np.diag(s) @ Q2.T
    rc, Fo, st, _ = O.model_update(tag, words, colmajor(F), np.zeros(9), pstate=[1, 1, 0])
    assert np.allclose(from_colmajor(Fo), Q1 @ Q2.T, atol=1e-12)
    assert np.isclose(st[1], 1 + np.linalg.norm(np.log(s)), rtol=1e-12)
    # plastic: compressive with large deviator
    e = np.array([0.05, -0.03, -0.04])
    assert np.linalg.norm(e - e.sum() / 3) + k * e.sum() * alpha > 0.01   # outside the cone
    F = Q1 @ np.diag(np.exp(e)) @ Q2.T
    rc, Fo, st, _ = O.model_update(tag, words, colmajor(F), np.zeros(9), pstate=[1, 1, 0])
    Fo = from_colmajor(Fo)
    so = np.linalg.svd(Fo, compute_uv=False)
    eo = np.log(so)
    dev = eo - eo.sum() / 3
    assert np.isclose(eo.sum(), e.sum(), atol=1e-12)
    assert abs(np.linalg.norm(dev) + k * (eo.sum() - P["coh"]) * alpha) < 1e-12
    gamma = np.linalg.norm(e - e.sum() / 3) + k * e.sum() * alpha
    assert np.isclose(st[1], 1 + gamma, rtol=1e-10)
    assert np.isclose(st[0], 1.0, rtol=1e-10) and abs(st[2]) < 1e-12
    # same principal directions: Fo = Q1 diag(.) Q2^T
    assert np.allclose(Q1.T @ Fo @ Q2, np.diag(np.diag(Q1.T @ Fo @ Q2)), atol=1e-12)


# ---------------------------------------------------------------------------- colliders

def box_boundary_closest(he, p):
    """Brute force: closest point over the six faces (clamped rectangles)."""
    best, bp = None, None
    for axis in range(3):
        for sgn in (-1, 1):
            q = np.clip(p, -he, he)
            q[axis] = sgn * he[axis]
            d = np.linalg.norm(q - p)
            if best is None or d < best - 1e-15:
                best, bp = d, q
    return best, bp


def test_cuboid_projection_brute_force():
    for _ in range(300):
        he = RNG.uniform(0.1, 1.0, size=3)
        q = RNG.normal(size=4)
        q /= np.linalg.norm(q)
        t = RNG.normal(size=3)
        col = scenes.cuboid(t, he, rotation=q)
        c = col[0]
        x, y, z, w = q.astype(np.float32).astype(np.float64)
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        local = RNG.uniform(-1.5, 1.5, size=3) * he
        p = R @ local + c["translation"].astype(np.float64)
        proj, inside = O.project_cuboid(col, p, 64)
        he64 = c["half_extents"].astype(np.float64)
        assert inside == bool(np.all(np.abs(local) <= he64))
        d, _ = box_boundary_closest(he64, local)
        assert np.isclose(np.linalg.norm(proj - p), d, atol=1e-6)


def test_boundary_conditions_on_a_floor():
    """Node inside a floor slab (grid_update_collide.rs:30-183 + boundary_condition.rs:77-114):
    stick -> 0; slip -> tangential part; separate(mu) -> tangent shrunk by mu * v_n."""
    h, dt = 1.0 / 64, float(np.float32(1e-4))
    sim = np.array([0, -9.81, 0, dt, h], np.float32)
    node = np.array([0.5, -0.25 * h, 0.5])
    vel = np.array([0.3, -0.4, 0.0])
    for bc, fr, expect in ((scenes.BC_STICK, 0.0, [0, 0, 0]), (scenes.BC_SLIP, 0.0, [0.3, 0, 0]),
                           (scenes.BC_SEPARATE, 0.5, [0.3 - 0.5 * 0.4, 0, 0]),
                           (scenes.BC_SEPARATE, 1.0, [0, 0, 0])):
        floor = scenes.cuboid([0.5, -1.0, 0.5], [2.0, 1.0, 2.0], bc=bc, friction=fr)
        out = O.grid_node_bc(floor, sim, node, vel, 64)
        assert np.allclose(out, expect, atol=1e-12), (bc, fr, out)
    # far from the collider: untouched; approaching within reach: normal speed limited
    floor = scenes.cuboid([0.5, -1.0, 0.5], [2.0, 1.0, 2.0], bc=scenes.BC_STICK)
    far = O.grid_node_bc(floor, sim, np.array([0.5, 0.5, 0.5]), vel, 64)
    assert np.allclose(far, vel)
    near = np.array([0.5, 1.2 * h, 0.5])
    v_fast = np.array([0.0, -50.0, 0.0])
    out = O.grid_node_bc(floor, sim, near, v_fast, 64)
    assert np.isclose(out[1], -(1.2 * h - h) / dt, rtol=1e-9)


# ---------------------------------------------------------------------------- whole substep

def affine_scene(n_side=12, B=None, a=None, model=None):
    # Dyadic positions, coefficients and mass: v = a + B x and C = m B are exact in FP32,
    # so the particle field is exactly affine.
    sc = scenes.elastic_cube(n_side, seed=99, side=0.1, center=(0.5, 0.5, 0.5))
    sc.x = (np.round(sc.x.astype(np.float64) * 2 ** 14) / 2 ** 14).astype(np.float32)
    sc.mass[:] = 2.0 ** -16
    B = np.array([[0.0, 0.1875, -0.125], [0.0625, 0.0625, 0.3125], [-0.1875, 0.125, -0.0625]]) if B is None else B
    a = np.array([0.125, -0.1875, 0.0625]) if a is None else a
    sc.v = (a + sc.x.astype(np.float64) @ B.T).astype(np.float32)
    assert np.array_equal(sc.v.astype(np.float64), a + sc.x.astype(np.float64) @ B.T)
    sc.C = (sc.mass[:, None, None] * B[None]).transpose(0, 2, 1).reshape(-1, 9).astype(np.float32)
    sc.colliders = np.zeros(0, dtype=scenes.CUBOID_DTYPE)
    if model is not None:
        tag, words = model
        sc.tag[:] = tag
        sc.model[:] = words
    return sc, a, B


def test_apic_affine_field_is_reproduced_exactly():
    """APIC exactness: particles carrying v = a + B x with affine momentum C = m B make
    the grid velocity exactly a + B x_i (+ g dt); G2P returns v_p = a + g dt + B x_p
    and grad v = B.  With F = I the neo-Hookean stress of F' = I + B dt gives
    C' = m B - tau V0 (4/h^2) dt and x' = x + v' dt."""
    sc, a, B = affine_scene()
    o = O.Oracle(sc, 64)
    x0 = o.s.x.copy()
    o.step(1)
    g = np.array([0.0, float(np.float32(-9.81)), 0.0])   # the simulation parameters are FP32
    dt = float(np.float32(sc.dt))
    v_exp = a + g * dt + x0 @ B.T
    assert np.allclose(o.s.v, v_exp, rtol=0, atol=1e-12)
    assert np.allclose(o.s.x, x0 + v_exp * dt, atol=1e-14)
    Fexp = np.eye(3) + B * dt
    assert np.allclose(from_colmajor(o.s.F[0]), Fexp, atol=1e-14)
    tag, words = scenes.model_neo_hookean(1e5, 0.3)
    _, _, _, tau = O.model_update(tag, words, colmajor(Fexp), np.zeros(9))
    h = float(np.float32(sc.cell_width))
    for i in range(0, sc.n, 97):
        Cexp = o.s.mass[i] * B - from_colmajor(tau) * (o.s.vol0[i] * 4 / (h * h) * dt)
        assert np.allclose(from_colmajor(o.s.C[i]), Cexp, rtol=1e-9, atol=1e-12)


def test_p2g_conserves_mass_momentum_and_angular_momentum():
    """Sum_i m_i = Sum_p m_p, Sum_i m_i v_i = Sum_p m_p v_p and, for quadratic splines,
    Sum_i x_i x (m v)_i = Sum_p [x_p x m_p v_p] + (h^2/4) eps_abc C_cb (grid moments of the
    weights: Sum w dpt = 0, Sum w dpt dpt^T = h^2/4 I)."""
    sc, a, B = affine_scene(10)
    sc.C = (RNG.normal(size=(sc.n, 9)) * 1e-3).astype(np.float32)
    o = O.Oracle(sc, 64)
    x = o.s.x.copy()
    m = o.s.mass.copy()
    v = o.s.v.copy()
    C = o.s.C.copy()
    o.step(1)
    gr = o.grid()
    b = o.blocks()
    h = float(np.float32(sc.cell_width))
    assert np.isclose(gr["m"].sum(), m.sum(), rtol=1e-13)
    assert np.allclose(gr["mv"].reshape(-1, 3).sum(0), (m[:, None] * v).sum(0), rtol=1e-11, atol=1e-15)
    nodes = []
    for blk in range(b["vid"].shape[0]):
        for t in range(64):
            nodes.append((b["vid"][blk] * 4 + np.array([t & 3, (t >> 2) & 3, t >> 4])) * h)
    nodes = np.array(nodes)
    L_grid = np.cross(nodes, gr["mv"].reshape(-1, 3)).sum(0)
    Cm = np.array([from_colmajor(c) for c in C])
    eps_term = (h * h / 4) * np.array([Cm[:, 2, 1] - Cm[:, 1, 2], Cm[:, 0, 2] - Cm[:, 2, 0], Cm[:, 1, 0] - Cm[:, 0, 1]]).sum(1)
    L_part = np.cross(x, m[:, None] * v).sum(0) + eps_term
    assert np.allclose(L_grid, L_part, rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("bits", [32, 64])
def test_free_fall_closed_form(bits):
    """Elastic cube far from its walls, 100 substeps: v = -g n dt and
    COM_y(n) = COM_y(0) - g dt^2 n (n + 1) / 2 (symplectic Euler)."""
    sc = scenes.elastic_cube(16)
    o = O.Oracle(sc, bits)
    com0 = o.s.x.astype(np.float64).mean(0)
    n = 100
    o.step(n, keep_debug=False)
    dt = float(np.float32(sc.dt))
    g = -float(np.float32(-9.81))   # the simulation parameters are FP32
    tol = 1e-12 if bits == 64 else 3e-6
    assert np.allclose(o.s.v[:, 1].astype(np.float64), -g * n * dt, rtol=tol)
    assert abs(o.s.x[:, 1].astype(np.float64).mean() - (com0[1] - g * dt * dt * n * (n + 1) / 2)) < (1e-12 if bits == 64 else 2e-7)


def test_fp32_mirror_tracks_fp64():
    """The FP32 mirror stays within FP32 rounding of the FP64 oracle over one substep
    of each material."""
    for sc in (scenes.sand_column(4000, seed=2, lo=(0.4, 0.4, 0.4), hi=(0.45, 0.5, 0.45)),
               scenes.fluid_clumps(5000, seed=3, n_clumps=2, cell_width=1 / 256)):
        a, b = O.Oracle(sc, 32), O.Oracle(sc, 64)
        a.step(1)
        b.step(1)
        assert np.max(np.abs(a.s.x - b.s.x)) < 1e-6
        assert np.max(np.abs(a.s.F - b.s.F)) < 1e-5


# ---------------------------------------------------------------------------- CFL bound

def _one_particle(model, density=1000.0, radius=0.01, velocity=(0.0, 0.0, 0.0), h=1.0 / 64):
    parts = scenes.particles(np.array([[0.5, 0.5, 0.5]]), radius, density, model,
                             velocity=np.array([velocity], np.float64))
    return scenes.make_scene(parts, h, 1e-4)


@pytest.mark.parametrize("young,poisson", [(1e5, 0.3), (3.5e5, 0.25), (2e6, 0.45)])
def test_timestep_bound_is_p_wave_cfl(young, poisson):
    """At rest the bound is cfl h / c with c the P-wave speed sqrt(M / rho), where the
    P-wave modulus is the textbook M = E (1 - nu) / ((1 + nu)(1 - 2 nu)) (not the
    lambda/mu route the model takes).  Neo-Hookean (cfl word 0.5) and sand (elastic
    backbone) agree; the result is in 1e-12 s units."""
    h = 1.0 / 64
    M = young * (1 - poisson) / ((1 + poisson) * (1 - 2 * poisson))
    for model, rho in ((scenes.model_neo_hookean(young, poisson), 1000.0),
                       (scenes.model_sand(young, poisson, friction_deg=30.0), 2200.0)):
        sc = _one_particle(model, density=rho, h=h)
        o = O.Oracle(sc, 64)
        cfl = 0.5
        expect = cfl * h / math.sqrt(M / rho)
        u, secs = O.timestep_bound(o.s)
        assert abs(u - expect * 1e12) <= 1e-6 * expect * 1e12 + 1, (u, expect * 1e12)


def test_timestep_bound_snow_hardening():
    """Snow: the P-wave speed of the elastic backbone with both moduli scaled by the
    hardening factor exp(xi (1 - det Fp)), so c grows by the factor's square root: a
    compacted particle (det Fp = 0.8) gets a bound sqrt(exp(0.2 xi)) times smaller."""
    h = 1.0 / 64
    young, poisson, rho = 1.4e5, 0.2, 400.0
    M = young * (1 - poisson) / ((1 + poisson) * (1 - 2 * poisson))
    sc = _one_particle(scenes.model_snow(young, poisson), density=rho, h=h)
    o = O.Oracle(sc, 64)
    expect = 0.5 * h / math.sqrt(M / rho)
    assert abs(O.timestep_bound(o.s)[0] - expect * 1e12) <= 1e-6 * expect * 1e12 + 1
    o.s.pstate[0, 0] = 0.8
    expect = 0.5 * h / math.sqrt(M * math.exp(10.0 * 0.2) / rho)
    assert abs(O.timestep_bound(o.s)[0] - expect * 1e12) <= 1e-6 * expect * 1e12 + 1


def test_timestep_bound_fluid_tait_sound_speed_and_velocity():
    """Fluid: the Tait law p = K ((rho/rho0)^gamma - 1) has dp/drho = K gamma / rho0 at
    rest, so c = sqrt(K gamma / rho0).  A particle faster than c bounds dt by cfl h / |v|
    (cfl = 0.5 < 1, so the sound-speed term is the smaller one); compression (det F < 1)
    raises the density and lowers c."""
    h = 1.0 / 128
    K, gamma, rho = 1e4, 7.0, 1000.0
    c = math.sqrt(K * gamma / rho)
    sc = _one_particle(scenes.model_fluid(K, gamma), density=rho, h=h)
    o = O.Oracle(sc, 64)
    assert abs(O.timestep_bound(o.s)[0] - 0.5 * h / c * 1e12) <= 1e-6 * 0.5 * h / c * 1e12 + 1
    v = 3.0 * c
    sc = _one_particle(scenes.model_fluid(K, gamma), density=rho, h=h, velocity=(v, 0.0, 0.0))
    o = O.Oracle(sc, 64)
    assert abs(O.timestep_bound(o.s)[0] - 0.5 * h / v * 1e12) <= 1e-6 * 0.5 * h / v * 1e12 + 1
    sc = _one_particle(scenes.model_fluid(K, gamma), density=rho, h=h)
    o = O.Oracle(sc, 64)
    J = 0.5
    o.s.F[0] = np.diag([J, 1.0, 1.0]).reshape(-1)
    expect = 0.5 * h / math.sqrt(K * gamma / (rho / J))
    assert abs(O.timestep_bound(o.s)[0] - expect * 1e12) <= 1e-6 * expect * 1e12 + 1


def test_timestep_bound_minimum_and_disabled():
    """The bound is the minimum over enabled particles; disabled particles are skipped
    and an empty set leaves the reset value u32::MAX."""
    h = 1.0 / 64
    a = _one_particle(scenes.model_neo_hookean(1e5, 0.3), h=h)
    b = _one_particle(scenes.model_neo_hookean(4e5, 0.3), h=h)
    both = scenes.make_scene(scenes.concat([
        {k: getattr(a, k) for k in ("x", "v", "F", "C", "mass", "vol0", "radius", "damping", "enabled", "fixed",
                                    "tag", "model")},
        {k: getattr(b, k) for k in ("x", "v", "F", "C", "mass", "vol0", "radius", "damping", "enabled", "fixed",
                                    "tag", "model")}]), h, 1e-4)
    ua = O.timestep_bound(O.Oracle(a, 64).s)[0]
    ub = O.timestep_bound(O.Oracle(b, 64).s)[0]
    o = O.Oracle(both, 64)
    assert O.timestep_bound(o.s)[0] == min(ua, ub) == ub      # stiffer: faster sound, smaller dt
    o.s.enabled[1] = 0
    assert O.timestep_bound(o.s)[0] == ua
    o.s.enabled[:] = 0
    assert O.timestep_bound(o.s)[0] == 0xFFFFFFFF


def test_adaptive_schedule_is_min_of_dt_and_the_bound():
    """The oracle's adaptive substep (the schedule the GPU's adaptive scenes are checked
    against): a block at rest takes the P-wave CFL bound cfl h / sqrt(M / rho) when dt is
    larger, and dt itself when dt is smaller."""
    from nexus_amd import scenes
    h = 1 / 64
    idx = np.stack(np.meshgrid(*(np.arange(6),) * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = 0.5 + (idx + 0.5) * h / 2
    lam, mu = scenes.lame_lambda_mu(1e5, 0.3)
    cfl_dt = 0.5 * h / np.sqrt((float(lam) + 2 * float(mu)) / 1000.0)
    for dt_max, expect in ((1e-2, cfl_dt), (1e-4, 1e-4)):
        sc = scenes.make_scene(scenes.particles(pos, h / 4, 1000.0, scenes.model_neo_hookean(1e5, 0.3)), h, dt_max,
                               gravity=(0.0, 0.0, 0.0))
        o = O.Oracle(sc, 64)
        got = o.step_adaptive(sc.dt)
        assert abs(got - expect) <= 1e-6 * expect, (got, expect)
        assert o.s.sim[3] == np.float32(got)
