"""GPU parity: the HIP substep against the CPU oracle on the same seeded inputs.

The oracle runs the reference algorithm in FP32 (op-for-op mirror) and FP64.  The
sort is compared bit-exactly in its canonical form; particle fields x, v, C, F within
1e-5 of the field's largest value (max-norm), or within 2x of the FP32 mirror's own distance to FP64.
"""
import numpy as np
import pytest

from nexus_amd import mpm, scenes

from parity import assert_field_close, compare_sort, oracle_pair

pytestmark = pytest.mark.gpu


def run_gpu(pipeline, scene, steps=1, debug=False):
    st = mpm.MpmState.from_scene(pipeline, scene)
    if debug:
        st.set_debug_grid(True)
    for _ in range(steps):
        pipeline.step(st)
    st.synchronize()
    return st


def check_state(gp, o32, o64):
    s32, s64 = o32.s, o64.s
    out = {}
    for name, g, a, b in (("x", gp["x"], s32.x, s64.x), ("v", gp["v"], s32.v, s64.v),
                          ("C", gp["C"], s32.C, s64.C), ("F", gp["F"], s32.F, s64.F)):
        out[name] = assert_field_close(name, g, a, b, s64)
    assert np.array_equal(gp["enabled"], s32.enabled)
    return out


def small_sand(n=20000, seed=11):
    sc = scenes.sand_column(n, seed=seed, lo=(0.4, 0.1, 0.4), hi=(0.5, 0.3, 0.5), n_colliders=16)
    return sc


def small_fluid(n=30000, seed=12):
    return scenes.fluid_clumps(n, seed=seed, n_clumps=4, cell_width=1 / 256, sigma=(0.01, 0.03))


@pytest.mark.parametrize("make", [lambda: scenes.elastic_cube(16), small_sand, small_fluid],
                         ids=["neo_hookean_cube", "sand_colliders", "fluid_clumps"])
def test_one_substep_matches_oracle(pipeline, make):
    scene = make()
    x0 = scene.x.copy()
    st = run_gpu(pipeline, scene, 1, debug=True)
    o32, o64 = oracle_pair(scene, 1)
    compare_sort(st.read_sort_canonical(), o32, x0, scene.cell_width)
    errs = check_state(st.read_particles(), o32, o64)
    print({k: f"gpu {v[0]:.2e} fp32 {v[1]:.2e} scale {v[2]:.2e}" for k, v in errs.items()})


def test_grid_after_p2g_matches_oracle(pipeline):
    """Grid mass and momentum after P2G and velocity after the grid update (gravity,
    clamp, 16 cuboid colliders), node by node, against FP64 with the parity bar of the
    particle fields (within 2x of the FP32 mirror's distance to FP64)."""
    scene = small_sand(8000, seed=3)
    st = run_gpu(pipeline, scene, 1, debug=True)
    o32, o64 = oracle_pair(scene, 1)
    can = st.read_sort_canonical()
    grid = st.read_grid(momentum=True)
    g64, g32 = o64.grid(), o32.grid()
    b = o64.blocks()
    b32 = o32.blocks()
    assert np.array_equal(b["vid"], b32["vid"])
    gidx = {tuple(map(int, can["vid"][i])): i for i in range(can["vid"].shape[0])}
    order = np.array([gidx[tuple(map(int, v))] for v in b["vid"]])
    assert_field_close("mass", grid["m"][order], g32["m"], g64["m"])
    assert_field_close("momentum", grid["mv"][order], g32["mv"], g64["mv"])
    # Velocity wherever the GPU node has mass.  A node whose total weight is below the
    # P2G fixed-point resolution (2^-62 np of the block's largest contribution) may read
    # as massless; the oracle's mass there must then be negligible.
    gm, om = grid["m"][order], g64["m"]
    has = gm > 0
    assert np.all(om[~has] <= 1e-12 * om.max())
    assert_field_close("velocity", grid["vel"][order][has], g32["vel"][has], g64["vel"][has])


def test_free_fall_closed_form(pipeline):
    """Elastic cube, 100 substeps of free fall: v = -g n dt, COM_y(n) = COM_y(0) - g dt^2 n(n+1)/2."""
    scene = scenes.elastic_cube(16)
    st = run_gpu(pipeline, scene, 100)
    p = st.read_particles()
    dt = float(np.float32(scene.dt))
    g = -float(np.float32(-9.81))
    n = 100
    assert np.allclose(p["v"][:, 1].astype(np.float64), -g * n * dt, rtol=2e-5, atol=0)
    com0 = scene.x.astype(np.float64).mean(0)
    com = p["x"].astype(np.float64).mean(0)
    assert abs(com[1] - (com0[1] - g * dt * dt * n * (n + 1) / 2)) < 2e-7
    assert np.max(np.abs(p["F"] - np.eye(3, dtype=np.float32).reshape(9))) < 1e-5


def test_batch_matches_alone(pipeline):
    sc_a = scenes.elastic_cube(12)
    sc_b = small_sand(6000, seed=5)
    sc_c = small_fluid(9000, seed=6)
    alone = []
    for sc in (sc_a, sc_b, sc_c):
        st = run_gpu(pipeline, sc, 5)
        alone.append(st.read_particles())
    batch = mpm.Batch(pipeline)
    states = [mpm.MpmState.from_scene(pipeline, sc) for sc in (sc_a, sc_b, sc_c)]
    for s in states:
        batch.add(s)
    batch.run(5)
    batch.synchronize()
    # P2G sums in exact fixed point and the other stages work per particle or per node,
    # so a scene's result does not depend on its batch: bit-identical.
    for s, ref in zip(states, alone):
        got = s.read_particles()
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(got[f], ref[f]), f
    batch.close()


def test_graph_replay_matches_eager(pipeline):
    """Each substep replayed from a captured graph equals the eagerly launched substep, on
    every replay, with host reads (synchronizing copies) between the replays and a
    recapture when a scene joins the batch."""
    sc = small_sand(8000, seed=21)
    eager = mpm.MpmState.from_scene(pipeline, sc)
    graphed = mpm.MpmState.from_scene(pipeline, sc)
    graphed.set_graph(True)
    for k in range(6):
        pipeline.step(eager)
        pipeline.step(graphed)
        assert np.array_equal(graphed.read_positions(), eager.read_positions()), k
    a, b = eager.read_particles(), graphed.read_particles()
    for f in ("x", "v", "C", "F", "enabled"):
        assert np.array_equal(a[f], b[f]), f
    eager.close()
    graphed.close()

    extra = scenes.elastic_cube(8)
    ref1 = run_gpu(pipeline, sc, 6).read_particles()
    ref2 = run_gpu(pipeline, extra, 2).read_particles()
    batch = mpm.Batch(pipeline)
    batch.set_graph(True)
    s1 = mpm.MpmState.from_scene(pipeline, sc)
    batch.add(s1)
    batch.run(2)
    s2 = mpm.MpmState.from_scene(pipeline, extra)
    batch.add(s2)                      # new tables: recapture
    batch.run(2)
    got2 = s2.read_particles()         # 2 substeps in the batch
    batch.run(2)
    got1 = s1.read_particles()         # 6 substeps
    for f in ("x", "v", "C", "F"):
        assert np.array_equal(got1[f], ref1[f]), f
        assert np.array_equal(got2[f], ref2[f]), f
    batch.close()


def small_snow(n_side=14, seed=43):
    rng = np.random.Generator(np.random.PCG64(seed))
    h = 1 / 64
    s = h / 2
    idx = np.stack(np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = 0.5 - n_side * s / 2 + (idx + 0.5) * s + rng.uniform(-0.25 * s, 0.25 * s, idx.shape)
    vel = rng.normal(0.0, 0.5, pos.shape)
    parts = scenes.particles(pos, s / 2, 400.0, scenes.model_snow(1.4e5, 0.2), velocity=vel)
    return scenes.make_scene(parts, h, 1e-4)


@pytest.mark.parametrize("make", [lambda: scenes.elastic_cube(12), lambda: small_sand(8000, seed=41),
                                  lambda: small_fluid(12000, seed=42), small_snow],
                         ids=["cube", "sand", "fluid", "snow"])
def test_timestep_bound_matches_oracle(pipeline, make):
    """The CFL bound of the GPU state equals the oracle's FP32 bound of the same state
    (the reference's 1e-12 s fixed point, within FP32 rounding), alone and in a batch."""
    import oracle as O
    sc = make()
    st = run_gpu(pipeline, sc, 3)
    p = st.read_particles()
    o = O.Oracle(sc, 32)
    for f in ("v", "C", "F", "mass", "enabled"):
        getattr(o.s, f)[...] = p[f]
    snow = p["tag"] == scenes.MODEL_SNOW
    o.s.pstate[snow, 0] = p["model"][snow, 0].view(np.float32)
    _, secs = O.timestep_bound(o.s)
    got = st.timestep_bound()
    assert abs(got - secs) <= 1e-5 * secs, (got, secs)
    other = mpm.MpmState.from_scene(pipeline, scenes.elastic_cube(6))
    batch = mpm.Batch(pipeline)
    batch.add(other)
    batch.add(st)
    assert st.timestep_bound() == got
    batch.close()
    other.close()
    st.close()


def _swap_remove_host(arrs: dict, ids):
    """GpuParticles::swap_remove on host arrays: descending ids, last moves into each."""
    n = arrs["x"].shape[0]
    moves = []
    for i in sorted(set(int(k) for k in ids), reverse=True):
        if i >= n:
            continue
        last = n - 1
        if i != last:
            for k in arrs:
                arrs[k][i] = arrs[k][last]
            moves.append((last, i))
        n -= 1
    return {k: v[:n].copy() for k, v in arrs.items()}, moves


def test_swap_remove_then_step(pipeline):
    """Removing particles mid-run equals building the reduced scene from the same state:
    the moves follow swap_remove, and the next substeps agree bit for bit."""
    sc = small_sand(6000, seed=51)
    st = run_gpu(pipeline, sc, 2)
    p = st.read_particles()
    ids = [5, 17, 5, 4321, sc.n - 1, sc.n + 10, 0]
    keys = ("x", "v", "C", "F", "mass", "vol0", "radius", "damping", "enabled", "fixed", "tag", "model")
    arrs = {k: np.array(p[k]) for k in keys}       # model words carry the plastic state
    reduced, moves = _swap_remove_host(arrs, ids)
    got_moves = st.swap_remove(ids)
    assert got_moves == moves
    assert len(st) == sc.n - 5
    ref_scene = scenes.Scene(cell_width=sc.cell_width, dt=sc.dt, gravity=sc.gravity, colliders=sc.colliders,
                             **reduced)
    ref = run_gpu(pipeline, ref_scene, 2)
    for _ in range(2):
        pipeline.step(st)
    a, b = st.read_particles(), ref.read_particles()
    for f in ("x", "v", "C", "F", "enabled"):
        assert np.array_equal(a[f], b[f]), f
    st.close()
    ref.close()


def test_append_then_step(pipeline):
    """GpuParticles::append on a resident scene mid-run (src_mpm/solver/particle.rs:533):
    the appended particles join the next substep (whose block keys the last G2P did not
    make for them), bit for bit as a scene built from the same state with them, and in
    graph replay as eagerly."""
    sc = small_sand(6000, seed=52)
    extra = small_sand(1500, seed=53)
    for graph in (False, True):
        st = mpm.MpmState.from_scene(pipeline, sc)
        st.set_graph(graph)
        for _ in range(3):
            pipeline.step(st)
        p = st.read_particles()
        st.add_particles(extra)
        for _ in range(4):
            pipeline.step(st)
        keys = ("x", "v", "C", "F", "mass", "vol0", "radius", "damping", "enabled", "fixed", "tag", "model")
        both = {k: np.concatenate([np.array(p[k]), np.array(getattr(extra, k))]) for k in keys}
        ref_scene = scenes.Scene(cell_width=sc.cell_width, dt=sc.dt, gravity=sc.gravity, colliders=sc.colliders, **both)
        ref = run_gpu(pipeline, ref_scene, 4)
        a, b = st.read_particles(), ref.read_particles()
        assert len(st) == sc.n + extra.n
        for f in ("x", "v", "C", "F", "enabled"):
            assert np.array_equal(a[f], b[f]), (graph, f)
        st.close()
        ref.close()


def test_launch_shapes_do_not_change_results(pipeline):
    """P2G sums in exact fixed point and the other stages work per particle or per node,
    so any launch shape gives the same bits: fixed extreme shapes and an autotuned run
    (which switches shapes between substeps) all agree."""
    mix = [scenes.elastic_cube(10), small_sand(5000, seed=61), small_fluid(7000, seed=62)]
    cu = None
    results = []
    for mode in ("default", "small", "large", "autotune"):
        batch = mpm.Batch(pipeline)
        states = [mpm.MpmState.from_scene(pipeline, sc) for sc in mix]
        for s in states:
            batch.add(s)
        if cu is None:
            cu = batch.launch_config()[0] // 32
        if mode == "small":
            batch.set_launch_config(cu * 16, 4, cu * 4)
        elif mode == "large":
            batch.set_launch_config(cu * 512, 64, cu * 64)   # the largest shapes the tuner tries
        elif mode == "autotune":
            batch.set_autotune(True)
        batch.run(32)
        results.append([s.read_particles() for s in states])
        if mode == "autotune":
            g, k, gg = batch.launch_config()
            assert g in (cu * 32, cu * 128, cu * 256, cu * 512) and k in (4, 16, 64)
            assert gg in (cu * 4, cu * 8, cu * 16, cu * 32, cu * 64)
        batch.close()
        for s in states:
            s.close()
    for other in results[1:]:
        for a, b in zip(results[0], other):
            for f in ("x", "v", "C", "F", "enabled"):
                assert np.array_equal(a[f], b[f]), f


def _moving_floor_scene(seed=71):
    sc = small_sand(6000, seed=seed)
    q = np.array([np.sin(0.05), 0.0, 0.0, np.cos(0.05)], np.float32)
    floor = scenes.cuboid((0.45, 0.06, 0.45), (0.2, 0.03, 0.2), rotation=q, bc=scenes.BC_SEPARATE, friction=0.6,
                          linvel=(0.3, 0.4, -0.2), angvel=(0.5, 2.0, -1.0))
    sc.colliders = np.concatenate([sc.colliders[:3], floor])
    return sc


def test_moving_colliders_integrate_like_the_oracle(pipeline):
    """Coupled bodies are integrated by MPM every substep (src_mpm/pipeline.rs:510-520):
    the moving, turning floor's pose and the sand on it after 30 substeps against the
    oracle."""
    sc = _moving_floor_scene()
    st = run_gpu(pipeline, sc, 30)
    cols = st.read_colliders()
    o32, o64 = oracle_pair(sc, 30)
    for f in ("translation", "rotation"):
        assert np.allclose(cols[f], o64.s.colliders[f], rtol=0, atol=2e-6), f
    assert not np.allclose(cols["translation"][-1], sc.colliders["translation"][-1])   # it moved
    check_state(st.read_particles(), o32, o64)
    st.close()


def test_dynamic_body_falls_through_the_sand_like_the_oracle(pipeline):
    """A dynamic body coupled two ways (nonzero inverse mass, off-centre centre of mass,
    turning): MPM adds gravity to its velocity every substep while the sand's impulses
    on it are compiled out (src_mpm_shaders/lib.rs:38), so it falls through the sand,
    which it pushes aside.  Pose, velocity and sand after 30 substeps against the
    oracle."""
    sc = small_sand(6000, seed=73)
    body = scenes.cuboid((0.45, 0.3, 0.45), (0.05, 0.05, 0.08), bc=scenes.BC_SLIP, linvel=(0.0, -1.0, 0.0),
                         angvel=(0.0, 2.0, 0.0), inv_mass=(0.5, 0.5, 0.5), local_com=(0.02, 0.0, 0.0))
    sc.colliders = np.concatenate([sc.colliders[:3], body])
    st = run_gpu(pipeline, sc, 30)
    cols = st.read_colliders()
    o32, o64 = oracle_pair(sc, 30)
    for f in ("translation", "rotation", "linvel"):
        assert np.allclose(cols[f], o64.s.colliders[f], rtol=0, atol=2e-6), f
    assert cols["linvel"][-1][1] < -1.0                 # gravity acted on the body
    check_state(st.read_particles(), o32, o64)
    st.close()


def test_moving_colliders_survive_batch_rebuilds(pipeline):
    """The integrated poses live on the device: a scene that joins a batch after stepping
    alone, and whose batch is rebuilt when another scene joins, continues from its moved
    colliders, bit for bit as if it had stepped alone."""
    sc = _moving_floor_scene(seed=72)
    ref = run_gpu(pipeline, sc, 10)
    a, ca = ref.read_particles(), ref.read_colliders()
    st = run_gpu(pipeline, sc, 3)                      # 3 substeps in its own world
    batch = mpm.Batch(pipeline)
    batch.add(st)                                      # moves into the batch's world
    batch.run(4)
    other = mpm.MpmState.from_scene(pipeline, scenes.elastic_cube(8))
    batch.add(other)                                   # table rebuild with the poses kept
    batch.run(3)
    b, cb = st.read_particles(), st.read_colliders()
    for f in ("translation", "rotation"):
        assert np.array_equal(ca[f], cb[f]), f
    for f in ("x", "v", "C", "F", "enabled"):
        assert np.array_equal(a[f], b[f]), f
    batch.close()


def test_particle_chunks_on_the_device(pipeline):
    """NexusState's particle chunks (src/state.rs:665-774) over a resident scene: chunks
    added in separate regions, one removed whole and one in part between substeps; every
    remaining particle's chunk (followed through the library's swap-remove moves) is the
    region it lies in, and the counts add up."""
    from nexus_amd.chunks import ParticleChunks
    h = 1.0 / 64
    model = scenes.model_neo_hookean(1e5, 0.3)

    def block(x0, seed, n_side=6):
        rng = np.random.default_rng(seed)
        pos = np.array([x0, 0.5, 0.5]) + rng.random((n_side ** 3, 3)) * 0.06
        return scenes.make_scene(scenes.particles(pos, h / 4, 1000.0, model), h, 1e-4)

    base = block(0.1, 1)
    st = mpm.MpmState.from_scene(pipeline, base)
    ch = ParticleChunks(st)
    a = ch.add_particles(block(0.3, 2))
    b = ch.add_particles(block(0.5, 3))
    pipeline.step(st)
    ch.extend_chunk(a, block(0.3, 4, 4))
    c = ch.add_particles(block(0.7, 5))
    pipeline.step(st)
    ch.remove_chunk(b)
    assert ch.remove_particles_from_chunk(a, 50) == 50
    pipeline.step(st)
    x = st.read_particles()["x"]
    region = {-1: 0.1, a: 0.3, c: 0.7}
    assert len(st) == base.n + 216 + 64 - 50 + 216 == len(ch.slot2chunk)
    for slot, chunk in enumerate(ch.slot2chunk):
        assert chunk in region and region[chunk] - 0.01 < x[slot, 0] < region[chunk] + 0.07, (slot, chunk)
    assert ch.sizes == {a: 216 + 64 - 50, c: 216}
    st.close()
