#!/usr/bin/env python3
"""Benchmark of the MLS-MPM substep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sand|cube|fluid]

Default workload: the Drucker-Prager sand column, 1,048,576 particles, h = 1/256,
16 seeded cuboid colliders, dt = 5e-5 (the headline configuration); K = 1000 substeps
are timed after W warm-up substeps (the launch-shape autotuner runs on the warm-up).  Two further measurements go into the same line:
"batched" (256 mixed scenes stepped as one batch, against the plain per-scene loop) and
"served" (Poisson arrivals served with continuous batching: p50/p99 substep latency and
sustained throughput), and "cpic" (the sand column at 64x scale with the CPIC
transfers).  Particle state is resident in HBM before the
timed region starts.  Prints one JSON line with the whole-job throughput
(particle-substeps/s), the roofline figure of the dominant kernel (timed live with HIP
events on the scene's stream, on one timed substep in ten) and a CPU baseline (the oracle's
FP32 restatement of the reference algorithm on the host cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s at 1M FP32 MLS-MPM; batched scenes/s + p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
# VALU issue peak without packed FP32: one wave64 VALU instruction per CU per cycle (four
# 16-lane SIMDs, four cycles each), 256 CUs at 2.4 GHz = 614.4 G wave-instructions/s.
VALU_PEAK_G = 256 * 2.4
SEEDS = {"cube": 0x4E584D01, "sand": 0x4E584D02, "fluid": 0x4E584D03}
# The kernel variants the bench workloads launch (no CPIC, no adaptive substeps), under the
# names rocprofv3 records them with in profiles/counters_<workload>.json.
PROFILE_KERNELS = {"k_p2g": "k_p2g<false>", "k_g2p_update": "k_g2p_update<false, false>"}


def build_scene(name: str):
    from nexus_amd import scenes
    if name == "sand":
        sc = scenes.sand_column()
        desc = ("Drucker-Prager sand column, 1,048,576 particles (friction 30 deg, E=3.5e5, nu=0.3, "
                "density 2200), h=1/256, 16 seeded cuboid colliders, dt=5e-5")
    elif name == "cube":
        sc = scenes.elastic_cube()
        desc = "neo-Hookean cube, 32,768 particles, h=1/64, six stick walls, dt=1e-4"
    elif name == "fluid":
        sc = scenes.fluid_clumps()
        desc = "weakly-compressible fluid, 8,388,608 particles in 64 Gaussian clumps, h=1/512, dt=2e-5"
    else:
        raise SystemExit(f"unknown workload {name}")
    return sc, desc


def p2g_algorithmic_bytes(n_particles: int, n_blocks: int) -> int:
    """Compulsory traffic of one P2G launch: per particle x, v, C, m read (12 + 12 + 36 + 4
    bytes); per node of every active block its momentum and mass written (16 bytes)."""
    return 64 * n_particles + 16 * 64 * n_blocks


def g2p_algorithmic_bytes(n_particles: int, n_blocks: int, plastic_words: int) -> int:
    """Compulsory traffic of one G2P + particle update launch: per particle x, F, V0 read
    (12 + 36 + 4 bytes) and x, v, C, F written (12 + 12 + 36 + 36), the plastic state read
    and written (4 bytes a word: 3 for Drucker-Prager sand), the next substep's block key
    written (4); per node of every active block its velocity read (16 bytes)."""
    return (52 + 96 + 8 * plastic_words + 4) * n_particles + 16 * 64 * n_blocks


def finish_tuning(obj, step, limit: int = 400) -> dict:
    """Steps ``obj`` (a scene or a batch, through ``step``) until its launch-shape tuner
    has kept a shape, so that no trial shape and no tuner synchronization falls inside a
    timed region.  Returns the tuner's state for the JSON line."""
    extra = 0
    tuning, used = obj.autotune_state()
    while tuning and extra < limit:
        step()
        extra += 1
        tuning, used = obj.autotune_state()
    return {"finished": not tuning, "substeps": used, "extra_warmup_substeps": extra}


def cpu_baseline(scene, budget_s: float = 15.0, max_steps: int = 30):
    """The oracle's FP32 restatement of the reference substep, OpenMP on the host cores,
    on the same scene: substeps until the budget is spent (at least 2)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(scene, 32)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        o.step(1, keep_debug=False)
        steps += 1
        el = time.perf_counter() - t0
        if steps >= 2 and el >= budget_s:
            break
    el = time.perf_counter() - t0
    return {
        "value": scene.n * steps / el,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} substeps of the full {scene.n}-particle scene, FP32 oracle (C restatement of the "
                  f"reference algorithm), {threads} OpenMP threads, {el:.1f} s",
    }


def bench_cpic(pipe, steps: int, warmup: int = 40, autotune: bool = True) -> dict:
    """The sand column at 64x scale (h = 0.25) with the CPIC transfers instead of the
    grid collider pass: particle-substeps/s and stage times (the contact distance field
    kernels are counted in the P2G stage)."""
    from nexus_amd import mpm, scenes
    sc = scenes.sand_column_cpic()
    res = {}
    for cpic in (True, False):
        run = sc.copy()
        run.cpic = cpic
        st = mpm.MpmState.from_scene(pipe, run)
        if autotune:
            st.set_autotune(True)
        for _ in range(warmup):
            pipe.step(st)
        if autotune:
            finish_tuning(st, lambda: pipe.step(st))
        st.synchronize()
        st.profile_begin(every=10)
        t0 = time.perf_counter()
        for _ in range(steps):
            pipe.step(st)
        st.synchronize()
        el = time.perf_counter() - t0
        stages, _ = st.profile_end()
        aff = st.read_cdf()["affinity"] if cpic else None
        cfg = dict(zip(("p2g_groups", "p2g_block_cost", "grid_groups"), st.launch_config()))
        res[cpic] = (el, stages, st.sort_size(), aff, cfg)
        st.close()
    el, stages, nb, aff, cfg = res[True]
    return {
        "workload": "Drucker-Prager sand column, 1,048,576 particles, 64x scale (h=0.25, dt=3.2e-3), 16 seeded "
                    "cuboid colliders, CPIC",
        "particle_substeps_per_s": sc.n * steps / el,
        "ms_per_step": 1e3 * el / steps,
        "steps": steps,
        "active_blocks": nb,
        "particles_with_affinity": int((aff & 0xFFFF != 0).sum()),
        "launch_config": cfg,
        "stages_ms": {"sort": stages.sort_ms, "cdf_and_p2g": stages.p2g_ms, "grid_update": stages.grid_update_ms,
                      "g2p_update": stages.g2p_update_ms},
        "same_scene_without_cpic": {"particle_substeps_per_s": sc.n * steps / res[False][0],
                                    "stages_ms": {"sort": res[False][1].sort_ms, "p2g": res[False][1].p2g_ms,
                                                  "grid_update": res[False][1].grid_update_ms,
                                                  "g2p_update": res[False][1].g2p_update_ms}},
    }


def bench_batched(pipe, count: int, substeps: int, loop_substeps: int) -> dict:
    """Many-scene workload: ``count`` independent scenes (log-uniform sizes, mixed
    materials and cell widths) stepped together as one batch, against the plain loop
    that steps each scene on its own through the single-scene API."""
    from nexus_amd import mpm, scenes
    scs = scenes.batch_scenes(count, substeps=substeps)
    total = sum(sc.n for sc in scs)

    states = [mpm.MpmState.from_scene(pipe, sc) for sc in scs]
    for st in states:                      # upload and warm up
        pipe.step(st)
    for st in states:
        st.synchronize()
    t0 = time.perf_counter()
    for _ in range(loop_substeps):
        for st in states:
            pipe.step(st)
    for st in states:
        st.synchronize()
    loop_s = time.perf_counter() - t0
    for st in states:
        st.close()
    loop_rate = total * loop_substeps / loop_s

    states = [mpm.MpmState.from_scene(pipe, sc) for sc in scs]
    batch = mpm.Batch(pipe)
    for st in states:
        batch.add(st)
    batch.run(2)
    batch.synchronize()
    batch.profile_begin()
    batch.run(4)
    stages, _ = batch.profile_end()
    t0 = time.perf_counter()
    batch.run(substeps)
    batch.synchronize()
    el = time.perf_counter() - t0
    # the same batch with each substep replayed as one captured graph
    batch.set_graph(True)
    batch.run(2)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(substeps)
    batch.synchronize()
    el_graph = time.perf_counter() - t0
    # autotuned launch shapes (the tuner runs on live substeps), then graph replay
    batch.set_graph(False)
    batch.set_autotune(True)
    tuner = finish_tuning(batch, lambda: batch.run(1))
    batch.set_graph(True)
    batch.run(2)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(substeps)
    batch.synchronize()
    el_tuned = time.perf_counter() - t0
    tuned_cfg = batch.launch_config()
    batch.close()
    for st in states:
        st.close()
    rate = total * substeps / el
    rate_graph = total * substeps / el_graph
    rate_tuned = total * substeps / el_tuned
    best = max(rate, rate_graph, rate_tuned)
    el_best = min(el, el_graph, el_tuned)
    return {
        "scenes": count, "particles": total, "substeps": substeps,
        "particle_substeps_per_s": best,
        "scenes_per_s": count / el_best,                  # full horizons completed per second
        "ms_per_batch_substep": 1e3 * el_best / substeps,
        "eager_particle_substeps_per_s": rate,
        "graph_particle_substeps_per_s": rate_graph,
        "autotuned_graph_particle_substeps_per_s": rate_tuned,
        "autotuned_launch_config": {"p2g_groups": tuned_cfg[0], "p2g_block_cost": tuned_cfg[1],
                                    "grid_groups": tuned_cfg[2]},
        "tuner": tuner,
        "stages_ms": {"sort": stages.sort_ms, "p2g": stages.p2g_ms, "grid_update": stages.grid_update_ms,
                      "g2p_update": stages.g2p_update_ms},
        "plain_loop": {"particle_substeps_per_s": loop_rate, "substeps": loop_substeps,
                       "ms_per_substep_all_scenes": 1e3 * loop_s / loop_substeps},
        "speedup_vs_plain_loop": best / loop_rate,
        "workload": f"{count} scenes, N log-uniform in [2e3, 2e5], materials uniform over neo-Hookean / "
                    f"Drucker-Prager / fluid, h = 1/64..1/128, dt = 1e-4",
    }


def bench_served(pipe, seconds: float, rate: float, budget: float) -> dict:
    """Serving workload: Poisson arrivals, continuous batching, HBM-budget admission."""
    from nexus_amd import scenes, serve
    specs = scenes.served_requests(seconds, rate=rate)
    # Requests carry their particle data (built here, before the clock starts); the server's
    # ingest thread turns each into a scene when it arrives.
    reqs = [serve.Request(sp.arrival, sp.build(), sp.horizon, rid=sp.index) for sp in specs]
    srv = serve.Server(pipe, hbm_budget_bytes=budget)
    rep = srv.run(reqs, clock="wall")
    srv.close()
    out = rep.summary()
    out["workload"] = (f"Poisson arrivals at {rate:g} scenes/s for {seconds:g} s ({len(reqs)} requests), "
                       f"N log-uniform in [1e3, 5e5], horizons uniform in [50, 500] substeps, "
                       f"HBM budget {budget / 1e9:g} GB")
    out["offered_particle_substeps_per_s"] = sum(sp.n * sp.horizon for sp in specs) / seconds
    return out


def counters_from_profiles(workload: str) -> dict:
    """Per-kernel counter figures recorded by rocprofv3 for this workload
    (profiles/counters_<workload>.json: HBM bytes per launch from FETCH_SIZE / WRITE_SIZE
    scaled by a copy calibration, VALU wave-instructions per launch from SQ_INSTS_VALU),
    keyed by kernel name; empty when no profile is committed."""
    path = os.path.join(ROOT, "profiles", f"counters_{workload}.json")
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return {}
    kernels = d.get("kernels", {})
    source = f"profiles/counters_{workload}.json (recorded at commit {d.get('commit', '?')})"
    return {k: dict(v, source=source) for k, v in kernels.items()}


def measured_copy_gbs(nbytes: int = 1 << 30, reps: int = 10) -> dict:
    """HBM bandwidth a plain device-to-device copy reaches on this GPU (torch `copy_` of a
    1 GiB buffer, timed with events on the current stream): bytes read + bytes written per
    copy / its average time, beside the datasheet's 8 TB/s."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        y.copy_(x)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    del x, y
    return {"value": 2 * nbytes / (ms * 1e-3) / 1e9, "unit": "GB/s", "bytes_per_copy": 2 * nbytes, "ms_per_copy": ms,
            "copies": reps}


def copy_in_child() -> dict | None:
    """measured_copy_gbs in a child process: torch's HIP runtime and the library's do not
    share one process."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--measure-copy"], capture_output=True,
                           text=True, timeout=300)
        lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
        return json.loads(lines[-1]) if r.returncode == 0 and lines else None
    except (OSError, ValueError, subprocess.SubprocessError):
        return None


def calibrate_copy(nbytes: int = 1 << 30, reps: int = 4):
    """A plain device-to-device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": nbytes, "reps": reps}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--graph", action="store_true", help="replay each timed substep as one captured HIP graph")
    ap.add_argument("--workload", default="sand", choices=["sand", "cube", "fluid"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the many-scene batch measurement")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-substeps", type=int, default=500)
    ap.add_argument("--loop-substeps", type=int, default=10)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-cpic", action="store_true", help="skip the CPIC measurement")
    ap.add_argument("--profile-every", type=int, default=10,
                    help="record the stage events on one timed substep in this many")
    ap.add_argument("--cpic-steps", type=int, default=300)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--measure-copy", action="store_true",
                    help="print only the measured device-copy bandwidth (run by the bench in a child process)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (calibration of the HBM byte counters)")
    args = ap.parse_args(argv)

    if args.calibrate_copy:
        return calibrate_copy()
    if args.measure_copy:
        print(json.dumps(measured_copy_gbs()), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (scenes shard across GPUs "
              f"without communication, one process per GPU, so N GPUs step N times the scenes)", file=sys.stderr)
        return 2

    from nexus_amd import mpm

    scene, desc = build_scene(args.workload)
    pipe = mpm.MpmPipeline(0)
    st = mpm.MpmState.from_scene(pipe, scene)
    if not args.no_autotune:
        st.set_autotune(True)       # the tuner runs on live substeps (28 of them)
    for _ in range(args.warmup):
        pipe.step(st)
    # Whatever --warmup is, the clock starts only once the tuner has kept a shape.
    tuner = finish_tuning(st, lambda: pipe.step(st)) if not args.no_autotune else {"finished": False,
                                                                                   "substeps": 0, "off": True}
    if args.graph:
        st.set_graph(True)          # substeps with stage events still run uncaptured
        for _ in range(2):
            pipe.step(st)
    st.synchronize()

    # Stage events on one substep in ten: an event record between kernels costs GPU time of
    # its own (about 7% of the substep when recorded on every one).
    st.profile_begin(every=args.profile_every)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pipe.step(st)
    st.synchronize()
    t1 = time.perf_counter()
    stages, nprof = st.profile_end()
    elapsed = t1 - t0
    n = scene.n
    value = n * args.steps / elapsed
    nblocks = st.sort_size()
    launch_cfg = st.launch_config()

    # The roofline of the dominant kernel (the longer of P2G and G2P + update), the other
    # beside it.  Each kernel's launch time is the average of its own start/end timestamps
    # over the sampled timed substeps.
    plastic = 3 if args.workload == "sand" else 0
    counters = counters_from_profiles(args.workload)
    figures, valu = {}, {}
    for kernel, kname, kernel_ms, stage_ms, nbytes in (
            ("k_p2g", PROFILE_KERNELS["k_p2g"], stages.p2g_kernel_ms, stages.p2g_ms, p2g_algorithmic_bytes(n, nblocks)),
            ("k_g2p_update", PROFILE_KERNELS["k_g2p_update"], stages.g2p_update_kernel_ms, stages.g2p_update_ms,
             g2p_algorithmic_bytes(n, nblocks, plastic))):
        achieved = nbytes / (kernel_ms * 1e-3) / 1e9
        prof = counters.get(kname, {})
        if counters and not prof:
            print(f"bench: no profile figures for {kname} in profiles/counters_{args.workload}.json",
                  file=sys.stderr)
        figures[kernel] = {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                           "traffic": prof.get("traffic_bytes"), "traffic_source": prof.get("source"),
                           "algorithmic_bytes": nbytes, "avg_launch_ms": kernel_ms, "stage_ms": stage_ms,
                           "launches_timed": nprof}
        if prof.get("valu_insts"):
            rate = prof["valu_insts"] / (kernel_ms * 1e-3) / 1e9
            valu[kernel] = {"kernel": kernel, "bound": "valu", "achieved": rate, "peak": VALU_PEAK_G,
                            "unit": "G VALU wave-instructions/s", "frac": rate / VALU_PEAK_G,
                            "valu_insts_per_launch": prof["valu_insts"], "avg_launch_ms": kernel_ms,
                            "source": prof.get("source")}
    dominant = max(figures, key=lambda k: figures[k]["avg_launch_ms"])
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": desc, "particles": n, "cell_width": scene.cell_width, "dt": scene.dt,
                   "colliders": int(scene.colliders.size), "active_blocks": nblocks, "parallelism": "single GPU",
                   "seed": SEEDS[args.workload], "generator": "numpy PCG64 (nexus_amd/scenes.py)"},
        "launch_config": dict(zip(("p2g_groups", "p2g_block_cost", "grid_groups"), launch_cfg)),
        "tuner": tuner,
        "graph_replay": bool(args.graph),
        "stages_ms": {"sort": stages.sort_ms, "p2g": stages.p2g_ms, "grid_update": stages.grid_update_ms,
                      "g2p_update": stages.g2p_update_ms},
        "roofline": figures[dominant],
        "roofline_other": figures[[k for k in figures if k != dominant][0]],
    }
    if valu:
        out["roofline_valu"] = valu
    st.close()
    # The copy bandwidth measured on this GPU, beside the datasheet peak the rooflines use.
    copy = copy_in_child()
    out["hbm_copy"] = copy
    if copy:
        for fig in (out["roofline"], out["roofline_other"]):
            fig["frac_of_copy"] = fig["achieved"] / copy["value"]
    if not args.no_cpic:
        out["cpic"] = bench_cpic(pipe, args.cpic_steps, autotune=not args.no_autotune)
    if not args.no_batch:
        out["batched"] = bench_batched(pipe, args.batch_scenes, args.batch_substeps, args.loop_substeps)
    if not args.no_serve:
        out["served"] = bench_served(pipe, args.serve_seconds, args.serve_rate, args.serve_budget_gb * 1e9)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, budget_s=args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
