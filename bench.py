"""Benchmark of the per-step AVBD solve on one MI355X.

Headline (`value`): steps/s of the 32^3 Kuhn cube (35,937 vertices, 196,608
tets, 12,288 surface triangles) dropped from 0.5 m onto the ground plane,
dt = 1/60, every step iterated until the relative force residual
||r|| / ||M g|| is <= --tol (1e-4), warm-started by the learned network, with
the particle state in FP64 on the device (an FP32 state cannot resolve a
residual of 1e-4 on this mesh).  The plain (inertial-start) leg runs the
same steps.  The same two legs run on the 512-scene batch (`batch`), where
every scene must reach the tolerance on its own.  A step is timed from its
launch to its completion, inputs resident in HBM; `steps` steps are timed
after `warmup` untimed ones.

Also reported: the reference's fixed count of 10 iterations per step with an
FP32 state (`fixed_iterations`), the roofline of the dominant kernel (the
per-colour vertex-block solve of the headline leg: algorithmic bytes per
launch over its average duration, measured with HIP events on the
library's stream), and the CPU baseline: the FP64 C++ restatement of the
reference's multithreaded CPU path on this machine's host cores, fixed-count
and residual-targeted, with the thread count and a one-thread figure.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "sim steps/sec at fixed force residual, warm-start vs plain; vertex-updates/sec"
HBM_PEAK_GBS = 8000.0
CPU_SHARE = 16   # host CPUs allotted to one GPU on the measurement machines


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_scene(name: str):
    from avbd_amd import scenes
    if name == "dropped_cube_32":
        return scenes.dropped_cube(n=32)
    if name == "dropped_cube_32_self":   # with the surface's vertex-face self collision (margin 1 mm)
        return scenes.dropped_cube(n=32, self_collision_margin=1e-3)
    if name == "hanging_cube_8":
        return scenes.hanging_cube(n=8)
    if name == "cube_batch_512":
        return scenes.cube_batch(n_scenes=512)
    if name == "rod_grid":
        return scenes.rod_grid()
    if name == "rod_grid_jointed":   # tip spheres as bodies held by two-sided fixed joints
        return scenes.rod_grid(tip_joint=True)
    if name == "rod_grid_contacts":  # ... and sphere-sphere contacts through the LBVH
        return scenes.rod_grid(tip_joint=True, contact_margin=1e-3)
    if name in ("ball_on_block", "rod_on_block", "hail_on_block"):
        return getattr(scenes, name)()
    raise SystemExit(f"unknown workload {name}")


def free_counts(sc):
    """(free vertices, their tet incidences, ground records on free vertices)."""
    n_items = n_inc = n_g = 0
    for m in sc.meshes:
        free = np.ones(m.rest.shape[0], bool) if m.fixed is None else (m.fixed == 0)
        val = np.bincount(m.tets.ravel().astype(np.int64), minlength=m.rest.shape[0])
        n_items += int(free.sum())
        n_inc += int(val[free].sum())
        n_g += int(free.sum()) if m.ground else 0
    return n_items, n_inc, n_g


def solve_bytes_per_launch(sc, n_colors: int, fp64: bool) -> float:
    """Average algorithmic bytes of one colour-phase launch of the vertex-block
    solve, with the same terms in both state precisions: own state (x read
    + write, y, mass), per incident tet its 64 B record (indices, rest shape,
    material) and the three other corners' positions, 48 B per ground record.
    FP32 state: 40 B + 100 B per incidence; FP64 state: 80 B + 136 B."""
    n_items, n_inc, n_g = free_counts(sc)
    if fp64:
        total = 80 * n_items + 136 * n_inc + 48 * n_g
    else:
        total = 40 * n_items + 100 * n_inc + 48 * n_g
    return total / n_colors


def residual_leg(workload: str, args, launch: dict, params=None, budget=None):
    """Every step iterated to the residual target; `warmup` untimed steps,
    then `steps` steps each timed from launch to completion."""
    from avbd_amd.sim import Simulation
    t0 = time.perf_counter()
    budget = budget or args.max_iters
    sim = Simulation(make_scene(workload), residual_target=(args.tol, budget, args.check_every),
                     use_graph=not args.res_eager, fp64_state=not args.fp32, **launch)
    sim.setup()
    sim.setup_colors = sim.ctx.coloring()[0].copy()
    if params is not None:
        sim.ctx.mlp_set_params(params)
        sim.ctx.set_warm_start(True)
    t_setup = time.perf_counter() - t0
    sim.ctx.step(args.warmup)
    sim.ctx.synchronize()
    s0 = sim.ctx.stats()
    per_ms, per_it = [], []
    prev_it = s0.iterations_total
    multi = len(sim.scene.meshes) > 1
    scene_budget_hits, scene_iters = 0, 0
    for _ in range(args.steps):
        t = time.perf_counter()
        sim.ctx.step(1)
        sim.ctx.synchronize()
        per_ms.append((time.perf_counter() - t) * 1e3)
        st = sim.ctx.stats()
        per_it.append(int(st.iterations_total - prev_it))
        prev_it = st.iterations_total
        if multi:
            its = sim.ctx.scene_iterations()
            scene_budget_hits += int((its >= budget).sum())
            scene_iters += int(its.sum())
    s1 = sim.ctx.stats()
    total_s = sum(per_ms) * 1e-3
    out = {
        "steps_per_s": round(args.steps / total_s, 3),
        "ms_per_step": round(total_s / args.steps * 1e3, 4),
        "mean_iterations": round(float(np.mean(per_it)), 2),
        "steps_at_budget": int(s1.steps_at_budget - s0.steps_at_budget),
        "vertex_updates_per_s": round((s1.vertex_updates - s0.vertex_updates) / total_s, 1),
        "iterations_per_step": per_it,
        "ms_per_step_each": [round(v, 3) for v in per_ms],
        "final_residual_rel": s1.residual_rel,
        "last_step_residual_history": [float(f"{v:.4g}") for v in sim.ctx.residual_history()][-40:],
        "setup_s": round(t_setup, 2),
    }
    if multi:
        n_sc = len(sim.scene.meshes)
        out["scene_steps_at_budget"] = scene_budget_hits
        out["scene_steps"] = n_sc * args.steps
        out["mean_scene_iterations"] = round(scene_iters / (n_sc * args.steps), 2)
        out["last_step_worst_scene_history"] = [float(f"{v:.4g}") for v in
                                                sim.ctx.residual_history(worst_scene=True)][-40:]
    if params is not None:
        out["predictions"] = int(s1.warm_start_predictions - s0.warm_start_predictions)
        out["fallbacks_to_inertial_start"] = int(s1.warm_start_fallbacks - s0.warm_start_fallbacks)
    lanes, block, batch = sim.ctx.launch_config()
    out["launch_config"] = {"lanes_per_vertex": lanes, "block_size": block, "gather_batch": batch,
                            "fused_chebyshev": sim.ctx.fused_chebyshev()}
    return sim, out


def warm_params(args):
    from avbd_amd import warm_start
    wpath = Path(args.warm_weights) if args.warm_weights else warm_start.DEFAULT_PATH
    if not wpath.exists():
        return None, None
    rel = str(wpath.relative_to(ROOT)) if wpath.is_relative_to(ROOT) else str(wpath)
    return warm_start.load(wpath), rel


def contact_aware_step(sim) -> bool:
    """Whether the next step of a dynamic-contact simulation launches the
    contact-aware solve variants: it does unless no contact is active and the
    work list is still the set-up one (same colours, nothing conflicted), when
    the step is the static one."""
    if any(len(sim.ctx.contacts(k)) for k in range(4)) or sim.ctx.conflicted().any():
        return True
    return not np.array_equal(sim.ctx.coloring()[0], sim.setup_colors)


def headline_pair(workload: str, args, launch: dict, budget=None):
    """Plain and warm-started residual legs; returns (plain, warm, the warm
    leg's simulation for kernel timing)."""
    sim_p, plain = residual_leg(workload, args, launch, budget=budget)
    log(f"{workload} residual leg (plain): {plain['steps_per_s']:.2f} steps/s, mean {plain['mean_iterations']} "
        f"iterations, {plain['steps_at_budget']} at budget")
    params, wrel = warm_params(args)
    warm, sim_w = None, None
    if params is not None:
        sim_w, warm = residual_leg(workload, args, launch, params, budget=budget)
        warm["weights"] = wrel
        log(f"{workload} residual leg (warm start): {warm['steps_per_s']:.2f} steps/s, mean "
            f"{warm['mean_iterations']} iterations, {warm['fallbacks_to_inertial_start']} fallbacks")
    sim_p.ctx.close()
    return plain, warm, sim_w


def cpu_info():
    model, phys = None, set()
    try:
        cur = {}
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if ":" in line:
                k, v = (t.strip() for t in line.split(":", 1))
                if k == "model name" and model is None:
                    model = v
                if k in ("physical id", "core id"):
                    cur[k] = v
            elif cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0))
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None,
            "usable_cpus": usable}


def timed_median(fn, runs=3):
    vals = [fn() for _ in range(runs)]
    return statistics.median(vals), vals


def cpu_baseline(sc, args, gpu_plain):
    """FP64 oracle (restatement of the reference's SPMD CPU path,
    simulation/Simulation.cpp:34 sizes it to the hardware threads): setup
    excluded, 5 warm-up steps, median of 3 bounded runs; fixed-count steps
    with all usable CPUs of this GPU's share and with one thread, then the
    residual-targeted steps of the headline window."""
    import dataclasses
    from oracle import oracle_py
    info = cpu_info()
    threads = max(1, min(CPU_SHARE, info["usable_cpus"]))
    note = ""
    if sc.self_collision and not (sc.rods or sc.spheres):
        # the reference's pipeline on this scene: its broad phase runs over
        # the surface triangles and its triangle-triangle narrow phase is
        # empty, so it runs the cube without self contacts (which the GPU run
        # does not find in these steps either); the SPMD path of the port
        sc = dataclasses.replace(sc, contact_margin=None, self_collision=False)
        note = ("; the reference's pipeline: broad phase over the surface triangles, empty triangle-triangle "
                "narrow phase")

    def fixed_rate(nthreads, budget):
        o = oracle_py.from_scene(sc)
        o.step(5, nthreads)

        def run():
            t0, n = time.perf_counter(), 0
            while True:
                o.step(1, nthreads)
                n += 1
                el = time.perf_counter() - t0
                if el >= budget or n >= 200:
                    return n / el
        return timed_median(run)

    fixed_mt, fixed_mt_runs = fixed_rate(threads, args.cpu_budget / 10)
    fixed_1, fixed_1_runs = fixed_rate(1, args.cpu_budget / 15)
    log(f"cpu fixed-count: {fixed_mt:.2f} steps/s ({threads} threads), {fixed_1:.3f} steps/s (1 thread)")
    # residual-targeted: the same warm-up, then the headline window's steps
    # until the budget is spent
    o = oracle_py.from_scene(sc)
    o.set_residual_target(args.tol, args.max_iters, args.check_every)
    o.step(args.warmup, threads)
    t_sum, its, k = 0.0, [], 0
    while k < args.steps and t_sum < args.cpu_budget / 2:
        t0 = time.perf_counter()
        o.step(1, threads)
        t_sum += time.perf_counter() - t0
        its.append(o.last_iters())
        k += 1
        log(f"cpu residual-targeted step {k}: {its[-1]} iterations, {t_sum:.1f} s so far")
    cpu_res = k / t_sum
    gpu_same = None
    if gpu_plain is not None:
        ms = gpu_plain["ms_per_step_each"][:k]
        gpu_same = round(k / (sum(ms) * 1e-3), 3)
    log(f"cpu residual-targeted: {cpu_res:.3f} steps/s over {k} steps ({sum(its)} iterations)")
    n_v = sum(m.rest.shape[0] for m in sc.meshes)
    return {
        "value": round(cpu_res, 4),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sc.name}: steps {args.warmup}..{args.warmup + k - 1} iterated to relative force residual "
                   f"{args.tol:g} (check every {args.check_every}), FP64, {threads} SPMD threads with a spin barrier "
                   f"per phase, after {args.warmup} warm-up steps at the same target; {sum(its)} iterations{note}"),
        "iterations_per_step": its,
        "gpu_plain_steps_per_s_same_steps": gpu_same,
        "fixed_iterations": {"iterations": sc.iters, "steps_per_s": round(fixed_mt, 3), "runs": [round(v, 3) for v in fixed_mt_runs],
                             "vertex_updates_per_s": round(fixed_mt * sc.iters * n_v, 1),
                             "one_thread_steps_per_s": round(fixed_1, 4),
                             "one_thread_runs": [round(v, 4) for v in fixed_1_runs]},
        "threads_rule": (f"all usable CPUs of this GPU's share (min of the process affinity, "
                         f"{info['usable_cpus']}, and the {CPU_SHARE} CPUs allotted per GPU)"),
        **info,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the residual-targeted legs")
    ap.add_argument("--warmup", type=int, default=5)
    # the 32^3 cube as configured: ground contact and the surface's
    # vertex-face self collision through the LBVH every step
    ap.add_argument("--workload", default="dropped_cube_32_self")
    ap.add_argument("--batch-workload", default="cube_batch_512")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch legs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=30.0, help="seconds of CPU work of the baseline")
    ap.add_argument("--tol", type=float, default=1e-4, help="relative force residual target")
    ap.add_argument("--max-iters", type=int, default=50000, help="iteration budget per step")
    ap.add_argument("--batch-max-iters", type=int, default=20000, help="iteration budget per step of the batch legs")
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--fp32", action="store_true", help="FP32 particle state in the residual legs")
    ap.add_argument("--fixed-steps", type=int, default=200, help="timed steps of the fixed-iteration leg")
    ap.add_argument("--no-fixed", action="store_true", help="skip the fixed-iteration leg")
    ap.add_argument("--no-residual", action="store_true", help="skip the residual-targeted legs of the workload")
    ap.add_argument("--lanes-per-vertex", type=int, default=0, help="0: autotuned at set-up")
    ap.add_argument("--block-size", type=int, default=0, help="0: autotuned at set-up")
    ap.add_argument("--gather-batch", type=int, default=0, help="0: autotuned at set-up")
    ap.add_argument("--res-eager", action="store_true",
                    help="launch the residual legs eagerly instead of replaying step graphs")
    ap.add_argument("--warm-weights", default=None,
                    help="trained warm-start parameters (.npz); default: avbd_amd/data/warm_start_mlp.npz")
    ap.add_argument("--calibrate", type=int, default=0,
                    help="also run one plain copy kernel of this many bytes (HBM counter calibration)")
    ap.add_argument("--solve-only", type=int, default=0,
                    help="profiling run: set up the headline workload, take the warm-up steps, time this many "
                         "rounds of its colour-phase solve launches (the roofline's kernel) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the solver on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    from avbd_amd.sim import Simulation

    sc = make_scene(args.workload)
    log(f"workload {sc.name}: {sc.n_vertices} vertices, {sc.n_tets} tets, {sc.n_ground} ground contacts")
    launch = dict(lanes_per_vertex=args.lanes_per_vertex, block_size=args.block_size, gather_batch=args.gather_batch)
    dynamic = sc.contact_margin is not None
    fp64 = not args.fp32

    if args.solve_only:
        sim = Simulation(sc, residual_target=(args.tol, args.max_iters, args.check_every), fp64_state=fp64,
                         use_graph=False, **launch)
        sim.setup()
        sim.ctx.step(args.warmup)
        sim.ctx.synchronize()
        avg_ms, launches = sim.ctx.time_solve_launches(args.solve_only)
        if args.calibrate:
            ms = sim.ctx.calibrate_copy(args.calibrate, 10)
            log(f"copy kernel: {args.calibrate} bytes in {ms:.4f} ms ({2 * args.calibrate / (ms * 1e-3) / 1e9:.1f} GB/s "
                f"read + write)")
        _, n_colors = sim.ctx.coloring()
        lanes, block, batch = sim.ctx.launch_config()
        nb = solve_bytes_per_launch(sc, n_colors, fp64)
        log(f"solve launches: {launches} x {avg_ms * 1e3:.3f} us, {nb:.0f} algorithmic bytes each "
            f"({nb / (avg_ms * 1e-3) / 1e9:.1f} GB/s); lanes {lanes} block {block} gather batch {batch}")
        sim.ctx.close()
        return

    # headline: residual-targeted legs
    res, roof, sim_w = None, None, None
    # residual legs: static scenes, and dynamic ones whose contacts the FP64
    # state takes (vertex-face self collision between triangles)
    res_ok = not dynamic or args.fp32 or not (sc.rods or sc.spheres)
    if res_ok and not args.no_residual:
        plain, warm, sim_w = headline_pair(args.workload, args, launch)
        best = warm if warm is not None else plain
        res = {"tol": args.tol, "max_iters": args.max_iters, "check_every": args.check_every,
               "state": "fp64" if fp64 else "fp32", "steps": args.steps, "warmup": args.warmup,
               "plain": plain, "warm_start": warm,
               "speedup_warm_vs_plain": round(warm["steps_per_s"] / plain["steps_per_s"], 3) if warm else None,
               "steps_per_s": best["steps_per_s"], "ms_per_step": best["ms_per_step"],
               "vertex_updates_per_s": best["vertex_updates_per_s"]}
        # dominant kernel of the headline leg: the per-colour vertex solve
        ksim = sim_w
        if ksim is not None:
            avg_ms, launches = ksim.ctx.time_solve_launches(20)
            _, n_colors = ksim.ctx.coloring()
            bytes_launch = solve_bytes_per_launch(sc, n_colors, fp64)
            achieved = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
            kernel = "solve_color_f64_kernel" if fp64 else "solve_color_kernel"
            roof = {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_launch_us": round(avg_ms * 1e3, 3), "algorithmic_bytes_per_launch": int(bytes_launch),
                    "launches_timed": launches}
            # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE
            # runs, profiles/), used only when they were taken on the variant
            # timed here
            lanes, _, batch = ksim.ctx.launch_config()
            damp = any(m.kd > 0 for m in sc.meshes)
            if dynamic and contact_aware_step(ksim):   # one incidence gathered per lane in FP64
                variant = (f"solve_color_f64_kernel<{lanes}, {str(damp).lower()}, 1, true>" if fp64 else
                           f"solve_color_kernel<{lanes}, {str(damp).lower()}, true, {batch}>")
            else:
                variant = (f"solve_color_f64_kernel<{lanes}, {str(damp).lower()}, {batch}, false>" if fp64 else
                           f"solve_color_kernel<{lanes}, {str(damp).lower()}, false, {batch}>")
            roof["variant"] = variant
            # the workload's own profile, else the one of its scene (the same
            # cube without the self-collision pass runs the same static kernel)
            for stem in (args.workload, sc.name):
                prof = ROOT / "profiles" / f"{kernel}_{stem}_traffic.json"
                t = json.loads(prof.read_text()) if prof.exists() else {}
                if t.get("variant") == variant:
                    roof["traffic"] = t.get("bytes_per_launch")
                    roof["traffic_source"] = t.get("source")
                    break
            ksim.ctx.close()

    # the reference's fixed iteration count, FP32 state
    fixed, calib, lbvh_ms = None, None, 0.0
    if not args.no_fixed:
        t_setup = time.perf_counter()
        sim = Simulation(sc, **launch)
        sim.setup()
        t_setup = time.perf_counter() - t_setup
        sim.ctx.step(args.warmup)
        sim.ctx.synchronize()
        t0 = time.perf_counter()
        sim.ctx.step(args.fixed_steps)
        sim.ctx.synchronize()
        el = time.perf_counter() - t0
        st = sim.ctx.stats()
        n_detected = len(sim.ctx.detected()[0]) if dynamic else None
        fixed_rate = args.fixed_steps / el
        vups = st.vertex_updates / max(st.steps_taken, 1) * fixed_rate
        sim.ctx.set_timing(True)
        sim.ctx.step(2)
        lbvh_ms = sim.ctx.broadphase_timing()
        sim.ctx.set_timing(False)
        f_avg_ms, f_launches = sim.ctx.time_solve_launches(20)
        _, n_colors = sim.ctx.coloring()
        f_bytes = solve_bytes_per_launch(sc, n_colors, False)
        lanes, block, batch = sim.ctx.launch_config()
        fixed = {"iterations": sc.iters, "state": "fp32", "steps": args.fixed_steps,
                 "steps_per_s": round(fixed_rate, 3), "ms_per_step": round(el / args.fixed_steps * 1e3, 4),
                 "vertex_updates_per_s": round(vups, 1), "final_residual_rel": st.residual_rel,
                 "setup_s": round(t_setup, 2),
                 "contacts_detected_last_step": n_detected,
                 "solve_kernel": {"kernel": "solve_color_kernel", "avg_launch_us": round(f_avg_ms * 1e3, 3),
                                  "algorithmic_bytes_per_launch": int(f_bytes),
                                  "achieved_GBps": round(f_bytes / (f_avg_ms * 1e-3) / 1e9, 1) if f_avg_ms > 0 else 0},
                 "launch_config": {"lanes_per_vertex": lanes, "block_size": block, "gather_batch": batch,
                                   "fused_chebyshev": sim.ctx.fused_chebyshev()}}
        log(f"fixed-count leg: {fixed_rate:.1f} steps/s, residual_rel {st.residual_rel:.3e}")
        if roof is None and f_launches:
            roof = {"kernel": "solve_color_kernel", "bound": "hbm", "achieved": fixed["solve_kernel"]["achieved_GBps"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(fixed["solve_kernel"]["achieved_GBps"] / HBM_PEAK_GBS, 4),
                    "traffic": None, "avg_launch_us": round(f_avg_ms * 1e3, 3),
                    "algorithmic_bytes_per_launch": int(f_bytes), "launches_timed": f_launches}
        calib = None
        if args.calibrate:
            ms = sim.ctx.calibrate_copy(args.calibrate, 10)
            calib = {"bytes": args.calibrate, "avg_ms": ms, "GBps_read_plus_write": 2 * args.calibrate / (ms * 1e-3) / 1e9}
        sim.ctx.close()

    # the same legs on the scene batch
    batch_res = None
    if not args.no_batch and args.batch_workload != args.workload and not (sc.rods or sc.spheres):
        bsc = make_scene(args.batch_workload)
        log(f"batch {bsc.name}: {len(bsc.meshes)} scenes, {bsc.n_vertices} vertices, {bsc.n_tets} tets")
        bplain, bwarm, bsim = headline_pair(args.batch_workload, args, launch, budget=args.batch_max_iters)
        if bsim is not None:
            bsim.ctx.close()
        batch_res = {"workload": bsc.name, "scenes": len(bsc.meshes), "vertices": bsc.n_vertices,
                     "tets": bsc.n_tets, "tol": args.tol, "max_iters": args.batch_max_iters,
                     "rule": "every scene reaches the tolerance on its own; a scene that has stops iterating",
                     "plain": bplain, "warm_start": bwarm,
                     "speedup_warm_vs_plain": round(bwarm["steps_per_s"] / bplain["steps_per_s"], 3) if bwarm else None}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(sc, args, res["plain"] if res else None)

    n_tris = sum(len(m.triangles) for m in sc.meshes if m.triangles is not None) if len(sc.meshes) == 1 else 0
    head_value = res["steps_per_s"] if res else fixed["steps_per_s"]
    head_ms = res["ms_per_step"] if res else fixed["ms_per_step"]
    out = {
        "metric": METRIC,
        "value": head_value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps if res else args.fixed_steps,
        "warmup": args.warmup,
        "ms_per_step": head_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if fp64 else "f32") if res else "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "vertices": sc.n_vertices,
            "tets": sc.n_tets,
            "ground_contacts": sc.n_ground,
            "dt": sc.dt,
            "mode": (f"AVBD iterating each step to relative force residual {args.tol:g} (budget {args.max_iters}, "
                     f"checked every {args.check_every}); {'FP64' if fp64 else 'FP32'} particle state; "
                     + ("warm-started by the learned MLP" if res and res["warm_start"] else "plain start")) if res
                    else f"plain AVBD, {sc.iters} iterations per step",
            "broad_phase": (f"LBVH over the {n_tris:,} surface triangles (widened by the {sc.contact_margin:g} m "
                            f"margin) rebuilt and self-traversed, vertex-face narrow phase, every step"
                            if n_tris and sc.self_collision else
                            f"LBVH over the {n_tris:,} surface triangles rebuilt and self-traversed every step"
                            if n_tris else "none"),
            "self_collision": bool(sc.self_collision),
            "parallelism": "single GPU",
        },
        "fixed_residual": res,
        "batch": batch_res,
        "fixed_iterations": fixed,
        "vertex_updates_per_s": res["vertex_updates_per_s"] if res else round(vups, 1),
        "step_breakdown_us": ({"detection_round_host_wall": round(lbvh_ms * 1e3, 2)} if dynamic else
                              {"broad_phase_concurrent": round(lbvh_ms * 1e3, 2)}),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if calib:
        out["copy_calibration"] = calib
    if sc.rods or sc.spheres:
        tips = [r for r in sc.rods if r.tip_sphere is not None]
        out["config"].update({
            "workload": args.workload,
            "rods": len(sc.rods), "rod_nodes": sum(r.elements + 1 for r in sc.rods),
            "rigid_spheres": len(sc.spheres) + sum(1 for r in tips if r.tip_joint),
            "tip_spheres": ("bodies held by two-sided fixed joints" if tips and all(r.tip_joint for r in tips)
                            else "welded into the tip nodes" if tips else "none"),
            "broad_phase": (f"LBVH over the spheres' collision primitives every step (margin {sc.contact_margin:g})"
                            if dynamic else "none"),
        })
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
