"""Benchmark of the per-step AVBD solve on one MI355X.

Workload (default): the 32^3 Kuhn cube (35,937 vertices, 196,608 tets) dropped
from 0.5 m onto the ground plane with AL contact on every vertex, dt = 1/60,
10 VBD iterations per step, synthetic seeded input.  A step is one full
solve_Parallel on the device (hipGraph replay); inputs are resident in HBM.

Prints one JSON line.  `roofline` is for the dominant kernel (the per-colour
vertex-block solve): algorithmic bytes per launch (FP32 gather model: 40 B own
state + 144 B per incident tet + 48 B per ground record, per vertex block)
divided by its average launch duration measured with HIP events on the
library's stream.  `cpu_baseline` times the FP64 C++ restatement of the
reference's multithreaded CPU path on this machine's host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "sim steps/sec at fixed force residual, warm-start vs plain; vertex-updates/sec"
HBM_PEAK_GBS = 8000.0


def make_scene(name: str):
    from avbd_amd import scenes
    if name == "dropped_cube_32":
        return scenes.dropped_cube(n=32)
    if name == "hanging_cube_8":
        return scenes.hanging_cube(n=8)
    if name == "cube_batch_512":
        return scenes.cube_batch(n_scenes=512)
    if name == "rod_grid":
        return scenes.rod_grid()
    if name in ("ball_on_block", "rod_on_block", "hail_on_block"):
        return getattr(scenes, name)()
    raise SystemExit(f"unknown workload {name}")


def solve_bytes_per_launch(sim) -> float:
    """Average algorithmic bytes of one colour-phase launch."""
    sc = sim.scene
    n_items = 0
    n_inc = 0
    n_g = 0
    for m in sc.meshes:
        free = np.ones(m.rest.shape[0], bool) if m.fixed is None else (m.fixed == 0)
        val = np.bincount(m.tets.ravel().astype(np.int64), minlength=m.rest.shape[0])
        n_items += int(free.sum())
        n_inc += int(val[free].sum())
        n_g += int(free.sum()) if m.ground else 0
    _, k = sim.ctx.coloring()
    total = 40 * n_items + 144 * n_inc + 48 * n_g
    return total / k


def step_bytes(sc) -> int:
    """Algorithmic bytes of one whole step (FP32 gather model)."""
    from avbd_amd.scenes import algorithmic_bytes_per_step
    n_v = sum(int((np.ones(m.rest.shape[0]) if m.fixed is None else (m.fixed == 0)).sum()) for m in sc.meshes)
    return algorithmic_bytes_per_step(n_v, sc.n_tets, sc.n_ground, sc.iters)


def cpu_baseline_batch(sc, budget_s: float, threads: int, steps_per_scene: int = 5):
    """A batch of independent scenes on the CPU path: the reference runs one
    Simulation per scene, so each scene gets its own oracle, stepped with
    `threads` SPMD threads; scenes are timed in order until the budget is
    spent and the time of one batch step is extrapolated by tet count."""
    from oracle import oracle_py
    t_sum, tets_done, k = 0.0, 0, 0
    total_tets = sum(m.tets.shape[0] for m in sc.meshes)
    t_start = time.perf_counter()
    for k, m in enumerate(sc.meshes):
        o = oracle_py.from_scene(sc, mesh_index=k)
        o.step(1, threads)
        t0 = time.perf_counter()
        o.step(steps_per_scene, threads)
        t_sum += (time.perf_counter() - t0) / steps_per_scene
        tets_done += m.tets.shape[0]
        if time.perf_counter() - t_start >= budget_s:
            break
    t_step = t_sum * total_tets / tets_done
    return {
        "value": 1.0 / t_step,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sc.name}: {k + 1} of {len(sc.meshes)} scenes, one oracle per scene (one reference Simulation "
                  f"each), {steps_per_scene} whole steps ({sc.iters} iterations each) per scene after 1 warm-up "
                  f"step, FP64, {threads} SPMD threads; batch step time extrapolated by tet count "
                  f"({tets_done} of {total_tets} tets timed)",
        "vertex_updates_per_s": sc.iters * sum(m.rest.shape[0] for m in sc.meshes) / t_step,
    }


def cpu_baseline(sc, budget_s: float = 12.0):
    """FP64 oracle (restatement of the reference's SPMD CPU path) on the host
    cores: setup excluded, 1 warm-up step, then whole steps until ~budget_s."""
    from oracle import oracle_py
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    if sc.contact_margin is not None:
        threads = 1   # the reference runs detection, lifecycle and recolouring on one worker; the port steps serially
    if getattr(sc, "batch", True) and len(sc.meshes) > 1:
        return cpu_baseline_batch(sc, budget_s, threads)
    o = oracle_py.from_scene(sc)
    o.step(1, threads)
    t0 = time.perf_counter()
    n = 0
    while True:
        o.step(1, threads)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 400:
            break
    return {
        "value": n / el,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sc.name}: {n} whole steps ({sc.iters} iterations each) after 1 warm-up step, "
                  + (f"FP64, {threads} SPMD threads with a spin barrier per phase" if threads > 1
                     else "FP64, one thread"),
        "vertex_updates_per_s": n * sc.iters * sum(m.rest.shape[0] for m in sc.meshes) / el,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dropped_cube_32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--tol", type=float, default=0.1, help="relative force residual target of the headline leg")
    ap.add_argument("--max-iters", type=int, default=400, help="iteration budget per step of the residual leg")
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--res-steps", type=int, default=60, help="timed steps of the residual-targeted leg")
    ap.add_argument("--lanes-per-vertex", type=int, default=0, help="0: autotuned at set-up")
    ap.add_argument("--block-size", type=int, default=0, help="0: autotuned at set-up")
    ap.add_argument("--gather-batch", type=int, default=0, help="0: autotuned at set-up")
    ap.add_argument("--res-eager", action="store_true",
                    help="launch the residual-targeted leg eagerly instead of replaying its step graphs "
                         "(for kernel traces of that leg; not the measured configuration)")
    ap.add_argument("--warm-weights", default=None,
                    help="trained warm-start parameters (.npz); default: avbd_amd/data/warm_start_mlp.npz if present")
    ap.add_argument("--calibrate", type=int, default=0,
                    help="also run one plain copy kernel of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the solver on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    from avbd_amd.sim import Simulation

    sc = make_scene(args.workload)
    print(f"[bench] workload {sc.name}: {sc.n_vertices} vertices, {sc.n_tets} tets, "
          f"{sc.n_ground} ground contacts, {sc.iters} iterations", file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    launch = dict(lanes_per_vertex=args.lanes_per_vertex, block_size=args.block_size, gather_batch=args.gather_batch)
    sim = Simulation(sc, **launch)
    sim.setup()
    t_setup = time.perf_counter() - t_setup
    lanes, block, batch = sim.ctx.launch_config()
    print(f"[bench] setup {t_setup:.2f}s, launch config lanes={lanes} block={block} batch={batch}", file=sys.stderr, flush=True)

    dynamic = sc.contact_margin is not None
    if dynamic:
        args.res_steps = 0   # residual targets are not available with dynamic contacts
    sim.ctx.step(args.warmup)
    sim.ctx.synchronize()
    t0 = time.perf_counter()
    sim.ctx.step(args.steps)
    sim.ctx.synchronize()
    el = time.perf_counter() - t0
    steps_per_s = args.steps / el
    st = sim.ctx.stats()
    # vertex-block solves per second (positional and oriented work items)
    vups = st.vertex_updates / max(st.steps_taken, 1) * steps_per_s
    print(f"[bench] {args.steps} steps in {el * 1e3:.1f} ms -> {steps_per_s:.1f} steps/s, "
          f"residual_rel {st.residual_rel:.3e}", file=sys.stderr, flush=True)

    # live kernel timing of the dominant kernel (HIP events on the ctx stream)
    sim.ctx.set_timing(True)
    sim.ctx.step(2)
    avg_ms, launches = sim.ctx.solve_timing()
    lbvh_ms = sim.ctx.broadphase_timing()
    sim.ctx.set_timing(False)
    pgrid = sim.ctx.persistent_grid()
    if not pgrid:
        # average solve-launch duration from back-to-back launches (the per-step
        # timing above brackets each eager launch and includes its dispatch gap)
        avg_ms, launches = sim.ctx.time_solve_launches(20)
    if pgrid:
        kernel = "step_persistent_kernel"        # one launch = one whole solve
        bytes_launch = step_bytes(sc)
    else:
        kernel = "solve_color_kernel"            # one launch = one colour phase
        bytes_launch = solve_bytes_per_launch(sim)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roof = {
        "kernel": kernel,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_us": round(avg_ms * 1e3, 3),
        "algorithmic_bytes_per_launch": int(bytes_launch),
        "launches_timed": launches,
    }
    if launches == 0:
        # no positional vertex blocks (an all-oriented scene such as the rod
        # grid): the colour-phase kernel never runs, so there is no roofline
        roof = None
    prof = ROOT / "profiles" / f"{kernel}_{sc.name}_traffic.json"
    if roof is not None and prof.exists():
        try:
            roof["traffic"] = json.loads(prof.read_text()).get("bytes_per_launch")
        except Exception:
            pass

    # headline leg: AVBD iterating each step to the residual target, plain
    # and (when trained parameters exist) warm-started by the network
    def residual_leg(params=None):
        rsim = Simulation(make_scene(args.workload), residual_target=(args.tol, args.max_iters, args.check_every),
                          use_graph=not args.res_eager, **launch)
        rsim.setup()
        if params is not None:
            rsim.ctx.mlp_set_params(params)
            rsim.ctx.set_warm_start(True)
        rsim.ctx.step(args.warmup)
        rsim.ctx.synchronize()
        s0 = rsim.ctx.stats()
        t0 = time.perf_counter()
        rsim.ctx.step(args.res_steps)
        rsim.ctx.synchronize()
        rel = time.perf_counter() - t0
        s1 = rsim.ctx.stats()
        iters = s1.iterations_total - s0.iterations_total
        out = {
            "steps_per_s": round(args.res_steps / rel, 3),
            "ms_per_step": round(rel / args.res_steps * 1e3, 4),
            "mean_iterations": round(iters / args.res_steps, 2),
            "steps_at_budget": int(s1.steps_at_budget - s0.steps_at_budget),
            "vertex_updates_per_s": round((s1.vertex_updates - s0.vertex_updates) / rel, 1),
            "last_step_iterations": s1.last_step_iterations,
            "last_step_residual_history": [float(f"{v:.4g}") for v in rsim.ctx.residual_history()],
            "final_residual_rel": s1.residual_rel,
        }
        if params is not None:
            out["predictions"] = int(s1.warm_start_predictions - s0.warm_start_predictions)
            out["fallbacks_to_inertial_start"] = int(s1.warm_start_fallbacks - s0.warm_start_fallbacks)
        rsim.ctx.close()
        return out

    res = None
    if args.res_steps > 0:
        plain = residual_leg()
        print(f"[bench] residual leg (plain): {plain['steps_per_s']:.1f} steps/s, mean {plain['mean_iterations']} "
              f"iterations, final residual {plain['final_residual_rel']:.3e}", file=sys.stderr, flush=True)
        warm = None
        from avbd_amd import warm_start
        wpath = Path(args.warm_weights) if args.warm_weights else warm_start.DEFAULT_PATH
        if wpath.exists():
            warm = residual_leg(warm_start.load(wpath))
            warm["weights"] = str(wpath.relative_to(ROOT)) if wpath.is_relative_to(ROOT) else str(wpath)
            print(f"[bench] residual leg (warm start): {warm['steps_per_s']:.1f} steps/s, mean "
                  f"{warm['mean_iterations']} iterations, {warm['fallbacks_to_inertial_start']} fallbacks",
                  file=sys.stderr, flush=True)
        best = warm if warm is not None else plain
        res = {"tol": args.tol, "max_iters": args.max_iters, "check_every": args.check_every,
               "steps": args.res_steps, "plain": plain, "warm_start": warm,
               "speedup_warm_vs_plain": round(warm["steps_per_s"] / plain["steps_per_s"], 3) if warm else None,
               "steps_per_s": best["steps_per_s"], "ms_per_step": best["ms_per_step"],
               "vertex_updates_per_s": best["vertex_updates_per_s"]}

    calib = None
    if args.calibrate:
        ms = sim.ctx.calibrate_copy(args.calibrate, 1)
        calib = {"bytes": args.calibrate, "avg_ms": ms, "GBps_read_plus_write": 2 * args.calibrate / (ms * 1e-3) / 1e9}

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sc, args.cpu_budget)

    _, n_colors = sim.ctx.coloring()
    solve_us = avg_ms * 1e3 * n_colors * sc.iters
    n_tris = sum(len(m.triangles) for m in sc.meshes if m.triangles is not None) if len(sc.meshes) == 1 else 0
    if dynamic:
        keys, _, _ = sim.ctx.detected()
        n_prims = n_tris + sum(r.elements for r in sc.rods) + len(sc.spheres)
        broad = (f"LBVH over {n_prims:,} primitives (triangles, rod segments, spheres), FP64 narrow phase, "
                 f"host contact lifecycle and incremental recolouring, every step")
        breakdown = {"detection_round_host_wall": round(lbvh_ms * 1e3, 2), "solve": round(solve_us, 2)}
        dyn_info = {"contacts_last_step": int(len(keys)), "margin": sc.contact_margin}
    else:
        broad = (f"LBVH over the {n_tris:,} surface triangles rebuilt and self-traversed every step" if n_tris
                 else "none (no collision triangles)")
        breakdown = {"broad_phase_concurrent": round(lbvh_ms * 1e3, 2), "solve": round(solve_us, 2)}
        dyn_info = None
    head_value = res["steps_per_s"] if res else round(steps_per_s, 3)
    head_ms = res["ms_per_step"] if res else round(el / args.steps * 1e3, 4)
    out = {
        "metric": METRIC,
        "value": head_value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.res_steps if res else args.steps,
        "warmup": args.warmup,
        "ms_per_step": head_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": sc.name,
            "vertices": sc.n_vertices,
            "tets": sc.n_tets,
            "ground_contacts": sc.n_ground,
            "iterations": sc.iters,
            "dt": sc.dt,
            "mode": (f"AVBD iterating each step to relative force residual {args.tol:g} "
                     f"(budget {args.max_iters}, checked every {args.check_every}); "
                     + ("warm-started by the learned MLP" if res and res["warm_start"] else "plain start")) if res
                    else "plain AVBD, fixed iteration count",
            "broad_phase": broad,
            "self_collision": False,
            "parallelism": "single GPU",
        },
        "fixed_residual": res,
        "fixed_iterations": {"iterations": sc.iters, "steps": args.steps, "steps_per_s": round(steps_per_s, 3),
                             "ms_per_step": round(el / args.steps * 1e3, 4),
                             "vertex_updates_per_s": round(vups, 1), "final_residual_rel": st.residual_rel},
        "vertex_updates_per_s": res["vertex_updates_per_s"] if res else round(vups, 1),
        "step_breakdown_us": breakdown,
        "roofline": roof,
        "cpu_baseline": cpu,
        "launch_config": {"lanes_per_vertex": lanes, "block_size": block, "gather_batch": batch,
                          "persistent_grid": pgrid, "fused_chebyshev": sim.ctx.fused_chebyshev()},
    }
    if dyn_info:
        out["dynamic_contacts"] = dyn_info
    if calib:
        out["copy_calibration"] = calib
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
