"""Learned warm start: training-data generation with the plain GPU solver,
training of the per-vertex MLP on the GPU (the library's BF16 MFMA forward and
backward kernels, Adam on FP32 master weights), and storage of the trained
parameters.

The network predicts, per free vertex, the correction from the inertial /
adaptive initial guess to the converged end-of-step position, in units of
dt^2 g; the solver starts from the prediction when it lowers the scene's
incremental potential and then runs plain VBD iterations as usual.

    python -m avbd_amd.warm_start --out avbd_amd/data/warm_start_mlp.npz
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path
from typing import List, Tuple

import numpy as np

from . import scenes
from .scenes import SplitMix64

DEFAULT_PATH = Path(__file__).resolve().parent / "data" / "warm_start_mlp.npz"


def training_scenes(n_scenes: int, seed: int, n_min: int = 8, n_max: int = 14, log_e: Tuple[float, float] = (5.0, 7.0),
                    nu: Tuple[float, float] = (0.3, 0.48), height: Tuple[float, float] = (0.1, 0.6),
                    field_sigma: float = 0.0, field_fraction: float = 0.0) -> List[scenes.Scene]:
    """Dropped Kuhn cubes in the style of the 32^3 workload: random size
    (n_min..n_max cells per axis), Young's modulus (log-uniform over log_e),
    Poisson ratio, rotation and drop height (uniform over the ranges); a
    fraction field_fraction of them also carries a seeded Gaussian per-vertex
    external acceleration field of standard deviation field_sigma (N/kg)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(n_scenes):
        n = rng.randint(n_min, n_max)
        E = float(10 ** (log_e[0] + (log_e[1] - log_e[0]) * rng.uniform()))
        nu_i = nu[0] + (nu[1] - nu[0]) * rng.uniform()
        h = height[0] + (height[1] - height[0]) * rng.uniform()
        with_field = field_sigma > 0 and rng.uniform() < field_fraction
        sc = scenes.dropped_cube(n=n, seed=1000 + 7 * seed + i, height=h, E=E, nu=nu_i)
        if with_field:
            m = sc.meshes[0]
            m.ext_accel = scenes.gaussian_force_field(m.rest.shape[0], field_sigma, seed=5000 + 7 * seed + i)
        out.append(sc)
    return out


def generate(scene_list, steps: int, tol: float, max_iters: int, check_every: int = 10, keep: float = 0.25,
             seed: int = 0, log=print, fp64: bool = False, per_batch: int = 1,
             chunk: int = 16) -> Tuple[np.ndarray, np.ndarray]:
    """Run the scenes for `steps` steps with the plain solver iterating every
    step to the residual target, recording features and targets; keep a
    random fraction of the rows.  `per_batch` scenes run together as one
    batch (each still stops at its own target); the steps are recorded
    `chunk` at a time."""
    from .sim import Simulation
    rng = np.random.default_rng(seed)
    Xs, Ts = [], []
    for b0 in range(0, len(scene_list), per_batch):
        group = scene_list[b0:b0 + per_batch]
        sc = group[0] if len(group) == 1 else scenes.Scene(
            name=f"train_batch_{b0}", meshes=[m for g in group for m in g.meshes], iters=group[0].iters,
            rho=group[0].rho, dt=group[0].dt, g=group[0].g, steps=steps)
        sim = Simulation(sc, residual_target=(tol, max_iters, check_every), fp64_state=fp64)
        sim.setup()
        n_items = int(sum(int((m.fixed == 0).sum()) if m.fixed is not None else m.rest.shape[0] for m in sc.meshes))
        kept, its, at_budget = 0, 0, 0
        for k0 in range(0, steps, chunk):
            n = min(chunk, steps - k0)
            X, T = sim.ctx.record_steps(n, n_items)
            sel = rng.random(X.shape[0]) < keep
            Xs.append(X[sel])
            Ts.append(T[sel])
            kept += int(sel.sum())
        st = sim.ctx.stats()
        log(f"[warm-start] scenes {b0}..{b0 + len(group) - 1}: {sc.n_vertices} vertices, {steps} steps, "
            f"{st.iterations_total} iterations, {st.steps_at_budget} steps at the budget, {kept} rows kept")
        sim.ctx.close()
    return np.concatenate(Xs), np.concatenate(Ts)


def normalization(X: np.ndarray, T: np.ndarray) -> np.ndarray:
    fm, fs = X.mean(0), X.std(0)
    fs = np.where(fs > 1e-6, fs, 1.0)
    tm, ts = T.mean(0), T.std(0)
    ts = np.where(ts > 1e-9, ts, 1.0)
    return np.concatenate([fm, fs, tm, ts]).astype(np.float32)


def train(ctx, X: np.ndarray, T: np.ndarray, epochs: int = 8, batch: int = 4096, lr: float = 1e-3,
          seed: int = 0) -> Tuple[np.ndarray, float]:
    """Train on the device; returns the parameter vector and the last epoch's loss."""
    ctx.mlp_init(seed)
    ctx.mlp_set_normalization(normalization(X, T))
    batch = min(batch, (X.shape[0] // 16) * 16)
    loss = ctx.mlp_train(X, T, epochs, batch, lr, seed)
    return ctx.mlp_params(), loss


def save(path: Path, params: np.ndarray, **meta) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=params.astype(np.float32), **{k: np.asarray(v) for k, v in meta.items()})


def load(path: Path = DEFAULT_PATH) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_PATH))
    ap.add_argument("--scenes", type=int, default=128)
    ap.add_argument("--steps", type=int, default=160)
    ap.add_argument("--per-batch", type=int, default=64, help="scenes stepped together as one batch")
    ap.add_argument("--keep", type=float, default=0.1)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--max-iters", type=int, default=4000)
    ap.add_argument("--fp32", action="store_true", help="FP32 particle state while generating (default FP64)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--n-min", type=int, default=8)
    ap.add_argument("--n-max", type=int, default=16)
    ap.add_argument("--log-e", type=float, nargs=2, default=(5.5, 6.5))
    ap.add_argument("--nu", type=float, nargs=2, default=(0.3, 0.45))
    ap.add_argument("--height", type=float, nargs=2, default=(0.1, 0.6))
    ap.add_argument("--field-sigma", type=float, default=10.0,
                    help="standard deviation (N/kg) of the Gaussian per-vertex force fields")
    ap.add_argument("--field-fraction", type=float, default=0.5, help="fraction of scenes that carry a field")
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    t0 = time.perf_counter()
    sc = training_scenes(args.scenes, args.seed, args.n_min, args.n_max, tuple(args.log_e), tuple(args.nu),
                         tuple(args.height), args.field_sigma, args.field_fraction)
    X, T = generate(sc, args.steps, args.tol, args.max_iters, keep=args.keep, log=log, fp64=not args.fp32,
                    per_batch=args.per_batch)
    log(f"[warm-start] {X.shape[0]} rows in {time.perf_counter() - t0:.1f}s")
    from .sim import SimulationContext
    ctx = SimulationContext(dt=1 / 60, solver_iters=1)
    ctx.add_particles(np.zeros((1, 3)), mass=np.ones(1))
    ctx.build_graph()
    t1 = time.perf_counter()
    params, loss = train(ctx, X, T, args.epochs, args.batch, args.lr, args.seed)
    log(f"[warm-start] trained {args.epochs} epochs in {time.perf_counter() - t1:.1f}s, final loss {loss:.4f}")
    save(Path(args.out), params, rows=X.shape[0], loss=loss, tol=args.tol, scenes=args.scenes, steps=args.steps,
         max_iters=args.max_iters, epochs=args.epochs, seed=args.seed, n_range=(args.n_min, args.n_max),
         log_e=tuple(args.log_e), nu=tuple(args.nu), height=tuple(args.height), keep=args.keep,
         field_sigma=args.field_sigma, field_fraction=args.field_fraction, fp64=not args.fp32,
         per_batch=args.per_batch, scene_steps=args.scenes * args.steps)
    log(f"[warm-start] wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
