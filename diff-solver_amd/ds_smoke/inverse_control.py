"""Inverse control: Adam on the initial velocity of the plume so that the density after `steps`
steps matches a target (the loss 0.5 ||rho_T - target||^2 of ds_loss_gradient).

  python -m ds_smoke.inverse_control --grid 128 --iterations 200 --steps 64 --k 8 --lr 1e-2 \\
         --warm-start previous --out result.json

Each iteration is one ds_loss_gradient call (steps forward steps, then the checkpointed reverse
sweep with a checkpoint every k steps) and one Adam update (beta = (0.9, 0.999), eps = 1e-8) of
the three face-velocity fields, done in PyTorch on the GPU. The target is the 4-blob Gaussian
target of the plume scene with its own seed. Prints one JSON line per `--report` iterations and a
final JSON line with the loss history, the wall time and the pressure-solve iteration counts.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def run(grid: int = 128, iterations: int = 200, steps: int = 64, k: int = 8, lr: float = 1e-2, warm_start: str = "none",
        weights: str | None = None, target_seed: int = 5, tolerance: float = 1e-6, report: int = 10, device: int = 0, log=print):
    import torch

    from . import abi, scenes
    from .solver import Context

    modes = {"none": abi.WARM_START_NONE, "previous": abi.WARM_START_PREVIOUS, "extrapolate": abi.WARM_START_EXTRAPOLATE, "projection": abi.WARM_START_PROJECTION,
             "cnn_alone": abi.WARM_START_CNN_ALONE, "cnn": abi.WARM_START_CNN}
    sc = scenes.plume(grid, target_seed=target_seed)
    ctx = Context(sc.config(tolerance=tolerance, warm_start=modes[warm_start]), abi.MODE_DIFFERENTIABLE, device)
    if warm_start in ("cnn", "cnn_alone"):
        from . import warm_start as ws

        ws.load_into(ctx, weights)
    s0 = ctx.state_from_host(scenes.initial_state(sc))
    control = ctx.control_from_host(scenes.control(sc))
    params = ctx.allocate_parameters(sc.params)
    target = ctx.to_device(sc.target)
    u = [v.clone() for v in s0["velocity"]]
    m = [torch.zeros_like(v) for v in u]
    v2 = [torch.zeros_like(v) for v in u]
    b1, b2, eps = 0.9, 0.999, 1e-8
    history = []
    solves = {"forward": [0, 0], "recompute": [0, 0], "adjoint": [0, 0]}
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for it in range(1, iterations + 1):
        state = dict(s0, velocity=tuple(u))
        loss, grad, info = ctx.loss_gradient(state, control, params, steps, k, target)
        for key, (n_solves, n_it) in (("forward", (info.forward_solves, info.forward_iterations)),
                                      ("recompute", (info.recompute_solves, info.recompute_iterations)),
                                      ("adjoint", (info.adjoint_solves, info.adjoint_iterations))):
            solves[key][0] += n_solves
            solves[key][1] += n_it
        history.append(loss)
        with torch.cuda.stream(ctx.stream):
            for a in range(3):
                g = grad["velocity"][a]
                m[a].mul_(b1).add_(g, alpha=1.0 - b1)
                v2[a].mul_(b2).addcmul_(g, g, value=1.0 - b2)
                m_hat = m[a] / (1.0 - b1 ** it)
                v_hat = v2[a] / (1.0 - b2 ** it)
                u[a].sub_(lr * m_hat / (v_hat.sqrt() + eps))
        if it % report == 0 or it == 1 or it == iterations:
            ctx.synchronize()
            log(json.dumps({"iteration": it, "loss": loss, "seconds": round(time.perf_counter() - t0, 2)}))
    ctx.synchronize()
    wall = time.perf_counter() - t0
    per_solve = {key: (n_it / n_s if n_s else 0.0) for key, (n_s, n_it) in solves.items()}
    ctx.close()
    return {"grid": grid, "iterations": iterations, "steps": steps, "checkpoint_interval": k, "lr": lr, "warm_start": warm_start,
            "initial_loss": history[0], "final_loss": history[-1], "loss_every_10": history[::10] + [history[-1]], "losses": history,
            "wall_seconds": wall, "seconds_per_iteration": wall / iterations, "cg_iterations_per_solve": per_solve}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--iterations", type=int, default=200)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--warm-start", choices=["none", "previous", "extrapolate", "projection", "cnn_alone", "cnn"], default="none")
    ap.add_argument("--weights", default=None, help="warm-start network (a file written by ds_smoke.train_warm_start)")
    ap.add_argument("--report", type=int, default=10)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    if args.warm_start in ("cnn", "cnn_alone") and not args.weights:
        import os

        args.weights = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "warm_start_plume.npz")
    res = run(args.grid, args.iterations, args.steps, args.k, args.lr, args.warm_start, args.weights, report=args.report,
              log=lambda s: print(s, flush=True))
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
