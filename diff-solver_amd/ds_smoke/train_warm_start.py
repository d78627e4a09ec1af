"""Train the warm-start network on the GPU with the library's own kernels, on pairs the solver
itself generates, and measure what it saves.

  python -m ds_smoke.train_warm_start --grid 64 --runs 50 --steps 100 --minutes 8 --out weights.npz

Data (the 64^3 plume of the forward benchmark with a jittered source, seeded noise and a seeded
4-blob target per run): every forward step's pressure p_t and, from the reverse sweep of the
density loss, every adjoint solution l_t, each solved to `--tolerance`. Four samples per step, each
an error e whose residual r = A e the trainer forms (exact pairs by construction):
  forward, plain:       e = p_t                      (the guess from b alone)
  forward, correction:  e = p_t - x_ext              (x_ext: the quadratic extrapolation of p)
  adjoint, plain:       e = l_t
  adjoint, correction:  e = l_t - l_{t+1}            (the previous adjoint solution of the sweep)
Training: Adam (ds_unet_trainer, BF16 MFMA convolutions, FP32 master weights) on the L2 error of
the guess the solver forms, ||gamma g - e||^2 / ||e||^2 (or the energy loss), over shuffled
batches, with a cosine-decayed learning rate.
Evaluation: held-out plumes at `--eval-grid`, CG iterations to 1e-6 from x0 = 0, from the
network's guess, from the sequence guess and from the network's correction of it.
Prints one JSON line per phase; torch only holds device memory.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

KINDS = ("forward", "forward_correction", "adjoint", "adjoint_correction")


def _extrapolate(hist):
    """The sequence guess of DS_WARM_START_EXTRAPOLATE from the last solutions (newest first)."""
    if len(hist) >= 3:
        return 3.0 * (hist[0] - hist[1]) + hist[2]
    if len(hist) == 2:
        return 2.0 * hist[0] - hist[1]
    return hist[0]


def jittered_plume(n, run, seed, jitter_cells, solids=0, solid_seed=4):
    from . import scenes

    rng = np.random.default_rng(seed * 1000 + run)
    off = rng.uniform(-jitter_cells, jitter_cells, 3) / n
    centre = (0.5 + off[0], 0.125 + abs(off[1]) * 0.5, 0.5 + off[2])
    return scenes.plume(n, noise_seed=seed + 1 + run, target_seed=seed + 5000 + run, source_centre=centre, solids=solids,
                        solid_seed=solid_seed)


def generate_pairs(n, runs, steps, seed=100, tolerance=1e-7, jitter_cells=8.0, device=0, log=print, solids=0, adjoint=True,
                   max_pairs=None):
    """Errors (P, nz, ny, nx) float32 on the GPU (0 off the unknowns), the kind of each sample (P,)
    and the unknown code."""
    import torch

    from . import abi, scenes
    from .solver import Context

    R, K = [], []
    code = None
    t0 = time.time()
    for run in range(runs):
        sc = jittered_plume(n, run, seed, jitter_cells, solids)
        ctx = Context(sc.config(tolerance=tolerance, warm_start=abi.WARM_START_EXTRAPOLATE), abi.MODE_DIFFERENTIABLE, device)
        if code is None:
            code = ctx.masks()[1]
        params = ctx.allocate_parameters(sc.params)
        ctl = ctx.control_from_host(scenes.control(sc))
        states = [ctx.state_from_host(scenes.initial_state(sc))]
        caches = []
        hist = []

        def add(e, kind):
            R.append(e)
            K.append(kind)

        for t in range(steps):
            nxt, cache, info = ctx.forward_step(states[-1], ctl, params)
            p = ctx.last_solution(0)
            with torch.cuda.stream(ctx.stream):
                add(p, 0)
                if hist:
                    add((p - _extrapolate(hist)).contiguous(), 1)
                hist = [p] + hist[:2]
            states.append(nxt)
            caches.append(cache)
        if adjoint:
            adj = ctx.allocate_state_adjoint()
            with torch.cuda.stream(ctx.stream):
                adj["density"].copy_(states[-1]["density"] - ctx.to_device(sc.target))
            prev_l = None
            for t in range(steps - 1, -1, -1):
                adj, _ = ctx.vjp_step(states[t], ctl, params, states[t + 1], caches[t], adj)
                lam = ctx.last_solution(1)
                with torch.cuda.stream(ctx.stream):
                    add(lam, 2)
                    if prev_l is not None:
                        add((lam - prev_l).contiguous(), 3)
                prev_l = lam
        ctx.synchronize()
        del states, caches, hist
        ctx.close()
        log(f"run {run + 1}/{runs}: {len(R)} pairs, {time.time() - t0:.0f} s")
        if max_pairs and len(R) >= max_pairs:
            break
    torch.cuda.synchronize()
    Rt = torch.stack(R)
    del R
    ok = torch.isfinite(Rt).flatten(1).all(1) & (Rt.flatten(1).abs().amax(1) > 0)
    return Rt[ok].contiguous(), np.asarray(K)[ok.cpu().numpy()], code


def train(ctx, R, params, batch=8, lr=1e-3, minutes=8.0, epochs=1000, seed=0, loss="l2", log=print):
    """Adam over shuffled batches until the time budget or the epochs run out (cosine-decayed rate
    over the budget). Returns the parameters (host FP32) and the loss history."""
    import torch

    from .unet import Trainer

    tr = Trainer(ctx, batch, params)
    P = R.shape[0]
    gen = torch.Generator(device="cpu").manual_seed(seed)
    with torch.cuda.stream(ctx.stream):
        rb = torch.empty((batch,) + tuple(R.shape[1:]), dtype=torch.float32, device=R.device)
    budget = 60.0 * minutes
    t0 = time.time()
    history = []
    step = 0
    losses = []
    done = False
    for ep in range(epochs):
        perm = torch.randperm(P, generator=gen)
        for i in range(0, P - batch + 1, batch):
            frac = min(1.0, (time.time() - t0) / budget)
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * frac))
            idx = perm[i:i + batch].to(R.device)
            with torch.cuda.stream(ctx.stream):
                torch.index_select(R, 0, idx, out=rb)
            lv = tr.step(rb, rate, loss=loss)
            with torch.cuda.stream(ctx.stream):
                losses.append(lv.clone())
            step += 1
            if step % 50 == 0:
                ctx.synchronize()
                mean = float(torch.cat(losses).mean())
                losses = []
                history.append((step, time.time() - t0, mean))
                log(f"step {step} (epoch {ep + 1}): loss {mean:.4f}, lr {rate:.2e}, {time.time() - t0:.0f} s")
            if time.time() - t0 > budget:
                done = True
                break
        if done:
            break
    ctx.synchronize()
    out = tr.read("params").cpu().numpy()
    tr.close()
    return out, history, step


def evaluate(weights, grid, runs, steps, seed, device=0, log=print, solids=0, levels=0, cg_variant=None):
    """Held-out plumes at `grid`: for each forward step's problem A x = b (b = A p, p the step's
    pressure to 1e-8), CG iterations to 1e-6 from x0 = 0, the network's guess, the extrapolation of
    the previous solutions and the network's correction of it. The iterations are counted on
    Jacobi-preconditioned CG (the s-step form) unless cg_variant names another: the network was
    trained to save those; the multigrid preconditioner leaves it little to save."""
    import torch

    from . import abi, scenes
    from .solver import Context

    variant = abi.CG_S_STEP if cg_variant is None else cg_variant
    it = {"plain": [], "network": [], "extrapolate": [], "extrapolate_network": []}
    t_net = []
    for run in range(runs):
        sc = jittered_plume(grid, run, seed, 8.0 * grid / 64.0, solids)
        ctx = Context(sc.config(tolerance=1e-8, warm_start=abi.WARM_START_EXTRAPOLATE), abi.MODE_FORWARD, device)
        net = Context(sc.config(tolerance=1e-6, warm_start=abi.WARM_START_CNN_ALONE), abi.MODE_FORWARD, device)
        net.warm_start_load(weights, 1.0, levels=levels)
        plain = Context(sc.config(tolerance=1e-6, cg_variant=variant), abi.MODE_FORWARD, device)
        params = ctx.allocate_parameters(sc.params)
        ctl = ctx.control_from_host(scenes.control(sc))
        state = ctx.state_from_host(scenes.initial_state(sc))
        hist = []
        for t in range(steps):
            state, _, _ = ctx.forward_step(state, ctl, params)
            p = ctx.last_solution(0)
            b = ctx.apply_matrix(p)
            ctx.synchronize()
            _, i0, _ = plain.pcg_solve(b)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            g = net.warm_start_guess(b)
            t_net.append(time.perf_counter() - t1)
            _, i1, _ = plain.pcg_solve(b, x0=g)
            it["plain"].append(i0.iterations)
            it["network"].append(i1.iterations)
            if hist:
                with torch.cuda.stream(ctx.stream):
                    xe = _extrapolate(hist).contiguous()
                ctx.synchronize()
                _, i2, _ = plain.pcg_solve(b, x0=xe)
                gc = net.warm_start_guess(b, base=xe)
                _, i3, _ = plain.pcg_solve(b, x0=gc)
                it["extrapolate"].append(i2.iterations)
                it["extrapolate_network"].append(i3.iterations)
            hist = [p] + hist[:2]
        for c in (ctx, net, plain):
            c.close()
        log(f"eval run {run + 1}: " + ", ".join(f"{k} {np.mean(v):.0f}" for k, v in it.items() if v))
    out = {f"iterations_{k}": float(np.mean(v)) for k, v in it.items() if v}
    out["solves"] = len(it["plain"])
    out["ms_network_guess"] = 1e3 * float(np.median(t_net))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=64, help="training grid")
    ap.add_argument("--runs", type=int, default=50)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--tolerance", type=float, default=1e-7, help="relative residual of the solves that make the pairs")
    ap.add_argument("--minutes", type=float, default=8.0)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--init", default=None, help="start from these parameters (a file written by this script)")
    ap.add_argument("--solids", type=int, default=0, help="seeded solid spheres in every run (the same placement)")
    ap.add_argument("--no-adjoint", action="store_true", help="forward pairs only")
    ap.add_argument("--loss", choices=["l2", "energy"], default="l2")
    ap.add_argument("--eval-grid", type=int, default=128)
    ap.add_argument("--eval-runs", type=int, default=1)
    ap.add_argument("--eval-steps", type=int, default=24)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    from . import abi, scenes, warm_start
    from .solver import Context
    from .unet import init_parameters

    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.time()
    R, K, code = generate_pairs(args.grid, args.runs, args.steps, seed=args.seed, tolerance=args.tolerance, log=log,
                                   solids=args.solids, adjoint=not args.no_adjoint)
    counts = {KINDS[k]: int((K == k).sum()) for k in range(4)}
    print(json.dumps({"phase": "data", "pairs": int(R.shape[0]), "by_kind": counts, "seconds": round(time.time() - t0, 1)}), flush=True)
    sc = scenes.plume(args.grid, with_target=False, solids=args.solids)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_FORWARD)
    init = warm_start.load(args.init)[0] if args.init else init_parameters(args.seed)
    t1 = time.time()
    flat, history, steps = train(ctx, R, init, batch=args.batch, lr=args.lr, minutes=args.minutes, seed=args.seed, loss=args.loss,
                                 log=log)
    meta = {"grid": args.grid, "solids": args.solids, "runs": args.runs, "steps": args.steps, "pairs": int(R.shape[0]),
            "pairs_by_kind": counts, "label_tolerance": args.tolerance, "minutes": args.minutes, "batch": args.batch, "lr": args.lr,
            "optimizer_steps": steps, "loss": args.loss, "loss_history": history[-20:], "trainer": "ds_unet_trainer (HIP)"}
    warm_start.save(args.out, flat, 1.0, meta)
    print(json.dumps({"phase": "train", "optimizer_steps": steps, "samples_per_s": steps * args.batch / (time.time() - t1),
                      "final_loss": history[-1][2] if history else None, "seconds": round(time.time() - t1, 1)}), flush=True)
    ctx.close()
    del R
    if args.eval_runs > 0:
        ev = evaluate(flat, args.eval_grid, args.eval_runs, args.eval_steps, seed=900, log=log, solids=args.solids,
                      levels=warm_start.levels_for({"grid": args.grid}, (args.eval_grid,) * 3))
        print(json.dumps({"phase": "eval", "grid": args.eval_grid, "levels": warm_start.levels_for({"grid": args.grid}, (args.eval_grid,) * 3),
                          **ev}), flush=True)


if __name__ == "__main__":
    main()
