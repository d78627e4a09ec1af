"""Train the warm-start network on solver-generated pairs and measure what it saves.

  python -m ds_smoke.train_warm_start --grid 64 --eval-grid 128 --runs 16 --steps 64 --minutes 7 --out weights.npz

Generates (b, p) pairs from forward steps of jittered plumes (pressure solved to 1e-8) at the
training grid, trains the U-Net with Adam, then compares PCG iterations to 1e-6 from x0 = 0 and
from the network's guess on held-out plumes (other noise seeds and source positions) at the
evaluation grid. The network is fully convolutional and the library scales its guess to the
A-norm-optimal multiple, so a network trained on a coarser grid applies to a finer one (the
pressure's magnitude grows with the resolution; its shape does not change much). Prints one JSON
line per phase.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def evaluate(weights, output_scale, grid, runs, steps, seed, device=0, log=print, solids=0):
    import torch

    from . import abi, scenes
    from .solver import Context
    from .warm_start import generate_pairs

    bs, ps, unknown, _ = generate_pairs(grid, runs, steps, seed=seed, tolerance=1e-8, device=device, log=log, solids=solids)
    sc = scenes.plume(grid, with_target=False, solids=solids)
    plain = Context(sc.config(tolerance=1e-6), abi.MODE_FORWARD, device)
    cnn = Context(sc.config(tolerance=1e-6, warm_start=abi.WARM_START_CNN), abi.MODE_FORWARD, device)
    cnn.warm_start_load(weights, output_scale)
    it_plain, it_cnn, rel_err, rejected, t_plain, t_cnn = [], [], [], 0, 0.0, 0.0
    it_prev, it_prev_cnn = [], []
    for k, (b, p) in enumerate(zip(bs, ps)):
        if k % steps:   # the previous step's solution as the guess, alone and corrected by the network
            p_prev = ps[k - 1].contiguous()
            _, iv, _ = plain.pcg_solve(b, x0=p_prev)
            r = (b - plain.apply_matrix(p_prev)).contiguous()
            plain.synchronize()
            x0 = (p_prev + cnn.warm_start_infer(r)).contiguous()
            cnn.synchronize()
            _, ivc, _ = plain.pcg_solve(b, x0=x0)
            it_prev.append(iv.iterations)
            it_prev_cnn.append(ivc.iterations)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, ip, _ = plain.pcg_solve(b)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, ic, _ = cnn.pcg_solve(b)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        g = cnn.warm_start_infer(b)
        cnn.synchronize()
        rel_err.append(float(torch.linalg.vector_norm((g - p).double()) / torch.linalg.vector_norm(p.double()).clamp(min=1e-30)))
        it_plain.append(ip.iterations)
        it_cnn.append(ic.iterations)
        rejected += int(ic.guess_rejected)
        t_plain += t1 - t0
        t_cnn += t2 - t1
    return {"pairs": len(it_plain), "iterations_plain": float(np.mean(it_plain)), "iterations_cnn": float(np.mean(it_cnn)),
            "iterations_previous": float(np.mean(it_prev)), "iterations_previous_plus_cnn": float(np.mean(it_prev_cnn)),
            "guess_relative_error": float(np.median(rel_err)), "guesses_rejected": rejected,
            "ms_per_solve_plain": 1e3 * t_plain / len(it_plain), "ms_per_solve_cnn": 1e3 * t_cnn / len(it_cnn)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=64, help="training grid")
    ap.add_argument("--eval-grid", type=int, default=128)
    ap.add_argument("--runs", type=int, default=16)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--minutes", type=float, default=8.0)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-runs", type=int, default=2)
    ap.add_argument("--fp32", action="store_true", help="train without BF16 autocast")
    ap.add_argument("--init", default=None, help="start from these parameters (a file written by this script)")
    ap.add_argument("--solids", type=int, default=0, help="seeded solid spheres in every run (the same placement)")
    ap.add_argument("--crop", type=int, default=0, help="train on random crop^3 windows of the pairs (0: whole grid)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    from . import warm_start

    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.time()
    bs, ps, unknown, diag = warm_start.generate_pairs(args.grid, args.runs, args.steps, seed=100, log=log, solids=args.solids)
    print(json.dumps({"phase": "data", "pairs": len(bs), "seconds": round(time.time() - t0, 1)}), flush=True)
    init = warm_start.load(args.init)[0] if args.init else None
    model, output_scale = warm_start.train(bs, ps, unknown, epochs=1000, batch=args.batch, lr=args.lr,
                                           time_budget=60.0 * args.minutes, log=log, diag=diag, amp=not args.fp32, init=init,
                                           crop=args.crop)
    flat = warm_start.flatten_parameters(model)
    meta = {"grid": args.grid, "eval_grid": args.eval_grid, "solids": args.solids, "runs": args.runs, "steps": args.steps, "pairs": len(bs),
            "minutes": args.minutes, "crop": args.crop,
            "batch": args.batch, "lr": args.lr, "label_tolerance": 1e-8}
    warm_start.save(args.out, flat, output_scale, meta)
    print(json.dumps({"phase": "train", "output_scale": output_scale, "seconds": round(time.time() - t0, 1)}), flush=True)
    del bs, ps
    ev = evaluate(flat, output_scale, args.eval_grid, args.eval_runs, min(args.steps, 32), seed=900, log=log, solids=args.solids)
    print(json.dumps({"phase": "eval", **ev}), flush=True)


if __name__ == "__main__":
    main()
