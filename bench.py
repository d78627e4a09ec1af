#!/usr/bin/env python
"""Benchmark of the MI355X smoke solver's hot path.

Workload (one bench step): the 128^3 smoke plume, 64 forward steps, then the VJP of
0.5 ||rho_64 - target||^2 with respect to the initial state by a reverse sweep with a checkpoint
every 8 steps (recompute in between). Each warm-start mode of the pressure solves runs in turn in
its own context: x0 = 0 ("plain"), the sequence guess (quadratic extrapolation of the forward
solves, previous adjoint solution), the warm-start network alone, and the network correcting the
sequence guess ("cnn", the headline when its weights exist). value = simulated fwd+VJP steps per
second of the headline mode over the timed bench steps, inputs resident in HBM. Also reported: PCG
iterations per solve (forward, recompute, adjoint) for every mode, the network-vs-plain
comparison, the HBM roofline of the dominant kernel (HIP-event timing of sampled launches inside
the timed region), the MFMA roofline of the network's inference, and the FP64 CPU oracle timed on
the host cores on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diff-solver_amd"))

METRIC = "fwd+VJP steps/s at 128^3 and CG iters/solve to 1e-6: CNN warm-start vs plain"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell of one launch of each PCG iteration kernel: unknowns read/write
# their FP32 vectors and the u8 code once; other cells only read the code (and the direction
# kernel writes p = 0 there).
# pcg_fused_kernel (single-reduction form, one launch per iteration): reads code, r, w, s, p and
# the solution increment, writes r, w, s, p and the increment on unknowns; other cells read the
# code and write zeros into the next r and w.
BYTES_UNKNOWN = {"pcg_direction_kernel": 4 + 4 + 1 + 4, "pcg_update_kernel": 4 + 4 + 1 + 8 + 4 + 8,
                 "pcg_fused_kernel": 1 + 4 * 4 + 4 + 4 * 4 + 4}
BYTES_OTHER = {"pcg_direction_kernel": 1 + 4, "pcg_update_kernel": 1, "pcg_fused_kernel": 1 + 4 + 4}
# the resident kernel performs the streamed kernel's iteration with its vectors held on chip: the
# algorithmic bytes of an iteration are the same (its timing unit is one iteration)
BYTES_UNKNOWN["pcg_resident_kernel"] = BYTES_UNKNOWN["pcg_fused_kernel"]
BYTES_OTHER["pcg_resident_kernel"] = BYTES_OTHER["pcg_fused_kernel"]
# pcg_s2_kernel (s-step form, two iterations per launch): every cell reads the code, the five
# basis vectors and the solution increment and writes the new basis and the increment
BYTES_UNKNOWN["pcg_s2_kernel"] = BYTES_OTHER["pcg_s2_kernel"] = 1 + 5 * 4 + 4 + 5 * 4 + 4
# pcg_mg_iteration (multigrid-preconditioned form; the timed unit is one iteration, thirteen
# kernels: the V-cycle, the direction and the update). Level 0 per unknown: pre-smoothing reads
# code and r and writes x (9), restriction reads code, r, x and writes an eighth of a coarse value
# (9.5), the first post sweep reads code, r, x, an eighth of the coarse result and writes y (13.5),
# the second reads code, r, y and writes z (13), the direction reads code, z, p_old and writes p
# (13), the update reads code, p, r and the solution pair and writes r and the pair (29); other
# cells read the code and write the zeros of x, y, z, p (22.5 with the restriction's share). Each
# cell of a coarser level but the last: diagonal, its reciprocal, three couplings and the
# right-hand side read by the down and the up kernel, an eighth of a value written down, an eighth
# read and the result written up (53); the coarsest level's FP64 inverse is read once (mg_bytes).
BYTES_UNKNOWN["pcg_mg_iteration"] = 9 + 9.5 + 13.5 + 13 + 13 + 29
BYTES_OTHER["pcg_mg_iteration"] = 5 + 1.5 + 5 + 5 + 5 + 1
MG_COARSE_BYTES_PER_CELL = 53
MG_COARSEST = 512
# advect_velocity_vjp_kernel, one launch per face axis: 75 B per face of the axis (the reference's
# decomposition: the face's output adjoint, the three velocity components of its trace and the
# three adjoint components it scatters into, each array streamed once)
FACE_BYTES = {"advect_velocity_vjp_kernel": 75}
# composite timing units (several kernels), reported but not a kernel's roofline
COMPOSITE = {"pcg_mg_iteration"}


def mg_bytes(res):
    """Bytes of the coarser levels and the coarsest inverse per multigrid iteration: 2 x 2 x 2
    aggregation (extents halved, rounded up) until at most MG_COARSEST cells."""
    dims, coarse = list(res), 0
    while dims[0] * dims[1] * dims[2] > MG_COARSEST:
        dims = [(d + 1) // 2 for d in dims]
        coarse += dims[0] * dims[1] * dims[2]
    last = dims[0] * dims[1] * dims[2]
    return MG_COARSE_BYTES_PER_CELL * (coarse - (last if coarse else 0)) + 4 * last * 2 + 8 * last * last


def host_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(grid, sim_steps, k, sample_steps, threads=0):
    """FP64 oracle (same discretisation and stopping rule) on a bounded sample of the workload, on
    `threads` host threads (0: all the process may use)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from ds_smoke import scenes

    threads = threads or host_threads()
    O.set_threads(threads)
    sc = scenes.plume(grid)
    orc = O.Oracle(sc, tolerance=1e-6)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    t0 = time.perf_counter()
    out = orc.loss_gradient(s0, ctl, sc.params, sample_steps, min(k, sample_steps), sc.target)
    dt = time.perf_counter() - t0
    info = out["info"]
    return {
        "value": sample_steps / dt, "unit": "fwd+VJP steps/s", "cores": threads, "kind": "port",
        "sample": f"FP64 oracle, {grid}^3 plume, {sample_steps} forward steps + VJP (checkpoint {min(k, sample_steps)}), "
                  f"PCG to 1e-6; {dt:.1f} s; forward/adjoint iterations per solve "
                  f"{info.forward_iterations / max(1, info.forward_solves):.0f}/"
                  f"{info.adjoint_iterations / max(1, info.adjoint_solves):.0f}",
    }


def calibrate_copy(mib):
    """Device copies of a known size, each launched 10 times: torch's copy (16-byte accesses per
    lane) and the library's stream_copy_kernel (4-byte accesses per lane, as the PCG kernels).
    FETCH_SIZE / WRITE_SIZE of these launches give the counter-to-bytes factor that the PCG
    kernels' counters are scaled by."""
    import torch

    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ctx = Context(scenes.plume(16).config(), abi.MODE_FORWARD)
    out = {"copy_mib": mib, "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4}
    for name, copy in (("torch_copy", lambda: dst.copy_(src)),
                       ("stream_copy_kernel", lambda: ctx.stream_copy(src, dst))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            copy()
        ctx.synchronize()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 10
        out[name] = {"ms_per_copy": dt * 1e3, "gbs": 2 * n * 4 / dt / 1e9}
    ctx.close()
    print(json.dumps(out))


# Multiply-adds per fine voxel of one inference of the warm-start network (3-level U-Net,
# 16/32/64 channels): level 1 (every voxel): 3x3x3 convolutions 2->16, 16->16, 32->16, 16->16, the
# 32->16 transposed convolution (one 2x2x2 tap per fine voxel) and the 16->1 head; level 2 (one
# voxel per 8): 16->32, 32->32, 64->32, 32->32 and the 64->32 transposed convolution; level 3
# (one per 64): 32->64, 64->64.
UNET_MACS_PER_VOXEL = (27 * (2 * 16 + 16 * 16 + 32 * 16 + 16 * 16) + 32 * 16 + 16
                       + (27 * (16 * 32 + 32 * 32 + 64 * 32 + 32 * 32) + 64 * 32) / 8
                       + 27 * (32 * 64 + 64 * 64) / 64)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def unet_timing(ctx, sc, levels, repeats=20):
    """MFMA roofline of the network's guess at the workload's grid: algorithmic FLOPs
    (2 x UNET_MACS_PER_VOXEL per voxel of the grid the network runs on, the workload's grid
    coarsened `levels` times) / the mean time of `repeats` whole guesses (restriction, network,
    prolongation and step length; HIP events on the context's stream, after one warm-up)."""
    import numpy as np
    import torch

    _, code = ctx.masks()
    rng = np.random.default_rng(0)
    b = ctx.to_device(np.where(code > 0, rng.standard_normal(code.shape), 0.0).astype(np.float32))
    ctx.warm_start_infer(b)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ctx.synchronize()
    start.record(ctx.stream)
    for _ in range(repeats):
        ctx.warm_start_guess(b)
    stop.record(ctx.stream)
    stop.synchronize()
    ms = start.elapsed_time(stop) / repeats
    net_grid = [int(n) >> levels for n in sc.resolution]
    flops = 2.0 * UNET_MACS_PER_VOXEL * int(np.prod(net_grid))
    tf = flops / (ms * 1e-3) / 1e12
    return {"bound": "mfma", "kernel": "warm-start network guess (13 layers)", "network_grid": net_grid, "achieved": round(tf, 2),
            "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4),
            "flops_per_inference": flops, "ms_per_inference": round(ms, 4), "dtype": "bf16 (fp32 accumulation)"}


def network_only(grid, levels, weights):
    """The warm-start network alone: its guess on the grid^3 plume's pressure system (random right-hand
    side on the unknowns), the network run on the grid coarsened `levels` times."""
    from ds_smoke import abi, scenes, warm_start
    from ds_smoke.solver import Context

    sc = scenes.plume(grid)
    ctx = Context(sc.config(tolerance=1e-6, warm_start=abi.WARM_START_CNN, cg_variant=abi.CG_MULTIGRID), abi.MODE_DIFFERENTIABLE)
    flat, scale, _ = warm_start.load(weights)
    ctx.warm_start_load(flat, scale, levels=levels)
    print(json.dumps({"workload": f"plume{grid}_network_guess", "levels": levels, **unet_timing(ctx, sc, levels)}), flush=True)


def traffic_from_profiles(kernel):
    """HBM bytes per launch from the committed rocprofv3 counter summary, when present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_plume128_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed bench steps (each a 64-step fwd+VJP trajectory)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--sim-steps", type=int, default=64)
    ap.add_argument("--checkpoint", type=int, default=8)
    ap.add_argument("--profile-every", type=int, default=8)
    ap.add_argument("--cpu-sample-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--warm-start", choices=["all", "none", "previous", "extrapolate", "projection", "cnn_alone", "cnn"], default="all",
                    help="initial guess of each pressure solve: zero, the previous solve's solution, the extrapolated "
                         "sequence, the projection onto the previous eight, the warm-start network alone, or the network "
                         "correcting the extrapolated sequence; all = none, extrapolate, cnn_alone and cnn in turn (the "
                         "network modes when its weights exist)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "diff-solver_amd", "data", "warm_start_plume.npz"),
                    help="warm-start network parameters (for --warm-start cnn)")
    ap.add_argument("--cg-variant", choices=["single", "two", "resident", "s2", "mg"], default="mg",
                    help="PCG recurrence: single-reduction (one fused kernel per iteration), resident (the same "
                         "recurrence as one cooperative launch per replacement interval where the grid allows), "
                         "two-reduction, s2 (s-step: two iterations per fused launch), or mg (preconditioned by an "
                         "aggregation multigrid V-cycle instead of Jacobi)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 10 device-to-device copies of MIB MiB (the reference for HBM counter scaling)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="host threads of the CPU baseline (0: all)")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only time the CPU baseline (the FP64 oracle on --cpu-sample-steps steps) and print it as one JSON line")
    ap.add_argument("--network-only", type=int, default=-1, metavar="LEVELS",
                    help="only time the warm-start network's guess on the --grid plume, the grid coarsened LEVELS "
                         "times (0, 1 or 2), and print its MFMA roofline as one JSON line")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.network_only >= 0:
        network_only(args.grid, args.network_only, args.weights)
        return
    if args.cpu_only:
        print(json.dumps({"cpu_baseline": cpu_baseline(args.grid, args.sim_steps, args.checkpoint, args.cpu_sample_steps,
                                                       args.cpu_threads)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the hot path is single-GPU here); got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    import torch
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    sc = scenes.plume(args.grid)
    variant = {"single": abi.CG_SINGLE_REDUCTION, "two": abi.CG_TWO_REDUCTION,
               "resident": abi.CG_SINGLE_REDUCTION_RESIDENT, "s2": abi.CG_S_STEP, "mg": abi.CG_MULTIGRID}[args.cg_variant]
    if args.warm_start == "all":
        modes = ["none", "extrapolate"] + (["cnn_alone", "cnn"] if os.path.exists(args.weights) else [])
    else:
        modes = [args.warm_start]

    def run_mode(mode):
        """Warm-up, then the timed bench steps of one warm-start mode in a fresh context."""
        warm = {"none": abi.WARM_START_NONE, "previous": abi.WARM_START_PREVIOUS, "extrapolate": abi.WARM_START_EXTRAPOLATE,
                "projection": abi.WARM_START_PROJECTION, "cnn_alone": abi.WARM_START_CNN_ALONE, "cnn": abi.WARM_START_CNN}[mode]
        ctx = Context(sc.config(tolerance=1e-6, warm_start=warm, cg_variant=variant), abi.MODE_DIFFERENTIABLE)
        unet = None
        if mode in ("cnn", "cnn_alone"):
            from ds_smoke import warm_start

            meta = warm_start.load_into(ctx, args.weights)
            unet = unet_timing(ctx, sc, warm_start.levels_for(meta, sc.resolution))
        s0 = ctx.state_from_host(scenes.initial_state(sc))
        ctl = ctx.control_from_host(scenes.control(sc))
        params = ctx.allocate_parameters(sc.params)
        target = ctx.to_device(sc.target)
        grad = ctx.allocate_state_adjoint()

        def one():
            with torch.cuda.stream(ctx.stream):
                for t in (grad["density"], grad["temperature"]) + tuple(grad["velocity"]):
                    t.zero_()
            return ctx.loss_gradient(s0, ctl, params, args.sim_steps, args.checkpoint, target, initial_adj=grad)

        for _ in range(args.warmup):
            one()
        torch.cuda.synchronize()
        ctx.profile_enable(args.profile_every)
        infos = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            loss, _, info = one()
            infos.append(info)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        prof = ctx.profile_read()
        ctx.profile_enable(0)
        info = infos[-1]
        per_solve = {
            "forward": info.forward_iterations / max(1, info.forward_solves),
            "recompute": info.recompute_iterations / max(1, info.recompute_solves),
            "adjoint": info.adjoint_iterations / max(1, info.adjoint_solves),
        }
        n, nf = ctx.domain.cell_count, ctx.domain.unknown_count
        kern = {}
        faces = sum(ctx.domain.face_count[a] for a in range(3)) / 3.0   # per launch (one axis), averaged
        for name, (launches, samples, avg_ms) in prof.items():
            if name in FACE_BYTES:
                b = FACE_BYTES[name] * faces
            else:
                b = BYTES_UNKNOWN[name] * nf + BYTES_OTHER[name] * (n - nf) + (mg_bytes(sc.resolution) if name == "pcg_mg_iteration" else 0)
            kern[name] = {"launches": launches, "samples": samples, "avg_us": avg_ms * 1e3, "algorithmic_bytes": b,
                          "gbs": b / (avg_ms * 1e-3) / 1e9 if samples else None}
        ctx.close()
        iters = (info.forward_iterations + info.recompute_iterations + info.adjoint_iterations) / args.sim_steps
        return {"value": args.steps * args.sim_steps / elapsed, "ms_per_step": elapsed / args.steps * 1e3, "loss": loss,
                "per_solve": per_solve, "kernels": kern, "not_converged": info.not_converged,
                "cells": n, "unknowns": nf, "cg_iterations_per_sim_step": iters, "unet": unet}

    results = {m: run_mode(m) for m in modes}
    # A mode whose solves did not all converge, or whose loss is not finite, is not a valid
    # measurement (a non-finite residual ends PCG at once, which would look fast).
    excluded = {m: ("not converged: %d solves" % results[m]["not_converged"]) if results[m]["not_converged"] else "non-finite loss"
                for m in modes if results[m]["not_converged"] or not math.isfinite(results[m]["loss"])}
    qualified = [m for m in modes if m not in excluded]
    if not qualified:
        print(json.dumps({"metric": METRIC, "error": "no warm-start mode converged", "excluded": excluded}))
        sys.exit(1)
    # the metric is the network's warm start against x0 = 0: the network's mode is the headline
    # when it ran and qualified, else the fastest qualified mode (named in config.warm_start)
    best = "cnn" if "cnn" in qualified else max(qualified, key=lambda m: results[m]["value"])
    r = results[best]
    kern = r["kernels"]
    kernels_only = [k for k in kern if k not in COMPOSITE]
    dom = max(kernels_only, key=lambda k: kern[k]["launches"] * (kern[k]["avg_us"] or 0.0)) if kernels_only else None
    roofline = None
    if dom:
        d = kern[dom]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(dom),
                    "algorithmic_bytes_per_launch": d["algorithmic_bytes"], "avg_launch_us": round(d["avg_us"], 2)}
    # whole fwd+VJP step at the reference's operator decomposition: 311 B/cell forward, 311 B/cell
    # recompute, 605 B/cell reverse, 47 B per unknown per CG iteration (the two-kernel textbook PCG)
    step_s = r["ms_per_step"] * 1e-3 / args.sim_steps
    step_bytes = r["cells"] * (311 + 311 + 605) + 47 * r["unknowns"] * r["cg_iterations_per_sim_step"]
    step_gbs = step_bytes / step_s / 1e9
    step_roofline = {"bound": "hbm", "bytes_per_step": round(step_bytes), "cg_iterations_per_step": round(r["cg_iterations_per_sim_step"], 1),
                     "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)}
    label = {"none": "plain", "previous": "previous", "extrapolate": "extrapolate", "projection": "projection", "cnn_alone": "cnn_alone",
             "cnn": "cnn"}
    # the metric's own comparison: the network's modes against x0 = 0 at the same 1e-6 residual
    cnn_vs_plain = None
    if "none" in results:
        base = results["none"]
        cnn_vs_plain = {}
        for m in ("none", "cnn_alone", "extrapolate", "cnn"):
            if m not in results:
                continue
            rm = results[m]
            cnn_vs_plain[label[m]] = {
                "steps_per_s": round(rm["value"], 4), "speedup_vs_plain": round(rm["value"] / base["value"], 4),
                "forward_cg_per_solve": round(rm["per_solve"]["forward"], 1), "adjoint_cg_per_solve": round(rm["per_solve"]["adjoint"], 1),
                "cg_iterations_vs_plain": round(rm["cg_iterations_per_sim_step"] / base["cg_iterations_per_sim_step"], 4),
                "excluded": excluded.get(m)}
    result = {
        "metric": METRIC,
        "value": round(r["value"], 4),
        "unit": "fwd+VJP steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plume{args.grid}_fwd{args.sim_steps}_vjp_checkpoint{args.checkpoint}",
                   "grid": args.grid, "sim_steps_per_bench_step": args.sim_steps,
                   "checkpoint_interval": args.checkpoint, "cg_tolerance": 1e-6, "warm_start": label[best],
                   "cg_variant": args.cg_variant, "pressure_tape": True},
        "cg_iterations_per_solve": {label[m]: {k: round(v, 1) for k, v in results[m]["per_solve"].items()} for m in modes},
        "steps_per_s_by_warm_start": {label[m]: round(results[m]["value"], 4) for m in modes},
        "loss": r["loss"],
        "cnn_vs_plain": cnn_vs_plain,
        "roofline": roofline,
        "solver_iteration": next(({"unit": k, "kernels_per_iteration": 13, "avg_us": round(kern[k]["avg_us"], 2),
                                   "algorithmic_bytes": kern[k]["algorithmic_bytes"], "achieved": round(kern[k]["gbs"], 1),
                                   "peak": HBM_PEAK_GBS, "unit_bw": "GB/s", "frac": round(kern[k]["gbs"] / HBM_PEAK_GBS, 4)}
                                  for k in COMPOSITE if k in kern and kern[k]["gbs"]), None),
        "unet_roofline": next((results[m]["unet"] for m in ("cnn", "cnn_alone") if m in results and results[m]["unet"]), None),
        "step_roofline": step_roofline,
        "kernels": {k: {kk: (round(vv, 2) if isinstance(vv, float) else vv) for kk, vv in v.items()} for k, v in kern.items()},
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(args.grid, args.sim_steps, args.checkpoint, args.cpu_sample_steps, args.cpu_threads)
        if "none" in results:
            # the GPU from x0 = 0, as the oracle starts every solve; the oracle iterates Jacobi-PCG,
            # the GPU's default PCG form is preconditioned by a multigrid V-cycle, so the ratio holds
            # the algorithm's saving too: the sample names both iteration counts
            rn = results["none"]
            cb["sample"] += (f"; against the GPU from x0 = 0 ({rn['value']:.2f} steps/s, PCG form {args.cg_variant}, "
                             f"{rn['per_solve']['forward']:.1f}/{rn['per_solve']['adjoint']:.1f} iterations per solve): "
                             f"{rn['value'] / cb['value']:.0f}x")
        result["cpu_baseline"] = cb
    print(json.dumps(result))


if __name__ == "__main__":
    main()
