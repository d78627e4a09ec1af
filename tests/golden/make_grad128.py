"""Writes tests/golden/grad128.npz: the benchmark gradient at full size from the FP64 oracle.

Workload: the 128^3 plume (h = 1/128, dt = h/2, no-slip box, source sphere of radius 12 cells at
(64, 16, 64) cells, velocity noise from splitmix64 seed 1), 64 forward steps, then the gradient of
0.5 ||rho_64 - target||^2 (target: 4 Gaussian blobs, seed 2) with respect to the initial state by
the checkpointed reverse sweep (checkpoint every 8 steps), every pressure solve to a relative
residual of 1e-9 (solver.ixx:115-150 for the sweep's semantics).

Stored (the full gradient is 24 MiB; these pin it at a few hundred KB):
* loss, the four parameter adjoints and the CG iterations of the run;
* the gradient with respect to the initial velocity at 32 DOFs drawn uniformly over the 6,340,608
  face DOFs (numpy default_rng(20261202).choice, sorted);
* <g, v_j> for two directions v_j = default_rng(seed_j).standard_normal per face array (FP32),
  seeds 11 and 12, and the L2 norm of each axis' gradient;
* two x-y planes of each axis' gradient (FP32): z = 40, off the scene's mirror planes, and z = 64,
  the mirror plane x = z = 64 of the source crosses, where face velocities across it are near 0 and
  traces end on the sampling planes (where the trilinear derivative jumps).

Usage: python tests/golden/make_grad128.py [--threads T]   (about 20 minutes on 8 cores)
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "diff-solver_amd")]

import oracle as O  # noqa: E402
from ds_smoke import scenes  # noqa: E402

DOF_SEED = 20261202
DIRECTION_SEEDS = (11, 12)
PLANE_Z = 40
MIRROR_Z = 64
TOLERANCE = 1e-9


def directions(shapes, seed):
    rng = np.random.default_rng(seed)
    return tuple(rng.standard_normal(s).astype(np.float32) for s in shapes)


def dof_sample(sizes):
    return np.sort(np.random.default_rng(DOF_SEED).choice(sum(sizes), 32, replace=False))


def dof_values(g, sizes, dofs):
    flat = np.concatenate([np.asarray(v).reshape(-1) for v in g])
    assert flat.size == sum(sizes)
    return flat[dofs]


def main(threads):
    sc = scenes.plume(128)
    O.set_threads(threads)
    o = O.Oracle(sc, tolerance=TOLERANCE)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    t0 = time.time()
    out = o.loss_gradient(s0, ctl, sc.params, 64, 8, sc.target)
    dt = time.time() - t0
    info = out["info"]
    g = out["grad"]["velocity"]
    shapes = [v.shape for v in s0["velocity"]]
    sizes = [int(np.prod(s)) for s in shapes]
    dofs = dof_sample(sizes)
    dots = [float(sum(np.dot(gv.reshape(-1), vv.reshape(-1).astype(np.float64)) for gv, vv in zip(g, directions(shapes, sd))))
            for sd in DIRECTION_SEEDS]
    np.savez_compressed(
        os.path.join(HERE, "grad128.npz"),
        loss=np.float64(out["loss"]), params=out["params"].astype(np.float64),
        dofs=dofs.astype(np.int64), dof_gradient=dof_values(g, sizes, dofs).astype(np.float64),
        direction_seeds=np.array(DIRECTION_SEEDS, dtype=np.int64), direction_dots=np.array(dots),
        axis_norms=np.array([np.linalg.norm(v) for v in g]),
        plane_z=np.int64(PLANE_Z),
        plane_x=g[0][PLANE_Z].astype(np.float32), plane_y=g[1][PLANE_Z].astype(np.float32), plane_z_faces=g[2][PLANE_Z].astype(np.float32),
        mirror_z=np.int64(MIRROR_Z),
        mirror_x=g[0][MIRROR_Z].astype(np.float32), mirror_y=g[1][MIRROR_Z].astype(np.float32),
        mirror_z_faces=g[2][MIRROR_Z].astype(np.float32),
        tolerance=np.float64(TOLERANCE),
        cg_iterations=np.array([info.forward_iterations, info.recompute_iterations, info.adjoint_iterations], dtype=np.int64),
        seconds=np.float64(dt), threads=np.int64(threads))
    print(f"loss {out['loss']:.10e}; {dt:.0f} s on {threads} threads; CG forward/recompute/adjoint "
          f"{info.forward_iterations}/{info.recompute_iterations}/{info.adjoint_iterations}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    main(ap.parse_args().threads)
