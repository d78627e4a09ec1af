"""Warm-start network of the pressure solves: definition, data generation and training.

The network is a 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions with ReLU, 2x2x2 max
pooling, 2x2x2 stride-2 transposed convolutions, a 1x1x1 head; 339,681 parameters). Input: the
right-hand side b divided by s = rms(b) over the unknowns, and the unknown indicator. Output:
the guess x0 = head * s * output_scale (0 off the unknowns).

Inference inside the solver runs in the HIP library (csrc/unet.hip, BF16 MFMA implicit-GEMM
convolutions); this module is the training harness. It trains in PyTorch on (b, p) pairs that the
solver itself produces: every forward step of jittered plume runs gives the converged pressure p
(to 1e-8) and b = A p. The parameters are saved in PyTorch state_dict order, flattened to FP32,
which is the order ds_warm_start_load expects.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import abi, scenes

LAYERS = ("enc1a", "enc1b", "enc2a", "enc2b", "bota", "botb", "up2", "dec2a", "dec2b", "up1", "dec1a", "dec1b", "head")


def build_unet():
    import torch.nn as nn

    class UNet(nn.Module):
        def __init__(self):
            super().__init__()
            conv = lambda ci, co: nn.Conv3d(ci, co, 3, padding=1)
            self.enc1a, self.enc1b = conv(2, 16), conv(16, 16)
            self.enc2a, self.enc2b = conv(16, 32), conv(32, 32)
            self.bota, self.botb = conv(32, 64), conv(64, 64)
            self.up2 = nn.ConvTranspose3d(64, 32, 2, stride=2)
            self.dec2a, self.dec2b = conv(64, 32), conv(32, 32)
            self.up1 = nn.ConvTranspose3d(32, 16, 2, stride=2)
            self.dec1a, self.dec1b = conv(32, 16), conv(16, 16)
            self.head = nn.Conv3d(16, 1, 1)

        def forward(self, x):
            import torch
            import torch.nn.functional as F

            r = F.relu
            s1 = r(self.enc1b(r(self.enc1a(x))))
            s2 = r(self.enc2b(r(self.enc2a(F.max_pool3d(s1, 2)))))
            c = r(self.botb(r(self.bota(F.max_pool3d(s2, 2)))))
            d2 = r(self.dec2b(r(self.dec2a(torch.cat([self.up2(c), s2], 1)))))
            d1 = r(self.dec1b(r(self.dec1a(torch.cat([self.up1(d2), s1], 1)))))
            return self.head(d1)

    return UNet()


def flatten_parameters(model) -> np.ndarray:
    return np.concatenate([t.detach().float().cpu().numpy().ravel() for t in model.state_dict().values()]).astype(np.float32)


def load_parameters(model, flat: np.ndarray):
    import torch

    sd = model.state_dict()
    off = 0
    for k, t in sd.items():
        n = t.numel()
        sd[k] = torch.from_numpy(np.asarray(flat[off:off + n], dtype=np.float32).reshape(t.shape))
        off += n
    assert off == flat.size, "parameter count mismatch"
    model.load_state_dict(sd)


def network_input(b, unknown):
    """(b / s, unknown) as a (1, 2, nz, ny, nx) tensor, and s = rms of b over the unknowns."""
    import torch

    nu = unknown.sum().clamp(min=1)
    s = torch.sqrt((b.double() ** 2).sum() / nu).clamp(min=1e-30)
    x = torch.stack([torch.where(unknown, b / s.float(), torch.zeros_like(b)), unknown.float()], 0)[None]
    return x, s


def save(path: str, flat: np.ndarray, output_scale: float, meta: dict):
    np.savez(path, params=flat, output_scale=np.float64(output_scale), meta=json.dumps(meta))


def load(path: str):
    d = np.load(path, allow_pickle=False)
    return d["params"], float(d["output_scale"]), json.loads(str(d["meta"]))


def generate_pairs(n: int, runs: int, steps: int, seed: int = 100, jitter_cells: float = 8.0, tolerance: float = 1e-8,
                   device: int = 0, log=print, solids: int = 0, solid_seed: int = 4):
    """(b, p) pairs from forward steps of jittered plumes at n^3, kept on the GPU. With solids, every
    run has the same `solids` seeded spheres (seed solid_seed), so the unknowns are shared."""
    import torch

    from .solver import Context

    rng = np.random.default_rng(seed)
    bs, ps = [], []
    unknown = None
    for r in range(runs):
        off = rng.uniform(-jitter_cells, jitter_cells, 3) / n
        centre = (0.5 + off[0], 0.125 + abs(off[1]) * 0.5, 0.5 + off[2])
        sc = scenes.plume(n, noise_seed=seed + 1 + r, source_centre=centre, with_target=False, solids=solids, solid_seed=solid_seed)
        ctx = Context(sc.config(tolerance=tolerance, max_iterations=6000), abi.MODE_FORWARD, device)
        state = ctx.state_from_host(scenes.initial_state(sc))
        ctl = ctx.control_from_host(scenes.control(sc))
        params = ctx.allocate_parameters(sc.params)
        cache = ctx.allocate_step_cache(pressure=True)
        nxt = ctx.allocate_state()
        if unknown is None:
            code = ctx.masks()[1]
            unknown = torch.from_numpy(code > 0).to(ctx.device)
        dropped = 0
        for k in range(steps):
            _, _, info = ctx.forward_step(state, ctl, params, next_state=nxt, cache=cache, check=False)
            ctx.synchronize()
            p = (cache["pressure"][0].double() + cache["pressure"][1].double()).float()
            bad = int((~torch.isfinite(p)).sum())
            if bad or not info.converged:
                # a pair whose label is not a converged finite solution would teach the network noise
                log(f"run {r + 1} step {k}: pair dropped ({bad} non-finite pressure values, "
                    f"{bad - int((~torch.isfinite(p[unknown])).sum())} of them off the unknowns; converged {bool(info.converged)}, "
                    f"{info.iterations} iterations)")
                dropped += 1
            else:
                p = torch.where(unknown, p, torch.zeros_like(p))
                b = ctx.apply_matrix(p)
                ctx.synchronize()
                bs.append(b.clone())
                ps.append(p.clone())
            state, nxt = nxt, state
        ctx.close()
        log(f"run {r + 1}/{runs}: {len(bs)} pairs" + (f" ({dropped} dropped)" if dropped else ""))
    return torch.stack(bs), torch.stack(ps), unknown, torch.from_numpy(code.astype(np.float32))


def apply_matrix(x, diag, per=(False, False, False)):
    """A x for the pressure matrix on a (..., nz, ny, nx) tensor that is 0 off the unknowns
    (diag = the per-cell diagonal count, 0 off the unknowns)."""
    import torch

    y = diag * x
    for dim, axis in ((-1, 0), (-2, 1), (-3, 2)):
        if per[axis]:
            y = y - torch.roll(x, 1, dim) - torch.roll(x, -1, dim)
        else:
            n = x.shape[dim]
            y = y - torch.nn.functional.pad(x.narrow(dim, 0, n - 1), _pad(dim, 1, 0))
            y = y - torch.nn.functional.pad(x.narrow(dim, 1, n - 1), _pad(dim, 0, 1))
    return torch.where(diag > 0, y, torch.zeros_like(y))


def _pad(dim, before, after):
    # F.pad takes (last-dim-before, last-dim-after, second-to-last-before, ...)
    k = -dim - 1
    return (0, 0) * k + (before, after)


def energy_loss(g, b, p, diag):
    """1 - cos_A(g, p)^2 = 1 - <g, b>^2 / (<g, A g> <b, p>): the share of the solution's energy
    that the best multiple of g leaves (scale-invariant; 0 for g parallel to p). It weighs the
    error modes as CG's convergence does."""
    ag = apply_matrix(g, diag)
    gb = (g * b).flatten(1).sum(1)
    gag = (g * ag).flatten(1).sum(1).clamp(min=1e-30)
    bp = (b * p).flatten(1).sum(1).clamp(min=1e-30)
    return 1.0 - gb * gb / (gag * bp)


def crop_pair(p, unknown, diag, crop: int, gen):
    """A random crop^3 window of a solution p, as an exact pair of the window's own problem:
    A_w (zero outside the window) applied to the cropped p gives its right-hand side, so the
    energy loss stays the error's A-norm share for that problem."""
    import torch

    nz, ny, nx = p.shape
    o = [int(torch.randint(0, m - crop + 1, (1,), generator=gen)) for m in (nz, ny, nx)]
    sl = (slice(o[0], o[0] + crop), slice(o[1], o[1] + crop), slice(o[2], o[2] + crop))
    unk, dg = unknown[sl], diag[0][sl]
    pc = torch.where(unk, p[sl], torch.zeros_like(p[sl]))
    return apply_matrix(pc, dg), pc, unk, dg


def train(bs, ps, unknown, epochs: int = 20, batch: int = 2, lr: float = 1e-3, seed: int = 0, time_budget: float = 600.0, log=print,
          diag=None, amp: bool = True, init=None, crop: int = 0):
    """Adam on the energy loss of the network's guess (see energy_loss); init: flat parameters to
    start from; crop > 0: train on random crop^3 windows of the pairs (crop_pair) instead of the
    whole grid."""
    import torch

    torch.manual_seed(seed)
    model = build_unet()
    if init is not None:
        load_parameters(model, init)
    model = model.to(bs.device).to(memory_format=torch.channels_last_3d)
    diag = diag.to(bs.device).float()[None]
    # output scale: p / s is O(output_scale)
    with torch.no_grad():
        ratios = []
        for i in range(0, len(bs), max(1, len(bs) // 16)):
            _, s = network_input(bs[i], unknown)
            ratios.append(float(ps[i][unknown].std() / s))
        ratios = [v for v in ratios if np.isfinite(v) and v > 0.0]
        output_scale = float(np.median(ratios)) if ratios else 1.0
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    gen = torch.Generator().manual_seed(seed)
    n = len(bs)
    t0 = time.time()
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n, device=bs.device)
        total = 0.0
        for i in range(0, n, batch):
            idx = perm[i:i + batch]
            if crop:
                wins = [crop_pair(ps[j], unknown, diag, crop, gen) for j in idx.tolist()]
                bw, pw = [w[0] for w in wins], [w[1] for w in wins]
                unk_b, diag_b = torch.stack([w[2] for w in wins]), torch.stack([w[3] for w in wins])
            else:
                bw, pw = [bs[j] for j in idx.tolist()], [ps[j] for j in idx.tolist()]
                unk_b, diag_b = unknown[None], diag
            xs, ss = zip(*[network_input(b, u) for b, u in zip(bw, unk_b if crop else [unknown] * len(bw))])
            x = torch.cat(xs).contiguous(memory_format=torch.channels_last_3d)
            try:
                with torch.autocast("cuda", dtype=torch.bfloat16, enabled=amp):
                    out = model(x)
            except RuntimeError as e:   # no BF16 kernels for these convolutions: train in FP32
                if not amp:
                    raise
                log(f"BF16 autocast failed ({e}); training in FP32")
                amp = False
                out = model(x)
            g = out.float()[:, 0] * unk_b.float()
            bn = torch.stack([b / float(s) for b, s in zip(bw, ss)])
            pn = torch.stack([p / float(s) for p, s in zip(pw, ss)])
            loss = energy_loss(g, bn, pn, diag_b).mean()
            opt.zero_grad()
            loss.backward()
            opt.step()
            total += loss.item() * len(idx)
            step += 1
            if step % 20 == 0:
                log(f"step {step}: batch energy loss {loss.item():.4f} ({time.time() - t0:.0f} s)")
            if time.time() - t0 > time_budget:
                break
        log(f"epoch {ep + 1}: energy loss {total / max(1, min(n, i + batch)):.4f} ({time.time() - t0:.0f} s)")
        if time.time() - t0 > time_budget:
            break
    return model, output_scale
