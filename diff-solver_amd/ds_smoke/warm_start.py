"""Warm-start network of the pressure solves: the flat parameter format, the saved-weights file,
and a PyTorch definition of the same network used only as an independent cross-check of the HIP
inference (tests/test_gpu_warm_start.py) and of the training loss (tests/test_warm_start_cpu.py).

The network is a 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions with ReLU, 2x2x2 max
pooling, 2x2x2 stride-2 transposed convolutions, a 1x1x1 head; 339,681 parameters). Input: a
right-hand side (or residual) r divided by s = rms(r) over the unknowns, and the unknown indicator.
Output: a direction the solver applies at its A-norm-best multiple. Inference runs in the HIP
library (csrc/unet.hip) and so does training (csrc/unet_train.hip, driven by
ds_smoke/train_warm_start.py); the parameters are one flat FP32 vector in PyTorch state_dict order,
the order ds_warm_start_load expects.
"""
from __future__ import annotations

import json

import numpy as np

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


def levels_for(meta: dict, resolution) -> int:
    """Coarsenings (ds_warm_start_configure) that bring a grid to the network's training grid:
    log2(nx / training nx) when that is 0, 1 or 2, else 0."""
    trained = int(meta.get("grid", resolution[0]))
    ratio = int(resolution[0]) // max(1, trained)
    return {1: 0, 2: 1, 4: 2}.get(ratio, 0) if ratio * trained == int(resolution[0]) else 0


def load_into(ctx, path: str) -> dict:
    """Load the weights file into a context, coarsening its grid to the training grid."""
    flat, scale, meta = load(path)
    ctx.warm_start_load(flat, scale, levels=levels_for(meta, ctx.n))
    return meta


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
