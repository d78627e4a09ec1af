"""Host side of the warm-start network's GPU training (csrc/unet_train.hip) over the C ABI.

The network is a 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions with ReLU, 2x2x2 max
pooling, 2x2x2 stride-2 transposed convolutions, a 1x1x1 head). Its parameters are one flat FP32
vector in PyTorch state_dict order (per layer the weight, then the bias; Conv3d weights
[co][ci][kz][ky][kx], ConvTranspose3d weights [ci][co][kz][ky][kx]), the order ds_warm_start_load
and ds_unet_trainer_create take. Every computation runs in the HIP library; torch tensors only hold
device memory (BF16 tensors are int16 tensors carrying the BF16 bits).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .solver import _ptr, load_library

# (kind, cin, cout): kind 0 = 3x3x3 conv, 1 = 2x2x2 stride-2 transposed conv, 2 = 1x1x1 head
LAYERS = ((0, 2, 16), (0, 16, 16), (0, 16, 32), (0, 32, 32), (0, 32, 64), (0, 64, 64), (1, 64, 32),
          (0, 64, 32), (0, 32, 32), (1, 32, 16), (0, 32, 16), (0, 16, 16), (2, 16, 1))
TAPS = {0: 27, 1: 8, 2: 1}
NAMES = ("enc1a", "enc1b", "enc2a", "enc2b", "bota", "botb", "up2", "dec2a", "dec2b", "up1", "dec1a", "dec1b", "head")


def weight_shape(kind, cin, cout):
    if kind == 0:
        return (cout, cin, 3, 3, 3)
    if kind == 1:
        return (cin, cout, 2, 2, 2)
    return (cout, cin, 1, 1, 1)


def parameter_count() -> int:
    return sum(cin * cout * TAPS[k] + cout for k, cin, cout in LAYERS)


def split(flat):
    """[(weight, bias)] views of a flat parameter vector, per layer."""
    out, off = [], 0
    for k, cin, cout in LAYERS:
        nw = cin * cout * TAPS[k]
        out.append((flat[off:off + nw].reshape(weight_shape(k, cin, cout)), flat[off + nw:off + nw + cout]))
        off += nw + cout
    return out


def init_parameters(seed: int = 0) -> np.ndarray:
    """PyTorch's default initialisation of Conv3d / ConvTranspose3d (weights and biases uniform in
    +-1/sqrt(fan_in), fan_in = weight.size(1) * taps), drawn from numpy's PCG64 with `seed`."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, cin, cout in LAYERS:
        shape = weight_shape(k, cin, cout)
        fan_in = shape[1] * TAPS[k]
        bound = 1.0 / np.sqrt(fan_in)
        parts.append(rng.uniform(-bound, bound, int(np.prod(shape))))
        parts.append(rng.uniform(-bound, bound, cout))
    return np.concatenate(parts).astype(np.float32)


def bf16_bits(a) -> np.ndarray:
    """FP32 -> BF16 bits (round to nearest even), as uint16."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(h) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(a) -> np.ndarray:
    return from_bf16_bits(bf16_bits(a))


def _s(stream):
    return C.c_void_p(stream.cuda_stream) if stream is not None else None


class Ops:
    """The network's operators (ds_unet_*). Tensors: int16 (BF16 bits) or float32 torch tensors on
    the device, channels last, batch stacked along z; grid = the operator's grid as documented in
    include/ds/ds_smoke.h."""

    def __init__(self, stream=None):
        self.lib = load_library()
        self.stream = stream

    def _check(self, st, where):
        if st != 0:
            raise RuntimeError(f"{where}: status {st}: {self.lib.ds_last_error().decode()}")

    def conv3_forward(self, batch, grid, cin1, cin2, cout, relu, in1, in2, weight, bias, out):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_conv3_forward(batch, nx, ny, nz, cin1, cin2, cout, int(relu), _ptr(in1), _ptr(in2), _ptr(weight),
                                                   _ptr(bias), _ptr(out), _s(self.stream)), "ds_unet_conv3_forward")

    def conv3_dgrad(self, batch, grid, cin1, cin2, cout, dz, weight, input_relu, din1, din2=None):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_conv3_dgrad(batch, nx, ny, nz, cin1, cin2, cout, _ptr(dz), _ptr(weight), _ptr(input_relu), _ptr(din1),
                                                 _ptr(din2), _s(self.stream)), "ds_unet_conv3_dgrad")

    def conv3_wgrad(self, batch, grid, cin1, cin2, cout, in1, in2, dz, dweight, dbias):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_conv3_wgrad(batch, nx, ny, nz, cin1, cin2, cout, _ptr(in1), _ptr(in2), _ptr(dz), _ptr(dweight),
                                                 _ptr(dbias), _s(self.stream)), "ds_unet_conv3_wgrad")

    def tconv2_forward(self, batch, grid, cin, cout, x, weight, bias, out):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_tconv2_forward(batch, nx, ny, nz, cin, cout, _ptr(x), _ptr(weight), _ptr(bias), _ptr(out),
                                                    _s(self.stream)), "ds_unet_tconv2_forward")

    def tconv2_dgrad(self, batch, grid, cin, cout, dout, weight, input_relu, din):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_tconv2_dgrad(batch, nx, ny, nz, cin, cout, _ptr(dout), _ptr(weight), _ptr(input_relu), _ptr(din),
                                                  _s(self.stream)), "ds_unet_tconv2_dgrad")

    def tconv2_wgrad(self, batch, grid, cin, cout, x, dout, dweight, dbias):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_tconv2_wgrad(batch, nx, ny, nz, cin, cout, _ptr(x), _ptr(dout), _ptr(dweight), _ptr(dbias),
                                                  _s(self.stream)), "ds_unet_tconv2_wgrad")

    def maxpool_forward(self, batch, grid, c, x, out):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_maxpool_forward(batch, nx, ny, nz, c, _ptr(x), _ptr(out), _s(self.stream)), "ds_unet_maxpool_forward")

    def maxpool_backward(self, batch, grid, c, x, dout, skip, din):
        nx, ny, nz = grid
        self._check(self.lib.ds_unet_maxpool_backward(batch, nx, ny, nz, c, _ptr(x), _ptr(dout), _ptr(skip), _ptr(din), _s(self.stream)),
                    "ds_unet_maxpool_backward")


class Trainer:
    """ds_unet_trainer: Adam on the network's loss over batches of errors (a solution, or a solution
    minus a base guess) on the context's grid and matrix, on the context's stream."""

    def __init__(self, ctx, batch: int, params: np.ndarray):
        self.ctx = ctx
        self.lib = ctx.lib
        self.batch = int(batch)
        flat = np.ascontiguousarray(params, dtype=np.float32).ravel()
        assert flat.size == parameter_count()
        h = C.c_void_p()
        st = self.lib.ds_unet_trainer_create(ctx.h, self.batch, flat.ctypes.data_as(C.c_void_p), flat.size, C.byref(h))
        ctx._check(st, "ds_unet_trainer_create")
        self.h = h
        self.loss = ctx._zeros((self.batch,), dtype=ctx.torch.float64)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ds_unet_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, error, lr: float, loss: str = "l2"):
        """error: (batch, nz, ny, nx) float32, 0 off the unknowns; lr = 0: gradient only; loss:
        "energy" or "l2". Returns the per-sample loss (device float64 tensor, valid after the stream
        drains)."""
        kind = {"energy": 0, "l2": 1}[loss]
        self.ctx._check(self.lib.ds_unet_trainer_step(self.h, _ptr(error), kind, float(lr), _ptr(self.loss)), "ds_unet_trainer_step")
        return self.loss

    def read(self, what: str):
        t = self.ctx.torch
        code = {"params": 0, "gradient": 1, "output": 2}[what]
        n = self.batch * int(np.prod(self.ctx.n)) if what == "output" else parameter_count()
        out = self.ctx._zeros((n,), dtype=t.float32)
        self.ctx._check(self.lib.ds_unet_trainer_read(self.h, code, _ptr(out)), "ds_unet_trainer_read")
        return out
