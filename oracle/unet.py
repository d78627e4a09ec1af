"""ctypes binding of the FP64 warm-start-network oracle (oracle/unet_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP training kernels. Arrays
are numpy float64, channels last: (B, nz, ny, nx, C) for activations; weights in PyTorch's
layout: conv (cout, cin, 3, 3, 3), transposed conv (cin, cout, 2, 2, 2), flattened as stored.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libunet_oracle.so")
_lib = None
DP = C.POINTER(C.c_double)
I = C.c_int


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB_PATH)
        for name in ("unet_conv3_dgrad", "unet_tconv2_dgrad"):
            getattr(L, name).argtypes = [I, I, I, I, I, I, DP, DP, DP]
        for name in ("unet_conv3_wgrad", "unet_tconv2_wgrad"):
            getattr(L, name).argtypes = [I, I, I, I, I, I, DP, DP, DP, DP]
        L.unet_conv3_forward.argtypes = [I, I, I, I, I, I, DP, DP, DP, I, DP]
        L.unet_tconv2_forward.argtypes = [I, I, I, I, I, I, DP, DP, DP, DP]
        L.unet_maxpool_forward.argtypes = [I, I, I, I, I, DP, DP]
        L.unet_maxpool_backward.argtypes = [I, I, I, I, I, DP, DP, DP]
        L.unet_parameter_count.restype = C.c_int64
        L.unet_network.argtypes = [I, I, I, I, DP, DP, I, DP, DP, DP]
        L.unet_energy_loss.argtypes = [I, I, I, I, C.c_void_p, C.c_uint, DP, DP, DP, DP, DP]
        L.unet_l2_loss.argtypes = [I, I, I, I, C.c_void_p, C.c_uint, DP, DP, DP, DP]
        L.unet_adam.argtypes = [C.c_int64, DP, DP, DP, DP, I, C.c_double, C.c_double, C.c_double, C.c_double]
        for f in ("unet_conv3_forward", "unet_conv3_dgrad", "unet_conv3_wgrad", "unet_tconv2_forward", "unet_tconv2_dgrad",
                  "unet_tconv2_wgrad", "unet_maxpool_forward", "unet_maxpool_backward", "unet_network", "unet_energy_loss",
                  "unet_l2_loss", "unet_adam"):
            getattr(L, f).restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(DP) if a is not None else None


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dims(x):
    B, nz, ny, nx, c = x.shape
    return B, nx, ny, nz, c


def conv3_forward(x, w, bias=None, relu=False):
    B, nx, ny, nz, cin = _dims(x)
    cout = w.shape[0]
    out = np.zeros((B, nz, ny, nx, cout))
    lib().unet_conv3_forward(B, nx, ny, nz, cin, cout, _p(_f(x)), _p(_f(w)), _p(_f(bias) if bias is not None else None), int(relu),
                             _p(out))
    return out


def conv3_dgrad(dout, w):
    B, nx, ny, nz, cout = _dims(dout)
    cin = w.shape[1]
    din = np.zeros((B, nz, ny, nx, cin))
    lib().unet_conv3_dgrad(B, nx, ny, nz, cin, cout, _p(_f(dout)), _p(_f(w)), _p(din))
    return din


def conv3_wgrad(x, dout):
    B, nx, ny, nz, cin = _dims(x)
    cout = dout.shape[-1]
    dw = np.zeros((cout, cin, 3, 3, 3))
    db = np.zeros(cout)
    lib().unet_conv3_wgrad(B, nx, ny, nz, cin, cout, _p(_f(x)), _p(_f(dout)), _p(dw), _p(db))
    return dw, db


def tconv2_forward(x, w, bias=None):
    B, nx, ny, nz, cin = _dims(x)
    cout = w.shape[1]
    out = np.zeros((B, 2 * nz, 2 * ny, 2 * nx, cout))
    lib().unet_tconv2_forward(B, nx, ny, nz, cin, cout, _p(_f(x)), _p(_f(w)), _p(_f(bias) if bias is not None else None), _p(out))
    return out


def tconv2_dgrad(dout, w):
    B, nz2, ny2, nx2, cout = dout.shape
    cin = w.shape[0]
    din = np.zeros((B, nz2 // 2, ny2 // 2, nx2 // 2, cin))
    lib().unet_tconv2_dgrad(B, nx2 // 2, ny2 // 2, nz2 // 2, cin, cout, _p(_f(dout)), _p(_f(w)), _p(din))
    return din


def tconv2_wgrad(x, dout):
    B, nx, ny, nz, cin = _dims(x)
    cout = dout.shape[-1]
    dw = np.zeros((cin, cout, 2, 2, 2))
    db = np.zeros(cout)
    lib().unet_tconv2_wgrad(B, nx, ny, nz, cin, cout, _p(_f(x)), _p(_f(dout)), _p(dw), _p(db))
    return dw, db


def maxpool_forward(x):
    B, nx, ny, nz, c = _dims(x)
    out = np.zeros((B, nz // 2, ny // 2, nx // 2, c))
    lib().unet_maxpool_forward(B, nx // 2, ny // 2, nz // 2, c, _p(_f(x)), _p(out))
    return out


def maxpool_backward(x, dout, din=None):
    B, nx, ny, nz, c = _dims(x)
    din = np.zeros_like(_f(x)) if din is None else _f(din).copy()
    lib().unet_maxpool_backward(B, nx // 2, ny // 2, nz // 2, c, _p(_f(x)), _p(_f(dout)), _p(din))
    return din


def parameter_count():
    return int(lib().unet_parameter_count())


def network(params, x2, bf16=False, dout=None):
    """Forward (and with dout the parameter gradient of <dout, out>) of the whole network.
    x2: (B, nz, ny, nx, 2). Returns out (B, nz, ny, nx) [, grad (flat)]."""
    B, nx, ny, nz, c = _dims(x2)
    assert c == 2
    out = np.zeros((B, nz, ny, nx))
    grad = np.zeros(parameter_count()) if dout is not None else None
    lib().unet_network(B, nx, ny, nz, _p(_f(params)), _p(_f(x2)), int(bf16), _p(out), _p(_f(dout) if dout is not None else None),
                       _p(grad))
    return (out, grad) if dout is not None else out


def energy_loss(code, g, r, e, periodic=0):
    """Per-sample loss 1 - <g,r>^2 / (<g,Ag> <e,r>) and its gradient with respect to g.
    code: (nz, ny, nx) uint8; g, r, e: (B, nz, ny, nx)."""
    B, nz, ny, nx = g.shape
    code = np.ascontiguousarray(code, dtype=np.uint8)
    loss = np.zeros(B)
    dg = np.zeros((B, nz, ny, nx))
    lib().unet_energy_loss(B, nx, ny, nz, code.ctypes.data_as(C.c_void_p), periodic, _p(_f(g)), _p(_f(r)), _p(_f(e)), _p(loss), _p(dg))
    return loss, dg


def l2_loss(code, g, e, periodic=0):
    """Per-sample loss ||gamma g - e||^2 / ||e||^2 (gamma = <g, A e> / <g, A g>) and its gradient with
    respect to g. code: (nz, ny, nx) uint8; g, e: (B, nz, ny, nx)."""
    B, nz, ny, nx = g.shape
    code = np.ascontiguousarray(code, dtype=np.uint8)
    loss = np.zeros(B)
    dg = np.zeros((B, nz, ny, nx))
    lib().unet_l2_loss(B, nx, ny, nz, code.ctypes.data_as(C.c_void_p), periodic, _p(_f(g)), _p(_f(e)), _p(loss), _p(dg))
    return loss, dg


def adam(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in place on float64 arrays p, m, v."""
    lib().unet_adam(p.size, _p(p), _p(_f(g)), _p(m), _p(v), t, lr, b1, b2, eps)
