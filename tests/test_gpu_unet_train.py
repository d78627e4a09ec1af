"""The warm-start network's training kernels (csrc/unet_train.hip) against the FP64 oracle
(oracle/unet_oracle.cpp, pinned against PyTorch's CPU convolutions in tests/test_oracle_unet.py).

Per operator, on BF16 inputs (the oracle computes in FP64 from the same BF16 values): every
3x3x3 convolution shape of the network (forward, input gradient with each of its epilogues:
ReLU mask, plain, the concatenation split; weight and bias gradients), the two transposed
convolutions (forward, input gradient, weight gradient), max pooling and its backward with the
skip connection's gradient. Then the whole network: the trainer's output, loss and parameter
gradient on a batch of two exact pairs (r = A e) against the oracle's network in its BF16 mode,
and one Adam step against the oracle's Adam.

Tolerances: weight gradients are FP32 sums of exact BF16 products (measured ~1e-7 relative);
BF16 outputs carry one rounding (relative L2 ~1e-3); the network compounds roundings layer by
layer on both sides, where an FP32 and an FP64 sum can round to neighbouring BF16 values.
"""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu

CONV_SHAPES = [(2, 0, 16), (16, 0, 16), (16, 0, 32), (32, 0, 32), (32, 0, 64), (64, 0, 64), (32, 32, 32), (16, 16, 16)]
GRID = (32, 8, 6)   # nx % 16 == 0, ny % 4 == 0
B = 2


def _t():
    import torch

    return torch


def dev_bf16(a):
    """numpy float array -> (device int16 tensor of its BF16 bits, float64 copy of the BF16 values)."""
    import ds_smoke.unet as UN

    bits = UN.bf16_bits(a)
    t = _t().from_numpy(bits.view(np.int16).copy()).to("cuda")
    return t, UN.from_bf16_bits(bits).astype(np.float64)


def dev_f32(a):
    return _t().from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")


def host_bf16(t):
    import ds_smoke.unet as UN

    _t().cuda.synchronize()
    return UN.from_bf16_bits(t.cpu().numpy().view(np.uint16)).astype(np.float64)


def host_f32(t):
    _t().cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


def empty_bf16(shape):
    return _t().zeros(shape, dtype=_t().int16, device="cuda")


def empty_f32(shape):
    return _t().zeros(shape, dtype=_t().float32, device="cuda")


def relu_like(rng, shape):
    """A ReLU output: about half the values 0, the rest positive."""
    v = rng.standard_normal(shape)
    return np.where(v > 0, v, 0.0)


@pytest.fixture(scope="module")
def ops():
    from ds_smoke.unet import Ops

    return Ops()


def _stored_input(x, c1):
    """The input layer stores its 2 channels with stride 4 (channels 2, 3 zero)."""
    if c1 != 2:
        return x
    pad = np.zeros(x.shape[:-1] + (4,))
    pad[..., :2] = x
    return pad


@pytest.mark.parametrize("shape", CONV_SHAPES, ids=[f"{a}+{b}->{c}" for a, b, c in CONV_SHAPES])
def test_conv3_forward_and_weight_gradient(ops, shape):
    import unet as U

    c1, c2, cout = shape
    nx, ny, nz = GRID
    rng = np.random.default_rng(100 + c1 + c2 + cout)
    x1 = rng.standard_normal((B, nz, ny, nx, c1))
    x2 = rng.standard_normal((B, nz, ny, nx, c2)) if c2 else None
    w = rng.standard_normal((cout, c1 + c2, 3, 3, 3)) / np.sqrt(27 * (c1 + c2))
    bias = rng.standard_normal(cout) * 0.1
    t1, x1b = dev_bf16(_stored_input(x1, c1))
    t2, x2b = dev_bf16(x2) if c2 else (None, None)
    xcat = np.concatenate([x1b[..., :c1], x2b], -1) if c2 else x1b[..., :c1]
    out = empty_bf16((B, nz, ny, nx, cout))
    ops.conv3_forward(B, GRID, c1, c2, cout, True, t1, t2, dev_f32(w), dev_f32(bias), out)
    ref = U.conv3_forward(xcat, _bf16(w), bias, relu=True)
    e_fwd = rel(host_bf16(out), ref)
    # weight and bias gradient from a BF16 output gradient
    dz = relu_like(rng, (B, nz, ny, nx, cout))
    tdz, dzb = dev_bf16(dz)
    dw = empty_f32(w.shape)
    db = empty_f32((cout,))
    ops.conv3_wgrad(B, GRID, c1, c2, cout, t1, t2, tdz, dw, db)
    rdw, rdb = U.conv3_wgrad(xcat, dzb)
    e_dw, e_db = rel(host_f32(dw), rdw), rel(host_f32(db), rdb)
    print(f"conv3 {shape}: forward {e_fwd:.2e}, weight gradient {e_dw:.2e}, bias gradient {e_db:.2e}")
    assert e_fwd < 4e-3
    assert e_dw < 1e-5 and e_db < 1e-6


def _bf16(a):
    import ds_smoke.unet as UN

    return UN.bf16_round(a).astype(np.float64)


@pytest.mark.parametrize("shape", CONV_SHAPES[1:], ids=[f"{a}+{b}->{c}" for a, b, c in CONV_SHAPES[1:]])
def test_conv3_input_gradient(ops, shape):
    import unet as U

    c1, c2, cout = shape
    nx, ny, nz = GRID
    rng = np.random.default_rng(200 + c1 + c2 + cout)
    w = rng.standard_normal((cout, c1 + c2, 3, 3, 3)) / np.sqrt(27 * cout)
    dz = relu_like(rng, (B, nz, ny, nx, cout))
    tdz, dzb = dev_bf16(dz)
    ref = U.conv3_dgrad(dzb, _bf16(w))
    if c2:
        d1 = empty_bf16((B, nz, ny, nx, c1))
        d2 = empty_f32((B, nz, ny, nx, c2))
        ops.conv3_dgrad(B, GRID, c1, c2, cout, tdz, dev_f32(w), None, d1, d2)
        e1, e2 = rel(host_bf16(d1), ref[..., :c1]), rel(host_f32(d2), ref[..., c1:])
        print(f"conv3 dgrad {shape}: concatenation split {e1:.2e} (BF16), {e2:.2e} (FP32)")
        assert e1 < 4e-3 and e2 < 1e-5
        return
    mask = relu_like(rng, (B, nz, ny, nx, c1))
    tmask, maskb = dev_bf16(mask)
    d1 = empty_bf16((B, nz, ny, nx, c1))
    ops.conv3_dgrad(B, GRID, c1, 0, cout, tdz, dev_f32(w), tmask, d1)
    e_mask = rel(host_bf16(d1), np.where(maskb > 0, ref, 0.0))
    ops.conv3_dgrad(B, GRID, c1, 0, cout, tdz, dev_f32(w), None, d1)
    e_plain = rel(host_bf16(d1), ref)
    print(f"conv3 dgrad {shape}: masked {e_mask:.2e}, plain {e_plain:.2e}")
    assert e_mask < 4e-3 and e_plain < 4e-3


@pytest.mark.parametrize("shape", [(64, 32), (32, 16)], ids=["64->32", "32->16"])
def test_tconv2_forward_and_gradients(ops, shape):
    import unet as U

    cin, cout = shape
    cg = (16, 4, 3)   # coarse grid
    nx, ny, nz = cg
    rng = np.random.default_rng(300 + cin)
    x = relu_like(rng, (B, nz, ny, nx, cin))
    w = rng.standard_normal((cin, cout, 2, 2, 2)) / np.sqrt(8 * cin)
    bias = rng.standard_normal(cout) * 0.1
    tx, xb = dev_bf16(x)
    out = empty_bf16((B, 2 * nz, 2 * ny, 2 * nx, cout))
    ops.tconv2_forward(B, cg, cin, cout, tx, dev_f32(w), dev_f32(bias), out)
    e_fwd = rel(host_bf16(out), U.tconv2_forward(xb, _bf16(w), bias))
    dout = rng.standard_normal((B, 2 * nz, 2 * ny, 2 * nx, cout))
    tdo, dob = dev_bf16(dout)
    din = empty_bf16((B, nz, ny, nx, cin))
    ops.tconv2_dgrad(B, cg, cin, cout, tdo, dev_f32(w), tx, din)
    e_d = rel(host_bf16(din), np.where(xb > 0, U.tconv2_dgrad(dob, _bf16(w)), 0.0))
    dw = empty_f32(w.shape)
    db = empty_f32((cout,))
    ops.tconv2_wgrad(B, cg, cin, cout, tx, tdo, dw, db)
    rdw, rdb = U.tconv2_wgrad(xb, dob)
    e_dw, e_db = rel(host_f32(dw), rdw), rel(host_f32(db), rdb)
    print(f"tconv2 {shape}: forward {e_fwd:.2e}, input gradient {e_d:.2e}, weight gradient {e_dw:.2e}, bias {e_db:.2e}")
    assert e_fwd < 4e-3 and e_d < 4e-3
    assert e_dw < 1e-5 and e_db < 1e-6


@pytest.mark.parametrize("c", [16, 32])
def test_maxpool_and_backward(ops, c):
    import unet as U

    cg = (16, 4, 3)
    nx, ny, nz = cg
    rng = np.random.default_rng(400 + c)
    s = relu_like(rng, (B, 2 * nz, 2 * ny, 2 * nx, c))
    s[0, 0, 0, 0, :] = s[0, 0, 0, 1, :] = 1.5   # ties: the first maximum in x order takes the gradient
    ts, sb = dev_bf16(s)
    pooled = empty_bf16((B, nz, ny, nx, c))
    ops.maxpool_forward(B, cg, c, ts, pooled)
    np.testing.assert_array_equal(host_bf16(pooled), U.maxpool_forward(sb))
    dp = rng.standard_normal((B, nz, ny, nx, c))
    tdp, dpb = dev_bf16(dp)
    skip = rng.standard_normal(s.shape).astype(np.float32)
    din = empty_bf16(s.shape)
    ops.maxpool_backward(B, cg, c, ts, tdp, dev_f32(skip), din)
    ref = U.maxpool_backward(sb, dpb, skip.astype(np.float64))
    ref = np.where(sb > 0, _bf16(ref), 0.0)
    got = host_bf16(din)
    print(f"maxpool {c}: backward max |diff| {np.abs(got - ref).max():.2e}")
    np.testing.assert_allclose(got, ref, rtol=8e-3, atol=0)


def _pack_input(r, unknown):
    """The trainer's network input, replicated: (bf16(r * (1/s)), unknown) with s the FP64 rms of r
    over the unknowns and 1/s, r * (1/s) in FP32."""
    out = np.zeros(r.shape + (2,))
    for b in range(r.shape[0]):
        rb = r[b].astype(np.float32)
        s = np.sqrt(np.sum(rb.astype(np.float64)[unknown] ** 2) / unknown.sum())
        inv = np.float32(1.0 / s)
        out[b, ..., 0] = np.where(unknown, _bf16(rb * inv), 0.0)
        out[b, ..., 1] = unknown
    return out


@pytest.mark.parametrize("loss_kind", ["l2", "energy"])
def test_network_gradient_and_adam_match_the_oracle(loss_kind):
    import unet as U
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context
    from ds_smoke.unet import Trainer, init_parameters

    sc = scenes.plume(64, with_target=False)
    sc.resolution = (64, 16, 8)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_FORWARD)
    _, code = ctx.masks()
    unknown = code > 0
    rng = np.random.default_rng(7)
    params = init_parameters(3)
    e = (rng.standard_normal((B,) + code.shape) * unknown).astype(np.float32)
    et = ctx.to_device(e)
    # the residuals the trainer forms (same kernel arithmetic as ds_warm_start_pair) and <e, A e>
    rt = ctx._zeros((B,) + code.shape)
    ent = ctx._zeros((B,), dtype=ctx.torch.float64)
    for b in range(B):
        ctx.warm_start_pair(et[b], residual=rt[b], energy=ent[b:b + 1])
    r = ctx.to_host(rt).astype(np.float64)
    np.testing.assert_allclose(ctx.to_host(ent), [np.dot(e[b].ravel().astype(np.float64), r[b].ravel()) for b in range(B)], rtol=1e-6)
    tr = Trainer(ctx, B, params)
    loss = ctx.to_host(tr.step(et, 0.0, loss=loss_kind)).copy()
    g = ctx.to_host(tr.read("output")).reshape((B,) + code.shape).astype(np.float64)
    grad = ctx.to_host(tr.read("gradient")).astype(np.float64)
    x2 = _pack_input(r.astype(np.float32), unknown)
    out = U.network(params.astype(np.float64), x2, bf16=True)
    g_ref = out * unknown
    if loss_kind == "energy":
        loss_ref, dg_ref = U.energy_loss(code, g_ref, r, e.astype(np.float64))
    else:
        loss_ref, dg_ref = U.l2_loss(code, g_ref, e.astype(np.float64))
    _, grad_ref = U.network(params.astype(np.float64), x2, bf16=True, dout=dg_ref / B)
    e_out = rel(g, g_ref)
    e_grad = rel(grad, grad_ref)
    per_layer = []
    off = 0
    for k, cin, cout in U_layers():
        nw = cin * cout * {0: 27, 1: 8, 2: 1}[k]
        per_layer.append(rel(grad[off:off + nw], grad_ref[off:off + nw]))
        off += nw + cout
    print(f"network ({loss_kind} loss): output {e_out:.2e}, loss {loss} vs {loss_ref}, gradient {e_grad:.2e}; per layer "
          + " ".join(f"{v:.1e}" for v in per_layer))
    # measured (energy loss): output 5.4e-4, loss 4e-7 absolute, gradient 2.0e-3 (layers 2e-3 .. 1.2e-2)
    assert e_out < 5e-3
    np.testing.assert_allclose(loss, loss_ref, rtol=0, atol=2e-4)
    assert e_grad < 2e-2
    assert max(per_layer) < 6e-2
    # one Adam step against the oracle's (same gradient in)
    lr = 1e-3
    tr.step(et, lr, loss=loss_kind)
    p1 = ctx.to_host(tr.read("params")).astype(np.float64)
    g1 = ctx.to_host(tr.read("gradient")).astype(np.float64)
    p_ref, m, v = params.astype(np.float64).copy(), np.zeros(params.size), np.zeros(params.size)
    U.adam(p_ref, g1, m, v, 1, lr=lr)
    np.testing.assert_allclose(p1, p_ref, rtol=1e-6, atol=1e-7)
    tr.close()
    ctx.close()


def U_layers():
    from ds_smoke.unet import LAYERS

    return LAYERS
