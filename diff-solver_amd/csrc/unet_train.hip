// unet_train.hip — training of the warm-start network on the GPU with this library's own kernels
// (unet_kernels.hpp): the forward pass over a batch, the loss, the backward pass (input and weight
// gradients of every layer) and Adam, plus per-operator entry points that tests compare with the
// FP64 oracle. No counterpart in the reference, which has no learned part.
//
// A training sample is an error e on the solve's unknowns (a solution, or a solution minus a base
// guess); its residual r = A e is what the solver would hand the network. The network sees
// (r / rms(r), unknown indicator) and returns g, which the solver applies at the A-norm-best step
// gamma = <g, r> / <g, A g>. Losses (q = <g, r>, D = <g, A g>, E = <e, r>):
//   kind 0, energy: L = 1 - q^2 / (D E), the share of ||e||_A^2 that gamma g leaves;
//     dL/dg = -(2 q / (D E)) (r - (q / D) A g);
//   kind 1, L2: L = ||gamma g - e||^2 / ||e||^2, the error the guess leaves in the 2-norm, which
//     weighs the smooth error modes that set CG's iteration count more than the A-norm does;
//     with f = gamma g - e: dL/dg = (2 <f, g> (r - 2 gamma A g) / D + 2 gamma f) / ||e||^2.
//
// Precision: master weights, their gradients and Adam's moments are FP32; the convolutions take
// BF16 weights (repacked after every update) and BF16 activations and activation gradients, with
// FP32 accumulation; the skip connections' share of a concatenation's input gradient stays FP32
// until it meets the pooled path's gradient; every reduction over voxels runs in FP64.
#include <algorithm>
#include <string>
#include <vector>

#include "context.hpp"
#include "unet_kernels.hpp"

namespace ds {

using namespace unet;

namespace {

struct LayerShape {
    int kind;   // 0 = 3x3x3 conv, 1 = 2x2x2 transposed conv, 2 = 1x1x1 head
    int cin, cout;
};
// PyTorch state_dict order of the network (per layer: weight, then bias).
constexpr LayerShape kNet[13] = {
    {0, 2, 16},  {0, 16, 16}, {0, 16, 32}, {0, 32, 32}, {0, 32, 64}, {0, 64, 64}, {1, 64, 32},
    {0, 64, 32}, {0, 32, 32}, {1, 32, 16}, {0, 32, 16}, {0, 16, 16}, {2, 16, 1},
};
constexpr int kTaps[3] = {27, 8, 1};

uint64_t weights_of(const LayerShape& L) { return (uint64_t)L.cin * L.cout * kTaps[L.kind]; }
uint64_t packed_elems(int cin, int cout, int taps) { return (uint64_t)taps * ((cin + 15) / 16) * (cout / 16) * 64 * 4; }

constexpr int kMT = 4;   // rows per wave of the convolution kernels

uint32_t waves_blocks(uint64_t waves) { return (uint32_t)((waves + 3) / 4); }

// ---- launchers over the network's layer shapes --------------------------------------------------

template <int C1, int C2, int COUT, int EPI, int CU = 0>
void conv_launch(const Dims& d, int batch, const uint16_t* in1, const uint16_t* in2, const uint16_t* wpk, const ConvOut& o, hipStream_t s) {
    const uint64_t waves = (uint64_t)(d.nx / 16) * (d.ny / kMT) * d.nz * batch;
    conv3_kernel<C1, C2, COUT, kMT, EPI, CU><<<waves_blocks(waves), 256, 0, s>>>(d, batch, in1, in2,
                                                                                reinterpret_cast<const bf16x4_t*>(wpk), o);
}

// Forward 3x3x3 convolution of one of the network's shapes (c1 + c2 -> cout), with bias and
// optional ReLU.
void conv_forward(int c1, int c2, int cout, bool relu, const Dims& d, int batch, const uint16_t* in1, const uint16_t* in2,
                  const uint16_t* wpk, const float* bias, uint16_t* out, hipStream_t s) {
    ConvOut o;
    o.out = out;
    o.bias = bias;
#define DS_CONV_FWD(A, B, C)                                                                         \
    if (c1 == A && c2 == B && cout == C) {                                                           \
        if (relu) conv_launch<A, B, C, kEpiBiasRelu>(d, batch, in1, in2, wpk, o, s);               \
        else conv_launch<A, B, C, kEpiBias>(d, batch, in1, in2, wpk, o, s);                        \
        return;                                                                                      \
    }
    DS_CONV_FWD(2, 0, 16)
    DS_CONV_FWD(16, 0, 16)
    DS_CONV_FWD(16, 0, 32)
    DS_CONV_FWD(32, 0, 32)
    DS_CONV_FWD(32, 0, 64)
    DS_CONV_FWD(64, 0, 64)
    DS_CONV_FWD(32, 32, 32)
    DS_CONV_FWD(16, 16, 16)
#undef DS_CONV_FWD
    throw Failure{DS_ERROR_UNSUPPORTED, "3x3x3 convolution shape not in the network"};
}

// Input gradient of a forward convolution (c1 + c2 -> cout) from its output gradient dz (cout
// channels): c2 = 0 and mask: din1 = mask > 0 ? dgrad : 0; c2 = 0 without mask: din1 = dgrad;
// c2 > 0: din1 = the first c1 channels (BF16), din2 = the last c2 (FP32). wpk: kind-2 packing.
void conv_dgrad(int c1, int c2, int cout, const Dims& d, int batch, const uint16_t* dz, const uint16_t* wpk, const uint16_t* mask,
                uint16_t* din1, float* din2, hipStream_t s) {
    ConvOut o;
    o.out = din1;
    o.mask = mask;
    o.out_f32 = din2;
#define DS_CONV_DGRAD(A, C)                                                                          \
    if (c1 == A && c2 == 0 && cout == C) {                                                           \
        if (mask) conv_launch<C, 0, A, kEpiMask>(d, batch, dz, nullptr, wpk, o, s);                \
        else conv_launch<C, 0, A, kEpiPlain>(d, batch, dz, nullptr, wpk, o, s);                    \
        return;                                                                                      \
    }
    DS_CONV_DGRAD(16, 16)
    DS_CONV_DGRAD(16, 32)
    DS_CONV_DGRAD(32, 32)
    DS_CONV_DGRAD(32, 64)
    DS_CONV_DGRAD(64, 64)
#undef DS_CONV_DGRAD
    if (c1 == 32 && c2 == 32 && cout == 32) return conv_launch<32, 0, 64, kEpiSplit, 32>(d, batch, dz, nullptr, wpk, o, s);
    if (c1 == 16 && c2 == 16 && cout == 16) return conv_launch<16, 0, 32, kEpiSplit, 16>(d, batch, dz, nullptr, wpk, o, s);
    throw Failure{DS_ERROR_UNSUPPORTED, "3x3x3 convolution input gradient: shape not in the network"};
}

// Weight-gradient partial slots: waves and groups per wave for `groups` 16-voxel groups, with the
// slots of one call capped at kWgradSlotFloats floats.
constexpr uint64_t kWgradSlotFloats = 32ull << 20;
struct WgradPlan {
    uint32_t blocks, waves, per_wave;
};
WgradPlan wgrad_plan(uint64_t groups, int taps, int mc, int nc) {
    // waves are launched in blocks of 4: the cap is a multiple of 4, so the launched waves fit too
    uint64_t cap = kWgradSlotFloats / ((uint64_t)taps * mc * nc);
    cap = std::max<uint64_t>(4, std::min<uint64_t>(cap, 1024)) & ~3ull;
    uint64_t waves = std::min<uint64_t>(groups, cap);
    const uint64_t per = (groups + waves - 1) / waves;
    waves = (groups + per - 1) / per;
    WgradPlan p;
    p.blocks = waves_blocks(waves);
    p.waves = p.blocks * 4;
    p.per_wave = (uint32_t)per;
    return p;
}

// dW (PyTorch layout) and db of a 3x3x3 convolution (c1 + c2 -> cout) from its input [in1, in2]
// and pre-activation gradient dz. slots: at least kWgradSlotFloats floats; cpart: 256 * 64 doubles.
template <int C>
void channel_sum(uint64_t n, const uint16_t* t, double* cpart, float* out, hipStream_t s) {
    channel_sum_kernel<C><<<256, 256, 0, s>>>((uint32_t)n, t, cpart);
    channel_sum_finish_kernel<<<1, 64, 0, s>>>(cpart, 256, C, out);
}

void channel_sum_any(int c, uint64_t n, const uint16_t* t, double* cpart, float* out, hipStream_t s) {
    switch (c) {
        case 16: return channel_sum<16>(n, t, cpart, out, s);
        case 32: return channel_sum<32>(n, t, cpart, out, s);
        case 64: return channel_sum<64>(n, t, cpart, out, s);
        default: throw Failure{DS_ERROR_UNSUPPORTED, "channel count not in the network"};
    }
}

void conv_wgrad(int c1, int c2, int cout, const Dims& d, int batch, const uint16_t* in1, const uint16_t* in2, const uint16_t* dz,
                float* slots, double* cpart, float* dw, float* db, hipStream_t s) {
    const uint64_t groups = (uint64_t)(d.nx / 16) * d.ny * d.nz * batch;
    const int nc = ((c1 + c2 + 15) / 16) * 16;
    const WgradPlan p = wgrad_plan(groups, 27, cout, nc);
    const dim3 grid(p.blocks, 27);
    bool done = false;
#define DS_WGRAD(A, B, C)                                                                                    \
    if (!done && c1 == A && c2 == B && cout == C) {                                                          \
        wgrad_kernel<0, C, A, B><<<grid, 256, 0, s>>>(d, batch, dz, in1, in2, p.per_wave, slots);          \
        done = true;                                                                                         \
    }
    DS_WGRAD(2, 0, 16)
    DS_WGRAD(16, 0, 16)
    DS_WGRAD(16, 0, 32)
    DS_WGRAD(32, 0, 32)
    DS_WGRAD(32, 0, 64)
    DS_WGRAD(64, 0, 64)
    DS_WGRAD(32, 32, 32)
    DS_WGRAD(16, 16, 16)
#undef DS_WGRAD
    if (!done) throw Failure{DS_ERROR_UNSUPPORTED, "3x3x3 convolution weight gradient: shape not in the network"};
    const uint32_t per = 27u * cout * nc;
    wgrad_reduce_kernel<<<blocks_for(per), 256, 0, s>>>(slots, p.waves, 27, cout, nc, c1 + c2, dw);
    channel_sum_any(cout, (uint64_t)d.nx * d.ny * d.nz * batch, dz, cpart, db, s);
}

void tconv_forward(int cin, int cout, const Dims& dc, int batch, const uint16_t* in, const uint16_t* wpk, const float* bias, uint16_t* out,
                   hipStream_t s) {
    const uint64_t waves = (uint64_t)(dc.nx / 16) * dc.ny * dc.nz * batch;
    const bf16x4_t* w = reinterpret_cast<const bf16x4_t*>(wpk);
    if (cin == 64 && cout == 32) tconv2_kernel<64, 32><<<waves_blocks(waves), 256, 0, s>>>(dc, batch, in, w, bias, out);
    else if (cin == 32 && cout == 16) tconv2_kernel<32, 16><<<waves_blocks(waves), 256, 0, s>>>(dc, batch, in, w, bias, out);
    else throw Failure{DS_ERROR_UNSUPPORTED, "transposed convolution shape not in the network"};
}

void tconv_dgrad(int cin, int cout, const Dims& dc, int batch, const uint16_t* dout, const uint16_t* wpk, const uint16_t* mask,
                 uint16_t* din, hipStream_t s) {
    const uint64_t waves = (uint64_t)(dc.nx / 16) * dc.ny * dc.nz * batch;
    const bf16x4_t* w = reinterpret_cast<const bf16x4_t*>(wpk);
    if (cin == 64 && cout == 32) tconv2_dgrad_kernel<64, 32><<<waves_blocks(waves), 256, 0, s>>>(dc, batch, dout, w, mask, din);
    else if (cin == 32 && cout == 16) tconv2_dgrad_kernel<32, 16><<<waves_blocks(waves), 256, 0, s>>>(dc, batch, dout, w, mask, din);
    else throw Failure{DS_ERROR_UNSUPPORTED, "transposed convolution shape not in the network"};
}

void tconv_wgrad(int cin, int cout, const Dims& dc, int batch, const uint16_t* in, const uint16_t* dout, float* slots, double* cpart,
                 float* dw, float* db, hipStream_t s) {
    const uint64_t groups = (uint64_t)(dc.nx / 16) * dc.ny * dc.nz * batch;
    const WgradPlan p = wgrad_plan(groups, 8, cin, cout);
    const dim3 grid(p.blocks, 8);
    if (cin == 64 && cout == 32) wgrad_kernel<1, 64, 32, 0><<<grid, 256, 0, s>>>(dc, batch, in, dout, nullptr, p.per_wave, slots);
    else if (cin == 32 && cout == 16) wgrad_kernel<1, 32, 16, 0><<<grid, 256, 0, s>>>(dc, batch, in, dout, nullptr, p.per_wave, slots);
    else throw Failure{DS_ERROR_UNSUPPORTED, "transposed convolution shape not in the network"};
    const uint32_t per = 8u * cin * cout;
    wgrad_reduce_kernel<<<blocks_for(per), 256, 0, s>>>(slots, p.waves, 8, cin, cout, cout, dw);
    channel_sum_any(cout, (uint64_t)dc.nx * dc.ny * dc.nz * batch * 8, dout, cpart, db, s);
}

void pool_forward(int c, const Dims& dc, int batch, const uint16_t* in, uint16_t* out, hipStream_t s) {
    const uint64_t n = (uint64_t)dc.nx * dc.ny * dc.nz * batch;
    if (c == 16) maxpool2_kernel<16><<<blocks_for(n * 4), 256, 0, s>>>(dc, (uint32_t)n, in, out);
    else if (c == 32) maxpool2_kernel<32><<<blocks_for(n * 8), 256, 0, s>>>(dc, (uint32_t)n, in, out);
    else throw Failure{DS_ERROR_UNSUPPORTED, "pooling channel count not in the network"};
}

void pool_backward(int c, const Dims& dc, int batch, const uint16_t* in, const uint16_t* dp, const float* skip, uint16_t* din, hipStream_t s) {
    const uint64_t n = (uint64_t)dc.nx * dc.ny * dc.nz * batch;
    if (c == 16) maxpool2_backward_kernel<16><<<blocks_for(n * 16), 256, 0, s>>>(dc, (uint32_t)n, in, dp, skip, din);
    else if (c == 32) maxpool2_backward_kernel<32><<<blocks_for(n * 32), 256, 0, s>>>(dc, (uint32_t)n, in, dp, skip, din);
    else throw Failure{DS_ERROR_UNSUPPORTED, "pooling channel count not in the network"};
}

void pack(const float* w, int cin, int cout, int taps, int kind, uint16_t* pk, hipStream_t s) {
    const uint64_t e = packed_elems(cin, cout, taps);
    pack_weights_kernel<<<blocks_for(e), 256, 0, s>>>(w, cin, cout, taps, kind, pk);
}

// ---- training-specific kernels -------------------------------------------------------------------

// g = head(d1) on the unknowns, 0 elsewhere (FP32).
__global__ __launch_bounds__(256) void head_forward_kernel(uint32_t n, uint32_t total, const uint16_t* __restrict__ d1,
                                                           const float* __restrict__ w, const float* __restrict__ bias,
                                                           const uint8_t* __restrict__ code, float* __restrict__ g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    float v = bias[0];
#pragma unroll
    for (int c = 0; c < 16; ++c) v += w[c] * from_bf16(d1[16 * (size_t)i + c]);
    g[i] = code[i % n] != 0u ? v : 0.0f;
}

// r = A e for every sample of the batch (blockIdx.y = sample; 0 off the unknowns).
__global__ __launch_bounds__(256) void batch_residual_kernel(Geo geo, uint32_t per, const uint8_t* __restrict__ code,
                                                            const float* __restrict__ e, float* __restrict__ r) {
    const uint32_t n = geo.n;
    const float* eb = e + (size_t)blockIdx.y * n;
    float* rb = r + (size_t)blockIdx.y * n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t d = code[i];
        if (!d) {
            rb[i] = 0.0f;
            continue;
        }
        int x, y, z;
        decode_cell(geo, i, x, y, z);
        const int c[3]{x, y, z}, ext[3]{geo.nx, geo.ny, geo.nz};
        double s = (double)d * (double)eb[i];
        for (int a = 0; a < 3; ++a)
            for (int dir = -1; dir <= 1; dir += 2) {
                int q[3]{c[0], c[1], c[2]};
                q[a] += dir;
                if (q[a] < 0 || q[a] >= ext[a]) {
                    if (!(per & (1u << a))) continue;
                    q[a] = q[a] < 0 ? ext[a] - 1 : 0;
                }
                const uint32_t j = cidx(geo, q[0], q[1], q[2]);
                if (code[j]) s -= (double)eb[j];
            }
        rb[i] = (float)s;
    }
}

constexpr int kLossSums = 6;   // <g, r>, <g, A g>, <g, g>, <g, e>, <e, e>, <e, r>

// A g (the pressure matrix on the unknowns, periodic bit a wraps axis a; g = 0 off the unknowns)
// and the per-sample partial sums of the loss: partials[(b * G + block) * kLossSums + k].
__global__ __launch_bounds__(256) void loss_partials_kernel(Geo geo, uint32_t per, const uint8_t* __restrict__ code,
                                                            const float* __restrict__ g, const float* __restrict__ r,
                                                            const float* __restrict__ e, float* __restrict__ ag,
                                                            double* __restrict__ partials) {
    __shared__ double lds[kLossSums * kMaxWaves];
    const uint32_t n = geo.n, b = blockIdx.y;
    const float* gb = g + (size_t)b * n;
    const float* rb = r + (size_t)b * n;
    const float* eb = e + (size_t)b * n;
    float* agb = ag + (size_t)b * n;
    double v[kLossSums]{};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t d = code[i];
        if (!d) {
            agb[i] = 0.0f;
            continue;
        }
        int x, y, z;
        decode_cell(geo, i, x, y, z);
        const int c[3]{x, y, z}, ext[3]{geo.nx, geo.ny, geo.nz};
        double s = (double)d * (double)gb[i];
        for (int a = 0; a < 3; ++a)
            for (int dir = -1; dir <= 1; dir += 2) {
                int q[3]{c[0], c[1], c[2]};
                q[a] += dir;
                if (q[a] < 0 || q[a] >= ext[a]) {
                    if (!(per & (1u << a))) continue;
                    q[a] = q[a] < 0 ? ext[a] - 1 : 0;
                }
                s -= (double)gb[cidx(geo, q[0], q[1], q[2])];
            }
        agb[i] = (float)s;
        const double gi = gb[i], ri = rb[i], ei = eb[i];
        v[0] += gi * ri;
        v[1] += gi * s;
        v[2] += gi * gi;
        v[3] += gi * ei;
        v[4] += ei * ei;
        v[5] += ei * ri;
    }
    block_sum<kLossSums>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kLossSums; ++k) partials[((size_t)b * gridDim.x + blockIdx.x) * kLossSums + k] = v[k];
}

// Per-sample loss of `kind` and dL/dg / batch (the gradient of the batch mean) on the unknowns.
__global__ __launch_bounds__(256) void loss_gradient_kernel(uint32_t n, uint32_t batch, int kind, const uint8_t* __restrict__ code,
                                                            const float* __restrict__ g, const float* __restrict__ r,
                                                            const float* __restrict__ e, const float* __restrict__ ag,
                                                            const double* __restrict__ partials, uint32_t G, float* __restrict__ dg,
                                                            double* __restrict__ loss) {
    __shared__ double lds[kLossSums * kMaxWaves];
    const uint32_t b = blockIdx.y;
    double v[kLossSums]{};
    for (uint32_t k = threadIdx.x; k < G; k += blockDim.x)
#pragma unroll
        for (int j = 0; j < kLossSums; ++j) v[j] += partials[((size_t)b * G + k) * kLossSums + j];
    block_sum<kLossSums>(v, lds);
    const double q = v[0], D = v[1], GG = v[2], P = v[3], E2 = v[4], E = v[5];
    const bool ok = D > 0.0 && E > 0.0 && E2 > 0.0 && isfinite(q) && isfinite(D) && isfinite(GG);
    const double gamma = ok ? q / D : 0.0;
    // dg = c_r r + c_ag A g + c_g g + c_e e
    double c_r = 0.0, c_ag = 0.0, c_g = 0.0, c_e = 0.0, L = 1.0;
    if (ok && kind == 0) {
        L = 1.0 - q * q / (D * E);
        c_r = -2.0 * q / (D * E);
        c_ag = -c_r * gamma;
    } else if (ok) {
        const double fg = gamma * GG - P;   // <f, g>
        L = (gamma * gamma * GG - 2.0 * gamma * P + E2) / E2;
        c_r = 2.0 * fg / D / E2;
        c_ag = -2.0 * gamma * c_r;
        c_g = 2.0 * gamma * gamma / E2;
        c_e = -2.0 * gamma / E2;
    }
    const double inv_b = 1.0 / (double)batch;
    if (blockIdx.x == 0 && threadIdx.x == 0) loss[b] = L;
    const size_t off = (size_t)b * n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dg[off + i] = code[i] ? (float)(inv_b * (c_r * (double)r[off + i] + c_ag * (double)ag[off + i] + c_g * (double)g[off + i] +
                                                 c_e * (double)e[off + i]))
                              : 0.0f;
}

// Head backward: g_d1[v][c] = d1 > 0 ? bf16(dg[v] w[c]) : 0, and per-block partial sums of
// dW[c] = sum_v dg[v] d1[v][c] (c < 16) and db = sum_v dg[v] (slot 16): partials[block * 17 + k].
__global__ __launch_bounds__(256) void head_backward_kernel(uint32_t total, const float* __restrict__ dg, const uint16_t* __restrict__ d1,
                                                            const float* __restrict__ w, uint16_t* __restrict__ gd1,
                                                            double* __restrict__ partials) {
    __shared__ double lds[17 * kMaxWaves];
    double v[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) v[k] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const float gi = dg[i];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint16_t h = d1[16 * (size_t)i + c];
            v[c] += (double)gi * (double)from_bf16(h);
            gd1[16 * (size_t)i + c] = bf16_positive(h) ? to_bf16(gi * w[c]) : (uint16_t)0;
        }
        v[16] += (double)gi;
    }
    block_sum<17>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 17; ++k) partials[(size_t)blockIdx.x * 17 + k] = v[k];
}

__global__ __launch_bounds__(64) void head_backward_finish_kernel(const double* __restrict__ partials, uint32_t blocks,
                                                                  float* __restrict__ dw, float* __restrict__ db) {
    const int k = (int)threadIdx.x;
    if (k >= 17) return;
    double s = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) s += partials[(size_t)b * 17 + k];
    if (k < 16) dw[k] = (float)s;
    else db[0] = (float)s;
}

// Adam (Kingma & Ba, with bias correction): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// p -= lr (m / c1) / (sqrt(v / c2) + eps), c1 = 1 - b1^t, c2 = 1 - b2^t.
__global__ __launch_bounds__(256) void adam_kernel(uint32_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                                   float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// r = A e on the unknowns and per-block partials of <e, A e>.
__global__ __launch_bounds__(256) void pair_kernel(Geo geo, uint32_t per, const uint8_t* __restrict__ code, const float* __restrict__ e,
                                                   float* __restrict__ r, double* __restrict__ partials) {
    __shared__ double lds[kMaxWaves];
    double v[1]{0.0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < geo.n; i += gridDim.x * blockDim.x) {
        const uint32_t d = code[i];
        if (!d) {
            r[i] = 0.0f;
            continue;
        }
        int x, y, z;
        decode_cell(geo, i, x, y, z);
        const int c[3]{x, y, z}, ext[3]{geo.nx, geo.ny, geo.nz};
        double s = (double)d * (double)e[i];
        for (int a = 0; a < 3; ++a)
            for (int dir = -1; dir <= 1; dir += 2) {
                int q[3]{c[0], c[1], c[2]};
                q[a] += dir;
                if (q[a] < 0 || q[a] >= ext[a]) {
                    if (!(per & (1u << a))) continue;
                    q[a] = q[a] < 0 ? ext[a] - 1 : 0;
                }
                const uint32_t j = cidx(geo, q[0], q[1], q[2]);
                if (code[j]) s -= (double)e[j];
            }
        r[i] = (float)s;
        v[0] += (double)e[i] * s;
    }
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(256) void pair_finish_kernel(const double* __restrict__ partials, uint32_t G, double* __restrict__ energy) {
    __shared__ double lds[kMaxWaves];
    const double* parts[1]{partials};
    double red[1];
    reduce_partials<1>(parts, G, red, lds);
    if (threadIdx.x == 0) *energy = red[0];
}

constexpr uint32_t kLossBlocks = 256;   // blocks per sample of the loss / input kernels

}  // namespace

// ---- the trainer -----------------------------------------------------------------------------------

}  // namespace ds

struct ds_unet_trainer {
    ds_context* ctx = nullptr;
    uint32_t batch = 0;
    ds::unet::Dims d1{}, d2{}, d3{};
    uint64_t n1 = 0, n2 = 0, n3 = 0;   // voxels per sample per level
    uint64_t count = 0;                // parameters
    uint64_t off_w[13]{}, off_b[13]{};
    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    uint32_t step = 0;
    uint16_t* wf[13]{};   // forward packing
    uint16_t* wd[13]{};   // input-gradient packing
    // activations
    uint16_t *x = nullptr, *a1 = nullptr, *s1 = nullptr, *p1 = nullptr, *a2 = nullptr, *s2 = nullptr, *p2 = nullptr, *a3 = nullptr,
             *c = nullptr, *u2 = nullptr, *a4 = nullptr, *dd2 = nullptr, *u1 = nullptr, *a5 = nullptr, *dd1 = nullptr;
    float *g = nullptr, *dg = nullptr, *ag = nullptr, *r = nullptr;
    // gradients (BF16; g_s*cat FP32)
    uint16_t *gd1 = nullptr, *ga5 = nullptr, *gu1 = nullptr, *gd2 = nullptr, *ga4 = nullptr, *gu2 = nullptr, *gc = nullptr,
             *ga3 = nullptr, *gp2 = nullptr, *gs2 = nullptr, *ga2 = nullptr, *gp1 = nullptr, *gs1 = nullptr, *ga1 = nullptr;
    float *gs1cat = nullptr, *gs2cat = nullptr;
    float* slots = nullptr;
    double *cpart = nullptr, *lpart = nullptr, *npart = nullptr, *hpart = nullptr;
    std::vector<void*> allocations;
    ~ds_unet_trainer() {
        for (void* p : allocations) (void)hipFree(p);
    }
    template <typename T>
    T* alloc(size_t count_) {
        void* p = nullptr;
        DS_HIP(hipMalloc(&p, std::max<size_t>(count_ * sizeof(T), 1)));
        allocations.push_back(p);
        return static_cast<T*>(p);
    }
};

namespace ds {
namespace {

void repack_all(ds_unet_trainer* t) {
    hipStream_t s = t->ctx->stream;
    for (int k = 0; k < 12; ++k) {
        const LayerShape& L = kNet[k];
        const float* w = t->params + t->off_w[k];
        if (L.kind == 0) {
            pack(w, L.cin, L.cout, 27, 0, t->wf[k], s);
            if (k > 0) pack(w, L.cout, L.cin, 27, 2, t->wd[k], s);
        } else {
            pack(w, L.cin, L.cout, 8, 1, t->wf[k], s);
            pack(w, L.cout, L.cin, 8, 3, t->wd[k], s);
        }
    }
    check_launch("pack_weights_kernel");
}

void trainer_forward(ds_unet_trainer* t, const float* r) {
    ds_context* c = t->ctx;
    hipStream_t s = c->stream;
    const int B = (int)t->batch;
    const uint32_t n = c->geo.n;
    const float* P = t->params;
    auto bias = [&](int k) { return P + t->off_b[k]; };
    residual_norm_kernel<<<dim3(kLossBlocks, B), 256, 0, s>>>(n, r, c->code, t->npart);
    pack_batch_input_kernel<<<dim3(kLossBlocks, B), 256, 0, s>>>(n, c->unknowns, r, c->code, t->npart, kLossBlocks, t->x);
    conv_forward(2, 0, 16, true, t->d1, B, t->x, nullptr, t->wf[0], bias(0), t->a1, s);
    conv_forward(16, 0, 16, true, t->d1, B, t->a1, nullptr, t->wf[1], bias(1), t->s1, s);
    pool_forward(16, t->d2, B, t->s1, t->p1, s);
    conv_forward(16, 0, 32, true, t->d2, B, t->p1, nullptr, t->wf[2], bias(2), t->a2, s);
    conv_forward(32, 0, 32, true, t->d2, B, t->a2, nullptr, t->wf[3], bias(3), t->s2, s);
    pool_forward(32, t->d3, B, t->s2, t->p2, s);
    conv_forward(32, 0, 64, true, t->d3, B, t->p2, nullptr, t->wf[4], bias(4), t->a3, s);
    conv_forward(64, 0, 64, true, t->d3, B, t->a3, nullptr, t->wf[5], bias(5), t->c, s);
    tconv_forward(64, 32, t->d3, B, t->c, t->wf[6], bias(6), t->u2, s);
    conv_forward(32, 32, 32, true, t->d2, B, t->u2, t->s2, t->wf[7], bias(7), t->a4, s);
    conv_forward(32, 0, 32, true, t->d2, B, t->a4, nullptr, t->wf[8], bias(8), t->dd2, s);
    tconv_forward(32, 16, t->d2, B, t->dd2, t->wf[9], bias(9), t->u1, s);
    conv_forward(16, 16, 16, true, t->d1, B, t->u1, t->s1, t->wf[10], bias(10), t->a5, s);
    conv_forward(16, 0, 16, true, t->d1, B, t->a5, nullptr, t->wf[11], bias(11), t->dd1, s);
    const uint32_t total = n * (uint32_t)B;
    head_forward_kernel<<<blocks_for(total), 256, 0, s>>>(n, total, t->dd1, P + t->off_w[12], bias(12), c->code, t->g);
    check_launch("warm-start network forward (training)");
}

void trainer_backward(ds_unet_trainer* t) {
    ds_context* c = t->ctx;
    hipStream_t s = c->stream;
    const int B = (int)t->batch;
    const uint32_t total = c->geo.n * (uint32_t)B;
    float* G = t->grad;
    auto gw = [&](int k) { return G + t->off_w[k]; };
    auto gb = [&](int k) { return G + t->off_b[k]; };
    head_backward_kernel<<<256, 256, 0, s>>>(total, t->dg, t->dd1, t->params + t->off_w[12], t->gd1, t->hpart);
    head_backward_finish_kernel<<<1, 64, 0, s>>>(t->hpart, 256, gw(12), gb(12));
    // level 1 decoder
    conv_wgrad(16, 0, 16, t->d1, B, t->a5, nullptr, t->gd1, t->slots, t->cpart, gw(11), gb(11), s);
    conv_dgrad(16, 0, 16, t->d1, B, t->gd1, t->wd[11], t->a5, t->ga5, nullptr, s);
    conv_wgrad(16, 16, 16, t->d1, B, t->u1, t->s1, t->ga5, t->slots, t->cpart, gw(10), gb(10), s);
    conv_dgrad(16, 16, 16, t->d1, B, t->ga5, t->wd[10], nullptr, t->gu1, t->gs1cat, s);
    tconv_wgrad(32, 16, t->d2, B, t->dd2, t->gu1, t->slots, t->cpart, gw(9), gb(9), s);
    tconv_dgrad(32, 16, t->d2, B, t->gu1, t->wd[9], t->dd2, t->gd2, s);
    // level 2 decoder
    conv_wgrad(32, 0, 32, t->d2, B, t->a4, nullptr, t->gd2, t->slots, t->cpart, gw(8), gb(8), s);
    conv_dgrad(32, 0, 32, t->d2, B, t->gd2, t->wd[8], t->a4, t->ga4, nullptr, s);
    conv_wgrad(32, 32, 32, t->d2, B, t->u2, t->s2, t->ga4, t->slots, t->cpart, gw(7), gb(7), s);
    conv_dgrad(32, 32, 32, t->d2, B, t->ga4, t->wd[7], nullptr, t->gu2, t->gs2cat, s);
    tconv_wgrad(64, 32, t->d3, B, t->c, t->gu2, t->slots, t->cpart, gw(6), gb(6), s);
    tconv_dgrad(64, 32, t->d3, B, t->gu2, t->wd[6], t->c, t->gc, s);
    // bottleneck
    conv_wgrad(64, 0, 64, t->d3, B, t->a3, nullptr, t->gc, t->slots, t->cpart, gw(5), gb(5), s);
    conv_dgrad(64, 0, 64, t->d3, B, t->gc, t->wd[5], t->a3, t->ga3, nullptr, s);
    conv_wgrad(32, 0, 64, t->d3, B, t->p2, nullptr, t->ga3, t->slots, t->cpart, gw(4), gb(4), s);
    conv_dgrad(32, 0, 64, t->d3, B, t->ga3, t->wd[4], nullptr, t->gp2, nullptr, s);
    pool_backward(32, t->d3, B, t->s2, t->gp2, t->gs2cat, t->gs2, s);
    // level 2 encoder
    conv_wgrad(32, 0, 32, t->d2, B, t->a2, nullptr, t->gs2, t->slots, t->cpart, gw(3), gb(3), s);
    conv_dgrad(32, 0, 32, t->d2, B, t->gs2, t->wd[3], t->a2, t->ga2, nullptr, s);
    conv_wgrad(16, 0, 32, t->d2, B, t->p1, nullptr, t->ga2, t->slots, t->cpart, gw(2), gb(2), s);
    conv_dgrad(16, 0, 32, t->d2, B, t->ga2, t->wd[2], nullptr, t->gp1, nullptr, s);
    pool_backward(16, t->d2, B, t->s1, t->gp1, t->gs1cat, t->gs1, s);
    // level 1 encoder
    conv_wgrad(16, 0, 16, t->d1, B, t->a1, nullptr, t->gs1, t->slots, t->cpart, gw(1), gb(1), s);
    conv_dgrad(16, 0, 16, t->d1, B, t->gs1, t->wd[1], t->a1, t->ga1, nullptr, s);
    conv_wgrad(2, 0, 16, t->d1, B, t->x, nullptr, t->ga1, t->slots, t->cpart, gw(0), gb(0), s);
    check_launch("warm-start network backward (training)");
}

}  // namespace

namespace {

// Temporary device buffers of one operator call, released after the stream has drained.
struct Scratch {
    hipStream_t s;
    std::vector<void*> ptrs;
    explicit Scratch(hipStream_t st) : s(st) {}
    template <typename T>
    T* get(size_t n) {
        void* p = nullptr;
        DS_HIP(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 1)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~Scratch() {
        (void)hipStreamSynchronize(s);
        for (void* p : ptrs) (void)hipFree(p);
    }
};

void check_dims(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, bool conv) {
    require(batch > 0 && nx > 0 && ny > 0 && nz > 0, "batch and grid extents must be positive");
    require(nx % 16 == 0, "nx must be a multiple of 16");
    if (conv) require(ny % kMT == 0, "ny must be a multiple of 4");
    require((uint64_t)batch * nx * ny * nz * 8 * 64 < (1ull << 32), "tensor too large for 32-bit element indices");
}

}  // namespace
}  // namespace ds

using namespace ds;

extern "C" {

ds_status ds_unet_trainer_create(ds_context* c, uint32_t batch, const float* params, uint64_t count, ds_unet_trainer** out) {
    if (!c || !params || !out || batch == 0) {
        set_error("ds_unet_trainer_create: null argument or zero batch");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    return guard([&] {
        if (count != unet_parameter_count())
            throw Failure{DS_ERROR_INVALID_ARGUMENT, "ds_unet_trainer_create: parameter count does not match the network (" +
                                                         std::to_string(unet_parameter_count()) + " expected)"};
        const Geo& g = c->geo;
        if (g.nx % 64 || g.ny % 16 || g.nz % 4)
            throw Failure{DS_ERROR_UNSUPPORTED, "ds_unet_trainer_create: the network needs nx % 64 == 0, ny % 16 == 0 and nz % 4 == 0"};
        if ((uint64_t)g.n * batch * 64 >= (1ull << 32))
            throw Failure{DS_ERROR_INVALID_ARGUMENT, "ds_unet_trainer_create: batch x cells too large for 32-bit element indices"};
        auto* t = new ds_unet_trainer();
        try {
            t->ctx = c;
            t->batch = batch;
            t->d1 = Dims{g.nx, g.ny, g.nz};
            t->d2 = Dims{g.nx / 2, g.ny / 2, g.nz / 2};
            t->d3 = Dims{g.nx / 4, g.ny / 4, g.nz / 4};
            t->n1 = g.n;
            t->n2 = g.n / 8;
            t->n3 = g.n / 64;
            t->count = count;
            uint64_t off = 0;
            for (int k = 0; k < 13; ++k) {
                t->off_w[k] = off;
                off += weights_of(kNet[k]);
                t->off_b[k] = off;
                off += kNet[k].cout;
            }
            t->params = t->alloc<float>(count);
            t->grad = t->alloc<float>(count);
            t->m = t->alloc<float>(count);
            t->v = t->alloc<float>(count);
            hipStream_t s = c->stream;
            DS_HIP(hipMemcpyAsync(t->params, params, count * sizeof(float), hipMemcpyHostToDevice, s));
            DS_HIP(hipMemsetAsync(t->grad, 0, count * sizeof(float), s));
            DS_HIP(hipMemsetAsync(t->m, 0, count * sizeof(float), s));
            DS_HIP(hipMemsetAsync(t->v, 0, count * sizeof(float), s));
            for (int k = 0; k < 12; ++k) {
                const LayerShape& L = kNet[k];
                const int taps = kTaps[L.kind];
                t->wf[k] = t->alloc<uint16_t>(packed_elems(L.cin, L.cout, taps));
                if (k > 0) t->wd[k] = t->alloc<uint16_t>(packed_elems(L.cout, L.cin, taps));
            }
            const size_t B = batch, N1 = t->n1 * B, N2 = t->n2 * B, N3 = t->n3 * B;
            t->x = t->alloc<uint16_t>(4 * N1);
            t->a1 = t->alloc<uint16_t>(16 * N1);
            t->s1 = t->alloc<uint16_t>(16 * N1);
            t->p1 = t->alloc<uint16_t>(16 * N2);
            t->a2 = t->alloc<uint16_t>(32 * N2);
            t->s2 = t->alloc<uint16_t>(32 * N2);
            t->p2 = t->alloc<uint16_t>(32 * N3);
            t->a3 = t->alloc<uint16_t>(64 * N3);
            t->c = t->alloc<uint16_t>(64 * N3);
            t->u2 = t->alloc<uint16_t>(32 * N2);
            t->a4 = t->alloc<uint16_t>(32 * N2);
            t->dd2 = t->alloc<uint16_t>(32 * N2);
            t->u1 = t->alloc<uint16_t>(16 * N1);
            t->a5 = t->alloc<uint16_t>(16 * N1);
            t->dd1 = t->alloc<uint16_t>(16 * N1);
            t->g = t->alloc<float>(N1);
            t->dg = t->alloc<float>(N1);
            t->ag = t->alloc<float>(N1);
            t->r = t->alloc<float>(N1);
            t->gd1 = t->alloc<uint16_t>(16 * N1);
            t->ga5 = t->alloc<uint16_t>(16 * N1);
            t->gu1 = t->alloc<uint16_t>(16 * N1);
            t->gs1cat = t->alloc<float>(16 * N1);
            t->gd2 = t->alloc<uint16_t>(32 * N2);
            t->ga4 = t->alloc<uint16_t>(32 * N2);
            t->gu2 = t->alloc<uint16_t>(32 * N2);
            t->gs2cat = t->alloc<float>(32 * N2);
            t->gc = t->alloc<uint16_t>(64 * N3);
            t->ga3 = t->alloc<uint16_t>(64 * N3);
            t->gp2 = t->alloc<uint16_t>(32 * N3);
            t->gs2 = t->alloc<uint16_t>(32 * N2);
            t->ga2 = t->alloc<uint16_t>(32 * N2);
            t->gp1 = t->alloc<uint16_t>(16 * N2);
            t->gs1 = t->alloc<uint16_t>(16 * N1);
            t->ga1 = t->alloc<uint16_t>(16 * N1);
            t->slots = t->alloc<float>(kWgradSlotFloats);
            t->cpart = t->alloc<double>(256 * 64);
            t->lpart = t->alloc<double>((size_t)kLossBlocks * B * kLossSums);
            t->npart = t->alloc<double>((size_t)kLossBlocks * B);
            t->hpart = t->alloc<double>(256 * 17);
            repack_all(t);
            DS_HIP(hipStreamSynchronize(s));
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

ds_status ds_unet_trainer_destroy(ds_unet_trainer* t) {
    if (!t) return DS_OK;
    return guard([&] {
        (void)hipStreamSynchronize(t->ctx->stream);
        delete t;
    });
}

ds_status ds_unet_trainer_step(ds_unet_trainer* t, const float* error, uint32_t loss_kind, float learning_rate, double* loss) {
    if (!t || !error || !loss || loss_kind > 1) {
        set_error("ds_unet_trainer_step: null argument or unknown loss");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!(learning_rate >= 0.0f) || !std::isfinite(learning_rate)) {
        set_error("ds_unet_trainer_step: learning_rate must be finite and >= 0");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        ds_context* c = t->ctx;
        hipStream_t s = c->stream;
        const uint32_t B = t->batch;
        const uint32_t per = pcg_periodic_bits(c);
        batch_residual_kernel<<<dim3(kLossBlocks, B), 256, 0, s>>>(c->geo, per, c->code, error, t->r);
        trainer_forward(t, t->r);
        loss_partials_kernel<<<dim3(kLossBlocks, B), 256, 0, s>>>(c->geo, per, c->code, t->g, t->r, error, t->ag, t->lpart);
        loss_gradient_kernel<<<dim3(kLossBlocks, B), 256, 0, s>>>(c->geo.n, B, (int)loss_kind, c->code, t->g, t->r, error, t->ag,
                                                                  t->lpart, kLossBlocks, t->dg, loss);
        check_launch("warm-start loss");
        trainer_backward(t);
        if (learning_rate > 0.0f) {
            ++t->step;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = (float)(1.0 - std::pow(0.9, (double)t->step)), c2 = (float)(1.0 - std::pow(0.999, (double)t->step));
            adam_kernel<<<blocks_for(t->count), 256, 0, s>>>((uint32_t)t->count, t->params, t->grad, t->m, t->v, learning_rate, b1, b2,
                                                             1e-8f, c1, c2);
            check_launch("adam_kernel");
            repack_all(t);
        }
    });
}

ds_status ds_unet_trainer_read(ds_unet_trainer* t, uint32_t what, float* out) {
    if (!t || !out || what > 2) {
        set_error("ds_unet_trainer_read: null argument or unknown field");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        const float* src = what == 0 ? t->params : what == 1 ? t->grad : t->g;
        const size_t n = what == 2 ? (size_t)t->n1 * t->batch : (size_t)t->count;
        DS_HIP(hipMemcpyAsync(out, src, n * sizeof(float), hipMemcpyDeviceToDevice, t->ctx->stream));
    });
}

ds_status ds_warm_start_pair(ds_context* c, const float* error, float* residual, double* energy) {
    if (!c || !error || !residual || !energy) {
        set_error("ds_warm_start_pair: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        hipStream_t s = c->stream;
        pair_kernel<<<kPcgBlocks, 256, 0, s>>>(c->geo, pcg_periodic_bits(c), c->code, error, residual, c->pcg.partials);
        pair_finish_kernel<<<1, 256, 0, s>>>(c->pcg.partials, kPcgBlocks, energy);
        check_launch("pair_kernel");
    });
}


ds_status ds_unet_conv3_forward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin1, uint32_t cin2, uint32_t cout,
                                uint32_t relu, const uint16_t* in1, const uint16_t* in2, const float* weight, const float* bias,
                                uint16_t* out, void* stream) {
    return guard([&] {
        require(in1 && weight && bias && out && (cin2 == 0 || in2), "ds_unet_conv3_forward: null argument");
        check_dims(batch, nx, ny, nz, true);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Scratch sc(s);
        uint16_t* pk = sc.get<uint16_t>(packed_elems(cin1 + cin2, cout, 27));
        pack(weight, cin1 + cin2, cout, 27, 0, pk, s);
        conv_forward((int)cin1, (int)cin2, (int)cout, relu != 0, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, in1, in2, pk, bias, out, s);
        check_launch("ds_unet_conv3_forward");
    });
}

ds_status ds_unet_conv3_dgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin1, uint32_t cin2, uint32_t cout,
                              const uint16_t* dz, const float* weight, const uint16_t* input_relu, uint16_t* din1, float* din2,
                              void* stream) {
    return guard([&] {
        require(dz && weight && din1 && (cin2 == 0 || din2), "ds_unet_conv3_dgrad: null argument");
        require(cin1 >= 16, "ds_unet_conv3_dgrad: the input layer has no input gradient");
        require(cin2 == 0 || input_relu == nullptr, "ds_unet_conv3_dgrad: a concatenated input takes no ReLU mask");
        check_dims(batch, nx, ny, nz, true);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Scratch sc(s);
        uint16_t* pk = sc.get<uint16_t>(packed_elems(cout, cin1 + cin2, 27));
        pack(weight, cout, cin1 + cin2, 27, 2, pk, s);
        conv_dgrad((int)cin1, (int)cin2, (int)cout, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, dz, pk, input_relu, din1, din2, s);
        check_launch("ds_unet_conv3_dgrad");
    });
}

ds_status ds_unet_conv3_wgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin1, uint32_t cin2, uint32_t cout,
                              const uint16_t* in1, const uint16_t* in2, const uint16_t* dz, float* dweight, float* dbias, void* stream) {
    return guard([&] {
        require(in1 && dz && dweight && dbias && (cin2 == 0 || in2), "ds_unet_conv3_wgrad: null argument");
        check_dims(batch, nx, ny, nz, false);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Scratch sc(s);
        float* slots = sc.get<float>(kWgradSlotFloats);
        double* cpart = sc.get<double>(256 * 64);
        conv_wgrad((int)cin1, (int)cin2, (int)cout, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, in1, in2, dz, slots, cpart, dweight, dbias, s);
        check_launch("ds_unet_conv3_wgrad");
    });
}

ds_status ds_unet_tconv2_forward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin, uint32_t cout, const uint16_t* in,
                                 const float* weight, const float* bias, uint16_t* out, void* stream) {
    return guard([&] {
        require(in && weight && bias && out, "ds_unet_tconv2_forward: null argument");
        check_dims(batch, nx, ny, nz, false);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Scratch sc(s);
        uint16_t* pk = sc.get<uint16_t>(packed_elems(cin, cout, 8));
        pack(weight, cin, cout, 8, 1, pk, s);
        tconv_forward((int)cin, (int)cout, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, in, pk, bias, out, s);
        check_launch("ds_unet_tconv2_forward");
    });
}

ds_status ds_unet_tconv2_dgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin, uint32_t cout, const uint16_t* dout,
                               const float* weight, const uint16_t* input_relu, uint16_t* din, void* stream) {
    return guard([&] {
        require(dout && weight && input_relu && din, "ds_unet_tconv2_dgrad: null argument");
        check_dims(batch, nx, ny, nz, false);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Scratch sc(s);
        uint16_t* pk = sc.get<uint16_t>(packed_elems(cout, cin, 8));
        pack(weight, cout, cin, 8, 3, pk, s);
        tconv_dgrad((int)cin, (int)cout, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, dout, pk, input_relu, din, s);
        check_launch("ds_unet_tconv2_dgrad");
    });
}

ds_status ds_unet_tconv2_wgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin, uint32_t cout, const uint16_t* in,
                               const uint16_t* dout, float* dweight, float* dbias, void* stream) {
    return guard([&] {
        require(in && dout && dweight && dbias, "ds_unet_tconv2_wgrad: null argument");
        check_dims(batch, nx, ny, nz, false);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Scratch sc(s);
        float* slots = sc.get<float>(kWgradSlotFloats);
        double* cpart = sc.get<double>(256 * 64);
        tconv_wgrad((int)cin, (int)cout, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, in, dout, slots, cpart, dweight, dbias, s);
        check_launch("ds_unet_tconv2_wgrad");
    });
}

ds_status ds_unet_maxpool_forward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t c, const uint16_t* in, uint16_t* out,
                                  void* stream) {
    return guard([&] {
        require(in && out && batch > 0 && nx > 0 && ny > 0 && nz > 0, "ds_unet_maxpool_forward: null argument or empty grid");
        pool_forward((int)c, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, in, out, static_cast<hipStream_t>(stream));
        check_launch("ds_unet_maxpool_forward");
    });
}

ds_status ds_unet_maxpool_backward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t c, const uint16_t* in,
                                   const uint16_t* dout, const float* skip_gradient, uint16_t* din, void* stream) {
    return guard([&] {
        require(in && dout && din && batch > 0 && nx > 0 && ny > 0 && nz > 0, "ds_unet_maxpool_backward: null argument or empty grid");
        pool_backward((int)c, Dims{(int)nx, (int)ny, (int)nz}, (int)batch, in, dout, skip_gradient, din, static_cast<hipStream_t>(stream));
        check_launch("ds_unet_maxpool_backward");
    });
}

}  // extern "C"
