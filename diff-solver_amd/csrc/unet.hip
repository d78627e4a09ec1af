// unet.hip — the warm-start network of the pressure solves: a 3-level 3-D U-Net (16/32/64
// channels, 3x3x3 convolutions with ReLU, 2x2x2 max pooling, 2x2x2 stride-2 transposed
// convolutions, a 1x1x1 head) mapping the normalised right-hand side and the unknown mask to an
// initial guess of the solution. Inference only; training runs in PyTorch on solver-generated
// pairs (ds_smoke/warm_start.py) and writes the parameters this file loads.
//
// Activations are BF16, channels last (NDHWC, x fastest among voxels); every product accumulates
// in FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16) as an implicit GEMM: M = 16 voxels
// consecutive in x, N = 16 output channels, K = 16 input channels of one tap. The A fragment of a
// lane is 4 channels of one voxel (one 8-byte load); weights are repacked at load time into the
// fragment order, so a B fragment is one 8-byte load too.
//
// Fragment layout (16x16x16 BF16): lane l holds A[m = l%16][k = 4(l/16) + j], B[k = 4(l/16) + j]
// [n = l%16] (j = 0..3) and D[m = 4(l/16) + i][n = l%16] (i = 0..3).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

namespace ds {

namespace {

__host__ __device__ __forceinline__ uint16_t to_bf16(float f) {   // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

struct Dims {
    int nx, ny, nz;
};

__device__ __forceinline__ uint32_t vox(const Dims& d, int x, int y, int z) { return (uint32_t)x + (uint32_t)d.nx * ((uint32_t)y + (uint32_t)d.ny * (uint32_t)z); }

// 3x3x3 convolution, stride 1, zero padding, over the channel concatenation [in1 (C1), in2 (C2)].
// One wave computes MT rows (y .. y+MT-1) of 16 voxels in x and all COUT channels; a block holds
// 4 waves. C1 < 16 only without in2 (the 4-channel input layer: channels >= C1 load as 0).
template <int C1, int C2, int COUT, int MT, bool RELU>
__global__ __launch_bounds__(256) void conv3_kernel(Dims d, const uint16_t* __restrict__ in1, const uint16_t* __restrict__ in2,
                                                    const bf16x4_t* __restrict__ wpk, const float* __restrict__ bias,
                                                    uint16_t* __restrict__ out) {
    constexpr int NCH1 = (C1 + 15) / 16, NCH = NCH1 + C2 / 16, NB = COUT / 16;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int tiles_x = d.nx / 16, rows = d.ny / MT;
    const uint32_t total = (uint32_t)tiles_x * rows * d.nz;
    if (wave >= total) return;
    const int tx = (int)(wave % (uint32_t)tiles_x);
    const int ry = (int)((wave / (uint32_t)tiles_x) % (uint32_t)rows);
    const int z = (int)(wave / ((uint32_t)tiles_x * rows));
    const int x0 = tx * 16, y0 = ry * MT;
    const int m = lane & 15, kq = lane >> 4;
    floatx4_t acc[MT][NB];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[t][nb] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tap = 0; tap < 27; ++tap) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const int xx = x0 + m + dx, zz = z + dz;
        const bool okxz = xx >= 0 && xx < d.nx && zz >= 0 && zz < d.nz;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            bf16x4_t b[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) b[nb] = wpk[(((size_t)tap * NCH + ch) * NB + nb) * 64 + lane];
            const bool first = ch < NCH1;
            const int cstride = first ? (C1 < 4 ? 4 : C1) : C2;
            const int cbase = (first ? ch * 16 : (ch - NCH1) * 16) + 4 * kq;
            const bool okc = first ? (cbase < (C1 < 4 ? 4 : C1)) : true;
            const uint16_t* src = first ? in1 : in2;
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                const int yy = y0 + t + dy;
                bf16x4_t a = bf16x4_t{0, 0, 0, 0};
                if (okxz && okc && yy >= 0 && yy < d.ny)
                    a = *reinterpret_cast<const bf16x4_t*>(src + (size_t)vox(d, xx, yy, zz) * cstride + cbase);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[t][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nb], acc[t][nb], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int co = nb * 16 + m;
            const float bv = bias[co];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[t][nb][i] + bv;
                if (RELU) v = fmaxf(v, 0.0f);
                out[(size_t)vox(d, x0 + 4 * kq + i, y0 + t, z) * COUT + co] = to_bf16(v);
            }
        }
}

// 2x2x2 stride-2 transposed convolution from the coarse grid dc to the fine grid (2 dc): fine voxel
// (2x+kx, 2y+ky, 2z+kz) = W[kz,ky,kx]^T in(x, y, z) + bias. One wave: 16 coarse voxels in x.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void tconv2_kernel(Dims dc, const uint16_t* __restrict__ in, const bf16x4_t* __restrict__ wpk,
                                                     const float* __restrict__ bias, uint16_t* __restrict__ out) {
    constexpr int NCH = CIN / 16, NB = COUT / 16;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int tiles_x = dc.nx / 16;
    const uint32_t total = (uint32_t)tiles_x * dc.ny * dc.nz;
    if (wave >= total) return;
    const int tx = (int)(wave % (uint32_t)tiles_x);
    const int y = (int)((wave / (uint32_t)tiles_x) % (uint32_t)dc.ny);
    const int z = (int)(wave / ((uint32_t)tiles_x * dc.ny));
    const int x0 = tx * 16, m = lane & 15, kq = lane >> 4;
    bf16x4_t a[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) a[ch] = *reinterpret_cast<const bf16x4_t*>(in + (size_t)vox(dc, x0 + m, y, z) * CIN + ch * 16 + 4 * kq);
    const Dims df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    for (int tap = 0; tap < 8; ++tap) {
        const int kz = tap >> 2, ky = (tap >> 1) & 1, kx = tap & 1;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            floatx4_t acc = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ch], wpk[(((size_t)tap * NCH + ch) * NB + nb) * 64 + lane], acc, 0, 0, 0);
            const int co = nb * 16 + m;
            const float bv = bias[co];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xc = x0 + 4 * kq + i;
                out[(size_t)vox(df, 2 * xc + kx, 2 * y + ky, 2 * z + kz) * COUT + co] = to_bf16(acc[i] + bv);
            }
        }
    }
}

// 2x2x2 max pooling; one thread per (coarse voxel, 4-channel group).
template <int C>
__global__ __launch_bounds__(256) void maxpool2_kernel(Dims dc, const uint16_t* __restrict__ in, uint16_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t groups = C / 4;
    const uint32_t n = (uint32_t)dc.nx * dc.ny * dc.nz;
    if (i >= n * groups) return;
    const uint32_t v = i / groups, g = i - v * groups;
    const int x = (int)(v % (uint32_t)dc.nx), y = (int)((v / (uint32_t)dc.nx) % (uint32_t)dc.ny), z = (int)(v / ((uint32_t)dc.nx * dc.ny));
    const Dims df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    float mx[4]{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; ++c) {
        const uint16_t* p = in + (size_t)vox(df, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), 2 * z + (c >> 2)) * C + 4 * g;
        for (int j = 0; j < 4; ++j) mx[j] = fmaxf(mx[j], from_bf16(p[j]));
    }
    for (int j = 0; j < 4; ++j) out[(size_t)v * C + 4 * g + j] = to_bf16(mx[j]);
}

// ||b||^2 partials of the solve's set-up (partial set 0) -> the input scale s = rms of b over the
// unknowns (1 when b = 0), kept on the device.
__global__ __launch_bounds__(256) void input_scale_kernel(const double* __restrict__ partials, uint32_t G, uint64_t unknowns,
                                                         double* __restrict__ scale) {
    __shared__ double lds[3 * kMaxWaves];
    const double* parts[1]{partials + kPartBB * kPcgMaxBlocks};
    double red[1];
    reduce_partials<1>(parts, G, red, lds);
    if (threadIdx.x == 0) {
        const double s = unknowns ? std::sqrt(red[0] / (double)unknowns) : 0.0;
        scale[0] = s > 0.0 ? s : 1.0;
    }
}

// Network input: (b / s, unknown indicator, 0, 0) per voxel.
__global__ __launch_bounds__(256) void pack_input_kernel(uint32_t n, const float* __restrict__ b, const uint8_t* __restrict__ code,
                                                         const double* __restrict__ scale, uint16_t* __restrict__ x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float inv = (float)(1.0 / scale[0]);
    const bool u = code[i] != 0u;
    x[4 * (size_t)i + 0] = to_bf16(u ? b[i] * inv : 0.0f);
    x[4 * (size_t)i + 1] = to_bf16(u ? 1.0f : 0.0f);
    x[4 * (size_t)i + 2] = 0;
    x[4 * (size_t)i + 3] = 0;
}

// 1x1x1 head (16 -> 1), times s * output_scale; 0 off the unknowns.
__global__ __launch_bounds__(256) void head_kernel(uint32_t n, const uint16_t* __restrict__ h, const float* __restrict__ w,
                                                   const float* __restrict__ bias, const uint8_t* __restrict__ code,
                                                   const double* __restrict__ scale, double output_scale, float* __restrict__ guess) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = bias[0];
#pragma unroll
    for (int c = 0; c < 16; ++c) v += w[c] * from_bf16(h[16 * (size_t)i + c]);
    guess[i] = code[i] != 0u ? (float)((double)v * scale[0] * output_scale) : 0.0f;
}

// A-norm-optimal scale of the guess g: c = <g, b> / <g, A g> minimises ||c g - x*||_A, so the
// scaled guess is never worse than x0 = 0 in the norm CG minimises. Pass 1: per-block partials of
// <g, b> and <g, A g> (A g gathered on the fly, g = 0 off the unknowns); pass 2: every block
// reduces them in the same order and scales its share of g.
__global__ __launch_bounds__(256) void guess_scale_partials_kernel(Geo g, uint32_t per, const uint8_t* __restrict__ code,
                                                                   const float* __restrict__ x, const float* __restrict__ b,
                                                                   double* __restrict__ partials) {
    __shared__ double lds[3 * kMaxWaves];
    double gb = 0.0, gag = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += gridDim.x * blockDim.x) {
        const uint32_t d = code[i];
        if (d == 0u) continue;
        int xx, yy, zz;
        decode_cell(g, i, xx, yy, zz);
        const double xi = x[i];
        double s = 0.0;
        const int c[3]{xx, yy, zz}, n[3]{g.nx, g.ny, g.nz};
        for (int a = 0; a < 3; ++a)
            for (int dir = -1; dir <= 1; dir += 2) {
                int q[3]{c[0], c[1], c[2]};
                q[a] += dir;
                if (q[a] < 0 || q[a] >= n[a]) {
                    if (!(per & (1u << a))) continue;
                    q[a] = q[a] < 0 ? n[a] - 1 : 0;
                }
                s += (double)x[cidx(g, q[0], q[1], q[2])];
            }
        gb += xi * (double)b[i];
        gag += xi * ((double)d * xi - s);
    }
    double v[2]{gb, gag};
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = v[0];
        partials[gridDim.x + blockIdx.x] = v[1];
    }
}

__global__ __launch_bounds__(256) void guess_scale_apply_kernel(uint32_t n, const double* __restrict__ partials, uint32_t G,
                                                                float* __restrict__ x) {
    __shared__ double lds[3 * kMaxWaves];
    const double* parts[2]{partials, partials + G};
    double red[2];
    reduce_partials<2>(parts, G, red, lds);
    const double c = red[1] > 0.0 && std::isfinite(red[0] / red[1]) ? red[0] / red[1] : 0.0;
    const float cf = (float)c;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] *= cf;
}

}  // namespace

// Parameters in PyTorch state_dict order (weight then bias per layer; Conv3d weights [co][ci][kz][ky][kx],
// ConvTranspose3d weights [ci][co][kz][ky][kx]).
struct UNetLayer {
    int kind;   // 0 = 3x3x3 conv, 1 = 2x2x2 transposed conv, 2 = 1x1x1 head
    int cin, cout;
};
constexpr UNetLayer kUNetLayers[] = {
    {0, 2, 16},  {0, 16, 16}, {0, 16, 32}, {0, 32, 32}, {0, 32, 64}, {0, 64, 64}, {1, 64, 32},
    {0, 64, 32}, {0, 32, 32}, {1, 32, 16}, {0, 32, 16}, {0, 16, 16}, {2, 16, 1},
};
constexpr int kUNetLayerCount = sizeof(kUNetLayers) / sizeof(kUNetLayers[0]);

uint64_t unet_parameter_count() {
    uint64_t n = 0;
    for (const UNetLayer& L : kUNetLayers) n += (uint64_t)L.cin * L.cout * (L.kind == 0 ? 27 : L.kind == 1 ? 8 : 1) + L.cout;
    return n;
}

struct UNet {
    std::vector<void*> allocations;
    bf16x4_t* wpk[kUNetLayerCount]{};
    float* bias[kUNetLayerCount]{};
    float* head_w = nullptr;
    double output_scale = 1.0;
    // activations (fine: N voxels, mid: N/8, coarse: N/64)
    uint16_t *in4 = nullptr, *f1 = nullptr, *f2 = nullptr, *skip1 = nullptr;
    uint16_t *m1 = nullptr, *m2 = nullptr, *skip2 = nullptr;
    uint16_t *c1 = nullptr, *c2 = nullptr;
    double* scale = nullptr;
    float* guess = nullptr;
    double* scale_partials = nullptr;   // 2 * kPcgBlocks
};

void unet_release(UNet* u) {
    if (!u) return;
    for (void* p : u->allocations) (void)hipFree(p);
    delete u;
}

namespace {

template <typename T>
T* unet_alloc(UNet* u, size_t count) {
    void* p = nullptr;
    DS_HIP(hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 1));
    u->allocations.push_back(p);
    return static_cast<T*>(p);
}

// Conv weights -> fragment order: pk[((tap * NCH + ch) * NB + nb) * 64 + l][j] =
// W[co = 16 nb + l%16][ci = 16 ch + 4 (l/16) + j][tap] (0 beyond cin); transposed convolutions
// use their [ci][co][tap] weights the same way with 8 taps.
std::vector<uint16_t> pack_weights(const float* w, int cin, int cout, int taps, bool transposed) {
    const int nch = (cin + 15) / 16, nb = cout / 16;
    std::vector<uint16_t> pk((size_t)taps * nch * nb * 64 * 4, 0);
    for (int tap = 0; tap < taps; ++tap)
        for (int ch = 0; ch < nch; ++ch)
            for (int b = 0; b < nb; ++b)
                for (int l = 0; l < 64; ++l)
                    for (int j = 0; j < 4; ++j) {
                        const int co = 16 * b + (l & 15), ci = 16 * ch + 4 * (l >> 4) + j;
                        if (ci >= cin) continue;
                        const float v = transposed ? w[((size_t)ci * cout + co) * taps + tap] : w[((size_t)co * cin + ci) * taps + tap];
                        pk[((((size_t)tap * nch + ch) * nb + b) * 64 + l) * 4 + j] = to_bf16(v);
                    }
    return pk;
}

uint32_t wave_blocks(uint64_t waves) { return (uint32_t)((waves + 3) / 4); }

}  // namespace

void unet_load(ds_context* c, const float* params, uint64_t count, double output_scale) {
    require(params != nullptr, "warm-start parameters are null");
    if (count != unet_parameter_count())
        throw Failure{DS_ERROR_INVALID_ARGUMENT, "warm-start parameter count does not match the network (" +
                                                     std::to_string(unet_parameter_count()) + " expected)"};
    const Geo& g = c->geo;
    if (g.nx % 64 || g.ny % 16 || g.nz % 4)
        throw Failure{DS_ERROR_UNSUPPORTED, "the warm-start network needs nx % 64 == 0, ny % 16 == 0 and nz % 4 == 0"};
    UNet* u = c->unet ? c->unet : new UNet();
    c->unet = u;
    const float* p = params;
    for (int k = 0; k < kUNetLayerCount; ++k) {
        const UNetLayer& L = kUNetLayers[k];
        const int taps = L.kind == 0 ? 27 : L.kind == 1 ? 8 : 1;
        const size_t nw = (size_t)L.cin * L.cout * taps;
        if (!u->bias[k]) u->bias[k] = unet_alloc<float>(u, L.cout);
        if (L.kind == 2) {
            if (!u->head_w) u->head_w = unet_alloc<float>(u, L.cin);
            DS_HIP(hipMemcpyAsync(u->head_w, p, nw * sizeof(float), hipMemcpyHostToDevice, c->stream));
        } else {
            const std::vector<uint16_t> pk = pack_weights(p, L.cin, L.cout, taps, L.kind == 1);
            if (!u->wpk[k]) u->wpk[k] = unet_alloc<bf16x4_t>(u, pk.size() / 4);
            DS_HIP(hipMemcpyAsync(u->wpk[k], pk.data(), pk.size() * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream));
        }
        DS_HIP(hipMemcpyAsync(u->bias[k], p + nw, L.cout * sizeof(float), hipMemcpyHostToDevice, c->stream));
        p += nw + L.cout;
    }
    u->output_scale = output_scale;
    if (!u->in4) {
        const size_t n = g.n, nm = n / 8, nc = n / 64;
        u->in4 = unet_alloc<uint16_t>(u, 4 * n);
        u->f1 = unet_alloc<uint16_t>(u, 16 * n);
        u->f2 = unet_alloc<uint16_t>(u, 16 * n);
        u->skip1 = unet_alloc<uint16_t>(u, 16 * n);
        u->m1 = unet_alloc<uint16_t>(u, 32 * nm);
        u->m2 = unet_alloc<uint16_t>(u, 32 * nm);
        u->skip2 = unet_alloc<uint16_t>(u, 32 * nm);
        u->c1 = unet_alloc<uint16_t>(u, 64 * nc);
        u->c2 = unet_alloc<uint16_t>(u, 64 * nc);
        u->scale = unet_alloc<double>(u, 1);
        u->guess = unet_alloc<float>(u, n);
        u->scale_partials = unet_alloc<double>(u, 2 * (size_t)kPcgBlocks);
    }
    DS_HIP(hipStreamSynchronize(c->stream));   // the host staging vectors go out of scope
}

// Runs the network on b (the solve's set-up kernel has left the ||b||^2 partials in partial set
// 0), scales its output to the A-norm-optimal multiple, and returns the device guess buffer.
const float* unet_infer(ds_context* c, const float* b) {
    UNet* u = c->unet;
    require(u != nullptr && u->in4 != nullptr, "no warm-start network loaded (ds_warm_start_load)");
    const Geo& g = c->geo;
    const Dims d1{g.nx, g.ny, g.nz}, d2{g.nx / 2, g.ny / 2, g.nz / 2}, d3{g.nx / 4, g.ny / 4, g.nz / 4};
    const uint64_t n1 = g.n, n2 = n1 / 8, n3 = n1 / 64;
    hipStream_t s = c->stream;
    input_scale_kernel<<<1, kBlock, 0, s>>>(c->pcg.partials, c->pcg.blocks, c->unknowns, u->scale);
    pack_input_kernel<<<blocks_for(n1), kBlock, 0, s>>>((uint32_t)n1, b, c->code, u->scale, u->in4);
    constexpr int MT = 4;
    auto conv_waves = [&](const Dims& d) { return wave_blocks((uint64_t)(d.nx / 16) * (d.ny / MT) * d.nz); };
    // encoder
    conv3_kernel<2, 0, 16, MT, true><<<conv_waves(d1), 256, 0, s>>>(d1, u->in4, nullptr, u->wpk[0], u->bias[0], u->f1);
    conv3_kernel<16, 0, 16, MT, true><<<conv_waves(d1), 256, 0, s>>>(d1, u->f1, nullptr, u->wpk[1], u->bias[1], u->skip1);
    maxpool2_kernel<16><<<blocks_for(n2 * 4), kBlock, 0, s>>>(d2, u->skip1, u->m1);
    conv3_kernel<16, 0, 32, MT, true><<<conv_waves(d2), 256, 0, s>>>(d2, u->m1, nullptr, u->wpk[2], u->bias[2], u->m2);
    conv3_kernel<32, 0, 32, MT, true><<<conv_waves(d2), 256, 0, s>>>(d2, u->m2, nullptr, u->wpk[3], u->bias[3], u->skip2);
    maxpool2_kernel<32><<<blocks_for(n3 * 8), kBlock, 0, s>>>(d3, u->skip2, u->c1);
    conv3_kernel<32, 0, 64, MT, true><<<conv_waves(d3), 256, 0, s>>>(d3, u->c1, nullptr, u->wpk[4], u->bias[4], u->c2);
    conv3_kernel<64, 0, 64, MT, true><<<conv_waves(d3), 256, 0, s>>>(d3, u->c2, nullptr, u->wpk[5], u->bias[5], u->c1);
    // decoder
    tconv2_kernel<64, 32><<<wave_blocks((uint64_t)(d3.nx / 16) * d3.ny * d3.nz), 256, 0, s>>>(d3, u->c1, u->wpk[6], u->bias[6], u->m1);
    conv3_kernel<32, 32, 32, MT, true><<<conv_waves(d2), 256, 0, s>>>(d2, u->m1, u->skip2, u->wpk[7], u->bias[7], u->m2);
    conv3_kernel<32, 0, 32, MT, true><<<conv_waves(d2), 256, 0, s>>>(d2, u->m2, nullptr, u->wpk[8], u->bias[8], u->m1);
    tconv2_kernel<32, 16><<<wave_blocks((uint64_t)(d2.nx / 16) * d2.ny * d2.nz), 256, 0, s>>>(d2, u->m1, u->wpk[9], u->bias[9], u->f1);
    conv3_kernel<16, 16, 16, MT, true><<<conv_waves(d1), 256, 0, s>>>(d1, u->f1, u->skip1, u->wpk[10], u->bias[10], u->f2);
    conv3_kernel<16, 0, 16, MT, true><<<conv_waves(d1), 256, 0, s>>>(d1, u->f2, nullptr, u->wpk[11], u->bias[11], u->f1);
    head_kernel<<<blocks_for(n1), kBlock, 0, s>>>((uint32_t)n1, u->f1, u->head_w, u->bias[12], c->code, u->scale, u->output_scale,
                                                  u->guess);
    guess_scale_partials_kernel<<<kPcgBlocks, kBlock, 0, s>>>(g, pcg_periodic_bits(c), c->code, u->guess, b, u->scale_partials);
    guess_scale_apply_kernel<<<kPcgBlocks, kBlock, 0, s>>>((uint32_t)n1, u->scale_partials, kPcgBlocks, u->guess);
    check_launch("warm-start network");
    return u->guess;
}

}  // namespace ds
