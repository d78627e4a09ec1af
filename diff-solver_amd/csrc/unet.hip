// unet.hip — the warm-start network of the pressure solves at inference: a 3-level 3-D U-Net
// (16/32/64 channels, 3x3x3 convolutions with ReLU, 2x2x2 max pooling, 2x2x2 stride-2 transposed
// convolutions, a 1x1x1 head) mapping a normalised residual and the unknown mask to a direction d
// in which the solve's error lies. The kernels are those of unet_kernels.hpp (BF16 activations,
// FP32 accumulation on the matrix cores); the network is trained by unet_train.hip.
//
// The guess for A x = b from a base guess x_b (or x_b = 0): r = b - A x_b, d = N(r / rms(r)), and
// x0 = x_b + gamma d with gamma = <d, r> / <d, A d>, the step along d that minimises the A-norm
// error (the norm CG minimises), so x0 is never further from x* in that norm than x_b. The network
// is trained for exactly this use: its loss is the error of x_b + gamma d (unet_train.hip).
//
// With `levels` L > 0 the network runs on the grid coarsened 2^L times per axis: r is restricted
// (the sum over each coarse cell's children; a coarse cell is an unknown when any child is), the
// network's coarse output is prolonged back by cell-centred trilinear interpolation (clamped at the
// walls) and masked to the fine unknowns, and gamma is formed on the fine grid. A network trained
// on 64^3 problems then sees 128^3 (L = 1) or 256^3 (L = 2) problems at the size it was trained
// on, its receptive field covers the same share of the domain, and it costs 8^L times less.
#include <cmath>
#include <vector>

#include "context.hpp"
#include "unet_kernels.hpp"

namespace ds {

using namespace unet;

namespace {

// r = b - A x_b on the unknowns (0 elsewhere).
__global__ __launch_bounds__(256) void base_residual_kernel(Geo geo, uint32_t per, const uint8_t* __restrict__ code,
                                                           const float* __restrict__ b, const float* __restrict__ xb,
                                                           float* __restrict__ r) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < geo.n; i += gridDim.x * blockDim.x) {
        // every code and value is loaded up front (the neighbours are cells of the grid), the codes
        // select afterwards, so the loads issue together
        const uint32_t d = code[i];
        int x, y, z;
        decode_cell(geo, i, x, y, z);
        const int c[3]{x, y, z}, ext[3]{geo.nx, geo.ny, geo.nz};
        double s = (double)d * (double)xb[i];
        const float bi = b[i];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int dir = -1; dir <= 1; dir += 2) {
                int q[3]{c[0], c[1], c[2]};
                q[a] += dir;
                if (q[a] < 0 || q[a] >= ext[a]) {
                    if (!(per & (1u << a))) continue;
                    q[a] = q[a] < 0 ? ext[a] - 1 : 0;
                }
                const uint32_t j = cidx(geo, q[0], q[1], q[2]);
                const double xj = (double)xb[j];
                s -= code[j] ? xj : 0.0;
            }
        r[i] = d ? (float)((double)bi - s) : 0.0f;
    }
}

// rc[I] = sum of r over the f^3 children of coarse cell I (f = 2^levels).
__global__ __launch_bounds__(256) void restrict_kernel(Geo g, int f, Dims dc, const float* __restrict__ r, float* __restrict__ rc) {
    const uint32_t nc = (uint32_t)dc.nx * dc.ny * dc.nz;
    for (uint32_t I = blockIdx.x * blockDim.x + threadIdx.x; I < nc; I += gridDim.x * blockDim.x) {
        const int X = (int)(I % (uint32_t)dc.nx), Y = (int)((I / (uint32_t)dc.nx) % (uint32_t)dc.ny), Z = (int)(I / ((uint32_t)dc.nx * dc.ny));
        float s = 0.0f;
        for (int k = 0; k < f; ++k)
            for (int j = 0; j < f; ++j)
                for (int i = 0; i < f; ++i) s += r[cidx(g, f * X + i, f * Y + j, f * Z + k)];
        rc[I] = s;
    }
}

// ccode[I] = 1 when any child of coarse cell I is an unknown (the network's unknown indicator), and
// per-block counts of such cells: counts[block].
__global__ __launch_bounds__(256) void coarse_code_kernel(Geo g, int f, Dims dc, const uint8_t* __restrict__ code, uint8_t* __restrict__ ccode,
                                                          double* __restrict__ counts) {
    __shared__ double lds[kMaxWaves];
    const uint32_t nc = (uint32_t)dc.nx * dc.ny * dc.nz;
    double v[1]{0.0};
    for (uint32_t I = blockIdx.x * blockDim.x + threadIdx.x; I < nc; I += gridDim.x * blockDim.x) {
        const int X = (int)(I % (uint32_t)dc.nx), Y = (int)((I / (uint32_t)dc.nx) % (uint32_t)dc.ny), Z = (int)(I / ((uint32_t)dc.nx * dc.ny));
        uint8_t any = 0;
        for (int k = 0; k < f; ++k)
            for (int j = 0; j < f; ++j)
                for (int i = 0; i < f; ++i) any |= code[cidx(g, f * X + i, f * Y + j, f * Z + k)] != 0u ? 1u : 0u;
        ccode[I] = any;
        v[0] += any;
    }
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) counts[blockIdx.x] = v[0];
}

// d[fine] = cell-centred trilinear interpolation of dc (coarse position (x + 1/2) / f - 1/2 per
// axis, clamped to the coarse grid), 0 off the fine unknowns.
__device__ __forceinline__ void lin_weights(int x, int f, int n, int& i0, int& i1, float& t) {
    const float X = ((float)x + 0.5f) / (float)f - 0.5f;
    const float fl = floorf(X);
    t = X - fl;
    const int a = (int)fl;
    i0 = min(max(a, 0), n - 1);
    i1 = min(max(a + 1, 0), n - 1);
}
__global__ __launch_bounds__(256) void prolong_kernel(Geo g, int f, Dims dc, const uint8_t* __restrict__ code, const float* __restrict__ dcv,
                                                      float* __restrict__ d) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += gridDim.x * blockDim.x) {
        if (!code[i]) {
            d[i] = 0.0f;
            continue;
        }
        int x, y, z;
        decode_cell(g, i, x, y, z);
        int x0, x1, y0, y1, z0, z1;
        float tx, ty, tz;
        lin_weights(x, f, dc.nx, x0, x1, tx);
        lin_weights(y, f, dc.ny, y0, y1, ty);
        lin_weights(z, f, dc.nz, z0, z1, tz);
        auto at = [&](int X, int Y, int Z) { return dcv[(uint32_t)X + (uint32_t)dc.nx * ((uint32_t)Y + (uint32_t)dc.ny * (uint32_t)Z)]; };
        const float c00 = at(x0, y0, z0) + tx * (at(x1, y0, z0) - at(x0, y0, z0));
        const float c10 = at(x0, y1, z0) + tx * (at(x1, y1, z0) - at(x0, y1, z0));
        const float c01 = at(x0, y0, z1) + tx * (at(x1, y0, z1) - at(x0, y0, z1));
        const float c11 = at(x0, y1, z1) + tx * (at(x1, y1, z1) - at(x0, y1, z1));
        const float c0 = c00 + ty * (c10 - c00), c1 = c01 + ty * (c11 - c01);
        d[i] = c0 + tz * (c1 - c0);
    }
}

// 1x1x1 head (16 -> 1); 0 off the unknowns.
__global__ __launch_bounds__(256) void head_kernel(uint32_t n, const uint16_t* __restrict__ h, const float* __restrict__ w,
                                                   const float* __restrict__ bias, const uint8_t* __restrict__ code,
                                                   float* __restrict__ d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = bias[0];
#pragma unroll
    for (int c = 0; c < 16; ++c) v += w[c] * from_bf16(h[16 * (size_t)i + c]);
    d[i] = code[i] != 0u ? v : 0.0f;
}

// Per-block partials of the step along d: <d, r> and <d, A d> (A d gathered on the fly; d = 0 off
// the unknowns): partials[k * G + block], k = 0, 1. A thread takes kStepCells cells a block apart
// (coalesced), their codes loaded first, so a wave keeps several cells' loads in flight.
constexpr int kStepCells = 4;
__global__ __launch_bounds__(256) void step_partials_kernel(Geo geo, uint32_t per, const uint8_t* __restrict__ code,
                                                           const float* __restrict__ r, const float* __restrict__ d,
                                                           double* __restrict__ partials) {
    __shared__ double lds[2 * kMaxWaves];
    double v[2]{0.0, 0.0};
    const uint32_t i0 = blockIdx.x * (blockDim.x * kStepCells) + threadIdx.x;
    uint32_t cd[kStepCells];
#pragma unroll
    for (int k = 0; k < kStepCells; ++k) {
        const uint32_t i = i0 + k * blockDim.x;
        cd[k] = i < geo.n ? code[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kStepCells; ++k) {
        // the loads of a cell outside the grid read its last cell; its code 0 drops the terms
        const uint32_t i = min(i0 + k * blockDim.x, geo.n - 1u);
        int x, y, z;
        decode_cell(geo, i, x, y, z);
        const int c[3]{x, y, z}, ext[3]{geo.nx, geo.ny, geo.nz};
        const double di = d[i];
        const float ri = r[i];
        double s = (double)cd[k] * di;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int dir = -1; dir <= 1; dir += 2) {
                int q[3]{c[0], c[1], c[2]};
                q[a] += dir;
                if (q[a] < 0 || q[a] >= ext[a]) {
                    if (!(per & (1u << a))) continue;
                    q[a] = q[a] < 0 ? ext[a] - 1 : 0;
                }
                s -= (double)d[cidx(geo, q[0], q[1], q[2])];
            }
        if (cd[k]) {
            v[0] += di * (double)ri;
            v[1] += di * s;
        }
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = v[0];
        partials[gridDim.x + blockIdx.x] = v[1];
    }
}

// x0 = x_b + gamma d (x_b = null: 0), gamma = <d, r> / <d, A d> (every block reduces the partials in
// the same order; gamma = 0 when not finite).
__global__ __launch_bounds__(256) void step_apply_kernel(uint32_t n, const double* __restrict__ partials, uint32_t G,
                                                        const float* __restrict__ xb, const float* __restrict__ d,
                                                        float* __restrict__ x0) {
    __shared__ double lds[2 * kMaxWaves];
    const double* parts[2]{partials, partials + G};
    double t[2];
    reduce_partials<2>(parts, G, t, lds);
    double gamma = t[1] > 0.0 ? t[0] / t[1] : 0.0;
    if (!isfinite(gamma)) gamma = 0.0;
    const float gf = (float)gamma;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x0[i] = (xb ? xb[i] : 0.0f) + gf * d[i];
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
    std::vector<void*> allocations;     // parameters
    std::vector<void*> activations;     // sized for the network's grid (released when levels change)
    int levels = 0;                     // 2^levels coarsening of the solve's grid before the network
    Dims dn{};                          // the network's grid
    uint8_t* ccode = nullptr;           // the network grid's unknown indicator (levels > 0)
    uint64_t net_unknowns = 0;          // its unknown count
    float* rc = nullptr;                // restricted input (levels > 0)
    float* dc = nullptr;                // network output on the network grid (levels > 0)
    double* count_partials = nullptr;   // kPcgBlocks + 1
    bf16x4_t* wpk[kUNetLayerCount]{};
    uint4* wpk32[kUNetLayerCount]{};     // 3x3x3 convolutions: the 16x16x32 packing (pack_weights32)
    float* bias[kUNetLayerCount]{};
    float* head_w = nullptr;
    // activations (fine: N voxels, mid: N/8, coarse: N/64)
    uint16_t *in4 = nullptr, *f1 = nullptr, *f2 = nullptr, *skip1 = nullptr;
    uint16_t *m1 = nullptr, *m2 = nullptr, *skip2 = nullptr;
    uint16_t *c1 = nullptr, *c2 = nullptr;
    float* resid = nullptr;              // network input r (b - A x_b)
    float* dir = nullptr;                // network output d
    float* guess = nullptr;              // x0
    double* norm_partials = nullptr;     // kPcgBlocks
    double* step_partials = nullptr;     // 2 * step_blocks
    uint32_t step_blocks = 0;            // blocks of step_partials_kernel
};

void unet_release(UNet* u) {
    if (!u) return;
    for (void* p : u->allocations) (void)hipFree(p);
    for (void* p : u->activations) (void)hipFree(p);
    delete u;
}

namespace {

template <typename T>
T* unet_alloc(UNet* u, size_t count, bool activation = false) {
    void* p = nullptr;
    DS_HIP(hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 1));
    (activation ? u->activations : u->allocations).push_back(p);
    return static_cast<T*>(p);
}

__global__ void count_finish_kernel(const double* __restrict__ partials, uint32_t G, double* __restrict__ out) {
    __shared__ double lds[kMaxWaves];
    const double* parts[1]{partials};
    double red[1];
    reduce_partials<1>(parts, G, red, lds);
    if (threadIdx.x == 0) out[0] = red[0];
}

// Activation buffers and the coarse unknown indicator for the current levels.
void unet_prepare(ds_context* c, UNet* u) {
    if (u->in4) return;
    const Geo& g = c->geo;
    const int L = u->levels, f = 1 << L;
    u->dn = Dims{g.nx >> L, g.ny >> L, g.nz >> L};
    const size_t n = (size_t)u->dn.nx * u->dn.ny * u->dn.nz, nm = n / 8, nc = n / 64, nf = g.n;
    u->in4 = unet_alloc<uint16_t>(u, 4 * n, true);
    u->f1 = unet_alloc<uint16_t>(u, 16 * n, true);
    u->f2 = unet_alloc<uint16_t>(u, 16 * n, true);
    u->skip1 = unet_alloc<uint16_t>(u, 16 * n, true);
    u->m1 = unet_alloc<uint16_t>(u, 32 * nm, true);
    u->m2 = unet_alloc<uint16_t>(u, 32 * nm, true);
    u->skip2 = unet_alloc<uint16_t>(u, 32 * nm, true);
    u->c1 = unet_alloc<uint16_t>(u, 64 * nc, true);
    u->c2 = unet_alloc<uint16_t>(u, 64 * nc, true);
    u->resid = unet_alloc<float>(u, nf, true);
    u->dir = unet_alloc<float>(u, nf, true);
    u->guess = unet_alloc<float>(u, nf, true);
    u->norm_partials = unet_alloc<double>(u, (size_t)kPcgBlocks, true);
    u->step_blocks = (uint32_t)((nf + (size_t)kBlock * kStepCells - 1) / ((size_t)kBlock * kStepCells));
    u->step_partials = unet_alloc<double>(u, 2 * (size_t)u->step_blocks, true);
    if (L > 0) {
        u->ccode = unet_alloc<uint8_t>(u, n, true);
        u->rc = unet_alloc<float>(u, n, true);
        u->dc = unet_alloc<float>(u, n, true);
        u->count_partials = unet_alloc<double>(u, (size_t)kPcgBlocks + 1, true);
        coarse_code_kernel<<<kPcgBlocks, kBlock, 0, c->stream>>>(g, f, u->dn, c->code, u->ccode, u->count_partials);
        count_finish_kernel<<<1, kBlock, 0, c->stream>>>(u->count_partials, kPcgBlocks, u->count_partials + kPcgBlocks);
        check_launch("coarse_code_kernel");
        double cnt = 0.0;
        DS_HIP(hipMemcpyAsync(&cnt, u->count_partials + kPcgBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        DS_HIP(hipStreamSynchronize(c->stream));
        u->net_unknowns = (uint64_t)cnt;
    } else {
        u->net_unknowns = c->unknowns;
    }
}

// Conv weights -> fragment order: pk[((tap * NCH + ch) * NB + nb) * 64 + l][j] =
// W[co = 16 nb + l%16][ci = 16 ch + 4 (l/16) + j][tap] (0 beyond cin); transposed convolutions
// use their [ci][co][tap] weights the same way with 8 taps (pack_weights_kernel, kinds 0 and 1).
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

// 3x3x3 conv weights -> the order of conv3_tile_kernel's steps: pk[((s * NB + nb) * 64 + l) * 8 + j] =
// W[co = 16 nb + l%16][ci][tap] with kk = 32 s + 8 (l/16) + j = tap * cs + ci (cs: the input
// channels as stored, cin padded to 4 for the input layer), 0 for ci >= cin and tap >= 27.
std::vector<uint16_t> pack_weights32(const float* w, int cin, int cs, int cout) {
    const int steps = (27 * cs + 31) / 32, nb = cout / 16;
    std::vector<uint16_t> pk((size_t)steps * nb * 64 * 8, 0);
    for (int st = 0; st < steps; ++st)
        for (int b = 0; b < nb; ++b)
            for (int l = 0; l < 64; ++l)
                for (int j = 0; j < 8; ++j) {
                    const int kk = 32 * st + 8 * (l >> 4) + j, tap = kk / cs, ci = kk % cs, co = 16 * b + (l & 15);
                    if (tap >= 27 || ci >= cin) continue;
                    pk[(((size_t)st * nb + b) * 64 + l) * 8 + j] = to_bf16(w[((size_t)co * cin + ci) * 27 + tap]);
                }
    return pk;
}

uint32_t wave_blocks(uint64_t waves) { return (uint32_t)((waves + 3) / 4); }

}  // namespace

void unet_configure(ds_context* c, uint32_t levels) {
    require(levels <= 2, "warm-start levels must be 0, 1 or 2");
    const Geo& g = c->geo;
    const int f = 1 << levels;
    if (g.nx % (64 * f) || g.ny % (16 * f) || g.nz % (4 * f))
        throw Failure{DS_ERROR_UNSUPPORTED, "the warm-start network needs nx % 64 == 0, ny % 16 == 0 and nz % 4 == 0 on its grid "
                                            "(the solve's grid coarsened 2^levels times)"};
    UNet* u = c->unet ? c->unet : new UNet();
    c->unet = u;
    if (u->levels == (int)levels) return;
    DS_HIP(hipStreamSynchronize(c->stream));
    for (void* p : u->activations) (void)hipFree(p);
    u->activations.clear();
    u->in4 = nullptr;
    u->ccode = nullptr;
    u->rc = u->dc = nullptr;
    u->levels = (int)levels;
    if (u->head_w) unet_prepare(c, u);   // a loaded network: its buffers now, not inside a solve
}

void unet_load(ds_context* c, const float* params, uint64_t count, double output_scale) {
    (void)output_scale;   // the guess is applied at its A-norm-best multiple, so a scale cancels
    require(params != nullptr, "warm-start parameters are null");
    if (count != unet_parameter_count())
        throw Failure{DS_ERROR_INVALID_ARGUMENT, "warm-start parameter count does not match the network (" +
                                                     std::to_string(unet_parameter_count()) + " expected)"};
    {
        const Geo& g = c->geo;
        const int f = 1 << (c->unet ? c->unet->levels : 0);
        if (g.nx % (64 * f) || g.ny % (16 * f) || g.nz % (4 * f))
            throw Failure{DS_ERROR_UNSUPPORTED, "the warm-start network needs nx % 64 == 0, ny % 16 == 0 and nz % 4 == 0 on its grid"};
    }
    UNet* u = c->unet ? c->unet : new UNet();
    c->unet = u;
    const float* p = params;
    std::vector<std::vector<uint16_t>> staging;   // packed weights, alive until the copies are done
    staging.reserve(2 * kUNetLayerCount);
    for (int k = 0; k < kUNetLayerCount; ++k) {
        const UNetLayer& L = kUNetLayers[k];
        const int taps = L.kind == 0 ? 27 : L.kind == 1 ? 8 : 1;
        const size_t nw = (size_t)L.cin * L.cout * taps;
        if (!u->bias[k]) u->bias[k] = unet_alloc<float>(u, L.cout);
        if (L.kind == 2) {
            if (!u->head_w) u->head_w = unet_alloc<float>(u, L.cin);
            DS_HIP(hipMemcpyAsync(u->head_w, p, nw * sizeof(float), hipMemcpyHostToDevice, c->stream));
        } else {
            const std::vector<uint16_t>& pk = staging.emplace_back(pack_weights(p, L.cin, L.cout, taps, L.kind == 1));
            if (!u->wpk[k]) u->wpk[k] = unet_alloc<bf16x4_t>(u, pk.size() / 4);
            DS_HIP(hipMemcpyAsync(u->wpk[k], pk.data(), pk.size() * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream));
            if (L.kind == 0) {
                const std::vector<uint16_t>& pk32 = staging.emplace_back(pack_weights32(p, L.cin, L.cin < 4 ? 4 : L.cin, L.cout));
                if (!u->wpk32[k]) u->wpk32[k] = unet_alloc<uint4>(u, pk32.size() / 8);
                DS_HIP(hipMemcpyAsync(u->wpk32[k], pk32.data(), pk32.size() * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream));
            }
        }
        DS_HIP(hipMemcpyAsync(u->bias[k], p + nw, L.cout * sizeof(float), hipMemcpyHostToDevice, c->stream));
        p += nw + L.cout;
    }
    unet_prepare(c, u);
    DS_HIP(hipStreamSynchronize(c->stream));   // the host staging vectors go out of scope
}

bool unet_loaded(const ds_context* c) { return c->unet != nullptr && c->unet->head_w != nullptr; }

// The network's guess for A x = b (device b, 0 off the unknowns), corrected from the base guess x_b
// when given: returns the device guess buffer.
const float* unet_infer(ds_context* c, const float* b, const float* xb) {
    UNet* u = c->unet;
    require(unet_loaded(c), "no warm-start network loaded (ds_warm_start_load)");
    const Geo& g = c->geo;
    unet_prepare(c, u);
    const int L = u->levels, f = 1 << L;
    const Dims d1 = u->dn, d2{d1.nx / 2, d1.ny / 2, d1.nz / 2}, d3{d1.nx / 4, d1.ny / 4, d1.nz / 4};
    const uint64_t n1 = (uint64_t)d1.nx * d1.ny * d1.nz, n2 = n1 / 8, n3 = n1 / 64;
    const uint32_t per = pcg_periodic_bits(c);
    hipStream_t s = c->stream;
    const float* r = b;
    if (xb) {
        // one cell per thread: the stencil's dependent loads of a thread are not repeated in a loop
        base_residual_kernel<<<blocks_for(g.n), kBlock, 0, s>>>(g, per, c->code, b, xb, u->resid);
        r = u->resid;
    }
    // the network's input: r itself, or r restricted to the network's grid
    const float* rn = r;
    const uint8_t* coden = c->code;
    if (L > 0) {
        restrict_kernel<<<blocks_for(n1), kBlock, 0, s>>>(g, f, d1, r, u->rc);
        rn = u->rc;
        coden = u->ccode;
    }
    residual_norm_kernel<<<dim3(kPcgBlocks, 1), kBlock, 0, s>>>((uint32_t)n1, rn, coden, u->norm_partials);
    pack_batch_input_kernel<<<dim3(kPcgBlocks, 1), kBlock, 0, s>>>((uint32_t)n1, u->net_unknowns, rn, coden, u->norm_partials,
                                                                   kPcgBlocks, u->in4);
    auto fwd = [&](int k, uint16_t* out) {
        ConvOut o;
        o.out = out;
        o.bias = u->bias[k];
        return o;
    };
    // a 3x3x3 convolution with ReLU: the LDS-tiled kernel with bricks of 16 x 8 x 2 voxels (16 output
    // channels), 16 x 4 x 2 (32) or 16 x 2 x 1 (64) where the grid divides into them; else the
    // direct-load kernel with 4 rows and every channel tile per wave where that gives the grid at
    // least kConvWaves waves, else 1 row per wave, else 1 row and 1 channel tile per wave
    constexpr uint64_t kConvWaves = 2048;
    auto conv = [&]<int C1, int C2, int COUT>(const Dims& d, const uint16_t* in1, const uint16_t* in2, int k, uint16_t* out) {
        constexpr int TY = COUT == 16 ? 8 : COUT == 32 ? 4 : 2, TZ = COUT == 64 ? 1 : 2, NBW = COUT == 64 ? 4 : COUT / 16;
        constexpr int WK = COUT == 64 ? 4 : COUT == 32 ? 2 : 1, RING = COUT == 16 ? 1 : 4;
        if (d.nx % 16 == 0 && d.ny % TY == 0 && d.nz % TZ == 0) {
            const uint64_t bricks = (uint64_t)(d.nx / 16) * (d.ny / TY) * (d.nz / TZ);
            conv3_tile_kernel<C1, C2, COUT, TY, TZ, NBW, WK, RING, kEpiBiasRelu><<<(uint32_t)((bricks + 7) & ~7ull), 256, 0, s>>>(
                d, 1, in1, in2, u->wpk32[k], fwd(k, out));
            return;
        }
        const uint64_t units = (uint64_t)(d.nx / 16) * d.ny * d.nz;   // 16-voxel x rows
        if (d.ny % 4 == 0 && units / 4 >= kConvWaves)
            conv3_kernel<C1, C2, COUT, 4, kEpiBiasRelu><<<wave_blocks(units / 4), 256, 0, s>>>(d, 1, in1, in2, u->wpk[k], fwd(k, out));
        else if (units >= kConvWaves || COUT == 16)
            conv3_kernel<C1, C2, COUT, 1, kEpiBiasRelu><<<wave_blocks(units), 256, 0, s>>>(d, 1, in1, in2, u->wpk[k], fwd(k, out));
        else
            conv3_kernel<C1, C2, COUT, 1, kEpiBiasRelu, 0, 1><<<wave_blocks(units * (COUT / 16)), 256, 0, s>>>(d, 1, in1, in2, u->wpk[k],
                                                                                                          fwd(k, out));
    };
    // encoder
    conv.template operator()<2, 0, 16>(d1, u->in4, nullptr, 0, u->f1);
    conv.template operator()<16, 0, 16>(d1, u->f1, nullptr, 1, u->skip1);
    maxpool2_kernel<16><<<blocks_for(n2 * 4), kBlock, 0, s>>>(d2, (uint32_t)n2, u->skip1, u->m1);
    conv.template operator()<16, 0, 32>(d2, u->m1, nullptr, 2, u->m2);
    conv.template operator()<32, 0, 32>(d2, u->m2, nullptr, 3, u->skip2);
    maxpool2_kernel<32><<<blocks_for(n3 * 8), kBlock, 0, s>>>(d3, (uint32_t)n3, u->skip2, u->c1);
    conv.template operator()<32, 0, 64>(d3, u->c1, nullptr, 4, u->c2);
    conv.template operator()<64, 0, 64>(d3, u->c2, nullptr, 5, u->c1);
    // decoder
    tconv2_kernel<64, 32><<<wave_blocks((uint64_t)(d3.nx / 16) * d3.ny * d3.nz), 256, 0, s>>>(d3, 1, u->c1, u->wpk[6], u->bias[6], u->m1);
    conv.template operator()<32, 32, 32>(d2, u->m1, u->skip2, 7, u->m2);
    conv.template operator()<32, 0, 32>(d2, u->m2, nullptr, 8, u->m1);
    tconv2_kernel<32, 16><<<wave_blocks((uint64_t)(d2.nx / 16) * d2.ny * d2.nz), 256, 0, s>>>(d2, 1, u->m1, u->wpk[9], u->bias[9], u->f1);
    conv.template operator()<16, 16, 16>(d1, u->f1, u->skip1, 10, u->f2);
    conv.template operator()<16, 0, 16>(d1, u->f2, nullptr, 11, u->f1);
    head_kernel<<<blocks_for(n1), kBlock, 0, s>>>((uint32_t)n1, u->f1, u->head_w, u->bias[12], coden, L > 0 ? u->dc : u->dir);
    if (L > 0) prolong_kernel<<<blocks_for(g.n), kBlock, 0, s>>>(g, f, d1, c->code, u->dc, u->dir);
    step_partials_kernel<<<u->step_blocks, kBlock, 0, s>>>(g, per, c->code, r, u->dir, u->step_partials);
    step_apply_kernel<<<kPcgBlocks, kBlock, 0, s>>>(g.n, u->step_partials, u->step_blocks, xb, u->dir, u->guess);
    check_launch("warm-start network");
    return u->guess;
}

}  // namespace ds
