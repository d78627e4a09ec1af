// unet_kernels.hpp — device kernels of the warm-start network (a 3-level 3-D U-Net: 3x3x3
// convolutions, 2x2x2 max pooling, 2x2x2 stride-2 transposed convolutions, a 1x1x1 head), shared
// by inference (unet.hip) and training (unet_train.hip).
//
// Activations and activation gradients are BF16, channels last, `batch` samples stacked along z
// (sample b occupies planes [b nz, (b + 1) nz) of a grid whose x/y extents are the layer's), so a
// voxel's address is x + nx (y + ny zg) with zg the stacked plane; the kernels test the in-sample
// plane for the zero padding. Every product runs on the matrix cores as an implicit GEMM with FP32
// accumulation, v_mfma_f32_16x16x16_bf16: lane l holds A[m = l%16][k = 4(l/16) + j],
// B[k = 4(l/16) + j][n = l%16] (j = 0..3) and D[m = 4(l/16) + i][n = l%16] (i = 0..3).
//
// * forward / input-gradient convolutions: M = 16 voxels consecutive in x, N = 16 output channels,
//   K = 16 input channels of one tap; an A fragment is one 8-byte load of 4 channels of one voxel,
//   a B fragment one 8-byte load of weights repacked into fragment order. The input gradient of a
//   convolution is the convolution of the output gradient with the weights transposed (in <-> out)
//   and the taps mirrored; of a transposed convolution, a stride-2 gather of the 8 fine voxels.
// * weight gradients: M and N are channels and K runs over voxels, 16 per step; a fragment gathers
//   4 voxels of one channel (2-byte loads), each wave leaves its partial sum in its own slot, and a
//   second kernel adds the slots in a fixed order (deterministic, no atomics).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "device.hpp"

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

namespace ds {
namespace unet {
// internal linkage: each translation unit that includes this header gets its own kernels
namespace {

__host__ __device__ __forceinline__ uint16_t to_bf16(float f) {   // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);   // inf / nan keep their class
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
// BF16 value > 0 (sign bit clear, not +0)
__device__ __forceinline__ bool bf16_positive(uint16_t h) { return (h & 0x8000u) == 0u && h != 0u; }

struct Dims {   // per-sample grid of a layer
    int nx, ny, nz;
};

__device__ __forceinline__ uint32_t vox(const Dims& d, int x, int y, int zg) {
    return (uint32_t)x + (uint32_t)d.nx * ((uint32_t)y + (uint32_t)d.ny * (uint32_t)zg);
}

__device__ __forceinline__ floatx4_t mfma(bf16x4_t a, bf16x4_t b, floatx4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// v_mfma_f32_16x16x32_bf16 (gfx950): 8 BF16 per lane and operand, lane l holds A[m = l%16][k = 8(l/16)
// + j] and B[k = 8(l/16) + j][n = l%16] (j = 0..7); D as the 16x16x16 form.
__device__ __forceinline__ floatx4_t mfma32(uint4 a, uint4 b, floatx4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

// Output stage of a 3x3x3 convolution kernel.
enum ConvEpilogue : int {
    kEpiBias = 0,       // out = bf16(acc + bias)
    kEpiBiasRelu = 1,   // out = bf16(max(acc + bias, 0))
    kEpiMask = 2,       // out = mask > 0 ? bf16(acc) : 0   (gradient through the input's ReLU)
    kEpiPlain = 3,      // out = bf16(acc)
    kEpiSplit = 4,      // channels < CU: out = bf16(acc); the rest: out_f32 = acc (FP32)
};

struct ConvOut {
    uint16_t* out = nullptr;          // BF16 (kEpiSplit: the first CU channels)
    const float* bias = nullptr;      // kEpiBias*
    const uint16_t* mask = nullptr;   // kEpiMask: BF16 tensor with the output's channels
    float* out_f32 = nullptr;         // kEpiSplit: channels CU .. COUT-1
};

// Writes one MFMA D fragment of a convolution: acc[i] is output channel co of voxel (x + i, y, zg).
template <int EPI, int COUT, int CU>
__device__ __forceinline__ void conv_store(const ConvOut& o, const Dims& d, const floatx4_t& acc, int x, int y, int zg, int co) {
    const float bv = (EPI == kEpiBias || EPI == kEpiBiasRelu) ? o.bias[co] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const size_t v = vox(d, x + i, y, zg);
        float val = acc[i] + bv;
        if (EPI == kEpiBiasRelu) val = fmaxf(val, 0.0f);
        if (EPI == kEpiMask) val = bf16_positive(o.mask[v * COUT + co]) ? val : 0.0f;
        if (EPI == kEpiSplit) {
            if (co < CU) o.out[v * CU + co] = to_bf16(val);
            else o.out_f32[v * (COUT - CU) + (co - CU)] = val;
        } else {
            o.out[v * COUT + co] = to_bf16(val);
        }
    }
}

// 3x3x3 convolution, stride 1, zero padding, over the channel concatenation [in1 (C1), in2 (C2)],
// with the weights in fragment order wpk[((tap * NCH + ch) * NB + nb) * 64 + lane]. One wave
// computes MT rows (y .. y+MT-1) of 16 voxels in x and NBW of the COUT / 16 channel tiles (all by
// default; fewer, and one row, give the small grids of the coarse levels enough waves); a block
// holds 4 waves. C1 < 4 only without in2 (the input layer: in1 is stored with 4 channels, the rest
// load as 0).
template <int C1, int C2, int COUT, int MT, int EPI, int CU = 0, int NBW = COUT / 16>
__global__ __launch_bounds__(256) void conv3_kernel(Dims d, int batch, const uint16_t* __restrict__ in1,
                                                    const uint16_t* __restrict__ in2, const bf16x4_t* __restrict__ wpk, ConvOut o) {
    constexpr int NCH1 = (C1 + 15) / 16, NCH = NCH1 + C2 / 16, NB = COUT / 16;
    constexpr int GROUPS = NB / NBW;
    static_assert(NB % NBW == 0, "channel tiles per wave must divide the output tiles");
    constexpr int S1 = C1 < 4 ? 4 : C1;   // channel stride of in1
    const int lane = threadIdx.x & 63;
    const uint32_t wave_all = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int tiles_x = d.nx / 16, rows = d.ny / MT;
    const uint32_t total = (uint32_t)tiles_x * rows * d.nz * batch * GROUPS;
    if (wave_all >= total) return;
    const int nb0 = (int)(wave_all % (uint32_t)GROUPS) * NBW;
    const uint32_t wave = wave_all / (uint32_t)GROUPS;
    const int tx = (int)(wave % (uint32_t)tiles_x);
    const int ry = (int)((wave / (uint32_t)tiles_x) % (uint32_t)rows);
    const int zg = (int)(wave / ((uint32_t)tiles_x * rows));
    const int zl = zg % d.nz;
    const int x0 = tx * 16, y0 = ry * MT;
    const int m = lane & 15, kq = lane >> 4;
    floatx4_t acc[MT][NBW];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) acc[t][nb] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tap = 0; tap < 27; ++tap) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const int xx = x0 + m + dx;
        const bool okxz = xx >= 0 && xx < d.nx && zl + dz >= 0 && zl + dz < d.nz;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            bf16x4_t b[NBW];
#pragma unroll
            for (int nb = 0; nb < NBW; ++nb) b[nb] = wpk[(((size_t)tap * NCH + ch) * NB + nb0 + nb) * 64 + lane];
            const bool first = ch < NCH1;
            const int cstride = first ? S1 : C2;
            const int cbase = (first ? ch * 16 : (ch - NCH1) * 16) + 4 * kq;
            const bool okc = first ? (cbase < S1) : true;
            const uint16_t* src = first ? in1 : in2;
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                const int yy = y0 + t + dy;
                bf16x4_t a = bf16x4_t{0, 0, 0, 0};
                if (okxz && okc && yy >= 0 && yy < d.ny)
                    a = *reinterpret_cast<const bf16x4_t*>(src + (size_t)vox(d, xx, yy, zg + dz) * cstride + cbase);
#pragma unroll
                for (int nb = 0; nb < NBW; ++nb) acc[t][nb] = mfma(a, b[nb], acc[t][nb]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) conv_store<EPI, COUT, CU>(o, d, acc[t][nb], x0 + 4 * kq, y0 + t, zg, (nb0 + nb) * 16 + m);
}

// 3x3x3 convolution as conv3_kernel, with the block's input tile staged in LDS and the products on
// v_mfma_f32_16x16x32_bf16. A block computes a 16 x TY x TZ brick of output voxels (rows of 16 in x)
// over all COUT channels: it loads the brick plus a one-voxel halo, every input channel, into LDS
// once (zeros outside the sample: the padding), and the 27 taps read it from there. K runs over
// (tap, input channel), channel fastest, kk = tap * CS + ci with CS the stored input channels (the
// input layer's in1 holds 4, the concatenation [in1, in2] C1 + C2), 32 per MFMA step; the taps
// 27..31 of a last partial step read zeros. The 4 waves split the work into WK interleaved parts
// of the steps x NG groups of NBW channel tiles x RG groups of RW rows; with WK > 1 the partial
// sums meet in LDS and the first part's waves add them in a fixed order. Each wave keeps the
// weights of RING steps in flight (RING = 1: the next step's, loaded before a step's products)
// and, with RING > 1, reads the next step's LDS operands during a step's products. wpk: weights in
// the order of the steps, wpk[(step * NB + nb) * 64 + lane] = 8 BF16 (pack_weights32). The grid is
// the brick count rounded up to a multiple of 8 (XCD remap).
template <int C1, int C2, int COUT, int TY, int TZ, int NBW, int WK, int RING, int EPI, int CU = 0>
__global__ __launch_bounds__(256) void conv3_tile_kernel(Dims d, int batch, const uint16_t* __restrict__ in1,
                                                         const uint16_t* __restrict__ in2, const uint4* __restrict__ wpk, ConvOut o) {
    static_assert(C1 >= 4 || C2 == 0, "an input layer (fewer than 4 channels) has no second input");
    constexpr int S1 = C1 < 4 ? 4 : C1;          // channel stride of in1
    constexpr int CS = S1 + C2;                  // channels per voxel in K
    static_assert(CS == 4 || (S1 % 8 == 0 && C2 % 8 == 0), "channel groups of 8");
    constexpr int VS = CS == 4 ? 4 : CS + 8;     // LDS voxel stride (BF16): 16 lanes' reads hit distinct banks
    constexpr int NB = COUT / 16, NG = NB / NBW, R = TY * TZ;
    static_assert(NB % NBW == 0 && 4 % (WK * NG) == 0, "wave split");
    constexpr int RG = 4 / (WK * NG);
    static_assert(R % RG == 0, "row split");
    constexpr int RW = R / RG;
    constexpr int BX = 18, BY = TY + 2, BZ = TZ + 2, BV = BX * BY * BZ;
    constexpr int STEPS = (27 * CS + 31) / 32;
    constexpr int kRing = RING;
    static_assert(RING >= 1, "weights in flight");
    // the tile, then (WK > 1) the partial sums of the parts after the first, in the same memory
    constexpr int TILE_BYTES = BV * VS * 2, RED_BYTES = (WK - 1) * NG * RG * RW * NBW * 64 * 16;
    __shared__ __align__(16) uint8_t smem[TILE_BYTES > RED_BYTES ? TILE_BYTES : RED_BYTES];
    uint16_t* tile = reinterpret_cast<uint16_t*>(smem);

    const int tiles_x = d.nx / 16, tiles_y = d.ny / TY, tiles_z = d.nz / TZ;
    uint32_t blk = xcd_block();
    if (blk >= (uint32_t)tiles_x * tiles_y * tiles_z * batch) return;   // padding blocks (whole block)
    const int x0 = (int)(blk % (uint32_t)tiles_x) * 16;
    blk /= (uint32_t)tiles_x;
    const int y0 = (int)(blk % (uint32_t)tiles_y) * TY;
    blk /= (uint32_t)tiles_y;
    const int z0 = (int)(blk % (uint32_t)tiles_z) * TZ;
    const int zs = (int)(blk / (uint32_t)tiles_z) * d.nz;   // first plane of the sample

    // stage the brick and its halo
    if constexpr (CS == 4) {
        for (int u = threadIdx.x; u < BV; u += 256) {
            const int bx = u % BX, by = (u / BX) % BY, bz = u / (BX * BY);
            const int x = x0 - 1 + bx, y = y0 - 1 + by, z = z0 - 1 + bz;
            uint2 v = make_uint2(0u, 0u);
            if (x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz)
                v = *reinterpret_cast<const uint2*>(in1 + (size_t)vox(d, x, y, zs + z) * 4);
            *reinterpret_cast<uint2*>(tile + u * VS) = v;
        }
    } else {
        constexpr int U = CS / 8;   // 16-byte units per voxel
        for (int u = threadIdx.x; u < BV * U; u += 256) {
            const int bv = u / U, c = (u % U) * 8;
            const int bx = bv % BX, by = (bv / BX) % BY, bz = bv / (BX * BY);
            const int x = x0 - 1 + bx, y = y0 - 1 + by, z = z0 - 1 + bz;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz) {
                const size_t vx = vox(d, x, y, zs + z);
                v = c < S1 ? *reinterpret_cast<const uint4*>(in1 + vx * S1 + c) : *reinterpret_cast<const uint4*>(in2 + vx * C2 + (c - S1));
            }
            *reinterpret_cast<uint4*>(tile + bv * VS + c) = v;
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int ks = w % WK, grp = w / WK;          // part of the steps; row x channel-tile group
    const int nb0 = (grp % NG) * NBW, r0 = (grp / NG) * RW;
    int rowoff[RW];   // LDS voxel of (x = m - 1, the row's y - 1, z - 1): tap (0, 0, 0) of the row
#pragma unroll
    for (int i = 0; i < RW; ++i) rowoff[i] = m + BX * (((r0 + i) % TY) + BY * ((r0 + i) / TY));
    floatx4_t acc[RW][NBW];
#pragma unroll
    for (int i = 0; i < RW; ++i)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) acc[i][nb] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    auto tap_off = [](int tap) { return tap % 3 + BX * ((tap / 3) % 3 + BY * (tap / 9)); };
    // the A operands of step s (zeros for the taps past 26)
    auto load_a = [&](int s, uint4 (&a)[RW]) {
        const int kk = 32 * s + 8 * g;
        if constexpr (CS == 4) {
            const int t0 = kk >> 2, o0 = tap_off(t0), o1 = tap_off(t0 + 1);
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                uint2 lo = make_uint2(0u, 0u), hi = make_uint2(0u, 0u);
                if (t0 < 27) lo = *reinterpret_cast<const uint2*>(tile + (rowoff[i] + o0) * VS);
                if (t0 + 1 < 27) hi = *reinterpret_cast<const uint2*>(tile + (rowoff[i] + o1) * VS);
                a[i] = make_uint4(lo.x, lo.y, hi.x, hi.y);
            }
        } else {
            const int tap = kk / CS, c0 = kk % CS, o0 = tap_off(tap);
#pragma unroll
            for (int i = 0; i < RW; ++i)
                a[i] = tap < 27 ? *reinterpret_cast<const uint4*>(tile + (rowoff[i] + o0) * VS + c0) : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    // this wave's steps: ks, ks + WK, ...; n = their count
    constexpr int NMAX = (STEPS + WK - 1) / WK;
    const int n = (STEPS - ks + WK - 1) / WK;
    uint4 ring[kRing][NBW];
#pragma unroll
    for (int p = 0; p < kRing; ++p)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb)
            ring[p][nb] = p < n ? wpk[((size_t)(ks + WK * p) * NB + nb0 + nb) * 64 + lane] : make_uint4(0u, 0u, 0u, 0u);
    uint4 a[RW];
    if constexpr (RING == 1) {
        // fully unrolled: the compiler overlaps a step's LDS reads with the previous step's products
#pragma unroll
        for (int t = 0; t < NMAX; ++t) {
            if (t < n) {
                uint4 bn[NBW];
#pragma unroll
                for (int nb = 0; nb < NBW; ++nb)
                    bn[nb] = t + 1 < n ? wpk[((size_t)(ks + WK * (t + 1)) * NB + nb0 + nb) * 64 + lane] : ring[0][nb];
                load_a(ks + WK * t, a);
#pragma unroll
                for (int i = 0; i < RW; ++i)
#pragma unroll
                    for (int nb = 0; nb < NBW; ++nb) acc[i][nb] = mfma32(a[i], ring[0][nb], acc[i][nb]);
#pragma unroll
                for (int nb = 0; nb < NBW; ++nb) ring[0][nb] = bn[nb];
            }
        }
    } else {
        load_a(ks, a);
#pragma unroll 1
        for (int t0 = 0; t0 < NMAX; t0 += kRing) {
#pragma unroll
            for (int p = 0; p < kRing; ++p) {
                const int t = t0 + p;
                if (t < n) {
                    uint4 an[RW];
                    if (t + 1 < n) load_a(ks + WK * (t + 1), an);
#pragma unroll
                    for (int i = 0; i < RW; ++i)
#pragma unroll
                        for (int nb = 0; nb < NBW; ++nb) acc[i][nb] = mfma32(a[i], ring[p][nb], acc[i][nb]);
                    if (t + kRing < n)
#pragma unroll
                        for (int nb = 0; nb < NBW; ++nb) ring[p][nb] = wpk[((size_t)(ks + WK * (t + kRing)) * NB + nb0 + nb) * 64 + lane];
#pragma unroll
                    for (int i = 0; i < RW; ++i) a[i] = an[i];
                }
            }
        }
    }
    if constexpr (WK > 1) {
        floatx4_t* red = reinterpret_cast<floatx4_t*>(smem);   // [(ks - 1)][grp][i][nb][lane]
        constexpr int GR = NG * RG;
        __syncthreads();   // every wave is done with the tile
        if (ks > 0)
#pragma unroll
            for (int i = 0; i < RW; ++i)
#pragma unroll
                for (int nb = 0; nb < NBW; ++nb) red[((((ks - 1) * GR + grp) * RW + i) * NBW + nb) * 64 + lane] = acc[i][nb];
        __syncthreads();
        if (ks > 0) return;
        for (int k = 1; k < WK; ++k)
#pragma unroll
            for (int i = 0; i < RW; ++i)
#pragma unroll
                for (int nb = 0; nb < NBW; ++nb) acc[i][nb] += red[((((k - 1) * GR + grp) * RW + i) * NBW + nb) * 64 + lane];
    }
#pragma unroll
    for (int i = 0; i < RW; ++i)
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb)
            conv_store<EPI, COUT, CU>(o, d, acc[i][nb], x0 + 4 * g, y0 + (r0 + i) % TY, zs + z0 + (r0 + i) / TY, (nb0 + nb) * 16 + m);
}

// 2x2x2 stride-2 transposed convolution from the coarse grid dc to the fine grid (2 dc): fine voxel
// (2x+kx, 2y+ky, 2z+kz) = W[kz,ky,kx]^T in(x, y, z) + bias. One wave: 16 coarse voxels in x.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void tconv2_kernel(Dims dc, int batch, const uint16_t* __restrict__ in,
                                                     const bf16x4_t* __restrict__ wpk, const float* __restrict__ bias,
                                                     uint16_t* __restrict__ out) {
    constexpr int NCH = CIN / 16, NB = COUT / 16;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int tiles_x = dc.nx / 16;
    const uint32_t total = (uint32_t)tiles_x * dc.ny * dc.nz * batch;
    if (wave >= total) return;
    const int tx = (int)(wave % (uint32_t)tiles_x);
    const int y = (int)((wave / (uint32_t)tiles_x) % (uint32_t)dc.ny);
    const int zg = (int)(wave / ((uint32_t)tiles_x * dc.ny));
    const int x0 = tx * 16, m = lane & 15, kq = lane >> 4;
    bf16x4_t a[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) a[ch] = *reinterpret_cast<const bf16x4_t*>(in + (size_t)vox(dc, x0 + m, y, zg) * CIN + ch * 16 + 4 * kq);
    const Dims df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    for (int tap = 0; tap < 8; ++tap) {
        const int kz = tap >> 2, ky = (tap >> 1) & 1, kx = tap & 1;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            floatx4_t acc = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) acc = mfma(a[ch], wpk[(((size_t)tap * NCH + ch) * NB + nb) * 64 + lane], acc);
            const int co = nb * 16 + m;
            const float bv = bias[co];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xc = x0 + 4 * kq + i;
                out[(size_t)vox(df, 2 * xc + kx, 2 * y + ky, 2 * zg + kz) * COUT + co] = to_bf16(acc[i] + bv);
            }
        }
    }
}

// Input gradient of tconv2_kernel: din(x, y, z)[ci] = sum_t sum_co W[ci][co][t] dout(2x + t)[co],
// times the ReLU mask of the transposed convolution's input (mask > 0). M = 16 coarse voxels,
// N = 16 input channels, K = 16 output channels of one of the 8 fine positions.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void tconv2_dgrad_kernel(Dims dc, int batch, const uint16_t* __restrict__ dout,
                                                           const bf16x4_t* __restrict__ wpk, const uint16_t* __restrict__ mask,
                                                           uint16_t* __restrict__ din) {
    constexpr int NCH = COUT / 16, NB = CIN / 16;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int tiles_x = dc.nx / 16;
    const uint32_t total = (uint32_t)tiles_x * dc.ny * dc.nz * batch;
    if (wave >= total) return;
    const int tx = (int)(wave % (uint32_t)tiles_x);
    const int y = (int)((wave / (uint32_t)tiles_x) % (uint32_t)dc.ny);
    const int zg = (int)(wave / ((uint32_t)tiles_x * dc.ny));
    const int x0 = tx * 16, m = lane & 15, kq = lane >> 4;
    const Dims df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    floatx4_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tap = 0; tap < 8; ++tap) {
        const int kz = tap >> 2, ky = (tap >> 1) & 1, kx = tap & 1;
        const size_t fv = vox(df, 2 * (x0 + m) + kx, 2 * y + ky, 2 * zg + kz);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const bf16x4_t a = *reinterpret_cast<const bf16x4_t*>(dout + fv * COUT + ch * 16 + 4 * kq);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[nb] = mfma(a, wpk[(((size_t)tap * NCH + ch) * NB + nb) * 64 + lane], acc[nb]);
        }
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int ci = nb * 16 + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const size_t v = vox(dc, x0 + 4 * kq + i, y, zg);
            din[v * CIN + ci] = bf16_positive(mask[v * CIN + ci]) ? to_bf16(acc[nb][i]) : (uint16_t)0;
        }
    }
}

// 2x2x2 max pooling; one thread per (coarse voxel, 4-channel group). n = coarse voxels * batch.
template <int C>
__global__ __launch_bounds__(256) void maxpool2_kernel(Dims dc, uint32_t n, const uint16_t* __restrict__ in, uint16_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr uint32_t groups = C / 4;
    if (i >= n * groups) return;
    const uint32_t v = i / groups, g = i - v * groups;
    const int x = (int)(v % (uint32_t)dc.nx), y = (int)((v / (uint32_t)dc.nx) % (uint32_t)dc.ny), zg = (int)(v / ((uint32_t)dc.nx * dc.ny));
    const Dims df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    float mx[4]{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; ++c) {
        const uint16_t* p = in + (size_t)vox(df, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), 2 * zg + (c >> 2)) * C + 4 * g;
        for (int j = 0; j < 4; ++j) mx[j] = fmaxf(mx[j], from_bf16(p[j]));
    }
    for (int j = 0; j < 4; ++j) out[(size_t)v * C + 4 * g + j] = to_bf16(mx[j]);
}

// Gradient through a pooled ReLU output s: the coarse gradient dp goes to the first maximum of each
// 2x2x2 group in the order t = dx + 2 dy + 4 dz, is added to the skip connection's FP32 gradient
// (null: none), and the sum passes s's ReLU (s > 0) and is rounded to BF16.
// One thread per (coarse voxel, channel); n = coarse voxels * batch.
template <int C>
__global__ __launch_bounds__(256) void maxpool2_backward_kernel(Dims dc, uint32_t n, const uint16_t* __restrict__ s,
                                                                const uint16_t* __restrict__ dp, const float* __restrict__ skip,
                                                                uint16_t* __restrict__ ds) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * C) return;
    const uint32_t v = i / C, ch = i - v * C;
    const int x = (int)(v % (uint32_t)dc.nx), y = (int)((v / (uint32_t)dc.nx) % (uint32_t)dc.ny), zg = (int)(v / ((uint32_t)dc.nx * dc.ny));
    const Dims df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    size_t fv[8];
    float best = -INFINITY;
    int arg = 0;
    for (int t = 0; t < 8; ++t) {
        fv[t] = vox(df, 2 * x + (t & 1), 2 * y + ((t >> 1) & 1), 2 * zg + (t >> 2));
        const float val = from_bf16(s[fv[t] * C + ch]);
        if (val > best) {
            best = val;
            arg = t;
        }
    }
    const float g = from_bf16(dp[(size_t)v * C + ch]);
    for (int t = 0; t < 8; ++t) {
        const size_t e = fv[t] * C + ch;
        float val = (skip ? skip[e] : 0.0f) + (t == arg ? g : 0.0f);
        ds[e] = bf16_positive(s[e]) ? to_bf16(val) : (uint16_t)0;
    }
}

// Weight gradient of a 3x3x3 convolution (KIND 0) or a 2x2x2 stride-2 transposed convolution
// (KIND 1), one tap per blockIdx.y:
//   KIND 0: dW[co][ci][tap] = sum_v dz(v)[co] x(v + off(tap))[ci]   (M = co: dz, N = ci: x = [in1, in2])
//   KIND 1: dW[ci][co][t]  = sum_v in(v)[ci] dout(2 v + t)[co]       (M = ci: in1, N = co: dout)
// The M tensor has MC channels at the wave's voxels (16 consecutive in x per step); the N tensor has
// C1 + C2 channels (KIND 0; in1 stored with stride max(C1, 4)) or C1 (KIND 1). Wave w of the grid
// leaves its sum in partials[(w * taps + tap) * MC * NC + m * NC + n] (NC = N channels rounded up to
// 16); wgrad_reduce_kernel adds the waves' slots.
template <int KIND, int MC, int C1, int C2>
__global__ __launch_bounds__(256) void wgrad_kernel(Dims d, int batch, const uint16_t* __restrict__ mt, const uint16_t* __restrict__ in1,
                                                    const uint16_t* __restrict__ in2, uint32_t groups_per_wave,
                                                    float* __restrict__ partials) {
    constexpr int NC = ((C1 + C2 + 15) / 16) * 16, NBM = MC / 16, NBN = NC / 16;
    constexpr int NCH1 = (C1 + 15) / 16;
    constexpr int S1 = C1 < 4 ? 4 : C1;
    constexpr int TAPS = KIND == 0 ? 27 : 8;
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int tap = blockIdx.y;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int tiles_x = d.nx / 16;
    const uint32_t groups = (uint32_t)tiles_x * d.ny * d.nz * batch;
    int dx = 0, dy = 0, dz = 0;
    if (KIND == 0) {
        dz = tap / 9 - 1;
        dy = (tap / 3) % 3 - 1;
        dx = tap % 3 - 1;
    } else {
        dz = tap >> 2;
        dy = (tap >> 1) & 1;
        dx = tap & 1;
    }
    const Dims df{2 * d.nx, 2 * d.ny, 2 * d.nz};
    floatx4_t acc[NBM][NBN];
#pragma unroll
    for (int a = 0; a < NBM; ++a)
#pragma unroll
        for (int b = 0; b < NBN; ++b) acc[a][b] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t g0 = wave * groups_per_wave;
    const uint32_t g1 = min(groups, g0 + groups_per_wave);
    for (uint32_t gi = g0; gi < g1; ++gi) {
        const int tx = (int)(gi % (uint32_t)tiles_x);
        const int y = (int)((gi / (uint32_t)tiles_x) % (uint32_t)d.ny);
        const int zg = (int)(gi / ((uint32_t)tiles_x * d.ny));
        const int zl = zg % d.nz;
        const int xb = tx * 16 + 4 * kq;   // this lane's 4 voxels: xb .. xb + 3
        bf16x4_t a[NBM], b[NBN];
#pragma unroll
        for (int ab = 0; ab < NBM; ++ab) {
            const uint16_t* p = mt + (size_t)vox(d, xb, y, zg) * MC + ab * 16 + m;
            a[ab] = bf16x4_t{(short)p[0], (short)p[MC], (short)p[2 * MC], (short)p[3 * MC]};
        }
        if (KIND == 0) {
            const int yy = y + dy, zz = zl + dz;
            const bool okyz = yy >= 0 && yy < d.ny && zz >= 0 && zz < d.nz;
#pragma unroll
            for (int bb = 0; bb < NBN; ++bb) {
                const bool first = bb < NCH1;
                const int c = (first ? bb * 16 : (bb - NCH1) * 16) + m;
                const int cs = first ? S1 : C2;
                const uint16_t* src = first ? in1 : in2;
                const bool okc = first ? c < C1 : c < C2;
                short v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int xx = xb + j + dx;
                    v[j] = (okyz && okc && xx >= 0 && xx < d.nx) ? (short)src[(size_t)vox(d, xx, yy, zg + dz) * cs + c] : (short)0;
                }
                b[bb] = bf16x4_t{v[0], v[1], v[2], v[3]};
            }
        } else {
#pragma unroll
            for (int bb = 0; bb < NBN; ++bb) {
                const int c = bb * 16 + m;
                short v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = (short)in1[(size_t)vox(df, 2 * (xb + j) + dx, 2 * y + dy, 2 * zg + dz) * C1 + c];
                b[bb] = bf16x4_t{v[0], v[1], v[2], v[3]};
            }
        }
#pragma unroll
        for (int ab = 0; ab < NBM; ++ab)
#pragma unroll
            for (int bb = 0; bb < NBN; ++bb) acc[ab][bb] = mfma(a[ab], b[bb], acc[ab][bb]);
    }
    float* out = partials + ((size_t)wave * TAPS + tap) * MC * NC;
#pragma unroll
    for (int ab = 0; ab < NBM; ++ab)
#pragma unroll
        for (int bb = 0; bb < NBN; ++bb)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[(size_t)(ab * 16 + 4 * kq + i) * NC + bb * 16 + m] = acc[ab][bb][i];
}

// dW[mi][ni][tap] (mi < MC, ni < NREAL) = sum over the waves' slots, in wave order (FP64 sums).
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float* __restrict__ partials, uint32_t waves, int taps, int mc, int nc,
                                                           int nreal, float* __restrict__ dw) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t per = (uint32_t)taps * mc * nc;
    if (i >= per) return;
    const int tap = (int)(i / ((uint32_t)mc * nc));
    const int mi = (int)((i / (uint32_t)nc) % (uint32_t)mc);
    const int ni = (int)(i % (uint32_t)nc);
    double s = 0.0;
    for (uint32_t w = 0; w < waves; ++w) s += (double)partials[(size_t)w * per + i];
    if (ni < nreal) dw[((size_t)mi * nreal + ni) * taps + tap] = (float)s;
}

// Per-block FP64 channel sums of a BF16 tensor with C channels over n voxels:
// partials[block * C + c]. The block's threads cover 256 / C voxels per step.
template <int C>
__global__ __launch_bounds__(256) void channel_sum_kernel(uint32_t n, const uint16_t* __restrict__ t, double* __restrict__ partials) {
    static_assert(256 % C == 0, "channel count must divide 256");
    __shared__ double lds[256];
    constexpr uint32_t per = 256 / C;
    const uint32_t c = threadIdx.x % C, vo = threadIdx.x / C;
    double s = 0.0;
    for (uint32_t v = blockIdx.x * per + vo; v < n; v += gridDim.x * per) s += (double)from_bf16(t[(size_t)v * C + c]);
    lds[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        double tot = 0.0;
        for (uint32_t k = 0; k < per; ++k) tot += lds[k * C + threadIdx.x];
        partials[(size_t)blockIdx.x * C + threadIdx.x] = tot;
    }
}

// out[c] = sum over `blocks` partial rows (fixed order).
__global__ __launch_bounds__(256) void channel_sum_finish_kernel(const double* __restrict__ partials, uint32_t blocks, int c,
                                                                 float* __restrict__ out) {
    const int i = (int)threadIdx.x;
    if (i >= c) return;
    double s = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) s += partials[(size_t)b * c + i];
    out[i] = (float)s;
}

// Per-sample sum of r^2 over the unknowns: partials[b * G + block] (blockIdx.y = sample b).
__global__ __launch_bounds__(256) void residual_norm_kernel(uint32_t n, const float* __restrict__ r, const uint8_t* __restrict__ code,
                                                           double* __restrict__ partials) {
    __shared__ double lds[kMaxWaves];
    const float* rb = r + (size_t)blockIdx.y * n;
    double v[1]{0.0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (code[i]) v[0] += (double)rb[i] * (double)rb[i];
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

// Network input (r / s, unknown, 0, 0) with s = rms of r over the unknowns (1 when r = 0).
__global__ __launch_bounds__(256) void pack_batch_input_kernel(uint32_t n, uint64_t unknowns, const float* __restrict__ r,
                                                               const uint8_t* __restrict__ code, const double* __restrict__ partials,
                                                               uint32_t G, uint16_t* __restrict__ x) {
    __shared__ double lds[kMaxWaves];
    const uint32_t b = blockIdx.y;
    const double* parts[1]{partials + (size_t)b * G};
    double red[1];
    reduce_partials<1>(parts, G, red, lds);
    const double sc = unknowns ? sqrt(red[0] / (double)unknowns) : 0.0;
    const float inv = (float)(1.0 / (sc > 0.0 ? sc : 1.0));
    const float* rb = r + (size_t)b * n;
    uint16_t* xb = x + (size_t)b * n * 4;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const bool u = code[i] != 0u;
        xb[4 * (size_t)i + 0] = to_bf16(u ? rb[i] * inv : 0.0f);
        xb[4 * (size_t)i + 1] = to_bf16(u ? 1.0f : 0.0f);
        xb[4 * (size_t)i + 2] = 0;
        xb[4 * (size_t)i + 3] = 0;
    }
}

// Packs FP32 weights (PyTorch layout) into the BF16 fragment order of the convolution kernels:
// pk[((tap * NCH + ch) * NB + nb) * 64 + l][j] = W'[co = 16 nb + l%16][ci = 16 ch + 4 (l/16) + j][tap]
// (0 for ci >= cin), where W' is the weight of the convolution the kernel computes:
//   kind 0 (3x3x3 forward): W'[co][ci][tap] = w[co][ci][tap]
//   kind 1 (2x2x2 transposed, forward): W'[co][ci][t] = w[ci][co][t]
//   kind 2 (3x3x3 input gradient): W'[co][ci][tap] = w[ci][co][26 - tap]   (cin, cout of W')
//   kind 3 (2x2x2 transposed, input gradient): W'[co][ci][t] = w[co][ci][t]
__global__ __launch_bounds__(256) void pack_weights_kernel(const float* __restrict__ w, int cin, int cout, int taps, int kind,
                                                           uint16_t* __restrict__ pk) {
    const int nch = (cin + 15) / 16, nb = cout / 16;
    const uint32_t total = (uint32_t)taps * nch * nb * 64 * 4;
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int j = (int)(e & 3u), l = (int)((e >> 2) & 63u);
    uint32_t r = e >> 8;
    const int b = (int)(r % (uint32_t)nb);
    r /= (uint32_t)nb;
    const int ch = (int)(r % (uint32_t)nch);
    const int tap = (int)(r / (uint32_t)nch);
    const int co = 16 * b + (l & 15), ci = 16 * ch + 4 * (l >> 4) + j;
    float v = 0.0f;
    if (ci < cin) {
        if (kind == 0 || kind == 3) v = w[((size_t)co * cin + ci) * taps + tap];
        else v = w[((size_t)ci * cout + co) * taps + (kind == 2 ? taps - 1 - tap : tap)];
    }
    pk[e] = to_bf16(v);
}

}  // namespace
}  // namespace unet
}  // namespace ds
