// pcg_resident.hpp — the single-reduction Jacobi-PCG of pcg_fused.hpp as one cooperative launch
// per residual-replacement interval, each cell's vectors held on chip (DS_CG_SINGLE_REDUCTION_RESIDENT).
//
// The streamed kernel (one launch per iteration) re-reads and re-writes r, w, s, p and x_d every
// iteration and pays a launch's fixed cost (wave dispatch, partial reduction, tail) each time. Here
// all blocks of the launch are resident at once (cooperative launch), a block keeps the
// (64 x kTY x Z) cells of its plane tile on chip for the whole interval, and iterations are
// separated by a grid barrier instead of a launch. Per iteration a block
//   * loads the partials of the previous iteration (one partial set per thread) and the raw
//     (r, w, s, code) of its halo: tile rims (spread over all threads) and the planes below and
//     above its chunk (the own column of each thread);
//   * reduces the partials to alpha, beta in the same order in every block (the values agree bit
//     for bit, so every block takes the same convergence decision);
//   * updates its own cells as the streamed kernel does (same FMAs), forms u_{i+1} of the halo
//     from the raw halo values, then w_{i+1} = A u_{i+1} for all its planes through LDS;
//   * writes its partials and publishes r, w, s of the cells that are another block's halo (the
//     tile edge columns and the first and last plane), then meets the grid barrier.
// At exit (the interval done, or convergence) every own cell is written back, so the vectors,
// partials, history and status are what the same number of streamed launches leave.
//
// Data that crosses blocks inside a launch (published halo values, partials) is written and read
// at the device coherence point (buffer accesses with sc0 | sc1, agent-scope atomics for the
// partials): the L2 of one XCD does not see another XCD's stores, and a whole-L2 write-back and
// invalidate per block per iteration cost more than the launch it replaces (measured 76 us per
// iteration at 128^3, against 34 us this way). The grid barrier: every wave waits for its stores
// to be acknowledged, thread 0 of each block adds 1 to a counter and polls it until every block
// has arrived. Every poll is bounded: after kSpinLimit polls (far longer than any healthy wait) the
// block raises the abort flag, which ends every other block's wait too, so the launch always
// drains; the host then reports the failure.
//
// Measured on MI355X (Jacobi-PCG to 1e-6 on the plume): 64^3, 13.7 us per iteration resident
// against 14.6 us streamed; 128^3, 33.9 us against 27.3 us. The barrier, the partial broadcast
// and the coherent halo traffic cost more than the streamed launch saves once a block holds 8
// planes, so the default stays the streamed kernel.
#pragma once

#include "pcg_fused.hpp"

namespace ds {

constexpr int kResidentZ = 8;                      // own planes per block (registers per vector)
constexpr int kRimPerPlane = 2 * kTY + 2 * kTX;    // x-rim cells (2 columns) then y-rim cells (2 rows)
constexpr uint32_t kSpinLimit = 1u << 21;

// Sets as separate members, chosen with a select: a runtime index into an array member of a
// by-value kernel argument would move the argument to scratch memory.
struct ResidentVecs {
    float *r0, *r1;   // residual sets (set c read, c^1 published)
    float *w0, *w1;   // w = A u sets
    float *s0, *s1;   // s = A p sets
    float* p;         // search direction (own cells)
    float* xd;        // solution increment (own cells)
};

struct ResidentSync {
    uint32_t arrive;   // blocks arrived, over the launch (iteration i completes at (i+1) * blocks)
    uint32_t abort;    // set when a wait exceeded its bound
};

__device__ __forceinline__ bool resident_barrier(ResidentSync* sync, uint32_t target, uint32_t* flag) {
    // every wave waits until its coherent stores are acknowledged before the block arrives
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(&sync->arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t ok = 1u;
        uint32_t spins = 0;
        while (__hip_atomic_load(&sync->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&sync->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++spins > kSpinLimit) {
                __hip_atomic_store(&sync->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0u;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *flag = ok;
    }
    __syncthreads();
    return *flag != 0u;
}

// Element i of an FP32 array through a 32-bit byte offset (valid below 2^30 elements): the
// access can then use the array's base in scalar registers plus one 32-bit offset register,
// instead of a 64-bit address per access.
__device__ __forceinline__ float& at(float* a, uint32_t i) { return *reinterpret_cast<float*>(reinterpret_cast<char*>(a) + (i << 2)); }
__device__ __forceinline__ float at(const float* a, uint32_t i) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(a) + (i << 2));
}
__device__ __forceinline__ double at(const double* a, uint32_t i) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(a) + (i << 3));
}

// Raw buffer access: the array base lives in a scalar resource and each access needs one 32-bit
// offset register (plus a scalar offset for the plane), where a flat address would take two
// registers per cell. Word 3 is the gfx9 raw-buffer format (32-bit data, no swizzle); the record
// count bounds the access (a load past it returns 0).
constexpr int kBufferWord3 = 0x00020000;
using Buf = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Buf make_buf(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kBufferWord3);
}
__device__ __forceinline__ float buf_load(Buf b, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(b, voff, soff, 0));
}
__device__ __forceinline__ void buf_store(Buf b, uint32_t voff, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), b, voff, soff, 0);
}
constexpr int kCoherent = 1 | 16;   // sc0 | sc1: served at the device coherence point
__device__ __forceinline__ float buf_load_c(Buf b, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(b, voff, soff, kCoherent));
}
__device__ __forceinline__ void buf_store_c(Buf b, uint32_t voff, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), b, voff, soff, kCoherent);
}
__device__ __forceinline__ uint32_t buf_load_u8(Buf b, uint32_t voff, uint32_t soff) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(b, voff, soff, 0);
}

__device__ __forceinline__ FHalo resident_load(uint32_t j, Buf r, Buf w, Buf s, Buf code) {
    FHalo h;
    h.r = buf_load_c(r, j << 2, 0u);
    h.w = buf_load_c(w, j << 2, 0u);
    h.s = buf_load_c(s, j << 2, 0u);
    h.d = (uint8_t)buf_load_u8(code, j, 0u);
    return h;
}

// u_{i+1} of a halo cell from its raw (r_i, w_i, s_{i-1}), rounded as its owner rounds it.
__device__ __forceinline__ float resident_u(float r, float w, float s, uint32_t d, float alpha, float beta, bool next) {
    const float si = next ? __fmaf_rn(beta, s, w) : w;
    return __fmaf_rn(-alpha, si, r) * inv_diag(d);
}

// Iterations k0 .. k0+count-1 (fewer if the recursive residual meets tol2 * ||b||^2 first),
// reading set `set0` and the partials parts[set0] as launch k0 of the streamed form would. Needs
// full tiles (nx % 64 == 0, ny % kTY == 0) and at most Z planes per block.
//
// Storage per thread: r and s of its Z cells in registers; w, p and x_d in LDS (own slots); u_{i+1}
// of the whole chunk with its tile rim in an LDS tile, so the stencil of every plane follows one
// barrier and the rim u values are written by the threads that load the rim cells.
template <int Z>
__global__ __launch_bounds__(kPcgThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void pcg_resident_kernel(
    Geo g, uint32_t per, PlaneTiling pt, const uint8_t* __restrict__ code, ResidentVecs v, double* parts0, double* parts1,
    const double* __restrict__ bb_parts, double* hist, PcgStatus* status, ResidentSync* sync, double tol2, uint32_t k0, uint32_t count,
    uint32_t set0) {
    __shared__ float utile[Z][kTY + 2][kTX + 2];
    __shared__ float own_w[Z][kPcgThreads];
    __shared__ float own_p[Z][kPcgThreads];
    __shared__ float own_xd[Z][kPcgThreads];
    __shared__ double lds[7 * kMaxWaves];
    __shared__ uint32_t flag;
    if (status->stop) return;   // queued after a launch that converged (block-uniform)
    const uint32_t lb = xcd_block();
    int x0, y0, z0, z1;
    plane_block(g, pt, lb, x0, y0, z0, z1);
    const uint32_t t = threadIdx.x;
    const int lx = (int)(t & 63u), ly = (int)(t >> 6);
    const int x = x0 + lx, y = y0 + ly;
    const int nown = z1 - z0;   // block-uniform, at most Z (below 1 for the padding blocks)
    // columns on the tile edge are halo of the neighbouring tiles (periodic wrap included: with
    // full tiles the domain faces are tile edges); every column's first and last planes are
    // halo of the chunks below and above
    const bool edge_col = lx == 0 || lx == kTX - 1 || ly == 0 || ly == kTY - 1;
    const uint32_t i0 = nown > 0 ? cidx(g, x, y, z0) : 0u;
    const uint32_t plane = (uint32_t)g.nx * (uint32_t)g.ny;
    const uint32_t vbytes = (g.n + 1u) * 4u;   // FP32 vectors carry a zero pad element n
    auto vec = [&](const float* a) { return make_buf(a, vbytes); };
    const Buf bcode = make_buf(code, g.n + 1u);
    const Buf bp = vec(v.p), bxd = vec(v.xd);
    const uint32_t voff = i0 << 2;   // own cell j: voff + j * plane * 4 (a scalar offset)

    float r[Z], s[Z];
    uint32_t dpack[(Z + 3) / 4];   // the code of the own cells, 4 per register
    uint32_t set = set0;
#pragma unroll
    for (int q = 0; q < (Z + 3) / 4; ++q) dpack[q] = 0u;
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        r[j] = s[j] = 0.0f;
        float wj = 0.0f, pj = 0.0f, xj = 0.0f;
        if (j < nown) {
            const uint32_t so = (uint32_t)j * plane * 4u;
            r[j] = buf_load(vec(set ? v.r1 : v.r0), voff, so);
            wj = buf_load(vec(set ? v.w1 : v.w0), voff, so);
            s[j] = buf_load(vec(set ? v.s1 : v.s0), voff, so);
            pj = buf_load(bp, voff, so);
            xj = buf_load(bxd, voff, so);
            dpack[j / 4] |= buf_load_u8(bcode, i0, (uint32_t)j * plane) << (8 * (j % 4));
        }
        own_w[j][t] = wj;
        own_p[j][t] = pj;
        own_xd[j][t] = xj;
    }
    auto dcode = [&](int j) -> uint32_t { return (dpack[j / 4] >> (8 * (j % 4))) & 0xffu; };
    auto write_back = [&](uint32_t st) {
#pragma unroll
        for (int j = 0; j < Z; ++j)
            if (j < nown) {
                const uint32_t so = (uint32_t)j * plane * 4u;
                buf_store(vec(st ? v.r1 : v.r0), voff, so, r[j]);
                buf_store(vec(st ? v.w1 : v.w0), voff, so, own_w[j][t]);
                buf_store(vec(st ? v.s1 : v.s0), voff, so, s[j]);
                buf_store(bp, voff, so, own_p[j][t]);
                buf_store(bxd, voff, so, own_xd[j][t]);
            }
    };
    double bb = bb_parts ? 0.0 : status->bb;
    double gamma_prev = k0 > 0 ? hist[2 * (k0 - 1)] : 1.0;
    const bool has_part = t < gridDim.x;
    const uint32_t tp = has_part ? t : 0u;
    constexpr int kSlots[6]{kPartRR, kPartRZ, kPartPQ, kPartWP, kPartSU, kPartSP};
    constexpr int kRims = Z * kRimPerPlane;
    constexpr int kRimLoads = (kRims + kPcgThreads - 1) / kPcgThreads;
    // the halo cells this thread loads: rim cells q = t + m * threads (tile slot and cell index),
    // and the planes below and above its own column
    int rim_slot[kRimLoads];
    uint32_t rim_cell[kRimLoads];
#pragma unroll
    for (int m = 0; m < kRimLoads; ++m) {
        const int q = (int)t + m * (int)kPcgThreads;
        int64_t jj = -1;
        int slot = -1;
        if (q < kRims) {
            const int j = q / kRimPerPlane, e = q - j * kRimPerPlane;
            const int zz = z0 + j;
            if (zz < z1) {
                if (e < 2 * kTY) {
                    const int side = e / kTY, row = e - side * kTY;
                    jj = wrapped_index(g, per, side ? x0 + kTX : x0 - 1, y0 + row, zz);
                    slot = (j * (kTY + 2) + row + 1) * (kTX + 2) + (side ? kTX + 1 : 0);
                } else {
                    const int e2 = e - 2 * kTY, side = e2 / kTX, col = e2 - side * kTX;
                    jj = wrapped_index(g, per, x0 + col, side ? y0 + kTY : y0 - 1, zz);
                    slot = (j * (kTY + 2) + (side ? kTY + 1 : 0)) * (kTX + 2) + col + 1;
                }
            }
        }
        rim_slot[m] = slot;
        rim_cell[m] = halo_index(g, jj);
    }
    const uint32_t below_cell = halo_index(g, wrapped_index(g, per, x, y, z0 - 1));
    const uint32_t above_cell = halo_index(g, wrapped_index(g, per, x, y, z1));
    float* const ut = &utile[0][0][0];

    for (uint32_t it = 0; it < count; ++it) {
        const uint32_t k = k0 + it;
        const bool next = k > 0;
        const double* in_parts = set ? parts1 : parts0;
        double* out_parts = set ? parts0 : parts1;
        const Buf vr = vec(set ? v.r1 : v.r0);
        const Buf vw = vec(set ? v.w1 : v.w0);
        const Buf vs = vec(set ? v.s1 : v.s0);
        // independent loads, issued together: partials, rim cells, the planes below and above
        double part[7];
#pragma unroll
        for (int q = 0; q < 6; ++q)
            part[q] = __hip_atomic_load(&in_parts[kSlots[q] * kPcgMaxBlocks + tp], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        part[6] = (it == 0 && bb_parts) ? at(bb_parts, kPartBB * kPcgMaxBlocks + tp) : 0.0;
        FHalo rl[kRimLoads];
#pragma unroll
        for (int m = 0; m < kRimLoads; ++m) rl[m] = resident_load(rim_cell[m], vr, vw, vs, bcode);
        const FHalo hb = resident_load(below_cell, vr, vw, vs, bcode);
        const FHalo ht = resident_load(above_cell, vr, vw, vs, bcode);

        double red[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) red[q] = has_part ? part[q] : 0.0;
        block_sum<7>(red, lds);
        if (it == 0 && bb_parts) bb = red[6];
        const double rho0 = red[0], gamma = red[1], delta = red[2];
        if (!(rho0 > tol2 * bb)) {   // converged (NaN also stops; the true residual reports it)
            if (lb == 0 && t == 0) {
                if (it == 0 && bb_parts) status->bb = bb;
                status->rr = rho0;
                status->iterations = k;
                status->stop = 1u;
            }
            write_back(set);
            return;
        }
        double b64 = 0.0, a64;
        if (next) {
            b64 = gamma / gamma_prev;
            // explicit FMAs: the rounding does not depend on the compiler contracting products
            a64 = gamma / __fma_rn(b64 * b64, red[5], __fma_rn(b64, red[3] + red[4], delta));
        } else {
            a64 = gamma / delta;
        }
        if (lb == 0 && t == 0) {
            if (it == 0 && bb_parts) status->bb = bb;
            hist[2 * k] = gamma;
            hist[2 * k + 1] = a64;
        }
        gamma_prev = gamma;
        const float alpha = (float)a64, beta = (float)b64;

        // u_{i+1} of the rim cells this thread loaded, into their tile slots
#pragma unroll
        for (int m = 0; m < kRimLoads; ++m)
            if (rim_slot[m] >= 0) ut[rim_slot[m]] = resident_u(rl[m].r, rl[m].w, rl[m].s, rl[m].d, alpha, beta, next);
        // own cells: s_i, p_i, x_d, r_{i+1}, u_{i+1} (the streamed kernel's own_update)
        double rho = 0.0, gam = 0.0, del = 0.0, wp = 0.0, su = 0.0, sp = 0.0;
#pragma unroll
        for (int j = 0; j < Z; ++j) {
            if (j < nown) {
                const uint32_t d = dcode(j);
                const float inv = inv_diag(d);
                const float wj = own_w[j][t];
                const float si = next ? __fmaf_rn(beta, s[j], wj) : wj;
                const float rn = __fmaf_rn(-alpha, si, r[j]);
                const float un = rn * inv;
                if (d != 0u) {
                    const float ui = r[j] * inv;
                    const float pi = next ? __fmaf_rn(beta, own_p[j][t], ui) : ui;
                    own_p[j][t] = pi;
                    own_xd[j][t] = __fmaf_rn(alpha, pi, own_xd[j][t]);
                    s[j] = si;
                    su = __fma_rn((double)si, (double)un, su);
                    sp = __fma_rn((double)si, (double)pi, sp);
                    r[j] = rn;
                } else {
                    r[j] = 0.0f;
                }
                rho = __fma_rn((double)rn, (double)rn, rho);
                gam = __fma_rn((double)rn, (double)un, gam);
                utile[j][ly + 1][lx + 1] = un;
            }
        }
        const float u_below = resident_u(hb.r, hb.w, hb.s, hb.d, alpha, beta, next);
        const float u_above = resident_u(ht.r, ht.w, ht.s, ht.d, alpha, beta, next);
        __syncthreads();
        // w_{i+1} = A u_{i+1} for every own plane (the streamed kernel's stencil, same order)
#pragma unroll
        for (int j = 0; j < Z; ++j) {
            if (j < nown) {
                const uint32_t d = dcode(j);
                const float un = utile[j][ly + 1][lx + 1];
                const float dn = j > 0 ? utile[j - 1][ly + 1][lx + 1] : u_below;
                const float up = j + 1 < nown ? utile[j + 1 < Z ? j + 1 : j][ly + 1][lx + 1] : u_above;
                float wv = 0.0f;
                if (d != 0u)
                    wv = __fmaf_rn((float)d, un,
                                   -((((utile[j][ly + 1][lx] + utile[j][ly + 1][lx + 2]) + (utile[j][ly][lx + 1] + utile[j][ly + 2][lx + 1])) +
                                      dn) +
                                     up));
                own_w[j][t] = wv;
                del = __fma_rn((double)wv, (double)un, del);
                wp = __fma_rn((double)wv, (double)(d != 0u ? own_p[j][t] : 0.0f), wp);
            }
        }

        double red6[6]{rho, gam, del, wp, su, sp};
        block_sum<6>(red6, lds);
        if (t == 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q)
                __hip_atomic_store(&out_parts[kSlots[q] * kPcgMaxBlocks + lb], red6[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        set ^= 1u;
        if (it + 1 == count) break;
        // publish the cells that are another block's halo into the new set
        const Buf pr = vec(set ? v.r1 : v.r0);
        const Buf pw = vec(set ? v.w1 : v.w0);
        const Buf ps = vec(set ? v.s1 : v.s0);
#pragma unroll
        for (int j = 0; j < Z; ++j)
            if (j < nown && (edge_col || j == 0 || j == nown - 1)) {
                const uint32_t so = (uint32_t)j * plane * 4u;
                buf_store_c(pr, voff, so, r[j]);
                buf_store_c(pw, voff, so, own_w[j][t]);
                buf_store_c(ps, voff, so, s[j]);
            }
        if (!resident_barrier(sync, (it + 1) * gridDim.x, &flag)) {   // the host reports the failure
            if (t == 0) {
                status->aborted = 1u;
                status->stop = 1u;
            }
            return;
        }
    }
    write_back(set);
}

}  // namespace ds
