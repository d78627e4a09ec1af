// pcg_s2.hpp — s-step Jacobi-PCG with s = 2: two CG iterations per kernel launch and per global
// reduction.
//
// With T = D^-1 A (D the Jacobi diagonal) and the preconditioned residual z = D^-1 r, the direction,
// residual and solution increment of the next two CG steps lie in the span of the monomial basis
//   V = [P0, P1, P2, Z0, Z1] = [p, T p, T^2 p, z, T z].
// Every inner product CG takes is a quadratic form of the D-weighted Gram matrix G = V^T D V
// (<r, z> = z^T D z and <p, A p> = p^T D (T p)), and T acts on basis coordinates as the shift
// B: P0 -> P1 -> P2, Z0 -> Z1. One reduction of G (15 sums) and of rho = ||D Z0||^2 = ||r||^2
// therefore gives the scalars of both steps (Chronopoulos & Gear's s-step CG, J. Comput. Appl.
// Math. 25, 1989; in the form of Carson's communication-avoiding CG, "Communication-avoiding Krylov
// subspace methods in theory and practice", 2015, with the basis stored instead of recomputed).
//
// A launch runs the two steps in coordinates (FP64, every thread alike), then per cell
//   stage A (pointwise, plane t):   p'' = V a, z'' = V c, x_d += V e    (own cell and tile rim 2)
//   stage B (stencil, plane t-1):   P1'' = T p''                         (own cell and tile rim 1)
//   stage C (stencil, plane t-2):   P2'' = T P1'', Z1'' = T z''; stores the new basis V'' and the
//                                   partial sums of the next G and rho   (own cell)
// The tile is the plane march of the one-step kernels: 64 x kTY (x, y) own cells per block, z
// planes of one chunk; the tile rim of width 2 (in x, y and z) is rebuilt from the stored basis,
// so no value crosses blocks inside a launch. Traffic per unknown and launch: the basis and x_d
// read, the new basis and x_d written, the code byte: 49 B for two iterations (the one-step
// single-reduction kernel moves 41 B per iteration).
//
// A cell's values are computed by the same expressions whether a block owns it or rebuilds it as
// rim, so the stored basis does not depend on the tiling.
#pragma once

#include "pcg_plane.hpp"

namespace ds {

// Partial sums per block (block-major, 16 doubles per block): G upper triangle row-major
// (00 01 02 03 04 11 12 13 14 22 23 24 33 34 44), then rho.
constexpr int kS2Parts = 16;
constexpr int kS2Rho = 15;
// Planes whose loads are in flight ahead of stage A.
constexpr int kS2Depth = 3;
// Steps of the plane march per unrolled loop pass: one load slot each.
constexpr int kS2Slots = kS2Depth + 1;

static_assert(kTX == 64 && kTY >= 6, "tile rim roles assume 64-wide tiles of at least 6 rows");

struct S2Vecs {
    const float* v[5];   // basis of this launch (set c): P0, P1, P2, Z0, Z1; element n is a zero pad
    float* o[5];         // basis for the next launch (set c^1)
    float* xd;           // solution increment since the last fold into the pair (own cells, in place)
};

// The same sets as one allocation each: vector m of a set starts at element m * (n + 1). The
// iteration kernel addresses them through buffer resources (a 32-bit byte offset per cell shared
// by the five vectors), so a set must stay below 4 GiB: 5 (n + 1) 4 bytes < 2^32.
struct S2Sets {
    const float* in;     // set c
    float* out;          // set c^1
    float* xd;
};
constexpr uint32_t kS2BufferFlags = 0x00020000u;   // dword 3 of a raw (stride 0) gfx9 buffer resource

__device__ __forceinline__ __amdgpu_buffer_rsrc_t s2_resource(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, (int)kS2BufferFlags);
}
__device__ __forceinline__ float s2_load(__amdgpu_buffer_rsrc_t r, uint32_t byte, uint32_t base) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte, (int)base, 0));
}
__device__ __forceinline__ void s2_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t byte, uint32_t base) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)byte, (int)base, 0);
}

// The two (or one) CG steps in basis coordinates. a, c, e: coordinates of the new direction,
// residual and solution increment; gamma = z^T D z before; ok = false when a denominator is not
// positive and finite (the host then checks the true residual).
struct S2Coef {
    double a[5], c[5], e[5];
    bool ok;
};

__device__ __forceinline__ double s2_form(const double (&G)[5][5], const double (&u)[5], const double (&w)[5]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 5; ++j) t = __fma_rn(G[i][j], w[j], t);
        s = __fma_rn(u[i], t, s);
    }
    return s;
}

__device__ __forceinline__ S2Coef s2_coefficients(const double (&G)[5][5], uint32_t steps) {
    S2Coef k{};
    double p[5]{1.0, 0.0, 0.0, 0.0, 0.0}, z[5]{0.0, 0.0, 0.0, 1.0, 0.0}, x[5]{};
    double gamma = s2_form(G, z, z);
    k.ok = gamma > 0.0 && gamma < INFINITY;
#pragma unroll
    for (uint32_t j = 0; j < 2; ++j) {
        if (j < steps) {
            // B p: the shift of the basis (p never holds P2 or Z1 coordinates before its last step)
            const double bp[5]{0.0, p[0], p[1], 0.0, p[3]};
            const double den = s2_form(G, p, bp);
            k.ok = k.ok && den > 0.0 && den < INFINITY;
            const double alpha = gamma / den;
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                x[i] = __fma_rn(alpha, p[i], x[i]);
                z[i] = __fma_rn(-alpha, bp[i], z[i]);
            }
            const double gamma_next = s2_form(G, z, z);
            const double beta = gamma_next / gamma;
#pragma unroll
            for (int i = 0; i < 5; ++i) p[i] = __fma_rn(beta, p[i], z[i]);
            gamma = gamma_next;
        }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        k.a[i] = p[i];
        k.c[i] = z[i];
        k.e[i] = x[i];
        // coefficients that do not fit FP32 would put non-finite values into the basis and the
        // solution: stop instead, and let the true residual decide
        k.ok = k.ok && fabs(p[i]) < 1e30 && fabs(z[i]) < 1e30 && fabs(x[i]) < 1e30;
    }
    return k;
}

// T x at a cell: (d x - sum of the 6 neighbours) / d, the neighbour sum in the fixed order
// ((x- + x+) + (y- + y+)) + z- + z+ (0 off the unknowns through inv = 0).
__device__ __forceinline__ float s2_apply(float c, float xm, float xp, float ym, float yp, float zm, float zp, float d, float inv) {
    const float s = (((xm + xp) + (ym + yp)) + zm) + zp;
    return __fmaf_rn(d, c, -s) * inv;
}

// V u for 5 basis values (the FMA chain fixes the rounding).
__device__ __forceinline__ float s2_combine(const float (&u)[5], const float (&v)[5]) {
    float s = u[0] * v[0];
#pragma unroll
    for (int i = 1; i < 5; ++i) s = __fmaf_rn(u[i], v[i], s);
    return s;
}

// Cell coordinate of a tile slot along one axis, wrapped on a periodic axis: [-2, n + 1] maps
// into [0, n); anything else (or a non-periodic axis outside [0, n)) is -1.
__device__ __forceinline__ int s2_wrap(int x, int n, bool periodic) {
    if (x >= 0 && x < n) return x;
    if (!periodic || x < -2 || x > n + 1) return -1;
    return x < 0 ? x + n : x - n;
}

// Wave sum of 16 values by reduce-scatter: after the exchanges over lane bits 5..2 each lane holds
// one value index summed over 16 lanes, the last two exchanges finish the sum over the wave.
// Returns, in lane l, the wave total of value (l >> 2) & 15.
__device__ __forceinline__ double s2_wave_reduce16(const double (&v)[16]) {
    const int lane = (int)(threadIdx.x & 63u);
    double a[8];
    const bool b5 = lane & 32;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const double keep = b5 ? v[8 + i] : v[i], give = b5 ? v[i] : v[8 + i];
        a[i] = keep + __shfl_xor(give, 32, 64);
    }
    double b[4];
    const bool b4 = lane & 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double keep = b4 ? a[4 + i] : a[i], give = b4 ? a[i] : a[4 + i];
        b[i] = keep + __shfl_xor(give, 16, 64);
    }
    double c[2];
    const bool b3 = lane & 8;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double keep = b3 ? b[2 + i] : b[i], give = b3 ? b[i] : b[2 + i];
        c[i] = keep + __shfl_xor(give, 8, 64);
    }
    const bool b2 = lane & 4;
    double d = (b2 ? c[1] : c[0]) + __shfl_xor(b2 ? c[0] : c[1], 4, 64);
    d += __shfl_xor(d, 2, 64);
    d += __shfl_xor(d, 1, 64);
    return d;
}

// Block totals of 16 per-thread values (a fixed order: the same in every block), written to
// out[0..15] by threads 0..15. lds: 16 * kMaxWaves doubles.
__device__ __forceinline__ void s2_block_partials(const double (&v)[16], double* lds, double* __restrict__ out) {
    constexpr int kWaves = (int)(kPcgThreads / 64);
    const int lane = (int)(threadIdx.x & 63u), wid = (int)(threadIdx.x >> 6);
    const double w = s2_wave_reduce16(v);
    if ((lane & 3) == 0) lds[wid * 16 + (lane >> 2)] = w;
    __syncthreads();
    if (threadIdx.x < 16) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) t += lds[q * 16 + threadIdx.x];
        out[threadIdx.x] = t;
    }
}

// Totals of the producer's per-block partials (block-major, 16 per block, at most 512 blocks),
// identical in every consumer block: thread t sums slot t & 15 over blocks t >> 4, +32, ...
// (s2_load_partials, issued before anything else of the launch), then over the 4 lanes of a wave
// with that slot, then over the waves in a fixed order (s2_totals, which returns with tot[0..15]
// in LDS after a barrier).
constexpr int kS2PartLoads = (512 * 16) / kPcgThreads;
__device__ __forceinline__ void s2_load_partials(const double* __restrict__ parts, uint32_t blocks, double (&pl)[kS2PartLoads]) {
    const uint32_t q = threadIdx.x & 15u, b0 = threadIdx.x >> 4;
#pragma unroll
    for (int j = 0; j < kS2PartLoads; ++j) {
        const uint32_t b = b0 + (uint32_t)j * (kPcgThreads / 16);
        pl[j] = b < blocks ? parts[b * 16u + q] : 0.0;
    }
}
__device__ __forceinline__ void s2_totals(const double (&pl)[kS2PartLoads], double* lds, double* tot) {
    constexpr int kWaves = (int)(kPcgThreads / 64);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kS2PartLoads; ++j) s += pl[j];
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    const int lane = (int)(threadIdx.x & 63u), wid = (int)(threadIdx.x >> 6);
    if (lane < 16) lds[wid * 16 + lane] = s;
    __syncthreads();
    if (threadIdx.x < 16) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += lds[w * 16 + threadIdx.x];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
}

struct S2Slot {
    float v[5];    // own cell: basis values
    float xd;      // own cell: solution increment
    float rv[5];   // rim cell: basis values
    uint8_t d;     // own cell: diagonal count (0 off the unknowns)
    uint8_t rd;    // rim-1 cell: diagonal count
};

// One launch: steps (1 or 2) CG iterations from iteration k. in_parts: the producer's partial
// sums (16 per block, block-major); status->bb holds ||b||^2. check_only: convergence test only.
__global__ __launch_bounds__(kPcgThreads) void pcg_s2_kernel(Geo g, uint32_t per, PlaneTiling pt, const uint8_t* __restrict__ code,
                                                             S2Sets v, const double* __restrict__ in_parts, uint32_t in_blocks,
                                                             double* __restrict__ out_parts, PcgStatus* status, double tol2, uint32_t k,
                                                             uint32_t steps, uint32_t check_only) {
    __shared__ float lp[2][kTY + 4][kTX + 4];    // p'' of a plane, tile rim 2
    __shared__ float lq[2][kTY + 2][kTX + 2];    // P1'' of a plane, tile rim 1
    __shared__ float lz[2][kTY + 2][kTX + 2];    // z'' of a plane, tile rim 1
    __shared__ float linv[8];                    // 1/d
    __shared__ double lds[16 * kMaxWaves];
    __shared__ double tot[16];
    // the producer's partials first: every later step of the launch waits for them. A launch that
    // stops (or follows one that stopped) writes -inf into its rho partial, so the launches queued
    // behind a stop return after this reduction without touching the vectors.
    double pl[kS2PartLoads];
    s2_load_partials(in_parts, in_blocks, pl);
    const uint32_t lb = xcd_block();
    if (threadIdx.x < 8) linv[threadIdx.x] = inv_diag(threadIdx.x);

    int x0, y0, z0, z1;
    plane_block(g, pt, lb, x0, y0, z0, z1);
    const int lx = (int)(threadIdx.x & 63u), ly = (int)(threadIdx.x >> 6);
    const int wave = __builtin_amdgcn_readfirstlane(ly);
    const bool pxp = per & 1u, pyp = per & 2u, pzp = per & 4u;
    const uint32_t n = g.n, sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
    auto column = [&](int cx, int cy) -> uint32_t {
        const int wx = s2_wrap(cx, g.nx, pxp), wy = s2_wrap(cy, g.ny, pyp);
        return wx < 0 || wy < 0 ? n : (uint32_t)wx + sy * (uint32_t)wy;
    };
    // own slot: the cell (x0 + lx, y0 + ly); lanes past the domain hold its periodic image or pad
    const bool own = x0 + lx < g.nx && y0 + ly < g.ny;
    const uint32_t col = column(x0 + lx, y0 + ly);
    // rim slot: rows y0-1 and y0+kTY (rim 1) on the first and last wave, rows y0-2 and y0+kTY+1
    // (rim 2) on the second and second-to-last, the rim columns on waves 2 (rim 1) and 3 (rim 2 and
    // the four rim-1 corners, which only stage B reads)
    int rx = 0, ry = 0;
    bool r1 = false, r2 = false;
    if (wave == 0 || wave == kTY - 1) {
        r1 = true;
        rx = lx;
        ry = wave == 0 ? -1 : kTY;
    } else if (wave == 1 || wave == kTY - 2) {
        r2 = true;
        rx = lx;
        ry = wave == 1 ? -2 : kTY + 1;
    } else if (wave == 2 && lx < 2 * kTY) {
        r1 = true;
        rx = lx < kTY ? -1 : kTX;
        ry = lx < kTY ? lx : lx - kTY;
    } else if (wave == 3 && lx < 2 * kTY + 4) {
        r2 = true;
        if (lx < 2 * kTY) {
            rx = lx < kTY ? -2 : kTX + 1;
            ry = lx < kTY ? lx : lx - kTY;
        } else {
            rx = (lx & 1) ? kTX : -1;
            ry = lx < 2 * kTY + 2 ? -1 : kTY;
        }
    }
    const bool rim = r1 || r2;
    const uint32_t rcol = rim ? column(x0 + rx, y0 + ry) : n;
    // plane offsets: out-of-domain planes of a non-periodic axis read the pad
    auto plane_off = [&](int zz) -> uint32_t {
        if (zz >= 0 && zz < g.nz) return sz * (uint32_t)zz;
        if (!pzp) return n;
        const int w = ((zz % g.nz) + g.nz) % g.nz;
        return sz * (uint32_t)w;
    };
    // buffer resources of the two sets, the increment and the code; vector m of a set at byte
    // m * vstride
    const uint32_t vstride = (n + 1u) * 4u;
    const __amdgpu_buffer_rsrc_t rin = s2_resource(v.in, 5u * vstride), rout = s2_resource(v.out, 5u * vstride);
    const __amdgpu_buffer_rsrc_t rxd = s2_resource(v.xd, vstride), rcode = s2_resource(code, n + 1u);
    auto load = [&](int t, S2Slot& S) {
        const uint32_t off = plane_off(t);
        const uint32_t io = min(col + off, n);
#pragma unroll
        for (int m = 0; m < 5; ++m) S.v[m] = s2_load(rin, io * 4u, (uint32_t)m * vstride);
        S.d = __builtin_amdgcn_raw_buffer_load_b8(rcode, (int)io, 0, 0);
        S.xd = t >= z0 && t < z1 ? s2_load(rxd, io * 4u, 0u) : 0.0f;   // block-uniform
        if (rim) {
            const uint32_t ir = min(rcol + off, n);
#pragma unroll
            for (int m = 0; m < 5; ++m) S.rv[m] = s2_load(rin, ir * 4u, (uint32_t)m * vstride);
            S.rd = r1 ? __builtin_amdgcn_raw_buffer_load_b8(rcode, (int)ir, 0, 0) : (uint8_t)0;
        }
    };

    S2Slot slot[kS2Slots];
    const int t0 = z0 - 2;
    if (z0 < z1) {
#pragma unroll
        for (int j = 0; j < kS2Slots; ++j) load(t0 + j, slot[j]);
    }

    s2_totals(pl, lds, tot);   // its barriers also publish linv
    const double rho = tot[kS2Rho];
    const double bb = status->bb;
    double G[5][5];
    {
        int q = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i)
#pragma unroll
            for (int j = i; j < 5; ++j) {
                G[i][j] = tot[q];
                G[j][i] = tot[q];
                ++q;
            }
    }
    const S2Coef cf = s2_coefficients(G, steps);
    if (!(rho > tol2 * bb) || check_only || !cf.ok) {   // converged (NaN also stops; the true residual decides)
        const bool stopping = !(rho > tol2 * bb) || !cf.ok;
        if (lb == 0 && threadIdx.x == 0) {
            // the first launch that stops records the state; later ones only pass the marker on
            if (!status->stop) {
                status->rr = rho;
                status->iterations = k;
                if (stopping) status->stop = 1u;
            }
        }
        if (stopping && threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 16; ++q) out_parts[16u * lb + q] = q == kS2Rho ? -INFINITY : 0.0;
        }
        return;
    }
    // the coefficients are the same in every thread: keep them in scalar registers
    auto uniform = [](double x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)x))); };
    float ca[5], cc[5], ce[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        ca[i] = uniform(cf.a[i]);
        cc[i] = uniform(cf.c[i]);
        ce[i] = uniform(cf.e[i]);
    }

    double acc[kS2Parts];
#pragma unroll
    for (int q = 0; q < kS2Parts; ++q) acc[q] = 0.0;
    if (z0 < z1) {
        // values along z kept by the thread at the start of step t: own column p'' at t-2, t-1;
        // P1'' at t-3, t-2; z'' at t-3, t-2, t-1; d and 1/d at t-2, t-1; the rim cell's p'' at
        // t-2, t-1, z'' at t-1, d and 1/d at t-1
        float pw0 = 0.0f, pw1 = 0.0f;
        float qw0 = 0.0f, qw1 = 0.0f;
        float zw0 = 0.0f, zw1 = 0.0f, zw2 = 0.0f;
        float dw0 = 0.0f, dw1 = 0.0f, iw0 = 0.0f, iw1 = 0.0f;
        float rp0 = 0.0f, rp1 = 0.0f, rzw = 0.0f, rdw = 0.0f, riw = 0.0f;
        const int ox = lx + 2, oy = ly + 2;              // own cell in lp
        const int rxp = rx + 2, ryp = ry + 2;            // rim cell in lp
        // step t: stage A on plane t, B on t-1, C on t-2; then the loads of plane t + kS2Slots
        auto step = [&](S2Slot& S, int t) {
            const int b0 = t & 1, b1 = (t - 1) & 1;
            // stage A
            const float pa = s2_combine(ca, S.v);
            const float za = s2_combine(cc, S.v);
            const float da = (float)S.d;
            const float ia = linv[S.d & 7u];
            if (own && t >= z0 && t < z1) {
                const float xn = __fmaf_rn(ce[3], S.v[3], __fmaf_rn(ce[1], S.v[1], __fmaf_rn(ce[0], S.v[0], S.xd)));
                s2_store(xn, rxd, (col + plane_off(t)) * 4u, 0u);
            }
            lp[b0][oy][ox] = pa;
            lz[b1][ly + 1][lx + 1] = zw2;   // z'' of plane t-1, read by stage C of plane t-1
            float rpa = 0.0f, rza = 0.0f, rda = 0.0f, ria = 0.0f;
            if (rim) {
                rpa = s2_combine(ca, S.rv);
                lp[b0][ryp][rxp] = rpa;
                if (r1) {
                    rza = s2_combine(cc, S.rv);
                    rda = (float)S.rd;
                    ria = linv[S.rd & 7u];
                    lz[b1][ry + 1][rx + 1] = rzw;
                }
            }
            // stage B on plane t-1 (lp of plane t-1 was written by the previous step)
            const float qb = s2_apply(pw1, lp[b1][oy][ox - 1], lp[b1][oy][ox + 1], lp[b1][oy - 1][ox], lp[b1][oy + 1][ox], pw0, pa, dw1, iw1);
            lq[b1][ly + 1][lx + 1] = qb;
            if (r1) {
                const float rqb = s2_apply(rp1, lp[b1][ryp][rxp - 1], lp[b1][ryp][rxp + 1], lp[b1][ryp - 1][rxp], lp[b1][ryp + 1][rxp], rp0, rpa, rdw,
                                           riw);
                lq[b1][ry + 1][rx + 1] = rqb;
            }
            // stage C on plane t-2 (lq and lz of plane t-2 were written by the previous step)
            const int tc = t - 2;
            if (own && tc >= z0 && tc < z1) {   // tc >= z0 also keeps the warm-up steps out
                const int cx = lx + 1, cy = ly + 1;
                const float P2 = s2_apply(qw1, lq[b0][cy][cx - 1], lq[b0][cy][cx + 1], lq[b0][cy - 1][cx], lq[b0][cy + 1][cx], qw0, qb, dw0, iw0);
                const float Z1 = s2_apply(zw1, lz[b0][cy][cx - 1], lz[b0][cy][cx + 1], lz[b0][cy - 1][cx], lz[b0][cy + 1][cx], zw0, zw2, dw0, iw0);
                const uint32_t i = col + plane_off(tc);
                const float nv[5]{pw0, qw1, P2, zw1, Z1};
#pragma unroll
                for (int m = 0; m < 5; ++m) s2_store(nv[m], rout, i * 4u, (uint32_t)m * vstride);
                const double dd = (double)dw0;
                double q[5], wq[5];
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    q[m] = (double)nv[m];
                    wq[m] = dd * q[m];
                }
                int s = 0;
#pragma unroll
                for (int a = 0; a < 5; ++a)
#pragma unroll
                    for (int c = a; c < 5; ++c) {
                        acc[s] = __fma_rn(wq[a], q[c], acc[s]);
                        ++s;
                    }
                acc[kS2Rho] = __fma_rn(wq[3], wq[3], acc[kS2Rho]);
            }
            load(t + kS2Slots, S);
            __syncthreads();
            pw0 = pw1; pw1 = pa;
            qw0 = qw1; qw1 = qb;
            zw0 = zw1; zw1 = zw2; zw2 = za;
            dw0 = dw1; dw1 = da;
            iw0 = iw1; iw1 = ia;
            rp0 = rp1; rp1 = rpa;
            rzw = rza; rdw = rda; riw = ria;
        };
        // steps t0 .. z1 + 1; the last pass of the unrolled loop may run up to kS2Slots - 1 steps
        // past it, which only load (pad or periodic images) and skip the stores
        for (int t = t0; t <= z1 + 1; t += kS2Slots) {
#pragma unroll
            for (int j = 0; j < kS2Slots; ++j) step(slot[j], t + j);
        }
    }
    s2_block_partials(acc, lds, out_parts + 16u * lb);
}

}  // namespace ds
