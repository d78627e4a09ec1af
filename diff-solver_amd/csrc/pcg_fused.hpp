// pcg_fused.hpp — single-reduction Jacobi-PCG: one kernel launch per iteration.
//
// The Chronopoulos-Gear arrangement of preconditioned CG (Chronopoulos & Gear, "s-step iterative
// methods for symmetric linear systems", J. Comput. Appl. Math. 25, 1989) carries w = A u
// (u = D^-1 r) and s = A p as vectors, so both inner products of an iteration,
// gamma = <r, u> and delta = <w, u>, come from the same vectors and need one global reduction:
//
//   beta_i  = gamma_i / gamma_{i-1}                      (0 for i = 0)
//   alpha_i = gamma_i / <s_i, p_i>                        (gamma_0 / delta_0 for i = 0)
//   p_i = u_i + beta_i p_{i-1};   s_i = w_i + beta_i s_{i-1}
//   x += alpha_i p_i;   r_{i+1} = r_i - alpha_i s_i;   u_{i+1} = D^-1 r_{i+1};   w_{i+1} = A u_{i+1}
//   partials of gamma_{i+1}, delta_{i+1}, rho_{i+1} = ||r_{i+1}||^2 and of the three products
//   <w_{i+1}, p_i>, <s_i, u_{i+1}>, <s_i, p_i> that expand the next denominator:
//   <s_{i+1}, p_{i+1}> = delta_{i+1} + beta (<w_{i+1}, p_i> + <s_i, u_{i+1}>) + beta^2 <s_i, p_i>.
// Chronopoulos & Gear write the denominator as delta_i - beta_i gamma_i / alpha_{i-1}, which holds
// only in exact arithmetic; with FP32 vectors it let the recursive residual run ahead of the true
// one (solves that met 1e-6 recursively failed the true-residual check and restarted). The
// expanded form is the product of the vectors the update actually uses, as in the textbook form.
//
// In exact arithmetic the iterates equal those of the two-reduction form. One kernel does all of
// iteration i: every quantity is pointwise except w_{i+1} = A u_{i+1}, and u_{i+1} at a neighbour
// is a pointwise function of that neighbour's r_i, w_i, s_{i-1} and code, so each block rebuilds
// the u_{i+1} of its tile rim and halo planes from those vectors. r, w and s are therefore
// double-buffered (set c is read, set c^1 written); p and x are read and written by their own
// cell only and stay single. The per-block FP64 partials are double-buffered the same way.
//
// The solution is the compensated pair x_hi + x_lo plus an FP32 increment x_d that the iterations
// update (x_d += alpha p); pcg_fold_kernel adds x_d into the pair and clears it at every residual
// replacement and at exit. x_d holds at most one replacement interval of steps, so its rounding is
// relative to those steps, not to x, and the pair keeps the accuracy it needs (FP32 x alone stalls
// above 1e-6); each iteration touches 4 bytes of solution instead of the pair's 8.
//
// Layout: the plane march of the two-kernel path — a block owns a 64 x (threads / 64) (x, y)
// tile and a chunk of z planes; u of the plane being differentiated sits in LDS (own cell plus
// the tile rim), its z neighbours in registers. The raw loads of plane z+1 are consumed entirely
// while plane z is computed and those of plane z+1+kFusedDepth are issued then; the z loop is
// unrolled over kFusedDepth load slots (compile-time indices, so a slot is never copied) and the
// loads stay in flight for kFusedDepth plane steps.
#pragma once

#include "pcg_plane.hpp"

namespace ds {

// Planes whose loads are in flight ahead of the plane being computed. Measured on the 128^3
// plume: depth 2, 28.3 us per launch; depth 4 (122 VGPRs, still two blocks per CU), 28.7 us; at
// 256^3 both stream about 4.2 TB/s, so the 128^3 launch is bound by its fixed cost (launch,
// partial reduction, tail), not by the prefetch depth.
constexpr int kFusedDepth = 2;

// Kernel modes: the first iteration after set-up (no previous direction), a general iteration,
// and the priming pass w = A u, partials gamma, delta, rho (after set-up or a residual replacement).
enum FusedMode : int { kFusedFirst = 0, kFusedNext = 1, kFusedPrime = 2 };

struct FusedVecs {
    const float* r;       // r_i          (set c)
    const float* w;       // w_i = A u_i  (set c)
    const float* s;       // s_{i-1}      (set c)
    float* r_out;         // r_{i+1}      (set c^1)
    float* w_out;         // w_{i+1}      (set c^1; the priming pass writes set c)
    float* s_out;         // s_i          (set c^1)
    float* p;             // p_{i-1} -> p_i (in place)
    float* xd;            // FP32 increment of the solution since the last fold into the pair (in place)
};

// Halo values of one cell: enough to form u_{i+1} there.
struct FHalo {
    float r, w, s;
    uint8_t d;   // kept as loaded: a widening right after the load would wait for it
};
// Own-cell values only the owning thread needs (d: the solution increment; s: the priming
// pass's s_{i-1}).
struct FOwn {
    float p, d, s;
};
struct FPlane {
    FHalo c, x1, y2;
    FOwn o;
};

// Loaded values are not touched before the plane step that consumes them: a select or a
// divergent branch on a load result makes the compiler wait for that load on the spot, which
// would drain the prefetch. Cells outside the domain (and the rim slots of lanes that are not on
// the tile edge) therefore read element n of every vector and of the code, a pad that is always 0.
__device__ __forceinline__ uint32_t halo_index(const Geo& g, int64_t j) { return j >= 0 ? (uint32_t)j : (uint32_t)g.n; }

template <int Mode>
__device__ __forceinline__ FHalo fused_load_halo(uint32_t j, const FusedVecs& v, const uint8_t* __restrict__ code) {
    FHalo h;
    h.r = v.r[j];
    h.d = code[j];
    h.w = Mode != kFusedPrime ? v.w[j] : 0.0f;
    h.s = Mode == kFusedNext ? v.s[j] : 0.0f;
    return h;
}

// s_i = w_i + beta s_{i-1} and r_{i+1} = r_i - alpha s_i with explicit FMAs, so that the owner of
// a cell and a block that rebuilds it as halo round identically.
template <int Mode>
__device__ __forceinline__ float fused_s(const FHalo& h, float beta) {
    return Mode == kFusedNext ? __fmaf_rn(beta, h.s, h.w) : h.w;
}
template <int Mode>
__device__ __forceinline__ float fused_r(const FHalo& h, float alpha, float beta) {
    return Mode == kFusedPrime ? h.r : __fmaf_rn(-alpha, fused_s<Mode>(h, beta), h.r);
}
template <int Mode>
__device__ __forceinline__ float fused_u(const FHalo& h, float alpha, float beta) {
    return fused_r<Mode>(h, alpha, beta) * inv_diag(h.d);
}

// Iteration k (or the priming pass). in_parts / out_parts: partial sets (kPartCount * kPcgMaxBlocks
// doubles each, slots RR = rho, RZ = gamma, PQ = delta, WP, SU, SP; BB read from bb_parts on the
// first launch of a solve). hist[2k] = gamma_k, hist[2k+1] = alpha_k. check_only: convergence
// test only. with_ps (priming pass): p_{i-1} and s_{i-1} exist, so form WP, SU, SP too.
template <int Mode>
__global__ __launch_bounds__(kPcgThreads) void pcg_fused_kernel(Geo g, uint32_t per, PlaneTiling pt, const uint8_t* __restrict__ code,
                                                         FusedVecs v, const double* __restrict__ in_parts,
                                                         const double* __restrict__ bb_parts, double* __restrict__ out_parts,
                                                         double* __restrict__ hist, PcgStatus* status, double tol2, uint32_t k,
                                                         uint32_t check_only, uint32_t respect_stop, uint32_t with_ps) {
    __shared__ float tile[2][kTY + 2][kTX + 2];
    __shared__ double lds[7 * kMaxWaves];
    // the stop flag, gamma_{i-1} and the producer's partials (below) are independent loads issued
    // together; the plane loads wait for the stop flag (launches queued after convergence must
    // not stream the vectors), gamma_{i-1} is not waited for until beta is formed
    const uint32_t stop = (Mode != kFusedPrime || respect_stop) ? status->stop : 0u;
    const double gamma_prev = Mode == kFusedNext ? hist[2 * (k - 1)] : 1.0;
    const uint32_t lb = xcd_block();
    int x0, y0, z0, z1;
    plane_block(g, pt, lb, x0, y0, z0, z1);
    const int lx = (int)(threadIdx.x & 63u), ly = (int)(threadIdx.x >> 6);
    const int x = x0 + lx, y = y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const Slots sl = make_slots(x0, y0, lx, ly);
    // the y rim belongs to whole waves (rows 0 and kTY-1): a wave-uniform branch
    const int wave_row = __builtin_amdgcn_readfirstlane(ly);
    const bool row_edge = wave_row == 0 || wave_row == kTY - 1;
    constexpr bool kUpdate = Mode != kFusedPrime;

    // this thread's share of the producer's partials, loaded before the planes so that waiting
    // for them leaves the plane loads in flight (gridDim.x <= blockDim.x)
    const bool has_part = threadIdx.x < gridDim.x;
    const uint32_t tp = has_part ? threadIdx.x : 0u;
    const double* bbp = bb_parts ? bb_parts : in_parts;
    constexpr int kSlots[7]{kPartRR, kPartRZ, kPartPQ, kPartWP, kPartSU, kPartSP, kPartBB};
    double part[7]{};
    if (Mode != kFusedPrime) {
#pragma unroll
        for (int q = 0; q < 6; ++q) part[q] = in_parts[kSlots[q] * kPcgMaxBlocks + tp];
        part[6] = bbp[kPartBB * kPcgMaxBlocks + tp];
    }

    if (stop) return;   // block-uniform, before any barrier

    // loads of plane zz: the own column up to the top halo plane z1 (planes beyond read the pad);
    // rims for the chunk planes; the own-cell update values for every plane (used only where the
    // thread owns a cell of a chunk plane)
    auto load_plane = [&](int zz, FPlane& P) {
        const bool chunk = zz < z1;   // block-uniform
        P.x1 = fused_load_halo<Mode>(chunk && sl.has1 ? halo_index(g, wrapped_index(g, per, sl.sx1, y, zz)) : (uint32_t)g.n, v, code);
        if (chunk && row_edge)
            P.y2 = fused_load_halo<Mode>(halo_index(g, wrapped_index(g, per, x, sl.sy2, zz)), v, code);
        else
            P.y2 = FHalo{0.0f, 0.0f, 0.0f, 0u};
        const int64_t j = zz <= z1 ? wrapped_index(g, per, x, y, zz) : -1;
        P.c = fused_load_halo<Mode>(halo_index(g, j), v, code);
        const uint32_t jo = j >= 0 ? (uint32_t)j : 0u;
        if (kUpdate) {
            P.o.d = v.xd[jo];
            P.o.p = Mode == kFusedNext ? v.p[jo] : 0.0f;
            P.o.s = 0.0f;
        } else if (with_ps) {   // block-uniform
            P.o = FOwn{v.p[jo], 0.0f, v.s[jo]};
        } else {
            P.o = FOwn{0.0f, 0.0f, 0.0f};
        }
    };

    FHalo below_raw{0.0f, 0.0f, 0.0f, 0u};
    FPlane cur, slots[kFusedDepth];
    if (z0 < z1) {
        below_raw = fused_load_halo<Mode>(halo_index(g, wrapped_index(g, per, x, y, z0 - 1)), v, code);
        load_plane(z0, cur);
#pragma unroll
        for (int j = 0; j < kFusedDepth; ++j) load_plane(z0 + 1 + j, slots[j]);
    }

    float alpha = 0.0f, beta = 0.0f;
    if (Mode != kFusedPrime) {
        double red[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) red[q] = has_part ? part[q] : 0.0;
        block_sum<7>(red, lds);
        const double rho = red[0], gamma = red[1], delta = red[2];
        const double bb = bb_parts ? red[6] : status->bb;
        if (!(rho > tol2 * bb) || check_only) {   // converged (NaN also stops; the true residual reports it)
            if (lb == 0 && threadIdx.x == 0) {
                if (bb_parts) status->bb = bb;
                status->rr = rho;
                status->iterations = k;
                if (!(rho > tol2 * bb)) status->stop = 1u;
            }
            return;
        }
        double b64 = 0.0, a64;
        if (Mode == kFusedNext) {
            b64 = gamma / gamma_prev;
            // explicit FMAs: the rounding does not depend on the compiler contracting products
            a64 = gamma / __fma_rn(b64 * b64, red[5], __fma_rn(b64, red[3] + red[4], delta));
        } else {
            a64 = gamma / delta;
        }
        if (lb == 0 && threadIdx.x == 0) {
            if (bb_parts) status->bb = bb;
            hist[2 * k] = gamma;
            hist[2 * k + 1] = a64;
        }
        alpha = (float)a64;
        beta = (float)b64;
    }

    double rho = 0.0, gam = 0.0, del = 0.0, wp = 0.0, su = 0.0, sp = 0.0;
    if (z0 < z1) {
        // pointwise part of plane zz for the own cell: updates, stores, partials; returns u_{i+1}
        // and leaves p_i of the cell in pz (for <w_{i+1}, p_i> at the plane's stencil step)
        auto own_update = [&](const FPlane& P, int zz, float& pz) -> float {
            const float u = fused_u<Mode>(P.c, alpha, beta);
            pz = 0.0f;
            if (own && zz < z1) {
                const uint32_t i = cidx(g, x, y, zz);
                const float rn = fused_r<Mode>(P.c, alpha, beta);
                if (kUpdate && P.c.d != 0u) {
                    const float si = fused_s<Mode>(P.c, beta);
                    const float ui = P.c.r * inv_diag(P.c.d);
                    const float pi = Mode == kFusedNext ? __fmaf_rn(beta, P.o.p, ui) : ui;
                    v.p[i] = pi;
                    v.xd[i] = __fmaf_rn(alpha, pi, P.o.d);
                    v.s_out[i] = si;
                    pz = pi;
                    su = __fma_rn((double)si, (double)u, su);
                    sp = __fma_rn((double)si, (double)pi, sp);
                }
                if (!kUpdate && P.c.d != 0u) {   // priming: the existing p_{i-1}, s_{i-1}
                    pz = P.o.p;
                    su = __fma_rn((double)P.o.s, (double)u, su);
                    sp = __fma_rn((double)P.o.s, (double)P.o.p, sp);
                }
                if (kUpdate) v.r_out[i] = P.c.d != 0u ? rn : 0.0f;
                rho = __fma_rn((double)rn, (double)rn, rho);
                gam = __fma_rn((double)rn, (double)u, gam);
            }
            return u;
        };
        auto stage = [&](const FPlane& P, int b) {
            if (sl.has1) tile[b][ly + 1][sl.t1x] = fused_u<Mode>(P.x1, alpha, beta);
            if (sl.has2) tile[b][sl.t2y][lx + 1] = fused_u<Mode>(P.y2, alpha, beta);
        };
        float below = fused_u<Mode>(below_raw, alpha, beta);
        float pc = 0.0f;
        float centre = own_update(cur, z0, pc);
        uint8_t d = cur.c.d;
        tile[0][ly + 1][lx + 1] = centre;
        stage(cur, 0);
        __syncthreads();
        int buf = 0;
        // one plane step: consume the slot holding plane z+1, finish plane z, refill the slot with
        // z+1+kFusedDepth. A step with z >= z1 (past the end of a chunk whose length is not a
        // multiple of the depth) only loads and waits, so every step of the unrolled loop runs
        // unconditionally.
        auto step = [&](FPlane& S, int z) {
            float pn = 0.0f;
            const float next = own_update(S, z + 1, pn);
            if (z + 1 < z1) {
                tile[buf ^ 1][ly + 1][lx + 1] = next;
                stage(S, buf ^ 1);
            }
            if (own && z < z1) {
                const uint32_t i = cidx(g, x, y, z);
                float wv = 0.0f;
                if (d != 0u)
                    wv = __fmaf_rn((float)d, centre,
                                   -((((tile[buf][ly + 1][lx] + tile[buf][ly + 1][lx + 2]) +
                                       (tile[buf][ly][lx + 1] + tile[buf][ly + 2][lx + 1])) +
                                      below) +
                                     next));
                v.w_out[i] = wv;
                del = __fma_rn((double)wv, (double)centre, del);
                wp = __fma_rn((double)wv, (double)pc, wp);
            }
            const uint8_t dn = S.c.d;
            load_plane(z + 1 + kFusedDepth, S);
            __syncthreads();
            below = centre;
            centre = next;
            pc = pn;
            d = dn;
            buf ^= 1;
        };
        for (int z = z0; z < z1; z += kFusedDepth) {
#pragma unroll
            for (int j = 0; j < kFusedDepth; ++j) step(slots[j], z + j);
        }
    }
    double red[6]{rho, gam, del, wp, su, sp};
    block_sum<6>(red, lds);
    if (threadIdx.x == 0) {
        out_parts[kPartRR * kPcgMaxBlocks + lb] = red[0];
        out_parts[kPartRZ * kPcgMaxBlocks + lb] = red[1];
        out_parts[kPartPQ * kPcgMaxBlocks + lb] = red[2];
        out_parts[kPartWP * kPcgMaxBlocks + lb] = red[3];
        out_parts[kPartSU * kPcgMaxBlocks + lb] = red[4];
        out_parts[kPartSP * kPcgMaxBlocks + lb] = red[5];
    }
}

}  // namespace ds
