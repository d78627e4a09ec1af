// pcg.hip — Jacobi-preconditioned conjugate gradients on the smoke solver's pressure matrix.
//
// The matrix is the one the reference's red-black Gauss-Seidel iterates on (pressure_neighbor /
// rbgs_kernel, solvers/smoke/operators.cu:852-902): for an unknown i, A_ii = number of connected
// fluid neighbours (periodic wrap included) + number of fixed-value faces, A_ij = -1 for each
// connected neighbour that is itself an unknown; anchors and solids are not unknowns (p = 0).
// The per-cell connectivity code holds A_ii (0 = not an unknown). Every work vector is zero off
// the unknowns, so a stencil sums its in-domain neighbours without looking at their codes.
//
// Precision: FP32 vectors and stencils; FP64 dot products, norms and true residuals. The
// solution is a compensated FP32 pair x = hi + lo (the rounding of x alone to FP32 leaves a
// residual above 1e-6 ||b|| on smooth right-hand sides). Stopping rule ||b - A x|| <= tol ||b||
// with the true residual recomputed every `replacement_interval` iterations and at exit.
//
// Two recurrences (ds_config.cg_variant):
//   single reduction (default, pcg_fused.hpp): one fused kernel per iteration;
//   two reductions (pcg_plane.hpp): per iteration
//     direction: p_k = r/d + beta p_{k-1}, q = A p_k, partial <p_k, q>   (reads r, p_{k-1}; writes p_k)
//     update:    q = A p_k again, r -= alpha q, x += alpha p_k, partials <r,z>, <r,r>
// Each consumer finishes its producer's per-block FP64 partials in a fixed order, so every
// block derives bit-identical scalars and the iteration is deterministic without atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "context.hpp"
#include "pcg_fused.hpp"
#include "pcg_plane.hpp"
#include "pcg_resident.hpp"
#include "pcg_s2.hpp"
#include "rhs.hpp"

namespace ds {

namespace {

// Contiguous cell range of this (XCD-remapped) block; the chunk is a whole number of
// kPcgUnroll * blockDim cells, so every thread runs the same unrolled trip count.
__device__ __forceinline__ void block_range(uint32_t n, uint32_t& lb, uint32_t& start, uint32_t& end) {
    lb = xcd_block();
    const uint32_t step = kPcgUnroll * blockDim.x;
    uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    chunk = (chunk + step - 1) / step * step;
    start = min(n, lb * chunk);
    end = min(n, start + chunk);
}

// The 6 neighbours of cell i = (x, y, z) as clamped indices and 0/1 weights: an out-of-domain
// direction (non-periodic axis) points at i itself with weight 0, a periodic axis wraps. All
// loads stay in bounds and the gather has no branches, so the loads of several cells issue
// back to back.
struct Stencil {
    uint32_t j[6];
    float w[6];
};

__device__ __forceinline__ Stencil stencil(const Geo& g, uint32_t per, int x, int y, int z, uint32_t i) {
    const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
    const bool px = per & 1u, py = per & 2u, pz = per & 4u;
    Stencil s;
    s.j[0] = x > 0 ? i - 1u : (px ? i + (uint32_t)(g.nx - 1) : i);
    s.w[0] = (x > 0 || px) ? 1.0f : 0.0f;
    s.j[1] = x < g.nx - 1 ? i + 1u : (px ? i - (uint32_t)(g.nx - 1) : i);
    s.w[1] = (x < g.nx - 1 || px) ? 1.0f : 0.0f;
    s.j[2] = y > 0 ? i - sy : (py ? i + (uint32_t)(g.ny - 1) * sy : i);
    s.w[2] = (y > 0 || py) ? 1.0f : 0.0f;
    s.j[3] = y < g.ny - 1 ? i + sy : (py ? i - (uint32_t)(g.ny - 1) * sy : i);
    s.w[3] = (y < g.ny - 1 || py) ? 1.0f : 0.0f;
    s.j[4] = z > 0 ? i - sz : (pz ? i + (uint32_t)(g.nz - 1) * sz : i);
    s.w[4] = (z > 0 || pz) ? 1.0f : 0.0f;
    s.j[5] = z < g.nz - 1 ? i + sz : (pz ? i - (uint32_t)(g.nz - 1) * sz : i);
    s.w[5] = (z < g.nz - 1 || pz) ? 1.0f : 0.0f;
    return s;
}

// Sum of v over the 6 neighbours of cell i = (x, y, z); periodic axes wrap (bit k of per).
template <typename T, typename F>
__device__ __forceinline__ T neighbour_sum(const Geo& g, uint32_t per, int x, int y, int z, uint32_t i, F&& v) {
    const Stencil s = stencil(g, per, x, y, z, i);
    T sum = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) sum += (T)s.w[k] * v(s.j[k]);
    return sum;
}

// Setup: b (computed or given), x = x0 (or 0), r = b - A x0 (FP64), partials ||b||^2, ||r||^2, <r,z>.
template <int Kind>
__global__ __launch_bounds__(kPcgThreads) void pcg_setup_kernel(Geo g, uint32_t per, const uint8_t* __restrict__ code,
                                                        const uint8_t* __restrict__ mask, RhsIn in,
                                                        const float* __restrict__ fixed_term, float* __restrict__ b,
                                                        const float* x0h, const float* x0l, uint32_t x_is_guess,
                                                        float* xh, float* xl, float* __restrict__ r,
                                                        double* __restrict__ partials, PcgStatus* status) {
    __shared__ double lds[3 * kMaxWaves];
    uint32_t lb, start, end;
    block_range(g.n, lb, start, end);
    double bb = 0.0, rr = 0.0, rz = 0.0;
#pragma unroll 2
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t d = code[i];
        int x, y, z;
        decode_cell(g, i, x, y, z);
        float bi = 0.0f;
        if (d != 0u) bi = Kind == kRhsGiven ? b[i] : cell_rhs<Kind>(g, in, mask, fixed_term, i, x, y, z);
        if (Kind != kRhsGiven || d == 0u) b[i] = bi;
        float ri = bi;
        if (x0h != nullptr) {
            if (d != 0u) {
                // loaded unconditionally (j is a cell of the grid), selected by the code
                auto xv = [&](uint32_t j) -> double {
                    const double v = (double)x0h[j] + (x0l ? (double)x0l[j] : 0.0);
                    return code[j] ? v : 0.0;
                };
                const double ax = (double)d * xv(i) - neighbour_sum<double>(g, per, x, y, z, i, xv);
                ri = (float)((double)bi - ax);
            }
            if (!x_is_guess) {
                xh[i] = d ? x0h[i] : 0.0f;
                xl[i] = d && x0l ? x0l[i] : 0.0f;
            }
        } else {
            xh[i] = 0.0f;
            xl[i] = 0.0f;
        }
        if (d == 0u) ri = 0.0f;
        r[i] = ri;
        const float zi = ri * inv_diag(d);
        bb += (double)bi * bi;
        rr += (double)ri * ri;
        rz += (double)ri * zi;
    }
    double v[3]{bb, rr, rz};
    block_sum<3>(v, lds);
    if (threadIdx.x == 0) {
        partials[kPartBB * kPcgMaxBlocks + lb] = v[0];
        partials[kPartRR * kPcgMaxBlocks + lb] = v[1];
        partials[kPartRZ * kPcgMaxBlocks + lb] = v[2];
        if (lb == 0) {
            status->stop = 0u;
            status->iterations = 0u;
            status->rejected = 0u;
            status->aborted = 0u;
        }
    }
}

// True residual r = b - A (x_hi + x_lo) in FP64; replaces the recursive residual.
__global__ __launch_bounds__(kPcgThreads) void pcg_true_residual_kernel(Geo g, uint32_t per, const uint8_t* __restrict__ code,
                                                                const float* __restrict__ b, const float* __restrict__ xh,
                                                                const float* __restrict__ xl, float* __restrict__ r,
                                                                double* __restrict__ partials, PcgStatus* status,
                                                                uint32_t respect_stop) {
    __shared__ double lds[3 * kMaxWaves];
    if (respect_stop && status->stop) return;
    uint32_t lb, start, end;
    block_range(g.n, lb, start, end);
    const uint32_t G = gridDim.x;
    double rr = 0.0, rz = 0.0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t d = code[i];
        if (d == 0u) continue;
        int x, y, z;
        decode_cell(g, i, x, y, z);
        auto xv = [&](uint32_t j) -> double { return (double)xh[j] + (double)xl[j]; };
        const double ax = (double)d * xv(i) - neighbour_sum<double>(g, per, x, y, z, i, xv);
        const double ri = (double)b[i] - ax;
        const float rf = (float)ri;
        r[i] = rf;
        rr += ri * ri;
        rz += (double)rf * (rf * inv_diag(d));
    }
    double v[2]{rr, rz};
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[kPartRR * kPcgMaxBlocks + lb] = v[0];
        partials[kPartRZ * kPcgMaxBlocks + lb] = v[1];
    }
}

// Residual replacement of the single-reduction form: r = b - A (x_hi + x_lo) and s = A p in FP64,
// rounded to FP32 (with_s = 0 when no direction exists yet); partial ||r||^2 into out_parts.
__global__ __launch_bounds__(kPcgThreads) void pcg_fused_replace_kernel(Geo g, uint32_t per, const uint8_t* __restrict__ code,
                                                                const float* __restrict__ b, const float* __restrict__ xh,
                                                                const float* __restrict__ xl, const float* __restrict__ p,
                                                                float* __restrict__ r, float* __restrict__ s, uint32_t with_s,
                                                                double* __restrict__ out_parts, PcgStatus* status,
                                                                uint32_t respect_stop) {
    __shared__ double lds[3 * kMaxWaves];
    if (respect_stop && status->stop) return;
    uint32_t lb, start, end;
    block_range(g.n, lb, start, end);
    double rr = 0.0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t d = code[i];
        if (d == 0u) continue;
        int x, y, z;
        decode_cell(g, i, x, y, z);
        auto xv = [&](uint32_t j) -> double { return (double)xh[j] + (double)xl[j]; };
        const double ri = (double)b[i] - ((double)d * xv(i) - neighbour_sum<double>(g, per, x, y, z, i, xv));
        r[i] = (float)ri;
        rr += ri * ri;
        if (with_s) {
            auto pv = [&](uint32_t j) -> double { return code[j] ? (double)p[j] : 0.0; };
            s[i] = (float)((double)d * pv(i) - neighbour_sum<double>(g, per, x, y, z, i, pv));
        }
    }
    double v[1]{rr};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) out_parts[kPartRR * kPcgMaxBlocks + lb] = v[0];
}

// x_hi + x_lo += x_d (compensated), x_d = 0: folds the single-reduction iterations' increment
// into the solution pair.
__global__ __launch_bounds__(kPcgThreads) void pcg_fold_kernel(uint32_t n, float* __restrict__ xh, float* __restrict__ xl,
                                                               float* __restrict__ xd, PcgStatus* status, uint32_t respect_stop) {
    if (respect_stop && status->stop) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = xd[i];
        if (d == 0.0f) continue;
        float h = xh[i], l = xl[i];
        compensated_add(h, l, d);
        xh[i] = h;
        xl[i] = l;
        xd[i] = 0.0f;
    }
}

// s-step priming, pass 1 (one cell per thread, neighbours gathered): from the residual r and the
// direction P0 of set v (WithP; else p = z), P1 = T p, Z0 = z = D^-1 r and Z1 = T z, written into
// the same set (P0 too when p = z). T is evaluated exactly as in pcg_s2_kernel.
template <bool WithP>
__global__ __launch_bounds__(kBlock) void pcg_s2_prime_kernel(Geo g, uint32_t per, const uint8_t* __restrict__ code,
                                                              const float* __restrict__ r, S2Vecs v, PcgStatus* status,
                                                              uint32_t respect_stop) {
    if (respect_stop && status->stop) return;
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const uint32_t d = code[i];
    float* P0 = v.o[0];
    if (d == 0u) {
        if (!WithP) P0[i] = 0.0f;
        v.o[1][i] = 0.0f;
        v.o[3][i] = 0.0f;
        v.o[4][i] = 0.0f;
        return;
    }
    int x, y, z;
    decode_cell(g, i, x, y, z);
    const Stencil st = stencil(g, per, x, y, z, i);
    auto zval = [&](uint32_t j) { return r[j] * inv_diag(code[j]); };
    float zn[6], pn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        zn[q] = st.w[q] * zval(st.j[q]);
        pn[q] = WithP ? st.w[q] * P0[st.j[q]] : zn[q];
    }
    const float df = (float)d, inv = inv_diag(d);
    const float zc = zval(i);
    const float pc = WithP ? P0[i] : zc;
    const float tz = s2_apply(zc, zn[0], zn[1], zn[2], zn[3], zn[4], zn[5], df, inv);
    const float tp = WithP ? s2_apply(pc, pn[0], pn[1], pn[2], pn[3], pn[4], pn[5], df, inv) : tz;
    if (!WithP) P0[i] = zc;
    v.o[1][i] = tp;
    v.o[3][i] = zc;
    v.o[4][i] = tz;
}

// s-step priming, pass 2: P2 = T P1 and the partial sums of the basis' D-weighted Gram matrix and
// of rho = ||r||^2 (the residual that pass 1 read), in the s-step kernel's slot order.
__global__ __launch_bounds__(kPcgThreads) void pcg_s2_gram_kernel(Geo g, uint32_t per, const uint8_t* __restrict__ code,
                                                                  const float* __restrict__ r, S2Vecs v, double* __restrict__ out_parts,
                                                                  PcgStatus* status, uint32_t respect_stop) {
    __shared__ double lds[16 * kMaxWaves];
    if (respect_stop && status->stop) return;
    uint32_t lb, start, end;
    block_range(g.n, lb, start, end);
    double acc[kS2Parts];
#pragma unroll
    for (int q = 0; q < kS2Parts; ++q) acc[q] = 0.0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t d = code[i];
        float P2 = 0.0f;
        if (d != 0u) {
            int x, y, z;
            decode_cell(g, i, x, y, z);
            const Stencil st = stencil(g, per, x, y, z, i);
            float qn[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) qn[q] = st.w[q] * v.o[1][st.j[q]];
            P2 = s2_apply(v.o[1][i], qn[0], qn[1], qn[2], qn[3], qn[4], qn[5], (float)d, inv_diag(d));
        }
        v.o[2][i] = P2;
        if (d == 0u) continue;
        const float nv[5]{v.o[0][i], v.o[1][i], P2, v.o[3][i], v.o[4][i]};
        const double dd = (double)d;
        double q[5], wq[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            q[m] = (double)nv[m];
            wq[m] = dd * q[m];
        }
        int sidx = 0;
#pragma unroll
        for (int a = 0; a < 5; ++a)
#pragma unroll
            for (int c = a; c < 5; ++c) {
                acc[sidx] = __fma_rn(wq[a], q[c], acc[sidx]);
                ++sidx;
            }
        acc[kS2Rho] = __fma_rn((double)r[i], (double)r[i], acc[kS2Rho]);
    }
    s2_block_partials(acc, lds, out_parts + 16u * lb);
}

// One block: reduce the ||b||^2 (optional) and ||r||^2 partials into the status.
__global__ __launch_bounds__(256) void pcg_finalize_kernel(const double* __restrict__ partials, uint32_t G,
                                                           PcgStatus* status, uint32_t with_bb) {
    __shared__ double lds[3 * kMaxWaves];
    const double* parts[2]{partials + kPartBB * kPcgMaxBlocks, partials + kPartRR * kPcgMaxBlocks};
    double red[2];
    reduce_partials<2>(parts, G, red, lds);
    if (threadIdx.x == 0) {
        if (with_bb) status->bb = red[0];
        status->rr = red[1];
    }
}

__global__ void pcg_clear_stop_kernel(PcgStatus* status) {
    if (threadIdx.x == 0) status->stop = 0u;
}

// dst = src with one 4-byte load and one 4-byte store per lane and iteration, the access width of
// the PCG iteration kernels.
__global__ __launch_bounds__(256) void stream_copy_kernel(const float* __restrict__ src, float* __restrict__ dst, uint64_t count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void apply_matrix_kernel(Geo g, uint32_t per, const uint8_t* __restrict__ code,
                                                           const float* __restrict__ x, float* __restrict__ y) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const uint32_t d = code[i];
    if (d == 0u) {
        y[i] = 0.0f;
        return;
    }
    int xx, yy, zz;
    decode_cell(g, i, xx, yy, zz);
    const double s = (double)d * x[i] -
                     neighbour_sum<double>(g, per, xx, yy, zz, i, [&](uint32_t j) { return code[j] ? (double)x[j] : 0.0; });
    y[i] = (float)s;
}

uint32_t periodic_bits(const ds_context* c);

// Launches the direction kernel of iteration k (k == 0 has no previous direction).
void launch_direction(ds_context* c, uint32_t per, double tol2, uint32_t k, bool first) {
    PcgWork& w = c->pcg;
    float* p_new = w.p[k & 1u];
    const float* p_old = w.p[(k + 1u) & 1u];
    if (k == 0)
        pcg_direction_kernel<false><<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, w.tiling, c->code, w.r, p_old, p_new, w.partials,
                                                                         w.hist, w.status, tol2, k, first ? 1u : 0u);
    else
        pcg_direction_kernel<true><<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, w.tiling, c->code, w.r, p_old, p_new, w.partials,
                                                                        w.hist, w.status, tol2, k, first ? 1u : 0u);
}

uint32_t periodic_bits(const ds_context* c) {
    return (c->pressure_periodic[0] ? 1u : 0u) | (c->pressure_periodic[1] ? 2u : 0u) | (c->pressure_periodic[2] ? 4u : 0u);
}

void launch_setup(ds_context* c, int kind, const RhsIn& in, const float* x0h, const float* x0l, bool x_is_guess,
                  float* xh, float* xl) {
    const uint32_t per = periodic_bits(c);
    PcgWork& w = c->pcg;
    switch (kind) {
        case kRhsGiven:
            pcg_setup_kernel<kRhsGiven><<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, c->mask, in, c->fixed_term, w.b,
                                                                             x0h, x0l, x_is_guess, xh, xl, w.r, w.partials, w.status);
            break;
        case kRhsForward:
        case kRhsTangent:
            pcg_setup_kernel<kRhsForward><<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(
                c->geo, per, c->code, c->mask, in, kind == kRhsTangent ? nullptr : c->fixed_term, w.b, x0h, x0l, x_is_guess, xh, xl, w.r,
                w.partials, w.status);
            break;
        default:
            pcg_setup_kernel<kRhsAdjoint><<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, c->mask, in, c->fixed_term,
                                                                               w.b, x0h, x0l, x_is_guess, xh, xl, w.r, w.partials, w.status);
            break;
    }
    check_launch("pcg_setup_kernel");
}

PcgStatus read_status(ds_context* c) {
    DS_HIP(hipMemcpyAsync(&c->pcg.host_status[0], c->pcg.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, c->stream));
    DS_HIP(hipStreamSynchronize(c->stream));
    return c->pcg.host_status[0];
}

template <int Mode>
void launch_fused(ds_context* c, uint32_t per, const FusedVecs& v, const double* in_parts, const double* bb_parts,
                  double* out_parts, double tol2, uint32_t k, uint32_t check_only, uint32_t respect_stop, uint32_t with_ps = 0) {
    PcgWork& w = c->pcg;
    pcg_fused_kernel<Mode><<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, w.tiling, c->code, v, in_parts, bb_parts, out_parts,
                                                                    w.hist, w.status, tol2, k, check_only, respect_stop, with_ps);
}

// One cooperative launch of iterations k0 .. k0+count-1 of the single-reduction form, reading set
// `set` (every block must be co-resident: pcg_resident_prepare checked the occupancy).
void launch_resident(ds_context* c, uint32_t per, int set, double* const parts[2], const double* bb_parts, double tol2, uint32_t k0,
                     uint32_t count) {
    PcgWork& w = c->pcg;
    fill_zero(c, reinterpret_cast<float*>(w.resident_sync), 2);
    Geo g = c->geo;
    PlaneTiling pt = w.tiling;
    const uint8_t* code = c->code;
    ResidentVecs v{w.r, w.r_alt, w.wv[0], w.wv[1], w.sv[0], w.sv[1], w.p[0], w.xd};
    double* p0 = parts[0];
    double* p1 = parts[1];
    double* hist = w.hist;
    PcgStatus* status = w.status;
    ResidentSync* sync = reinterpret_cast<ResidentSync*>(w.resident_sync);
    uint32_t set0 = (uint32_t)set;
    void* args[]{&g, &per, &pt, &code, &v, &p0, &p1, &bb_parts, &hist, &status, &sync, &tol2, &k0, &count, &set0};
    DS_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&pcg_resident_kernel<kResidentZ>), dim3(w.blocks), dim3(kPcgThreads),
                                      args, 0, c->stream));
}

// Single-reduction iteration from the state left by launch_setup (r in set 0, BB/RR/RZ partials
// in partial set 0) until the true residual meets tol or max_it.
ds_solve_info pcg_iterate_fused(ds_context* c, float* xh, float* xl, double tol, uint32_t max_it, bool warm, bool rejected) {
    PcgWork& w = c->pcg;
    const uint32_t per = periodic_bits(c);
    const double tol2 = tol * tol;
    const uint32_t R = c->replacement_interval;
    float* rv[2]{w.r, w.r_alt};
    double* parts[2]{w.partials, w.partials + (size_t)kPartCount * kPcgMaxBlocks};
    auto vecs = [&](int set, bool prime) {
        FusedVecs v;
        v.r = rv[set];
        v.w = w.wv[set];
        v.s = w.sv[set];
        v.r_out = rv[set ^ 1];
        v.w_out = prime ? w.wv[set] : w.wv[set ^ 1];
        v.s_out = w.sv[set ^ 1];
        v.p = w.p[0];
        v.xd = w.xd;
        return v;
    };
    // with_ps: a previous direction exists (k > 0), so the denominator products are formed too
    auto prime = [&](int set, uint32_t respect_stop, bool with_ps) {
        FusedVecs v = vecs(set, true);
        v.s = w.sv[set];   // s_{i-1}: the replaced s of this set
        launch_fused<kFusedPrime>(c, per, v, nullptr, nullptr, parts[set], tol2, 0u, 0u, respect_stop, with_ps ? 1u : 0u);
    };
    auto fold = [&](uint32_t respect_stop) {
        pcg_fold_kernel<<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo.n, xh, xl, w.xd, w.status, respect_stop);
    };
    auto replace = [&](int set, bool with_s, uint32_t respect_stop) {
        fold(respect_stop);
        pcg_fused_replace_kernel<<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, w.b, xh, xl, w.p[0], rv[set], w.sv[set],
                                                                          with_s ? 1u : 0u, parts[set], w.status, respect_stop);
    };
    ds_solve_info info{};
    info.warm_started = warm ? 1u : 0u;
    info.guess_rejected = rejected ? 1u : 0u;
    int set = 0;
    fill_zero(c, w.xd, c->geo.n);
    prime(0, 0u, false);
    check_launch("pcg_fused_kernel (prime)");
    uint32_t k = 0;
    bool first = true;
    uint32_t restarts = 0;
    Profiler& prof = c->profiler;
    std::vector<std::pair<size_t, uint32_t>> resident_recs;   // (profiler record, first iteration)
    while (true) {
        uint32_t batch = 1;
        int slot = 0, pending = -1;
        bool stop_seen = false;
        while (k < max_it && !stop_seen) {
            const uint32_t nb = std::min(batch, max_it - k);
            if (w.resident) {
                // one cooperative launch per replacement interval; `batch` of them per status check
                for (uint32_t j = 0; j < nb && k < max_it; ++j) {
                    const uint32_t count = std::min(R - k % R, max_it - k);
                    hipEvent_t e0 = nullptr, e1 = nullptr;
                    if (prof.every && prof.used + 2 <= 6000) {
                        e0 = prof.take();
                        e1 = prof.take();
                    }
                    if (e0) DS_HIP(hipEventRecord(e0, c->stream));
                    launch_resident(c, per, set, parts, first ? parts[0] : nullptr, tol2, k, count);
                    first = false;
                    if (e1) {
                        DS_HIP(hipEventRecord(e1, c->stream));
                        prof.recs.push_back({3, e0, e1, count});
                        resident_recs.push_back({prof.recs.size() - 1, k});
                    }
                    if (prof.every) prof.launches[3] += count;
                    k += count;
                    if (count & 1u) set ^= 1;
                    if (k % R == 0u && k < max_it) {
                        replace(set, true, 1u);
                        prime(set, 1u, true);
                    }
                }
                check_launch("pcg_resident_kernel");
            } else
            for (uint32_t j = 0; j < nb; ++j, ++k) {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (prof.every && k % prof.every == 0 && prof.used + 2 <= 6000) {
                    e0 = prof.take();
                    e1 = prof.take();
                }
                if (e0) DS_HIP(hipEventRecord(e0, c->stream));
                const FusedVecs v = vecs(set, false);
                const double* bbp = first ? parts[0] : nullptr;
                if (k == 0)
                    launch_fused<kFusedFirst>(c, per, v, parts[set], bbp, parts[set ^ 1], tol2, k, 0u, 1u);
                else
                    launch_fused<kFusedNext>(c, per, v, parts[set], bbp, parts[set ^ 1], tol2, k, 0u, 1u);
                first = false;
                if (e1) {
                    DS_HIP(hipEventRecord(e1, c->stream));
                    prof.recs.push_back({2, e0, e1});
                }
                if (prof.every) prof.launches[2]++;
                set ^= 1;
                if ((k + 1u) % R == 0u && k + 1u < max_it) {
                    replace(set, true, 1u);
                    prime(set, 1u, true);
                }
            }
            check_launch("pcg_fused_kernel");
            DS_HIP(hipMemcpyAsync(&w.host_status[slot], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, c->stream));
            DS_HIP(hipEventRecord(w.batch_event[slot], c->stream));
            if (pending >= 0) {
                DS_HIP(hipEventSynchronize(w.batch_event[pending]));
                if (w.host_status[pending].stop) stop_seen = true;
            }
            pending = slot;
            slot ^= 1;
            batch = std::min<uint32_t>(batch * 2u, 32u);
        }
        PcgStatus st = read_status(c);
        if (st.aborted) {
            // a grid barrier wait ran past its bound: the vectors are incomplete. Later solves of
            // this context use one launch per iteration.
            w.resident = false;
            throw Failure{DS_ERROR_RUNTIME, "pcg_resident_kernel: a grid barrier wait exceeded its bound"};
        }
        if (!st.stop && k >= max_it) {
            // one more convergence test on the last iteration's residual
            if (k == 0)
                launch_fused<kFusedFirst>(c, per, vecs(set, false), parts[set], first ? parts[0] : nullptr, parts[set ^ 1], tol2, k, 1u, 1u);
            else
                launch_fused<kFusedNext>(c, per, vecs(set, false), parts[set], first ? parts[0] : nullptr, parts[set ^ 1], tol2, k, 1u, 1u);
            first = false;
            check_launch("pcg_fused_kernel (check)");
            st = read_status(c);
        }
        // the launches queued after the stopping iteration returned at once: `set` is one flip per
        // launch ahead; the stopping iteration read set (done & 1) relative to this pass's start
        const uint32_t done = st.stop ? st.iterations : max_it;
        set = (int)((uint32_t)set ^ ((k - done) & 1u));
        k = done;
        for (const auto& rr : resident_recs) {   // the iterations each timed resident launch performed
            Profiler::Rec& rec = prof.recs[rr.first];
            rec.units = done > rr.second ? std::min(rec.units, done - rr.second) : 0u;
        }
        resident_recs.clear();
        replace(set, done > 0, 0u);
        pcg_finalize_kernel<<<1, kBlock, 0, c->stream>>>(parts[set], w.blocks, w.status, 0u);
        check_launch("pcg_fused_replace_kernel");
        const PcgStatus fin = read_status(c);
        info.iterations = done;
        info.rhs_norm = std::sqrt(fin.bb);
        info.final_residual = std::sqrt(std::max(0.0, fin.rr));
        const bool finite = std::isfinite(fin.rr);
        const bool ok = finite && fin.rr <= tol2 * fin.bb;
        info.converged = ok ? 1u : 0u;
        if (!finite) info.final_residual = fin.rr;
        if (ok || !finite || !st.stop || done >= max_it || ++restarts > 8) break;
        // the recursive residual met the tolerance but the true one does not: continue from the
        // true residual (r and s replaced above) with the same search direction
        prime(set, 0u, done > 0);
        pcg_clear_stop_kernel<<<1, 64, 0, c->stream>>>(w.status);
        check_launch("pcg_clear_stop_kernel");
    }
    return info;
}

// s-step iteration (DS_CG_S_STEP) from the state left by launch_setup (r, BB partials in set 0 of
// w.partials) until the true residual meets tol or max_it. Each launch performs two iterations (one
// when a single iteration is left before max_it); every replacement interval (rounded up to even)
// folds the increment, recomputes the true residual and primes a new basis with the same direction.
ds_solve_info pcg_iterate_s2(ds_context* c, float* xh, float* xl, double tol, uint32_t max_it, bool warm, bool rejected) {
    PcgWork& w = c->pcg;
    const uint32_t per = periodic_bits(c);
    const double tol2 = tol * tol;
    const uint32_t R = (c->replacement_interval + 1u) & ~1u;
    double* parts[2]{w.s2_parts, w.s2_parts + (size_t)16 * kPcgMaxBlocks};
    auto vecs = [&](int set) {
        S2Vecs v;
        for (int m = 0; m < 5; ++m) {
            v.v[m] = w.s2v[set][m];
            v.o[m] = w.s2v[set ^ 1][m];
        }
        v.xd = w.xd;
        return v;
    };
    auto prime = [&](int set, bool with_p, uint32_t respect_stop) {
        S2Vecs v = vecs(set ^ 1);   // v.o = set
        const uint32_t nb = xcd_blocks_for(c->geo.n);
        if (with_p)
            pcg_s2_prime_kernel<true><<<nb, kBlock, 0, c->stream>>>(c->geo, per, c->code, w.r, v, w.status, respect_stop);
        else
            pcg_s2_prime_kernel<false><<<nb, kBlock, 0, c->stream>>>(c->geo, per, c->code, w.r, v, w.status, respect_stop);
        pcg_s2_gram_kernel<<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, w.r, v, parts[set], w.status, respect_stop);
    };
    auto fold = [&](uint32_t respect_stop) {
        pcg_fold_kernel<<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo.n, xh, xl, w.xd, w.status, respect_stop);
    };
    auto true_residual = [&](uint32_t respect_stop) {
        pcg_true_residual_kernel<<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, w.b, xh, xl, w.r, w.partials, w.status,
                                                                          respect_stop);
    };
    auto launch = [&](int set, uint32_t k, uint32_t steps, uint32_t check_only) {
        const S2Sets sets{w.s2set[set], w.s2set[set ^ 1], w.xd};
        pcg_s2_kernel<<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, w.tiling, c->code, sets, parts[set], w.blocks,
                                                               parts[set ^ 1], w.status, tol2, k, steps, check_only);
    };
    ds_solve_info info{};
    info.warm_started = warm ? 1u : 0u;
    info.guess_rejected = rejected ? 1u : 0u;
    fill_zero(c, w.xd, c->geo.n);
    // ||b||^2 (and ||r0||^2) from the set-up's partials into the status
    pcg_finalize_kernel<<<1, kBlock, 0, c->stream>>>(w.partials, w.blocks, w.status, 1u);
    int set = 0;
    prime(0, false, 0u);
    check_launch("pcg_s2_prime_kernel");
    uint32_t k = 0;
    uint32_t restarts = 0;
    Profiler& prof = c->profiler;
    // (iteration at launch, set read) of this pass's launches, to find the state a stop left
    std::vector<std::pair<uint32_t, int>> launched;
    while (true) {
        uint32_t batch = 1;
        int slot = 0, pending = -1;
        bool stop_seen = false;
        launched.clear();
        while (k < max_it && !stop_seen) {
            for (uint32_t j = 0; j < batch && k < max_it; ++j) {
                const uint32_t steps = std::min<uint32_t>(2u, max_it - k);
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (prof.every && (k / 2u) % prof.every == 0 && prof.used + 2 <= 6000) {
                    e0 = prof.take();
                    e1 = prof.take();
                }
                if (e0) DS_HIP(hipEventRecord(e0, c->stream));
                launch(set, k, steps, 0u);
                if (e1) {
                    DS_HIP(hipEventRecord(e1, c->stream));
                    prof.recs.push_back({4, e0, e1});
                }
                if (prof.every) prof.launches[4]++;
                launched.push_back({k, set});
                k += steps;
                set ^= 1;
                if (k % R == 0u && k < max_it) {
                    fold(1u);
                    true_residual(1u);
                    prime(set, true, 1u);
                }
            }
            check_launch("pcg_s2_kernel");
            DS_HIP(hipMemcpyAsync(&w.host_status[slot], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, c->stream));
            DS_HIP(hipEventRecord(w.batch_event[slot], c->stream));
            if (pending >= 0) {
                DS_HIP(hipEventSynchronize(w.batch_event[pending]));
                if (w.host_status[pending].stop) stop_seen = true;
            }
            pending = slot;
            slot ^= 1;
            batch = std::min<uint32_t>(batch * 2u, 16u);
        }
        PcgStatus st = read_status(c);
        if (!st.stop && k >= max_it) {
            // one more convergence test on the last launch's residual
            launch(set, k, 2u, 1u);
            check_launch("pcg_s2_kernel (check)");
            st = read_status(c);
            launched.push_back({k, set});
        }
        // the state a stop left is the set its launch read (the launches queued after it returned
        // at once)
        const uint32_t done = st.stop ? st.iterations : k;
        for (const auto& l : launched)
            if (l.first == done) set = l.second;
        k = done;
        fold(0u);
        true_residual(0u);
        pcg_finalize_kernel<<<1, kBlock, 0, c->stream>>>(w.partials, w.blocks, w.status, 0u);
        check_launch("pcg_true_residual_kernel");
        const PcgStatus fin = read_status(c);
        info.iterations = done;
        info.rhs_norm = std::sqrt(fin.bb);
        info.final_residual = std::sqrt(std::max(0.0, fin.rr));
        const bool finite = std::isfinite(fin.rr);
        const bool ok = finite && fin.rr <= tol2 * fin.bb;
        info.converged = ok ? 1u : 0u;
        if (!finite) info.final_residual = fin.rr;
        if (ok || !finite || !st.stop || done >= max_it || ++restarts > 8) break;
        // the recursive residual met the tolerance but the true one does not: continue from the
        // true residual (computed above) with the same direction
        prime(set, done > 0, 0u);
        pcg_clear_stop_kernel<<<1, 64, 0, c->stream>>>(w.status);
        check_launch("pcg_clear_stop_kernel");
    }
    return info;
}

// Multigrid-preconditioned iteration (DS_CG_MULTIGRID, pcg_mg.hip) from the state left by
// launch_setup (r; BB, RR partials) until the true residual meets tol or max_it. Iteration k: the
// V-cycle z = M^-1 r, the direction kernel (stopping test, p, <p, A p>), the update kernel (which
// adds to the solution pair directly); the true residual is recomputed every replacement interval.
ds_solve_info pcg_iterate_mg(ds_context* c, float* xh, float* xl, double tol, uint32_t max_it, bool warm, bool rejected) {
    PcgWork& w = c->pcg;
    const uint32_t per = periodic_bits(c);
    const double tol2 = tol * tol;
    const uint32_t R = c->replacement_interval;
    float* z = c->mg.z;
    auto true_residual = [&](uint32_t respect_stop) {
        pcg_true_residual_kernel<<<w.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, w.b, xh, xl, w.r, w.partials, w.status,
                                                                          respect_stop);
    };
    ds_solve_info info{};
    info.warm_started = warm ? 1u : 0u;
    info.guess_rejected = rejected ? 1u : 0u;
    // ||b||^2 (and ||r0||^2) from the set-up's partials into the status; iteration 0 next
    pcg_finalize_kernel<<<1, kBlock, 0, c->stream>>>(w.partials, w.blocks, w.status, 1u);
    mg_set_iteration(c, ~0u);
    uint32_t k = 0;
    uint32_t restarts = 0;
    Profiler& prof = c->profiler;
    while (true) {
        // Batches are queued one ahead of the status the host reads, and the iterations queued
        // behind the stop cost their launches: the first batch runs to a little before the
        // previous solve's count (a V-cycle's convergence rate hardly changes between the solves
        // of a trajectory), the later ones two iterations at a time.
        uint32_t batch = w.mg_expect > 4u ? w.mg_expect - 3u : 2u;
        int slot = 0, pending = -1;
        bool stop_seen = false;
        PcgStatus st{};
        while (k < max_it && !stop_seen) {
            const uint32_t nb = std::min(batch, max_it - k);
            for (uint32_t j = 0; j < nb; ++j, ++k) {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (prof.every && k % prof.every == 0 && prof.used + 2 <= 6000) {
                    e0 = prof.take();
                    e1 = prof.take();
                }
                if (e0) DS_HIP(hipEventRecord(e0, c->stream));
                mg_iteration(c, xh, xl, tol2);
                if (e1) {
                    DS_HIP(hipEventRecord(e1, c->stream));
                    prof.recs.push_back({5, e0, e1});
                }
                if (prof.every) prof.launches[5]++;
                if ((k + 1u) % R == 0u && k + 1u < max_it) true_residual(1u);
            }
            check_launch("pcg multigrid iteration");
            DS_HIP(hipMemcpyAsync(&w.host_status[slot], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, c->stream));
            DS_HIP(hipEventRecord(w.batch_event[slot], c->stream));
            if (pending >= 0) {
                DS_HIP(hipEventSynchronize(w.batch_event[pending]));
                if (w.host_status[pending].stop) {
                    stop_seen = true;
                    st = w.host_status[pending];
                }
            }
            pending = slot;
            slot ^= 1;
            batch = 2u;
        }
        // the stopping launch's status when the host saw it, else the last batch's
        if (!stop_seen) st = read_status(c);
        if (!st.stop && k >= max_it) {
            // one more stopping test on the last update's residual (its V-cycle gives the RZ
            // partials the test reads but does not use)
            mg_precondition(c, w.r, z);
            mg_direction(c, z, tol2, true);
            check_launch("mg_direction_kernel");
            st = read_status(c);
        }
        const uint32_t done = st.stop ? st.iterations : max_it;
        k = done;
        true_residual(0u);
        pcg_finalize_kernel<<<1, kBlock, 0, c->stream>>>(w.partials, w.blocks, w.status, 0u);
        check_launch("pcg_true_residual_kernel");
        const PcgStatus fin = read_status(c);
        if (st.stop) w.mg_expect = done;
        info.iterations = done;
        info.rhs_norm = std::sqrt(fin.bb);
        info.final_residual = std::sqrt(std::max(0.0, fin.rr));
        const bool finite = std::isfinite(fin.rr);
        const bool ok = finite && fin.rr <= tol2 * fin.bb;
        info.converged = ok ? 1u : 0u;
        if (!finite) info.final_residual = fin.rr;
        if (ok || !finite || !st.stop || done >= max_it || ++restarts > 8) break;
        // the recursive residual met the tolerance but the true one (now in r) does not: continue
        // from it with the previous direction discarded
        fill_zero(c, w.p[(k + 1u) & 1u], c->geo.n);
        mg_set_iteration(c, k - 1u);
        pcg_clear_stop_kernel<<<1, 64, 0, c->stream>>>(w.status);
        check_launch("pcg_clear_stop_kernel");
    }
    return info;
}

// Iterates from the state left by launch_setup until the true residual meets tol or max_it.
ds_solve_info pcg_iterate(ds_context* c, float* xh, float* xl, double tol, uint32_t max_it, bool warm, bool rejected) {
    PcgWork& w = c->pcg;
    const uint32_t per = periodic_bits(c);
    const double tol2 = tol * tol;
    const uint32_t R = c->replacement_interval;
    ds_solve_info info{};
    info.warm_started = warm ? 1u : 0u;
    info.guess_rejected = rejected ? 1u : 0u;
    uint32_t k = 0;
    bool first = true;
    uint32_t restarts = 0;
    PcgStatus st{};
    while (true) {
        uint32_t batch = 1;
        int slot = 0, pending = -1;
        bool stop_seen = false;
        while (k < max_it && !stop_seen) {
            const uint32_t nb = std::min(batch, max_it - k);
            for (uint32_t j = 0; j < nb; ++j, ++k) {
                Profiler& prof = c->profiler;
                hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
                if (prof.every && k % prof.every == 0 && prof.used + 3 <= 6000) {
                    e0 = prof.take();
                    e1 = prof.take();
                    e2 = prof.take();
                }
                if (e0) DS_HIP(hipEventRecord(e0, c->stream));
                launch_direction(c, per, tol2, k, first);
                first = false;
                if (e1) DS_HIP(hipEventRecord(e1, c->stream));
                pcg_update_kernel<<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, w.tiling, c->code, w.p[k & 1u], w.r, xh, xl, w.partials,
                                                                        w.hist, w.status, k);
                if (e2) {
                    DS_HIP(hipEventRecord(e2, c->stream));
                    prof.recs.push_back({0, e0, e1});
                    prof.recs.push_back({1, e1, e2});
                }
                if (prof.every) {
                    prof.launches[0]++;
                    prof.launches[1]++;
                }
                if ((k + 1u) % R == 0u && k + 1u < max_it)
                    pcg_true_residual_kernel<<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, w.b, xh, xl, w.r, w.partials,
                                                                                   w.status, 1u);
            }
            check_launch("pcg iteration");
            DS_HIP(hipMemcpyAsync(&w.host_status[slot], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, c->stream));
            DS_HIP(hipEventRecord(w.batch_event[slot], c->stream));
            if (pending >= 0) {
                DS_HIP(hipEventSynchronize(w.batch_event[pending]));
                if (w.host_status[pending].stop) stop_seen = true;
            }
            pending = slot;
            slot ^= 1;
            batch = std::min<uint32_t>(batch * 2u, 32u);
        }
        if (!stop_seen && k >= max_it) {
            // one more convergence test on the last update's residual
            launch_direction(c, per, tol2, k, first);
            first = false;
            check_launch("pcg_direction_kernel");
        }
        st = read_status(c);
        const uint32_t done = st.stop ? st.iterations : max_it;
        pcg_true_residual_kernel<<<c->pcg.blocks, kPcgThreads, 0, c->stream>>>(c->geo, per, c->code, w.b, xh, xl, w.r, w.partials,
                                                                       w.status, 0u);
        pcg_finalize_kernel<<<1, kBlock, 0, c->stream>>>(w.partials, c->pcg.blocks, w.status, 0u);
        check_launch("pcg_true_residual_kernel");
        const PcgStatus fin = read_status(c);
        info.iterations = done;
        info.rhs_norm = std::sqrt(fin.bb);
        info.final_residual = std::sqrt(std::max(0.0, fin.rr));
        const bool finite = std::isfinite(fin.rr);
        const bool ok = finite && fin.rr <= tol2 * fin.bb;
        info.converged = ok ? 1u : 0u;
        if (!finite) info.final_residual = fin.rr;
        if (ok || !finite || !st.stop || done >= max_it || ++restarts > 8) break;
        // the recursive residual met the tolerance but the true one does not: continue from the
        // true residual (residual replacement) with the same search direction
        pcg_clear_stop_kernel<<<1, 64, 0, c->stream>>>(w.status);
        check_launch("pcg_clear_stop_kernel");
        k = done;
        first = false;
    }
    return info;
}

}  // namespace

uint32_t pcg_periodic_bits(const ds_context* c) { return periodic_bits(c); }

void pcg_resident_prepare(ds_context* c) {
    PcgWork& w = c->pcg;
    w.resident = false;
    const Geo& g = c->geo;
    if (g.nx % kTX != 0 || g.ny % kTY != 0 || w.tiling.zc > (uint32_t)kResidentZ || w.blocks > kPcgThreads) return;
    int coop = 0, cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device) != hipSuccess || !coop) return;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) return;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&pcg_resident_kernel<kResidentZ>),
                                                     kPcgThreads, 0) != hipSuccess)
        return;
    if ((uint64_t)per_cu * (uint64_t)cus < w.blocks) return;
    w.resident_sync = reinterpret_cast<uint32_t*>(device_alloc<float>(c, 2));
    fill_zero(c, reinterpret_cast<float*>(w.resident_sync), 2);
    w.resident = true;
}

// Solves A x = b for the given right-hand-side kind. Returns the solve info; x = (xh, xl).
ds_solve_info pcg_solve_kind(ds_context* c, int kind, const float* const u[3], const float* const extra[3], float* xh,
                             float* xl, const float* x0h, const float* x0l, double tol, uint32_t max_it, bool correct) {
    if (c->config.projection_solver == DS_PROJECTION_RBGS && kind != kRhsGiven) return rbgs_solve_kind(c, kind, u, extra, xh, xl);
    RhsIn in{};
    for (int a = 0; a < 3; ++a) {
        in.u[a] = u ? u[a] : nullptr;
        in.extra[a] = extra ? extra[a] : nullptr;
    }
    max_it = std::min(max_it, c->max_iterations);   // the per-iteration history is sized for this cap
    bool network_guess = false;
    // a solve with a given guess (the recompute pass's taped pressure) is not part of the sequence
    SolveWindow* win = x0h == nullptr ? window_of(c, kind) : nullptr;
    if (win && win->count > 0) {
        // the projection onto the previous solutions needs b: set-up without a guess, then again
        // from the projected guess with b given
        launch_setup(c, kind, in, nullptr, nullptr, false, xh, xl);
        x0h = window_guess(c, *win, c->pcg.b);
        x0l = nullptr;
        kind = kRhsGiven;
    } else if ((c->config.warm_start == DS_WARM_START_CNN || c->config.warm_start == DS_WARM_START_CNN_ALONE) &&
               (x0h == nullptr || (correct && c->config.warm_start == DS_WARM_START_CNN))) {
        // the network's guess: b first (set-up without a guess), then the set-up again from the guess;
        // a given guess (the sequence's) is the base the network corrects, combined to FP32 first
        const float* base = nullptr;
        if (x0h) {
            if (x0l) {
                combine_pair(c, x0h, x0l, c->guess);
                base = c->guess;
            } else {
                base = x0h;
            }
        }
        launch_setup(c, kind, in, nullptr, nullptr, false, xh, xl);
        x0h = unet_infer(c, c->pcg.b, base);
        x0l = nullptr;
        kind = kRhsGiven;
        network_guess = true;
    }
    const bool warm = x0h != nullptr;
    bool rejected = false;
    launch_setup(c, kind, in, x0h, x0l, x0h == xh, xh, xl);
    double rr0 = 0.0, bb = 0.0;
    if (warm) {
        // Reject a non-finite guess, and a given guess whose residual exceeds ||b|| (worse than
        // x0 = 0). The network's guess is scaled to its A-norm-optimal multiple, which never has
        // more energy error than x0 = 0 (the norm CG minimises), though its residual can be larger.
        pcg_finalize_kernel<<<1, kBlock, 0, c->stream>>>(c->pcg.partials, c->pcg.blocks, c->pcg.status, 1u);
        check_launch("pcg_finalize_kernel");
        const PcgStatus s = read_status(c);
        rr0 = s.rr;
        bb = s.bb;
        if (!std::isfinite(s.rr) || (!network_guess && !(s.rr <= s.bb))) {
            rejected = true;
            launch_setup(c, kRhsGiven, in, nullptr, nullptr, false, xh, xl);
        }
    }
    ds_solve_info info = c->pcg.variant == DS_CG_SINGLE_REDUCTION ? pcg_iterate_fused(c, xh, xl, tol, max_it, warm, rejected)
                         : c->pcg.variant == DS_CG_S_STEP        ? pcg_iterate_s2(c, xh, xl, tol, max_it, warm, rejected)
                         : c->pcg.variant == DS_CG_MULTIGRID     ? pcg_iterate_mg(c, xh, xl, tol, max_it, warm, rejected)
                                                                 : pcg_iterate(c, xh, xl, tol, max_it, warm, rejected);
    info.initial_residual = warm && !rejected ? std::sqrt(std::max(0.0, rr0)) : info.rhs_norm;
    (void)bb;
    if (win && info.converged) window_add(c, *win, xh, xl);
    return info;
}

ds_solve_info pcg_solve(ds_context* c, float* x_hi, float* x_lo, const float* x0, double tol, uint32_t max_it) {
    return pcg_solve_kind(c, kRhsGiven, nullptr, nullptr, x_hi, x_lo, x0, nullptr, tol, max_it);
}

void launch_apply_matrix(ds_context* c, const float* x, float* y) {
    apply_matrix_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, periodic_bits(c), c->code, x, y);
    check_launch("apply_matrix_kernel");
}

}  // namespace ds

using namespace ds;

extern "C" {

ds_status ds_pcg_solve(ds_context* c, const float* b, float* x, const float* x0, double tolerance, uint32_t max_iterations,
                       ds_solve_info* info) {
    if (!c || !b || !x) {
        set_error("ds_pcg_solve: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ds_solve_info si{};
    ds_status st = guard([&] {
        DS_HIP(hipMemcpyAsync(c->pcg.b, b, (size_t)c->geo.n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        const double tol = tolerance > 0.0 ? tolerance : c->tolerance;
        const uint32_t mi = max_iterations ? max_iterations : c->max_iterations;
        // solve into the context's pressure pair, then round hi + lo into x
        si = pcg_solve_kind(c, kRhsGiven, nullptr, nullptr, c->pressure_hi, c->pressure_lo, x0, nullptr, tol, mi);
        c->pressure_valid = false;
        c->pressure_prev_count = 0;
    });
    if (st != DS_OK) return st;
    st = guard([&] {
        combine_pair(c, c->pressure_hi, c->pressure_lo, x);
        DS_HIP(hipStreamSynchronize(c->stream));
    });
    if (info) *info = si;
    if (st != DS_OK) return st;
    if (!si.converged) {
        set_error(std::isfinite(si.final_residual) ? "ds_pcg_solve: not converged" : "ds_pcg_solve: non-finite residual");
        return std::isfinite(si.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    return DS_OK;
}

ds_status ds_profile_enable(ds_context* c, uint32_t sample_every) {
    if (!c) {
        set_error("ds_profile_enable: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    c->profiler.every = sample_every;
    for (auto& l : c->profiler.launches) l = 0;
    return DS_OK;
}

ds_status ds_profile_read(ds_context* c, ds_kernel_timing* out, uint32_t capacity, uint32_t* count) {
    if (!c || (capacity && !out)) {
        set_error("ds_profile_read: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        DS_HIP(hipStreamSynchronize(c->stream));
        Profiler& prof = c->profiler;
        double total[kProfiledKernels]{};
        uint64_t samples[kProfiledKernels]{};
        for (const Profiler::Rec& r : prof.recs) {
            if (r.units == 0) continue;   // a resident launch queued after convergence
            float ms = 0.0f;
            DS_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            total[r.kernel] += ms;
            samples[r.kernel] += r.units;
        }
        // the resident launch is reported per iteration: launches and samples count iterations
        const char* names[kProfiledKernels]{"pcg_direction_kernel", "pcg_update_kernel", "pcg_fused_kernel", "pcg_resident_kernel",
                                            "pcg_s2_kernel", "pcg_mg_iteration", "advect_velocity_vjp_kernel"};
        uint32_t n = 0;
        for (int k = 0; k < kProfiledKernels; ++k) {   // kernels that ran
            if (prof.launches[k] == 0 || n >= capacity) continue;
            std::snprintf(out[n].name, sizeof(out[n].name), "%s", names[k]);
            out[n].launches = prof.launches[k];
            out[n].samples = samples[k];
            out[n].total_ms = total[k];
            ++n;
        }
        if (count) *count = n;
        prof.recs.clear();
        prof.used = 0;
        for (auto& l : prof.launches) l = 0;
    });
}

uint64_t ds_warm_start_parameter_count(void) { return unet_parameter_count(); }

ds_status ds_warm_start_load(ds_context* c, const float* params, uint64_t count, double output_scale) {
    if (!c || !params) {
        set_error("ds_warm_start_load: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!(output_scale > 0.0) || !std::isfinite(output_scale)) {
        set_error("ds_warm_start_load: output_scale must be positive and finite");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { unet_load(c, params, count, output_scale); });
}

ds_status ds_warm_start_configure(ds_context* c, uint32_t levels) {
    if (!c) {
        set_error("ds_warm_start_configure: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { unet_configure(c, levels); });
}

ds_status ds_warm_start_infer(ds_context* c, const float* b, float* x) {
    if (!c || !b || !x) {
        set_error("ds_warm_start_infer: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        DS_HIP(hipMemcpyAsync(c->pcg.b, b, (size_t)c->geo.n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        RhsIn in{};
        launch_setup(c, kRhsGiven, in, nullptr, nullptr, false, c->pressure_hi, c->pressure_lo);   // ||b||^2 partials
        c->pressure_valid = false;
        c->pressure_prev_count = 0;
        const float* g = unet_infer(c, c->pcg.b, nullptr);
        DS_HIP(hipMemcpyAsync(x, g, (size_t)c->geo.n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        DS_HIP(hipStreamSynchronize(c->stream));
    });
}

ds_status ds_warm_start_guess(ds_context* c, const float* b, const float* base, float* x) {
    if (!c || !b || !x) {
        set_error("ds_warm_start_guess: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        const float* g = unet_infer(c, b, base);
        DS_HIP(hipMemcpyAsync(x, g, (size_t)c->geo.n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        DS_HIP(hipStreamSynchronize(c->stream));
    });
}

ds_status ds_apply_pressure_matrix(ds_context* c, const float* x, float* y) {
    if (!c || !x || !y) {
        set_error("ds_apply_pressure_matrix: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { launch_apply_matrix(c, x, y); });
}

ds_status ds_stream_copy(ds_context* c, const float* src, float* dst, uint64_t count) {
    if (!c || (count && (!src || !dst))) {
        set_error("ds_stream_copy: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        if (count == 0) return;
        stream_copy_kernel<<<2048, 256, 0, c->stream>>>(src, dst, count);
        check_launch("stream_copy_kernel");
    });
}

}  // extern "C"
