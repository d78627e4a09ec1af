// pcg_mg.hip — conjugate gradients preconditioned by an aggregation multigrid V-cycle
// (DS_CG_MULTIGRID): the same pressure system as the Jacobi forms (the reference solves it with
// Jacobi-preconditioned CG, solvers/smoke/operators.cu), to the same FP64 stopping test on the true
// residual, in a few tens of iterations instead of about a thousand at 128^3.
//
// Hierarchy: each coarser level aggregates 2 x 2 x 2 cells of the finer one (an odd extent leaves
// a last aggregate one cell wide; an extent of 1 stays 1). The prolongation P copies a coarse value
// to the aggregate's cells and the restriction is its transpose (the cells' sum), so the coarse
// operator A_c = P^T A P sums the fine couplings over the aggregates: a 7-point operator again, held
// as its diagonal and the per-axis couplings to the + neighbour, built once per context on the host
// from the connectivity code (periodic wraps included). The coarsest level (at most kMgCoarsest
// cells) is solved exactly with its inverse (FP64 Cholesky on the host, applied in FP64).
//
// V-cycle (symmetric, so a valid CG preconditioner): damped Jacobi sweeps from zero (two on level 0,
// one on the coarser levels), the residual restricted, the coarse correction scaled up (piecewise-
// constant aggregation under-corrects smooth errors; scaling it is the usual remedy), then as many
// damped Jacobi sweeps again. Values a sweep needs at the neighbours are formed on the fly: level 0
// costs four launches (pre-smoothing, restriction, two post sweeps), a coarser level two (down: the
// restriction of the residual after its pre-smoothing; up: prolongation plus its post sweep).
// Without the over-correction an FP64 CPU model of the 128^3 plume needs about twice the CG
// iterations (25 against 13 on V(2,2)); Jacobi-PCG needs 1015.
//
// The outer loop is textbook PCG with two reductions per iteration: the direction kernel forms
// p = z + beta p_old and <p, A p>; the update kernel adds alpha p to the solution increment, updates
// the residual and forms ||r||^2; the V-cycle's last kernel forms <r, z>. Every scalar comes from
// FP64 per-block partials that each consumer block reduces in the same order. Every kernel returns at
// once after the stopping test has fired, so the iterations queued behind it cost only launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "context.hpp"

namespace ds {

namespace {

// Level 0 runs V(2,2) (two damped Jacobi sweeps before and after the coarse correction), the coarser
// levels V(1,1): an FP64 CPU model of the 128^3 plume with random right-hand sides needs 14 CG
// iterations to 1e-6 for these parameters, 13 with V(2,2) on every level (four launches per level
// instead of two), 17 with V(1,1) everywhere.
constexpr float kOmega0 = 0.9f;   // level-0 Jacobi damping
constexpr float kOver0 = 1.8f;    // scale of the coarse correction prolonged to level 0
constexpr float kOmegaC = 0.9f;   // coarser levels
constexpr float kOverC = 1.5f;
constexpr uint32_t kMgThreads = 512;   // the kernels over contiguous ranges of level 0
constexpr uint32_t kMgUnroll = 4;      // cells per thread whose loads are in flight together there

__device__ __forceinline__ void mg_coords(const MgLevel& L, uint32_t i, int& x, int& y, int& z) {
    const uint32_t t = fdiv(i, L.dx);
    x = (int)(i - t * (uint32_t)L.nx);
    const uint32_t u = fdiv(t, L.dy);
    y = (int)(t - u * (uint32_t)L.ny);
    z = (int)u;
}

// omega / d for a level-0 diagonal count (the hardware reciprocal: every use of a cell's value is
// the same, so the smoother stays symmetric). Counts of 0 (not an unknown, or the zero pad) give
// omega; the values they scale are 0 there.
__device__ __forceinline__ float mg_wdinv(uint32_t d) { return kOmega0 * __builtin_amdgcn_rcpf((float)max(d, 1u)); }

// Level-0 neighbours of cell i = (x, y, z) in the order -x, +x, -y, +y, -z, +z. A periodic axis
// wraps (on an extent of 1 both neighbours are i itself, as in the matrix-vector product); a
// neighbour beyond a non-periodic axis is index n, the zero pad every level-0 vector and the code
// carry, so it reads 0 and adds nothing: no branch per neighbour.
struct Nb0 {
    uint32_t j[6];
};
__device__ __forceinline__ Nb0 mg_nb0(const MgLevel& L, uint32_t per, int x, int y, int z, uint32_t i) {
    const uint32_t n = L.n, sy = (uint32_t)L.nx, sz = (uint32_t)L.nx * (uint32_t)L.ny;
    Nb0 s;
    s.j[0] = x > 0 ? i - 1u : ((per & 1u) ? i + (uint32_t)(L.nx - 1) : n);
    s.j[1] = x + 1 < L.nx ? i + 1u : ((per & 1u) ? i - (uint32_t)(L.nx - 1) : n);
    s.j[2] = y > 0 ? i - sy : ((per & 2u) ? i + (uint32_t)(L.ny - 1) * sy : n);
    s.j[3] = y + 1 < L.ny ? i + sy : ((per & 2u) ? i - (uint32_t)(L.ny - 1) * sy : n);
    s.j[4] = z > 0 ? i - sz : ((per & 4u) ? i + (uint32_t)(L.nz - 1) * sz : n);
    s.j[5] = z + 1 < L.nz ? i + sz : ((per & 4u) ? i - (uint32_t)(L.nz - 1) * sz : n);
    return s;
}

// Level-1 aggregates of cell (x, y, z)'s level-0 neighbours, in mg_nb0's order (a neighbour beyond
// a non-periodic axis gets the cell's own aggregate: its value is not used), and of the cell itself.
__device__ __forceinline__ uint32_t mg_nb0_parents(const MgLevel& L, const MgLevel& C, uint32_t per, int x, int y, int z, uint32_t (&P)[6]) {
    const int X = x >> 1, Y = y >> 1, Z = z >> 1;
    const uint32_t sy = (uint32_t)C.nx, sz = (uint32_t)C.nx * (uint32_t)C.ny;
    const uint32_t P0 = (uint32_t)X + sy * (uint32_t)Y + sz * (uint32_t)Z;
    int q;
    q = x > 0 ? x - 1 : ((per & 1u) ? L.nx - 1 : x);
    P[0] = P0 + (uint32_t)((q >> 1) - X);
    q = x + 1 < L.nx ? x + 1 : ((per & 1u) ? 0 : x);
    P[1] = P0 + (uint32_t)((q >> 1) - X);
    q = y > 0 ? y - 1 : ((per & 2u) ? L.ny - 1 : y);
    P[2] = P0 + (uint32_t)((q >> 1) - Y) * sy;
    q = y + 1 < L.ny ? y + 1 : ((per & 2u) ? 0 : y);
    P[3] = P0 + (uint32_t)((q >> 1) - Y) * sy;
    q = z > 0 ? z - 1 : ((per & 4u) ? L.nz - 1 : z);
    P[4] = P0 + (uint32_t)((q >> 1) - Z) * sz;
    q = z + 1 < L.nz ? z + 1 : ((per & 4u) ? 0 : z);
    P[5] = P0 + (uint32_t)((q >> 1) - Z) * sz;
    return P0;
}

// Calls f(j, q, w) for every neighbour j of cell i = (x, y, z) on level L: q its coordinates, w the
// coupling A_ij (level 0: -1, to be applied to values that are 0 off the unknowns; coarser levels:
// i's + coupling, or the - neighbour's own). A periodic axis wraps (on an extent of 1 both
// neighbours are i itself, as in the level-0 matrix-vector product); a non-periodic one ends.
template <bool Fine, typename F>
__device__ __forceinline__ void mg_for_neighbours(const MgLevel& L, uint32_t per, int x, int y, int z, uint32_t i, F&& f) {
    const int ext[3]{L.nx, L.ny, L.nz};
    const uint32_t stride[3]{1u, (uint32_t)L.nx, (uint32_t)L.nx * (uint32_t)L.ny};
    const int c[3]{x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool pa = per & (1u << a);
        if (c[a] + 1 < ext[a] || pa) {
            int q[3]{x, y, z};
            q[a] = c[a] + 1 < ext[a] ? c[a] + 1 : 0;
            const uint32_t j = i + (uint32_t)(q[a] - c[a]) * stride[a];
            f(j, q, Fine ? -1.0f : L.wp[a][i]);
        }
        if (c[a] > 0 || pa) {
            int q[3]{x, y, z};
            q[a] = c[a] > 0 ? c[a] - 1 : ext[a] - 1;
            const uint32_t j = i + (uint32_t)(q[a] - c[a]) * stride[a];
            f(j, q, Fine ? -1.0f : L.wp[a][j]);
        }
    }
}

// Off-diagonal part of (A v)_i, v given by value(j, q) (0 off the level's unknowns).
template <bool Fine, typename F>
__device__ __forceinline__ float mg_offdiag(const MgLevel& L, uint32_t per, int x, int y, int z, uint32_t i, F&& value) {
    float s = 0.0f;
    mg_for_neighbours<Fine>(L, per, x, y, z, i, [&](uint32_t j, const int (&q)[3], float w) {
        if constexpr (Fine) s -= value(j, q);
        else s = __fmaf_rn(w, value(j, q), s);
    });
    return s;
}

// Index on level C of the aggregate holding the finer cell with coordinates q.
__device__ __forceinline__ uint32_t mg_parent(const MgLevel& C, const int (&q)[3]) {
    return (uint32_t)(q[0] >> 1) + (uint32_t)C.nx * ((uint32_t)(q[1] >> 1) + (uint32_t)C.ny * (uint32_t)(q[2] >> 1));
}

// Contiguous cell range of this block (XCD-remapped) for the kernels that reduce over level 0.
__device__ __forceinline__ void mg_range(uint32_t n, uint32_t& lb, uint32_t& start, uint32_t& end) {
    lb = xcd_block();
    uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    chunk = (chunk + blockDim.x - 1) / blockDim.x * blockDim.x;
    start = min(n, lb * chunk);
    end = min(n, start + chunk);
}

// Level 0, two damped Jacobi sweeps from zero for A x = r: x1 = w D^-1 r, x = x1 + w D^-1 (r - A x1),
// with x1 at the neighbours formed on the fly.
// Advances the iteration index (the first kernel of every V-cycle; nothing else reads it here).
__global__ __launch_bounds__(kBlock) void mg_pre0_kernel(MgLevel L, uint32_t per, const float* __restrict__ r, float* __restrict__ x,
                                                         PcgStatus* status) {
    if (status->stop) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) status->k = status->k + 1u;
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    const uint32_t d = L.code[i];
    if (d == 0u) {
        x[i] = 0.0f;
        return;
    }
    int cx, cy, cz;
    mg_coords(L, i, cx, cy, cz);
    const Nb0 nb = mg_nb0(L, per, cx, cy, cz, i);
    float s1 = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) s1 += mg_wdinv(L.code[nb.j[k]]) * r[nb.j[k]];
    const float ri = r[i], w = mg_wdinv(d);
    const float xi1 = w * ri;
    const float ax1 = __fmaf_rn((float)d, xi1, -s1);
    x[i] = __fmaf_rn(w, ri - ax1, xi1);
}

// Restriction on the coarser levels with one thread per cell (there are few aggregates to spread
// the work over): thread t handles child k = t mod 8 of aggregate I = t / 8
// (k = 1 x-bit + 2 y-bit + 4 z-bit), so the 8 lanes of an aligned group share I and their values are
// summed with three butterfly shuffles (the same association in every lane). Sets the child's cell
// coordinates and index; false when the child lies beyond an odd extent (it contributes 0).
__device__ __forceinline__ bool mg_child(const MgLevel& L, const MgLevel& C, uint32_t t, uint32_t& I, int (&q)[3], uint32_t& i) {
    I = t >> 3;
    const uint32_t k = t & 7u;
    int X, Y, Z;
    mg_coords(C, I, X, Y, Z);
    q[0] = 2 * X + (int)(k & 1u);
    q[1] = 2 * Y + (int)((k >> 1) & 1u);
    q[2] = 2 * Z + (int)(k >> 2);
    i = (uint32_t)q[0] + (uint32_t)L.nx * ((uint32_t)q[1] + (uint32_t)L.ny * (uint32_t)q[2]);
    return q[0] < L.nx && q[1] < L.ny && q[2] < L.nz;
}

__device__ __forceinline__ float mg_group8_sum(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    return v;
}

// Level 0 to 1: rc_I = sum of (r - A x)_i over the cells i of aggregate I (x: the pre-smoothed
// iterate), one thread per aggregate (consecutive lanes read consecutive pairs of cells of each
// row: coalesced, unlike one lane per cell of a 2 x 2 x 2 block). Children beyond an odd extent are
// clamped onto the last cell and masked out, so every load is issued unconditionally.
__global__ __launch_bounds__(kBlock) void mg_restrict0_kernel(MgLevel L, MgLevel C, uint32_t per, const float* __restrict__ r,
                                                              const float* __restrict__ x, float* __restrict__ rc, const PcgStatus* status) {
    if (status->stop) return;
    const uint32_t I = xcd_block() * blockDim.x + threadIdx.x;
    if (I >= C.n) return;
    int X, Y, Z;
    mg_coords(C, I, X, Y, Z);
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int qx = 2 * X + (k & 1), qy = 2 * Y + ((k >> 1) & 1), qz = 2 * Z + (k >> 2);
        const bool inside = qx < L.nx && qy < L.ny && qz < L.nz;
        const int cx = min(qx, L.nx - 1), cy = min(qy, L.ny - 1), cz = min(qz, L.nz - 1);
        const uint32_t i = (uint32_t)cx + (uint32_t)L.nx * ((uint32_t)cy + (uint32_t)L.ny * (uint32_t)cz);
        const uint32_t d = L.code[i];
        const Nb0 nb = mg_nb0(L, per, cx, cy, cz, i);
        float sx = 0.0f;
#pragma unroll
        for (int m = 0; m < 6; ++m) sx += x[nb.j[m]];
        const float v = r[i] - __fmaf_rn((float)d, x[i], -sx);
        s += (inside && d != 0u) ? v : 0.0f;
    }
    rc[I] = s;
}

// A coarser level l >= 1 on the way down (V(1,1)): the pre-smoothing x1 = w D^-1 r is formed on the
// fly, so one launch gives rc_I = sum of (r - A x1)_i over the cells i of aggregate I on level l + 1.
// Grid: 8 threads per aggregate.
__global__ __launch_bounds__(kBlock) void mg_down_kernel(MgLevel L, MgLevel C, uint32_t per, const float* __restrict__ r,
                                                         float* __restrict__ rc, const PcgStatus* status) {
    if (status->stop) return;
    const uint32_t t = xcd_block() * blockDim.x + threadIdx.x;
    if ((t >> 3) >= C.n) return;   // whole groups of 8 leave together
    uint32_t I, i;
    int q[3];
    float v = 0.0f;
    if (mg_child(L, C, t, I, q, i)) {
        const float aii = L.diag[i];
        if (aii > 0.0f) {
            auto x1 = [&](uint32_t j, const int (&)[3]) { return kOmegaC * L.idiag[j] * r[j]; };
            const float ax = __fmaf_rn(aii, x1(i, q), mg_offdiag<false>(L, per, q[0], q[1], q[2], i, x1));
            v = r[i] - ax;
        }
    }
    v = mg_group8_sum(v);
    if ((t & 7u) == 0u) rc[I] = v;
}

// A coarser level l >= 1 on the way up: xc = w D^-1 r + over P e (the pre-smoothed iterate plus the
// prolonged correction, both formed on the fly), then one sweep z = xc + w D^-1 (r - A xc).
__global__ __launch_bounds__(kBlock) void mg_up_kernel(MgLevel L, MgLevel C, uint32_t per, const float* __restrict__ r,
                                                       const float* __restrict__ e, float* __restrict__ z, const PcgStatus* status) {
    if (status->stop) return;
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    const float aii = L.diag[i];
    if (!(aii > 0.0f)) {
        z[i] = 0.0f;
        return;
    }
    int cx, cy, cz;
    mg_coords(L, i, cx, cy, cz);
    auto xc = [&](uint32_t j, const int (&q)[3]) {
        const float inv = L.idiag[j];
        return inv > 0.0f ? __fmaf_rn(kOverC, e[mg_parent(C, q)], kOmegaC * inv * r[j]) : 0.0f;
    };
    const int c[3]{cx, cy, cz};
    const float xi = xc(i, c);
    const float axc = __fmaf_rn(aii, xi, mg_offdiag<false>(L, per, cx, cy, cz, i, xc));
    z[i] = __fmaf_rn(kOmegaC * L.idiag[i], r[i] - axc, xi);
}

// Level 0 on the way up, first sweep, with the prolonged correction applied on the fly:
// xc = x + over P e, y = xc + w D^-1 (r - A xc).
__global__ __launch_bounds__(kBlock) void mg_post0_kernel(MgLevel L, MgLevel C, uint32_t per, const float* __restrict__ r,
                                                          const float* __restrict__ x, const float* __restrict__ e, float* __restrict__ y,
                                                          const PcgStatus* status) {
    if (status->stop) return;
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    const uint32_t d = L.code[i];
    if (d == 0u) {
        y[i] = 0.0f;
        return;
    }
    int cx, cy, cz;
    mg_coords(L, i, cx, cy, cz);
    const Nb0 nb = mg_nb0(L, per, cx, cy, cz, i);
    uint32_t P[6];
    const uint32_t Pi = mg_nb0_parents(L, C, per, cx, cy, cz, P);
    float sx = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t j = nb.j[k];
        sx += L.code[j] ? __fmaf_rn(kOver0, e[P[k]], x[j]) : 0.0f;
    }
    const float xi = __fmaf_rn(kOver0, e[Pi], x[i]);
    const float axc = __fmaf_rn((float)d, xi, -sx);
    y[i] = __fmaf_rn(mg_wdinv(d), r[i] - axc, xi);
}

// Level 0, second sweep z = y + w D^-1 (r - A y), over contiguous block ranges, with the partials of
// <r, z>.
__global__ __launch_bounds__(kMgThreads) void mg_post0_last_kernel(MgLevel L, uint32_t per, const float* __restrict__ r,
                                                                   const float* __restrict__ y, float* __restrict__ z,
                                                                   double* __restrict__ partials, const PcgStatus* status) {
    __shared__ double lds[kMaxWaves];
    if (status->stop) return;
    uint32_t lb, start, end;
    mg_range(L.n, lb, start, end);
    double rz = 0.0;
    for (uint32_t base = start + threadIdx.x; base < end; base += kMgUnroll * blockDim.x) {
#pragma unroll
        for (int m = 0; m < kMgUnroll; ++m) {
            const uint32_t i0 = base + (uint32_t)m * blockDim.x;
            const bool valid = i0 < end;
            const uint32_t i = min(i0, L.n - 1u);
            const uint32_t d = L.code[i];
            int cx, cy, cz;
            mg_coords(L, i, cx, cy, cz);
            const Nb0 nb = mg_nb0(L, per, cx, cy, cz, i);
            float sy = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) sy += y[nb.j[k]];
            const float yi = y[i], ri = r[i];
            const float zi = d ? __fmaf_rn(mg_wdinv(d), ri - __fmaf_rn((float)d, yi, -sy), yi) : 0.0f;
            if (valid) z[i] = zi;
            rz = valid ? __fma_rn((double)ri, (double)zi, rz) : rz;
        }
    }
    double v[1]{rz};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[kPartRZ * kPcgMaxBlocks + lb] = v[0];
}

// Coarsest level: x = A^-1 r in FP64, one block per row of the inverse. When level 0 is the coarsest
// this is the V-cycle's first kernel and advances the iteration index.
__global__ __launch_bounds__(kBlock) void mg_dense_kernel(uint32_t n, const double* __restrict__ inv, const float* __restrict__ r,
                                                          float* __restrict__ x, PcgStatus* status, uint32_t advance) {
    __shared__ double lds[kMaxWaves];
    if (status->stop) return;
    if (advance && blockIdx.x == 0 && threadIdx.x == 0) status->k = status->k + 1u;
    const uint32_t I = blockIdx.x;
    double s = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) s = __fma_rn(inv[(size_t)I * n + j], (double)r[j], s);
    double v[1]{s};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) x[I] = (float)v[0];
}

// Partials of <r, z> on level 0 (when level 0 is the coarsest and the dense solve produced z).
__global__ __launch_bounds__(kMgThreads) void mg_dot_kernel(uint32_t n, const float* __restrict__ r, const float* __restrict__ z,
                                                            double* __restrict__ partials, const PcgStatus* status) {
    __shared__ double lds[kMaxWaves];
    if (status->stop) return;
    uint32_t lb, start, end;
    mg_range(n, lb, start, end);
    double rz = 0.0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) rz = __fma_rn((double)r[i], (double)z[i], rz);
    double v[1]{rz};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[kPartRZ * kPcgMaxBlocks + lb] = v[0];
}

// PCG direction of iteration k: stopping test on ||r||^2 (RR partials) against tol^2 ||b||^2 (BB
// partials on the first launch of a solve, then the status), beta = rz_k / rz_{k-1} (RZ partials of
// the V-cycle), p = z + beta p_old (p = z when k = 0), partials of <p, A p>; A p is formed from p
// at the neighbours on the fly.
__global__ __launch_bounds__(kMgThreads) void mg_direction_kernel(MgLevel L, uint32_t per, const float* __restrict__ z,
                                                                  float* __restrict__ p0, float* __restrict__ p1,
                                                                  double* __restrict__ partials, double* __restrict__ hist,
                                                                  PcgStatus* status, double tol2, uint32_t check_only) {
    __shared__ double lds[2 * kMaxWaves];
    if (status->stop) return;
    const uint32_t k = status->k;
    const double* parts[2]{partials + kPartRR * kPcgMaxBlocks, partials + kPartRZ * kPcgMaxBlocks};
    double red[2];
    reduce_partials<2>(parts, gridDim.x, red, lds);
    const double bb = status->bb, rho = red[0], rz = red[1];
    uint32_t lb, start, end;
    mg_range(L.n, lb, start, end);
    const bool converged = !(rho > tol2 * bb);
    if (converged || check_only) {
        if (lb == 0 && threadIdx.x == 0) {
            status->rr = rho;
            status->iterations = k;
            if (converged) status->stop = 1u;
        }
        return;
    }
    if (lb == 0 && threadIdx.x == 0) hist[2 * k] = rz;
    float* __restrict__ p_new = (k & 1u) ? p1 : p0;
    const float* __restrict__ p_old = (k & 1u) ? p0 : p1;
    const bool with_old = k > 0;
    const float beta = with_old ? (float)(rz / hist[2 * (k - 1)]) : 0.0f;
    const uint8_t* __restrict__ code = L.code;
    // z and p_old are 0 off the unknowns, so p_new is too
    auto pn = [&](uint32_t j) -> float { return with_old ? __fmaf_rn(beta, p_old[j], z[j]) : z[j]; };
    double pq = 0.0;
    for (uint32_t base = start + threadIdx.x; base < end; base += kMgUnroll * blockDim.x) {
#pragma unroll
        for (int m = 0; m < kMgUnroll; ++m) {
            const uint32_t i0 = base + (uint32_t)m * blockDim.x;
            const bool valid = i0 < end;
            const uint32_t i = min(i0, L.n - 1u);
            const uint32_t d = code[i];
            int cx, cy, cz;
            mg_coords(L, i, cx, cy, cz);
            const float pi = pn(i);   // 0 off the unknowns
            if (valid) p_new[i] = pi;
            // A p as the update forms it (exactly), so alpha and the residual recurrence agree
            const Nb0 nb = mg_nb0(L, per, cx, cy, cz, i);
            double sp = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) sp += (double)pn(nb.j[k]);
            const double q = (double)d * (double)pi - sp;
            pq = valid ? __fma_rn((double)pi, q, pq) : pq;
        }
    }
    double v[1]{pq};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[kPartPQ * kPcgMaxBlocks + lb] = v[0];
}

// PCG update of iteration k: alpha = rz_k / <p, A p>; x += alpha p; r -= alpha A p; partials of
// ||r||^2. The solution pair (hi + lo, about 48 bits) takes alpha p in FP64 and A p is exact, so
// the recursive residual follows the true one to FP32 rounding of r itself: with the near-null
// mode of the matrix the solution is some 10^3 ||b|| large, and an FP32 increment (or an FP32
// A p) would leave a gap of about 1e-4 ||b|| between them after the first, largest steps.
__global__ __launch_bounds__(kMgThreads) void mg_update_kernel(MgLevel L, uint32_t per, const float* __restrict__ p0,
                                                               const float* __restrict__ p1, float* __restrict__ r, float* __restrict__ xh,
                                                               float* __restrict__ xl, double* __restrict__ partials,
                                                               const double* __restrict__ hist, const PcgStatus* status) {
    __shared__ double lds[kMaxWaves];
    if (status->stop) return;
    const uint32_t k = status->k;
    const float* __restrict__ p = (k & 1u) ? p1 : p0;
    const double* parts[1]{partials + kPartPQ * kPcgMaxBlocks};
    double red[1];
    reduce_partials<1>(parts, gridDim.x, red, lds);
    const double alpha = hist[2 * k] / red[0];
    uint32_t lb, start, end;
    mg_range(L.n, lb, start, end);
    double rr = 0.0;
    for (uint32_t base = start + threadIdx.x; base < end; base += kMgUnroll * blockDim.x) {
#pragma unroll
        for (int m = 0; m < kMgUnroll; ++m) {
            const uint32_t i0 = base + (uint32_t)m * blockDim.x;
            const uint32_t i = min(i0, L.n - 1u);
            const uint32_t d = L.code[i];
            const bool live = i0 < end && d != 0u;
            int cx, cy, cz;
            mg_coords(L, i, cx, cy, cz);
            const Nb0 nb = mg_nb0(L, per, cx, cy, cz, i);
            const float pi = p[i];
            double sp = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) sp += (double)p[nb.j[k]];
            const double q = (double)d * (double)pi - sp;
            const double xv = __fma_rn(alpha, (double)pi, (double)xh[i] + (double)xl[i]);
            const float hi = (float)xv;
            const float ri = (float)__fma_rn(-alpha, q, (double)r[i]);
            if (live) {
                xh[i] = hi;
                xl[i] = (float)(xv - (double)hi);
                r[i] = ri;
            }
            rr = live ? __fma_rn((double)ri, (double)ri, rr) : rr;
        }
    }
    double v[1]{rr};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[kPartRR * kPcgMaxBlocks + lb] = v[0];
}

// ---- hierarchy (host) ------------------------------------------------------------------------

struct HostLevel {
    int nx = 0, ny = 0, nz = 0;
    std::vector<double> diag;    // A_ii
    std::vector<double> wp[3];   // coupling to the + neighbour per axis (wrapped on a periodic axis)
    size_t cells() const { return (size_t)nx * ny * nz; }
};

// Index of the + neighbour of (x, y, z) along a, or SIZE_MAX outside a non-periodic axis.
size_t plus_neighbour(const HostLevel& L, const bool per[3], int x, int y, int z, int a) {
    int q[3]{x, y, z};
    const int e[3]{L.nx, L.ny, L.nz};
    if (++q[a] >= e[a]) {
        if (!per[a]) return SIZE_MAX;
        q[a] = 0;
    }
    return (size_t)q[0] + (size_t)L.nx * ((size_t)q[1] + (size_t)L.ny * (size_t)q[2]);
}

// Galerkin operator P^T A P of the 2 x 2 x 2 aggregation of level f. Each + edge (i, j) of weight w
// adds 2 w to the aggregate's diagonal when i and j share an aggregate (the edge's two off-diagonal
// entries), else w to the + coupling of i's aggregate along the edge's axis (a periodic wrap of the
// fine level is a wrap of the coarse level too).
HostLevel coarsen(const HostLevel& f, const bool per[3]) {
    HostLevel c;
    c.nx = (f.nx + 1) / 2;
    c.ny = (f.ny + 1) / 2;
    c.nz = (f.nz + 1) / 2;
    const size_t nc = c.cells();
    c.diag.assign(nc, 0.0);
    for (auto& w : c.wp) w.assign(nc, 0.0);
    auto agg = [&](size_t j) {
        const size_t x = j % (size_t)f.nx, t = j / (size_t)f.nx;
        const size_t y = t % (size_t)f.ny, z = t / (size_t)f.ny;
        return x / 2 + (size_t)c.nx * (y / 2 + (size_t)c.ny * (z / 2));
    };
    for (int z = 0; z < f.nz; ++z)
        for (int y = 0; y < f.ny; ++y)
            for (int x = 0; x < f.nx; ++x) {
                const size_t i = (size_t)x + (size_t)f.nx * ((size_t)y + (size_t)f.ny * (size_t)z);
                const size_t I = (size_t)(x / 2) + (size_t)c.nx * ((size_t)(y / 2) + (size_t)c.ny * (size_t)(z / 2));
                c.diag[I] += f.diag[i];
                for (int a = 0; a < 3; ++a) {
                    const double w = f.wp[a][i];
                    if (w == 0.0) continue;
                    const size_t j = plus_neighbour(f, per, x, y, z, a);
                    if (j == SIZE_MAX) continue;
                    if (agg(j) == I) c.diag[I] += 2.0 * w;
                    else c.wp[a][I] += w;
                }
            }
    return c;
}

// Inverse of level L's operator over its unknowns (diag > 0), FP64 Cholesky, as a symmetric
// n x n matrix with zero rows and columns off the unknowns.
std::vector<double> dense_inverse(const HostLevel& L, const bool per[3]) {
    const size_t n = L.cells();
    std::vector<int64_t> pos(n, -1);
    std::vector<size_t> unk;
    for (size_t i = 0; i < n; ++i)
        if (L.diag[i] > 0.0) {
            pos[i] = (int64_t)unk.size();
            unk.push_back(i);
        }
    const size_t m = unk.size();
    std::vector<double> A(m * m, 0.0);
    for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
            for (int x = 0; x < L.nx; ++x) {
                const size_t i = (size_t)x + (size_t)L.nx * ((size_t)y + (size_t)L.ny * (size_t)z);
                if (pos[i] < 0) continue;
                const size_t pi = (size_t)pos[i];
                A[pi * m + pi] += L.diag[i];
                for (int a = 0; a < 3; ++a) {
                    const double w = L.wp[a][i];
                    if (w == 0.0) continue;
                    const size_t j = plus_neighbour(L, per, x, y, z, a);
                    if (j == SIZE_MAX || pos[j] < 0) continue;
                    const size_t pj = (size_t)pos[j];
                    A[pi * m + pj] += w;
                    A[pj * m + pi] += w;
                }
            }
    // A = G G^T, G lower triangular (in place)
    for (size_t j = 0; j < m; ++j) {
        double s = A[j * m + j];
        for (size_t k = 0; k < j; ++k) s -= A[j * m + k] * A[j * m + k];
        if (!(s > 0.0)) throw Failure{DS_ERROR_RUNTIME, "multigrid: the coarsest operator is not positive definite"};
        const double d = std::sqrt(s);
        A[j * m + j] = d;
        for (size_t i = j + 1; i < m; ++i) {
            double t = A[i * m + j];
            for (size_t k = 0; k < j; ++k) t -= A[i * m + k] * A[j * m + k];
            A[i * m + j] = t / d;
        }
    }
    // column c of A^-1: G y = e_c, then G^T x = y
    std::vector<double> inv(m * m);
    std::vector<double> col(m);
    for (size_t c = 0; c < m; ++c) {
        for (size_t i = 0; i < m; ++i) {
            double t = i == c ? 1.0 : 0.0;
            for (size_t k = 0; k < i; ++k) t -= A[i * m + k] * col[k];
            col[i] = t / A[i * m + i];
        }
        for (size_t i = m; i-- > 0;) {
            double t = col[i];
            for (size_t k = i + 1; k < m; ++k) t -= A[k * m + i] * col[k];
            col[i] = t / A[i * m + i];
        }
        for (size_t i = 0; i < m; ++i) inv[i * m + c] = col[i];
    }
    std::vector<double> out(n * n, 0.0);
    for (size_t a = 0; a < m; ++a)
        for (size_t b = 0; b < m; ++b) out[unk[a] * n + unk[b]] = 0.5 * (inv[a * m + b] + inv[b * m + a]);
    return out;
}

uint32_t periodic_mask(const ds_context* c) {
    return (c->pressure_periodic[0] ? 1u : 0u) | (c->pressure_periodic[1] ? 2u : 0u) | (c->pressure_periodic[2] ? 4u : 0u);
}

}  // namespace

void mg_prepare(ds_context* c, const std::vector<uint8_t>& code) {
    MgWork& w = c->mg;
    const Geo& g = c->geo;
    const bool per[3]{c->pressure_periodic[0], c->pressure_periodic[1], c->pressure_periodic[2]};
    require(code.size() >= (size_t)g.n, "multigrid: code shorter than the grid");
    // level 0 as diagonal + couplings: A_ii = code_i, -1 between neighbouring unknowns (on an
    // extent-1 periodic axis a cell neighbours itself twice, as in the matrix-vector product)
    std::vector<HostLevel> levels(1);
    {
        HostLevel& f = levels[0];
        f.nx = g.nx;
        f.ny = g.ny;
        f.nz = g.nz;
        const size_t n = f.cells();
        f.diag.resize(n);
        for (auto& v : f.wp) v.assign(n, 0.0);
        for (int z = 0; z < g.nz; ++z)
            for (int y = 0; y < g.ny; ++y)
                for (int x = 0; x < g.nx; ++x) {
                    const size_t i = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * (size_t)z);
                    f.diag[i] = (double)code[i];
                    if (!code[i]) continue;
                    for (int a = 0; a < 3; ++a) {
                        const size_t j = plus_neighbour(f, per, x, y, z, a);
                        if (j != SIZE_MAX && code[j]) f.wp[a][i] = -1.0;
                    }
                }
    }
    while (levels.back().cells() > kMgCoarsest) {
        if ((int)levels.size() == kMgMaxLevels) throw Failure{DS_ERROR_UNSUPPORTED, "multigrid: too many levels"};
        HostLevel next = coarsen(levels.back(), per);
        levels.push_back(std::move(next));
    }
    w.levels = (int)levels.size();
    for (int l = 0; l < w.levels; ++l) {
        const HostLevel& H = levels[l];
        MgLevel& L = w.lv[l];
        L.nx = H.nx;
        L.ny = H.ny;
        L.nz = H.nz;
        L.n = (uint32_t)H.cells();
        L.dx = make_fastdiv((uint32_t)H.nx);
        L.dy = make_fastdiv((uint32_t)H.ny);
        L.code = l == 0 ? c->code : nullptr;
        L.diag = nullptr;
        L.idiag = nullptr;
        for (int a = 0; a < 3; ++a) L.wp[a] = nullptr;
        if (l > 0) {
            std::vector<float> buf(L.n);
            auto put = [&](const std::vector<double>& v) {
                for (uint32_t i = 0; i < L.n; ++i) buf[i] = (float)v[i];
                float* d = device_alloc<float>(c, L.n);
                DS_HIP(hipMemcpy(d, buf.data(), L.n * sizeof(float), hipMemcpyHostToDevice));
                return d;
            };
            L.diag = put(H.diag);
            std::vector<double> inv(L.n);
            for (uint32_t i = 0; i < L.n; ++i) inv[i] = H.diag[i] > 0.0 ? 1.0 / H.diag[i] : 0.0;
            L.idiag = put(inv);
            for (int a = 0; a < 3; ++a) L.wp[a] = put(H.wp[a]);
            // right-hand side and result
            w.vec[l][0] = device_alloc<float>(c, L.n);
            w.vec[l][3] = device_alloc<float>(c, L.n);
        } else {
            // the pre-smoothed iterate and the iterate after the first post sweep, each with the
            // zero pad element n that out-of-domain neighbours read
            w.vec[l][1] = device_alloc<float>(c, (size_t)L.n + 1);
            w.vec[l][2] = device_alloc<float>(c, (size_t)L.n + 1);
        }
        for (int k = 0; k < 4; ++k) fill_zero(c, w.vec[l][k], (size_t)L.n + (l == 0 ? 1 : 0));
    }
    const std::vector<double> inv = dense_inverse(levels.back(), per);
    w.dense = device_alloc<double>(c, inv.size());
    DS_HIP(hipMemcpy(w.dense, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice));
    w.z = device_alloc<float>(c, (size_t)g.n + 1);
    fill_zero(c, w.z, (size_t)g.n + 1);
}

// z = M^-1 r, the V-cycle from level 0 (r: level 0's right-hand side), with the partials of <r, z>.
// Level 0: pre-smoothing, restriction; each coarser level but the last: one launch down (its
// right-hand side to the next level's); the coarsest: the dense solve; back up one launch per
// coarser level (its result from the next level's); level 0: the two post sweeps.
void mg_precondition(ds_context* c, const float* r, float* z) {
    MgWork& w = c->mg;
    PcgWork& pw = c->pcg;
    const uint32_t per = periodic_mask(c);
    const int Lc = w.levels - 1;
    hipStream_t s = c->stream;
    PcgStatus* st = pw.status;
    if (Lc == 0) {
        mg_dense_kernel<<<w.lv[0].n, kBlock, 0, s>>>(w.lv[0].n, w.dense, r, z, st, 1u);
        mg_dot_kernel<<<pw.blocks, kMgThreads, 0, s>>>(w.lv[0].n, r, z, pw.partials, st);
        check_launch("multigrid direct solve");
        return;
    }
    const MgLevel& L0 = w.lv[0];
    mg_pre0_kernel<<<xcd_blocks_for(L0.n), kBlock, 0, s>>>(L0, per, r, w.vec[0][1], st);
    mg_restrict0_kernel<<<xcd_blocks_for(w.lv[1].n), kBlock, 0, s>>>(L0, w.lv[1], per, r, w.vec[0][1], w.vec[1][0], st);
    for (int l = 1; l < Lc; ++l)
        mg_down_kernel<<<xcd_blocks_for(8ull * w.lv[l + 1].n), kBlock, 0, s>>>(w.lv[l], w.lv[l + 1], per, w.vec[l][0], w.vec[l + 1][0],
                                                                                st);
    mg_dense_kernel<<<w.lv[Lc].n, kBlock, 0, s>>>(w.lv[Lc].n, w.dense, w.vec[Lc][0], w.vec[Lc][3], st, 0u);
    for (int l = Lc - 1; l >= 1; --l)
        mg_up_kernel<<<xcd_blocks_for(w.lv[l].n), kBlock, 0, s>>>(w.lv[l], w.lv[l + 1], per, w.vec[l][0], w.vec[l + 1][3], w.vec[l][3], st);
    mg_post0_kernel<<<xcd_blocks_for(L0.n), kBlock, 0, s>>>(L0, w.lv[1], per, r, w.vec[0][1], w.vec[1][3], w.vec[0][2], st);
    mg_post0_last_kernel<<<pw.blocks, kMgThreads, 0, s>>>(L0, per, r, w.vec[0][2], z, pw.partials, st);
    check_launch("multigrid V-cycle");
}

void mg_direction(ds_context* c, const float* z, double tol2, bool check_only) {
    PcgWork& w = c->pcg;
    mg_direction_kernel<<<w.blocks, kMgThreads, 0, c->stream>>>(c->mg.lv[0], periodic_mask(c), z, w.p[0], w.p[1], w.partials, w.hist,
                                                               w.status, tol2, check_only ? 1u : 0u);
}

void mg_update(ds_context* c, float* xh, float* xl) {
    PcgWork& w = c->pcg;
    mg_update_kernel<<<w.blocks, kMgThreads, 0, c->stream>>>(c->mg.lv[0], periodic_mask(c), w.p[0], w.p[1], w.r, xh, xl, w.partials,
                                                            w.hist, w.status);
}

namespace {

__global__ void mg_set_iteration_kernel(PcgStatus* status, uint32_t k) {
    if (threadIdx.x == 0) status->k = k;
}

constexpr size_t kMgMaxGraphs = 8;

}  // namespace

void mg_set_iteration(ds_context* c, uint32_t k_before) {
    mg_set_iteration_kernel<<<1, 64, 0, c->stream>>>(c->pcg.status, k_before);
    check_launch("mg_set_iteration_kernel");
}

void mg_iteration(ds_context* c, float* xh, float* xl, double tol2) {
    MgWork& w = c->mg;
    hipGraphExec_t exec = nullptr;
    for (const MgWork::Graph& g : w.graphs)
        if (g.xh == xh && g.xl == xl && g.tol2 == tol2) exec = g.exec;
    if (!exec) {
        if (w.graphs.size() == kMgMaxGraphs) {
            (void)hipGraphExecDestroy(w.graphs.front().exec);
            w.graphs.erase(w.graphs.begin());
        }
        hipGraph_t graph = nullptr;
        DS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        try {
            mg_precondition(c, c->pcg.r, w.z);
            mg_direction(c, w.z, tol2, false);
            mg_update(c, xh, xl);
            check_launch("multigrid iteration (capture)");
        } catch (...) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(c->stream, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            throw;
        }
        DS_HIP(hipStreamEndCapture(c->stream, &graph));
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        DS_HIP(e);
        w.graphs.push_back({xh, xl, tol2, exec});
    }
    DS_HIP(hipGraphLaunch(exec, c->stream));
}

void mg_release(ds_context* c) {
    for (const MgWork::Graph& g : c->mg.graphs) (void)hipGraphExecDestroy(g.exec);
    c->mg.graphs.clear();
}

}  // namespace ds
