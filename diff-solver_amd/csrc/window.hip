// window.hip — DS_WARM_START_PROJECTION: the initial guess of a pressure solve as the A-norm-best
// combination of the last kWindow solutions of the same sequence (forward solves in step order,
// adjoint solves in reverse-step order).
//
// For a new right-hand side b with exact solution x*, the guess x0 = X c with c solving
// (X^T A X) c = X^T b minimises ||x* - X c||_A over the span of the stored solutions X: the
// Galerkin projection that CG itself would start from (P. F. Fischer, "Projection techniques for
// iterative solution of Ax = b with successive right-hand sides", Comput. Methods Appl. Mech.
// Engrg. 163, 1998). Kept per sequence: X and A X in FP32 (ring of kWindow slots) and the Gram
// matrix G = X^T A X in FP64, updated one row per solve. Per solve: X^T b (one pass over the
// window), a Cholesky solve of G on one thread, x0 = X c (one pass), and after the solve the new
// solution, its image (one stencil pass) and its Gram row (one pass). A guess whose residual
// exceeds ||b|| is dropped by the solver like any other.
#include <algorithm>

#include "kernels.hpp"
#include "rhs.hpp"

namespace ds {
namespace {

struct WinVecs {
    const float* v[kWindow];
};

// partials[j * G + block] = <a, v_j> over the cells, j < count
__global__ __launch_bounds__(256) void window_dots_kernel(uint32_t n, const float* __restrict__ a, WinVecs w, uint32_t count,
                                                          double* __restrict__ partials) {
    __shared__ double lds[kWindow * kMaxWaves];
    double acc[kWindow]{};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ai = a[i];
#pragma unroll
        for (int j = 0; j < kWindow; ++j)
            if ((uint32_t)j < count) acc[j] = __fma_rn(ai, (double)w.v[j][i], acc[j]);
    }
    block_sum<kWindow>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < kWindow; ++j) partials[(size_t)j * gridDim.x + blockIdx.x] = acc[j];
}

// Finish X^T b and solve G c = X^T b (Cholesky with a relative ridge; a pivot that is not
// positive ends the factorisation there and the remaining coefficients are 0). `order` lists the
// slots from newest to oldest, so a breakdown keeps the most recent solutions.
__global__ __launch_bounds__(256) void window_solve_kernel(const double* __restrict__ partials, uint32_t G, uint32_t count,
                                                           const double* __restrict__ gram, WinOrder order, float* __restrict__ coef) {
    __shared__ double lds[kWindow * kMaxWaves];
    const double* parts[kWindow];
#pragma unroll
    for (int j = 0; j < kWindow; ++j) parts[j] = partials + (size_t)j * G;
    double d[kWindow];
    reduce_partials<kWindow>(parts, G, d, lds);
    if (threadIdx.x != 0) return;
    const int m = (int)count;
    double a[kWindow][kWindow], y[kWindow], c[kWindow];
    double tr = 0.0;
    for (int p = 0; p < m; ++p) tr += gram[order.slot[p] * kWindow + order.slot[p]];
    const double ridge = 1e-10 * tr / (m > 0 ? m : 1);
    for (int p = 0; p < m; ++p)
        for (int q = 0; q < m; ++q) a[p][q] = gram[order.slot[p] * kWindow + order.slot[q]] + (p == q ? ridge : 0.0);
    int rank = 0;
    for (int p = 0; p < m; ++p) {   // in-place Cholesky, lower triangle
        double s = a[p][p];
        for (int q = 0; q < p; ++q) s -= a[p][q] * a[p][q];
        if (!(s > 0.0)) break;
        a[p][p] = sqrt(s);
        for (int r = p + 1; r < m; ++r) {
            double t = a[r][p];
            for (int q = 0; q < p; ++q) t -= a[r][q] * a[p][q];
            a[r][p] = t / a[p][p];
        }
        rank = p + 1;
    }
    for (int p = 0; p < rank; ++p) {   // L y = d (in newest-first order)
        double t = d[order.slot[p]];
        for (int q = 0; q < p; ++q) t -= a[p][q] * y[q];
        y[p] = t / a[p][p];
    }
    for (int p = rank - 1; p >= 0; --p) {   // L^T c = y
        double t = y[p];
        for (int q = p + 1; q < rank; ++q) t -= a[q][p] * c[q];
        c[p] = t / a[p][p];
    }
    for (int j = 0; j < kWindow; ++j) coef[j] = 0.0f;
    for (int p = 0; p < rank; ++p) coef[order.slot[p]] = (float)c[p];
}

// x0 = sum_j coef_j v_j
__global__ __launch_bounds__(256) void window_combine_kernel(uint32_t n, WinVecs w, uint32_t count, const float* __restrict__ coef,
                                                             float* __restrict__ x0) {
    float cf[kWindow];
#pragma unroll
    for (int j = 0; j < kWindow; ++j) cf[j] = (uint32_t)j < count ? coef[j] : 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < kWindow; ++j)
            if ((uint32_t)j < count) s = __fmaf_rn(cf[j], w.v[j][i], s);
        x0[i] = s;
    }
}

// slot = hi + lo
__global__ __launch_bounds__(256) void window_store_kernel(uint32_t n, const float* __restrict__ hi, const float* __restrict__ lo,
                                                           float* __restrict__ slot) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) slot[i] = hi[i] + lo[i];
}

// The Gram row of the new slot s: gram[s][j] = gram[j][s] = <A x_s, x_j>
__global__ __launch_bounds__(256) void window_gram_kernel(const double* __restrict__ partials, uint32_t G, uint32_t count, uint32_t s,
                                                          double* __restrict__ gram) {
    __shared__ double lds[kWindow * kMaxWaves];
    const double* parts[kWindow];
#pragma unroll
    for (int j = 0; j < kWindow; ++j) parts[j] = partials + (size_t)j * G;
    double d[kWindow];
    reduce_partials<kWindow>(parts, G, d, lds);
    if (threadIdx.x == 0)
        for (uint32_t j = 0; j < count; ++j) {
            gram[s * kWindow + j] = d[j];
            gram[j * kWindow + s] = d[j];
        }
}

constexpr uint32_t kWindowBlocks = 512;

WinVecs vecs_of(float* const* v) {
    WinVecs w{};
    for (int j = 0; j < kWindow; ++j) w.v[j] = v[j];
    return w;
}

WinOrder order_of(const SolveWindow& w) {
    WinOrder o{};
    for (uint32_t p = 0; p < w.count; ++p) o.slot[p] = (int)((w.head + kWindow - 1 - p) % kWindow);
    return o;
}

}  // namespace

void window_reset(ds_context* c) {
    for (SolveWindow& w : c->window) {
        w.count = 0;
        w.head = 0;
    }
}

SolveWindow* window_of(ds_context* c, int kind) {
    if (c->config.warm_start != DS_WARM_START_PROJECTION) return nullptr;
    SolveWindow* w = kind == kRhsForward ? &c->window[0] : kind == kRhsAdjoint ? &c->window[1] : nullptr;
    return w && w->x[0] ? w : nullptr;
}

const float* window_guess(ds_context* c, SolveWindow& w, const float* b) {
    if (w.count == 0) return nullptr;
    const uint32_t n = c->geo.n;
    window_dots_kernel<<<kWindowBlocks, kBlock, 0, c->stream>>>(n, b, vecs_of(w.x), w.count, w.partials);
    window_solve_kernel<<<1, kBlock, 0, c->stream>>>(w.partials, kWindowBlocks, w.count, w.gram, order_of(w), w.coef);
    window_combine_kernel<<<2048, kBlock, 0, c->stream>>>(n, vecs_of(w.x), w.count, w.coef, c->guess);
    check_launch("window_guess");
    return c->guess;
}

void window_add(ds_context* c, SolveWindow& w, const float* hi, const float* lo) {
    const uint32_t n = c->geo.n;
    const uint32_t s = w.head;
    window_store_kernel<<<2048, kBlock, 0, c->stream>>>(n, hi, lo, w.x[s]);
    launch_apply_matrix(c, w.x[s], w.ax[s]);
    w.head = (w.head + 1) % kWindow;
    w.count = std::min<uint32_t>(w.count + 1, kWindow);
    // <A x_s, x_j> over the live slots (slot indices are below count once the ring is full, and
    // the live slots are 0 .. count-1 before it is)
    window_dots_kernel<<<kWindowBlocks, kBlock, 0, c->stream>>>(n, w.ax[s], vecs_of(w.x), w.count, w.partials);
    window_gram_kernel<<<1, kBlock, 0, c->stream>>>(w.partials, kWindowBlocks, w.count, s, w.gram);
    check_launch("window_add");
}

void window_allocate(ds_context* c, SolveWindow& w) {
    const uint32_t n = c->geo.n;
    for (int j = 0; j < kWindow; ++j) {
        w.x[j] = device_alloc<float>(c, n);
        w.ax[j] = device_alloc<float>(c, n);
        fill_zero(c, w.x[j], n);
        fill_zero(c, w.ax[j], n);
    }
    w.partials = device_alloc<double>(c, (size_t)kWindow * kWindowBlocks);
    w.gram = device_alloc<double>(c, (size_t)kWindow * kWindow);
    fill_zero_d(c, w.gram, (size_t)kWindow * kWindow);
    w.coef = device_alloc<float>(c, kWindow);
    w.count = 0;
    w.head = 0;
}

}  // namespace ds
