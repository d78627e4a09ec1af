// vorticity.hpp — device functions of vorticity confinement (solvers/smoke/operators.cu:662-856).
//
// Forward, per fluid cell: cell-centred velocity c = mean of the two faces (0 in solids), vorticity
// w = curl c by central differences with neighbours clamped to the grid, smoothed magnitude
// |w| = sqrt(w.w + eps^2) (eps in solids), normal N = grad|w| / len with len = sqrt(|grad|w||^2 +
// eps^2), force += confinement * h * (N x w).
//
// The reference runs the reverse as five scatter kernels with atomics into zeroed FP64 scratch.
// Here every reverse stencil is a gather: the transpose of the clamped central difference
// D v(c) = v[clamp(c+1)] - v[clamp(c-1)] is
//   D^T v(c) = v[c-1] (c > 0) - v[c+1] (c < n-1) + v[c] (c = n-1) - v[c] (c = 0),
// so the normal, curl and centring reverses read their neighbours and write only their own cell
// or face: deterministic, no atomics and no scratch to clear.
#pragma once

#include "kernels.hpp"

namespace ds {

// smooth_epsilon, operators.cu:12
constexpr float kSmoothEpsilon = 1.0e-6f;

struct VortFields {   // per-cell vorticity (3) and smoothed magnitude
    const float* w[3];
    const float* mag;
};

// Neighbour offsets of cell i along dimension d, clamped to i at the grid's ends.
struct Clamp1 {
    uint32_t lo, hi;
    float slo, shi;   // signs of the transposed difference at (lo, hi)
};
__device__ __forceinline__ Clamp1 clamp_neighbours(uint32_t i, int coord, int size, uint32_t stride) {
    Clamp1 r;
    const bool has_lo = coord > 0, has_hi = coord < size - 1;
    r.lo = has_lo ? i - stride : i;
    r.hi = has_hi ? i + stride : i;
    r.slo = has_lo ? 1.0f : -1.0f;
    r.shi = has_hi ? -1.0f : 1.0f;
    return r;
}

__device__ __forceinline__ uint32_t cell_stride(const Geo& g, int d) {
    return d == 0 ? 1u : d == 1 ? (uint32_t)g.nx : (uint32_t)g.nx * (uint32_t)g.ny;
}

// Centred velocity component A of cell q = (x, y, z) (centered_velocity_kernel, operators.cu:678-687).
template <int A>
__device__ __forceinline__ float centred(const Geo& g, const uint8_t* __restrict__ mask, const CF3& u, uint32_t q, int x, int y,
                                         int z) {
    const uint32_t f = fidx(g, A, x, y, z);
    const uint32_t step = A == 0 ? 1u : A == 1 ? (uint32_t)g.nx : (uint32_t)g.nx * (uint32_t)g.ny;
    const float v = 0.5f * (u.p[A][f] + u.p[A][f + step]);
    return mask[q] != 0 ? 0.0f : v;
}

// curl_magnitude_kernel (operators.cu:689-701) at fluid cell (x, y, z).
__device__ __forceinline__ void curl_at(const Geo& g, const uint8_t* __restrict__ mask, const CF3& u, int x, int y, int z, float w[3]) {
    const int xm = max(x - 1, 0), xp = min(x + 1, g.nx - 1);
    const int ym = max(y - 1, 0), yp = min(y + 1, g.ny - 1);
    const int zm = max(z - 1, 0), zp = min(z + 1, g.nz - 1);
    const float s = 0.5f * g.inv_h;
    const float cz_yp = centred<2>(g, mask, u, cidx(g, x, yp, z), x, yp, z), cz_ym = centred<2>(g, mask, u, cidx(g, x, ym, z), x, ym, z);
    const float cy_zp = centred<1>(g, mask, u, cidx(g, x, y, zp), x, y, zp), cy_zm = centred<1>(g, mask, u, cidx(g, x, y, zm), x, y, zm);
    const float cx_zp = centred<0>(g, mask, u, cidx(g, x, y, zp), x, y, zp), cx_zm = centred<0>(g, mask, u, cidx(g, x, y, zm), x, y, zm);
    const float cz_xp = centred<2>(g, mask, u, cidx(g, xp, y, z), xp, y, z), cz_xm = centred<2>(g, mask, u, cidx(g, xm, y, z), xm, y, z);
    const float cy_xp = centred<1>(g, mask, u, cidx(g, xp, y, z), xp, y, z), cy_xm = centred<1>(g, mask, u, cidx(g, xm, y, z), xm, y, z);
    const float cx_yp = centred<0>(g, mask, u, cidx(g, x, yp, z), x, yp, z), cx_ym = centred<0>(g, mask, u, cidx(g, x, ym, z), x, ym, z);
    w[0] = (cz_yp - cz_ym - cy_zp + cy_zm) * s;
    w[1] = (cx_zp - cx_zm - cz_xp + cz_xm) * s;
    w[2] = (cy_xp - cy_xm - cx_yp + cx_ym) * s;
}

// vorticity_normal_kernel (operators.cu:703-720) at fluid cell i = (x, y, z): N and len.
__device__ __forceinline__ float normal_at(const Geo& g, const float* __restrict__ mag, uint32_t i, int x, int y, int z, float N[3]) {
    const float s = 0.5f * g.inv_h;
    const int c[3]{x, y, z};
    float grad[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const Clamp1 k = clamp_neighbours(i, c[d], dim_size(g, d), cell_stride(g, d));
        grad[d] = (mag[k.hi] - mag[k.lo]) * s;
    }
    const float len = sqrtf(grad[0] * grad[0] + grad[1] * grad[1] + grad[2] * grad[2] + kSmoothEpsilon * kSmoothEpsilon);
    N[0] = grad[0] / len;
    N[1] = grad[1] / len;
    N[2] = grad[2] / len;
    return len;
}

// add_vorticity_force_kernel (operators.cu:722-729): f += scale * (N x w), scale = confinement * h.
__device__ __forceinline__ void add_confinement_force(const VortFields& v, const float* __restrict__ mag, const Geo& g, float scale,
                                                      uint32_t i, int x, int y, int z, float f[3]) {
    float N[3];
    (void)normal_at(g, mag, i, x, y, z, N);
    const float w0 = v.w[0][i], w1 = v.w[1][i], w2 = v.w[2][i];
    f[0] += scale * (N[1] * w2 - N[2] * w1);
    f[1] += scale * (N[2] * w0 - N[0] * w2);
    f[2] += scale * (N[0] * w1 - N[1] * w0);
}

// vorticity_force_reverse_kernel + the per-cell part of normal_reverse_kernel (operators.cu:781-812)
// at fluid cell i with force adjoint fa: returns h * <N x w, fa> (the confinement adjoint term),
// writes the magnitude-gradient adjoint ga = (N_bar - N (N . N_bar)) / len with N_bar =
// scale (w x fa), and the direct vorticity adjoint t = scale (fa x N).
__device__ __forceinline__ double confinement_reverse_at(const Geo& g, const VortFields& v, uint32_t i, int x, int y, int z, float scale,
                                                         const float fa[3], float ga[3], float t[3]) {
    float N[3];
    const float len = normal_at(g, v.mag, i, x, y, z, N);
    const float w0 = v.w[0][i], w1 = v.w[1][i], w2 = v.w[2][i];
    const float c0 = N[1] * w2 - N[2] * w1, c1 = N[2] * w0 - N[0] * w2, c2 = N[0] * w1 - N[1] * w0;
    const double part = (double)g.h * ((double)c0 * fa[0] + (double)c1 * fa[1] + (double)c2 * fa[2]);
    const float nb0 = scale * (w1 * fa[2] - w2 * fa[1]);
    const float nb1 = scale * (w2 * fa[0] - w0 * fa[2]);
    const float nb2 = scale * (w0 * fa[1] - w1 * fa[0]);
    const float proj = N[0] * nb0 + N[1] * nb1 + N[2] * nb2;
    ga[0] = (nb0 - N[0] * proj) / len;
    ga[1] = (nb1 - N[1] * proj) / len;
    ga[2] = (nb2 - N[2] * proj) / len;
    t[0] = scale * (fa[1] * N[2] - fa[2] * N[1]);
    t[1] = scale * (fa[2] * N[0] - fa[0] * N[2]);
    t[2] = scale * (fa[0] * N[1] - fa[1] * N[0]);
    return part;
}

// D^T of component array v along dimension d at cell i (coordinate coord).
__device__ __forceinline__ float transposed_difference(const Geo& g, const float* __restrict__ v, uint32_t i, int coord, int d) {
    const Clamp1 k = clamp_neighbours(i, coord, dim_size(g, d), cell_stride(g, d));
    return k.slo * v[k.lo] + k.shi * v[k.hi];
}

}  // namespace ds
