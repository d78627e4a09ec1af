// rbgs.hip — the reference's own pressure solve on gfx950 (DS_PROJECTION_RBGS): red-black
// Gauss-Seidel from p = 0 for `pressure_iterations` sweeps of each colour (projection_forward,
// solvers/smoke/operators.cu:1004-1013; rbgs_kernel, :885-902), its unrolled reverse
// (projection_vjp, :1020-1031; rbgs_reverse_kernel, :935-955) and its tangent (projection_jvp,
// :1015-1018: the same sweeps with the fixed boundary values zeroed).
//
// A sweep updates one colour, so one launch covers only that colour's cells (x advances by 2
// within a row). The reverse of a sweep of colour P scatters p_bar[i] / diag_i (i of colour P) to
// i's connected neighbours, adds it to rhs_bar[i] and clears p_bar[i]; here it is a gather: the
// neighbours (colour 1 - P) read p_bar[i] / diag_i, and the clearing and the rhs_bar update of
// colour P are deferred into the next launch, which runs over colour P. So each launch writes
// only its own colour: no atomics, deterministic, 2 * pressure_iterations + 1 launches.
//
// A periodic axis with an odd cell count puts cells of one colour side by side across the seam;
// the reference's kernel then reads a neighbour that the same launch writes, and so does this one.
#include <hip/hip_runtime.h>

#include "context.hpp"
#include "kernels.hpp"
#include "rhs.hpp"

namespace ds {

namespace {

// pressure_neighbor (operators.cu:852-873): the neighbour of (x, y, z) along dimension d in
// direction dir: a connected fluid cell (index j), or a fixed-value face (value), or neither.
__device__ __forceinline__ void rbgs_neighbour(const Geo& g, const SBound& b, const uint8_t* __restrict__ mask, int x, int y, int z, int d,
                                               int dir, uint32_t& j, bool& connected, bool& fixed, float& fixed_value) {
    int c[3]{x, y, z};
    c[d] += dir;
    const int size = dim_size(g, d);
    connected = false;
    fixed = false;
    if (c[d] < 0 || c[d] >= size) {
        if (b.modes[2 * d] == 2u && b.modes[2 * d + 1] == 2u) {
            c[d] = c[d] < 0 ? c[d] + size : c[d] - size;
        } else {
            const int face = 2 * d + (dir > 0 ? 1 : 0);
            fixed = b.modes[face] == 0u;
            fixed_value = b.values[face];
            return;
        }
    }
    j = cidx(g, c[0], c[1], c[2]);
    connected = mask[j] == 0;
}

// Diagonal of every cell (connected neighbours + fixed faces), 0 for solids, the anchor and cells
// with no neighbour: the cells the sweeps leave alone.
__global__ __launch_bounds__(256) void rbgs_diag_kernel(Geo g, const SBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                        uint32_t anchor, uint8_t* __restrict__ diag) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (mask[i] != 0 || i == anchor) {
        diag[i] = 0;
        return;
    }
    int x, y, z;
    decode_cell(g, i, x, y, z);
    uint32_t count = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int dir = -1; dir <= 1; dir += 2) {
            uint32_t j;
            bool connected, fixed;
            float fv;
            rbgs_neighbour(g, *bp, mask, x, y, z, d, dir, j, connected, fixed, fv);
            count += (connected || fixed) ? 1u : 0u;
        }
    diag[i] = (uint8_t)count;
}

// Cell (x, y, z) of colour `parity` handled by thread t of a colour launch: rows of (nx + 1) / 2.
__device__ __forceinline__ bool colour_cell(const Geo& g, uint32_t t, uint32_t parity, int& x, int& y, int& z) {
    const uint32_t half = ((uint32_t)g.nx + 1u) / 2u;
    const uint32_t row = t / half;
    const uint32_t k = t - row * half;
    y = (int)(row % (uint32_t)g.ny);
    z = (int)(row / (uint32_t)g.ny);
    if (z >= g.nz) return false;
    x = (int)(2u * k + ((parity + (uint32_t)y + (uint32_t)z) & 1u));
    return x < g.nx;
}

// Right-hand side of the sweeps: b = -h^2 div(u) / dt on fluid non-anchor cells, 0 elsewhere
// (pressure_rhs_kernel, operators.cu:875-883); with Adjoint, the projection reverse as a gather
// (project_velocity_reverse_kernel, operators.cu:922-933) on fluid cells. Also zeroes x.
template <int Kind>
__global__ __launch_bounds__(256) void rbgs_rhs_kernel(Geo g, const uint8_t* __restrict__ mask, uint32_t anchor, RhsIn in,
                                                       float* __restrict__ b, float* __restrict__ x0, float* __restrict__ x1) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    x0[i] = 0.0f;
    x1[i] = 0.0f;
    const bool skip = mask[i] != 0 || (Kind != kRhsAdjoint && i == anchor);
    if (skip) {
        b[i] = 0.0f;
        return;
    }
    int x, y, z;
    decode_cell(g, i, x, y, z);
    b[i] = cell_rhs<Kind>(g, in, mask, nullptr, i, x, y, z);
}

// rbgs_kernel (operators.cu:885-902) over the cells of one colour.
__global__ __launch_bounds__(256) void rbgs_sweep_kernel(Geo g, const SBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                         const uint8_t* __restrict__ diag, const float* __restrict__ rhs, float* p,
                                                         uint32_t parity) {
    int x, y, z;
    if (!colour_cell(g, blockIdx.x * blockDim.x + threadIdx.x, parity, x, y, z)) return;
    const uint32_t i = cidx(g, x, y, z);
    const uint32_t d = diag[i];
    if (d == 0u) return;
    float sum = rhs[i];
#pragma unroll
    for (int dim = 0; dim < 3; ++dim)
#pragma unroll
        for (int dir = -1; dir <= 1; dir += 2) {
            uint32_t j;
            bool connected, fixed;
            float fv = 0.0f;
            rbgs_neighbour(g, *bp, mask, x, y, z, dim, dir, j, connected, fixed, fv);
            if (connected) sum += p[j];
            else if (fixed) sum += fv;
        }
    p[i] = sum / (float)d;
}

// One launch of the reverse over the cells of colour `parity`: with Finish, the deferred second
// half of the previous reverse sweep (that colour's rhs_bar += p_bar / diag, p_bar = 0); with
// Gather, the first half of the next one (p_bar += sum over connected neighbours of p_bar / diag).
template <bool Finish, bool Gather>
__global__ __launch_bounds__(256) void rbgs_reverse_kernel(Geo g, const SBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                           const uint8_t* __restrict__ diag, float* pbar, float* __restrict__ rbar,
                                                           uint32_t parity) {
    int x, y, z;
    if (!colour_cell(g, blockIdx.x * blockDim.x + threadIdx.x, parity, x, y, z)) return;
    const uint32_t i = cidx(g, x, y, z);
    float own = pbar[i];
    if (Finish) {
        const uint32_t d = diag[i];
        if (d != 0u) {
            rbar[i] += own / (float)d;
            own = 0.0f;
        }
    }
    if (Gather && mask[i] == 0) {
        float s = 0.0f;
#pragma unroll
        for (int dim = 0; dim < 3; ++dim)
#pragma unroll
            for (int dir = -1; dir <= 1; dir += 2) {
                uint32_t j;
                bool connected, fixed;
                float fv;
                rbgs_neighbour(g, *bp, mask, x, y, z, dim, dir, j, connected, fixed, fv);
                if (connected) {
                    const uint32_t dj = diag[j];
                    if (dj != 0u) s += pbar[j] / (float)dj;
                }
            }
        own += s;
    }
    pbar[i] = own;
}

uint32_t colour_blocks(const Geo& g) {
    const uint64_t threads = (uint64_t)(((uint32_t)g.nx + 1u) / 2u) * (uint64_t)g.ny * (uint64_t)g.nz;
    return (uint32_t)((threads + kBlock - 1) / kBlock);
}

}  // namespace

void rbgs_prepare(ds_context* c) {
    rbgs_diag_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->d_pb, c->mask, (uint32_t)c->anchor, c->rbgs_diag);
    check_launch("rbgs_diag_kernel");
}

// Forward (kind 1) or tangent (kind 3): p = RBGS(rhs of u) from 0, into xh (xl = 0). Adjoint
// (kind 2): xh = rhs_bar of the unrolled reverse for the projection reverse of u (+ extra), xl = 0.
ds_solve_info rbgs_solve_kind(ds_context* c, int kind, const float* const u[3], const float* const extra[3], float* xh, float* xl) {
    const Geo& g = c->geo;
    RhsIn in{};
    for (int a = 0; a < 3; ++a) {
        in.u[a] = u[a];
        in.extra[a] = extra ? extra[a] : nullptr;
    }
    float* b = c->pcg.b;
    const uint32_t G = xcd_blocks_for(g.n);
    const uint32_t anchor = (uint32_t)c->anchor;
    if (kind == kRhsAdjoint) rbgs_rhs_kernel<kRhsAdjoint><<<G, kBlock, 0, c->stream>>>(g, c->mask, anchor, in, b, xh, xl);
    else rbgs_rhs_kernel<kRhsForward><<<G, kBlock, 0, c->stream>>>(g, c->mask, anchor, in, b, xh, xl);
    check_launch("rbgs_rhs_kernel");
    const uint32_t iterations = c->config.pressure_iterations;
    const uint32_t CB = colour_blocks(g);
    if (kind != kRhsAdjoint) {
        const SBound* bp = kind == kRhsTangent ? c->d_pb_zero : c->d_pb;
        for (uint32_t it = 0; it < iterations; ++it) {
            rbgs_sweep_kernel<<<CB, kBlock, 0, c->stream>>>(g, bp, c->mask, c->rbgs_diag, b, xh, 0u);
            rbgs_sweep_kernel<<<CB, kBlock, 0, c->stream>>>(g, bp, c->mask, c->rbgs_diag, b, xh, 1u);
        }
        check_launch("rbgs_sweep_kernel");
    } else if (iterations > 0) {
        // b holds p_bar; xh accumulates rhs_bar. Reverse order of the sweeps: colour 1, then 0, ...
        rbgs_reverse_kernel<false, true><<<CB, kBlock, 0, c->stream>>>(g, c->d_pb, c->mask, c->rbgs_diag, b, xh, 0u);
        for (uint32_t k = 1; k < 2 * iterations; ++k)
            rbgs_reverse_kernel<true, true><<<CB, kBlock, 0, c->stream>>>(g, c->d_pb, c->mask, c->rbgs_diag, b, xh, k & 1u);
        rbgs_reverse_kernel<true, false><<<CB, kBlock, 0, c->stream>>>(g, c->d_pb, c->mask, c->rbgs_diag, b, xh, 0u);
        check_launch("rbgs_reverse_kernel");
    }
    ds_solve_info si{};
    si.iterations = iterations;
    si.converged = 1;
    return si;
}

}  // namespace ds
