// rhs.hpp — right-hand sides of the pressure solves, shared by the PCG (pcg.hip) and the
// reference's red-black Gauss-Seidel mode (rbgs.hip).
#pragma once

#include "context.hpp"

namespace ds {

struct RhsIn {
    const float* u[3];        // forward: advected velocity; adjoint: projected-velocity adjoint
    const float* extra[3];    // adjoint only: scalar-advection contributions (may be null)
};

// kRhsTangent: the forward right-hand side of a tangent velocity, without the fixed-face term
// (the tangent of a fixed pressure value is 0, operators.cu:1015-1018); a host-side kind only.
enum RhsKind : int { kRhsGiven = 0, kRhsForward = 1, kRhsAdjoint = 2, kRhsTangent = 3 };

// Right-hand side of one cell.
// forward (pressure_rhs_kernel, operators.cu:875-883): b = -h^2 div(u)/dt + fixed faces;
// adjoint (project_velocity_reverse_kernel, operators.cu:922-933, as a gather): p_bar_i =
//   (dt/h) * sum_a [u_bar(i+e_a) - u_bar(i)] over projectable faces (both cells fluid, in domain).
template <int Kind>
__device__ __forceinline__ float cell_rhs(const Geo& g, const RhsIn& in, const uint8_t* __restrict__ mask,
                                          const float* __restrict__ fixed_term, uint32_t i, int x, int y, int z) {
    const uint32_t ex0 = (uint32_t)g.nx + 1u, ey1 = (uint32_t)g.ny + 1u;
    const uint32_t f0 = (uint32_t)x + ex0 * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
    const uint32_t f1 = (uint32_t)x + (uint32_t)g.nx * ((uint32_t)y + ey1 * (uint32_t)z);
    const uint32_t f2 = i;
    const uint32_t sz = (uint32_t)g.nx * (uint32_t)g.ny;
    if (Kind == kRhsForward) {
        const float sum = (in.u[0][f0 + 1u] - in.u[0][f0]) + (in.u[1][f1 + (uint32_t)g.nx] - in.u[1][f1]) +
                          (in.u[2][f2 + sz] - in.u[2][f2]);
        const float divergence = sum * g.inv_h;
        float b = -g.h * g.h * divergence / g.dt;
        if (fixed_term) b += fixed_term[i];
        return b;
    } else {
        auto ua = [&](int a, uint32_t f) { return in.u[a][f] + (in.extra[a] ? in.extra[a][f] : 0.0f); };
        float s = 0.0f;
        if (x + 1 < g.nx && mask[i + 1u] == 0) s += ua(0, f0 + 1u);
        if (x > 0 && mask[i - 1u] == 0) s -= ua(0, f0);
        if (y + 1 < g.ny && mask[i + (uint32_t)g.nx] == 0) s += ua(1, f1 + (uint32_t)g.nx);
        if (y > 0 && mask[i - (uint32_t)g.nx] == 0) s -= ua(1, f1);
        if (z + 1 < g.nz && mask[i + sz] == 0) s += ua(2, f2 + sz);
        if (z > 0 && mask[i - sz] == 0) s -= ua(2, f2);
        return g.dt * s * g.inv_h;
    }
}

}  // namespace ds
