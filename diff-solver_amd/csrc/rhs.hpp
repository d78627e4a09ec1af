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
        // every face value and neighbour flag is loaded unconditionally (the six faces of a cell
        // exist; a neighbour outside the domain reads the cell itself and is masked off), so the
        // loads issue together; an excluded face adds zero
        const bool xp = x + 1 < g.nx, xm = x > 0, yp = y + 1 < g.ny, ym = y > 0, zp = z + 1 < g.nz, zm = z > 0;
        const uint8_t mxp = mask[xp ? i + 1u : i], mxm = mask[xm ? i - 1u : i];
        const uint8_t myp = mask[yp ? i + (uint32_t)g.nx : i], mym = mask[ym ? i - (uint32_t)g.nx : i];
        const uint8_t mzp = mask[zp ? i + sz : i], mzm = mask[zm ? i - sz : i];
        const float uxp = ua(0, f0 + 1u), uxm = ua(0, f0), uyp = ua(1, f1 + (uint32_t)g.nx), uym = ua(1, f1);
        const float uzp = ua(2, f2 + sz), uzm = ua(2, f2);
        float s = 0.0f;
        s += xp && mxp == 0 ? uxp : 0.0f;
        s -= xm && mxm == 0 ? uxm : 0.0f;
        s += yp && myp == 0 ? uyp : 0.0f;
        s -= ym && mym == 0 ? uym : 0.0f;
        s += zp && mzp == 0 ? uzp : 0.0f;
        s -= zm && mzm == 0 ? uzm : 0.0f;
        return g.dt * s * g.inv_h;
    }
}

}  // namespace ds
