// kernels.hpp — kernel argument bundles and the launch wrappers shared by the operator,
// step and trajectory translation units.
#pragma once

#include "context.hpp"

namespace ds {

struct F3 {
    float* p[3];
};
struct CF3 {
    const float* p[3];
};
struct Params {   // device pointers to one float each (Parameters, data.ixx:129-134)
    const float* ambient;
    const float* alpha_rho;
    const float* alpha_t;
    const float* confinement;   // null unless vorticity confinement is on
};

// Force adjoint of fluid cell i from the forced-velocity adjoint (integrate_vjp_kernel,
// operators.cu:366-379, as a gather): each face adds dt * u_bar / count to each in-domain fluid
// cell beside it, count = number of such cells.
__device__ __forceinline__ void force_gather(const Geo& g, const uint8_t* __restrict__ mask, const CF3& fa, uint32_t i, int x, int y,
                                             int z, float f[3]) {
    const int c[3]{x, y, z};
    const uint32_t stride[3]{1u, (uint32_t)g.nx, (uint32_t)g.nx * (uint32_t)g.ny};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int n_a = dim_size(g, a);
        const uint32_t lo = fidx(g, a, x, y, z);
        const uint32_t hi = fidx(g, a, x + (a == 0), y + (a == 1), z + (a == 2));
        const float cnt_lo = 1.0f + ((c[a] > 0 && mask[i - stride[a]] == 0) ? 1.0f : 0.0f);
        const float cnt_hi = 1.0f + ((c[a] + 1 < n_a && mask[i + stride[a]] == 0) ? 1.0f : 0.0f);
        f[a] = g.dt * fa.p[a][lo] / cnt_lo + g.dt * fa.p[a][hi] / cnt_hi;
    }
}

// adjacent_solid, operators.cu:441-450
template <int A>
__device__ __forceinline__ bool adjacent_solid(const Geo& g, const uint8_t* __restrict__ mask, int x, int y, int z) {
    const int coord = A == 0 ? x : A == 1 ? y : z;
    const int n_a = dim_size(g, A);
    if (coord < n_a && mask[cidx(g, x, y, z)] != 0) return true;
    if (coord > 0 && mask[cidx(g, x - (A == 0), y - (A == 1), z - (A == 2))] != 0) return true;
    return false;
}

// Result of the constraint of face (A, x, y, z) (constrain_velocity_forward_kernel,
// operators.cu:466-481): 0 = keep the advected value, 1 = fixed value in `value`,
// 2 = periodic duplicate (use face 0 of the axis).
template <int A>
__device__ __forceinline__ int constraint(const Geo& g, const VBound& b, const uint8_t* __restrict__ mask, const CF3& collider,
                                          bool has_solids, int x, int y, int z, uint32_t local, float& value) {
    const int coord = A == 0 ? x : A == 1 ? y : z;
    const int n_a = dim_size(g, A);
    if (coord == 0 || coord == n_a) {
        const bool hi = coord == n_a;
        const uint32_t mode = hi ? b.modes[2 * A + 1] : b.modes[2 * A];
        if (mode == 0u || mode == 2u) {
            value = hi ? b.values[(2 * A + 1) * 3 + A] : b.values[(2 * A) * 3 + A];
            return 1;
        }
    }
    if (has_solids && adjacent_solid<A>(g, mask, x, y, z)) {
        value = collider.p[A] ? collider.p[A][local] : 0.0f;
        return 1;
    }
    if (vperiodic(b, A) && coord == n_a) return 2;
    return 0;
}

Params params_of(const ds_parameters* p);
CF3 collider_of(const ds_context* c);

void launch_source_buoyancy(ds_context* c, const float* rho, const float* T, const float* s_rho, const float* s_t, CF3 e,
                            const Params& prm, float* rho_s, float* t_s, F3 f);
void launch_integrate(ds_context* c, CF3 u, CF3 f, F3 out);
void launch_advect_velocity(ds_context* c, CF3 u, F3 out, bool constrain);
// constrain_velocity_forward (operators.cu:466-481) with the given boundary copy and collider
// velocity (null: 0), as the tangent constraint uses it (operators.cu:633-638).
void launch_constrain(ds_context* c, const VBound* bp, CF3 collider, CF3 in, F3 out);
void launch_project(ds_context* c, CF3 u, const float* ph, const float* pl, F3 out);
void launch_advect_scalars(ds_context* c, int nf, int first_field, const float* src0, const float* src1, CF3 u, float* out0,
                           float* out1);
void copy_pair(ds_context* c, const float* hi, const float* lo, float* out_hi, float* out_lo);

// Model::jvp_step (jvp.hip): next_t = the tangent of the next state for the tangents of the state,
// the control and the parameters (pt); state / next / cache as forward_step left them.
ds_solve_info jvp_step(ds_context* c, const ds_parameters* p, const ds_parameters* pt, const ds_state* state, const ds_state* next,
                       const ds_step_cache* cache, const ds_state* state_t, const ds_control* control_t, ds_state* next_t);

// Vorticity confinement (vorticity.hip). The curl pass writes c->vort_w / c->vort_mag from the
// face velocity u; the force and reverse passes read them.
void launch_vorticity_curl(ds_context* c, CF3 u);
// out += the confinement force (fluid cells).
void launch_vorticity_force(ds_context* c, const Params& prm, F3 out);
// Reverse from a cell force adjoint fa (after launch_vorticity_curl on the same velocity):
// u_adj += the velocity adjoint, *conf_adj += the confinement adjoint (conf_adj may be null).
void launch_vorticity_vjp(ds_context* c, const Params& prm, CF3 fa, F3 u_adj, double* conf_adj);
// The magnitude and curl/centring reverses, after the fused force reverse filled c->vort_ga / c->vort_t.
void launch_vorticity_vjp_tail(ds_context* c, F3 u_adj);

}  // namespace ds
