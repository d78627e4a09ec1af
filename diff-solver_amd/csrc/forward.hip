// forward.hip — forward operators of the smoke step on gfx950 (solvers/smoke/operators.cu:286-507,
// 904-920): sources, buoyancy, face integration, semi-Lagrangian velocity self-advection with the
// boundary/collider constraint fused in, projection, and density/temperature advection fused
// into one trace. Face kernels run over the concatenated (x, y, z) face index space in one launch.
#include <hip/hip_runtime.h>

#include "context.hpp"
#include "kernels.hpp"
#include "vorticity.hpp"

namespace ds {

namespace {

// source_forward_kernel, operators.cu:286-290
__global__ __launch_bounds__(256) void source_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ s,
                                                     const float* __restrict__ src, float* __restrict__ out) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    out[i] = mask[i] == 0 ? s[i] + g.dt * src[i] : 0.0f;
}

// buoyancy_forward_kernel, operators.cu:304-313
__global__ __launch_bounds__(256) void buoyancy_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ rho,
                                                       const float* __restrict__ T, CF3 e, Params prm, F3 f) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (mask[i] != 0) {
        f.p[0][i] = f.p[1][i] = f.p[2][i] = 0.0f;
        return;
    }
    const float amb = *prm.ambient, ar = *prm.alpha_rho, at = *prm.alpha_t;
    f.p[0][i] = e.p[0][i];
    f.p[1][i] = e.p[1][i] + ar * rho[i] + at * (T[i] - amb);
    f.p[2][i] = e.p[2][i];
}

// source (rho and T) and buoyancy fused: one pass over the cell fields (Model::forward_step
// steps 1-3, model.cpp:342-344). Vort: the confinement force is added in the same pass
// (ForceAssemblyOperator::forward, operators.cpp:125-128), from the vorticity fields that the
// curl pass wrote for the step's input velocity.
template <bool Vort>
__global__ __launch_bounds__(256) void source_buoyancy_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ rho,
                                                              const float* __restrict__ T, const float* __restrict__ s_rho,
                                                              const float* __restrict__ s_t, CF3 e, Params prm, VortFields vort,
                                                              float* __restrict__ rho_s, float* __restrict__ t_s, F3 f) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (mask[i] != 0) {
        rho_s[i] = 0.0f;
        t_s[i] = 0.0f;
        f.p[0][i] = f.p[1][i] = f.p[2][i] = 0.0f;
        return;
    }
    const float amb = *prm.ambient, ar = *prm.alpha_rho, at = *prm.alpha_t;
    const float rs = rho[i] + g.dt * s_rho[i];
    const float ts = T[i] + g.dt * s_t[i];
    rho_s[i] = rs;
    t_s[i] = ts;
    float fv[3]{e.p[0][i], e.p[1][i] + ar * rs + at * (ts - amb), e.p[2][i]};
    if constexpr (Vort) {
        int x, y, z;
        decode_cell(g, i, x, y, z);
        add_confinement_force(vort, vort.mag, g, *prm.confinement * g.h, i, x, y, z, fv);
    }
    f.p[0][i] = fv[0];
    f.p[1][i] = fv[1];
    f.p[2][i] = fv[2];
}

// averaged_center_force + integrate_kernel, operators.cu:339-364
__global__ __launch_bounds__(256) void integrate_kernel(Geo g, const uint8_t* __restrict__ mask, CF3 u, CF3 f, F3 out) {
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        const int coord = A == 0 ? x : A == 1 ? y : z;
        const int n_a = dim_size(g, A);
        const float* fa = f.p[A];
        // both neighbour cells' flag and force loaded at once: a side outside the domain reads the
        // other side's cell and is not counted
        const bool h0 = coord > 0, h1 = coord < n_a;
        const uint32_t cl = cidx(g, x - (A == 0 && h0), y - (A == 1 && h0), z - (A == 2 && h0));   // low side, else the cell
        const uint32_t cc = h1 ? cidx(g, x, y, z) : cl;
        const uint32_t c0 = h0 ? cl : cc;
        const uint8_t m0 = mask[c0], m1 = mask[cc];
        const float f0 = fa[c0], f1 = fa[cc], ua = u.p[A][local];
        float sum = 0.0f, count = 0.0f;
        if (h0 && m0 == 0) { sum += f0; count += 1.0f; }
        if (h1 && m1 == 0) { sum += f1; count += 1.0f; }
        out.p[A][local] = ua + g.dt * (count == 0.0f ? 0.0f : sum / count);
    });
}

// advect_velocity_forward_kernel (operators.cu:381-388), optionally with the constraint fused:
// constrained faces skip the trace; a periodic duplicate face traces face 0's position.
// 8 waves per SIMD (63 registers, nothing spilled): the trace's dependent gathers want waves.
template <bool Constrain, bool HasSolids>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void advect_velocity_kernel(Geo g, const VBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                              Vel3 u, CF3 collider, F3 out) {
    const VBound& b = *bp;
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        if (Constrain) {
            float value;
            const int kind = constraint<A>(g, b, mask, collider, HasSolids, x, y, z, local, value);
            if (kind == 1) {
                out.p[A][local] = value;
                return;
            }
            if (kind == 2) {
                if (A == 0) x = 0;
                if (A == 1) y = 0;
                if (A == 2) z = 0;
            }
        }
        const float sx = (x + (A == 0 ? 0.0f : 0.5f)) * g.h;
        const float sy = (y + (A == 1 ? 0.0f : 0.5f)) * g.h;
        const float sz = (z + (A == 2 ? 0.0f : 0.5f)) * g.h;
        const TraceOut t = trace_rk2<HasSolids>(sx, sy, sz, u, mask, g, b);
        out.p[A][local] = sample_face<A, false>(u.u[A], t.px, t.py, t.pz, g, b).value;
    });
}

// constrain_velocity_forward_kernel, operators.cu:466-481
__global__ __launch_bounds__(256) void constrain_kernel(Geo g, const VBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                        uint32_t has_solids, CF3 collider, CF3 in, F3 out) {
    const VBound& b = *bp;
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        float value;
        const int kind = constraint<A>(g, b, mask, collider, has_solids != 0, x, y, z, local, value);
        if (kind == 1) out.p[A][local] = value;
        else if (kind == 2) out.p[A][local] = in.p[A][fidx(g, A, A == 0 ? 0 : x, A == 1 ? 0 : y, A == 2 ? 0 : z)];
        else out.p[A][local] = in.p[A][local];
    });
}

// projectable_face + project_velocity_kernel, operators.cu:904-920; the pressure is the
// compensated pair hi + lo (lo may be null): the difference is taken per part.
__global__ __launch_bounds__(256) void project_kernel(Geo g, const uint8_t* __restrict__ mask, CF3 u, const float* __restrict__ ph,
                                                      const float* __restrict__ pl, F3 out) {
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        const int coord = A == 0 ? x : A == 1 ? y : z;
        const int n_a = dim_size(g, A);
        const float ua = u.p[A][local];
        // flags and pressures of both cells loaded at once (a domain face reads cell 0 and keeps ua)
        const bool edge = coord == 0 || coord == n_a;
        const uint32_t stride = A == 0 ? 1u : A == 1 ? (uint32_t)g.nx : (uint32_t)g.nx * (uint32_t)g.ny;
        const uint32_t c1 = edge ? 0u : cidx(g, x, y, z);
        const uint32_t c0 = edge ? 0u : c1 - stride;
        const uint8_t m0 = mask[c0], m1 = mask[c1];
        const float p0 = ph[c0], p1 = ph[c1];
        const float l0 = pl ? pl[c0] : 0.0f, l1 = pl ? pl[c1] : 0.0f;
        if (edge || m0 != 0 || m1 != 0) {
            out.p[A][local] = ua;
            return;
        }
        float dp = p1 - p0;
        if (pl) dp += l1 - l0;
        out.p[A][local] = ua - g.dt * dp / g.h;
    });
}

// advect_scalar_forward_kernel (operators.cu:499-507) for NF fields sharing one RK2 trace.
// 8 waves per SIMD, as advect_velocity_kernel.
template <int NF, bool HasSolids>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void advect_scalar_kernel(Geo g, const VBound* __restrict__ vbp, const SBound* __restrict__ sb0p,
                                                            const SBound* __restrict__ sb1p, const uint8_t* __restrict__ mask,
                                                            const float* __restrict__ col0, const float* __restrict__ col1,
                                                            const float* __restrict__ src0, const float* __restrict__ src1, Vel3 u,
                                                            float* __restrict__ out0, float* __restrict__ out1) {
    const VBound& vb = *vbp;
    const SBound& sb0 = *sb0p;
    const SBound& sb1 = *sb1p;
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (HasSolids && mask[i] != 0) {
        out0[i] = col0 ? col0[i] : 0.0f;
        if (NF == 2) out1[i] = col1 ? col1[i] : 0.0f;
        return;
    }
    int x, y, z;
    decode_cell(g, i, x, y, z);
    const TraceOut t = trace_rk2<HasSolids>((x + 0.5f) * g.h, (y + 0.5f) * g.h, (z + 0.5f) * g.h, u, mask, g, vb);
    out0[i] = sample_scalar<false>(src0, t.px, t.py, t.pz, g, sb0).value;
    if (NF == 2) out1[i] = sample_scalar<false>(src1, t.px, t.py, t.pz, g, sb1).value;
}

__global__ __launch_bounds__(256) void combine_pair_kernel(uint32_t n, const float* __restrict__ hi, const float* __restrict__ lo,
                                                           float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = hi[i] + lo[i];
}

__global__ __launch_bounds__(256) void copy_pair_kernel(uint32_t n, const float* __restrict__ hi, const float* __restrict__ lo,
                                                        float* __restrict__ out_hi, float* __restrict__ out_lo) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out_hi[i] = hi[i];
        out_lo[i] = lo[i];
    }
}

}  // namespace

void combine_pair(ds_context* c, const float* hi, const float* lo, float* out) {
    combine_pair_kernel<<<blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo.n, hi, lo, out);
    check_launch("combine_pair_kernel");
}

void copy_pair(ds_context* c, const float* hi, const float* lo, float* out_hi, float* out_lo) {
    copy_pair_kernel<<<blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo.n, hi, lo, out_hi, out_lo);
    check_launch("copy_pair_kernel");
}

Params params_of(const ds_parameters* p) {
    require(p && p->ambient_temperature && p->density_buoyancy && p->temperature_buoyancy, "parameters must be device pointers");
    return Params{p->ambient_temperature, p->density_buoyancy, p->temperature_buoyancy, p->vorticity_confinement};
}

CF3 collider_of(const ds_context* c) { return CF3{{c->collider_velocity[0], c->collider_velocity[1], c->collider_velocity[2]}}; }

void launch_source_buoyancy(ds_context* c, const float* rho, const float* T, const float* s_rho, const float* s_t, CF3 e,
                            const Params& prm, float* rho_s, float* t_s, F3 f) {
    const uint32_t G = xcd_blocks_for(c->geo.n);
    if (c->vorticity) {
        require(prm.confinement != nullptr, "parameters: vorticity_confinement must be a device pointer with confinement on");
        const VortFields v{{c->vort_w[0], c->vort_w[1], c->vort_w[2]}, c->vort_mag};
        source_buoyancy_kernel<true><<<G, kBlock, 0, c->stream>>>(c->geo, c->mask, rho, T, s_rho, s_t, e, prm, v, rho_s, t_s, f);
    } else {
        source_buoyancy_kernel<false><<<G, kBlock, 0, c->stream>>>(c->geo, c->mask, rho, T, s_rho, s_t, e, prm, VortFields{}, rho_s, t_s, f);
    }
    check_launch("source_buoyancy_kernel");
}

void launch_integrate(ds_context* c, CF3 u, CF3 f, F3 out) {
    integrate_kernel<<<xcd_blocks_for(c->geo.foff[3]), kBlock, 0, c->stream>>>(c->geo, c->mask, u, f, out);
    check_launch("integrate_kernel");
}

void launch_advect_velocity(ds_context* c, CF3 u, F3 out, bool constrain) {
    const Vel3 v{{u.p[0], u.p[1], u.p[2]}};
    const uint32_t G = xcd_blocks_for(c->geo.foff[3]);
    const CF3 col = collider_of(c);
    if (constrain) {
        if (c->has_solids) advect_velocity_kernel<true, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->mask, v, col, out);
        else advect_velocity_kernel<true, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->mask, v, col, out);
    } else {
        if (c->has_solids) advect_velocity_kernel<false, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->mask, v, col, out);
        else advect_velocity_kernel<false, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->mask, v, col, out);
    }
    check_launch("advect_velocity_kernel");
}

void launch_constrain(ds_context* c, const VBound* bp, CF3 collider, CF3 in, F3 out) {
    constrain_kernel<<<xcd_blocks_for(c->geo.foff[3]), kBlock, 0, c->stream>>>(c->geo, bp, c->mask, c->has_solids ? 1u : 0u, collider, in,
                                                                                out);
    check_launch("constrain_kernel");
}

void launch_project(ds_context* c, CF3 u, const float* ph, const float* pl, F3 out) {
    project_kernel<<<xcd_blocks_for(c->geo.foff[3]), kBlock, 0, c->stream>>>(c->geo, c->mask, u, ph, pl, out);
    check_launch("project_kernel");
}

void launch_advect_scalars(ds_context* c, int nf, int first_field, const float* src0, const float* src1, CF3 u, float* out0,
                           float* out1) {
    const Vel3 v{{u.p[0], u.p[1], u.p[2]}};
    const uint32_t G = xcd_blocks_for(c->geo.n);
    const SBound* b0 = first_field == 0 ? c->d_db : c->d_tb;
    const float* col0 = first_field == 0 ? c->collider_density : c->collider_temperature;
    if (nf == 2) {
        if (c->has_solids)
            advect_scalar_kernel<2, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_db, c->d_tb, c->mask, c->collider_density,
                                                                      c->collider_temperature, src0, src1, v, out0, out1);
        else
            advect_scalar_kernel<2, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_db, c->d_tb, c->mask, c->collider_density,
                                                                       c->collider_temperature, src0, src1, v, out0, out1);
    } else {
        if (c->has_solids)
            advect_scalar_kernel<1, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, b0, b0, c->mask, col0, col0, src0, src0, v, out0,
                                                                      out0);
        else
            advect_scalar_kernel<1, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, b0, b0, c->mask, col0, col0, src0, src0, v, out0,
                                                                       out0);
    }
    check_launch("advect_scalar_kernel");
}

}  // namespace ds

using namespace ds;

extern "C" {

ds_status ds_source_forward(ds_context* c, const float* state, const float* source, float* out) {
    if (!c || !state || !source || !out) {
        set_error("ds_source_forward: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        source_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, state, source, out);
        check_launch("source_kernel");
    });
}

ds_status ds_buoyancy_forward(ds_context* c, const float* density, const float* temperature, float* const e[3],
                              const ds_parameters* params, float* force[3]) {
    if (!c || !density || !temperature || !e || !force) {
        set_error("ds_buoyancy_forward: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        const Params prm = params_of(params);
        buoyancy_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, density, temperature,
                                                                             CF3{{e[0], e[1], e[2]}}, prm,
                                                                             F3{{force[0], force[1], force[2]}});
        check_launch("buoyancy_kernel");
    });
}

ds_status ds_integrate_velocity_forward(ds_context* c, float* const velocity[3], float* const force[3], float* out[3]) {
    if (!c || !velocity || !force || !out) {
        set_error("ds_integrate_velocity_forward: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_integrate(c, CF3{{velocity[0], velocity[1], velocity[2]}}, CF3{{force[0], force[1], force[2]}},
                         F3{{out[0], out[1], out[2]}});
    });
}

ds_status ds_advect_velocity_forward(ds_context* c, float* const velocity[3], float* out[3]) {
    if (!c || !velocity || !out) {
        set_error("ds_advect_velocity_forward: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_advect_velocity(c, CF3{{velocity[0], velocity[1], velocity[2]}}, F3{{out[0], out[1], out[2]}}, false);
    });
}

ds_status ds_constrain_velocity_forward(ds_context* c, float* const velocity[3], float* out[3]) {
    if (!c || !velocity || !out) {
        set_error("ds_constrain_velocity_forward: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_constrain(c, c->d_vb, collider_of(c), CF3{{velocity[0], velocity[1], velocity[2]}}, F3{{out[0], out[1], out[2]}});
    });
}

ds_status ds_advect_scalar_forward(ds_context* c, uint32_t field, const float* source, float* const velocity[3], float* out) {
    if (!c || !source || !velocity || !out || field > 1u) {
        set_error("ds_advect_scalar_forward: invalid argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_advect_scalars(c, 1, (int)field, source, source, CF3{{velocity[0], velocity[1], velocity[2]}}, out, out);
    });
}

ds_status ds_projection_forward(ds_context* c, float* const velocity[3], float* pressure, float* rhs, const float* guess,
                                float* out[3], ds_solve_info* info) {
    if (!c || !velocity || !out) {
        set_error("ds_projection_forward: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ds_solve_info si{};
    const ds_status st = guard([&] {
        const float* u[3]{velocity[0], velocity[1], velocity[2]};
        si = pcg_solve_kind(c, 1, u, nullptr, c->pressure_hi, c->pressure_lo, guess, nullptr, c->tolerance, c->max_iterations);
        c->pressure_valid = true;
        launch_project(c, CF3{{velocity[0], velocity[1], velocity[2]}}, c->pressure_hi, c->pressure_lo, F3{{out[0], out[1], out[2]}});
        if (pressure) combine_pair(c, c->pressure_hi, c->pressure_lo, pressure);
        if (rhs) DS_HIP(hipMemcpyAsync(rhs, c->pcg.b, (size_t)c->geo.n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    });
    if (info) *info = si;
    if (st != DS_OK) return st;
    if (!si.converged) {
        set_error("ds_projection_forward: pressure solve did not converge");
        return std::isfinite(si.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    return DS_OK;
}

}  // extern "C"
