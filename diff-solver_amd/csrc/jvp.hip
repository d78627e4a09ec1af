// jvp.hip — forward mode of the smoke step on gfx950: Model::jvp_step (solvers/smoke/model.cpp:351-359)
// and its tangent kernels (operators.cu:315-324, 390-413, 509-535; the forwards reused on tangents,
// :575-638, 1015-1018; vorticity, :731-779, 981-988).
//
// Tangents are FP32. One per-cell pass forms the source, buoyancy and confinement-force tangents;
// the integration tangent is the forward face kernel on tangents; the velocity-advection tangent
// has the tangent constraint fused in (constrained faces get 0, a periodic duplicate face takes
// face 0's tangent); the projection tangent is a PCG solve of the tangent's right-hand side (no
// fixed-face term) followed by the forward projection; the density and temperature tangents share
// one RK2 trace per cell. Tangent samples read boundaries whose fixed values are zeroed.
#include <hip/hip_runtime.h>

#include <cmath>

#include "context.hpp"
#include "kernels.hpp"
#include "vorticity.hpp"

namespace ds {

namespace {

struct VortTangent {      // inputs of the confinement-force tangent
    VortFields v;         // w and |w| of the step's input velocity
    const float* wt[3];   // vorticity tangent
    const float* mt;      // magnitude tangent (0 in solids)
};

// normal_tangent_kernel + vorticity_force_tangent_kernel (operators.cu:737-779) at fluid cell i:
// f_t += h (conf_t (N x w) + conf (N_t x w + N x w_t)), N_t = (g_t - N (N . g_t)) / len, g_t the
// clamped central difference of the magnitude tangent.
__device__ __forceinline__ void add_confinement_force_tangent(const Geo& g, const VortTangent& vt, float conf, float conf_t, uint32_t i,
                                                              int x, int y, int z, float f[3]) {
    float N[3];
    const float len = normal_at(g, vt.v.mag, i, x, y, z, N);
    const float s = 0.5f * g.inv_h;
    const int c[3]{x, y, z};
    float gt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const Clamp1 k = clamp_neighbours(i, c[d], dim_size(g, d), cell_stride(g, d));
        gt[d] = (vt.mt[k.hi] - vt.mt[k.lo]) * s;
    }
    const float proj = N[0] * gt[0] + N[1] * gt[1] + N[2] * gt[2];
    const float Nt[3]{(gt[0] - N[0] * proj) / len, (gt[1] - N[1] * proj) / len, (gt[2] - N[2] * proj) / len};
    const float w[3]{vt.v.w[0][i], vt.v.w[1][i], vt.v.w[2][i]};
    const float wt[3]{vt.wt[0][i], vt.wt[1][i], vt.wt[2][i]};
    const float cross[3]{N[1] * w[2] - N[2] * w[1], N[2] * w[0] - N[0] * w[2], N[0] * w[1] - N[1] * w[0]};
    const float cross_t[3]{Nt[1] * w[2] + N[1] * wt[2] - Nt[2] * w[1] - N[2] * wt[1],
                           Nt[2] * w[0] + N[2] * wt[0] - Nt[0] * w[2] - N[0] * wt[2],
                           Nt[0] * w[1] + N[0] * wt[1] - Nt[1] * w[0] - N[1] * wt[0]};
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] += g.h * (conf_t * cross[a] + conf * cross_t[a]);
}

// The tangent curl and magnitude (operators.cu:981-984 with magnitude_tangent_kernel, :731-735):
// w_t = curl of the centred tangent velocity, |w|_t = (w . w_t) / |w| on fluid cells, 0 in solids.
__global__ __launch_bounds__(256) void vorticity_curl_tangent_kernel(Geo g, const uint8_t* __restrict__ mask, CF3 ut, VortFields v, F3 wt,
                                                                     float* __restrict__ mt) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (mask[i] != 0) {
        wt.p[0][i] = wt.p[1][i] = wt.p[2][i] = 0.0f;
        mt[i] = 0.0f;
        return;
    }
    int x, y, z;
    decode_cell(g, i, x, y, z);
    float w_t[3];
    curl_at(g, mask, ut, x, y, z, w_t);
    wt.p[0][i] = w_t[0];
    wt.p[1][i] = w_t[1];
    wt.p[2][i] = w_t[2];
    mt[i] = (v.w[0][i] * w_t[0] + v.w[1][i] * w_t[1] + v.w[2][i] * w_t[2]) / v.mag[i];
}

// source_jvp (the source forward on tangents, operators.cu:575-577), buoyancy_jvp_kernel
// (operators.cu:315-324) and, with Vort, the confinement-force tangent, fused per cell.
template <bool Vort>
__global__ __launch_bounds__(256) void source_buoyancy_jvp_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ rho_t,
                                                                  const float* __restrict__ T_t, const float* __restrict__ s_rho_t,
                                                                  const float* __restrict__ s_T_t, CF3 e_t, const float* __restrict__ rho_s,
                                                                  const float* __restrict__ T_s, Params prm, Params prm_t, VortTangent vt,
                                                                  float* __restrict__ rho_s_t, float* __restrict__ T_s_t, F3 f_t) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (mask[i] != 0) {
        rho_s_t[i] = T_s_t[i] = 0.0f;
        f_t.p[0][i] = f_t.p[1][i] = f_t.p[2][i] = 0.0f;
        return;
    }
    const float rst = rho_t[i] + g.dt * s_rho_t[i];
    const float tst = T_t[i] + g.dt * s_T_t[i];
    rho_s_t[i] = rst;
    T_s_t[i] = tst;
    const float amb = *prm.ambient, ar = *prm.alpha_rho, at = *prm.alpha_t;
    const float amb_t = *prm_t.ambient, ar_t = *prm_t.alpha_rho, at_t = *prm_t.alpha_t;
    float f[3]{e_t.p[0][i], e_t.p[1][i] + ar * rst + ar_t * rho_s[i] + at * (tst - amb_t) + at_t * (T_s[i] - amb), e_t.p[2][i]};
    if constexpr (Vort) {
        int x, y, z;
        decode_cell(g, i, x, y, z);
        add_confinement_force_tangent(g, vt, *prm.confinement, *prm_t.confinement, i, x, y, z, f);
    }
    f_t.p[0][i] = f[0];
    f_t.p[1][i] = f[1];
    f_t.p[2][i] = f[2];
}

// buoyancy_jvp_kernel (operators.cu:315-324) alone, for the per-operator entry point.
__global__ __launch_bounds__(256) void buoyancy_jvp_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ rho,
                                                           const float* __restrict__ T, const float* __restrict__ rho_t,
                                                           const float* __restrict__ T_t, CF3 e_t, Params prm, Params prm_t, F3 f_t) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (mask[i] != 0) {
        f_t.p[0][i] = f_t.p[1][i] = f_t.p[2][i] = 0.0f;
        return;
    }
    const float amb = *prm.ambient, ar = *prm.alpha_rho, at = *prm.alpha_t;
    const float amb_t = *prm_t.ambient, ar_t = *prm_t.alpha_rho, at_t = *prm_t.alpha_t;
    f_t.p[0][i] = e_t.p[0][i];
    f_t.p[1][i] = e_t.p[1][i] + ar * rho_t[i] + ar_t * rho[i] + at * (T_t[i] - amb_t) + at_t * (T[i] - amb);
    f_t.p[2][i] = e_t.p[2][i];
}

// The shared part of advect_velocity_jvp_kernel and advect_scalar_jvp_kernel (operators.cu:390-412,
// 509-533): the trace from s, and the tangent of its end point.
template <bool HasSolids>
__device__ __forceinline__ TraceOut trace_tangent(float sx, float sy, float sz, const Vel3& u, const Vel3& ut,
                                                  const uint8_t* __restrict__ mask, const Geo& g, const VBound& b, const VBound& bz,
                                                  float pos_t[3]) {
    Sample mid[3];
    const TraceOut t = trace_rk2<HasSolids, true>(sx, sy, sz, u, mask, g, b, mid);
    const float hdt = -0.5f * g.dt;
    const float mtx = hdt * sample_face<0, false>(ut.u[0], sx, sy, sz, g, bz).value;
    const float mty = hdt * sample_face<1, false>(ut.u[1], sx, sy, sz, g, bz).value;
    const float mtz = hdt * sample_face<2, false>(ut.u[2], sx, sy, sz, g, bz).value;
    const Sample& s0 = mid[0];
    const Sample& s1 = mid[1];
    const Sample& s2 = mid[2];
    const float v1x = sample_face<0, false>(ut.u[0], t.mx, t.my, t.mz, g, bz).value + s0.gx * mtx + s0.gy * mty + s0.gz * mtz;
    const float v1y = sample_face<1, false>(ut.u[1], t.mx, t.my, t.mz, g, bz).value + s1.gx * mtx + s1.gy * mty + s1.gz * mtz;
    const float v1z = sample_face<2, false>(ut.u[2], t.mx, t.my, t.mz, g, bz).value + s2.gx * mtx + s2.gy * mty + s2.gz * mtz;
    pos_t[0] = -t.dx * g.dt * v1x;
    pos_t[1] = -t.dy * g.dt * v1y;
    pos_t[2] = -t.dz * g.dt * v1z;
    return t;
}

// advect_velocity_jvp_kernel (operators.cu:390-413); Constrain fuses constrain_velocity_jvp
// (operators.cu:633-638): constrained faces get 0, a periodic duplicate face traces face 0's position.
template <bool Constrain, bool HasSolids>
__global__ __launch_bounds__(256) void advect_velocity_jvp_kernel(Geo g, const VBound* __restrict__ bp, const VBound* __restrict__ bzp,
                                                                  const uint8_t* __restrict__ mask, Vel3 u, Vel3 ut, F3 out) {
    const VBound& b = *bp;
    const VBound& bz = *bzp;
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        if (Constrain) {
            float value;
            const int kind = constraint<A>(g, bz, mask, CF3{{nullptr, nullptr, nullptr}}, HasSolids, x, y, z, local, value);
            if (kind == 1) {
                out.p[A][local] = 0.0f;
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
        float pt[3];
        const TraceOut t = trace_tangent<HasSolids>(sx, sy, sz, u, ut, mask, g, b, bz, pt);
        const Sample src = sample_face<A, true>(u.u[A], t.px, t.py, t.pz, g, b);
        out.p[A][local] = sample_face<A, false>(ut.u[A], t.px, t.py, t.pz, g, bz).value + src.gx * pt[0] + src.gy * pt[1] + src.gz * pt[2];
    });
}

// advect_scalar_jvp_kernel (operators.cu:509-535) for NF fields sharing one trace; solid cells get 0.
template <int NF, bool HasSolids>
__global__ __launch_bounds__(256) void advect_scalar_jvp_kernel(Geo g, const VBound* __restrict__ vbp, const VBound* __restrict__ vbzp,
                                                                const SBound* __restrict__ sb0p, const SBound* __restrict__ sb1p,
                                                                const SBound* __restrict__ sbz0p, const SBound* __restrict__ sbz1p,
                                                                const uint8_t* __restrict__ mask, const float* __restrict__ src0,
                                                                const float* __restrict__ src1, const float* __restrict__ src0_t,
                                                                const float* __restrict__ src1_t, Vel3 u, Vel3 ut, float* __restrict__ out0,
                                                                float* __restrict__ out1) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (HasSolids && mask[i] != 0) {
        out0[i] = 0.0f;
        if (NF == 2) out1[i] = 0.0f;
        return;
    }
    int x, y, z;
    decode_cell(g, i, x, y, z);
    float pt[3];
    const TraceOut t = trace_tangent<HasSolids>((x + 0.5f) * g.h, (y + 0.5f) * g.h, (z + 0.5f) * g.h, u, ut, mask, g, *vbp, *vbzp, pt);
    const Sample s0 = sample_scalar<true>(src0, t.px, t.py, t.pz, g, *sb0p);
    out0[i] = sample_scalar<false>(src0_t, t.px, t.py, t.pz, g, *sbz0p).value + s0.gx * pt[0] + s0.gy * pt[1] + s0.gz * pt[2];
    if (NF == 2) {
        const Sample s1 = sample_scalar<true>(src1, t.px, t.py, t.pz, g, *sb1p);
        out1[i] = sample_scalar<false>(src1_t, t.px, t.py, t.pz, g, *sbz1p).value + s1.gx * pt[0] + s1.gy * pt[1] + s1.gz * pt[2];
    }
}

// The confinement-force tangent alone (vorticity_force_tangent_kernel, operators.cu:756-779): f_t +=.
__global__ __launch_bounds__(256) void vorticity_force_tangent_kernel(Geo g, const uint8_t* __restrict__ mask, VortTangent vt,
                                                                      const float* __restrict__ conf, const float* __restrict__ conf_t,
                                                                      F3 f_t) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n || mask[i] != 0) return;
    int x, y, z;
    decode_cell(g, i, x, y, z);
    float f[3]{0.0f, 0.0f, 0.0f};
    add_confinement_force_tangent(g, vt, *conf, *conf_t, i, x, y, z, f);
    f_t.p[0][i] += f[0];
    f_t.p[1][i] += f[1];
    f_t.p[2][i] += f[2];
}

VortTangent vort_tangent_of(const ds_context* c) {
    return VortTangent{VortFields{{c->vort_w[0], c->vort_w[1], c->vort_w[2]}, c->vort_mag}, {c->vort_wt[0], c->vort_wt[1], c->vort_wt[2]},
                       c->vort_mt};
}

void launch_vorticity_curl_tangent(ds_context* c, CF3 ut) {
    const VortTangent vt = vort_tangent_of(c);
    vorticity_curl_tangent_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, ut, vt.v,
                                                                                      F3{{c->vort_wt[0], c->vort_wt[1], c->vort_wt[2]}},
                                                                                      c->vort_mt);
    check_launch("vorticity_curl_tangent_kernel");
}

void vorticity_force_tangent(ds_context* c, const float* conf, const float* conf_t, F3 f_t) {
    vorticity_force_tangent_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, vort_tangent_of(c), conf, conf_t, f_t);
    check_launch("vorticity_force_tangent_kernel");
}

void launch_advect_velocity_jvp(ds_context* c, CF3 u, CF3 ut, F3 out, bool constrain) {
    const Vel3 v{{u.p[0], u.p[1], u.p[2]}}, vt{{ut.p[0], ut.p[1], ut.p[2]}};
    const uint32_t G = xcd_blocks_for(c->geo.foff[3]);
    if (constrain) {
        if (c->has_solids) advect_velocity_jvp_kernel<true, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, c->mask, v, vt, out);
        else advect_velocity_jvp_kernel<true, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, c->mask, v, vt, out);
    } else {
        if (c->has_solids) advect_velocity_jvp_kernel<false, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, c->mask, v, vt, out);
        else advect_velocity_jvp_kernel<false, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, c->mask, v, vt, out);
    }
    check_launch("advect_velocity_jvp_kernel");
}

void launch_advect_scalar_jvp(ds_context* c, int nf, int first_field, const float* src0, const float* src1, const float* src0_t,
                              const float* src1_t, CF3 u, CF3 ut, float* out0, float* out1) {
    const Vel3 v{{u.p[0], u.p[1], u.p[2]}}, vt{{ut.p[0], ut.p[1], ut.p[2]}};
    const uint32_t G = xcd_blocks_for(c->geo.n);
    if (nf == 2) {
        if (c->has_solids)
            advect_scalar_jvp_kernel<2, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, c->d_db, c->d_tb, c->d_db_zero,
                                                                          c->d_tb_zero, c->mask, src0, src1, src0_t, src1_t, v, vt, out0, out1);
        else
            advect_scalar_jvp_kernel<2, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, c->d_db, c->d_tb, c->d_db_zero,
                                                                           c->d_tb_zero, c->mask, src0, src1, src0_t, src1_t, v, vt, out0, out1);
    } else {
        const SBound* b0 = first_field == 0 ? c->d_db : c->d_tb;
        const SBound* bz0 = first_field == 0 ? c->d_db_zero : c->d_tb_zero;
        if (c->has_solids)
            advect_scalar_jvp_kernel<1, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, b0, b0, bz0, bz0, c->mask, src0, src0,
                                                                          src0_t, src0_t, v, vt, out0, out0);
        else
            advect_scalar_jvp_kernel<1, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_vb_zero, b0, b0, bz0, bz0, c->mask, src0,
                                                                           src0, src0_t, src0_t, v, vt, out0, out0);
    }
    check_launch("advect_scalar_jvp_kernel");
}

// projection_jvp (operators.cu:1015-1018): the tangent pressure solve (warm-started from the
// previous tangent solve under every warm start other than DS_WARM_START_NONE: the tangent
// sequence has no extrapolation, window or network of its own) and the forward projection.
ds_solve_info project_tangent(ds_context* c, CF3 ut, F3 out) {
    const bool warm = c->config.warm_start != DS_WARM_START_NONE && c->tangent_valid;
    const float* u[3]{ut.p[0], ut.p[1], ut.p[2]};
    const ds_solve_info si = pcg_solve_kind(c, 3, u, nullptr, c->tangent_hi, c->tangent_lo, warm ? c->tangent_hi : nullptr,
                                            warm ? c->tangent_lo : nullptr, c->tolerance, c->max_iterations);
    c->tangent_valid = true;
    launch_project(c, ut, c->tangent_hi, c->tangent_lo, out);
    return si;
}

void check_tangent_state(const ds_state* s, const char* what) {
    require(s && s->density && s->temperature && s->velocity[0] && s->velocity[1] && s->velocity[2], what);
}

}  // namespace

// Model::jvp_step (model.cpp:351-359): source and force tangents (with the confinement tangent),
// integrate -> advect -> constrain tangents, projection tangent, scalar advection tangents.
ds_solve_info jvp_step(ds_context* c, const ds_parameters* p, const ds_parameters* pt, const ds_state* state, const ds_state* next,
                       const ds_step_cache* cache, const ds_state* state_t, const ds_control* control_t, ds_state* next_t) {
    require(c->mode == DS_MODE_DIFFERENTIABLE, "jvp_step: the context was not created in differentiable mode");
    check_tangent_state(next, "jvp_step: next_state has a null field");
    check_tangent_state(state_t, "jvp_step: state_tangent has a null field");
    check_tangent_state(next_t, "jvp_step: next_state_tangent has a null field");
    require(control_t && control_t->density_source && control_t->temperature_source && control_t->external_acceleration[0] &&
                control_t->external_acceleration[1] && control_t->external_acceleration[2],
            "jvp_step: control_tangent has a null field");
    require(cache && cache->sourced_density && cache->sourced_temperature && cache->forced_velocity[0] && cache->forced_velocity[1] &&
                cache->forced_velocity[2],
            "jvp_step: the step cache needs sourced fields and the forced velocity");
    const Params prm = params_of(p);
    const Params prm_t = params_of(pt);
    const Geo& g = c->geo;
    const uint32_t G = xcd_blocks_for(g.n);
    const CF3 e_t{{control_t->external_acceleration[0], control_t->external_acceleration[1], control_t->external_acceleration[2]}};
    const F3 f_t{{c->force_t[0], c->force_t[1], c->force_t[2]}};
    if (c->vorticity) {
        require(state && state->velocity[0] && state->velocity[1] && state->velocity[2],
                "jvp_step: vorticity confinement needs the step's input velocity (state)");
        require(prm.confinement && prm_t.confinement, "jvp_step: vorticity_confinement must be a device pointer in both parameter sets");
        launch_vorticity_curl(c, CF3{{state->velocity[0], state->velocity[1], state->velocity[2]}});
        launch_vorticity_curl_tangent(c, CF3{{state_t->velocity[0], state_t->velocity[1], state_t->velocity[2]}});
        source_buoyancy_jvp_kernel<true><<<G, kBlock, 0, c->stream>>>(g, c->mask, state_t->density, state_t->temperature,
                                                                     control_t->density_source, control_t->temperature_source, e_t,
                                                                     cache->sourced_density, cache->sourced_temperature, prm, prm_t,
                                                                     vort_tangent_of(c), c->sourced_t[0], c->sourced_t[1], f_t);
    } else {
        source_buoyancy_jvp_kernel<false><<<G, kBlock, 0, c->stream>>>(g, c->mask, state_t->density, state_t->temperature,
                                                                      control_t->density_source, control_t->temperature_source, e_t,
                                                                      cache->sourced_density, cache->sourced_temperature, prm, prm_t,
                                                                      VortTangent{}, c->sourced_t[0], c->sourced_t[1], f_t);
    }
    check_launch("source_buoyancy_jvp_kernel");
    const F3 us_t{{c->forced_t[0], c->forced_t[1], c->forced_t[2]}};
    launch_integrate(c, CF3{{state_t->velocity[0], state_t->velocity[1], state_t->velocity[2]}}, CF3{{f_t.p[0], f_t.p[1], f_t.p[2]}}, us_t);
    const F3 ua_t{{c->advected_t[0], c->advected_t[1], c->advected_t[2]}};
    launch_advect_velocity_jvp(c, CF3{{cache->forced_velocity[0], cache->forced_velocity[1], cache->forced_velocity[2]}},
                               CF3{{us_t.p[0], us_t.p[1], us_t.p[2]}}, ua_t, true);
    const ds_solve_info si = project_tangent(c, CF3{{ua_t.p[0], ua_t.p[1], ua_t.p[2]}}, F3{{next_t->velocity[0], next_t->velocity[1], next_t->velocity[2]}});
    launch_advect_scalar_jvp(c, 2, 0, cache->sourced_density, cache->sourced_temperature, c->sourced_t[0], c->sourced_t[1],
                             CF3{{next->velocity[0], next->velocity[1], next->velocity[2]}},
                             CF3{{next_t->velocity[0], next_t->velocity[1], next_t->velocity[2]}}, next_t->density, next_t->temperature);
    return si;
}

}  // namespace ds

using namespace ds;

namespace {
bool need_tangent_context(ds_context* c, const char* name) {
    if (c->mode != DS_MODE_DIFFERENTIABLE) {
        set_error(std::string(name) + ": the context was not created in differentiable mode");
        return false;
    }
    return true;
}
}  // namespace

extern "C" {

ds_status ds_jvp_step(ds_context* c, const ds_state* state, const ds_control* control, const ds_parameters* params,
                      const ds_state* next_state, const ds_step_cache* cache, const ds_state_tangent* state_tangent,
                      const ds_control_tangent* control_tangent, const ds_parameter_tangent* param_tangent,
                      ds_state_tangent* next_state_tangent, ds_step_info* info) {
    (void)control;
    if (!c) {
        set_error("ds_jvp_step: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ds_solve_info si{};
    const ds_status st = guard([&] {
        si = jvp_step(c, params, param_tangent, state, next_state, cache, state_tangent, control_tangent, next_state_tangent);
    });
    if (info) info->projection = si;
    if (st != DS_OK) return st;
    if (!si.converged) {
        set_error("ds_jvp_step: tangent pressure solve did not converge");
        return std::isfinite(si.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    return DS_OK;
}

ds_status ds_buoyancy_jvp(ds_context* c, const float* density, const float* temperature, const float* density_tangent,
                          const float* temperature_tangent, float* const external_acceleration_tangent[3], const ds_parameters* params,
                          const ds_parameter_tangent* param_tangent, float* force_tangent[3]) {
    if (!c || !density || !temperature || !density_tangent || !temperature_tangent || !external_acceleration_tangent || !force_tangent) {
        set_error("ds_buoyancy_jvp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        buoyancy_jvp_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(
            c->geo, c->mask, density, temperature, density_tangent, temperature_tangent,
            CF3{{external_acceleration_tangent[0], external_acceleration_tangent[1], external_acceleration_tangent[2]}}, params_of(params),
            params_of(param_tangent), F3{{force_tangent[0], force_tangent[1], force_tangent[2]}});
        check_launch("buoyancy_jvp_kernel");
    });
}

ds_status ds_vorticity_jvp(ds_context* c, float* const velocity[3], float* const velocity_tangent[3], const ds_parameters* params,
                           const ds_parameter_tangent* param_tangent, float* force_tangent[3]) {
    if (!c || !velocity || !velocity_tangent || !params || !param_tangent || !force_tangent) {
        set_error("ds_vorticity_jvp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!c->vorticity) {
        set_error("ds_vorticity_jvp: the context was created with vorticity_confinement_enabled = 0");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!need_tangent_context(c, "ds_vorticity_jvp")) return DS_ERROR_INVALID_ARGUMENT;
    return guard([&] {
        require(params->vorticity_confinement && param_tangent->vorticity_confinement,
                "ds_vorticity_jvp: vorticity_confinement must be a device pointer in both parameter sets");
        launch_vorticity_curl(c, CF3{{velocity[0], velocity[1], velocity[2]}});
        launch_vorticity_curl_tangent(c, CF3{{velocity_tangent[0], velocity_tangent[1], velocity_tangent[2]}});
        vorticity_force_tangent(c, params->vorticity_confinement, param_tangent->vorticity_confinement,
                                F3{{force_tangent[0], force_tangent[1], force_tangent[2]}});
    });
}

ds_status ds_advect_velocity_jvp(ds_context* c, float* const velocity[3], float* const velocity_tangent[3], float* out_tangent[3]) {
    if (!c || !velocity || !velocity_tangent || !out_tangent) {
        set_error("ds_advect_velocity_jvp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_advect_velocity_jvp(c, CF3{{velocity[0], velocity[1], velocity[2]}},
                                   CF3{{velocity_tangent[0], velocity_tangent[1], velocity_tangent[2]}},
                                   F3{{out_tangent[0], out_tangent[1], out_tangent[2]}}, false);
    });
}

ds_status ds_constrain_velocity_jvp(ds_context* c, float* const velocity_tangent[3], float* out_tangent[3]) {
    if (!c || !velocity_tangent || !out_tangent) {
        set_error("ds_constrain_velocity_jvp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_constrain(c, c->d_vb_zero, CF3{{nullptr, nullptr, nullptr}}, CF3{{velocity_tangent[0], velocity_tangent[1], velocity_tangent[2]}},
                         F3{{out_tangent[0], out_tangent[1], out_tangent[2]}});
    });
}

ds_status ds_advect_scalar_jvp(ds_context* c, uint32_t field, const float* source, const float* source_tangent, float* const velocity[3],
                               float* const velocity_tangent[3], float* out_tangent) {
    if (!c || !source || !source_tangent || !velocity || !velocity_tangent || !out_tangent || field > 1u) {
        set_error("ds_advect_scalar_jvp: invalid argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_advect_scalar_jvp(c, 1, (int)field, source, source, source_tangent, source_tangent, CF3{{velocity[0], velocity[1], velocity[2]}},
                                 CF3{{velocity_tangent[0], velocity_tangent[1], velocity_tangent[2]}}, out_tangent, out_tangent);
    });
}

ds_status ds_projection_jvp(ds_context* c, float* const velocity_tangent[3], float* out_tangent[3], ds_solve_info* info) {
    if (!c || !velocity_tangent || !out_tangent) {
        set_error("ds_projection_jvp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!need_tangent_context(c, "ds_projection_jvp")) return DS_ERROR_INVALID_ARGUMENT;
    ds_solve_info si{};
    const ds_status st = guard([&] {
        si = project_tangent(c, CF3{{velocity_tangent[0], velocity_tangent[1], velocity_tangent[2]}},
                             F3{{out_tangent[0], out_tangent[1], out_tangent[2]}});
    });
    if (info) *info = si;
    if (st != DS_OK) return st;
    if (!si.converged) {
        set_error("ds_projection_jvp: tangent pressure solve did not converge");
        return std::isfinite(si.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    return DS_OK;
}

}  // extern "C"
