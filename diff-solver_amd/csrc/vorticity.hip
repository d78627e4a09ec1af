// vorticity.hip — vorticity confinement on gfx950 (solvers/smoke/operators.cu:662-856, launchers
// :973-1002; wired in ForceAssemblyOperator, operators.cpp:125-138).
//
// Forward: one curl pass (centred velocity, curl and smoothed magnitude fused; the centred
// velocity is never stored) writes w and |w|; the force itself is added inside the fused
// source/buoyancy kernel of the step (forward.hip), or by vorticity_force_kernel for the
// per-operator entry point. Reverse: the force and normal reverse run per cell on the force
// adjoint (fused into the step's force/source reverse, adjoint.hip, or vorticity_force_vjp_kernel
// here), then two gathers: the magnitude reverse per cell and the curl + centring reverse per face.
#include <hip/hip_runtime.h>

#include "context.hpp"
#include "kernels.hpp"
#include "vorticity.hpp"

namespace ds {

namespace {

// centered_velocity_kernel + curl_magnitude_kernel (operators.cu:678-701).
__global__ __launch_bounds__(256) void vorticity_curl_kernel(Geo g, const uint8_t* __restrict__ mask, CF3 u, F3 w, float* __restrict__ mag) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (mask[i] != 0) {
        w.p[0][i] = w.p[1][i] = w.p[2][i] = 0.0f;
        mag[i] = kSmoothEpsilon;
        return;
    }
    int x, y, z;
    decode_cell(g, i, x, y, z);
    float v[3];
    curl_at(g, mask, u, x, y, z, v);
    w.p[0][i] = v[0];
    w.p[1][i] = v[1];
    w.p[2][i] = v[2];
    mag[i] = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + kSmoothEpsilon * kSmoothEpsilon);
}

// vorticity_normal_kernel + add_vorticity_force_kernel (operators.cu:703-729): out += the force.
__global__ __launch_bounds__(256) void vorticity_force_kernel(Geo g, const uint8_t* __restrict__ mask, VortFields v,
                                                              const float* __restrict__ confinement, F3 out) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n || mask[i] != 0) return;
    int x, y, z;
    decode_cell(g, i, x, y, z);
    float f[3]{0.0f, 0.0f, 0.0f};
    add_confinement_force(v, v.mag, g, *confinement * g.h, i, x, y, z, f);
    out.p[0][i] += f[0];
    out.p[1][i] += f[1];
    out.p[2][i] += f[2];
}

// vorticity_force_reverse_kernel + the per-cell part of normal_reverse_kernel (operators.cu:781-812)
// from a cell force adjoint; the confinement adjoint as FP64 block partials.
__global__ __launch_bounds__(256) void vorticity_force_vjp_kernel(Geo g, const uint8_t* __restrict__ mask, VortFields v, CF3 fa,
                                                                  const float* __restrict__ confinement, F3 ga, F3 t,
                                                                  double* __restrict__ partials) {
    __shared__ double lds[kMaxWaves];
    const uint32_t lb = xcd_block();
    uint32_t chunk = (g.n + gridDim.x - 1) / gridDim.x;
    chunk = (chunk + 255u) & ~255u;
    const uint32_t start = min(g.n, lb * chunk), end = min(g.n, start + chunk);
    const float scale = *confinement * g.h;
    double part = 0.0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        if (mask[i] != 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) ga.p[a][i] = t.p[a][i] = 0.0f;
            continue;
        }
        int x, y, z;
        decode_cell(g, i, x, y, z);
        const float f[3]{fa.p[0][i], fa.p[1][i], fa.p[2][i]};
        float gv[3], tv[3];
        part += confinement_reverse_at(g, v, i, x, y, z, scale, f, gv, tv);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ga.p[a][i] = gv[a];
            t.p[a][i] = tv[a];
        }
    }
    double r[1]{part};
    block_sum<1>(r, lds);
    if (threadIdx.x == 0) partials[lb] = r[0];
}

// Gather form of the rest of normal_reverse_kernel and of magnitude_reverse_kernel
// (operators.cu:812-828): |w|_bar = 0.5/h sum_d D_d^T ga_d, then t += |w|_bar / |w| * w in place
// (t becomes the full vorticity adjoint). Solid cells keep t = 0.
__global__ __launch_bounds__(256) void vorticity_magnitude_vjp_kernel(Geo g, const uint8_t* __restrict__ mask, VortFields v, CF3 ga,
                                                                      F3 t) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n || mask[i] != 0) return;
    int x, y, z;
    decode_cell(g, i, x, y, z);
    const float mbar = 0.5f * g.inv_h *
                       (transposed_difference(g, ga.p[0], i, x, 0) + transposed_difference(g, ga.p[1], i, y, 1) +
                        transposed_difference(g, ga.p[2], i, z, 2));
    const float k = mbar / v.mag[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) t.p[a][i] += k * v.w[a][i];
}

// Centred-velocity adjoint of component A at fluid cell i (curl_reverse_kernel, operators.cu:830-844,
// as a gather): with B = A+1, C = A+2 (mod 3), c_bar_A = 0.5/h (D_C^T wa_B - D_B^T wa_C).
template <int A>
__device__ __forceinline__ float centred_adjoint(const Geo& g, CF3 wa, uint32_t i, int x, int y, int z) {
    constexpr int B = (A + 1) % 3, Cc = (A + 2) % 3;
    const int c[3]{x, y, z};
    return 0.5f * g.inv_h * (transposed_difference(g, wa.p[B], i, c[Cc], Cc) - transposed_difference(g, wa.p[Cc], i, c[B], B));
}

// centered_velocity_reverse_kernel (operators.cu:846-856) as a gather per face: u_bar += 0.5 *
// (c_bar(lo cell) + c_bar(hi cell)) over the in-domain fluid cells beside the face.
__global__ __launch_bounds__(256) void vorticity_velocity_vjp_kernel(Geo g, const uint8_t* __restrict__ mask, CF3 wa, F3 ua) {
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        const int coord = A == 0 ? x : A == 1 ? y : z;
        float sum = 0.0f;
        if (coord > 0) {
            const int xl = x - (A == 0), yl = y - (A == 1), zl = z - (A == 2);
            const uint32_t q = cidx(g, xl, yl, zl);
            if (mask[q] == 0) sum += centred_adjoint<A>(g, wa, q, xl, yl, zl);
        }
        if (coord < dim_size(g, A)) {
            const uint32_t q = cidx(g, x, y, z);
            if (mask[q] == 0) sum += centred_adjoint<A>(g, wa, q, x, y, z);
        }
        ua.p[A][local] += 0.5f * sum;
    });
}

// One block: *out += sum of the partials (fixed order).
__global__ __launch_bounds__(256) void partial_accumulate_kernel(const double* __restrict__ partials, uint32_t G, double* out) {
    __shared__ double lds[kMaxWaves];
    const double* parts[1]{partials};
    double red[1];
    reduce_partials<1>(parts, G, red, lds);
    if (threadIdx.x == 0) *out += red[0];
}

VortFields fields_of(const ds_context* c) { return VortFields{{c->vort_w[0], c->vort_w[1], c->vort_w[2]}, c->vort_mag}; }

}  // namespace

void launch_vorticity_curl(ds_context* c, CF3 u) {
    vorticity_curl_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, u,
                                                                              F3{{c->vort_w[0], c->vort_w[1], c->vort_w[2]}}, c->vort_mag);
    check_launch("vorticity_curl_kernel");
}

void launch_vorticity_force(ds_context* c, const Params& prm, F3 out) {
    vorticity_force_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, fields_of(c), prm.confinement, out);
    check_launch("vorticity_force_kernel");
}

void launch_vorticity_vjp_tail(ds_context* c, F3 u_adj) {
    const F3 t{{c->vort_t[0], c->vort_t[1], c->vort_t[2]}};
    vorticity_magnitude_vjp_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(
        c->geo, c->mask, fields_of(c), CF3{{c->vort_ga[0], c->vort_ga[1], c->vort_ga[2]}}, t);
    check_launch("vorticity_magnitude_vjp_kernel");
    vorticity_velocity_vjp_kernel<<<xcd_blocks_for(c->geo.foff[3]), kBlock, 0, c->stream>>>(
        c->geo, c->mask, CF3{{t.p[0], t.p[1], t.p[2]}}, u_adj);
    check_launch("vorticity_velocity_vjp_kernel");
}

void launch_vorticity_vjp(ds_context* c, const Params& prm, CF3 fa, F3 u_adj, double* conf_adj) {
    double* parts = c->param_partials + 3 * (size_t)kPcgBlocks;
    vorticity_force_vjp_kernel<<<kPcgBlocks, kBlock, 0, c->stream>>>(c->geo, c->mask, fields_of(c), fa, prm.confinement,
                                                                     F3{{c->vort_ga[0], c->vort_ga[1], c->vort_ga[2]}},
                                                                     F3{{c->vort_t[0], c->vort_t[1], c->vort_t[2]}}, parts);
    check_launch("vorticity_force_vjp_kernel");
    if (conf_adj) {
        partial_accumulate_kernel<<<1, kBlock, 0, c->stream>>>(parts, kPcgBlocks, conf_adj);
        check_launch("partial_accumulate_kernel");
    }
    launch_vorticity_vjp_tail(c, u_adj);
}

}  // namespace ds

using namespace ds;

namespace {
const float* confinement_of(const ds_parameters* p) {
    require(p && p->vorticity_confinement, "parameters: vorticity_confinement must be a device pointer");
    return p->vorticity_confinement;
}
}  // namespace

extern "C" {

ds_status ds_vorticity_forward(ds_context* c, float* const velocity[3], const ds_parameters* params, float* force[3]) {
    if (!c || !velocity || !force || !velocity[0] || !velocity[1] || !velocity[2] || !force[0] || !force[1] || !force[2]) {
        set_error("ds_vorticity_forward: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!c->vorticity) {
        set_error("ds_vorticity_forward: the context was created with vorticity_confinement_enabled = 0");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        Params prm{};
        prm.confinement = confinement_of(params);
        launch_vorticity_curl(c, CF3{{velocity[0], velocity[1], velocity[2]}});
        launch_vorticity_force(c, prm, F3{{force[0], force[1], force[2]}});
    });
}

ds_status ds_vorticity_vjp(ds_context* c, float* const velocity[3], const ds_parameters* params, float* const force_adj[3],
                           float* velocity_adj[3], const ds_parameter_adjoint* param_adj) {
    if (!c || !velocity || !force_adj || !velocity_adj || !velocity[0] || !velocity[1] || !velocity[2] || !force_adj[0] ||
        !force_adj[1] || !force_adj[2] || !velocity_adj[0] || !velocity_adj[1] || !velocity_adj[2]) {
        set_error("ds_vorticity_vjp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!c->vorticity || c->mode != DS_MODE_DIFFERENTIABLE) {
        set_error("ds_vorticity_vjp: needs a differentiable context created with vorticity_confinement_enabled = 1");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        Params prm{};
        prm.confinement = confinement_of(params);
        launch_vorticity_curl(c, CF3{{velocity[0], velocity[1], velocity[2]}});
        launch_vorticity_vjp(c, prm, CF3{{force_adj[0], force_adj[1], force_adj[2]}}, F3{{velocity_adj[0], velocity_adj[1], velocity_adj[2]}},
                             param_adj ? param_adj->vorticity_confinement : nullptr);
    });
}

}  // extern "C"
