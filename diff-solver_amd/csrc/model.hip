// model.hip — Model::forward_step / vjp_step (solvers/smoke/model.cpp:332-389) on gfx950 and the
// trajectory entry points (solvers/solver.ixx:53-150) driven by the checkpointed templates of
// xayah/solver.hpp through a thin adapter over the C-ABI views.
#include <hip/hip_runtime.h>

#include <cmath>
#include <functional>
#include <utility>
#include <vector>

#include "context.hpp"
#include "kernels.hpp"
#include "xayah/solver.hpp"

namespace ds {

void launch_advect_scalar_vjp(ds_context* c, int nf, int first_field, const float* src0, const float* src1, CF3 u, const float* oa0,
                              const float* oa1, float* sa0, float* sa1, F3 va);
void launch_advect_velocity_vjp(ds_context* c, CF3 u, CF3 oa, F3 va);
void launch_projection_vjp_faces(ds_context* c, const float* const ua[3], const float* const extra[3], const float* lh, const float* ll,
                                 F3 out, bool constrain);
void launch_face_accumulate(ds_context* c, CF3 in, F3 out);
void launch_cell_accumulate(ds_context* c, const float* in, float* out);
void launch_force_source_vjp(ds_context* c, CF3 forced_adj, const float* rho_s, const float* t_s, const float* rho_s_adj,
                             const float* t_s_adj, const Params& prm, F3 e_adj, float* prev_rho, float* prev_t, float* s_rho_adj,
                             float* s_t_adj, const ds_parameter_adjoint* param_adj);

namespace {

// Several buffers zeroed in one launch (fill kernels instead of memsets keep a step capturable).
struct ZeroList {
    float* p[8];
    uint32_t n[8];
    uint32_t count;
};

__global__ __launch_bounds__(256) void zero_list_kernel(ZeroList z) {
    for (uint32_t k = 0; k < z.count; ++k) {
        float* p = z.p[k];
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < z.n[k]; i += gridDim.x * blockDim.x) p[i] = 0.0f;
    }
}

// L = 0.5 ||rho - target||^2 (FP64 block partials) and rho_bar = rho - target.
__global__ __launch_bounds__(256) void density_loss_kernel(uint32_t n, const float* __restrict__ rho, const float* __restrict__ target,
                                                           float* __restrict__ rho_adj, double* __restrict__ partials) {
    __shared__ double lds[3 * kMaxWaves];
    double l = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = rho[i] - target[i];
        rho_adj[i] = d;
        l += 0.5 * (double)d * d;
    }
    double v[1]{l};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ partials, uint32_t count, double* out) {
    __shared__ double lds[3 * kMaxWaves];
    const double* parts[1]{partials};
    double red[1];
    reduce_partials<1>(parts, count, red, lds);
    if (threadIdx.x == 0) *out = red[0];
}

void zero_buffers(ds_context* c, std::initializer_list<std::pair<float*, uint32_t>> list) {
    ZeroList z{};
    for (const auto& e : list) {
        if (!e.first || e.second == 0) continue;
        if (z.count == 8) {
            zero_list_kernel<<<2048, kBlock, 0, c->stream>>>(z);
            z = ZeroList{};
        }
        z.p[z.count] = e.first;
        z.n[z.count] = e.second;
        ++z.count;
    }
    if (z.count) zero_list_kernel<<<2048, kBlock, 0, c->stream>>>(z);
    check_launch("zero_list_kernel");
}

// The extrapolated guess from the last solution x = hi + lo and the earlier ones in prev (p1 at
// [0, n), p2 at [n, 2n)): order 1, 2 x - p1; order 2, 3 (x - p1) + p2 (the polynomial through the
// last two or three solutions, at the next step); order 0 only records x. The history moves on by
// one: p2 = p1, p1 = x.
__global__ __launch_bounds__(256) void extrapolate_kernel(uint32_t n, int order, const float* __restrict__ hi, const float* __restrict__ lo,
                                                          float* __restrict__ prev, float* __restrict__ guess) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float x = hi[i] + lo[i];
        const float p1 = prev[i];
        if (order == 2) guess[i] = 3.0f * (x - p1) + prev[n + i];
        else if (order == 1) guess[i] = 2.0f * x - p1;
        prev[n + i] = p1;
        prev[i] = x;
    }
}

// The initial guess of a solve whose last solution is (hi, lo), valid or not, under
// DS_WARM_START_PREVIOUS (the last solution, in place) or DS_WARM_START_EXTRAPOLATE (the quadratic
// extrapolation of the last three solutions of the sweep, the linear one of the last two, the last
// solution). Returns the guess pair (null: x0 = 0) and moves the history on. prev = null: the
// previous solution only.
std::pair<const float*, const float*> sequence_guess(ds_context* c, float* hi, float* lo, bool valid, float* prev, int& prev_count,
                                                     bool previous_under_network = false) {
    uint32_t mode = c->config.warm_start;
    // the network corrects the extrapolated forward guess; adjoint solves keep the previous solution
    if (mode == DS_WARM_START_CNN) mode = previous_under_network ? DS_WARM_START_PREVIOUS : DS_WARM_START_EXTRAPOLATE;
    if (!valid || (mode != DS_WARM_START_PREVIOUS && mode != DS_WARM_START_EXTRAPOLATE)) return {nullptr, nullptr};
    if (mode == DS_WARM_START_EXTRAPOLATE && prev) {
        const int order = prev_count;
        extrapolate_kernel<<<2048, kBlock, 0, c->stream>>>(c->geo.n, order, hi, lo, prev, c->guess);
        check_launch("extrapolate_kernel");
        prev_count = std::min(prev_count + 1, 2);
        if (order > 0) return {c->guess, nullptr};
    }
    return {hi, lo};
}

void check_state(const ds_state* s, const char* what) {
    require(s && s->density && s->temperature && s->velocity[0] && s->velocity[1] && s->velocity[2], what);
}

// ds_state and ds_state_adjoint have the same layout (five FP32 field pointers).
ds_state as_state(const ds_state_adjoint& a) { return {a.density, a.temperature, {a.velocity[0], a.velocity[1], a.velocity[2]}}; }
ds_state_adjoint as_adjoint(const ds_state& a) { return {a.density, a.temperature, {a.velocity[0], a.velocity[1], a.velocity[2]}}; }

}  // namespace

// Model::forward_step (model.cpp:341-349): S(rho), S(T), F, V = integrate -> advect -> constrain,
// P (PCG), A_s(rho), A_s(T). Returns the pressure solve info.
ds_solve_info forward_step(ds_context* c, const ds_state* s, const ds_control* k, const ds_parameters* p, ds_state* next,
                           ds_step_cache* cache) {
    check_state(s, "forward_step: state has a null field");
    check_state(next, "forward_step: next_state has a null field");
    require(k && k->density_source && k->temperature_source && k->external_acceleration[0] && k->external_acceleration[1] &&
                k->external_acceleration[2],
            "forward_step: control has a null field");
    require(cache && cache->sourced_density && cache->sourced_temperature && cache->forced_velocity[0] && cache->forced_velocity[1] &&
                cache->forced_velocity[2],
            "forward_step: the step cache needs sourced fields and the forced velocity");
    const Params prm = params_of(p);
    const uint32_t n = c->geo.n;
    F3 f{{cache->force[0] ? cache->force[0] : c->force[0], cache->force[1] ? cache->force[1] : c->force[1],
          cache->force[2] ? cache->force[2] : c->force[2]}};
    F3 us{{cache->forced_velocity[0], cache->forced_velocity[1], cache->forced_velocity[2]}};
    F3 ua{{cache->advected_velocity[0] ? cache->advected_velocity[0] : c->advected[0],
           cache->advected_velocity[1] ? cache->advected_velocity[1] : c->advected[1],
           cache->advected_velocity[2] ? cache->advected_velocity[2] : c->advected[2]}};
    if (c->vorticity) launch_vorticity_curl(c, CF3{{s->velocity[0], s->velocity[1], s->velocity[2]}});
    launch_source_buoyancy(c, s->density, s->temperature, k->density_source, k->temperature_source,
                           CF3{{k->external_acceleration[0], k->external_acceleration[1], k->external_acceleration[2]}}, prm,
                           cache->sourced_density, cache->sourced_temperature, f);
    launch_integrate(c, CF3{{s->velocity[0], s->velocity[1], s->velocity[2]}}, CF3{{f.p[0], f.p[1], f.p[2]}}, us);
    launch_advect_velocity(c, CF3{{us.p[0], us.p[1], us.p[2]}}, ua, true);
    const float* x0h = nullptr;
    const float* x0l = nullptr;
    if (cache->pressure_guess) {
        x0h = cache->pressure_guess;
        x0l = cache->pressure_guess + n;
        c->pressure_prev_count = 0;   // a solve out of the step sequence
    } else {
        const auto g = sequence_guess(c, c->pressure_hi, c->pressure_lo, c->pressure_valid, c->pressure_prev, c->pressure_prev_count);
        x0h = g.first;
        x0l = g.second;
    }
    const float* uac[3]{ua.p[0], ua.p[1], ua.p[2]};
    const ds_solve_info si = pcg_solve_kind(c, 1, uac, nullptr, c->pressure_hi, c->pressure_lo, x0h, x0l, c->tolerance, c->max_iterations,
                                            c->config.warm_start == DS_WARM_START_CNN && !cache->pressure_guess);
    c->pressure_valid = true;
    launch_project(c, CF3{{ua.p[0], ua.p[1], ua.p[2]}}, c->pressure_hi, c->pressure_lo, F3{{next->velocity[0], next->velocity[1], next->velocity[2]}});
    launch_advect_scalars(c, 2, 0, cache->sourced_density, cache->sourced_temperature,
                          CF3{{next->velocity[0], next->velocity[1], next->velocity[2]}}, next->density, next->temperature);
    if (cache->pressure) copy_pair(c, c->pressure_hi, c->pressure_lo, cache->pressure, cache->pressure + n);
    return si;
}

// Model::vjp_step (model.cpp:361-389): scalar advection reverse (rho and T in one pass), projection
// reverse (adjoint PCG), constraint + velocity advection + integration reverse, buoyancy and source
// reverse, and with confinement on its reverse (vorticity recomputed from the input velocity of
// the step, state->velocity). prev_adj / control_adj / param_adj accumulate.
ds_solve_info vjp_step(ds_context* c, const ds_parameters* p, const ds_state* state, const ds_state* next, const ds_step_cache* cache,
                       const ds_state_adjoint* next_adj, ds_state_adjoint* prev_adj, ds_control_adjoint* control_adj,
                       const ds_parameter_adjoint* param_adj) {
    require(c->mode == DS_MODE_DIFFERENTIABLE, "vjp_step: the context was not created in differentiable mode");
    check_state(next, "vjp_step: next_state has a null field");
    if (c->vorticity)
        require(state && state->velocity[0] && state->velocity[1] && state->velocity[2],
                "vjp_step: vorticity confinement needs the step's input velocity (state)");
    require(next_adj && next_adj->density && next_adj->temperature && next_adj->velocity[0] && next_adj->velocity[1] &&
                next_adj->velocity[2],
            "vjp_step: next_state_adjoint has a null field");
    require(prev_adj && prev_adj->density && prev_adj->temperature && prev_adj->velocity[0] && prev_adj->velocity[1] &&
                prev_adj->velocity[2],
            "vjp_step: previous_state_adjoint has a null field");
    require(cache && cache->sourced_density && cache->sourced_temperature && cache->forced_velocity[0] && cache->forced_velocity[1] &&
                cache->forced_velocity[2],
            "vjp_step: the step cache needs sourced fields and the forced velocity");
    const Params prm = params_of(p);
    const Geo& g = c->geo;
    zero_buffers(c, {{c->sourced_adj[0], g.n},
                     {c->sourced_adj[1], g.n},
                     {c->projected_extra[0], g.nf[0]},
                     {c->projected_extra[1], g.nf[1]},
                     {c->projected_extra[2], g.nf[2]},
                     {c->forced_adj[0], g.nf[0]},
                     {c->forced_adj[1], g.nf[1]},
                     {c->forced_adj[2], g.nf[2]}});
    const CF3 nv{{next->velocity[0], next->velocity[1], next->velocity[2]}};
    launch_advect_scalar_vjp(c, 2, 0, cache->sourced_density, cache->sourced_temperature, nv, next_adj->density, next_adj->temperature,
                             c->sourced_adj[0], c->sourced_adj[1], F3{{c->projected_extra[0], c->projected_extra[1], c->projected_extra[2]}});
    const float* na[3]{next_adj->velocity[0], next_adj->velocity[1], next_adj->velocity[2]};
    const float* ex[3]{c->projected_extra[0], c->projected_extra[1], c->projected_extra[2]};
    // adjoint solves start from the previous adjoint solution under both sequence modes (on the
    // 128^3 plume the extrapolated adjoint guess took more iterations, 606 per solve, than the
    // previous solution, 566). Under DS_WARM_START_CNN the network does not correct it: on the same
    // trajectory the corrected guess took 575 iterations per adjoint solve against the previous
    // solution's 565, and each correction costs an inference (0.55 ms at 128^3 through the 64^3 grid)
    int no_prev = 0;
    const auto lg = sequence_guess(c, c->lambda_hi, c->lambda_lo, c->lambda_valid, nullptr, no_prev, true);
    const ds_solve_info si = pcg_solve_kind(c, 2, na, ex, c->lambda_hi, c->lambda_lo, lg.first, lg.second, c->tolerance, c->max_iterations,
                                            false);
    c->lambda_valid = true;
    launch_projection_vjp_faces(c, na, ex, c->lambda_hi, c->lambda_lo, F3{{c->raw_adj[0], c->raw_adj[1], c->raw_adj[2]}}, true);
    launch_advect_velocity_vjp(c, CF3{{cache->forced_velocity[0], cache->forced_velocity[1], cache->forced_velocity[2]}},
                               CF3{{c->raw_adj[0], c->raw_adj[1], c->raw_adj[2]}}, F3{{c->forced_adj[0], c->forced_adj[1], c->forced_adj[2]}});
    const CF3 fa{{c->forced_adj[0], c->forced_adj[1], c->forced_adj[2]}};
    launch_face_accumulate(c, fa, F3{{prev_adj->velocity[0], prev_adj->velocity[1], prev_adj->velocity[2]}});
    F3 ea{{nullptr, nullptr, nullptr}};
    float* sra = nullptr;
    float* sta = nullptr;
    if (control_adj) {
        ea = F3{{control_adj->external_acceleration[0], control_adj->external_acceleration[1], control_adj->external_acceleration[2]}};
        if (!ea.p[0] || !ea.p[1] || !ea.p[2]) ea = F3{{nullptr, nullptr, nullptr}};
        sra = control_adj->density_source;
        sta = control_adj->temperature_source;
    }
    if (c->vorticity) launch_vorticity_curl(c, CF3{{state->velocity[0], state->velocity[1], state->velocity[2]}});
    launch_force_source_vjp(c, fa, cache->sourced_density, cache->sourced_temperature, c->sourced_adj[0], c->sourced_adj[1], prm, ea,
                            prev_adj->density, prev_adj->temperature, sra, sta, param_adj);
    if (c->vorticity) launch_vorticity_vjp_tail(c, F3{{prev_adj->velocity[0], prev_adj->velocity[1], prev_adj->velocity[2]}});
    return si;
}

namespace {

// ---- adapter of the C-ABI views to the templates of xayah/solver.hpp ----------------------------
template <typename T>
struct DBuf {   // stream-ordered device buffer
    T* p = nullptr;
    hipStream_t s = nullptr;
    DBuf() = default;
    DBuf(hipStream_t stream, size_t n) : s(stream) {
        void* q = nullptr;
        DS_HIP(hipMallocAsync(&q, n * sizeof(T) > 0 ? n * sizeof(T) : 1, stream));
        p = static_cast<T*>(q);
    }
    ~DBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(std::exchange(o.p, nullptr)), s(o.s) {}
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipFreeAsync(p, s);
            p = std::exchange(o.p, nullptr);
            s = o.s;
        }
        return *this;
    }
};

struct AState {   // owning (or borrowing) five-field state
    ds_state v{};
    std::vector<DBuf<float>> own;
};
struct ACache {
    DBuf<float> rs, ts, us[3];
    ds_step_cache v{};
};
struct ATape {
    DBuf<float> buf;
    float* data = nullptr;
};
struct ACtx {
    ds_context* c;
    ds_trajectory_info* info;
    ds_solve_info last{};
    ds_status soft = DS_OK;   // NOT_CONVERGED is recorded and the sweep continues
};

void record(ds_trajectory_info* t, const ds_solve_info& si, int kind) {
    if (!t) return;
    if (kind == 0) t->forward_solves++, t->forward_iterations += si.iterations;
    else if (kind == 1) t->recompute_solves++, t->recompute_iterations += si.iterations;
    else t->adjoint_solves++, t->adjoint_iterations += si.iterations;
    t->max_iterations = std::max(t->max_iterations, si.iterations);
    if (!si.converged) t->not_converged++;
    if (si.guess_rejected) t->guesses_rejected++;
    if (si.rhs_norm > 0.0) t->max_relative_residual = std::max(t->max_relative_residual, si.final_residual / si.rhs_norm);
}

struct AbiModel {
    AState allocate_state(ACtx& x) const {
        const Geo& g = x.c->geo;
        AState s;
        s.own.emplace_back(x.c->stream, g.n);
        s.own.emplace_back(x.c->stream, g.n);
        for (int a = 0; a < 3; ++a) s.own.emplace_back(x.c->stream, g.nf[a]);
        s.v.density = s.own[0].p;
        s.v.temperature = s.own[1].p;
        for (int a = 0; a < 3; ++a) s.v.velocity[a] = s.own[2 + a].p;
        return s;
    }
    ACache allocate_step_cache(ACtx& x) const {
        const Geo& g = x.c->geo;
        ACache k;
        k.rs = DBuf<float>(x.c->stream, g.n);
        k.ts = DBuf<float>(x.c->stream, g.n);
        for (int a = 0; a < 3; ++a) k.us[a] = DBuf<float>(x.c->stream, g.nf[a]);
        k.v.sourced_density = k.rs.p;
        k.v.sourced_temperature = k.ts.p;
        for (int a = 0; a < 3; ++a) k.v.forced_velocity[a] = k.us[a].p;
        return k;
    }
    AState allocate_state_adjoint(ACtx& x) const {
        AState s = allocate_state(x);
        zero_state_adjoint(s, x);
        return s;
    }
    ds_control_adjoint allocate_control_adjoint(ACtx&) const { return {}; }
    ds_parameter_adjoint allocate_parameter_adjoint(ACtx&) const { return {}; }
    ATape allocate_pressure_tape(ACtx& x) const {
        ATape t;
        t.buf = DBuf<float>(x.c->stream, 2 * (size_t)x.c->geo.n);
        t.data = t.buf.p;
        return t;
    }
    static void set_pressure_out(ACache& k, float* p) { k.v.pressure = p; }
    static void set_pressure_guess(ACache& k, const float* p) { k.v.pressure_guess = p; }
    void copy_state(const AState& s, AState& d, ACtx& x) const {
        const Geo& g = x.c->geo;
        DS_HIP(hipMemcpyAsync(d.v.density, s.v.density, g.n * sizeof(float), hipMemcpyDeviceToDevice, x.c->stream));
        DS_HIP(hipMemcpyAsync(d.v.temperature, s.v.temperature, g.n * sizeof(float), hipMemcpyDeviceToDevice, x.c->stream));
        for (int a = 0; a < 3; ++a)
            DS_HIP(hipMemcpyAsync(d.v.velocity[a], s.v.velocity[a], g.nf[a] * sizeof(float), hipMemcpyDeviceToDevice, x.c->stream));
    }
    void copy_state_adjoint(const AState& s, AState& d, ACtx& x) const { copy_state(s, d, x); }
    void zero_state_adjoint(AState& s, ACtx& x) const {
        const Geo& g = x.c->geo;
        zero_buffers(x.c, {{s.v.density, g.n}, {s.v.temperature, g.n}, {s.v.velocity[0], g.nf[0]}, {s.v.velocity[1], g.nf[1]},
                           {s.v.velocity[2], g.nf[2]}});
    }
    void accumulate_state_adjoint(const AState& s, AState& d, ACtx& x) const {
        launch_cell_accumulate(x.c, s.v.density, d.v.density);
        launch_cell_accumulate(x.c, s.v.temperature, d.v.temperature);
        launch_face_accumulate(x.c, CF3{{s.v.velocity[0], s.v.velocity[1], s.v.velocity[2]}},
                               F3{{d.v.velocity[0], d.v.velocity[1], d.v.velocity[2]}});
    }
    void forward_step(const AState& s, const ds_control& k, const ds_parameters& p, AState& n, ACache& cache, ACtx& x) const {
        x.last = ds::forward_step(x.c, &s.v, &k, &p, &n.v, &cache.v);
    }
    // forward mode: tangents have the layout of states (ds_state_tangent), controls and parameters
    AState allocate_state_tangent(ACtx& x) const { return allocate_state(x); }
    ds_control allocate_control_tangent(ACtx&) const { return {}; }
    ds_parameters allocate_parameter_tangent(ACtx&) const { return {}; }
    void copy_state_tangent(const AState& s, AState& d, ACtx& x) const { copy_state(s, d, x); }
    void jvp_step(const AState& s, const ds_control&, const ds_parameters& p, const AState& n, const ACache& cache, const AState& st,
                  const ds_control& kt, const ds_parameters& pt, AState& nt, ACtx& x) const {
        x.last = ds::jvp_step(x.c, &p, &pt, &s.v, &n.v, &cache.v, &st.v, &kt, &nt.v);
    }
    static void after_jvp(ACtx& x) {
        record(x.info, x.last, 2);
        if (!x.last.converged) x.soft = std::isfinite(x.last.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    void vjp_step(const AState& s, const ds_control&, const ds_parameters& p, const AState& n, const ACache& cache, const AState& na,
                  AState& pa, ds_control_adjoint& ca, ds_parameter_adjoint& qa, ACtx& x) const {
        const ds_state_adjoint nav = as_adjoint(na.v);
        ds_state_adjoint pav = as_adjoint(pa.v);
        const bool any_param = qa.density_buoyancy || qa.ambient_temperature || qa.temperature_buoyancy || qa.vorticity_confinement;
        x.last = ds::vjp_step(x.c, &p, &s.v, &n.v, &cache.v, &nav, &pav, ca.density_source || ca.external_acceleration[0] ? &ca : nullptr,
                              any_param ? &qa : nullptr);
    }
    static void after_forward(ACtx& x, int kind) {
        record(x.info, x.last, kind);
        if (!x.last.converged) x.soft = std::isfinite(x.last.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    static void after_vjp(ACtx& x) {
        record(x.info, x.last, 2);
        if (!x.last.converged) x.soft = std::isfinite(x.last.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
};

AState borrow(const ds_state& v) {
    AState s;
    s.v = v;
    return s;
}

}  // namespace

}  // namespace ds

using namespace ds;

extern "C" {

ds_status ds_copy_state(ds_context* c, const ds_state* src, ds_state* dst) {
    if (!c || !src || !dst) {
        set_error("ds_copy_state: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        check_state(src, "ds_copy_state: null field");
        check_state(dst, "ds_copy_state: null field");
        ACtx x{c, nullptr};
        AState d = borrow(*dst);
        AbiModel{}.copy_state(borrow(*src), d, x);
    });
}

ds_status ds_copy_state_adjoint(ds_context* c, const ds_state_adjoint* src, ds_state_adjoint* dst) {
    if (!src || !dst) {
        set_error("ds_copy_state_adjoint: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    const ds_state s = as_state(*src);
    ds_state d = as_state(*dst);
    return ds_copy_state(c, &s, &d);
}

ds_status ds_accumulate_state_adjoint(ds_context* c, const ds_state_adjoint* src, ds_state_adjoint* dst) {
    if (!c || !src || !dst) {
        set_error("ds_accumulate_state_adjoint: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        const ds_state s = as_state(*src);
        const ds_state dv = as_state(*dst);
        check_state(&s, "ds_accumulate_state_adjoint: null field");
        check_state(&dv, "ds_accumulate_state_adjoint: null field");
        ACtx x{c, nullptr};
        AState d = borrow(dv);
        AbiModel{}.accumulate_state_adjoint(borrow(s), d, x);
    });
}

ds_status ds_forward_step(ds_context* c, const ds_state* state, const ds_control* control, const ds_parameters* params,
                          ds_state* next_state, ds_step_cache* cache, ds_step_info* info) {
    if (!c) {
        set_error("ds_forward_step: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ds_solve_info si{};
    const ds_status st = guard([&] { si = forward_step(c, state, control, params, next_state, cache); });
    if (info) info->projection = si;
    if (st != DS_OK) return st;
    if (!si.converged) {
        set_error("ds_forward_step: pressure solve did not converge");
        return std::isfinite(si.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    return DS_OK;
}

ds_status ds_vjp_step(ds_context* c, const ds_state* state, const ds_control* control, const ds_parameters* params,
                      const ds_state* next_state, const ds_step_cache* cache, const ds_state_adjoint* next_adj, ds_state_adjoint* prev_adj,
                      ds_control_adjoint* control_adj, const ds_parameter_adjoint* param_adj, ds_step_info* info) {
    (void)control;
    if (!c) {
        set_error("ds_vjp_step: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ds_solve_info si{};
    const ds_status st = guard([&] { si = vjp_step(c, params, state, next_state, cache, next_adj, prev_adj, control_adj, param_adj); });
    if (info) info->projection = si;
    if (st != DS_OK) return st;
    if (!si.converged) {
        set_error("ds_vjp_step: adjoint pressure solve did not converge");
        return std::isfinite(si.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    return DS_OK;
}

ds_status ds_trajectory_jvp(ds_context* c, const ds_state* initial, const ds_control* controls, uint32_t control_count,
                            const ds_parameters* params, uint32_t steps, const ds_state_tangent* initial_tangent,
                            const ds_control_tangent* control_tangents, const ds_parameter_tangent* param_tangent, ds_state* final_state,
                            ds_state_tangent* final_tangent, ds_trajectory_info* info) {
    if (!c || !initial || !controls || !params || !initial_tangent || !control_tangents || !param_tangent || !final_state ||
        !final_tangent || (control_count != 1 && control_count != steps)) {
        set_error("ds_trajectory_jvp: invalid argument (control_count must be 1 or steps)");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (c->mode != DS_MODE_DIFFERENTIABLE) {
        set_error("ds_trajectory_jvp: the context was not created in differentiable mode");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ACtx x{c, info};
    if (info) {
        *info = ds_trajectory_info{};
        info->steps = steps;
    }
    const ds_status st = guard([&] {
        check_state(initial, "ds_trajectory_jvp: initial state has a null field");
        check_state(initial_tangent, "ds_trajectory_jvp: initial tangent has a null field");
        check_state(final_state, "ds_trajectory_jvp: final state has a null field");
        check_state(final_tangent, "ds_trajectory_jvp: final tangent has a null field");
        AbiModel m;
        std::function<const ds_control&(uint32_t)> ctrl = [&](uint32_t s) -> const ds_control& {
            return controls[control_count == 1 ? 0 : s];
        };
        std::function<const ds_control&(uint32_t)> ctrl_t = [&](uint32_t s) -> const ds_control& {
            return control_tangents[control_count == 1 ? 0 : s];
        };
        auto r = xayah::solver::jvp_streaming<AbiModel, ACtx, AState, ds_control, ds_parameters, AState, ds_control, ds_parameters>(
            m, x, borrow(*initial), ctrl, *params, steps, borrow(*initial_tangent), ctrl_t, *param_tangent);
        AState out = borrow(*final_state), out_t = borrow(*final_tangent);
        m.copy_state(r.first, out, x);
        m.copy_state(r.second, out_t, x);
        DS_HIP(hipStreamSynchronize(c->stream));
    });
    if (st != DS_OK) return st;
    if (x.soft != DS_OK) set_error("ds_trajectory_jvp: a pressure solve did not converge");
    return x.soft;
}

ds_status ds_simulate(ds_context* c, const ds_state* initial, const ds_control* controls, uint32_t control_count,
                      const ds_parameters* params, uint32_t steps, ds_state* final_state, ds_trajectory_info* info) {
    if (!c || !initial || !controls || !params || !final_state || (control_count != 1 && control_count != steps)) {
        set_error("ds_simulate: invalid argument (control_count must be 1 or steps)");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ACtx x{c, info};
    if (info) {
        *info = ds_trajectory_info{};
        info->steps = steps;
        info->checkpoint_interval = steps;
    }
    const ds_status st = guard([&] {
        check_state(initial, "ds_simulate: initial state has a null field");
        check_state(final_state, "ds_simulate: final state has a null field");
        c->pressure_prev_count = 0;   // an extrapolated or projected guess spans solves of one sweep only
        window_reset(c);
        AbiModel m;
        std::function<const ds_control&(uint32_t)> ctrl = [&](uint32_t s) -> const ds_control& {
            return controls[control_count == 1 ? 0 : s];
        };
        auto t = xayah::solver::simulate_checkpointed<AbiModel, ACtx, AState, ds_control, ds_parameters>(
            m, x, borrow(*initial), ctrl, *params, steps, steps + 1u, false);
        AState out = borrow(*final_state);
        m.copy_state(*t.final_state, out, x);
        DS_HIP(hipStreamSynchronize(c->stream));
    });
    if (st != DS_OK) return st;
    if (x.soft != DS_OK) set_error("ds_simulate: a pressure solve did not converge");
    return x.soft;
}

}  // extern "C"

namespace {

// Checkpointed forward pass and reverse sweep (solver.ixx:115-150 with a trajectory adjoint that
// is zero except at the final state). The terminal adjoint is either the density loss
// 0.5 ||rho_T - target||^2 (target_density) or given (terminal_adj).
ds_status trajectory_gradient(const char* who, ds_context* c, const ds_state* initial, const ds_control* controls,
                              uint32_t control_count, const ds_parameters* params, uint32_t steps, uint32_t checkpoint_interval,
                              const float* target_density, const ds_state_adjoint* terminal_adj, ds_state* final_state,
                              ds_state_adjoint* initial_adj, ds_control_adjoint* control_adj, const ds_parameter_adjoint* param_adj,
                              ds_trajectory_info* info) {
    if (!c || !initial || !controls || !params || !(target_density || terminal_adj) || !initial_adj || steps == 0 ||
        (control_count != 1 && control_count != steps)) {
        set_error(std::string(who) + ": invalid argument (control_count must be 1 or steps; steps > 0)");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (c->mode != DS_MODE_DIFFERENTIABLE) {
        set_error(std::string(who) + ": the context was not created in differentiable mode");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ds_trajectory_info local{};
    ds_trajectory_info* ti = info ? info : &local;
    *ti = ds_trajectory_info{};
    ti->steps = steps;
    ti->checkpoint_interval = checkpoint_interval ? checkpoint_interval : 1u;
    ACtx x{c, ti};
    const ds_status st = guard([&] {
        check_state(initial, "initial state has a null field");
        const ds_state init_view = as_state(*initial_adj);
        check_state(&init_view, "initial adjoint has a null field");
        if (terminal_adj) {
            const ds_state tv = as_state(*terminal_adj);
            check_state(&tv, "final-state adjoint has a null field");
        }
        AbiModel m;
        std::function<const ds_control&(uint32_t)> ctrl = [&](uint32_t s) -> const ds_control& {
            return controls[control_count == 1 ? 0 : s];
        };
        c->pressure_prev_count = 0;   // an extrapolated or projected guess spans solves of one sweep only
        window_reset(c);
        auto t = xayah::solver::simulate_checkpointed<AbiModel, ACtx, AState, ds_control, ds_parameters>(
            m, x, borrow(*initial), ctrl, *params, steps, ti->checkpoint_interval, true);
        // terminal adjoint (and loss)
        AState final_adj = m.allocate_state_adjoint(x);
        if (target_density) {
            DBuf<double> partials(c->stream, 1024), loss(c->stream, 1);
            density_loss_kernel<<<1024, kBlock, 0, c->stream>>>(c->geo.n, t.final_state->v.density, target_density, final_adj.v.density,
                                                                partials.p);
            sum_partials_kernel<<<1, kBlock, 0, c->stream>>>(partials.p, 1024, loss.p);
            check_launch("density_loss_kernel");
            DS_HIP(hipMemcpyAsync(&ti->loss, loss.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        } else {
            m.copy_state_adjoint(borrow(as_state(*terminal_adj)), final_adj, x);
        }
        if (final_state) {
            AState out = borrow(*final_state);
            m.copy_state(*t.final_state, out, x);
        }
        ds_control_adjoint none{};
        std::function<ds_control_adjoint&(uint32_t)> cadj = [&](uint32_t s) -> ds_control_adjoint& {
            if (!control_adj) return none;
            return control_adj[control_count == 1 ? 0 : s];
        };
        ds_parameter_adjoint qa = param_adj ? *param_adj : ds_parameter_adjoint{};
        AState init = borrow(init_view);
        xayah::solver::vjp_checkpointed<AbiModel, ACtx, AState, ds_control, ds_parameters, AState, ds_control_adjoint, ds_parameter_adjoint>(
            m, x, t, ctrl, *params, final_adj, cadj, qa, init);
        DS_HIP(hipStreamSynchronize(c->stream));
    });
    if (st != DS_OK) return st;
    if (x.soft != DS_OK) set_error(std::string(who) + ": a pressure solve did not converge");
    return x.soft;
}

}  // namespace

extern "C" {

ds_status ds_loss_gradient(ds_context* c, const ds_state* initial, const ds_control* controls, uint32_t control_count,
                           const ds_parameters* params, uint32_t steps, uint32_t checkpoint_interval, const float* target_density,
                           ds_state* final_state, ds_state_adjoint* initial_adj, ds_control_adjoint* control_adj,
                           const ds_parameter_adjoint* param_adj, ds_trajectory_info* info) {
    if (!target_density) {
        set_error("ds_loss_gradient: null target");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return trajectory_gradient("ds_loss_gradient", c, initial, controls, control_count, params, steps, checkpoint_interval,
                               target_density, nullptr, final_state, initial_adj, control_adj, param_adj, info);
}

ds_status ds_trajectory_vjp(ds_context* c, const ds_state* initial, const ds_control* controls, uint32_t control_count,
                            const ds_parameters* params, uint32_t steps, uint32_t checkpoint_interval,
                            const ds_state_adjoint* final_adj, ds_state* final_state, ds_state_adjoint* initial_adj,
                            ds_control_adjoint* control_adj, const ds_parameter_adjoint* param_adj, ds_trajectory_info* info) {
    if (!final_adj) {
        set_error("ds_trajectory_vjp: null final-state adjoint");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return trajectory_gradient("ds_trajectory_vjp", c, initial, controls, control_count, params, steps, checkpoint_interval,
                               nullptr, final_adj, final_state, initial_adj, control_adj, param_adj, info);
}

}  // extern "C"
