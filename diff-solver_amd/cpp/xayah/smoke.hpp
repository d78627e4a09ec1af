// xayah/smoke.hpp — C++ API of the MI355X smoke solver, shaped like the reference's
// xayah.smoke modules (solvers/smoke/data.ixx, model.ixx) and xayah.cuda (solvers/cuda.ixx):
// Configuration, State/Control/Parameters with their tangents and adjoints, Model with
// allocate_context / allocate_* / copy_* / accumulate_state_adjoint / forward_step / jvp_step /
// vjp_step.
//
// Header-only over the C ABI in ds/ds_smoke.h. Differences from the reference, by design:
//   * adjoint fields are FP32 (adjoint_t), parameter adjoints FP64;
//   * Configuration adds the pressure solver fields (projection_solver, cg_tolerance,
//     cg_max_iterations, warm_start, cg_variant); vorticity confinement defaults to on, as in
//     the reference (data.ixx:82);
//   * errors surface as std::runtime_error carrying ds_last_error() (as check_launch did);
//   * StepCache carries an optional pressure pair used to warm-start a recompute.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "ds/ds_smoke.h"

namespace xayah {

namespace hip {

inline void check(ds_status status, const char* what) {
    if (status != DS_OK) throw std::runtime_error(std::string(what) + ": " + ds_last_error());
}

// Owns a ds_context (stream + domain + scratch); the role of cuda::Resource (cuda.ixx:7-28).
struct Resource {
    explicit Resource(ds_context* handle) : handle(handle) {}
    ~Resource() noexcept {
        if (handle) ds_context_destroy(handle);
    }
    Resource(const Resource&) = delete;
    Resource& operator=(const Resource&) = delete;

    [[nodiscard]] void* allocate(std::size_t bytes) {
        void* p = nullptr;
        check(ds_allocate(handle, bytes, &p), "ds_allocate");
        return p;
    }
    void release(void* p) noexcept { (void)ds_release(handle, p); }
    void copy_from_host(void* d, const void* h, std::size_t bytes) { check(ds_copy_from_host(handle, d, h, bytes), "ds_copy_from_host"); }
    void copy_to_host(void* h, const void* d, std::size_t bytes) { check(ds_copy_to_host(handle, h, d, bytes), "ds_copy_to_host"); }
    void copy_device(void* d, const void* s, std::size_t bytes) { check(ds_copy_device(handle, d, s, bytes), "ds_copy_device"); }
    void zero(void* d, std::size_t bytes) { check(ds_zero(handle, d, bytes), "ds_zero"); }
    void synchronize() { check(ds_synchronize(handle), "ds_synchronize"); }
    [[nodiscard]] void* native_stream() const { return ds_context_stream(handle); }

    ds_context* handle;
};

// Move-only device buffer with stream-ordered release (cuda.ixx:30-59).
template <typename Element>
struct Buffer {
    Buffer() noexcept : data(nullptr), size(0) {}
    Buffer(std::shared_ptr<Resource> resource, std::size_t n)
        : data(n == 0 ? nullptr : static_cast<Element*>(resource->allocate(n * sizeof(Element)))), size(n), resource_(std::move(resource)) {}
    ~Buffer() noexcept {
        if (data) resource_->release(data);
    }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    Buffer(Buffer&& o) noexcept : data(std::exchange(o.data, nullptr)), size(std::exchange(o.size, 0u)), resource_(std::move(o.resource_)) {}
    Buffer& operator=(Buffer&& o) noexcept {
        if (this != &o) {
            if (data) resource_->release(data);
            resource_ = std::move(o.resource_);
            data = std::exchange(o.data, nullptr);
            size = std::exchange(o.size, 0u);
        }
        return *this;
    }
    Element* data;
    std::size_t size;

private:
    std::shared_ptr<Resource> resource_;
};

}  // namespace hip

namespace smoke {

using adjoint_t = float;

struct Vector3 {
    float x, y, z;
};

enum class ScalarBoundaryMode : std::uint32_t { fixed_value, zero_gradient, periodic };
enum class VelocityBoundaryMode : std::uint32_t { fixed_value, zero_gradient, normal_fixed_tangent_zero_gradient, periodic };

struct ScalarBoundaryFace {
    ScalarBoundaryMode mode{ScalarBoundaryMode::zero_gradient};
    float value{};
};
struct VelocityBoundaryFace {
    VelocityBoundaryMode mode{VelocityBoundaryMode::zero_gradient};
    Vector3 value{};
};
struct ScalarBoundary {
    ScalarBoundaryFace x_min{}, x_max{}, y_min{}, y_max{}, z_min{}, z_max{};
};
struct VelocityBoundary {
    VelocityBoundaryFace x_min{}, x_max{}, y_min{}, y_max{}, z_min{}, z_max{};
};
struct Ellipsoid {
    Vector3 center{};
    Vector3 radius{1.0F, 1.0F, 1.0F};
};
struct Box {
    Vector3 center{};
    Vector3 half_extent{1.0F, 1.0F, 1.0F};
};
struct Collider {
    std::variant<Ellipsoid, Box> shape{Ellipsoid{}};
    Vector3 velocity{};
    float density{};
    float temperature{};
};

enum class ProjectionSolver : std::uint32_t { pcg = DS_PROJECTION_PCG, rbgs = DS_PROJECTION_RBGS };
enum class WarmStart : std::uint32_t {
    none = DS_WARM_START_NONE,
    previous = DS_WARM_START_PREVIOUS,
    cnn = DS_WARM_START_CNN,
    extrapolate = DS_WARM_START_EXTRAPOLATE,
    projection = DS_WARM_START_PROJECTION
};
enum class CgVariant : std::uint32_t {
    single_reduction = DS_CG_SINGLE_REDUCTION,
    two_reduction = DS_CG_TWO_REDUCTION,
    single_reduction_resident = DS_CG_SINGLE_REDUCTION_RESIDENT,
    s_step = DS_CG_S_STEP,
    multigrid = DS_CG_MULTIGRID
};

struct Configuration {
    std::uint32_t resolution[3]{32u, 48u, 32u};
    float cell_size{1.0F / 32.0F};
    float time_step{1.0F / 60.0F};
    std::uint32_t pressure_iterations{80u};
    VelocityBoundary velocity_boundary{};
    ScalarBoundary pressure_boundary{};
    ScalarBoundary density_boundary{};
    ScalarBoundary temperature_boundary{};
    std::vector<Collider> colliders{};
    bool vorticity_confinement_enabled{true};   // data.ixx:82
    ProjectionSolver projection_solver{ProjectionSolver::pcg};
    double cg_tolerance{1e-6};
    std::uint32_t cg_max_iterations{0u};
    WarmStart warm_start{WarmStart::none};
    CgVariant cg_variant{CgVariant::multigrid};
};

struct ScalarField {
    hip::Buffer<float> values;
};
struct CenteredVectorField {
    ScalarField x, y, z;
};
struct StaggeredVectorField {
    hip::Buffer<float> x, y, z;
};
struct ScalarAdjointField {
    hip::Buffer<adjoint_t> values;
};
struct CenteredVectorAdjointField {
    ScalarAdjointField x, y, z;
};
struct StaggeredVectorAdjointField {
    hip::Buffer<adjoint_t> x, y, z;
};
struct State {
    ScalarField density;
    ScalarField temperature;
    StaggeredVectorField velocity;
};
struct Control {
    ScalarField density_source;
    ScalarField temperature_source;
    CenteredVectorField external_acceleration;
};
struct Parameters {
    hip::Buffer<float> ambient_temperature, density_buoyancy, temperature_buoyancy, vorticity_confinement;
};
// StateTangent / ControlTangent / ParameterTangent (data.ixx:136-153): FP32 fields with the
// members of their primal twins.
using StateTangent = State;
using ControlTangent = Control;
using ParameterTangent = Parameters;
struct StateAdjoint {
    ScalarAdjointField density;
    ScalarAdjointField temperature;
    StaggeredVectorAdjointField velocity;
};
struct ControlAdjoint {
    ScalarAdjointField density_source;
    ScalarAdjointField temperature_source;
    CenteredVectorAdjointField external_acceleration;
};
struct ParameterAdjoint {
    hip::Buffer<double> ambient_temperature, density_buoyancy, temperature_buoyancy, vorticity_confinement;
};
struct StepCache {
    ScalarField sourced_density;
    ScalarField sourced_temperature;
    CenteredVectorField force;
    StaggeredVectorField forced_velocity;
    StaggeredVectorField advected_velocity;
    float* pressure = nullptr;               // optional out: compensated pair, 2 * cell_count floats
    const float* pressure_guess = nullptr;   // optional in: initial guess pair of the pressure solve
};

enum class ExecutionMode : std::uint32_t { forward = DS_MODE_FORWARD, differentiable = DS_MODE_DIFFERENTIABLE };

struct ExecutionContext {
    std::shared_ptr<hip::Resource> resource;
    ds_domain_info domain{};
    ds_step_info last_step{};   // solve statistics of the last forward_step / vjp_step
};

// ---- views for the C ABI ------------------------------------------------------------------
inline ds_state view(const State& s) {
    return {s.density.values.data, s.temperature.values.data, {s.velocity.x.data, s.velocity.y.data, s.velocity.z.data}};
}
inline ds_control view(const Control& c) {
    return {c.density_source.values.data,
            c.temperature_source.values.data,
            {c.external_acceleration.x.values.data, c.external_acceleration.y.values.data, c.external_acceleration.z.values.data}};
}
inline ds_parameters view(const Parameters& p) {
    return {p.ambient_temperature.data, p.density_buoyancy.data, p.temperature_buoyancy.data, p.vorticity_confinement.data};
}
inline ds_state_adjoint view(const StateAdjoint& s) {
    return {s.density.values.data, s.temperature.values.data, {s.velocity.x.data, s.velocity.y.data, s.velocity.z.data}};
}
inline ds_control_adjoint view(const ControlAdjoint& c) {
    return {c.density_source.values.data,
            c.temperature_source.values.data,
            {c.external_acceleration.x.values.data, c.external_acceleration.y.values.data, c.external_acceleration.z.values.data}};
}
inline ds_parameter_adjoint view(const ParameterAdjoint& p) {
    return {p.ambient_temperature.data, p.density_buoyancy.data, p.temperature_buoyancy.data, p.vorticity_confinement.data};
}
inline ds_step_cache view(const StepCache& c) {
    return {c.sourced_density.values.data,
            c.sourced_temperature.values.data,
            {c.force.x.values.data, c.force.y.values.data, c.force.z.values.data},
            {c.forced_velocity.x.data, c.forced_velocity.y.data, c.forced_velocity.z.data},
            {c.advected_velocity.x.data, c.advected_velocity.y.data, c.advected_velocity.z.data},
            c.pressure,
            c.pressure_guess};
}

inline ds_config to_abi(const Configuration& cfg, std::vector<ds_collider>& colliders) {
    ds_config out{};
    for (int k = 0; k < 3; ++k) out.resolution[k] = cfg.resolution[k];
    out.cell_size = cfg.cell_size;
    out.time_step = cfg.time_step;
    out.pressure_iterations = cfg.pressure_iterations;
    const VelocityBoundaryFace vf[6]{cfg.velocity_boundary.x_min, cfg.velocity_boundary.x_max, cfg.velocity_boundary.y_min,
                                     cfg.velocity_boundary.y_max, cfg.velocity_boundary.z_min, cfg.velocity_boundary.z_max};
    for (int f = 0; f < 6; ++f) {   // operators.cpp:95-105
        out.velocity_boundary.modes[f] = static_cast<std::uint32_t>(vf[f].mode);
        out.velocity_boundary.values[3 * f + 0] = vf[f].value.x;
        out.velocity_boundary.values[3 * f + 1] = vf[f].value.y;
        out.velocity_boundary.values[3 * f + 2] = vf[f].value.z;
    }
    auto pack = [](const ScalarBoundary& b, ds_scalar_boundary& o) {   // operators.cpp:85-93
        const ScalarBoundaryFace f[6]{b.x_min, b.x_max, b.y_min, b.y_max, b.z_min, b.z_max};
        for (int k = 0; k < 6; ++k) {
            o.modes[k] = static_cast<std::uint32_t>(f[k].mode);
            o.values[k] = f[k].value;
        }
    };
    pack(cfg.pressure_boundary, out.pressure_boundary);
    pack(cfg.density_boundary, out.density_boundary);
    pack(cfg.temperature_boundary, out.temperature_boundary);
    colliders.clear();
    for (const Collider& c : cfg.colliders) {
        ds_collider d{};
        if (const Ellipsoid* e = std::get_if<Ellipsoid>(&c.shape)) {
            d.shape = DS_SHAPE_ELLIPSOID;
            d.center[0] = e->center.x, d.center[1] = e->center.y, d.center[2] = e->center.z;
            d.size[0] = e->radius.x, d.size[1] = e->radius.y, d.size[2] = e->radius.z;
        } else {
            const Box& b = std::get<Box>(c.shape);
            d.shape = DS_SHAPE_BOX;
            d.center[0] = b.center.x, d.center[1] = b.center.y, d.center[2] = b.center.z;
            d.size[0] = b.half_extent.x, d.size[1] = b.half_extent.y, d.size[2] = b.half_extent.z;
        }
        d.velocity[0] = c.velocity.x, d.velocity[1] = c.velocity.y, d.velocity[2] = c.velocity.z;
        d.density = c.density;
        d.temperature = c.temperature;
        colliders.push_back(d);
    }
    out.colliders = colliders.empty() ? nullptr : colliders.data();
    out.collider_count = static_cast<std::uint32_t>(colliders.size());
    out.vorticity_confinement_enabled = cfg.vorticity_confinement_enabled ? 1u : 0u;
    out.projection_solver = static_cast<std::uint32_t>(cfg.projection_solver);
    out.cg_max_iterations = cfg.cg_max_iterations;
    out.cg_tolerance = cfg.cg_tolerance;
    out.warm_start = static_cast<std::uint32_t>(cfg.warm_start);
    out.cg_variant = static_cast<std::uint32_t>(cfg.cg_variant);
    return out;
}

// xayah::smoke::Model (model.ixx:51-96).
struct Model {
    const Configuration configuration;

    explicit Model(Configuration c) : configuration(std::move(c)) {}

    [[nodiscard]] ExecutionContext allocate_context(ExecutionMode mode, void* stream = nullptr) const {
        std::vector<ds_collider> colliders;
        const ds_config cfg = to_abi(configuration, colliders);
        ds_context* handle = nullptr;
        hip::check(ds_context_create(&cfg, static_cast<std::uint32_t>(mode), stream, &handle), "ds_context_create");
        ExecutionContext ctx;
        ctx.resource = std::make_shared<hip::Resource>(handle);
        hip::check(ds_context_domain_info(handle, &ctx.domain), "ds_context_domain_info");
        return ctx;
    }

    [[nodiscard]] State allocate_state(ExecutionContext& c) const {
        State s{scalar(c), scalar(c), staggered(c)};
        zero(c, s);
        return s;
    }
    [[nodiscard]] Control allocate_control(ExecutionContext& c) const {
        Control v{scalar(c), scalar(c), {scalar(c), scalar(c), scalar(c)}};
        for (auto* f : {&v.density_source, &v.temperature_source, &v.external_acceleration.x, &v.external_acceleration.y,
                        &v.external_acceleration.z})
            c.resource->zero(f->values.data, f->values.size * sizeof(float));
        return v;
    }
    [[nodiscard]] Parameters allocate_parameters(ExecutionContext& c) const {
        Parameters p{hip::Buffer<float>(c.resource, 1), hip::Buffer<float>(c.resource, 1), hip::Buffer<float>(c.resource, 1),
                     hip::Buffer<float>(c.resource, 1)};
        for (auto* b : {&p.ambient_temperature, &p.density_buoyancy, &p.temperature_buoyancy, &p.vorticity_confinement})
            c.resource->zero(b->data, sizeof(float));
        return p;
    }
    // lean = true leaves out the buffers the reverse step does not read (force, advected velocity)
    [[nodiscard]] StepCache allocate_step_cache(ExecutionContext& c, bool lean = false) const {
        StepCache s{scalar(c), scalar(c), {}, staggered(c), {}};
        if (!lean) {
            s.force = {scalar(c), scalar(c), scalar(c)};
            s.advected_velocity = staggered(c);
        }
        return s;
    }
    // hooks of the checkpointed driver (xayah/solver.hpp)
    [[nodiscard]] hip::Buffer<float> allocate_pressure_tape(ExecutionContext& c) const { return hip::Buffer<float>(c.resource, 2 * cells(c)); }
    static void set_pressure_out(StepCache& cache, float* p) { cache.pressure = p; }
    static void set_pressure_guess(StepCache& cache, const float* p) { cache.pressure_guess = p; }
    void zero_state_adjoint(StateAdjoint& a, ExecutionContext& c) const { zero(c, a); }
    void zero_state(State& s, ExecutionContext& c) const { zero(c, s); }
    static void after_forward(ExecutionContext&, int) {}
    static void after_vjp(ExecutionContext&) {}
    [[nodiscard]] StateAdjoint allocate_state_adjoint(ExecutionContext& c) const {
        StateAdjoint s{{hip::Buffer<adjoint_t>(c.resource, cells(c))},
                       {hip::Buffer<adjoint_t>(c.resource, cells(c))},
                       {hip::Buffer<adjoint_t>(c.resource, faces(c, 0)), hip::Buffer<adjoint_t>(c.resource, faces(c, 1)),
                        hip::Buffer<adjoint_t>(c.resource, faces(c, 2))}};
        zero(c, s);
        return s;
    }
    [[nodiscard]] ControlAdjoint allocate_control_adjoint(ExecutionContext& c) const {
        ControlAdjoint v{{hip::Buffer<adjoint_t>(c.resource, cells(c))},
                         {hip::Buffer<adjoint_t>(c.resource, cells(c))},
                         {{hip::Buffer<adjoint_t>(c.resource, cells(c))},
                          {hip::Buffer<adjoint_t>(c.resource, cells(c))},
                          {hip::Buffer<adjoint_t>(c.resource, cells(c))}}};
        for (auto* f : {&v.density_source, &v.temperature_source, &v.external_acceleration.x, &v.external_acceleration.y,
                        &v.external_acceleration.z})
            c.resource->zero(f->values.data, f->values.size * sizeof(adjoint_t));
        return v;
    }
    [[nodiscard]] ParameterAdjoint allocate_parameter_adjoint(ExecutionContext& c) const {
        ParameterAdjoint p{hip::Buffer<double>(c.resource, 1), hip::Buffer<double>(c.resource, 1), hip::Buffer<double>(c.resource, 1),
                           hip::Buffer<double>(c.resource, 1)};
        for (auto* b : {&p.ambient_temperature, &p.density_buoyancy, &p.temperature_buoyancy, &p.vorticity_confinement})
            c.resource->zero(b->data, sizeof(double));
        return p;
    }

    void copy_state(const State& src, State& dst, ExecutionContext& c) const {
        const ds_state s = view(src);
        ds_state d = view(dst);
        hip::check(ds_copy_state(c.resource->handle, &s, &d), "ds_copy_state");
    }
    void copy_state_adjoint(const StateAdjoint& src, StateAdjoint& dst, ExecutionContext& c) const {
        const ds_state_adjoint s = view(src);
        ds_state_adjoint d = view(dst);
        hip::check(ds_copy_state_adjoint(c.resource->handle, &s, &d), "ds_copy_state_adjoint");
    }
    void accumulate_state_adjoint(const StateAdjoint& src, StateAdjoint& dst, ExecutionContext& c) const {
        const ds_state_adjoint s = view(src);
        ds_state_adjoint d = view(dst);
        hip::check(ds_accumulate_state_adjoint(c.resource->handle, &s, &d), "ds_accumulate_state_adjoint");
    }

    // Model::forward_step (model.cpp:341-349).
    void forward_step(const State& state, const Control& control, const Parameters& parameters, State& next_state, StepCache& cache,
                      ExecutionContext& c) const {
        const ds_state s = view(state);
        const ds_control k = view(control);
        const ds_parameters p = view(parameters);
        ds_state n = view(next_state);
        ds_step_cache ch = view(cache);
        hip::check(ds_forward_step(c.resource->handle, &s, &k, &p, &n, &ch, &c.last_step), "ds_forward_step");
    }

    // forward mode (model.ixx:74-76,82)
    [[nodiscard]] StateTangent allocate_state_tangent(ExecutionContext& c) const { return allocate_state(c); }
    [[nodiscard]] ControlTangent allocate_control_tangent(ExecutionContext& c) const { return allocate_control(c); }
    [[nodiscard]] ParameterTangent allocate_parameter_tangent(ExecutionContext& c) const { return allocate_parameters(c); }
    void copy_state_tangent(const StateTangent& src, StateTangent& dst, ExecutionContext& c) const { copy_state(src, dst, c); }

    // Model::jvp_step (model.cpp:351-359).
    void jvp_step(const State& state, const Control& control, const Parameters& parameters, const State& next_state,
                  const StepCache& cache, const StateTangent& state_tangent, const ControlTangent& control_tangent,
                  const ParameterTangent& parameter_tangent, StateTangent& next_state_tangent, ExecutionContext& c) const {
        const ds_state s = view(state);
        const ds_control k = view(control);
        const ds_parameters p = view(parameters);
        const ds_state n = view(next_state);
        const ds_step_cache ch = view(cache);
        const ds_state_tangent st = view(state_tangent);
        const ds_control_tangent kt = view(control_tangent);
        const ds_parameter_tangent pt = view(parameter_tangent);
        ds_state_tangent nt = view(next_state_tangent);
        hip::check(ds_jvp_step(c.resource->handle, &s, &k, &p, &n, &ch, &st, &kt, &pt, &nt, &c.last_step), "ds_jvp_step");
    }

    // Model::vjp_step (model.cpp:361-389).
    void vjp_step(const State& state, const Control& control, const Parameters& parameters, const State& next_state,
                  const StepCache& cache, const StateAdjoint& next_adj, StateAdjoint& prev_adj, ControlAdjoint& control_adj,
                  ParameterAdjoint& param_adj, ExecutionContext& c) const {
        const ds_state s = view(state);
        const ds_control k = view(control);
        const ds_parameters p = view(parameters);
        const ds_state n = view(next_state);
        const ds_step_cache ch = view(cache);
        const ds_state_adjoint na = view(next_adj);
        ds_state_adjoint pa = view(prev_adj);
        ds_control_adjoint ca = view(control_adj);
        const ds_parameter_adjoint qa = view(param_adj);
        hip::check(ds_vjp_step(c.resource->handle, &s, &k, &p, &n, &ch, &na, &pa, &ca, &qa, &c.last_step), "ds_vjp_step");
    }

private:
    static std::size_t cells(const ExecutionContext& c) { return static_cast<std::size_t>(c.domain.cell_count); }
    static std::size_t faces(const ExecutionContext& c, int a) { return static_cast<std::size_t>(c.domain.face_count[a]); }
    ScalarField scalar(ExecutionContext& c) const { return {hip::Buffer<float>(c.resource, cells(c))}; }
    StaggeredVectorField staggered(ExecutionContext& c) const {
        return {hip::Buffer<float>(c.resource, faces(c, 0)), hip::Buffer<float>(c.resource, faces(c, 1)),
                hip::Buffer<float>(c.resource, faces(c, 2))};
    }
    template <typename S>
    static void zero(ExecutionContext& c, S& s) {
        c.resource->zero(s.density.values.data, s.density.values.size * sizeof(*s.density.values.data));
        c.resource->zero(s.temperature.values.data, s.temperature.values.size * sizeof(*s.temperature.values.data));
        c.resource->zero(s.velocity.x.data, s.velocity.x.size * sizeof(*s.velocity.x.data));
        c.resource->zero(s.velocity.y.data, s.velocity.y.size * sizeof(*s.velocity.y.data));
        c.resource->zero(s.velocity.z.data, s.velocity.z.size * sizeof(*s.velocity.z.data));
    }
};

}  // namespace smoke
}  // namespace xayah
