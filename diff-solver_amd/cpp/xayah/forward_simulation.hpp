// xayah/forward_simulation.hpp — the reference's forward example (examples/smoke/forward/
// simulation.ixx, simulation.cpp): a pulsed double-jet thermal plume with vorticity confinement
// in a box open at the top, stepped by ForwardSimulation with per-step metrics.
//
// Header-only over the C ABI (ds_example_write_control, ds_example_reduce_metrics) and the
// C++ facade in xayah/smoke.hpp. Differences from the reference, by design:
//   * options add the pressure solver: projection_solver (the reference's red-black
//     Gauss-Seidel with pressure_iterations sweeps by default, or PCG to cg_tolerance) and the
//     PCG warm start (the previous step's pressure by default);
//   * the metric sums are deterministic (fixed-order block reduction instead of FP64 atomics);
//   * metrics.projection holds the last pressure solve's statistics.
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <utility>

#include "xayah/smoke.hpp"

namespace xayah::smoke::examples::forward {

// ForwardSimulationOptions (simulation.ixx:9-27), same fields and defaults.
struct ForwardSimulationOptions {
    std::array<std::uint32_t, 3u> resolution{128u, 192u, 128u};
    float cell_size{1.0F / 128.0F};
    float time_step{1.0F / 120.0F};
    std::uint32_t pressure_iterations{160u};
    float density_source_rate{4.0F};
    float temperature_source_rate{8.0F};
    float ambient_temperature{};
    float density_buoyancy{-0.1F};
    float temperature_buoyancy{1.0F};
    float vorticity_confinement{2.0F};
    Vector3 left_source_center{0.25F, 0.10F, 0.36F};
    Vector3 right_source_center{0.75F, 0.10F, 0.64F};
    float source_radius{0.055F};
    Vector3 left_acceleration{3.5F, 5.0F, 1.8F};
    Vector3 right_acceleration{-3.5F, 5.0F, -1.8F};
    float pulse_period{0.9F};
    bool host_metrics{true};
    // additions: the pressure solver
    ProjectionSolver projection_solver{ProjectionSolver::rbgs};
    double cg_tolerance{1e-6};
    WarmStart warm_start{WarmStart::previous};
};

// ForwardSimulationMetrics (simulation.ixx:29-42).
struct ForwardSimulationMetrics {
    std::uint64_t step{};
    double physical_time{};
    double density_mass{};
    double density_maximum{};
    double temperature_maximum{};
    double maximum_velocity{};
    double cfl{};
    double pre_projection_divergence_rms{};
    double post_projection_divergence_rms{};
    double divergence_ratio{};
    double step_milliseconds{};
    double average_step_milliseconds{};
    ds_solve_info projection{};
};

// make_configuration (simulation.cpp:17-35): normal-fixed / tangent-free walls (value 0) on
// five sides, an open top (zero-gradient velocity, pressure / density / temperature fixed at 0
// on y_max), vorticity confinement on.
inline Configuration make_configuration(const ForwardSimulationOptions& o) {
    Configuration c;
    for (int a = 0; a < 3; ++a) c.resolution[a] = o.resolution[a];
    c.cell_size = o.cell_size;
    c.time_step = o.time_step;
    c.pressure_iterations = o.pressure_iterations;
    constexpr auto wall = VelocityBoundaryMode::normal_fixed_tangent_zero_gradient;
    c.velocity_boundary.x_min.mode = wall;
    c.velocity_boundary.x_max.mode = wall;
    c.velocity_boundary.y_min.mode = wall;
    c.velocity_boundary.y_max.mode = VelocityBoundaryMode::zero_gradient;
    c.velocity_boundary.z_min.mode = wall;
    c.velocity_boundary.z_max.mode = wall;
    c.pressure_boundary.y_max.mode = ScalarBoundaryMode::fixed_value;
    c.density_boundary.y_max.mode = ScalarBoundaryMode::fixed_value;
    c.temperature_boundary.y_max.mode = ScalarBoundaryMode::fixed_value;
    c.vorticity_confinement_enabled = true;
    c.projection_solver = o.projection_solver;
    c.cg_tolerance = o.cg_tolerance;
    c.warm_start = o.projection_solver == ProjectionSolver::pcg ? o.warm_start : WarmStart::none;
    return c;
}

inline ds_jet_control jet_control(const ForwardSimulationOptions& o) {
    ds_jet_control j{};
    j.pulse_period = o.pulse_period;
    const Vector3* v[4]{&o.left_source_center, &o.right_source_center, &o.left_acceleration, &o.right_acceleration};
    float* dst[4]{j.left_source_center, j.right_source_center, j.left_acceleration, j.right_acceleration};
    for (int k = 0; k < 4; ++k) {
        dst[k][0] = v[k]->x;
        dst[k][1] = v[k]->y;
        dst[k][2] = v[k]->z;
    }
    j.source_radius = o.source_radius;
    j.density_source_rate = o.density_source_rate;
    j.temperature_source_rate = o.temperature_source_rate;
    return j;
}

// The derived metrics of one step from the six raw values (simulation.cpp:104-119).
inline void finish_metrics(const ForwardSimulationOptions& o, const double (&v)[6], double step_ms, ForwardSimulationMetrics& m) {
    const double previous_steps = static_cast<double>(m.step - 1u);
    const double h = o.cell_size;
    m.density_mass = v[0] * h * h * h;
    m.density_maximum = v[1];
    m.temperature_maximum = v[2];
    m.maximum_velocity = v[3];
    m.cfl = v[3] * static_cast<double>(o.time_step) / h;
    const double cells = static_cast<double>(o.resolution[0]) * o.resolution[1] * o.resolution[2];
    m.pre_projection_divergence_rms = std::sqrt(v[4] / cells);
    m.post_projection_divergence_rms = std::sqrt(v[5] / cells);
    m.divergence_ratio = m.post_projection_divergence_rms / m.pre_projection_divergence_rms;
    m.step_milliseconds = step_ms;
    m.average_step_milliseconds = (previous_steps * m.average_step_milliseconds + step_ms) / static_cast<double>(m.step);
}

// ForwardSimulation (simulation.ixx:44-68, simulation.cpp:61-120).
struct ForwardSimulation {
    ForwardSimulationOptions options;
    Model model;
    ExecutionContext context;
    State current_state;
    ForwardSimulationMetrics metrics;
    hip::Buffer<double> device_metrics;

    explicit ForwardSimulation(ForwardSimulationOptions next_options = {})
        : options(next_options),
          model(make_configuration(options)),
          context(model.allocate_context(ExecutionMode::forward)),
          current_state(model.allocate_state(context)),
          metrics{},
          device_metrics(context.resource, 6u),
          next_state_(model.allocate_state(context)),
          control_(model.allocate_control(context)),
          parameters_(model.allocate_parameters(context)),
          step_cache_(model.allocate_step_cache(context)),
          jet_(jet_control(options)) {
        auto& r = *context.resource;
        r.copy_from_host(parameters_.ambient_temperature.data, &options.ambient_temperature, sizeof(float));
        r.copy_from_host(parameters_.density_buoyancy.data, &options.density_buoyancy, sizeof(float));
        r.copy_from_host(parameters_.temperature_buoyancy.data, &options.temperature_buoyancy, sizeof(float));
        r.copy_from_host(parameters_.vorticity_confinement.data, &options.vorticity_confinement, sizeof(float));
        r.synchronize();
        reset();
    }

    ForwardSimulation(const ForwardSimulation&) = delete;
    ForwardSimulation(ForwardSimulation&&) = delete;
    ForwardSimulation& operator=(const ForwardSimulation&) = delete;
    ForwardSimulation& operator=(ForwardSimulation&&) = delete;

    // reset (simulation.cpp:74-88): zero states and metrics.
    void reset() {
        model.zero_state(current_state, context);
        model.zero_state(next_state_, context);
        context.resource->zero(device_metrics.data, 6u * sizeof(double));
        context.resource->synchronize();
        metrics = {};
    }

    // step (simulation.cpp:90-120): write the jets' control for this step's time, one forward
    // step, reduce the metrics over the fluid cells (pre-projection divergence from the step
    // cache's constrained advected velocity), swap the states.
    void step() {
        const auto begin = options.host_metrics ? std::chrono::steady_clock::now() : std::chrono::steady_clock::time_point{};
        ds_context* h = context.resource->handle;
        float* accel[3]{control_.external_acceleration.x.values.data, control_.external_acceleration.y.values.data,
                        control_.external_acceleration.z.values.data};
        hip::check(ds_example_write_control(h, metrics.step, &jet_, control_.density_source.values.data,
                                            control_.temperature_source.values.data, accel),
                   "ds_example_write_control");
        model.forward_step(current_state, control_, parameters_, next_state_, step_cache_, context);
        const float* pre[3]{step_cache_.advected_velocity.x.data, step_cache_.advected_velocity.y.data, step_cache_.advected_velocity.z.data};
        const float* post[3]{next_state_.velocity.x.data, next_state_.velocity.y.data, next_state_.velocity.z.data};
        hip::check(ds_example_reduce_metrics(h, next_state_.density.values.data, next_state_.temperature.values.data, pre, post,
                                             device_metrics.data),
                   "ds_example_reduce_metrics");
        double values[6]{};
        if (options.host_metrics) {
            context.resource->copy_to_host(values, device_metrics.data, sizeof(values));
            context.resource->synchronize();
        }
        std::swap(current_state, next_state_);
        ++metrics.step;
        metrics.physical_time = static_cast<double>(metrics.step) * options.time_step;
        metrics.projection = context.last_step.projection;
        if (!options.host_metrics) return;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - begin).count();
        finish_metrics(options, values, ms, metrics);
    }

    // The control written by the last step (for inspection and tests).
    [[nodiscard]] const Control& control() const { return control_; }

private:
    State next_state_;
    Control control_;
    Parameters parameters_;
    StepCache step_cache_;
    ds_jet_control jet_;
};

}  // namespace xayah::smoke::examples::forward
