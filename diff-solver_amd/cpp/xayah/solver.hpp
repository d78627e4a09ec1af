// xayah/solver.hpp — generic trajectory driver, shaped like the reference's xayah.solver module
// (solvers/solver.ixx:5-150): TapeMode, Trajectory, GradientResult, simulate, jvp and vjp over a
// model satisfying ForwardModel / JvpModel / VjpModel, plus a checkpoint-every-k reverse sweep
// and a one-sweep forward mode that the reference does not have (its recompute mode keeps every
// state, solver.ixx:61-79,130-141; its jvp needs the stored trajectory, :83-113).
//
// Checkpointed sweep: the forward pass keeps states 0, k, 2k, ... and (optionally) every
// step's converged pressure pair. The reverse pass recomputes one segment of k steps from its
// checkpoint, storing its states and step caches, then runs vjp_step backwards over it. Peak
// memory is (steps/k + k) states + k caches (+ steps pressure pairs with the pressure tape).
// With the tape, the recompute's pressure solves start from the converged forward pressure of
// the same step and stop at their first residual check.
#pragma once

#include <algorithm>
#include <concepts>
#include <cstdint>
#include <functional>
#include <optional>
#include <span>
#include <utility>
#include <vector>

namespace xayah::solver {

enum class TapeMode { store_all, recompute_step_cache };

template <typename State, typename StepCache>
struct Trajectory {
    TapeMode tape_mode;
    std::vector<State> states;
    std::vector<StepCache> step_caches;
};

template <typename StateAdjoint, typename ControlAdjoint, typename ParameterAdjoint>
struct GradientResult {
    StateAdjoint initial_state;
    std::vector<ControlAdjoint> controls;
    ParameterAdjoint parameters;
};

template <typename Model, typename Context, typename State, typename Control, typename Parameters>
concept ForwardModel = requires(const Model& m, Context& c, const State& s, const Control& k, const Parameters& p,
                                decltype(std::declval<const Model&>().allocate_state(c))& n,
                                decltype(std::declval<const Model&>().allocate_step_cache(c))& cache) {
    { m.allocate_state(c) } -> std::same_as<State>;
    m.allocate_step_cache(c);
    { m.copy_state(s, n, c) } -> std::same_as<void>;
    { m.forward_step(s, k, p, n, cache, c) } -> std::same_as<void>;
};

template <typename Model, typename Context, typename State, typename Control, typename Parameters>
concept VjpModel = ForwardModel<Model, Context, State, Control, Parameters> &&
                   requires(const Model& m, Context& c, const State& s, const Control& k, const Parameters& p, const State& n,
                            const decltype(std::declval<const Model&>().allocate_step_cache(c))& cache,
                            const decltype(std::declval<const Model&>().allocate_state_adjoint(c))& na,
                            decltype(std::declval<const Model&>().allocate_state_adjoint(c))& pa,
                            decltype(std::declval<const Model&>().allocate_control_adjoint(c))& ca,
                            decltype(std::declval<const Model&>().allocate_parameter_adjoint(c))& qa) {
                       m.allocate_state_adjoint(c);
                       m.allocate_control_adjoint(c);
                       m.allocate_parameter_adjoint(c);
                       { m.copy_state_adjoint(na, pa, c) } -> std::same_as<void>;
                       { m.accumulate_state_adjoint(na, pa, c) } -> std::same_as<void>;
                       { m.vjp_step(s, k, p, n, cache, na, pa, ca, qa, c) } -> std::same_as<void>;
                   };

template <typename Model, typename Context, typename State, typename Control, typename Parameters>
concept JvpModel = ForwardModel<Model, Context, State, Control, Parameters> &&
                   requires(const Model& m, Context& c, const State& s, const Control& k, const Parameters& p, const State& n,
                            const decltype(std::declval<const Model&>().allocate_step_cache(c))& cache,
                            const decltype(std::declval<const Model&>().allocate_state_tangent(c))& st,
                            decltype(std::declval<const Model&>().allocate_state_tangent(c))& nt,
                            const decltype(std::declval<const Model&>().allocate_control_tangent(c))& kt,
                            const decltype(std::declval<const Model&>().allocate_parameter_tangent(c))& pt) {
                       m.allocate_state_tangent(c);
                       m.allocate_control_tangent(c);
                       m.allocate_parameter_tangent(c);
                       { m.copy_state_tangent(st, nt, c) } -> std::same_as<void>;
                       { m.jvp_step(s, k, p, n, cache, st, kt, pt, nt, c) } -> std::same_as<void>;
                   };

// solver::simulate, solver.ixx:53-81.
template <typename Model, typename Context, typename State, typename Control, typename Parameters>
    requires ForwardModel<Model, Context, State, Control, Parameters>
[[nodiscard]] auto simulate(const Model& model, Context& context, const State& initial_state, std::span<const Control> controls,
                            const Parameters& parameters, TapeMode tape_mode)
    -> Trajectory<State, decltype(model.allocate_step_cache(context))> {
    using Cache = decltype(model.allocate_step_cache(context));
    Trajectory<State, Cache> t{tape_mode, {}, {}};
    t.states.reserve(controls.size() + 1);
    State first = model.allocate_state(context);
    model.copy_state(initial_state, first, context);
    t.states.push_back(std::move(first));
    std::optional<Cache> reusable;
    if (tape_mode == TapeMode::recompute_step_cache && !controls.empty()) reusable.emplace(model.allocate_step_cache(context));
    for (const Control& control : controls) {
        State next = model.allocate_state(context);
        if (tape_mode == TapeMode::store_all) {
            Cache cache = model.allocate_step_cache(context);
            model.forward_step(t.states.back(), control, parameters, next, cache, context);
            t.step_caches.push_back(std::move(cache));
        } else {
            model.forward_step(t.states.back(), control, parameters, next, *reusable, context);
        }
        t.states.push_back(std::move(next));
    }
    return t;
}

// solver::vjp, solver.ixx:115-150.
template <typename Model, typename Context, typename State, typename StepCache, typename Control, typename Parameters, typename StateAdjoint>
    requires VjpModel<Model, Context, State, Control, Parameters>
[[nodiscard]] auto vjp(const Model& model, Context& context, const Trajectory<State, StepCache>& trajectory, std::span<const Control> controls,
                       const Parameters& parameters, std::span<const StateAdjoint> trajectory_adjoint)
    -> GradientResult<StateAdjoint, decltype(model.allocate_control_adjoint(context)), decltype(model.allocate_parameter_adjoint(context))> {
    GradientResult<StateAdjoint, decltype(model.allocate_control_adjoint(context)), decltype(model.allocate_parameter_adjoint(context))> r{
        model.allocate_state_adjoint(context), {}, model.allocate_parameter_adjoint(context)};
    r.controls.reserve(controls.size());
    for (std::size_t s = 0; s < controls.size(); ++s) r.controls.push_back(model.allocate_control_adjoint(context));
    StateAdjoint running = model.allocate_state_adjoint(context);
    model.copy_state_adjoint(trajectory_adjoint.back(), running, context);
    std::optional<State> recomputed_state;
    std::optional<StepCache> recomputed_cache;
    if (trajectory.tape_mode == TapeMode::recompute_step_cache && !controls.empty()) {
        recomputed_state.emplace(model.allocate_state(context));
        recomputed_cache.emplace(model.allocate_step_cache(context));
    }
    for (std::size_t step = controls.size(); step-- > 0;) {
        const StepCache* cache;
        if (trajectory.tape_mode == TapeMode::store_all) cache = &trajectory.step_caches[step];
        else {
            model.forward_step(trajectory.states[step], controls[step], parameters, *recomputed_state, *recomputed_cache, context);
            cache = &*recomputed_cache;
        }
        StateAdjoint previous = model.allocate_state_adjoint(context);
        model.vjp_step(trajectory.states[step], controls[step], parameters, trajectory.states[step + 1], *cache, running, previous,
                       r.controls[step], r.parameters, context);
        model.accumulate_state_adjoint(trajectory_adjoint[step], previous, context);
        running = std::move(previous);
    }
    r.initial_state = std::move(running);
    return r;
}

// solver::jvp, solver.ixx:83-113: the state tangent after every step of a stored trajectory
// (element 0 is a copy of the initial tangent).
template <typename Model, typename Context, typename State, typename StepCache, typename Control, typename Parameters, typename StateTangent,
          typename ControlTangent, typename ParameterTangent>
    requires JvpModel<Model, Context, State, Control, Parameters>
[[nodiscard]] auto jvp(const Model& model, Context& context, const Trajectory<State, StepCache>& trajectory, std::span<const Control> controls,
                       const Parameters& parameters, const StateTangent& initial_state_tangent,
                       std::span<const ControlTangent> control_tangents, const ParameterTangent& parameter_tangent) -> std::vector<StateTangent> {
    std::vector<StateTangent> result;
    result.reserve(trajectory.states.size());
    StateTangent first = model.allocate_state_tangent(context);
    model.copy_state_tangent(initial_state_tangent, first, context);
    result.push_back(std::move(first));
    std::optional<State> recomputed_state;
    std::optional<StepCache> recomputed_cache;
    if (trajectory.tape_mode == TapeMode::recompute_step_cache && !controls.empty()) {
        recomputed_state.emplace(model.allocate_state(context));
        recomputed_cache.emplace(model.allocate_step_cache(context));
    }
    for (std::size_t step = 0; step < controls.size(); ++step) {
        const StepCache* cache;
        if (trajectory.tape_mode == TapeMode::store_all) cache = &trajectory.step_caches[step];
        else {
            model.forward_step(trajectory.states[step], controls[step], parameters, *recomputed_state, *recomputed_cache, context);
            cache = &*recomputed_cache;
        }
        StateTangent next = model.allocate_state_tangent(context);
        model.jvp_step(trajectory.states[step], controls[step], parameters, trajectory.states[step + 1], *cache, result.back(),
                       control_tangents[step], parameter_tangent, next, context);
        result.push_back(std::move(next));
    }
    return result;
}

// Forward mode in one sweep (new): each step runs forward_step and then jvp_step on that step's
// cache, so nothing is stored (two states, two tangents, one cache). Returns (state_T, its
// tangent). control(step) / control_tangent(step) return a step's control and its tangent.
// Optional statistics hooks: after_forward(c, 0) and after_jvp(c).
template <typename Model, typename Context, typename State, typename Control, typename Parameters, typename StateTangent,
          typename ControlTangent, typename ParameterTangent>
    requires JvpModel<Model, Context, State, Control, Parameters>
auto jvp_streaming(const Model& model, Context& context, const State& initial_state,
                   const std::function<const Control&(std::uint32_t)>& control, const Parameters& parameters, std::uint32_t steps,
                   const StateTangent& initial_tangent, const std::function<const ControlTangent&(std::uint32_t)>& control_tangent,
                   const ParameterTangent& parameter_tangent) -> std::pair<State, StateTangent> {
    auto cache = model.allocate_step_cache(context);
    State a = model.allocate_state(context), b = model.allocate_state(context);
    StateTangent at = model.allocate_state_tangent(context), bt = model.allocate_state_tangent(context);
    model.copy_state(initial_state, a, context);
    model.copy_state_tangent(initial_tangent, at, context);
    for (std::uint32_t s = 0; s < steps; ++s) {
        model.forward_step(a, control(s), parameters, b, cache, context);
        if constexpr (requires { model.after_forward(context, 0); }) model.after_forward(context, 0);
        model.jvp_step(a, control(s), parameters, b, cache, at, control_tangent(s), parameter_tangent, bt, context);
        if constexpr (requires { model.after_jvp(context); }) model.after_jvp(context);
        std::swap(a, b);
        std::swap(at, bt);
    }
    return {std::move(a), std::move(at)};
}

// ---- checkpoint-every-k (new) ---------------------------------------------------------------
// Models used by the checkpointed driver also provide:
//   allocate_pressure_tape(c)            -> buffer for one step's pressure pair (has .data)
//   set_pressure_out(cache, float*)       where forward_step stores the converged pressure
//   set_pressure_guess(cache, const float*) the initial guess of forward_step's pressure solve
//   zero_state_adjoint(a, c)              clears an adjoint that vjp_step accumulates into
//   after_forward(c, kind) / after_vjp(c) statistics hooks (kind: 0 forward, 1 recompute)
template <typename Model, typename Context, typename State, typename Control, typename Parameters>
struct CheckpointedTrajectory {
    std::uint32_t interval = 1;
    std::uint32_t steps = 0;
    std::vector<State> checkpoints;   // states 0, k, 2k, ...
    std::optional<State> final_state;
    std::vector<decltype(std::declval<const Model&>().allocate_pressure_tape(std::declval<Context&>()))> pressure_tape;
};

// Forward pass keeping a checkpoint every `interval` steps. control(step) returns the control of
// a step (the same object for a constant control).
template <typename Model, typename Context, typename State, typename Control, typename Parameters>
auto simulate_checkpointed(const Model& model, Context& context, const State& initial_state,
                           const std::function<const Control&(std::uint32_t)>& control, const Parameters& parameters,
                           std::uint32_t steps, std::uint32_t interval, bool tape_pressure)
    -> CheckpointedTrajectory<Model, Context, State, Control, Parameters> {
    CheckpointedTrajectory<Model, Context, State, Control, Parameters> t;
    t.interval = std::max<std::uint32_t>(1u, interval);
    t.steps = steps;
    auto cache = model.allocate_step_cache(context);
    State a = model.allocate_state(context), b = model.allocate_state(context);
    model.copy_state(initial_state, a, context);
    for (std::uint32_t s = 0; s < steps; ++s) {
        if (s % t.interval == 0) {
            State cp = model.allocate_state(context);
            model.copy_state(a, cp, context);
            t.checkpoints.push_back(std::move(cp));
        }
        if (tape_pressure) {
            t.pressure_tape.push_back(model.allocate_pressure_tape(context));
            model.set_pressure_out(cache, t.pressure_tape.back().data);
        }
        model.set_pressure_guess(cache, nullptr);
        model.forward_step(a, control(s), parameters, b, cache, context);
        model.after_forward(context, 0);
        std::swap(a, b);
    }
    t.final_state.emplace(std::move(a));
    return t;
}

// Reverse sweep over a checkpointed trajectory. final_adjoint is dL/d(final state);
// control_adjoint(step) returns where step's control adjoint accumulates (the same object to
// sum over steps); the result accumulates into initial_adjoint and parameter_adjoint.
template <typename Model, typename Context, typename State, typename Control, typename Parameters, typename StateAdjoint,
          typename ControlAdjoint, typename ParameterAdjoint>
void vjp_checkpointed(const Model& model, Context& context, const CheckpointedTrajectory<Model, Context, State, Control, Parameters>& t,
                      const std::function<const Control&(std::uint32_t)>& control, const Parameters& parameters,
                      const StateAdjoint& final_adjoint, const std::function<ControlAdjoint&(std::uint32_t)>& control_adjoint,
                      ParameterAdjoint& parameter_adjoint, StateAdjoint& initial_adjoint) {
    const std::uint32_t k = t.interval;
    const std::uint32_t seg_len = std::min(k, t.steps);
    std::vector<State> states;   // states[i] = state s0 + i of the current segment, i in [0, seg_len]
    std::vector<decltype(model.allocate_step_cache(context))> caches;
    for (std::uint32_t i = 0; i <= seg_len; ++i) states.push_back(model.allocate_state(context));
    for (std::uint32_t i = 0; i < seg_len; ++i) caches.push_back(model.allocate_step_cache(context));
    StateAdjoint running = model.allocate_state_adjoint(context), previous = model.allocate_state_adjoint(context);
    model.copy_state_adjoint(final_adjoint, running, context);
    const std::uint32_t segments = (t.steps + k - 1) / k;
    for (std::uint32_t seg = segments; seg-- > 0;) {
        const std::uint32_t s0 = seg * k, s1 = std::min(t.steps, s0 + k);
        model.copy_state(t.checkpoints[seg], states[0], context);
        for (std::uint32_t s = s0; s < s1; ++s) {   // recompute the segment's states and caches
            auto& cache = caches[s - s0];
            model.set_pressure_out(cache, nullptr);
            model.set_pressure_guess(cache, t.pressure_tape.empty() ? nullptr : t.pressure_tape[s].data);
            model.forward_step(states[s - s0], control(s), parameters, states[s + 1 - s0], cache, context);
            model.after_forward(context, 1);
        }
        for (std::uint32_t r = s1; r-- > s0;) {   // reverse over the segment
            model.zero_state_adjoint(previous, context);
            model.vjp_step(states[r - s0], control(r), parameters, states[r + 1 - s0], caches[r - s0], running, previous,
                           control_adjoint(r), parameter_adjoint, context);
            model.after_vjp(context);
            std::swap(running, previous);
        }
    }
    model.accumulate_state_adjoint(running, initial_adjoint, context);
}

}  // namespace xayah::solver
