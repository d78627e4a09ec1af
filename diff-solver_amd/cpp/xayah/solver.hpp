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

namespace detail {

// The step caches of a stored trajectory, in any order: the taped cache of a step under
// store_all; under recompute_step_cache the step is replayed from its stored input state into
// one scratch state and one scratch cache, which the returned reference then points at (valid
// until the next call).
template <typename Model, typename Context, typename State, typename StepCache, typename Control, typename Parameters>
class CacheReplay {
public:
    CacheReplay(const Model& model, Context& context, const Trajectory<State, StepCache>& trajectory, std::span<const Control> controls,
                const Parameters& parameters)
        : model_(model), context_(context), trajectory_(trajectory), controls_(controls), parameters_(parameters) {
        if (trajectory.tape_mode == TapeMode::recompute_step_cache && !controls.empty()) {
            scratch_state_.emplace(model.allocate_state(context));
            scratch_cache_.emplace(model.allocate_step_cache(context));
        }
    }
    const StepCache& operator()(std::size_t step) {
        if (!scratch_cache_) return trajectory_.step_caches[step];
        model_.forward_step(trajectory_.states[step], controls_[step], parameters_, *scratch_state_, *scratch_cache_, context_);
        return *scratch_cache_;
    }

private:
    const Model& model_;
    Context& context_;
    const Trajectory<State, StepCache>& trajectory_;
    std::span<const Control> controls_;
    const Parameters& parameters_;
    std::optional<State> scratch_state_;
    std::optional<StepCache> scratch_cache_;
};

template <typename T, typename Make>
std::vector<T> make_n(std::size_t n, Make&& make) {
    std::vector<T> v;
    v.reserve(n);
    for (std::size_t i = 0; i < n; ++i) v.push_back(make());
    return v;
}

}  // namespace detail

// solver::simulate (solver.ixx:53-81): the n + 1 states of the trajectory (element 0 a copy of the
// initial state) and, under store_all, the n step caches. Every buffer is allocated before the
// first step; under recompute_step_cache one cache serves every step and is dropped at the end.
template <typename Model, typename Context, typename State, typename Control, typename Parameters>
    requires ForwardModel<Model, Context, State, Control, Parameters>
[[nodiscard]] auto simulate(const Model& model, Context& context, const State& initial_state, std::span<const Control> controls,
                            const Parameters& parameters, TapeMode tape_mode)
    -> Trajectory<State, decltype(model.allocate_step_cache(context))> {
    using Cache = decltype(model.allocate_step_cache(context));
    const std::size_t n = controls.size();
    const bool taped = tape_mode == TapeMode::store_all;
    std::vector<State> states = detail::make_n<State>(n + 1, [&] { return model.allocate_state(context); });
    std::vector<Cache> caches = detail::make_n<Cache>(taped ? n : std::min<std::size_t>(n, 1),
                                                      [&] { return model.allocate_step_cache(context); });
    model.copy_state(initial_state, states.front(), context);
    for (std::size_t step = 0; step < n; ++step)
        model.forward_step(states[step], controls[step], parameters, states[step + 1], caches[taped ? step : 0], context);
    if (!taped) caches.clear();
    return {tape_mode, std::move(states), std::move(caches)};
}

// solver::vjp (solver.ixx:115-150): trajectory_adjoint holds dL/d(state i) for i = 0..n. The
// reverse sweep carries the adjoint of state step + 1 in one buffer and builds the adjoint of
// state step in a second one: seeded with the loss adjoint of that state, then vjp_step adds the
// step's pullback (vjp_step accumulates into its output). The two buffers swap roles each step.
// Control and parameter adjoints accumulate over the steps as vjp_step adds into them.
template <typename Model, typename Context, typename State, typename StepCache, typename Control, typename Parameters, typename StateAdjoint>
    requires VjpModel<Model, Context, State, Control, Parameters>
[[nodiscard]] auto vjp(const Model& model, Context& context, const Trajectory<State, StepCache>& trajectory, std::span<const Control> controls,
                       const Parameters& parameters, std::span<const StateAdjoint> trajectory_adjoint)
    -> GradientResult<StateAdjoint, decltype(model.allocate_control_adjoint(context)), decltype(model.allocate_parameter_adjoint(context))> {
    using ControlAdjoint = decltype(model.allocate_control_adjoint(context));
    const std::size_t n = controls.size();
    std::vector<ControlAdjoint> control_adjoints =
        detail::make_n<ControlAdjoint>(n, [&] { return model.allocate_control_adjoint(context); });
    auto parameter_adjoint = model.allocate_parameter_adjoint(context);
    StateAdjoint later = model.allocate_state_adjoint(context), earlier = model.allocate_state_adjoint(context);
    model.copy_state_adjoint(trajectory_adjoint[n], later, context);
    detail::CacheReplay replay(model, context, trajectory, controls, parameters);
    for (std::size_t remaining = n; remaining > 0; --remaining) {
        const std::size_t step = remaining - 1;
        const StepCache& cache = replay(step);
        model.copy_state_adjoint(trajectory_adjoint[step], earlier, context);
        model.vjp_step(trajectory.states[step], controls[step], parameters, trajectory.states[step + 1], cache, later, earlier,
                       control_adjoints[step], parameter_adjoint, context);
        std::swap(later, earlier);
    }
    return {std::move(later), std::move(control_adjoints), std::move(parameter_adjoint)};
}

// solver::jvp (solver.ixx:83-113): the state tangent after every step of a stored trajectory
// (element 0 is a copy of the initial tangent); all n + 1 tangents are allocated up front and
// jvp_step writes tangent step + 1 from tangent step.
template <typename Model, typename Context, typename State, typename StepCache, typename Control, typename Parameters, typename StateTangent,
          typename ControlTangent, typename ParameterTangent>
    requires JvpModel<Model, Context, State, Control, Parameters>
[[nodiscard]] auto jvp(const Model& model, Context& context, const Trajectory<State, StepCache>& trajectory, std::span<const Control> controls,
                       const Parameters& parameters, const StateTangent& initial_state_tangent,
                       std::span<const ControlTangent> control_tangents, const ParameterTangent& parameter_tangent) -> std::vector<StateTangent> {
    const std::size_t n = controls.size();
    std::vector<StateTangent> tangents = detail::make_n<StateTangent>(n + 1, [&] { return model.allocate_state_tangent(context); });
    model.copy_state_tangent(initial_state_tangent, tangents.front(), context);
    detail::CacheReplay replay(model, context, trajectory, controls, parameters);
    for (std::size_t step = 0; step < n; ++step)
        model.jvp_step(trajectory.states[step], controls[step], parameters, trajectory.states[step + 1], replay(step), tangents[step],
                       control_tangents[step], parameter_tangent, tangents[step + 1], context);
    return tangents;
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
