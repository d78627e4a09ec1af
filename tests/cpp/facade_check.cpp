// facade_check — drives the smoke solver through the C++ facade exactly as code written against
// the reference's xayah::smoke::Model and xayah::solver would (model.ixx:51-96,
// solver.ixx:53-150): a closed-box plume trajectory with solver::simulate (store_all), its
// gradient with solver::vjp, the same gradient with the checkpointed pair, and the tangent of the
// final state with solver::jvp over the stored trajectory and with the one-sweep jvp_streaming.
//
// usage: facade_check <dir> <n> <steps> <k>
//   reads  <dir>/u0_{x,y,z}.f32 (initial face velocity), source.f32 (density and temperature
//          source), target.f32 (target density), v_{x,y,z}.f32 (initial velocity tangent)
//   writes <dir>/final_density.f32, grad_store_{x,y,z}.f32, grad_ckpt_{x,y,z}.f32, loss.txt,
//          tangent_store.f32, tangent_stream.f32 (the final density tangent of each driver)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <span>
#include <string>
#include <vector>

#include "xayah/smoke.hpp"
#include "xayah/solver.hpp"

using namespace xayah;

namespace {

std::vector<float> read_f32(const std::string& path, std::size_t count) {
    std::vector<float> v(count);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(count * sizeof(float))))
        throw std::runtime_error("cannot read " + path);
    return v;
}

void write_f32(const std::string& path, const std::vector<float>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(float)));
}

template <typename T>
std::vector<float> download(smoke::ExecutionContext& c, const hip::Buffer<T>& b) {
    std::vector<float> v(b.size);
    c.resource->copy_to_host(v.data(), b.data, b.size * sizeof(float));
    c.resource->synchronize();
    return v;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <dir> <n> <steps> <k>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const std::uint32_t n = static_cast<std::uint32_t>(std::atoi(argv[2]));
    const std::uint32_t steps = static_cast<std::uint32_t>(std::atoi(argv[3]));
    const std::uint32_t k = static_cast<std::uint32_t>(std::atoi(argv[4]));
    try {
        smoke::Configuration cfg;
        cfg.resolution[0] = cfg.resolution[1] = cfg.resolution[2] = n;
        cfg.cell_size = 1.0F / static_cast<float>(n);
        cfg.time_step = cfg.cell_size * 0.5F;
        for (smoke::VelocityBoundaryFace* f : {&cfg.velocity_boundary.x_min, &cfg.velocity_boundary.x_max, &cfg.velocity_boundary.y_min,
                                               &cfg.velocity_boundary.y_max, &cfg.velocity_boundary.z_min, &cfg.velocity_boundary.z_max})
            f->mode = smoke::VelocityBoundaryMode::fixed_value;   // closed no-slip box
        cfg.cg_tolerance = 1e-6;
        cfg.vorticity_confinement_enabled = false;   // the benchmark plume runs without confinement
        const smoke::Model model{cfg};
        smoke::ExecutionContext ctx = model.allocate_context(smoke::ExecutionMode::differentiable);
        const std::size_t cells = ctx.domain.cell_count;

        smoke::State s0 = model.allocate_state(ctx);
        hip::Buffer<float>* u0[3] = {&s0.velocity.x, &s0.velocity.y, &s0.velocity.z};
        const char* axis = "xyz";
        for (int a = 0; a < 3; ++a) {
            const std::vector<float> v = read_f32(dir + "/u0_" + axis[a] + ".f32", u0[a]->size);
            ctx.resource->copy_from_host(u0[a]->data, v.data(), v.size() * sizeof(float));
        }
        const std::vector<float> source = read_f32(dir + "/source.f32", cells);
        const std::vector<float> target = read_f32(dir + "/target.f32", cells);
        std::vector<smoke::Control> controls;
        for (std::uint32_t s = 0; s < steps; ++s) {
            controls.push_back(model.allocate_control(ctx));
            ctx.resource->copy_from_host(controls.back().density_source.values.data, source.data(), cells * sizeof(float));
            ctx.resource->copy_from_host(controls.back().temperature_source.values.data, source.data(), cells * sizeof(float));
        }
        smoke::Parameters par = model.allocate_parameters(ctx);
        const float alpha_rho = -0.1F, alpha_t = 1.0F;
        ctx.resource->copy_from_host(par.density_buoyancy.data, &alpha_rho, sizeof(float));
        ctx.resource->copy_from_host(par.temperature_buoyancy.data, &alpha_t, sizeof(float));

        // store_all trajectory + solver::vjp with a loss adjoint on the final state only
        auto traj = solver::simulate(model, ctx, s0, std::span<const smoke::Control>(controls), par, solver::TapeMode::store_all);
        std::vector<float> final_density = download(ctx, traj.states.back().density.values);
        write_f32(dir + "/final_density.f32", final_density);
        std::vector<float> residual(cells);
        double loss = 0.0;
        for (std::size_t i = 0; i < cells; ++i) {
            residual[i] = final_density[i] - target[i];
            loss += 0.5 * static_cast<double>(residual[i]) * residual[i];
        }
        std::vector<smoke::StateAdjoint> dl;
        for (std::size_t i = 0; i < traj.states.size(); ++i) dl.push_back(model.allocate_state_adjoint(ctx));
        ctx.resource->copy_from_host(dl.back().density.values.data, residual.data(), cells * sizeof(float));
        auto g = solver::vjp(model, ctx, traj, std::span<const smoke::Control>(controls), par, std::span<const smoke::StateAdjoint>(dl));
        const hip::Buffer<float>* gs[3] = {&g.initial_state.velocity.x, &g.initial_state.velocity.y, &g.initial_state.velocity.z};
        for (int a = 0; a < 3; ++a) write_f32(dir + "/grad_store_" + axis[a] + ".f32", download(ctx, *gs[a]));

        // checkpointed pair, pressure taped
        std::function<const smoke::Control&(std::uint32_t)> ctl = [&](std::uint32_t s) -> const smoke::Control& { return controls[s]; };
        auto t = solver::simulate_checkpointed<smoke::Model, smoke::ExecutionContext, smoke::State, smoke::Control, smoke::Parameters>(
            model, ctx, s0, ctl, par, steps, k, true);
        smoke::StateAdjoint final_adj = model.allocate_state_adjoint(ctx), init_adj = model.allocate_state_adjoint(ctx);
        ctx.resource->copy_from_host(final_adj.density.values.data, residual.data(), cells * sizeof(float));
        smoke::ControlAdjoint ctl_adj = model.allocate_control_adjoint(ctx);
        smoke::ParameterAdjoint par_adj = model.allocate_parameter_adjoint(ctx);
        std::function<smoke::ControlAdjoint&(std::uint32_t)> ctl_adj_of = [&](std::uint32_t) -> smoke::ControlAdjoint& { return ctl_adj; };
        solver::vjp_checkpointed(model, ctx, t, ctl, par, final_adj, ctl_adj_of, par_adj, init_adj);
        const hip::Buffer<float>* gc[3] = {&init_adj.velocity.x, &init_adj.velocity.y, &init_adj.velocity.z};
        for (int a = 0; a < 3; ++a) write_f32(dir + "/grad_ckpt_" + axis[a] + ".f32", download(ctx, *gc[a]));
        // forward mode: solver::jvp over the stored trajectory, and the one-sweep driver
        smoke::StateTangent v0 = model.allocate_state_tangent(ctx);
        hip::Buffer<float>* vt[3] = {&v0.velocity.x, &v0.velocity.y, &v0.velocity.z};
        for (int a = 0; a < 3; ++a) {
            const std::vector<float> v = read_f32(dir + "/v_" + axis[a] + ".f32", vt[a]->size);
            ctx.resource->copy_from_host(vt[a]->data, v.data(), v.size() * sizeof(float));
        }
        std::vector<smoke::ControlTangent> ctl_t;
        for (std::uint32_t s = 0; s < steps; ++s) ctl_t.push_back(model.allocate_control_tangent(ctx));
        const smoke::ParameterTangent par_t = model.allocate_parameter_tangent(ctx);
        auto tangents = solver::jvp(model, ctx, traj, std::span<const smoke::Control>(controls), par, v0,
                                    std::span<const smoke::ControlTangent>(ctl_t), par_t);
        write_f32(dir + "/tangent_store.f32", download(ctx, tangents.back().density.values));
        std::function<const smoke::ControlTangent&(std::uint32_t)> ctl_t_of = [&](std::uint32_t s) -> const smoke::ControlTangent& {
            return ctl_t[s];
        };
        auto streamed = solver::jvp_streaming(model, ctx, s0, ctl, par, steps, v0, ctl_t_of, par_t);
        write_f32(dir + "/tangent_stream.f32", download(ctx, streamed.second.density.values));
        std::ofstream(dir + "/loss.txt") << std::to_string(loss) << "\n";
        std::printf("facade_check: n=%u steps=%u k=%u loss=%.9g\n", n, steps, k, loss);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_check failed: %s\n", e.what());
        return 1;
    }
    return 0;
}
