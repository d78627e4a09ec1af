// smoke_forward — the reference's forward example program (examples/smoke/forward/main.cpp):
// the 128 x 192 x 128 double-jet thermal smoke for 960 steps, a metrics line every 60 steps.
//
//   smoke_forward [--steps N] [--report R] [--solver rbgs|pcg] [--resolution NX NY NZ]
//                 [--cell-size H] [--dump DIR]
//
// --dump DIR writes every step's metrics to DIR/metrics.csv and the final density (FP32, reference
// layout) to DIR/final_density.f32. Exit status 1 with a message on failure, as main.cpp:24-27.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "xayah/forward_simulation.hpp"

namespace fwd = xayah::smoke::examples::forward;

namespace {

std::size_t free_memory() {
    std::size_t free_bytes = 0, total = 0;
    if (hipMemGetInfo(&free_bytes, &total) != hipSuccess) throw std::runtime_error("hipMemGetInfo failed");
    return free_bytes;
}

void print_metrics(const fwd::ForwardSimulationMetrics& m) {
    std::printf("step %3llu  time %.3f s  density mass %.6f max %.5f  temperature max %.5f  velocity max %.5f  CFL %.4f  "
                "divergence %.4e -> %.4e (%.4f)  step %.3f ms  average %.3f ms\n",
                static_cast<unsigned long long>(m.step), m.physical_time, m.density_mass, m.density_maximum, m.temperature_maximum,
                m.maximum_velocity, m.cfl, m.pre_projection_divergence_rms, m.post_projection_divergence_rms, m.divergence_ratio,
                m.step_milliseconds, m.average_step_milliseconds);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
    try {
        fwd::ForwardSimulationOptions options{};
        std::uint32_t steps = 960u, report = 60u;
        std::string dump;
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            auto next = [&]() -> const char* {
                if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
                return argv[++i];
            };
            if (a == "--steps") steps = static_cast<std::uint32_t>(std::strtoul(next(), nullptr, 10));
            else if (a == "--report") report = static_cast<std::uint32_t>(std::strtoul(next(), nullptr, 10));
            else if (a == "--solver") {
                const std::string s = next();
                if (s == "rbgs") options.projection_solver = xayah::smoke::ProjectionSolver::rbgs;
                else if (s == "pcg") options.projection_solver = xayah::smoke::ProjectionSolver::pcg;
                else throw std::runtime_error("--solver takes rbgs or pcg");
            } else if (a == "--resolution") {
                for (int k = 0; k < 3; ++k) options.resolution[k] = static_cast<std::uint32_t>(std::strtoul(next(), nullptr, 10));
            } else if (a == "--cell-size") options.cell_size = std::strtof(next(), nullptr);
            else if (a == "--dump") dump = next();
            else throw std::runtime_error("unknown option " + a);
        }
        if (report == 0u) report = 1u;
        const std::size_t before = free_memory();
        fwd::ForwardSimulation simulation{options};
        const std::size_t after = free_memory();
        std::printf("%u x %u x %u double-jet thermal smoke (%s)\n", options.resolution[0], options.resolution[1], options.resolution[2],
                    options.projection_solver == xayah::smoke::ProjectionSolver::rbgs ? "red-black Gauss-Seidel" : "PCG");
        std::printf("GPU allocation: %.1f MiB\n", static_cast<double>(before - after) / (1024.0 * 1024.0));
        std::FILE* csv = nullptr;
        if (!dump.empty()) {
            csv = std::fopen((dump + "/metrics.csv").c_str(), "w");
            if (!csv) throw std::runtime_error("cannot write " + dump + "/metrics.csv");
            std::fprintf(csv, "step,physical_time,density_mass,density_maximum,temperature_maximum,maximum_velocity,cfl,"
                              "pre_projection_divergence_rms,post_projection_divergence_rms,divergence_ratio,iterations\n");
        }
        for (std::uint32_t s = 0u; s < steps; ++s) {
            simulation.step();
            const fwd::ForwardSimulationMetrics& m = simulation.metrics;
            if (csv)
                std::fprintf(csv, "%llu,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%u\n", static_cast<unsigned long long>(m.step),
                             m.physical_time, m.density_mass, m.density_maximum, m.temperature_maximum, m.maximum_velocity, m.cfl,
                             m.pre_projection_divergence_rms, m.post_projection_divergence_rms, m.divergence_ratio, m.projection.iterations);
            if (m.step % report == 0u) print_metrics(m);
        }
        if (csv) std::fclose(csv);
        if (!dump.empty()) {
            const auto& d = simulation.current_state.density.values;
            std::vector<float> host(d.size);
            simulation.context.resource->copy_to_host(host.data(), d.data, d.size * sizeof(float));
            simulation.context.resource->synchronize();
            std::FILE* f = std::fopen((dump + "/final_density.f32").c_str(), "wb");
            if (!f || std::fwrite(host.data(), sizeof(float), host.size(), f) != host.size()) throw std::runtime_error("cannot write final_density.f32");
            std::fclose(f);
        }
        return 0;
    } catch (const std::exception& error) {
        std::fprintf(stderr, "smoke_forward failed: %s\n", error.what());
        return 1;
    }
}
