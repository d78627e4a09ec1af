// sanitize_check.cpp — runs the FP64 oracle (test infrastructure) under AddressSanitizer and
// UndefinedBehaviorSanitizer: a forward step and its VJP on small grids that reach every boundary
// kind (fixed-value, zero-gradient, normal-fixed and periodic faces), both collider shapes,
// vorticity confinement and both projection solvers. Built and run by `make -C oracle sanitize`;
// exits non-zero on a failed call or a non-finite output (the sanitizers abort on their own finds).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "ds/ds_smoke.h"

extern "C" {
const char* oracle_last_error(void);
int oracle_set_threads(int threads);
void* oracle_create(const ds_config* cfg);
void oracle_destroy(void* h);
int oracle_forward_step(void* h, const double* const s[5], const double* const c[5], const double* params, double* const next[5],
                        double* const cache[12], const double* guess, ds_solve_info* info);
int oracle_vjp_step(void* h, const double* const state[5], const double* const next[5], const double* const cache[12],
                    const double* params, const double* const next_adj[5], double* const prev_adj[5], double* const control_adj[5],
                    double* param_adj, ds_solve_info* info);
}

namespace {

uint64_t g_seed = 0x9e3779b97f4a7c15ull;
double uniform() {   // xorshift64*, in [-1, 1)
    g_seed ^= g_seed >> 12;
    g_seed ^= g_seed << 25;
    g_seed ^= g_seed >> 27;
    return (double)((g_seed * 0x2545f4914f6cdd1dull) >> 11) / 4503599627370496.0 - 1.0;
}

struct Fields {   // density, temperature (cells), velocity (faces of x, y, z)
    std::vector<double> a[5];
    Fields(const uint32_t n[3], double scale) {
        const size_t cells = (size_t)n[0] * n[1] * n[2];
        const size_t sizes[5] = {cells, cells, (size_t)(n[0] + 1) * n[1] * n[2], (size_t)n[0] * (n[1] + 1) * n[2],
                                 (size_t)n[0] * n[1] * (n[2] + 1)};
        for (int k = 0; k < 5; ++k) {
            a[k].resize(sizes[k]);
            for (double& v : a[k]) v = scale * uniform();
        }
    }
    void ptrs(double* p[5]) {
        for (int k = 0; k < 5; ++k) p[k] = a[k].data();
    }
    bool finite() const {
        for (const auto& v : a)
            for (double x : v)
                if (!std::isfinite(x)) return false;
        return true;
    }
};

int run(const char* name, const ds_config& cfg) {
    void* h = oracle_create(&cfg);
    if (!h) {
        std::fprintf(stderr, "%s: oracle_create failed: %s\n", name, oracle_last_error());
        return 1;
    }
    const uint32_t* n = cfg.resolution;
    const size_t cells = (size_t)n[0] * n[1] * n[2];
    Fields state(n, 0.5), control(n, 0.1), next(n, 0.0), next_adj(n, 1.0), prev_adj(n, 0.0), control_adj(n, 0.0);
    for (double& v : state.a[0]) v = std::fabs(v);   // densities are non-negative
    // cache: sourced density, temperature, force (3 cells), forced and advected velocity, pressure
    std::vector<std::vector<double>> cache(12);
    for (int k = 0; k < 5; ++k) cache[k].assign(cells, 0.0);
    for (int k = 5; k < 11; ++k) cache[k].assign(next.a[2 + (k - 5) % 3].size(), 0.0);
    cache[11].assign(cells, 0.0);
    double* s[5];
    double* c[5];
    double* nx[5];
    double* na[5];
    double* pa[5];
    double* ca[5];
    double* ch[12];
    state.ptrs(s);
    control.ptrs(c);
    next.ptrs(nx);
    next_adj.ptrs(na);
    prev_adj.ptrs(pa);
    control_adj.ptrs(ca);
    for (int k = 0; k < 12; ++k) ch[k] = cache[k].data();
    const double params[4] = {0.1, -0.2, 1.0, 0.3};
    double param_adj[4] = {0.0, 0.0, 0.0, 0.0};
    ds_solve_info info{};
    int rc = oracle_forward_step(h, s, c, params, nx, ch, nullptr, &info);
    if (rc != 0 || !next.finite()) {
        std::fprintf(stderr, "%s: forward step failed (%d): %s\n", name, rc, oracle_last_error());
        oracle_destroy(h);
        return 1;
    }
    rc = oracle_vjp_step(h, s, nx, ch, params, na, pa, ca, param_adj, &info);
    bool ok = rc == 0 && prev_adj.finite() && control_adj.finite();
    for (double v : param_adj) ok = ok && std::isfinite(v);
    if (!ok) std::fprintf(stderr, "%s: VJP step failed (%d): %s\n", name, rc, oracle_last_error());
    oracle_destroy(h);
    std::printf("%s: forward + VJP ok (%u solver iterations in the adjoint projection)\n", name, info.iterations);
    return ok ? 0 : 1;
}

ds_config base(uint32_t nx, uint32_t ny, uint32_t nz) {
    ds_config c{};
    c.resolution[0] = nx;
    c.resolution[1] = ny;
    c.resolution[2] = nz;
    c.cell_size = 1.0F / (float)nx;
    c.time_step = 0.5F / (float)nx;
    for (int f = 0; f < 6; ++f) c.pressure_boundary.modes[f] = DS_SCALAR_ZERO_GRADIENT;
    c.cg_tolerance = 1e-10;
    return c;
}

}  // namespace

int main() {
    oracle_set_threads(2);
    int failures = 0;
    {
        // closed box (fixed-value velocity faces), defaults otherwise
        const ds_config c = base(8, 8, 8);
        failures += run("closed", c);
    }
    {
        // mixed faces, a fixed-value pressure face, both collider shapes, confinement on
        ds_config c = base(10, 8, 6);
        const uint32_t vm[6] = {DS_VELOCITY_NORMAL_FIXED_TANGENT_ZERO_GRADIENT, DS_VELOCITY_NORMAL_FIXED_TANGENT_ZERO_GRADIENT,
                                DS_VELOCITY_ZERO_GRADIENT, DS_VELOCITY_FIXED_VALUE, DS_VELOCITY_FIXED_VALUE, DS_VELOCITY_ZERO_GRADIENT};
        for (int f = 0; f < 6; ++f) c.velocity_boundary.modes[f] = vm[f];
        c.velocity_boundary.values[3 * 4 + 0] = 0.05F;
        c.pressure_boundary.modes[3] = DS_SCALAR_FIXED_VALUE;
        c.pressure_boundary.values[3] = 0.1F;
        c.density_boundary.modes[0] = DS_SCALAR_FIXED_VALUE;
        c.density_boundary.values[0] = 0.3F;
        for (int f = 1; f < 6; ++f) c.density_boundary.modes[f] = DS_SCALAR_ZERO_GRADIENT;
        for (int f = 0; f < 6; ++f) c.temperature_boundary.modes[f] = DS_SCALAR_ZERO_GRADIENT;
        const ds_collider cols[2] = {{DS_SHAPE_BOX, {0.3F, 0.3F, 0.3F}, {0.1F, 0.12F, 0.1F}, {0.2F, -0.1F, 0.05F}, 0.5F, 0.25F},
                                     {DS_SHAPE_ELLIPSOID, {0.7F, 0.5F, 0.4F}, {0.12F, 0.1F, 0.14F}, {0.0F, 0.0F, 0.0F}, 0.0F, 0.0F}};
        c.colliders = cols;
        c.collider_count = 2;
        c.vorticity_confinement_enabled = 1;
        failures += run("solids", c);
    }
    {
        // periodic in x and y (velocity and scalars), red-black Gauss-Seidel projection
        ds_config c = base(8, 6, 6);
        for (int f = 0; f < 4; ++f) {
            c.velocity_boundary.modes[f] = DS_VELOCITY_PERIODIC;
            c.pressure_boundary.modes[f] = DS_SCALAR_PERIODIC;
            c.density_boundary.modes[f] = DS_SCALAR_PERIODIC;
            c.temperature_boundary.modes[f] = DS_SCALAR_PERIODIC;
        }
        c.projection_solver = DS_PROJECTION_RBGS;
        c.pressure_iterations = 40;
        failures += run("periodic_rbgs", c);
    }
    if (failures) std::fprintf(stderr, "%d case(s) failed\n", failures);
    return failures ? 1 : 0;
}
