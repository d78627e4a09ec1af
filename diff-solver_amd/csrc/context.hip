// context.hip — context creation (Model::allocate_context, solvers/smoke/model.cpp:32-182), the
// device resource layer (xayah::cuda::Resource, solvers/cuda.cpp:11-66) and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"

namespace ds {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& message) { g_last_error = message; }

void check_launch(const char* name) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Failure{DS_ERROR_LAUNCH, std::string(name) + ": " + hipGetErrorString(e)};
}

__global__ void fill_zero_kernel(float* __restrict__ p, uint64_t count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}
__global__ void fill_zero_d_kernel(double* __restrict__ p, uint64_t count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// Zeroing uses a fill kernel rather than hipMemsetAsync so that a caller may capture the step
// into a HIP graph.
void fill_zero(ds_context* c, float* p, uint64_t count) {
    if (!p || count == 0) return;
    const uint32_t g = std::min<uint64_t>(blocks_for(count), 4096);
    fill_zero_kernel<<<g, kBlock, 0, c->stream>>>(p, count);
    check_launch("fill_zero_kernel");
}
void fill_zero_d(ds_context* c, double* p, uint64_t count) {
    if (!p || count == 0) return;
    const uint32_t g = std::min<uint64_t>(blocks_for(count), 4096);
    fill_zero_d_kernel<<<g, kBlock, 0, c->stream>>>(p, count);
    check_launch("fill_zero_d_kernel");
}

namespace {

// model.cpp:17-28, evaluated in FP32 like the reference.
bool contains(const ds_collider& col, float px, float py, float pz) {
    if (col.shape == DS_SHAPE_ELLIPSOID) {
        const float x = (px - col.center[0]) / col.size[0];
        const float y = (py - col.center[1]) / col.size[1];
        const float z = (pz - col.center[2]) / col.size[2];
        return x * x + y * y + z * z <= 1.0F;
    }
    return std::fabs(px - col.center[0]) <= col.size[0] && std::fabs(py - col.center[1]) <= col.size[1] &&
           std::fabs(pz - col.center[2]) <= col.size[2];
}

struct HostDomain {
    std::vector<uint8_t> mask, code;
    std::vector<float> fixed_term;
    std::vector<float> collider_velocity[3];
    std::vector<float> collider_density, collider_temperature;
    uint64_t anchor = 0;
    uint32_t extra_anchors = 0;
    uint64_t unknowns = 0, solids = 0;
    bool fixed_pressure = false;
    bool any_fixed_term = false;
};

void validate(const ds_config& cfg, uint32_t mode) {
    auto bad = [](const std::string& m) { throw Failure{DS_ERROR_INVALID_ARGUMENT, m}; };
    for (int k = 0; k < 3; ++k)
        if (cfg.resolution[k] == 0 || cfg.resolution[k] > 1024) bad("resolution must be in [1, 1024] per axis");
    if (!(cfg.cell_size > 0.0F) || !std::isfinite(cfg.cell_size)) bad("cell_size must be positive and finite");
    if (!(cfg.time_step > 0.0F) || !std::isfinite(cfg.time_step)) bad("time_step must be positive and finite");
    if (mode != DS_MODE_FORWARD && mode != DS_MODE_DIFFERENTIABLE) bad("mode must be forward or differentiable");
    for (int f = 0; f < 6; ++f) {
        if (cfg.velocity_boundary.modes[f] > 3u) bad("velocity boundary mode out of range");
        if (cfg.pressure_boundary.modes[f] > 2u || cfg.density_boundary.modes[f] > 2u ||
            cfg.temperature_boundary.modes[f] > 2u)
            bad("scalar boundary mode out of range");
    }
    if (cfg.collider_count > 0 && cfg.colliders == nullptr) bad("colliders is null with collider_count > 0");
    for (uint32_t k = 0; k < cfg.collider_count; ++k) {
        const ds_collider& col = cfg.colliders[k];
        if (col.shape > 1u) bad("collider shape out of range");
        // an ellipsoid divides by its radii (model.cpp:17-22); a box's half extent may be 0 (it
        // then holds the cells whose centre lies on its mid-plane, model.cpp:24-26)
        for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(col.size[a])) bad("collider size must be finite");
            if (col.shape == DS_SHAPE_ELLIPSOID && !(col.size[a] > 0.0F)) bad("ellipsoid radii must be positive");
            if (col.shape == DS_SHAPE_BOX && !(col.size[a] >= 0.0F)) bad("box half extents must be non-negative");
        }
    }
    if (cfg.projection_solver > DS_PROJECTION_RBGS) bad("projection_solver out of range");
    if (cfg.warm_start > DS_WARM_START_CNN_ALONE) bad("warm_start out of range");
    if ((cfg.warm_start == DS_WARM_START_CNN || cfg.warm_start == DS_WARM_START_CNN_ALONE) && (cfg.resolution[0] % 64 || cfg.resolution[1] % 16 || cfg.resolution[2] % 4))
        throw Failure{DS_ERROR_UNSUPPORTED, "the warm-start network needs nx % 64 == 0, ny % 16 == 0 and nz % 4 == 0"};
    if (!(cfg.cg_tolerance >= 0.0) || !std::isfinite(cfg.cg_tolerance)) bad("cg_tolerance must be >= 0 and finite");
    if (cfg.cg_variant > DS_CG_MULTIGRID) bad("cg_variant out of range");
}

// model.cpp:56-182 plus the pressure connectivity (operators.cu:852-873): solid mask from cell
// centres, anchor, face collider velocity, per-cell diagonal count, and an anchor per fluid
// pocket that is cut off from the main anchor and from every fixed-value face.
HostDomain build_domain(const ds_config& cfg) {
    HostDomain d;
    const int nx = (int)cfg.resolution[0], ny = (int)cfg.resolution[1], nz = (int)cfg.resolution[2];
    const uint64_t n = (uint64_t)nx * ny * nz;
    auto cell = [&](int x, int y, int z) { return (uint64_t)x + (uint64_t)nx * ((uint64_t)y + (uint64_t)ny * z); };
    d.mask.assign(n, 0);
    d.collider_density.assign(n, 0.0F);
    d.collider_temperature.assign(n, 0.0F);
    std::vector<float> cvel[3];
    if (cfg.collider_count)
        for (auto& v : cvel) v.assign(n, 0.0F);
    for (int z = 0; z < nz && cfg.collider_count; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const float px = ((float)x + 0.5F) * cfg.cell_size;
                const float py = ((float)y + 0.5F) * cfg.cell_size;
                const float pz = ((float)z + 0.5F) * cfg.cell_size;
                const uint64_t i = cell(x, y, z);
                for (uint32_t k = 0; k < cfg.collider_count; ++k) {   // later colliders win (model.cpp:60-69)
                    const ds_collider& col = cfg.colliders[k];
                    if (!contains(col, px, py, pz)) continue;
                    d.mask[i] = 1;
                    for (int a = 0; a < 3; ++a) cvel[a][i] = col.velocity[a];
                    d.collider_density[i] = col.density;
                    d.collider_temperature[i] = col.temperature;
                }
            }
    d.solids = 0;
    for (uint64_t i = 0; i < n; ++i) d.solids += d.mask[i];
    if (d.solids == n) throw Failure{DS_ERROR_INVALID_ARGUMENT, "every cell is solid"};
    d.fixed_pressure = false;
    for (int f = 0; f < 6; ++f)
        if (cfg.pressure_boundary.modes[f] == DS_SCALAR_FIXED_VALUE) d.fixed_pressure = true;
    if (d.fixed_pressure) d.anchor = n;   // model.cpp:73-75
    else {
        d.anchor = 0;
        while (d.mask[d.anchor] != 0) ++d.anchor;
    }
    if (cfg.collider_count) {   // model.cpp:79-150
        for (int a = 0; a < 3; ++a) {
            const int ex = nx + (a == 0), ey = ny + (a == 1), ez = nz + (a == 2);
            d.collider_velocity[a].assign((uint64_t)ex * ey * ez, 0.0F);
            for (int z = 0; z < ez; ++z)
                for (int y = 0; y < ey; ++y)
                    for (int x = 0; x < ex; ++x) {
                        float sum = 0.0F, count = 0.0F;
                        const int c0[3]{x - (a == 0), y - (a == 1), z - (a == 2)};
                        const int c1[3]{x, y, z};
                        for (const int* c : {c0, c1}) {
                            if (c[0] < 0 || c[0] >= nx || c[1] < 0 || c[1] >= ny || c[2] < 0 || c[2] >= nz) continue;
                            const uint64_t ci = cell(c[0], c[1], c[2]);
                            if (d.mask[ci]) { sum += cvel[a][ci]; count += 1.0F; }
                        }
                        d.collider_velocity[a][(uint64_t)x + (uint64_t)ex * ((uint64_t)y + (uint64_t)ey * z)] =
                            count == 0.0F ? 0.0F : sum / count;
                    }
        }
    }
    const int dims[3]{nx, ny, nz};
    bool per[3];
    for (int k = 0; k < 3; ++k)
        per[k] = cfg.pressure_boundary.modes[2 * k] == DS_SCALAR_PERIODIC && cfg.pressure_boundary.modes[2 * k + 1] == DS_SCALAR_PERIODIC;
    // neighbour of (x,y,z) in direction (dim, dir): cell index, or -1 with fixed / value set
    auto neighbour = [&](int x, int y, int z, int dim, int dir, bool& fixed, float& value) -> int64_t {
        int c[3]{x, y, z};
        c[dim] += dir;
        fixed = false;
        if (c[dim] < 0 || c[dim] >= dims[dim]) {
            if (per[dim]) c[dim] = (c[dim] % dims[dim] + dims[dim]) % dims[dim];
            else {
                const int f = dim * 2 + (dir > 0);
                fixed = cfg.pressure_boundary.modes[f] == DS_SCALAR_FIXED_VALUE;
                value = cfg.pressure_boundary.values[f];
                return -1;
            }
        }
        return (int64_t)cell(c[0], c[1], c[2]);
    };
    std::vector<uint8_t> diag(n, 0), touches_fixed(n, 0);
    d.fixed_term.assign(n, 0.0F);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const uint64_t i = cell(x, y, z);
                if (d.mask[i]) continue;
                int dg = 0;
                float ft = 0.0F;
                for (int dim = 0; dim < 3; ++dim)
                    for (int dir = -1; dir <= 1; dir += 2) {
                        bool fixed;
                        float v = 0.0F;
                        const int64_t j = neighbour(x, y, z, dim, dir, fixed, v);
                        if (j >= 0 && d.mask[j] == 0) ++dg;
                        else if (j < 0 && fixed) { ++dg; ft += v; touches_fixed[i] = 1; }
                    }
                diag[i] = (uint8_t)dg;
                d.fixed_term[i] = ft;
            }
    // connected components of the fluid; a component without the anchor or a fixed face is
    // singular and gets its lowest-index cell as its own anchor
    std::vector<int32_t> comp(n, -1);
    std::vector<uint64_t> stack;
    std::vector<uint64_t> extra;
    int32_t ncomp = 0;
    for (uint64_t seed = 0; seed < n; ++seed) {
        if (d.mask[seed] || comp[seed] >= 0) continue;
        bool grounded = false;
        stack.assign(1, seed);
        comp[seed] = ncomp;
        while (!stack.empty()) {
            const uint64_t i = stack.back();
            stack.pop_back();
            if (i == d.anchor || touches_fixed[i]) grounded = true;
            const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((uint64_t)nx * ny));
            for (int dim = 0; dim < 3; ++dim)
                for (int dir = -1; dir <= 1; dir += 2) {
                    bool fixed;
                    float v;
                    const int64_t j = neighbour(x, y, z, dim, dir, fixed, v);
                    if (j >= 0 && d.mask[j] == 0 && comp[j] < 0) {
                        comp[j] = ncomp;
                        stack.push_back((uint64_t)j);
                    }
                }
        }
        if (!grounded) extra.push_back(seed);
        ++ncomp;
    }
    std::vector<uint8_t> is_anchor(n, 0);
    if (d.anchor < n) is_anchor[d.anchor] = 1;
    for (uint64_t a : extra) is_anchor[a] = 1;
    d.extra_anchors = (uint32_t)extra.size();
    d.code.assign(n, 0);
    d.unknowns = 0;
    d.any_fixed_term = false;
    for (uint64_t i = 0; i < n; ++i) {
        if (d.mask[i] || is_anchor[i] || diag[i] == 0) {
            d.code[i] = 0;
            d.fixed_term[i] = 0.0F;
            continue;
        }
        d.code[i] = diag[i];
        ++d.unknowns;
        if (d.fixed_term[i] != 0.0F) d.any_fixed_term = true;
    }
    return d;
}

template <typename T>
T* upload(ds_context* c, const std::vector<T>& host) {
    T* p = device_alloc<T>(c, host.size());
    DS_HIP(hipMemcpyAsync(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return p;
}

}  // namespace

}  // namespace ds

using namespace ds;

extern "C" {

const char* ds_last_error(void) { return g_last_error.c_str(); }

uint64_t ds_config_cell_count(const ds_config* config) {
    if (!config) return 0;
    return (uint64_t)config->resolution[0] * config->resolution[1] * config->resolution[2];
}

uint64_t ds_config_face_count(const ds_config* config, uint32_t axis) {
    if (!config || axis > 2) return 0;
    uint64_t n = 1;
    for (uint32_t k = 0; k < 3; ++k) n *= config->resolution[k] + (k == axis ? 1u : 0u);
    return n;
}

ds_status ds_context_create(const ds_config* config, uint32_t mode, void* stream, ds_context** out) {
    if (out) *out = nullptr;
    if (!config || !out) {
        set_error("ds_context_create: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    ds_context* c = new (std::nothrow) ds_context();
    if (!c) {
        set_error("host allocation failed");
        return DS_ERROR_OUT_OF_MEMORY;
    }
    const ds_status st = guard([&] {
        validate(*config, mode);
        c->config = *config;
        c->colliders.assign(config->colliders, config->colliders + config->collider_count);
        c->config.colliders = c->colliders.empty() ? nullptr : c->colliders.data();
        c->mode = mode;
        DS_HIP(hipGetDevice(&c->device));
        if (stream) c->stream = static_cast<hipStream_t>(stream);
        else {
            DS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));   // cuda.cpp:27-29
            c->own_stream = true;
        }
        const ds_config& cfg = c->config;
        Geo& g = c->geo;
        g.nx = (int)cfg.resolution[0];
        g.ny = (int)cfg.resolution[1];
        g.nz = (int)cfg.resolution[2];
        g.h = cfg.cell_size;
        g.inv_h = 1.0f / cfg.cell_size;
        g.h_pow2 = is_pow2(cfg.cell_size) ? 1u : 0u;
        g.dt = cfg.time_step;
        g.n = (uint32_t)ds_config_cell_count(&cfg);
        g.foff[0] = 0;
        for (int a = 0; a < 3; ++a) {
            g.nf[a] = (uint32_t)ds_config_face_count(&cfg, a);
            g.foff[a + 1] = g.foff[a] + g.nf[a];
            g.fx[a] = make_fastdiv((uint32_t)(g.nx + (a == 0)));
            g.fy[a] = make_fastdiv((uint32_t)(g.ny + (a == 1)));
        }
        g.cx = make_fastdiv((uint32_t)g.nx);
        g.cy = make_fastdiv((uint32_t)g.ny);
        for (int f = 0; f < 6; ++f) {
            c->vb.modes[f] = cfg.velocity_boundary.modes[f];
            c->pb.modes[f] = cfg.pressure_boundary.modes[f];
            c->db.modes[f] = cfg.density_boundary.modes[f];
            c->tb.modes[f] = cfg.temperature_boundary.modes[f];
            c->pb.values[f] = cfg.pressure_boundary.values[f];
            c->db.values[f] = cfg.density_boundary.values[f];
            c->tb.values[f] = cfg.temperature_boundary.values[f];
        }
        for (int k = 0; k < 18; ++k) c->vb.values[k] = cfg.velocity_boundary.values[k];
        for (int k = 0; k < 3; ++k) c->pressure_periodic[k] = c->pb.modes[2 * k] == 2u && c->pb.modes[2 * k + 1] == 2u;
        c->tolerance = cfg.cg_tolerance > 0.0 ? cfg.cg_tolerance : 1e-6;
        c->max_iterations = cfg.cg_max_iterations ? cfg.cg_max_iterations
                                                  : 4u * (cfg.resolution[0] + cfg.resolution[1] + cfg.resolution[2]) + 1000u;
        c->replacement_interval = cfg.residual_replacement_interval ? cfg.residual_replacement_interval : 50u;

        {
            c->vb_zero = c->vb;
            c->db_zero = c->db;
            c->tb_zero = c->tb;
            c->pb_zero = c->pb;
            for (float& v : c->vb_zero.values) v = 0.0f;
            for (float& v : c->db_zero.values) v = 0.0f;
            for (float& v : c->tb_zero.values) v = 0.0f;
            for (float& v : c->pb_zero.values) v = 0.0f;
            VBound* v = device_alloc<VBound>(c, 2);
            SBound* sb = device_alloc<SBound>(c, 6);
            DS_HIP(hipMemcpyAsync(sb + 4, &c->pb, sizeof(SBound), hipMemcpyHostToDevice, c->stream));
            DS_HIP(hipMemcpyAsync(sb + 5, &c->pb_zero, sizeof(SBound), hipMemcpyHostToDevice, c->stream));
            c->d_pb = sb + 4;
            c->d_pb_zero = sb + 5;
            DS_HIP(hipMemcpyAsync(v, &c->vb, sizeof(VBound), hipMemcpyHostToDevice, c->stream));
            DS_HIP(hipMemcpyAsync(v + 1, &c->vb_zero, sizeof(VBound), hipMemcpyHostToDevice, c->stream));
            DS_HIP(hipMemcpyAsync(sb, &c->db, sizeof(SBound), hipMemcpyHostToDevice, c->stream));
            DS_HIP(hipMemcpyAsync(sb + 1, &c->tb, sizeof(SBound), hipMemcpyHostToDevice, c->stream));
            DS_HIP(hipMemcpyAsync(sb + 2, &c->db_zero, sizeof(SBound), hipMemcpyHostToDevice, c->stream));
            DS_HIP(hipMemcpyAsync(sb + 3, &c->tb_zero, sizeof(SBound), hipMemcpyHostToDevice, c->stream));
            c->d_vb = v;
            c->d_vb_zero = v + 1;
            c->d_db = sb;
            c->d_tb = sb + 1;
            c->d_db_zero = sb + 2;
            c->d_tb_zero = sb + 3;
        }
        const HostDomain d = build_domain(cfg);
        c->anchor = d.anchor;
        c->extra_anchors = d.extra_anchors;
        c->unknowns = d.unknowns;
        c->solids = d.solids;
        c->has_solids = d.solids > 0;
        c->has_fixed_pressure = d.fixed_pressure;
        c->mask = upload(c, d.mask);
        {
            // element n is a pad of 0 that out-of-domain stencil loads read
            std::vector<uint8_t> padded(d.code);
            padded.push_back(0);
            c->code = upload(c, padded);
        }
        if (d.any_fixed_term) c->fixed_term = upload(c, d.fixed_term);
        if (cfg.collider_count) {
            for (int a = 0; a < 3; ++a) c->collider_velocity[a] = upload(c, d.collider_velocity[a]);
            c->collider_density = upload(c, d.collider_density);
            c->collider_temperature = upload(c, d.collider_temperature);
        }
        const uint64_t n = g.n;
        for (int a = 0; a < 3; ++a) {
            c->force[a] = device_alloc<float>(c, n);
            c->advected[a] = device_alloc<float>(c, g.nf[a]);
        }
        c->pressure_hi = device_alloc<float>(c, n);
        if (cfg.warm_start == DS_WARM_START_EXTRAPOLATE || cfg.warm_start == DS_WARM_START_CNN) {
            c->pressure_prev = device_alloc<float>(c, 2 * (size_t)n);
            c->guess = device_alloc<float>(c, n);
        }
        if (cfg.warm_start == DS_WARM_START_PROJECTION) {
            c->guess = device_alloc<float>(c, n);
            window_allocate(c, c->window[0]);
            if (mode == DS_MODE_DIFFERENTIABLE) window_allocate(c, c->window[1]);
        }
        c->pressure_lo = device_alloc<float>(c, n);
        // PCG work vectors
        c->pcg.b = device_alloc<float>(c, n);
        // iteration vectors carry a zero pad element n (read by out-of-domain stencil loads)
        c->pcg.r = device_alloc<float>(c, n + 1);
        c->pcg.p[0] = device_alloc<float>(c, n + 1);
        c->pcg.p[1] = device_alloc<float>(c, n + 1);
        for (float* v : {c->pcg.r, c->pcg.p[0], c->pcg.p[1]}) fill_zero(c, v, n + 1);
        // the resident single-reduction variant runs the same recurrence
        c->pcg.variant = cfg.cg_variant == DS_CG_TWO_REDUCTION || cfg.cg_variant == DS_CG_S_STEP || cfg.cg_variant == DS_CG_MULTIGRID
                              ? cfg.cg_variant
                              : DS_CG_SINGLE_REDUCTION;
        if (c->pcg.variant == DS_CG_S_STEP) {
            // every basis vector carries the zero pad element n (read by out-of-domain rim loads);
            // a set is addressed with 32-bit byte offsets
            if (5ull * (n + 1) * sizeof(float) >= (1ull << 32)) throw Failure{DS_ERROR_UNSUPPORTED, "grid too large for the s-step PCG"};
            for (int k = 0; k < 2; ++k) {
                c->pcg.s2set[k] = device_alloc<float>(c, 5 * (size_t)(n + 1));
                fill_zero(c, c->pcg.s2set[k], 5 * (size_t)(n + 1));
                for (int m = 0; m < 5; ++m) c->pcg.s2v[k][m] = c->pcg.s2set[k] + (size_t)m * (n + 1);
            }
            c->pcg.xd = device_alloc<float>(c, n + 1);
            fill_zero(c, c->pcg.xd, n + 1);
            c->pcg.s2_parts = device_alloc<double>(c, 2 * (size_t)16 * kPcgMaxBlocks);
        }
        if (c->pcg.variant == DS_CG_SINGLE_REDUCTION) {
            // s is written on unknowns only; its other cells must read as 0 (as halo values)
            c->pcg.r_alt = device_alloc<float>(c, n + 1);
            fill_zero(c, c->pcg.r_alt, n + 1);
            c->pcg.xd = device_alloc<float>(c, n + 1);
            fill_zero(c, c->pcg.xd, n + 1);
            for (int k = 0; k < 2; ++k) {
                c->pcg.wv[k] = device_alloc<float>(c, n + 1);
                c->pcg.sv[k] = device_alloc<float>(c, n + 1);
                fill_zero(c, c->pcg.wv[k], n + 1);
                fill_zero(c, c->pcg.sv[k], n + 1);
            }
        }
        {
            // about kPcgTargetBlocks blocks (kPcgTargetBlocksS2 for the s-step form): zc = tiles * nz / target
            // rounded up to a power of two in [4, 64]
            PlaneTiling& t = c->pcg.tiling;
            t.tiles_x = (uint32_t)((g.nx + kPlaneTileX - 1) / kPlaneTileX);
            t.tiles_y = (uint32_t)((g.ny + kPlaneTileY - 1) / kPlaneTileY);
            const uint64_t tiles = (uint64_t)t.tiles_x * t.tiles_y;
            uint32_t zc = 4;
            // the fused kernel reads one producer partial per thread: blocks <= kPcgThreads
            const uint64_t cap = std::min<uint64_t>(kPcgMaxBlocks, kPcgThreads);
            const uint64_t target = cfg.cg_variant == DS_CG_S_STEP ? kPcgTargetBlocksS2 : kPcgTargetBlocks;
            while (zc < 64 && tiles * (uint64_t)g.nz > target * zc) zc *= 2;
            while (tiles * (((uint64_t)g.nz + zc - 1) / zc) > cap) zc *= 2;
            t.zc = zc;
            const uint64_t G = tiles * (((uint64_t)g.nz + zc - 1) / zc);
            c->pcg.blocks = (uint32_t)((G + 7) & ~7ull);
            // the fused kernel gives each thread at most one producer partial
            if (c->pcg.blocks > kPcgThreads) throw Failure{DS_ERROR_UNSUPPORTED, "grid too large for the PCG tiling"};
            // the multigrid form has no fused kernel: its reducing kernels (and the set-up, true
            // residual and finalisation that share their partial slots) take one block per 2048
            // cells, up to the slots' capacity, for more waves in flight
            if (cfg.cg_variant == DS_CG_MULTIGRID) {
                const uint64_t b = std::min<uint64_t>(kPcgMaxBlocks, std::max<uint64_t>(8, ((uint64_t)g.n + 2047) / 2048));
                c->pcg.blocks = (uint32_t)((b + 7) & ~7ull);
            }
        }
        c->pcg.partials = device_alloc<double>(c, 2 * (size_t)kPartCount * kPcgMaxBlocks);   // two sets
        c->pcg.hist = device_alloc<double>(c, 2 * ((size_t)c->max_iterations + 4));
        c->pcg.status = device_alloc<PcgStatus>(c, 1);
        if (cfg.cg_variant == DS_CG_SINGLE_REDUCTION_RESIDENT) pcg_resident_prepare(c);
        if (cfg.cg_variant == DS_CG_MULTIGRID) mg_prepare(c, d.code);
        c->metric_partials = device_alloc<double>(c, 6 * (size_t)kMetricBlocks);
        void* hs = nullptr;
        DS_HIP(hipHostMalloc(&hs, 2 * sizeof(PcgStatus), hipHostMallocDefault));
        c->pcg.host_status = static_cast<PcgStatus*>(hs);
        DS_HIP(hipEventCreateWithFlags(&c->pcg.batch_event[0], hipEventDisableTiming));
        DS_HIP(hipEventCreateWithFlags(&c->pcg.batch_event[1], hipEventDisableTiming));
        if (mode == DS_MODE_DIFFERENTIABLE) {
            for (int k = 0; k < 2; ++k) c->sourced_adj[k] = device_alloc<float>(c, n);
            for (int a = 0; a < 3; ++a) {
                c->projected_extra[a] = device_alloc<float>(c, g.nf[a]);
                c->forced_adj[a] = device_alloc<float>(c, g.nf[a]);
                c->raw_adj[a] = device_alloc<float>(c, g.nf[a]);
            }
            {
                const uint32_t nmax = std::max({g.nf[0], g.nf[1], g.nf[2]});
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) c->vjp_start[a][b] = device_alloc<float>(c, nmax);
            }
            c->lambda_hi = device_alloc<float>(c, n);
            c->lambda_lo = device_alloc<float>(c, n);
            c->param_partials = device_alloc<double>(c, 4 * (size_t)std::max(kPcgBlocks, kForceBlocks));
            for (int k = 0; k < 2; ++k) c->sourced_t[k] = device_alloc<float>(c, n);
            for (int a = 0; a < 3; ++a) {
                c->force_t[a] = device_alloc<float>(c, n);
                c->forced_t[a] = device_alloc<float>(c, g.nf[a]);
                c->advected_t[a] = device_alloc<float>(c, g.nf[a]);
            }
            c->tangent_hi = device_alloc<float>(c, n);
            c->tangent_lo = device_alloc<float>(c, n);
        }
        if (cfg.projection_solver == DS_PROJECTION_RBGS) {
            c->rbgs_diag = device_alloc<uint8_t>(c, n);
            rbgs_prepare(c);
        }
        if (cfg.vorticity_confinement_enabled) {
            c->vorticity = true;
            for (int a = 0; a < 3; ++a) c->vort_w[a] = device_alloc<float>(c, n);
            c->vort_mag = device_alloc<float>(c, n);
            if (mode == DS_MODE_DIFFERENTIABLE) {
                for (int a = 0; a < 3; ++a) {
                    c->vort_ga[a] = device_alloc<float>(c, n);
                    c->vort_t[a] = device_alloc<float>(c, n);
                    c->vort_wt[a] = device_alloc<float>(c, n);
                }
                c->vort_mt = device_alloc<float>(c, n);
            }
        }
        DS_HIP(hipStreamSynchronize(c->stream));   // model.cpp:181
    });
    if (st != DS_OK) {
        ds_context_destroy(c);
        return st;
    }
    *out = c;
    return DS_OK;
}

ds_status ds_context_destroy(ds_context* c) {
    if (!c) return DS_OK;
    ds_status st = DS_OK;
    if (c->stream && hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("ds_context_destroy: stream synchronisation failed");
        st = DS_ERROR_RUNTIME;
    }
    mg_release(c);
    for (void* p : c->allocations) (void)hipFree(p);
    unet_release(c->unet);
    if (c->pcg.host_status) (void)hipHostFree(c->pcg.host_status);
    for (auto& e : c->pcg.batch_event)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->profiler.pool) (void)hipEventDestroy(e);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return st;
}

void* ds_context_stream(ds_context* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

ds_status ds_context_domain_info(ds_context* c, ds_domain_info* info) {
    if (!c || !info) {
        set_error("ds_context_domain_info: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    info->cell_count = c->geo.n;
    for (int a = 0; a < 3; ++a) info->face_count[a] = c->geo.nf[a];
    info->unknown_count = c->unknowns;
    info->solid_count = c->solids;
    info->pressure_anchor = c->anchor;
    info->extra_anchor_count = c->extra_anchors;
    info->has_fixed_pressure = c->has_fixed_pressure;
    return DS_OK;
}

ds_status ds_context_domain_masks(ds_context* c, uint8_t* mask, uint8_t* code) {
    if (!c) {
        set_error("ds_context_domain_masks: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        if (mask) DS_HIP(hipMemcpyAsync(mask, c->mask, c->geo.n, hipMemcpyDeviceToHost, c->stream));
        if (code) DS_HIP(hipMemcpyAsync(code, c->code, c->geo.n, hipMemcpyDeviceToHost, c->stream));
        DS_HIP(hipStreamSynchronize(c->stream));
    });
}

ds_status ds_context_last_solution(ds_context* c, uint32_t which, float* out) {
    if (!c || !out || which > 1) {
        set_error("ds_context_last_solution: null argument or unknown sequence");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        const bool valid = which == 0 ? c->pressure_valid : c->lambda_valid;
        require(valid && (which == 0 || c->lambda_hi), "ds_context_last_solution: no solve of that sequence yet");
        combine_pair(c, which == 0 ? c->pressure_hi : c->lambda_hi, which == 0 ? c->pressure_lo : c->lambda_lo, out);
        check_launch("combine_pair_kernel");
    });
}

ds_status ds_allocate(ds_context* c, size_t bytes, void** out) {
    if (!c || !out) {
        set_error("ds_allocate: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        *out = nullptr;
        DS_HIP(hipMallocAsync(out, bytes ? bytes : 1, c->stream));   // cuda.cpp:39-43 (stream-ordered pool)
    });
}

ds_status ds_release(ds_context* c, void* p) {
    if (!c) {
        set_error("ds_release: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!p) return DS_OK;
    return guard([&] { DS_HIP(hipFreeAsync(p, c->stream)); });
}

ds_status ds_copy_from_host(ds_context* c, void* device, const void* host, size_t bytes) {
    if (!c || (bytes && (!device || !host))) {
        set_error("ds_copy_from_host: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { DS_HIP(hipMemcpyAsync(device, host, bytes, hipMemcpyHostToDevice, c->stream)); });
}

ds_status ds_copy_to_host(ds_context* c, void* host, const void* device, size_t bytes) {
    if (!c || (bytes && (!device || !host))) {
        set_error("ds_copy_to_host: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { DS_HIP(hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, c->stream)); });
}

ds_status ds_copy_device(ds_context* c, void* dst, const void* src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) {
        set_error("ds_copy_device: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { DS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream)); });
}

ds_status ds_zero(ds_context* c, void* device, size_t bytes) {
    if (!c || (bytes && !device)) {
        set_error("ds_zero: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { DS_HIP(hipMemsetAsync(device, 0, bytes, c->stream)); });
}

ds_status ds_synchronize(ds_context* c) {
    if (!c) {
        set_error("ds_synchronize: null context");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] { DS_HIP(hipStreamSynchronize(c->stream)); });
}

}  // extern "C"
