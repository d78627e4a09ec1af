// smoke_oracle.cpp — TEST INFRASTRUCTURE ONLY. FP64 CPU restatement of the reference smoke
// solver's discretisation (Xayah-Graphics/diff-solver, solvers/smoke/). It is the checker for the
// HIP path and the timed host-core baseline; the product never links, loads or calls it.
//
// Every function cites the reference file:line whose algorithm it restates. Arithmetic is FP64
// except where the reference defines a quantity in FP32 on the host (collider rasterisation,
// model.cpp:17-28,56-71, and the collider face velocity, model.cpp:79-150), which is kept in
// FP32 so that the solid mask is the reference's mask.
//
// One deliberate change from the reference (the build's pressure solver): the pressure solve is
// Jacobi-preconditioned CG to a relative residual on the matrix the reference's red-black
// Gauss-Seidel iterates on (operators.cu:852-902). Fluid pockets that neither contain the
// anchor nor touch a fixed-value pressure face make that matrix singular; each such pocket is
// given its own anchor (its lowest-index cell), mirroring the reference's rule for the main
// anchor (model.cpp:72-77). RBGS with a fixed sweep count is also restated as a reference mode.
//
// Parity pins (tests/test_oracle_*.py): closed forms per element, exact trilinear reproduction,
// exact shift of a uniform-velocity trace, dense symmetric-positive-definite checks of the
// pressure matrix against numpy, PCG against dense solves and manufactured solutions,
// projection divergence and idempotence, RBGS convergence to the PCG solution, and FP64
// central finite differences of the full step and of the checkpointed trajectory gradient.

#include "ds/ds_smoke.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using i64 = std::int64_t;

struct V3 {
    double x, y, z;
};
struct Sample {
    double value;
    V3 gradient;
};
struct Trace {
    V3 position;
    V3 derivative;
};
struct SB {
    std::uint32_t modes[6];
    double values[6];
};
struct VB {
    std::uint32_t modes[6];
    double values[18];
};

struct Oracle {
    int nx = 0, ny = 0, nz = 0;
    double h = 0.0, dt = 0.0;
    float hf = 0.0F;
    VB vb{};
    SB pb{}, db{}, tb{};
    i64 n = 0;
    i64 nf[3]{};
    std::vector<std::uint8_t> mask;   // 1 = solid
    std::vector<std::uint8_t> code;   // diagonal count of an unknown, 0 = not an unknown
    std::vector<double> fixed_term;   // sum of fixed-value pressure faces (A p = b + fixed_term)
    std::vector<double> collider_velocity[3];
    std::vector<double> collider_density, collider_temperature;
    i64 anchor = 0;
    std::uint32_t extra_anchors = 0;
    bool fixed_pressure = false;
    bool vorticity = false;            // vorticity_confinement_enabled (data.ixx:82)
    i64 unknowns = 0, solids = 0;
    std::uint32_t projection_solver = DS_PROJECTION_PCG;
    std::uint32_t rbgs_iterations = 80;
    double tolerance = 1e-6;
    std::uint32_t max_iterations = 1000;
    // scratch
    std::vector<double> pressure, rhs;
};

thread_local std::string g_error;

// ---- indexing (operators.cu:32-54,105-113) ------------------------------------------------
inline int extent(const Oracle& o, int axis, int dim) {
    const int base = dim == 0 ? o.nx : dim == 1 ? o.ny : o.nz;
    return base + (axis == dim ? 1 : 0);
}
inline i64 idx3(i64 x, i64 y, i64 z, i64 ex, i64 ey) { return x + ex * (y + ey * z); }
inline i64 cell(const Oracle& o, int x, int y, int z) { return idx3(x, y, z, o.nx, o.ny); }
inline int size_of(const Oracle& o, int dim) { return dim == 0 ? o.nx : dim == 1 ? o.ny : o.nz; }

// operators.cu:80-96
inline bool vperiodic(const VB& b, int d) { return b.modes[2 * d] == 3u && b.modes[2 * d + 1] == 3u; }
inline bool speriodic(const SB& b, int d) { return b.modes[2 * d] == 2u && b.modes[2 * d + 1] == 2u; }
inline int wrap(int v, int p) {
    const int r = v % p;
    return r < 0 ? r + p : r;
}
inline int map_coordinate(int v, int size, bool is_periodic, int period) {
    if (is_periodic) return wrap(v, period);
    return std::max(0, std::min(size - 1, v));
}

// load_scalar, operators.cu:115-123: fixed-value sides return the face value (first hit in face
// order), otherwise wrap (periodic) or clamp.
double load_scalar(const Oracle& o, const double* v, int x, int y, int z, const SB& b) {
    if (x < 0 && b.modes[0] == 0u) return b.values[0];
    if (x >= o.nx && b.modes[1] == 0u) return b.values[1];
    if (y < 0 && b.modes[2] == 0u) return b.values[2];
    if (y >= o.ny && b.modes[3] == 0u) return b.values[3];
    if (z < 0 && b.modes[4] == 0u) return b.values[4];
    if (z >= o.nz && b.modes[5] == 0u) return b.values[5];
    x = map_coordinate(x, o.nx, speriodic(b, 0), o.nx);
    y = map_coordinate(y, o.ny, speriodic(b, 1), o.ny);
    z = map_coordinate(z, o.nz, speriodic(b, 2), o.nz);
    return v[cell(o, x, y, z)];
}

inline bool scalar_constant(const Oracle& o, int x, int y, int z, const SB& b) {
    return (x < 0 && b.modes[0] == 0u) || (x >= o.nx && b.modes[1] == 0u) || (y < 0 && b.modes[2] == 0u) ||
           (y >= o.ny && b.modes[3] == 0u) || (z < 0 && b.modes[4] == 0u) || (z >= o.nz && b.modes[5] == 0u);
}

inline i64 scalar_index(const Oracle& o, int x, int y, int z, const SB& b) {
    x = map_coordinate(x, o.nx, speriodic(b, 0), o.nx);
    y = map_coordinate(y, o.ny, speriodic(b, 1), o.ny);
    z = map_coordinate(z, o.nz, speriodic(b, 2), o.nz);
    return cell(o, x, y, z);
}

// mapped_face_index, operators.cu:105-113 (periodic wrap uses the cell count as the period)
inline i64 face_index(const Oracle& o, int axis, int x, int y, int z, const VB& b) {
    const int ex = extent(o, axis, 0), ey = extent(o, axis, 1), ez = extent(o, axis, 2);
    x = map_coordinate(x, ex, vperiodic(b, 0), o.nx);
    y = map_coordinate(y, ey, vperiodic(b, 1), o.ny);
    z = map_coordinate(z, ez, vperiodic(b, 2), o.nz);
    return idx3(x, y, z, ex, ey);
}

// load_face, operators.cu:125-134
bool face_constant(const Oracle& o, int axis, const int c[3], const VB& b, double* value) {
    for (int d = 0; d < 3; ++d) {
        const int size = extent(o, axis, d);
        if (c[d] >= 0 && c[d] < size) continue;
        const int f = 2 * d + (c[d] >= size);
        if (b.modes[f] == 0u || (b.modes[f] == 2u && axis == d)) {
            if (value) *value = b.values[3 * f + axis];
            return true;
        }
    }
    return false;
}

double load_face(const Oracle& o, const double* v, int axis, int x, int y, int z, const VB& b) {
    const int c[3]{x, y, z};
    double value = 0.0;
    if (face_constant(o, axis, c, b, &value)) return value;
    return v[face_index(o, axis, x, y, z, b)];
}

// sample_scalar / sample_face, operators.cu:136-194: trilinear value and its analytic gradient.
Sample trilinear(const double vals[2][2][2], double tx, double ty, double tz, double h) {
    const double wx[2]{1.0 - tx, tx}, wy[2]{1.0 - ty, ty}, wz[2]{1.0 - tz, tz};
    Sample s{0.0, {0.0, 0.0, 0.0}};
    for (int dz = 0; dz < 2; ++dz) {
        for (int dy = 0; dy < 2; ++dy) {
            for (int dx = 0; dx < 2; ++dx) s.value += wx[dx] * wy[dy] * wz[dz] * vals[dz][dy][dx];
            s.gradient.x += wy[dy] * wz[dz] * (vals[dz][dy][1] - vals[dz][dy][0]) / h;
        }
        for (int dx = 0; dx < 2; ++dx) s.gradient.y += wx[dx] * wz[dz] * (vals[dz][1][dx] - vals[dz][0][dx]) / h;
    }
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) s.gradient.z += wx[dx] * wy[dy] * (vals[1][dy][dx] - vals[0][dy][dx]) / h;
    return s;
}

Sample sample_scalar(const Oracle& o, const double* v, V3 p, const SB& b) {
    const double gx = p.x / o.h - 0.5, gy = p.y / o.h - 0.5, gz = p.z / o.h - 0.5;
    const int x0 = (int)std::floor(gx), y0 = (int)std::floor(gy), z0 = (int)std::floor(gz);
    double vals[2][2][2];
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) vals[dz][dy][dx] = load_scalar(o, v, x0 + dx, y0 + dy, z0 + dz, b);
    return trilinear(vals, gx - x0, gy - y0, gz - z0, o.h);
}

Sample sample_face(const Oracle& o, const double* v, int axis, V3 p, const VB& b) {
    const double gx = p.x / o.h - (axis == 0 ? 0.0 : 0.5);
    const double gy = p.y / o.h - (axis == 1 ? 0.0 : 0.5);
    const double gz = p.z / o.h - (axis == 2 ? 0.0 : 0.5);
    const int x0 = (int)std::floor(gx), y0 = (int)std::floor(gy), z0 = (int)std::floor(gz);
    double vals[2][2][2];
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) vals[dz][dy][dx] = load_face(o, v, axis, x0 + dx, y0 + dy, z0 + dz, b);
    return trilinear(vals, gx - x0, gy - y0, gz - z0, o.h);
}

inline void atomic_add(double* target, double value) {
#pragma omp atomic
    *target += value;
}

// scatter_scalar, operators.cu:196-215: transpose of sample_scalar's value; constants skipped.
void scatter_scalar(const Oracle& o, double* v, V3 p, double adj, const SB& b) {
    const double gx = p.x / o.h - 0.5, gy = p.y / o.h - 0.5, gz = p.z / o.h - 0.5;
    const int x0 = (int)std::floor(gx), y0 = (int)std::floor(gy), z0 = (int)std::floor(gz);
    const double wx[2]{1.0 - (gx - x0), gx - x0}, wy[2]{1.0 - (gy - y0), gy - y0}, wz[2]{1.0 - (gz - z0), gz - z0};
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int x = x0 + dx, y = y0 + dy, z = z0 + dz;
                if (scalar_constant(o, x, y, z, b)) continue;
                atomic_add(v + scalar_index(o, x, y, z, b), adj * wx[dx] * wy[dy] * wz[dz]);
            }
}

// scatter_face, operators.cu:217-240
void scatter_face(const Oracle& o, double* v, int axis, V3 p, double adj, const VB& b) {
    const double gx = p.x / o.h - (axis == 0 ? 0.0 : 0.5);
    const double gy = p.y / o.h - (axis == 1 ? 0.0 : 0.5);
    const double gz = p.z / o.h - (axis == 2 ? 0.0 : 0.5);
    const int x0 = (int)std::floor(gx), y0 = (int)std::floor(gy), z0 = (int)std::floor(gz);
    const double wx[2]{1.0 - (gx - x0), gx - x0}, wy[2]{1.0 - (gy - y0), gy - y0}, wz[2]{1.0 - (gz - z0), gz - z0};
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int c[3]{x0 + dx, y0 + dy, z0 + dz};
                if (face_constant(o, axis, c, b, nullptr)) continue;
                atomic_add(v + face_index(o, axis, c[0], c[1], c[2], b), adj * wx[dx] * wy[dy] * wz[dz]);
            }
}

// sample_velocity_value, operators.cu:242-244
V3 sample_velocity(const Oracle& o, const double* const u[3], V3 p) {
    return {sample_face(o, u[0], 0, p, o.vb).value, sample_face(o, u[1], 1, p, o.vb).value,
            sample_face(o, u[2], 2, p, o.vb).value};
}

// solid_at, operators.cu:246-254
bool solid_at(const Oracle& o, V3 p) {
    const double lx = o.nx * o.h, ly = o.ny * o.h, lz = o.nz * o.h;
    if (vperiodic(o.vb, 0)) p.x -= std::floor(p.x / lx) * lx;
    if (vperiodic(o.vb, 1)) p.y -= std::floor(p.y / ly) * ly;
    if (vperiodic(o.vb, 2)) p.z -= std::floor(p.z / lz) * lz;
    const int x = std::max(0, std::min(o.nx - 1, (int)std::floor(p.x / o.h)));
    const int y = std::max(0, std::min(o.ny - 1, (int)std::floor(p.y / o.h)));
    const int z = std::max(0, std::min(o.nz - 1, (int)std::floor(p.z / o.h)));
    return o.mask[cell(o, x, y, z)] != 0;
}

// trace_rk2, operators.cu:256-276: RK2 midpoint backtrace, clamp to the box on non-periodic
// axes (derivative 0 there), 8-step bisection toward the start when the end is inside a solid
// (derivative scaled by the bisection fraction, which is treated as constant).
Trace trace_rk2(const Oracle& o, V3 start, const double* const u[3]) {
    const V3 v0 = sample_velocity(o, u, start);
    const V3 mid{start.x - 0.5 * o.dt * v0.x, start.y - 0.5 * o.dt * v0.y, start.z - 0.5 * o.dt * v0.z};
    const V3 v1 = sample_velocity(o, u, mid);
    V3 raw{start.x - o.dt * v1.x, start.y - o.dt * v1.y, start.z - o.dt * v1.z};
    V3 d{1.0, 1.0, 1.0};
    const V3 mx{o.nx * o.h, o.ny * o.h, o.nz * o.h};
    if (!vperiodic(o.vb, 0) && (raw.x < 0.0 || raw.x > mx.x)) { raw.x = std::min(std::max(raw.x, 0.0), mx.x); d.x = 0.0; }
    if (!vperiodic(o.vb, 1) && (raw.y < 0.0 || raw.y > mx.y)) { raw.y = std::min(std::max(raw.y, 0.0), mx.y); d.y = 0.0; }
    if (!vperiodic(o.vb, 2) && (raw.z < 0.0 || raw.z > mx.z)) { raw.z = std::min(std::max(raw.z, 0.0), mx.z); d.z = 0.0; }
    if (!solid_at(o, raw)) return {raw, d};
    double lo = 0.0, hi = 1.0;
    for (int it = 0; it < 8; ++it) {
        const double f = 0.5 * (lo + hi);
        const V3 t{start.x + f * (raw.x - start.x), start.y + f * (raw.y - start.y), start.z + f * (raw.z - start.z)};
        if (solid_at(o, t)) hi = f;
        else lo = f;
    }
    return {{start.x + lo * (raw.x - start.x), start.y + lo * (raw.y - start.y), start.z + lo * (raw.z - start.z)},
            {lo * d.x, lo * d.y, lo * d.z}};
}

// operators.cu:278-284
inline V3 face_position(const Oracle& o, int axis, int x, int y, int z) {
    return {(x + (axis == 0 ? 0.0 : 0.5)) * o.h, (y + (axis == 1 ? 0.0 : 0.5)) * o.h, (z + (axis == 2 ? 0.0 : 0.5)) * o.h};
}
inline V3 cell_position(const Oracle& o, int x, int y, int z) { return {(x + 0.5) * o.h, (y + 0.5) * o.h, (z + 0.5) * o.h}; }

inline bool in_domain(const Oracle& o, int x, int y, int z) {
    return x >= 0 && x < o.nx && y >= 0 && y < o.ny && z >= 0 && z < o.nz;
}

// ---- operators ------------------------------------------------------------------------------
// source_forward_kernel, operators.cu:286-290
void source_forward(const Oracle& o, const double* s, const double* src, double* out) {
#pragma omp parallel for
    for (i64 i = 0; i < o.n; ++i) out[i] = o.mask[i] == 0 ? s[i] + o.dt * src[i] : 0.0;
}

// source_vjp_kernel, operators.cu:297-302
void source_vjp(const Oracle& o, const double* out_adj, double* s_adj, double* src_adj) {
#pragma omp parallel for
    for (i64 i = 0; i < o.n; ++i) {
        if (o.mask[i] != 0) continue;
        s_adj[i] += out_adj[i];
        if (src_adj) src_adj[i] += o.dt * out_adj[i];
    }
}

// buoyancy_forward_kernel, operators.cu:304-313. params = {ambient, alpha_rho, alpha_T, conf}.
void buoyancy_forward(const Oracle& o, const double* rho, const double* T, const double* const e[3],
                      const double* params, double* const f[3]) {
#pragma omp parallel for
    for (i64 i = 0; i < o.n; ++i) {
        if (o.mask[i] != 0) { f[0][i] = f[1][i] = f[2][i] = 0.0; continue; }
        f[0][i] = e[0][i];
        f[1][i] = e[1][i] + params[1] * rho[i] + params[2] * (T[i] - params[0]);
        f[2][i] = e[2][i];
    }
}

// buoyancy_vjp_kernel, operators.cu:326-337
void buoyancy_vjp(const Oracle& o, const double* rho, const double* T, const double* params,
                  const double* const fa[3], double* rho_adj, double* T_adj, double* const e_adj[3],
                  double* param_adj) {
    double amb = 0.0, ar = 0.0, at = 0.0;
#pragma omp parallel for reduction(+ : amb, ar, at)
    for (i64 i = 0; i < o.n; ++i) {
        if (o.mask[i] != 0) continue;
        if (e_adj) {
            e_adj[0][i] += fa[0][i];
            e_adj[1][i] += fa[1][i];
            e_adj[2][i] += fa[2][i];
        }
        rho_adj[i] += params[1] * fa[1][i];
        T_adj[i] += params[2] * fa[1][i];
        amb += -params[2] * fa[1][i];
        ar += rho[i] * fa[1][i];
        at += (T[i] - params[0]) * fa[1][i];
    }
    if (param_adj) {
        param_adj[0] += amb;
        param_adj[1] += ar;
        param_adj[2] += at;
    }
}

// ---- vorticity confinement (operators.cu:662-846, launchers :973-1002) ----------------------
// smooth_epsilon, operators.cu:12 (an FP32 constant there; its FP32 value is used here).
constexpr double kSmoothEpsilon = (double)1.0e-6F;

// centered_load / centered_scatter, operators.cu:662-674: neighbours clamp to the grid (never
// wrap, whatever the boundary modes).
inline i64 clamped_cell(const Oracle& o, int x, int y, int z) {
    x = std::max(0, std::min(o.nx - 1, x));
    y = std::max(0, std::min(o.ny - 1, y));
    z = std::max(0, std::min(o.nz - 1, z));
    return cell(o, x, y, z);
}

// Cached fields of vorticity_forward (the reference's VorticityCache): cell-centred velocity,
// vorticity, smoothed magnitude sqrt(|w|^2 + eps^2), normal N = grad|w| / len and its
// normaliser len = sqrt(|grad|w||^2 + eps^2). Solid cells: c = w = N = 0, magnitude = len = eps.
struct VorticityFields {
    std::vector<double> c[3], w[3], mag, nrm[3], len;
};

void vorticity_fields(const Oracle& o, const double* const u[3], VorticityFields& v) {
    const i64 n = o.n;
    for (int a = 0; a < 3; ++a) {
        v.c[a].assign(n, 0.0);
        v.w[a].assign(n, 0.0);
        v.nrm[a].assign(n, 0.0);
    }
    v.mag.assign(n, kSmoothEpsilon);
    v.len.assign(n, kSmoothEpsilon);
    // centered_velocity_kernel, operators.cu:678-687
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                v.c[0][i] = 0.5 * (u[0][idx3(x, y, z, o.nx + 1, o.ny)] + u[0][idx3(x + 1, y, z, o.nx + 1, o.ny)]);
                v.c[1][i] = 0.5 * (u[1][idx3(x, y, z, o.nx, o.ny + 1)] + u[1][idx3(x, y + 1, z, o.nx, o.ny + 1)]);
                v.c[2][i] = 0.5 * (u[2][idx3(x, y, z, o.nx, o.ny)] + u[2][idx3(x, y, z + 1, o.nx, o.ny)]);
            }
    const double s = 0.5 / o.h;
    auto C = [&](int a, int x, int y, int z) { return v.c[a][clamped_cell(o, x, y, z)]; };
    // curl_magnitude_kernel, operators.cu:689-701
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                const double wx = (C(2, x, y + 1, z) - C(2, x, y - 1, z) - C(1, x, y, z + 1) + C(1, x, y, z - 1)) * s;
                const double wy = (C(0, x, y, z + 1) - C(0, x, y, z - 1) - C(2, x + 1, y, z) + C(2, x - 1, y, z)) * s;
                const double wz = (C(1, x + 1, y, z) - C(1, x - 1, y, z) - C(0, x, y + 1, z) + C(0, x, y - 1, z)) * s;
                v.w[0][i] = wx;
                v.w[1][i] = wy;
                v.w[2][i] = wz;
                v.mag[i] = std::sqrt(wx * wx + wy * wy + wz * wz + kSmoothEpsilon * kSmoothEpsilon);
            }
    // vorticity_normal_kernel, operators.cu:703-720 (solid neighbours contribute their eps)
    auto M = [&](int x, int y, int z) { return v.mag[clamped_cell(o, x, y, z)]; };
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                const double gx = (M(x + 1, y, z) - M(x - 1, y, z)) * s;
                const double gy = (M(x, y + 1, z) - M(x, y - 1, z)) * s;
                const double gz = (M(x, y, z + 1) - M(x, y, z - 1)) * s;
                const double len = std::sqrt(gx * gx + gy * gy + gz * gz + kSmoothEpsilon * kSmoothEpsilon);
                v.nrm[0][i] = gx / len;
                v.nrm[1][i] = gy / len;
                v.nrm[2][i] = gz / len;
                v.len[i] = len;
            }
}

// vorticity_forward (operators.cu:973-979): f += confinement * h * (N x w) on fluid cells
// (add_vorticity_force_kernel, :722-729). u is the step's input velocity.
void vorticity_forward(const Oracle& o, const double* const u[3], double confinement, double* const f[3]) {
    VorticityFields v;
    vorticity_fields(o, u, v);
    const double scale = confinement * o.h;
    for (i64 i = 0; i < o.n; ++i) {
        if (o.mask[i] != 0) continue;
        const double N0 = v.nrm[0][i], N1 = v.nrm[1][i], N2 = v.nrm[2][i];
        const double w0 = v.w[0][i], w1 = v.w[1][i], w2 = v.w[2][i];
        f[0][i] += scale * (N1 * w2 - N2 * w1);
        f[1][i] += scale * (N2 * w0 - N0 * w2);
        f[2][i] += scale * (N0 * w1 - N1 * w0);
    }
}

// vorticity_vjp (operators.cu:990-1002): force reverse (:781-800), normal reverse (:802-819),
// magnitude reverse (:821-828), curl reverse (:830-844), centered-velocity reverse (:846-856).
// Accumulates into the velocity adjoint ua (faces) and conf_adj (one double).
void vorticity_vjp(const Oracle& o, const double* const u[3], double confinement, const double* const fa[3],
                   double* const ua[3], double* conf_adj) {
    VorticityFields v;
    vorticity_fields(o, u, v);
    const i64 n = o.n;
    std::vector<double> wa[3], na[3], ca[3], ma(n, 0.0);
    for (int a = 0; a < 3; ++a) {
        wa[a].assign(n, 0.0);
        na[a].assign(n, 0.0);
        ca[a].assign(n, 0.0);
    }
    const double scale = confinement * o.h;
    double ctot = 0.0;
    for (i64 i = 0; i < n; ++i) {
        if (o.mask[i] != 0) continue;
        const double N0 = v.nrm[0][i], N1 = v.nrm[1][i], N2 = v.nrm[2][i];
        const double w0 = v.w[0][i], w1 = v.w[1][i], w2 = v.w[2][i];
        const double f0 = fa[0][i], f1 = fa[1][i], f2 = fa[2][i];
        ctot += o.h * ((N1 * w2 - N2 * w1) * f0 + (N2 * w0 - N0 * w2) * f1 + (N0 * w1 - N1 * w0) * f2);
        na[0][i] += scale * (w1 * f2 - w2 * f1);   // w x f_bar
        na[1][i] += scale * (w2 * f0 - w0 * f2);
        na[2][i] += scale * (w0 * f1 - w1 * f0);
        wa[0][i] += scale * (f1 * N2 - f2 * N1);   // f_bar x N
        wa[1][i] += scale * (f2 * N0 - f0 * N2);
        wa[2][i] += scale * (f0 * N1 - f1 * N0);
    }
    if (conf_adj) *conf_adj += ctot;
    const double s = 0.5 / o.h;
    auto scatter = [&](std::vector<double>& dst, int x, int y, int z, double value) { dst[clamped_cell(o, x, y, z)] += value; };
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                const double proj = v.nrm[0][i] * na[0][i] + v.nrm[1][i] * na[1][i] + v.nrm[2][i] * na[2][i];
                const double gx = (na[0][i] - v.nrm[0][i] * proj) / v.len[i];
                const double gy = (na[1][i] - v.nrm[1][i] * proj) / v.len[i];
                const double gz = (na[2][i] - v.nrm[2][i] * proj) / v.len[i];
                scatter(ma, x + 1, y, z, s * gx);
                scatter(ma, x - 1, y, z, -s * gx);
                scatter(ma, x, y + 1, z, s * gy);
                scatter(ma, x, y - 1, z, -s * gy);
                scatter(ma, x, y, z + 1, s * gz);
                scatter(ma, x, y, z - 1, -s * gz);
            }
    for (i64 i = 0; i < n; ++i) {
        if (o.mask[i] != 0) continue;
        const double k = ma[i] / v.mag[i];
        for (int a = 0; a < 3; ++a) wa[a][i] += k * v.w[a][i];
    }
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                const double ax = wa[0][i] * s, ay = wa[1][i] * s, az = wa[2][i] * s;
                scatter(ca[2], x, y + 1, z, ax);
                scatter(ca[2], x, y - 1, z, -ax);
                scatter(ca[1], x, y, z + 1, -ax);
                scatter(ca[1], x, y, z - 1, ax);
                scatter(ca[0], x, y, z + 1, ay);
                scatter(ca[0], x, y, z - 1, -ay);
                scatter(ca[2], x + 1, y, z, -ay);
                scatter(ca[2], x - 1, y, z, ay);
                scatter(ca[1], x + 1, y, z, az);
                scatter(ca[1], x - 1, y, z, -az);
                scatter(ca[0], x, y + 1, z, -az);
                scatter(ca[0], x, y - 1, z, az);
            }
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                ua[0][idx3(x, y, z, o.nx + 1, o.ny)] += 0.5 * ca[0][i];
                ua[0][idx3(x + 1, y, z, o.nx + 1, o.ny)] += 0.5 * ca[0][i];
                ua[1][idx3(x, y, z, o.nx, o.ny + 1)] += 0.5 * ca[1][i];
                ua[1][idx3(x, y + 1, z, o.nx, o.ny + 1)] += 0.5 * ca[1][i];
                ua[2][idx3(x, y, z, o.nx, o.ny)] += 0.5 * ca[2][i];
                ua[2][idx3(x, y, z + 1, o.nx, o.ny)] += 0.5 * ca[2][i];
            }
}

// averaged_center_force + integrate_kernel, operators.cu:339-364
void integrate_forward(const Oracle& o, const double* const u[3], const double* const f[3], double* const out[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
#pragma omp parallel for
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const i64 fi = idx3(x, y, z, ex, ey);
                    double sum = 0.0, count = 0.0;
                    const int c0[3]{x - (a == 0), y - (a == 1), z - (a == 2)};
                    const int c1[3]{x, y, z};
                    for (const int* c : {c0, c1}) {
                        if (!in_domain(o, c[0], c[1], c[2])) continue;
                        const i64 ci = cell(o, c[0], c[1], c[2]);
                        if (o.mask[ci] == 0) { sum += f[a][ci]; count += 1.0; }
                    }
                    out[a][fi] = u[a][fi] + o.dt * (count == 0.0 ? 0.0 : sum / count);
                }
    }
}

// integrate_vjp_kernel, operators.cu:366-379
void integrate_vjp(const Oracle& o, const double* const oa[3], double* const ua[3], double* const fa[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
#pragma omp parallel for
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const i64 fi = idx3(x, y, z, ex, ey);
                    ua[a][fi] += oa[a][fi];
                    const int c0[3]{x - (a == 0), y - (a == 1), z - (a == 2)};
                    const int c1[3]{x, y, z};
                    int count = 0;
                    for (const int* c : {c0, c1})
                        if (in_domain(o, c[0], c[1], c[2]) && o.mask[cell(o, c[0], c[1], c[2])] == 0) ++count;
                    if (count == 0 || fa == nullptr) continue;
                    for (const int* c : {c0, c1})
                        if (in_domain(o, c[0], c[1], c[2]) && o.mask[cell(o, c[0], c[1], c[2])] == 0)
                            atomic_add(fa[a] + cell(o, c[0], c[1], c[2]), o.dt * oa[a][fi] / count);
                }
    }
}

// advect_velocity_forward_kernel, operators.cu:381-388
void advect_velocity_forward(const Oracle& o, const double* const u[3], double* const out[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
#pragma omp parallel for
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const Trace t = trace_rk2(o, face_position(o, a, x, y, z), u);
                    out[a][idx3(x, y, z, ex, ey)] = sample_face(o, u[a], a, t.position, o.vb).value;
                }
    }
}

// Shared reverse of trace_rk2 + final sample, operators.cu:415-439 and 537-561: scatter the
// output adjoint at the trace point, then the chain through the midpoint and start samples.
void trace_vjp(const Oracle& o, V3 start, const double* const u[3], const Trace& t, const V3& source_gradient,
               double out_adj, double* const ua[3]) {
    const V3 v0 = sample_velocity(o, u, start);
    const V3 mid{start.x - 0.5 * o.dt * v0.x, start.y - 0.5 * o.dt * v0.y, start.z - 0.5 * o.dt * v0.z};
    const double v1a[3]{-(t.derivative.x * o.dt * source_gradient.x) * out_adj,
                        -(t.derivative.y * o.dt * source_gradient.y) * out_adj,
                        -(t.derivative.z * o.dt * source_gradient.z) * out_adj};
    Sample ms[3];
    for (int c = 0; c < 3; ++c) ms[c] = sample_face(o, u[c], c, mid, o.vb);
    for (int c = 0; c < 3; ++c) scatter_face(o, ua[c], c, mid, v1a[c], o.vb);
    const double mx = ms[0].gradient.x * v1a[0] + ms[1].gradient.x * v1a[1] + ms[2].gradient.x * v1a[2];
    const double my = ms[0].gradient.y * v1a[0] + ms[1].gradient.y * v1a[1] + ms[2].gradient.y * v1a[2];
    const double mz = ms[0].gradient.z * v1a[0] + ms[1].gradient.z * v1a[1] + ms[2].gradient.z * v1a[2];
    scatter_face(o, ua[0], 0, start, -0.5 * o.dt * mx, o.vb);
    scatter_face(o, ua[1], 1, start, -0.5 * o.dt * my, o.vb);
    scatter_face(o, ua[2], 2, start, -0.5 * o.dt * mz, o.vb);
}

// advect_velocity_vjp_kernel, operators.cu:415-439
void advect_velocity_vjp(const Oracle& o, const double* const u[3], const double* const oa[3], double* const ua[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
#pragma omp parallel for
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const double adj = oa[a][idx3(x, y, z, ex, ey)];
                    if (adj == 0.0) continue;
                    const V3 start = face_position(o, a, x, y, z);
                    const Trace t = trace_rk2(o, start, u);
                    const Sample s = sample_face(o, u[a], a, t.position, o.vb);
                    scatter_face(o, ua[a], a, t.position, adj, o.vb);
                    trace_vjp(o, start, u, t, s.gradient, adj, ua);
                }
    }
}

// adjacent_solid / constrained_boundary / boundary_velocity_value, operators.cu:441-464
bool adjacent_solid(const Oracle& o, int a, int x, int y, int z) {
    const int c0[3]{x - (a == 0), y - (a == 1), z - (a == 2)};
    const int c1[3]{x, y, z};
    for (const int* c : {c0, c1})
        if (in_domain(o, c[0], c[1], c[2]) && o.mask[cell(o, c[0], c[1], c[2])] != 0) return true;
    return false;
}
bool constrained_boundary(const Oracle& o, int a, int x, int y, int z, double* value) {
    const int coord = a == 0 ? x : a == 1 ? y : z;
    const int maximum = size_of(o, a);
    if (coord != 0 && coord != maximum) return false;
    const int f = a * 2 + (coord == maximum);
    const std::uint32_t mode = o.vb.modes[f];
    if (value) *value = o.vb.values[f * 3 + a];
    return mode == 0u || mode == 2u;
}

// constrain_velocity_forward_kernel, operators.cu:466-481
void constrain_forward(const Oracle& o, const double* const u[3], double* const out[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
#pragma omp parallel for
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const i64 fi = idx3(x, y, z, ex, ey);
                    double bv = 0.0;
                    if (constrained_boundary(o, a, x, y, z, &bv)) { out[a][fi] = bv; continue; }
                    if (adjacent_solid(o, a, x, y, z)) { out[a][fi] = o.collider_velocity[a][fi]; continue; }
                    const int coord = a == 0 ? x : a == 1 ? y : z;
                    if (vperiodic(o.vb, a) && coord == size_of(o, a)) {
                        out[a][fi] = u[a][idx3(a == 0 ? 0 : x, a == 1 ? 0 : y, a == 2 ? 0 : z, ex, ey)];
                        continue;
                    }
                    out[a][fi] = u[a][fi];
                }
    }
}

// constrain_velocity_vjp_kernel, operators.cu:483-497
void constrain_vjp(const Oracle& o, const double* const oa[3], double* const ua[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const i64 fi = idx3(x, y, z, ex, ey);
                    if (constrained_boundary(o, a, x, y, z, nullptr) || adjacent_solid(o, a, x, y, z)) continue;
                    const int coord = a == 0 ? x : a == 1 ? y : z;
                    if (vperiodic(o.vb, a) && coord == size_of(o, a)) {
                        ua[a][idx3(a == 0 ? 0 : x, a == 1 ? 0 : y, a == 2 ? 0 : z, ex, ey)] += oa[a][fi];
                        continue;
                    }
                    ua[a][fi] += oa[a][fi];
                }
    }
}

const SB& scalar_bc(const Oracle& o, int field) { return field == 0 ? o.db : o.tb; }
const std::vector<double>& collider_value(const Oracle& o, int field) {
    return field == 0 ? o.collider_density : o.collider_temperature;
}

// advect_scalar_forward_kernel, operators.cu:499-507
void advect_scalar_forward(const Oracle& o, int field, const double* src, const double* const u[3], double* out) {
    const SB& b = scalar_bc(o, field);
    const std::vector<double>& cv = collider_value(o, field);
#pragma omp parallel for
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) { out[i] = cv[i]; continue; }
                const Trace t = trace_rk2(o, cell_position(o, x, y, z), u);
                out[i] = sample_scalar(o, src, t.position, b).value;
            }
}

// advect_scalar_vjp_kernel, operators.cu:537-561
void advect_scalar_vjp(const Oracle& o, int field, const double* src, const double* const u[3], const double* oa,
                       double* src_adj, double* const ua[3]) {
    const SB& b = scalar_bc(o, field);
#pragma omp parallel for
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0 || oa[i] == 0.0) continue;
                const V3 start = cell_position(o, x, y, z);
                const Trace t = trace_rk2(o, start, u);
                const Sample s = sample_scalar(o, src, t.position, b);
                scatter_scalar(o, src_adj, t.position, oa[i], b);
                trace_vjp(o, start, u, t, s.gradient, oa[i], ua);
            }
}

// ---- pressure system (operators.cu:848-902, model.cpp:72-77) --------------------------------
// pressure_neighbor, operators.cu:852-873. Returns the neighbour cell (or -1) and whether the
// direction is a fixed-value face (with its value).
inline i64 pressure_neighbor(const Oracle& o, int x, int y, int z, int d, int dir, bool& fixed, double& fixed_value) {
    int c[3]{x, y, z};
    c[d] += dir;
    fixed = false;
    const int size = size_of(o, d);
    if (c[d] < 0 || c[d] >= size) {
        if (speriodic(o.pb, d)) c[d] = wrap(c[d], size);
        else {
            const int f = d * 2 + (dir > 0);
            fixed = o.pb.modes[f] == 0u;
            fixed_value = o.pb.values[f];
            return -1;
        }
    }
    return cell(o, c[0], c[1], c[2]);
}

// pressure_rhs_kernel, operators.cu:875-883: b = -h^2 div(u)/dt on fluid non-anchor cells.
void pressure_rhs(const Oracle& o, const double* const u[3], double* rhs) {
#pragma omp parallel for
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0 || i == o.anchor) { rhs[i] = 0.0; continue; }
                const double div = (u[0][idx3(x + 1, y, z, o.nx + 1, o.ny)] - u[0][idx3(x, y, z, o.nx + 1, o.ny)] +
                                    u[1][idx3(x, y + 1, z, o.nx, o.ny + 1)] - u[1][idx3(x, y, z, o.nx, o.ny + 1)] +
                                    u[2][idx3(x, y, z + 1, o.nx, o.ny)] - u[2][idx3(x, y, z, o.nx, o.ny)]) / o.h;
                rhs[i] = -o.h * o.h * div / o.dt;
            }
}

// y = A x on unknowns, A_ii = diag (connected + fixed faces), A_ij = -1 for connected unknowns.
void apply_matrix(const Oracle& o, const double* xv, double* y) {
#pragma omp parallel for
    for (int z = 0; z < o.nz; ++z)
        for (int yy = 0; yy < o.ny; ++yy)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, yy, z);
                if (o.code[i] == 0) { y[i] = 0.0; continue; }
                double s = o.code[i] * xv[i];
                for (int d = 0; d < 3; ++d)
                    for (int dir = -1; dir <= 1; dir += 2) {
                        bool fixed;
                        double fv;
                        const i64 j = pressure_neighbor(o, x, yy, z, d, dir, fixed, fv);
                        if (j >= 0 && o.code[j] != 0) s -= xv[j];
                    }
                y[i] = s;
            }
}

double dot(const Oracle& o, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (i64 i = 0; i < o.n; ++i) s += a[i] * b[i];
    return s;
}

// Jacobi-PCG on A x = b (b zero off the unknowns). Stops when ||b - A x|| <= tol ||b||.
int pcg(const Oracle& o, const double* b, const double* x0, double* x, double tol, std::uint32_t max_it, ds_solve_info* info) {
    const i64 n = o.n;
    std::vector<double> r(n), z(n), p(n), q(n);
    for (i64 i = 0; i < n; ++i) x[i] = (x0 && o.code[i] != 0) ? x0[i] : 0.0;
    apply_matrix(o, x, q.data());
    for (i64 i = 0; i < n; ++i) r[i] = o.code[i] ? b[i] - q[i] : 0.0;
    const double bn = std::sqrt(dot(o, b, b));
    double rn = std::sqrt(dot(o, r.data(), r.data()));
    ds_solve_info local{};
    local.rhs_norm = bn;
    local.initial_residual = rn;
    local.warm_started = x0 != nullptr;
    std::uint32_t it = 0;
    if (bn == 0.0) {
        std::fill(x, x + n, 0.0);
        local.converged = 1;
        local.final_residual = 0.0;
        if (info) *info = local;
        return DS_OK;
    }
    for (i64 i = 0; i < n; ++i) z[i] = o.code[i] ? r[i] / o.code[i] : 0.0;
    p = z;
    double rz = dot(o, r.data(), z.data());
    while (rn > tol * bn && it < max_it) {
        apply_matrix(o, p.data(), q.data());
        const double alpha = rz / dot(o, p.data(), q.data());
#pragma omp parallel for
        for (i64 i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = o.code[i] ? r[i] / o.code[i] : 0.0;
        }
        ++it;
        const double rz_new = dot(o, r.data(), z.data());
        rn = std::sqrt(dot(o, r.data(), r.data()));
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for
        for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    apply_matrix(o, x, q.data());
    for (i64 i = 0; i < n; ++i) r[i] = o.code[i] ? b[i] - q[i] : 0.0;
    local.final_residual = std::sqrt(dot(o, r.data(), r.data()));
    local.iterations = it;
    local.converged = local.final_residual <= tol * bn * (1.0 + 1e-9) || rn <= tol * bn;
    if (info) *info = local;
    return DS_OK;
}

// rbgs_kernel, operators.cu:885-902 with the sweep loop of operators.cu:1004-1010 (reference mode).
void rbgs(const Oracle& o, const double* rhs, double* p, std::uint32_t iterations) {
    std::fill(p, p + o.n, 0.0);
    for (std::uint32_t it = 0; it < iterations; ++it)
        for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for
            for (int z = 0; z < o.nz; ++z)
                for (int y = 0; y < o.ny; ++y)
                    for (int x = 0; x < o.nx; ++x) {
                        const i64 i = cell(o, x, y, z);
                        if (((x + y + z) & 1) != parity || o.mask[i] != 0 || i == o.anchor) continue;
                        double sum = rhs[i], diag = 0.0;
                        for (int d = 0; d < 3; ++d)
                            for (int dir = -1; dir <= 1; dir += 2) {
                                bool fixed;
                                double fv = 0.0;
                                const i64 j = pressure_neighbor(o, x, y, z, d, dir, fixed, fv);
                                if (j >= 0 && o.mask[j] == 0) { sum += p[j]; diag += 1.0; }
                                else if (j < 0 && fixed) { sum += fv; diag += 1.0; }
                            }
                        p[i] = diag == 0.0 ? 0.0 : sum / diag;
                    }
        }
}

// projectable_face + project_velocity_kernel, operators.cu:904-920
bool projectable(const Oracle& o, int a, int x, int y, int z, i64& first, i64& second) {
    const int f[3]{x - (a == 0), y - (a == 1), z - (a == 2)};
    if (!in_domain(o, f[0], f[1], f[2]) || !in_domain(o, x, y, z)) return false;
    first = cell(o, f[0], f[1], f[2]);
    second = cell(o, x, y, z);
    return o.mask[first] == 0 && o.mask[second] == 0;
}

void project(const Oracle& o, const double* const u[3], const double* p, double* const out[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
#pragma omp parallel for
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const i64 fi = idx3(x, y, z, ex, ey);
                    i64 c0, c1;
                    if (!projectable(o, a, x, y, z, c0, c1)) { out[a][fi] = u[a][fi]; continue; }
                    out[a][fi] = u[a][fi] - o.dt * (p[c1] - p[c0]) / o.h;
                }
    }
}

// projection_forward, operators.cu:1004-1013, with the PCG solve (or RBGS in reference mode).
int projection_forward(Oracle& o, const double* const u[3], double* p, double* rhs, const double* guess,
                       double* const out[3], ds_solve_info* info) {
    pressure_rhs(o, u, rhs);
    if (o.projection_solver == DS_PROJECTION_RBGS) {
        rbgs(o, rhs, p, o.rbgs_iterations);
        if (info) *info = ds_solve_info{o.rbgs_iterations, 1, 0, 0, 0.0, 0.0, 0.0};
    } else {
        std::vector<double> b(o.n);
        for (i64 i = 0; i < o.n; ++i) b[i] = o.code[i] ? rhs[i] + o.fixed_term[i] : 0.0;
        const int st = pcg(o, b.data(), guess, p, o.tolerance, o.max_iterations, info);
        if (st != DS_OK) return st;
    }
    project(o, u, p, out);
    return DS_OK;
}

// projection_vjp, operators.cu:1020-1031 with the unrolled RBGS reverse replaced by the adjoint
// solve A lambda = p_bar (A symmetric): project reverse (operators.cu:922-933), solve, then
// pressure_rhs_reverse (operators.cu:957-969).
int projection_vjp(Oracle& o, const double* const oa[3], double* const ua[3], ds_solve_info* info) {
    std::vector<double> pbar(o.n, 0.0), lambda(o.n, 0.0);
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const i64 fi = idx3(x, y, z, ex, ey);
                    ua[a][fi] += oa[a][fi];
                    i64 c0, c1;
                    if (!projectable(o, a, x, y, z, c0, c1)) continue;
                    const double scale = o.dt * oa[a][fi] / o.h;
                    pbar[c0] += scale;
                    pbar[c1] -= scale;
                }
    }
    if (o.projection_solver == DS_PROJECTION_RBGS) {
        // Reverse of the unrolled sweeps (rbgs_reverse_kernel, operators.cu:935-955,1025-1028).
        std::vector<double>& pa = pbar;
        std::vector<double> ra(o.n, 0.0);
        for (std::uint32_t it = 0; it < o.rbgs_iterations; ++it)
            for (int parity = 1; parity >= 0; --parity) {
                for (int z = 0; z < o.nz; ++z)
                    for (int y = 0; y < o.ny; ++y)
                        for (int x = 0; x < o.nx; ++x) {
                            const i64 i = cell(o, x, y, z);
                            if (((x + y + z) & 1) != parity || o.mask[i] != 0 || i == o.anchor) continue;
                            double diag = 0.0;
                            i64 nb[6];
                            int cnt = 0;
                            for (int d = 0; d < 3; ++d)
                                for (int dir = -1; dir <= 1; dir += 2) {
                                    bool fixed;
                                    double fv = 0.0;
                                    const i64 j = pressure_neighbor(o, x, y, z, d, dir, fixed, fv);
                                    if (j >= 0 && o.mask[j] == 0) { nb[cnt++] = j; diag += 1.0; }
                                    else if (j < 0 && fixed) diag += 1.0;
                                }
                            const double adj = diag == 0.0 ? 0.0 : pa[i] / diag;
                            pa[i] = 0.0;
                            ra[i] += adj;
                            for (int k = 0; k < cnt; ++k) pa[nb[k]] += adj;
                        }
            }
        lambda = ra;
        if (info) *info = ds_solve_info{o.rbgs_iterations, 1, 0, 0, 0.0, 0.0, 0.0};
    } else {
        for (i64 i = 0; i < o.n; ++i)
            if (o.code[i] == 0) pbar[i] = 0.0;
        const int st = pcg(o, pbar.data(), nullptr, lambda.data(), o.tolerance, o.max_iterations, info);
        if (st != DS_OK) return st;
    }
    // pressure_rhs_reverse_kernel, operators.cu:957-969: only non-anchor fluid cells carry b.
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0 || i == o.anchor) continue;
                if (o.projection_solver == DS_PROJECTION_PCG && o.code[i] == 0) continue;
                const double scale = -o.h * lambda[i] / o.dt;
                ua[0][idx3(x + 1, y, z, o.nx + 1, o.ny)] += scale;
                ua[0][idx3(x, y, z, o.nx + 1, o.ny)] -= scale;
                ua[1][idx3(x, y + 1, z, o.nx, o.ny + 1)] += scale;
                ua[1][idx3(x, y, z, o.nx, o.ny + 1)] -= scale;
                ua[2][idx3(x, y, z + 1, o.nx, o.ny)] += scale;
                ua[2][idx3(x, y, z, o.nx, o.ny)] -= scale;
            }
    return DS_OK;
}

// ---- forward mode: the *_jvp kernels and the forwards they reuse ---------------------------
// The tangent of a fixed boundary value is 0: tangent samples and solves use a copy of the
// boundary with every value zeroed (operators.cu:396-397, 516-519, 633-638, 1015-1018).
VB zero_values(const VB& b) {
    VB t = b;
    for (double& v : t.values) v = 0.0;
    return t;
}
SB zero_values(const SB& b) {
    SB t = b;
    for (double& v : t.values) v = 0.0;
    return t;
}

V3 sample_velocity_in(const Oracle& o, const double* const u[3], V3 p, const VB& b) {
    return {sample_face(o, u[0], 0, p, b).value, sample_face(o, u[1], 1, p, b).value, sample_face(o, u[2], 2, p, b).value};
}

// buoyancy_jvp_kernel, operators.cu:315-324. pt = parameter tangents (same order as params).
void buoyancy_jvp(const Oracle& o, const double* rho, const double* T, const double* rho_t, const double* T_t,
                  const double* const e_t[3], const double* params, const double* pt, double* const f_t[3]) {
    for (i64 i = 0; i < o.n; ++i) {
        if (o.mask[i] != 0) { f_t[0][i] = f_t[1][i] = f_t[2][i] = 0.0; continue; }
        f_t[0][i] = e_t[0][i];
        f_t[1][i] = e_t[1][i] + params[1] * rho_t[i] + pt[1] * rho[i] + params[2] * (T_t[i] - pt[0]) + pt[2] * (T[i] - params[0]);
        f_t[2][i] = e_t[2][i];
    }
}

// vorticity_jvp (operators.cu:981-988): the tangent vorticity is the curl of the centred tangent
// velocity (the same linear kernels), magnitude_tangent_kernel (:731-735), normal_tangent_kernel
// (:737-754) and vorticity_force_tangent_kernel (:756-779). Adds to f_t.
void vorticity_jvp(const Oracle& o, const double* const u[3], const double* const u_t[3], double confinement,
                   double confinement_t, double* const f_t[3]) {
    VorticityFields v, vt;
    vorticity_fields(o, u, v);
    vorticity_fields(o, u_t, vt);   // vt.w = curl of the centred tangent velocity
    std::vector<double> m_t(o.n, 0.0);
    for (i64 i = 0; i < o.n; ++i)
        if (o.mask[i] == 0) m_t[i] = (v.w[0][i] * vt.w[0][i] + v.w[1][i] * vt.w[1][i] + v.w[2][i] * vt.w[2][i]) / v.mag[i];
    const double s = 0.5 / o.h;
    auto M = [&](int x, int y, int z) { return m_t[clamped_cell(o, x, y, z)]; };
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                const double gt[3]{(M(x + 1, y, z) - M(x - 1, y, z)) * s, (M(x, y + 1, z) - M(x, y - 1, z)) * s,
                                   (M(x, y, z + 1) - M(x, y, z - 1)) * s};
                const double N[3]{v.nrm[0][i], v.nrm[1][i], v.nrm[2][i]};
                const double proj = N[0] * gt[0] + N[1] * gt[1] + N[2] * gt[2];
                const double Nt[3]{(gt[0] - N[0] * proj) / v.len[i], (gt[1] - N[1] * proj) / v.len[i], (gt[2] - N[2] * proj) / v.len[i]};
                const double w[3]{v.w[0][i], v.w[1][i], v.w[2][i]};
                const double wt[3]{vt.w[0][i], vt.w[1][i], vt.w[2][i]};
                const double cross[3]{N[1] * w[2] - N[2] * w[1], N[2] * w[0] - N[0] * w[2], N[0] * w[1] - N[1] * w[0]};
                const double cross_t[3]{Nt[1] * w[2] + N[1] * wt[2] - Nt[2] * w[1] - N[2] * wt[1],
                                        Nt[2] * w[0] + N[2] * wt[0] - Nt[0] * w[2] - N[0] * wt[2],
                                        Nt[0] * w[1] + N[0] * wt[1] - Nt[1] * w[0] - N[1] * wt[0]};
                for (int a = 0; a < 3; ++a) f_t[a][i] += o.h * (confinement_t * cross[a] + confinement * cross_t[a]);
            }
}

// Tangent of the RK2 velocity at the trace's midpoint and the end-point tangent (the shared part
// of advect_velocity_jvp_kernel, operators.cu:390-413, and advect_scalar_jvp_kernel, :509-535).
V3 trace_tangent(const Oracle& o, V3 start, const double* const u[3], const double* const u_t[3], const VB& tb, Trace& trace) {
    const V3 v0 = sample_velocity_in(o, u, start, o.vb);
    const V3 t0 = sample_velocity_in(o, u_t, start, tb);
    const V3 mid{start.x - 0.5 * o.dt * v0.x, start.y - 0.5 * o.dt * v0.y, start.z - 0.5 * o.dt * v0.z};
    const V3 mid_t{-0.5 * o.dt * t0.x, -0.5 * o.dt * t0.y, -0.5 * o.dt * t0.z};
    const Sample s[3]{sample_face(o, u[0], 0, mid, o.vb), sample_face(o, u[1], 1, mid, o.vb), sample_face(o, u[2], 2, mid, o.vb)};
    const V3 st = sample_velocity_in(o, u_t, mid, tb);
    const double v1t[3]{st.x + s[0].gradient.x * mid_t.x + s[0].gradient.y * mid_t.y + s[0].gradient.z * mid_t.z,
                        st.y + s[1].gradient.x * mid_t.x + s[1].gradient.y * mid_t.y + s[1].gradient.z * mid_t.z,
                        st.z + s[2].gradient.x * mid_t.x + s[2].gradient.y * mid_t.y + s[2].gradient.z * mid_t.z};
    trace = trace_rk2(o, start, u);
    return {-trace.derivative.x * o.dt * v1t[0], -trace.derivative.y * o.dt * v1t[1], -trace.derivative.z * o.dt * v1t[2]};
}

// advect_velocity_jvp_kernel, operators.cu:390-413 (every face, like the forward).
void advect_velocity_jvp(const Oracle& o, const double* const u[3], const double* const u_t[3], double* const out_t[3]) {
    const VB tb = zero_values(o.vb);
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
#pragma omp parallel for
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    Trace t;
                    const V3 pt = trace_tangent(o, face_position(o, a, x, y, z), u, u_t, tb, t);
                    const Sample src = sample_face(o, u[a], a, t.position, o.vb);
                    out_t[a][idx3(x, y, z, ex, ey)] = sample_face(o, u_t[a], a, t.position, tb).value + src.gradient.x * pt.x +
                                                      src.gradient.y * pt.y + src.gradient.z * pt.z;
                }
    }
}

// constrain_velocity_jvp, operators.cu:633-638: the constraint with zero boundary values and no
// collider velocity.
void constrain_jvp(const Oracle& o, const double* const u_t[3], double* const out_t[3]) {
    for (int a = 0; a < 3; ++a) {
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    const i64 fi = idx3(x, y, z, ex, ey);
                    double bv = 0.0;
                    if (constrained_boundary(o, a, x, y, z, &bv) || adjacent_solid(o, a, x, y, z)) { out_t[a][fi] = 0.0; continue; }
                    const int coord = a == 0 ? x : a == 1 ? y : z;
                    if (vperiodic(o.vb, a) && coord == size_of(o, a)) {
                        out_t[a][fi] = u_t[a][idx3(a == 0 ? 0 : x, a == 1 ? 0 : y, a == 2 ? 0 : z, ex, ey)];
                        continue;
                    }
                    out_t[a][fi] = u_t[a][fi];
                }
    }
}

// advect_scalar_jvp_kernel, operators.cu:509-535 (solid cells: 0).
void advect_scalar_jvp(const Oracle& o, int field, const double* src, const double* src_t, const double* const u[3],
                       const double* const u_t[3], double* out_t) {
    const SB& b = scalar_bc(o, field);
    const SB sbt = zero_values(b);
    const VB tb = zero_values(o.vb);
#pragma omp parallel for
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) { out_t[i] = 0.0; continue; }
                Trace t;
                const V3 pt = trace_tangent(o, cell_position(o, x, y, z), u, u_t, tb, t);
                const Sample s = sample_scalar(o, src, t.position, b);
                out_t[i] = sample_scalar(o, src_t, t.position, sbt).value + s.gradient.x * pt.x + s.gradient.y * pt.y + s.gradient.z * pt.z;
            }
}

// projection_jvp, operators.cu:1015-1018: the projection of the tangent with zero pressure
// boundary values (PCG mode: the right-hand side without the fixed-face term; RBGS mode: the
// sweeps with zeroed boundary values).
int projection_jvp(Oracle& o, const double* const u_t[3], double* const out_t[3], ds_solve_info* info) {
    std::vector<double> rhs(o.n), p(o.n, 0.0);
    pressure_rhs(o, u_t, rhs.data());
    if (o.projection_solver == DS_PROJECTION_RBGS) {
        const SB saved = o.pb;
        o.pb = zero_values(o.pb);
        rbgs(o, rhs.data(), p.data(), o.rbgs_iterations);
        o.pb = saved;
        if (info) *info = ds_solve_info{o.rbgs_iterations, 1, 0, 0, 0.0, 0.0, 0.0};
    } else {
        for (i64 i = 0; i < o.n; ++i)
            if (!o.code[i]) rhs[i] = 0.0;
        const int st = pcg(o, rhs.data(), nullptr, p.data(), o.tolerance, o.max_iterations, info);
        if (st != DS_OK) return st;
    }
    project(o, u_t, p.data(), out_t);
    return DS_OK;
}

// ---- model steps (model.cpp:341-389) --------------------------------------------------------
// state/next: {rho, T, ux, uy, uz}; control: {s_rho, s_T, ex, ey, ez}; params: {ambient,
// alpha_rho, alpha_T, confinement}; cache: {rho_s, T_s, fx, fy, fz, u*x, u*y, u*z, u^x, u^y, u^z,
// pressure} (any of force / advected / pressure may be null).
struct Fields {
    std::vector<double> v[12];
};

int forward_step(Oracle& o, const double* const s[5], const double* const c[5], const double* params,
                 double* const next[5], double* const cache[12], const double* guess, ds_solve_info* info) {
    const i64 n = o.n;
    std::vector<double> tmp_cell[6];
    std::vector<double> tmp_face[6];
    auto cell_buf = [&](int k, double* given) -> double* {
        if (given) return given;
        tmp_cell[k].assign(n, 0.0);
        return tmp_cell[k].data();
    };
    auto face_buf = [&](int k, int a, double* given) -> double* {
        if (given) return given;
        tmp_face[k].assign(o.nf[a], 0.0);
        return tmp_face[k].data();
    };
    double* rs = cell_buf(0, cache[0]);
    double* ts = cell_buf(1, cache[1]);
    double* f[3]{cell_buf(2, cache[2]), cell_buf(3, cache[3]), cell_buf(4, cache[4])};
    double* us[3]{face_buf(0, 0, cache[5]), face_buf(1, 1, cache[6]), face_buf(2, 2, cache[7])};
    double* ua[3]{face_buf(3, 0, cache[8]), face_buf(4, 1, cache[9]), face_buf(5, 2, cache[10])};
    std::vector<double> raw[3];
    for (int a = 0; a < 3; ++a) raw[a].assign(o.nf[a], 0.0);
    double* rawp[3]{raw[0].data(), raw[1].data(), raw[2].data()};
    std::vector<double> p_local;
    double* p = cache[11];
    if (!p) {
        p_local.assign(n, 0.0);
        p = p_local.data();
    }
    std::vector<double> rhs(n);
    source_forward(o, s[0], c[0], rs);
    source_forward(o, s[1], c[1], ts);
    const double* e[3]{c[2], c[3], c[4]};
    buoyancy_forward(o, rs, ts, e, params, f);
    const double* u[3]{s[2], s[3], s[4]};
    if (o.vorticity) vorticity_forward(o, u, params[3], f);   // ForceAssemblyOperator, operators.cpp:125-128
    const double* fc[3]{f[0], f[1], f[2]};
    integrate_forward(o, u, fc, us);
    const double* usc[3]{us[0], us[1], us[2]};
    advect_velocity_forward(o, usc, rawp);
    const double* rawc[3]{rawp[0], rawp[1], rawp[2]};
    constrain_forward(o, rawc, ua);
    const double* uac[3]{ua[0], ua[1], ua[2]};
    double* nv[3]{next[2], next[3], next[4]};
    const int st = projection_forward(o, uac, p, rhs.data(), guess, nv, info);
    if (st != DS_OK) return st;
    const double* nvc[3]{next[2], next[3], next[4]};
    advect_scalar_forward(o, 0, rs, nvc, next[0]);
    advect_scalar_forward(o, 1, ts, nvc, next[1]);
    return DS_OK;
}

// Model::vjp_step, model.cpp:361-389 (reverse order of forward_step). state (the step's input) is
// read only for its velocity, and only with vorticity confinement on.
int vjp_step(Oracle& o, const double* const state[5], const double* const next[5], const double* const cache[12],
             const double* params, const double* const next_adj[5], double* const prev_adj[5], double* const control_adj[5],
             double* param_adj, ds_solve_info* info) {
    if (o.vorticity && (!state || !state[2] || !state[3] || !state[4])) {
        g_error = "vjp_step: vorticity confinement needs the step's input velocity";
        return DS_ERROR_INVALID_ARGUMENT;
    }
    const i64 n = o.n;
    std::vector<double> rsa(n, 0.0), tsa(n, 0.0), fa[3], pua[3], aua[3], rua[3], fua[3];
    for (int a = 0; a < 3; ++a) {
        fa[a].assign(n, 0.0);
        pua[a].assign(next_adj[2 + a], next_adj[2 + a] + o.nf[a]);
        aua[a].assign(o.nf[a], 0.0);
        rua[a].assign(o.nf[a], 0.0);
        fua[a].assign(o.nf[a], 0.0);
    }
    const double* nv[3]{next[2], next[3], next[4]};
    double* puap[3]{pua[0].data(), pua[1].data(), pua[2].data()};
    advect_scalar_vjp(o, 0, cache[0], nv, next_adj[0], rsa.data(), puap);
    advect_scalar_vjp(o, 1, cache[1], nv, next_adj[1], tsa.data(), puap);
    const double* puac[3]{puap[0], puap[1], puap[2]};
    double* auap[3]{aua[0].data(), aua[1].data(), aua[2].data()};
    const int st = projection_vjp(o, puac, auap, info);
    if (st != DS_OK) return st;
    const double* auac[3]{auap[0], auap[1], auap[2]};
    double* ruap[3]{rua[0].data(), rua[1].data(), rua[2].data()};
    constrain_vjp(o, auac, ruap);
    const double* ruac[3]{ruap[0], ruap[1], ruap[2]};
    double* fuap[3]{fua[0].data(), fua[1].data(), fua[2].data()};
    const double* us[3]{cache[5], cache[6], cache[7]};
    advect_velocity_vjp(o, us, ruac, fuap);
    const double* fuac[3]{fuap[0], fuap[1], fuap[2]};
    double* pv[3]{prev_adj[2], prev_adj[3], prev_adj[4]};
    double* fap[3]{fa[0].data(), fa[1].data(), fa[2].data()};
    integrate_vjp(o, fuac, pv, fap);
    const double* fac[3]{fap[0], fap[1], fap[2]};
    double* ea[3]{control_adj ? control_adj[2] : nullptr, control_adj ? control_adj[3] : nullptr,
                  control_adj ? control_adj[4] : nullptr};
    std::vector<double> e_sink[3];
    if (!control_adj)
        for (int a = 0; a < 3; ++a) {
            e_sink[a].assign(n, 0.0);
            ea[a] = e_sink[a].data();
        }
    buoyancy_vjp(o, cache[0], cache[1], params, fac, rsa.data(), tsa.data(), ea, param_adj);
    if (o.vorticity) {   // ForceAssemblyOperator::vjp, operators.cpp:135-138
        const double* u[3]{state[2], state[3], state[4]};
        vorticity_vjp(o, u, params[3], fac, pv, param_adj ? param_adj + 3 : nullptr);
    }
    source_vjp(o, rsa.data(), prev_adj[0], control_adj ? control_adj[0] : nullptr);
    source_vjp(o, tsa.data(), prev_adj[1], control_adj ? control_adj[1] : nullptr);
    return DS_OK;
}

// Model::jvp_step, model.cpp:351-359. state_t / control_t / params_t are the tangents; state,
// next and cache come from forward_step on the same inputs (cache: sourced fields and forced
// velocity). Writes next_t.
int jvp_step(Oracle& o, const double* const state[5], const double* const next[5], const double* const cache[12],
             const double* params, const double* const state_t[5], const double* const control_t[5], const double* params_t,
             double* const next_t[5], ds_solve_info* info) {
    const i64 n = o.n;
    std::vector<double> rs_t(n), ts_t(n), f_t[3], us_t[3], raw_t[3], ua_t[3];
    for (int a = 0; a < 3; ++a) {
        f_t[a].assign(n, 0.0);
        us_t[a].assign(o.nf[a], 0.0);
        raw_t[a].assign(o.nf[a], 0.0);
        ua_t[a].assign(o.nf[a], 0.0);
    }
    source_forward(o, state_t[0], control_t[0], rs_t.data());   // source_jvp = the forward on tangents
    source_forward(o, state_t[1], control_t[1], ts_t.data());
    double* fp[3]{f_t[0].data(), f_t[1].data(), f_t[2].data()};
    const double* e_t[3]{control_t[2], control_t[3], control_t[4]};
    buoyancy_jvp(o, cache[0], cache[1], rs_t.data(), ts_t.data(), e_t, params, params_t, fp);
    const double* ut[3]{state_t[2], state_t[3], state_t[4]};
    if (o.vorticity) {
        const double* u[3]{state[2], state[3], state[4]};
        vorticity_jvp(o, u, ut, params[3], params_t[3], fp);
    }
    const double* fc[3]{fp[0], fp[1], fp[2]};
    double* usp[3]{us_t[0].data(), us_t[1].data(), us_t[2].data()};
    integrate_forward(o, ut, fc, usp);   // integrate_velocity_jvp = the forward on tangents
    const double* us[3]{cache[5], cache[6], cache[7]};
    const double* usc[3]{usp[0], usp[1], usp[2]};
    double* rawp[3]{raw_t[0].data(), raw_t[1].data(), raw_t[2].data()};
    advect_velocity_jvp(o, us, usc, rawp);
    const double* rawc[3]{rawp[0], rawp[1], rawp[2]};
    double* uap[3]{ua_t[0].data(), ua_t[1].data(), ua_t[2].data()};
    constrain_jvp(o, rawc, uap);
    const double* uac[3]{uap[0], uap[1], uap[2]};
    double* nv_t[3]{next_t[2], next_t[3], next_t[4]};
    const int st = projection_jvp(o, uac, nv_t, info);
    if (st != DS_OK) return st;
    const double* nv[3]{next[2], next[3], next[4]};
    const double* nvt[3]{next_t[2], next_t[3], next_t[4]};
    advect_scalar_jvp(o, 0, cache[0], rs_t.data(), nv, nvt, next_t[0]);
    advect_scalar_jvp(o, 1, cache[1], ts_t.data(), nv, nvt, next_t[1]);
    return DS_OK;
}

void record(ds_trajectory_info* ti, const ds_solve_info& si, int kind) {
    if (!ti) return;
    if (kind == 0) { ti->forward_solves++; ti->forward_iterations += si.iterations; }
    if (kind == 1) { ti->recompute_solves++; ti->recompute_iterations += si.iterations; }
    if (kind == 2) { ti->adjoint_solves++; ti->adjoint_iterations += si.iterations; }
    ti->max_iterations = std::max(ti->max_iterations, si.iterations);
    if (!si.converged) ti->not_converged++;
    if (si.rhs_norm > 0.0) ti->max_relative_residual = std::max(ti->max_relative_residual, si.final_residual / si.rhs_norm);
}

// solver::vjp (solver.ixx:115-150) with a checkpoint every k steps: the forward pass keeps
// states 0, k, 2k, ...; the reverse sweep recomputes one segment's states and caches from its
// checkpoint, then runs vjp_step backwards over it. The loss is 0.5 ||rho_T - target||^2.
// solver::jvp (solver.ixx:83-113) in one forward sweep: each step runs forward_step and then
// jvp_step on its cache, so no trajectory is stored. The control and its tangent are the same at
// every step. Writes the final state and the final state tangent.
int trajectory_jvp(Oracle& o, const double* const init[5], const double* const control[5], const double* params,
                   std::uint32_t steps, const double* const init_t[5], const double* const control_t[5], const double* params_t,
                   double* const final_state[5], double* const final_t[5], ds_trajectory_info* ti) {
    const i64 n = o.n;
    auto alloc_state = [&]() {
        std::vector<std::vector<double>> s(5);
        s[0].assign(n, 0.0);
        s[1].assign(n, 0.0);
        for (int a = 0; a < 3; ++a) s[2 + a].assign(o.nf[a], 0.0);
        return s;
    };
    auto ptrs = [](std::vector<std::vector<double>>& s, double** out) {
        for (int i = 0; i < (int)s.size(); ++i) out[i] = s[i].data();
    };
    auto cur = alloc_state(), cur_t = alloc_state();
    for (int i = 0; i < 5; ++i) {
        std::copy(init[i], init[i] + cur[i].size(), cur[i].begin());
        std::copy(init_t[i], init_t[i] + cur_t[i].size(), cur_t[i].begin());
    }
    if (ti) {
        *ti = ds_trajectory_info{};
        ti->steps = steps;
    }
    std::vector<double> c_rs(n), c_ts(n), c_us[3];
    for (int a = 0; a < 3; ++a) c_us[a].assign(o.nf[a], 0.0);
    for (std::uint32_t s = 0; s < steps; ++s) {
        auto nxt = alloc_state(), nxt_t = alloc_state();
        double *sp[5], *np[5], *tp[5], *ntp[5];
        ptrs(cur, sp);
        ptrs(nxt, np);
        ptrs(cur_t, tp);
        ptrs(nxt_t, ntp);
        double* cache[12]{};
        cache[0] = c_rs.data();
        cache[1] = c_ts.data();
        for (int a = 0; a < 3; ++a) cache[5 + a] = c_us[a].data();
        ds_solve_info si{};
        const double* spc[5]{sp[0], sp[1], sp[2], sp[3], sp[4]};
        int st = forward_step(o, spc, control, params, np, cache, nullptr, &si);
        if (st != DS_OK) return st;
        record(ti, si, 0);
        const double* cc[12]{};
        for (int i = 0; i < 12; ++i) cc[i] = cache[i];
        const double* npc[5]{np[0], np[1], np[2], np[3], np[4]};
        const double* tpc[5]{tp[0], tp[1], tp[2], tp[3], tp[4]};
        st = jvp_step(o, spc, npc, cc, params, tpc, control_t, params_t, ntp, &si);
        if (st != DS_OK) return st;
        record(ti, si, 2);   // the linearised (tangent) solves are reported as the adjoint ones
        cur = std::move(nxt);
        cur_t = std::move(nxt_t);
    }
    for (int i = 0; i < 5; ++i) {
        std::copy(cur[i].begin(), cur[i].end(), final_state[i]);
        std::copy(cur_t[i].begin(), cur_t[i].end(), final_t[i]);
    }
    return DS_OK;
}

int loss_gradient(Oracle& o, const double* const init[5], const double* const control[5], const double* params,
                  std::uint32_t steps, std::uint32_t k, const double* target, double* const final_state[5],
                  double* loss, double* const init_adj[5], double* const control_adj[5], double* param_adj,
                  ds_trajectory_info* ti) {
    if (k == 0) k = 1;
    const i64 n = o.n;
    auto alloc_state = [&]() {
        std::vector<std::vector<double>> s(5);
        s[0].assign(n, 0.0);
        s[1].assign(n, 0.0);
        for (int a = 0; a < 3; ++a) s[2 + a].assign(o.nf[a], 0.0);
        return s;
    };
    auto ptrs = [](std::vector<std::vector<double>>& s, double** out) {
        for (int i = 0; i < (int)s.size(); ++i) out[i] = s[i].data();
    };
    std::vector<std::vector<std::vector<double>>> checkpoints;
    std::vector<std::vector<std::vector<double>>> pressure_tape(steps);
    auto cur = alloc_state();
    for (int i = 0; i < 5; ++i) std::copy(init[i], init[i] + cur[i].size(), cur[i].begin());
    if (ti) {
        *ti = ds_trajectory_info{};
        ti->steps = steps;
        ti->checkpoint_interval = k;
    }
    for (std::uint32_t s = 0; s < steps; ++s) {
        if (s % k == 0) checkpoints.push_back(cur);
        auto nxt = alloc_state();
        double* sp[5];
        double* np[5];
        ptrs(cur, sp);
        ptrs(nxt, np);
        std::vector<double> pr(n, 0.0);
        double* cache[12]{};
        cache[11] = pr.data();
        ds_solve_info si{};
        const double* spc[5]{sp[0], sp[1], sp[2], sp[3], sp[4]};
        const int st = forward_step(o, spc, control, params, np, cache, nullptr, &si);
        if (st != DS_OK) return st;
        record(ti, si, 0);
        pressure_tape[s].push_back(std::move(pr));
        cur = std::move(nxt);
    }
    double l = 0.0;
    std::vector<double> rho_adj(n);
    for (i64 i = 0; i < n; ++i) {
        const double d = cur[0][i] - target[i];
        l += 0.5 * d * d;
        rho_adj[i] = d;
    }
    if (loss) *loss = l;
    if (ti) ti->loss = l;
    if (final_state)
        for (int i = 0; i < 5; ++i)
            if (final_state[i]) std::copy(cur[i].begin(), cur[i].end(), final_state[i]);
    auto running = alloc_state();
    running[0] = rho_adj;
    const std::uint32_t segments = (steps + k - 1) / k;
    for (std::uint32_t seg = segments; seg-- > 0;) {
        const std::uint32_t s0 = seg * k, s1 = std::min(steps, s0 + k);
        std::vector<std::vector<std::vector<double>>> states;
        std::vector<std::vector<std::vector<double>>> caches;
        states.push_back(checkpoints[seg]);
        for (std::uint32_t s = s0; s < s1; ++s) {
            auto nxt = alloc_state();
            std::vector<std::vector<double>> cache(12);
            cache[0].assign(n, 0.0);
            cache[1].assign(n, 0.0);
            for (int a = 0; a < 3; ++a) cache[5 + a].assign(o.nf[a], 0.0);
            cache[11].assign(n, 0.0);
            double* sp[5];
            double* np[5];
            ptrs(states.back(), sp);
            ptrs(nxt, np);
            double* cp[12]{};
            for (int i = 0; i < 12; ++i) cp[i] = cache[i].empty() ? nullptr : cache[i].data();
            ds_solve_info si{};
            const double* spc[5]{sp[0], sp[1], sp[2], sp[3], sp[4]};
            const int st = forward_step(o, spc, control, params, np, cp, pressure_tape[s][0].data(), &si);
            if (st != DS_OK) return st;
            record(ti, si, 1);
            states.push_back(std::move(nxt));
            caches.push_back(std::move(cache));
        }
        for (std::uint32_t s = s1; s-- > s0;) {
            auto prev = alloc_state();
            double* np[5];
            double* ra[5];
            double* pa[5];
            ptrs(states[s - s0 + 1], np);
            ptrs(running, ra);
            ptrs(prev, pa);
            const double* cp[12]{};
            for (int i = 0; i < 12; ++i) cp[i] = caches[s - s0][i].empty() ? nullptr : caches[s - s0][i].data();
            ds_solve_info si{};
            const double* npc[5]{np[0], np[1], np[2], np[3], np[4]};
            const double* rac[5]{ra[0], ra[1], ra[2], ra[3], ra[4]};
            double* sp[5];
            ptrs(states[s - s0], sp);
            const double* spc[5]{sp[0], sp[1], sp[2], sp[3], sp[4]};
            const int st = vjp_step(o, spc, npc, cp, params, rac, pa, control_adj, param_adj, &si);
            if (st != DS_OK) return st;
            record(ti, si, 2);
            running = std::move(prev);
        }
    }
    for (int i = 0; i < 5; ++i)
        if (init_adj[i])
            for (std::size_t j = 0; j < running[i].size(); ++j) init_adj[i][j] += running[i][j];
    return DS_OK;
}

// ---- domain set-up (model.cpp:17-182) -------------------------------------------------------
bool contains(const ds_collider& c, float px, float py, float pz) {
    if (c.shape == DS_SHAPE_ELLIPSOID) {   // model.cpp:17-22
        const float x = (px - c.center[0]) / c.size[0];
        const float y = (py - c.center[1]) / c.size[1];
        const float z = (pz - c.center[2]) / c.size[2];
        return x * x + y * y + z * z <= 1.0F;
    }
    return std::fabs(px - c.center[0]) <= c.size[0] && std::fabs(py - c.center[1]) <= c.size[1] &&   // :24-26
           std::fabs(pz - c.center[2]) <= c.size[2];
}

int build(Oracle& o, const ds_config& cfg) {
    if (cfg.resolution[0] == 0 || cfg.resolution[1] == 0 || cfg.resolution[2] == 0) {
        g_error = "resolution must be positive";
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!(cfg.cell_size > 0.0F) || !(cfg.time_step > 0.0F)) {
        g_error = "cell_size and time_step must be positive";
        return DS_ERROR_INVALID_ARGUMENT;
    }
    o.nx = (int)cfg.resolution[0];
    o.ny = (int)cfg.resolution[1];
    o.nz = (int)cfg.resolution[2];
    o.hf = cfg.cell_size;
    o.h = cfg.cell_size;
    o.dt = cfg.time_step;
    for (int f = 0; f < 6; ++f) {
        o.vb.modes[f] = cfg.velocity_boundary.modes[f];
        o.pb.modes[f] = cfg.pressure_boundary.modes[f];
        o.db.modes[f] = cfg.density_boundary.modes[f];
        o.tb.modes[f] = cfg.temperature_boundary.modes[f];
        o.pb.values[f] = cfg.pressure_boundary.values[f];
        o.db.values[f] = cfg.density_boundary.values[f];
        o.tb.values[f] = cfg.temperature_boundary.values[f];
    }
    for (int k = 0; k < 18; ++k) o.vb.values[k] = cfg.velocity_boundary.values[k];
    o.projection_solver = cfg.projection_solver;
    o.rbgs_iterations = cfg.pressure_iterations;
    o.vorticity = cfg.vorticity_confinement_enabled != 0;
    o.tolerance = cfg.cg_tolerance > 0.0 ? cfg.cg_tolerance : 1e-6;
    o.max_iterations = cfg.cg_max_iterations ? cfg.cg_max_iterations : 4u * (cfg.resolution[0] + cfg.resolution[1] + cfg.resolution[2]) + 1000u;
    o.n = (i64)o.nx * o.ny * o.nz;
    for (int a = 0; a < 3; ++a) o.nf[a] = (i64)extent(o, a, 0) * extent(o, a, 1) * extent(o, a, 2);
    o.mask.assign(o.n, 0);
    o.collider_density.assign(o.n, 0.0);
    o.collider_temperature.assign(o.n, 0.0);
    std::vector<float> cvel[3];
    for (int a = 0; a < 3; ++a) cvel[a].assign(o.n, 0.0F);
    // model.cpp:56-71: cell centres in FP32; later colliders overwrite earlier ones.
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                const float px = ((float)x + 0.5F) * cfg.cell_size;
                const float py = ((float)y + 0.5F) * cfg.cell_size;
                const float pz = ((float)z + 0.5F) * cfg.cell_size;
                for (std::uint32_t k = 0; k < cfg.collider_count; ++k) {
                    const ds_collider& c = cfg.colliders[k];
                    if (!contains(c, px, py, pz)) continue;
                    o.mask[i] = 1;
                    for (int a = 0; a < 3; ++a) cvel[a][i] = c.velocity[a];
                    o.collider_density[i] = c.density;
                    o.collider_temperature[i] = c.temperature;
                }
            }
    // model.cpp:72-77: anchor = none with a fixed-value pressure face, else first fluid cell.
    o.fixed_pressure = false;
    for (int f = 0; f < 6; ++f)
        if (o.pb.modes[f] == 0u) o.fixed_pressure = true;
    o.solids = 0;
    for (i64 i = 0; i < o.n; ++i) o.solids += o.mask[i];
    if (o.solids == o.n) {
        g_error = "every cell is solid";
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (o.fixed_pressure) o.anchor = o.n;
    else {
        o.anchor = 0;
        while (o.mask[o.anchor] != 0) ++o.anchor;
    }
    // model.cpp:79-150: face collider velocity = mean over adjacent solid cells (FP32).
    for (int a = 0; a < 3; ++a) {
        o.collider_velocity[a].assign(o.nf[a], 0.0);
        const int ex = extent(o, a, 0), ey = extent(o, a, 1), ez = extent(o, a, 2);
        for (int z = 0; z < ez; ++z)
            for (int y = 0; y < ey; ++y)
                for (int x = 0; x < ex; ++x) {
                    float sum = 0.0F, count = 0.0F;
                    const int c0[3]{x - (a == 0), y - (a == 1), z - (a == 2)};
                    const int c1[3]{x, y, z};
                    for (const int* c : {c0, c1}) {
                        if (!in_domain(o, c[0], c[1], c[2])) continue;
                        const i64 ci = cell(o, c[0], c[1], c[2]);
                        if (o.mask[ci] != 0) { sum += cvel[a][ci]; count += 1.0F; }
                    }
                    o.collider_velocity[a][idx3(x, y, z, ex, ey)] = count == 0.0F ? 0.0 : (double)(sum / count);
                }
    }
    // Pressure matrix: diagonal = connected fluid neighbours + fixed-value faces; pockets
    // without the anchor or a fixed face get their own anchor (lowest index).
    o.code.assign(o.n, 0);
    o.fixed_term.assign(o.n, 0.0);
    std::vector<std::uint8_t> diag(o.n, 0), touches_fixed(o.n, 0);
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = cell(o, x, y, z);
                if (o.mask[i] != 0) continue;
                int d = 0;
                double ft = 0.0;
                for (int dd = 0; dd < 3; ++dd)
                    for (int dir = -1; dir <= 1; dir += 2) {
                        bool fixed;
                        double fv = 0.0;
                        const i64 j = pressure_neighbor(o, x, y, z, dd, dir, fixed, fv);
                        if (j >= 0 && o.mask[j] == 0) ++d;
                        else if (j < 0 && fixed) { ++d; ft += fv; touches_fixed[i] = 1; }
                    }
                diag[i] = (std::uint8_t)d;
                o.fixed_term[i] = ft;
            }
    std::vector<std::int32_t> comp(o.n, -1);
    std::vector<i64> stack;
    std::vector<i64> extra;
    std::int32_t ncomp = 0;
    for (i64 seed = 0; seed < o.n; ++seed) {
        if (o.mask[seed] != 0 || comp[seed] >= 0) continue;
        bool grounded = false;
        stack.assign(1, seed);
        comp[seed] = ncomp;
        while (!stack.empty()) {
            const i64 i = stack.back();
            stack.pop_back();
            if (i == o.anchor || touches_fixed[i]) grounded = true;
            const int x = (int)(i % o.nx), y = (int)((i / o.nx) % o.ny), z = (int)(i / ((i64)o.nx * o.ny));
            for (int dd = 0; dd < 3; ++dd)
                for (int dir = -1; dir <= 1; dir += 2) {
                    bool fixed;
                    double fv = 0.0;
                    const i64 j = pressure_neighbor(o, x, y, z, dd, dir, fixed, fv);
                    if (j >= 0 && o.mask[j] == 0 && comp[j] < 0) {
                        comp[j] = ncomp;
                        stack.push_back(j);
                    }
                }
        }
        if (!grounded) extra.push_back(seed);
        ++ncomp;
    }
    std::vector<std::uint8_t> is_anchor(o.n, 0);
    if (o.anchor < o.n) is_anchor[o.anchor] = 1;
    for (i64 a : extra) is_anchor[a] = 1;
    o.extra_anchors = (std::uint32_t)extra.size();
    o.unknowns = 0;
    for (i64 i = 0; i < o.n; ++i) {
        if (o.mask[i] != 0 || is_anchor[i] || diag[i] == 0) {
            o.code[i] = 0;
            o.fixed_term[i] = 0.0;
            continue;
        }
        o.code[i] = diag[i];
        ++o.unknowns;
    }
    return DS_OK;
}

// ---- forward example harness (examples/smoke/forward/simulation.cu) -------------------------------
struct Jet {
    double pulse_period;
    double lc[3], rc[3];
    double radius, rho_rate, t_rate;
    double la[3], ra[3];
};

// write_control_kernel, simulation.cu:35-57 (gaussian :24-29), in FP64: two Gaussian jets with
// amplitudes 0.75 + 0.25 sin(2 pi t / T) and 0.75 + 0.25 sin(2 pi t / T + pi), t = step * dt.
void example_write_control(const Oracle& o, std::uint64_t step, const Jet& j, double* s_rho, double* s_t, double* const e[3]) {
    const double two_pi = 6.28318530717958647692;
    const double time = (double)step * (double)o.dt;
    const double phase = two_pi * time / j.pulse_period;
    const double amp_l = 0.75 + 0.25 * std::sin(phase);
    const double amp_r = 0.75 + 0.25 * std::sin(phase + 0.5 * two_pi);
    auto gauss = [&](double x, double y, double z, const double c[3]) {
        const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
        return std::exp(-0.5 * (dx * dx + dy * dy + dz * dz) / (j.radius * j.radius));
    };
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = x + (i64)o.nx * (y + (i64)o.ny * z);
                const double px = (x + 0.5) * o.h, py = (y + 0.5) * o.h, pz = (z + 0.5) * o.h;
                const double wl = amp_l * gauss(px, py, pz, j.lc), wr = amp_r * gauss(px, py, pz, j.rc);
                s_rho[i] = j.rho_rate * (wl + wr);
                s_t[i] = j.t_rate * (wl + wr);
                for (int a = 0; a < 3; ++a) e[a][i] = wl * j.la[a] + wr * j.ra[a];
            }
}

// reduce_metrics_kernel, simulation.cu:59-85, over fluid cells (mask == 0): out = { sum rho,
// max(0, rho), max(0, T), max(0, |cell-centred post velocity|), sum div_pre^2, sum div_post^2 }.
// The fields are FP32 in the reference, so each face value is taken as float, the face pair of
// the cell velocity is summed in FP32 and halved in FP64, and the divergence is the FP32 sum
// x1 - x0 + y1 - y0 + z1 - z0 in that order, scaled by 1/h (h as float) in FP64.
void example_reduce_metrics(const Oracle& o, const double* rho, const double* T, const double* const pre[3], const double* const post[3],
                            double out[6]) {
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    const double inv_h = 1.0 / (double)o.hf;
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const i64 i = x + (i64)o.nx * (y + (i64)o.ny * z);
                if (o.mask[i] != 0) continue;
                const i64 x0 = x + (i64)(o.nx + 1) * (y + (i64)o.ny * z), x1 = x0 + 1;
                const i64 y0 = x + (i64)o.nx * (y + (i64)(o.ny + 1) * z), y1 = y0 + o.nx;
                const i64 z0 = i, z1 = i + (i64)o.nx * o.ny;
                auto f = [](const double* a, i64 k) { return (float)a[k]; };
                const double vx = 0.5 * (double)(f(post[0], x0) + f(post[0], x1));
                const double vy = 0.5 * (double)(f(post[1], y0) + f(post[1], y1));
                const double vz = 0.5 * (double)(f(post[2], z0) + f(post[2], z1));
                float dpre = f(pre[0], x1) - f(pre[0], x0);
                dpre = dpre + f(pre[1], y1);
                dpre = dpre - f(pre[1], y0);
                dpre = dpre + f(pre[2], z1);
                dpre = dpre - f(pre[2], z0);
                float dpost = f(post[0], x1) - f(post[0], x0);
                dpost = dpost + f(post[1], y1);
                dpost = dpost - f(post[1], y0);
                dpost = dpost + f(post[2], z1);
                dpost = dpost - f(post[2], z0);
                const double dv_pre = (double)dpre * inv_h, dv_post = (double)dpost * inv_h;
                const double speed = std::sqrt(vx * vx + vy * vy + vz * vz);
                out[0] += (double)(float)rho[i];
                if ((double)(float)rho[i] > out[1]) out[1] = (double)(float)rho[i];
                if ((double)(float)T[i] > out[2]) out[2] = (double)(float)T[i];
                if (speed > out[3]) out[3] = speed;
                out[4] += dv_pre * dv_pre;
                out[5] += dv_post * dv_post;
            }
}

}  // namespace

// ---- C API for ctypes (test infrastructure) ---------------------------------------------------
extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

int oracle_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
    return omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

void* oracle_create(const ds_config* cfg) {
    if (!cfg) return nullptr;
    Oracle* o = new Oracle();
    if (build(*o, *cfg) != DS_OK) {
        delete o;
        return nullptr;
    }
    return o;
}

void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }

void oracle_domain_info(void* h, ds_domain_info* info) {
    const Oracle& o = *static_cast<Oracle*>(h);
    info->cell_count = (std::uint64_t)o.n;
    for (int a = 0; a < 3; ++a) info->face_count[a] = (std::uint64_t)o.nf[a];
    info->unknown_count = (std::uint64_t)o.unknowns;
    info->solid_count = (std::uint64_t)o.solids;
    info->pressure_anchor = (std::uint64_t)o.anchor;
    info->extra_anchor_count = o.extra_anchors;
    info->has_fixed_pressure = o.fixed_pressure;
}

void oracle_masks(void* h, std::uint8_t* mask, std::uint8_t* code, double* fixed_term) {
    const Oracle& o = *static_cast<Oracle*>(h);
    if (mask) std::copy(o.mask.begin(), o.mask.end(), mask);
    if (code) std::copy(o.code.begin(), o.code.end(), code);
    if (fixed_term) std::copy(o.fixed_term.begin(), o.fixed_term.end(), fixed_term);
}

void oracle_collider_fields(void* h, double* velocity[3], double* density, double* temperature) {
    const Oracle& o = *static_cast<Oracle*>(h);
    for (int a = 0; a < 3; ++a)
        if (velocity && velocity[a]) std::copy(o.collider_velocity[a].begin(), o.collider_velocity[a].end(), velocity[a]);
    if (density) std::copy(o.collider_density.begin(), o.collider_density.end(), density);
    if (temperature) std::copy(o.collider_temperature.begin(), o.collider_temperature.end(), temperature);
}

void oracle_set_solver(void* h, std::uint32_t solver, double tolerance, std::uint32_t max_iterations, std::uint32_t rbgs_iterations) {
    Oracle& o = *static_cast<Oracle*>(h);
    o.projection_solver = solver;
    if (tolerance > 0.0) o.tolerance = tolerance;
    if (max_iterations) o.max_iterations = max_iterations;
    if (rbgs_iterations) o.rbgs_iterations = rbgs_iterations;
}

void oracle_source_forward(void* h, const double* s, const double* src, double* out) {
    source_forward(*static_cast<Oracle*>(h), s, src, out);
}
void oracle_source_vjp(void* h, const double* oa, double* sa, double* srca) {
    source_vjp(*static_cast<Oracle*>(h), oa, sa, srca);
}
void oracle_buoyancy_forward(void* h, const double* rho, const double* T, const double* const e[3], const double* params, double* const f[3]) {
    buoyancy_forward(*static_cast<Oracle*>(h), rho, T, e, params, f);
}
void oracle_buoyancy_vjp(void* h, const double* rho, const double* T, const double* params, const double* const fa[3],
                         double* rho_adj, double* T_adj, double* const e_adj[3], double* param_adj) {
    buoyancy_vjp(*static_cast<Oracle*>(h), rho, T, params, fa, rho_adj, T_adj, e_adj, param_adj);
}
void oracle_vorticity_forward(void* h, const double* const u[3], double confinement, double* const f[3]) {
    vorticity_forward(*static_cast<Oracle*>(h), u, confinement, f);
}
void oracle_vorticity_vjp(void* h, const double* const u[3], double confinement, const double* const fa[3],
                          double* const ua[3], double* conf_adj) {
    vorticity_vjp(*static_cast<Oracle*>(h), u, confinement, fa, ua, conf_adj);
}
void oracle_integrate_forward(void* h, const double* const u[3], const double* const f[3], double* const out[3]) {
    integrate_forward(*static_cast<Oracle*>(h), u, f, out);
}
void oracle_integrate_vjp(void* h, const double* const oa[3], double* const ua[3], double* const fa[3]) {
    integrate_vjp(*static_cast<Oracle*>(h), oa, ua, fa);
}
void oracle_advect_velocity_forward(void* h, const double* const u[3], double* const out[3]) {
    advect_velocity_forward(*static_cast<Oracle*>(h), u, out);
}
void oracle_advect_velocity_vjp(void* h, const double* const u[3], const double* const oa[3], double* const ua[3]) {
    advect_velocity_vjp(*static_cast<Oracle*>(h), u, oa, ua);
}
void oracle_constrain_forward(void* h, const double* const u[3], double* const out[3]) {
    constrain_forward(*static_cast<Oracle*>(h), u, out);
}
void oracle_constrain_vjp(void* h, const double* const oa[3], double* const ua[3]) {
    constrain_vjp(*static_cast<Oracle*>(h), oa, ua);
}
void oracle_advect_scalar_forward(void* h, int field, const double* src, const double* const u[3], double* out) {
    advect_scalar_forward(*static_cast<Oracle*>(h), field, src, u, out);
}
void oracle_advect_scalar_vjp(void* h, int field, const double* src, const double* const u[3], const double* oa,
                              double* src_adj, double* const ua[3]) {
    advect_scalar_vjp(*static_cast<Oracle*>(h), field, src, u, oa, src_adj, ua);
}
void oracle_pressure_rhs(void* h, const double* const u[3], double* rhs) { pressure_rhs(*static_cast<Oracle*>(h), u, rhs); }
void oracle_apply_matrix(void* h, const double* x, double* y) { apply_matrix(*static_cast<Oracle*>(h), x, y); }
int oracle_pcg(void* h, const double* b, const double* x0, double* x, double tol, std::uint32_t max_it, ds_solve_info* info) {
    return pcg(*static_cast<Oracle*>(h), b, x0, x, tol, max_it, info);
}
void oracle_rbgs(void* h, const double* rhs, double* p, std::uint32_t iterations) { rbgs(*static_cast<Oracle*>(h), rhs, p, iterations); }
void oracle_project(void* h, const double* const u[3], const double* p, double* const out[3]) {
    project(*static_cast<Oracle*>(h), u, p, out);
}
int oracle_projection_forward(void* h, const double* const u[3], double* p, double* rhs, const double* guess,
                              double* const out[3], ds_solve_info* info) {
    return projection_forward(*static_cast<Oracle*>(h), u, p, rhs, guess, out, info);
}
int oracle_projection_vjp(void* h, const double* const oa[3], double* const ua[3], ds_solve_info* info) {
    return projection_vjp(*static_cast<Oracle*>(h), oa, ua, info);
}
int oracle_forward_step(void* h, const double* const s[5], const double* const c[5], const double* params,
                        double* const next[5], double* const cache[12], const double* guess, ds_solve_info* info) {
    return forward_step(*static_cast<Oracle*>(h), s, c, params, next, cache, guess, info);
}
int oracle_vjp_step(void* h, const double* const state[5], const double* const next[5], const double* const cache[12],
                    const double* params, const double* const next_adj[5], double* const prev_adj[5],
                    double* const control_adj[5], double* param_adj, ds_solve_info* info) {
    return vjp_step(*static_cast<Oracle*>(h), state, next, cache, params, next_adj, prev_adj, control_adj, param_adj, info);
}
void oracle_buoyancy_jvp(void* h, const double* rho, const double* T, const double* rho_t, const double* T_t,
                         const double* const e_t[3], const double* params, const double* pt, double* const f_t[3]) {
    buoyancy_jvp(*static_cast<Oracle*>(h), rho, T, rho_t, T_t, e_t, params, pt, f_t);
}
void oracle_vorticity_jvp(void* h, const double* const u[3], const double* const u_t[3], double confinement, double confinement_t,
                          double* const f_t[3]) {
    vorticity_jvp(*static_cast<Oracle*>(h), u, u_t, confinement, confinement_t, f_t);
}
void oracle_advect_velocity_jvp(void* h, const double* const u[3], const double* const u_t[3], double* const out_t[3]) {
    advect_velocity_jvp(*static_cast<Oracle*>(h), u, u_t, out_t);
}
void oracle_constrain_jvp(void* h, const double* const u_t[3], double* const out_t[3]) {
    constrain_jvp(*static_cast<Oracle*>(h), u_t, out_t);
}
void oracle_advect_scalar_jvp(void* h, int field, const double* src, const double* src_t, const double* const u[3],
                              const double* const u_t[3], double* out_t) {
    advect_scalar_jvp(*static_cast<Oracle*>(h), field, src, src_t, u, u_t, out_t);
}
int oracle_projection_jvp(void* h, const double* const u_t[3], double* const out_t[3], ds_solve_info* info) {
    return projection_jvp(*static_cast<Oracle*>(h), u_t, out_t, info);
}
int oracle_jvp_step(void* h, const double* const state[5], const double* const next[5], const double* const cache[12],
                    const double* params, const double* const state_t[5], const double* const control_t[5],
                    const double* params_t, double* const next_t[5], ds_solve_info* info) {
    return jvp_step(*static_cast<Oracle*>(h), state, next, cache, params, state_t, control_t, params_t, next_t, info);
}
int oracle_trajectory_jvp(void* h, const double* const init[5], const double* const control[5], const double* params,
                          std::uint32_t steps, const double* const init_t[5], const double* const control_t[5],
                          const double* params_t, double* const final_state[5], double* const final_t[5],
                          ds_trajectory_info* info) {
    return trajectory_jvp(*static_cast<Oracle*>(h), init, control, params, steps, init_t, control_t, params_t, final_state,
                          final_t, info);
}
int oracle_loss_gradient(void* h, const double* const init[5], const double* const control[5], const double* params,
                         std::uint32_t steps, std::uint32_t k, const double* target, double* const final_state[5],
                         double* loss, double* const init_adj[5], double* const control_adj[5], double* param_adj,
                         ds_trajectory_info* info) {
    return loss_gradient(*static_cast<Oracle*>(h), init, control, params, steps, k, target, final_state, loss, init_adj,
                         control_adj, param_adj, info);
}

// jet: pulse_period, left centre[3], right centre[3], radius, density rate, temperature rate,
// left acceleration[3], right acceleration[3] (16 values, the order of ds_jet_control)
void oracle_example_write_control(void* h, std::uint64_t step, const double* jet, double* s_rho, double* s_t, double* const e[3]) {
    Jet j{};
    j.pulse_period = jet[0];
    for (int a = 0; a < 3; ++a) {
        j.lc[a] = jet[1 + a];
        j.rc[a] = jet[4 + a];
        j.la[a] = jet[10 + a];
        j.ra[a] = jet[13 + a];
    }
    j.radius = jet[7];
    j.rho_rate = jet[8];
    j.t_rate = jet[9];
    example_write_control(*static_cast<Oracle*>(h), step, j, s_rho, s_t, e);
}
void oracle_example_reduce_metrics(void* h, const double* rho, const double* T, const double* const pre[3], const double* const post[3],
                                   double* out) {
    example_reduce_metrics(*static_cast<Oracle*>(h), rho, T, pre, post, out);
}

}  // extern "C"
