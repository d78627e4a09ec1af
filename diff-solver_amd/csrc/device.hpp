// device.hpp — gfx950 device helpers of the smoke solver: grid geometry, boundary-aware loads,
// trilinear sampling and its transpose, RK2 backtrace, wave64 FP64 reductions.
//
// The discretisation restates solvers/smoke/operators.cu:32-284 (Xayah-Graphics/diff-solver) in
// FP32: indexing x + nx*(y + ny*z), face extents n_a + 1 on the own axis, boundary loads
// (fixed-value faces return their value, periodic wraps, otherwise clamp), trilinear sampling
// with its analytic gradient, and the RK2 midpoint backtrace with box clamp and solid bisection.
// Indices are 32-bit (grids up to 1024^3); divisions by runtime extents use multiply-high.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ds {

constexpr int kBlock = 256;
constexpr int kWave = 64;

// Division by a runtime constant d >= 1 with a 32-bit multiply-high (round-up method):
// q = (umulhi(n, m) + n) >> s, valid for every 32-bit n.
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0u, 0u};
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    f.s = s;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1ull);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    const uint64_t t = (uint64_t)__umulhi(n, f.m) + n;
    return (uint32_t)(t >> f.s);
}

// Geometry passed by value to every kernel (Grid, cuda_types.h:13-19, plus derived sizes).
struct Geo {
    int nx, ny, nz;
    float h, inv_h, dt;
    uint32_t h_pow2;       // 1 when h is a power of two: p / h is then exactly p * inv_h
    uint32_t n;            // cells
    uint32_t nf[3];        // faces per axis
    uint32_t foff[4];      // offsets of axis 0/1/2 in the concatenated face index space, total
    FastDiv cx, cy;        // cell decode: divide by nx, ny
    FastDiv fx[3], fy[3];  // face decode per axis: divide by extent x, extent y
};

struct VBound {   // VelocityBoundaryData, cuda_types.h:26-29
    uint32_t modes[6];
    float values[18];
};
struct SBound {   // ScalarBoundaryData, cuda_types.h:21-24
    uint32_t modes[6];
    float values[6];
};

__device__ __forceinline__ int ext(const Geo& g, int axis, int dim) {
    const int base = dim == 0 ? g.nx : dim == 1 ? g.ny : g.nz;
    return base + (axis == dim ? 1 : 0);
}
__device__ __forceinline__ int dim_size(const Geo& g, int dim) { return dim == 0 ? g.nx : dim == 1 ? g.ny : g.nz; }
__device__ __forceinline__ uint32_t cidx(const Geo& g, int x, int y, int z) {
    return (uint32_t)x + (uint32_t)g.nx * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
}
__device__ __forceinline__ uint32_t fidx(const Geo& g, int axis, int x, int y, int z) {
    return (uint32_t)x + (uint32_t)ext(g, axis, 0) * ((uint32_t)y + (uint32_t)ext(g, axis, 1) * (uint32_t)z);
}
__device__ __forceinline__ void decode_cell(const Geo& g, uint32_t i, int& x, int& y, int& z) {
    const uint32_t t = fdiv(i, g.cx);
    x = (int)(i - t * (uint32_t)g.nx);
    const uint32_t u = fdiv(t, g.cy);
    y = (int)(t - u * (uint32_t)g.ny);
    z = (int)u;
}
// Local face index of axis A -> (x, y, z).
template <int A>
__device__ __forceinline__ void decode_face_axis(const Geo& g, uint32_t local, int& x, int& y, int& z) {
    const uint32_t ex = (uint32_t)ext(g, A, 0), ey = (uint32_t)ext(g, A, 1);
    const uint32_t t = fdiv(local, g.fx[A]);
    x = (int)(local - t * ex);
    const uint32_t u = fdiv(t, g.fy[A]);
    y = (int)(t - u * ey);
    z = (int)u;
}

// Runs body.template operator()<A>(local) for the axis A owning concatenated face index f
// (blocks straddle at most one axis boundary, so the branch is uniform in almost every wave).
template <typename Body>
__device__ __forceinline__ void for_face(const Geo& g, uint32_t f, Body&& body) {
    if (f < g.foff[1]) body.template operator()<0>(f);
    else if (f < g.foff[2]) body.template operator()<1>(f - g.foff[1]);
    else body.template operator()<2>(f - g.foff[2]);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8). Remapping
// gives each XCD a contiguous range of logical blocks, so the neighbour reuse of a stencil or
// a gather stays in that XCD's L2. The grid must be a multiple of 8.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t b = blockIdx.x, G = gridDim.x;
    return (b & 7u) * (G >> 3) + (b >> 3);
}

__device__ __forceinline__ bool vperiodic(const VBound& b, int d) { return b.modes[2 * d] == 3u && b.modes[2 * d + 1] == 3u; }
__device__ __forceinline__ bool speriodic(const SBound& b, int d) { return b.modes[2 * d] == 2u && b.modes[2 * d + 1] == 2u; }
__device__ __forceinline__ int wrapi(int v, int p) {
    const int r = v % p;
    return r < 0 ? r + p : r;
}
__device__ __forceinline__ int mapc(int v, int size, bool per, int period) {
    return per ? wrapi(v, period) : max(0, min(size - 1, v));
}

// load_scalar constant test (operators.cu:115-121 / scatter skip at :212).
__device__ __forceinline__ bool scalar_const(const Geo& g, int x, int y, int z, const SBound& b, float& value) {
    if (x < 0 && b.modes[0] == 0u) { value = b.values[0]; return true; }
    if (x >= g.nx && b.modes[1] == 0u) { value = b.values[1]; return true; }
    if (y < 0 && b.modes[2] == 0u) { value = b.values[2]; return true; }
    if (y >= g.ny && b.modes[3] == 0u) { value = b.values[3]; return true; }
    if (z < 0 && b.modes[4] == 0u) { value = b.values[4]; return true; }
    if (z >= g.nz && b.modes[5] == 0u) { value = b.values[5]; return true; }
    return false;
}
__device__ __forceinline__ uint32_t scalar_mapped(const Geo& g, int x, int y, int z, const SBound& b) {
    x = mapc(x, g.nx, speriodic(b, 0), g.nx);
    y = mapc(y, g.ny, speriodic(b, 1), g.ny);
    z = mapc(z, g.nz, speriodic(b, 2), g.nz);
    return cidx(g, x, y, z);
}
// load_face constant test (operators.cu:125-132 / scatter skip at :230-237). The axis is a
// template parameter so that every boundary-array index is a compile-time constant (a runtime
// index into a by-value kernel argument would move it to scratch memory).
template <int A>
__device__ __forceinline__ bool face_const(const Geo& g, int x, int y, int z, const VBound& b, float& value) {
    const int c[3]{x, y, z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int size = ext(g, A, d);
        if (c[d] < 0) {
            const uint32_t m = b.modes[2 * d];
            if (m == 0u || (m == 2u && A == d)) { value = b.values[3 * (2 * d) + A]; return true; }
        } else if (c[d] >= size) {
            const uint32_t m = b.modes[2 * d + 1];
            if (m == 0u || (m == 2u && A == d)) { value = b.values[3 * (2 * d + 1) + A]; return true; }
        }
    }
    return false;
}
// mapped_face_index, operators.cu:105-113 (period = cell count of the dimension).
template <int A>
__device__ __forceinline__ uint32_t face_mapped(const Geo& g, int x, int y, int z, const VBound& b) {
    x = mapc(x, ext(g, A, 0), vperiodic(b, 0), g.nx);
    y = mapc(y, ext(g, A, 1), vperiodic(b, 1), g.ny);
    z = mapc(z, ext(g, A, 2), vperiodic(b, 2), g.nz);
    return fidx(g, A, x, y, z);
}

struct Corner8 {   // the 2x2x2 trilinear stencil: base cell and fractional offsets
    int x0, y0, z0;
    float tx, ty, tz;
};

// p / h: grid coordinates divide by h as the reference does (operators.cu:137-139,167-172); with a
// cell size that is not a power of two, p * (1/h) can round to the other side of a cell boundary.
// For a power of two the product is the same number and skips the division sequence (the branch
// is uniform: g is a kernel argument).
__device__ __forceinline__ float over_h(const Geo& g, float p) { return g.h_pow2 ? p * g.inv_h : p / g.h; }

__device__ __forceinline__ Corner8 corners(const Geo& g, float px, float py, float pz, float ox, float oy, float oz) {
    const float gx = over_h(g, px) - ox, gy = over_h(g, py) - oy, gz = over_h(g, pz) - oz;
    Corner8 c;
    const float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    // A non-finite or huge position (a diverged solve upstream) must not reach the index
    // arithmetic: clamped to +-2^24 cells (NaN to the lower bound) before the conversion, the
    // corner lies far outside the domain, fails every interior test without overflowing (x0 + 1
    // stays representable) and takes the boundary-mapped loads; finite positions within 2^24
    // cells of the domain are unchanged.
    constexpr float kFar = 16777216.0f;
    c.x0 = (int)fminf(fmaxf(fx, -kFar), kFar);
    c.y0 = (int)fminf(fmaxf(fy, -kFar), kFar);
    c.z0 = (int)fminf(fmaxf(fz, -kFar), kFar);
    c.tx = gx - fx;
    c.ty = gy - fy;
    c.tz = gz - fz;
    return c;
}

struct Sample {
    float value, gx, gy, gz;
};

// Trilinear value and analytic gradient of 8 corner values (operators.cu:150-163).
template <bool Grad>
__device__ __forceinline__ Sample trilerp(const float v[8], const Corner8& c, float inv_h) {
    const float wx0 = 1.0f - c.tx, wx1 = c.tx, wy0 = 1.0f - c.ty, wy1 = c.ty, wz0 = 1.0f - c.tz, wz1 = c.tz;
    // v index: dx + 2*dy + 4*dz
    Sample s;
    s.value = wz0 * (wy0 * (wx0 * v[0] + wx1 * v[1]) + wy1 * (wx0 * v[2] + wx1 * v[3])) +
              wz1 * (wy0 * (wx0 * v[4] + wx1 * v[5]) + wy1 * (wx0 * v[6] + wx1 * v[7]));
    if (Grad) {
        s.gx = (wz0 * (wy0 * (v[1] - v[0]) + wy1 * (v[3] - v[2])) + wz1 * (wy0 * (v[5] - v[4]) + wy1 * (v[7] - v[6]))) * inv_h;
        s.gy = (wz0 * (wx0 * (v[2] - v[0]) + wx1 * (v[3] - v[1])) + wz1 * (wx0 * (v[6] - v[4]) + wx1 * (v[7] - v[5]))) * inv_h;
        s.gz = (wy0 * (wx0 * (v[4] - v[0]) + wx1 * (v[5] - v[1])) + wy1 * (wx0 * (v[6] - v[2]) + wx1 * (v[7] - v[3]))) * inv_h;
    } else {
        s.gx = s.gy = s.gz = 0.0f;
    }
    return s;
}

// True when all 8 corners of a face stencil are in range and map to themselves. A periodic
// axis maps its own-axis face n to face 0 (wrap with period n, operators.cu:109-111), so the
// fast path stops one face earlier there.
template <int A>
__device__ __forceinline__ bool face_interior(const Geo& g, const Corner8& c, const VBound& b) {
    const int lx = A == 0 && !vperiodic(b, 0) ? g.nx + 1 : g.nx;
    const int ly = A == 1 && !vperiodic(b, 1) ? g.ny + 1 : g.ny;
    const int lz = A == 2 && !vperiodic(b, 2) ? g.nz + 1 : g.nz;
    return c.x0 >= 0 && c.x0 + 1 < lx && c.y0 >= 0 && c.y0 + 1 < ly && c.z0 >= 0 && c.z0 + 1 < lz;
}

// sample_face (operators.cu:166-194) with an interior fast path.
template <int A, bool Grad>
__device__ __forceinline__ Sample sample_face(const float* __restrict__ v, float px, float py, float pz, const Geo& g,
                                              const VBound& b) {
    const Corner8 c = corners(g, px, py, pz, A == 0 ? 0.0f : 0.5f, A == 1 ? 0.0f : 0.5f, A == 2 ? 0.0f : 0.5f);
    const int ex = ext(g, A, 0), ey = ext(g, A, 1);
    float vals[8];
    if (face_interior<A>(g, c, b)) {
        const uint32_t base = (uint32_t)c.x0 + (uint32_t)ex * ((uint32_t)c.y0 + (uint32_t)ey * (uint32_t)c.z0);
        const uint32_t sy = (uint32_t)ex, sz = (uint32_t)ex * (uint32_t)ey;
        vals[0] = v[base];
        vals[1] = v[base + 1];
        vals[2] = v[base + sy];
        vals[3] = v[base + sy + 1];
        vals[4] = v[base + sz];
        vals[5] = v[base + sz + 1];
        vals[6] = v[base + sz + sy];
        vals[7] = v[base + sz + sy + 1];
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int x = c.x0 + (k & 1), y = c.y0 + ((k >> 1) & 1), z = c.z0 + (k >> 2);
            float cv;
            vals[k] = face_const<A>(g, x, y, z, b, cv) ? cv : v[face_mapped<A>(g, x, y, z, b)];
        }
    }
    return trilerp<Grad>(vals, c, g.inv_h);
}

// sample_scalar (operators.cu:136-164) with an interior fast path.
template <bool Grad>
__device__ __forceinline__ Sample sample_scalar(const float* __restrict__ v, float px, float py, float pz, const Geo& g,
                                                const SBound& b) {
    const Corner8 c = corners(g, px, py, pz, 0.5f, 0.5f, 0.5f);
    float vals[8];
    if (c.x0 >= 0 && c.x0 + 1 < g.nx && c.y0 >= 0 && c.y0 + 1 < g.ny && c.z0 >= 0 && c.z0 + 1 < g.nz) {
        const uint32_t base = cidx(g, c.x0, c.y0, c.z0);
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        vals[0] = v[base];
        vals[1] = v[base + 1];
        vals[2] = v[base + sy];
        vals[3] = v[base + sy + 1];
        vals[4] = v[base + sz];
        vals[5] = v[base + sz + 1];
        vals[6] = v[base + sz + sy];
        vals[7] = v[base + sz + sy + 1];
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int x = c.x0 + (k & 1), y = c.y0 + ((k >> 1) & 1), z = c.z0 + (k >> 2);
            float cv;
            vals[k] = scalar_const(g, x, y, z, b, cv) ? cv : v[scalar_mapped(g, x, y, z, b)];
        }
    }
    return trilerp<Grad>(vals, c, g.inv_h);
}

// Sink of a scatter: adds value to face (x, y, z) of axis A (flat index i). The default adds to
// the global array with an FP32 atomic.
struct GlobalSink {
    float* v;
    template <int A>
    __device__ __forceinline__ void add(int, int, int, uint32_t i, float value) const { atomicAdd(v + i, value); }
};

// scatter_face (operators.cu:217-240): transpose of sample_face's value.
template <int A, typename Sink>
__device__ __forceinline__ void scatter_face_to(const Sink& sink, float px, float py, float pz, float adj, const Geo& g, const VBound& b) {
    if (adj == 0.0f) return;
    const Corner8 c = corners(g, px, py, pz, A == 0 ? 0.0f : 0.5f, A == 1 ? 0.0f : 0.5f, A == 2 ? 0.0f : 0.5f);
    const int ex = ext(g, A, 0), ey = ext(g, A, 1);
    const float w[2][3] = {{1.0f - c.tx, 1.0f - c.ty, 1.0f - c.tz}, {c.tx, c.ty, c.tz}};
    if (face_interior<A>(g, c, b)) {
        const uint32_t base = (uint32_t)c.x0 + (uint32_t)ex * ((uint32_t)c.y0 + (uint32_t)ey * (uint32_t)c.z0);
        const uint32_t sy = (uint32_t)ex, sz = (uint32_t)ex * (uint32_t)ey;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
            const float wk = w[dx][0] * w[dy][1] * w[dz][2];
            if (wk != 0.0f) sink.template add<A>(c.x0 + dx, c.y0 + dy, c.z0 + dz, base + dx + dy * sy + dz * sz, adj * wk);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
            int x = c.x0 + dx, y = c.y0 + dy, z = c.z0 + dz;
            float cv;
            if (face_const<A>(g, x, y, z, b, cv)) continue;
            const float wk = w[dx][0] * w[dy][1] * w[dz][2];
            if (wk == 0.0f) continue;
            x = mapc(x, ext(g, A, 0), vperiodic(b, 0), g.nx);
            y = mapc(y, ext(g, A, 1), vperiodic(b, 1), g.ny);
            z = mapc(z, ext(g, A, 2), vperiodic(b, 2), g.nz);
            sink.template add<A>(x, y, z, fidx(g, A, x, y, z), adj * wk);
        }
    }
}

template <int A>
__device__ __forceinline__ void scatter_face(float* v, float px, float py, float pz, float adj, const Geo& g, const VBound& b) {
    scatter_face_to<A>(GlobalSink{v}, px, py, pz, adj, g, b);
}

// scatter_scalar with a sink: the same weights and skips, the adds go to sink.add<W>(x, y, z, i, value)
// with (x, y, z) the mapped cell and i its index (W names the sink's target array).
template <int W, typename Sink>
__device__ __forceinline__ void scatter_scalar_to(const Sink& sink, float px, float py, float pz, float adj, const Geo& g,
                                                  const SBound& b) {
    if (adj == 0.0f) return;
    const Corner8 c = corners(g, px, py, pz, 0.5f, 0.5f, 0.5f);
    const float w[2][3] = {{1.0f - c.tx, 1.0f - c.ty, 1.0f - c.tz}, {c.tx, c.ty, c.tz}};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        int x = c.x0 + dx, y = c.y0 + dy, z = c.z0 + dz;
        float cv;
        if (scalar_const(g, x, y, z, b, cv)) continue;
        const float wk = w[dx][0] * w[dy][1] * w[dz][2];
        if (wk == 0.0f) continue;
        x = mapc(x, g.nx, speriodic(b, 0), g.nx);
        y = mapc(y, g.ny, speriodic(b, 1), g.ny);
        z = mapc(z, g.nz, speriodic(b, 2), g.nz);
        sink.template add<W>(x, y, z, cidx(g, x, y, z), adj * wk);
    }
}

// scatter_scalar (operators.cu:196-215).
__device__ __forceinline__ void scatter_scalar(float* v, float px, float py, float pz, float adj, const Geo& g,
                                               const SBound& b) {
    if (adj == 0.0f) return;
    const Corner8 c = corners(g, px, py, pz, 0.5f, 0.5f, 0.5f);
    const float w[2][3] = {{1.0f - c.tx, 1.0f - c.ty, 1.0f - c.tz}, {c.tx, c.ty, c.tz}};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const int x = c.x0 + dx, y = c.y0 + dy, z = c.z0 + dz;
        float cv;
        if (scalar_const(g, x, y, z, b, cv)) continue;
        const float wk = w[dx][0] * w[dy][1] * w[dz][2];
        if (wk != 0.0f) atomicAdd(v + scalar_mapped(g, x, y, z, b), adj * wk);
    }
}

struct Vel3 {
    const float* __restrict__ u[3];
};

// solid_at, operators.cu:246-254 (mask: u8 per cell, 1 = solid).
__device__ __forceinline__ bool solid_at(float px, float py, float pz, const uint8_t* __restrict__ mask, const Geo& g,
                                         const VBound& b) {
    if (vperiodic(b, 0)) { const float L = g.nx * g.h; px -= floorf(px / L) * L; }
    if (vperiodic(b, 1)) { const float L = g.ny * g.h; py -= floorf(py / L) * L; }
    if (vperiodic(b, 2)) { const float L = g.nz * g.h; pz -= floorf(pz / L) * L; }
    // clamped as floats, so a non-finite position maps to a cell of the domain (NaN to 0)
    const int x = (int)fminf(fmaxf(floorf(over_h(g, px)), 0.0f), (float)(g.nx - 1));
    const int y = (int)fminf(fmaxf(floorf(over_h(g, py)), 0.0f), (float)(g.ny - 1));
    const int z = (int)fminf(fmaxf(floorf(over_h(g, pz)), 0.0f), (float)(g.nz - 1));
    return mask[cidx(g, x, y, z)] != 0;
}

struct TraceOut {
    float px, py, pz;   // end point
    float dx, dy, dz;   // derivative scale per axis (0 when clamped, bisection fraction when bisected)
    float mx, my, mz;   // midpoint
};

// trace_rk2, operators.cu:256-276. has_solids = false skips the mask lookup (no colliders).
// MidGrad: also return the three velocity samples at the midpoint with their gradients in mid
// (the reverse and tangent sweeps need them; the values are the ones the trace uses).
template <bool HasSolids, bool MidGrad = false>
__device__ __forceinline__ TraceOut trace_rk2(float sx, float sy, float sz, const Vel3& u, const uint8_t* __restrict__ mask,
                                              const Geo& g, const VBound& b, Sample* mid = nullptr) {
    TraceOut t;
    const float v0x = sample_face<0, false>(u.u[0], sx, sy, sz, g, b).value;
    const float v0y = sample_face<1, false>(u.u[1], sx, sy, sz, g, b).value;
    const float v0z = sample_face<2, false>(u.u[2], sx, sy, sz, g, b).value;
    t.mx = sx - 0.5f * g.dt * v0x;
    t.my = sy - 0.5f * g.dt * v0y;
    t.mz = sz - 0.5f * g.dt * v0z;
    const Sample m0 = sample_face<0, MidGrad>(u.u[0], t.mx, t.my, t.mz, g, b);
    const Sample m1 = sample_face<1, MidGrad>(u.u[1], t.mx, t.my, t.mz, g, b);
    const Sample m2 = sample_face<2, MidGrad>(u.u[2], t.mx, t.my, t.mz, g, b);
    if (MidGrad) {
        mid[0] = m0;
        mid[1] = m1;
        mid[2] = m2;
    }
    const float v1x = m0.value, v1y = m1.value, v1z = m2.value;
    float rx = sx - g.dt * v1x, ry = sy - g.dt * v1y, rz = sz - g.dt * v1z;
    t.dx = t.dy = t.dz = 1.0f;
    const float Lx = g.nx * g.h, Ly = g.ny * g.h, Lz = g.nz * g.h;
    if (!vperiodic(b, 0) && (rx < 0.0f || rx > Lx)) { rx = fminf(fmaxf(rx, 0.0f), Lx); t.dx = 0.0f; }
    if (!vperiodic(b, 1) && (ry < 0.0f || ry > Ly)) { ry = fminf(fmaxf(ry, 0.0f), Ly); t.dy = 0.0f; }
    if (!vperiodic(b, 2) && (rz < 0.0f || rz > Lz)) { rz = fminf(fmaxf(rz, 0.0f), Lz); t.dz = 0.0f; }
    if (HasSolids && solid_at(rx, ry, rz, mask, g, b)) {
        float lo = 0.0f, hi = 1.0f;
        for (int it = 0; it < 8; ++it) {
            const float f = 0.5f * (lo + hi);
            if (solid_at(sx + f * (rx - sx), sy + f * (ry - sy), sz + f * (rz - sz), mask, g, b)) hi = f;
            else lo = f;
        }
        rx = sx + lo * (rx - sx);
        ry = sy + lo * (ry - sy);
        rz = sz + lo * (rz - sz);
        t.dx *= lo;
        t.dy *= lo;
        t.dz *= lo;
    }
    t.px = rx;
    t.py = ry;
    t.pz = rz;
    return t;
}

// ---- reductions: FP64 sums over a block of up to 1024 threads (wave64 shuffles, then LDS) ---
constexpr int kMaxWaves = 16;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums N values per thread over the block; every thread receives the totals in a fixed order.
// lds: N * kMaxWaves doubles.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int waves = (int)(blockDim.x >> 6);
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();   // lds may still be read by a previous call
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) lds[k * kMaxWaves + wid] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
        for (int w = 0; w < waves; ++w) s += lds[k * kMaxWaves + w];
        v[k] = s;
    }
}

// Deterministic sum of `count` partials (same order in every block): returns the total to all
// threads. Used by a consumer kernel to finish the reduction its producer left per block.
template <int N>
__device__ __forceinline__ void reduce_partials(const double* const (&parts)[N], uint32_t count, double (&out)[N], double* lds) {
    double v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] += parts[k][i];
    block_sum<N>(v, lds);
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = v[k];
}

}  // namespace ds
