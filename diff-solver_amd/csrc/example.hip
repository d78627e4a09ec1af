// example.hip — the two kernels of the reference's forward example (examples/smoke/forward/
// simulation.cu): the pulsed double-jet control writer and the per-step metric reduction.
//
// write_control: one thread per cell, 20 B written per cell (HBM-write bound).
// reduce_metrics: ~33 B read per cell (rho, T, mask, one face pair per axis of the pre- and
// post-projection velocity; the x+1 / y+1 / z+1 faces come out of L2). The reference accumulates
// with FP64 atomics, so its sums vary from run to run in the last bits; here every block leaves
// six partials in a fixed slot and one block finishes them in a fixed order, so the result is
// the same on every run.
#include <mutex>

#include "kernels.hpp"

namespace ds {
namespace {

constexpr float kTwoPi = 6.28318530717958647692f;

struct Jet {   // the by-value control arguments of launch_write_control (simulation.h:10)
    float pulse_period;
    float lc[3], rc[3];
    float radius, rho_rate, t_rate;
    float la[3], ra[3];
};

// gaussian, simulation.cu:24-29
__device__ __forceinline__ float jet_gaussian(float x, float y, float z, const float c[3], float radius) {
    const float dx = x - c[0], dy = y - c[1], dz = z - c[2];
    return expf(-0.5f * (dx * dx + dy * dy + dz * dz) / (radius * radius));
}

// write_control_kernel, simulation.cu:35-57: two Gaussian jets whose amplitudes pulse in
// antiphase with period pulse_period; the density and temperature sources are the summed
// weight times their rates, the acceleration is each weight times its jet's vector.
__global__ __launch_bounds__(256) void write_control_kernel(Geo g, uint64_t step, Jet j, float* __restrict__ s_rho,
                                                            float* __restrict__ s_t, F3 e) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const uint32_t yz = fdiv(i, g.cx);
    const uint32_t xi = i - yz * (uint32_t)g.nx;
    const uint32_t zi = fdiv(yz, g.cy);
    const uint32_t yi = yz - zi * (uint32_t)g.ny;
    const float x = ((float)xi + 0.5f) * g.h;
    const float y = ((float)yi + 0.5f) * g.h;
    const float z = ((float)zi + 0.5f) * g.h;
    const float time = (float)step * g.dt;
    const float phase = kTwoPi * time / j.pulse_period;
    const float left_amp = 0.75f + 0.25f * sinf(phase);
    const float right_amp = 0.75f + 0.25f * sinf(phase + 0.5f * kTwoPi);
    const float wl = left_amp * jet_gaussian(x, y, z, j.lc, j.radius);
    const float wr = right_amp * jet_gaussian(x, y, z, j.rc, j.radius);
    const float w = wl + wr;
    s_rho[i] = j.rho_rate * w;
    s_t[i] = j.t_rate * w;
    e.p[0][i] = wl * j.la[0] + wr * j.ra[0];
    e.p[1][i] = wl * j.la[1] + wr * j.ra[1];
    e.p[2][i] = wl * j.la[2] + wr * j.ra[2];
}

// Metric slots, as the reference's device_metrics (simulation.cpp:95-117): 0 sum of rho,
// 1 max rho, 2 max T, 3 max |cell-centred velocity|, 4 sum of squared pre-projection
// divergence, 5 the same after projection. Maxima start at 0, as the reference's zeroed buffer.
constexpr int kMetrics = 6;

__device__ __forceinline__ bool metric_is_max(int k) { return k >= 1 && k <= 3; }

__device__ __forceinline__ double wave_combine(double v, bool is_max) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double other = __shfl_xor(v, o, 64);
        v = is_max ? fmax(v, other) : v + other;
    }
    return v;
}

// Combines the six per-thread values over the block in a fixed order; thread 0 gets the totals.
__device__ __forceinline__ void block_metrics(double (&v)[kMetrics], double* lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int waves = (int)(blockDim.x >> 6);
#pragma unroll
    for (int k = 0; k < kMetrics; ++k) v[k] = wave_combine(v[k], metric_is_max(k));
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < kMetrics; ++k) lds[k * kMaxWaves + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kMetrics; ++k) {
            double s = lds[k * kMaxWaves];
            for (int w = 1; w < waves; ++w) s = metric_is_max(k) ? fmax(s, lds[k * kMaxWaves + w]) : s + lds[k * kMaxWaves + w];
            v[k] = s;
        }
}

// reduce_metrics_kernel, simulation.cu:59-85, as a fixed grid-stride pass over the cells:
// fluid cells only (mask == 0); the divergence is the FP32 sum of the six face differences in
// the reference's order, scaled by 1/h in FP64; the cell velocity is the FP64 half of each FP32
// face-pair sum. Block b leaves its six values at partials[k * gridDim.x + b].
// A null mask means no solid cells.
__global__ __launch_bounds__(256) void metrics_partial_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ rho,
                                                              const float* __restrict__ T, CF3 pre, CF3 post,
                                                              double* __restrict__ partials) {
    __shared__ double lds[kMetrics * kMaxWaves];
    double v[kMetrics]{};
    const double inv_h = 1.0 / (double)g.h;
    const uint32_t plane = (uint32_t)g.nx * (uint32_t)g.ny;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += gridDim.x * blockDim.x) {
        if (mask && mask[i] != 0) continue;
        const uint32_t yz = fdiv(i, g.cx);
        const uint32_t x = i - yz * (uint32_t)g.nx;
        const uint32_t z = fdiv(yz, g.cy);
        const uint32_t y = yz - z * (uint32_t)g.ny;
        const uint32_t x0 = x + ((uint32_t)g.nx + 1u) * (y + (uint32_t)g.ny * z), x1 = x0 + 1u;
        const uint32_t y0 = x + (uint32_t)g.nx * (y + ((uint32_t)g.ny + 1u) * z), y1 = y0 + (uint32_t)g.nx;
        const uint32_t z0 = i, z1 = i + plane;
        const double vx = 0.5 * (double)(post.p[0][x0] + post.p[0][x1]);
        const double vy = 0.5 * (double)(post.p[1][y0] + post.p[1][y1]);
        const double vz = 0.5 * (double)(post.p[2][z0] + post.p[2][z1]);
        const float dpre = pre.p[0][x1] - pre.p[0][x0] + pre.p[1][y1] - pre.p[1][y0] + pre.p[2][z1] - pre.p[2][z0];
        const float dpost = post.p[0][x1] - post.p[0][x0] + post.p[1][y1] - post.p[1][y0] + post.p[2][z1] - post.p[2][z0];
        const double div_pre = (double)dpre * inv_h, div_post = (double)dpost * inv_h;
        v[0] += (double)rho[i];
        v[1] = fmax(v[1], (double)rho[i]);
        v[2] = fmax(v[2], (double)T[i]);
        v[3] = fmax(v[3], sqrt(vx * vx + vy * vy + vz * vz));
        v[4] += div_pre * div_pre;
        v[5] += div_post * div_post;
    }
    block_metrics(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kMetrics; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = v[k];
}

__global__ __launch_bounds__(256) void metrics_finalize_kernel(const double* __restrict__ partials, uint32_t count,
                                                               double* __restrict__ out) {
    __shared__ double lds[kMetrics * kMaxWaves];
    double v[kMetrics]{};
    for (uint32_t b = threadIdx.x; b < count; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < kMetrics; ++k) {
            const double p = partials[(size_t)k * count + b];
            v[k] = metric_is_max(k) ? fmax(v[k], p) : v[k] + p;
        }
    block_metrics(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kMetrics; ++k) out[k] = v[k];
}

// Partials of ds_example_reduce_metrics_static, which allocates nothing: one slot per block and
// metric (kMetrics * kMetricBlocks FP64 = 24 KiB per GPU, held until the process ends). The launch
// always uses kMetricBlocks blocks, so every grid the kernels can index fits in it. Host code
// never passes this name to a kernel: it stands for a host-side shadow there; the device address
// comes from hipGetSymbolAddress, per device.
__device__ __align__(256) double scratch[kMetrics * kMetricBlocks];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kScratchDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
double* metrics_scratch() {
    int dev = 0;
    DS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) throw Failure{DS_ERROR_UNSUPPORTED, "device ordinal beyond the scratch table"};
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        DS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

}  // namespace
}  // namespace ds

using namespace ds;

extern "C" {

ds_status ds_example_write_control(ds_context* c, uint64_t step, const ds_jet_control* jet, float* density_source,
                                   float* temperature_source, float* const external_acceleration[3]) {
    if (!c || !jet || !density_source || !temperature_source || !external_acceleration || !external_acceleration[0] ||
        !external_acceleration[1] || !external_acceleration[2]) {
        set_error("ds_example_write_control: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!(jet->source_radius > 0.0f) || !(jet->pulse_period != 0.0f)) {
        set_error("ds_example_write_control: source_radius must be > 0 and pulse_period nonzero");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        Jet j{};
        j.pulse_period = jet->pulse_period;
        j.radius = jet->source_radius;
        j.rho_rate = jet->density_source_rate;
        j.t_rate = jet->temperature_source_rate;
        for (int a = 0; a < 3; ++a) {
            j.lc[a] = jet->left_source_center[a];
            j.rc[a] = jet->right_source_center[a];
            j.la[a] = jet->left_acceleration[a];
            j.ra[a] = jet->right_acceleration[a];
        }
        write_control_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(
            c->geo, step, j, density_source, temperature_source,
            F3{{external_acceleration[0], external_acceleration[1], external_acceleration[2]}});
        check_launch("write_control_kernel");
    });
}

ds_status ds_example_reduce_metrics(ds_context* c, const float* density, const float* temperature,
                                    const float* const pre_projection_velocity[3], const float* const post_projection_velocity[3],
                                    double* metrics) {
    if (!c || !density || !temperature || !pre_projection_velocity || !post_projection_velocity || !metrics ||
        !pre_projection_velocity[0] || !pre_projection_velocity[1] || !pre_projection_velocity[2] || !post_projection_velocity[0] ||
        !post_projection_velocity[1] || !post_projection_velocity[2]) {
        set_error("ds_example_reduce_metrics: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        metrics_partial_kernel<<<kMetricBlocks, kBlock, 0, c->stream>>>(
            c->geo, c->mask, density, temperature, CF3{{pre_projection_velocity[0], pre_projection_velocity[1], pre_projection_velocity[2]}},
            CF3{{post_projection_velocity[0], post_projection_velocity[1], post_projection_velocity[2]}}, c->metric_partials);
        check_launch("metrics_partial_kernel");
        metrics_finalize_kernel<<<1, kBlock, 0, c->stream>>>(c->metric_partials, kMetricBlocks, metrics);
        check_launch("metrics_finalize_kernel");
    });
}

ds_status ds_example_reduce_metrics_static(const ds_config* config, const uint8_t* solid_mask, const float* density,
                                           const float* temperature, const float* const pre_projection_velocity[3],
                                           const float* const post_projection_velocity[3], double* metrics, void* stream) {
    if (!config || !density || !temperature || !pre_projection_velocity || !post_projection_velocity || !metrics ||
        !pre_projection_velocity[0] || !pre_projection_velocity[1] || !pre_projection_velocity[2] || !post_projection_velocity[0] ||
        !post_projection_velocity[1] || !post_projection_velocity[2]) {
        set_error("ds_example_reduce_metrics_static: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    const uint64_t nx = config->resolution[0], ny = config->resolution[1], nz = config->resolution[2];
    if (!nx || !ny || !nz || !(config->cell_size > 0.0f)) {
        set_error("ds_example_reduce_metrics_static: resolution must be positive and cell_size > 0");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    // the kernels index cells and faces in 32 bits
    if ((nx + 1) * (ny + 1) * (nz + 1) >= (1ull << 32)) {
        set_error("ds_example_reduce_metrics_static: grid too large (cell and face indices must fit in 32 bits)");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        double* partials = metrics_scratch();
        Geo g{};
        g.nx = (int)nx;
        g.ny = (int)ny;
        g.nz = (int)nz;
        g.h = config->cell_size;
        g.inv_h = 1.0f / config->cell_size;
        g.h_pow2 = is_pow2(config->cell_size) ? 1u : 0u;
        g.dt = config->time_step;
        g.n = (uint32_t)(nx * ny * nz);
        g.cx = make_fastdiv((uint32_t)nx);
        g.cy = make_fastdiv((uint32_t)ny);
        hipStream_t s = static_cast<hipStream_t>(stream);
        metrics_partial_kernel<<<kMetricBlocks, kBlock, 0, s>>>(
            g, solid_mask, density, temperature, CF3{{pre_projection_velocity[0], pre_projection_velocity[1], pre_projection_velocity[2]}},
            CF3{{post_projection_velocity[0], post_projection_velocity[1], post_projection_velocity[2]}}, partials);
        check_launch("metrics_partial_kernel");
        metrics_finalize_kernel<<<1, kBlock, 0, s>>>(partials, kMetricBlocks, metrics);
        check_launch("metrics_finalize_kernel");
    });
}

}  // extern "C"
