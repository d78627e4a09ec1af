// context.hpp — host-side execution context of the smoke solver (ExecutionContext + Resource of
// the reference: solvers/smoke/model.ixx:15-49, solvers/cuda.ixx:7-28) and the helpers every
// C-ABI entry point shares: status codes with a thread-local message, launch checks, scratch.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "ds/ds_smoke.h"
#include "device.hpp"

namespace ds {

void set_error(const std::string& message);

// Throwing is internal only: every C-ABI function catches and converts to a ds_status.
struct Failure {
    ds_status status;
    std::string message;
};

// A failed runtime call becomes a Failure; the runtime's last-error slot, which the failed call
// also set, is cleared so that the next check_launch does not report it again as a launch error.
#define DS_HIP(call)                                                                                  \
    do {                                                                                              \
        const hipError_t ds_hip_status_ = (call);                                                     \
        if (ds_hip_status_ != hipSuccess) {                                                           \
            (void)hipGetLastError();                                                                  \
            throw ::ds::Failure{ds_hip_status_ == hipErrorOutOfMemory ? DS_ERROR_OUT_OF_MEMORY : DS_ERROR_RUNTIME, \
                                std::string(#call) + ": " + hipGetErrorString(ds_hip_status_)};       \
        }                                                                                             \
    } while (0)

// check_launch (operators.cu:28-30): a failed launch becomes DS_ERROR_LAUNCH with the kernel name.
void check_launch(const char* name);

// Runs body; converts a Failure (or a C++ exception) to a status code plus ds_last_error().
template <typename F>
ds_status guard(F&& body) {
    try {
        body();
        return DS_OK;
    } catch (const Failure& f) {
        set_error(f.message);
        return f.status;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return DS_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_error(e.what());
        return DS_ERROR_RUNTIME;
    }
}

inline void require(bool ok, const char* message) {
    if (!ok) throw Failure{DS_ERROR_INVALID_ARGUMENT, message};
}

// True for a positive normal power of two (then p / h == p * (1 / h) exactly).
inline bool is_pow2(float h) {
    int e = 0;
    return h > 0.0f && std::frexp(h, &e) == 0.5f && h >= 1.17549435e-38f && 1.0f / h < 3.4e38f;
}

inline uint32_t blocks_for(uint64_t count, uint32_t block = kBlock) { return (uint32_t)((count + block - 1) / block); }
// Block count rounded up to a multiple of 8 for kernels that remap blocks per XCD.
inline uint32_t xcd_blocks_for(uint64_t count, uint32_t block = kBlock) { return (blocks_for(count, block) + 7u) & ~7u; }

// Reduction layout: kernels of kPcgBlocks (256-thread) blocks leave one FP64 partial per block
// per quantity; consumers finish the sum in a fixed order (deterministic, no atomics).
constexpr uint32_t kPcgBlocks = 512;
// Blocks of the force and source reverse (per-cell work with FP64 parameter partials): enough waves
// to keep its dependent gathers in flight.
constexpr uint32_t kForceBlocks = 2048;
// Blocks of the example metric reduction (one partial per block per metric, finished in order).
constexpr uint32_t kMetricBlocks = 512;
// PCG iteration kernels: 1024-thread blocks, each over a contiguous range of whole multiples of
// kPcgUnroll * 1024 cells; the block count (<= kPcgMaxBlocks, a multiple of 8) is per context.
// (measured on the 128^3 plume: 512 threads / ~512 blocks, i.e. two independent 8-wave blocks per
// CU, 29.3 us per fused iteration; 1024 / 256: 32.1 us; 512 / 256: 30.8 us; 1024 / 512: 44.3 us)
constexpr uint32_t kPcgThreads = 512;
// Block count the plane-marching iteration kernels aim for (the z chunk is sized to it).
constexpr uint64_t kPcgTargetBlocks = 512;
// the s-step kernel: one 512-thread block per CU, each with a longer z chunk and deeper prefetch
constexpr uint64_t kPcgTargetBlocksS2 = 256;
constexpr uint32_t kPcgUnroll = 4;
constexpr uint32_t kPcgMaxBlocks = 1024;

enum Partial : int {
    kPartBB = 0, kPartRR = 1, kPartRZ = 2, kPartPQ = 3,
    // single-reduction form only: <w_{i+1}, p_i>, <s_i, u_{i+1}>, <s_i, p_i>
    kPartWP = 4, kPartSU = 5, kPartSP = 6,
    kPartCount = 8
};

// Device status of one PCG solve; written by block 0 of the kernel that decides.
struct PcgStatus {
    uint32_t stop;          // 1 once converged (or the guess made the residual small enough)
    uint32_t iterations;    // iterations performed when stop was set
    uint32_t rejected;
    uint32_t aborted;       // 1 when a resident launch's grid barrier exceeded its bound
    double bb;              // ||b||^2
    double rr;              // last reduced ||r||^2 (recursive or true)
    uint32_t k;             // DS_CG_MULTIGRID: the iteration whose V-cycle ran last (advanced on the device)
    uint32_t unused;
};

// HIP-event timing of sampled PCG kernel launches (kernel 0 = direction, 1 = update, 2 = the
// fused single-reduction iteration).
// 3 = the resident single-reduction launch (units: iterations), 4 = the s-step launch, 5 = one
// multigrid-preconditioned iteration (V-cycle, direction and update), 6 = one launch of the
// velocity advection VJP (one face axis)
constexpr int kProfiledKernels = 7;
struct Profiler {
    uint32_t every = 0;
    uint64_t launches[kProfiledKernels]{};
    std::vector<hipEvent_t> pool;   // created lazily, reused
    size_t used = 0;
    struct Rec {
        int kernel;
        hipEvent_t a, b;
        uint32_t units = 1;   // iterations the timed launch performed (resident launches)
    };
    std::vector<Rec> recs;
    hipEvent_t take() {
        if (used == pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
};

// Plane-marching tiling of the PCG iteration kernels: 64 x 16 (x, y) tiles, zc planes per block.
constexpr int kPlaneTileX = 64;
constexpr int kPlaneTileY = (int)(kPcgThreads / 64);   // one wave per tile row
struct PlaneTiling {
    uint32_t tiles_x, tiles_y, zc;
};

struct UNet;   // warm-start network (unet.hip)

// DS_WARM_START_PROJECTION (window.hip): the last kWindow solutions of one solve sequence, their
// images under A and their Gram matrix, as a ring.
constexpr int kWindow = 8;
struct SolveWindow {
    float* x[kWindow]{};
    float* ax[kWindow]{};
    double* gram = nullptr;       // kWindow x kWindow, <x_i, A x_j>
    double* partials = nullptr;   // kWindow x window blocks
    float* coef = nullptr;        // kWindow
    uint32_t count = 0, head = 0; // live slots; the next slot to write
};
struct WinOrder {
    int slot[kWindow];            // live slots from newest to oldest
};

struct PcgWork {
    PlaneTiling tiling{1, 1, 4};
    float* b = nullptr;          // right-hand side (zero off the unknowns)
    float* r = nullptr;          // residual
    float* p[2]{};               // search directions, ping-pong between iterations
    uint32_t blocks = 8;         // blocks of the PCG kernels
    uint32_t variant = DS_CG_SINGLE_REDUCTION;
    // single-reduction form: r (sets r, r_alt), w = A D^-1 r and s = A p double-buffered; p = p[0]
    float* r_alt = nullptr;
    float* wv[2]{};
    float* sv[2]{};
    float* xd = nullptr;         // solution increment since the last fold into the pair
    // s-step form (DS_CG_S_STEP): the basis P0, P1, P2, Z0, Z1 double-buffered; its partial sums
    float* s2set[2]{};           // each 5 (n + 1) floats: P0, P1, P2, Z0, Z1
    float* s2v[2][5]{};          // the vectors of each set
    double* s2_parts = nullptr;  // two sets of 16 * kPcgMaxBlocks
    double* partials = nullptr;  // two sets of kPartCount * kPcgMaxBlocks
    double* hist = nullptr;      // per-iteration scalars: rz[k], pq[k] (2 * (max_it + 2))
    PcgStatus* status = nullptr; // device
    PcgStatus* host_status = nullptr;  // pinned, 2 slots (double-buffered batches)
    hipEvent_t batch_event[2]{};
    // single-reduction form run as one cooperative launch per replacement interval with the
    // vectors on chip (pcg_resident.hpp; DS_CG_SINGLE_REDUCTION_RESIDENT), when the grid allows it
    bool resident = false;
    uint32_t* resident_sync = nullptr;   // ResidentSync (2 words)
    uint32_t mg_expect = 0;      // DS_CG_MULTIGRID: iterations of the last converged solve
};

// DS_CG_MULTIGRID (pcg_mg.hip): one level of the aggregation hierarchy. Level 0 is the pressure
// matrix itself (code: the diagonal count, -1 between neighbouring unknowns); a coarser level holds
// its Galerkin operator as the diagonal and, per axis, the coupling to the + neighbour (a periodic
// axis wraps). Cells with a zero diagonal are not unknowns of their level.
struct MgLevel {
    int nx, ny, nz;
    uint32_t n;
    FastDiv dx, dy;                 // cell decode: divide by nx, ny
    const uint8_t* code;            // level 0
    const float* diag;              // coarser levels
    const float* idiag;             // 1 / diag on the unknowns, 0 elsewhere (coarser levels)
    const float* wp[3];
};
constexpr int kMgMaxLevels = 32;
constexpr uint32_t kMgCoarsest = 512;   // at most this many cells on the level solved directly
struct MgWork {
    int levels = 0;
    MgLevel lv[kMgMaxLevels]{};
    // per level: right-hand side (levels > 0), pre-smoothed iterate (level 0), iterate after the
    // first post sweep (level 0), result (levels > 0)
    float* vec[kMgMaxLevels][4]{};
    double* dense = nullptr;        // inverse of the coarsest operator, n x n (symmetric)
    float* z = nullptr;             // the preconditioned residual (n + 1, zero pad)
    // one iteration (V-cycle, direction, update) captured as a HIP graph per solution pair and
    // tolerance: a kernel launch costs the host about as long as the kernels of a coarse level
    // take on the device, a graph launch replays all thirteen for about the cost of one
    struct Graph {
        const float* xh;
        const float* xl;
        double tol2;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
};

}  // namespace ds

struct ds_context {
    ds_config config{};
    std::vector<ds_collider> colliders;
    uint32_t mode = DS_MODE_FORWARD;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;

    ds::Geo geo{};
    ds::VBound vb{};
    ds::SBound pb{}, db{}, tb{};
    // device copies read by the kernels (a runtime index into a by-value kernel argument would
    // move the argument to scratch memory)
    const ds::VBound* d_vb = nullptr;
    const ds::SBound* d_db = nullptr;
    const ds::SBound* d_tb = nullptr;
    bool has_solids = false;
    bool has_fixed_pressure = false;
    bool pressure_periodic[3]{};
    uint64_t anchor = 0;
    uint32_t extra_anchors = 0;
    uint64_t unknowns = 0, solids = 0;
    double tolerance = 1e-6;
    uint32_t max_iterations = 1000;
    uint32_t replacement_interval = 50;

    // domain (device)
    uint8_t* mask = nullptr;          // 1 = solid
    uint8_t* code = nullptr;          // diagonal count of an unknown, 0 = not an unknown
    float* fixed_term = nullptr;      // sum of fixed-value pressure faces; null when none
    float* collider_velocity[3]{};    // face collider velocity; null when no colliders
    float* collider_density = nullptr;
    float* collider_temperature = nullptr;

    // forward scratch (device)
    float* force[3]{};
    float* advected[3]{};
    float* pressure_hi = nullptr;     // solution of the last forward solve (hi + lo)
    float* pressure_lo = nullptr;
    bool pressure_valid = false;
    // DS_WARM_START_EXTRAPOLATE: the two solutions before the last one (p_{n-2} at [0, n), p_{n-3}
    // at [n, 2n)), how many of them hold a solution of the current sweep, and the extrapolated guess
    float* pressure_prev = nullptr;
    int pressure_prev_count = 0;
    float* guess = nullptr;
    ds::SolveWindow window[2];        // DS_WARM_START_PROJECTION: forward and adjoint sequences

    // adjoint scratch (device, differentiable mode)
    float* sourced_adj[2]{};          // adjoints of rho_s, T_s
    float* projected_extra[3]{};      // scalar-advection contributions to the projected-velocity adjoint
    float* forced_adj[3]{};           // adjoint of the forced velocity u*
    float* raw_adj[3]{};              // adjoint of the raw advected velocity
    // velocity advection reverse (differentiable mode, power-of-two cell size): the start point's
    // velocity adjoint per face of axis A (first index), one array per velocity component
    float* vjp_start[3][3]{};
    float* lambda_hi = nullptr;       // adjoint pressure solve solution (hi + lo)
    float* lambda_lo = nullptr;
    bool lambda_valid = false;
    double* param_partials = nullptr; // 4 * kForceBlocks
    double* metric_partials = nullptr; // 6 * kMetricBlocks (ds_example_reduce_metrics)

    // forward-mode (tangent) scratch (device, differentiable mode)
    float* sourced_t[2]{};            // tangents of rho_s, T_s
    float* force_t[3]{};              // force tangent
    float* forced_t[3]{};             // tangent of the forced velocity u*
    float* advected_t[3]{};           // tangent of the advected, constrained velocity
    float* tangent_hi = nullptr;      // pressure tangent (hi + lo)
    float* tangent_lo = nullptr;
    bool tangent_valid = false;
    // boundaries with every fixed value zeroed: the tangent of a constant is 0
    ds::VBound vb_zero{};
    ds::SBound db_zero{}, tb_zero{};
    const ds::VBound* d_vb_zero = nullptr;
    const ds::SBound* d_db_zero = nullptr;
    const ds::SBound* d_tb_zero = nullptr;
    // pressure boundary (device) for the red-black Gauss-Seidel mode, and its zeroed copy
    const ds::SBound* d_pb = nullptr;
    const ds::SBound* d_pb_zero = nullptr;
    ds::SBound pb_zero{};
    uint8_t* rbgs_diag = nullptr;     // RBGS diagonal per cell, 0 = not updated (DS_PROJECTION_RBGS)

    // vorticity confinement (device; allocated when vorticity_confinement_enabled)
    bool vorticity = false;
    float* vort_w[3]{};               // vorticity of the step's input velocity
    float* vort_mag = nullptr;        // smoothed magnitude sqrt(|w|^2 + eps^2)
    float* vort_ga[3]{};              // adjoint of the magnitude gradient (differentiable mode)
    float* vort_t[3]{};               // vorticity adjoint (differentiable mode)
    float* vort_wt[3]{};              // vorticity tangent (differentiable mode)
    float* vort_mt = nullptr;         // magnitude tangent (differentiable mode)

    ds::PcgWork pcg;
    ds::MgWork mg;                    // DS_CG_MULTIGRID hierarchy
    ds::Profiler profiler;
    ds::UNet* unet = nullptr;         // warm-start network, loaded by ds_warm_start_load

    std::vector<void*> allocations;   // everything above, released at destroy
};

namespace ds {

template <typename T>
T* device_alloc(ds_context* c, size_t count) {
    void* p = nullptr;
    DS_HIP(hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 1));
    c->allocations.push_back(p);
    return static_cast<T*>(p);
}

void fill_zero(ds_context* c, float* p, uint64_t count);
void fill_zero_d(ds_context* c, double* p, uint64_t count);

// PCG on the context's pressure matrix: solves A x = b where b is already in c->pcg.b.
// x is the compensated pair (hi, lo). x0 may be null. Returns the solve info.
ds_solve_info pcg_solve(ds_context* c, float* x_hi, float* x_lo, const float* x0, double tol, uint32_t max_it);

// Solves A x = b with b built from `u` (kind 1: forward divergence of the advected velocity;
// kind 2: adjoint, u = projected-velocity adjoint plus `extra`; kind 3: tangent, the forward
// divergence without the fixed-face term) or given in c->pcg.b (kind 0).
// correct: under DS_WARM_START_CNN the network corrects the given guess (x0h, x0l).
ds_solve_info pcg_solve_kind(ds_context* c, int kind, const float* const u[3], const float* const extra[3], float* xh,
                             float* xl, const float* x0h, const float* x0l, double tol, uint32_t max_it, bool correct = false);
uint32_t pcg_periodic_bits(const ds_context* c);
// DS_WARM_START_PROJECTION (window.hip): the window of a right-hand-side kind (null when the mode
// is off or the kind has none), the guess for right-hand side b (null while the window is empty),
// adding a solution, clearing both windows, allocating one.
SolveWindow* window_of(ds_context* c, int kind);
const float* window_guess(ds_context* c, SolveWindow& w, const float* b);
void window_add(ds_context* c, SolveWindow& w, const float* hi, const float* lo);
void window_reset(ds_context* c);
void window_allocate(ds_context* c, SolveWindow& w);
// Decides whether the single-reduction solve runs resident (full 64 x kTY tiles, at most 8 planes
// per block, cooperative launch possible with every block co-resident) and allocates its state.
void pcg_resident_prepare(ds_context* c);
// DS_CG_MULTIGRID (pcg_mg.hip): builds the hierarchy from the host copy of the code; then per
// iteration z = M^-1 r (the V-cycle, <r, z> partials into pcg.partials), the direction kernel
// (convergence test, p = z + beta p_old, <p, A p> partials) and the update kernel (solution pair
// x = (xh, xl), residual, ||r||^2 partials).
// The kernels take the iteration index from the status (PcgStatus::k, advanced by the V-cycle's
// first kernel), so one captured iteration replays for every k: mg_set_iteration sets it (the
// V-cycle of iteration k runs after k - 1), mg_iteration launches the captured iteration (captured
// on first use for the pair and tolerance), mg_release destroys the captured graphs.
void mg_prepare(ds_context* c, const std::vector<uint8_t>& code);
void mg_precondition(ds_context* c, const float* r, float* z);
void mg_direction(ds_context* c, const float* z, double tol2, bool check_only);
void mg_update(ds_context* c, float* xh, float* xl);
void mg_set_iteration(ds_context* c, uint32_t k_before);
void mg_iteration(ds_context* c, float* xh, float* xl, double tol2);
void mg_release(ds_context* c);
// DS_PROJECTION_RBGS (rbgs.hip): the diagonal table, then the solve of a right-hand-side kind
// (1 forward, 2 the unrolled reverse, 3 tangent) into xh with xl = 0.
void rbgs_prepare(ds_context* c);
ds_solve_info rbgs_solve_kind(ds_context* c, int kind, const float* const u[3], const float* const extra[3], float* xh, float* xl);
// out = hi + lo (rounded to FP32).
void combine_pair(ds_context* c, const float* hi, const float* lo, float* out);

// Launch wrappers shared between translation units.
void launch_apply_matrix(ds_context* c, const float* x, float* y);

// Warm-start network (unet.hip).
uint64_t unet_parameter_count();
void unet_load(ds_context* c, const float* params, uint64_t count, double output_scale);
// Runs the network on the solve's grid coarsened 2^levels times (0, 1, 2).
void unet_configure(ds_context* c, uint32_t levels);
// The network's guess for A x = b (device b), corrected from the base guess xb when not null.
const float* unet_infer(ds_context* c, const float* b, const float* xb);
bool unet_loaded(const ds_context* c);
void unet_release(UNet* u);

}  // namespace ds
