/*
 * ds_smoke.h — C ABI of the MI355X (gfx950) smoke solver: forward step, hand-written VJP,
 * preconditioned-CG pressure projection and its adjoint, checkpointed trajectory gradient.
 *
 * Every entry point takes plain pointers and sizes. Field pointers are DEVICE pointers in the
 * reference's layout (cuda_types.h:7-129, operators.cu:36-47): one allocation per component,
 * x-fastest indexing x + nx*(y + ny*z); axis-a face arrays have extent n_a + 1 on their own axis.
 * Cell fields hold nx*ny*nz floats, face component a holds ds_face_count(a) floats.
 *
 * Element types follow the build's precision rule: primal fields, stencils and adjoint fields
 * are FP32 (the reference's adjoints are FP64, data.ixx:101-115); CG dot products, residual
 * norms and the four scalar parameter adjoints are FP64.
 *
 * Errors: every function returns a ds_status (0 = success). ds_last_error() returns a
 * thread-local description of the last failure (replaces the std::runtime_error thrown by
 * check_launch, operators.cu:28-30, and by xayah::cuda::Resource, cuda.cpp:13-66).
 *
 * Ownership (operators.h:40-72, model.cpp:185-339): the caller owns states, controls,
 * parameters, step caches and adjoints. The context owns its scratch (PCG vectors, adjoint
 * scratch, domain masks). No step function allocates. Adjoint outputs are ACCUMULATED (+=)
 * into caller buffers, as in the reference (model.cpp:279-287, operators.cu:297-302).
 * All work is stream-ordered on the context stream; one context per host thread.
 */
#ifndef DS_SMOKE_H
#define DS_SMOKE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ds_status;
enum {
    DS_OK = 0,
    DS_ERROR_INVALID_ARGUMENT = 1,
    DS_ERROR_OUT_OF_MEMORY = 2,
    DS_ERROR_LAUNCH = 3,
    DS_ERROR_NOT_CONVERGED = 4,
    DS_ERROR_NON_FINITE = 5,
    DS_ERROR_RUNTIME = 6,
    DS_ERROR_UNSUPPORTED = 7
};

/* Boundary modes, values as the reference enums (data.ixx:14-25). Face order for modes[] and
 * values[]: x_min, x_max, y_min, y_max, z_min, z_max (operators.cpp:85-105). Velocity value of
 * face f, component a is values[3*f + a]. */
enum { DS_SCALAR_FIXED_VALUE = 0, DS_SCALAR_ZERO_GRADIENT = 1, DS_SCALAR_PERIODIC = 2 };
enum {
    DS_VELOCITY_FIXED_VALUE = 0,
    DS_VELOCITY_ZERO_GRADIENT = 1,
    DS_VELOCITY_NORMAL_FIXED_TANGENT_ZERO_GRADIENT = 2,
    DS_VELOCITY_PERIODIC = 3
};

typedef struct ds_scalar_boundary {   /* cuda_types.h:21-24 */
    uint32_t modes[6];
    float values[6];
} ds_scalar_boundary;

typedef struct ds_velocity_boundary { /* cuda_types.h:26-29 */
    uint32_t modes[6];
    float values[18];
} ds_velocity_boundary;

/* Collider (data.ixx:55-70): shape 0 = ellipsoid (size = radii), 1 = box (size = half extents). */
enum { DS_SHAPE_ELLIPSOID = 0, DS_SHAPE_BOX = 1 };
typedef struct ds_collider {
    uint32_t shape;
    float center[3];
    float size[3];
    float velocity[3];
    float density;
    float temperature;
} ds_collider;

/* Pressure solver. PCG is the build's solver (preconditioned CG on the reference's matrix,
 * operators.cu:852-902; the preconditioner per cg_variant); RBGS is the reference's fixed-count red-black Gauss-Seidel
 * (operators.cu:885-902,1004-1013), kept as a reference mode. */
enum { DS_PROJECTION_PCG = 0, DS_PROJECTION_RBGS = 1 };

/* PCG form: multigrid-preconditioned (what the C++ facade and the Python host select by default),
 * or Jacobi-preconditioned in one of four recurrences: s-step (two iterations per fused kernel and
 * per global reduction), single-reduction (Chronopoulos-Gear: one fused kernel and one global
 * reduction per iteration; the value of a zero-initialised ds_config), its resident variant, or the
 * two-reduction textbook form (two kernels per iteration). The Jacobi forms differ only in rounding;
 * every form solves the same system to the same true-residual test. */
/* DS_CG_SINGLE_REDUCTION: one FP64 reduction and one fused kernel launch per iteration.
 * DS_CG_SINGLE_REDUCTION_RESIDENT: the same recurrence as one cooperative launch per
 * residual-replacement interval with the vectors held on chip, where the grid allows it (nx a
 * multiple of 64, ny of 8, at most 8 planes per block, every block co-resident), else as the
 * default; faster on small grids (64^3), slower at 128^3. DS_CG_TWO_REDUCTION: the classic
 * two-kernel form. */
/* DS_CG_S_STEP: the same Jacobi-preconditioned CG two iterations per kernel launch and per
 * reduction (s-step CG, s = 2: the basis [p, T p, T^2 p, z, T z], T = D^-1 A, is stored and its
 * D-weighted Gram matrix gives both iterations' scalars; about 1.7x the one-step form's iterations
 * per second at 128^3). The stopping test runs every second iteration, so iteration counts are
 * even except at an odd iteration cap. */
/* DS_CG_MULTIGRID: CG preconditioned by an aggregation multigrid V-cycle instead of Jacobi (2 x 2 x 2
 * aggregates, Galerkin coarse operators built at context creation, two damped Jacobi sweeps before
 * and after each coarse correction, the coarsest level of at most 512 cells solved exactly). Same
 * stopping test on the true residual; a few tens of iterations where the Jacobi forms need about a
 * thousand (128^3 plume). One iteration is captured as a HIP graph on the context's stream. No
 * counterpart in the reference, which iterates red-black Gauss-Seidel (DS_PROJECTION_RBGS). */
enum { DS_CG_SINGLE_REDUCTION = 0, DS_CG_TWO_REDUCTION = 1, DS_CG_SINGLE_REDUCTION_RESIDENT = 2, DS_CG_S_STEP = 3, DS_CG_MULTIGRID = 4 };

/* Initial guess of the forward and adjoint pressure solves. */
/* DS_WARM_START_EXTRAPOLATE: each forward pressure solve of a sweep starts from the quadratic
 * extrapolation 3 (p_{n-1} - p_{n-2}) + p_{n-3} of its last three solutions (the linear one,
 * 2 p_{n-1} - p_{n-2}, or the previous solution while fewer exist); adjoint solves start from the
 * previous adjoint solution, as under DS_WARM_START_PREVIOUS. Like every given guess it is dropped
 * (x0 = 0) when its residual exceeds ||b||. A solve started from the step cache's pressure_guess
 * restarts the sequence, and so does every trajectory entry point. */
/* DS_WARM_START_PROJECTION: each forward and each adjoint solve starts from the A-norm-best
 * combination of the last 8 solutions of its sequence (Galerkin projection onto their span). */
/* DS_WARM_START_CNN: the warm-start network corrects the forward solves' sequence guess of
 * DS_WARM_START_EXTRAPOLATE: with x_b that guess, the network maps the residual r = b - A x_b to a
 * direction d and the solve starts from x_b + gamma d, gamma = <d, r> / <d, A d> (the A-norm-best
 * step along d), never worse in the A-norm than x_b alone. Adjoint solves start from the previous
 * adjoint solution, uncorrected. A solve without a base guess (the first of a sequence, the
 * per-operator projection entry points) starts from the network's direction for b alone, at its
 * best multiple.
 * DS_WARM_START_CNN_ALONE: every solve starts from the network's direction for b alone. */
enum {
    DS_WARM_START_NONE = 0,
    DS_WARM_START_PREVIOUS = 1,
    DS_WARM_START_CNN = 2,
    DS_WARM_START_EXTRAPOLATE = 3,
    DS_WARM_START_PROJECTION = 4,
    DS_WARM_START_CNN_ALONE = 5
};

/* smoke::Configuration (data.ixx:72-83) plus the solver fields this build adds. */
typedef struct ds_config {
    uint32_t resolution[3];
    float cell_size;
    float time_step;
    uint32_t pressure_iterations;           /* RBGS sweeps (reference mode only) */
    ds_velocity_boundary velocity_boundary;
    ds_scalar_boundary pressure_boundary;
    ds_scalar_boundary density_boundary;
    ds_scalar_boundary temperature_boundary;
    const ds_collider* colliders;           /* host array, read at context creation only */
    uint32_t collider_count;
    uint32_t vorticity_confinement_enabled; /* adds the confinement force (data.ixx:82) */
    uint32_t projection_solver;             /* DS_PROJECTION_* */
    uint32_t cg_max_iterations;             /* 0 selects 4 * (nx + ny + nz) + 1000 */
    double cg_tolerance;                    /* ||b - A x||_2 <= tol * ||b||_2, FP64 norms; 0 selects 1e-6 */
    uint32_t warm_start;                    /* DS_WARM_START_* */
    uint32_t residual_replacement_interval; /* true-residual refresh period, 0 selects 50 */
    uint32_t cg_variant;                    /* DS_CG_* recurrence of the PCG iteration */
} ds_config;

/* ExecutionMode (model.ixx:10-13). */
enum { DS_MODE_FORWARD = 0, DS_MODE_DIFFERENTIABLE = 1 };

/* Field views: device pointers (data.ixx:117-134, 155-196). */
typedef struct ds_state {
    float* density;
    float* temperature;
    float* velocity[3];
} ds_state;

typedef struct ds_control {
    float* density_source;
    float* temperature_source;
    float* external_acceleration[3];
} ds_control;

typedef struct ds_parameters {          /* each a device pointer to ONE float */
    float* ambient_temperature;
    float* density_buoyancy;
    float* temperature_buoyancy;
    float* vorticity_confinement;
} ds_parameters;

/* StepCache (data.ixx:189-196) without the vorticity cache (the reverse recomputes the vorticity
 * from the step's input velocity, one stencil pass). force and advected_velocity may be
 * NULL (context scratch is used; the VJP does not read them). The pressure is kept as a
 * compensated FP32 pair (hi then lo, 2 * cell_count floats): pressure (may be NULL) receives
 * the converged pair of the forward step; pressure_guess (may be NULL) is read as the initial
 * guess of its pressure solve, e.g. the pair a first pass over the same step stored. */
typedef struct ds_step_cache {
    float* sourced_density;
    float* sourced_temperature;
    float* force[3];
    float* forced_velocity[3];
    float* advected_velocity[3];
    float* pressure;
    const float* pressure_guess;
} ds_step_cache;

typedef struct ds_state_adjoint {
    float* density;
    float* temperature;
    float* velocity[3];
} ds_state_adjoint;

typedef struct ds_control_adjoint {
    float* density_source;
    float* temperature_source;
    float* external_acceleration[3];
} ds_control_adjoint;

/* StateTangent / ControlTangent / ParameterTangent (data.ixx:136-153): FP32 fields with the layout
 * of their primal twins. */
typedef ds_state ds_state_tangent;
typedef ds_control ds_control_tangent;
typedef ds_parameters ds_parameter_tangent;

typedef struct ds_parameter_adjoint {   /* each a device pointer to ONE double */
    double* ambient_temperature;
    double* density_buoyancy;
    double* temperature_buoyancy;
    double* vorticity_confinement;
} ds_parameter_adjoint;

typedef struct ds_solve_info {
    uint32_t iterations;
    uint32_t converged;
    uint32_t warm_started;              /* 1 when a non-zero initial guess was used */
    uint32_t guess_rejected;            /* 1 when the guess was non-finite or worse than x0 = 0 */
    double rhs_norm;                    /* ||b||_2 */
    double initial_residual;            /* ||b - A x0||_2 */
    double final_residual;              /* true residual at exit */
} ds_solve_info;

typedef struct ds_step_info {
    ds_solve_info projection;
} ds_step_info;

typedef struct ds_trajectory_info {
    uint32_t steps;
    uint32_t checkpoint_interval;
    uint32_t forward_solves, recompute_solves, adjoint_solves;
    uint32_t not_converged;
    uint64_t forward_iterations, recompute_iterations, adjoint_iterations;
    uint32_t max_iterations;
    uint32_t guesses_rejected;
    double max_relative_residual;
    double loss;
} ds_trajectory_info;

typedef struct ds_domain_info {
    uint64_t cell_count;
    uint64_t face_count[3];
    uint64_t unknown_count;             /* PCG unknowns: fluid cells minus anchors */
    uint64_t solid_count;
    uint64_t pressure_anchor;           /* reference anchor (model.cpp:72-77); cell_count if none */
    uint32_t extra_anchor_count;        /* anchors of fluid pockets cut off from the main anchor */
    uint32_t has_fixed_pressure;
} ds_domain_info;

typedef struct ds_context ds_context;

/* ---- error reporting ------------------------------------------------------------------- */
const char* ds_last_error(void);
/* Number of cells / faces of a configuration without a context (sizes for the caller's buffers). */
uint64_t ds_config_cell_count(const ds_config* config);
uint64_t ds_config_face_count(const ds_config* config, uint32_t axis);

/* ---- context and device resource (cuda.ixx:7-28, model.cpp:32-182) --------------------- */
/* Model::allocate_context: validates the configuration, rasterises colliders, picks the
 * pressure anchor, builds the per-cell connectivity code and allocates scratch. stream may be
 * NULL (the context creates its own non-blocking stream) or a hipStream_t to run on. */
ds_status ds_context_create(const ds_config* config, uint32_t mode, void* stream, ds_context** out);
ds_status ds_context_destroy(ds_context* context);
void* ds_context_stream(ds_context* context);
ds_status ds_context_domain_info(ds_context* context, ds_domain_info* info);
/* Copies the u8 solid mask and the u8 connectivity code (diagonal count, 0 = not an unknown)
 * to host arrays of cell_count bytes. Either pointer may be NULL. Synchronises. */
ds_status ds_context_domain_masks(ds_context* context, uint8_t* mask, uint8_t* code);
/* Resource::allocate / release / copy_* / zero / synchronize (cuda.cpp:39-66). */
/* The last solution (hi + lo, rounded to FP32; DEVICE cell field) of the forward pressure
 * solves (which = 0) or of the adjoint solves of ds_vjp_step (which = 1): the training pairs of
 * the warm-start network are formed from them. No counterpart in the reference. */
ds_status ds_context_last_solution(ds_context* context, uint32_t which, float* out);
ds_status ds_allocate(ds_context* context, size_t bytes, void** out);
ds_status ds_release(ds_context* context, void* pointer);
ds_status ds_copy_from_host(ds_context* context, void* device, const void* host, size_t bytes);
ds_status ds_copy_to_host(ds_context* context, void* host, const void* device, size_t bytes);
ds_status ds_copy_device(ds_context* context, void* destination, const void* source, size_t bytes);
ds_status ds_zero(ds_context* context, void* device, size_t bytes);
ds_status ds_synchronize(ds_context* context);

/* ---- operators (cuda_kernels launchers, operators.h:40-72) ----------------------------- */
/* source_forward (operators.cu:286-290): out = mask ? 0 : state + dt * source. */
ds_status ds_source_forward(ds_context* c, const float* state, const float* source, float* out);
/* source_vjp (operators.cu:297-302): state_adj += out_adj, source_adj += dt * out_adj (fluid). */
ds_status ds_source_vjp(ds_context* c, const float* out_adj, float* state_adj, float* source_adj);
/* buoyancy_forward (operators.cu:304-313). */
ds_status ds_buoyancy_forward(ds_context* c, const float* density, const float* temperature,
                              float* const external_acceleration[3], const ds_parameters* params,
                              float* force[3]);
/* buoyancy_vjp (operators.cu:326-337); parameter adjoints are FP64 dot products. */
ds_status ds_buoyancy_vjp(ds_context* c, const float* density, const float* temperature,
                          const ds_parameters* params, float* const force_adj[3], float* density_adj,
                          float* temperature_adj, float* external_acceleration_adj[3],
                          const ds_parameter_adjoint* param_adj);
/* vorticity_forward (operators.cu:973-979): force += confinement * h * (N x w) on fluid cells,
 * w = curl of the cell-centred velocity, N = normalised gradient of the smoothed |w|. Needs a
 * context created with vorticity_confinement_enabled = 1; reads params->vorticity_confinement. */
ds_status ds_vorticity_forward(ds_context* c, float* const velocity[3], const ds_parameters* params,
                               float* force[3]);
/* vorticity_vjp (operators.cu:990-1002): velocity_adj += the velocity adjoint of the force above
 * for the cell force adjoint force_adj; param_adj->vorticity_confinement (FP64) += its
 * confinement adjoint when param_adj is not NULL. Differentiable context with confinement on. */
ds_status ds_vorticity_vjp(ds_context* c, float* const velocity[3], const ds_parameters* params,
                           float* const force_adj[3], float* velocity_adj[3],
                           const ds_parameter_adjoint* param_adj);
/* The tangent operators (the *_jvp launchers, operators.h:40-72). source_jvp and
 * integrate_velocity_jvp are the forwards on tangents (operators.cu:575-577, 604-606): call
 * ds_source_forward / ds_integrate_velocity_forward with the tangents. */
/* buoyancy_jvp (operators.cu:315-324, 589-592). */
ds_status ds_buoyancy_jvp(ds_context* c, const float* density, const float* temperature,
                          const float* density_tangent, const float* temperature_tangent,
                          float* const external_acceleration_tangent[3], const ds_parameters* params,
                          const ds_parameter_tangent* param_tangent, float* force_tangent[3]);
/* vorticity_jvp (operators.cu:981-988): force_tangent += the confinement-force tangent.
 * Differentiable context with confinement on. */
ds_status ds_vorticity_jvp(ds_context* c, float* const velocity[3], float* const velocity_tangent[3],
                           const ds_parameters* params, const ds_parameter_tangent* param_tangent,
                           float* force_tangent[3]);
/* advect_velocity_jvp (operators.cu:390-413, 618-621): raw self-advection tangent. */
ds_status ds_advect_velocity_jvp(ds_context* c, float* const velocity[3],
                                 float* const velocity_tangent[3], float* out_tangent[3]);
/* constrain_velocity_jvp (operators.cu:633-638): the constraint with zero boundary values and no
 * collider velocity. */
ds_status ds_constrain_velocity_jvp(ds_context* c, float* const velocity_tangent[3],
                                    float* out_tangent[3]);
/* advect_scalar_jvp (operators.cu:509-535, 650-653); field as in ds_advect_scalar_forward. */
ds_status ds_advect_scalar_jvp(ds_context* c, uint32_t field, const float* source,
                               const float* source_tangent, float* const velocity[3],
                               float* const velocity_tangent[3], float* out_tangent);
/* projection_jvp (operators.cu:1015-1018) with the PCG solve: the tangent right-hand side has no
 * fixed-face term. Differentiable context. info may be NULL. */
ds_status ds_projection_jvp(ds_context* c, float* const velocity_tangent[3], float* out_tangent[3],
                            ds_solve_info* info);
/* integrate_velocity_forward (operators.cu:339-364). */
ds_status ds_integrate_velocity_forward(ds_context* c, float* const velocity[3], float* const force[3],
                                        float* out[3]);
/* integrate_velocity_vjp (operators.cu:366-379). */
ds_status ds_integrate_velocity_vjp(ds_context* c, float* const out_adj[3], float* velocity_adj[3],
                                    float* force_adj[3]);
/* advect_velocity_forward (operators.cu:381-388): raw self-advection, no constraint. */
ds_status ds_advect_velocity_forward(ds_context* c, float* const velocity[3], float* out[3]);
/* advect_velocity_vjp (operators.cu:415-439). */
ds_status ds_advect_velocity_vjp(ds_context* c, float* const velocity[3], float* const out_adj[3],
                                 float* velocity_adj[3]);
/* constrain_velocity_forward (operators.cu:466-481). */
ds_status ds_constrain_velocity_forward(ds_context* c, float* const velocity[3], float* out[3]);
/* constrain_velocity_vjp (operators.cu:483-497). */
ds_status ds_constrain_velocity_vjp(ds_context* c, float* const out_adj[3], float* velocity_adj[3]);
/* advect_scalar_forward (operators.cu:499-507); field 0 = density, 1 = temperature selects the
 * boundary (operators.cpp:170-172) and the collider value. */
ds_status ds_advect_scalar_forward(ds_context* c, uint32_t field, const float* source,
                                   float* const velocity[3], float* out);
/* advect_scalar_vjp (operators.cu:537-561). */
ds_status ds_advect_scalar_vjp(ds_context* c, uint32_t field, const float* source,
                               float* const velocity[3], const float* out_adj, float* source_adj,
                               float* velocity_adj[3]);
/* projection_forward (operators.cu:1004-1013) with the pressure solve replaced by Jacobi-PCG to
 * the configured relative residual. pressure and rhs are cell fields (rhs receives b of A p = b).
 * guess may be NULL. info may be NULL. */
ds_status ds_projection_forward(ds_context* c, float* const velocity[3], float* pressure, float* rhs,
                                const float* guess, float* out[3], ds_solve_info* info);
/* projection_vjp (operators.cu:1020-1031) with the unrolled RBGS reverse replaced by the adjoint
 * PCG solve A lambda = p_bar on the same SPD matrix. velocity_adj += (transpose of projection). */
ds_status ds_projection_vjp(ds_context* c, float* const out_adj[3], float* velocity_adj[3],
                            ds_solve_info* info);
/* Jacobi-PCG on the projection matrix A (operators.cu:852-902): solves A x = b. x0 may be NULL. */
ds_status ds_pcg_solve(ds_context* c, const float* b, float* x, const float* x0, double tolerance,
                       uint32_t max_iterations, ds_solve_info* info);
/* y = A x on unknowns (0 elsewhere); for tests. */
ds_status ds_apply_pressure_matrix(ds_context* c, const float* x, float* y);
/* dst = src for count floats on the context's stream, one 4-byte load and store per lane (the
 * access width of the PCG kernels): a launch of known traffic that scales the hardware's memory
 * counters to bytes. No counterpart in the reference. */
ds_status ds_stream_copy(ds_context* c, const float* src, float* dst, uint64_t count);

/* ---- warm-start network (DS_WARM_START_CNN; no counterpart in the reference, whose projection
 * starts every solve from p = 0, operators.cu:1005) ------------------------------------------ */
/* Number of FP32 parameters of the network (a 3-level 3-D U-Net, 16/32/64 channels). */
uint64_t ds_warm_start_parameter_count(void);
/* Loads the network's parameters (FP32, PyTorch state_dict order of the 3-level U-Net: per layer
 * the weight, then the bias; ds_unet_trainer_read writes them in this order). Needs nx % 64 == 0,
 * ny % 16 == 0, nz % 4 == 0. output_scale (> 0) is kept for compatibility: every guess is applied
 * at its A-norm-best multiple, so a scale of the network's output cancels. A non-finite guess is
 * rejected (x0 = 0). */
ds_status ds_warm_start_load(ds_context* c, const float* params, uint64_t count, double output_scale);
/* The grid the network runs on: the solve's grid coarsened 2^levels times per axis (levels 0, 1
 * or 2; default 0). With levels > 0 the network's input is restricted (sum over each coarse cell's
 * children) and its output prolonged back (cell-centred trilinear), so a network trained on 64^3
 * problems serves 128^3 (levels 1) or 256^3 (levels 2). The network's grid needs nx % 64 == 0,
 * ny % 16 == 0, nz % 4 == 0. Callable before or after ds_warm_start_load. */
ds_status ds_warm_start_configure(ds_context* c, uint32_t levels);
/* The network's guess x for A x = b from b alone (device cell fields; b is 0 off the unknowns). */
ds_status ds_warm_start_infer(ds_context* c, const float* b, float* x);
/* The network's correction of a base guess: x = base + gamma N(b - A base) at the A-norm-best
 * gamma; base = NULL is ds_warm_start_infer. */
ds_status ds_warm_start_guess(ds_context* c, const float* b, const float* base, float* x);

/* ---- training of the warm-start network on the GPU (no counterpart in the reference) ---------
 * A training sample is an error e on the unknowns (a solution, or a solution minus a base guess);
 * its residual r = A e is the network's input (as r / rms(r), with the unknown indicator) and the
 * network's output g is applied at the A-norm-best step gamma = <g, r> / <g, A g>. Losses: 0 =
 * energy, 1 - <g, r>^2 / (<g, A g> <e, r>); 1 = L2, ||gamma g - e||^2 / ||e||^2. */
/* r = A e (0 off the unknowns) and energy = <e, A e> (DEVICE double) for the context's matrix. */
ds_status ds_warm_start_pair(ds_context* c, const float* error, float* residual, double* energy);
typedef struct ds_unet_trainer ds_unet_trainer;
/* A trainer over batches of `batch` samples on the context's grid (nx % 64 == 0, ny % 16 == 0,
 * nz % 4 == 0) and matrix, on the context's stream, starting from params (HOST, FP32,
 * ds_warm_start_parameter_count values). Adam state starts at zero. */
ds_status ds_unet_trainer_create(ds_context* c, uint32_t batch, const float* params, uint64_t count, ds_unet_trainer** out);
ds_status ds_unet_trainer_destroy(ds_unet_trainer* t);
/* One step on `batch` errors (DEVICE FP32, batch x cells, sample-major, 0 off the unknowns):
 * residuals, forward pass, loss of loss_kind (DEVICE, batch doubles, per sample), gradient of the
 * batch mean by the backward pass; with learning_rate > 0 then one Adam step (beta1 0.9, beta2
 * 0.999, eps 1e-8) and the repacking of the BF16 weights. */
ds_status ds_unet_trainer_step(ds_unet_trainer* t, const float* error, uint32_t loss_kind, float learning_rate, double* loss);
/* Copies (DEVICE FP32) what = 0: the parameters, 1: the gradient of the last step (both
 * ds_warm_start_parameter_count values), 2: the network output g of the last step (batch x cells). */
ds_status ds_unet_trainer_read(ds_unet_trainer* t, uint32_t what, float* out);

/* The network's operators on BF16 channels-last tensors (DEVICE; `batch` samples stacked along z;
 * FP32 weights in PyTorch layout), for the network's layer shapes only: 3x3x3 convolutions
 * (cin1 + cin2 -> cout) 2->16, 16->16, 16->32, 32->32, 32->64, 64->64, 32+32->32, 16+16->16 (in2,
 * cin2: the concatenated skip connection); 2x2x2 stride-2 transposed convolutions 64->32, 32->16
 * (grid = the coarse input grid); 2x2x2 max pooling of 16 or 32 channels (grid = the pooled grid).
 * nx % 16 == 0 (and ny % 4 == 0 for the 3x3x3 convolution). The 2-channel input is stored with 4
 * channels (the last two 0). Results are BF16 (round to nearest even) unless stated. Each call
 * allocates its packing and partial-sum buffers and waits for its stream. */
ds_status ds_unet_conv3_forward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin1, uint32_t cin2, uint32_t cout,
                                uint32_t relu, const uint16_t* in1, const uint16_t* in2, const float* weight, const float* bias,
                                uint16_t* out, void* stream);
/* Input gradient from the output gradient dz: cin2 = 0: din1 = (input_relu > 0 or no mask when
 * NULL) * dgrad; cin2 > 0: din1 = the first cin1 channels (BF16), din2 = the last cin2 (FP32). */
ds_status ds_unet_conv3_dgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin1, uint32_t cin2, uint32_t cout,
                              const uint16_t* dz, const float* weight, const uint16_t* input_relu, uint16_t* din1, float* din2,
                              void* stream);
/* Weight (PyTorch layout) and bias gradients (FP32, overwritten). */
ds_status ds_unet_conv3_wgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin1, uint32_t cin2, uint32_t cout,
                              const uint16_t* in1, const uint16_t* in2, const uint16_t* dz, float* dweight, float* dbias, void* stream);
ds_status ds_unet_tconv2_forward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin, uint32_t cout, const uint16_t* in,
                                 const float* weight, const float* bias, uint16_t* out, void* stream);
/* din = (input_relu > 0) * dgrad. */
ds_status ds_unet_tconv2_dgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin, uint32_t cout, const uint16_t* dout,
                               const float* weight, const uint16_t* input_relu, uint16_t* din, void* stream);
ds_status ds_unet_tconv2_wgrad(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t cin, uint32_t cout, const uint16_t* in,
                               const uint16_t* dout, float* dweight, float* dbias, void* stream);
ds_status ds_unet_maxpool_forward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t c, const uint16_t* in, uint16_t* out,
                                  void* stream);
/* din = (in > 0) * (skip_gradient (FP32, NULL: 0) + dout routed to the first maximum of each
 * 2x2x2 group in the order dx + 2 dy + 4 dz). */
ds_status ds_unet_maxpool_backward(uint32_t batch, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t c, const uint16_t* in,
                                   const uint16_t* dout, const float* skip_gradient, uint16_t* din, void* stream);

/* ---- kernel timing (no counterpart in the reference, which times whole steps with
 * std::chrono, examples/smoke/forward/simulation.cpp:91,107-119) -------------------------- */
/* Brackets every `sample_every`-th launch of the two PCG iteration kernels with HIP events on
 * the context stream (0 disables). */
typedef struct ds_kernel_timing {
    char name[48];
    uint64_t launches;   /* launches since profiling was enabled */
    uint64_t samples;    /* launches bracketed by events */
    double total_ms;     /* summed event time of the sampled launches */
} ds_kernel_timing;
ds_status ds_profile_enable(ds_context* c, uint32_t sample_every);
/* Synchronises, writes up to `capacity` entries and their number, and clears the samples. */
ds_status ds_profile_read(ds_context* c, ds_kernel_timing* out, uint32_t capacity, uint32_t* count);

/* ---- model steps (model.ixx:57-92, model.cpp:332-389) ---------------------------------- */
ds_status ds_copy_state(ds_context* c, const ds_state* source, ds_state* destination);
ds_status ds_copy_state_adjoint(ds_context* c, const ds_state_adjoint* source, ds_state_adjoint* destination);
ds_status ds_accumulate_state_adjoint(ds_context* c, const ds_state_adjoint* source, ds_state_adjoint* destination);
/* Model::forward_step (model.cpp:341-349). info may be NULL. */
ds_status ds_forward_step(ds_context* c, const ds_state* state, const ds_control* control,
                          const ds_parameters* params, ds_state* next_state, ds_step_cache* cache,
                          ds_step_info* info);
/* Model::vjp_step (model.cpp:361-389). Reads next_state->velocity and the cache's sourced fields
 * and forced velocity; with vorticity confinement on also state->velocity (the vorticity fields
 * are recomputed from it instead of being cached). prev_adj, control_adj and param_adj are
 * accumulated (+=). */
ds_status ds_vjp_step(ds_context* c, const ds_state* state, const ds_control* control,
                      const ds_parameters* params, const ds_state* next_state,
                      const ds_step_cache* cache, const ds_state_adjoint* next_adj,
                      ds_state_adjoint* prev_adj, ds_control_adjoint* control_adj,
                      const ds_parameter_adjoint* param_adj, ds_step_info* info);
/* Model::jvp_step (model.cpp:351-359): next_state_tangent = the tangent of the next state for
 * the tangents of the state, the control and the parameters (param_tangent: four device floats,
 * as ds_parameters). state, next_state and cache are those of ds_forward_step on the same
 * inputs (the cache's sourced fields and forced velocity are read; state->velocity with
 * vorticity confinement on). Needs a differentiable context. The tangent pressure solve starts
 * from the previous tangent solution when warm_start is DS_WARM_START_PREVIOUS. */
ds_status ds_jvp_step(ds_context* c, const ds_state* state, const ds_control* control,
                      const ds_parameters* params, const ds_state* next_state,
                      const ds_step_cache* cache, const ds_state_tangent* state_tangent,
                      const ds_control_tangent* control_tangent,
                      const ds_parameter_tangent* param_tangent,
                      ds_state_tangent* next_state_tangent, ds_step_info* info);

/* ---- trajectory driver (solver.ixx:53-81,115-150) --------------------------------------- */
/* Runs `steps` forward steps from `initial` into `final_state`. controls holds control_count
 * views: 1 (the same control every step) or `steps`. */
ds_status ds_simulate(ds_context* c, const ds_state* initial, const ds_control* controls,
                      uint32_t control_count, const ds_parameters* params, uint32_t steps,
                      ds_state* final_state, ds_trajectory_info* info);
/* Loss 0.5 * ||density_T - target||^2 after `steps` steps and its gradient by a reverse sweep
 * with a checkpoint every `checkpoint_interval` steps and recompute in between (peak memory
 * (steps/k + k) states). initial_adj receives dL/d(initial state) (+=); control_adj (may be
 * NULL) receives the sum over steps of dL/d(control) when control_count == 1, else it points
 * to `steps` views; param_adj (may be NULL) accumulates dL/d(parameters). final_state may be
 * NULL. */
ds_status ds_loss_gradient(ds_context* c, const ds_state* initial, const ds_control* controls,
                           uint32_t control_count, const ds_parameters* params, uint32_t steps,
                           uint32_t checkpoint_interval, const float* target_density,
                           ds_state* final_state, ds_state_adjoint* initial_adj,
                           ds_control_adjoint* control_adj, const ds_parameter_adjoint* param_adj,
                           ds_trajectory_info* info);

/* Reverse sweep for a given adjoint of the final state (dL/d state_T), as solver::vjp
 * (solver.ixx:115-150) with a trajectory adjoint that is zero except at the final state:
 * recomputes the forward pass with a checkpoint every `checkpoint_interval` steps and
 * accumulates (+=) dL/d(initial state) into initial_adj, the control adjoints (as in
 * ds_loss_gradient) and the parameter adjoints. final_state (may be NULL) receives state_T. */
ds_status ds_trajectory_vjp(ds_context* c, const ds_state* initial, const ds_control* controls,
                            uint32_t control_count, const ds_parameters* params, uint32_t steps,
                            uint32_t checkpoint_interval, const ds_state_adjoint* final_adj,
                            ds_state* final_state, ds_state_adjoint* initial_adj,
                            ds_control_adjoint* control_adj, const ds_parameter_adjoint* param_adj,
                            ds_trajectory_info* info);

/* Forward mode over a trajectory, as solver::jvp (solver.ixx:83-113) returns its last element:
 * one sweep that runs ds_forward_step and ds_jvp_step per step, so no trajectory is stored
 * (memory: two states, two tangents, one cache). controls and control_tangents each hold
 * control_count views (1 or `steps`). Writes state_T into final_state and its tangent into
 * final_tangent. The tangent solves are counted in info's adjoint_* fields. */
ds_status ds_trajectory_jvp(ds_context* c, const ds_state* initial, const ds_control* controls,
                            uint32_t control_count, const ds_parameters* params, uint32_t steps,
                            const ds_state_tangent* initial_tangent,
                            const ds_control_tangent* control_tangents,
                            const ds_parameter_tangent* param_tangent, ds_state* final_state,
                            ds_state_tangent* final_tangent, ds_trajectory_info* info);

/* ---- forward example harness (examples/smoke/forward/simulation.{ixx,cpp,cu}) ------------- */

/* The control arguments of launch_write_control (simulation.h:10, ForwardSimulationOptions at
 * simulation.ixx:9-27): two Gaussian jets (centres and radius in domain units) whose amplitudes
 * 0.75 + 0.25 sin(2 pi t / pulse_period [+ pi]) pulse in antiphase. */
typedef struct ds_jet_control {
    float pulse_period;
    float left_source_center[3];
    float right_source_center[3];
    float source_radius;
    float density_source_rate;
    float temperature_source_rate;
    float left_acceleration[3];
    float right_acceleration[3];
} ds_jet_control;

/* write_control_kernel (simulation.cu:35-57) for step `step` (time step * dt): overwrites the
 * density and temperature sources (rate * summed jet weight) and the cell-centred external
 * acceleration (each jet's weight times its vector) on every cell. */
ds_status ds_example_write_control(ds_context* c, uint64_t step, const ds_jet_control* jet, float* density_source,
                                   float* temperature_source, float* const external_acceleration[3]);

/* reduce_metrics_kernel (simulation.cu:59-85) over the fluid cells: metrics (DEVICE, 6 doubles,
 * overwritten) = { sum rho, max rho, max T, max |cell-centred post-projection velocity|,
 * sum div(pre)^2, sum div(post)^2 }, divergence = face differences / h. The maxima are taken
 * with 0 (the reference zeroes the buffer first). Deterministic: a fixed-order two-pass
 * reduction (the reference uses FP64 atomics, simulation.cu:78-83). */
ds_status ds_example_reduce_metrics(ds_context* c, const float* density, const float* temperature,
                                    const float* const pre_projection_velocity[3],
                                    const float* const post_projection_velocity[3], double* metrics);

/* ds_example_reduce_metrics without a context, for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * It allocates nothing: its per-block partials live in a fixed device array of the library
 * (6 x 512 FP64 = 24 KiB per GPU, held until the process ends), which covers every grid whose
 * cell and face indices fit in 32 bits; larger grids are rejected (DS_ERROR_INVALID_ARGUMENT).
 * Only config->resolution and config->cell_size are read; solid_mask (DEVICE, one byte per cell,
 * nonzero = solid, as ds_context_domain_masks returns) may be NULL for no solids; stream may be
 * NULL (the null stream). The result equals ds_example_reduce_metrics's bit for bit. Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller. */
ds_status ds_example_reduce_metrics_static(const ds_config* config, const uint8_t* solid_mask, const float* density,
                                           const float* temperature, const float* const pre_projection_velocity[3],
                                           const float* const post_projection_velocity[3], double* metrics, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DS_SMOKE_H */
