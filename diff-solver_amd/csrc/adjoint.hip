// adjoint.hip — reverse (VJP) operators of the smoke step on gfx950 (solvers/smoke/operators.cu:
// 297-302, 326-337, 366-379, 415-439, 483-497, 537-561, 922-969).
//
// Adjoint fields are FP32. Scatters that follow a data-dependent trace (the transposes of the
// trilinear samples) use hardware FP32 atomics; every transpose with a fixed stencil is
// rewritten as a gather so that it needs no atomics: the projection reverse (p_bar per cell),
// the divergence reverse and the constraint transpose (one face kernel), and the integration
// reverse (force adjoint per cell, fused with the buoyancy and source reverses). Parameter
// adjoints are FP64 block partials reduced in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>

#include "context.hpp"
#include "kernels.hpp"
#include "vorticity.hpp"

namespace ds {

namespace {

// Sink of the three velocity-adjoint components: plain global atomics.
struct VelSink {
    F3 va;
    template <int B>
    __device__ __forceinline__ void add(int, int, int, uint32_t i, float value) const { atomicAdd(va.p[B] + i, value); }
};

// Reverse of trace_rk2 + final sample (operators.cu:422-438 / 544-560): given the adjoint of the
// trace end point (pbx, pby, pbz) = -deriv * dt * grad(sample) * out_adj, scatter through the
// midpoint and start samples of the velocity; mid: the velocity samples at the midpoint with their
// gradients (from trace_rk2<.., true>).
template <typename Sink>
__device__ __forceinline__ void trace_reverse(const Geo& g, const VBound& b, const Sink& sink, float sx, float sy, float sz,
                                              const TraceOut& t, const Sample (&mid)[3], float v1x, float v1y, float v1z) {
    const Sample& m0 = mid[0];
    const Sample& m1 = mid[1];
    const Sample& m2 = mid[2];
    scatter_face_to<0>(sink, t.mx, t.my, t.mz, v1x, g, b);
    scatter_face_to<1>(sink, t.mx, t.my, t.mz, v1y, g, b);
    scatter_face_to<2>(sink, t.mx, t.my, t.mz, v1z, g, b);
    const float mx = m0.gx * v1x + m1.gx * v1y + m2.gx * v1z;
    const float my = m0.gy * v1x + m1.gy * v1y + m2.gy * v1z;
    const float mz = m0.gz * v1x + m1.gz * v1y + m2.gz * v1z;
    const float s = -0.5f * g.dt;
    scatter_face_to<0>(sink, sx, sy, sz, s * mx, g, b);
    scatter_face_to<1>(sink, sx, sy, sz, s * my, g, b);
    scatter_face_to<2>(sink, sx, sy, sz, s * mz, g, b);
}

// advect_scalar_vjp_kernel (operators.cu:537-561) for NF fields that share one trace: the
// velocity adjoint of both fields is summed before it is scattered (the map is linear). A block
// owns a kSvBX x kSvBY x kSvBZ brick of cells and privatises its scatters as the velocity reverse
// does: the NF fields' and the three velocity components' adds that land in a window around the
// brick (halo kSvHalo) are LDS atomics, the rest (long traces, periodic wraps) global atomics; the
// windows are flushed once.
constexpr int kSvBX = 32, kSvBY = 8, kSvBZ = 2, kSvHalo = 1;
constexpr int kSwX = kSvBX + 2 * kSvHalo + 1, kSwY = kSvBY + 2 * kSvHalo + 1, kSwZ = kSvBZ + 2 * kSvHalo + 1;
constexpr int kSw = kSwX * kSwY * kSwZ;   // per target array: velocity components 0-2, fields 3-4

struct ScalarBrickSink {
    float* win;        // LDS, 5 * kSw
    int ox, oy, oz;    // coordinates of window element (0, 0, 0)
    F3 va;
    float* s0;
    float* s1;
    template <int W>
    __device__ __forceinline__ void add(int x, int y, int z, uint32_t i, float value) const {
        const unsigned lx = (unsigned)(x - ox), ly = (unsigned)(y - oy), lz = (unsigned)(z - oz);
        if (lx < (unsigned)kSwX && ly < (unsigned)kSwY && lz < (unsigned)kSwZ)
            __hip_atomic_fetch_add(win + W * kSw + ((int)lz * kSwY + (int)ly) * kSwX + (int)lx, value, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            atomicAdd((W < 3 ? va.p[W < 3 ? W : 0] : W == 3 ? s0 : s1) + i, value);
    }
};

// 4 waves per SIMD (registers capped at 128, the rest spilled), the most its 38.5 KB of LDS allows:
// the dependent gathers of the trace want waves more than registers.
template <int NF, bool HasSolids>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void advect_scalar_vjp_kernel(Geo g, const VBound* __restrict__ vbp, const SBound* __restrict__ sb0p,
                                                                const SBound* __restrict__ sb1p, const uint8_t* __restrict__ mask,
                                                                const float* __restrict__ src0, const float* __restrict__ src1, Vel3 u,
                                                                const float* __restrict__ oa0, const float* __restrict__ oa1,
                                                                float* src_adj0, float* src_adj1, F3 va) {
    static_assert(kSvBX * kSvBY == 256, "one plane of the brick per pass of the block");
    __shared__ float win[5 * kSw];
    const VBound& vb = *vbp;
    const SBound& sb0 = *sb0p;
    const SBound& sb1 = *sb1p;
    const uint32_t nbx = (uint32_t)(g.nx + kSvBX - 1) / kSvBX, nby = (uint32_t)(g.ny + kSvBY - 1) / kSvBY;
    const uint32_t bricks = nbx * nby * (uint32_t)((g.nz + kSvBZ - 1) / kSvBZ);
    const uint32_t lb = xcd_block();
    if (lb >= bricks) return;   // padding blocks of the XCD-rounded grid (whole block)
    const int x0 = (int)(lb % nbx) * kSvBX, y0 = (int)((lb / nbx) % nby) * kSvBY, z0 = (int)(lb / (nbx * nby)) * kSvBZ;
    for (int k = threadIdx.x; k < 5 * kSw; k += blockDim.x) win[k] = 0.0f;
    __syncthreads();
    const ScalarBrickSink sink{win, x0 - kSvHalo, y0 - kSvHalo, z0 - kSvHalo, va, src_adj0, NF == 2 ? src_adj1 : src_adj0};
    const int x = x0 + (int)(threadIdx.x % kSvBX), y = y0 + (int)(threadIdx.x / kSvBX);
    for (int f = 0; f < kSvBZ; ++f) {
        const int z = z0 + f;
        if (x >= g.nx || y >= g.ny || z >= g.nz) continue;
        const uint32_t i = cidx(g, x, y, z);
        if (HasSolids && mask[i] != 0) continue;
        const float a0 = oa0[i];
        const float a1 = NF == 2 ? oa1[i] : 0.0f;
        if (a0 == 0.0f && a1 == 0.0f) continue;
        const float sx = (x + 0.5f) * g.h, sy = (y + 0.5f) * g.h, sz = (z + 0.5f) * g.h;
        Sample mid[3];
        const TraceOut t = trace_rk2<HasSolids, true>(sx, sy, sz, u, mask, g, vb, mid);
        const Sample s0 = sample_scalar<true>(src0, t.px, t.py, t.pz, g, sb0);
        scatter_scalar_to<3>(sink, t.px, t.py, t.pz, a0, g, sb0);
        float gx = s0.gx * a0, gy = s0.gy * a0, gz = s0.gz * a0;
        if (NF == 2) {
            const Sample s1 = sample_scalar<true>(src1, t.px, t.py, t.pz, g, sb1);
            scatter_scalar_to<4>(sink, t.px, t.py, t.pz, a1, g, sb1);
            gx += s1.gx * a1;
            gy += s1.gy * a1;
            gz += s1.gz * a1;
        }
        const float v1x = -(t.dx * g.dt) * gx, v1y = -(t.dy * g.dt) * gy, v1z = -(t.dz * g.dt) * gz;
        trace_reverse(g, vb, sink, sx, sy, sz, t, mid, v1x, v1y, v1z);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 5 * kSw; k += blockDim.x) {
        const float v = win[k];
        if (v == 0.0f) continue;
        const int W = k / kSw, r = k - W * kSw;
        const int wx = x0 - kSvHalo + r % kSwX, wy = y0 - kSvHalo + (r / kSwX) % kSwY, wz = z0 - kSvHalo + r / (kSwX * kSwY);
        // only in-range targets receive window adds; the test keeps the flush in bounds regardless
        if (W < 3) {
            if (wx < 0 || wy < 0 || wz < 0 || wx >= ext(g, W, 0) || wy >= ext(g, W, 1) || wz >= ext(g, W, 2)) continue;
            float* dst = W == 0 ? va.p[0] : W == 1 ? va.p[1] : va.p[2];
            atomicAdd(dst + fidx(g, W, wx, wy, wz), v);
        } else {
            if (wx < 0 || wy < 0 || wz < 0 || wx >= g.nx || wy >= g.ny || wz >= g.nz) continue;
            atomicAdd((W == 3 ? src_adj0 : src_adj1) + cidx(g, wx, wy, wz), v);
        }
    }
}

// advect_velocity_vjp_kernel (operators.cu:415-439) for the faces of axis A, with its scatters
// privatised per block: a block owns a kVjpBX x kVjpBY x kVjpBZ brick of A-faces, whose traces
// move a few cells at most, so their 56 scatters per face land in a small box around the brick.
// Scatters inside the box (brick plus a halo of kVjpHalo faces, each component) are LDS atomics;
// those outside it (long traces, periodic wraps) stay global atomics. The box is flushed once.
constexpr int kVjpBX = 32, kVjpBY = 8, kVjpBZ = 4, kVjpHalo = 1;
constexpr int kWinX = kVjpBX + 2 * kVjpHalo + 1, kWinY = kVjpBY + 2 * kVjpHalo + 1, kWinZ = kVjpBZ + 2 * kVjpHalo + 1;
constexpr int kWin = kWinX * kWinY * kWinZ;   // per component

struct WindowSink {
    float* win;        // LDS, 3 * kWin
    int ox, oy, oz;    // face coordinates of window element (0, 0, 0)
    F3 va;
    template <int B>
    __device__ __forceinline__ void add(int x, int y, int z, uint32_t i, float value) const {
        const unsigned lx = (unsigned)(x - ox), ly = (unsigned)(y - oy), lz = (unsigned)(z - oz);
        // workgroup scope for the LDS add: it keeps the two atomics distinct (merged into one
        // atomic on a selected pointer they would both become flat atomics)
        if (lx < (unsigned)kWinX && ly < (unsigned)kWinY && lz < (unsigned)kWinZ)
            __hip_atomic_fetch_add(win + B * kWin + ((int)lz * kWinY + (int)ly) * kWinX + (int)lx, value, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            atomicAdd(va.p[B] + i, value);
    }
};

__device__ __forceinline__ uint32_t vjp_bricks(const Geo& g, int A, int& nbx, int& nby) {
    nbx = (ext(g, A, 0) + kVjpBX - 1) / kVjpBX;
    nby = (ext(g, A, 1) + kVjpBY - 1) / kVjpBY;
    return (uint32_t)nbx * nby * ((ext(g, A, 2) + kVjpBZ - 1) / kVjpBZ);
}

// The start point of a face's trace is the face itself, so the trace reverse's last three scatters
// (the start sample's adjoint, trace_reverse) have a fixed stencil: with a power-of-two cell size
// the component A corner is the face (weight 1) and a component B != A has the four B-faces
// {a - 1, a} x {b, b + 1} x {c} (along A, B and the third axis; weight 1/4 each). A face whose
// stencil needs no boundary value and no periodic mapping (start_gatherable) leaves those values in
// `start` instead of scattering them, and start_gather_kernel (after the three axes) adds them to
// each target face as a gather: 9 of the about 41 LDS atomics per face, which bound the kernel.
template <int A>
__device__ __forceinline__ bool start_gatherable(const Geo& g, uint32_t per, int x, int y, int z) {
    const int c[3]{x, y, z}, n[3]{g.nx, g.ny, g.nz};
    bool ok = c[A] >= 1 && c[A] <= n[A] - 1;
#pragma unroll
    for (int B = 0; B < 3; ++B)
        if (B != A && (per & (1u << B))) ok = ok && c[B] + 1 < n[B];
    return ok;
}

// 4 waves per SIMD (registers capped at 128, the rest spilled), the most its 32 KB window allows.
template <int A, bool HasSolids, bool GatherStart>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void advect_velocity_vjp_kernel(Geo g, const VBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                                  Vel3 u, const float* __restrict__ oa, F3 va, F3 start, uint32_t per) {
    __shared__ float win[3 * kWin];
    const VBound& b = *bp;
    int nbx, nby;
    const uint32_t bricks = vjp_bricks(g, A, nbx, nby);
    const uint32_t lb = xcd_block();
    if (lb >= bricks) return;   // padding blocks of the XCD-rounded grid (whole block)
    const int x0 = (int)(lb % (uint32_t)nbx) * kVjpBX;
    const int y0 = (int)((lb / (uint32_t)nbx) % (uint32_t)nby) * kVjpBY;
    const int z0 = (int)(lb / ((uint32_t)nbx * nby)) * kVjpBZ;
    for (int i = threadIdx.x; i < 3 * kWin; i += blockDim.x) win[i] = 0.0f;
    __syncthreads();
    const WindowSink sink{win, x0 - kVjpHalo, y0 - kVjpHalo, z0 - kVjpHalo, va};
    const int ex = ext(g, A, 0), ey = ext(g, A, 1), ez = ext(g, A, 2);
    const int x = x0 + (int)(threadIdx.x % kVjpBX), y = y0 + (int)(threadIdx.x / kVjpBX);
    for (int f = 0; f < kVjpBZ; ++f) {
        const int z = z0 + f;
        if (x >= ex || y >= ey || z >= ez) continue;
        const uint32_t fi = fidx(g, A, x, y, z);
        const float a = oa[fi];
        const bool gather = GatherStart && start_gatherable<A>(g, per, x, y, z);
        float st[3]{0.0f, 0.0f, 0.0f};   // the start point's adjoints, gathered later
        if (a != 0.0f) {
            const float sx = (x + (A == 0 ? 0.0f : 0.5f)) * g.h;
            const float sy = (y + (A == 1 ? 0.0f : 0.5f)) * g.h;
            const float sz = (z + (A == 2 ? 0.0f : 0.5f)) * g.h;
            Sample mid[3];
            const TraceOut t = trace_rk2<HasSolids, true>(sx, sy, sz, u, mask, g, b, mid);
            const Sample s = sample_face<A, true>(u.u[A], t.px, t.py, t.pz, g, b);
            scatter_face_to<A>(sink, t.px, t.py, t.pz, a, g, b);
            const float v1x = -(t.dx * g.dt) * (s.gx * a), v1y = -(t.dy * g.dt) * (s.gy * a), v1z = -(t.dz * g.dt) * (s.gz * a);
            // trace_reverse, with the start point's three adjoints either kept for the gather or
            // scattered here
            scatter_face_to<0>(sink, t.mx, t.my, t.mz, v1x, g, b);
            scatter_face_to<1>(sink, t.mx, t.my, t.mz, v1y, g, b);
            scatter_face_to<2>(sink, t.mx, t.my, t.mz, v1z, g, b);
            const float sc = -0.5f * g.dt;
            st[0] = sc * (mid[0].gx * v1x + mid[1].gx * v1y + mid[2].gx * v1z);
            st[1] = sc * (mid[0].gy * v1x + mid[1].gy * v1y + mid[2].gy * v1z);
            st[2] = sc * (mid[0].gz * v1x + mid[1].gz * v1y + mid[2].gz * v1z);
            if (!gather) {
                scatter_face_to<0>(sink, sx, sy, sz, st[0], g, b);
                scatter_face_to<1>(sink, sx, sy, sz, st[1], g, b);
                scatter_face_to<2>(sink, sx, sy, sz, st[2], g, b);
                st[0] = st[1] = st[2] = 0.0f;
            }
        }
        if (GatherStart) {
            start.p[0][fi] = st[0];
            start.p[1][fi] = st[1];
            start.p[2][fi] = st[2];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * kWin; i += blockDim.x) {
        const float v = win[i];
        if (v == 0.0f) continue;
        const int B = i / kWin, r = i - B * kWin;
        const int wx = x0 - kVjpHalo + r % kWinX, wy = y0 - kVjpHalo + (r / kWinX) % kWinY, wz = z0 - kVjpHalo + r / (kWinX * kWinY);
        // only in-range faces receive window adds; the test keeps the flush in bounds regardless
        if (wx < 0 || wy < 0 || wz < 0 || wx >= ext(g, B, 0) || wy >= ext(g, B, 1) || wz >= ext(g, B, 2)) continue;
        float* dst = B == 0 ? va.p[0] : B == 1 ? va.p[1] : va.p[2];
        atomicAdd(dst + fidx(g, B, wx, wy, wz), v);
    }
}

// The transpose of the start stencils of the three axes' faces, for every face of component B (one
// thread per face of the three components, concatenated), after the three axes' reverse kernels:
// adds the gathered start adjoints. start[A][B]: component B's start adjoints of axis A's faces.
struct StartAdj {
    const float* p[3][3];
};
__global__ __launch_bounds__(256) void start_gather_kernel(Geo g, uint32_t per, StartAdj start, F3 va) {
    const uint32_t f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= g.foff[3]) return;
    for_face(g, f, [&]<int B>(uint32_t local) {
        int t[3];
        decode_face_axis<B>(g, local, t[0], t[1], t[2]);
        const int n[3]{g.nx, g.ny, g.nz};
        float s = 0.0f;
        auto from_axis = [&]<int A>() {
            if (A == B) {
                if (start_gatherable<A>(g, per, t[0], t[1], t[2])) s += start.p[A][A][local];
            } else {
#pragma unroll
                for (int da = 0; da < 2; ++da)
#pragma unroll
                    for (int db = -1; db <= 0; ++db) {
                        int q[3]{t[0], t[1], t[2]};
                        q[A] += da;
                        q[B] += db;
                        // an axis-A face: its own axis has n + 1 faces, the others n
                        if (q[A] < 0 || q[A] > n[A] || q[B] < 0 || q[B] >= n[B]) continue;
                        if (!start_gatherable<A>(g, per, q[0], q[1], q[2])) continue;
                        s += 0.25f * start.p[A][B][fidx(g, A, q[0], q[1], q[2])];
                    }
            }
        };
        from_axis.template operator()<0>();
        from_axis.template operator()<1>();
        from_axis.template operator()<2>();
        if (s != 0.0f) va.p[B][local] += s;
    });
}

// constrained_boundary, operators.cu:452-458
template <int A>
__device__ __forceinline__ bool constrained_boundary(const Geo& g, const VBound& b, int coord) {
    const int n_a = dim_size(g, A);
    if (coord != 0 && coord != n_a) return false;
    const uint32_t mode = coord == n_a ? b.modes[2 * A + 1] : b.modes[2 * A];
    return mode == 0u || mode == 2u;
}

// Adjoint of the advected velocity at one face from the projected-velocity adjoint and lambda
// (projection reverse + pressure_rhs_reverse as a gather, operators.cu:922-933,957-969):
//   adj = u_bar(f) - (h/dt) * (lambda(c0) - lambda(c1)), lambda = 0 off the unknowns.
struct ProjAdjIn {
    const float* ua[3];
    const float* extra[3];
    const float* lh;
    const float* ll;
};

template <int A>
__device__ __forceinline__ float projected_face_adjoint(const Geo& g, const ProjAdjIn& in, int x, int y, int z, uint32_t local) {
    float a = in.ua[A][local];
    if (in.extra[A]) a += in.extra[A][local];
    const int coord = A == 0 ? x : A == 1 ? y : z;
    const int n_a = dim_size(g, A);
    float l0 = 0.0f, l1 = 0.0f;
    if (coord > 0) {
        const uint32_t c0 = cidx(g, x - (A == 0), y - (A == 1), z - (A == 2));
        l0 = in.lh[c0] + in.ll[c0];
    }
    if (coord < n_a) {
        const uint32_t c1 = cidx(g, x, y, z);
        l1 = in.lh[c1] + in.ll[c1];
    }
    return a - g.h * (l0 - l1) / g.dt;
}

// Projection reverse with the constraint transpose fused (constrain_velocity_vjp_kernel,
// operators.cu:483-497, as a gather): constrained, solid-adjacent and periodic-duplicate faces
// receive 0; face 0 of a periodic axis also receives its duplicate's adjoint.
// Constrain = false: out += projected adjoint (the per-operator projection VJP).
template <bool Constrain>
__global__ __launch_bounds__(256) void projection_vjp_face_kernel(Geo g, const VBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                                  uint32_t has_solids, ProjAdjIn in, F3 out) {
    const VBound& b = *bp;
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        if (!Constrain) {
            out.p[A][local] += projected_face_adjoint<A>(g, in, x, y, z, local);
            return;
        }
        const int coord = A == 0 ? x : A == 1 ? y : z;
        const int n_a = dim_size(g, A);
        const bool per = vperiodic(b, A);
        if (constrained_boundary<A>(g, b, coord) || (has_solids && adjacent_solid<A>(g, mask, x, y, z)) || (per && coord == n_a)) {
            out.p[A][local] = 0.0f;
            return;
        }
        float a = projected_face_adjoint<A>(g, in, x, y, z, local);
        if (per && coord == 0) {
            const int dx = A == 0 ? n_a : x, dy = A == 1 ? n_a : y, dz = A == 2 ? n_a : z;
            if (!(has_solids && adjacent_solid<A>(g, mask, dx, dy, dz)))
                a += projected_face_adjoint<A>(g, in, dx, dy, dz, fidx(g, A, dx, dy, dz));
        }
        out.p[A][local] = a;
    });
}

// constrain_velocity_vjp_kernel (operators.cu:483-497) as a gather: out += C^T in.
__global__ __launch_bounds__(256) void constrain_vjp_kernel(Geo g, const VBound* __restrict__ bp, const uint8_t* __restrict__ mask,
                                                            uint32_t has_solids, CF3 in, F3 out) {
    const VBound& b = *bp;
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) {
        int x, y, z;
        decode_face_axis<A>(g, local, x, y, z);
        const int coord = A == 0 ? x : A == 1 ? y : z;
        const int n_a = dim_size(g, A);
        const bool per = vperiodic(b, A);
        if (constrained_boundary<A>(g, b, coord) || (has_solids && adjacent_solid<A>(g, mask, x, y, z)) || (per && coord == n_a))
            return;
        float a = in.p[A][local];
        if (per && coord == 0) {
            const int dx = A == 0 ? n_a : x, dy = A == 1 ? n_a : y, dz = A == 2 ? n_a : z;
            if (!(has_solids && adjacent_solid<A>(g, mask, dx, dy, dz))) a += in.p[A][fidx(g, A, dx, dy, dz)];
        }
        out.p[A][local] += a;
    });
}

// out += in over the concatenated faces (the velocity pass-through of integrate_vjp_kernel,
// operators.cu:370, and accumulate_kernel, operators.cu:292-295).
__global__ __launch_bounds__(256) void face_accumulate_kernel(Geo g, CF3 in, F3 out) {
    const uint32_t fi = xcd_block() * blockDim.x + threadIdx.x;
    if (fi >= g.foff[3]) return;
    for_face(g, fi, [&]<int A>(uint32_t local) { out.p[A][local] += in.p[A][local]; });
}

__global__ __launch_bounds__(256) void cell_accumulate_kernel(uint32_t n, const float* __restrict__ in, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += in[i];
}

// Integration reverse (force part), buoyancy reverse (operators.cu:326-337) and source reverse
// (operators.cu:297-302) fused per fluid cell. Parameter adjoints: FP64 block partials.
// Vort: the confinement force reverse runs on the same force adjoint (vorticity_force_reverse +
// the per-cell part of normal_reverse, operators.cu:781-812), writing vort.ga / vort.t (zeros in
// solids) and the confinement adjoint as partial set 3; the vorticity fields in vort.in come
// from the curl pass over the step's input velocity.
struct VortReverse {
    VortFields in;
    F3 ga, t;
};

template <bool Vort>
__global__ __launch_bounds__(256) void force_source_vjp_kernel(Geo g, const uint8_t* __restrict__ mask, CF3 forced_adj,
                                                               const float* __restrict__ rho_s, const float* __restrict__ t_s,
                                                               const float* __restrict__ rho_s_adj, const float* __restrict__ t_s_adj,
                                                               Params prm, F3 e_adj, float* __restrict__ prev_rho,
                                                               float* __restrict__ prev_t, float* __restrict__ s_rho_adj,
                                                               float* __restrict__ s_t_adj, VortReverse vort, double* __restrict__ partials) {
    constexpr int kParts = Vort ? 4 : 3;
    __shared__ double lds[kParts * kMaxWaves];
    const uint32_t lb = xcd_block();
    uint32_t chunk = (g.n + gridDim.x - 1) / gridDim.x;
    chunk = (chunk + 255u) & ~255u;
    const uint32_t start = min(g.n, lb * chunk), end = min(g.n, start + chunk);
    const float amb = *prm.ambient, ar = *prm.alpha_rho, at = *prm.alpha_t;
    const float scale = Vort ? *prm.confinement * g.h : 0.0f;
    double p_amb = 0.0, p_ar = 0.0, p_at = 0.0, p_conf = 0.0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        if (mask[i] != 0) {
            if constexpr (Vort) {
#pragma unroll
                for (int a = 0; a < 3; ++a) vort.ga.p[a][i] = vort.t.p[a][i] = 0.0f;
            }
            continue;
        }
        int x, y, z;
        decode_cell(g, i, x, y, z);
        float f[3];
        force_gather(g, mask, forced_adj, i, x, y, z, f);
        if constexpr (Vort) {
            float ga[3], t[3];
            p_conf += confinement_reverse_at(g, vort.in, i, x, y, z, scale, f, ga, t);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                vort.ga.p[a][i] = ga[a];
                vort.t.p[a][i] = t[a];
            }
        }
        if (e_adj.p[0]) {
            e_adj.p[0][i] += f[0];
            e_adj.p[1][i] += f[1];
            e_adj.p[2][i] += f[2];
        }
        const float rsa = (rho_s_adj ? rho_s_adj[i] : 0.0f) + ar * f[1];
        const float tsa = (t_s_adj ? t_s_adj[i] : 0.0f) + at * f[1];
        p_amb += -(double)at * f[1];
        p_ar += (double)rho_s[i] * f[1];
        p_at += (double)(t_s[i] - amb) * f[1];
        prev_rho[i] += rsa;
        prev_t[i] += tsa;
        if (s_rho_adj) s_rho_adj[i] += g.dt * rsa;
        if (s_t_adj) s_t_adj[i] += g.dt * tsa;
    }
    double v[kParts];
    v[0] = p_amb;
    v[1] = p_ar;
    v[2] = p_at;
    if constexpr (Vort) v[3] = p_conf;
    block_sum<kParts>(v, lds);
    if (threadIdx.x == 0 && partials) {
#pragma unroll
        for (int k = 0; k < kParts; ++k) partials[k * gridDim.x + lb] = v[k];
    }
}

// buoyancy_vjp_kernel (operators.cu:326-337) with FP64 block partials for the parameters.
__global__ __launch_bounds__(256) void buoyancy_vjp_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ rho,
                                                           const float* __restrict__ T, Params prm, CF3 fa, float* __restrict__ rho_adj,
                                                           float* __restrict__ t_adj, F3 e_adj, double* __restrict__ partials) {
    __shared__ double lds[3 * kMaxWaves];
    const uint32_t lb = xcd_block();
    uint32_t chunk = (g.n + gridDim.x - 1) / gridDim.x;
    chunk = (chunk + 255u) & ~255u;
    const uint32_t start = min(g.n, lb * chunk), end = min(g.n, start + chunk);
    const float amb = *prm.ambient, ar = *prm.alpha_rho, at = *prm.alpha_t;
    double p_amb = 0.0, p_ar = 0.0, p_at = 0.0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        if (mask[i] != 0) continue;
        const float fy = fa.p[1][i];
        if (e_adj.p[0]) {
            e_adj.p[0][i] += fa.p[0][i];
            e_adj.p[1][i] += fy;
            e_adj.p[2][i] += fa.p[2][i];
        }
        rho_adj[i] += ar * fy;
        t_adj[i] += at * fy;
        p_amb += -(double)at * fy;
        p_ar += (double)rho[i] * fy;
        p_at += (double)(T[i] - amb) * fy;
    }
    double v[3]{p_amb, p_ar, p_at};
    block_sum<3>(v, lds);
    if (threadIdx.x == 0) {
        partials[0 * gridDim.x + lb] = v[0];
        partials[1 * gridDim.x + lb] = v[1];
        partials[2 * gridDim.x + lb] = v[2];
    }
}

// One block: param_adj[k] += sum of the partials of parameter k (fixed order), for the K first
// parameters (ambient, density buoyancy, temperature buoyancy, confinement).
struct ParamOut {
    double* p[4];
};
template <int K>
__global__ __launch_bounds__(256) void param_finalize_kernel(const double* __restrict__ partials, uint32_t G, ParamOut out) {
    __shared__ double lds[K * kMaxWaves];
    const double* parts[K];
#pragma unroll
    for (int k = 0; k < K; ++k) parts[k] = partials + (size_t)k * G;
    double red[K];
    reduce_partials<K>(parts, G, red, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (out.p[k]) *out.p[k] += red[k];
    }
}

ParamOut param_out(const ds_parameter_adjoint* a) {
    return ParamOut{{a->ambient_temperature, a->density_buoyancy, a->temperature_buoyancy, a->vorticity_confinement}};
}

// source_vjp_kernel, operators.cu:297-302
__global__ __launch_bounds__(256) void source_vjp_kernel(Geo g, const uint8_t* __restrict__ mask, const float* __restrict__ oa,
                                                         float* __restrict__ sa, float* __restrict__ srca) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n || mask[i] != 0) return;
    sa[i] += oa[i];
    if (srca) srca[i] += g.dt * oa[i];
}

// Force part of integrate_vjp_kernel as a gather (per-operator entry point).
__global__ __launch_bounds__(256) void integrate_force_vjp_kernel(Geo g, const uint8_t* __restrict__ mask, CF3 fa, F3 force_adj) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= g.n || mask[i] != 0) return;
    int x, y, z;
    decode_cell(g, i, x, y, z);
    float f[3];
    force_gather(g, mask, fa, i, x, y, z, f);
    force_adj.p[0][i] += f[0];
    force_adj.p[1][i] += f[1];
    force_adj.p[2][i] += f[2];
}

}  // namespace

void launch_advect_scalar_vjp(ds_context* c, int nf, int first_field, const float* src0, const float* src1, CF3 u, const float* oa0,
                              const float* oa1, float* sa0, float* sa1, F3 va) {
    const Vel3 v{{u.p[0], u.p[1], u.p[2]}};
    const Geo& g = c->geo;
    const uint64_t bricks = (uint64_t)((g.nx + kSvBX - 1) / kSvBX) * ((g.ny + kSvBY - 1) / kSvBY) * ((g.nz + kSvBZ - 1) / kSvBZ);
    const uint32_t G = (uint32_t)((bricks + 7) & ~7ull);
    if (nf == 2) {
        if (c->has_solids)
            advect_scalar_vjp_kernel<2, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_db, c->d_tb, c->mask, src0, src1, v, oa0, oa1, sa0,
                                                                          sa1, va);
        else
            advect_scalar_vjp_kernel<2, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->d_db, c->d_tb, c->mask, src0, src1, v, oa0, oa1,
                                                                           sa0, sa1, va);
    } else {
        const SBound* b0 = first_field == 0 ? c->d_db : c->d_tb;
        if (c->has_solids)
            advect_scalar_vjp_kernel<1, true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, b0, b0, c->mask, src0, src0, v, oa0, oa0, sa0, sa0,
                                                                          va);
        else
            advect_scalar_vjp_kernel<1, false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, b0, b0, c->mask, src0, src0, v, oa0, oa0, sa0,
                                                                           sa0, va);
    }
    check_launch("advect_scalar_vjp_kernel");
}

void launch_advect_velocity_vjp(ds_context* c, CF3 u, CF3 oa, F3 va) {
    const Vel3 v{{u.p[0], u.p[1], u.p[2]}};
    const Geo& g = c->geo;
    uint32_t per = 0;   // bit d: axis d of the velocity boundary is periodic (vperiodic)
    for (int d = 0; d < 3; ++d)
        if (c->vb.modes[2 * d] == 3u && c->vb.modes[2 * d + 1] == 3u) per |= 1u << d;
    // the start stencils are gathered when the context holds the buffers and the fractions are exact
    const bool gather = c->vjp_start[0][0] != nullptr && g.h_pow2;
    // profiling: every sample_every-th call has each axis launch bracketed by HIP events
    Profiler& prof = c->profiler;
    const bool sample = prof.every && (prof.launches[6] / 3) % prof.every == 0 && prof.used + 6 <= 6000;
    auto launch = [&]<int A>() {
        const int e[3]{g.nx + (A == 0), g.ny + (A == 1), g.nz + (A == 2)};   // face extents of axis A
        const uint64_t bricks = (uint64_t)((e[0] + kVjpBX - 1) / kVjpBX) * ((e[1] + kVjpBY - 1) / kVjpBY) * ((e[2] + kVjpBZ - 1) / kVjpBZ);
        const uint32_t G = (uint32_t)((bricks + 7) & ~7ull);
        hipEvent_t e0 = sample ? prof.take() : nullptr, e1 = sample ? prof.take() : nullptr;
        if (e0) DS_HIP(hipEventRecord(e0, c->stream));
        const F3 st{{c->vjp_start[A][0], c->vjp_start[A][1], c->vjp_start[A][2]}};
        if (gather) {
            if (c->has_solids)
                advect_velocity_vjp_kernel<A, true, true><<<G, kBlock, 0, c->stream>>>(g, c->d_vb, c->mask, v, oa.p[A], va, st, per);
            else
                advect_velocity_vjp_kernel<A, false, true><<<G, kBlock, 0, c->stream>>>(g, c->d_vb, c->mask, v, oa.p[A], va, st, per);
        } else {
            if (c->has_solids)
                advect_velocity_vjp_kernel<A, true, false><<<G, kBlock, 0, c->stream>>>(g, c->d_vb, c->mask, v, oa.p[A], va, st, per);
            else
                advect_velocity_vjp_kernel<A, false, false><<<G, kBlock, 0, c->stream>>>(g, c->d_vb, c->mask, v, oa.p[A], va, st, per);
        }
        if (e0 && e1) {
            DS_HIP(hipEventRecord(e1, c->stream));
            prof.recs.push_back({6, e0, e1});
        }
        if (prof.every) prof.launches[6]++;
    };
    launch.template operator()<0>();
    launch.template operator()<1>();
    launch.template operator()<2>();
    if (gather) {
        StartAdj sa;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) sa.p[a][b] = c->vjp_start[a][b];
        start_gather_kernel<<<xcd_blocks_for(g.foff[3]), kBlock, 0, c->stream>>>(g, per, sa, va);
    }
    check_launch("advect_velocity_vjp_kernel");
}

void launch_projection_vjp_faces(ds_context* c, const float* const ua[3], const float* const extra[3], const float* lh, const float* ll,
                                 F3 out, bool constrain) {
    ProjAdjIn in{};
    for (int a = 0; a < 3; ++a) {
        in.ua[a] = ua[a];
        in.extra[a] = extra ? extra[a] : nullptr;
    }
    in.lh = lh;
    in.ll = ll;
    const uint32_t G = xcd_blocks_for(c->geo.foff[3]);
    if (constrain)
        projection_vjp_face_kernel<true><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->mask, c->has_solids ? 1u : 0u, in, out);
    else
        projection_vjp_face_kernel<false><<<G, kBlock, 0, c->stream>>>(c->geo, c->d_vb, c->mask, c->has_solids ? 1u : 0u, in, out);
    check_launch("projection_vjp_face_kernel");
}

void launch_face_accumulate(ds_context* c, CF3 in, F3 out) {
    face_accumulate_kernel<<<xcd_blocks_for(c->geo.foff[3]), kBlock, 0, c->stream>>>(c->geo, in, out);
    check_launch("face_accumulate_kernel");
}

void launch_cell_accumulate(ds_context* c, const float* in, float* out) {
    cell_accumulate_kernel<<<blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo.n, in, out);
    check_launch("cell_accumulate_kernel");
}

void launch_force_source_vjp(ds_context* c, CF3 forced_adj, const float* rho_s, const float* t_s, const float* rho_s_adj,
                             const float* t_s_adj, const Params& prm, F3 e_adj, float* prev_rho, float* prev_t, float* s_rho_adj,
                             float* s_t_adj, const ds_parameter_adjoint* param_adj) {
    if (c->vorticity) {
        require(prm.confinement != nullptr, "parameters: vorticity_confinement must be a device pointer with confinement on");
        const VortReverse vr{VortFields{{c->vort_w[0], c->vort_w[1], c->vort_w[2]}, c->vort_mag},
                             F3{{c->vort_ga[0], c->vort_ga[1], c->vort_ga[2]}}, F3{{c->vort_t[0], c->vort_t[1], c->vort_t[2]}}};
        force_source_vjp_kernel<true><<<kForceBlocks, kBlock, 0, c->stream>>>(c->geo, c->mask, forced_adj, rho_s, t_s, rho_s_adj, t_s_adj,
                                                                            prm, e_adj, prev_rho, prev_t, s_rho_adj, s_t_adj, vr,
                                                                            c->param_partials);
    } else {
        force_source_vjp_kernel<false><<<kForceBlocks, kBlock, 0, c->stream>>>(c->geo, c->mask, forced_adj, rho_s, t_s, rho_s_adj, t_s_adj,
                                                                             prm, e_adj, prev_rho, prev_t, s_rho_adj, s_t_adj, VortReverse{},
                                                                             c->param_partials);
    }
    check_launch("force_source_vjp_kernel");
    if (param_adj) {
        if (c->vorticity) param_finalize_kernel<4><<<1, kBlock, 0, c->stream>>>(c->param_partials, kForceBlocks, param_out(param_adj));
        else param_finalize_kernel<3><<<1, kBlock, 0, c->stream>>>(c->param_partials, kForceBlocks, param_out(param_adj));
        check_launch("param_finalize_kernel");
    }
}

}  // namespace ds

using namespace ds;

namespace {
bool need_differentiable(ds_context* c, const char* name) {
    if (c->mode != DS_MODE_DIFFERENTIABLE) {
        set_error(std::string(name) + ": the context was not created in differentiable mode");
        return false;
    }
    return true;
}
}  // namespace

extern "C" {

ds_status ds_source_vjp(ds_context* c, const float* out_adj, float* state_adj, float* source_adj) {
    if (!c || !out_adj || !state_adj) {
        set_error("ds_source_vjp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        source_vjp_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, out_adj, state_adj, source_adj);
        check_launch("source_vjp_kernel");
    });
}

ds_status ds_buoyancy_vjp(ds_context* c, const float* density, const float* temperature, const ds_parameters* params,
                          float* const force_adj[3], float* density_adj, float* temperature_adj, float* e_adj[3],
                          const ds_parameter_adjoint* param_adj) {
    if (!c || !density || !temperature || !force_adj || !density_adj || !temperature_adj) {
        set_error("ds_buoyancy_vjp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!need_differentiable(c, "ds_buoyancy_vjp")) return DS_ERROR_INVALID_ARGUMENT;
    return guard([&] {
        const Params prm = params_of(params);
        F3 ea{{e_adj ? e_adj[0] : nullptr, e_adj ? e_adj[1] : nullptr, e_adj ? e_adj[2] : nullptr}};
        buoyancy_vjp_kernel<<<kPcgBlocks, kBlock, 0, c->stream>>>(c->geo, c->mask, density, temperature, prm,
                                                                  CF3{{force_adj[0], force_adj[1], force_adj[2]}}, density_adj,
                                                                  temperature_adj, ea, c->param_partials);
        check_launch("buoyancy_vjp_kernel");
        if (param_adj) {
            ParamOut out = param_out(param_adj);
            out.p[3] = nullptr;
            param_finalize_kernel<3><<<1, kBlock, 0, c->stream>>>(c->param_partials, kPcgBlocks, out);
            check_launch("param_finalize_kernel");
        }
    });
}

ds_status ds_integrate_velocity_vjp(ds_context* c, float* const out_adj[3], float* velocity_adj[3], float* force_adj[3]) {
    if (!c || !out_adj || !velocity_adj) {
        set_error("ds_integrate_velocity_vjp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        const CF3 oa{{out_adj[0], out_adj[1], out_adj[2]}};
        launch_face_accumulate(c, oa, F3{{velocity_adj[0], velocity_adj[1], velocity_adj[2]}});
        if (force_adj) {
            integrate_force_vjp_kernel<<<xcd_blocks_for(c->geo.n), kBlock, 0, c->stream>>>(c->geo, c->mask, oa,
                                                                                           F3{{force_adj[0], force_adj[1], force_adj[2]}});
            check_launch("integrate_force_vjp_kernel");
        }
    });
}

ds_status ds_advect_velocity_vjp(ds_context* c, float* const velocity[3], float* const out_adj[3], float* velocity_adj[3]) {
    if (!c || !velocity || !out_adj || !velocity_adj) {
        set_error("ds_advect_velocity_vjp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_advect_velocity_vjp(c, CF3{{velocity[0], velocity[1], velocity[2]}}, CF3{{out_adj[0], out_adj[1], out_adj[2]}},
                                   F3{{velocity_adj[0], velocity_adj[1], velocity_adj[2]}});
    });
}

ds_status ds_constrain_velocity_vjp(ds_context* c, float* const out_adj[3], float* velocity_adj[3]) {
    if (!c || !out_adj || !velocity_adj) {
        set_error("ds_constrain_velocity_vjp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        constrain_vjp_kernel<<<xcd_blocks_for(c->geo.foff[3]), kBlock, 0, c->stream>>>(
            c->geo, c->d_vb, c->mask, c->has_solids ? 1u : 0u, CF3{{out_adj[0], out_adj[1], out_adj[2]}},
            F3{{velocity_adj[0], velocity_adj[1], velocity_adj[2]}});
        check_launch("constrain_vjp_kernel");
    });
}

ds_status ds_advect_scalar_vjp(ds_context* c, uint32_t field, const float* source, float* const velocity[3], const float* out_adj,
                               float* source_adj, float* velocity_adj[3]) {
    if (!c || !source || !velocity || !out_adj || !source_adj || !velocity_adj || field > 1u) {
        set_error("ds_advect_scalar_vjp: invalid argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    return guard([&] {
        launch_advect_scalar_vjp(c, 1, (int)field, source, source, CF3{{velocity[0], velocity[1], velocity[2]}}, out_adj, out_adj,
                                 source_adj, source_adj, F3{{velocity_adj[0], velocity_adj[1], velocity_adj[2]}});
    });
}

ds_status ds_projection_vjp(ds_context* c, float* const out_adj[3], float* velocity_adj[3], ds_solve_info* info) {
    if (!c || !out_adj || !velocity_adj) {
        set_error("ds_projection_vjp: null argument");
        return DS_ERROR_INVALID_ARGUMENT;
    }
    if (!need_differentiable(c, "ds_projection_vjp")) return DS_ERROR_INVALID_ARGUMENT;
    ds_solve_info si{};
    const ds_status st = guard([&] {
        const float* ua[3]{out_adj[0], out_adj[1], out_adj[2]};
        si = pcg_solve_kind(c, 2, ua, nullptr, c->lambda_hi, c->lambda_lo, nullptr, nullptr, c->tolerance, c->max_iterations);
        c->lambda_valid = true;
        launch_projection_vjp_faces(c, ua, nullptr, c->lambda_hi, c->lambda_lo, F3{{velocity_adj[0], velocity_adj[1], velocity_adj[2]}},
                                    false);
    });
    if (info) *info = si;
    if (st != DS_OK) return st;
    if (!si.converged) {
        set_error("ds_projection_vjp: adjoint pressure solve did not converge");
        return std::isfinite(si.final_residual) ? DS_ERROR_NOT_CONVERGED : DS_ERROR_NON_FINITE;
    }
    return DS_OK;
}

}  // extern "C"
