// pcg_plane.hpp — the two PCG iteration kernels as software-pipelined LDS plane marches.
//
// A 1024-thread block owns a 64 x 16 (x, y) tile and walks a chunk of z planes. The values of
// plane z+1 (own cell plus the tile rim: one x-rim and one y-rim cell for edge threads) are
// staged in LDS while plane z is computed from LDS; the z neighbours stay in registers
// (below / centre / next); the global loads of plane z+2 are issued one step ahead into
// registers, so memory latency overlaps a whole plane step. Every vector element is loaded
// once per block; the overhead is 2 halo planes per chunk plus the tile rim. Loads are
// 64-wide coalesced x rows. Blocks are ordered chunk-major, so after the XCD remap each XCD
// walks a contiguous z slab and tile-rim reuse stays in its L2.
#pragma once

#include "context.hpp"

namespace ds {

constexpr int kTX = kPlaneTileX;
constexpr int kTY = kPlaneTileY;

// 1/d for the diagonal counts 1..6 of the 7-point matrix, 0 off the unknowns: a chain of selects
// (a nested conditional expression compiles to divergent branches).
__device__ __forceinline__ float inv_diag(uint32_t d) {
    float v = 0.0f;
    v = d == 1u ? 1.0f : v;
    v = d == 2u ? 0.5f : v;
    v = d == 3u ? (1.0f / 3.0f) : v;
    v = d == 4u ? 0.25f : v;
    v = d == 5u ? 0.2f : v;
    v = d == 6u ? (1.0f / 6.0f) : v;
    return v;
}

// x_hi + x_lo += y with FP32 error-free transformations (Knuth two-sum, then renormalise).
__device__ __forceinline__ void compensated_add(float& hi, float& lo, float y) {
#pragma clang fp contract(off)
    const float s = hi + y;
    const float bp = s - hi;
    const float e = (hi - (s - bp)) + (y - bp);
    const float t = lo + e;
    const float h2 = s + t;
    lo = t - (h2 - s);
    hi = h2;
}

// Index of (x, y, z), or -1 when the cell is outside the domain (the stencil weight there is 0).
// A periodic axis maps -1 and n to the opposite side; coordinates further out (the unused lanes
// of a partial tile) also give -1, so no load leaves the arrays.
__device__ __forceinline__ int64_t wrapped_index(const Geo& g, uint32_t per, int x, int y, int z) {
    if (x < 0 || x >= g.nx) {
        if (!(per & 1u) || x < -1 || x > g.nx) return -1;
        x = x < 0 ? g.nx - 1 : 0;
    }
    if (y < 0 || y >= g.ny) {
        if (!(per & 2u) || y < -1 || y > g.ny) return -1;
        y = y < 0 ? g.ny - 1 : 0;
    }
    if (z < 0 || z >= g.nz) {
        if (!(per & 4u) || z < -1 || z > g.nz) return -1;
        z = z < 0 ? g.nz - 1 : 0;
    }
    return (int64_t)cidx(g, x, y, z);
}

// Block -> tile origin and z range.
__device__ __forceinline__ void plane_block(const Geo& g, const PlaneTiling& pt, uint32_t lb, int& x0, int& y0, int& z0, int& z1) {
    const uint32_t ntile = pt.tiles_x * pt.tiles_y;
    const uint32_t chunk = lb / ntile, t = lb - chunk * ntile;
    const uint32_t ty = t / pt.tiles_x;
    x0 = (int)((t - ty * pt.tiles_x) * kTX);
    y0 = (int)(ty * kTY);
    z0 = (int)(chunk * pt.zc);
    z1 = min(g.nz, z0 + (int)pt.zc);
}

// Per-thread slots of a plane: 0 = own cell, 1 = x rim (lane 0: x0-1, lane 63: x0+64),
// 2 = y rim (row 0: y0-1, row 15: y0+16). Slot k of plane z is the cell index or -1.
struct Slots {
    int sx1, sy2;   // rim coordinates (x of slot 1, y of slot 2); slot present when the thread is on that edge
    bool has1, has2;
    int t1x, t2y;   // LDS column of slot 1, LDS row of slot 2
};

__device__ __forceinline__ Slots make_slots(int x0, int y0, int lx, int ly) {
    Slots s;
    s.has1 = lx == 0 || lx == kTX - 1;
    s.sx1 = lx == 0 ? x0 - 1 : x0 + kTX;
    s.t1x = lx == 0 ? 0 : kTX + 1;
    s.has2 = ly == 0 || ly == kTY - 1;
    s.sy2 = ly == 0 ? y0 - 1 : y0 + kTY;
    s.t2y = ly == 0 ? 0 : kTY + 1;
    return s;
}

// Raw loads of the direction kernel for one cell: r, p_old and the code.
struct RawDir {
    float r, p;
    uint32_t d;
};

template <bool HasOld>
__device__ __forceinline__ RawDir load_dir(int64_t j, const float* __restrict__ r, const float* __restrict__ p_old,
                                           const uint8_t* __restrict__ code) {
    RawDir v{0.0f, 0.0f, 0u};
    if (j >= 0) {
        v.r = r[j];
        v.d = code[j];
        if (HasOld) v.p = p_old[j];
    }
    return v;
}

template <bool HasOld>
__device__ __forceinline__ float pv_of(const RawDir& v, float beta) {
    const float z = v.r * inv_diag(v.d);
    return HasOld ? z + beta * v.p : z;
}

// Planes whose loads are in flight ahead of the plane being computed.
constexpr int kPrefetch = 4;

// Iteration k, first kernel: convergence test on the reduced ||r||^2, then the new direction
// p_k = z_k + beta p_{k-1} and the partial <p_k, A p_k>.
template <bool HasOld>
__global__ __launch_bounds__(kPcgThreads) void pcg_direction_kernel(Geo g, uint32_t per, PlaneTiling pt, const uint8_t* __restrict__ code,
                                                             const float* __restrict__ r, const float* __restrict__ p_old,
                                                             float* __restrict__ p_new, double* __restrict__ partials,
                                                             double* __restrict__ hist, PcgStatus* status, double tol2, uint32_t k,
                                                             uint32_t first) {
    __shared__ float tile[2][kTY + 2][kTX + 2];
    __shared__ double lds[3 * kMaxWaves];
    if (status->stop) return;
    const uint32_t lb = xcd_block();
    int x0, y0, z0, z1;
    plane_block(g, pt, lb, x0, y0, z0, z1);
    const int lx = (int)(threadIdx.x & 63u), ly = (int)(threadIdx.x >> 6);
    const int x = x0 + lx, y = y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const Slots sl = make_slots(x0, y0, lx, ly);
    // loads of plane zz: own cell for every plane up to the top halo z1, rims for chunk planes
    auto load_plane = [&](int zz, RawDir (&v)[3]) {
        v[0] = v[1] = v[2] = RawDir{0.0f, 0.0f, 0u};
        if (zz > z1) return;
        v[0] = load_dir<HasOld>(wrapped_index(g, per, x, y, zz), r, p_old, code);
        if (zz < z1) {
            if (sl.has1) v[1] = load_dir<HasOld>(wrapped_index(g, per, sl.sx1, y, zz), r, p_old, code);
            if (sl.has2) v[2] = load_dir<HasOld>(wrapped_index(g, per, x, sl.sy2, zz), r, p_old, code);
        }
    };
    // the first planes' loads do not depend on the reduction: issue them before it
    RawDir below_raw{}, cur[3], pre[kPrefetch][3];
    if (z0 < z1) {
        below_raw = load_dir<HasOld>(wrapped_index(g, per, x, y, z0 - 1), r, p_old, code);
        load_plane(z0, cur);
#pragma unroll
        for (int m = 0; m < kPrefetch; ++m) load_plane(z0 + 1 + m, pre[m]);
    }
    const double* parts[3]{partials + kPartBB * kPcgMaxBlocks, partials + kPartRR * kPcgMaxBlocks, partials + kPartRZ * kPcgMaxBlocks};
    double red[3];
    reduce_partials<3>(parts, gridDim.x, red, lds);
    const double bb = first ? red[0] : status->bb;
    const double rr = red[1], rz = red[2];
    if (!(rr > tol2 * bb)) {   // converged (NaN also stops; the true-residual check reports it)
        if (lb == 0 && threadIdx.x == 0) {
            if (first) status->bb = bb;
            status->rr = rr;
            status->iterations = k;
            status->stop = 1u;
        }
        return;
    }
    if (lb == 0 && threadIdx.x == 0) {
        if (first) status->bb = bb;
        hist[2 * k] = rz;
    }
    const float beta = HasOld ? (float)(rz / hist[2 * (k - 1)]) : 0.0f;
    double pq = 0.0;
    if (z0 < z1) {
        float below = pv_of<HasOld>(below_raw, beta);
        float centre = pv_of<HasOld>(cur[0], beta);
        uint32_t d = cur[0].d;
        tile[0][ly + 1][lx + 1] = centre;
        if (sl.has1) tile[0][ly + 1][sl.t1x] = pv_of<HasOld>(cur[1], beta);
        if (sl.has2) tile[0][sl.t2y][lx + 1] = pv_of<HasOld>(cur[2], beta);
        int buf = 0;
        __syncthreads();
        for (int z = z0; z < z1; ++z) {
            const float next = pv_of<HasOld>(pre[0][0], beta);
            if (own) {
                const uint32_t i = cidx(g, x, y, z);
                const float q = (float)d * centre -
                                (((tile[buf][ly + 1][lx] + tile[buf][ly + 1][lx + 2]) + (tile[buf][ly][lx + 1] + tile[buf][ly + 2][lx + 1])) +
                                 below + next);
                p_new[i] = centre;
                pq += (double)centre * q;
            }
            if (z + 1 < z1) {
                tile[buf ^ 1][ly + 1][lx + 1] = next;
                if (sl.has1) tile[buf ^ 1][ly + 1][sl.t1x] = pv_of<HasOld>(pre[0][1], beta);
                if (sl.has2) tile[buf ^ 1][sl.t2y][lx + 1] = pv_of<HasOld>(pre[0][2], beta);
            }
            d = pre[0][0].d;
#pragma unroll
            for (int m = 0; m + 1 < kPrefetch; ++m) {
                pre[m][0] = pre[m + 1][0];
                pre[m][1] = pre[m + 1][1];
                pre[m][2] = pre[m + 1][2];
            }
            load_plane(z + 1 + kPrefetch, pre[kPrefetch - 1]);
            __syncthreads();
            below = centre;
            centre = next;
            buf ^= 1;
        }
    }
    double v[1]{pq};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[kPartPQ * kPcgMaxBlocks + lb] = v[0];
}

// Raw loads of the update kernel for the own cell (p is loaded for every slot).
struct RawUpd {
    float p, r, h, l;
    uint32_t d;
};

__device__ __forceinline__ RawUpd load_upd(int64_t j, bool full, const float* __restrict__ p, const float* __restrict__ r,
                                           const float* __restrict__ xh, const float* __restrict__ xl, const uint8_t* __restrict__ code) {
    RawUpd v{0.0f, 0.0f, 0.0f, 0.0f, 0u};
    if (j >= 0) {
        v.p = p[j];
        if (full) {
            v.d = code[j];
            v.r = r[j];
            v.h = xh[j];
            v.l = xl[j];
        }
    }
    return v;
}

__device__ __forceinline__ float load_p(int64_t j, const float* __restrict__ p) { return j >= 0 ? p[j] : 0.0f; }

// Iteration k, second kernel: alpha = <r,z>/<p,Ap>, x += alpha p, r -= alpha A p, partials.
__global__ __launch_bounds__(kPcgThreads) void pcg_update_kernel(Geo g, uint32_t per, PlaneTiling pt, const uint8_t* __restrict__ code,
                                                          const float* __restrict__ p, float* __restrict__ r, float* __restrict__ xh,
                                                          float* __restrict__ xl, double* __restrict__ partials,
                                                          double* __restrict__ hist, PcgStatus* status, uint32_t k) {
    __shared__ float tile[2][kTY + 2][kTX + 2];
    __shared__ double lds[3 * kMaxWaves];
    if (status->stop) return;
    const uint32_t lb = xcd_block();
    int x0, y0, z0, z1;
    plane_block(g, pt, lb, x0, y0, z0, z1);
    const int lx = (int)(threadIdx.x & 63u), ly = (int)(threadIdx.x >> 6);
    const int x = x0 + lx, y = y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const Slots sl = make_slots(x0, y0, lx, ly);
    struct Plane {
        RawUpd c;
        float p1, p2;
    };
    auto load_plane = [&](int zz, Plane& v) {
        v.c = RawUpd{0.0f, 0.0f, 0.0f, 0.0f, 0u};
        v.p1 = v.p2 = 0.0f;
        if (zz > z1) return;
        v.c = load_upd(wrapped_index(g, per, x, y, zz), own && zz < z1, p, r, xh, xl, code);
        if (zz < z1) {
            if (sl.has1) v.p1 = load_p(wrapped_index(g, per, sl.sx1, y, zz), p);
            if (sl.has2) v.p2 = load_p(wrapped_index(g, per, x, sl.sy2, zz), p);
        }
    };
    float below = 0.0f;
    Plane cur{}, pre[kPrefetch]{};
    if (z0 < z1) {
        below = load_p(wrapped_index(g, per, x, y, z0 - 1), p);
        load_plane(z0, cur);
#pragma unroll
        for (int m = 0; m < kPrefetch; ++m) load_plane(z0 + 1 + m, pre[m]);
    }
    const double* parts[1]{partials + kPartPQ * kPcgMaxBlocks};
    double red[1];
    reduce_partials<1>(parts, gridDim.x, red, lds);
    const double pq = red[0];
    const double rz = hist[2 * k];
    const float alpha = (float)(rz / pq);
    if (lb == 0 && threadIdx.x == 0) hist[2 * k + 1] = pq;
    double rr = 0.0, rzn = 0.0;
    if (z0 < z1) {
        tile[0][ly + 1][lx + 1] = cur.c.p;
        if (sl.has1) tile[0][ly + 1][sl.t1x] = cur.p1;
        if (sl.has2) tile[0][sl.t2y][lx + 1] = cur.p2;
        int buf = 0;
        __syncthreads();
        for (int z = z0; z < z1; ++z) {
            const RawUpd& c = cur.c;
            if (own && c.d != 0u) {
                const uint32_t i = cidx(g, x, y, z);
                const float q = (float)c.d * c.p -
                                (((tile[buf][ly + 1][lx] + tile[buf][ly + 1][lx + 2]) + (tile[buf][ly][lx + 1] + tile[buf][ly + 2][lx + 1])) +
                                 below + pre[0].c.p);
                const float ri = c.r - alpha * q;
                r[i] = ri;
                float h = c.h, l = c.l;
                compensated_add(h, l, alpha * c.p);
                xh[i] = h;
                xl[i] = l;
                const float zi = ri * inv_diag(c.d);
                rr += (double)ri * ri;
                rzn += (double)ri * zi;
            }
            if (z + 1 < z1) {
                tile[buf ^ 1][ly + 1][lx + 1] = pre[0].c.p;
                if (sl.has1) tile[buf ^ 1][ly + 1][sl.t1x] = pre[0].p1;
                if (sl.has2) tile[buf ^ 1][sl.t2y][lx + 1] = pre[0].p2;
            }
            below = c.p;
            cur = pre[0];
#pragma unroll
            for (int m = 0; m + 1 < kPrefetch; ++m) pre[m] = pre[m + 1];
            load_plane(z + 1 + kPrefetch, pre[kPrefetch - 1]);
            __syncthreads();
            buf ^= 1;
        }
    }
    double v[2]{rr, rzn};
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[kPartRR * kPcgMaxBlocks + lb] = v[0];
        partials[kPartRZ * kPcgMaxBlocks + lb] = v[1];
    }
}

}  // namespace ds
