// cwa_batch2d.hip — the batched 2D coupled step: many independent scenes of
// 2D SPH (Koschier cubic kernel, WCSPH gamma = 3, box SDF) each coupled to
// its own 1D wave column, advanced by one launch per kernel per step.
//
// Reference, per scene (SphWave2D/):
//   SphUgrid::Compute            StencilBuffer.cpp:150-179
//   UniformGridSph2D::CollisionQuery  UniformGridGpu2D.cpp:220-265 and
//                                uniform_grid_sph_cs.glsl:112-183 (hash, count,
//                                insert; here a stable counting sort)
//   ComputeDensityPressure       SphWaveKoschier2D_grid_cs.glsl:393-487
//   ComputeForces (+ Euler)      SphWaveKoschier2D_grid_cs.glsl:490-628
//   ImageStencil::Compute        StencilImage2D.cpp:141-164 with
//                                Wave1D_cs.glsl IterateWave :120-130, FreeBC :139-157
//
// Packing.  Scene s owns the slot range [pbeg_s, pbeg_s + n_s) and the
// global cells [s*C, (s+1)*C) (C = nx*ny cells of size H/2, y fastest), so
// one counting sort over all cells keeps every scene inside its own slot
// range.  A step is 7 launches for any number of scenes:
//   k2_hash      cell keys, wave-aggregated count atomics (run heads)
//   scan         exclusive offsets over all cells (cwa_kernels.hip)
//   k2_scatter   run records into each cell's segment
//   k2_reorder   stable rank from the run records; state A -> B, sorted
//   k2_density   persistent, tile queues; B -> q, (rho, p), (p/rho^2, 1/rho)
//   k2_force     persistent, tile queues; q -> A (forward Euler)
//   k2_wave      one workgroup per scene: column flux sums, the wave
//                substeps in LDS, flux source; levels updated in place
// The state ends every step in A, so a captured step replays as is.
//
// Tiles.  A tile is up to one block of consecutive slots of one scene.
// Scenes are bound to per-XCD queues by longest-processing-time
// assignment of their cost (n_s^2: neighbours per particle grow with n_s in
// a fixed box); blocks take tiles from their XCD's queue (workgroup i runs
// on XCD i mod 8) and steal from the other queues when it is empty.
//
// Arithmetic follows the float restatement of the oracle
// (oracle/cwa_oracle2d.hpp) expression by expression, with contraction
// off, so results are bit-identical to it and independent of the launch
// configuration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cwa2d.h"
#include "cwa_device.hpp"
#include "cwa_kernels.hpp"
#include "cwa_modules.hpp"

#pragma clang fp contract(off)

namespace cwa {

int guarded_call(void (*fn)(void*), void* ctx);

namespace b2d {

constexpr int kNumQueues = 8;   // XCDs of an MI355X
// Candidates loaded ahead in the density / force loops (the lanes of a tile
// walk the same candidates, so each load is one address per wave).  Measured
// on the 256-scene workload (MI355X), cells of H: density 1, 2, 4, 6, 8, 16 ->
// 1.42, 1.25, 1.16, 1.13, 1.12, 1.13 ms; force 1, 2, 3, 4, 6, 8 ->
// 2.63, 2.58, 2.51, 2.49, 2.45, 2.45 ms.  With the shorter columns of H/2
// cells: density 4 / 8 -> 0.921 / 0.926 ms; force 4, 5, 6, 8 -> 1.885,
// 1.884, 1.898, 1.980 ms.  Again after the cheaper pair arithmetic (refined
// hardware square roots and reciprocals): density 3, 4, 6 -> 907, 899,
// 899 us; force 3, 4, 6 -> 1555, 1532, 1546 us.
constexpr int kDensityBatch2D = 4;
constexpr int kForceBatch2D = 4;
constexpr int kRunShift = 26;   // run record: slot | (len - 1) << 26
constexpr unsigned kRunSlotMask = (1u << kRunShift) - 1u;

// Per-scene table (device).
struct Scene2D {
    int pbeg, n, wbeg, L;
    int nsub, model;
    float mass, lambda, dx;
};

// Kernel-side parameters (derived constants in float, as the oracle).
struct Dev2 {
    float H, HSQ, invH, T;          // T: sqrt_rn(r2) < H  <=>  r2 < T
    float k2, k2x2, c6k2, m6k2, c01, ga, gb;  // ga, gb: c6k2 / H, m6k2 / H
    float rho0, k, visc_pre, psi, gx, gy;
    float lox, loy, hix, hiy, walls[4];
    float thr, wref, wpress, wdy, wmask, atten, beta, band, kappa;
    float hc;            // grid cell edge H/2
    float eps;           // margin of the trimmed candidate rows
    int nx, ny, ncell;   // grid cells (H/2) per axis, per scene
    int nxc;             // flux columns (H wide): cell columns 2c, 2c + 1
    int nsph;
};

// ---------------------------------------------------------------------------
// device helpers (each mirrors the oracle function of the same name)
// ---------------------------------------------------------------------------
// Called with r = sqrt of an r2 < HSQ (pairs) or r = d < H (walls); both
// give q = r (1/H) <= 1 (the largest such r times the rounded 1/H rounds to
// at most 1), and at q == 1 the second piece is +0, as the shader's early
// return.  Hence no q >= 1 test here.
__device__ __forceinline__ float W_cubic(const Dev2& c, float r) {
    const float q = r * c.invH;
    if (q <= 0.5f) return c.k2 * (6.0f * q * q * (q - 1.0f) + 1.0f);
    const float q1 = 1.0f - q;
    return c.k2x2 * q1 * q1 * q1;
}

// Called with r < H only (pairs with r2 < T, walls with d < H).  Then
// q = r (1/H) <= 1: the largest float below H times the rounded 1/H is at
// most 1 before rounding.  q >= 1 would be q == 1 exactly, where the second
// piece gives -0 and the shader's early return +0; the sums that take W'
// never hold -0, so both add the same.  Hence no q >= 1 test here.
__device__ __forceinline__ float W_cubic_grad(const Dev2& c, float r) {
    const float q = r * c.invH;
    if (q <= 0.5f) return c.ga * q * (3.0f * q - 2.0f);
    const float q1 = 1.0f - q;
    return c.gb * (q1 * q1);
}

__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

__device__ __forceinline__ float box_sdf(const Dev2& c, float x, float y, float& nx, float& ny) {
    const float d0 = y - c.walls[0], d1 = x - c.walls[1], d2 = c.walls[2] - x,
                d3 = c.walls[3] - y;
    float d = d0;
    nx = 0.0f;
    ny = 1.0f;
    if (!(d < d1)) { d = d1; nx = 1.0f; ny = 0.0f; }
    if (!(d < d2)) { d = d2; nx = -1.0f; ny = 0.0f; }
    if (!(d < d3)) { d = d3; nx = 0.0f; ny = -1.0f; }
    return d;
}

// floor()ed coordinate clamped to [lo, hi], NaN -> lo (fmaxf drops NaN).
__device__ __forceinline__ float clampf_nan(float f, float lo, float hi) {
    return fminf(fmaxf(f, lo), hi);
}

__device__ __forceinline__ float wave_at(const Dev2& c, const float* __restrict__ w, int L,
                                         float sL, float x) {
    const float t = (x - c.lox) * sL - 0.5f;
    const float f0 = floorf(t);
    const float a = t - f0;
    const float f = clampf_nan(f0, -2.0f, (float)(L + 1));
    const int i0 = min(max((int)f, 0), L - 1);
    const int i1 = min(max((int)f + 1, 0), L - 1);
    return (1.0f - a) * w[i0] + a * w[i1];
}

// The wave's coupling velocity at x: uh/h for a shallow-water column (the
// reference's xvel = w[1]/w[0], :362), (h^n - h^(n-1))/dt_sub for a leapfrog
// column.
__device__ __forceinline__ float wave_velocity(const Dev2& c, const Scene2D& S,
                                               const float* __restrict__ w0,
                                               const float* __restrict__ w1, float sL, float x,
                                               float dt) {
    if (S.model == CWA2D_WAVE_SHALLOW)
        return wave_at(c, w1, S.L, sL, x) / wave_at(c, w0, S.L, sL, x);
    const float dts = dt / (float)S.nsub;
    return (wave_at(c, w0, S.L, sL, x) - wave_at(c, w1, S.L, sL, x)) / dts;
}

// Correctly rounded square root of x >= 0, as sqrtf: the hardware square
// root (within 1 ulp) followed by the two residual tests that round it --
// the sequence the compiler emits for sqrtf, without its input scaling for
// x < 2^-96 and its select for 0 and inf, which change nothing for
// x in [2^-96, inf) and at 0.  Inputs in (0, 2^-96) take sqrtf itself.
__device__ __forceinline__ float sqrt_rn(float x) {
    if (__builtin_expect(x < 0x1p-96f && x > 0.0f, 0)) return sqrtf(x);
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1);
    const float sp = __int_as_float(__float_as_int(s) + 1);
    float r = fmaf(-sm, s, x) <= 0.0f ? sm : s;
    r = fmaf(-sp, s, x) > 0.0f ? sp : r;
    return r;
}

// 1 / b correctly rounded (bit-identical to IEEE division) for
// 2^-40 <= b <= 2^40, from a start y within 3 ulp of it: two Newton-Raphson
// steps y + y (1 - b y), each with two fused multiply-adds.  These give the
// correctly rounded reciprocal for every b in that range but one kind, a
// mantissa of all ones, b = (2 - 2^-23) 2^e: there 1 / b lies just above
// the midpoint 2^-(e+1) (1 + 2^-24), the step lands exactly on it and
// rounds to the even 2^-(e+1), whose last mantissa bit is the one missing.
// The carry out of mantissa + 1 is that bit.  Checked
// exhaustively over all 81 binades for starts up to 3 ulp either side.
__device__ __forceinline__ float recip_rn(float b, float y) {
    y = fmaf(fmaf(-b, y, 1.0f), y, y);
    y = fmaf(fmaf(-b, y, 1.0f), y, y);
    const unsigned ub = __float_as_uint(b);
    return __uint_as_float(__float_as_uint(y) | (((ub & 0x7FFFFFu) + 1u) >> 23));
}

// sqrt(x) correctly rounded for 2^-96 <= x <= 2^96, without compares:
// the hardware reciprocal square root y, g = x y and h = y / 2, one coupled
// Newton step on both and Markstein's correction g + (x - g g) h.  2h is
// then within 2 ulp of 1 / sqrt(x) rounded (checked exhaustively), a start
// for recip_rn that saves the hardware reciprocal.
__device__ __forceinline__ float sqrt_rn_mid(float x, float& half_rsqrt) {
    const float y = __builtin_amdgcn_rsqf(x);
    float g = x * y, h = 0.5f * y;
    const float e = fmaf(-g, h, 0.5f);
    g = fmaf(g, e, g);
    h = fmaf(h, e, h);
    half_rsqrt = h;
    return fmaf(fmaf(-g, g, x), h, g);
}

// r = sqrt(max(x, 2^-80)) correctly rounded, for 0 <= x < 2^38: the density
// pass's r.  The clamp (r = 2^-40 for x <= 2^-80) leaves W(r) unchanged:
// with H >= 2^-16, q = r / H <= 2^-24 there, 6 q^2 (q - 1) stays below half
// an ulp of 1 and W rounds to k2 for every such r, 0 included.
__device__ __forceinline__ float sqrt_rn_clamped(float x) {
    // max on the bit patterns: the same as fmaxf for x >= 0, one instruction
    float h;
    return sqrt_rn_mid(__uint_as_float(max(__float_as_uint(x), 0x17800000u)), h);
}

__device__ __forceinline__ int cell2(float p, float lo, float H, int dim) {
    const float f = clampf_nan(floorf((p - lo) / H), -1.0f, (float)dim);
    return min(max((int)f, 0), dim - 1);
}

// Candidate slot range [b, e) of cell column cx for a query at (x, y) in
// cell column ix: the rows of the column the disc of radius H reaches,
// widened by a margin far above the rounding of the float bounds, or an
// empty range when the column lies H or more away -- the oracle's
// column_rows, operation for operation (so both see the same candidates,
// in the same order, also for a neighbour moved out of its cell by the
// density pass's un-penetration).
__device__ __forceinline__ void column_rows(const Dev2& P, const int* __restrict__ off, int cb,
                                            float x, float y, int ix, int cx, int& b, int& e) {
    float dx = 0.0f;
    if (cx < ix) dx = x - (P.lox + (float)(cx + 1) * P.hc);
    else if (cx > ix) dx = (P.lox + (float)cx * P.hc) - x;
    dx = fmaxf(dx - P.eps, 0.0f);
    b = e = 0;
    const float d2 = dx * dx;
    if (!(d2 < P.T)) return;
    const float ry = sqrt_rn(P.T - d2) + P.eps;
    const int ylo = cell2(y - ry, P.loy, P.hc, P.ny);
    const int yhi = cell2(y + ry, P.loy, P.hc, P.ny);
    b = off[cb + cx * P.ny + ylo];
    e = off[cb + cx * P.ny + yhi + 1];
}

// ---------------------------------------------------------------------------
// grid: hash / count
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k2_hash(Dev2 P, const Scene2D* __restrict__ sc, const int* __restrict__ blk_scene,
            const float4* __restrict__ xv, int ntot, int* __restrict__ key,
            int* __restrict__ rec, int* __restrict__ count, int* __restrict__ queues,
            int nqueue_words) {
    if (blockIdx.x == 0 && (int)threadIdx.x < nqueue_words) queues[threadIdx.x] = 0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < ntot;
    int c = -1;
    if (valid) {
        int s = blk_scene[blockIdx.x];
        while (i >= sc[s].pbeg + sc[s].n) ++s;
        const float4 p = xv[i];
        const int ix = cell2(p.x, P.lox, P.hc, P.nx);
        const int iy = cell2(p.y, P.loy, P.hc, P.ny);
        c = s * P.ncell + ix * P.ny + iy;
        key[i] = c;
    }
    // One atomicAdd per run of equal keys (slots of a wave are nearly
    // sorted), issued by the run head for the whole run.
    const RunInfo r = wave_runs(c);
    int base = 0;
    if (r.head && valid) base = atomicAdd(&count[c], r.run_end - lane);
    if (valid) rec[i] = r.head ? base : -(lane - r.my_head);
}

// Run heads write (slot, length) at their place in the cell's segment.
__global__ void __launch_bounds__(256)
    k2_scatter(const int* __restrict__ key, const int* __restrict__ rec,
               const int* __restrict__ offset, int ntot, unsigned* __restrict__ runs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < ntot;
    const int c = valid ? key[i] : -1;
    const RunInfo r = wave_runs(c);
    if (valid && r.head) {
        const unsigned len1 = (unsigned)(r.run_end - lane - 1);
        runs[offset[c] + rec[i]] = (unsigned)i | (len1 << kRunShift);
    }
}

// Stable destination: slots of the cell before this slot's run start come
// from the runs that start earlier; the run itself is contiguous.  Also
// clears the counts for the next build.
__global__ void __launch_bounds__(256)
    k2_reorder(const int* __restrict__ key, const int* __restrict__ rec,
               const int* __restrict__ offset, const unsigned* __restrict__ runs, int ntot,
               const float4* __restrict__ xvA, const float2* __restrict__ rpA,
               const int* __restrict__ idfA, float4* __restrict__ xvB,
               float2* __restrict__ rpB, int* __restrict__ idfB, int* __restrict__ count,
               int ncells) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ncells) count[s] = 0;
    if (s >= ntot) return;
    const int c = key[s];
    const int r = rec[s];
    const int a = r >= 0 ? s : s + r;
    const int b = offset[c], e = offset[c + 1];
    int rank = 0;
    for (int p = b; p < e;) {
        const unsigned u = runs[p];
        const int st = (int)(u & kRunSlotMask);
        const int len = (int)(u >> kRunShift) + 1;
        if (st < a) rank += len;
        p += len;
    }
    const int d = b + rank + (s - a);
    if (d < b || d >= e) return;  // only on inconsistent run records
    xvB[d] = xvA[s];
    rpB[d] = rpA[s];
    idfB[d] = idfA[s];
}

// ---------------------------------------------------------------------------
// tile queues
// ---------------------------------------------------------------------------
// Next tile for a block whose home queue is `home`: its own queue first,
// then the others in order.  Every queue is finite, so every block ends.
__device__ __forceinline__ int next_tile(int* __restrict__ queue, const int* __restrict__ qbeg,
                                         int nq, int home) {
    for (int k = 0; k < nq; ++k) {
        int q = home + k;
        if (q >= nq) q -= nq;
        const int cnt = qbeg[q + 1] - qbeg[q];
        if (cnt == 0) continue;
        if (__hip_atomic_load(&queue[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt)
            continue;
        const int idx = atomicAdd(&queue[q], 1);
        if (idx < cnt) return qbeg[q] + idx;
    }
    return -1;
}

struct TileRef {
    int scene, beg, end;
};

template <int BS>
__device__ __forceinline__ bool fetch_tile(int* __restrict__ queue, const int* __restrict__ qbeg,
                                           const int4* __restrict__ tiles, int nq,
                                           int* s_tile, TileRef& t) {
    if (threadIdx.x == 0) s_tile[0] = next_tile(queue, qbeg, nq, (int)(blockIdx.x % nq));
    __syncthreads();
    const int id = __builtin_amdgcn_readfirstlane(s_tile[0]);
    __syncthreads();
    if (id < 0) return false;
    const int4 d = tiles[id];
    t.scene = __builtin_amdgcn_readfirstlane(d.x);
    t.beg = __builtin_amdgcn_readfirstlane(d.y);
    t.end = __builtin_amdgcn_readfirstlane(d.z);
    return true;
}

// ---------------------------------------------------------------------------
// density / pressure (ComputeDensityPressure, :393-487)
// ---------------------------------------------------------------------------
struct DensityArgs {
    const Scene2D* sc;
    const int4* tiles;
    const int* qbeg;
    int* queue;
    int nq;
    const float4* xvB;
    const float2* rpB;
    const int* idfB;
    const int* off;
    const float* h0;
    const float* h1;
    float4* q;
    float2* rpA;
    float2* aj;
    int* idfA;
    float* flux;
};

template <int BS>
__global__ void __launch_bounds__(BS) k2_density(Dev2 P, float dt, DensityArgs A) {
    __shared__ int s_tile[1];
    TileRef t;
    while (fetch_tile<BS>(A.queue, A.qbeg, A.tiles, A.nq, s_tile, t)) {
        const int i = t.beg + (int)threadIdx.x;
        if (i >= t.end) continue;
        const Scene2D S = A.sc[t.scene];
        const float* w0 = A.h0 + S.wbeg;
        const float* w1 = A.h1 + S.wbeg;
        const float sL = (float)S.L / (P.hix - P.lox);
        const float4 p4 = A.xvB[i];
        float x = p4.x, y = p4.y, vx = p4.z, vy = p4.w;
        int idf = A.idfB[i];
        int en = (idf >> 30) & 1;
        // enable flag (:400-415); with a zero threshold an enabled particle
        // stays enabled, so the samples are skipped
        if (!(en == 1 && P.thr == 0.0f)) {
            const float vw = wave_velocity(P, S, w0, w1, sL, x, dt);
            if (en == 0 && fabsf(vw) > P.thr) en = 1;
            if (en == 1 && fabsf(vw) < P.thr) en = 0;
        }
        float bnx, bny;
        const float d = box_sdf(P, x, y, bnx, bny);
        if (d < 0.0f) {  // un-penetrate (:417-423)
            x -= 0.75f * d * bnx;
            y -= 0.75f * d * bny;
            const float vn = fminf(0.0f, vx * bnx + vy * bny);
            vx -= 1.75f * vn * bnx;
            vy -= 1.75f * vn * bny;
        }
        A.q[i] = make_float4(x, y, vx, vy);
        float f = 0.0f;
        if (P.band > 0.0f) {
            const float dd = fabsf(y - wave_at(P, w0, S.L, sL, x));
            if (dd < P.band) {
                const float qq = dd / P.band;
                const float tq = 1.0f - qq * qq;
                f = S.mass / P.rho0 * vy * tq * tq;
            }
        }
        A.flux[i] = f;
        float2 rp = A.rpB[i];
        float rho = rp.x, p = rp.y;
        if (!(y > P.hiy)) {
            const int ix = cell2(x, P.lox, P.hc, P.nx);
            const int cb = t.scene * P.ncell;
            float acc = 0.0f;
            for (int cx = max(ix - 2, 0); cx <= min(ix + 2, P.nx - 1); ++cx) {
                int b, e;
                column_rows(P, A.off, cb, x, y, ix, cx, b, e);
                auto cand = [&](const float4 pj) {
                    const float dx = x - pj.x, dy = y - pj.y;
                    const float r2 = fmaf(dy, dy, dx * dx);
                    if (r2 < P.HSQ) acc += W_cubic(P, sqrt_rn_clamped(r2));
                };
                // kDensityBatch2D candidates loaded before the first is
                // tested; tested in order
                int j = b;
                for (; j + kDensityBatch2D <= e; j += kDensityBatch2D) {
                    float4 pj[kDensityBatch2D];
#pragma unroll
                    for (int u = 0; u < kDensityBatch2D; ++u) pj[u] = A.xvB[j + u];
#pragma unroll
                    for (int u = 0; u < kDensityBatch2D; ++u) cand(pj[u]);
                }
                for (; j < e; ++j) cand(A.xvB[j]);
            }
            rho = S.mass * acc;
            if (d < P.H) rho += P.psi * W_cubic(P, fmaxf(0.0f, d));
            rho = fmaxf(P.rho0, rho);
            const float tr = rho / P.rho0;
            p = P.k * (tr * tr * tr - 1.0f);
        }
        A.rpA[i] = make_float2(rho, p);
        A.aj[i] = make_float2(p / (rho * rho), 1.0f / rho);
        A.idfA[i] = (idf & ~(1 << 30)) | (en << 30);
    }
}

// ---------------------------------------------------------------------------
// forces + forward Euler (ComputeForces, :490-628)
// ---------------------------------------------------------------------------
struct ForceArgs {
    const Scene2D* sc;
    const int4* tiles;
    const int* qbeg;
    int* queue;
    int nq;
    const float4* q;
    const float2* aj;
    const int* idfA;
    const int* off;
    const float* h0;
    const float* h1;
    float4* xvA;
};

template <int BS>
__global__ void __launch_bounds__(BS) k2_force(Dev2 P, float dt, ForceArgs A) {
    __shared__ int s_tile[1];
    TileRef t;
    while (fetch_tile<BS>(A.queue, A.qbeg, A.tiles, A.nq, s_tile, t)) {
        const int i = t.beg + (int)threadIdx.x;
        if (i >= t.end) continue;
        const float4 qi = A.q[i];
        const float x = qi.x, y = qi.y, vx = qi.z, vy = qi.w;
        const int idf = A.idfA[i];
        if (((idf >> 30) & 1) == 0) {  // disabled: pass through (:494-498)
            A.xvA[i] = qi;
            continue;
        }
        const Scene2D S = A.sc[t.scene];
        const float ai = A.aj[i].x;
        float apx = 0.0f, apy = 0.0f, avx = 0.0f, avy = 0.0f;
        const int ix = cell2(x, P.lox, P.hc, P.nx);
        const int cb = t.scene * P.ncell;
        for (int cx = max(ix - 2, 0); cx <= min(ix + 2, P.nx - 1); ++cx) {
            int b, e;
            column_rows(P, A.off, cb, x, y, ix, cx, b, e);
            auto pair = [&](const float4 qj, const float2 bj) {
                const float dx = x - qj.x, dy = y - qj.y;
                const float r2 = fmaf(dy, dy, dx * dx);
                if (r2 < P.T && r2 > 0.0f) {  // r2 > 0 also excludes j == i
                    // r2 < T <= 2^38 (h <= 2^19); below 2^-80 the library square
                    // root and division
                    const bool wide = r2 < 0x1p-80f;
                    float r, inv;
                    if (__builtin_expect(wide, 0)) {
                        r = sqrtf(r2);
                        inv = 1.0f / r;
                    } else {
                        float h;
                        r = sqrt_rn_mid(r2, h);
                        inv = recip_rn(r, h + h);  // 2h: within 2 ulp of 1 / r
                    }
                    const float wg = W_cubic_grad(P, r);
                    const float ux = dx * inv, uy = dy * inv;
                    const float s = (ai + bj.x) * wg;
                    apx -= s * ux;
                    apy -= s * uy;
                    const float vdr = (vx - qj.z) * dx + (vy - qj.w) * dy;
                    // dot(vij, rij) / (r^2 + 0.01 H^2) as dot * (1 / ...); the
                    // denominator lies in [2^-40, 2^40] for 2^-16 <= H <= 2^19
                    const float den = r2 + P.c01;
                    const float tv = bj.y * vdr * recip_rn(den, __builtin_amdgcn_rcpf(den)) * wg;
                    avx -= tv * ux;
                    avy -= tv * uy;
                }
            };
            // kForceBatch2D candidates (position / velocity and the pressure
            // and 1/rho pair) loaded before the first is tested; tested and
            // accumulated in order.  Loading the pair of every candidate, not
            // only of the neighbours, trades a few cached loads for the
            // latency each neighbour otherwise waits on.
            int j = b;
            for (; j + kForceBatch2D <= e; j += kForceBatch2D) {
                float4 qj[kForceBatch2D];
#pragma unroll
                for (int u = 0; u < kForceBatch2D; ++u) qj[u] = A.q[j + u];
                float2 bj[kForceBatch2D];
#pragma unroll
                for (int u = 0; u < kForceBatch2D; ++u) bj[u] = A.aj[j + u];
#pragma unroll
                for (int u = 0; u < kForceBatch2D; ++u) pair(qj[u], bj[u]);
            }
            for (; j < e; ++j) pair(A.q[j], A.aj[j]);
        }
        const float cv = P.visc_pre * S.mass;
        avx *= cv;
        avy *= cv;
        apx *= S.mass;
        apy *= S.mass;
        float bnx, bny;
        const float d = box_sdf(P, x, y, bnx, bny);
        if (d < P.H) {  // boundary pressure (:553-562)
            const float wg = W_cubic_grad(P, fmaxf(0.0f, d));
            apx += P.psi * ai * wg * -bnx;
            apy += P.psi * ai * wg * -bny;
        }
        float aty = 0.0f;
        if ((idf & 0x3FFFFFFF) % 5 < 4) {  // wave pressure (:586-603)
            const float sL = (float)S.L / (P.hix - P.lox);
            const float hw = wave_at(P, A.h0 + S.wbeg, S.L, sL, x);
            const float vw = wave_velocity(P, S, A.h0 + S.wbeg, A.h1 + S.wbeg, sL, x, dt);
            const float dh = hw - P.wref;
            const float mask = smoothstep(0.0f, P.wmask, fabsf(vw));
            float dyw = y - hw;
            dyw -= 0.5f * dh;
            aty = 0.0f - mask * P.wpress * smoothstep(0.0f, P.wdy, dyw);
        }
        const float ax = apx + avx + P.gx + 0.0f;
        const float ay = apy + avy + P.gy + aty;
        const float v1x = vx + dt * ax, v1y = vy + dt * ay;
        A.xvA[i] = make_float4(x + dt * v1x, y + dt * v1y, v1x, v1y);
    }
}

// ---------------------------------------------------------------------------
// wave column: flux sums, substeps in LDS, flux source (one block per scene)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024)
    k2_wave(Dev2 P, float dt, const Scene2D* __restrict__ sc, const int* __restrict__ off,
            const float* __restrict__ flux, float* __restrict__ h0, float* __restrict__ h1,
            int lds_stride) {
    extern __shared__ float sm[];
    const Scene2D S = sc[blockIdx.x];
    const int L = S.L, tid = threadIdx.x, bd = blockDim.x;
    float* A = sm;
    float* B = sm + lds_stride;
    float* O = sm + 2 * lds_stride;
    float* col = sm + 4 * lds_stride;
    float* g0 = h0 + S.wbeg;
    float* g1 = h1 + S.wbeg;
    for (int k = tid; k < L; k += bd) {
        A[k] = g0[k];
        B[k] = g1[k];
    }
    const bool couple = P.band > 0.0f;
    if (couple) {
        // column sums in the fixed order: lane-strided partials, xor butterfly
        const int lane = tid & 63, wv = tid >> 6, nw = bd >> 6;
        const int cb = blockIdx.x * P.ncell;
        for (int cx = wv; cx < P.nxc; cx += nw) {  // flux column = cell columns 2cx, 2cx + 1
            const int b = off[cb + 2 * cx * P.ny], e = off[cb + min(2 * cx + 2, P.nx) * P.ny];
            float s = 0.0f;
            for (int k = b + lane; k < e; k += 64) s += flux[k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) col[cx] = s;
        }
    }
    __syncthreads();
    const float lam = S.lambda;
    if (S.model == CWA2D_WAVE_SHALLOW) {
        // Shallow1D (Shallow1D_cs.glsl:133-167): A = h, B = uh, O = h at
        // i + 1/2, M = uh at i + 1/2; each phase ends with FreeBC's copy of
        // cells 1 and L-2 (all four channels) to the edges.
        float* M = sm + 3 * lds_stride;
        const float G = 9.8f, hg = 0.5f * G, l2 = lam / 2.0f, dx = S.dx;
        auto edges = [&] {
            if (tid == 0) {
                A[0] = A[1], B[0] = B[1], O[0] = O[1], M[0] = M[1];
                A[L - 1] = A[L - 2], B[L - 1] = B[L - 2], O[L - 1] = O[L - 2],
                      M[L - 1] = M[L - 2];
            }
        };
        for (int s = 0; s < S.nsub; ++s) {
            for (int k = 1 + tid; k < L - 1; k += bd) {
                const int e = k + 1;
                O[k] = (A[k] + A[e]) / 2.0f - l2 * (B[e] - B[k]) / dx;
                M[k] = (B[k] + B[e]) / 2.0f -
                       l2 * (B[e] * B[e] / A[e] + hg * A[e] * A[e] - B[k] * B[k] / A[k] -
                             hg * A[k] * A[k]) / dx;
            }
            __syncthreads();
            edges();
            __syncthreads();
            for (int k = 1 + tid; k < L - 1; k += bd) {
                const int w = k - 1;
                const float h = A[k] - lam * (M[k] - M[w]) / dx;
                const float uh = B[k] - lam * (M[k] * M[k] / O[k] + hg * O[k] * O[k] -
                                               M[w] * M[w] / O[w] - hg * O[w] * O[w]) / dx;
                A[k] = h;
                B[k] = uh;
            }
            __syncthreads();
            edges();
            __syncthreads();
        }
    } else {
        const float ca = 2.0f - 2.0f * lam - P.beta, cb1 = 1.0f - P.beta;
        for (int s = 0; s < S.nsub; ++s) {
            for (int k = 1 + tid; k < L - 1; k += bd) {
                const float v = P.atten * (ca * A[k] + lam * (A[k + 1] + A[k - 1]) - cb1 * B[k]);
                O[k] = v;
                if (k == 1) O[0] = v;          // free boundary (FreeBC :139-157)
                if (k == L - 2) O[L - 1] = v;
            }
            __syncthreads();
            float* tmp = B;
            B = A;
            A = O;
            O = tmp;
        }
    }
    if (couple) {
        const float cw = (P.hix - P.lox) / (float)L;
        const float kd = P.kappa * (dt * (float)P.nsph);
        for (int k = tid; k < L; k += bd) {
            const float gx = ((float)k + 0.5f) * cw * P.invH - 0.5f;
            const float f0 = floorf(gx);
            const float al = gx - f0;
            const float f = clampf_nan(f0, -2.0f, (float)(P.nxc + 1));
            const int c0 = min(max((int)f, 0), P.nxc - 1);
            const int c1 = min(max((int)f + 1, 0), P.nxc - 1);
            const float Sx = (1.0f - al) * col[c0] + al * col[c1];
            A[k] = A[k] + kd * Sx * P.invH;
        }
    }
    for (int k = tid; k < L; k += bd) {
        g0[k] = A[k];
        g1[k] = B[k];
    }
}

// Splash (Shallow1D_cs.glsl:101-115) on the shallow-water scenes (scene < 0)
// or on one: h += -0.1 V exp(-xc^2/0.005), uh += 0.2 |dh| sign(xc).
__global__ void __launch_bounds__(256) k2_splash(const Scene2D* __restrict__ sc, int only,
                                                 float view_height, float* __restrict__ h,
                                                 float* __restrict__ uh) {
    const int s = blockIdx.y;
    if (only >= 0 && s != only) return;
    const Scene2D S = sc[s];
    if (S.model != CWA2D_WAVE_SHALLOW) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < S.L; k += gridDim.x * blockDim.x) {
        const float x = (float)k / (float)(S.L - 1);
        const float xc = x - 0.5f;
        const float dh = -view_height * 0.1f * expf(-xc * xc / 0.005f);
        const float sg = xc > 0.0f ? 1.0f : (xc < 0.0f ? -1.0f : 0.0f);
        h[S.wbeg + k] += dh;
        uh[S.wbeg + k] += 0.2f * fabsf(dh) * sg;
    }
}

// Render copy: (x, y) of every slot into a caller buffer.
__global__ void __launch_bounds__(256) k2_points(const float4* __restrict__ xv, int n,
                                                 float2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 p = xv[i];
        out[i] = make_float2(p.x, p.y);
    }
}

inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

static float exact_r2_threshold(float h) {
    float x = h * h;
    while (x > 0.0f && std::sqrt(x) >= h) x = std::nextafter(x, 0.0f);
    while (!(std::sqrt(x) >= h)) x = std::nextafter(x, INFINITY);
    return x;
}

static int grid_dim(float lo, float hi, float h) {
    const float w = (hi - lo) / h;
    return std::max(1, (int)std::ceil(w));
}

static Dev2 make_dev2(const cwa2d_params& p) {
    Dev2 d{};
    d.H = p.h;
    d.HSQ = d.H * d.H;
    d.invH = 1.0f / d.H;
    d.T = exact_r2_threshold(p.h);
    d.k2 = 40.0f / (7.0f * (float)3.14159265 * d.HSQ);
    d.k2x2 = d.k2 * 2.0f;
    d.c6k2 = 6.0f * d.k2;
    d.m6k2 = -6.0f * d.k2;
    d.ga = d.c6k2 * d.invH;
    d.gb = d.m6k2 * d.invH;
    d.c01 = 0.01f * d.HSQ;
    d.rho0 = p.rest_rho;
    d.k = p.gas_const;
    d.visc_pre = -p.visc * 8.0f;
    d.psi = p.psi;
    d.gx = p.gravity[0];
    d.gy = p.gravity[1];
    d.lox = p.lo[0];
    d.loy = p.lo[1];
    d.hix = p.hi[0];
    d.hiy = p.hi[1];
    for (int i = 0; i < 4; ++i) d.walls[i] = p.walls[i];
    d.thr = p.enable_vel;
    d.wref = p.wave_ref;
    d.wpress = p.wave_press;
    d.wdy = p.wave_dy;
    d.wmask = p.wave_maskv;
    d.atten = p.atten;
    d.beta = p.beta;
    d.band = p.couple_band;
    d.kappa = p.couple_strength;
    d.hc = 0.5f * p.h;
    d.nx = grid_dim(p.lo[0], p.hi[0], d.hc);
    d.ny = grid_dim(p.lo[1], p.hi[1], d.hc);
    d.ncell = d.nx * d.ny;
    d.nxc = grid_dim(p.lo[0], p.hi[0], p.h);
    {   // margin of the trimmed rows: far above the rounding of the bounds
        const float ext = std::max({std::fabs(p.lo[0]), std::fabs(p.hi[0]), std::fabs(p.lo[1]),
                                    std::fabs(p.hi[1]), p.h});
        d.eps = 4e-6f * ext;
    }
    d.nsph = std::max(1, (int)p.sph_substeps);
    return d;
}

static void validate(const cwa2d_params& p) {
    auto fin = [](float v) { return std::isfinite(v); };
    // the pair loops' square root and reciprocals are exact in this range
    if (!(p.h >= 0x1p-16f) || !(p.h <= 0x1p19f)) throw Error(CWA_EINVAL, "h must be in [2^-16, 2^19]");
    if (!(p.rest_rho > 0.0f) || !fin(p.rest_rho)) throw Error(CWA_EINVAL, "rest_rho must be positive");
    if (!(p.hi[0] > p.lo[0]) || !(p.hi[1] > p.lo[1]) || !fin(p.lo[0]) || !fin(p.hi[0]) ||
        !fin(p.lo[1]) || !fin(p.hi[1]))
        throw Error(CWA_EINVAL, "domain needs hi > lo");
    if (!(p.wave_maskv > 0.0f) || !(p.wave_dy > 0.0f))
        throw Error(CWA_EINVAL, "wave_maskv and wave_dy must be positive");
    if (p.sph_substeps < 1 || p.sph_substeps > 64) throw Error(CWA_EINVAL, "sph_substeps in 1..64");
    const double cells = (double)grid_dim(p.lo[0], p.hi[0], 0.5f * p.h) *
                         grid_dim(p.lo[1], p.hi[1], 0.5f * p.h);
    if (cells > (double)(1 << 24)) throw Error(CWA_EINVAL, "grid too fine for the domain (cells per scene > 2^24)");
    if (grid_dim(p.lo[0], p.hi[0], p.h) > 4096) throw Error(CWA_EINVAL, "more than 4096 grid columns");
}

// ---------------------------------------------------------------------------
// Batch2D: the batch as a Module (SphUgrid + ImageStencil per scene)
// ---------------------------------------------------------------------------
enum K2Slot { B_HASH, B_SCAN, B_SCATTER, B_REORDER, B_DENSITY, B_FORCE, B_WAVE };
static const char* kNames[CWA2D_NUM_KERNELS] = {"hash", "scan", "scatter", "reorder",
                                                "density", "force", "wave"};

struct TileSet {
    DeviceBuffer<int4> tiles;
    DeviceBuffer<int> qbeg;  // nq + 1
    int ntiles = 0;
};

class Batch2D {
   public:
    Batch2D(const cwa2d_params& p, int nscenes, const cwa2d_scene* scenes) {
        validate(p);
        if (nscenes < 1) throw Error(CWA_EINVAL, "nscenes must be >= 1");
        params_ = p;
        dev_ = make_dev2(p);
        long long po = 0, wo = 0;
        host_sc_.resize(nscenes);
        lmax_ = 3;
        for (int s = 0; s < nscenes; ++s) {
            const cwa2d_scene& q = scenes[s];
            if (q.n < 1) throw Error(CWA_EINVAL, "scene " + std::to_string(s) + ": n < 1");
            if (q.wave_len < 3 || q.wave_len > CWA2D_MAX_WAVE_LEN)
                throw Error(CWA_EINVAL, "scene " + std::to_string(s) + ": wave_len out of range");
            if (q.wave_substeps < 1 || q.wave_substeps > 1024)
                throw Error(CWA_EINVAL, "scene " + std::to_string(s) + ": wave_substeps in 1..1024");
            if (!(q.mass > 0.0f) || !std::isfinite(q.mass) || !std::isfinite(q.lambda))
                throw Error(CWA_EINVAL, "scene " + std::to_string(s) + ": bad mass or lambda");
            if (q.wave_model != CWA2D_WAVE_LEAPFROG && q.wave_model != CWA2D_WAVE_SHALLOW)
                throw Error(CWA_EINVAL, "scene " + std::to_string(s) + ": unknown wave model");
            if (q.wave_model == CWA2D_WAVE_SHALLOW && (!(q.wave_dx > 0.0f) || !std::isfinite(q.wave_dx)))
                throw Error(CWA_EINVAL, "scene " + std::to_string(s) + ": wave_dx must be positive");
            Scene2D& d = host_sc_[s];
            d.pbeg = (int)po;
            d.n = q.n;
            d.wbeg = (int)wo;
            d.L = q.wave_len;
            d.nsub = q.wave_substeps;
            d.model = q.wave_model;
            d.mass = q.mass;
            d.lambda = q.lambda;
            d.dx = q.wave_dx;
            po += q.n;
            wo += q.wave_len;
            lmax_ = std::max(lmax_, q.wave_len);
        }
        if (po > CWA2D_MAX_PARTICLES) throw Error(CWA_EINVAL, "too many particles in the batch");
        ns_ = nscenes;
        ntot_ = (int)po;
        wtot_ = (int)wo;
        const long long cells = (long long)nscenes * dev_.ncell;
        if (cells + 1 > (1ll << 30)) throw Error(CWA_EINVAL, "too many cells in the batch");
        ncells_ = (int)cells;
        lds_bytes_ = (size_t)(4 * lmax_ + dev_.nxc) * sizeof(float);
        if (lds_bytes_ > 160 * 1024) throw Error(CWA_EINVAL, "wave column and grid exceed LDS");

        hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "stream");
        int dev = 0;
        hip_check(hipGetDevice(&dev), "device");
        hip_check(hipDeviceGetAttribute(&num_cu_, hipDeviceAttributeMultiprocessorCount, dev), "cu count");
        if (lds_bytes_ > 65536)
            hip_check(hipFuncSetAttribute((const void*)k2_wave,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds_bytes_),
                      "wave lds");

        sc_.resize(ns_);
        hip_check(hipMemcpy(sc_.get(), host_sc_.data(), sizeof(Scene2D) * ns_, hipMemcpyHostToDevice), "scenes");
        std::vector<int> bsc(blocks_for(ntot_, 256));
        {
            int s = 0;
            for (size_t b = 0; b < bsc.size(); ++b) {
                const long long i = (long long)b * 256;
                while (i >= host_sc_[s].pbeg + host_sc_[s].n) ++s;
                bsc[b] = s;
            }
        }
        blk_scene_.resize(bsc.size());
        hip_check(hipMemcpy(blk_scene_.get(), bsc.data(), sizeof(int) * bsc.size(), hipMemcpyHostToDevice), "block scenes");
        std::vector<int> pb(ns_), wb(ns_);
        for (int s = 0; s < ns_; ++s) {
            pb[s] = host_sc_[s].pbeg;
            wb[s] = host_sc_[s].wbeg;
        }
        pbeg_.resize(ns_);
        wbeg_.resize(ns_);
        hip_check(hipMemcpy(pbeg_.get(), pb.data(), sizeof(int) * ns_, hipMemcpyHostToDevice), "pbeg");
        hip_check(hipMemcpy(wbeg_.get(), wb.data(), sizeof(int) * ns_, hipMemcpyHostToDevice), "wbeg");

        for (int k = 0; k < 2; ++k) {
            xv_[k].resize(ntot_);
            rp_[k].resize(ntot_);
            idf_[k].resize(ntot_);
        }
        q_.resize(ntot_);
        aj_.resize(ntot_);
        flux_.resize(ntot_);
        key_.resize(ntot_);
        rec_.resize(ntot_);
        runs_.resize(ntot_);
        count_.resize(ncells_ + 1);
        offset_.resize(ncells_ + 1);
        ws_.resize(scan_workspace_ints(ncells_ + 1));
        h0_.resize(wtot_);
        h1_.resize(wtot_);
        queue_.resize(2 * kNumQueues);
        finite_.resize(1);
        hip_check(hipMemset(count_.get(), 0, sizeof(int) * (ncells_ + 1)), "count zero");
        hip_check(hipMemset(offset_.get(), 0, sizeof(int) * (ncells_ + 1)), "offset zero");
        hip_check(hipMemset(key_.get(), 0, sizeof(int) * ntot_), "key zero");
        hip_check(hipMemset(queue_.get(), 0, sizeof(int) * 2 * kNumQueues), "queue zero");

        tune_.block_density = 256;
        tune_.block_force = 256;
        tune_.grid_density = 0;
        tune_.grid_force = 0;
        tune_.xcd_queues = 1;
        tune_.wave_block = 256;
        tune_.ms_per_step = 0.0f;
        for (auto& e : ev_)
            for (auto& x : e) hip_check(hipEventCreate(&x), "event");
    }

    ~Batch2D() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        for (auto& e : ev_)
            for (auto& x : e)
                if (x) (void)hipEventDestroy(x);
        if (st_) {
            (void)hipStreamSynchronize(st_);
            (void)hipStreamDestroy(st_);
        }
    }

    void Reinit(const float* xv, const int32_t* en, const float* h0, const float* h1) {
        if (!xv) throw Error(CWA_EINVAL, "xv is required");
        std::vector<float2> rp(ntot_, make_float2(params_.rest_rho, 0.0f));
        std::vector<int> idf(ntot_);
        for (int s = 0; s < ns_; ++s)
            for (int i = 0; i < host_sc_[s].n; ++i) {
                const int g = host_sc_[s].pbeg + i;
                const int e = en ? (en[g] != 0) : 1;
                idf[g] = i | (e << 30);
            }
        for (long long i = 0; i < 4ll * ntot_; ++i)
            if (!std::isfinite(xv[i])) throw Error(CWA_EINVAL, "non-finite particle state");
        hip_check(hipStreamSynchronize(st_), "sync");
        hip_check(hipMemcpy(xv_[0].get(), xv, sizeof(float4) * ntot_, hipMemcpyHostToDevice), "xv");
        hip_check(hipMemcpy(rp_[0].get(), rp.data(), sizeof(float2) * ntot_, hipMemcpyHostToDevice), "rp");
        hip_check(hipMemcpy(idf_[0].get(), idf.data(), sizeof(int) * ntot_, hipMemcpyHostToDevice), "idf");
        std::vector<float> z(wtot_, 0.0f);
        hip_check(hipMemcpy(h0_.get(), h0 ? h0 : z.data(), sizeof(float) * wtot_, hipMemcpyHostToDevice), "h0");
        hip_check(hipMemcpy(h1_.get(), h1 ? h1 : z.data(), sizeof(float) * wtot_, hipMemcpyHostToDevice), "h1");
        hip_check(hipMemset(count_.get(), 0, sizeof(int) * (ncells_ + 1)), "count zero");
        hip_check(hipDeviceSynchronize(), "reinit");
    }

    void SetParams(const cwa2d_params& p) {
        validate(p);
        const Dev2 d = make_dev2(p);
        if (d.nx != dev_.nx || d.ny != dev_.ny)
            throw Error(CWA_EINVAL, "parameters change the grid dimensions of the batch");
        if ((size_t)(4 * lmax_ + d.nxc) * sizeof(float) > lds_bytes_)
            throw Error(CWA_EINVAL, "grid exceeds the wave pass's LDS");
        params_ = p;
        dev_ = d;
        InvalidateGraph();
    }
    const cwa2d_params& Params() const { return params_; }

    void SetTuning(const cwa2d_tuning& t) {
        auto okb = [](int b) { return b == 64 || b == 128 || b == 256 || b == 512 || b == 1024; };
        if (!okb(t.block_density) || !okb(t.block_force) || !okb(t.wave_block))
            throw Error(CWA_EINVAL, "block sizes must be 64, 128, 256, 512 or 1024");
        if (t.grid_density < 0 || t.grid_force < 0) throw Error(CWA_EINVAL, "grid must be >= 0");
        tune_ = t;
        tune_.xcd_queues = t.xcd_queues ? 1 : 0;
        InvalidateGraph();
    }
    const cwa2d_tuning& Tuning() const { return tune_; }

    // --- one step on the stream (optionally with events around kernels) ---
    void Step(float dt, bool profile) {
        const int nq = tune_.xcd_queues ? kNumQueues : 1;
        const TileSet& td = Tiles(tune_.block_density, nq);
        const TileSet& tf = Tiles(tune_.block_force, nq);
        auto mark = [&](int slot, int which) {
            if (profile) hip_check(hipEventRecord(ev_[slot][which], st_), "event");
        };
        for (int sub = 0; sub < dev_.nsph; ++sub) {
            const bool first = sub == 0;
            mark(B_HASH, first ? 0 : 2);
            k2_hash<<<blocks_for(ntot_, 256), 256, 0, st_>>>(
                dev_, sc_.get(), blk_scene_.get(), xv_[0].get(), ntot_, key_.get(), rec_.get(),
                count_.get(), queue_.get(), 2 * kNumQueues);
            hip_check(hipGetLastError(), "hash");
            mark(B_HASH, first ? 1 : 3);
            mark(B_SCAN, first ? 0 : 2);
            hip_check(launch_scan(count_.get(), offset_.get(), ncells_, ws_.get(), st_, nullptr), "scan");
            mark(B_SCAN, first ? 1 : 3);
            mark(B_SCATTER, first ? 0 : 2);
            k2_scatter<<<blocks_for(ntot_, 256), 256, 0, st_>>>(key_.get(), rec_.get(), offset_.get(),
                                                                ntot_, runs_.get());
            hip_check(hipGetLastError(), "scatter");
            mark(B_SCATTER, first ? 1 : 3);
            mark(B_REORDER, first ? 0 : 2);
            k2_reorder<<<blocks_for(std::max(ntot_, ncells_), 256), 256, 0, st_>>>(
                key_.get(), rec_.get(), offset_.get(), runs_.get(), ntot_, xv_[0].get(),
                rp_[0].get(), idf_[0].get(), xv_[1].get(), rp_[1].get(), idf_[1].get(),
                count_.get(), ncells_);
            hip_check(hipGetLastError(), "reorder");
            mark(B_REORDER, first ? 1 : 3);

            DensityArgs da{sc_.get(), td.tiles.get(), td.qbeg.get(), queue_.get(), nq,
                           xv_[1].get(), rp_[1].get(), idf_[1].get(), offset_.get(), h0_.get(),
                           h1_.get(), q_.get(), rp_[0].get(), aj_.get(), idf_[0].get(), flux_.get()};
            mark(B_DENSITY, first ? 0 : 2);
            LaunchDensity(td, da, dt);
            mark(B_DENSITY, first ? 1 : 3);
            ForceArgs fa{sc_.get(), tf.tiles.get(), tf.qbeg.get(), queue_.get() + kNumQueues, nq,
                         q_.get(), aj_.get(), idf_[0].get(), offset_.get(), h0_.get(), h1_.get(),
                         xv_[0].get()};
            mark(B_FORCE, first ? 0 : 2);
            LaunchForce(tf, fa, dt);
            mark(B_FORCE, first ? 1 : 3);
        }
        mark(B_WAVE, 0);
        k2_wave<<<ns_, tune_.wave_block, lds_bytes_, st_>>>(dev_, dt, sc_.get(), offset_.get(),
                                                             flux_.get(), h0_.get(), h1_.get(), lmax_);
        hip_check(hipGetLastError(), "wave");
        mark(B_WAVE, 1);
    }

    void StepN(float dt, int steps) {
        for (int k = 0; k < steps; ++k) Step(dt, false);
        if (params_.check_finite) CheckFinite();
    }

    void StepGraph(float dt, int steps) {
        if (!graph_exec_ || graph_dt_ != dt) {
            InvalidateGraph();
            // tile tables allocate: build them before the capture
            const int nq = tune_.xcd_queues ? kNumQueues : 1;
            Tiles(tune_.block_density, nq);
            Tiles(tune_.block_force, nq);
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "capture");
            try {
                Step(dt, false);
            } catch (...) {
                (void)hipStreamEndCapture(st_, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(st_, &g), "end capture");
            const hipError_t e = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(e, "graph instantiate");
            graph_dt_ = dt;
        }
        for (int k = 0; k < steps; ++k) hip_check(hipGraphLaunch(graph_exec_, st_), "graph launch");
        if (params_.check_finite) CheckFinite();
    }

    void StepProfiled(float dt, cwa2d_kernel_times* t) {
        Step(dt, true);
        hip_check(hipStreamSynchronize(st_), "sync");
        for (int k = 0; k < CWA2D_NUM_KERNELS; ++k) {
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, ev_[k][0], ev_[k][1]), "elapsed");
            if (dev_.nsph > 1 && k != B_WAVE) {
                float m2 = 0.0f;
                hip_check(hipEventElapsedTime(&m2, ev_[k][2], ev_[k][3]), "elapsed");
                ms += m2 * (float)(dev_.nsph - 1);  // later substeps: last one, scaled
            }
            if (t) t->ms[k] = ms;
        }
        if (params_.check_finite) CheckFinite();
    }

    void CheckFinite() {
        hip_check(hipMemsetAsync(finite_.get(), 0, sizeof(int), st_), "finite zero");
        hip_check(launch_check_finite(reinterpret_cast<const float*>(xv_[0].get()), 4ull * ntot_,
                                      finite_.get(), st_), "finite xv");
        hip_check(launch_check_finite(h0_.get(), (size_t)wtot_, finite_.get(), st_), "finite h0");
        int flag = 0;
        hip_check(hipMemcpyAsync(&flag, finite_.get(), sizeof(int), hipMemcpyDeviceToHost, st_), "finite copy");
        hip_check(hipStreamSynchronize(st_), "sync");
        if (flag) throw Error(CWA_ENONFINITE, "non-finite particle or wave state after the step");
    }

    void Splash(int scene, float view_height) {
        if (scene >= ns_) throw Error(CWA_EINVAL, "scene index out of range");
        dim3 grd(blocks_for(lmax_, 256), ns_);
        k2_splash<<<grd, 256, 0, st_>>>(sc_.get(), scene, view_height, h0_.get(), h1_.get());
        hip_check(hipGetLastError(), "splash");
    }

    void Render(float* points_xy, float* wave) {
        if (points_xy) {
            k2_points<<<blocks_for(ntot_, 256), 256, 0, st_>>>(
                xv_[0].get(), ntot_, reinterpret_cast<float2*>(points_xy));
            hip_check(hipGetLastError(), "render points");
        }
        if (wave)
            hip_check(hipMemcpyAsync(wave, h0_.get(), sizeof(float) * wtot_,
                                     hipMemcpyDeviceToDevice, st_),
                      "render wave");
    }

    void Sync() { hip_check(hipStreamSynchronize(st_), "sync"); }
    hipStream_t stream() const { return st_; }

    void Views(cwa2d_views* v) {
        v->nscenes = ns_;
        v->n_total = ntot_;
        v->wave_total = wtot_;
        v->grid_nx = dev_.nx;
        v->grid_ny = dev_.ny;
        v->cells_total = ncells_;
        v->xv = reinterpret_cast<float*>(xv_[0].get());
        v->rp = reinterpret_cast<float*>(rp_[0].get());
        v->idf = idf_[0].get();
        v->key = key_.get();
        v->offset = offset_.get();
        v->h0 = h0_.get();
        v->h1 = h1_.get();
        v->scene_begin = pbeg_.get();
        v->wave_begin = wbeg_.get();
    }

    void Readback(float* xv, float* rp, int32_t* en, float* h0, float* h1) {
        Sync();
        std::vector<float4> x(ntot_);
        std::vector<float2> r(ntot_);
        std::vector<int> idf(ntot_);
        hip_check(hipMemcpy(x.data(), xv_[0].get(), sizeof(float4) * ntot_, hipMemcpyDeviceToHost), "xv");
        hip_check(hipMemcpy(r.data(), rp_[0].get(), sizeof(float2) * ntot_, hipMemcpyDeviceToHost), "rp");
        hip_check(hipMemcpy(idf.data(), idf_[0].get(), sizeof(int) * ntot_, hipMemcpyDeviceToHost), "idf");
        for (int s = 0; s < ns_; ++s) {
            const Scene2D& S = host_sc_[s];
            for (int k = 0; k < S.n; ++k) {
                const int slot = S.pbeg + k;
                const int id = idf[slot] & 0x3FFFFFFF;
                if (id >= S.n) throw Error(CWA_ESTATE, "particle id out of range");
                const int g = S.pbeg + id;
                if (xv) std::memcpy(xv + 4ll * g, &x[slot], sizeof(float4));
                if (rp) std::memcpy(rp + 2ll * g, &r[slot], sizeof(float2));
                if (en) en[g] = (idf[slot] >> 30) & 1;
            }
        }
        if (h0) hip_check(hipMemcpy(h0, h0_.get(), sizeof(float) * wtot_, hipMemcpyDeviceToHost), "h0");
        if (h1) hip_check(hipMemcpy(h1, h1_.get(), sizeof(float) * wtot_, hipMemcpyDeviceToHost), "h1");
    }

    // --- autotuner -------------------------------------------------------
    cwa2d_tuning Tune(float dt, int steps) {
        if (steps < 1) throw Error(CWA_EINVAL, "tune steps must be >= 1");
        Snapshot snap(*this);
        const cwa2d_tuning start = tune_;
        const int blocks[] = {64, 128, 256, 512};
        const int mults[] = {0, 1, 2, 4, 8};  // persistent blocks per CU (0: one per tile)
        struct Cand {
            cwa2d_tuning t;
            double dens = 0, force = 0;
        };
        std::vector<Cand> cands;
        for (int xq = 0; xq < 2; ++xq)
            for (int b : blocks)
                for (int m : mults) {
                    Cand c;
                    c.t = start;
                    c.t.block_density = c.t.block_force = b;
                    c.t.grid_density = c.t.grid_force = m * num_cu_;
                    c.t.xcd_queues = xq;
                    cands.push_back(c);
                }
        cwa2d_kernel_times kt{};
        for (auto& c : cands) {
            snap.Restore(*this);
            SetTuning(c.t);
            Step(dt, false);  // warm caches and tiles for this setting
            snap.Restore(*this);
            for (int k = 0; k < steps; ++k) {
                StepProfiled(dt, &kt);
                c.dens += kt.ms[B_DENSITY];
                c.force += kt.ms[B_FORCE];
            }
        }
        cwa2d_tuning best = start;
        double bd = 1e30, bf = 1e30;
        for (auto& c : cands) {
            if (c.dens < bd) {
                bd = c.dens;
                best.block_density = c.t.block_density;
                best.grid_density = c.t.grid_density;
                best.xcd_queues = c.t.xcd_queues;
            }
        }
        for (auto& c : cands) {
            if (c.t.xcd_queues != best.xcd_queues) continue;  // one queue layout for both
            if (c.force < bf) {
                bf = c.force;
                best.block_force = c.t.block_force;
                best.grid_force = c.t.grid_force;
            }
        }
        double bw = 1e30;
        for (int wbk : {128, 256, 512, 1024}) {
            cwa2d_tuning t = best;
            t.wave_block = wbk;
            snap.Restore(*this);
            SetTuning(t);
            double w = 0;
            for (int k = 0; k < steps; ++k) {
                StepProfiled(dt, &kt);
                w += kt.ms[B_WAVE];
            }
            if (w < bw) {
                bw = w;
                best.wave_block = wbk;
            }
        }
        // the chosen setting, timed as a graph replay
        snap.Restore(*this);
        SetTuning(best);
        StepGraph(dt, 1);
        Sync();
        snap.Restore(*this);
        hipEvent_t a = ev_[0][0], b = ev_[0][1];
        hip_check(hipEventRecord(a, st_), "event");
        StepGraph(dt, steps);
        hip_check(hipEventRecord(b, st_), "event");
        Sync();
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, a, b), "elapsed");
        best.ms_per_step = ms / (float)steps;
        tune_.ms_per_step = best.ms_per_step;
        snap.Restore(*this);
        return tune_;
    }

   private:
    struct Snapshot {
        DeviceBuffer<float4> xv;
        DeviceBuffer<float2> rp;
        DeviceBuffer<int> idf;
        DeviceBuffer<float> h0, h1;
        explicit Snapshot(Batch2D& b) {
            b.Sync();
            xv.resize(b.ntot_);
            rp.resize(b.ntot_);
            idf.resize(b.ntot_);
            h0.resize(b.wtot_);
            h1.resize(b.wtot_);
            copy(b, false);
        }
        void copy(Batch2D& b, bool restore) {
            auto cp = [&](void* dst, const void* src, size_t bytes) {
                hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, b.st_), "snapshot");
            };
            if (!restore) {
                cp(xv.get(), b.xv_[0].get(), sizeof(float4) * b.ntot_);
                cp(rp.get(), b.rp_[0].get(), sizeof(float2) * b.ntot_);
                cp(idf.get(), b.idf_[0].get(), sizeof(int) * b.ntot_);
                cp(h0.get(), b.h0_.get(), sizeof(float) * b.wtot_);
                cp(h1.get(), b.h1_.get(), sizeof(float) * b.wtot_);
            } else {
                cp(b.xv_[0].get(), xv.get(), sizeof(float4) * b.ntot_);
                cp(b.rp_[0].get(), rp.get(), sizeof(float2) * b.ntot_);
                cp(b.idf_[0].get(), idf.get(), sizeof(int) * b.ntot_);
                cp(b.h0_.get(), h0.get(), sizeof(float) * b.wtot_);
                cp(b.h1_.get(), h1.get(), sizeof(float) * b.wtot_);
            }
            b.Sync();
        }
        void Restore(Batch2D& b) { copy(b, true); }
    };

    void InvalidateGraph() {
        if (graph_exec_) {
            (void)hipStreamSynchronize(st_);
            (void)hipGraphExecDestroy(graph_exec_);
            graph_exec_ = nullptr;
        }
    }

    // Tiles of `bs` slots per scene, scenes bound to nq queues by LPT over
    // n_s^2, tiles of each queue in decreasing scene cost.
    const TileSet& Tiles(int bs, int nq) {
        const int key = bs * 16 + nq;
        auto it = tiles_.find(key);
        if (it != tiles_.end()) return it->second;
        std::vector<int> order(ns_);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return host_sc_[a].n > host_sc_[b].n; });
        std::vector<double> load(nq, 0.0);
        std::vector<std::vector<int>> qs(nq);
        for (int s : order) {
            const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[q] += (double)host_sc_[s].n * (double)host_sc_[s].n;
            qs[q].push_back(s);
        }
        std::vector<int4> tl;
        std::vector<int> qb(nq + 1, 0);
        for (int q = 0; q < nq; ++q) {
            qb[q] = (int)tl.size();
            for (int s : qs[q]) {
                const Scene2D& S = host_sc_[s];
                for (int b = S.pbeg; b < S.pbeg + S.n; b += bs)
                    tl.push_back(make_int4(s, b, std::min(b + bs, S.pbeg + S.n), 0));
            }
        }
        qb[nq] = (int)tl.size();
        TileSet& t = tiles_[key];
        t.ntiles = (int)tl.size();
        t.tiles.resize(tl.size());
        t.qbeg.resize(nq + 1);
        hip_check(hipMemcpy(t.tiles.get(), tl.data(), sizeof(int4) * tl.size(), hipMemcpyHostToDevice), "tiles");
        hip_check(hipMemcpy(t.qbeg.get(), qb.data(), sizeof(int) * (nq + 1), hipMemcpyHostToDevice), "qbeg");
        return t;
    }

    int GridFor(const TileSet& t, int g) const {
        const int n = g > 0 ? std::min(g, t.ntiles) : t.ntiles;
        return std::max(n, 1);
    }

    void LaunchDensity(const TileSet& t, const DensityArgs& a, float dt) {
        const int g = GridFor(t, tune_.grid_density);
        switch (tune_.block_density) {
            case 64: k2_density<64><<<g, 64, 0, st_>>>(dev_, dt, a); break;
            case 128: k2_density<128><<<g, 128, 0, st_>>>(dev_, dt, a); break;
            case 256: k2_density<256><<<g, 256, 0, st_>>>(dev_, dt, a); break;
            case 512: k2_density<512><<<g, 512, 0, st_>>>(dev_, dt, a); break;
            default: k2_density<1024><<<g, 1024, 0, st_>>>(dev_, dt, a); break;
        }
        hip_check(hipGetLastError(), "density");
    }

    void LaunchForce(const TileSet& t, const ForceArgs& a, float dt) {
        const int g = GridFor(t, tune_.grid_force);
        switch (tune_.block_force) {
            case 64: k2_force<64><<<g, 64, 0, st_>>>(dev_, dt, a); break;
            case 128: k2_force<128><<<g, 128, 0, st_>>>(dev_, dt, a); break;
            case 256: k2_force<256><<<g, 256, 0, st_>>>(dev_, dt, a); break;
            case 512: k2_force<512><<<g, 512, 0, st_>>>(dev_, dt, a); break;
            default: k2_force<1024><<<g, 1024, 0, st_>>>(dev_, dt, a); break;
        }
        hip_check(hipGetLastError(), "force");
    }

    cwa2d_params params_{};
    Dev2 dev_{};
    int ns_ = 0, ntot_ = 0, wtot_ = 0, ncells_ = 0, lmax_ = 3, num_cu_ = 256;
    size_t lds_bytes_ = 0;
    std::vector<Scene2D> host_sc_;
    DeviceBuffer<Scene2D> sc_;
    DeviceBuffer<int> blk_scene_, pbeg_, wbeg_;
    DeviceBuffer<float4> xv_[2];
    DeviceBuffer<float2> rp_[2];
    DeviceBuffer<int> idf_[2];
    DeviceBuffer<float4> q_;
    DeviceBuffer<float2> aj_;
    DeviceBuffer<float> flux_;
    DeviceBuffer<int> key_, rec_, count_, offset_, ws_, queue_, finite_;
    DeviceBuffer<unsigned> runs_;
    DeviceBuffer<float> h0_, h1_;
    std::map<int, TileSet> tiles_;
    cwa2d_tuning tune_{};
    hipStream_t st_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    float graph_dt_ = 0.0f;
    hipEvent_t ev_[CWA2D_NUM_KERNELS][4] = {};
};

}  // namespace b2d
}  // namespace cwa

// ===========================================================================
// C ABI
// ===========================================================================
using cwa::Error;
using cwa::b2d::Batch2D;

struct cwa2d_batch {
    Batch2D* impl;
};

namespace {
template <class F>
int guarded2(F&& f) {
    return cwa::guarded_call([](void* c) { (*static_cast<F*>(c))(); }, &f);
}
Batch2D* impl(cwa2d_batch* b) {
    if (!b || !b->impl) throw Error(CWA_EINVAL, "null batch");
    return b->impl;
}
}  // namespace

extern "C" {

int cwa2d_abi_version(void) { return CWA2D_ABI_VERSION; }

int cwa2d_struct_size(int which) {
    switch (which) {
        case 0: return (int)sizeof(cwa2d_params);
        case 1: return (int)sizeof(cwa2d_scene);
        case 2: return (int)sizeof(cwa2d_tuning);
        case 3: return (int)sizeof(cwa2d_views);
        case 4: return (int)sizeof(cwa2d_kernel_times);
        default: return -1;
    }
}

const char* cwa2d_kernel_name(int k) {
    if (k < 0 || k >= CWA2D_NUM_KERNELS) return "";
    return cwa::b2d::kNames[k];
}

int cwa2d_params_reference(cwa2d_params* p) {
    return guarded2([&] {
        if (!p) throw Error(CWA_EINVAL, "null params");
        std::memset(p, 0, sizeof(*p));
        const float prad = 0.025f;                     // PARTICLE_RADIUS
        const float H = 4.0f * prad;                    // :120
        const float view = 2.0f * 4.8f;                 // VIEW_WIDTH/HEIGHT :132-133
        const float k2 = 40.0f / (7.0f * (float)3.14159265 * (H * H));
        p->h = H;
        p->rest_rho = 1000.0f;
        p->gas_const = 35000.0f;
        p->visc = 0.05f;
        p->psi = 1000.0f / (1.5f * k2);                 // :136
        p->gravity[0] = 0.0f;
        p->gravity[1] = -9.8f;
        p->lo[0] = p->lo[1] = 0.0f;
        p->hi[0] = p->hi[1] = view;
        p->walls[0] = 0.5f * view - 15.0f * prad + prad - 0.3f;  // plane0, bottom = 0.3
        p->walls[1] = prad;
        p->walls[2] = view - prad;
        p->walls[3] = view - prad;
        p->enable_vel = 0.05f;
        p->wave_ref = 0.5f * view;
        p->wave_press = 500000.0f;
        p->wave_dy = 10.0f;
        p->wave_maskv = 0.2f;
        p->atten = 0.9995f;
        p->beta = 0.001f;
        p->couple_band = 0.0f;
        p->couple_strength = 0.0f;
        p->sph_substeps = 2;                            // SphWave2D/Main.cpp
        p->check_finite = 0;
    });
}

int cwa2d_wave_init_reference(int32_t L, float* h0, float* h1) {
    return guarded2([&] {
        if (L < 3 || !h0 || !h1) throw Error(CWA_EINVAL, "wave_init: L >= 3 and two outputs");
        for (int mode = 0; mode < 2; ++mode) {
            float* out = mode == 0 ? h1 : h0;
            const int cen = (int)(0.25f * (float)L) + mode;
            for (int k = 0; k < L; ++k) {
                const int x = cen - k;
                out[k] = 0.0f + 0.1f * std::exp((float)(-x * x) / 5000.0f);
            }
        }
    });
}

int cwa2d_wave_init_shallow(int32_t L, float view_height, float* h, float* uh) {
    return guarded2([&] {
        if (L < 3 || !h || !uh) throw Error(CWA_EINVAL, "wave_init_shallow: L >= 3 and two outputs");
        for (int k = 0; k < L; ++k) {
            const float x = (float)k / (float)(L - 1);
            const float xc = x - 0.5f;
            const float h_free = view_height * 0.5f;
            const float bump = view_height * 0.1f * std::exp(-xc * xc / 0.005f);
            const float sg = xc > 0.0f ? 1.0f : (xc < 0.0f ? -1.0f : 0.0f);
            h[k] = h_free + bump;
            uh[k] = 0.5f * std::fabs(bump) * sg;
        }
    });
}

int cwa2d_batch_splash(cwa2d_batch* b, int32_t scene, float view_height) {
    return guarded2([&] { impl(b)->Splash(scene, view_height); });
}

int cwa2d_batch_render(cwa2d_batch* b, float* points_xy, float* wave) {
    return guarded2([&] { impl(b)->Render(points_xy, wave); });
}

int cwa2d_batch_create(const cwa2d_params* p, int32_t nscenes, const cwa2d_scene* scenes,
                       const float* xv, const int32_t* enabled, const float* h0, const float* h1,
                       cwa2d_batch** out) {
    return guarded2([&] {
        if (!p || !scenes || !xv || !out) throw Error(CWA_EINVAL, "null argument");
        *out = nullptr;
        auto* b = new cwa2d_batch{nullptr};
        try {
            b->impl = new Batch2D(*p, nscenes, scenes);
            b->impl->Reinit(xv, enabled, h0, h1);
        } catch (...) {
            delete b->impl;
            delete b;
            throw;
        }
        *out = b;
    });
}

int cwa2d_batch_destroy(cwa2d_batch* b) {
    return guarded2([&] {
        if (!b) return;
        delete b->impl;
        delete b;
    });
}

int cwa2d_batch_reinit(cwa2d_batch* b, const float* xv, const int32_t* enabled, const float* h0,
                       const float* h1) {
    return guarded2([&] { impl(b)->Reinit(xv, enabled, h0, h1); });
}

int cwa2d_batch_set_params(cwa2d_batch* b, const cwa2d_params* p) {
    return guarded2([&] {
        if (!p) throw Error(CWA_EINVAL, "null params");
        impl(b)->SetParams(*p);
    });
}

int cwa2d_batch_get_params(cwa2d_batch* b, cwa2d_params* p) {
    return guarded2([&] {
        if (!p) throw Error(CWA_EINVAL, "null params");
        *p = impl(b)->Params();
    });
}

int cwa2d_batch_step(cwa2d_batch* b, float dt, int32_t steps) {
    return guarded2([&] {
        if (!(dt > 0.0f) || !std::isfinite(dt) || steps < 0) throw Error(CWA_EINVAL, "dt > 0, steps >= 0");
        impl(b)->StepN(dt, steps);
    });
}

int cwa2d_batch_step_graph(cwa2d_batch* b, float dt, int32_t steps) {
    return guarded2([&] {
        if (!(dt > 0.0f) || !std::isfinite(dt) || steps < 0) throw Error(CWA_EINVAL, "dt > 0, steps >= 0");
        impl(b)->StepGraph(dt, steps);
    });
}

int cwa2d_batch_step_profiled(cwa2d_batch* b, float dt, cwa2d_kernel_times* t) {
    return guarded2([&] {
        if (!(dt > 0.0f) || !std::isfinite(dt)) throw Error(CWA_EINVAL, "dt > 0");
        impl(b)->StepProfiled(dt, t);
    });
}

int cwa2d_batch_sync(cwa2d_batch* b) {
    return guarded2([&] { impl(b)->Sync(); });
}

int cwa2d_batch_stream(cwa2d_batch* b, void** s) {
    return guarded2([&] {
        if (!s) throw Error(CWA_EINVAL, "null output");
        *s = (void*)impl(b)->stream();
    });
}

int cwa2d_batch_views(cwa2d_batch* b, cwa2d_views* v) {
    return guarded2([&] {
        if (!v) throw Error(CWA_EINVAL, "null output");
        impl(b)->Views(v);
    });
}

int cwa2d_batch_readback(cwa2d_batch* b, float* xv, float* rp, int32_t* enabled, float* h0,
                         float* h1) {
    return guarded2([&] { impl(b)->Readback(xv, rp, enabled, h0, h1); });
}

int cwa2d_batch_tune(cwa2d_batch* b, float dt, int32_t steps, cwa2d_tuning* best) {
    return guarded2([&] {
        if (!(dt > 0.0f) || !std::isfinite(dt)) throw Error(CWA_EINVAL, "dt > 0");
        const cwa2d_tuning t = impl(b)->Tune(dt, steps);
        if (best) *best = t;
    });
}

int cwa2d_batch_get_tuning(cwa2d_batch* b, cwa2d_tuning* t) {
    return guarded2([&] {
        if (!t) throw Error(CWA_EINVAL, "null output");
        *t = impl(b)->Tuning();
    });
}

int cwa2d_batch_set_tuning(cwa2d_batch* b, const cwa2d_tuning* t) {
    return guarded2([&] {
        if (!t) throw Error(CWA_EINVAL, "null tuning");
        impl(b)->SetTuning(*t);
    });
}

}  // extern "C"
