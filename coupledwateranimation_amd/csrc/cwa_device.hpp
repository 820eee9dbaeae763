// cwa_device.hpp — parameter block and inline helpers shared by the HIP
// kernels of the coupled SPH + heightfield-wave step (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cwa {

// Kernel-side parameters: the fields of cwa_params that the kernels read,
// converted once on the host, with the derived constants of every pass.
struct DevParams {
    // SPH
    float h;         // smoothing length
    float h2;        // h*h
    float r2_lt;     // sqrt_rn(r2) < h  <=>  r2 < r2_lt (exact, see host)
    float m;         // particle mass
    float m_poly6;   // mass * 315/(64 pi h^9)
    float rho0;      // rest density
    float k;         // gas constant
    float k_rp;      // gas constant * particle radius
    float rp;        // particle radius
    float spiky;     // -C/(pi h^6)
    float lap;       // +C/(pi h^6)
    float visc;      // mu
    float g[3];      // gravity
    float crest, brk_mass, brk_visc;
    float torque, drag, wforce, du;
    float vdamp, bounce, foam_speed, foam_f, foam_rho, foam_p, foam_v;
    float up[3], lo[3], ceil_extra;
    int wake;        // 0 < type < 1
    // heightfield placement
    float org[2], uvs[2];
    int W, H;        // heightfield texels
    // wave stencil
    float wc0;       // 2 - 2 lx - 2 lz - beta
    float wc1;       // 1 - beta
    float lx, lz, atten, wake_thr, wake_amt;
    // coupling
    float band;      // <= 0: reference behaviour (no gate, no flux)
    float kappa;
    float m_rho0;    // mass / rest density
    float inv_area;  // 1/(cell*cell)
    float ax, bx, az, bz;  // texel -> column-grid coordinate maps
    // grid
    float gmin[3];
    float cell;      // x and z cell edge
    float cell_y;    // y cell edge (cell / grid_ydiv)
    float grid_eps;  // absolute margin of the trimmed candidate rows
    int dims[3];
    int num_cells;
    // Device copy of this block.  Kernels read their parameters through it,
    // so a captured graph replays with whatever the host last uploaded
    // (parameter fits change them between replays).
    const struct DevParams* dev;
};

// Squared distance as the fused chain fma(dz, dz, fma(dy, dy, dx*dx)); the
// oracle evaluates the same chain, so both see the same neighbour set.
__device__ __forceinline__ float dist2_fma(float dx, float dy, float dz) {
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Cell of a point: clamp(floor((p - gmin)/cell), 0, dims-1); the division
// is IEEE correctly rounded (HIP default), matching the oracle bit for bit.
__device__ __forceinline__ int cell_axis(float p, float gmin, float cell, int dim) {
#pragma clang fp contract(off)
    float f = floorf((p - gmin) / cell);
    f = fminf(fmaxf(f, -1.0f), (float)dim);
    int q = (int)f;
    return min(max(q, 0), dim - 1);
}

// Cell linearisation of the particle grid: y fastest, so every vertical
// column of cells is one contiguous range; the columns in z-tiles of
// kTileZ, x-major within a tile (the last tile may be narrower):
//   column(ix, iz) = t*kTileZ*nx + ix*w + (iz - t*kTileZ),
//   t = iz / kTileZ, w = min(kTileZ, nz - t*kTileZ);  lin = column*ny + iy.
// A run of consecutive sorted particles (one XCD's share, one block's)
// then covers an x-range of 16-column z-strips instead of whole z-rows, so
// the neighbour columns at x +- 1 lie closer in memory and in L2 (MI355X,
// 1M-particle dam break: force 109.5 -> 106.8 us; tiles of 8 and 32 alike).
constexpr int kTileZ = 16;
__host__ __device__ __forceinline__ int grid_column(const int dims[3], int ix, int iz) {
    const int t = iz / kTileZ;
    const int rest = dims[2] - t * kTileZ;
    const int w = rest < kTileZ ? rest : kTileZ;
    return t * kTileZ * dims[0] + ix * w + (iz - t * kTileZ);
}
// (ix, iz) of memory column `col` (the inverse of grid_column).
__host__ __device__ __forceinline__ void grid_column_xz(const int dims[3], int col, int& ix,
                                                       int& iz) {
    const int ntile = (dims[2] + kTileZ - 1) / kTileZ;
    const int t0 = col / (kTileZ * dims[0]);
    const int t = t0 < ntile - 1 ? t0 : ntile - 1;
    const int rem = col - t * kTileZ * dims[0];
    const int rest = dims[2] - t * kTileZ;
    const int w = rest < kTileZ ? rest : kTileZ;
    ix = rem / w;
    iz = t * kTileZ + (rem - ix * w);
}
// Loads of data a kernel reads once (nt: streamed through the caches with
// low retention).  In the force pass they keep the L2 for the neighbour
// records it gathers and for the arrays the next step's kernels read (A/B on
// one box: step 0.2625 -> 0.2595 ms, with the force pass itself 2.5 us
// slower and the reorder, scatter, wave and density passes after it faster).
__device__ __forceinline__ float load_once(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ int load_once(const int* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float4 load_once(const float4* p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 load_once(const uint2* p) {
    typedef unsigned v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_uint2(v.x, v.y);
}

// Stores of data no pass of the next step reads (the visible density and
// pressure): nt, so they do not displace what it does read.
__device__ __forceinline__ void store_once(float* p, float v) { __builtin_nontemporal_store(v, p); }

__device__ __forceinline__ int lin_cell(const int dims[3], int ix, int iy, int iz) {
    return grid_column(dims, ix, iz) * dims[1] + iy;
}

// Candidate slot range of a query point p in column c of its 3 x 3 column
// neighbourhood (cx = ix + c/3 - 1, cz = iz + c%3 - 1): the rows of that
// column the support sphere can reach.  The column is skipped when its
// footprint lies at horizontal distance >= h, and its rows are cut to
// |y - py| < sqrt(h^2 - d^2) with d the horizontal distance to the
// footprint; both bounds are widened by grid_eps (well above the rounding
// of the float bounds), so every neighbour stays inside and the exact
// per-candidate test decides.  Density, force and the adjoint call this with
// the same positions, so their candidate ranges (and mask bits) agree.
__device__ __forceinline__ void column_range(const DevParams& P, const int* __restrict__ offset,
                                             float px, float py, float pz, int ix, int iz, int c,
                                             int& b, int& e) {
    const int cx = ix + c / 3 - 1, cz = iz + c % 3 - 1;
    b = e = 0;
    if (cx < 0 || cx >= P.dims[0] || cz < 0 || cz >= P.dims[2]) return;
    float dx = 0.0f, dz = 0.0f;
    if (cx < ix) dx = px - (P.gmin[0] + (float)(cx + 1) * P.cell);
    else if (cx > ix) dx = (P.gmin[0] + (float)cx * P.cell) - px;
    if (cz < iz) dz = pz - (P.gmin[2] + (float)(cz + 1) * P.cell);
    else if (cz > iz) dz = (P.gmin[2] + (float)cz * P.cell) - pz;
    dx = fmaxf(dx - P.grid_eps, 0.0f);
    dz = fmaxf(dz - P.grid_eps, 0.0f);
    const float d2 = dx * dx + dz * dz;
    if (!(d2 < P.r2_lt)) return;
    // the hardware square root (about 1 ulp) is enough for a bound that
    // carries the grid_eps margin, and it is the same in every kernel
    const float ry = __builtin_amdgcn_sqrtf(P.r2_lt - d2) + P.grid_eps;
    const int ylo = cell_axis(py - ry, P.gmin[1], P.cell_y, P.dims[1]);
    const int yhi = cell_axis(py + ry, P.gmin[1], P.cell_y, P.dims[1]);
    b = offset[lin_cell(P.dims, cx, ylo, cz)];
    e = offset[lin_cell(P.dims, cx, yhi, cz) + 1];
}

// ---------------------------------------------------------------------------
// Neighbour words: the exact neighbour set of a particle as recorded by the
// density pass, for the passes that visit it again (force, adjoint).  For
// each candidate column in order, each 32-candidate chunk of the column that
// holds a neighbour of particle i (other than i itself) gives one word
// (bits, base): bit k set when slot base + k is a neighbour.  The words of
// particle i are packed, in candidate order:
//   nbw[k*n + i]                 word k, k < count
//   nb_count(nbw, n)[i]          count, or -1: the particle has more words
//                                than kNbEntries or a column longer than
//                                kNbMaxColumn candidates, and its consumers
//                                re-search its neighbours with the exact
//                                distance test instead
// in one buffer of nb_buffer_uint2(n) uint2 (the counts follow the words).
// Only a particle's non-empty words are written and read back.
// ---------------------------------------------------------------------------
constexpr int kNbEntries = 16;
constexpr int kNbMaxColumn = 64;

__host__ __device__ constexpr size_t nb_buffer_uint2(size_t n) {
    return (size_t)kNbEntries * n + (n + 1) / 2;
}
__device__ __forceinline__ int* nb_count(uint2* nbw, int n) {
    return reinterpret_cast<int*>(nbw + (size_t)kNbEntries * n);
}
__device__ __forceinline__ const int* nb_count(const uint2* nbw, int n) {
    return reinterpret_cast<const int*>(nbw + (size_t)kNbEntries * n);
}

// Writer of one particle's words.  A lane's k-th word belongs in plane k,
// and lanes reach plane k at different columns, so a store issued per word
// would scatter each instruction over several planes.  The first kNbHeld
// words wait in LDS (slot k of the thread at held[k * stride]) and are
// stored plane by plane at the end, each store one contiguous run of the
// wave's particles; later words (rare) are stored as they come.
constexpr int kNbHeld = 8;

struct NbWriter {
    uint2* held;  // kNbHeld LDS slots of this thread, stride `stride`
    int stride;
    int w = 0;    // words so far
    __device__ __forceinline__ NbWriter(uint2* h, int st) : held(h), stride(st) {}
    // Appends the word (bits, base) of particle i when bits is non-zero.
    __device__ __forceinline__ void put(uint2* __restrict__ nbw, int n, int i, uint32_t bits,
                                        uint32_t base) {
        if (!bits) return;
        const uint2 v = make_uint2(bits, base);
        if (w < kNbHeld) held[w * stride] = v;
        else if (w < kNbEntries) nbw[(size_t)w * n + i] = v;
        ++w;
    }
    __device__ __forceinline__ void finish(uint2* __restrict__ nbw, int n, int i,
                                           bool long_column) {
#pragma unroll
        for (int k = 0; k < kNbHeld; ++k)
            if (k < w) nbw[(size_t)k * n + i] = held[k * stride];
        nb_count(nbw, n)[i] = (long_column || w > kNbEntries) ? -1 : w;
    }
};

// The words of one particle (at most kNbEntries), staged per thread in LDS
// (slot t of the thread at s[t * stride]) by stage_nb_words, then walked in
// order by NbCursor: one flat loop over all columns, so a wave runs as many
// iterations as its busiest lane has neighbours rather than the sum over
// columns of each column's busiest lane.  Room for every word also caps a
// 256-thread block at 32 KB of LDS, five blocks per CU (the force pass
// measured 109.4 us with 16 words a particle, five blocks per CU; 110.6 with
// 18, four blocks; 113.2 with 12, six blocks: more waves gathering neighbour
// records through one L1 cost more than they hide).
constexpr int kNbStage = kNbEntries;

// Number of words staged, or -1 when the particle must be re-searched.
__device__ __forceinline__ int stage_nb_words(const uint2* __restrict__ nbw, int n, int i,
                                              uint2* s, int stride) {
    const int nw = load_once(&nb_count(nbw, n)[i]);
    for (int k = 0; k < nw; ++k) s[k * stride] = load_once(&nbw[(size_t)k * n + i]);
    return nw;
}

struct NbCursor {
    const uint2* s;
    int stride, nw, t;
    uint32_t cur;  // bits left in the current word
    int base;

    __device__ __forceinline__ NbCursor(const uint2* st, int stride_, int nw_)
        : s(st), stride(stride_), nw(nw_), t(0), cur(0u), base(0) {}
    // Slot of the next neighbour, -1 when none is left.
    __device__ __forceinline__ int next() {
        if (cur == 0u && t < nw) {
            const uint2 w = s[t * stride];
            cur = w.x;
            base = (int)w.y;
            ++t;
        }
        if (cur == 0u) return -1;
        const int j = base + __builtin_ctz(cur);
        cur &= cur - 1u;
        return j;
    }
};

// Neighbours of particle i (slot i, position pi) in candidate order by the
// exact distance test, excluding i: the path for particles with a column
// of more than 64 candidates.
template <class F>
__device__ __forceinline__ void for_each_candidate_neighbour(const DevParams& P,
                                                             const int* __restrict__ offset,
                                                             const float4* __restrict__ pos,
                                                             int i, float4 pi, F&& f) {
    const int ix = cell_axis(pi.x, P.gmin[0], P.cell, P.dims[0]);
    const int iz = cell_axis(pi.z, P.gmin[2], P.cell, P.dims[2]);
#pragma unroll 1
    for (int c = 0; c < 9; ++c) {
        int b, e;
        column_range(P, offset, pi.x, pi.y, pi.z, ix, iz, c, b, e);
        for (int j = b; j < e; ++j) {
            const float4 pj = pos[j];
            const float r2 = dist2_fma(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
            if (r2 < P.r2_lt && j != i) f(j);
        }
    }
}

// ---------------------------------------------------------------------------
// GL_LINEAR / CLAMP_TO_EDGE sampling of a single-channel fp32 image.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fetch_clamped(const float* __restrict__ img, int w, int h, int i,
                                               int j) {
    i = min(max(i, 0), w - 1);
    j = min(max(j, 0), h - 1);
    return img[(size_t)j * (size_t)w + (size_t)i];
}

__device__ __forceinline__ float bilerp(const float* __restrict__ img, int w, int h, float x,
                                        float y) {
#pragma clang fp contract(off)  // each product and sum rounded, as the oracle's sampler
    float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    fx = fminf(fmaxf(fx, -2.0f), (float)(w + 1));
    fy = fminf(fmaxf(fy, -2.0f), (float)(h + 1));
    const int i0 = (int)fx, j0 = (int)fy;
    const float t00 = fetch_clamped(img, w, h, i0, j0);
    const float t10 = fetch_clamped(img, w, h, i0 + 1, j0);
    const float t01 = fetch_clamped(img, w, h, i0, j0 + 1);
    const float t11 = fetch_clamped(img, w, h, i0 + 1, j0 + 1);
    const float r0 = (1.0f - a) * t00 + a * t10;
    const float r1 = (1.0f - a) * t01 + a * t11;
    return (1.0f - b) * r0 + b * r1;
}

__device__ __forceinline__ float tex_sample(const float* __restrict__ img, int w, int h, float u,
                                            float v) {
#pragma clang fp contract(off)
    return bilerp(img, w, h, u * (float)w - 0.5f, v * (float)h - 0.5f);
}

// Runs of equal keys among the 64 lanes of a wave (lane order): whether the
// lane heads a run, the lane of its run's head and the end of its run
// (exclusive lane index).
struct RunInfo {
    bool head;
    int my_head;
    int run_end;
};

__device__ __forceinline__ RunInfo wave_runs(int c) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(c, 1, 64);
    RunInfo r;
    r.head = lane == 0 || c != prev;
    const unsigned long long heads = __ballot(r.head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long after = heads & ~upto;
    r.run_end = after ? __ffsll((long long)after) - 1 : 64;
    r.my_head = 63 - __clzll((long long)(heads & upto));
    return r;
}

// Blocks are dealt to the 8 XCDs round-robin by workgroup id; remap so that
// XCD k runs one contiguous 1/8 of the logical blocks (its L2 then holds one
// spatial slab of the sorted particles).  Bijective for any grid size.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

}  // namespace cwa
