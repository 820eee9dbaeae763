// cwa_kernels.hip — CDNA4 (gfx950) kernels of the coupled SPH +
// heightfield-wave step.
//
// Data layout in HBM (one scene):
//   particles  SoA float4 pos / vel / force + int id, in cell-sorted order;
//              two copies (A = state, B = this step's reorder target)
//   work       rho_w (float), pw (float2: pressure, 1/rho), flux (float)
//   grid       key, trank (int per particle), count (int per cell),
//              offset (int per cell + 1), tmp / list (int per particle)
//   wave       three W*H fp32 fields (newest, previous, write)
//
// Reference kernels replaced (paths relative to the reference root):
//   ugrid_particles_cs.glsl (hash / count / insert)  -> k_hash_count,
//       k_scatter_tmp, k_stable_reorder
//   prefix_sum_cs.glsl + ParallelScan.cpp            -> k_scan_*
//   rho_pres_comp.glsl                               -> k_density
//   force_comp.glsl + integrate_comp.glsl            -> k_force_integrate
//   wave_comp.glsl (EVOLVE, INIT)                    -> k_wave, k_init_wave
//   particle -> heightfield flux (no reference)      -> column blocks of
//                                                       k_force_integrate
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "cwa_device.hpp"
#include "cwa_kernels.hpp"

namespace cwa {

static std::atomic<long long> g_lane_split{kLaneSplitDefault};
long long lane_split_limit() { return g_lane_split.load(std::memory_order_relaxed); }
long long set_lane_split_limit(long long max_n) { return g_lane_split.exchange(max_n); }

// ---------------------------------------------------------------------------
// utilities
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Exclusive scan of int32 (replaces the Blelloch tree of
// prefix_sum_cs.glsl:18-43, which needs 2 log2(n) dispatches and a
// power-of-two n).  Reduce-then-scan over tiles of kScanTile elements:
//   k_scan_reduce  tile sums
//   (recursive scan of the tile sums)
//   k_scan_down    tile-local exclusive scan + tile prefix
// Block-level scans are wave64 shuffle scans plus one LDS exchange of the
// four wave totals.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048
constexpr int kScanDirectTiles = 1024;  // tile prefixes summed in-kernel up to here

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// The kScanItems (8) consecutive values of a thread: two 16-byte loads when
// the whole chunk is in range and the array is 16-byte aligned (a chunk
// starts at a multiple of 8 ints), element loads otherwise.
__device__ __forceinline__ bool aligned16(const void* p) {
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

__device__ __forceinline__ void load_items(const int* __restrict__ in, int base, int n,
                                           int v[kScanItems]) {
    if (base + kScanItems <= n && aligned16(in)) {
        const int4 a = *reinterpret_cast<const int4*>(in + base);
        const int4 b = *reinterpret_cast<const int4*>(in + base + 4);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
        v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = (base + k < n) ? in[base + k] : 0;
    }
}

__device__ __forceinline__ void store_items(int* __restrict__ out, int base, int n,
                                            const int v[kScanItems]) {
    if (base + kScanItems <= n && aligned16(out)) {
        *reinterpret_cast<int4*>(out + base) = make_int4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<int4*>(out + base + 4) = make_int4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) out[base + k] = v[k];
    }
}

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Exclusive scan of one value per thread across the block; returns the
// exclusive prefix and the block total.
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
    __shared__ int wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int incl = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const int s = wsum[w];
        before += (w < wid) ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + incl - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const int* __restrict__ in, int n,
                                                             const int* __restrict__ tile_prefix,
                                                             int* __restrict__ out,
                                                             int* __restrict__ zero_out) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    load_items(in, base, n, v);
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    int total;
    int run = block_excl_scan(s, total) + tile_prefix[blockIdx.x];
    int o[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        o[k] = run;
        run += v[k];
    }
    store_items(out, base, n, o);
    if (zero_out) {
        const int z[kScanItems] = {0, 0, 0, 0, 0, 0, 0, 0};
        store_items(zero_out, base, n, z);
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const int* __restrict__ in, int n,
                                                              int* __restrict__ part) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    load_items(in, base, n, v);
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    __shared__ int wsum[kScanThreads / 64];
    s = wave_sum_i32(s);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += wsum[w];
        part[blockIdx.x] = t;
    }
}

// Exclusive scan of one tile per block.  The tile's prefix is the sum of
// the tile sums before it (part[0 .. blockIdx.x), summed by the block
// itself; null part: single tile).  total_out (optional) receives the grand
// total, written by the last block.  zero_out (optional) is cleared over
// the same index range (the other half of a double-buffered counter).
__global__ void __launch_bounds__(kScanThreads) k_scan_down(const int* __restrict__ in, int n,
                                                            const int* __restrict__ part,
                                                            int* __restrict__ out,
                                                            int* __restrict__ total_out,
                                                            int* __restrict__ zero_out) {
    int prefix = 0;
    if (part) {
        int s = 0;
        for (int t = threadIdx.x; t < (int)blockIdx.x; t += kScanThreads) s += part[t];
        int tot;
        (void)block_excl_scan(s, tot);
        prefix = tot;
    }
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    load_items(in, base, n, v);
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    int total;
    int run = block_excl_scan(s, total) + prefix;
    int o[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        o[k] = run;
        run += v[k];
    }
    store_items(out, base, n, o);
    if (zero_out) {
        const int z[kScanItems] = {0, 0, 0, 0, 0, 0, 0, 0};
        store_items(zero_out, base, n, z);
    }
    if (total_out && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total_out = total + prefix;
}

__global__ void k_scan_total(const int* __restrict__ src, int* __restrict__ dst) {
    if (threadIdx.x == 0) *dst = *src;
}

// Workspace per recursion level with t tiles: tile sums (t) followed by
// their exclusive scan (t + 1, the last entry being the level total).
size_t scan_workspace_ints(int n) {
    size_t ws = 0;
    int m = n;
    while (m > kScanTile * kScanDirectTiles) {
        const int tiles = (m + kScanTile - 1) / kScanTile;
        ws += 2 * (size_t)tiles + 1;
        m = tiles;
    }
    return ws + (size_t)(m + kScanTile - 1) / kScanTile + 1;
}

// out has n+1 entries; out[n] = total.  ws holds scan_workspace_ints(n).
// Up to kScanDirectTiles tiles the tile prefixes are summed inside
// k_scan_down (two launches); larger inputs recurse on the tile sums.
hipError_t launch_scan(const int* in, int* out, int n, int* ws, hipStream_t st,
                       int* zero_out) {
    if (n <= 0) {
        k_fill_i32<<<1, 64, 0, st>>>(out, 1, 0);
        return hipGetLastError();
    }
    const int tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        k_scan_down<<<1, kScanThreads, 0, st>>>(in, n, nullptr, out, out + n, zero_out);
        return hipGetLastError();
    }
    int* part = ws;
    k_scan_reduce<<<tiles, kScanThreads, 0, st>>>(in, n, part);
    if (tiles <= kScanDirectTiles) {
        k_scan_down<<<tiles, kScanThreads, 0, st>>>(in, n, part, out, out + n, zero_out);
        return hipGetLastError();
    }
    int* part_ex = ws + tiles;  // tiles + 1 entries
    hipError_t e = launch_scan(part, part_ex, tiles, ws + 2 * (size_t)tiles + 1, st, nullptr);
    if (e != hipSuccess) return e;
    k_scan_apply<<<tiles, kScanThreads, 0, st>>>(in, n, part_ex, out, zero_out);
    k_scan_total<<<1, 64, 0, st>>>(part_ex + tiles, out + n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Uniform grid: a stable counting sort of keyed items (particle slots, or
// (box, cell) pairs).  Replaces the atomicAdd insertion of
// ugrid_particles_cs.glsl:90-117 (whose order within a cell is not
// deterministic) with the UniformGrid2D::Build order (UniformGrid2D.cpp:
// 75-91: items ascending within each cell):
//   count    one atomicAdd per run of equal keys in a wave (lane order),
//            issued by the run's first lane, its result unused (the wave
//            does not wait for it)
//   scan     exclusive scan of the counts -> offsets
//   scatter  each run head takes a provisional place in its cell's segment
//            from a fill counter and writes (first item, length) there; the
//            runs tile the segment
//   reorder  the run head walks its cell's run records (one per run, not one
//            per item) and counts the items of the runs that start before
//            its own; the lanes of the run take that rank plus their
//            distance to the head.  The particle positions and ids move to
//            their sorted slots in the same pass, which also clears the
//            fill counters for the next build.
// The wave layout (blocks of 256 over the same item range) is identical in
// the three item passes, so every pass sees the same runs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void count_runs(int c, bool valid, int* __restrict__ count) {
    const int lane = threadIdx.x & 63;
    const RunInfo r = wave_runs(c);
    if (r.head && valid) atomicAdd(&count[c], r.run_end - lane);
}

// Cell hash + count of particle slots (ComputeCellIndex / Index / ComputeCounts,
// ugrid_particles_cs.glsl:90-108, with the y-fastest linearisation).
__global__ void __launch_bounds__(256) k_hash_count(const float4* __restrict__ pos, int n,
                                                    GridDesc g, int* __restrict__ key,
                                                    int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int c = -1;
    if (valid) {
        const float4 p = pos[i];
        const int ix = cell_axis(p.x, g.gmin[0], g.cell, g.dims[0]);
        const int iy = cell_axis(p.y, g.gmin[1], g.cell_y, g.dims[1]);
        const int iz = cell_axis(p.z, g.gmin[2], g.cell, g.dims[2]);
        c = lin_cell(g.dims, ix, iy, iz);
        key[i] = c;
    }
    count_runs(c, valid, count);
}

// Count of items whose keys are already known.
__global__ void __launch_bounds__(256) k_key_count(const int* __restrict__ key, int n,
                                                   int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    count_runs(valid ? key[i] : -1, valid, count);
}

// Run records: (first item, length) at a provisional place in the cell's
// segment (fill: zero on entry).
__global__ void __launch_bounds__(256) k_scatter_runs(const int* __restrict__ key,
                                                      const int* __restrict__ offset,
                                                      int* __restrict__ fill, int n,
                                                      int2* __restrict__ runs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    const int c = valid ? key[i] : -1;
    const RunInfo r = wave_runs(c);
    if (valid && r.head) {
        const int len = r.run_end - lane;
        runs[offset[c] + atomicAdd(&fill[c], len)] = make_int2(i, len);
    }
}

// Stable destination of every item; list[d] = item_of[item] (or the item
// itself), and the particle positions and ids move to their sorted slots
// when pos_out is set (the passes that follow read velocities and forces in
// their pre-sort slots through list).  Threads below nfill clear fill.
__global__ void __launch_bounds__(256) k_stable_reorder(
    const int* __restrict__ key, const int* __restrict__ offset,
    const int2* __restrict__ runs, int n, const int* __restrict__ item_of,
    int* __restrict__ list, const float4* __restrict__ pos_in,
    const int* __restrict__ id_in, float4* __restrict__ pos_out, int* __restrict__ id_out,
    int* __restrict__ fill, int nfill) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    if (s < nfill) fill[s] = 0;
    const bool valid = s < n;
    const int c = valid ? key[s] : -1;
    const RunInfo r = wave_runs(c);
    int rank = 0, b = 0;
    if (valid && r.head) {
        b = offset[c];
        const int e = offset[c + 1];
        for (int p = b; p < e;) {
            const int2 u = runs[p];
            if (u.x < s) rank += u.y;
            p += max(u.y, 1);
        }
        rank += b;
    }
    const int d = __shfl(rank, r.my_head, 64) + (lane - r.my_head);
    if (!valid || d < 0 || d >= n) return;  // out of range only on inconsistent run records
    list[d] = item_of ? item_of[s] : s;
    if (pos_out) {
        // read once (nt): the previous force pass's output leaves the L2 to the
        // sorted copy this pass writes and the next passes gather from (A/B
        // on one box: step 0.2606 -> 0.2548 ms, the force pass 109.6 -> 104.6 us)
        pos_out[d] = load_once(&pos_in[s]);
        id_out[d] = load_once(&id_in[s]);
    }
}

// Boxes (UniformGridPbdGpu3D, uniform_grid_pbd_3D_cs.glsl:145-206): the
// cells a box overlaps are ComputeCellIndex(x - hw) .. ComputeCellIndex(x + hw)
// per axis, with a cell size per axis.
struct BoxGrid {
    float gmin[3];
    float cell[3];
    int dims[3];
};

__device__ __forceinline__ void box_cells(const BoxGrid& g, float4 x, float4 hw, int lo[3],
                                          int hi[3]) {
#pragma clang fp contract(off)
    const float a[3] = {x.x - hw.x, x.y - hw.y, x.z - hw.z};
    const float b[3] = {x.x + hw.x, x.y + hw.y, x.z + hw.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = cell_axis(a[k], g.gmin[k], g.cell[k], g.dims[k]);
        hi[k] = cell_axis(b[k], g.gmin[k], g.cell[k], g.dims[k]);
    }
}

// Number of cells each box overlaps (0 for a box whose corners cross).
__global__ void __launch_bounds__(256) k_box_cover(const float4* __restrict__ x,
                                                   const float4* __restrict__ hw, int n,
                                                   BoxGrid g, int* __restrict__ ncov) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo[3], hi[3];
    box_cells(g, x[i], hw[i], lo, hi);
    long long m = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) m *= (long long)max(hi[k] - lo[k] + 1, 0);
    ncov[i] = (int)m;
}

// 64-bit sum of n ints (one atomic per wave), for totals an int scan would wrap.
__global__ void __launch_bounds__(256) k_sum_i64(const int* __restrict__ a, int n,
                                                 unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long v = i < n ? (long long)a[i] : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, (unsigned long long)v);
}

// The (box, cell) pairs in box order, the cells of a box x-outer, then y,
// then z (InsertBox's loop order): pair p = first[i] + t.
__global__ void __launch_bounds__(256) k_box_pairs(const float4* __restrict__ x,
                                                   const float4* __restrict__ hw, int n,
                                                   BoxGrid g, const int* __restrict__ first,
                                                   int* __restrict__ pkey,
                                                   int* __restrict__ pbox) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo[3], hi[3];
    box_cells(g, x[i], hw[i], lo, hi);
    int p = first[i];
    for (int a = lo[0]; a <= hi[0]; ++a)
        for (int b = lo[1]; b <= hi[1]; ++b)
            for (int c = lo[2]; c <= hi[2]; ++c) {
                pkey[p] = (a * g.dims[2] + c) * g.dims[1] + b;  // untiled, y fastest
                pbox[p] = i;
                ++p;
            }
}

// ---------------------------------------------------------------------------
// Neighbour columns: candidate column c = 0..8 is (dx, dz) = (c/3 - 1,
// c%3 - 1) (dx outer, dz inner, the oracle's order); its candidates are the
// contiguous sorted range of cells iy-1..iy+1 of that column.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool coupled(const DevParams& P, float y, float H) {
    return !(P.band > 0.0f) || fabsf(y - H) < P.band;
}

// One density candidate: w = max(T - r2, 0) with T the exact threshold
// (r < h  <=>  r2 < T), so w > 0 exactly for neighbours; accumulates w^3 and
// sets mask bit k for a neighbour.
__device__ __forceinline__ void density_candidate(const float4 pi, const float4 pj, float T,
                                                  float& acc, uint32_t& word, int k) {
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float r2 = dist2_fma(dx, dy, dz);
    const float w = fmaxf(T - r2, 0.0f);
    acc = fmaf(w * w, w, acc);
    word |= min(__float_as_uint(w), 1u) << k;
}

// What both density kernels do with a particle's kernel sum: EOS, wave
// term, rest-density clamp, the neighbour record and the particle flux.
__device__ __forceinline__ void density_finish(const DevParams& P, int i, float4 pi, float4 vi,
                                               float acc, const float* __restrict__ H0,
                                               float* __restrict__ rho_w,
                                               float2* __restrict__ pw, float* __restrict__ flux,
                                               float4* __restrict__ nrec,
                                               float* __restrict__ rho_raw) {
    float rho = P.m_poly6 * acc;
    if (rho_raw) rho_raw[i] = rho;  // taped for the adjoint
    float pressure = fmaxf(P.k * (rho - P.rho0), 0.0f);
    const float u = (pi.x - P.org[0]) * P.uvs[0];
    const float v = (pi.z - P.org[1]) * P.uvs[1];
    const float height = tex_sample(H0, P.W, P.H, u, v);
    if (coupled(P, pi.y, height)) {
        const float wf = height * rho;
        pressure += wf;
        rho += wf / P.k_rp;
    }
    rho = fmaxf(P.rho0, rho);
    const float irho = 1.0f / rho;
    rho_w[i] = rho;
    if (pw) pw[i] = make_float2(pressure, irho);  // taped for the adjoint
    // the 32-byte record the force pass reads for this particle as a neighbour
    nrec[2 * (size_t)i] = make_float4(pi.x, pi.y, pi.z, pressure);
    nrec[2 * (size_t)i + 1] = make_float4(vi.x, vi.y, vi.z, irho);
    float fl = 0.0f;
    if (P.band > 0.0f) {
        const float d = fabsf(pi.y - height);
        if (d < P.band) {
            const float qq = d / P.band;
            const float w = (1.0f - qq * qq) * (1.0f - qq * qq);
            fl = P.m_rho0 * vi.y * w;
        }
    }
    flux[i] = fl;
}

// ---------------------------------------------------------------------------
// Density / pressure (rho_pres_comp.glsl:49-81) + particle flux term.
// Also records the exact neighbour set of every particle as neighbour words
// (the packed non-empty words of each candidate column's 32-candidate
// chunks, the particle itself left out; layout at kNbEntries).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_density(const DevParams* __restrict__ Pd, const float4* __restrict__ pos,
                                                 const float4* __restrict__ vel,
                                                 const int* __restrict__ list,
                                                 const int* __restrict__ offset, int n,
                                                 const float* __restrict__ H0,
                                                 float* __restrict__ rho_w,
                                                 float2* __restrict__ pw,
                                                 float* __restrict__ flux,
                                                 uint2* __restrict__ nbw,
                                                 float4* __restrict__ nrec,
                                                 float* __restrict__ rho_raw,
                                                 float4* __restrict__ vel_out,
                                                 float* __restrict__ vpad) {
    const DevParams& P = *Pd;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = pos[i];
    const float4 vi = vel[list[i]];  // the state's velocity, in its pre-sort slot
    if (vel_out) vel_out[i] = vi;
    vpad[i] = vi.w;
    const int ix = cell_axis(pi.x, P.gmin[0], P.cell, P.dims[0]);
    const int iz = cell_axis(pi.z, P.gmin[2], P.cell, P.dims[2]);
    const float T = P.r2_lt;
    float acc = 0.0f;
    __shared__ uint2 s_held[kNbHeld][256];
    NbWriter wr(&s_held[0][threadIdx.x], 256);
    bool tail = false;  // a column longer than kNbMaxColumn candidates
#pragma unroll 1
    for (int c = 0; c < 9; ++c) {
        int b, e;
        column_range(P, offset, pi.x, pi.y, pi.z, ix, iz, c, b, e);
        const float4* __restrict__ q = pos + b;
        const int len = e - b;
        uint32_t lo = 0, hi = 0;
        const int l1 = min(len, 32);
        int k = 0;
        for (; k + 1 < l1; k += 2) {
            const float4 a0 = q[k], a1 = q[k + 1];
            density_candidate(pi, a0, T, acc, lo, k);
            density_candidate(pi, a1, T, acc, lo, k + 1);
        }
        if (k < l1) density_candidate(pi, q[k], T, acc, lo, k);
        if (len > 32) {
            const int l2 = min(len, 64);
            for (k = 32; k < l2; ++k) density_candidate(pi, q[k], T, acc, hi, k - 32);
            uint32_t none = 0;
            for (k = 64; k < len; ++k) density_candidate(pi, q[k], T, acc, none, 0);
            tail |= len > kNbMaxColumn;
        }
        const uint32_t self = (uint32_t)(i - b);  // i is a candidate of one column
        if (self < 32u) lo &= ~(1u << self);
        else if (self < 64u) hi &= ~(1u << (self - 32u));
        wr.put(nbw, n, i, lo, (uint32_t)b);
        wr.put(nbw, n, i, hi, (uint32_t)b + 32u);
    }
    wr.finish(nbw, n, i, tail);
    density_finish(P, i, pi, vi, acc, H0, rho_w, pw, flux, nrec, rho_raw);
}

// Density for scenes of up to lane_split_limit() particles: kDensitySplit
// consecutive lanes per particle, each testing every kDensitySplit-th
// candidate of each column, so a scene of some ten thousand particles still
// keeps several waves per SIMD busy.  The group's neighbour bits are OR-ed
// together (the words equal k_density's), its partial sums added in a fixed
// tree, and the group's first lane finishes the particle.
constexpr int kDensitySplit = 4;
__global__ void __launch_bounds__(256) k_density_split(
    const DevParams* __restrict__ Pd, const float4* __restrict__ pos,
    const float4* __restrict__ vel, const int* __restrict__ list,
    const int* __restrict__ offset, int n, const float* __restrict__ H0,
    float* __restrict__ rho_w, float2* __restrict__ pw, float* __restrict__ flux,
    uint2* __restrict__ nbw, float4* __restrict__ nrec, float* __restrict__ rho_raw,
    float4* __restrict__ vel_out, float* __restrict__ vpad) {
    const DevParams& P = *Pd;
    const int sub = threadIdx.x & (kDensitySplit - 1);
    const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kDensitySplit);
    if (i >= n) return;  // whole groups: the lanes of a group share i
    const float4 pi = pos[i];
    const float4 vi = vel[list[i]];  // the state's velocity, in its pre-sort slot
    const int ix = cell_axis(pi.x, P.gmin[0], P.cell, P.dims[0]);
    const int iz = cell_axis(pi.z, P.gmin[2], P.cell, P.dims[2]);
    const float T = P.r2_lt;
    float acc = 0.0f;
    // words written by the group's first lane
    __shared__ uint2 s_held[kNbHeld][256 / kDensitySplit];
    NbWriter wr(&s_held[0][threadIdx.x / kDensitySplit], 256 / kDensitySplit);
    bool tail = false;  // a column longer than kNbMaxColumn candidates
#pragma unroll 1
    for (int c = 0; c < 9; ++c) {
        int b, e;
        column_range(P, offset, pi.x, pi.y, pi.z, ix, iz, c, b, e);
        const float4* __restrict__ q = pos + b;
        const int len = e - b;
        uint32_t lo = 0, hi = 0, none = 0;
        int k = sub;
        for (; k < len && k < 32; k += kDensitySplit) density_candidate(pi, q[k], T, acc, lo, k);
        for (; k < len && k < 64; k += kDensitySplit)
            density_candidate(pi, q[k], T, acc, hi, k - 32);
        for (; k < len; k += kDensitySplit) density_candidate(pi, q[k], T, acc, none, 0);
#pragma unroll
        for (int o = 1; o < kDensitySplit; o <<= 1) {
            lo |= __shfl_xor(lo, o, 64);
            hi |= __shfl_xor(hi, o, 64);
        }
        tail |= len > kNbMaxColumn;
        const uint32_t self = (uint32_t)(i - b);  // i is a candidate of one column
        if (self < 32u) lo &= ~(1u << self);
        else if (self < 64u) hi &= ~(1u << (self - 32u));
        if (sub == 0) {
            wr.put(nbw, n, i, lo, (uint32_t)b);
            wr.put(nbw, n, i, hi, (uint32_t)b + 32u);
        }
    }
#pragma unroll
    for (int o = 1; o < kDensitySplit; o <<= 1) acc += __shfl_xor(acc, o, 64);
    if (sub != 0) return;
    wr.finish(nbw, n, i, tail);
    if (vel_out) vel_out[i] = vi;
    vpad[i] = vi.w;
    density_finish(P, i, pi, vi, acc, H0, rho_w, pw, flux, nrec, rho_raw);
}

// Pressure + viscosity contribution of one neighbour (force_comp.glsl:85-86).
struct PairAcc {
    float pfx, pfy, pfz, vfx, vfy, vfz;
};

// Neighbour record written by the density pass: (x, y, z, p), (vx, vy, vz, 1/rho),
// one 32-byte block per particle so a neighbour costs one cache line.
struct Nbr {
    float4 a;
    float4 b;
};

__device__ __forceinline__ Nbr load_nbr(const float4* __restrict__ nrec, int j) {
    return Nbr{nrec[2 * (size_t)j], nrec[2 * (size_t)j + 1]};
}

__device__ __forceinline__ void force_pair(float h, float4 pi, float4 vi, float p_i,
                                           float m_spiky_half, float m_lap, const Nbr& nb,
                                           PairAcc& a) {
    const float dx = pi.x - nb.a.x, dy = pi.y - nb.a.y, dz = pi.z - nb.a.z;
    const float r2 = dist2_fma(dx, dy, dz);
    // an exactly coincident pair (possible after boundary clamping) has no
    // pressure direction: rinv = 0 drops that term
    const float rinv = r2 > 0.0f ? __builtin_amdgcn_rsqf(r2) : 0.0f;
    const float r = r2 * rinv;
    const float hr = h - r;
    const float s = m_spiky_half * (p_i + nb.a.w) * nb.b.w * hr * hr * rinv;
    a.pfx -= s * dx;
    a.pfy -= s * dy;
    a.pfz -= s * dz;
    const float t = m_lap * nb.b.w * hr;
    a.vfx += t * (nb.b.x - vi.x);
    a.vfy += t * (nb.b.y - vi.y);
    a.vfz += t * (nb.b.z - vi.z);
}

// ---------------------------------------------------------------------------
// Particle -> wave flux per vertical column of grid cells.  Each column is
// one contiguous sorted range; one wave64 sums memory column `wave`
// (lane-strided partial sums, then a fixed xor tree: deterministic) and
// stores it at [iz*nx + ix] of its (ix, iz).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void column_flux_sum(const int* __restrict__ offset,
                                                const float* __restrict__ flux, int nx, int ny,
                                                int nz, float* __restrict__ colflux, int wave,
                                                int lane) {
    if (wave >= nx * nz) return;  // uniform across the wave
    const int b = offset[wave * ny], e = offset[(wave + 1) * ny];
    float s = 0.0f;
    for (int k = b + lane; k < e; k += 64) s += flux[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        const int dims[3] = {nx, ny, nz};
        int ix, iz;
        grid_column_xz(dims, wave, ix, iz);
        colflux[iz * nx + ix] = s;
    }
}

// ---------------------------------------------------------------------------
// Forces (force_comp.glsl:62-139) fused with symplectic Euler + boundary
// (integrate_comp.glsl:56-92, 135-178).  Visits only the neighbours the
// density pass recorded (its neighbour words, in candidate order).  Reads
// the particle's own record from the density pass (position, pressure,
// velocity), its id from the sorted copy and its previous force from the
// state in its pre-sort slot (list[i]); writes the next state in the sorted
// slot order.  When `next` is set it also hashes the new position into the
// next step's grid (cell key, run record and count, ComputeCounts of
// ugrid_particles_cs.glsl:90-95), so that step needs no hash pass of its own.
// Blocks nfb.. of the grid sum the particle flux per column instead (one
// wave per column; they read only the density pass's output and run in
// the force blocks' tail).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_force_integrate(
    const DevParams* __restrict__ Pd, float dt, const float4* __restrict__ bpos,
    const int* __restrict__ bid, const float* __restrict__ vpad,
    const float4* __restrict__ force, const int* __restrict__ list,
    const int* __restrict__ offset, int n, const float* __restrict__ rho_w,
    const uint2* __restrict__ nbw, const float4* __restrict__ nrec,
    const float* __restrict__ H0, float4* __restrict__ pos_o,
    float4* __restrict__ vel_o, float4* __restrict__ force_o, int* __restrict__ id_o,
    float* __restrict__ rho_o, float* __restrict__ prs_o, float4* __restrict__ fpre, int nfb,
    const float* __restrict__ pflux, float* __restrict__ colflux, NextKeys next) {
    const DevParams& P = *Pd;
    if ((int)blockIdx.x >= nfb) {
        column_flux_sum(offset, pflux, P.dims[0], P.dims[1], P.dims[2], colflux,
                        (((int)blockIdx.x - nfb) * (int)blockDim.x + (int)threadIdx.x) >> 6,
                        threadIdx.x & 63);
        return;
    }
    const int i = xcd_block(blockIdx.x, nfb) * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int cell = -1;
    __shared__ uint2 s_nb[kNbStage][256];
    if (valid) {
        const float4 own_a = nrec[2 * (size_t)i];      // (x, y, z, p)
        const float4 own_b = nrec[2 * (size_t)i + 1];  // (vx, vy, vz, 1/rho)
        const int src = load_once(&list[i]);           // the particle's slot in the state
        // the previous force is used only after the neighbour loop; issued
        // here, its load is in flight while the loop runs
        float4 fp = load_once(&force[src]);
        const float4 pi = make_float4(own_a.x, own_a.y, own_a.z, load_once(&bpos[i].w));
        const float4 vi = make_float4(own_b.x, own_b.y, own_b.z, load_once(&vpad[i]));
        const float p_i = own_a.w;
        const float h = P.h;
        const float m_spiky_half = P.m * P.spiky * 0.5f;
        const float m_lap = P.m * P.lap;
        PairAcc a{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        uint2* st = &s_nb[0][threadIdx.x];
        const int nw = stage_nb_words(nbw, n, i, st, 256);
        if (nw >= 0) {
            // kForceBatch neighbours per iteration: every record is loaded
            // before any pair is evaluated (slots past the end of the list hold
            // a zero record and are evaluated with zero weight); pairs are
            // accumulated in order.  At four blocks per CU, with the particle's
            // own record and previous force also in flight, 4 measured fastest
            // (MI355X, 1M-particle dam break, same box: 2, 3, 4, 5, 6, 7, 8 ->
            // 116.6, 112.3, 110.1, 115.2, 115.1, 116.5, 115.9 us).
            constexpr int kForceBatch = 4;
            const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
            NbCursor nb(st, 256, nw);
            for (;;) {
                int j[kForceBatch];
                j[0] = nb.next();
                if (j[0] < 0) break;
#pragma unroll
                for (int u = 1; u < kForceBatch; ++u) j[u] = nb.next();
                Nbr r[kForceBatch];
#pragma unroll
                for (int u = 0; u < kForceBatch; ++u) {
                    r[u] = Nbr{zero, zero};
                    if (j[u] >= 0) r[u] = load_nbr(nrec, j[u]);
                }
#pragma unroll
                for (int u = 0; u < kForceBatch; ++u) {
                    const bool live = j[u] >= 0;
                    force_pair(h, pi, vi, p_i, live ? m_spiky_half : 0.0f, live ? m_lap : 0.0f, r[u],
                               a);
                }
            }
        } else {
            for_each_candidate_neighbour(P, offset, bpos, i, pi, [&](int j) {
                force_pair(h, pi, vi, p_i, m_spiky_half, m_lap, load_nbr(nrec, j), a);
            });
        }
        float vfx = a.vfx, vfy = a.vfy, vfz = a.vfz;
        if (pi.y > P.crest) {
            fp.x /= P.brk_mass;
            fp.y /= P.brk_mass;
            fp.z /= P.brk_mass;
            fp.w /= P.brk_mass;
            vfx *= P.brk_visc;
            vfy *= P.brk_visc;
            vfz *= P.brk_visc;
        }
        vfx *= P.visc;
        vfy *= P.visc;
        vfz *= P.visc;
        const float u = (pi.x - P.org[0]) * P.uvs[0];
        const float v = (pi.z - P.org[1]) * P.uvs[1];
        const float height = tex_sample(H0, P.W, P.H, u, v);
        float tqx = (pi.y * fp.z - pi.z * fp.y) * P.torque;
        float tqy = (pi.z * fp.x - pi.x * fp.z) * P.torque;
        float tqz = (pi.x * fp.y - pi.y * fp.x) * P.torque;
        float drx = 0.f, dry = 0.f, drz = 0.f, wfx = 0.f, wfy = 0.f, wfz = 0.f;
        if (coupled(P, pi.y, height)) {
            // WaveVelocity (force_comp.glsl:117-128)
            const float hx = tex_sample(H0, P.W, P.H, u + P.du, v);
            const float hy = tex_sample(H0, P.W, P.H, u, v + P.du);
            const float wvx = (hx - height) / dt, wvy = (hy - height) / dt, wvz = (hx - hy) / P.du;
            drx = -P.drag * (vi.x - wvx);
            dry = -P.drag * (vi.y - wvy);
            drz = -P.drag * (vi.z - wvz);
            // WaveNormal (force_comp.glsl:131-139) = (H(u+1)-H, H(v+1)-H, -1)
            const float nx = tex_sample(H0, P.W, P.H, u + 1.0f, v) - height;
            const float ny = tex_sample(H0, P.W, P.H, u, v + 1.0f) - height;
            wfx = -height * nx * P.wforce;
            wfy = -height * ny * P.wforce;
            wfz = -height * -1.0f * P.wforce;
        }
        float rho = load_once(&rho_w[i]);
        float prs = p_i;
        float Fx = a.pfx + vfx + rho * P.g[0] + tqx + drx + wfx;
        float Fy = a.pfy + vfy + rho * P.g[1] + tqy + dry + wfy;
        float Fz = a.pfz + vfz + rho * P.g[2] + tqz + drz + wfz;
        float Fw = fp.w;
        if (fpre) fpre[i] = make_float4(Fx, Fy, Fz, 0.0f);  // taped for the adjoint

        // integrate_comp.glsl:56-92
        float nvx = vi.x + dt * (Fx / rho), nvy = vi.y + dt * (Fy / rho), nvz = vi.z + dt * (Fz / rho);
        float npx = pi.x + dt * nvx, npy = pi.y + dt * nvy, npz = pi.z + dt * nvz;
        const float damp = 1.0f - P.vdamp * dt;
        nvx *= damp;
        nvy *= damp;
        nvz *= damp;
        if (sqrtf(dist2_fma(nvx, nvy, nvz)) > P.foam_speed) {
            Fx *= P.foam_f;
            Fy *= P.foam_f;
            Fz *= P.foam_f;
            Fw *= P.foam_f;
            rho *= P.foam_rho;
            prs *= P.foam_p;
            nvx *= P.foam_v;
            nvy *= P.foam_v;
            nvz *= P.foam_v;
        }
        const float th = tex_sample(H0, P.W, P.H, (npx - P.org[0]) * P.uvs[0],
                                    (npz - P.org[1]) * P.uvs[1]);
        if (npy < th) npy = th - P.rp;
        // CheckBoundary (integrate_comp.glsl:135-178)
        if (npx < P.lo[0]) { npx = P.lo[0]; nvx *= -P.bounce; }
        else if (npx > P.up[0]) { npx = P.up[0]; nvx *= -P.bounce; }
        if (npy < P.lo[1]) { npy = P.lo[1]; nvy *= -P.bounce; }
        else if (npy > P.up[1]) { npy = P.up[1]; nvy *= -P.bounce; }
        if (npz < P.lo[2]) { npz = P.lo[2]; nvz *= -P.bounce; }
        else if (npz > P.up[2]) { npz = P.up[2]; nvz *= -P.bounce; }
        if (P.wake) {
            if (npy > P.up[1] + P.ceil_extra) { npy = P.up[1] + P.ceil_extra; nvy *= -P.bounce; }
        }
        pos_o[i] = make_float4(npx, npy, npz, pi.w);
        vel_o[i] = make_float4(nvx, nvy, nvz, vi.w);
        force_o[i] = make_float4(Fx, Fy, Fz, Fw);
        id_o[i] = load_once(&bid[i]);
        store_once(&rho_o[i], rho);  // not read by the next step (A/B: step -0.7 %)
        store_once(&prs_o[i], prs);
        if (next.key) {
            const int ix = cell_axis(npx, P.gmin[0], P.cell, P.dims[0]);
            const int iy = cell_axis(npy, P.gmin[1], P.cell_y, P.dims[1]);
            const int iz = cell_axis(npz, P.gmin[2], P.cell, P.dims[2]);
            cell = lin_cell(P.dims, ix, iy, iz);
            next.key[i] = cell;
        }
    }  // valid
    if (next.key) count_runs(cell, valid, next.count);
}

// ---------------------------------------------------------------------------
// Neighbour count of every slot as the last density pass found it (the
// particle itself included, as the density sum of rho_pres_comp.glsl:60-68
// counts it): popcounts of its neighbour words, or the re-search by distance
// for a particle whose words are flagged incomplete.  pos / offset: that
// pass's sorted positions and cell offsets.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_neighbour_counts(const DevParams* __restrict__ Pd,
                                                          const float4* __restrict__ pos,
                                                          const int* __restrict__ offset,
                                                          const uint2* __restrict__ nbw, int n,
                                                          int* __restrict__ out) {
    const DevParams& P = *Pd;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = 1;  // the particle itself
    const int nw = nb_count(nbw, n)[i];
    if (nw < 0) {
        for_each_candidate_neighbour(P, offset, pos, i, pos[i], [&](int) { ++k; });
    } else {
        for (int w = 0; w < nw; ++w) k += __popc(nbw[(size_t)w * n + i].x);
    }
    out[i] = k;
}

// ---------------------------------------------------------------------------
// Wave EVOLVE (wave_comp.glsl:171-206), lambda split per axis, clamped
// neighbours, optional particle flux source, wake line.  Arithmetic is not
// contracted so the result matches the oracle bit for bit.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_wave(const DevParams* __restrict__ Pd, const float* __restrict__ h0,
                                              const float* __restrict__ h1,
                                              float* __restrict__ out,
                                              const float* __restrict__ colflux, float dt) {
#pragma clang fp contract(off)
    const DevParams& P = *Pd;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int W = P.W, H = P.H;
    if (i >= W || j >= H) return;
    const size_t c = (size_t)j * W + i;
    const float C0 = h0[c];
    const float C1 = h1[c];
    const float nn = h0[(size_t)min(j + 1, H - 1) * W + i];
    const float ss = h0[(size_t)max(j - 1, 0) * W + i];
    const float ee = h0[(size_t)j * W + min(i + 1, W - 1)];
    const float ww = h0[(size_t)j * W + max(i - 1, 0)];
    float o = P.wc0 * C0 + P.lx * (ee + ww) + P.lz * (nn + ss) - P.wc1 * C1;
    o *= P.atten;
    if (colflux) {
        const float gx = P.ax * (float)i + P.bx, gz = P.az * (float)j + P.bz;
        const float S = bilerp(colflux, P.dims[0], P.dims[2], gx, gz);
        o += P.kappa * dt * (S * P.inv_area);
    }
    if (P.wake && o > P.wake_thr && i == W / 2) o += P.wake_amt;
    out[c] = o;
}

// The same EVOLVE, four texels per thread along u (W % 4 == 0): float4
// loads of the centre, north and south rows and of h^(n-1), scalar loads of
// the west / east neighbours of the quad.  Per texel the arithmetic is that
// of k_wave, so both give identical results.
__device__ __forceinline__ float wave_texel(const DevParams& P, float C0, float C1, float nn,
                                            float ss, float ee, float ww, int i, bool flux,
                                            float S, float dt) {
#pragma clang fp contract(off)
    float o = P.wc0 * C0 + P.lx * (ee + ww) + P.lz * (nn + ss) - P.wc1 * C1;
    o *= P.atten;
    if (flux) o += P.kappa * dt * (S * P.inv_area);
    if (P.wake && o > P.wake_thr && i == P.W / 2) o += P.wake_amt;
    return o;
}

// Column-flux source of texels i0..i0+3 of row j: bilerp of the column grid
// at each texel's (gx, gz), with the same values and arithmetic as bilerp.
// The four samples share gz and, when they fall in at most two adjacent x
// cells (the column grid is much coarser than the texels), also their six
// grid values, loaded once.
__device__ __forceinline__ void colflux_quad(const DevParams& P, const float* __restrict__ cf,
                                             int i0, int j, float S[4]) {
#pragma clang fp contract(off)
    const int w = P.dims[0], h = P.dims[2];
    const float gz = P.az * (float)j + P.bz;
    float a[4];
    int ix[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float gx = P.ax * (float)(i0 + t) + P.bx;
        float fx = floorf(gx);
        a[t] = gx - fx;
        fx = fminf(fmaxf(fx, -2.0f), (float)(w + 1));
        ix[t] = (int)fx;
    }
    if (ix[3] - ix[0] > 1) {  // a grid finer than the texels: sample one by one
#pragma unroll
        for (int t = 0; t < 4; ++t) S[t] = bilerp(cf, w, h, P.ax * (float)(i0 + t) + P.bx, gz);
        return;
    }
    float fy = floorf(gz);
    const float b = gz - fy;
    fy = fminf(fmaxf(fy, -2.0f), (float)(h + 1));
    const int j0 = (int)fy;
    const size_t ra = (size_t)min(max(j0, 0), h - 1) * w;
    const size_t rb = (size_t)min(max(j0 + 1, 0), h - 1) * w;
    float r0[3], r1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ik = min(max(ix[0] + k, 0), w - 1);
        r0[k] = cf[ra + ik];
        r1[k] = cf[rb + ik];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const bool d = ix[t] != ix[0];
        const float t00 = d ? r0[1] : r0[0], t10 = d ? r0[2] : r0[1];
        const float t01 = d ? r1[1] : r1[0], t11 = d ? r1[2] : r1[1];
        const float q0 = (1.0f - a[t]) * t00 + a[t] * t10;
        const float q1 = (1.0f - a[t]) * t01 + a[t] * t11;
        S[t] = (1.0f - b) * q0 + b * q1;
    }
}

__global__ void __launch_bounds__(256) k_wave4(const DevParams* __restrict__ Pd, const float* __restrict__ h0,
                                               const float* __restrict__ h1,
                                               float* __restrict__ out,
                                               const float* __restrict__ colflux, float dt) {
    const DevParams& P = *Pd;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int W = P.W, H = P.H;
    if (i0 >= W || j >= H) return;
    const size_t row = (size_t)j * W;
    const float4 c0 = *reinterpret_cast<const float4*>(h0 + row + i0);
    // the older level is read once, here (nt; A/B on one box: step -0.75 %)
    const float4 c1 = load_once(reinterpret_cast<const float4*>(h1 + row + i0));
    const float4 nn = *reinterpret_cast<const float4*>(h0 + (size_t)min(j + 1, H - 1) * W + i0);
    const float4 ss = *reinterpret_cast<const float4*>(h0 + (size_t)max(j - 1, 0) * W + i0);
    const float west = h0[row + max(i0 - 1, 0)];
    const float east = h0[row + min(i0 + 4, W - 1)];
    const bool flux = colflux != nullptr;
    float S[4] = {0.f, 0.f, 0.f, 0.f};
    if (flux) colflux_quad(P, colflux, i0, j, S);
    float4 o;
    o.x = wave_texel(P, c0.x, c1.x, nn.x, ss.x, c0.y, west, i0, flux, S[0], dt);
    o.y = wave_texel(P, c0.y, c1.y, nn.y, ss.y, c0.z, c0.x, i0 + 1, flux, S[1], dt);
    o.z = wave_texel(P, c0.z, c1.z, nn.z, ss.z, c0.w, c0.y, i0 + 2, flux, S[2], dt);
    o.w = wave_texel(P, c0.w, c1.w, nn.w, ss.w, east, c0.z, i0 + 3, flux, S[3], dt);
    *reinterpret_cast<float4*>(out + row + i0) = o;
}

// InitWave (wave_comp.glsl:82-140): peak * smoothstep(5, 0, d), d = distance
// to the nearest of 2 or 3 centres.
__global__ void k_init_wave(float* __restrict__ out, int W, int H, int c0x, int c0y, int c1x,
                            int c1y, int c2x, int c2y, int three, float peak) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    auto dist = [&](int cx, int cy) {
        const float dx = (float)(i - cx), dy = (float)(j - cy);
        return sqrtf(dx * dx + dy * dy);
    };
    float d = fminf(dist(c0x, c0y), dist(c1x, c1y));
    if (three) d = fminf(d, dist(c2x, c2y));
    float t = (d - 5.0f) / (0.0f - 5.0f);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    out[(size_t)j * W + i] = peak * (t * t * (3.0f - 2.0f * t));
}

// Non-finite detector for the optional per-step check.
__global__ void k_check_finite(const float* __restrict__ a, size_t n, int* __restrict__ flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    int bad = 0;
    for (; i < n; i += stride) bad |= !isfinite(a[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline int blocks_for(long long n, int t) { return (int)((n + t - 1) / t); }

hipError_t launch_fill_i32(int* p, int n, int v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_fill_i32<<<blocks_for(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_hash_count(const float4* pos, int n, const GridDesc& g, int* key, int* count,
                             hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_hash_count<<<blocks_for(n, 256), 256, 0, st>>>(pos, n, g, key, count);
    return hipGetLastError();
}

hipError_t launch_key_count(const int* key, int n, int* count, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_key_count<<<blocks_for(n, 256), 256, 0, st>>>(key, n, count);
    return hipGetLastError();
}

hipError_t launch_scatter_runs(const int* key, const int* offset, int* fill, int n, int2* runs,
                               hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_scatter_runs<<<blocks_for(n, 256), 256, 0, st>>>(key, offset, fill, n, runs);
    return hipGetLastError();
}

hipError_t launch_stable_reorder(const int* key, const int* offset, const int2* runs, int n,
                                 const int* item_of, int* list, const float4* pos_in,
                                 const int* id_in, float4* pos_out, int* id_out, int* fill,
                                 int nfill, hipStream_t st) {
    const int m = std::max(n, nfill);
    if (m <= 0) return hipSuccess;
    k_stable_reorder<<<blocks_for(m, 256), 256, 0, st>>>(key, offset, runs, n, item_of, list,
                                                         pos_in, id_in, pos_out, id_out, fill,
                                                         nfill);
    return hipGetLastError();
}

hipError_t launch_box_cover(const float4* x, const float4* hw, int n, const float gmin[3],
                            const float cell[3], const int dims[3], int* ncov, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    BoxGrid g{};
    for (int k = 0; k < 3; ++k) {
        g.gmin[k] = gmin[k];
        g.cell[k] = cell[k];
        g.dims[k] = dims[k];
    }
    k_box_cover<<<blocks_for(n, 256), 256, 0, st>>>(x, hw, n, g, ncov);
    return hipGetLastError();
}

hipError_t launch_sum_i64(const int* a, int n, unsigned long long* out, hipStream_t st) {
    k_fill_i32<<<1, 64, 0, st>>>(reinterpret_cast<int*>(out), 2, 0);
    if (n > 0) k_sum_i64<<<blocks_for(n, 256), 256, 0, st>>>(a, n, out);
    return hipGetLastError();
}

hipError_t launch_box_pairs(const float4* x, const float4* hw, int n, const float gmin[3],
                            const float cell[3], const int dims[3], const int* first,
                            int* pkey, int* pbox, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    BoxGrid g{};
    for (int k = 0; k < 3; ++k) {
        g.gmin[k] = gmin[k];
        g.cell[k] = cell[k];
        g.dims[k] = dims[k];
    }
    k_box_pairs<<<blocks_for(n, 256), 256, 0, st>>>(x, hw, n, g, first, pkey, pbox);
    return hipGetLastError();
}

hipError_t launch_density(const DevParams& P, const float4* bpos, const float4* state_vel,
                          const int* list, const int* offset, int n, const float* H0,
                          float* rho_w, float2* pw, float* flux, uint2* nbw, float4* nrec,
                          float* rho_raw, float4* vel_out, float* vpad, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (n <= lane_split_limit())
        k_density_split<<<blocks_for((long long)n * kDensitySplit, 256), 256, 0, st>>>(
            P.dev, bpos, state_vel, list, offset, n, H0, rho_w, pw, flux, nbw, nrec, rho_raw,
            vel_out, vpad);
    else
        k_density<<<blocks_for(n, 256), 256, 0, st>>>(P.dev, bpos, state_vel, list, offset, n,
                                                      H0, rho_w, pw, flux, nbw, nrec, rho_raw,
                                                      vel_out, vpad);
    return hipGetLastError();
}

hipError_t launch_force_integrate(const DevParams& P, float dt, const float4* bpos,
                                  const int* bid, const ParticleArrays& S, const int* list,
                                  const int* offset, int n, const float* rho_w, const uint2* nbw,
                                  const float4* nrec, const float* H0, ParticleArrays& N,
                                  float* rho, float* prs, float4* fpre, const float* flux,
                                  float* colflux, const float* vpad, const NextKeys& next,
                                  hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int nfb = blocks_for(n, 256);
    const long long ncol = colflux ? (long long)P.dims[0] * P.dims[2] : 0;
    const int ncb = blocks_for(ncol * 64, 256);
    k_force_integrate<<<nfb + ncb, 256, 0, st>>>(P.dev, dt, bpos, bid, vpad, S.force, list,
                                                 offset, n, rho_w, nbw, nrec, H0, N.pos, N.vel,
                                                 N.force, N.id, rho, prs, fpre, nfb, flux,
                                                 colflux, next);
    return hipGetLastError();
}

hipError_t launch_neighbour_counts(const DevParams& P, const float4* pos, const int* offset,
                                  const uint2* nbw, int n, int* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_neighbour_counts<<<blocks_for(n, 256), 256, 0, st>>>(P.dev, pos, offset, nbw, n, out);
    return hipGetLastError();
}

hipError_t launch_wave(const DevParams& P, const float* h0, const float* h1, float* out,
                       const float* colflux, float dt, hipStream_t st) {
    if (P.W <= 0 || P.H <= 0) return hipSuccess;
    dim3 blk(64, 4);
    if (P.W % 4 == 0) {
        dim3 grd((P.W / 4 + 63) / 64, (P.H + 3) / 4);
        k_wave4<<<grd, blk, 0, st>>>(P.dev, h0, h1, out, colflux, dt);
    } else {
        dim3 grd((P.W + 63) / 64, (P.H + 3) / 4);
        k_wave<<<grd, blk, 0, st>>>(P.dev, h0, h1, out, colflux, dt);
    }
    return hipGetLastError();
}

hipError_t launch_init_wave(float* out, int W, int H, float type, hipStream_t st) {
    if (W <= 0 || H <= 0) return hipSuccess;
    int c0x, c0y, c1x, c1y, c2x = 0, c2y = 0, three = 1;
    float peak;
    if (type == 1.0f) {  // splash
        c0x = (int)(0.25f * (float)W);
        c0y = (int)(0.25f * (float)H);
        c1x = (int)(0.75f * (float)W);
        c1y = (int)(0.75f * (float)H);
        three = 0;
        peak = 0.5f;
    } else if (type == 0.0f) {  // wave
        c0x = (int)(0.25f * (float)W);
        c0y = H;
        c2x = (int)(0.5f * (float)W);
        c2y = H;
        c1x = (int)(0.75f * (float)W);
        c1y = H;
        peak = 1.0f;
    } else {  // wake
        c0x = (int)(0.5f * (float)W);
        c0y = (int)(0.1f * (float)H);
        c1x = c2x = c0x;
        c1y = c2y = c0y;
        peak = 0.1f;
    }
    dim3 blk(64, 4);
    dim3 grd((W + 63) / 64, (H + 3) / 4);
    k_init_wave<<<grd, blk, 0, st>>>(out, W, H, c0x, c0y, c1x, c1y, c2x, c2y, three, peak);
    return hipGetLastError();
}

hipError_t launch_check_finite(const float* a, size_t n, int* flag, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 2048);
    k_check_finite<<<blocks, 256, 0, st>>>(a, n, flag);
    return hipGetLastError();
}

}  // namespace cwa
