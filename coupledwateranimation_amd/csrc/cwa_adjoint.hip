// cwa_adjoint.hip — reverse-mode (adjoint) kernels of one coupled step.
//
// The forward step is differentiated as the code runs it, with every
// discrete decision frozen at its forward value: the cell sort and the
// neighbour sets (neighbour words of the forward density pass), the branches of
// max(k(rho - rho0), 0) and max(rho0, rho), the coupling band gate, foam,
// surface clamp, box clamps and the crest factor.  The reference's torque
// term is not differentiated (the adjoint requires torque_coeff == 0).
//
// One reverse step, in sorted order of the forward step:
//   k_adj_wave        texel gather: transpose of the 5-point stencil, the
//                     pass-through of h^n into h^(n-1), and the partials of
//                     lambda_x, lambda_z, atten, beta, kappa; its extra
//                     block rows: column gather, the transpose of the
//                     bilinear column-flux source
//   k_adj_local       per particle: integrate, boundary, surface clamp,
//                     gravity, drag, wave force, crest / viscosity scale
//   k_adj_pairs       neighbour gather: pressure and viscosity pair terms
//                     (both directions, the neighbour relation being
//                     symmetric), then density / EOS / wave term / flux
//   k_adj_rho_pairs   neighbour gather: the poly6 density sum, then the
//                     un-permutation to the pre-sort order
// Heightfield-sample adjoints (particles reading h^n) are scattered with
// float atomics; every particle-particle term is a gather.  Parameter
// partial sums are reduced per block and added in fp64.
#include <hip/hip_runtime.h>

#include "cwa_adjoint.hpp"
#include "cwa_device.hpp"
#include "cwa_kernels.hpp"

namespace cwa {

namespace {

// atomicAdd(p + idx, v) for every lane executing the call, combined per run
// of consecutive lanes with equal idx: the run's first lane adds the run's
// sum with one atomic.  The lanes of a wave hold neighbouring sorted
// particles, whose texture taps mostly coincide, so one contended address
// takes a few atomics per wave instead of up to 64.  Safe in divergent code:
// lanes not executing the call end runs.
__device__ __forceinline__ void atomic_add_runs(float* p, uint32_t idx, float v) {
    const int lane = threadIdx.x & 63;
    const unsigned long long act = __ballot(1);
    const uint32_t prev = __shfl_up(idx, 1, 64);
    const bool head = lane == 0 || !((act >> (lane - 1)) & 1ull) || prev != idx;
    const unsigned long long stop = __ballot(head) | ~act;
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long after = stop & ~upto;
    const int run_end = after ? __ffsll((long long)after) - 1 : 64;
    float s = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float t = __shfl_down(s, o, 64);
        if (lane + o < run_end) s += t;
    }
    if (head && s != 0.0f) atomicAdd(p + idx, s);
}

// Where a heightfield-sample adjoint adds its four weighted taps.
// TexScatterGlobal: run-aggregated global atomics into img_bar.
struct TexScatterGlobal {
    float* img_bar;
    int w;
    __device__ __forceinline__ void operator()(int ia, int ja, float v) const {
        atomic_add_runs(img_bar, (uint32_t)(ja * w + ia), v);
    }
};

// TexScatterWindow: the particles of a block are neighbours in sorted order,
// so most of their taps fall in one small patch of texels, and the taps one
// texture width away (the wave normal's H(u + 1, v), H(u, v + 1)) clamp onto
// the same short stretch of the last column or row.  Taps inside the block's
// kTexWin x kTexWin LDS window (origin ox, oy), or on the last column / row
// beside it, accumulate in LDS with LDS atomics, and tex_window_end adds
// them to img_bar once per block; the other taps take run-aggregated global
// atomics.
constexpr int kTexWin = 64;
constexpr int kTexWinLds = kTexWin * kTexWin + 2 * kTexWin;  // window, last column, last row
struct TexScatterWindow {
    float* img_bar;
    int w, h;
    float* win;  // [kTexWin^2] window, [kTexWin] last column, [kTexWin] last row
    int ox, oy;
    __device__ __forceinline__ void operator()(int ia, int ja, float v) const {
        const unsigned dx = (unsigned)(ia - ox), dy = (unsigned)(ja - oy);
        const unsigned n = (unsigned)kTexWin;
        if (dx < n && dy < n)
            atomicAdd(&win[dy * n + dx], v);
        else if (ia == w - 1 && dy < n)
            atomicAdd(&win[n * n + dy], v);
        else if (ja == h - 1 && dx < n)
            atomicAdd(&win[n * n + n + dx], v);
        else
            atomic_add_runs(img_bar, (uint32_t)(ja * w + ia), v);
    }
};

// Window of the block: origin kTexWin/2 texels before the texel of the
// block's first slot (which exists for every launched block), cleared.
// Called by every thread of the block.
__device__ __forceinline__ TexScatterWindow tex_window_begin(const DevParams& P,
                                                             const float4* __restrict__ pos,
                                                             float* img_bar, float* win,
                                                             int* org) {
    if (threadIdx.x == 0) {
        const float4 x0 = pos[(size_t)blockIdx.x * blockDim.x];
        const float u = (x0.x - P.org[0]) * P.uvs[0] * (float)P.W - 0.5f;
        const float v = (x0.z - P.org[1]) * P.uvs[1] * (float)P.H - 0.5f;
        org[0] = (int)floorf(fminf(fmaxf(u, -1e6f), 1e6f)) - kTexWin / 2;
        org[1] = (int)floorf(fminf(fmaxf(v, -1e6f), 1e6f)) - kTexWin / 2;
    }
    for (int t = threadIdx.x; t < kTexWinLds; t += blockDim.x) win[t] = 0.0f;
    __syncthreads();
    return TexScatterWindow{img_bar, P.W, P.H, win, org[0], org[1]};
}

// Add the block's LDS accumulators to img_bar (non-zero texels inside the
// image).  Called by every thread of the block.
__device__ __forceinline__ void tex_window_end(const DevParams& P, const TexScatterWindow& sc) {
    __syncthreads();
    for (int t = threadIdx.x; t < kTexWinLds; t += blockDim.x) {
        const float v = sc.win[t];
        int tx, ty;
        if (t < kTexWin * kTexWin) {
            tx = sc.ox + t % kTexWin;
            ty = sc.oy + t / kTexWin;
        } else if (t < kTexWin * kTexWin + kTexWin) {
            tx = P.W - 1;
            ty = sc.oy + (t - kTexWin * kTexWin);
        } else {
            tx = sc.ox + (t - kTexWin * kTexWin - kTexWin);
            ty = P.H - 1;
        }
        if (v != 0.0f && tx >= 0 && tx < P.W && ty >= 0 && ty < P.H)
            atomicAdd(&sc.img_bar[(size_t)ty * P.W + tx], v);
    }
}

// d(bilerp)/d(x, y) at continuous texel coordinates; scatters g times the
// four bilinear weights through `add` (if not null).  Coordinates and
// weights are rounded as the forward sampler rounds them.
template <class Scatter>
__device__ __forceinline__ void bilerp_adj(const float* __restrict__ img, const Scatter* add,
                                           int w, int h, float x, float y, float g, float& gx,
                                           float& gy) {
#pragma clang fp contract(off)
    float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    fx = fminf(fmaxf(fx, -2.0f), (float)(w + 1));
    fy = fminf(fmaxf(fy, -2.0f), (float)(h + 1));
    const int i0 = (int)fx, j0 = (int)fy;
    const int ia = min(max(i0, 0), w - 1), ib = min(max(i0 + 1, 0), w - 1);
    const int ja = min(max(j0, 0), h - 1), jb = min(max(j0 + 1, 0), h - 1);
    const float t00 = img[(size_t)ja * w + ia], t10 = img[(size_t)ja * w + ib];
    const float t01 = img[(size_t)jb * w + ia], t11 = img[(size_t)jb * w + ib];
    gx = g * ((1.0f - b) * (t10 - t00) + b * (t11 - t01));
    gy = g * ((1.0f - a) * (t01 - t00) + a * (t11 - t10));
    if (add) {
        (*add)(ia, ja, g * (1.0f - a) * (1.0f - b));
        (*add)(ib, ja, g * a * (1.0f - b));
        (*add)(ia, jb, g * (1.0f - a) * b);
        (*add)(ib, jb, g * a * b);
    }
}

// Adjoint of texture(img, (u, v)): returns d/du, d/dv times g.
template <class Scatter>
__device__ __forceinline__ void tex_adj(const float* __restrict__ img, const Scatter* add, int w,
                                        int h, float u, float v, float g, float& gu, float& gv) {
#pragma clang fp contract(off)
    float gx, gy;
    bilerp_adj(img, add, w, h, u * (float)w - 0.5f, v * (float)h - 0.5f, g, gx, gy);
    gu = gx * (float)w;
    gv = gy * (float)h;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-reduce NP partial sums and add them to this block's own fp64 slots
// acc[block * NP + k]; each block owns its slots, so steps accumulate in
// stream order without atomics and the final sum over blocks runs in a
// fixed order (k_reduce_partials).
template <int NP>
__device__ __forceinline__ void block_accumulate(float (&v)[NP], double* __restrict__ acc) {
    __shared__ float red[NP][4];  // blocks of 256 threads: 4 waves
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int l = tid & 63, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float s = wave_sum(v[k]);
        if (l == 0) red[k][w] = s;
    }
    __syncthreads();
    if (tid < NP) {
        float t = 0.0f;
        for (int q = 0; q < nw; ++q) t += red[tid][q];
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        acc[b * NP + tid] += (double)t;
    }
}

__device__ __forceinline__ bool coupled_(const DevParams& P, float y, float H) {
    return !(P.band > 0.0f) || fabsf(y - H) < P.band;
}

// Visit the frozen neighbours of particle i (the density pass's neighbour
// words, candidate order), excluding i: load(j) fetches what one pair needs
// and use(record) consumes it.  U records are loaded before the first of
// them is used; records are used in order.  st: the thread's LDS staging
// slots (kNbStage of them, stride 256).
// for_each_neighbour for lane sub of a group of G lanes sharing particle
// i: the lane visits the neighbours whose position in the walk order is
// sub mod G (G = 1: all of them).
template <int U, int G, class Load, class Use>
__device__ __forceinline__ void for_each_neighbour_of_lane(
    const DevParams& P, const int* __restrict__ offset, const uint2* __restrict__ nbw,
    const float4* __restrict__ pos, int n, int i, float4 pi, uint2* st, int sub, Load&& load,
    Use&& use) {
    const int nw = stage_nb_words(nbw, n, i, st, 256);
    if (nw < 0) {
        int t = 0;
        for_each_candidate_neighbour(P, offset, pos, i, pi, [&](int j) {
            if ((t++ % G) == sub) use(load(j));
        });
        return;
    }
    NbCursor nb(st, 256, nw);
    for (int s = 0; s < sub; ++s) (void)nb.next();
    const auto next = [&]() {
        const int j = nb.next();
#pragma unroll
        for (int s = 1; s < G; ++s) (void)nb.next();
        return j;
    };
    for (;;) {
        int j[U];
        j[0] = next();
        if (j[0] < 0) break;
#pragma unroll
        for (int u = 1; u < U; ++u) j[u] = next();
        decltype(load(0)) r[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j[u] >= 0) r[u] = load(j[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j[u] >= 0) use(r[u]);
    }
}

template <int U, class Load, class Use>
__device__ __forceinline__ void for_each_neighbour(const DevParams& P,
                                                   const int* __restrict__ offset,
                                                   const uint2* __restrict__ nbw,
                                                   const float4* __restrict__ pos, int n, int i,
                                                   float4 pi, uint2* st, Load&& load, Use&& use) {
    const int nw = stage_nb_words(nbw, n, i, st, 256);
    if (nw < 0) {
        for_each_candidate_neighbour(P, offset, pos, i, pi, [&](int j) { use(load(j)); });
        return;
    }
    NbCursor nb(st, 256, nw);
    for (;;) {
        int j[U];
        j[0] = nb.next();
        if (j[0] < 0) break;
#pragma unroll
        for (int u = 1; u < U; ++u) j[u] = nb.next();
        decltype(load(0)) r[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j[u] >= 0) r[u] = load(j[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j[u] >= 0) use(r[u]);
    }
}

// What the pair adjoint reads of neighbour j.
struct AdjNbr {
    float4 ra, rb;    // neighbour record (x, y, z, p), (vx, vy, vz, 1/rho)
    float4 Fbj, Vbj;  // j's adjoint record
};

}  // namespace

// ---------------------------------------------------------------------------
// Column-flux adjoint: Sbar[col] = kappa dt / cell^2 * sum over texels t of
// G0[t] * (bilinear weight of col in t's source sample).  One wave per
// column visits the texels whose sample can touch it (texel -> column map
// is affine), in a fixed order.  Runs as extra blocks of the wave adjoint's
// launch (both read only G0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void adj_colflux_column(const DevParams& P, float dt,
                                                   const float* __restrict__ G0,
                                                   float* __restrict__ Sbar, int col, int lane) {
    // one wave64 per column; lanes stride over the column's texel window
    const int nx = P.dims[0], nz = P.dims[2];
    if (col >= nx * nz) return;  // uniform across the wave
    const int cx = col % nx, cz = col / nx;  // Sbar / colflux layout [iz*nx + ix]
    const int W = P.W, H = P.H;
    // texel range whose gx lies in (cx - 1, cx + 1), widened by one texel;
    // edge columns also collect the texels clamped onto them
    int ilo = 0, ihi = W - 1, jlo = 0, jhi = H - 1;
    if (cx > 0) ilo = max(0, (int)floorf(((float)cx - 1.0f - P.bx) / P.ax) - 1);
    if (cx < nx - 1) ihi = min(W - 1, (int)ceilf(((float)cx + 1.0f - P.bx) / P.ax) + 1);
    if (cz > 0) jlo = max(0, (int)floorf(((float)cz - 1.0f - P.bz) / P.az) - 1);
    if (cz < nz - 1) jhi = min(H - 1, (int)ceilf(((float)cz + 1.0f - P.bz) / P.az) + 1);
    const int ni = ihi - ilo + 1, nj = jhi - jlo + 1;
    const long long total = (ni > 0 && nj > 0) ? (long long)ni * nj : 0;
    float acc = 0.0f;
    for (long long t = lane; t < total; t += 64) {
        const int j = jlo + (int)(t / ni), i = ilo + (int)(t % ni);
        const float gz = P.az * (float)j + P.bz;
        float fz = floorf(gz);
        const float bz = gz - fz;
        fz = fminf(fmaxf(fz, -2.0f), (float)(nz + 1));
        const int z0 = min(max((int)fz, 0), nz - 1), z1 = min(max((int)fz + 1, 0), nz - 1);
        const float wz = (z0 == cz ? 1.0f - bz : 0.0f) + (z1 == cz ? bz : 0.0f);
        const float gx = P.ax * (float)i + P.bx;
        float fx = floorf(gx);
        const float bx = gx - fx;
        fx = fminf(fmaxf(fx, -2.0f), (float)(nx + 1));
        const int x0 = min(max((int)fx, 0), nx - 1), x1 = min(max((int)fx + 1, 0), nx - 1);
        const float wx = (x0 == cx ? 1.0f - bx : 0.0f) + (x1 == cx ? bx : 0.0f);
        acc += G0[(size_t)j * W + i] * (wx * wz);
    }
    acc = wave_sum(acc);
    if (lane == 0) Sbar[col] = P.kappa * dt * P.inv_area * acc;
}

// ---------------------------------------------------------------------------
// Wave adjoint.  G0 = adjoint of h^(n+1) (the step's output), G1 = adjoint of
// h^n carried by the pass-through into h^(n-1) of the next step.  Writes
// G0n = adjoint of h^n (before particle-sample contributions) and
// G1n = adjoint of h^(n-1).  Block rows nby.. compute the column-flux
// adjoint Sbar instead (when Sbar is not null).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_wave(const DevParams* __restrict__ Pd, float dt,
                                                  const float* __restrict__ H0,
                                                  const float* __restrict__ H1,
                                                  const float* __restrict__ colflux,
                                                  const float* __restrict__ G0,
                                                  const float* __restrict__ G1,
                                                  float* __restrict__ G0n,
                                                  float* __restrict__ G1n,
                                                  double* __restrict__ acc, int nby,
                                                  float* __restrict__ Sbar) {
    const DevParams& P = *Pd;
    if ((int)blockIdx.y >= nby) {  // column-flux adjoint blocks
        const int tid = threadIdx.y * blockDim.x + threadIdx.x;
        const int b = ((int)blockIdx.y - nby) * (int)gridDim.x + (int)blockIdx.x;
        adj_colflux_column(P, dt, G0, Sbar, (b * 256 + tid) >> 6, tid & 63);
        return;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int W = P.W, H = P.H;
    float part[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (i < W && j < H) {
        const size_t s = (size_t)j * W + i;
        const float g = G0[s];
        // transpose gather: texel t contributes lambda * G0[t] to the texel
        // it reads as its east / west / north / south neighbour (clamped)
        float ew = 0.0f, ns = 0.0f;
        if (i >= 1) ew += G0[s - 1];
        if (i == W - 1) ew += g;
        if (i <= W - 2) ew += G0[s + 1];
        if (i == 0) ew += g;
        if (j >= 1) ns += G0[s - W];
        if (j == H - 1) ns += g;
        if (j <= H - 2) ns += G0[s + W];
        if (j == 0) ns += g;
        G0n[s] = G1[s] + P.atten * (P.wc0 * g + P.lx * ew + P.lz * ns);
        G1n[s] = -P.atten * P.wc1 * g;
        // parameter partials
        const float C0 = H0[s], C1 = H1[s];
        const float nn = H0[(size_t)min(j + 1, H - 1) * W + i];
        const float sv = H0[(size_t)max(j - 1, 0) * W + i];
        const float ee = H0[(size_t)j * W + min(i + 1, W - 1)];
        const float wv = H0[(size_t)j * W + max(i - 1, 0)];
        const float L = P.wc0 * C0 + P.lx * (ee + wv) + P.lz * (nn + sv) - P.wc1 * C1;
        part[0] = g * P.atten * (ee + wv - 2.0f * C0);  // lambda_x
        part[1] = g * P.atten * (nn + sv - 2.0f * C0);  // lambda_z
        part[2] = g * L;                                // atten
        part[3] = g * P.atten * (C1 - C0);              // beta
        if (colflux) {
            const float gx = P.ax * (float)i + P.bx, gz = P.az * (float)j + P.bz;
            const float S = bilerp(colflux, P.dims[0], P.dims[2], gx, gz);
            part[4] = g * dt * S * P.inv_area;          // kappa
        }
    }
    block_accumulate<5>(part, acc);  // lambda_x, lambda_z, atten, beta, kappa
}

// Sum the per-block parameter partials (np slots per block, nb blocks) in
// block order and add them to the gradient slots slot[0..np).
struct SlotMap {
    int s[8];
};
__global__ void k_reduce_partials(const double* __restrict__ acc, int nb, int np, SlotMap map,
                                  double* __restrict__ pgrad) {
    const int k = threadIdx.x;
    if (k >= np) return;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += acc[(size_t)b * np + k];
    pgrad[map.s[k]] += t;
}

// ---------------------------------------------------------------------------
// Per-particle adjoint of integrate + the force terms local to the particle.
// xb, vb: adjoints of the step outputs (sorted order of this step).
// Writes F-bar and V-bar (the viscosity sum's adjoint, w = crest factor),
// the direct rho-bar, and the local parts of x-bar / v-bar.  Heightfield
// taps go through the block's LDS window.
// ---------------------------------------------------------------------------
// The per-particle body of k_adj_local; texture taps go through `sc`.
__device__ __forceinline__ void adj_local_particle(
    const DevParams& P, float dt, int i, const float4* __restrict__ pos,
    const float4* __restrict__ vel, const float4* __restrict__ fpre,
    const float* __restrict__ rho_w, const float* __restrict__ H0,
    const float4* __restrict__ xb_in, const float4* __restrict__ vb_in,
    const TexScatterWindow& sc, float4* __restrict__ adjrec, float* __restrict__ rhob_d,
    float4* __restrict__ xb_loc, float4* __restrict__ vb_loc) {
    const float4 x = pos[i], v = vel[i], F = fpre[i];
    const float rho = rho_w[i];
    // --- forward recompute of integrate (branch decisions) ---
    const float ax = F.x / rho, ay = F.y / rho, az = F.z / rho;
    const float nvx = v.x + dt * ax, nvy = v.y + dt * ay, nvz = v.z + dt * az;
    const float npx = x.x + dt * nvx, npy = x.y + dt * nvy, npz = x.z + dt * nvz;
    const float damp = 1.0f - P.vdamp * dt;
    const float dvx = nvx * damp, dvy = nvy * damp, dvz = nvz * damp;
    const bool foam = sqrtf(dist2_fma(dvx, dvy, dvz)) > P.foam_speed;
    const float tu = (npx - P.org[0]) * P.uvs[0], tv = (npz - P.org[1]) * P.uvs[1];
    const float th = tex_sample(H0, P.W, P.H, tu, tv);
    const bool sclamp = npy < th;
    const float py2 = sclamp ? th - P.rp : npy;
    const float p3[3] = {npx, py2, npz};
    bool cl[3];
    for (int a = 0; a < 3; ++a) cl[a] = p3[a] < P.lo[a] || p3[a] > P.up[a];
    float yafter = py2;
    if (yafter < P.lo[1]) yafter = P.lo[1];
    else if (yafter > P.up[1]) yafter = P.up[1];
    const bool ceil_hit = P.wake && yafter > P.up[1] + P.ceil_extra;

    // --- reverse ---
    const float4 xo = xb_in[i], vo = vb_in[i];
    float xbx = xo.x, xby = xo.y, xbz = xo.z;
    float vbx = vo.x, vby = vo.y, vbz = vo.z;
    if (ceil_hit) {
        vby *= -P.bounce;
        xby = 0.0f;
    }
    if (cl[0]) { vbx *= -P.bounce; xbx = 0.0f; }
    if (cl[1]) { vby *= -P.bounce; xby = 0.0f; }
    if (cl[2]) { vbz *= -P.bounce; xbz = 0.0f; }
    if (sclamp) {
        const float thb = xby;
        xby = 0.0f;
        float gu, gv;
        tex_adj(H0, &sc, P.W, P.H, tu, tv, thb, gu, gv);
        xbx += gu * P.uvs[0];
        xbz += gv * P.uvs[1];
    }
    if (foam) {
        vbx *= P.foam_v;
        vby *= P.foam_v;
        vbz *= P.foam_v;
    }
    // nvd = nv * damp ; np = x + dt nv ; nv = v + dt a ; a = F / rho
    float nvbx = vbx * damp + dt * xbx, nvby = vby * damp + dt * xby, nvbz = vbz * damp + dt * xbz;
    float lxbx = xbx, lxby = xby, lxbz = xbz;
    float lvbx = nvbx, lvby = nvby, lvbz = nvbz;
    const float abx = dt * nvbx, aby = dt * nvby, abz = dt * nvbz;
    const float Fbx = abx / rho, Fby = aby / rho, Fbz = abz / rho;
    float rhob = -(abx * F.x + aby * F.y + abz * F.z) / (rho * rho);
    // gravity: F += rho G
    rhob += Fbx * P.g[0] + Fby * P.g[1] + Fbz * P.g[2];
    // drag and wave force (inside the coupling band)
    const float u = (x.x - P.org[0]) * P.uvs[0], vv = (x.z - P.org[1]) * P.uvs[1];
    const float Hs = tex_sample(H0, P.W, P.H, u, vv);
    if (coupled_(P, x.y, Hs)) {
        lvbx += -P.drag * Fbx;
        lvby += -P.drag * Fby;
        lvbz += -P.drag * Fbz;
        const float wvbx = P.drag * Fbx, wvby = P.drag * Fby, wvbz = P.drag * Fbz;
        const float hxb = wvbx / dt + wvbz / P.du;
        const float hyb = wvby / dt - wvbz / P.du;
        float hb = -wvbx / dt - wvby / dt;
        const float h1x = tex_sample(H0, P.W, P.H, u + 1.0f, vv);
        const float h1y = tex_sample(H0, P.W, P.H, u, vv + 1.0f);
        const float nx = h1x - Hs, ny = h1y - Hs;
        hb += -P.wforce * (Fbx * nx + Fby * ny - Fbz);
        const float nxb = -P.wforce * Hs * Fbx, nyb = -P.wforce * Hs * Fby;
        hb -= nxb + nyb;
        float gu, gv, su = 0.0f, sv = 0.0f;
        tex_adj(H0, &sc, P.W, P.H, u + P.du, vv, hxb, gu, gv);
        su += gu;
        sv += gv;
        tex_adj(H0, &sc, P.W, P.H, u, vv + P.du, hyb, gu, gv);
        su += gu;
        sv += gv;
        tex_adj(H0, &sc, P.W, P.H, u + 1.0f, vv, nxb, gu, gv);
        su += gu;
        sv += gv;
        tex_adj(H0, &sc, P.W, P.H, u, vv + 1.0f, nyb, gu, gv);
        su += gu;
        sv += gv;
        tex_adj(H0, &sc, P.W, P.H, u, vv, hb, gu, gv);
        su += gu;
        sv += gv;
        lxbx += su * P.uvs[0];
        lxbz += sv * P.uvs[1];
    }
    const float crest = x.y > P.crest ? P.brk_visc : 1.0f;
    const float vs = P.visc * crest;
    adjrec[2 * (size_t)i] = make_float4(Fbx, Fby, Fbz, 0.0f);
    adjrec[2 * (size_t)i + 1] = make_float4(vs * Fbx, vs * Fby, vs * Fbz, crest);
    rhob_d[i] = rhob;
    xb_loc[i] = make_float4(lxbx, lxby, lxbz, 0.0f);
    vb_loc[i] = make_float4(lvbx, lvby, lvbz, 0.0f);
}

__global__ void __launch_bounds__(256) k_adj_local(const DevParams* __restrict__ Pd, float dt, int n,
                                                   const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel,
                                                   const float4* __restrict__ fpre,
                                                   const float* __restrict__ rho_w,
                                                   const float* __restrict__ H0,
                                                   const float4* __restrict__ xb_in,
                                                   const float4* __restrict__ vb_in,
                                                   float* __restrict__ G0n,
                                                   float4* __restrict__ adjrec,
                                                   float* __restrict__ rhob_d,
                                                   float4* __restrict__ xb_loc,
                                                   float4* __restrict__ vb_loc) {
    const DevParams& P = *Pd;
    __shared__ float s_win[kTexWinLds];
    __shared__ int s_org[2];
    const TexScatterWindow sc = tex_window_begin(P, pos, G0n, s_win, s_org);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) adj_local_particle(P, dt, i, pos, vel, fpre, rho_w, H0, xb_in, vb_in, sc, adjrec,
                                  rhob_d, xb_loc, vb_loc);
    tex_window_end(P, sc);
}

// ---------------------------------------------------------------------------
// Pair adjoint of the pressure and viscosity sums, then the density-side
// adjoint of particle k (EOS, rest-density clamp, wave term, flux).
// G lanes per particle (consecutive lanes of a wave): each walks every G-th
// neighbour, the group's sums are added in a fixed tree and its first lane
// finishes the particle.  G = 4 for scenes of up to lane_split_limit()
// particles (taken when a tape is created), which would otherwise keep one
// wave per SIMD.
// ---------------------------------------------------------------------------
constexpr int kAdjSplit = 4;

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int G>
__global__ void __launch_bounds__(256) k_adj_pairs(
    const DevParams* __restrict__ Pd, float dt, int n, const float4* __restrict__ pos,
    const float4* __restrict__ vel, const int* __restrict__ offset,
    const uint2* __restrict__ nbw, const float4* __restrict__ nrec,
    const float* __restrict__ rho_w, const float2* __restrict__ pw,
    const float* __restrict__ rho_raw, const float* __restrict__ H0,
    const float* __restrict__ Sbar, const float4* __restrict__ adjrec,
    const float* __restrict__ rhob_d, float* __restrict__ G0n, float4* __restrict__ xb_loc,
    float4* __restrict__ vb_loc, float* __restrict__ rtb, double* __restrict__ acc) {
    const DevParams& P = *Pd;
    __shared__ uint2 s_nb[kNbStage][256];
    const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) / G);
    const int sub = (int)threadIdx.x % G;
    float part[3] = {0.f, 0.f, 0.f};  // visc, gas_const, rest_rho
    float pkb = 0.f, irkb = 0.f;
    float xbx = 0.f, xby = 0.f, xbz = 0.f, vbx = 0.f, vby = 0.f, vbz = 0.f;
    float Vx = 0.f, Vy = 0.f, Vz = 0.f;  // viscosity sum of k (for d/d visc)
    if (k < n) {  // whole groups: the lanes of a group share k
        const float4 xk = pos[k], vk = vel[k];
        const float2 qk = pw[k];
        const float pk = qk.x, irk = qk.y;
        const float4 Fb = adjrec[2 * (size_t)k], Vb = adjrec[2 * (size_t)k + 1];
        const float msh = P.m * P.spiky * 0.5f, mlap = P.m * P.lap;
        const auto load = [&](int j) {
            return AdjNbr{nrec[2 * (size_t)j], nrec[2 * (size_t)j + 1], adjrec[2 * (size_t)j],
                          adjrec[2 * (size_t)j + 1]};
        };
        for_each_neighbour_of_lane<2, G>(P, offset, nbw, pos, n, k, xk, &s_nb[0][threadIdx.x],
                                         sub, load, [&](const AdjNbr& q) {
            const float4 ra = q.ra, rb = q.rb, Fbj = q.Fbj, Vbj = q.Vbj;
            const float pj = ra.w, irj = rb.w;
            const float dx = xk.x - ra.x, dy = xk.y - ra.y, dz = xk.z - ra.z;
            const float r2 = dist2_fma(dx, dy, dz);
            if (!(r2 > 0.0f)) return;  // coincident pair: no term in the forward
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            const float hr = P.h - r;
            const float phi = hr * hr * rinv;                   // (h-r)^2 / r
            const float dphi = -hr * (P.h + r) * rinv * rinv;   // d phi / dr
            // pressure, own term: c = -msh (pk + pj) irj phi (Fb . d)
            const float Fd = Fb.x * dx + Fb.y * dy + Fb.z * dz;
            pkb += -msh * irj * phi * Fd;
            const float co = -msh * (pk + pj) * irj;
            float gx = co * (phi * Fb.x + Fd * dphi * dx * rinv);
            float gy = co * (phi * Fb.y + Fd * dphi * dy * rinv);
            float gz = co * (phi * Fb.z + Fd * dphi * dz * rinv);
            // pressure, j's term with k as neighbour: d_jk = -d
            const float Fdj = -(Fbj.x * dx + Fbj.y * dy + Fbj.z * dz);
            pkb += -msh * irk * phi * Fdj;
            irkb += -msh * (pj + pk) * phi * Fdj;
            const float cj = -msh * (pj + pk) * irk;
            gx -= cj * (phi * Fbj.x - Fdj * dphi * dx * rinv);
            gy -= cj * (phi * Fbj.y - Fdj * dphi * dy * rinv);
            gz -= cj * (phi * Fbj.z - Fdj * dphi * dz * rinv);
            // viscosity, own term: e = mlap irj hr (Vb . (vj - vk))
            const float dvx = rb.x - vk.x, dvy = rb.y - vk.y, dvz = rb.z - vk.z;
            const float Vd = Vb.x * dvx + Vb.y * dvy + Vb.z * dvz;
            const float wo = mlap * irj * hr;
            vbx -= wo * Vb.x;
            vby -= wo * Vb.y;
            vbz -= wo * Vb.z;
            Vx += wo * dvx;
            Vy += wo * dvy;
            Vz += wo * dvz;
            const float eo = -mlap * irj * Vd * rinv;  // de/dr * (1/r)
            gx += eo * dx;
            gy += eo * dy;
            gz += eo * dz;
            // viscosity, j's term: e = mlap irk hr (Vbj . (vk - vj)), d_jk = -d
            const float Vdj = -(Vbj.x * dvx + Vbj.y * dvy + Vbj.z * dvz);
            const float wj = mlap * irk * hr;
            vbx += wj * Vbj.x;
            vby += wj * Vbj.y;
            vbz += wj * Vbj.z;
            irkb += mlap * hr * Vdj;
            const float ej = -mlap * irk * Vdj * rinv;
            gx += ej * dx;
            gy += ej * dy;
            gz += ej * dz;
            xbx += gx;
            xby += gy;
            xbz += gz;
        });
        if (G > 1) {
            pkb = group_sum<G>(pkb);
            irkb = group_sum<G>(irkb);
            xbx = group_sum<G>(xbx);
            xby = group_sum<G>(xby);
            xbz = group_sum<G>(xbz);
            vbx = group_sum<G>(vbx);
            vby = group_sum<G>(vby);
            vbz = group_sum<G>(vbz);
            Vx = group_sum<G>(Vx);
            Vy = group_sum<G>(Vy);
            Vz = group_sum<G>(Vz);
        }
    }
    if (k < n && sub == 0) {
        const float4 xk = pos[k], vk = vel[k];
        const float irk = pw[k].y;
        const float4 Fb = adjrec[2 * (size_t)k], Vb = adjrec[2 * (size_t)k + 1];
        part[0] = Vb.w * (Fb.x * Vx + Fb.y * Vy + Fb.z * Vz);
        // ---- density side ----
        const float rhob = rhob_d[k] - irkb * irk * irk;
        const float rt = rho_raw[k];
        const float u = (xk.x - P.org[0]) * P.uvs[0], vv = (xk.z - P.org[1]) * P.uvs[1];
        const float Hs = tex_sample(H0, P.W, P.H, u, vv);
        const bool g = coupled_(P, xk.y, Hs);
        const bool a_br = P.k * (rt - P.rho0) > 0.0f;
        const float rhat = g ? rt + (Hs * rt) / P.k_rp : rt;
        const bool b_br = rhat > P.rho0;
        const float rhatb = b_br ? rhob : 0.0f;
        float r0b = b_br ? 0.0f : rhob;
        float rtb_k = (a_br ? P.k * pkb : 0.0f) + (g ? Hs * pkb : 0.0f) +
                      rhatb * (g ? 1.0f + Hs / P.k_rp : 1.0f);
        float kb = a_br ? (rt - P.rho0) * pkb : 0.0f;
        if (g) kb += rhatb * rt * Hs * (-P.rp / (P.k_rp * P.k_rp));
        r0b += a_br ? -P.k * pkb : 0.0f;
        float Hb = g ? rt * pkb + rhatb * rt / P.k_rp : 0.0f;
        if (P.band > 0.0f) {
            const float d = fabsf(xk.y - Hs);
            if (d < P.band) {
                const int ix = cell_axis(xk.x, P.gmin[0], P.cell, P.dims[0]);
                const int iz = cell_axis(xk.z, P.gmin[2], P.cell, P.dims[2]);
                const float sb = Sbar[iz * P.dims[0] + ix];
                const float q = d / P.band;
                const float om = 1.0f - q * q;
                const float fl = P.m_rho0 * vk.y * om * om;
                vby += sb * P.m_rho0 * om * om;
                r0b += sb * (-fl / P.rho0);
                const float qb = sb * P.m_rho0 * vk.y * 2.0f * om * (-2.0f * q);
                const float db = qb / P.band;
                const float sg = (xk.y - Hs) >= 0.0f ? 1.0f : -1.0f;
                xby += sg * db;
                Hb -= sg * db;
            }
        }
        if (Hb != 0.0f) {
            float gu, gv;
            const TexScatterGlobal sc{G0n, P.W};
            tex_adj(H0, &sc, P.W, P.H, u, vv, Hb, gu, gv);
            xbx += gu * P.uvs[0];
            xbz += gv * P.uvs[1];
        }
        part[1] = kb;
        part[2] = r0b;
        rtb[k] = rtb_k;
        const float4 xl = xb_loc[k], vl = vb_loc[k];
        xb_loc[k] = make_float4(xl.x + xbx, xl.y + xby, xl.z + xbz, 0.0f);
        vb_loc[k] = make_float4(vl.x + vbx, vl.y + vby, vl.z + vbz, 0.0f);
    }
    block_accumulate<3>(part, acc);  // visc, gas_const, rest_rho
}

// ---------------------------------------------------------------------------
// Pair adjoint of the poly6 density sum, rho~_k = m c6 sum_j w_kj^3 with
// w = T - r^2:  x-bar_k += sum_j m c6 3 w^2 (-2 d_kj) (rt-bar_k + rt-bar_j).
// Then the adjoint state moves back to the pre-sort slots: slot list[k]
// (the forward reorder sent slot list[k] to sorted position k).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_rho_pairs(
    const DevParams* __restrict__ Pd, int n, const float4* __restrict__ pos, const int* __restrict__ offset,
    const uint2* __restrict__ nbw, const float* __restrict__ rtb,
    const float4* __restrict__ xb_loc, const float4* __restrict__ vb_loc,
    const int* __restrict__ list, float4* __restrict__ xb_out, float4* __restrict__ vb_out) {
    const DevParams& P = *Pd;
    __shared__ uint2 s_nb[kNbStage][256];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 xk = pos[k];
    const float bk = rtb[k];
    const float c = P.m_poly6 * 3.0f * -2.0f;
    float4 xb = xb_loc[k];
    const auto load = [&](int j) { return make_float4(pos[j].x, pos[j].y, pos[j].z, rtb[j]); };
    for_each_neighbour<4>(P, offset, nbw, pos, n, k, xk, &s_nb[0][threadIdx.x], load,
                          [&](const float4& pj) {
        const float dx = xk.x - pj.x, dy = xk.y - pj.y, dz = xk.z - pj.z;
        const float r2 = dist2_fma(dx, dy, dz);
        const float w = fmaxf(P.r2_lt - r2, 0.0f);
        const float s = c * w * w * (bk + pj.w);
        xb.x += s * dx;
        xb.y += s * dy;
        xb.z += s * dz;
    });
    const int dst = list[k];
    xb_out[dst] = xb;
    vb_out[dst] = vb_loc[k];
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

int adj_wave_blocks(const DevParams& P) { return ((P.W + 63) / 64) * ((P.H + 3) / 4); }
int adj_pair_group(int n) { return n <= lane_split_limit() ? kAdjSplit : 1; }
int adj_pair_blocks(int n, int group) { return nblk((long long)n * group, 256); }

// Loss seed of a heightfield fit: L = mean((h - t)^2), dL/dh = 2 (h - t) / N.
// Per-block partial sums (double, fixed order: grid-stride per thread, then
// an LDS tree), summed by one thread in block order.
__global__ void __launch_bounds__(256) k_mse_partials(const float* __restrict__ h,
                                                      const float* __restrict__ t, int N,
                                                      float* __restrict__ G,
                                                      double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    const double scale = 2.0 / (double)N;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < N; k += gridDim.x * 256) {
        const double d = (double)h[k] - (double)t[k];  // exact for two floats
        G[k] = (float)(scale * d);
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_mse_total(const double* __restrict__ part, int nb, int N,
                            double* __restrict__ loss) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    *loss = s / (double)N;
}

hipError_t launch_mse_seed(const float* h, const float* target, int N, float* G, double* part,
                           int nblocks, double* loss, hipStream_t st) {
    if (N <= 0 || nblocks <= 0) return hipErrorInvalidValue;
    k_mse_partials<<<nblocks, 256, 0, st>>>(h, target, N, G, part);
    k_mse_total<<<1, 64, 0, st>>>(part, nblocks, N, loss);
    return hipGetLastError();
}

hipError_t launch_adj_wave(const DevParams& P, float dt, const float* H0, const float* H1,
                           const float* colflux, const float* G0, const float* G1, float* G0n,
                           float* G1n, double* acc, float* Sbar, hipStream_t st) {
    const int gx = (P.W + 63) / 64, nby = (P.H + 3) / 4;
    int rows = 0;
    if (Sbar) {
        const long long ncb = nblk((long long)P.dims[0] * P.dims[2] * 64, 256);
        rows = (int)((ncb + gx - 1) / gx);
    }
    dim3 blk(64, 4), grd(gx, nby + rows);
    k_adj_wave<<<grd, blk, 0, st>>>(P.dev, dt, H0, H1, colflux, G0, G1, G0n, G1n, acc, nby, Sbar);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* acc, int nb, int np, const int* slots,
                                  double* pgrad, hipStream_t st) {
    SlotMap m{};
    for (int k = 0; k < np && k < 8; ++k) m.s[k] = slots[k];
    k_reduce_partials<<<1, 64, 0, st>>>(acc, nb, np, m, pgrad);
    return hipGetLastError();
}

hipError_t launch_adj_local(const DevParams& P, float dt, int n, const float4* pos,
                            const float4* vel, const float4* fpre, const float* rho_w,
                            const float* H0, const float4* xb_in, const float4* vb_in,
                            float* G0n, float4* adjrec, float* rhob_d, float4* xb_loc,
                            float4* vb_loc, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_adj_local<<<nblk(n, 256), 256, 0, st>>>(P.dev, dt, n, pos, vel, fpre, rho_w, H0, xb_in, vb_in,
                                              G0n, adjrec, rhob_d, xb_loc, vb_loc);
    return hipGetLastError();
}

hipError_t launch_adj_pairs(const DevParams& P, float dt, int n, const float4* pos,
                            const float4* vel, const int* offset, const uint2* nbw,
                            const float4* nrec, const float* rho_w, const float2* pw,
                            const float* rho_raw, const float* H0, const float* Sbar,
                            const float4* adjrec, const float* rhob_d, float* G0n,
                            float4* xb_loc, float4* vb_loc, float* rtb, double* acc,
                            int group, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (group != 1 && group != kAdjSplit) return hipErrorInvalidValue;
    if (group == kAdjSplit)
        k_adj_pairs<kAdjSplit><<<adj_pair_blocks(n, group), 256, 0, st>>>(
            P.dev, dt, n, pos, vel, offset, nbw, nrec, rho_w, pw, rho_raw, H0, Sbar, adjrec,
            rhob_d, G0n, xb_loc, vb_loc, rtb, acc);
    else
        k_adj_pairs<1><<<adj_pair_blocks(n, group), 256, 0, st>>>(
            P.dev, dt, n, pos, vel, offset, nbw, nrec, rho_w, pw, rho_raw, H0, Sbar, adjrec,
            rhob_d, G0n, xb_loc, vb_loc, rtb, acc);
    return hipGetLastError();
}

hipError_t launch_adj_rho_pairs(const DevParams& P, int n, const float4* pos, const int* offset,
                                const uint2* nbw, const float* rtb, const float4* xb_loc,
                                const float4* vb_loc, const int* list, float4* xb_out,
                                float4* vb_out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_adj_rho_pairs<<<nblk(n, 256), 256, 0, st>>>(P.dev, n, pos, offset, nbw, rtb, xb_loc, vb_loc,
                                                  list, xb_out, vb_out);
    return hipGetLastError();
}

}  // namespace cwa
