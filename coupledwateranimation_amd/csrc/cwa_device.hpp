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
    float cell;
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

__device__ __forceinline__ int lin_cell(const int dims[3], int ix, int iy, int iz) {
    return (ix * dims[2] + iz) * dims[1] + iy;
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
    return bilerp(img, w, h, u * (float)w - 0.5f, v * (float)h - 0.5f);
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
