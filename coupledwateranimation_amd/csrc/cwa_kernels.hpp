// cwa_kernels.hpp — launch wrappers of the coupled-step kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "cwa_device.hpp"

namespace cwa {

struct GridDesc {
    float gmin[3];
    float cell;    // x and z cell edge
    float cell_y;  // y cell edge
    int dims[3];
};

// One copy of the particle state (SoA, device).
struct ParticleArrays {
    float4* pos = nullptr;
    float4* vel = nullptr;
    float4* force = nullptr;
    int* id = nullptr;
};

size_t scan_workspace_ints(int n);
// zero_out (optional, n ints) is cleared in the same pass.
hipError_t launch_scan(const int* in, int* out, int n, int* ws, hipStream_t st,
                       int* zero_out = nullptr);
hipError_t launch_fill_i32(int* p, int n, int v, hipStream_t st);
// Stable counting sort of keyed items (rec: n ints, runs: n int2 scratch).
hipError_t launch_hash_count(const float4* pos, int n, const GridDesc& g, int* key, int* rec,
                             int* count, hipStream_t st);
hipError_t launch_key_count(const int* key, int n, int* rec, int* count, hipStream_t st);
hipError_t launch_scatter_runs(const int* key, const int* rec, const int* offset, int n,
                               int2* runs, hipStream_t st);
// list[d] = item_of ? item_of[s] : s at each item's stable destination d;
// in/out may be null: only the list is produced.
hipError_t launch_stable_reorder(const int* key, const int* rec, const int* offset,
                                 const int2* runs, int n, const int* item_of, int* list,
                                 const ParticleArrays* in, ParticleArrays* out, hipStream_t st);
// Boxes: cells covered per box (ncov, n ints), then the (box, cell) pairs at
// first[i] = exclusive scan of ncov (pkey: cell, pbox: box index).
hipError_t launch_box_cover(const float4* x, const float4* hw, int n, const float gmin[3],
                            const float cell[3], const int dims[3], int* ncov, hipStream_t st);
// *out = sum of a[0..n) in 64 bits (non-negative values).
hipError_t launch_sum_i64(const int* a, int n, unsigned long long* out, hipStream_t st);
hipError_t launch_box_pairs(const float4* x, const float4* hw, int n, const float gmin[3],
                            const float cell[3], const int dims[3], const int* first,
                            int* pkey, int* pbox, hipStream_t st);
// nbw (nb_buffer_uint2(n) uint2: neighbour words and counts) and nrec (2*n
// float4) are written by the density pass
// and read by the force pass.  rho_raw (density before the wave term and the
// rest-density clamp) and fpre (force before the foam branch) are optional
// outputs for the adjoint tape (null: not written).
hipError_t launch_density(const DevParams& P, const ParticleArrays& B, const int* offset, int n,
                          const float* H0, float* rho_w, float2* pw, float* flux, uint2* nbw,
                          float4* nrec, float* rho_raw, hipStream_t st);
hipError_t launch_force_integrate(const DevParams& P, float dt, const ParticleArrays& B,
                                  const int* offset, int n, const float* rho_w, const float2* pw,
                                  const uint2* nbw, const float4* nrec, const float* H0,
                                  ParticleArrays& A, float* rho, float* prs, float4* fpre,
                                  const float* flux, float* colflux, hipStream_t st);
// Neighbour counts (self included) of the last density pass, per sorted slot.
hipError_t launch_neighbour_counts(const DevParams& P, const float4* pos, const int* offset,
                                   const uint2* nbw, int n, int* out, hipStream_t st);
hipError_t launch_wave(const DevParams& P, const float* h0, const float* h1, float* out,
                       const float* colflux, float dt, hipStream_t st);
hipError_t launch_init_wave(float* out, int W, int H, float type, hipStream_t st);
hipError_t launch_check_finite(const float* a, size_t n, int* flag, hipStream_t st);

}  // namespace cwa
