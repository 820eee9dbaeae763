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
// Stable counting sort of keyed items: count (keys of particle positions, or
// given keys) into a zeroed counter, scan, scatter of run records (runs: n
// int2; fill: num_cells ints, zero on entry), reorder (clears fill[0 ..
// nfill) for the next build).
hipError_t launch_hash_count(const float4* pos, int n, const GridDesc& g, int* key, int* count,
                             hipStream_t st);
hipError_t launch_key_count(const int* key, int n, int* count, hipStream_t st);
hipError_t launch_scatter_runs(const int* key, const int* offset, int* fill, int n, int2* runs,
                               hipStream_t st);
// list[d] = item_of ? item_of[s] : s at each item's stable destination d;
// with pos_out set, pos_out[d] = pos_in[s] and id_out[d] = id_in[s].
hipError_t launch_stable_reorder(const int* key, const int* offset, const int2* runs, int n,
                                 const int* item_of, int* list, const float4* pos_in,
                                 const int* id_in, float4* pos_out, int* id_out, int* fill,
                                 int nfill, hipStream_t st);
// Boxes: cells covered per box (ncov, n ints), then the (box, cell) pairs at
// first[i] = exclusive scan of ncov (pkey: cell, pbox: box index).
hipError_t launch_box_cover(const float4* x, const float4* hw, int n, const float gmin[3],
                            const float cell[3], const int dims[3], int* ncov, hipStream_t st);
// *out = sum of a[0..n) in 64 bits (non-negative values).
hipError_t launch_sum_i64(const int* a, int n, unsigned long long* out, hipStream_t st);
hipError_t launch_box_pairs(const float4* x, const float4* hw, int n, const float gmin[3],
                            const float cell[3], const int dims[3], const int* first,
                            int* pkey, int* pbox, hipStream_t st);
// The density pass: sorted positions bpos, the state's velocities read in
// their pre-sort slots state_vel[list[i]].  nbw (nb_buffer_uint2(n) uint2:
// neighbour words and counts) and nrec (2*n float4: position and pressure,
// velocity and 1/rho) and vpad (the velocity's fourth component, carried
// through the step) are written for the force pass.  pw (pressure, 1/rho),
// rho_raw (density before the wave term and the rest-density clamp) and
// vel_out (the sorted velocities) are optional outputs for the adjoint tape
// (null: not written).
// Scenes of up to this many particles run the density pass (and a tape's
// pair adjoint) with four lanes per particle, larger ones with one lane;
// read at each density launch and at tape creation.  Default kLaneSplitDefault.
constexpr long long kLaneSplitDefault = 1 << 18;
long long lane_split_limit();
long long set_lane_split_limit(long long max_n);  // returns the previous limit

hipError_t launch_density(const DevParams& P, const float4* bpos, const float4* state_vel,
                          const int* list, const int* offset, int n, const float* H0,
                          float* rho_w, float2* pw, float* flux, uint2* nbw, float4* nrec,
                          float* rho_raw, float4* vel_out, float* vpad, hipStream_t st);
// Where the force pass hashes the next step's grid (key, count; the counter
// zero on entry), or key == null for no fused hash.
struct NextKeys {
    int* key;
    int* count;
};
// Forces + integration: own record from nrec, id from bid and the velocity
// pad from vpad (sorted), previous force from the state S in its pre-sort
// slot list[i]; writes the next state N in sorted order.  fpre (force before
// the foam branch) is an optional output for the adjoint tape.
hipError_t launch_force_integrate(const DevParams& P, float dt, const float4* bpos,
                                  const int* bid, const ParticleArrays& S, const int* list,
                                  const int* offset, int n, const float* rho_w, const uint2* nbw,
                                  const float4* nrec, const float* H0, ParticleArrays& N,
                                  float* rho, float* prs, float4* fpre, const float* flux,
                                  float* colflux, const float* vpad, const NextKeys& next,
                                  hipStream_t st);
// Neighbour counts (self included) of the last density pass, per sorted slot.
hipError_t launch_neighbour_counts(const DevParams& P, const float4* pos, const int* offset,
                                   const uint2* nbw, int n, int* out, hipStream_t st);
hipError_t launch_wave(const DevParams& P, const float* h0, const float* h1, float* out,
                       const float* colflux, float dt, hipStream_t st);
hipError_t launch_init_wave(float* out, int W, int H, float type, hipStream_t st);
hipError_t launch_check_finite(const float* a, size_t n, int* flag, hipStream_t st);

}  // namespace cwa
