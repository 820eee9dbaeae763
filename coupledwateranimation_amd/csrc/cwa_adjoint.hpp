// cwa_adjoint.hpp — adjoint kernels of the coupled step (see cwa_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "cwa_device.hpp"

namespace cwa {

// Slots of the fp64 parameter-gradient accumulator (cwa_grad order).
enum GradSlot {
    GRAD_VISC = 0,
    GRAD_GAS_CONST,
    GRAD_REST_RHO,
    GRAD_LAMBDA_X,
    GRAD_LAMBDA_Z,
    GRAD_ATTEN,
    GRAD_KAPPA,
    GRAD_BETA,
    GRAD_NUM
};

// Parameter partials are accumulated per block (fp64, no atomics) into acc
// arrays of blocks x slots; launch_reduce_partials sums them in block order.
int adj_wave_blocks(const DevParams& P);  // blocks of launch_adj_wave (5 slots each)
int adj_pair_group(int n);                // lanes per particle of the pair adjoint now
int adj_pair_blocks(int n, int group);    // blocks of launch_adj_pairs (3 slots each)
// MSE of a heightfield against a target: seed G = 2 (h - t)/N, loss = mean.
hipError_t launch_mse_seed(const float* h, const float* target, int N, float* G, double* part,
                           int nblocks, double* loss, hipStream_t st);
hipError_t launch_reduce_partials(const double* acc, int nb, int np, const int* slots,
                                  double* pgrad, hipStream_t st);
// Sbar (optional): the same launch also writes the column-flux adjoint.
hipError_t launch_adj_wave(const DevParams& P, float dt, const float* H0, const float* H1,
                           const float* colflux, const float* G0, const float* G1, float* G0n,
                           float* G1n, double* acc, float* Sbar, hipStream_t st);
hipError_t launch_adj_local(const DevParams& P, float dt, int n, const float4* pos,
                            const float4* vel, const float4* fpre, const float* rho_w,
                            const float* H0, const float4* xb_in, const float4* vb_in,
                            float* G0n, float4* adjrec, float* rhob_d, float4* xb_loc,
                            float4* vb_loc, hipStream_t st);
hipError_t launch_adj_pairs(const DevParams& P, float dt, int n, const float4* pos,
                            const float4* vel, const int* offset, const uint2* nbw,
                            const float4* nrec, const float* rho_w, const float2* pw,
                            const float* rho_raw, const float* H0, const float* Sbar,
                            const float4* adjrec, const float* rhob_d, float* G0n,
                            float4* xb_loc, float4* vb_loc, float* rtb, double* acc,
                            int group, hipStream_t st);
hipError_t launch_adj_rho_pairs(const DevParams& P, int n, const float4* pos, const int* offset,
                                const uint2* nbw, const float* rtb, const float4* xb_loc,
                                const float4* vb_loc, const int* list, float4* xb_out,
                                float4* vb_out, hipStream_t st);

}  // namespace cwa
