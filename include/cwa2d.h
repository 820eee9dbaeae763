/* cwa2d.h — C ABI of the batched 2D coupled step: many independent scenes,
 * each a 2D SPH fluid (Koschier cubic kernel, WCSPH gamma = 3, box SDF)
 * coupled to its own 1D damped wave column, advanced together by one set
 * of kernel launches per step.
 *
 * Per scene this replaces the SphWave2D application's step:
 *   cwa2d_batch_step     SphUgrid::Compute (SphWave2D/StencilBuffer.cpp:150-179:
 *                        per SPH substep UniformGridSph2D::CollisionQuery
 *                        UniformGridGpu2D.cpp:220-265, then the density and
 *                        force passes of SphWaveKoschier2D_grid_cs.glsl:393-628)
 *                        followed by ImageStencil::Compute
 *                        (SphWave2D/StencilImage2D.cpp:141-164, Wave1D_cs.glsl
 *                        IterateWave :120-130 with FreeBC :139-157)
 *   cwa2d_batch_create   SphUgrid::Init (StencilBuffer.cpp:141-146) and
 *                        ImageStencil::Init/Reinit (StencilImage2D.cpp:88-108)
 *   cwa2d_batch_views    StencilBuffer::GetReadBuffer (StencilBuffer.h:14),
 *                        ImageStencil::GetReadImage
 *   cwa2d_wave_init_reference
 *                        Wave1D_cs.glsl InitWave :96-118 (modes 0 and 1)
 *   wave_model = CWA2D_WAVE_SHALLOW
 *                        Shallow1D_cs.glsl IterateWave :133-167, InitWave :117-131,
 *                        Splash :101-115 (the wave SphWave2D/Main.cpp:283 runs)
 * Status codes and cwa_last_error() are those of cwa.h.
 */
#ifndef CWA2D_H
#define CWA2D_H

#include <stddef.h>
#include <stdint.h>

#include "cwa.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CWA2D_ABI_VERSION 2
#define CWA2D_MAX_WAVE_LEN 8192      /* cells of one wave column (LDS-resident) */
#define CWA2D_MAX_PARTICLES (1 << 26) /* slots of the whole batch */

/* Batch-wide physics.  Shader line numbers refer to
 * SphWave2D/SphWaveKoschier2D_grid_cs.glsl unless noted. */
typedef struct cwa2d_params {
    float h;               /* support radius H (:120), 2^-16 <= H <= 2^19; cell = H */
    float rest_rho;        /* REST_DENS (:122)                                   */
    float gas_const;       /* GAS_CONST (:126); p = k((rho/rho0)^3 - 1) (:303)   */
    float visc;            /* VISC (:123)                                        */
    float psi;             /* boundary particle factor PSI (:136)                */
    float gravity[2];      /* G.xy                                               */
    float lo[2], hi[2];    /* domain: grid extent, wave x range; particles with
                              y > hi[1] keep their density (:428)                */
    float walls[4];        /* SDF planes: floor y, left x, right x, ceiling y
                              (boundary_sdf :205-230)                            */
    float enable_vel;      /* wave-velocity threshold of the enable flag
                              (:400-415, 0.05); 0 keeps every particle enabled   */
    float wave_ref;        /* VIEW_HEIGHT/2 in dh (:590)                         */
    float wave_press;      /* 500000 of the wave pressure (:603)                 */
    float wave_dy;         /* 10 of smoothstep(0, 10, dy) (:603)                 */
    float wave_maskv;      /* 0.2 of smoothstep(0, 0.2, |v|) (:592)              */
    float atten;           /* Wave1D_cs.glsl:19                                  */
    float beta;            /* Wave1D_cs.glsl:20                                  */
    float couple_band;     /* particle -> wave flux band; <= 0: one-way
                              coupling, as the reference                         */
    float couple_strength; /* kappa of the flux source                           */
    int32_t sph_substeps;  /* SPH substeps per step (SphUgrid::mSubsteps)        */
    int32_t check_finite;  /* nonzero: every step checks the state for NaN/Inf   */
} cwa2d_params;

/* Wave models of a scene's 1D column. */
#define CWA2D_WAVE_LEAPFROG 0 /* Wave1D_cs.glsl: levels h0 = h^n, h1 = h^(n-1)       */
#define CWA2D_WAVE_SHALLOW 1  /* Shallow1D_cs.glsl: h0 = water height h, h1 = uh     */

/* One scene of the batch. */
typedef struct cwa2d_scene {
    int32_t n;             /* particles (>= 1)                                   */
    int32_t wave_len;      /* wave cells, 3..CWA2D_MAX_WAVE_LEN                  */
    int32_t wave_substeps; /* ImageStencil::SetSubsteps (>= 1)                   */
    int32_t wave_model;    /* CWA2D_WAVE_LEAPFROG or CWA2D_WAVE_SHALLOW          */
    float mass;            /* MASS (:124)                                        */
    float lambda;          /* leapfrog: (c dt_sub / dx)^2 (Wave1D_cs.glsl:18);
                              shallow: the Lax-Wendroff time step
                              (Shallow1D_cs.glsl:22)                              */
    float wave_dx;         /* shallow: the scheme's cell size dx
                              (Shallow1D_cs.glsl:23); unused by the leapfrog     */
} cwa2d_scene;

/* Launch configuration (the autotuner's output). */
typedef struct cwa2d_tuning {
    int32_t block_density; /* threads per block = particles per tile: 64..1024  */
    int32_t block_force;
    int32_t grid_density;  /* persistent blocks; 0: one block per tile          */
    int32_t grid_force;
    int32_t xcd_queues;    /* 1: scenes bound to XCD-local tile queues (with
                              stealing); 0: one global queue                    */
    int32_t wave_block;    /* threads per scene in the wave pass: 64..1024      */
    float ms_per_step;     /* tuner: measured time of this setting              */
} cwa2d_tuning;

/* Device pointers into the batch (valid until the next step/reinit/destroy).
 * Particle arrays are in slot order: scene s owns slots
 * [scene_begin[s], scene_begin[s] + n_s), cell-sorted within the scene. */
typedef struct cwa2d_views {
    int32_t nscenes;
    int32_t n_total;       /* particles in the batch                            */
    int32_t wave_total;    /* wave cells in the batch                           */
    int32_t grid_nx, grid_ny; /* cells per scene axis (x outer, y fastest)      */
    int32_t cells_total;   /* nscenes * grid_nx * grid_ny                       */
    float* xv;             /* float4 (x, y, vx, vy) per slot                    */
    float* rp;             /* float2 (rho, p) per slot                          */
    int32_t* idf;          /* scene-local particle id | enabled << 30           */
    int32_t* key;          /* global cell of each slot (last sort)              */
    int32_t* offset;       /* cells_total + 1 exclusive offsets (last sort)     */
    float* h0;             /* newest wave level, scenes concatenated            */
    float* h1;             /* previous wave level                               */
    const int32_t* scene_begin; /* device: first slot of each scene             */
    const int32_t* wave_begin;  /* device: first wave cell of each scene        */
} cwa2d_views;

#define CWA2D_NUM_KERNELS 7
typedef struct cwa2d_kernel_times {
    float ms[CWA2D_NUM_KERNELS]; /* hash, scan, scatter, reorder, density,
                                    force, wave (summed over SPH substeps)       */
} cwa2d_kernel_times;

typedef struct cwa2d_batch cwa2d_batch;

int cwa2d_abi_version(void);
int cwa2d_struct_size(int which); /* 0 params, 1 scene, 2 tuning, 3 views, 4 times */
const char* cwa2d_kernel_name(int k);

/* The SphWave2D application's constants (H = 0.1, 9.6 x 9.6 view, floor at
 * 0.5 VIEW_HEIGHT - 14 PARTICLE_RADIUS + bottom, Wave1D defaults). */
int cwa2d_params_reference(cwa2d_params* p);

/* Wave1D InitWave (Wave1D_cs.glsl:96-118) as ImageStencil::Reinit runs it:
 * mode 0 then mode 1, so h1 (older) has the bump at int(0.25 L) and h0
 * (newer) one cell further.  Host computation. */
int cwa2d_wave_init_reference(int32_t L, float* h0, float* h1);
/* Shallow1D InitWave (Shallow1D_cs.glsl:117-131): h = view_height/2 +
 * 0.1 view_height exp(-xc^2/0.005), uh = 0.5 |bump| sign(xc), xc = k/(L-1) - 1/2.
 * Host computation. */
int cwa2d_wave_init_shallow(int32_t L, float view_height, float* h, float* uh);

/* xv: 4 floats per particle (x, y, vx, vy), scenes concatenated in order;
 * enabled: per particle 0/1 or null (all enabled, as InitParticle :375);
 * h0/h1: wave levels, scenes concatenated (null: zero). */
int cwa2d_batch_create(const cwa2d_params* p, int32_t nscenes, const cwa2d_scene* scenes,
                       const float* xv, const int32_t* enabled, const float* h0,
                       const float* h1, cwa2d_batch** out);
int cwa2d_batch_destroy(cwa2d_batch* b);
int cwa2d_batch_reinit(cwa2d_batch* b, const float* xv, const int32_t* enabled,
                       const float* h0, const float* h1);
int cwa2d_batch_set_params(cwa2d_batch* b, const cwa2d_params* p);
int cwa2d_batch_get_params(cwa2d_batch* b, cwa2d_params* p);

/* `steps` coupled steps of SPH time step dt (each: sph_substeps SPH
 * substeps, then every scene's wave_substeps wave substeps). */
int cwa2d_batch_step(cwa2d_batch* b, float dt, int32_t steps);
/* The same, replayed from a hipGraph of one step (captured on first use
 * and after any change of dt, parameters or tuning). */
int cwa2d_batch_step_graph(cwa2d_batch* b, float dt, int32_t steps);
/* One step with a HIP event pair around every kernel. */
int cwa2d_batch_step_profiled(cwa2d_batch* b, float dt, cwa2d_kernel_times* t);
int cwa2d_batch_sync(cwa2d_batch* b);
/* The "Splash" button (SphWave2D/Main.cpp:124-129, ImageStencil::ComputeFunc(1)):
 * Shallow1D's Splash (Shallow1D_cs.glsl:101-115) on every shallow-water scene
 * (scene < 0) or on one; view_height scales the bump.  Leapfrog scenes are
 * left unchanged. */
int cwa2d_batch_splash(cwa2d_batch* b, int32_t scene, float view_height);
/* Render interop: copies the particle positions (float2 per slot, slot
 * order) and the newest wave level into caller device memory, e.g. the
 * pointers of GL buffers mapped with hipGraphicsGLRegisterBuffer /
 * hipGraphicsResourceGetMappedPointer.  Either pointer may be null.
 * Ordered on the batch's stream. */
int cwa2d_batch_render(cwa2d_batch* b, float* points_xy, float* wave);
int cwa2d_batch_stream(cwa2d_batch* b, void** hip_stream);
int cwa2d_batch_views(cwa2d_batch* b, cwa2d_views* v);

/* Host copies in original order (scene-major, particle id within scene);
 * any pointer may be null.  xv: 4 floats, rp: 2 floats per particle. */
int cwa2d_batch_readback(cwa2d_batch* b, float* xv, float* rp, int32_t* enabled, float* h0,
                         float* h1);

/* Autotuner: times candidate launch configurations over `steps` steps each
 * from a snapshot of the state (restored afterwards), keeps the fastest
 * and returns it.  Results do not depend on the configuration. */
int cwa2d_batch_tune(cwa2d_batch* b, float dt, int32_t steps, cwa2d_tuning* best);
int cwa2d_batch_get_tuning(cwa2d_batch* b, cwa2d_tuning* t);
int cwa2d_batch_set_tuning(cwa2d_batch* b, const cwa2d_tuning* t);

#ifdef __cplusplus
}
#endif

#endif /* CWA2D_H */
