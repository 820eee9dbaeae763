/*
 * cwa.h — C ABI of the MI355X-native coupled SPH + heightfield-wave step.
 *
 * The reference drives one coupled step from C++ through a small object
 * API: a particle SSBO plus three compute programs (density/pressure,
 * forces, integration) and a triple-buffered heightfield Module
 * (StencilImage2DTripleBuffered) with Init / Reinit / Compute / PingPong /
 * GetReadImage.  This header is the drop-in boundary for that path: an
 * opaque scene handle owns the device buffers, `cwa_scene_step` is the
 * per-frame `idle()` dispatch sequence, and the buffer-handle getters play
 * the role of the SSBO / image handles that the renderer consumes.
 *
 * Conventions
 *   - every entry point returns an int status: CWA_OK (0) or a negative
 *     CWA_E* code; `cwa_last_error()` holds a message for the calling thread.
 *     No C++ exception crosses this boundary.
 *   - the library owns all device memory of a scene; pointers handed out by
 *     the getters stay valid until the next step / reinit / destroy.
 *   - one HIP stream per scene; calls on one handle are not thread-safe.
 *   - particle arrays on the device are SoA float4 (xyz + pad) in *cell
 *     sorted* order; `id[slot]` is the original index of the particle in
 *     that slot.  `cwa_scene_readback` returns arrays in original order.
 *   - heightfields are single-channel fp32, row-major, texel (i, j) at
 *     [j * width + i]; i runs along u (world x) and j along v (world z).
 *
 * Reference citations are relative to the reference repository root.
 */
#ifndef CWA_H
#define CWA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CWA_ABI_VERSION 1

/* status codes */
#define CWA_OK 0
#define CWA_EINVAL (-1)   /* invalid argument / parameter */
#define CWA_EHIP (-2)     /* HIP runtime error */
#define CWA_ENOMEM (-3)   /* device allocation failed */
#define CWA_ESTATE (-4)   /* call not valid in the current state */
#define CWA_ENONFINITE (-5) /* non-finite value detected by a finite check */

/* Wave initialisation modes, StencilImage2DTripleBuffered.h:20-23. */
#define CWA_MODE_INIT 0
#define CWA_MODE_INIT_FROM_TEXTURE 1
#define CWA_MODE_EVOLVE 2

/* Largest particle count of one scene (slot indices keep two spare bits). */
#define CWA_MAX_PARTICLES (1 << 29)

/*
 * Every constant of the reference shaders and uniform blocks, as a field.
 *   ConstantsUniform  Main.cpp:184-190   (mass, smoothing, visc, rest rho)
 *   BoundaryUniform   Main.cpp:192-198   (upper / lower bounds)
 *   WaveUniforms      Main.cpp:200-204   (lambda, attenuation, beta, type)
 *   #defines          rho_pres_comp.glsl:3-5,41-43; force_comp.glsl:3-9,52-55;
 *                     integrate_comp.glsl:3-8,51
 * `cwa_params_reference()` fills the reference's own values.
 */
typedef struct cwa_params {
    /* --- SPH --- */
    float mass;             /* ConstantsUniform.mass */
    float h;                /* smoothing length = smoothing_coeff * PARTICLE_RADIUS */
    float visc;             /* ConstantsUniform.visc (mu) */
    float rest_rho;         /* ConstantsUniform.resting_rho */
    float gas_const;        /* GAS_CONST, rho_pres_comp.glsl:41 (stiffness k) */
    float particle_radius;  /* PARTICLE_RADIUS */
    float spiky_coeff;      /* C in the spiky gradient -C/(pi h^6), force_comp.glsl:68 */
    float visc_coeff;       /* C in the viscosity laplacian +C/(pi h^6), force_comp.glsl:69 */
    float gravity[3];       /* G, force_comp.glsl:52 */
    float crest_threshold;  /* CREST_THRESHOLD, force_comp.glsl:5 */
    float breaking_mass;    /* BREAKING_MASS_FACTOR, force_comp.glsl:6 */
    float breaking_visc;    /* BREAKING_VISC_FACTOR, force_comp.glsl:7 */
    float torque_coeff;     /* rotational coefficient, force_comp.glsl:104 */
    float drag_coeff;       /* wave drag coefficient, force_comp.glsl:108 */
    float wave_force_coeff; /* wave normal force coefficient, force_comp.glsl:110 */
    float wave_vel_du;      /* uv offset in WaveVelocity, force_comp.glsl:119 */
    float vel_damping;      /* DAMPING in v *= 1 - DAMPING*dt, integrate_comp.glsl:66 */
    float bounce;           /* DAMPING in v *= -DAMPING at walls, integrate_comp.glsl:140 */
    float foam_speed;       /* |v| foam threshold, integrate_comp.glsl:69 */
    float foam_force;       /* force scale in foam, integrate_comp.glsl:72 */
    float foam_rho;         /* density scale in foam, integrate_comp.glsl:73 */
    float foam_pressure;    /* pressure scale in foam, integrate_comp.glsl:74 */
    float foam_vel;         /* velocity scale in foam, integrate_comp.glsl:75 */
    float upper[3];         /* BoundaryUniform.upper.xyz */
    float lower[3];         /* BoundaryUniform.lower.xyz */
    float ceiling_extra;    /* +0.1 ceiling in wake mode, integrate_comp.glsl:172 */
    /* --- heightfield placement: uv = (pos.xz - wave_origin) * wave_uv_scale
     *     (the reference samples at 2*pos.xz, rho_pres_comp.glsl:72) --- */
    float wave_origin[2];
    float wave_uv_scale[2];
    /* --- wave stencil, wave_comp.glsl:171-187 (attributes = lambda,
     *     attenuation, beta, type); lambda may differ per axis --- */
    float lambda_x;         /* along u (texel i, world x) */
    float lambda_z;         /* along v (texel j, world z) */
    float atten;
    float beta;
    float wave_type;        /* 1 splash, 0 wave, in (0,1) wake */
    float wake_threshold;   /* wave_comp.glsl:177 */
    float wake_amount;      /* wave_comp.glsl:182 */
    /* --- two-way coupling (particle -> wave has no reference counterpart) ---
     * couple_band > 0: wave -> particle terms act only where |y - H| < band and
     *   particles in that band push the surface (volume-flux source scaled
     *   by couple_strength).  couple_band <= 0: the reference's behaviour,
     *   wave -> particle terms on every particle, no particle -> wave term. */
    float couple_band;
    float couple_strength;
    /* --- uniform grid (Ugrid3D, UniformGridParticles3D.h:60-67):
     *     cell = clamp(floor((p - grid_min) / edge), 0, dims - 1), with
     *     edge = cell_size along x and z and cell_size / grid_ydiv along y --- */
    float grid_min[3];
    float cell_size;        /* x and z cell edge; must be >= h */
    int32_t grid_dims[3];   /* cells along x, y, z */
    int32_t grid_ydiv;      /* y rows per cell_size (1..16; 0 means 1).  Finer
                               rows let the neighbour search keep, in each of
                               the 9 columns, only the rows the support sphere
                               reaches */
    int32_t check_finite;   /* nonzero: step reports CWA_ENONFINITE on NaN/Inf */
} cwa_params;

/* Initial state handed to create / reinit (host memory). */
typedef struct cwa_init {
    int32_t n;              /* particle count, 0..CWA_MAX_PARTICLES */
    const float* pos;       /* 4*n floats, xyz + pad (Particle.pos, Main.cpp:169) */
    const float* vel;       /* 4*n floats or NULL for zero (Particle.vel) */
    const float* force;     /* 4*n floats or NULL for zero (Particle.force) */
    int32_t wave_w, wave_h; /* heightfield texels, u then v */
    const float* height0;   /* w*h floats, newest field h^n, or NULL for zero */
    const float* height1;   /* w*h floats, previous field h^(n-1), or NULL = height0 */
} cwa_init;

/* Device pointers of the particle buffers (replaces particles_ssbo): the
 * newest state, in the cell order of the last step.  The library alternates
 * two state buffers, so the pointers hold until the next step, reinit or
 * destroy.  The last step already hashed these positions into the next
 * step's cell keys: after writing positions through the view, call
 * cwa_scene_particles_written (or change the state with cwa_scene_reinit). */
typedef struct cwa_particle_view {
    int32_t n;
    float* pos;    /* float4[n] */
    float* vel;    /* float4[n] */
    float* force;  /* float4[n] */
    float* rho;    /* float[n]  (Particle.extras[0]) */
    float* prs;    /* float[n]  (Particle.extras[1]) */
    int32_t* id;   /* int[n], original particle index per slot */
} cwa_particle_view;

/* Device pointers of the uniform grid of the last step (cell order as
 * cwa_grid_build's). */
typedef struct cwa_grid_view {
    int32_t dims[3];
    int32_t num_cells;
    int32_t n;
    int32_t* key;     /* int[n]   cell of each slot before the step's reorder */
    int32_t* count;   /* int[num_cells] particles per cell */
    int32_t* offset;  /* int[num_cells+1] exclusive scan of count */
    int32_t* list;    /* int[n]   stable cell order: slot indices (pre-reorder) */
    float* flux;      /* float[dims[0]*dims[2]] column flux into the wave */
} cwa_grid_view;

/* Host arrays for readback, in original particle order. Any may be NULL. */
typedef struct cwa_host_state {
    float* pos;     /* 4*n */
    float* vel;     /* 4*n */
    float* force;   /* 4*n */
    float* rho;     /* n */
    float* prs;     /* n */
    float* height0; /* w*h, newest */
    float* height1; /* w*h, previous */
} cwa_host_state;

/* Average per-kernel device time over a profiled run (HIP events on the
 * scene stream).  Slots that did not run have launches 0 (the column flux
 * is summed inside the force launch and is timed with it). */
#define CWA_NUM_KERNELS 10
typedef struct cwa_kernel_times {
    int32_t steps;
    float ms[CWA_NUM_KERNELS];     /* average ms per launch */
    int32_t launches[CWA_NUM_KERNELS];
} cwa_kernel_times;

typedef struct cwa_scene cwa_scene;

/* ---- library ---- */
int cwa_abi_version(void);
/* sizeof of the ABI structs, for binding checks: 0 cwa_params, 1 cwa_init,
 * 2 cwa_particle_view, 3 cwa_grid_view, 4 cwa_host_state,
 * 5 cwa_kernel_times, 6 cwa_grad; -1 for an unknown index. */
int cwa_struct_size(int which);
/* Number of visible HIP devices (0 when none), or a negative status. */
int cwa_device_count(void);
const char* cwa_last_error(void);
/* Name of kernel slot k in cwa_kernel_times (e.g. "density"). */
const char* cwa_kernel_name(int k);

/* Lanes per particle (no reference counterpart: a tuning knob of this
 * implementation, also used by tests to reach both kernel forms on small
 * scenes).  Scenes of up to max_n particles run the density pass with four
 * lanes per particle and the pair adjoint of a tape created afterwards with
 * four; larger scenes one lane each.  Default 262144 (2^18); max_n < 0
 * restores it.  Process-wide; taken at each density launch (a step graph
 * keeps the form it was captured with) and at tape creation.  *previous
 * (optional) receives the limit before the call.  Both forms find the same
 * neighbour words; the density sums differ in summation order only. */
int cwa_set_lane_split_limit(long long max_n, long long* previous);

/* Reference defaults (Main.cpp:184-204 and the shader #defines) for a grid
 * over the reference box; the heightfield spans [0,0.5]^2 in xz. */
void cwa_params_reference(cwa_params* p);

/* ---- scene lifetime: Module::Init / sInitAll, init_particles
 *      (Main.cpp:764-794), StencilImage2DTripleBuffered::Init (.cpp:10-31) */
int cwa_scene_create(const cwa_params* p, const cwa_init* init, cwa_scene** out);
/* Keyboard reset 'R' (Main.cpp:677-684): new particles and fields. */
int cwa_scene_reinit(cwa_scene* s, const cwa_init* init);
int cwa_scene_destroy(cwa_scene* s);
/* Uniform-block uploads of display() (Main.cpp:366-382). */
int cwa_scene_set_params(cwa_scene* s, const cwa_params* p);
int cwa_scene_get_params(const cwa_scene* s, cwa_params* p);

/* ---- heightfield Module: StencilImage2DTripleBuffered ---- */
/* Reinit (.cpp:42-59): run MODE_INIT twice, both history levels equal. */
int cwa_scene_wave_reinit(cwa_scene* s);
/* ReinitFromTexture (.cpp:61-77): one MODE_INIT_FROM_TEXTURE dispatch into
 * the write image, then PingPong, so only the newest level gets texel
 * (i, j) of a single-channel fp32 host image read at (2*i, j) (texelFetch
 * at coord*(2,1), wave_comp.glsl:76-80) and the previous newest becomes the
 * previous level; fetches past the image read 0. */
int cwa_scene_wave_reinit_from_texture(cwa_scene* s, const float* tex, int tex_w, int tex_h);
/* Compute (.cpp:79-95): one EVOLVE dispatch then PingPong, with no
 * particle coupling. */
int cwa_scene_wave_compute(cwa_scene* s, float dt);
/* GetReadImage(age) (.h:14): age 0 = newest, 1 = previous; age 2 is the
 * write image (GetWriteImage, .h:13). pitch in floats. */
int cwa_scene_height(const cwa_scene* s, int age, float** dptr, int* w, int* h, int* pitch);

/* ---- the coupled step: idle() Main.cpp:530-562 ----
 * Per step: grid build (hash, count, scan, stable insert + reorder),
 * density/pressure, forces + symplectic Euler + boundary with the particle
 * -> wave column flux in the same launch, wave EVOLVE + PingPong. */
int cwa_scene_step(cwa_scene* s, float dt, int nsteps);
/* Same as cwa_scene_step with HIP events around every kernel; `out`
 * receives the per-kernel average device times. */
int cwa_scene_step_profiled(cwa_scene* s, float dt, int nsteps, cwa_kernel_times* out);
/* Replay-based execution: capture a 6-step cycle (the period of the double
 * buffers of the grid counters / keys and of the particle state x the wave
 * triple buffer) into a hipGraph and
 * launch it; nsteps that are not a multiple of 6 run the tail eagerly.
 * The graph is re-captured when dt changes. */
int cwa_scene_step_graph(cwa_scene* s, float dt, int nsteps);
/* Captures cwa_scene_step_graph's 6-step graph for dt now, without running
 * a step, so later cwa_scene_step_graph calls with this dt only replay it
 * (for loops that must not pay the capture inside a timed or latency-bound
 * region; no reference counterpart).  *captured (optional) = 1 when a graph
 * is ready, 0 when none could be captured yet: before the scene's first step
 * (a captured step takes its cell keys from the step before it) or with
 * check_finite set (those steps run eagerly). */
int cwa_scene_prepare_graph(cwa_scene* s, float dt, int* captured);
int cwa_scene_sync(cwa_scene* s);
/* The scene's hipStream_t (as void*). */
void* cwa_scene_stream(const cwa_scene* s);

/* ---- buffer handles ---- */
int cwa_scene_particles(const cwa_scene* s, cwa_particle_view* v);
/* Declares that the caller wrote particle state (positions in particular)
 * through cwa_scene_particles' pointers (the reference's particle SSBO,
 * Main.cpp:779-782, is GL storage any GL code may write between dispatches;
 * no function of the reference corresponds): the next step hashes the positions
 * afresh instead of using the cell keys the last step computed ahead.  Call
 * it after the writes are ordered before the scene's next step (on the
 * scene's stream, or synchronised), and before that step. */
int cwa_scene_particles_written(cwa_scene* s);
int cwa_scene_grid(const cwa_scene* s, cwa_grid_view* v);
int cwa_scene_readback(cwa_scene* s, cwa_host_state* h);
/* Neighbour count of every particle (original order, n int32 on the host)
 * as the last step's density pass found it: the j with |x_i - x_j| < h, the
 * particle itself included, over which rho_pres_comp.glsl:60-68 sums.
 * CWA_ESTATE before the first step and after a reinit or set_params. */
int cwa_scene_neighbour_counts(cwa_scene* s, int32_t* counts);

/* ---- standalone primitives on caller device buffers (stream may be NULL
 *      for the null stream) ---- */
/* ParallelScan::Compute (SphWave2D/ParallelScan.cpp:43-95): exclusive
 * prefix sum of n int32 values, any n >= 0 (the reference requires a power
 * of two). out[n] receives the total; out must hold n+1 ints.  in and out
 * may not alias.  Uses an internal device workspace per call. */
int cwa_scan_exclusive_i32(const int32_t* d_in, int32_t* d_out, int n, void* stream);
/* UgridParticles3D::BuildGrid (UniformGridParticles3D.cpp:170-211) with the
 * stable insertion order of UniformGrid2D::Build (UniformGrid2D.cpp:33-92):
 * key[i] = linear cell of pos[i], count, offset (num_cells+1), list = indices
 * sorted by (cell, index).  Linear cell = column(ix, iz)*dims_y + iy, y
 * fastest; columns in z-tiles of 16, x-major within a tile (the last tile
 * narrower when 16 does not divide dims_z): column(ix, iz) =
 * t*16*dims_x + ix*w + (iz - 16 t), t = iz/16, w = min(16, dims_z - 16 t).
 * The scene's grid (cwa_scene_grid) uses the same order. */
int cwa_grid_build(const float* d_pos4, int n, const float grid_min[3], float cell_size,
                   const int32_t dims[3], int32_t* d_key, int32_t* d_count,
                   int32_t* d_offset, int32_t* d_list, void* stream);

/* cwa_grid_build without allocating: no reference counterpart; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  Same arguments and results as cwa_grid_build, but its
 * scratch (run records, fill counters, the scan's workspace) is one
 * file-scope device array of the library (32 MB per GPU, held until the
 * process ends), so n <= 2^21 and dims[0]*dims[1]*dims[2] <= 2^22, else
 * CWA_EINVAL.  Asynchronous on `stream` (no synchronisation; results are
 * ready when the stream reaches them).  The first call on a device looks
 * up the array's address; make it outside graph capture.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams
 * or threads) race. */
int cwa_grid_build_static(const float* d_pos4, int n, const float grid_min[3], float cell_size,
                          const int32_t dims[3], int32_t* d_key, int32_t* d_count,
                          int32_t* d_offset, int32_t* d_list, void* stream);

/* UniformGridPbdGpu3D::CollisionQuery's grid build (UniformGridGpu3D.cpp:
 * 121-162; uniform_grid_pbd_3D_cs.glsl ComputeCounts / ComputeCellIndex /
 * Index / InsertBox :145-206): box i spans the cells
 * ComputeCellIndex(centre - halfw) .. ComputeCellIndex(centre + halfw) per
 * axis, cell = clamp(floor((p - grid_min) / cell_size), 0, dims - 1) with a
 * cell size per axis, linear cell (ix*dims_z + iz)*dims_y + iy.  count[c]
 * = boxes overlapping cell c, offset = its exclusive scan (num_cells+1),
 * list[offset[c] ..] = those boxes in ascending index order (the reference
 * inserts with atomicAdd, in no fixed order).  d_centre4 / d_halfw4: float4
 * per box (xyz used).  *total receives the number of (box, cell) pairs
 * (offset[num_cells]); list must hold that many (the reference sizes it 8
 * per box, boxes no larger than a cell), else CWA_EINVAL with *total set. */
int cwa_grid_build_boxes(const float* d_centre4, const float* d_halfw4, int n,
                         const float grid_min[3], const float cell_size[3],
                         const int32_t dims[3], int32_t* d_count, int32_t* d_offset,
                         int32_t* d_list, int64_t list_capacity, int64_t* total,
                         void* stream);

/* ---- gradients: checkpointed forward + adjoint (no reference
 *      counterpart; the differentiable form of the step above) ----
 * The adjoint differentiates the step as run, with every discrete decision
 * frozen at its forward value (cell sort, neighbour sets, EOS / rest
 * density clamps, coupling band, foam, surface and box clamps, crest).
 * Requires torque_coeff == 0.  Parameter gradients are returned for the
 * fields below; lambda = (c dt / dx)^2 gives d/dc by the chain rule. */
typedef struct cwa_grad {
    double visc;             /* d loss / d cwa_params.visc */
    double gas_const;        /* .gas_const */
    double rest_rho;         /* .rest_rho */
    double lambda_x;         /* .lambda_x */
    double lambda_z;         /* .lambda_z */
    double atten;            /* .atten (the wave damping) */
    double couple_strength;  /* .couple_strength */
    double beta;             /* .beta */
} cwa_grad;

typedef struct cwa_tape cwa_tape;

/* A tape over `nsteps` steps of scene s from its current state, with a
 * checkpoint every `ckpt_every` steps.  s must outlive the tape, and its
 * layout must stay as it was at creation: after cwa_scene_reinit, or a
 * cwa_scene_set_params that changes the grid, the sign of couple_band or
 * the wake mode, the tape's calls return CWA_ESTATE.  Device
 * memory: nsteps / ckpt_every + 1 checkpoints of the particle state and two
 * heightfields, and two step tapes of ckpt_every steps each (one recorded
 * while the adjoint walks the other). */
int cwa_tape_create(cwa_scene* s, int nsteps, int ckpt_every, cwa_tape** out);
/* Run the nsteps (the scene ends in the final state), storing checkpoints. */
int cwa_tape_forward(cwa_tape* t, float dt);
/* Adjoint sweep for the loss whose derivatives with respect to the final
 * state are the seeds (host arrays; NULL = zero): dheight0 / dheight1
 * (w*h, newest / previous heightfield), dpos / dvel (4*n, original particle
 * order, xyz used).  Writes the parameter gradients and, if not NULL, the
 * gradients with respect to the initial positions / velocities (4*n,
 * original order).  Recomputes each segment from its checkpoint (on the
 * scene stream, overlapped with the previous segment's adjoint on a second
 * stream the tape owns); the scene is left in the final state. */
int cwa_tape_backward(cwa_tape* t, const float* dheight0, const float* dheight1,
                      const float* dpos, const float* dvel, cwa_grad* grad, float* dpos0,
                      float* dvel0);
int cwa_tape_destroy(cwa_tape* t);

/* ---- batched heightfield fits (no reference counterpart) ----
 * For each of `nscenes` scenes: the checkpointed forward of `nsteps` steps
 * from the scene's state at creation, the loss mean((h0 - target)^2) of its
 * newest heightfield against `targets[s]` (H x W, row-major), and the
 * adjoint sweep.  All scenes are captured into one hipGraph (each scene on
 * its own stream) at creation; cwa_fit_run replays it and returns each
 * scene's loss and parameter gradients.  cwa_scene_set_params between runs
 * takes effect without a new capture as long as the grid, the coupling
 * band's sign and the wave mode stay as they were (otherwise, and after a
 * cwa_scene_reinit of a fit scene, cwa_fit_run returns CWA_ESTATE).
 * Scenes need torque_coeff == 0
 * and check_finite == 0, must not be stepped or destroyed while the fit
 * exists, and each scene may appear once. */
typedef struct cwa_fit cwa_fit;
int cwa_fit_create(int32_t nscenes, cwa_scene* const* scenes, const float* const* targets,
                   int32_t nsteps, int32_t ckpt_every, float dt, cwa_fit** out);
/* losses: nscenes doubles; grads: nscenes cwa_grad (either may be null). */
int cwa_fit_run(cwa_fit* f, double* losses, cwa_grad* grads);
/* Device bytes the fit allocated for one scene (tape, checkpoints, target). */
int cwa_fit_tape_bytes(cwa_fit* f, int32_t scene, int64_t* bytes);
int cwa_fit_destroy(cwa_fit* f);

/* ---- scene generators (host only, no device use) ----
 * Random numbers: SplitMix64 stream from `seed`; u = (x >> 40) * 2^-24.
 *
 * Jittered lattice.  Site k along axis a sits at lo[a] + (k + site_offset) *
 * spacing.  Sites are enumerated with axis order[0] fastest and order[2]
 * slowest; the two fast axes hold floor((hi-lo)/spacing) sites each, the
 * slowest axis is unbounded (layers are added until n sites exist).  Each
 * coordinate then moves by (2u - 1) * jitter * spacing, drawn x, y, z per
 * site in enumeration order.  Writes 4*n floats (w = 1), returns n. */
int cwa_gen_lattice(float* pos4, int n, const float lo[3], const float hi[3],
                    float spacing, float site_offset, float jitter,
                    const int32_t order[3], uint64_t seed);
/* Sum of Gaussian bumps on a w x h texel grid: centres U[0,w) x U[0,h),
 * amplitude U[amp_lo, amp_hi], sigma U[sig_lo, sig_hi] texels. */
int cwa_gen_bumps(float* height, int w, int h, int nbumps, float amp_lo, float amp_hi,
                  float sig_lo, float sig_hi, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* CWA_H */
