// cwa_scene.hip — host objects of the coupled step and the C ABI (cwa.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cwa.h"
#include "cwa_modules.hpp"

namespace cwa {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(CWA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// ---------------------------------------------------------------------------
// Module registry (Module.cpp:5-59)
// ---------------------------------------------------------------------------
Module::Module() { sAllModules().push_back(this); }
Module::~Module() { sAllModules().remove(this); }
std::list<Module*>& Module::sAllModules() {
    static std::list<Module*> all;
    return all;
}
void Module::sInitAll() {
    for (Module* p : sAllModules()) p->Init();
}
void Module::sReinitAll() {
    for (Module* p : sAllModules()) p->Reinit();
}
void Module::sComputeAll() {
    for (Module* p : sAllModules()) p->Compute();
}

// ---------------------------------------------------------------------------
// ParallelScan
// ---------------------------------------------------------------------------
void ParallelScan::Init(const int* input, int n) {
    if (n < 0) throw Error(CWA_EINVAL, "scan: n < 0");
    in_ = input;
    n_ = n;
    out_.resize((size_t)n + 1);
    ws_.resize(scan_workspace_ints(n));
}

void ParallelScan::Compute(hipStream_t st) {
    hip_check(launch_scan(in_, out_.get(), n_, ws_.get(), st), "scan");
}

// ---------------------------------------------------------------------------
// UniformGrid3D
// ---------------------------------------------------------------------------
void UniformGrid3D::Init(const GridDesc& g, int max_particles) {
    g_ = g;
    const size_t nc = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    const size_t np = (size_t)std::max(max_particles, 1);
    for (int k = 0; k < 2; ++k) {
        key_[k].resize(np);
        count_[k].resize(nc);
    }
    fill_.resize(nc);
    runs_.resize(np);
    list_.resize(np);
    offset_.resize(nc + 1);
    ws_.resize(scan_workspace_ints((int)nc));
    cur_ = last_ = 0;
    kc_ = last_kc_ = 0;
    keys_ready_ = false;
}

void UniformGrid3D::ClearCounters(hipStream_t st) {
    for (auto& c : count_) hip_check(launch_fill_i32(c.get(), num_cells(), 0, st), "clear counts");
    hip_check(launch_fill_i32(fill_.get(), num_cells(), 0, st), "clear fill");
    hip_check(launch_fill_i32(offset_.get(), num_cells() + 1, 0, st), "clear offsets");
    keys_ready_ = false;
}

void UniformGrid3D::BuildGrid(const ParticleArrays& A, int n, float4* bpos, int* bid,
                              hipStream_t st, hipEvent_t* ev, int* list_out, int* offset_out) {
    const int nc = num_cells();
    int* lst = list_out ? list_out : list_.get();
    int* off = offset_out ? offset_out : offset_.get();
    auto mark = [&](int slot, int which) {
        if (ev) hip_check(hipEventRecord(ev[2 * slot + which], st), "event");
    };
    int* cnt = count_[cur_].get();  // zero: cleared by the previous build's scan
    int* key = key_[kc_].get();
    mark(K_HASH, 0);
    if (!keys_ready_)  // else the last force pass hashed these positions
        hip_check(launch_hash_count(A.pos, n, g_, key, cnt, st), "hash");
    mark(K_HASH, 1);
    keys_ready_ = false;
    mark(K_SCAN, 0);
    hip_check(launch_scan(cnt, off, nc, ws_.get(), st, count_[cur_ ^ 1].get()), "scan");
    mark(K_SCAN, 1);
    last_ = cur_;
    cur_ ^= 1;
    last_kc_ = kc_;
    mark(K_SCATTER, 0);
    hip_check(launch_scatter_runs(key, off, fill_.get(), n, runs_.get(), st), "scatter");
    mark(K_SCATTER, 1);
    mark(K_REORDER, 0);
    hip_check(launch_stable_reorder(key, off, runs_.get(), n, nullptr, lst, A.pos, A.id, bpos, bid,
                                    fill_.get(), nc, st),
              "reorder");
    mark(K_REORDER, 1);
}

// ---------------------------------------------------------------------------
// HeightField2DTripleBuffered
// ---------------------------------------------------------------------------
void HeightField2DTripleBuffered::Allocate(int w, int h) {
    w_ = w;
    h_ = h;
    for (auto& b : img_) b.resize((size_t)w * h);
    read_[0] = 0;
    read_[1] = 1;
    write_ = 2;
}

void HeightField2DTripleBuffered::ReinitMode(float wave_type, hipStream_t st) {
    for (int k = 0; k < 2; ++k) {
        hip_check(launch_init_wave(GetWriteImage(), w_, h_, wave_type, st), "init wave");
        PingPong();
    }
}

void HeightField2DTripleBuffered::Upload(const float* h0, const float* h1, hipStream_t st) {
    const size_t bytes = (size_t)w_ * h_ * sizeof(float);
    if (bytes == 0) return;
    if (h0)
        hip_check(hipMemcpyAsync(GetReadImage(0), h0, bytes, hipMemcpyHostToDevice, st), "h0");
    else
        hip_check(hipMemsetAsync(GetReadImage(0), 0, bytes, st), "h0 zero");
    if (h1)
        hip_check(hipMemcpyAsync(GetReadImage(1), h1, bytes, hipMemcpyHostToDevice, st), "h1");
    else
        hip_check(hipMemcpyAsync(GetReadImage(1), GetReadImage(0), bytes, hipMemcpyDeviceToDevice,
                                 st),
                  "h1 copy");
    hip_check(hipMemsetAsync(GetWriteImage(), 0, bytes, st), "write zero");
    hip_check(hipStreamSynchronize(st), "upload sync");
}

void HeightField2DTripleBuffered::ReinitFromTexture(const float* tex, int tex_w, int tex_h,
                                                    hipStream_t st) {
    if (tex_w < 1 || tex_h < 1 || !tex) throw Error(CWA_EINVAL, "empty init texture");
    // texelFetch(uInitImage, (2x, y)) (wave_comp.glsl:76-80); a fetch outside
    // the texture reads 0, the robust-access result (GL leaves it undefined,
    // and the reference's 64^2 textures on its 64^2 field fetch past x = 31)
    std::vector<float> f((size_t)w_ * h_, 0.0f);
    for (int j = 0; j < h_ && j < tex_h; ++j)
        for (int i = 0; i < w_ && 2 * i < tex_w; ++i)
            f[(size_t)j * w_ + i] = tex[(size_t)j * tex_w + 2 * i];
    hip_check(hipMemcpyAsync(GetWriteImage(), f.data(), f.size() * sizeof(float),
                             hipMemcpyHostToDevice, st),
              "texture upload");
    hip_check(hipStreamSynchronize(st), "texture sync");
    PingPong();
}

void HeightField2DTripleBuffered::Evolve(const DevParams& P, const float* colflux, float dt,
                                         hipStream_t st) {
    hip_check(launch_wave(P, GetReadImage(0), GetReadImage(1), GetWriteImage(), colflux, dt, st),
              "wave");
}

// The texture units of the reference rotate so that the image just written
// is read as the newest and the newest becomes the previous
// (StencilImage2DTripleBuffered.cpp:33-40 with SwapUnits); period 3.
void HeightField2DTripleBuffered::PingPong() {
    const int w = write_;
    write_ = read_[1];
    read_[1] = read_[0];
    read_[0] = w;
}

// ---------------------------------------------------------------------------
// SphParticles3D
// ---------------------------------------------------------------------------
void SphParticles3D::Allocate(int n) {
    n_ = n;
    const size_t m = (size_t)std::max(n, 1);
    for (int k = 0; k < 2; ++k) {
        pos_[k].resize(m);
        vel_[k].resize(m);
        force_[k].resize(m);
        id_[k].resize(m);
    }
    bpos_.resize(m);
    bid_.resize(m);
    rho_.resize(m);
    prs_.resize(m);
    rho_w_.resize(m);
    flux_.resize(m);
    vpad_.resize(m);
    nbw_.resize(nb_buffer_uint2(m));
    nrec_.resize(2 * m);
    A = ParticleArrays{pos_[0].get(), vel_[0].get(), force_[0].get(), id_[0].get()};
    N = ParticleArrays{pos_[1].get(), vel_[1].get(), force_[1].get(), id_[1].get()};
    parity_ = 0;
}

void SphParticles3D::Upload(const float* pos, const float* vel, const float* force,
                            hipStream_t st) {
    if (n_ == 0) return;
    const size_t b4 = (size_t)n_ * sizeof(float4);
    hip_check(hipMemcpyAsync(A.pos, pos, b4, hipMemcpyHostToDevice, st), "pos");
    if (vel)
        hip_check(hipMemcpyAsync(A.vel, vel, b4, hipMemcpyHostToDevice, st), "vel");
    else
        hip_check(hipMemsetAsync(A.vel, 0, b4, st), "vel zero");
    if (force)
        hip_check(hipMemcpyAsync(A.force, force, b4, hipMemcpyHostToDevice, st), "force");
    else
        hip_check(hipMemsetAsync(A.force, 0, b4, st), "force zero");
    std::vector<int> ids(n_);
    std::iota(ids.begin(), ids.end(), 0);
    hip_check(hipMemcpyAsync(A.id, ids.data(), (size_t)n_ * sizeof(int), hipMemcpyHostToDevice,
                             st),
              "id");
    // Particle.extras = (rest rho, 0, ...) at init (Main.cpp:774)
    hip_check(hipMemsetAsync(prs_.get(), 0, (size_t)n_ * sizeof(float), st), "prs");
    hip_check(hipStreamSynchronize(st), "upload sync");
}

// ---------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------
const char* kernel_slot_name(int k) {
    static const char* names[K_NUM] = {"clear",   "hash",           "scan",
                                       "scatter", "reorder",        "density",
                                       "force_integrate", "column_flux", "wave",
                                       "finite"};
    return (k >= 0 && k < K_NUM) ? names[k] : "?";
}

void validate_params(const cwa_params& p) {
    auto bad = [](const char* m) { throw Error(CWA_EINVAL, m); };
    if (!(p.h > 0.0f) || !std::isfinite(p.h)) bad("h must be > 0");
    if (!(p.cell_size >= p.h)) bad("cell_size must be >= h (27-cell search)");
    if (p.grid_ydiv < 0 || p.grid_ydiv > 16) bad("grid_ydiv must be in 0..16");
    for (int a = 0; a < 3; ++a)
        if (p.grid_dims[a] < 1) bad("grid_dims must be >= 1");
    const double nc = (double)p.grid_dims[0] * p.grid_dims[1] * p.grid_dims[2];
    if (nc > (double)(1 << 30)) bad("grid has more than 2^30 cells");
    if (!(p.rest_rho > 0.0f)) bad("rest_rho must be > 0");
    if (!(p.gas_const * p.particle_radius != 0.0f)) bad("gas_const * particle_radius must be != 0");
    if (p.wave_uv_scale[0] == 0.0f || p.wave_uv_scale[1] == 0.0f) bad("wave_uv_scale must be != 0");
    if (p.wave_vel_du == 0.0f) bad("wave_vel_du must be != 0");
    if (p.breaking_mass == 0.0f) bad("breaking_mass must be != 0");
}

// Smallest float x with sqrtf(x) >= h, so that  sqrt_rn(r2) < h  <=>  r2 < x
// (sqrt_rn is monotone).  Host sqrtf is IEEE correctly rounded.
// grid_ydiv with 0 read as 1 (zero-initialised parameter blocks).
static int ydiv_of(const cwa_params& p) { return p.grid_ydiv <= 0 ? 1 : p.grid_ydiv; }

static float exact_r2_threshold(float h) {
    float x = h * h;
    while (x > 0.0f && std::sqrt(x) >= h) x = std::nextafter(x, 0.0f);
    while (!(std::sqrt(x) >= h)) x = std::nextafter(x, INFINITY);
    return x;
}

DevParams make_devparams(const cwa_params& p, int W, int H) {
    DevParams d{};
    const float pi = 3.141592741f;  // PI as written in the shaders
    d.h = p.h;
    d.h2 = p.h * p.h;
    d.r2_lt = exact_r2_threshold(p.h);
    d.m = p.mass;
    const float h2 = p.h * p.h, h6 = h2 * h2 * h2, h9 = h6 * h2 * p.h;
    const float poly6 = 315.0f / (64.0f * pi * h9);
    d.m_poly6 = p.mass * poly6;
    d.rho0 = p.rest_rho;
    d.k = p.gas_const;
    d.k_rp = p.gas_const * p.particle_radius;
    d.rp = p.particle_radius;
    d.spiky = -p.spiky_coeff / (pi * h6);
    d.lap = p.visc_coeff / (pi * h6);
    d.visc = p.visc;
    for (int a = 0; a < 3; ++a) d.g[a] = p.gravity[a];
    d.crest = p.crest_threshold;
    d.brk_mass = p.breaking_mass;
    d.brk_visc = p.breaking_visc;
    d.torque = p.torque_coeff;
    d.drag = p.drag_coeff;
    d.wforce = p.wave_force_coeff;
    d.du = p.wave_vel_du;
    d.vdamp = p.vel_damping;
    d.bounce = p.bounce;
    d.foam_speed = p.foam_speed;
    d.foam_f = p.foam_force;
    d.foam_rho = p.foam_rho;
    d.foam_p = p.foam_pressure;
    d.foam_v = p.foam_vel;
    for (int a = 0; a < 3; ++a) {
        d.up[a] = p.upper[a];
        d.lo[a] = p.lower[a];
        d.gmin[a] = p.grid_min[a];
        d.dims[a] = p.grid_dims[a];
    }
    d.ceil_extra = p.ceiling_extra;
    d.wake = (p.wave_type > 0.0f && p.wave_type < 1.0f) ? 1 : 0;
    for (int a = 0; a < 2; ++a) {
        d.org[a] = p.wave_origin[a];
        d.uvs[a] = p.wave_uv_scale[a];
    }
    d.W = W;
    d.H = H;
    d.wc0 = 2.0f - 2.0f * p.lambda_x - 2.0f * p.lambda_z - p.beta;
    d.wc1 = 1.0f - p.beta;
    d.lx = p.lambda_x;
    d.lz = p.lambda_z;
    d.atten = p.atten;
    d.wake_thr = p.wake_threshold;
    d.wake_amt = p.wake_amount;
    d.band = p.couple_band;
    d.kappa = p.couple_strength;
    d.m_rho0 = p.mass / p.rest_rho;
    d.cell = p.cell_size;
    d.cell_y = p.cell_size / (float)ydiv_of(p);
    {   // margin of the trimmed candidate rows: far above the rounding of
        // the float bounds at the grid's coordinate magnitudes
        float ext = 0.0f;
        for (int a = 0; a < 3; ++a) {
            const float edge = a == 1 ? d.cell_y : d.cell;
            ext = std::max(ext, std::fabs(p.grid_min[a]) + (float)p.grid_dims[a] * edge);
        }
        d.grid_eps = 4e-6f * ext;
    }
    d.inv_area = 1.0f / (p.cell_size * p.cell_size);
    // texel -> column-grid coordinates (computed in double, rounded once)
    const double Wd = W > 0 ? W : 1, Hd = H > 0 ? H : 1, cell = p.cell_size;
    d.ax = (float)(1.0 / (Wd * (double)p.wave_uv_scale[0] * cell));
    d.bx = (float)(((double)p.wave_origin[0] + 0.5 / (Wd * (double)p.wave_uv_scale[0]) -
                    (double)p.grid_min[0]) / cell - 0.5);
    d.az = (float)(1.0 / (Hd * (double)p.wave_uv_scale[1] * cell));
    d.bz = (float)(((double)p.wave_origin[1] + 0.5 / (Hd * (double)p.wave_uv_scale[1]) -
                    (double)p.grid_min[2]) / cell - 0.5);
    d.num_cells = p.grid_dims[0] * p.grid_dims[1] * p.grid_dims[2];
    return d;
}

static GridDesc grid_desc(const cwa_params& p) {
    GridDesc g{};
    g.cell_y = p.cell_size / (float)ydiv_of(p);
    for (int a = 0; a < 3; ++a) {
        g.gmin[a] = p.grid_min[a];
        g.dims[a] = p.grid_dims[a];
    }
    g.cell = p.cell_size;
    return g;
}

// ---------------------------------------------------------------------------
// CoupledScene
// ---------------------------------------------------------------------------
static void validate_init(const cwa_init& in) {
    if (in.n < 0) throw Error(CWA_EINVAL, "n < 0");
    if (in.n > CWA_MAX_PARTICLES) throw Error(CWA_EINVAL, "n > CWA_MAX_PARTICLES");
    if (in.n > 0 && !in.pos) throw Error(CWA_EINVAL, "pos is null");
    if (in.wave_w < 1 || in.wave_h < 1) throw Error(CWA_EINVAL, "wave size must be >= 1x1");
}

CoupledScene::CoupledScene(const cwa_params& p, const cwa_init& init) {
    validate_params(p);
    validate_init(init);
    hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "stream");
    params_ = p;
    flag_.resize(1);
    Reinit(init);
}

CoupledScene::~CoupledScene() {
    invalidate_graph();
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
}

void CoupledScene::invalidate_graph() {
    if (gexec_) {
        (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
    }
}

void CoupledScene::build_devparams() {
    dp_ = make_devparams(params_, wave.width(), wave.height());
    dpd_.resize(1);
    dp_.dev = dpd_.get();
    // synchronous: no kernel of this scene is in flight here (callers have
    // synchronised the stream), and the next launch must see the new block
    hip_check(hipMemcpy(dpd_.get(), &dp_, sizeof(DevParams), hipMemcpyHostToDevice), "params");
}

void CoupledScene::Reinit(const cwa_init& init) {
    validate_init(init);
    invalidate_graph();
    hip_check(hipStreamSynchronize(st_), "sync");
    stepped_ = false;
    ++layout_gen_;
    particles.Allocate(init.n);
    particles.Upload(init.pos, init.vel, init.force, st_);
    grid.Init(grid_desc(params_), init.n);
    grid.ClearCounters(st_);
    colflux.resize((size_t)params_.grid_dims[0] * params_.grid_dims[2]);
    hip_check(launch_fill_i32(reinterpret_cast<int*>(colflux.get()), (int)colflux.size(), 0, st_),
              "colflux zero");
    wave.Allocate(init.wave_w, init.wave_h);
    wave.Upload(init.height0, init.height1, st_);
    build_devparams();
    hip_check(hipStreamSynchronize(st_), "reinit sync");
}

void CoupledScene::ParticlesWritten() {
    grid.ClearCounters(st_);  // stream-ordered after the caller's writes
    stepped_ = false;         // the neighbour words describe the old positions
}

void CoupledScene::SetParams(const cwa_params& p) {
    validate_params(p);
    const bool grid_changed = p.grid_dims[0] != params_.grid_dims[0] ||
                              p.grid_dims[1] != params_.grid_dims[1] ||
                              p.grid_dims[2] != params_.grid_dims[2] ||
                              p.grid_min[0] != params_.grid_min[0] ||
                              p.grid_min[1] != params_.grid_min[1] ||
                              p.grid_min[2] != params_.grid_min[2] ||
                              p.cell_size != params_.cell_size || ydiv_of(p) != ydiv_of(params_);
    const bool path_changed = (p.couple_band > 0.0f) != (params_.couple_band > 0.0f) ||
                              make_devparams(p, 1, 1).wake != make_devparams(params_, 1, 1).wake;
    invalidate_graph();
    hip_check(hipStreamSynchronize(st_), "sync");
    params_ = p;
    stepped_ = false;  // the last step's neighbour records used the old parameters
    if (grid_changed || path_changed) ++layout_gen_;
    if (grid_changed) {
        grid.Init(grid_desc(params_), particles.n());
        grid.ClearCounters(st_);
        colflux.resize((size_t)params_.grid_dims[0] * params_.grid_dims[2]);
    }
    hip_check(launch_fill_i32(reinterpret_cast<int*>(colflux.get()), (int)colflux.size(), 0, st_),
              "colflux zero");
    build_devparams();
}

void CoupledScene::Step(float dt, hipEvent_t (*ev)[2], const TapeOut* tape) {
    auto mark = [&](int slot, int which) {
        if (ev) hip_check(hipEventRecord(ev[slot][which], st_), "event");
    };
    const int n = particles.n();
    const bool taped_heights = tape && tape->h_out;
    const float* H0 = taped_heights ? tape->h_in0 : wave.GetReadImage(0);
    const bool couple = n > 0 && params_.couple_band > 0.0f;
    // where this step's arrays live (the scene's own, or the tape's)
    auto pick = [&](auto* t, auto* own) { return (tape && t) ? t : own; };
    float4* bpos = pick(tape ? tape->pos : nullptr, particles.bpos());
    int* off = pick(tape ? tape->offset : nullptr, grid.offset());
    int* lst = pick(tape ? tape->list : nullptr, grid.list());
    uint2* nbm = pick(tape ? tape->nbw : nullptr, particles.nbw());
    float4* nrec = pick(tape ? tape->nrec : nullptr, particles.nrec());
    float* rho_w = pick(tape ? tape->rho : nullptr, particles.rho_w());
    float* cflux = pick(tape ? tape->colflux : nullptr, colflux.get());
    if (n > 0) {
        grid.BuildGrid(particles.A, n, bpos, particles.bid(), st_, ev ? &ev[0][0] : nullptr, lst,
                       off);
        mark(K_DENSITY, 0);
        hip_check(launch_density(dp_, bpos, particles.A.vel, lst, off, n, H0, rho_w,
                                 tape ? tape->pw : nullptr, particles.flux(), nbm, nrec,
                                 tape ? tape->rho_raw : nullptr, tape ? tape->vel : nullptr,
                                 particles.vpad(), st_),
                  "density");
        mark(K_DENSITY, 1);
        // the same launch sums the particle flux per column when coupled, and
        // hashes the new positions into the next step's grid
        mark(K_FORCE, 0);
        hip_check(launch_force_integrate(dp_, dt, bpos, particles.bid(), particles.A, lst, off, n,
                                         rho_w, nbm, nrec, H0, particles.N, particles.rho(),
                                         particles.prs(), tape ? tape->fpre : nullptr,
                                         particles.flux(), couple ? cflux : nullptr,
                                         particles.vpad(), grid.FusedNext(), st_),
                  "force/integrate");
        mark(K_FORCE, 1);
        particles.Flip();
        grid.KeysFused();
    }
    mark(K_WAVE, 0);
    if (taped_heights)
        hip_check(launch_wave(dp_, tape->h_in0, tape->h_in1, tape->h_out, couple ? cflux : nullptr,
                              dt, st_),
                  "wave");
    else
        wave.Evolve(dp_, couple ? cflux : nullptr, dt, st_);
    mark(K_WAVE, 1);
    wave.PingPong();
    // the scene's own neighbour records and sorted copy now describe this step
    // (a taped step writes them into the tape instead)
    stepped_ = n > 0 && !tape;
    // a taped step replays one the forward already checked
    if (params_.check_finite && !ev && !tape) CheckFinite();
}

void CoupledScene::CheckFinite() {
    hip_check(launch_fill_i32(flag_.get(), 1, 0, st_), "flag");
    const int n = particles.n();
    if (n > 0) {
        hip_check(launch_check_finite(reinterpret_cast<const float*>(particles.A.pos),
                                      4 * (size_t)n, flag_.get(), st_),
                  "finite pos");
        hip_check(launch_check_finite(reinterpret_cast<const float*>(particles.A.vel),
                                      4 * (size_t)n, flag_.get(), st_),
                  "finite vel");
    }
    hip_check(launch_check_finite(wave.GetReadImage(0), (size_t)wave.width() * wave.height(),
                                  flag_.get(), st_),
              "finite height");
    int h = 0;
    hip_check(hipMemcpyAsync(&h, flag_.get(), sizeof(int), hipMemcpyDeviceToHost, st_), "flag");
    hip_check(hipStreamSynchronize(st_), "flag sync");
    if (h) throw Error(CWA_ENONFINITE, "non-finite particle or height value after step");
}

// Records kGraphCycle steps into gexec_ without running them.  The host-side
// rotations have period kGraphCycle, so the capture leaves them where it
// found them, and the graph replays from this phase.
void CoupledScene::CaptureGraph(float dt) {
    invalidate_graph();
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "capture");
    try {
        for (int k = 0; k < kGraphCycle; ++k) Step(dt);
    } catch (...) {
        (void)hipStreamEndCapture(st_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(st_, &g), "end capture");
    const hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(e, "graph instantiate");
    graph_dt_ = dt;
    graph_phase_ = phase();
}

bool CoupledScene::PrepareGraph(float dt) {
    if (params_.check_finite) return false;  // StepGraph runs eagerly then
    // the captured steps take their keys from the step before them
    if (particles.n() > 0 && !grid.keys_ready()) return false;
    if (gexec_ && graph_dt_ == dt && graph_phase_ == phase()) return true;
    const bool stepped = stepped_;  // nothing runs: the last step's records stay current
    CaptureGraph(dt);
    stepped_ = stepped;
    return true;
}

void CoupledScene::StepGraph(float dt, int nsteps) {
    if (params_.check_finite) {  // the finite check synchronises: run eagerly
        for (int k = 0; k < nsteps; ++k) Step(dt);
        return;
    }
    const int cycle = kGraphCycle;
    int done = 0;
    // A captured step takes its keys from the step before; after a host-side
    // change of the particles the first step hashes them eagerly.
    if (nsteps > 0 && particles.n() > 0 && !grid.keys_ready()) {
        Step(dt);
        ++done;
    }
    // A captured graph replays the buffers of the phase it was captured at:
    // step eagerly until the rotation phase matches it again.
    if (gexec_ && graph_dt_ == dt) {
        while (done < nsteps && phase() != graph_phase_) {
            Step(dt);
            ++done;
        }
    }
    // A capture starts with the keys hashed ahead by the last force pass, so
    // that no captured step has a hash pass of its own and every replay can
    // follow the one before.
    if (nsteps - done > cycle && (!gexec_ || graph_dt_ != dt) && particles.n() > 0 &&
        !grid.keys_ready()) {
        Step(dt);
        ++done;
    }
    if (nsteps - done >= cycle && (!gexec_ || graph_dt_ != dt) &&
        (particles.n() == 0 || grid.keys_ready()))
        CaptureGraph(dt);
    while (gexec_ && graph_dt_ == dt && phase() == graph_phase_ && nsteps - done >= cycle) {
        hip_check(hipGraphLaunch(gexec_, st_), "graph launch");
        done += cycle;
    }
    for (; done < nsteps; ++done) Step(dt);
}

void CoupledScene::WaveOnly(float dt) {
    wave.Evolve(dp_, nullptr, dt, st_);
    wave.PingPong();
}

void CoupledScene::WaveReinit() {
    invalidate_graph();
    wave.ReinitMode(params_.wave_type, st_);
    hip_check(hipStreamSynchronize(st_), "sync");
}

void CoupledScene::WaveReinitFromTexture(const float* tex, int tw, int th) {
    invalidate_graph();
    hip_check(hipStreamSynchronize(st_), "sync");
    wave.ReinitFromTexture(tex, tw, th, st_);
}

void CoupledScene::NeighbourCounts(int32_t* out) {
    const int n = particles.n();
    if (n == 0) return;
    if (!stepped_)
        throw Error(CWA_ESTATE,
                    "neighbour counts exist after a step, until the next reinit or set_params");
    DeviceBuffer<int> d(n);
    hip_check(launch_neighbour_counts(dp_, particles.bpos(), grid.offset(), particles.nbw(), n,
                                      d.get(), st_),
              "neighbour counts");
    hip_check(hipStreamSynchronize(st_), "sync");
    std::vector<int> cnt(n), id(n);
    hip_check(hipMemcpy(cnt.data(), d.get(), n * sizeof(int), hipMemcpyDeviceToHost), "counts");
    hip_check(hipMemcpy(id.data(), particles.A.id, n * sizeof(int), hipMemcpyDeviceToHost), "id");
    for (int s = 0; s < n; ++s) out[id[s]] = cnt[s];
}

void CoupledScene::Readback(cwa_host_state* h) {
    const int n = particles.n();
    hip_check(hipStreamSynchronize(st_), "sync");
    if (n > 0) {
        std::vector<int> id(n);
        hip_check(hipMemcpy(id.data(), particles.A.id, n * sizeof(int), hipMemcpyDeviceToHost),
                  "id");
        auto un4 = [&](const float4* d, float* out) {
            if (!out) return;
            std::vector<float4> t(n);
            hip_check(hipMemcpy(t.data(), d, n * sizeof(float4), hipMemcpyDeviceToHost), "rb4");
            for (int s = 0; s < n; ++s) std::memcpy(out + 4 * (size_t)id[s], &t[s], sizeof(float4));
        };
        auto un1 = [&](const float* d, float* out) {
            if (!out) return;
            std::vector<float> t(n);
            hip_check(hipMemcpy(t.data(), d, n * sizeof(float), hipMemcpyDeviceToHost), "rb1");
            for (int s = 0; s < n; ++s) out[id[s]] = t[s];
        };
        un4(particles.A.pos, h->pos);
        un4(particles.A.vel, h->vel);
        un4(particles.A.force, h->force);
        un1(particles.rho(), h->rho);
        un1(particles.prs(), h->prs);
    }
    const size_t bytes = (size_t)wave.width() * wave.height() * sizeof(float);
    if (h->height0)
        hip_check(hipMemcpy(h->height0, wave.GetReadImage(0), bytes, hipMemcpyDeviceToHost), "h0");
    if (h->height1)
        hip_check(hipMemcpy(h->height1, wave.GetReadImage(1), bytes, hipMemcpyDeviceToHost), "h1");
}

}  // namespace cwa

// ===========================================================================
// Scratch of the allocation-free grid build (cwa_grid_build_static): run
// records for up to kStaticMaxItems items, fill counters for up to
// kStaticMaxCells cells and the cell scan's workspace, in one file-scope
// device array that every GPU running the library holds until the process
// ends (32 MB).  Host code never passes the array's name to a kernel: its
// device address, looked up once per device, is what the kernels get.
// ===========================================================================
namespace cwa {
constexpr int kStaticMaxItems = 1 << 21;
constexpr int kStaticMaxCells = 1 << 22;
constexpr size_t kStaticRunsBytes = (size_t)kStaticMaxItems * sizeof(int2);
constexpr size_t kStaticFillBytes = (size_t)kStaticMaxCells * sizeof(int);
constexpr size_t kStaticWsInts = 8192;  // scan_workspace_ints(kStaticMaxCells) is 4099
}  // namespace cwa

static __device__ __align__(256) unsigned char scratch[cwa::kStaticRunsBytes +
                                                       cwa::kStaticFillBytes +
                                                       cwa::kStaticWsInts * sizeof(int)];

namespace cwa {
// Device address of `scratch` on the current device (looked up on the first
// call per device, under a lock).
static unsigned char* static_scratch() {
    static std::mutex mu;
    static std::vector<unsigned char*> per_device;
    int dev = -1;
    hip_check(hipGetDevice(&dev), "current device");
    std::lock_guard<std::mutex> lock(mu);
    if ((int)per_device.size() <= dev) per_device.resize(dev + 1, nullptr);
    if (!per_device[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
        if (!p) throw Error(CWA_EHIP, "scratch address is null");
        per_device[dev] = static_cast<unsigned char*>(p);
    }
    return per_device[dev];
}
}  // namespace cwa

// ===========================================================================
// C ABI
// ===========================================================================
using namespace cwa;

struct cwa_scene {
    CoupledScene* impl;
};

static thread_local std::string g_last_error;

template <class F>
static int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return CWA_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return CWA_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CWA_EINVAL;
    }
}

namespace cwa {
CoupledScene* scene_impl(cwa_scene* s) {
    if (!s) throw Error(CWA_EINVAL, "null scene");
    return s->impl;
}
int guarded_call(void (*fn)(void*), void* ctx) {
    return guarded([&] { fn(ctx); });
}
}  // namespace cwa

#define CWA_REQUIRE(cond, msg)                  \
    do {                                        \
        if (!(cond)) throw Error(CWA_EINVAL, msg); \
    } while (0)

extern "C" {

int cwa_abi_version(void) { return CWA_ABI_VERSION; }

int cwa_struct_size(int which) {
    switch (which) {
        case 0: return (int)sizeof(cwa_params);
        case 1: return (int)sizeof(cwa_init);
        case 2: return (int)sizeof(cwa_particle_view);
        case 3: return (int)sizeof(cwa_grid_view);
        case 4: return (int)sizeof(cwa_host_state);
        case 5: return (int)sizeof(cwa_kernel_times);
        case 6: return (int)sizeof(cwa_grad);
        default: return -1;
    }
}

const char* cwa_last_error(void) { return g_last_error.c_str(); }

int cwa_device_count(void) {
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) {
        (void)hipGetLastError();
        return 0;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g_last_error = hipGetErrorString(e);
        return CWA_EHIP;
    }
    return n;
}

const char* cwa_kernel_name(int k) { return kernel_slot_name(k); }

int cwa_set_lane_split_limit(long long max_n, long long* previous) {
    const long long old = set_lane_split_limit(max_n < 0 ? kLaneSplitDefault : max_n);
    if (previous) *previous = old;
    return CWA_OK;
}

void cwa_params_reference(cwa_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    // ConstantsUniform (Main.cpp:184-190)
    p->mass = 0.02f;
    p->h = 2.0f * 0.005f;  // smoothing_coeff * PARTICLE_RADIUS
    p->visc = 3000.0f;
    p->rest_rho = 1000.0f;
    p->gas_const = 4000.0f;       // rho_pres_comp.glsl:41
    p->particle_radius = 0.005f;  // PARTICLE_RADIUS
    p->spiky_coeff = 20.0f;       // force_comp.glsl:68
    p->visc_coeff = 20.0f;        // laplacian = -spiky, force_comp.glsl:69
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9806.65f;  // force_comp.glsl:52
    p->gravity[2] = 0.0f;
    p->crest_threshold = 0.01f;  // force_comp.glsl:5
    p->breaking_mass = 0.25f;
    p->breaking_visc = 0.5f;
    p->torque_coeff = 0.25f;
    p->drag_coeff = 0.25f;
    p->wave_force_coeff = 0.5f;
    p->wave_vel_du = 0.01f;
    p->vel_damping = 0.3f;  // integrate_comp.glsl:8
    p->bounce = 0.3f;
    p->foam_speed = 25.0f;
    p->foam_force = 0.5f;
    p->foam_rho = 0.1f;
    p->foam_pressure = 0.25f;
    p->foam_vel = 0.1f;
    // BoundaryUniform (Main.cpp:192-198)
    p->upper[0] = 0.48f;
    p->upper[1] = 1.0f;
    p->upper[2] = 0.48f;
    p->lower[0] = 0.0f;
    p->lower[1] = -0.02f;
    p->lower[2] = 0.0f;
    p->ceiling_extra = 0.1f;
    p->wave_origin[0] = 0.0f;
    p->wave_origin[1] = 0.0f;
    p->wave_uv_scale[0] = 2.0f;  // coord = 2*pos.xz
    p->wave_uv_scale[1] = 2.0f;
    // WaveUniforms (Main.cpp:200-204)
    p->lambda_x = 0.01f;
    p->lambda_z = 0.01f;
    p->atten = 0.985f;
    p->beta = 0.001f;
    p->wave_type = 1.0f;
    p->wake_threshold = 0.0001f;
    p->wake_amount = 0.001f;
    p->couple_band = 0.0f;
    p->couple_strength = 0.0f;
    // grid over the box [lower, upper], cell = h
    p->cell_size = p->h;
    for (int a = 0; a < 3; ++a) {
        p->grid_min[a] = p->lower[a];
        p->grid_dims[a] =
            std::max(1, (int)std::ceil((p->upper[a] - p->lower[a]) / p->cell_size - 1e-4f));
    }
    p->grid_ydiv = 1;
    p->check_finite = 0;
}

int cwa_scene_create(const cwa_params* p, const cwa_init* init, cwa_scene** out) {
    return guarded([&] {
        CWA_REQUIRE(p && init && out, "null argument");
        *out = nullptr;
        auto* s = new cwa_scene{nullptr};
        try {
            s->impl = new CoupledScene(*p, *init);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int cwa_scene_reinit(cwa_scene* s, const cwa_init* init) {
    return guarded([&] {
        CWA_REQUIRE(s && init, "null argument");
        s->impl->Reinit(*init);
    });
}

int cwa_scene_destroy(cwa_scene* s) {
    return guarded([&] {
        if (!s) return;
        delete s->impl;
        delete s;
    });
}

int cwa_scene_set_params(cwa_scene* s, const cwa_params* p) {
    return guarded([&] {
        CWA_REQUIRE(s && p, "null argument");
        s->impl->SetParams(*p);
    });
}

int cwa_scene_get_params(const cwa_scene* s, cwa_params* p) {
    return guarded([&] {
        CWA_REQUIRE(s && p, "null argument");
        *p = s->impl->params();
    });
}

int cwa_scene_wave_reinit(cwa_scene* s) {
    return guarded([&] {
        CWA_REQUIRE(s, "null argument");
        s->impl->WaveReinit();
    });
}

int cwa_scene_wave_reinit_from_texture(cwa_scene* s, const float* tex, int tex_w, int tex_h) {
    return guarded([&] {
        CWA_REQUIRE(s && tex, "null argument");
        s->impl->WaveReinitFromTexture(tex, tex_w, tex_h);
    });
}

int cwa_scene_wave_compute(cwa_scene* s, float dt) {
    return guarded([&] {
        CWA_REQUIRE(s, "null argument");
        s->impl->WaveOnly(dt);
    });
}

int cwa_scene_height(const cwa_scene* s, int age, float** dptr, int* w, int* h, int* pitch) {
    return guarded([&] {
        CWA_REQUIRE(s && dptr, "null argument");
        CWA_REQUIRE(age >= 0 && age <= 2, "age must be 0, 1 or 2");
        const auto& wv = s->impl->wave;
        *dptr = age == 2 ? wv.GetWriteImage() : wv.GetReadImage(age);
        if (w) *w = wv.width();
        if (h) *h = wv.height();
        if (pitch) *pitch = wv.width();
    });
}

int cwa_scene_step(cwa_scene* s, float dt, int nsteps) {
    return guarded([&] {
        CWA_REQUIRE(s, "null argument");
        CWA_REQUIRE(nsteps >= 0, "nsteps < 0");
        CWA_REQUIRE(std::isfinite(dt), "dt must be finite");
        for (int k = 0; k < nsteps; ++k) s->impl->Step(dt);
    });
}

int cwa_scene_step_graph(cwa_scene* s, float dt, int nsteps) {
    return guarded([&] {
        CWA_REQUIRE(s, "null argument");
        CWA_REQUIRE(nsteps >= 0, "nsteps < 0");
        CWA_REQUIRE(std::isfinite(dt), "dt must be finite");
        s->impl->StepGraph(dt, nsteps);
    });
}

int cwa_scene_prepare_graph(cwa_scene* s, float dt, int* captured) {
    return guarded([&] {
        CWA_REQUIRE(s, "null argument");
        CWA_REQUIRE(std::isfinite(dt), "dt must be finite");
        const bool c = s->impl->PrepareGraph(dt);
        if (captured) *captured = c ? 1 : 0;
    });
}

int cwa_scene_step_profiled(cwa_scene* s, float dt, int nsteps, cwa_kernel_times* out) {
    return guarded([&] {
        CWA_REQUIRE(s && out, "null argument");
        CWA_REQUIRE(nsteps >= 0, "nsteps < 0");
        std::memset(out, 0, sizeof(*out));
        hipEvent_t ev[K_NUM][2];
        for (auto& e : ev)
            for (auto& x : e) hip_check(hipEventCreate(&x), "event create");
        struct Cleanup {
            hipEvent_t (*e)[2];
            ~Cleanup() {
                for (int k = 0; k < K_NUM; ++k)
                    for (int q = 0; q < 2; ++q) (void)hipEventDestroy(e[k][q]);
            }
        } cleanup{ev};
        CoupledScene& S = *s->impl;
        for (int k = 0; k < nsteps; ++k) {
            S.Step(dt, ev);
            hip_check(hipStreamSynchronize(S.stream()), "sync");
            for (int slot = 0; slot < K_NUM; ++slot) {
                const bool ran = slot == K_WAVE ||
                                 (S.particles.n() > 0 && slot >= K_HASH && slot <= K_FORCE);
                if (!ran) continue;
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, ev[slot][0], ev[slot][1]), "elapsed");
                out->ms[slot] += ms;
                out->launches[slot] += 1;
            }
        }
        for (int slot = 0; slot < K_NUM; ++slot)
            if (out->launches[slot]) out->ms[slot] /= (float)out->launches[slot];
        out->steps = nsteps;
    });
}

int cwa_scene_sync(cwa_scene* s) {
    return guarded([&] {
        CWA_REQUIRE(s, "null argument");
        hip_check(hipStreamSynchronize(s->impl->stream()), "sync");
    });
}

void* cwa_scene_stream(const cwa_scene* s) { return s ? (void*)s->impl->stream() : nullptr; }

int cwa_scene_particles(const cwa_scene* s, cwa_particle_view* v) {
    return guarded([&] {
        CWA_REQUIRE(s && v, "null argument");
        const auto& P = s->impl->particles;
        v->n = P.n();
        v->pos = reinterpret_cast<float*>(P.A.pos);
        v->vel = reinterpret_cast<float*>(P.A.vel);
        v->force = reinterpret_cast<float*>(P.A.force);
        v->rho = P.rho();
        v->prs = P.prs();
        v->id = P.A.id;
    });
}

int cwa_scene_particles_written(cwa_scene* s) {
    return guarded([&] {
        CWA_REQUIRE(s, "null argument");
        s->impl->ParticlesWritten();
    });
}

int cwa_scene_grid(const cwa_scene* s, cwa_grid_view* v) {
    return guarded([&] {
        CWA_REQUIRE(s && v, "null argument");
        const auto& G = s->impl->grid;
        for (int a = 0; a < 3; ++a) v->dims[a] = G.desc().dims[a];
        v->num_cells = G.num_cells();
        v->n = s->impl->particles.n();
        v->key = G.key();
        v->count = G.count();
        v->offset = G.offset();
        v->list = G.list();
        v->flux = s->impl->colflux.get();
    });
}

int cwa_scene_neighbour_counts(cwa_scene* s, int32_t* counts) {
    return guarded([&] {
        CWA_REQUIRE(s && counts, "null argument");
        s->impl->NeighbourCounts(counts);
    });
}

int cwa_scene_readback(cwa_scene* s, cwa_host_state* h) {
    return guarded([&] {
        CWA_REQUIRE(s && h, "null argument");
        s->impl->Readback(h);
    });
}

int cwa_scan_exclusive_i32(const int32_t* d_in, int32_t* d_out, int n, void* stream) {
    return guarded([&] {
        CWA_REQUIRE(n >= 0, "n < 0");
        CWA_REQUIRE(d_out && (n == 0 || d_in), "null argument");
        hipStream_t st = static_cast<hipStream_t>(stream);
        DeviceBuffer<int> ws(scan_workspace_ints(n));
        hip_check(launch_scan(d_in, d_out, n, ws.get(), st), "scan");
        hip_check(hipStreamSynchronize(st), "scan sync");
    });
}

int cwa_grid_build(const float* d_pos4, int n, const float grid_min[3], float cell_size,
                   const int32_t dims[3], int32_t* d_key, int32_t* d_count, int32_t* d_offset,
                   int32_t* d_list, void* stream) {
    return guarded([&] {
        CWA_REQUIRE(n >= 0, "n < 0");
        CWA_REQUIRE(grid_min && dims && d_count && d_offset, "null argument");
        CWA_REQUIRE(n == 0 || (d_pos4 && d_key && d_list), "null argument");
        CWA_REQUIRE(cell_size > 0.0f, "cell_size must be > 0");
        CWA_REQUIRE(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, "dims must be >= 1");
        CWA_REQUIRE((double)dims[0] * dims[1] * dims[2] <= (double)(1 << 30), "too many cells");
        hipStream_t st = static_cast<hipStream_t>(stream);
        GridDesc g{};
        for (int a = 0; a < 3; ++a) {
            g.gmin[a] = grid_min[a];
            g.dims[a] = dims[a];
        }
        g.cell = cell_size;
        g.cell_y = cell_size;
        const int nc = dims[0] * dims[1] * dims[2];
        DeviceBuffer<int> fill(nc), ws(scan_workspace_ints(nc));
        DeviceBuffer<int2> runs((size_t)std::max(n, 1));
        hip_check(launch_fill_i32(d_count, nc, 0, st), "clear");
        hip_check(launch_fill_i32(fill.get(), nc, 0, st), "clear fill");
        hip_check(launch_hash_count(reinterpret_cast<const float4*>(d_pos4), n, g, d_key, d_count,
                                    st),
                  "hash");
        hip_check(launch_scan(d_count, d_offset, nc, ws.get(), st), "scan");
        hip_check(launch_scatter_runs(d_key, d_offset, fill.get(), n, runs.get(), st), "scatter");
        hip_check(launch_stable_reorder(d_key, d_offset, runs.get(), n, nullptr, d_list, nullptr,
                                        nullptr, nullptr, nullptr, nullptr, 0, st),
                  "stable insert");
        hip_check(hipStreamSynchronize(st), "grid sync");
    });
}

int cwa_grid_build_static(const float* d_pos4, int n, const float grid_min[3], float cell_size,
                          const int32_t dims[3], int32_t* d_key, int32_t* d_count,
                          int32_t* d_offset, int32_t* d_list, void* stream) {
    return guarded([&] {
        CWA_REQUIRE(n >= 0, "n < 0");
        CWA_REQUIRE(grid_min && dims && d_count && d_offset, "null argument");
        CWA_REQUIRE(n == 0 || (d_pos4 && d_key && d_list), "null argument");
        CWA_REQUIRE(cell_size > 0.0f, "cell_size must be > 0");
        CWA_REQUIRE(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, "dims must be >= 1");
        CWA_REQUIRE(n <= kStaticMaxItems, "n exceeds the static scratch (2^21 particles)");
        CWA_REQUIRE((double)dims[0] * dims[1] * dims[2] <= (double)kStaticMaxCells,
                    "the grid exceeds the static scratch (2^22 cells)");
        const int nc = dims[0] * dims[1] * dims[2];
        if (scan_workspace_ints(nc) > kStaticWsInts)
            throw Error(CWA_EINVAL, "scan workspace exceeds the static scratch");
        hipStream_t st = static_cast<hipStream_t>(stream);
        GridDesc g{};
        for (int a = 0; a < 3; ++a) {
            g.gmin[a] = grid_min[a];
            g.dims[a] = dims[a];
        }
        g.cell = cell_size;
        g.cell_y = cell_size;
        unsigned char* base = static_scratch();
        int2* runs = reinterpret_cast<int2*>(base);
        int* fill = reinterpret_cast<int*>(base + kStaticRunsBytes);
        int* ws = reinterpret_cast<int*>(base + kStaticRunsBytes + kStaticFillBytes);
        // the array keeps the last call's contents: clear what must start at zero
        hip_check(launch_fill_i32(d_count, nc, 0, st), "clear");
        hip_check(launch_fill_i32(fill, nc, 0, st), "clear fill");
        hip_check(launch_hash_count(reinterpret_cast<const float4*>(d_pos4), n, g, d_key, d_count,
                                    st),
                  "hash");
        hip_check(launch_scan(d_count, d_offset, nc, ws, st), "scan");
        hip_check(launch_scatter_runs(d_key, d_offset, fill, n, runs, st), "scatter");
        hip_check(launch_stable_reorder(d_key, d_offset, runs, n, nullptr, d_list, nullptr,
                                        nullptr, nullptr, nullptr, nullptr, 0, st),
                  "stable insert");
    });
}

int cwa_grid_build_boxes(const float* d_centre4, const float* d_halfw4, int n,
                         const float grid_min[3], const float cell_size[3],
                         const int32_t dims[3], int32_t* d_count, int32_t* d_offset,
                         int32_t* d_list, int64_t list_capacity, int64_t* total,
                         void* stream) {
    return guarded([&] {
        CWA_REQUIRE(n >= 0, "n < 0");
        CWA_REQUIRE(grid_min && cell_size && dims && d_count && d_offset && total,
                    "null argument");
        CWA_REQUIRE(n == 0 || (d_centre4 && d_halfw4), "null argument");
        CWA_REQUIRE(list_capacity >= 0 && (list_capacity == 0 || d_list), "null list");
        for (int a = 0; a < 3; ++a) {
            CWA_REQUIRE(cell_size[a] > 0.0f && std::isfinite(cell_size[a]),
                        "cell_size must be > 0");
            CWA_REQUIRE(dims[a] >= 1, "dims must be >= 1");
        }
        CWA_REQUIRE((double)dims[0] * dims[1] * dims[2] <= (double)(1 << 30), "too many cells");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int nc = dims[0] * dims[1] * dims[2];
        const float4* x = reinterpret_cast<const float4*>(d_centre4);
        const float4* hw = reinterpret_cast<const float4*>(d_halfw4);
        // cells per box, then the pair index of each box's first pair
        DeviceBuffer<int> ncov((size_t)std::max(n, 1)), first((size_t)n + 1),
            ws(scan_workspace_ints(std::max(n, nc)));
        DeviceBuffer<unsigned long long> sum(1);
        hip_check(launch_box_cover(x, hw, n, grid_min, cell_size, dims, ncov.get(), st), "cover");
        hip_check(launch_sum_i64(ncov.get(), n, sum.get(), st), "pair total");
        unsigned long long npairs64 = 0;
        hip_check(hipMemcpyAsync(&npairs64, sum.get(), sizeof(npairs64), hipMemcpyDeviceToHost,
                                 st),
                  "pair count");
        hip_check(hipStreamSynchronize(st), "pair count sync");
        *total = (int64_t)npairs64;
        CWA_REQUIRE(npairs64 <= (unsigned long long)CWA_MAX_PARTICLES,
                    "more (box, cell) pairs than CWA_MAX_PARTICLES (returned in *total)");
        CWA_REQUIRE((int64_t)npairs64 <= list_capacity,
                    "list_capacity is below the number of (box, cell) pairs (returned in "
                    "*total)");
        const int npairs = (int)npairs64;
        hip_check(launch_scan(ncov.get(), first.get(), n, ws.get(), st), "cover scan");
        const size_t np = (size_t)std::max(npairs, 1);
        DeviceBuffer<int> pkey(np), pbox(np), fill(nc);
        DeviceBuffer<int2> runs(np);
        hip_check(launch_box_pairs(x, hw, n, grid_min, cell_size, dims, first.get(), pkey.get(),
                                   pbox.get(), st),
                  "pairs");
        hip_check(launch_fill_i32(d_count, nc, 0, st), "clear");
        hip_check(launch_fill_i32(fill.get(), nc, 0, st), "clear fill");
        hip_check(launch_key_count(pkey.get(), npairs, d_count, st), "count");
        hip_check(launch_scan(d_count, d_offset, nc, ws.get(), st), "scan");
        hip_check(launch_scatter_runs(pkey.get(), d_offset, fill.get(), npairs, runs.get(), st),
                  "scatter");
        hip_check(launch_stable_reorder(pkey.get(), d_offset, runs.get(), npairs, pbox.get(),
                                        d_list, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                        st),
                  "stable insert");
        hip_check(hipStreamSynchronize(st), "box grid sync");
    });
}

}  // extern "C"
