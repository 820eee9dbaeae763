// cwa_modules.hpp — host-side objects of the coupled step, shaped after the
// reference's Module framework:
//   Module                       Module.h:6-32 (virtual Init/Reinit/Compute,
//                                static registry, sComputeAll)
//   ParallelScan                 SphWave2D/ParallelScan.h (Init(input, n),
//                                Compute())
//   UniformGrid3D                Ugrid3D / UgridParticles3D
//                                (UniformGridParticles3D.h:50-100, BuildGrid)
//   HeightField2DTripleBuffered  StencilImage2DTripleBuffered.h:8-41
//   SphParticles3D               the particle SSBO + rho_pres / force /
//                                integrate programs of Main.cpp
//   CoupledScene                 the idle() step (Main.cpp:530-562)
// All device memory is owned here (RAII); kernels are in cwa_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <list>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cwa.h"
#include "cwa_device.hpp"
#include "cwa_kernels.hpp"

namespace cwa {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

// Owning device allocation.
template <class T>
class DeviceBuffer {
   public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    void resize(size_t n) {
        if (n == n_) return;
        release();
        if (n) {
            void* q = nullptr;
            const hipError_t e = hipMalloc(&q, n * sizeof(T));
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(CWA_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
            }
            p_ = static_cast<T*>(q);
        }
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    void swap(DeviceBuffer& o) noexcept {
        T* p = p_;
        p_ = o.p_;
        o.p_ = p;
        const size_t n = n_;
        n_ = o.n_;
        o.n_ = n;
    }

   private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Module.h:6-32.  Construction registers, destruction unregisters;
// sComputeAll runs Compute() in construction order.  The registry is
// process-wide like the reference's; it is not thread-safe.
class Module {
   public:
    Module();
    virtual ~Module();
    Module(const Module&) = delete;
    Module& operator=(const Module&) = delete;
    virtual void Init() {}
    virtual void Reinit() {}
    virtual void Compute() {}
    static std::list<Module*>& sAllModules();
    static void sInitAll();
    static void sReinitAll();
    static void sComputeAll();
};

// SphWave2D/ParallelScan.h: exclusive scan of an int buffer of n values
// into an owned output of n+1 values (any n >= 0).
class ParallelScan {
   public:
    void Init(const int* input, int n);
    void Compute(hipStream_t st);
    int* GetOutput() const { return out_.get(); }

   private:
    const int* in_ = nullptr;
    int n_ = 0;
    DeviceBuffer<int> out_, ws_;
};

// Ugrid3D / UgridParticles3D: per-step counting sort of particle slots by
// cell with a stable (slot-ordered) insertion, moving the positions and ids
// into sorted order.  The cell keys of a build come from its own hash pass,
// or from the force pass of the step before, which hashes the positions it
// writes (FusedNext / KeysFused); any change of the particles or counters
// between the two (ClearCounters) drops those keys.
class UniformGrid3D {
   public:
    void Init(const GridDesc& g, int max_particles);
    // BuildGrid (UniformGridParticles3D.cpp:170-211) over A.pos[0..n):
    // sorted positions / ids into bpos / bid.  list_out / offset_out
    // (optional) replace the grid's own list / offsets.
    void BuildGrid(const ParticleArrays& A, int n, float4* bpos, int* bid, hipStream_t st,
                   hipEvent_t* ev = nullptr, int* list_out = nullptr,
                   int* offset_out = nullptr);
    // Where the force pass of this step hashes the next step's keys.
    NextKeys FusedNext() const { return NextKeys{key_[kc_ ^ 1].get(), count_[cur_].get()}; }
    // The force pass has written the next step's keys and counts.
    void KeysFused() {
        kc_ ^= 1;
        keys_ready_ = true;
    }
    bool keys_ready() const { return keys_ready_; }
    const GridDesc& desc() const { return g_; }
    int num_cells() const { return g_.dims[0] * g_.dims[1] * g_.dims[2]; }
    // keys of the last build
    int* key() const { return key_[last_kc_].get(); }
    // counts of the last build (the other half of the double buffer is the
    // zeroed counter of the next build)
    int* count() const { return count_[last_].get(); }
    int* offset() const { return offset_.get(); }
    int* list() const { return list_.get(); }
    // Zero both counters (the counter of the next build must start at zero;
    // each build's scan clears the other one); drops keys hashed ahead.
    void ClearCounters(hipStream_t st);
    struct Parity {
        int cur, last, kc, last_kc, keys_ready;
    };
    int parity() const { return cur_ * 4 + kc_ * 2 + (keys_ready_ ? 1 : 0); }
    Parity get_parity() const { return Parity{cur_, last_, kc_, last_kc_, keys_ready_ ? 1 : 0}; }
    void set_parity(const Parity& p) {
        cur_ = p.cur;
        last_ = p.last;
        kc_ = p.kc;
        last_kc_ = p.last_kc;
        keys_ready_ = p.keys_ready != 0;
    }

   private:
    GridDesc g_{};
    DeviceBuffer<int> key_[2], count_[2], fill_, list_, offset_, ws_;
    DeviceBuffer<int2> runs_;
    int cur_ = 0, last_ = 0;       // counter double buffer
    int kc_ = 0, last_kc_ = 0;     // key double buffer
    bool keys_ready_ = false;      // key_[kc_] / count_[cur_] hold the next build's hash
};

// StencilImage2DTripleBuffered: three fp32 fields rotated by PingPong.
class HeightField2DTripleBuffered {
   public:
    static constexpr int MODE_INIT = CWA_MODE_INIT;
    static constexpr int MODE_INIT_FROM_TEXTURE = CWA_MODE_INIT_FROM_TEXTURE;
    static constexpr int MODE_EVOLVE = CWA_MODE_EVOLVE;

    void Allocate(int w, int h);
    // Reinit (.cpp:42-59): MODE_INIT twice with PingPong.
    void ReinitMode(float wave_type, hipStream_t st);
    // Upload both history levels (newest first).
    void Upload(const float* h0, const float* h1, hipStream_t st);
    // ReinitFromTexture (.cpp:61-77) on a host image.
    void ReinitFromTexture(const float* tex, int tex_w, int tex_h, hipStream_t st);
    // Compute (.cpp:79-95) without the PingPong.
    void Evolve(const DevParams& P, const float* colflux, float dt, hipStream_t st);
    // PingPong (.cpp:33-40): written image becomes read 0.
    void PingPong();
    // rotation state (which buffer is read 0 / read 1 / write), 0..2
    int phase() const { return read_[0]; }
    void get_rotation(int r[3]) const {
        r[0] = read_[0];
        r[1] = read_[1];
        r[2] = write_;
    }
    void set_rotation(const int r[3]) {
        read_[0] = r[0];
        read_[1] = r[1];
        write_ = r[2];
    }
    float* GetReadImage(int i) const { return img_[read_[i]].get(); }
    float* GetWriteImage() const { return img_[write_].get(); }
    int width() const { return w_; }
    int height() const { return h_; }

   private:
    int w_ = 0, h_ = 0;
    int read_[2] = {0, 1};
    int write_ = 2;
    DeviceBuffer<float> img_[3];
};

// The particle state in two copies (A: current, N: next; each step reads A
// and writes N in its sorted order, then they swap), the sorted positions
// and ids of the step (B), per-step work arrays and the visible density /
// pressure.
class SphParticles3D {
   public:
    void Allocate(int n);
    void Upload(const float* pos, const float* vel, const float* force, hipStream_t st);
    int n() const { return n_; }
    ParticleArrays A, N;
    float4* bpos() const { return bpos_.get(); }
    int* bid() const { return bid_.get(); }
    // N becomes the current state
    void Flip() {
        std::swap(A, N);
        parity_ ^= 1;
    }
    int parity() const { return parity_; }
    void set_parity(int p) {
        if (p != parity_) Flip();
    }
    float* rho() const { return rho_.get(); }
    float* prs() const { return prs_.get(); }
    float* rho_w() const { return rho_w_.get(); }
    float* flux() const { return flux_.get(); }
    float* vpad() const { return vpad_.get(); }
    uint2* nbw() const { return nbw_.get(); }
    float4* nrec() const { return nrec_.get(); }

   private:
    int n_ = 0;
    int parity_ = 0;
    DeviceBuffer<uint2> nbw_;  // neighbour words and counts of the last step
    DeviceBuffer<float4> nrec_;   // 2 x n neighbour records of the last step
    DeviceBuffer<float4> pos_[2], vel_[2], force_[2], bpos_;
    DeviceBuffer<int> id_[2], bid_;
    DeviceBuffer<float> rho_, prs_, rho_w_, flux_, vpad_;
};

// Optional per-step output locations used by the adjoint tape: when set,
// the step writes these arrays there instead of into the scene's own
// buffers (and reads them back from there within the step).
struct TapeOut {
    float* rho_raw = nullptr;  // density before the wave term and the clamp
    float4* fpre = nullptr;    // force before the foam branch
    float4* pos = nullptr;     // sorted state at the start of the step
    float4* vel = nullptr;
    int* list = nullptr;       // stable cell order of the step
    int* offset = nullptr;     // cell offsets of the step
    uint2* nbw = nullptr;      // neighbour words
    float4* nrec = nullptr;    // neighbour records
    float* rho = nullptr;      // density (after the clamp)
    float2* pw = nullptr;      // (pressure, 1/rho), written only for the tape
    float* colflux = nullptr;  // column flux
    // heightfields: when h_out is set the step reads h_in0 (newest) and
    // h_in1 (previous) instead of the scene's triple buffer and writes the
    // new heightfield to h_out (the scene's buffer is then left stale)
    const float* h_in0 = nullptr;
    const float* h_in1 = nullptr;
    float* h_out = nullptr;
};

// Kernel slots of cwa_kernel_times.
enum KernelSlot {
    K_CLEAR = 0,
    K_HASH,
    K_SCAN,
    K_SCATTER,
    K_REORDER,
    K_DENSITY,
    K_FORCE,
    K_COLFLUX,
    K_WAVE,
    K_FINITE,
    K_NUM
};
const char* kernel_slot_name(int k);

// The coupled scene is the Module: Compute() is one idle() step at the time
// step set by SetTimeStep (the reference bakes dt into the shaders).
class CoupledScene : public Module {
   public:
    CoupledScene(const cwa_params& p, const cwa_init& init);
    ~CoupledScene() override;
    using Module::Reinit;
    void Compute() override { Step(dt_); }
    void SetTimeStep(float dt) { dt_ = dt; }
    void Reinit(const cwa_init& init);
    void SetParams(const cwa_params& p);
    const cwa_params& params() const { return params_; }
    // idle() (Main.cpp:530-562).  events: optional [K_NUM][2] per step;
    // tape: optional extra outputs for the adjoint.
    void Step(float dt, hipEvent_t (*events)[2] = nullptr, const TapeOut* tape = nullptr);
    void StepGraph(float dt, int nsteps);
    // Capture StepGraph's replay graph for dt without running a step; false
    // (nothing captured) before the first step or with check_finite set.
    bool PrepareGraph(float dt);
    // Period of the host-side buffer rotations: the wave triple buffer (3) x
    // the double buffers (2) of the grid counters, the grid keys and the
    // particle state.
    static constexpr int kGraphCycle = 6;
    void WaveOnly(float dt);
    void WaveReinit();
    void WaveReinitFromTexture(const float* tex, int tw, int th);
    void CheckFinite();
    void Readback(cwa_host_state* h);
    // Neighbour count (self included) of every particle, original order, as
    // the last step's density pass found it.
    void NeighbourCounts(int32_t* out);
    // The caller wrote particle state through the device pointers: the cell
    // keys and counts the last step hashed ahead describe old positions, so
    // the next step hashes afresh.
    void ParticlesWritten();
    hipStream_t stream() const { return st_; }
    const DevParams& devparams() const { return dp_; }
    // Layout generation: advanced by every Reinit (particle count, heightfield
    // size) and by every SetParams that changes the grid, the sign of the
    // coupling band or the wake mode.  Objects that cache buffer sizes or
    // captured code paths of this scene (tapes, fits) record it and refuse
    // to run once it has moved.
    unsigned long long layout_generation() const { return layout_gen_; }
    // Host-side phase of the buffer rotations (wave triple buffer, grid
    // counter and key double buffers, particle state double buffer): in a
    // run of steps with fused keys the period is 6.
    int phase() const { return (wave.phase() * 8 + grid.parity()) * 2 + particles.parity(); }
    // The host-side rotation state itself: a captured sequence replays the
    // buffers of the rotation it was captured at, so its owner puts the
    // scene back into that rotation around replays.
    struct Rotation {
        int wave[3];
        UniformGrid3D::Parity grid;
        int particles;
    };
    Rotation rotation() const {
        Rotation r{};
        wave.get_rotation(r.wave);
        r.grid = grid.get_parity();
        r.particles = particles.parity();
        return r;
    }
    void set_rotation(const Rotation& r) {
        wave.set_rotation(r.wave);
        grid.set_parity(r.grid);
        particles.set_parity(r.particles);
    }

    SphParticles3D particles;
    UniformGrid3D grid;
    HeightField2DTripleBuffered wave;
    DeviceBuffer<float> colflux;

   private:
    void build_devparams();
    void invalidate_graph();
    void CaptureGraph(float dt);
    cwa_params params_{};
    DevParams dp_{};
    DeviceBuffer<DevParams> dpd_;  // device mirror read by the kernels
    hipStream_t st_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    float graph_dt_ = 0.0f;
    int graph_phase_ = -1;
    float dt_ = 0.0f;
    bool stepped_ = false;  // a step has run since the last (re)init
    unsigned long long layout_gen_ = 0;
    DeviceBuffer<int> flag_;
};

// Validation shared by create / set_params.
void validate_params(const cwa_params& p);
DevParams make_devparams(const cwa_params& p, int wave_w, int wave_h);

}  // namespace cwa
