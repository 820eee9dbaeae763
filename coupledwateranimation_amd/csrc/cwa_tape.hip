// cwa_tape.hip — checkpointed forward + adjoint sweep over many coupled
// steps (the gradient path), and its C ABI.
//
// Forward: the scene's state is saved every `ckpt_every` steps (particle
// SoA + id + the two newest heightfields) and once more at the end.
// Backward: for each segment, last to first, the segment's checkpoint is
// restored and its steps re-run with every intermediate the adjoint reads
// recorded per step (sorted state, permutation, offsets, neighbour words,
// densities, pressures, neighbour records, pre-foam force, heightfields,
// column flux); the adjoint then walks the segment's steps in reverse.
// The eager Backward pipelines the segments: the scene's stream re-runs
// segment s-1 into a second tape while a second stream runs the adjoint of
// segment s (BackwardSeeded, which a graph captures, keeps one stream).
// Memory: checkpoints of the whole run plus the tapes of one segment (two
// for the eager Backward).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <vector>

#include "../../include/cwa.h"
#include "cwa_adjoint.hpp"
#include "cwa_modules.hpp"

namespace cwa {

class Tape {
   public:
    // pipelined: allocate the second step tape, stream and events of the
    // eager Backward up front (tapes that only run BackwardSeeded pass false)
    Tape(CoupledScene& s, int nsteps, int every, bool pipelined);
    ~Tape();
    Tape(const Tape&) = delete;
    Tape& operator=(const Tape&) = delete;
    void Forward(float dt);
    void Backward(const float* dH0, const float* dH1, const float* dpos, const float* dvel,
                  double* grad, float* dpos0, float* dvel0);

    // Device-only passes (no host transfer or synchronisation, so a graph
    // can capture them): the forward from the saved initial state, the MSE
    // seed of the newest heightfield against a device target, and the
    // adjoint sweep leaving the parameter gradients in grad_dev().
    void SaveInitial();
    void ForwardFromInitial(float dt);
    void SeedMse(const float* target, double* partials, int nblocks, double* loss);
    void BackwardSeeded();
    const double* grad_dev() const { return pgrad_.get(); }
    int texels() const { return W_ * H_; }

   private:
    struct Ckpt {
        DeviceBuffer<float4> pos, vel, force;
        DeviceBuffer<int> id;
        DeviceBuffer<float> rho, prs, h0, h1;
    };
    struct StepTape {
        DeviceBuffer<float4> pos, vel, nrec, fpre;
        DeviceBuffer<int> list, offset;
        DeviceBuffer<uint2> nbw;
        DeviceBuffer<float> rho, rho_raw, hout, colflux;
        DeviceBuffer<float2> pw;
        // the two heightfields the step read: the segment's checkpoint or
        // earlier steps' hout (set while recording)
        const float* h0 = nullptr;
        const float* h1 = nullptr;
    };
    void alloc_step_tape(StepTape& t);
    void release_pipeline();
    void save(Ckpt& c);
    void restore(const Ckpt& c, bool heights = true);
    void record_step(std::vector<StepTape>& set, int k, const Ckpt& c);
    void adjoint_step(const StepTape& t, hipStream_t st);
    void reduce_partials(hipStream_t st);
    // The scene must still have the layout the tape was sized for.
    void check_layout() const {
        if (S_.layout_generation() != gen_)
            throw Error(CWA_ESTATE,
                        "the scene was reinitialised or its grid, coupling band or wave mode "
                        "changed after the tape was created");
    }
    template <class T>
    void copy(T* dst, const T* src, size_t count) {
        if (count)
            hip_check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, st_),
                      "tape copy");
    }

    CoupledScene& S_;
    hipStream_t st_;
    unsigned long long gen_;
    int nsteps_, every_, nseg_;
    int n_, W_, H_, ncell_, ncol_;
    int group_ = 1;  // lanes per particle of the pair adjoint, fixed at creation
    float dt_ = 0.0f;
    bool forward_done_ = false;
    std::vector<Ckpt> ck_;
    std::vector<StepTape> tp_, tp2_;  // tp2_: the eager Backward's second tape
    hipStream_t ast_ = nullptr;       // the eager Backward's adjoint stream
    hipEvent_t ev_[5] = {};           // recorded[2], adjoint done[2], setup / end
    DeviceBuffer<float4> xb_, vb_, xb2_, vb2_, xl_, vl_, adjrec_;
    DeviceBuffer<float> G0_, G1_, G0n_, G1n_, rhob_, rtb_, Sbar_;
    DeviceBuffer<double> pgrad_, acc_wave_, acc_pair_;
};

Tape::Tape(CoupledScene& s, int nsteps, int every, bool pipelined)
    : S_(s), st_(s.stream()), gen_(s.layout_generation()) {
    if (nsteps < 1) throw Error(CWA_EINVAL, "nsteps must be >= 1");
    if (every < 1) throw Error(CWA_EINVAL, "ckpt_every must be >= 1");
    if (s.params().torque_coeff != 0.0f)
        throw Error(CWA_EINVAL, "the adjoint requires torque_coeff == 0");
    nsteps_ = nsteps;
    every_ = every;
    nseg_ = (nsteps + every - 1) / every;
    n_ = s.particles.n();
    W_ = s.wave.width();
    H_ = s.wave.height();
    ncell_ = s.grid.num_cells();
    ncol_ = s.params().grid_dims[0] * s.params().grid_dims[2];
    const size_t n = (size_t)std::max(n_, 1), tex = (size_t)W_ * H_;
    ck_.resize(nseg_ + 1);
    for (auto& c : ck_) {
        c.pos.resize(n);
        c.vel.resize(n);
        c.force.resize(n);
        c.id.resize(n);
        c.rho.resize(n);
        c.prs.resize(n);
        c.h0.resize(tex);
        c.h1.resize(tex);
    }
    tp_.resize(std::min(every, nsteps));
    for (auto& t : tp_) alloc_step_tape(t);
    // the stream and events are created last, so an allocation failure
    // above leaves nothing behind (the buffers free themselves); the
    // constructor body after them releases them if it throws
    for (auto* b : {&xb_, &vb_, &xb2_, &vb2_, &xl_, &vl_}) b->resize(n);
    adjrec_.resize(2 * n);
    for (auto* b : {&G0_, &G1_, &G0n_, &G1n_}) b->resize(tex);
    rhob_.resize(n);
    rtb_.resize(n);
    Sbar_.resize(ncol_);
    pgrad_.resize(GRAD_NUM);
    acc_wave_.resize((size_t)adj_wave_blocks(s.devparams()) * 5);
    group_ = adj_pair_group(n_);
    acc_pair_.resize((size_t)adj_pair_blocks(std::max(n_, 1), group_) * 3);
    if (pipelined) {
        if (nseg_ > 1) {
            tp2_.resize(tp_.size());
            for (auto& t : tp2_) alloc_step_tape(t);
        }
        try {
            hip_check(hipStreamCreateWithFlags(&ast_, hipStreamNonBlocking), "adjoint stream");
            for (hipEvent_t& e : ev_)
                hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        } catch (...) {
            release_pipeline();
            throw;
        }
    }
}

Tape::~Tape() { release_pipeline(); }

void Tape::release_pipeline() {
    if (ast_) {
        (void)hipStreamSynchronize(ast_);
        (void)hipStreamDestroy(ast_);
        ast_ = nullptr;
    }
    for (hipEvent_t& e : ev_)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}

void Tape::alloc_step_tape(StepTape& t) {
    const size_t n = (size_t)std::max(n_, 1), tex = (size_t)W_ * H_;
    t.pos.resize(n);
    t.vel.resize(n);
    t.nrec.resize(2 * n);
    t.fpre.resize(n);
    t.list.resize(n);
    t.offset.resize((size_t)ncell_ + 1);
    t.nbw.resize(nb_buffer_uint2(n));
    t.rho.resize(n);
    t.rho_raw.resize(n);
    t.pw.resize(n);
    t.hout.resize(tex);
    t.colflux.resize(ncol_);
}

void Tape::save(Ckpt& c) {
    const auto& A = S_.particles.A;
    copy(c.pos.get(), A.pos, n_);
    copy(c.vel.get(), A.vel, n_);
    copy(c.force.get(), A.force, n_);
    copy(c.id.get(), A.id, n_);
    copy(c.rho.get(), S_.particles.rho(), n_);
    copy(c.prs.get(), S_.particles.prs(), n_);
    copy(c.h0.get(), S_.wave.GetReadImage(0), (size_t)W_ * H_);
    copy(c.h1.get(), S_.wave.GetReadImage(1), (size_t)W_ * H_);
}

void Tape::restore(const Ckpt& c, bool heights) {
    auto& A = S_.particles.A;
    copy(A.pos, c.pos.get(), n_);
    copy(A.vel, c.vel.get(), n_);
    copy(A.force, c.force.get(), n_);
    copy(A.id, c.id.get(), n_);
    copy(S_.particles.rho(), c.rho.get(), n_);
    copy(S_.particles.prs(), c.prs.get(), n_);
    if (heights) {
        copy(S_.wave.GetReadImage(0), c.h0.get(), (size_t)W_ * H_);
        copy(S_.wave.GetReadImage(1), c.h1.get(), (size_t)W_ * H_);
    }
    S_.grid.ClearCounters(st_);
}

void Tape::SaveInitial() {
    save(ck_[0]);
    hip_check(hipStreamSynchronize(st_), "tape initial state");
}

void Tape::ForwardFromInitial(float dt) {
    check_layout();
    dt_ = dt;
    restore(ck_[0]);
    int done = 0;
    for (int seg = 0; seg < nseg_; ++seg) {
        if (seg > 0) save(ck_[seg]);
        const int len = std::min(every_, nsteps_ - done);
        for (int k = 0; k < len; ++k) S_.Step(dt);
        done += len;
    }
    save(ck_[nseg_]);
    forward_done_ = true;
}

void Tape::SeedMse(const float* target, double* partials, int nblocks, double* loss) {
    const size_t n = (size_t)std::max(n_, 1);
    hip_check(launch_mse_seed(S_.wave.GetReadImage(0), target, W_ * H_, G0_.get(), partials,
                              nblocks, loss, st_),
              "mse seed");
    hip_check(launch_fill_i32(reinterpret_cast<int*>(G1_.get()), W_ * H_, 0, st_), "seed h1");
    hip_check(launch_fill_i32(reinterpret_cast<int*>(xb_.get()), (int)(4 * n), 0, st_), "seed x");
    hip_check(launch_fill_i32(reinterpret_cast<int*>(vb_.get()), (int)(4 * n), 0, st_), "seed v");
}

void Tape::BackwardSeeded() {
    check_layout();
    if (!forward_done_) throw Error(CWA_ESTATE, "tape backward before forward");
    auto zero = [&](DeviceBuffer<double>& b) {
        hip_check(launch_fill_i32(reinterpret_cast<int*>(b.get()), (int)(2 * b.size()), 0, st_),
                  "zero");
    };
    zero(pgrad_);
    zero(acc_wave_);
    zero(acc_pair_);
    for (int seg = nseg_ - 1; seg >= 0; --seg) {
        restore(ck_[seg], false);
        const int len = std::min(every_, nsteps_ - seg * every_);
        for (int k = 0; k < len; ++k) record_step(tp_, k, ck_[seg]);
        for (int k = len - 1; k >= 0; --k) adjoint_step(tp_[k], st_);
    }
    reduce_partials(st_);
    restore(ck_[nseg_]);
}

void Tape::reduce_partials(hipStream_t st) {
    const int wslots[5] = {GRAD_LAMBDA_X, GRAD_LAMBDA_Z, GRAD_ATTEN, GRAD_BETA, GRAD_KAPPA};
    const int pslots[3] = {GRAD_VISC, GRAD_GAS_CONST, GRAD_REST_RHO};
    hip_check(launch_reduce_partials(acc_wave_.get(), (int)(acc_wave_.size() / 5), 5, wslots,
                                     pgrad_.get(), st),
              "reduce");
    hip_check(launch_reduce_partials(acc_pair_.get(), (int)(acc_pair_.size() / 3), 3, pslots,
                                     pgrad_.get(), st),
              "reduce");
}

void Tape::Forward(float dt) {
    check_layout();
    dt_ = dt;
    int done = 0;
    for (int seg = 0; seg < nseg_; ++seg) {
        save(ck_[seg]);
        const int len = std::min(every_, nsteps_ - done);
        for (int k = 0; k < len; ++k) S_.Step(dt);
        done += len;
    }
    save(ck_[nseg_]);
    hip_check(hipStreamSynchronize(st_), "tape forward");
    forward_done_ = true;
}

// Step k of the segment replay from checkpoint c, recording what the
// adjoint reads: the step writes its intermediates straight into the tape
// slot, reads its two heightfields from the checkpoint (k = 0, 1) or the
// earlier slots' outputs, and writes the new heightfield into its own slot,
// so no heightfield is copied.
void Tape::record_step(std::vector<StepTape>& set, int k, const Ckpt& c) {
    StepTape& t = set[k];
    t.h0 = k == 0 ? c.h0.get() : set[k - 1].hout.get();
    t.h1 = k == 0 ? c.h1.get() : (k == 1 ? c.h0.get() : set[k - 2].hout.get());
    TapeOut out;
    out.h_in0 = t.h0;
    out.h_in1 = t.h1;
    out.h_out = t.hout.get();
    out.rho_raw = t.rho_raw.get();
    out.fpre = t.fpre.get();
    out.pos = t.pos.get();
    out.vel = t.vel.get();
    out.list = t.list.get();
    out.offset = t.offset.get();
    out.nbw = t.nbw.get();
    out.nrec = t.nrec.get();
    out.rho = t.rho.get();
    out.pw = t.pw.get();
    out.colflux = t.colflux.get();
    S_.Step(dt_, nullptr, &out);
}

void Tape::adjoint_step(const StepTape& t, hipStream_t st) {
    const DevParams& P = S_.devparams();
    const bool couple = n_ > 0 && P.band > 0.0f;
    hip_check(launch_adj_wave(P, dt_, t.h0, t.h1, couple ? t.colflux.get() : nullptr,
                              G0_.get(), G1_.get(), G0n_.get(), G1n_.get(), acc_wave_.get(),
                              couple ? Sbar_.get() : nullptr, st),
              "adj wave + column flux");
    if (n_ > 0) {
        hip_check(launch_adj_local(P, dt_, n_, t.pos.get(), t.vel.get(), t.fpre.get(),
                                   t.rho.get(), t.h0, xb_.get(), vb_.get(), G0n_.get(),
                                   adjrec_.get(), rhob_.get(), xl_.get(), vl_.get(), st),
                  "adj local");
        hip_check(launch_adj_pairs(P, dt_, n_, t.pos.get(), t.vel.get(), t.offset.get(),
                                   t.nbw.get(), t.nrec.get(), t.rho.get(), t.pw.get(),
                                   t.rho_raw.get(), t.h0, Sbar_.get(), adjrec_.get(),
                                   rhob_.get(), G0n_.get(), xl_.get(), vl_.get(), rtb_.get(),
                                   acc_pair_.get(), group_, st),
                  "adj pairs");
        hip_check(launch_adj_rho_pairs(P, n_, t.pos.get(), t.offset.get(), t.nbw.get(),
                                       rtb_.get(), xl_.get(), vl_.get(), t.list.get(),
                                       xb2_.get(), vb2_.get(), st),
                  "adj rho pairs");
        xb_.swap(xb2_);
        vb_.swap(vb2_);
    }
    G0_.swap(G0n_);
    G1_.swap(G1n_);
}

void Tape::Backward(const float* dH0, const float* dH1, const float* dpos, const float* dvel,
                    double* grad, float* dpos0, float* dvel0) {
    check_layout();
    if (!forward_done_) throw Error(CWA_ESTATE, "tape backward before forward");
    const size_t tex = (size_t)W_ * H_;
    // seeds: heights as given, particle seeds permuted into the final slots
    std::vector<int> id(std::max(n_, 1));
    if (n_ > 0)
        hip_check(hipMemcpy(id.data(), ck_[nseg_].id.get(), n_ * sizeof(int),
                            hipMemcpyDeviceToHost),
                  "ids");
    auto upload_particles = [&](const float* seed, DeviceBuffer<float4>& dst) {
        std::vector<float> h(4 * (size_t)std::max(n_, 1), 0.0f);
        if (seed)
            for (int s = 0; s < n_; ++s)
                for (int q = 0; q < 3; ++q) h[4 * (size_t)s + q] = seed[4 * (size_t)id[s] + q];
        if (n_ > 0)
            hip_check(hipMemcpy(dst.get(), h.data(), (size_t)n_ * sizeof(float4),
                                hipMemcpyHostToDevice),
                      "seed");
    };
    upload_particles(dpos, xb_);
    upload_particles(dvel, vb_);
    auto upload_field = [&](const float* seed, DeviceBuffer<float>& dst) {
        if (seed)
            hip_check(hipMemcpy(dst.get(), seed, tex * sizeof(float), hipMemcpyHostToDevice),
                      "seed field");
        else
            hip_check(launch_fill_i32(reinterpret_cast<int*>(dst.get()), (int)tex, 0, st_),
                      "zero field");
    };
    upload_field(dH0, G0_);
    upload_field(dH1, G1_);
    hip_check(hipMemsetAsync(pgrad_.get(), 0, GRAD_NUM * sizeof(double), st_), "grad zero");
    hip_check(hipMemsetAsync(acc_wave_.get(), 0, acc_wave_.size() * sizeof(double), st_), "acc");
    hip_check(hipMemsetAsync(acc_pair_.get(), 0, acc_pair_.size() * sizeof(double), st_), "acc");

    // Pipelined sweep: segment s is re-run on the scene's stream into tape
    // set s % 2 while the adjoint stream walks the previous segment's set;
    // a set is re-recorded only after the adjoint that read it is done.
    if (!ast_) throw Error(CWA_ESTATE, "tape created without the eager backward's resources");
    hip_check(hipEventRecord(ev_[4], st_), "event");
    hip_check(hipStreamWaitEvent(ast_, ev_[4], 0), "wait");
    for (int q = 0, seg = nseg_ - 1; seg >= 0; ++q, --seg) {
        std::vector<StepTape>& set = (q & 1) ? tp2_ : tp_;
        if (q >= 2) hip_check(hipStreamWaitEvent(st_, ev_[2 + (q & 1)], 0), "wait");
        restore(ck_[seg], false);
        const int len = std::min(every_, nsteps_ - seg * every_);
        for (int k = 0; k < len; ++k) record_step(set, k, ck_[seg]);
        hip_check(hipEventRecord(ev_[q & 1], st_), "event");
        hip_check(hipStreamWaitEvent(ast_, ev_[q & 1], 0), "wait");
        for (int k = len - 1; k >= 0; --k) adjoint_step(set[k], ast_);
        hip_check(hipEventRecord(ev_[2 + (q & 1)], ast_), "event");
    }
    reduce_partials(ast_);
    hip_check(hipEventRecord(ev_[4], ast_), "event");
    hip_check(hipStreamWaitEvent(st_, ev_[4], 0), "wait");
    restore(ck_[nseg_]);  // leave the scene in its final state
    hip_check(hipStreamSynchronize(st_), "backward");

    double g[GRAD_NUM];
    hip_check(hipMemcpy(g, pgrad_.get(), sizeof(g), hipMemcpyDeviceToHost), "grad");
    for (int k = 0; k < GRAD_NUM; ++k) grad[k] = g[k];
    // the adjoint state is now in the slot order of the initial state
    std::vector<int> id0(std::max(n_, 1));
    if (n_ > 0)
        hip_check(hipMemcpy(id0.data(), ck_[0].id.get(), n_ * sizeof(int),
                            hipMemcpyDeviceToHost),
                  "ids0");
    auto download = [&](const DeviceBuffer<float4>& src, float* out) {
        if (!out || n_ == 0) return;
        std::vector<float4> h(n_);
        hip_check(hipMemcpy(h.data(), src.get(), n_ * sizeof(float4), hipMemcpyDeviceToHost),
                  "grad particles");
        for (int s = 0; s < n_; ++s) {
            float* o = out + 4 * (size_t)id0[s];
            o[0] = h[s].x;
            o[1] = h[s].y;
            o[2] = h[s].z;
            o[3] = 0.0f;
        }
    };
    download(xb_, dpos0);
    download(vb_, dvel0);
}

// ---------------------------------------------------------------------------
// FitBatch: gradients of a heightfield fit for many scenes at once.  For
// every scene: the checkpointed forward from its saved initial state, the
// MSE of its newest heightfield against a target, and the adjoint sweep;
// all scenes are captured once into one hipGraph (each scene on its own
// stream, forked from and joined into one capture stream) and the graph is
// replayed per evaluation.  The kernels read their parameters from device
// memory, so cwa_scene_set_params between evaluations takes effect without
// a new capture (the grid and code paths must stay as captured).
// ---------------------------------------------------------------------------
class FitBatch {
   public:
    static constexpr int kMseBlocks = 256;

    FitBatch(std::vector<CoupledScene*> scenes, const float* const* targets, int nsteps,
             int every, float dt)
        : sc_(std::move(scenes)) {
        const int ns = (int)sc_.size();
        if (ns < 1) throw Error(CWA_EINVAL, "nscenes must be >= 1");
        if (!(dt > 0.0f)) throw Error(CWA_EINVAL, "dt must be positive");
        for (int s = 0; s < ns; ++s) {
            if (!sc_[s]) throw Error(CWA_EINVAL, "null scene");
            if (sc_[s]->params().check_finite)
                throw Error(CWA_EINVAL, "fit scenes must have check_finite == 0");
            if (!targets || !targets[s]) throw Error(CWA_EINVAL, "null target");
            for (int q = 0; q < s; ++q)
                if (sc_[q] == sc_[s]) throw Error(CWA_EINVAL, "a scene appears twice");
        }
        try {
            Build(targets, nsteps, every, dt);
        } catch (...) {
            release();
            throw;
        }
    }

    ~FitBatch() { release(); }

    void Run(double* losses, double* grads) {
        const int ns = (int)sc_.size();
        for (int s = 0; s < ns; ++s)
            if (sc_[s]->layout_generation() != gen_[s])
                throw Error(CWA_ESTATE,
                            "a fit scene was reinitialised or its grid, coupling band or wave "
                            "mode changed after the fit was created");
        hip_check(hipGraphLaunch(exec_, ms_), "fit graph");
        std::vector<double> g((size_t)ns * GRAD_NUM), l(ns);
        for (int s = 0; s < ns; ++s)
            hip_check(hipMemcpyAsync(g.data() + (size_t)s * GRAD_NUM, tapes_[s]->grad_dev(),
                                     sizeof(double) * GRAD_NUM, hipMemcpyDeviceToHost, ms_),
                      "grads");
        hip_check(hipMemcpyAsync(l.data(), loss_.get(), sizeof(double) * ns,
                                 hipMemcpyDeviceToHost, ms_),
                  "losses");
        hip_check(hipStreamSynchronize(ms_), "fit");
        // the device state now matches the rotation the capture ended in
        for (int s = 0; s < ns; ++s) sc_[s]->set_rotation(rot_end_[s]);
        if (losses) std::memcpy(losses, l.data(), sizeof(double) * ns);
        if (grads) std::memcpy(grads, g.data(), sizeof(double) * g.size());
    }

    int nscenes() const { return (int)sc_.size(); }
    long long TapeBytes(int s) const {
        if (s < 0 || s >= (int)sc_.size()) throw Error(CWA_EINVAL, "scene index out of range");
        return bytes_[s];
    }

   private:
    void Build(const float* const* targets, int nsteps, int every, float dt) {
        const int ns = (int)sc_.size();
        gen_.resize(ns);
        for (int s = 0; s < ns; ++s) gen_[s] = sc_[s]->layout_generation();
        hip_check(hipStreamCreateWithFlags(&ms_, hipStreamNonBlocking), "stream");
        hip_check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "event");
        ev_.assign(ns, nullptr);
        for (auto& e : ev_) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        tapes_.resize(ns);
        target_.resize(ns);
        bytes_.assign(ns, 0);
        for (int s = 0; s < ns; ++s) {
            size_t f0 = 0, f1 = 0, tot = 0;
            hip_check(hipMemGetInfo(&f0, &tot), "mem info");
            tapes_[s].reset(new Tape(*sc_[s], nsteps, every, false));
            const int tex = tapes_[s]->texels();
            target_[s].resize(tex);
            hip_check(hipMemcpy(target_[s].get(), targets[s], sizeof(float) * tex,
                                hipMemcpyHostToDevice),
                      "target");
            hip_check(hipMemGetInfo(&f1, &tot), "mem info");
            bytes_[s] = f0 > f1 ? (long long)(f0 - f1) : 0;
            tapes_[s]->SaveInitial();
        }
        part_.resize((size_t)ns * kMseBlocks);
        loss_.resize(ns);
        Capture(dt);
    }

    void release() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
        if (ms_) {
            (void)hipStreamSynchronize(ms_);
            (void)hipStreamDestroy(ms_);
            ms_ = nullptr;
        }
        for (auto& e : ev_)
            if (e) (void)hipEventDestroy(e);
        ev_.clear();
        if (fork_) (void)hipEventDestroy(fork_);
        fork_ = nullptr;
    }

    void Capture(float dt) {
        const int ns = (int)sc_.size();
        for (int s = 0; s < ns; ++s) hip_check(hipStreamSynchronize(sc_[s]->stream()), "sync");
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(ms_, hipStreamCaptureModeThreadLocal), "capture");
        try {
            hip_check(hipEventRecord(fork_, ms_), "fork");
            for (int s = 0; s < ns; ++s) {
                hipStream_t st = sc_[s]->stream();
                hip_check(hipStreamWaitEvent(st, fork_, 0), "fork wait");
                tapes_[s]->ForwardFromInitial(dt);
                tapes_[s]->SeedMse(target_[s].get(), part_.get() + (size_t)s * kMseBlocks,
                                   kMseBlocks, loss_.get() + s);
                tapes_[s]->BackwardSeeded();
                hip_check(hipEventRecord(ev_[s], st), "join");
                hip_check(hipStreamWaitEvent(ms_, ev_[s], 0), "join wait");
            }
        } catch (...) {
            (void)hipStreamEndCapture(ms_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(ms_, &g), "end capture");
        const hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "fit graph instantiate");
        rot_end_.resize(ns);
        for (int s = 0; s < ns; ++s) rot_end_[s] = sc_[s]->rotation();
    }

    std::vector<CoupledScene*> sc_;
    std::vector<std::unique_ptr<Tape>> tapes_;
    std::vector<DeviceBuffer<float>> target_;
    std::vector<long long> bytes_;
    std::vector<CoupledScene::Rotation> rot_end_;
    std::vector<unsigned long long> gen_;
    DeviceBuffer<double> part_, loss_;
    hipStream_t ms_ = nullptr;
    hipEvent_t fork_ = nullptr;
    std::vector<hipEvent_t> ev_;
    hipGraphExec_t exec_ = nullptr;
};

}  // namespace cwa

using namespace cwa;

struct cwa_tape {
    Tape* impl;
};

struct cwa_fit {
    FitBatch* impl;
};

// defined in cwa_scene.hip
namespace cwa {
CoupledScene* scene_impl(cwa_scene* s);
int guarded_call(void (*fn)(void*), void* ctx);
}  // namespace cwa

namespace {
template <class F>
int guarded_lambda(F&& f) {
    return cwa::guarded_call([](void* c) { (*static_cast<F*>(c))(); }, &f);
}

void copy_grad(const double* g, cwa_grad* out) {
    out->visc = g[GRAD_VISC];
    out->gas_const = g[GRAD_GAS_CONST];
    out->rest_rho = g[GRAD_REST_RHO];
    out->lambda_x = g[GRAD_LAMBDA_X];
    out->lambda_z = g[GRAD_LAMBDA_Z];
    out->atten = g[GRAD_ATTEN];
    out->couple_strength = g[GRAD_KAPPA];
    out->beta = g[GRAD_BETA];
}
}  // namespace

extern "C" {

int cwa_tape_create(cwa_scene* s, int nsteps, int ckpt_every, cwa_tape** out) {
    return guarded_lambda([&] {
        if (!s || !out) throw Error(CWA_EINVAL, "null argument");
        *out = nullptr;
        auto* t = new cwa_tape{nullptr};
        try {
            t->impl = new Tape(*scene_impl(s), nsteps, ckpt_every, true);
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

int cwa_tape_forward(cwa_tape* t, float dt) {
    return guarded_lambda([&] {
        if (!t) throw Error(CWA_EINVAL, "null argument");
        t->impl->Forward(dt);
    });
}

int cwa_tape_backward(cwa_tape* t, const float* dheight0, const float* dheight1,
                      const float* dpos, const float* dvel, cwa_grad* grad, float* dpos0,
                      float* dvel0) {
    return guarded_lambda([&] {
        if (!t || !grad) throw Error(CWA_EINVAL, "null argument");
        double g[GRAD_NUM];
        t->impl->Backward(dheight0, dheight1, dpos, dvel, g, dpos0, dvel0);
        copy_grad(g, grad);
    });
}

int cwa_fit_create(int32_t nscenes, cwa_scene* const* scenes, const float* const* targets,
                   int32_t nsteps, int32_t ckpt_every, float dt, cwa_fit** out) {
    return guarded_lambda([&] {
        if (!scenes || !targets || !out || nscenes < 1) throw Error(CWA_EINVAL, "bad argument");
        *out = nullptr;
        std::vector<CoupledScene*> sc(nscenes);
        for (int s = 0; s < nscenes; ++s) sc[s] = scene_impl(scenes[s]);
        auto* f = new cwa_fit{nullptr};
        try {
            f->impl = new FitBatch(std::move(sc), targets, nsteps, ckpt_every, dt);
        } catch (...) {
            delete f;
            throw;
        }
        *out = f;
    });
}

int cwa_fit_run(cwa_fit* f, double* losses, cwa_grad* grads) {
    return guarded_lambda([&] {
        if (!f || !f->impl) throw Error(CWA_EINVAL, "null fit");
        const int ns = f->impl->nscenes();
        std::vector<double> g((size_t)ns * GRAD_NUM);
        f->impl->Run(losses, g.data());
        if (grads)
            for (int s = 0; s < ns; ++s) copy_grad(g.data() + (size_t)s * GRAD_NUM, grads + s);
    });
}

int cwa_fit_tape_bytes(cwa_fit* f, int32_t scene, int64_t* bytes) {
    return guarded_lambda([&] {
        if (!f || !f->impl || !bytes) throw Error(CWA_EINVAL, "null argument");
        *bytes = f->impl->TapeBytes(scene);
    });
}

int cwa_fit_destroy(cwa_fit* f) {
    return guarded_lambda([&] {
        if (!f) return;
        delete f->impl;
        delete f;
    });
}

int cwa_tape_destroy(cwa_tape* t) {
    return guarded_lambda([&] {
        if (!t) return;
        delete t->impl;
        delete t;
    });
}

}  // extern "C"
