// cwa_oracle2d.cpp — C entry points of the batched 2D oracle
// (cwa_oracle2d.hpp), for the ctypes binding in oracle2d.py.
//
// TEST INFRASTRUCTURE ONLY.  Scenes are independent; a step runs them in
// parallel with OpenMP (largest scenes first), each scene serially.
#include <omp.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>

#include "cwa_oracle2d.hpp"

namespace {

struct BatchBase {
    virtual ~BatchBase() = default;
    cwa2d_params p{};
    int nscenes = 0;
    virtual void step(double dt, int steps, int threads) = 0;
    virtual int get(int s, double* xv, double* rp, int32_t* idf, double* h0, double* h1) = 0;
    virtual int grid(int s, int32_t* key, int32_t* offset) = 0;
    virtual int colsum(int s, double* out) = 0;
};

template <class R>
struct Batch : BatchBase {
    std::vector<cwao2::Scene2<R>> S;
    std::vector<int> order;  // scenes by decreasing n

    void step(double dt, int steps, int threads) override {
        const cwao2::K2<R> c(p);
        const R d = R(dt);
        for (int it = 0; it < steps; ++it) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (threads > 1)
            for (int k = 0; k < nscenes; ++k) cwao2::step_scene(c, S[order[k]], d);
        }
    }
    int get(int s, double* xv, double* rp, int32_t* idf, double* h0, double* h1) override {
        if (s < 0 || s >= nscenes) return -1;
        const auto& A = S[s];
        for (int i = 0; i < A.n; ++i) {
            if (xv) {
                xv[4 * i + 0] = (double)A.x[i];
                xv[4 * i + 1] = (double)A.y[i];
                xv[4 * i + 2] = (double)A.vx[i];
                xv[4 * i + 3] = (double)A.vy[i];
            }
            if (rp) {
                rp[2 * i + 0] = (double)A.rho[i];
                rp[2 * i + 1] = (double)A.prs[i];
            }
            if (idf) idf[i] = A.id[i] | (A.en[i] ? (1 << 30) : 0);
        }
        for (int k = 0; k < A.L; ++k) {
            if (h0) h0[k] = (double)A.h0[k];
            if (h1) h1[k] = (double)A.h1[k];
        }
        return 0;
    }
    int grid(int s, int32_t* key, int32_t* offset) override {
        if (s < 0 || s >= nscenes || S[s].offset.empty()) return -1;
        const auto& A = S[s];
        if (key) std::memcpy(key, A.key.data(), sizeof(int32_t) * A.key.size());
        if (offset) std::memcpy(offset, A.offset.data(), sizeof(int32_t) * A.offset.size());
        return 0;
    }
    int colsum(int s, double* out) override {
        if (s < 0 || s >= nscenes) return -1;
        const auto& A = S[s];
        for (size_t k = 0; k < A.colsum.size(); ++k) out[k] = (double)A.colsum[k];
        return 0;
    }
};

template <class R>
BatchBase* make(const cwa2d_params* p, int nscenes, const cwa2d_scene* sc, const float* xv,
                const int32_t* en, const float* h0, const float* h1) {
    auto* B = new Batch<R>();
    B->p = *p;
    B->nscenes = nscenes;
    B->S.resize(nscenes);
    size_t po = 0, wo = 0;
    for (int s = 0; s < nscenes; ++s) {
        auto& A = B->S[s];
        A.n = sc[s].n;
        A.L = sc[s].wave_len;
        A.nsub = std::max(1, (int)sc[s].wave_substeps);
        A.model = sc[s].wave_model;
        A.mass = R(sc[s].mass);
        A.lambda = R(sc[s].lambda);
        A.dx = R(sc[s].wave_dx);
        A.x.resize(A.n);
        A.y.resize(A.n);
        A.vx.resize(A.n);
        A.vy.resize(A.n);
        A.rho.assign(A.n, R(p->rest_rho));
        A.prs.assign(A.n, R(0));
        A.id.resize(A.n);
        A.en.resize(A.n);
        for (int i = 0; i < A.n; ++i) {
            const float* q = xv + 4 * (po + i);
            A.x[i] = R(q[0]);
            A.y[i] = R(q[1]);
            A.vx[i] = R(q[2]);
            A.vy[i] = R(q[3]);
            A.id[i] = i;
            A.en[i] = en ? (en[po + i] != 0) : 1;
        }
        A.h0.assign(A.L, R(0));
        A.h1.assign(A.L, R(0));
        for (int k = 0; k < A.L; ++k) {
            if (h0) A.h0[k] = R(h0[wo + k]);
            if (h1) A.h1[k] = R(h1[wo + k]);
        }
        po += A.n;
        wo += A.L;
    }
    B->order.resize(nscenes);
    std::iota(B->order.begin(), B->order.end(), 0);
    std::stable_sort(B->order.begin(), B->order.end(),
                     [&](int a, int b) { return B->S[a].n > B->S[b].n; });
    return B;
}

}  // namespace

extern "C" {

void* cwao2_create(const cwa2d_params* p, int nscenes, const cwa2d_scene* sc, const float* xv,
                   const int32_t* en, const float* h0, const float* h1, int use_double) {
    if (!p || nscenes <= 0 || !sc || !xv) return nullptr;
    for (int s = 0; s < nscenes; ++s)
        if (sc[s].n < 1 || sc[s].wave_len < 3 ||
            (sc[s].wave_model != CWA2D_WAVE_LEAPFROG && sc[s].wave_model != CWA2D_WAVE_SHALLOW))
            return nullptr;
    if (use_double) return make<double>(p, nscenes, sc, xv, en, h0, h1);
    return make<float>(p, nscenes, sc, xv, en, h0, h1);
}

void cwao2_destroy(void* b) { delete static_cast<BatchBase*>(b); }

int cwao2_set_params(void* b, const cwa2d_params* p) {
    static_cast<BatchBase*>(b)->p = *p;
    return 0;
}

int cwao2_step(void* b, double dt, int steps, int threads) {
    static_cast<BatchBase*>(b)->step(dt, steps, std::max(1, threads));
    return 0;
}

int cwao2_get(void* b, int s, double* xv, double* rp, int32_t* idf, double* h0, double* h1) {
    return static_cast<BatchBase*>(b)->get(s, xv, rp, idf, h0, h1);
}

int cwao2_grid(void* b, int s, int32_t* key, int32_t* offset) {
    return static_cast<BatchBase*>(b)->grid(s, key, offset);
}

int cwao2_colsum(void* b, int s, double* out) {
    return static_cast<BatchBase*>(b)->colsum(s, out);
}

// Grid cells (H/2) per axis.
void cwao2_dims(const cwa2d_params* p, int32_t* nx, int32_t* ny) {
    *nx = cwao2::grid_dim(p->lo[0], p->hi[0], 0.5f * p->h);
    *ny = cwao2::grid_dim(p->lo[1], p->hi[1], 0.5f * p->h);
}

// Pieces, for the closed-form checks (double).
double cwao2_W_d(const cwa2d_params* p, double r) {
    const cwao2::K2<double> c(*p);
    return cwao2::W_cubic(c, r);
}
double cwao2_Wgrad_d(const cwa2d_params* p, double r) {
    const cwao2::K2<double> c(*p);
    return cwao2::W_cubic_grad(c, r);
}
double cwao2_sdf_d(const cwa2d_params* p, double x, double y, double* n) {
    const cwao2::K2<double> c(*p);
    return cwao2::box_sdf(c, x, y, n[0], n[1]);
}
double cwao2_wave_at_d(const cwa2d_params* p, const double* w, int L, double x) {
    const cwao2::K2<double> c(*p);
    return cwao2::wave_at(c, w, L, x);
}
float cwao2_column_sum_f(const float* f, int n) { return cwao2::column_sum(f, 0, n); }

// nsub shallow-water substeps of (h, uh) in double, in place.
void cwao2_shallow_d(double* h, double* uh, int L, double lam, double dx, int nsub) {
    std::vector<double> H(h, h + L), U(uh, uh + L), Hm(L, 0.0), Um(L, 0.0);
    for (int s = 0; s < nsub; ++s) cwao2::shallow_substep(H, U, Hm, Um, lam, dx);
    std::copy(H.begin(), H.end(), h);
    std::copy(U.begin(), U.end(), uh);
}

}  // extern "C"
