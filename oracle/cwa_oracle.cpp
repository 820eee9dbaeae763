// cwa_oracle.cpp — C entry points of the CPU oracle (float and double).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the
// CPU-baseline leg of bench.py through ctypes.  Never linked into the
// shipped library.  Built with -ffp-contract=off so every expression rounds
// as written.
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "cwa_oracle.hpp"

using namespace cwao;

namespace {

struct Handle {
    int is_double = 0;
    std::unique_ptr<Scene<float>> f;
    std::unique_ptr<Scene<double>> d;
};

template <class R, class T>
void load_scene(Scene<R>& S, const cwa_params* P, int n, const T* pos, const T* vel,
                const T* force, int W, int H, const T* h0, const T* h1, int mode,
                int nthreads) {
    S.P = *P;
    S.c = make_consts<R>(*P, W, H);
    S.mode = mode;
    S.nthreads = nthreads < 1 ? 1 : nthreads;
    S.a.n = n;
    S.a.pos.assign(4 * (size_t)n, R(0));
    S.a.vel.assign(4 * (size_t)n, R(0));
    S.a.force.assign(4 * (size_t)n, R(0));
    for (size_t i = 0; i < 4 * (size_t)n; ++i) {
        S.a.pos[i] = R(pos[i]);
        if (vel) S.a.vel[i] = R(vel[i]);
        if (force) S.a.force[i] = R(force[i]);
    }
    S.a.id.resize(n);
    std::iota(S.a.id.begin(), S.a.id.end(), 0);
    S.a.rho.assign(n, R(0));
    S.a.prs.assign(n, R(0));
    S.a.flux.assign(n, R(0));
    S.W = W;
    S.H = H;
    const size_t cells = (size_t)W * (size_t)H;
    S.h0.assign(cells, R(0));
    S.h1.assign(cells, R(0));
    for (size_t i = 0; i < cells; ++i) {
        if (h0) S.h0[i] = R(h0[i]);
        S.h1[i] = h1 ? R(h1[i]) : S.h0[i];
    }
}

// Copy out in original particle order (inverse of the id permutation).
template <class R, class T>
void get_scene(const Scene<R>& S, T* pos, T* vel, T* force, T* rho, T* prs, T* flux, T* h0,
               T* h1, int32_t* slot_of_id) {
    const int n = S.a.n;
    for (int s = 0; s < n; ++s) {
        const int i = S.a.id[s];
        for (int q = 0; q < 4; ++q) {
            if (pos) pos[4 * i + q] = T(S.a.pos[4 * s + q]);
            if (vel) vel[4 * i + q] = T(S.a.vel[4 * s + q]);
            if (force) force[4 * i + q] = T(S.a.force[4 * s + q]);
        }
        if (rho) rho[i] = T(S.a.rho[s]);
        if (prs) prs[i] = T(S.a.prs[s]);
        if (flux) flux[i] = T(S.a.flux[s]);
        if (slot_of_id) slot_of_id[i] = s;
    }
    const size_t cells = (size_t)S.W * (size_t)S.H;
    for (size_t k = 0; k < cells; ++k) {
        if (h0) h0[k] = T(S.h0[k]);
        if (h1) h1[k] = T(S.h1[k]);
    }
}

}  // namespace

extern "C" {

int cwao_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void cwao_exclusive_scan(const int32_t* in, int n, int32_t* out) { exclusive_scan(in, n, out); }

void cwao_blelloch_scan(int32_t* x, int n) { blelloch_scan(x, n); }

long long cwao_grid_build_boxes(const float* x4, const float* hw4, int n, const float* gmin,
                                const float* cell, const int32_t* dims, int32_t* count,
                                int32_t* offset, int32_t* list, long long capacity) {
    return grid_build_boxes(x4, hw4, n, gmin, cell, dims, count, offset, list, capacity);
}

int cwao_grid_build(const cwa_params* P, const float* pos4, int n, int32_t* key, int32_t* count,
                    int32_t* offset, int32_t* list) {
    const Consts<float> c = make_consts<float>(*P, 1, 1);
    grid_build(c, pos4, n, key, count, offset, list);
    return 0;
}

// Neighbour counts (r < h, self included) per input index.
// mode 0: through the grid (27-cell search), mode 1: all pairs.
int cwao_neighbor_counts(const cwa_params* P, const float* pos4, int n, int mode,
                         int32_t* counts) {
    const Consts<float> c = make_consts<float>(*P, 1, 1);
    auto within = [&](int i, int j) {
        return dist3(pos4[4 * i] - pos4[4 * j], pos4[4 * i + 1] - pos4[4 * j + 1],
                     pos4[4 * i + 2] - pos4[4 * j + 2]) < c.h;
    };
    if (mode == 1) {
        for (int i = 0; i < n; ++i) {
            int k = 0;
            for (int j = 0; j < n; ++j) k += within(i, j) ? 1 : 0;
            counts[i] = k;
        }
        return 0;
    }
    const int nc = c.dims[0] * c.dims[1] * c.dims[2];
    std::vector<int32_t> key(n), count(nc), offset(nc + 1), list(n);
    grid_build(c, pos4, n, key.data(), count.data(), offset.data(), list.data());
    for (int i = 0; i < n; ++i) {
        int ix, iy, iz;
        cell_of(c, pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], ix, iy, iz);
        int k = 0;
        for_each_candidate(c, offset.data(), ix, iy, iz,
                           [&](int d) { k += within(i, list[d]) ? 1 : 0; });
        counts[i] = k;
    }
    return 0;
}

void cwao_init_wave_f(float type, float* out, int W, int H) { init_wave<float>(type, out, W, H); }
void cwao_init_wave_d(double type, double* out, int W, int H) {
    init_wave<double>(type, out, W, H);
}

// Texture sample for sampler tests.
float cwao_tex_sample_f(const float* img, int W, int H, float u, float v) {
    return tex_sample<float>(img, W, H, u, v);
}

// One EVOLVE of the wave without particles (col_flux may be null).
void cwao_wave_evolve_f(const cwa_params* P, const float* h0, const float* h1, float* out, int W,
                        int H, const float* col_flux, float dt) {
    const Consts<float> c = make_consts<float>(*P, W, H);
    wave_evolve(c, h0, h1, out, W, H, col_flux, dt, 1);
}
void cwao_wave_evolve_d(const cwa_params* P, const double* h0, const double* h1, double* out,
                        int W, int H, const double* col_flux, double dt) {
    const Consts<double> c = make_consts<double>(*P, W, H);
    wave_evolve(c, h0, h1, out, W, H, col_flux, dt, 1);
}

// ---- scenes: is_double selects the arithmetic; arrays are float for the
// float scene and double for the double scene ----
void* cwao_scene_create_f(const cwa_params* P, int n, const float* pos, const float* vel,
                          const float* force, int W, int H, const float* h0, const float* h1,
                          int mode, int nthreads) {
    auto* h = new Handle();
    h->f.reset(new Scene<float>());
    load_scene(*h->f, P, n, pos, vel, force, W, H, h0, h1, mode, nthreads);
    return h;
}

void* cwao_scene_create_d(const cwa_params* P, int n, const double* pos, const double* vel,
                          const double* force, int W, int H, const double* h0, const double* h1,
                          int mode, int nthreads) {
    auto* h = new Handle();
    h->is_double = 1;
    h->d.reset(new Scene<double>());
    load_scene(*h->d, P, n, pos, vel, force, W, H, h0, h1, mode, nthreads);
    return h;
}

void cwao_scene_destroy(void* hp) { delete static_cast<Handle*>(hp); }

int cwao_scene_set_params(void* hp, const cwa_params* P) {
    auto* h = static_cast<Handle*>(hp);
    if (h->is_double) {
        h->d->P = *P;
        h->d->c = make_consts<double>(*P, h->d->W, h->d->H);
    } else {
        h->f->P = *P;
        h->f->c = make_consts<float>(*P, h->f->W, h->f->H);
    }
    return 0;
}

int cwao_scene_step(void* hp, double dt, int nsteps) {
    auto* h = static_cast<Handle*>(hp);
    for (int k = 0; k < nsteps; ++k) {
        if (h->is_double)
            scene_step(*h->d, dt);
        else
            scene_step(*h->f, (float)dt);
    }
    return 0;
}

int cwao_scene_n(void* hp) {
    auto* h = static_cast<Handle*>(hp);
    return h->is_double ? h->d->a.n : h->f->a.n;
}

// Arrays in original particle order; pointer types follow the scene type.
int cwao_scene_get_f(void* hp, float* pos, float* vel, float* force, float* rho, float* prs,
                     float* flux, float* h0, float* h1, int32_t* slot_of_id) {
    auto* h = static_cast<Handle*>(hp);
    if (h->is_double) return -1;
    get_scene(*h->f, pos, vel, force, rho, prs, flux, h0, h1, slot_of_id);
    return 0;
}

int cwao_scene_get_d(void* hp, double* pos, double* vel, double* force, double* rho,
                     double* prs, double* flux, double* h0, double* h1, int32_t* slot_of_id) {
    auto* h = static_cast<Handle*>(hp);
    if (!h->is_double) return -1;
    get_scene(*h->d, pos, vel, force, rho, prs, flux, h0, h1, slot_of_id);
    return 0;
}

// Grid arrays and column flux of the last step (float scene only).
int cwao_scene_grid(void* hp, int32_t* key, int32_t* count, int32_t* offset, int32_t* list,
                    float* colflux) {
    auto* h = static_cast<Handle*>(hp);
    if (h->is_double) return -1;
    const Scene<float>& S = *h->f;
    const int n = S.a.n;
    const int nc = S.c.dims[0] * S.c.dims[1] * S.c.dims[2];
    if (key) std::copy(S.key.begin(), S.key.begin() + n, key);
    if (count) std::copy(S.count.begin(), S.count.begin() + nc, count);
    if (offset) std::copy(S.offset.begin(), S.offset.begin() + nc + 1, offset);
    if (list) std::copy(S.list.begin(), S.list.begin() + n, list);
    if (colflux) std::copy(S.colflux.begin(), S.colflux.end(), colflux);
    return 0;
}

}  // extern "C"
