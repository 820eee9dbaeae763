// cwa_oracle.hpp — CPU restatement of the reference's coupled SPH + wave step.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
// against; it is never part of the shipped library.  Only tests/, the smoke
// check and the CPU-baseline leg of bench.py may load it.
//
// Templated on the real type R (float: the reference's arithmetic; double:
// the FP64 twin used to report each field's FP32 error).  Every function
// cites the reference file:line it restates.  Where the reference is
// ambiguous or quirky the choice is stated in the comment beside it.
//
// Parity status (what pins each part; the tests name the checks):
//   texture sampler ........ pinned: affine fields reproduced, clamp at edges
//   wave stencil ........... pinned: zero/constant fields, plane-wave
//                            amplification vs the leapfrog characteristic
//                            polynomial, mirror symmetry
//   InitWave splash/wave/wake pinned: closed form at centres / radius 5
//   exclusive scan ......... pinned: reference KAT (ParallelScan.cpp:129-137)
//                            and the reference's Blelloch tree restated
//   grid build ............. pinned: permutation / stability invariants,
//                            the z-tiled column order spelled out,
//                            brute-force neighbour sets (also across tiles)
//   box grid build ......... pinned: brute-force interval overlap of every
//                            box with every cell, clamping at the edges
//   density / EOS .......... pinned: isolated particle, two-particle closed
//                            form, wave term closed form
//   forces ................. pinned: two-particle closed form, isolated
//                            particle (gravity + torque + drag + wave)
//   integrate / boundary ... pinned: closed forms per face, foam branch,
//                            surface clamp, the wake-mode ceiling's cases
//   particle->wave flux .... pinned: single-particle closed form (this term
//                            has no reference counterpart)
//   full trajectories ...... parity unpinned against the reference itself
//                            (the GLSL cannot run here); pinned only as the
//                            composition of the parts above.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/cwa.h"

namespace cwao {

// PI as written in the shaders (rho_pres_comp.glsl:8, force_comp.glsl:11).
static constexpr double kPI = 3.141592741;

template <class R>
struct V3 {
    R x, y, z;
};

template <class R>
static inline V3<R> v3(R x, R y, R z) {
    return V3<R>{x, y, z};
}

// ---------------------------------------------------------------------------
// GL_LINEAR + GL_CLAMP_TO_EDGE sampling of a single-channel image
// (GLSL texture(), GL 4.5 spec 8.14.2; the heightfield textures are created
// LINEAR / CLAMP_TO_EDGE in StencilImage2DTripleBuffered.cpp:25-26).
// Weights are computed in R; hardware fixed-point weights are not modelled.
// ---------------------------------------------------------------------------
template <class R>
static inline R fetch_clamped(const R* img, int w, int h, int i, int j) {
    i = std::min(std::max(i, 0), w - 1);
    j = std::min(std::max(j, 0), h - 1);
    return img[(size_t)j * (size_t)w + (size_t)i];
}

// Bilinear interpolation at continuous texel coordinates (x, y), texel
// centres at integers, edge texels repeated beyond the border.
template <class R>
static inline R bilerp(const R* img, int w, int h, R x, R y) {
    R fx = std::floor(x), fy = std::floor(y);
    // keep the integer conversion defined for far-out coordinates
    fx = std::min(std::max(fx, R(-2)), R(w + 1));
    fy = std::min(std::max(fy, R(-2)), R(h + 1));
    const int i0 = (int)fx, j0 = (int)fy;
    const R a = x - std::floor(x), b = y - std::floor(y);
    const R t00 = fetch_clamped(img, w, h, i0, j0);
    const R t10 = fetch_clamped(img, w, h, i0 + 1, j0);
    const R t01 = fetch_clamped(img, w, h, i0, j0 + 1);
    const R t11 = fetch_clamped(img, w, h, i0 + 1, j0 + 1);
    const R r0 = (R(1) - a) * t00 + a * t10;
    const R r1 = (R(1) - a) * t01 + a * t11;
    return (R(1) - b) * r0 + b * r1;
}

// texture(tex, uv).r for normalised uv.
template <class R>
static inline R tex_sample(const R* img, int w, int h, R u, R v) {
    return bilerp(img, w, h, u * R(w) - R(0.5), v * R(h) - R(0.5));
}

// ---------------------------------------------------------------------------
// Parameters converted to R with the derived constants every pass uses.
// ---------------------------------------------------------------------------
template <class R>
struct Consts {
    R m, h, visc, rho0, k, rp, spiky, lap, poly6;
    R g[3];
    R crest, brk_mass, brk_visc, torque, drag, wforce, du;
    R vdamp, bounce, foam_speed, foam_f, foam_rho, foam_p, foam_v;
    R up[3], lo[3], ceil_extra;
    R org[2], uvs[2];
    R lx, lz, atten, beta, type, wake_thr, wake_amt;
    R band, kappa;
    R gmin[3], cell, cell_y;
    int ydiv;
    int dims[3];
    // column grid -> wave texel affine maps (see wave_coupling_coords)
    R ax, bx, az, bz;
};

template <class R>
static Consts<R> make_consts(const cwa_params& P, int wave_w, int wave_h) {
    Consts<R> c;
    c.m = R(P.mass);
    c.h = R(P.h);
    c.visc = R(P.visc);
    c.rho0 = R(P.rest_rho);
    c.k = R(P.gas_const);
    c.rp = R(P.particle_radius);
    const R h2 = c.h * c.h;
    const R h6 = h2 * h2 * h2;
    const R h9 = h6 * h2 * c.h;
    // force_comp.glsl:68-69 : spiky = -20/(PI h^6), laplacian = -spiky
    c.spiky = -R(P.spiky_coeff) / (R(kPI) * h6);
    c.lap = R(P.visc_coeff) / (R(kPI) * h6);
    // rho_pres_comp.glsl:66 : mass*315/(64 PI h^9) (pow replaced by products)
    c.poly6 = R(315) / (R(64) * R(kPI) * h9);
    for (int a = 0; a < 3; ++a) c.g[a] = R(P.gravity[a]);
    c.crest = R(P.crest_threshold);
    c.brk_mass = R(P.breaking_mass);
    c.brk_visc = R(P.breaking_visc);
    c.torque = R(P.torque_coeff);
    c.drag = R(P.drag_coeff);
    c.wforce = R(P.wave_force_coeff);
    c.du = R(P.wave_vel_du);
    c.vdamp = R(P.vel_damping);
    c.bounce = R(P.bounce);
    c.foam_speed = R(P.foam_speed);
    c.foam_f = R(P.foam_force);
    c.foam_rho = R(P.foam_rho);
    c.foam_p = R(P.foam_pressure);
    c.foam_v = R(P.foam_vel);
    for (int a = 0; a < 3; ++a) {
        c.up[a] = R(P.upper[a]);
        c.lo[a] = R(P.lower[a]);
        c.gmin[a] = R(P.grid_min[a]);
        c.dims[a] = P.grid_dims[a];
    }
    c.ceil_extra = R(P.ceiling_extra);
    for (int a = 0; a < 2; ++a) {
        c.org[a] = R(P.wave_origin[a]);
        c.uvs[a] = R(P.wave_uv_scale[a]);
    }
    c.lx = R(P.lambda_x);
    c.lz = R(P.lambda_z);
    c.atten = R(P.atten);
    c.beta = R(P.beta);
    c.type = R(P.wave_type);
    c.wake_thr = R(P.wake_threshold);
    c.wake_amt = R(P.wake_amount);
    c.band = R(P.couple_band);
    c.kappa = R(P.couple_strength);
    c.cell = R(P.cell_size);
    c.ydiv = P.grid_ydiv <= 0 ? 1 : P.grid_ydiv;
    c.cell_y = R(P.cell_size / (float)c.ydiv);  // the float edge the kernels use
    // Texel i of the wave has world x = origin + (i + 1/2)/(W * scale); its
    // column-grid coordinate is (x - grid_min)/cell - 1/2.  Computed in
    // double and rounded once.
    const double W = wave_w > 0 ? wave_w : 1, H = wave_h > 0 ? wave_h : 1;
    const double cell = P.cell_size;
    c.ax = R(1.0 / (W * (double)P.wave_uv_scale[0] * cell));
    c.bx = R(((double)P.wave_origin[0] + 0.5 / (W * (double)P.wave_uv_scale[0]) -
              (double)P.grid_min[0]) / cell - 0.5);
    c.az = R(1.0 / (H * (double)P.wave_uv_scale[1] * cell));
    c.bz = R(((double)P.wave_origin[1] + 0.5 / (H * (double)P.wave_uv_scale[1]) -
              (double)P.grid_min[2]) / cell - 0.5);
    return c;
}

// uv of a world position, rho_pres_comp.glsl:72 (coord = 2*pos.xz).
template <class R>
static inline void world_uv(const Consts<R>& c, R x, R z, R& u, R& v) {
    u = (x - c.org[0]) * c.uvs[0];
    v = (z - c.org[1]) * c.uvs[1];
}

// ---------------------------------------------------------------------------
// Uniform grid (ugrid_particles_cs.glsl:97-108; UniformGrid2D.cpp:25-92).
// Linearisation is y fastest, so every vertical column of cells is one
// contiguous range (the reference's Index() uses (i*ny + j)*nx + k, which
// aliases cells on non-cubic grids); the columns in z-tiles of kTileZ,
// x-major within a tile, the last tile possibly narrower:
//   column(ix, iz) = t*kTileZ*nx + ix*w + (iz - t*kTileZ),
//   t = iz / kTileZ, w = min(kTileZ, nz - t*kTileZ);  lin = column*ny + iy.
// ---------------------------------------------------------------------------
template <class R>
static inline void cell_of(const Consts<R>& c, R px, R py, R pz, int& ix, int& iy, int& iz) {
    const R p[3] = {px, py, pz};
    int ci[3];
    for (int a = 0; a < 3; ++a) {
        R f = std::floor((p[a] - c.gmin[a]) / (a == 1 ? c.cell_y : c.cell));
        // clamp before converting so far-out points stay defined
        f = std::min(std::max(f, R(-1)), R(c.dims[a]));
        int q = (int)f;
        ci[a] = std::min(std::max(q, 0), c.dims[a] - 1);
    }
    ix = ci[0];
    iy = ci[1];
    iz = ci[2];
}

constexpr int kTileZ = 16;
static inline int grid_column(const int dims[3], int ix, int iz) {
    const int t = iz / kTileZ;
    const int w = std::min(kTileZ, dims[2] - t * kTileZ);
    return t * kTileZ * dims[0] + ix * w + (iz - t * kTileZ);
}

static inline int lin_cell(const int dims[3], int ix, int iy, int iz) {
    return grid_column(dims, ix, iz) * dims[1] + iy;
}

// The box grid keeps the plain order (ix*nz + iz)*ny + iy.
static inline int box_lin_cell(const int dims[3], int ix, int iy, int iz) {
    return (ix * dims[2] + iz) * dims[1] + iy;
}

// Serial exclusive scan, the `truth` loop of ParallelScanTest
// (SphWave2D/ParallelScan.cpp:132-137).  out has n+1 entries.
static inline void exclusive_scan(const int32_t* in, int n, int32_t* out) {
    int32_t s = 0;
    for (int i = 0; i < n; ++i) {
        out[i] = s;
        s += in[i];
    }
    out[n] = s;
}

// The reference's Blelloch scan (prefix_sum_cs.glsl:18-43 driven by
// ParallelScan::Compute, SphWave2D/ParallelScan.cpp:43-95), restated
// serially; n must be a power of two.  In-place on x.
static inline void blelloch_scan(int32_t* x, int n) {
    if (n < 2) {
        if (n == 1) x[0] = 0;
        return;
    }
    int pass = 0;
    int m = n / 2;
    for (;;) {  // up-sweep
        const int stride = 2 << pass;
        for (int gid = 0; gid < m; ++gid) {
            const int ka = (stride / 2 - 1) + gid * stride, kb = ka + stride / 2;
            if (kb >= n) continue;
            x[kb] = x[ka] + x[kb];
        }
        if (m == 1) break;
        m /= 2;
        pass += 1;
    }
    for (;;) {  // down-sweep
        const int stride = 2 << pass;
        for (int gid = 0; gid < m; ++gid) {
            const int ka = (stride / 2 - 1) + gid * stride, kb = ka + stride / 2;
            if (kb >= n) continue;
            if (stride == n) x[n - 1] = 0;
            const int32_t t = x[ka];
            x[ka] = x[kb];
            x[kb] = t + x[kb];
        }
        if (m == n / 2) break;
        m *= 2;
        pass -= 1;
    }
}

// Counting sort with the stable insertion order of UniformGrid2D::Build
// (count, exclusive prefix, insert in index order).  pos is float4/double4
// interleaved.  offset has num_cells+1 entries.
template <class R>
static void grid_build(const Consts<R>& c, const R* pos4, int n, int32_t* key, int32_t* count,
                       int32_t* offset, int32_t* list) {
    const int nc = c.dims[0] * c.dims[1] * c.dims[2];
    std::fill(count, count + nc, 0);
    for (int i = 0; i < n; ++i) {
        int ix, iy, iz;
        cell_of(c, pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], ix, iy, iz);
        key[i] = lin_cell(c.dims, ix, iy, iz);
        count[key[i]]++;
    }
    exclusive_scan(count, nc, offset);
    std::vector<int32_t> fill(nc, 0);
    for (int i = 0; i < n; ++i) {
        const int k = key[i];
        list[offset[k] + fill[k]] = i;
        fill[k]++;
    }
}

// The box grid of UniformGridPbdGpu3D::CollisionQuery (UniformGridGpu3D.cpp:
// 121-162): ComputeCounts and InsertBox (uniform_grid_pbd_3D_cs.glsl:145-165,
// 180-206) over the cells ComputeCellIndex(x - hw) .. ComputeCellIndex(x + hw)
// per axis (:167-173: floor((p - min) / cellSize) clamped to the grid, with a
// cell size per axis), the exclusive scan of the counts (ParallelScan), and
// the insertion made stable: boxes ascending within a cell, where the
// shader's atomicAdd leaves the order open.  The cell is linearised y
// fastest, (ix*nz + iz)*ny + iy, without the particle grid's z-tiles
// (Index() at :175-178 uses (i*ny + j)*nx + k).
// Arithmetic in float, as the shader: x - hw and x + hw rounded, then the
// correctly rounded division.  Returns the number of (box, cell) pairs;
// list is written only when it holds that many (capacity).
static inline int box_cell_axis(float p, float lo, float cell, int dim) {
    float f = std::floor((p - lo) / cell);
    f = std::min(std::max(f, -1.0f), (float)dim);
    return std::min(std::max((int)f, 0), dim - 1);
}

static long long grid_build_boxes(const float* x4, const float* hw4, int n, const float gmin[3],
                                  const float cell[3], const int dims[3], int32_t* count,
                                  int32_t* offset, int32_t* list, long long capacity) {
    const int nc = dims[0] * dims[1] * dims[2];
    std::fill(count, count + nc, 0);
    std::vector<int> lo(3 * (size_t)n), hi(3 * (size_t)n);
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        long long m = 1;
        for (int a = 0; a < 3; ++a) {
            const float p0 = x4[4 * i + a] - hw4[4 * i + a];
            const float p1 = x4[4 * i + a] + hw4[4 * i + a];
            lo[3 * i + a] = box_cell_axis(p0, gmin[a], cell[a], dims[a]);
            hi[3 * i + a] = box_cell_axis(p1, gmin[a], cell[a], dims[a]);
            m *= std::max(hi[3 * i + a] - lo[3 * i + a] + 1, 0);
        }
        total += m;
        for (int a = lo[3 * i]; a <= hi[3 * i]; ++a)
            for (int b = lo[3 * i + 1]; b <= hi[3 * i + 1]; ++b)
                for (int c = lo[3 * i + 2]; c <= hi[3 * i + 2]; ++c)
                    count[box_lin_cell(dims, a, b, c)]++;
    }
    exclusive_scan(count, nc, offset);
    if (total > capacity) return total;
    std::vector<int32_t> fill(nc, 0);
    for (int i = 0; i < n; ++i)
        for (int a = lo[3 * i]; a <= hi[3 * i]; ++a)
            for (int b = lo[3 * i + 1]; b <= hi[3 * i + 1]; ++b)
                for (int c = lo[3 * i + 2]; c <= hi[3 * i + 2]; ++c) {
                    const int k = box_lin_cell(dims, a, b, c);
                    list[offset[k] + fill[k]++] = i;
                }
    return total;
}

// Visit the candidates of particle slot i in sorted order: the 9 columns
// (dx outer, dz inner), each the contiguous row range iy-ydiv..iy+ydiv
// (one cell_size above and below; cell_size >= h).
template <class R, class F>
static inline void for_each_candidate(const Consts<R>& c, const int32_t* offset, int ix, int iy,
                                      int iz, F&& f) {
    const int ylo = std::max(iy - c.ydiv, 0), yhi = std::min(iy + c.ydiv, c.dims[1] - 1);
    for (int dx = -1; dx <= 1; ++dx) {
        const int cx = ix + dx;
        if (cx < 0 || cx >= c.dims[0]) continue;
        for (int dz = -1; dz <= 1; ++dz) {
            const int cz = iz + dz;
            if (cz < 0 || cz >= c.dims[2]) continue;
            const int b = offset[lin_cell(c.dims, cx, ylo, cz)];
            const int e = offset[lin_cell(c.dims, cx, yhi, cz) + 1];
            for (int j = b; j < e; ++j) f(j);
        }
    }
}

// ---------------------------------------------------------------------------
// One step's particle passes on arrays in a fixed slot order.
// ---------------------------------------------------------------------------
template <class R>
struct Arrays {
    int n = 0;
    std::vector<R> pos, vel, force;  // 4*n
    std::vector<R> rho, prs, flux;   // n
    std::vector<int32_t> id;         // n
};

// length(delta) < smoothing_length (rho_pres_comp.glsl:62-64).  GLSL leaves
// the rounding of length() to the implementation; here the squared length
// is the fused chain fma(dz, dz, fma(dy, dy, dx*dx)) (each step rounded
// once, as IEEE fma defines) and length = sqrt of it, so the neighbour set
// is fixed bit for bit.
template <class R>
static inline R dist3(R dx, R dy, R dz) {
    return std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx)));
}

// Coupling gate and particle -> wave volume flux (no reference counterpart):
// inside the band |y - H| < band the wave acts on the particle, and the
// particle contributes (m/rho0) * v_y * (1 - q^2)^2, q = |y - H|/band.
template <class R>
static inline bool coupled(const Consts<R>& c, R y, R H) {
    return !(c.band > R(0)) || std::fabs(y - H) < c.band;
}

// rho_pres_comp.glsl:49-81 for particle i; neighbours via `cand`.
// Also the flux term of the particle (0 when coupling is off).
template <class R, class Cand>
static inline void density_one(const Consts<R>& c, const R* pos, const R* vel, int i,
                               const R* H0, int W, int Hh, Cand&& cand, R& rho_out, R& p_out,
                               R& flux_out) {
    const R xi = pos[4 * i], yi = pos[4 * i + 1], zi = pos[4 * i + 2];
    // sum of (h^2 - r^2)^3 first, then one product with mass * 315/(64 pi h^9)
    R acc = 0;
    cand([&](int j) {
        const R r = dist3(xi - pos[4 * j], yi - pos[4 * j + 1], zi - pos[4 * j + 2]);
        if (r < c.h) {
            const R q = c.h * c.h - r * r;
            acc += q * q * q;
        }
    });
    R rho = (c.m * c.poly6) * acc;
    R pressure = std::max(c.k * (rho - c.rho0), R(0));  // :70
    R u, v;
    world_uv(c, xi, zi, u, v);
    const R height = tex_sample(H0, W, Hh, u, v);  // :72-73
    if (coupled(c, yi, height)) {
        const R wave_force = height * rho;  // :75
        pressure += wave_force;             // :76
        rho += wave_force / (c.k * c.rp);   // :77
    }
    rho_out = std::max(c.rho0, rho);  // :79
    p_out = pressure;                 // :80
    R fl = 0;
    if (c.band > R(0)) {
        const R d = std::fabs(yi - height);
        if (d < c.band) {
            const R q = d / c.band;
            const R w = (R(1) - q * q) * (R(1) - q * q);
            fl = (c.m / c.rho0) * vel[4 * i + 1] * w;
        }
    }
    flux_out = fl;
}

// WaveVelocity, force_comp.glsl:117-128.
template <class R>
static inline V3<R> wave_velocity(const Consts<R>& c, const R* H0, int W, int Hh, R u, R v, R dt) {
    const R height = tex_sample(H0, W, Hh, u, v);
    const R hx = tex_sample(H0, W, Hh, u + c.du, v);
    const R hy = tex_sample(H0, W, Hh, u, v + c.du);
    return v3((hx - height) / dt, (hy - height) / dt, (hx - hy) / c.du);
}

// WaveNormal, force_comp.glsl:131-139: cross(dy, dx) with dx = (1,0,H(uv+(1,0))-H),
// dy = (0,1,H(uv+(0,1))-H); the offsets are a full texture width.
template <class R>
static inline V3<R> wave_normal(const R* H0, int W, int Hh, R u, R v) {
    const R height = tex_sample(H0, W, Hh, u, v);
    const V3<R> dx = v3(R(1), R(0), tex_sample(H0, W, Hh, u + R(1), v) - height);
    const V3<R> dy = v3(R(0), R(1), tex_sample(H0, W, Hh, u, v + R(1)) - height);
    return v3(dy.y * dx.z - dy.z * dx.y, dy.z * dx.x - dy.x * dx.z, dy.x * dx.y - dy.y * dx.x);
}

// force_comp.glsl:62-115 for particle i.  Returns the new force.xyz; fw
// receives force.w, which the crest branch divides in place (:93 divides
// the whole vec4) and the final assignment (:114) leaves alone.
template <class R, class Cand>
static inline V3<R> force_one(const Consts<R>& c, const R* pos, const R* vel, const R* force,
                              const R* rho, const R* prs, int i, const R* H0, int W, int Hh, R dt,
                              Cand&& cand, R& fw) {
    const R xi = pos[4 * i], yi = pos[4 * i + 1], zi = pos[4 * i + 2];
    const R vxi = vel[4 * i], vyi = vel[4 * i + 1], vzi = vel[4 * i + 2];
    V3<R> pf = v3(R(0), R(0), R(0)), vf = v3(R(0), R(0), R(0));
    cand([&](int j) {
        if (j == i) return;  // :74-77
        const R dx = xi - pos[4 * j], dy = yi - pos[4 * j + 1], dz = zi - pos[4 * j + 2];
        const R r = dist3(dx, dy, dz);
        if (r < c.h) {
            // :85  pres -= m (p_i + p_j)/(2 rho_j) * spiky * (h-r)^2 * normalize(delta)
            // normalize(0) is undefined (NaN on GPUs); boundary clamping can
            // put two particles on the same point (a box corner), so an
            // exactly coincident pair exerts no pressure force.
            if (r > R(0)) {
                const R s =
                    c.m * (prs[i] + prs[j]) / (R(2) * rho[j]) * c.spiky * (c.h - r) * (c.h - r);
                pf.x -= s * (dx / r);
                pf.y -= s * (dy / r);
                pf.z -= s * (dz / r);
            }
            // :86  visc += m (v_j - v_i)/rho_j * laplacian * (h-r)
            const R t = c.m / rho[j];
            const R l = c.lap * (c.h - r);
            vf.x += t * (vel[4 * j] - vxi) * l;
            vf.y += t * (vel[4 * j + 1] - vyi) * l;
            vf.z += t * (vel[4 * j + 2] - vzi) * l;
        }
    });
    V3<R> fprev = v3(force[4 * i], force[4 * i + 1], force[4 * i + 2]);
    fw = force[4 * i + 3];
    if (yi > c.crest) {  // :91-95 (the stored force is divided, then read by the torque)
        fprev.x /= c.brk_mass;
        fprev.y /= c.brk_mass;
        fprev.z /= c.brk_mass;
        fw /= c.brk_mass;
        vf.x *= c.brk_visc;
        vf.y *= c.brk_visc;
        vf.z *= c.brk_visc;
    }
    vf.x *= c.visc;  // :97
    vf.y *= c.visc;
    vf.z *= c.visc;
    R u, v;
    world_uv(c, xi, zi, u, v);
    const R height = tex_sample(H0, W, Hh, u, v);  // :99-100
    // :103-104 torque = 0.25 * cross(pos, force)
    V3<R> tq = v3(yi * fprev.z - zi * fprev.y, zi * fprev.x - xi * fprev.z,
                  xi * fprev.y - yi * fprev.x);
    tq.x *= c.torque;
    tq.y *= c.torque;
    tq.z *= c.torque;
    V3<R> drag = v3(R(0), R(0), R(0)), wf = v3(R(0), R(0), R(0));
    if (coupled(c, yi, height)) {
        const V3<R> wv = wave_velocity(c, H0, W, Hh, u, v, dt);  // :107-108
        drag = v3(-c.drag * (vxi - wv.x), -c.drag * (vyi - wv.y), -c.drag * (vzi - wv.z));
        const V3<R> wn = wave_normal(H0, W, Hh, u, v);  // :110
        wf = v3(-height * wn.x * c.wforce, -height * wn.y * c.wforce, -height * wn.z * c.wforce);
    }
    const V3<R> grav = v3(rho[i] * c.g[0], rho[i] * c.g[1], rho[i] * c.g[2]);  // :113
    return v3(pf.x + vf.x + grav.x + tq.x + drag.x + wf.x,
              pf.y + vf.y + grav.y + tq.y + drag.y + wf.y,
              pf.z + vf.z + grav.z + tq.z + drag.z + wf.z);  // :114
}

// CheckBoundary, integrate_comp.glsl:135-178.
template <class R>
static inline void check_boundary(const Consts<R>& c, R* p, R* v) {
    for (int a = 0; a < 3; ++a) {
        if (p[a] < c.lo[a]) {
            p[a] = c.lo[a];
            v[a] *= -c.bounce;
        } else if (p[a] > c.up[a]) {
            p[a] = c.up[a];
            v[a] *= -c.bounce;
        }
    }
    if (c.type > R(0) && c.type < R(1)) {
        if (p[1] > c.up[1] + c.ceil_extra) {
            p[1] = c.up[1] + c.ceil_extra;
            v[1] *= -c.bounce;
        }
    }
}

// integrate_comp.glsl:56-92 for particle i (in place on the out arrays).
// F is the force of this step; rho/prs are this step's outputs.
template <class R>
static inline void integrate_one(const Consts<R>& c, R* pos_o, R* vel_o, R* force_o, R* rho,
                                 R* prs, const R* pos_i, const R* vel_i, V3<R> F, R fw, int i,
                                 const R* H0, int W, int Hh, R dt) {
    const R ax = F.x / rho[i], ay = F.y / rho[i], az = F.z / rho[i];  // :62
    R nv[3] = {vel_i[4 * i] + dt * ax, vel_i[4 * i + 1] + dt * ay, vel_i[4 * i + 2] + dt * az};
    R np[3] = {pos_i[4 * i] + dt * nv[0], pos_i[4 * i + 1] + dt * nv[1],
               pos_i[4 * i + 2] + dt * nv[2]};  // :63-64
    const R damp = R(1) - c.vdamp * dt;           // :66
    nv[0] *= damp;
    nv[1] *= damp;
    nv[2] *= damp;
    if (dist3(nv[0], nv[1], nv[2]) > c.foam_speed) {  // :69-76
        F.x *= c.foam_f;
        F.y *= c.foam_f;
        F.z *= c.foam_f;
        fw *= c.foam_f;
        rho[i] *= c.foam_rho;
        prs[i] *= c.foam_p;
        nv[0] *= c.foam_v;
        nv[1] *= c.foam_v;
        nv[2] *= c.foam_v;
    }
    R u, v;
    world_uv(c, np[0], np[2], u, v);
    const R th = tex_sample(H0, W, Hh, u, v);  // :79
    if (np[1] < th) np[1] = th - c.rp;          // :80-83
    check_boundary(c, np, nv);                  // :87
    pos_o[4 * i] = np[0];
    pos_o[4 * i + 1] = np[1];
    pos_o[4 * i + 2] = np[2];
    vel_o[4 * i] = nv[0];
    vel_o[4 * i + 1] = nv[1];
    vel_o[4 * i + 2] = nv[2];
    force_o[4 * i] = F.x;
    force_o[4 * i + 1] = F.y;
    force_o[4 * i + 2] = F.z;
    force_o[4 * i + 3] = fw;
}

// wave_comp.glsl:171-187 (EvolveWave + get_clamp), generalised to
// lambda_x along u and lambda_z along v; with lambda_x == lambda_z == lambda
// it is the reference stencil (2-4l-b)c0 + l(n+s+e+w) - (1-b)c1.  `src`
// (may be null) is the particle column flux on the column grid, added as
// kappa * dt * bilerp(src) / cell^2 after attenuation.
template <class R>
static void wave_evolve(const Consts<R>& c, const R* h0, const R* h1, R* out, int W, int H,
                        const R* col_flux, R dt, int nthreads) {
    const R cc = R(2) - R(2) * c.lx - R(2) * c.lz - c.beta;
    const R c1 = R(1) - c.beta;
    const bool wake = c.type > R(0) && c.type < R(1);
    const R inv_area = R(1) / (c.cell * c.cell);
    const int nxc = c.dims[0], nzc = c.dims[2];
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int j = 0; j < H; ++j) {
        for (int i = 0; i < W; ++i) {
            const R C0 = h0[(size_t)j * W + i];
            const R C1 = h1[(size_t)j * W + i];
            const R n = fetch_clamped(h0, W, H, i, j + 1);
            const R s = fetch_clamped(h0, W, H, i, j - 1);
            const R e = fetch_clamped(h0, W, H, i + 1, j);
            const R w = fetch_clamped(h0, W, H, i - 1, j);
            R o = cc * C0 + c.lx * (e + w) + c.lz * (n + s) - c1 * C1;
            o *= c.atten;
            if (col_flux) {
                const R gx = c.ax * R(i) + c.bx, gz = c.az * R(j) + c.bz;
                const R S = bilerp(col_flux, nxc, nzc, gx, gz);
                o += c.kappa * dt * (S * inv_area);
            }
            if (wake && o > c.wake_thr && i == W / 2) o += c.wake_amt;  // :177-184
            out[(size_t)j * W + i] = o;
        }
    }
}

// smoothstep as defined by GLSL 4.50 for any edge order:
// t = clamp((x-e0)/(e1-e0), 0, 1); t*t*(3-2t)  (smoothstep(5,0,d) in
// wave_comp.glsl:138 has e0 > e1).
template <class R>
static inline R smoothstep(R e0, R e1, R x) {
    R t = (x - e0) / (e1 - e0);
    t = std::min(std::max(t, R(0)), R(1));
    return t * t * (R(3) - R(2) * t);
}

// InitWave, wave_comp.glsl:82-140.  ivec2(0.25*size) truncates.
template <class R>
static void init_wave(R type, R* out, int W, int H) {
    int c0x, c0y, c1x, c1y, c2x, c2y;
    R peak;
    bool three = true;
    if (type == R(1)) {  // splash
        c0x = (int)(0.25f * (float)W);
        c0y = (int)(0.25f * (float)H);
        c1x = (int)(0.75f * (float)W);
        c1y = (int)(0.75f * (float)H);
        c2x = c2y = 0;
        peak = R(0.5);
        three = false;
    } else if (type == R(0)) {  // wave
        c0x = (int)(0.25f * (float)W);
        c0y = H;
        c2x = (int)(0.5f * (float)W);
        c2y = H;
        c1x = (int)(0.75f * (float)W);
        c1y = H;
        peak = R(1);
    } else {  // wake
        c0x = (int)(0.5f * (float)W);
        c0y = (int)(0.1f * (float)H);
        c1x = c2x = c0x;
        c1y = c2y = c0y;
        peak = R(0.1);
    }
    for (int j = 0; j < H; ++j) {
        for (int i = 0; i < W; ++i) {
            auto dist = [&](int cx, int cy) {
                const R dx = R(i - cx), dy = R(j - cy);
                return std::sqrt(dx * dx + dy * dy);
            };
            R d = std::min(dist(c0x, c0y), dist(c1x, c1y));
            if (three) d = std::min(d, dist(c2x, c2y));
            out[(size_t)j * W + i] = peak * smoothstep(R(5), R(0), d);
        }
    }
}

// ---------------------------------------------------------------------------
// The coupled step (Main.cpp:530-562 order: density, force, integrate, then
// the wave Module), on the cell-sorted arrays.  Per step:
//   1. grid over the current slots; stable counting sort; reorder slots
//   2. density / pressure / flux            (heightfield h^n)
//   3. force + integrate                    (heightfield h^n)
//   4. column flux sums
//   5. wave evolve with the flux source, then rotate h^n -> h^(n-1)
// mode 1 = all-pairs (the reference's literal O(N^2) loops, no reorder).
// The heightfield the particles see is the newest h^n (the reference's
// texture-unit binding makes it 0-2 steps stale; not reproduced).
// ---------------------------------------------------------------------------
template <class R>
struct Scene {
    cwa_params P{};
    Consts<R> c{};
    Arrays<R> a;
    int W = 0, H = 0;
    std::vector<R> h0, h1, hn;
    std::vector<int32_t> key, count, offset, list;
    std::vector<R> colflux;
    int nthreads = 1;
    int mode = 0;
};

template <class R>
static void scene_step(Scene<R>& S, R dt) {
    Arrays<R>& A = S.a;
    const Consts<R>& c = S.c;
    const int n = A.n;
    const int W = S.W, H = S.H;
    const int nthreads = S.nthreads;
    const int nc = c.dims[0] * c.dims[1] * c.dims[2];
    S.key.resize(n);
    S.count.resize(nc);
    S.offset.resize(nc + 1);
    S.list.resize(n);
    const R* H0 = S.h0.data();

    std::vector<R> pos_s, vel_s, force_s;
    std::vector<int32_t> id_s;
    if (S.mode == 0) {
        grid_build(c, A.pos.data(), n, S.key.data(), S.count.data(), S.offset.data(),
                   S.list.data());
        pos_s.resize(4 * (size_t)n);
        vel_s.resize(4 * (size_t)n);
        force_s.resize(4 * (size_t)n);
        id_s.resize(n);
        for (int d = 0; d < n; ++d) {
            const int s = S.list[d];
            for (int q = 0; q < 4; ++q) {
                pos_s[4 * d + q] = A.pos[4 * s + q];
                vel_s[4 * d + q] = A.vel[4 * s + q];
                force_s[4 * d + q] = A.force[4 * s + q];
            }
            id_s[d] = A.id[s];
        }
    } else {
        pos_s = A.pos;
        vel_s = A.vel;
        force_s = A.force;
        id_s = A.id;
    }
    A.rho.assign(n, R(0));
    A.prs.assign(n, R(0));
    A.flux.assign(n, R(0));
    std::vector<int> cix(n), ciy(n), ciz(n);
    for (int i = 0; i < n; ++i)
        cell_of(c, pos_s[4 * i], pos_s[4 * i + 1], pos_s[4 * i + 2], cix[i], ciy[i], ciz[i]);

    auto cand_for = [&](int i) {
        return [&, i](auto&& f) {
            if (S.mode == 0) {
                for_each_candidate(c, S.offset.data(), cix[i], ciy[i], ciz[i], f);
            } else {
                for (int j = 0; j < n; ++j) f(j);
            }
        };
    };

#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) if (nthreads > 1)
    for (int i = 0; i < n; ++i) {
        density_one(c, pos_s.data(), vel_s.data(), i, H0, W, H, cand_for(i), A.rho[i], A.prs[i],
                    A.flux[i]);
    }

    std::vector<R> pos_o(4 * (size_t)n), vel_o(4 * (size_t)n), force_o(force_s);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) if (nthreads > 1)
    for (int i = 0; i < n; ++i) {
        R fw;
        V3<R> F = force_one(c, pos_s.data(), vel_s.data(), force_s.data(), A.rho.data(),
                            A.prs.data(), i, H0, W, H, dt, cand_for(i), fw);
        integrate_one(c, pos_o.data(), vel_o.data(), force_o.data(), A.rho.data(), A.prs.data(),
                      pos_s.data(), vel_s.data(), F, fw, i, H0, W, H, dt);
        pos_o[4 * i + 3] = pos_s[4 * i + 3];
        vel_o[4 * i + 3] = vel_s[4 * i + 3];
    }

    // Column flux: column (ix, iz) is the contiguous cell range
    // [col*ny, (col + 1)*ny), col = grid_column(ix, iz), so its particles are
    // one contiguous sorted range, summed in slot order.
    // Stored at [iz*nx + ix] (x fastest, like the heightfield).
    const bool couple = c.band > R(0) && S.mode == 0;
    const int nxc = c.dims[0], nzc = c.dims[2];
    S.colflux.assign((size_t)nxc * nzc, R(0));
    if (couple) {
        for (int ix = 0; ix < nxc; ++ix) {
            for (int iz = 0; iz < nzc; ++iz) {
                const int col = grid_column(c.dims, ix, iz);
                const int b = S.offset[col * c.dims[1]];
                const int e = S.offset[(col + 1) * c.dims[1]];
                R sum = 0;
                for (int d = b; d < e; ++d) sum += A.flux[d];
                S.colflux[(size_t)iz * nxc + ix] = sum;
            }
        }
    }

    if (W > 0 && H > 0) {
        S.hn.resize((size_t)W * H);
        wave_evolve(c, S.h0.data(), S.h1.data(), S.hn.data(), W, H,
                    couple ? S.colflux.data() : nullptr, dt, nthreads);
        // PingPong, StencilImage2DTripleBuffered.cpp:33-40: newest -> read 0
        std::swap(S.h1, S.h0);
        std::swap(S.h0, S.hn);
    }

    A.pos.swap(pos_o);
    A.vel.swap(vel_o);
    A.force.swap(force_o);
    A.id.swap(id_s);
}

}  // namespace cwao
