// cwa_oracle2d.hpp — CPU restatement of the 2D coupled step of the
// reference's SphWave2D application, for a batch of independent scenes.
//
// TEST INFRASTRUCTURE ONLY: the checker of the batched HIP path.
//
// Followed, per scene:
//   SphWave2D/SphWaveKoschier2D_grid_cs.glsl
//     W_cubic / W_cubic_grad :233-262, k2 = 40/(7 pi H^2) :135, PSI :136
//     boundary_sdf :205-230 (opU keeps the later plane on ties, :191-194)
//     ComputeDensityPressure :393-487, ComputeForces :490-628
//   SphWave2D/Wave1D_cs.glsl IterateWave :120-130, FreeBC :139-157
//   SphWave2D/Shallow1D_cs.glsl IterateWave :133-167 (two-phase Lax-Wendroff,
//     FreeBC copies the full cell state to the edges after each phase)
//   SphWave2D/StencilBuffer.cpp:150-179 (SphUgrid::Compute: grid rebuilt
//     every SPH substep), StencilImage2D.cpp:141-164 (wave substeps)
//   UniformGrid2D/UniformGrid2D.cpp:33-92 (counting sort, stable by index)
//
// Float restatement.  Every expression below is evaluated in type R exactly
// as written (the build uses -ffp-contract=off), and the HIP kernels
// evaluate the same expressions in the same order, so the float build and
// the GPU agree bit for bit.  Where GLSL leaves the arithmetic open (the
// precision of '/', sqrt, pow, texture filtering) the restatement fixes it:
//   r/H                -> r * (1/H)              (GLSL allows x * (1/y))
//   6 k2 q (3q-2) / H, -6 k2 q1^2 / H (W_cubic_grad) -> (6 k2 (1/H)) q (3q-2),
//                         (-6 k2 (1/H)) q1^2     (constants folded)
//   pow(rho/rho0, 3)   -> (t*t)*t
//   length(rij)        -> sqrt(fma(dy, dy, dx*dx)), correctly rounded
//   rij / r            -> rij * (1/r)
//   dot(vij, rij) / (r*r + 0.01 HSQ) -> dot * (1/(r2 + 0.01 HSQ))
//   r*r in viscosity   -> the same fma chain r2
//   texture(1D, LINEAR, CLAMP) -> fp32 weights, t = u*L - 1/2
// Additions the reference leaves undefined (recorded as decisions):
//   * The coupling velocity of the enable flag and the wave pressure: for a
//     shallow-water column the reference's own xvel = w[1]/w[0] = uh/h
//     (:362); for a leapfrog column, whose second channel is no momentum,
//     the wave's vertical velocity (h^n - h^(n-1)) / dt_sub at x.
//   * Particle -> wave flux (couple_band > 0): (m/rho0) v_y (1 - q^2)^2 of
//     particles within the band of the surface, summed per grid column in a
//     fixed order (64 lane-strided partial sums, then a xor butterfly) and
//     added to the newest level after the substeps as kappa dt S(x) / H.
//   * Pair terms skip coincident particles (r = 0), where rij/r is 0/0.
//   * The grid's cells are H/2 wide: the query visits the 5 cell columns
//     ix-2..ix+2 around the particle's column (column order, slots ascending
//     within a column), each cut to the rows its disc of radius H reaches
//     (column_rows: float bounds with a correctly rounded square root and a
//     margin, the same operations as the kernels), the same neighbours
//     within H as the reference's (pos -+ H) cell box on its 0.3-wide
//     cells, in another order.  The particle -> wave flux is summed per
//     column pair (H wide).
// Pinned (tests/test_oracle2d.py): kernel normalisation by quadrature and its
// gradient by differences; viscosity as nu * Laplacian of a quadratic
// velocity field on a lattice, radial and dissipative per pair; the psi
// boundary terms at known wall distances and the boundary pressure as the
// difference gradient of the boundary density; the wave pressure's gate and
// smoothstep edges; SDF distances and ties; the wave's closed forms
// (uniform field, CFL-1 shift, free boundary, attenuation); Shallow1D at
// rest, mass telescoping and pulse speed; the stable sort; flux
// conservation.  Full trajectories against the reference: parity unpinned
// (the GLSL cannot run here).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../include/cwa2d.h"

namespace cwao2 {

// M_PI as written in SphWaveKoschier2D_grid_cs.glsl:131.
static constexpr double kPI2 = 3.14159265;

// Cells per axis: ceil((hi - lo) / h) in float, at least 1.
static inline int grid_dim(float lo, float hi, float h) {
    const float w = (hi - lo) / h;
    return std::max(1, (int)std::ceil(w));
}

template <class R>
struct K2 {
    R H, HSQ, invH, k2, k2x2, c6k2, m6k2, c01, ga, gb;  // ga, gb: c6k2 / H, m6k2 / H
    R rho0, k, visc_pre, psi, gx, gy;
    R lox, loy, hix, hiy, walls[4];
    R thr, wref, wpress, wdy, wmask, atten, beta, band, kappa;
    R hc;          // grid cell edge H/2
    int nx, ny;    // grid cells (H/2) per axis
    int nxc;       // flux columns (H wide): cell columns 2c and 2c + 1
    int nsph;
    // candidate-row trimming, in float whatever R is (the same float
    // arithmetic as the kernels): the domain and the cell edge as floats,
    // T = the smallest float with sqrt(T) >= H, the bounds' margin
    float flox, floy, fhc, fT, feps;
    explicit K2(const cwa2d_params& p) {
        H = R(p.h);
        HSQ = H * H;
        invH = R(1) / H;
        k2 = R(40) / (R(7) * R(kPI2) * HSQ);
        k2x2 = k2 * R(2);
        c6k2 = R(6) * k2;
        m6k2 = R(-6) * k2;
        ga = c6k2 * invH;
        gb = m6k2 * invH;
        c01 = R(0.01f) * HSQ;
        rho0 = R(p.rest_rho);
        k = R(p.gas_const);
        visc_pre = -R(p.visc) * R(8);
        psi = R(p.psi);
        gx = R(p.gravity[0]);
        gy = R(p.gravity[1]);
        lox = R(p.lo[0]);
        loy = R(p.lo[1]);
        hix = R(p.hi[0]);
        hiy = R(p.hi[1]);
        for (int i = 0; i < 4; ++i) walls[i] = R(p.walls[i]);
        thr = R(p.enable_vel);
        wref = R(p.wave_ref);
        wpress = R(p.wave_press);
        wdy = R(p.wave_dy);
        wmask = R(p.wave_maskv);
        atten = R(p.atten);
        beta = R(p.beta);
        band = R(p.couple_band);
        kappa = R(p.couple_strength);
        hc = R(0.5f * p.h);
        nx = grid_dim(p.lo[0], p.hi[0], 0.5f * p.h);
        ny = grid_dim(p.lo[1], p.hi[1], 0.5f * p.h);
        nxc = grid_dim(p.lo[0], p.hi[0], p.h);
        nsph = std::max(1, (int)p.sph_substeps);
        flox = p.lo[0];
        floy = p.lo[1];
        fhc = 0.5f * p.h;
        fT = p.h * p.h;
        while (fT > 0.0f && std::sqrt(fT) >= p.h) fT = std::nextafter(fT, 0.0f);
        while (!(std::sqrt(fT) >= p.h)) fT = std::nextafter(fT, INFINITY);
        const float ext = std::max({std::fabs(p.lo[0]), std::fabs(p.hi[0]), std::fabs(p.lo[1]),
                                    std::fabs(p.hi[1]), p.h});
        feps = 4e-6f * ext;
    }
};

template <class R>
static inline R W_cubic(const K2<R>& c, R r) {
    const R q = r * c.invH;
    if (q >= R(1)) return R(0);
    if (q <= R(0.5)) return c.k2 * (R(6) * q * q * (q - R(1)) + R(1));
    const R q1 = R(1) - q;
    return c.k2x2 * q1 * q1 * q1;
}

template <class R>
static inline R W_cubic_grad(const K2<R>& c, R r) {
    const R q = r * c.invH;
    if (q >= R(1)) return R(0);
    if (q <= R(0.5)) return c.ga * q * (R(3) * q - R(2));
    const R q1 = R(1) - q;
    return c.gb * (q1 * q1);
}

// GLSL smoothstep (GLSL 4.50 spec 8.3): t = clamp((x-e0)/(e1-e0), 0, 1).
template <class R>
static inline R smoothstep(R e0, R e1, R x) {
    R t = (x - e0) / (e1 - e0);
    t = std::min(std::max(t, R(0)), R(1));
    return t * t * (R(3) - R(2) * t);
}

// boundary_sdf (:205-230): floor, left, right, ceiling; opU(res, d) keeps
// res only when res.z < d.z, so ties go to the later plane.
template <class R>
static inline R box_sdf(const K2<R>& c, R x, R y, R& nx, R& ny) {
    const R d[4] = {y - c.walls[0], x - c.walls[1], c.walls[2] - x, c.walls[3] - y};
    static const int nn[4][2] = {{0, 1}, {1, 0}, {-1, 0}, {0, -1}};
    int b = 0;
    for (int k = 1; k < 4; ++k)
        if (!(d[b] < d[k])) b = k;
    nx = R(nn[b][0]);
    ny = R(nn[b][1]);
    return d[b];
}

// clamp of a floor()ed coordinate to [lo, hi] with NaN -> lo.
template <class R>
static inline R clampf_nan(R f, R lo, R hi) {
    if (!(f >= lo)) return lo;
    return f > hi ? hi : f;
}

// 1D LINEAR + CLAMP_TO_EDGE sample of a wave level at world x.
template <class R>
static inline R wave_at(const K2<R>& c, const R* w, int L, R x) {
    const R sL = R(L) / (c.hix - c.lox);
    const R t = (x - c.lox) * sL - R(0.5);
    const R f0 = std::floor(t);
    const R a = t - f0;
    const R f = clampf_nan(f0, R(-2), R(L + 1));
    const int i0 = std::min(std::max((int)f, 0), L - 1);
    const int i1 = std::min(std::max((int)f + 1, 0), L - 1);
    return (R(1) - a) * w[i0] + a * w[i1];
}

// clamp(floor((p - lo)/H), 0, dim-1), as the grid hash.
template <class R>
static inline int cell_axis2(R p, R lo, R H, int dim) {
    const R f = clampf_nan(std::floor((p - lo) / H), R(-1), R(dim));
    return std::min(std::max((int)f, 0), dim - 1);
}

template <class R>
struct Scene2 {
    int n = 0, L = 0, nsub = 1;
    int model = CWA2D_WAVE_LEAPFROG;
    R mass = 0, lambda = 0, dx = 1;
    // slot-ordered state (slots are cell-sorted after each grid build)
    std::vector<R> x, y, vx, vy, rho, prs;
    std::vector<int32_t> id;   // scene-local id
    std::vector<int32_t> en;   // enabled flag
    std::vector<R> h0, h1;     // leapfrog: newest / previous level; shallow: h, uh
    // last grid build: local cell of each (pre-sort) slot, offsets
    std::vector<int32_t> key, offset;
    std::vector<R> colsum;
};

// The fixed-order column sum: 64 lane-strided sequential partial sums, then
// s[l] += s[l ^ o] for o = 32, 16, ..., 1; the value of lane 0.
template <class R>
static inline R column_sum(const R* f, int b, int e) {
    R s[64];
    for (int l = 0; l < 64; ++l) {
        R acc = R(0);
        for (int k = b + l; k < e; k += 64) acc += f[k];
        s[l] = acc;
    }
    for (int o = 32; o > 0; o >>= 1) {
        R t[64];
        for (int l = 0; l < 64; ++l) t[l] = s[l] + s[l ^ o];
        for (int l = 0; l < 64; ++l) s[l] = t[l];
    }
    return s[0];
}

// Counting sort of the slots by cell (UniformGrid2D::Build semantics:
// count, exclusive prefix in cell order, insert in slot order = stable),
// then the state is permuted into sorted order.
template <class R>
static void grid_sort(const K2<R>& c, Scene2<R>& S) {
    const int n = S.n, nc = c.nx * c.ny;
    S.key.assign(n, 0);
    std::vector<int32_t> cnt(nc, 0), fill(nc, 0), lst(n);
    for (int i = 0; i < n; ++i) {
        const int ix = cell_axis2(S.x[i], c.lox, c.hc, c.nx);
        const int iy = cell_axis2(S.y[i], c.loy, c.hc, c.ny);
        S.key[i] = ix * c.ny + iy;
        cnt[S.key[i]]++;
    }
    S.offset.assign(nc + 1, 0);
    for (int k = 0; k < nc; ++k) S.offset[k + 1] = S.offset[k] + cnt[k];
    for (int i = 0; i < n; ++i) lst[S.offset[S.key[i]] + fill[S.key[i]]++] = i;
    auto perm = [&](auto& a) {
        auto t = a;
        for (int d = 0; d < n; ++d) t[d] = a[lst[d]];
        a.swap(t);
    };
    perm(S.x);
    perm(S.y);
    perm(S.vx);
    perm(S.vy);
    perm(S.rho);
    perm(S.prs);
    perm(S.id);
    perm(S.en);
}

// The candidates of a query at (px, py): the cell columns ix-2..ix+2 (cells
// H/2) around its column, in order, each cut to the rows the disc of radius
// H reaches (bounds widened by a margin, computed in float from the float
// coordinates), each one contiguous slot range.  Every point within H of
// the query lies in them, except a neighbour that its density pass moved
// out of its cell (un-penetration): the cell lists hold the positions the
// grid was built from, as the reference's do.  Pair terms accumulate only
// the neighbours, in this order.
template <class R>
static inline bool column_rows(const K2<R>& c, float x, float y, int ix, int cx, int& ylo,
                               int& yhi) {
    float dx = 0.0f;
    if (cx < ix) dx = x - (c.flox + (float)(cx + 1) * c.fhc);
    else if (cx > ix) dx = (c.flox + (float)cx * c.fhc) - x;
    dx = std::max(dx - c.feps, 0.0f);
    const float d2 = dx * dx;
    if (!(d2 < c.fT)) return false;
    const float ry = std::sqrt(c.fT - d2) + c.feps;
    ylo = cell_axis2(y - ry, c.floy, c.fhc, c.ny);
    yhi = cell_axis2(y + ry, c.floy, c.fhc, c.ny);
    return true;
}

template <class R, class F>
static inline void for_candidates(const K2<R>& c, const Scene2<R>& S, R px, R py, F&& f) {
    const float fx = (float)px, fy = (float)py;
    const int ix = cell_axis2(fx, c.flox, c.fhc, c.nx);
    for (int cx = std::max(ix - 2, 0); cx <= std::min(ix + 2, c.nx - 1); ++cx) {
        int ylo, yhi;
        if (!column_rows(c, fx, fy, ix, cx, ylo, yhi)) continue;
        const int b = S.offset[cx * c.ny + ylo], e = S.offset[cx * c.ny + yhi + 1];
        for (int j = b; j < e; ++j) f(j);
    }
}

// The wave's coupling velocity at x (see the header: uh/h for shallow
// water, (h^n - h^(n-1))/dt_sub for the leapfrog).
template <class R>
static inline R wave_velocity(const K2<R>& c, const Scene2<R>& S, R x, R dts) {
    if (S.model == CWA2D_WAVE_SHALLOW)
        return wave_at(c, S.h1.data(), S.L, x) / wave_at(c, S.h0.data(), S.L, x);
    return (wave_at(c, S.h0.data(), S.L, x) - wave_at(c, S.h1.data(), S.L, x)) / dts;
}

// Shallow1D substep (Shallow1D_cs.glsl:133-167): the half-step values at
// i + 1/2, the full step, each phase followed by FreeBC's copy of cells
// 1 and L-2 (all four channels) to the edges.
template <class R>
static void shallow_substep(std::vector<R>& H, std::vector<R>& U, std::vector<R>& Hm,
                            std::vector<R>& Um, R lam, R dx) {
    const int L = (int)H.size();
    const R G = R(9.8f), hg = R(0.5) * G, l2 = lam / R(2);
    for (int k = 1; k < L - 1; ++k) {
        const int e = k + 1;
        Hm[k] = (H[k] + H[e]) / R(2) - l2 * (U[e] - U[k]) / dx;
        Um[k] = (U[k] + U[e]) / R(2) -
                l2 * (U[e] * U[e] / H[e] + hg * H[e] * H[e] - U[k] * U[k] / H[k] -
                      hg * H[k] * H[k]) / dx;
    }
    auto edges = [&] {
        H[0] = H[1], U[0] = U[1], Hm[0] = Hm[1], Um[0] = Um[1];
        H[L - 1] = H[L - 2], U[L - 1] = U[L - 2], Hm[L - 1] = Hm[L - 2], Um[L - 1] = Um[L - 2];
    };
    edges();
    for (int k = 1; k < L - 1; ++k) {
        const int w = k - 1;
        const R h = H[k] - lam * (Um[k] - Um[w]) / dx;
        const R uh = U[k] - lam * (Um[k] * Um[k] / Hm[k] + hg * Hm[k] * Hm[k] -
                                   Um[w] * Um[w] / Hm[w] - hg * Hm[w] * Hm[w]) / dx;
        H[k] = h;
        U[k] = uh;
    }
    edges();
}

// One coupled step of one scene: sph_substeps x (grid, density, forces),
// then the wave substeps and the flux source.
template <class R>
static void step_scene(const K2<R>& c, Scene2<R>& S, R dt) {
    const int n = S.n, L = S.L;
    const R m = S.mass;
    const R dts = dt / R(S.nsub);
    const R m_rho0 = m / c.rho0;
    const R cv = c.visc_pre * m;
    std::vector<R> qx(n), qy(n), qvx(n), qvy(n), aa(n), ir(n), flux(n);
    for (int sub = 0; sub < c.nsph; ++sub) {
        grid_sort(c, S);
        // ---- density / pressure (:393-487) ----
        for (int i = 0; i < n; ++i) {
            R x = S.x[i], y = S.y[i], vx = S.vx[i], vy = S.vy[i];
            {   // enable flag from the wave velocity at the raw x (:400-415)
                const R vw = wave_velocity(c, S, x, dts);
                if (S.en[i] == 0 && std::fabs(vw) > c.thr) S.en[i] = 1;
                if (S.en[i] == 1 && std::fabs(vw) < c.thr) S.en[i] = 0;
            }
            R nx, ny;
            const R d = box_sdf(c, x, y, nx, ny);
            if (d < R(0)) {  // un-penetrate (:417-423)
                x -= R(0.75) * d * nx;
                y -= R(0.75) * d * ny;
                const R vn = std::min(R(0), vx * nx + vy * ny);
                vx -= R(1.75) * vn * nx;
                vy -= R(1.75) * vn * ny;
            }
            qx[i] = x;
            qy[i] = y;
            qvx[i] = vx;
            qvy[i] = vy;
            R f = R(0);
            if (c.band > R(0)) {
                const R dd = std::fabs(y - wave_at(c, S.h0.data(), L, x));
                if (dd < c.band) {
                    const R qq = dd / c.band;
                    const R t = R(1) - qq * qq;
                    f = m_rho0 * vy * t * t;
                }
            }
            flux[i] = f;
            R rho = S.rho[i], p = S.prs[i];
            if (!(y > c.hiy)) {  // particles above the view keep rho, p (:428)
                R acc = R(0);
                for_candidates(c, S, x, y, [&](int j) {
                    const R dx = x - S.x[j], dy = y - S.y[j];
                    const R r2 = std::fma(dy, dy, dx * dx);
                    if (r2 < c.HSQ) acc += W_cubic(c, std::sqrt(r2));
                });
                rho = m * acc;
                if (d < c.H) rho += c.psi * W_cubic(c, std::max(R(0), d));
                rho = std::max(c.rho0, rho);
                const R t = rho / c.rho0;
                p = c.k * (t * t * t - R(1));
            }
            S.rho[i] = rho;
            S.prs[i] = p;
            aa[i] = p / (rho * rho);
            ir[i] = R(1) / rho;
        }
        // ---- forces + forward Euler (:490-628) ----
        std::vector<R> nx_(n), ny_(n), nvx(n), nvy(n);
        for (int i = 0; i < n; ++i) {
            const R x = qx[i], y = qy[i], vx = qvx[i], vy = qvy[i];
            if (S.en[i] == 0) {  // disabled: the density pass's state (:494-498)
                nx_[i] = x;
                ny_[i] = y;
                nvx[i] = vx;
                nvy[i] = vy;
                continue;
            }
            const R ai = aa[i];
            R apx = 0, apy = 0, avx = 0, avy = 0;
            for_candidates(c, S, x, y, [&](int j) {
                if (j == i) return;
                const R dx = x - qx[j], dy = y - qy[j];
                const R r2 = std::fma(dy, dy, dx * dx);
                const R r = std::sqrt(r2);
                if (r < c.H && r > R(0)) {
                    const R wg = W_cubic_grad(c, r);
                    const R inv = R(1) / r;
                    const R ux = dx * inv, uy = dy * inv;
                    const R s = (ai + aa[j]) * wg;
                    apx -= s * ux;
                    apy -= s * uy;
                    const R vdr = (vx - qvx[j]) * dx + (vy - qvy[j]) * dy;
                    const R t = ir[j] * vdr * (R(1) / (r2 + c.c01)) * wg;
                    avx -= t * ux;
                    avy -= t * uy;
                }
            });
            avx *= cv;
            avy *= cv;
            apx *= m;
            apy *= m;
            R bnx, bny;
            const R d = box_sdf(c, x, y, bnx, bny);
            if (d < c.H) {  // boundary pressure (:553-562)
                const R wg = W_cubic_grad(c, std::max(R(0), d));
                apx += c.psi * ai * wg * -bnx;
                apy += c.psi * ai * wg * -bny;
            }
            R aty = R(0);
            if (S.id[i] % 5 < 4) {  // wave pressure (:586-603)
                const R hw = wave_at(c, S.h0.data(), L, x);
                const R vw = wave_velocity(c, S, x, dts);
                const R dh = hw - c.wref;
                const R mask = smoothstep(R(0), c.wmask, std::fabs(vw));
                R dyw = y - hw;
                dyw -= R(0.5) * dh;
                aty = R(0) - mask * c.wpress * smoothstep(R(0), c.wdy, dyw);
            }
            const R ax = apx + avx + c.gx + R(0);
            const R ay = apy + avy + c.gy + aty;
            const R v1x = vx + dt * ax, v1y = vy + dt * ay;
            nvx[i] = v1x;
            nvy[i] = v1y;
            nx_[i] = x + dt * v1x;
            ny_[i] = y + dt * v1y;
        }
        S.x.swap(nx_);
        S.y.swap(ny_);
        S.vx.swap(nvx);
        S.vy.swap(nvy);
    }
    // ---- column flux of the last substep ----
    S.colsum.assign(c.nxc, R(0));
    if (c.band > R(0))
        for (int cx = 0; cx < c.nxc; ++cx)
            S.colsum[cx] = column_sum(flux.data(), S.offset[2 * cx * c.ny],
                                      S.offset[std::min(2 * cx + 2, c.nx) * c.ny]);
    // ---- wave substeps, free boundary (Wave1D_cs.glsl:120-157 or
    //      Shallow1D_cs.glsl:133-167) ----
    std::vector<R> a(S.h0), b(S.h1), o(L);
    const R lam = S.lambda;
    if (S.model == CWA2D_WAVE_SHALLOW) {
        std::vector<R> hm(L, R(0)), um(L, R(0));
        for (int s = 0; s < S.nsub; ++s) shallow_substep(a, b, hm, um, lam, S.dx);
    } else {
        const R ca = R(2) - R(2) * lam - c.beta, cb = R(1) - c.beta;
        for (int s = 0; s < S.nsub; ++s) {
            for (int k = 1; k < L - 1; ++k)
                o[k] = c.atten * (ca * a[k] + lam * (a[k + 1] + a[k - 1]) - cb * b[k]);
            o[0] = o[1];
            o[L - 1] = o[L - 2];
            b.swap(a);  // b <- a
            a.swap(o);  // a <- o
        }
    }
    if (c.band > R(0)) {
        const R cw = (c.hix - c.lox) / R(L);
        const R kd = c.kappa * (dt * R(c.nsph));
        for (int k = 0; k < L; ++k) {
            const R gx = (R(k) + R(0.5)) * cw * c.invH - R(0.5);
            const R f0 = std::floor(gx);
            const R al = gx - f0;
            const R f = clampf_nan(f0, R(-2), R(c.nxc + 1));
            const int c0 = std::min(std::max((int)f, 0), c.nxc - 1);
            const int c1 = std::min(std::max((int)f + 1, 0), c.nxc - 1);
            const R Sx = (R(1) - al) * S.colsum[c0] + al * S.colsum[c1];
            a[k] = a[k] + kd * Sx * c.invH;
        }
    }
    S.h0.swap(a);
    S.h1.swap(b);
}

}  // namespace cwao2
