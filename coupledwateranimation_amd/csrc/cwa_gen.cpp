// cwa_gen.cpp — seeded scene generators (host only).  The CPU and GPU
// paths both consume the arrays these produce.
//
// Random numbers: SplitMix64 (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014); uniforms keep the top 24
// bits so every draw is exact in fp32.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/cwa.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // u in [0, 1) with 24 significant bits
    float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
};

}  // namespace

extern "C" {

int cwa_gen_lattice(float* pos4, int n, const float lo[3], const float hi[3], float spacing,
                    float site_offset, float jitter, const int32_t order[3], uint64_t seed) {
    if (n < 0 || !lo || !hi || !order || (n > 0 && !pos4)) return CWA_EINVAL;
    if (!(spacing > 0.0f)) return CWA_EINVAL;
    const int a0 = order[0], a1 = order[1], a2 = order[2];
    if (a0 < 0 || a0 > 2 || a1 < 0 || a1 > 2 || a2 < 0 || a2 > 2 || a0 == a1 || a1 == a2 ||
        a0 == a2)
        return CWA_EINVAL;
    // sites along the two fast axes; the small slack absorbs the rounding
    // of (hi - lo)/spacing for exact multiples
    const int n0 = (int)std::floor((double)(hi[a0] - lo[a0]) / spacing + 1e-4);
    const int n1 = (int)std::floor((double)(hi[a1] - lo[a1]) / spacing + 1e-4);
    if (n > 0 && (n0 < 1 || n1 < 1)) return CWA_EINVAL;
    SplitMix64 rng(seed);
    for (int k = 0; k < n; ++k) {
        const int i0 = k % n0;
        const int i1 = (k / n0) % n1;
        const int i2 = k / (n0 * n1);
        int idx[3];
        idx[a0] = i0;
        idx[a1] = i1;
        idx[a2] = i2;
        for (int a = 0; a < 3; ++a) {
            const float site = lo[a] + ((float)idx[a] + site_offset) * spacing;
            const float u = rng.uniform();
            pos4[4 * (size_t)k + a] = site + (2.0f * u - 1.0f) * jitter * spacing;
        }
        pos4[4 * (size_t)k + 3] = 1.0f;
    }
    return n;
}

int cwa_gen_bumps(float* height, int w, int h, int nbumps, float amp_lo, float amp_hi,
                  float sig_lo, float sig_hi, uint64_t seed) {
    if (w < 1 || h < 1 || nbumps < 0 || !height) return CWA_EINVAL;
    SplitMix64 rng(seed);
    struct Bump {
        double cx, cy, a, s;
    };
    std::vector<Bump> b(nbumps);
    for (auto& q : b) {
        q.cx = (double)rng.uniform() * w;
        q.cy = (double)rng.uniform() * h;
        q.a = amp_lo + (double)rng.uniform() * (amp_hi - amp_lo);
        q.s = sig_lo + (double)rng.uniform() * (sig_hi - sig_lo);
    }
    for (int j = 0; j < h; ++j) {
        for (int i = 0; i < w; ++i) {
            double v = 0.0;
            for (const auto& q : b) {
                const double dx = i + 0.5 - q.cx, dy = j + 0.5 - q.cy;
                const double r2 = dx * dx + dy * dy, s2 = 2.0 * q.s * q.s;
                if (r2 < 64.0 * s2) v += q.a * std::exp(-r2 / s2);
            }
            height[(size_t)j * w + i] = (float)v;
        }
    }
    return 0;
}

}  // extern "C"
