/* Host check of the square root of the batched 2D pair loops (sqrt_rn_mid /
 * sqrt_rn_clamped in coupledwateranimation_amd/csrc/cwa_batch2d.hip): from a
 * reciprocal square root y within K ulp of RN(1/sqrt x) (the hardware
 * estimate's place), g = x y, h = y / 2, one coupled Newton step and the
 * correction g + (x - g g) h give IEEE sqrtf(x) exactly.  With the estimate
 * taken at max(x, 2^-80), x = 0 returns +0.
 *
 * The force pass seeds the reciprocal of r = sqrt(x) with 2h; the check
 * also reports how far 2h lies from IEEE 1.0f / r (in ulp), which must stay
 * inside the 3-ulp window the reciprocal steps are checked for.
 *
 * usage: sqrt_rn_check E0 E1 K  -> every float in [2^E0, 2^E1), starts -K..K
 * ulp; prints "checked mismatches zero_ok max_seed_ulp", exits 1 on any
 * mismatch.  The
 * steps scale exactly by powers of four, so [1, 4) stands for every binade
 * pair inside [2^-96, 2^96]. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float f_of(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static uint32_t u_of(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }

static float sqrt_steps(float x, float y, float* half_rsqrt) {
    float g = x * y, h = 0.5f * y;
    const float e = fmaf(-g, h, 0.5f);
    g = fmaf(g, e, g);
    h = fmaf(h, e, h);
    *half_rsqrt = h;
    return fmaf(fmaf(-g, g, x), h, g);
}

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    const int e0 = atoi(argv[1]), e1 = atoi(argv[2]), k = atoi(argv[3]);
    if (e0 < -96 || e1 > 96 || e0 >= e1 || k < 0 || k > 8) return 2;
    long checked = 0, bad = 0;
    int max_seed = 0;
    float h;
    const uint32_t u0 = (uint32_t)(e0 + 127) << 23, u1 = (uint32_t)(e1 + 127) << 23;
    for (uint32_t u = u0; u < u1; ++u) {
        const float x = f_of(u);
        const float want = sqrtf(x);
        const float rs = (float)(1.0 / sqrt((double)x));
        for (int s = -k; s <= k; ++s) {
            const float got = sqrt_steps(x, f_of(u_of(rs) + (uint32_t)s), &h);
            ++checked;
            bad += u_of(got) != u_of(want);
            const int d = abs((int)u_of(2.0f * h) - (int)u_of(1.0f / got));
            if (d > max_seed) max_seed = d;
        }
    }
    /* x = 0 with the estimate of the clamped input */
    int zero_ok = 1;
    const float rz = (float)(1.0 / sqrt(0x1p-80));
    for (int s = -k; s <= k; ++s) {
        const float got = sqrt_steps(0.0f, f_of(u_of(rz) + (uint32_t)s), &h);
        zero_ok &= u_of(got) == 0u;
    }
    printf("%ld %ld %d %d\n", checked, bad, zero_ok, max_seed);
    return bad != 0 || !zero_ok;
}
