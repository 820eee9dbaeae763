// cwa_consumer.cpp — a C++ host that drives the coupled step through the C
// ABI alone, the way the reference application's Main.cpp would after the
// change shown in the integration notes: reference constants
// (Main.cpp:184-204 via cwa_params_reference), the make_cube particle block
// (Main.cpp:735-759: 64 x 5 x 64 sites at i * 2 * 0.85 * 0.005, z fastest,
// then y, then x), a 64 x 64 heightfield initialised by the wave's own
// Reinit (splash), then idle()'s step at the shaders' dt = 5e-5.
//
//   cwa_consumer check            no device work: reference parameters,
//                                 struct sizes, the invalid-argument path
//   cwa_consumer run <steps> <f>  run the scene, write the final state to
//                                 <f> (raw float32: pos 4n, vel 4n, rho n,
//                                 prs n, height0 64*64)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cwa.h"

namespace {

constexpr int kNx = 64, kNy = 5, kNz = 64, kN = kNx * kNy * kNz, kW = 64, kH = 64;

std::vector<float> make_cube() {
    const float s = (float)(2.0 * 0.85 * 0.005);
    std::vector<float> pos(4 * (size_t)kN, 0.0f);
    int k = 0;
    for (int i = 0; i < kNx; ++i)
        for (int j = 0; j < kNy; ++j)
            for (int l = 0; l < kNz; ++l, ++k) {
                pos[4 * k + 0] = (float)i * s;
                pos[4 * k + 1] = (float)j * s;
                pos[4 * k + 2] = (float)l * s;
                pos[4 * k + 3] = 1.0f;
            }
    return pos;
}

int fail(const char* what, int rc) {
    std::fprintf(stderr, "cwa_consumer: %s failed (%d): %s\n", what, rc, cwa_last_error());
    return 1;
}

int check() {
    if (cwa_abi_version() != CWA_ABI_VERSION) return fail("abi version", cwa_abi_version());
    if (cwa_struct_size(0) != (int)sizeof(cwa_params)) return fail("sizeof cwa_params", 0);
    if (cwa_struct_size(1) != (int)sizeof(cwa_init)) return fail("sizeof cwa_init", 0);
    if (cwa_struct_size(4) != (int)sizeof(cwa_host_state)) return fail("sizeof host state", 0);
    cwa_params p;
    cwa_params_reference(&p);
    if (!(p.h > 0.0f) || !(p.mass > 0.0f)) return fail("reference params", 0);
    // an invalid scene is rejected before any device work
    std::vector<float> pos = make_cube();
    cwa_init in{};
    in.n = kN;
    in.pos = pos.data();
    in.wave_w = 0;
    in.wave_h = kH;
    cwa_scene* s = nullptr;
    const int rc = cwa_scene_create(&p, &in, &s);
    if (rc != CWA_EINVAL || s != nullptr) return fail("rejecting a 0-wide heightfield", rc);
    std::printf("cwa_consumer check: abi %d, h %.6g, mass %.6g, invalid scene -> %d (%s)\n",
                cwa_abi_version(), p.h, p.mass, rc, cwa_last_error());
    return 0;
}

int run(int steps, const char* out) {
    cwa_params p;
    cwa_params_reference(&p);
    std::vector<float> pos = make_cube();
    cwa_init in{};
    in.n = kN;
    in.pos = pos.data();
    in.wave_w = kW;
    in.wave_h = kH;
    cwa_scene* s = nullptr;
    int rc = cwa_scene_create(&p, &in, &s);
    if (rc != CWA_OK) return fail("cwa_scene_create", rc);
    if ((rc = cwa_scene_wave_reinit(s)) != CWA_OK) return fail("cwa_scene_wave_reinit", rc);
    if ((rc = cwa_scene_step(s, 0.00005f, steps)) != CWA_OK) return fail("cwa_scene_step", rc);
    std::vector<float> v(4 * (size_t)kN), f(4 * (size_t)kN), rho(kN), prs(kN), h0(kW * kH),
        h1(kW * kH);
    std::vector<float> p4(4 * (size_t)kN);
    cwa_host_state st{p4.data(), v.data(), f.data(), rho.data(), prs.data(), h0.data(),
                      h1.data()};
    if ((rc = cwa_scene_readback(s, &st)) != CWA_OK) return fail("cwa_scene_readback", rc);
    cwa_scene_destroy(s);
    double rs = 0.0;
    bool finite = true;
    for (int i = 0; i < kN; ++i) {
        rs += rho[i];
        for (int a = 0; a < 3; ++a) finite = finite && std::isfinite(p4[4 * i + a]);
    }
    FILE* fp = std::fopen(out, "wb");
    if (!fp) return fail("opening the output file", 0);
    const bool ok = std::fwrite(p4.data(), sizeof(float), p4.size(), fp) == p4.size() &&
                    std::fwrite(v.data(), sizeof(float), v.size(), fp) == v.size() &&
                    std::fwrite(rho.data(), sizeof(float), rho.size(), fp) == rho.size() &&
                    std::fwrite(prs.data(), sizeof(float), prs.size(), fp) == prs.size() &&
                    std::fwrite(h0.data(), sizeof(float), h0.size(), fp) == h0.size();
    std::fclose(fp);
    if (!ok) return fail("writing the output file", 0);
    std::printf("cwa_consumer run: %d particles, %d steps, mean rho %.6g, finite %d\n", kN,
                steps, rs / kN, (int)finite);
    return finite ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "check") == 0) return check();
    if (argc >= 4 && std::strcmp(argv[1], "run") == 0) return run(std::atoi(argv[2]), argv[3]);
    std::fprintf(stderr, "usage: cwa_consumer check | run <steps> <out.bin>\n");
    return 2;
}
