// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitize.py with g++/gcc, no GPU): the initial-condition generators of libastro
// (astro_amd/csrc/ic.cpp) with ordinary and edge arguments, and the CPU oracle
// (oracle/oracle.c) building trees and stepping small and degenerate systems.  Prints "ok".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/astro/astro.h"
#include "../../oracle/oracle.h"

static int fails = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::printf("check failed: %s (line %d)\n", #c, __LINE__); \
            fails++;                                               \
        }                                                          \
    } while (0)

static void step_system(std::vector<PositionMass>& p, std::vector<VelocityRadius>& v,
                        std::vector<ColorType>& c, uint32_t bits, uint32_t multipole) {
    astro_params P;
    orc_params_physics(&P, bits);
    P.multipole = multipole;
    const uint32_t n = (uint32_t)p.size();
    std::vector<float> f(4 * (size_t)n);
    orc_stats st{};
    CHECK(orc_step(p.data(), v.data(), c.data(), n, n, &P, 1e-3f, f.data(), &st, 1, nullptr,
                   nullptr) == 0);
    for (float x : f) CHECK(std::isfinite(x));
}

int main() {
    // generators: ordinary sizes, one particle, and refused arguments
    for (uint32_t n : {1u, 2u, 1000u, 4097u}) {
        std::vector<PositionMass> p(n);
        std::vector<VelocityRadius> v(n);
        std::vector<ColorType> c(n);
        CHECK(astro_ic_plummer(n, 42, p.data(), v.data(), c.data()) == 0);
        step_system(p, v, c, 21, 0);
        CHECK(astro_ic_plummer_background(n, 7, 0.1f, 20.0f, 2.35f, 0.1f, 10.0f, p.data(),
                                          v.data(), c.data()) == 0);
        step_system(p, v, c, 21, 2);
        std::vector<PositionMass> p2(2 * n);
        std::vector<VelocityRadius> v2(2 * n);
        std::vector<ColorType> c2(2 * n);
        CHECK(astro_ic_two_disks(n, 3, 10.0f, 1.0f, p2.data(), v2.data(), c2.data()) == 0);
        step_system(p2, v2, c2, 7, 0);
    }
    std::vector<PositionMass> p(10);
    CHECK(astro_ic_plummer(0, 1, p.data(), nullptr, nullptr) == ASTRO_ERR_ARG);
    CHECK(astro_ic_two_disks(5, 1, 1.0f, 2.0f, p.data(), nullptr, nullptr) == ASTRO_ERR_ARG);
    CHECK(astro_ic_plummer_background(10, 1, 1.5f, 20.0f, 2.35f, 0.1f, 10.0f, p.data(), nullptr,
                                      nullptr) == ASTRO_ERR_ARG);
    std::vector<uint32_t> sizes(64);
    CHECK(astro_ic_batch_sizes(64, 1024, 65536, 7, sizes.data()) == 0);
    CHECK(astro_ic_batch_sizes(4, 10, 5, 7, sizes.data()) == ASTRO_ERR_ARG);
    // degenerate systems through the oracle: every particle on one point, pairs of equal
    // keys, the reference's compat parameters with its last-leaf quirk, the smallest tree
    {
        std::vector<PositionMass> q(300);
        std::vector<VelocityRadius> w(300);
        std::vector<ColorType> c(300);
        for (auto& x : q) {
            x.position[0] = x.position[1] = x.position[2] = 0.125f;
            x.mass = 1.0f;
        }
        step_system(q, w, c, 21, 2);
        for (size_t i = 0; i < q.size(); i++) q[i].position[0] = (float)(i / 2);
        step_system(q, w, c, 1, 0);
        astro_params P;
        orc_params_compat(&P);
        for (size_t i = 0; i < q.size(); i++) {
            q[i].position[0] = (float)(i * 997 % 4000) * 1000.0f;
            c[i].type = i < 5 ? 0 : (i < 205 ? 1 : 2);
        }
        std::vector<float> f(4 * q.size());
        orc_stats st{};
        CHECK(orc_step(q.data(), w.data(), c.data(), 300, 205, &P, 1.0f / 30, f.data(), &st, 1,
                       nullptr, nullptr) == 0);
    }
    std::printf("%s\n", fails ? "failed" : "ok");
    return fails ? 1 : 0;
}
