// Checks astro::accept_rmin against the division it replaces: for side = side0 * 2^l,
//   RN(side / r) < theta   <=>   r >= ldexp(accept_rmin(side0), l).
// Prints the number of disagreements (expected 0) and of near-boundary cases tried.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../astro_amd/csrc/accept.hpp"

int main() {
    std::mt19937_64 rng(12345);
    const float thetas[] = {0.7f, 0.5f, 0.3f, 1.0f, 0.7000001f, 0.9999999f, 0.25f, 1.7f};
    long bad = 0, near = 0, total = 0;
    for (float th : thetas) {
        const astro::AcceptRule rule = astro::make_accept_rule(th);
        for (int it = 0; it < 200000; it++) {
            const int e0 = (int)(rng() % 60) - 40;        // side0 = 2^e0
            const float side0 = std::ldexp(1.0f, e0);
            const int l = (int)(rng() % 22);
            const float side = std::ldexp(side0, l);
            float r;
            if (it & 1) {
                r = std::ldexp((float)(0.1 + (rng() % 100000) / 1000.0), e0 + l);
            } else {  // within a few ulps of side / theta
                r = (float)((double)side / (double)th);
                uint32_t b = astro::bits_from_f32(r) + (uint32_t)((int)(rng() % 9) - 4);
                r = astro::f32_from_bits(b);
                near++;
            }
            const float q = side / r;
            const bool want = q < th;
            const bool got = r >= std::ldexp(astro::accept_rmin(side0, rule), l);
            bad += want != got;
            total++;
        }
    }
    std::printf("%ld %ld %ld\n", bad, near, total);
    return 0;
}
