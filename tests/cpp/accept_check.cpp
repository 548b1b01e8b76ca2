// Checks astro::accept_rmin against the division it replaces: for side = side0 * 2^l,
//   RN(side / r) < theta   <=>   r >= ldexp(accept_rmin(side0), l),
// and astro::accept_r2min against the square root it replaces: for r = RN(sqrt(r2)),
//   RN(side / r) < theta   <=>   r2 >= ldexp(accept_r2min(accept_rmin(side0)), 2 l).
// Prints the disagreements of the first form (expected 0), the near-boundary cases and all
// cases tried, then the disagreements and cases of the second form.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../astro_amd/csrc/accept.hpp"

int main() {
    std::mt19937_64 rng(12345);
    const float thetas[] = {0.7f, 0.5f, 0.55f, 0.6f, 0.3f, 1.0f, 0.7000001f, 0.9999999f, 0.25f, 1.7f};
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
    long bad2 = 0, total2 = 0;
    for (float th : thetas) {
        const astro::AcceptRule rule = astro::make_accept_rule(th);
        for (int it = 0; it < 200000; it++) {
            const int e0 = (int)(rng() % 60) - 40;
            const float side0 = std::ldexp((float)(1.0 + (rng() % 1000) / 1000.0), e0);
            const int l = (int)(rng() % 22);
            const float side = std::ldexp(side0, l);
            const float rmin0 = astro::accept_rmin(side0, rule);
            if (!astro::accept_r2_ok(rmin0)) continue;
            const float c = std::ldexp(astro::accept_r2min(rmin0), 2 * l);
            float r2;
            if (it & 1) {
                r2 = std::ldexp((float)(0.1 + (rng() % 100000) / 1000.0), 2 * (e0 + l));
            } else {  // within a few ulps of the squared threshold
                r2 = astro::f32_from_bits(astro::bits_from_f32(c) + (uint32_t)((int)(rng() % 17) - 8));
            }
            if (!(r2 >= 0x1p-96f) || !std::isfinite(r2)) continue;
            const float r = std::sqrt(r2);
            const bool want = side / r < th;
            const bool by_r = r >= std::ldexp(rmin0, l);
            const bool got = r2 >= c;
            bad2 += (want != got) || (by_r != got);
            total2++;
        }
    }
    std::printf("%ld %ld %ld %ld %ld\n", bad, near, total, bad2, total2);
    return 0;
}
