// accept.hpp -- the opening test of the traversal without a division (host and device).
//
// The reference opens a node unless  side / r < theta  (BarnesHut.metal:75), evaluated here as
// RN(side / r) < theta with side = side0 * 2^layer.  Because RN is monotone, the test holds
// exactly when  side < m * r,  or side == m * r with the tie rounding down, where m is the
// midpoint between theta and its float predecessor a (ties round to even, so the tie counts
// when a's last mantissa bit is 0).  m has at most 26 significant bits and r 24, so m * r is
// exact in double and the comparison is exact.  The test is monotone in r and scales exactly
// with 2^layer, so for each target it reduces to  r >= rmin0 * 2^layer,  where rmin0 is the
// smallest float r that passes for side0.  The traversal therefore decides every node with one
// float compare and takes exactly the oracle's decisions.
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstring>

#if defined(__HIPCC__)
#define ASTRO_HD __host__ __device__
#else
#define ASTRO_HD
#endif

namespace astro {

struct AcceptRule {
    double thr;    // midpoint between theta and its float predecessor
    uint32_t tie;  // the midpoint itself passes
    uint32_t mode; // 0 normal, 1 never accept (theta <= 0 or NaN), 2 always (theta = +inf)
};

inline AcceptRule make_accept_rule(float theta) {
    AcceptRule r{0.0, 0u, 0u};
    if (!(theta > 0.0f)) {
        r.mode = 1;
        return r;
    }
    if (std::isinf(theta)) {
        r.mode = 2;
        return r;
    }
    const float pred = std::nextafter(theta, 0.0f);
    uint32_t bits;
    std::memcpy(&bits, &pred, sizeof(bits));
    r.thr = ((double)pred + (double)theta) * 0.5;
    r.tie = (bits & 1u) == 0u ? 1u : 0u;
    return r;
}

ASTRO_HD inline float f32_from_bits(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, sizeof(f));
    return f;
}
ASTRO_HD inline uint32_t bits_from_f32(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, sizeof(u));
    return u;
}

// Exact  RN(side / r) < theta  for side, r > 0 finite (normal).
ASTRO_HD inline bool accept_exact(float side, float r, const AcceptRule& a) {
    if (a.mode == 1) return false;
    if (a.mode == 2) return true;
    const double P = a.thr * (double)r;
    const double S = (double)side;
    return S < P || (a.tie && S == P);
}

// Smallest positive float r with accept_exact(side0, r); +inf when nothing passes.
ASTRO_HD inline float accept_rmin(float side0, const AcceptRule& a) {
    if (a.mode == 1) return INFINITY;
    if (a.mode == 2) return 0.0f;
    float r = (float)((double)side0 / a.thr);  // within an ulp or two of the answer
    for (int k = 0; k < 8 && r > 0.0f; k++) {
        const float down = f32_from_bits(bits_from_f32(r) - 1u);
        if (!accept_exact(side0, down, a)) break;
        r = down;
    }
    for (int k = 0; k < 8 && !accept_exact(side0, r, a); k++) r = f32_from_bits(bits_from_f32(r) + 1u);
    return r;
}

// Smallest float x with RN(sqrt(x)) >= T, for T = 0, +inf or T >= 2^-62 finite.  RN(sqrt(x))
// reaches T exactly when sqrt(x) exceeds the midpoint m of T and its float predecessor (the
// square root of a float never equals m: m has 25 significant bits ending in a 1, so m^2 would
// need more than 48), i.e. when x > m^2, i.e. x >= RU(m^2); m^2 is exact in double.  Scaling
// T by 2^l scales the result by 4^l exactly while both stay normal, so for a traversal with
// eps2 >= 2^-96 the test  RN(sqrt(r2)) >= rmin0 * 2^l  is  r2 >= accept_r2min(rmin0) * 4^l.
// Returns 0 for T = 0 and +inf for T = +inf or when RU(m^2) exceeds the float range.
ASTRO_HD inline float accept_r2min(float T) {
    if (T == 0.0f) return 0.0f;
    if (!(T < INFINITY)) return INFINITY;
    const float pred = f32_from_bits(bits_from_f32(T) - 1u);
    const double m = ((double)pred + (double)T) * 0.5;
    const double m2 = m * m;
    float c = (float)m2;
    if ((double)c < m2) c = f32_from_bits(bits_from_f32(c) + 1u);
    return c;
}
// The r2 form of the opening test applies to rmin0 (accept_r2min's domain).
ASTRO_HD inline bool accept_r2_ok(float rmin0) {
    return rmin0 == 0.0f || !(rmin0 < INFINITY) || rmin0 >= 0x1p-62f;
}

}  // namespace astro
