// morton.hpp -- the Morton key of a position (MortonCode.metal:26-52), shared by the key kernel
// and the seeded sort, which computes the keys itself while it places the sources.
#pragma once

#include "common.hpp"

namespace astro {

// Spread the low 21 bits of v to every third bit (bit i -> bit 3i).
__device__ inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
// float -> u32: NaN and negatives to 0, saturating, truncating (the oracle's definition).
__device__ inline uint32_t cvt_u32_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

// The key of position x: quantised against its system's cube (cubes[sys]: {lo, side}, read only
// by the cube quantizer) or by the offset-wrap quantizer, interleaved, top 3B bits kept.
__device__ inline uint64_t morton_key(float4 x, const float4* __restrict__ cubes, uint32_t sys,
                                      const astro_params& p) {
    uint32_t q[3];
    if (p.quantizer == ASTRO_QUANT_OFFSET_WRAP) {
        q[0] = cvt_u32_sat(x.x + p.quant_offset);
        q[1] = cvt_u32_sat(x.y + p.quant_offset);
        q[2] = cvt_u32_sat(x.z + p.quant_offset);
    } else {
        float4 c = cubes[sys];
        int e;
        frexpf(c.w, &e);
        const float scale = ldexpf(1.0f, 21 - (e - 1));
        const float v[3] = {x.x, x.y, x.z}, lo[3] = {c.x, c.y, c.z};
        for (int a = 0; a < 3; a++) {
            uint32_t t = cvt_u32_sat((v[a] - lo[a]) * scale);
            q[a] = t > 0x1FFFFFu ? 0x1FFFFFu : t;
        }
    }
    return (spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2)) >>
           (63 - 3 * p.bits_per_axis);
}

}  // namespace astro
