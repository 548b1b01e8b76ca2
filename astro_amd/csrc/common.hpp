// common.hpp -- device-side types and helpers shared by the libastro kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/astro/bridge.h"

namespace astro {

// Compact traversal node, 32 B, stored as two parallel 16-B arrays: the centre and mass
// (float4 {x, y, z, m}) and the links (TLink).  Children of a node are a contiguous run of
// the layer below (the build writes each layer in ascending key order), so a (last, count)
// pair replaces the reference's children[8] (Astro/bridge.h:62).  Split, a traversal step
// reads 16 B from each array; the 64 lanes of a wave then touch fewer distinct cache lines
// per load than with 32-B nodes, and the two loads share one index-to-offset shift.
//
// `next` threads the tree in the reference's visiting order (BarnesHut.metal:22-101 pushes
// children 0..7 on a LIFO stack, so the last child is visited first): it is the node visited
// after this node's subtree, i.e. the previous sibling, or for a first child its parent's
// next (kNextEnd after the root).  A traversal then needs no stack: an opened node goes to its
// last child, any other to its next.  Every node that starts at sorted position s has the
// same next: the node of layer h(s) (the highest layer at which s starts a node) that holds
// position s - 1, or kNextEnd where s starts a system.
struct alignas(16) TLink {
    uint32_t last;      // absolute index of the last child (leaves: ASTRO_INVALID_CHILD)
    uint32_t info;      // layer | nchild << 5 | head level << 9
    uint32_t next;      // the node after this subtree in visiting order (kNextEnd: none)
    uint32_t start;     // sorted position of the node's first particle (key = sorted key >> 3 layer)
};
static_assert(sizeof(TLink) == 16, "TLink must be 16 B");
// One node in registers (the build's working form).
struct TNode {
    float x, y, z, m;   // centre of mass, total mass
    uint32_t last, info, next, start;
};
constexpr uint32_t kNextEnd = 0xFFFFFFFFu;
// Quadrupole trees of fewer than kLinkPackNodes nodes also keep a node's links in the two
// free words of its second moment row (Sxz, Syz, last | layer << 27, next; leaves:
// kPackedNoChild), so that a quadrupole traversal step reads three 16-B rows, not four.
constexpr uint64_t kLinkPackNodes = 1ull << 27;
constexpr uint32_t kPackedIndexMask = 0x07FFFFFFu;
constexpr uint32_t kPackedNoChild = kPackedIndexMask;

__host__ __device__ inline uint32_t tn_info(uint32_t layer, uint32_t nchild, uint32_t h) {
    return layer | (nchild << 5) | (h << 9);
}
__host__ __device__ inline uint32_t tn_layer(uint32_t info) { return info & 31u; }
__host__ __device__ inline uint32_t tn_nchild(uint32_t info) { return (info >> 5) & 15u; }
__host__ __device__ inline uint32_t tn_level(uint32_t info) { return (info >> 9) & 31u; }

// Error word bits raised by kernels (read back by astro_check).
enum : uint32_t {
    ERR_SPIN = 1u,       // reserved (no kernel waits on another block)
    ERR_STACK = 2u,      // stage-path traversal stack overflow
    ERR_CAPACITY = 4u,   // a layer had more nodes than its capacity
    ERR_NODE_INDEX = 8u, // a traversal met a node index outside the tree
};

constexpr uint32_t kMaxLayers = 24;   // B <= 21 -> 22 layers

// Popcount of the lanes below this one in a 64-lane mask.
__device__ inline uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A workgroup barrier that orders LDS only.  __syncthreads also waits for the block's global
// stores to complete; a kernel that never reads its own stores back lets them drain in the
// background across this barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Correctly rounded sqrt for normal x >= 2^-96: the hardware estimate corrected by one ulp
// with two fused residuals -- the same correction the compiler emits for sqrtf, without the
// rescaling it adds for tiny inputs.
__device__ inline float sqrt_rn_normal(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rm = __builtin_fmaf(-sm, s, x);
    const float rp = __builtin_fmaf(-sp, s, x);
    s = rm <= 0.0f ? sm : s;
    s = rp > 0.0f ? sp : s;
    return s;
}

// RN(sqrt(x)) for finite x >= 2^-96 in five operations: one Goldschmidt step from the hardware
// reciprocal square root y ~ 1/sqrt(x) (g = x y, residual x - g^2 in one fma, g + residual y/2).
// Not an error bound but an exhaustive result: on gfx950 it equals the correctly rounded square
// root for every float of [2^-96, FLT_MAX] (k_check_sqrt checks all of them).  x = +inf gives
// NaN, so callers use it only where x is bounded.
__device__ inline float sqrt_rn_rsq(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float g = x * y, h = 0.5f * y;
    const float r = __builtin_fmaf(-g, g, x);
    return __builtin_fmaf(r, h, g);
}

// Inclusive scan across the 64 lanes of a wave.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// The NW wave totals tmp[0 .. NW) (NW <= 64) replaced by their exclusive scan, and their sum
// stored in tmp[NW], by the block's first wave in one shuffle scan.
template <int NW>
__device__ inline void scan_wave_totals(uint32_t* tmp) {
    static_assert(NW >= 1 && NW <= 64, "one wave scans the totals");
    if ((threadIdx.x >> 6) == 0) {
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t t = lane < (uint32_t)NW ? tmp[lane] : 0u;
        const uint32_t inc = wave_inclusive_scan(t);
        if (lane < (uint32_t)NW) tmp[lane] = inc - t;
        if (lane == (uint32_t)NW - 1u) tmp[NW] = inc;
    }
}

// Exclusive scan over a block of NT threads (NT multiple of 64, <= 1024).  `tmp` is an LDS
// array of NT/64 + 1 words.  Returns the exclusive prefix; *total receives the block sum.
template <int NT>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) tmp[wave] = inc;
    __syncthreads();
    scan_wave_totals<NW>(tmp);
    __syncthreads();
    uint32_t r = tmp[wave] + inc - v;
    *total = tmp[NW];
    __syncthreads();
    return r;
}

// One pass, two results: the block-wide sum of r, and the exclusive scan (and total) of v.
template <int NT>
__device__ inline uint32_t block_sum_and_scan(uint32_t r, uint32_t v, uint32_t* tmp,
                                              uint32_t* rsum, uint32_t* total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    const uint32_t rinc = wave_inclusive_scan(r);
    if (lane == 63) {
        tmp[wave] = inc;
        tmp[NW + 1 + wave] = rinc;
    }
    __syncthreads();
    if (wave == 0) {  // both totals by the first wave: a scan of v's, a sum of r's
        const uint32_t t = lane < NW ? tmp[lane] : 0u;
        const uint32_t rt = lane < NW ? tmp[NW + 1 + lane] : 0u;
        const uint32_t ti = wave_inclusive_scan(t), ri = wave_inclusive_scan(rt);
        if (lane < NW) tmp[lane] = ti - t;
        if (lane == NW - 1) {
            tmp[NW] = ti;
            tmp[2 * NW + 1] = ri;
        }
    }
    __syncthreads();
    const uint32_t out = tmp[wave] + inc - v;
    *total = tmp[NW];
    *rsum = tmp[2 * NW + 1];
    __syncthreads();
    return out;
}

// Float <-> order-preserving unsigned key, for atomic min/max of floats.
__device__ inline uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float ord2f(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

}  // namespace astro
