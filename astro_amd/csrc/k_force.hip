// k_force.hip -- Barnes-Hut force traversal and integrator for gfx950.
//
// k_force_compact: the product traversal over 32-B compact nodes.  Targets are processed in
// Morton order (sorted position j), so the 64 lanes of a wavefront walk nearly the same part
// of the tree; blocks are mapped so each XCD (and its L2) takes one contiguous stretch of the
// Morton curve.  The stack holds (first child, remaining count) runs, one per open level,
// so its depth is bounded by the layer count (<= 22) instead of 7 per level; it lives in LDS
// in a [depth][thread] layout.  Visit order equals the reference's LIFO order (children
// pushed 0..7, popped 7..0), so per-target sums are bit-identical to the oracle.
//
// k_force_reference: the reference kernel shape (BarnesHut.metal:22-101) over the 128-B
// OctreeNode array with children[8], for callers that built the tree with the stage calls.
#include "launch.hpp"

namespace astro {

constexpr uint32_t kForceThreads = 256;
constexpr uint32_t kStackDepth = 24;

struct Acc3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
};

template <bool KAHAN>
__device__ inline void accumulate(Acc3& F, float s, float dx, float dy, float dz) {
    if (KAHAN) {
        float y = __builtin_fmaf(s, dx, -F.cx);
        float t = F.x + y;
        F.cx = (t - F.x) - y;
        F.x = t;
        y = __builtin_fmaf(s, dy, -F.cy);
        t = F.y + y;
        F.cy = (t - F.y) - y;
        F.y = t;
        y = __builtin_fmaf(s, dz, -F.cz);
        t = F.z + y;
        F.cz = (t - F.z) - y;
        F.z = t;
    } else {
        F.x = __builtin_fmaf(s, dx, F.x);
        F.y = __builtin_fmaf(s, dy, F.y);
        F.z = __builtin_fmaf(s, dz, F.z);
    }
}

__device__ inline float clampf(float v, float c) {
    if (!(c > 0.0f)) return v;
    return v < -c ? -c : (v > c ? c : v);
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

template <bool KAHAN, bool STATS>
__global__ __launch_bounds__(kForceThreads) void k_force_compact(ForceArgs a) {
    __shared__ uint32_t s_stack[kStackDepth * kForceThreads];

    // XCD-aware mapping: hardware places block b on XCD b % 8; give XCD x the x-th
    // contiguous eighth of the target tiles.
    const uint32_t nb = gridDim.x, b = blockIdx.x, xcd = b & 7u, q = nb >> 3, r = nb & 7u;
    const uint32_t tile = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    const uint32_t tid = threadIdx.x;
    const uint32_t j = tile * kForceThreads + tid;

    unsigned long long visits = 0, leafs = 0, monos = 0, opens = 0;
    if (j < a.ns) {
        const uint32_t pidx = j < a.nt ? a.sorted_idx[j] : j;
        const float4 p = a.pos[pidx];
        const uint32_t type = a.type[pidx];
        const uint32_t self =
            (type != 2u && j < a.nt) ? (uint32_t)(a.off0 + a.leaf_of_sorted[j]) : 0xFFFFFFFFu;
        const uint32_t sys = a.sys_of ? a.sys_of[j] : 0u;
        float side0;
        if (a.p.side_rule == ASTRO_SIDE_SHIFTED) side0 = 2097152.0f;  // 2^21
        else side0 = ldexpf(a.cubes[sys].w, -(int)a.p.bits_per_axis);
        const float eps2 = a.p.eps2, r2f = a.p.r2_floor, rf = a.p.r_floor;
        const float cut = a.p.leaf_cutoff_r2, theta = a.p.theta, G = a.p.G;

        Acc3 F;
        uint32_t top_first = (uint32_t)(a.root_off + sys), top_cnt = 1, sp = 0;
        while (true) {
            if (top_cnt == 0) {
                if (sp == 0) break;
                sp--;
                const uint32_t e = s_stack[sp * kForceThreads + tid];
                top_first = e >> 3;
                top_cnt = (e & 7u) + 1u;
            }
            const uint32_t n = top_first + top_cnt - 1u;
            top_cnt--;
            const float4* np = reinterpret_cast<const float4*>(a.tnodes + n);
            const float4 cm = np[0];
            const uint4 li = reinterpret_cast<const uint4*>(np)[1];
            if (STATS) visits++;
            const float dx = cm.x - p.x, dy = cm.y - p.y, dz = cm.z - p.z;
            float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) + eps2;
            r2 = fmaxf(r2, r2f);
            const float rr = fmaxf(__fsqrt_rn(r2), rf);
            const uint32_t layer = tn_layer(li.y);
            bool interact;
            if (layer == 0u) {
                interact = (n != self) && (r2 > cut);
                if (STATS) leafs += interact;
            } else {
                const float side = ldexpf(side0, (int)layer);
                interact = __fdiv_rn(side, rr) < theta;
                if (interact) {
                    if (STATS) monos++;
                } else {
                    if (STATS) opens++;
                    if (top_cnt > 0u) {
                        if (sp >= kStackDepth) {
                            atomicOr(a.err, ERR_STACK);
                            break;
                        }
                        s_stack[sp * kForceThreads + tid] = (top_first << 3) | (top_cnt - 1u);
                        sp++;
                    }
                    top_first = li.x;
                    top_cnt = tn_nchild(li.y);
                }
            }
            if (interact) {
                const float gm = G * cm.w;
                const float s = __fdiv_rn(gm, r2 * rr);
                accumulate<KAHAN>(F, s, dx, dy, dz);
            }
        }
        const float c = a.p.force_clamp;
        a.force[pidx] = make_float4(clampf(F.x, c), clampf(F.y, c), clampf(F.z, c), 0.0f);
    }
    if (STATS) {
        visits = wave_sum_u64(visits);
        leafs = wave_sum_u64(leafs);
        monos = wave_sum_u64(monos);
        opens = wave_sum_u64(opens);
        if ((tid & 63u) == 0u) {
            atomicAdd(&a.stats[0], visits);
            atomicAdd(&a.stats[1], leafs);
            atomicAdd(&a.stats[2], monos);
            atomicAdd(&a.stats[3], opens);
        }
    }
}

void launch_force_compact(const ForceArgs& a, hipStream_t s) {
    if (!a.ns) return;
    const uint32_t blocks = (a.ns + kForceThreads - 1) / kForceThreads;
    const bool st = a.stats != nullptr;
    if (a.p.kahan) {
        if (st) hipLaunchKernelGGL((k_force_compact<true, true>), dim3(blocks), dim3(kForceThreads), 0, s, a);
        else hipLaunchKernelGGL((k_force_compact<true, false>), dim3(blocks), dim3(kForceThreads), 0, s, a);
    } else {
        if (st) hipLaunchKernelGGL((k_force_compact<false, true>), dim3(blocks), dim3(kForceThreads), 0, s, a);
        else hipLaunchKernelGGL((k_force_compact<false, false>), dim3(blocks), dim3(kForceThreads), 0, s, a);
    }
}

// BarnesHut.metal:22-101 over OctreeNode (children[8], 128 B), one target per thread in
// particle order.  Stack depth 7*(layers-1)+1 entries per thread in dynamic LDS.
__global__ __launch_bounds__(64) void k_force_reference(ForceArgs a, uint32_t depth) {
    extern __shared__ uint32_t s_stack[];
    const uint32_t tid = threadIdx.x;
    const uint32_t gid = blockIdx.x * 64u + tid;
    if (gid >= a.ns) return;
    const float4 p = a.pos[gid];
    const uint32_t type = a.type[gid];
    const bool has_self = type != 2u && gid < a.nt;
    const uint64_t mycode = has_self ? a.codes[gid] : ASTRO_INVALID_MORTON_CODE;
    const float side0 = a.p.side_rule == ASTRO_SIDE_SHIFTED
                            ? 2097152.0f
                            : ldexpf(a.cubes[0].w, -(int)a.p.bits_per_axis);
    Acc3 F;
    uint32_t sp = 0;
    s_stack[tid] = (uint32_t)a.root_off;
    sp = 1;
    while (sp > 0) {
        sp--;
        const uint32_t ni = s_stack[sp * 64u + tid];
        const OctreeNode* nd = &a.onodes[ni];
        if (nd->mortonCode == ASTRO_INVALID_MORTON_CODE) continue;
        const float dx = nd->centerOfMass[0] - p.x;
        const float dy = nd->centerOfMass[1] - p.y;
        const float dz = nd->centerOfMass[2] - p.z;
        float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) + a.p.eps2;
        r2 = fmaxf(r2, a.p.r2_floor);
        const float rr = fmaxf(__fsqrt_rn(r2), a.p.r_floor);
        const uint32_t layer = nd->layer;
        bool interact;
        if (layer == 0u) {
            const bool self = has_self && mycode == nd->mortonCode;
            interact = !self && r2 > a.p.leaf_cutoff_r2;
        } else {
            const float side = ldexpf(side0, (int)layer);
            interact = __fdiv_rn(side, rr) < a.p.theta;
            if (!interact) {
                for (int c = 0; c < 8; c++) {
                    const uint32_t ch = nd->children[c];
                    if (ch == ASTRO_INVALID_CHILD) continue;
                    if (sp < depth) s_stack[sp++ * 64u + tid] = ch;
                    else atomicOr(a.err, ERR_STACK);
                }
            }
        }
        if (interact) {
            const float s = __fdiv_rn(a.p.G * nd->totalMass, r2 * rr);
            if (a.p.kahan) accumulate<true>(F, s, dx, dy, dz);
            else accumulate<false>(F, s, dx, dy, dz);
        }
    }
    const float c = a.p.force_clamp;
    a.force[gid] = make_float4(clampf(F.x, c), clampf(F.y, c), clampf(F.z, c), 0.0f);
}

void launch_force_reference(const ForceArgs& a, hipStream_t s) {
    if (!a.ns) return;
    const uint32_t layers = a.p.bits_per_axis + 1u;
    const uint32_t depth = 7u * (layers - 1u) + 2u;
    hipLaunchKernelGGL(k_force_reference, dim3((a.ns + 63) / 64), dim3(64),
                       depth * 64u * sizeof(uint32_t), s, a, depth);
}

// BarnesHut.metal:169-196 (updateSpheres): a = F*scale[type]; v = v*damping + a*dt;
// x = x + v*dt -- separate roundings, no contraction.
__global__ __launch_bounds__(256) void k_update(float4* __restrict__ pos, float4* __restrict__ vel,
                                                const float4* __restrict__ force,
                                                const uint32_t* __restrict__ type, uint32_t n,
                                                float dt, astro_params p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = type[i];
    const float sc = t < 3u ? p.type_scale[t] : 1.0f;
    const float4 f = force[i];
    float4 v = vel[i];
    float4 x = pos[i];
    float ax = f.x * sc, ay = f.y * sc, az = f.z * sc;
    v.x = v.x * p.damping;
    v.y = v.y * p.damping;
    v.z = v.z * p.damping;
    v.x = v.x + ax * dt;
    v.y = v.y + ay * dt;
    v.z = v.z + az * dt;
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    vel[i] = v;
    pos[i] = x;
}

void launch_update(float4* pos, float4* vel, const float4* force, const uint32_t* type,
                   uint32_t n, float dt, const astro_params& p, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_update, dim3((n + 255) / 256), dim3(256), 0, s, pos, vel, force, type, n,
                       dt, p);
}

}  // namespace astro
