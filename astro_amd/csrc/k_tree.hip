// k_tree.hip -- bounding cube, Morton keys and the layered leaf-to-root octree for gfx950.
//
// The reference builds each layer with clears, an identity re-sort, mark/scan/scatter of
// segment heads, a count copy and an aggregation dispatch (Renderer.swift:570-755).  Here a
// layer is one kernel (plus a count kernel for the leaf layer): a node starts a parent when its
// key differs from its predecessor's above bit 3l (its "head level" byte is >= l); a tile
// prefix and scan give the parent index, and the head thread aggregates its (at most 8)
// contiguous children in the reference's serial order.  Aggregation order and arithmetic match
// the oracle exactly.
#include "launch.hpp"
#include "morton.hpp"

namespace astro {

// ------------------------------------------------------------------------- bounding cube
__global__ void k_bbox_init(uint32_t* mm, uint32_t nsys) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsys * 6) mm[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}
void launch_bbox_init(uint32_t* minmax, uint32_t nsys, hipStream_t s) {
    hipLaunchKernelGGL(k_bbox_init, dim3((nsys * 6 + 255) / 256), dim3(256), 0, s, minmax, nsys);
}

constexpr uint32_t kBoxTile = 4096;
// positions per block of the single-system cube (1,024 and 2,048 measured slower: the
// finishing atomics of more blocks cost more than the read they spread)
constexpr uint32_t kCubeTile = 4096;

__device__ inline float wave_min(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

// Per-system min/max of the positions in one tile; one atomic per (tile, system, value).
__global__ __launch_bounds__(256) void k_bbox(const float4* __restrict__ pos, uint32_t n,
                                              const uint32_t* __restrict__ sys_of,
                                              uint32_t* mm) {
    __shared__ float s_red[6][4];
    const uint64_t base = (uint64_t)blockIdx.x * kBoxTile;
    if (base >= n) return;
    const uint64_t end = base + kBoxTile < n ? base + kBoxTile : n;
    const uint32_t s0 = sys_of ? sys_of[base] : 0u;
    const uint32_t s1 = sys_of ? sys_of[end - 1] : 0u;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t sys = s0; sys <= s1; sys++) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint64_t e = base + threadIdx.x; e < end; e += 256) {
            if (sys_of && sys_of[e] != sys) continue;
            float4 p = pos[e];
            lo[0] = fminf(lo[0], p.x);
            lo[1] = fminf(lo[1], p.y);
            lo[2] = fminf(lo[2], p.z);
            hi[0] = fmaxf(hi[0], p.x);
            hi[1] = fmaxf(hi[1], p.y);
            hi[2] = fmaxf(hi[2], p.z);
        }
        for (int a = 0; a < 3; a++) {
            lo[a] = wave_min(lo[a]);
            hi[a] = wave_max(hi[a]);
        }
        if (lane == 0)
            for (int a = 0; a < 3; a++) {
                s_red[a][wave] = lo[a];
                s_red[3 + a][wave] = hi[a];
            }
        __syncthreads();
        if (threadIdx.x < 6) {
            const int a = threadIdx.x;
            float v = s_red[a][0];
            for (int w = 1; w < 4; w++) v = a < 3 ? fminf(v, s_red[a][w]) : fmaxf(v, s_red[a][w]);
            if (a < 3) {
                if (v != INFINITY) atomicMin(&mm[sys * 6 + a], f2ord(v));
            } else {
                if (v != -INFINITY) atomicMax(&mm[sys * 6 + a], f2ord(v));
            }
        }
        __syncthreads();
    }
}
void launch_bbox(const float4* pos, uint32_t n, const uint32_t* sys_of, uint32_t* minmax,
                 hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_bbox, dim3((n + kBoxTile - 1) / kBoxTile), dim3(256), 0, s, pos, n,
                       sys_of, minmax);
}

// cube = {lo.x, lo.y, lo.z, side} from the ordered min/max keys k[0..2] / k[3..5]:
// side = smallest power of two >= the largest extent; an empty set gives {0, 0, 0, 1}.
__device__ inline float4 cube_from_keys(const uint32_t* k) {
    float lo[3], hi[3];
    bool empty = false;
    for (int a = 0; a < 3; a++) {
        if (k[a] == 0xFFFFFFFFu || k[3 + a] == 0u) empty = true;
        lo[a] = ord2f(k[a]);
        hi[a] = ord2f(k[3 + a]);
    }
    if (empty)
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = 0.0f;
    float ext = 0.0f;
    for (int a = 0; a < 3; a++) {
        float e = hi[a] - lo[a];
        ext = e > ext ? e : ext;
    }
    float side = 1.0f;
    if (ext > 0.0f && isfinite(ext)) {
        int e;
        float m = frexpf(ext, &e);
        side = ldexpf(1.0f, m == 0.5f ? e - 1 : e);
    }
    return make_float4(lo[0], lo[1], lo[2], side);
}

__global__ void k_cube_finalize(const uint32_t* mm, float4* cubes, uint32_t nsys) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    uint32_t k[6];
    for (int a = 0; a < 6; a++) k[a] = mm[s * 6 + a];
    cubes[s] = cube_from_keys(k);
}
void launch_cube_finalize(const uint32_t* minmax, float4* cubes, uint32_t nsys, hipStream_t s) {
    hipLaunchKernelGGL(k_cube_finalize, dim3((nsys + 63) / 64), dim3(64), 0, s, minmax, cubes,
                       nsys);
}

__device__ inline uint32_t wave_umin(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, d, 64));
    return v;
}
__device__ inline uint32_t wave_umax(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
    return v;
}
// Min (k[0..2]) / max (k[3..5]) of the block's ordered keys into s_red[6][4]; result in k of
// thread 0.
__device__ inline void block_minmax(uint32_t* k, uint32_t (*s_red)[4]) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int a = 0; a < 3; a++) {
        k[a] = wave_umin(k[a]);
        k[3 + a] = wave_umax(k[3 + a]);
    }
    if (lane == 0)
        for (int a = 0; a < 6; a++) s_red[a][wave] = k[a];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int a = 0; a < 6; a++)
            for (int w = 1; w < 4; w++)
                k[a] = a < 3 ? min(k[a], s_red[a][w]) : max(k[a], s_red[a][w]);
}

// One system's cube in one launch: each block reduces a tile of kBoxTile positions to ordered
// min/max keys (the same total order as the atomics of k_bbox) and stores them in part; the
// last block to finish (counted on *counter, zero before the launch and zeroed again by that
// block) reduces the partials and writes cubes[0].  Plain stores and one atomic per block
// replace the 6 same-line atomics per tile of k_bbox.
__global__ __launch_bounds__(256) void k_cube_single(const float4* __restrict__ pos, uint32_t n,
                                                     uint32_t* part, uint32_t* counter,
                                                     float4* cubes) {
    __shared__ uint32_t s_red[6][4];
    __shared__ uint32_t s_last;
    const uint64_t base = (uint64_t)blockIdx.x * kCubeTile;
    const uint64_t end = base + kCubeTile < n ? base + kCubeTile : n;
    uint32_t k[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    // every load of the tile in flight at once (the kernel is one memory round trip long)
#pragma unroll 16
    for (uint64_t e = base + threadIdx.x; e < end; e += 256) {
        const float4 p = pos[e];
        const uint32_t x = f2ord(p.x), y = f2ord(p.y), z = f2ord(p.z);
        k[0] = min(k[0], x);
        k[1] = min(k[1], y);
        k[2] = min(k[2], z);
        k[3] = max(k[3], x);
        k[4] = max(k[4], y);
        k[5] = max(k[5], z);
    }
    block_minmax(k, s_red);
    if (threadIdx.x == 0) {
        for (int a = 0; a < 6; a++) part[blockIdx.x * 8 + a] = k[a];
        const uint32_t done =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = done + 1 == gridDim.x ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    for (int a = 0; a < 6; a++) k[a] = a < 3 ? 0xFFFFFFFFu : 0u;
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += 256)
        for (int a = 0; a < 6; a++) {
            const uint32_t v = part[b * 8 + a];
            k[a] = a < 3 ? min(k[a], v) : max(k[a], v);
        }
    __syncthreads();  // s_red is reused
    block_minmax(k, s_red);
    if (threadIdx.x == 0) {
        cubes[0] = cube_from_keys(k);
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
uint32_t cube_single_blocks(uint32_t n) { return n ? (n + kCubeTile - 1) / kCubeTile : 1u; }
void launch_cube_single(const float4* pos, uint32_t n, uint32_t* part, uint32_t* counter,
                        float4* cubes, hipStream_t s) {
    hipLaunchKernelGGL(k_cube_single, dim3(cube_single_blocks(n)), dim3(256), 0, s, pos, n, part,
                       counter, cubes);
}

// ------------------------------------------------------------------------------- Morton
// MortonCode.metal:26-52 (generateMortonCodes) + Octree.metal:39-51 (fillIndices), fused:
// key per tree source, written to the per-particle code array and the sort input, plus iota.
__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ pos, uint32_t n,
                                                const uint32_t* __restrict__ sys_of,
                                                const float4* __restrict__ cubes,
                                                astro_params p, uint64_t* __restrict__ codes,
                                                uint64_t* __restrict__ codes2,
                                                uint32_t* __restrict__ idx) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t key = morton_key(pos[i], cubes, sys_of ? sys_of[i] : 0u, p);
    if (codes) codes[i] = key;
    if (codes2) codes2[i] = key;
    if (idx) idx[i] = i;
}
void launch_morton(const float4* pos, uint32_t n, const uint32_t* sys_of, const float4* cubes,
                   const astro_params& p, uint64_t* codes, uint64_t* codes2, uint32_t* idx,
                   hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, s, pos, n, sys_of, cubes,
                       p, codes, codes2, idx);
}

// --------------------------------------------------------------------------- aggregation
// Octree.metal:84-94 / 157-172: mass += m; com = fma(x, m, com); COM = com / mass.
// A node's two halves (centre and mass; links) from / to their parallel arrays.
__device__ inline TNode load_node(const TreeArgs& a, uint64_t i) {
    const float4 p = a.tpos[i];
    const TLink k = a.tlink[i];
    return TNode{p.x, p.y, p.z, p.w, k.last, k.info, k.next, k.start};
}
__device__ inline void store_node(const TreeArgs& a, uint64_t i, const TNode& t) {
    a.tpos[i] = make_float4(t.x, t.y, t.z, t.m);
    a.tlink[i] = TLink{t.last, t.info, t.next, t.start};
}

struct Agg32 {
    float m = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    __device__ void add(float x, float y, float z, float w) {
        m = m + w;
        cx = __builtin_fmaf(x, w, cx);
        cy = __builtin_fmaf(y, w, cy);
        cz = __builtin_fmaf(z, w, cz);
    }
    __device__ void fin(float* o) const {
        o[3] = m;
        o[0] = m > 0.0f ? cx / m : 0.0f;
        o[1] = m > 0.0f ? cy / m : 0.0f;
        o[2] = m > 0.0f ? cz / m : 0.0f;
    }
};
struct Agg64 {
    double m = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    __device__ void add(float x, float y, float z, float w) {
        double dw = w;
        m += dw;
        cx = __builtin_fma((double)x, dw, cx);
        cy = __builtin_fma((double)y, dw, cy);
        cz = __builtin_fma((double)z, dw, cz);
    }
    __device__ void fin(float* o) const {
        o[3] = (float)m;
        o[0] = m > 0.0 ? (float)(cx / m) : 0.0f;
        o[1] = m > 0.0 ? (float)(cy / m) : 0.0f;
        o[2] = m > 0.0 ? (float)(cz / m) : 0.0f;
    }
};

// emittedColor / emittedColorCenter hold the second moments S (xx, yy, zz, xy | xz, yz) in
// quadrupole mode and stay zero otherwise: the physics reads nothing else from them.
__device__ inline void onode_fields(float4* q, uint64_t code, const float* cm, uint32_t first,
                                    uint32_t nchild, uint32_t layer, const float* S) {
    q[0] = make_float4(__uint_as_float((uint32_t)code), __uint_as_float((uint32_t)(code >> 32)),
                       0.0f, 0.0f);
    q[1] = make_float4(cm[0], cm[1], cm[2], 0.0f);
    q[2] = make_float4(cm[3], 0.0f, 0.0f, 0.0f);
    q[3] = S ? make_float4(S[0], S[1], S[2], S[3]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    q[4] = S ? make_float4(S[4], S[5], 0.0f, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t ch[8];
#pragma unroll
    for (int c = 0; c < 8; c++) ch[c] = (uint32_t)c < nchild ? first + c : ASTRO_INVALID_CHILD;
    q[5] = make_float4(__uint_as_float(ch[0]), __uint_as_float(ch[1]), __uint_as_float(ch[2]),
                       __uint_as_float(ch[3]));
    q[6] = make_float4(__uint_as_float(ch[4]), __uint_as_float(ch[5]), __uint_as_float(ch[6]),
                       __uint_as_float(ch[7]));
    q[7] = make_float4(__uint_as_float(layer), 0.0f, 0.0f, 0.0f);
}
__device__ inline void write_onode(OctreeNode* o, uint64_t code, const float* cm,
                                   uint32_t first, uint32_t nchild, uint32_t layer,
                                   const float* S = nullptr) {
    float4 f[8];
    onode_fields(f, code, cm, first, nchild, layer, S);
    float4* q = reinterpret_cast<float4*>(o);
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = f[k];
}

// Second moments about the node's float centre of mass P: with d = x - P,
// S_e = S_e + fma(m*d_i, d_j, Sc_e) for e = xx, yy, zz, xy, xz, yz (the oracle's order).
struct Mom32 {
    float s[6] = {0, 0, 0, 0, 0, 0};
    __device__ void add(float x, float y, float z, float w, const float* P, const float* Sc) {
        const float dx = x - P[0], dy = y - P[1], dz = z - P[2];
        const float mx = w * dx, my = w * dy, mz = w * dz;
        s[0] = s[0] + __builtin_fmaf(mx, dx, Sc ? Sc[0] : 0.0f);
        s[1] = s[1] + __builtin_fmaf(my, dy, Sc ? Sc[1] : 0.0f);
        s[2] = s[2] + __builtin_fmaf(mz, dz, Sc ? Sc[2] : 0.0f);
        s[3] = s[3] + __builtin_fmaf(mx, dy, Sc ? Sc[3] : 0.0f);
        s[4] = s[4] + __builtin_fmaf(mx, dz, Sc ? Sc[4] : 0.0f);
        s[5] = s[5] + __builtin_fmaf(my, dz, Sc ? Sc[5] : 0.0f);
    }
    __device__ void fin(float* o) const {
        for (int e = 0; e < 6; e++) o[e] = s[e];
    }
};
struct Mom64 {
    double s[6] = {0, 0, 0, 0, 0, 0};
    __device__ void add(float x, float y, float z, float w, const float* P, const float* Sc) {
        const double dx = (double)x - (double)P[0], dy = (double)y - (double)P[1],
                     dz = (double)z - (double)P[2];
        const double m = w;
        s[0] = s[0] + __builtin_fma(m * dx, dx, Sc ? (double)Sc[0] : 0.0);
        s[1] = s[1] + __builtin_fma(m * dy, dy, Sc ? (double)Sc[1] : 0.0);
        s[2] = s[2] + __builtin_fma(m * dz, dz, Sc ? (double)Sc[2] : 0.0);
        s[3] = s[3] + __builtin_fma(m * dx, dy, Sc ? (double)Sc[3] : 0.0);
        s[4] = s[4] + __builtin_fma(m * dx, dz, Sc ? (double)Sc[4] : 0.0);
        s[5] = s[5] + __builtin_fma(m * dy, dz, Sc ? (double)Sc[5] : 0.0);
    }
    __device__ void fin(float* o) const {
        for (int e = 0; e < 6; e++) o[e] = (float)s[e];
    }
};

// The moments of node n, and in a packed tree (a.pack) its links in the row's free words.
__device__ inline void store_moments(const TreeArgs& a, uint64_t n, const float* S,
                                     uint32_t last, uint32_t layer, uint32_t next) {
    const uint32_t w2 = a.pack ? ((last == ASTRO_INVALID_CHILD ? kPackedNoChild : last) |
                                  (layer << 27))
                               : 0u;
    const uint32_t w3 = a.pack ? next : 0u;
    a.qmom[2 * n] = make_float4(S[0], S[1], S[2], S[3]);
    a.qmom[2 * n + 1] = make_float4(S[4], S[5], __uint_as_float(w2), __uint_as_float(w3));
}

__device__ inline uint32_t head_level(uint64_t a, uint64_t b, uint32_t bits) {
    uint64_t x = a ^ b;  // nonzero for distinct keys
    uint32_t msb = 63u - (uint32_t)__clzll((long long)x);
    uint32_t h = msb / 3u;
    return h < bits ? h : bits;
}

// ------------------------------------------------------------------------------ leaf layer
// Octree.metal:282-318 + 53-117 fused: heads of equal-key runs in the sorted keys, their
// leaf index by tile scan, and one leaf per head aggregated serially in sorted order.
__device__ inline bool is_leaf_head(const TreeArgs& a, uint32_t j) {
    if (j == 0) return true;
    if (a.segmented && a.sys_of[j] != a.sys_of[j - 1]) return true;
    return a.sorted_keys[j] != a.sorted_keys[j - 1];
}

template <typename AGG, bool QUAD>
__device__ uint32_t emit_leaf(const TreeArgs& a, uint32_t j, uint32_t u) {
    uint32_t end = j + 1;
    while (end < a.nt && !is_leaf_head(a, end)) end++;
    // Octree.metal:71: the reference ends the last leaf's run at numUnique (= u + 1).
    if (a.quirk_last_leaf && end == a.nt) end = u + 1;
    AGG g;
    for (uint32_t i = j; i < end; i++) {
        float4 x = a.spos[i];
        g.add(x.x, x.y, x.z, x.w);
    }
    float cm[4];
    g.fin(cm);
    const uint64_t key = a.sorted_keys[j];
    uint32_t h = a.bits;
    if (j > 0 && !(a.segmented && a.sys_of[j] != a.sys_of[j - 1]))
        h = head_level(key, a.sorted_keys[j - 1], a.bits);
    const uint64_t n = a.off[0] + u;
    TNode t;
    t.x = cm[0];
    t.y = cm[1];
    t.z = cm[2];
    t.m = cm[3];
    t.last = ASTRO_INVALID_CHILD;
    t.info = tn_info(0, 0, h);
    t.next = a.pnext[j];
    t.start = j;
    store_node(a, n, t);
    a.levels[n] = (uint8_t)h;
    if (a.sysfirst && h == a.bits) a.sysfirst[(size_t)a.sys_of[j] * kMaxLayers] = u;
    float S[6];
    if (QUAD) {
        if (sizeof(AGG) == sizeof(Agg64)) {
            Mom64 q;
            for (uint32_t i = j; i < end; i++) {
                float4 x = a.spos[i];
                q.add(x.x, x.y, x.z, x.w, cm, nullptr);
            }
            q.fin(S);
        } else {
            Mom32 q;
            for (uint32_t i = j; i < end; i++) {
                float4 x = a.spos[i];
                q.add(x.x, x.y, x.z, x.w, cm, nullptr);
            }
            q.fin(S);
        }
        store_moments(a, n, S, ASTRO_INVALID_CHILD, 0u, t.next);
    }
    return h;
}

// ------------------------------------------------------------------------- layer kernels
// Every layer is built over tiles of kLayerTile candidates (sorted particles for the leaf
// layer, nodes of layer l-1 otherwise).  A tile's node offset is the sum of the head counts of
// the supertiles (kSuper tiles) before it plus those of the tiles before it in its own
// supertile; the block then scans its tile and writes one node per head.  k_layer_count
// counts the leaf layer's heads; each emit kernel counts the next layer's heads among the
// nodes it writes (their level bytes), so every layer above the leaves costs one launch.
// No block waits on another.
constexpr uint32_t kLayerItems = 1;
constexpr uint32_t kLayerTile = 256 * kLayerItems;
constexpr uint32_t kSuper = 64;
// Supertile sums sit one per 128-byte line: device-scope atomics to one line are serialized
// (about 10 ns each on MI355X), and every block of a layer adds to some supertile sum, so sums
// packed into one line would queue thousands of atomics behind each other.
constexpr uint32_t kSuperStride = 32;

__device__ inline uint32_t layer_candidates(const TreeArgs& a, uint32_t l) {
    return l == 0 ? a.nt : a.counts[l - 1];
}
__device__ inline bool layer_head(const TreeArgs& a, uint32_t l, uint32_t c) {
    if (l == 0) return is_leaf_head(a, c);
    return a.levels[a.off[l - 1] + c] >= l;
}

__global__ __launch_bounds__(256) void k_layer_count(TreeArgs a, uint32_t l,
                                                     uint32_t* tile_cnt, uint32_t* super) {
    __shared__ uint32_t s_tmp[8];
    const uint32_t n = layer_candidates(a, l);
    const uint64_t base = (uint64_t)blockIdx.x * kLayerTile;
    if (base >= n) return;
    const uint32_t c0 = (uint32_t)base + threadIdx.x * kLayerItems;
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t i = 0; i < kLayerItems; i++) {
        const uint32_t c = c0 + i;
        if (c < n && layer_head(a, l, c)) cnt++;
    }
    uint32_t total;
    block_exclusive_scan<256>(cnt, s_tmp, &total);
    if (threadIdx.x == 0) {
        tile_cnt[blockIdx.x] = total;
        if (total) atomicAdd(&super[(blockIdx.x / kSuper) * kSuperStride], total);
    }
}

// Adds the heads of layer l+1 among the parents this block writes (positions
// [first, first + total), levels in lev[]) to the next layer's tile counts and supertile sums.
// The positions span at most three tiles of kLayerTile.
__device__ inline void count_next_heads(uint32_t l, uint32_t first, uint32_t total,
                                        const uint32_t* pos, const uint32_t* lev, uint32_t k,
                                        uint32_t* s_cnt, uint32_t* next_cnt,
                                        uint32_t* next_super) {
    if (threadIdx.x < 3) s_cnt[threadIdx.x] = 0;
    lds_barrier();  // the node stores before it stay in flight
    const uint32_t t0 = first / kLayerTile;
    for (uint32_t i = 0; i < k; i++)
        if (lev[i] >= l + 1) atomicAdd(&s_cnt[pos[i] / kLayerTile - t0], 1u);
    lds_barrier();
    if (threadIdx.x < 3 && total) {
        const uint32_t c = s_cnt[threadIdx.x];
        const uint32_t t = t0 + threadIdx.x;
        if (c) {
            atomicAdd(&next_cnt[t], c);
            atomicAdd(&next_super[(t / kSuper) * kSuperStride], c);
        }
    }
}

// The leaf layer: heads of equal-key runs among the sorted particles, their leaf index by
// tile prefix + scan, one leaf per head.  next_cnt / next_super: layer 1's tile counts and
// supertile sums (zeroed), or null when the leaf layer is the top layer.
template <bool QUAD>
__global__ __launch_bounds__(256) void k_leaf_emit(TreeArgs a, const uint32_t* __restrict__ tile_cnt,
                                                   const uint32_t* __restrict__ super,
                                                   uint32_t* next_cnt, uint32_t* next_super) {
    __shared__ uint32_t s_tmp[16];
    __shared__ uint32_t s_cnt[3];
    const uint32_t n = a.nt;
    const uint32_t tile = blockIdx.x;
    const uint64_t base = (uint64_t)tile * kLayerTile;
    if (base >= n) return;
    const uint32_t sup = tile / kSuper;
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < sup; t += 256) part += super[t * kSuperStride];
    if (threadIdx.x < tile - sup * kSuper) part += tile_cnt[sup * kSuper + threadIdx.x];
    const uint32_t c0 = (uint32_t)base + threadIdx.x * kLayerItems;
    uint32_t hmask = 0, cnt = 0;
#pragma unroll
    for (uint32_t i = 0; i < kLayerItems; i++) {
        const uint32_t c = c0 + i;
        if (c < n && is_leaf_head(a, c)) {
            hmask |= 1u << i;
            cnt++;
        }
    }
    uint32_t prefix, total;
    uint32_t p = block_sum_and_scan<256>(part, cnt, s_tmp, &prefix, &total);
    p += prefix;
    if (threadIdx.x == 0 && base + kLayerTile >= n) {
        a.counts[0] = prefix + total;  // the last tile knows the layer's size
        if (prefix + total > a.cap[0]) atomicOr(a.err, ERR_CAPACITY);
    }
    uint32_t pos[kLayerItems], lev[kLayerItems], k = 0;
    const bool fp64 = a.fp64_from == 0;
    for (uint32_t i = 0; i < kLayerItems; i++) {
        const uint32_t j = c0 + i;
        if (j >= n) break;
        if (hmask & (1u << i)) {
            const uint32_t h = fp64 ? emit_leaf<Agg64, QUAD>(a, j, p) : emit_leaf<Agg32, QUAD>(a, j, p);
            pos[k] = p;
            lev[k++] = h;
            p++;
        }
        a.leaf_of_sorted[j] = p - 1;
    }
    if (next_cnt)
        count_next_heads(0, prefix, total, pos, lev, k, s_cnt, next_cnt, next_super);
}

// One parent of layer l >= 1: candidate c (a head at level >= l) of layer l - 1 starts it and
// it takes index p.  u0 is candidate c's node and lv[0..7] the level bytes of candidates c ..
// c + 7 (read as 0 past the layer's capacity); n is the candidate count.  The children are c
// and the candidates after it up to the next head (at most 8); they are aggregated serially in
// order (Octree.metal:157-172).  Returns the parent's head level.
template <bool QUAD, bool PRE = false>
__device__ __forceinline__ uint32_t emit_parent(const TreeArgs& a, uint32_t l, uint32_t c,
                                                uint32_t p, const uint32_t* lv, const TNode& u0,
                                                uint32_t n, const float4* pre = nullptr) {
    const uint64_t cb = a.off[l - 1];
    const bool fp64 = l >= a.fp64_from;
    uint32_t kc = 1;
#pragma unroll
    for (uint32_t j = 1; j < 8; j++) kc += (kc == j && c + j < n && lv[j] < l) ? 1u : 0u;
    // the children's centres: preloaded by the caller (PRE: pre[j], j < 8, indexed with
    // constants once the loops unroll, so the array stays in registers), or read here
    const float4* gch = a.tpos + cb + c;
    auto ch = [&](uint32_t j) { return PRE ? pre[j] : gch[j]; };  // (constant j when PRE)
    float cm[4];
    if (fp64) {
        Agg64 g;
        g.add(u0.x, u0.y, u0.z, u0.m);
        if constexpr (PRE) {
#pragma unroll
            for (uint32_t j = 1; j < 8; j++) {  // unrolled: the array stays in registers
                if (j >= kc) break;
                g.add(pre[j].x, pre[j].y, pre[j].z, pre[j].w);
            }
        } else {
            for (uint32_t j = 1; j < kc; j++) {
                const float4 v = gch[j];
                g.add(v.x, v.y, v.z, v.w);
            }
        }
        g.fin(cm);
    } else {
        Agg32 g;
        g.add(u0.x, u0.y, u0.z, u0.m);
        if constexpr (PRE) {
#pragma unroll
            for (uint32_t j = 1; j < 8; j++) {  // unrolled: the array stays in registers
                if (j >= kc) break;
                g.add(pre[j].x, pre[j].y, pre[j].z, pre[j].w);
            }
        } else {
            for (uint32_t j = 1; j < kc; j++) {
                const float4 v = gch[j];
                g.add(v.x, v.y, v.z, v.w);
            }
        }
        g.fin(cm);
    }
    const uint32_t level = lv[0];
    const uint64_t nd = a.off[l] + p;
    TNode t;
    t.x = cm[0];
    t.y = cm[1];
    t.z = cm[2];
    t.m = cm[3];
    t.last = (uint32_t)(cb + c + kc - 1);
    t.info = tn_info(l, kc, level);
    t.next = u0.next;  // every node starting at u0.start has the same next (pnext of it)
    t.start = u0.start;
    store_node(a, nd, t);
    a.levels[nd] = (uint8_t)level;
    if (a.sysfirst && level == a.bits)  // a system's first node of this layer
        a.sysfirst[(size_t)a.sys_of[u0.start] * kMaxLayers + l] = p;
    if (QUAD) {
        float S[6];
        const float4* cq = a.qmom + 2 * (cb + c);
        if (fp64) {
            Mom64 q;
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                if (j >= kc) break;  // (unrolled; kc <= 8)
                const float4 q0 = cq[2 * j], q1 = cq[2 * j + 1];
                const float Sc[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
                const float4 v = ch(j);
                q.add(v.x, v.y, v.z, v.w, cm, Sc);
            }
            q.fin(S);
        } else {
            Mom32 q;
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                if (j >= kc) break;  // (unrolled; kc <= 8)
                const float4 q0 = cq[2 * j], q1 = cq[2 * j + 1];
                const float Sc[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
                const float4 v = ch(j);
                q.add(v.x, v.y, v.z, v.w, cm, Sc);
            }
            q.fin(S);
        }
        store_moments(a, nd, S, t.last, l, t.next);
    }
    return level;
}

// Layers l >= 1.  Every global load whose address depends only on the block and thread
// position (the candidate count, the prefix sums, the candidates' level bytes with the seven
// after them, each candidate node) is issued up front, so a block
// waits on one round of memory latency instead of a chain of them; only the further children
// of a parent with more than one child are loaded afterwards.  Reads past the candidate count
// stay inside the layer's capacity and are masked out.
template <bool QUAD>
__global__ __launch_bounds__(256) void k_layer_emit_int(TreeArgs a, uint32_t l,
                                                        const uint32_t* __restrict__ tile_cnt,
                                                        const uint32_t* __restrict__ super,
                                                        uint32_t* next_cnt, uint32_t* next_super) {
    constexpr uint32_t NI = kLayerItems, NL = kLayerItems + 7;
    __shared__ uint32_t s_tmp[16];
    __shared__ uint32_t s_cnt[3];
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * kLayerTile;
    const uint32_t capc = (uint32_t)a.cap[l - 1];
    const uint64_t cb = a.off[l - 1];
    const uint32_t c0 = base + threadIdx.x * NI;
    const uint32_t n = a.counts[l - 1];
    const uint32_t sup = tile / kSuper;
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < sup; t += 256) part += super[t * kSuperStride];
    if (threadIdx.x < tile - sup * kSuper) part += tile_cnt[sup * kSuper + threadIdx.x];
    uint32_t lv[NL];
#pragma unroll
    for (uint32_t i = 0; i < NL; i++) lv[i] = c0 + i < capc ? a.levels[cb + c0 + i] : 0u;
    TNode u0[NI];
#pragma unroll
    for (uint32_t i = 0; i < NI; i++) u0[i] = load_node(a, cb + (c0 + i < capc ? c0 + i : 0u));
    if (base >= n) return;
    uint32_t hmask = 0, cnt = 0;
#pragma unroll
    for (uint32_t i = 0; i < NI; i++)
        if (c0 + i < n && lv[i] >= l) {
            hmask |= 1u << i;
            cnt++;
        }
    uint32_t prefix, total;
    uint32_t p = block_sum_and_scan<256>(part, cnt, s_tmp, &prefix, &total);
    p += prefix;
    if (threadIdx.x == 0 && base + kLayerTile >= n) {
        a.counts[l] = prefix + total;  // the last tile knows the layer's size
        if (prefix + total > a.cap[l]) atomicOr(a.err, ERR_CAPACITY);
    }
    uint32_t pos[NI], lev[NI], k = 0;
#pragma unroll
    for (uint32_t i = 0; i < NI; i++) {
        if (!(hmask & (1u << i))) continue;
        if (p < a.cap[l]) {
            pos[k] = p;
            lev[k++] = emit_parent<QUAD>(a, l, c0 + i, p, &lv[i], u0[i], n);
        }
        p++;
    }
    if (next_cnt)
        count_next_heads(l, prefix, total, pos, lev, k, s_cnt, next_cnt, next_super);
}

// The top layers in one launch of one block: for small layers the per-layer launches cost
// more than the work.  Layer l's candidates (layer l - 1's nodes) are taken 1,024 at a time;
// a block scan of their head flags, carried across the chunks, gives each parent's index.
// Layer l's nodes are complete (and visible to the block) before layer l + 1 reads them.
constexpr uint32_t kTopThreads = 1024;
constexpr uint64_t kTopCandidates = 32768;  // largest candidate capacity of a layer built here
template <bool QUAD>
__global__ __launch_bounds__(kTopThreads) void k_layers_top(TreeArgs a, uint32_t l0) {
    __shared__ uint32_t s_tmp[kTopThreads / 64 + 1];
    uint32_t n = a.counts[l0 - 1];
    for (uint32_t l = l0; l < a.layers; l++) {
        const uint64_t cb = a.off[l - 1];
        const uint32_t capc = (uint32_t)a.cap[l - 1];
        uint32_t carry = 0;
        for (uint32_t base = 0; base < n; base += kTopThreads) {
            const uint32_t c = base + threadIdx.x;
            // every load of the round up front: levels, the candidate and (with quadrupoles, whose
            // children's rows make a long chain otherwise) the seven nodes after it, so a layer
            // waits on one round of latency; only candidates below the count are read.
            // Monopoles measured faster loading the further children in emit_parent.
            uint32_t lv[8];
            float4 ch[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                const bool in = c + j < n && c + j < capc;  // capc: a count past capacity
                lv[j] = in ? a.levels[cb + c + j] : 0u;
                if (QUAD) ch[j] = in ? a.tpos[cb + c + j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            const TNode u0 = load_node(a, cb + (c < n && c < capc ? c : 0u));
            const bool head = c < n && lv[0] >= l;
            uint32_t total;
            const uint32_t p = carry + block_exclusive_scan<kTopThreads>(head ? 1u : 0u, s_tmp,
                                                                        &total);
            if (head && p < a.cap[l]) emit_parent<QUAD, QUAD>(a, l, c, p, lv, u0, n, ch);
            carry += total;
        }
        if (threadIdx.x == 0) {
            a.counts[l] = carry;
            if (carry > a.cap[l]) atomicOr(a.err, ERR_CAPACITY);
        }
        n = carry;
        __syncthreads();  // layer l is written before layer l + 1 reads it
    }
}

// ----------------------------------------------------------------- fused lower layers
// The lower layers (0 .. lf-1, where nearly every node holds a handful of particles) in three
// launches instead of one per layer.  A node of layer l starts at the sorted position j of its
// first particle, and j starts nodes exactly at the layers l <= h(j), its head level (the
// level byte the layer kernels carry; 0xFF for a particle whose key equals its predecessor's).
//  * k_fused_heads: h(j) for every position, and per tile of kFusedTile positions the number
//    of nodes of each layer < lf starting in it;
//  * k_fused_scan: per layer, the exclusive scan of those counts over the tiles (each tile's
//    first node index) and the layer's size;
//  * k_fused_build: one block per tile builds every layer of its positions in LDS, leaves from
//    the particles and each layer from the one below, and writes the nodes that start in its
//    tile, with their next pointers (from its own ranks of every layer below the top, in place
//    of k_next_fill, which runs only without fused layers).  It reads past its tile up to the next head at layer lf-1 (at most kFusedHalo
//    positions), so the nodes it owns are complete; a node that runs past that halo is left to
//  * fused_fix: the fused build's last block to finish builds those nodes, layer by layer,
//    from the finished layer below.
// Aggregation order and arithmetic are those of the per-layer kernels (emit_leaf, emit_parent),
// so the nodes are bit-identical.
constexpr uint32_t kFusedTile = 1024;
constexpr uint32_t kFusedHalo = 1024;
constexpr uint32_t kFusedRange = kFusedTile + kFusedHalo;
constexpr uint32_t kFusedThreads = 1024;
constexpr uint8_t kNoHead = 0xFF;

uint32_t fused_tiles(uint32_t nt) { return (nt + kFusedTile - 1) / kFusedTile; }

// block_exclusive_scan with LDS-only barriers (tmp: NT / 64 + 1 words).  Without the closing
// barrier the caller must pass another barrier before tmp is written again.
template <int NT, bool CLOSE = true>
__device__ inline uint32_t lds_block_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) tmp[wave] = inc;
    lds_barrier();
    scan_wave_totals<NW>(tmp);
    lds_barrier();
    const uint32_t r = tmp[wave] + inc - v;
    *total = tmp[NW];
    if (CLOSE) lds_barrier();
    return r;
}

// Head level of every sorted position (kNoHead inside a leaf's run) and, per tile, the
// nodes of each layer below the top that start in it (layers 0 .. nl-1, nl = bits).
__global__ __launch_bounds__(kFusedThreads) void k_fused_heads(TreeArgs a, uint8_t* plevel,
                                                               uint32_t* fcnt, uint32_t* ffix,
                                                               uint32_t lf, uint32_t nl,
                                                               uint32_t tiles, uint32_t* clr,
                                                               uint32_t nclr) {
    __shared__ uint32_t s_hist[kMaxLayers];
    // the per-layer kernels' tile counts and supertile sums start at zero
    for (uint32_t i = blockIdx.x * kFusedThreads + threadIdx.x; i < nclr; i += tiles * kFusedThreads)
        clr[i] = 0u;
    if (threadIdx.x < kMaxLayers) s_hist[threadIdx.x] = 0;
    if (threadIdx.x < lf) ffix[threadIdx.x * tiles + blockIdx.x] = 0xFFFFFFFFu;  // no leftover
    __syncthreads();
    const uint32_t j = blockIdx.x * kFusedTile + threadIdx.x;
    if (j < a.nt) {
        uint32_t h = kNoHead;
        if (j == 0 || (a.segmented && a.sys_of[j] != a.sys_of[j - 1])) {
            h = a.bits;
        } else {
            const uint64_t k = a.sorted_keys[j], kp = a.sorted_keys[j - 1];
            if (k != kp) h = head_level(k, kp, a.bits);
        }
        plevel[j] = (uint8_t)h;
        if (h != kNoHead) atomicAdd(&s_hist[h < nl - 1 ? h : nl - 1], 1u);
    }
    __syncthreads();
    if (threadIdx.x < nl) {
        uint32_t c = 0;
        for (uint32_t b = threadIdx.x; b < nl; b++) c += s_hist[b];
        fcnt[threadIdx.x * tiles + blockIdx.x] = c;
    }
}

// Block l: exclusive scan of layer l's per-tile node counts; for the fused layers (l < lf)
// also the layer's size.
__global__ __launch_bounds__(kFusedThreads) void k_fused_scan(TreeArgs a, const uint32_t* fcnt,
                                                              uint32_t* fbase, uint32_t lf,
                                                              uint32_t tiles) {
    __shared__ uint32_t s_tmp[kFusedThreads / 64 + 1];
    const uint32_t l = blockIdx.x;
    const uint32_t* in = fcnt + (uint64_t)l * tiles;
    uint32_t* out = fbase + (uint64_t)l * tiles;
    // the block of layer lf - 1 also decides the lowest layer to write (block 0 without fused
    // layers): k_fused_build's range of tile t - 1 ends at the first head at layer >= lf - 1
    // in tile t, so a full tile t without one (no node of layer lf - 1 starts in it) leaves a
    // node for fused_fix, which needs every layer below: then every layer is written
    const bool decide = l == (lf >= 2 ? lf - 1 : 0u);
    bool trunc = false;
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < tiles; t0 += kFusedThreads) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t v = t < tiles ? in[t] : 0u;
        if (decide && t >= 1 && t < tiles && v == 0u &&
            (uint64_t)t * kFusedTile + kFusedTile < a.nt)
            trunc = true;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<kFusedThreads>(v, s_tmp, &total);
        if (t < tiles) out[t] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && l < lf) {
        a.counts[l] = carry;
        if (carry > a.cap[l]) atomicOr(a.err, ERR_CAPACITY);
    }
    if (decide) {
        trunc = __syncthreads_or(trunc) != 0;
        if (threadIdx.x == 0) {
            uint32_t W = 0;
            if (a.skip_ok && lf >= 2 && !trunc) {
                // layers 0 .. K-1 pass the traversal's opening test RN(sqrt(r2)) >= rmin0 2^l
                // for every finite target (r2 >= eps2, and RN(sqrt) is monotone): their nodes
                // are never opened, and the walks reach layer K - 1 at the lowest
                const float side0 =
                    a.side_shifted ? 2097152.0f : ldexpf(a.cubes[0].w, -(int)a.bits);
                const float rmin0 = accept_rmin(side0, a.acc);
                const float re = sqrt_rn_normal(a.eps2);
                uint32_t K = 0;
                while (K < lf && re >= ldexpf(rmin0, (int)K)) K++;
                W = K > 0 ? K - 1 : 0u;
            }
            a.wlow[0] = W;
            a.wlow[1] = trunc ? 1u : 0u;  // fused_fix has nodes to build
        }
    }
}

// The next pointer of the nodes starting at each head position s (TNode): with H = h(s) below
// the top, the node of layer H holding s - 1, whose index is off[H] plus the heads at level
// >= H before s, minus one; kNextEnd where s starts a system.  The heads at level >= H before
// s in the tile are counted with one ballot per layer: a wave's own count from the ballot of
// its lane's level, the waves before it from their per-layer totals in LDS.
__global__ __launch_bounds__(kFusedThreads) void k_next_fill(TreeArgs a,
                                                             const uint8_t* __restrict__ plevel,
                                                             const uint32_t* __restrict__ fbase,
                                                             uint32_t nl, uint32_t tiles) {
    constexpr uint32_t NW = kFusedThreads / 64;
    __shared__ uint32_t s_cnt[NW][kMaxLayers];
    const uint32_t t = blockIdx.x, lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t j = t * kFusedTile + threadIdx.x;
    const uint32_t h = j < a.nt ? plevel[j] : kNoHead;
    const int hv = h == kNoHead ? -1 : (int)h;
    uint64_t mine = 0;
    for (uint32_t L = 0; L < nl; L++) {
        const uint64_t m = __ballot(hv >= (int)L);
        mine = (uint32_t)hv == L ? m : mine;
        if (lane == 0) s_cnt[wave][L] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (hv < 0 || j >= a.nt) return;
    if (h >= nl) {  // a system's first position (h = bits)
        a.pnext[j] = kNextEnd;
        return;
    }
    uint32_t before = lanes_below(mine);
    for (uint32_t w = 0; w < wave; w++) before += s_cnt[w][h];
    // the first position of s's system, a head at every layer, precedes s: the index of the
    // node holding s - 1 is at least off[h]
    a.pnext[j] = (uint32_t)(a.off[h] + fbase[(uint64_t)h * tiles + t] + before - 1u);
}

__device__ inline uint32_t block_min_u32(uint32_t v, uint32_t* tmp) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, d, 64));
    if ((threadIdx.x & 63u) == 0u) tmp[threadIdx.x >> 6] = v;
    lds_barrier();
    uint32_t m = tmp[0];
    for (uint32_t w = 1; w < blockDim.x / 64u; w++) m = min(m, tmp[w]);
    lds_barrier();
    return m;
}

template <typename AGG>
__device__ inline void agg_particles(const TreeArgs& a, uint32_t j0, uint32_t j1, float* cm) {
    AGG g;
    for (uint32_t i = j0; i < j1; i++) {
        const float4 x = a.spos[i];
        g.add(x.x, x.y, x.z, x.w);
    }
    g.fin(cm);
}
template <typename AGG>
__device__ inline void agg_nodes(const float4* v, uint32_t c0, uint32_t c1, float* cm) {
    AGG g;
    for (uint32_t i = c0; i < c1; i++) g.add(v[i].x, v[i].y, v[i].z, v[i].w);
    g.fin(cm);
}
// Second moments about cm, in emit_leaf's and emit_parent's order.
template <typename MOM>
__device__ inline void mom_particles(const TreeArgs& a, uint32_t j0, uint32_t j1, const float* cm,
                                     float* S) {
    MOM q;
    for (uint32_t i = j0; i < j1; i++) {
        const float4 x = a.spos[i];
        q.add(x.x, x.y, x.z, x.w, cm, nullptr);
    }
    q.fin(S);
}
template <typename MOM>
__device__ inline void mom_nodes(const float4* v, const float* sm, uint32_t c0, uint32_t c1,
                                 const float* cm, float* S) {
    MOM q;
    for (uint32_t i = c0; i < c1; i++) q.add(v[i].x, v[i].y, v[i].z, v[i].w, cm, sm + 6 * i);
    q.fin(S);
}

// QUAD: the second moments too, aggregated in the per-layer kernels' order (emit_leaf,
// emit_parent) from the layer below kept in LDS beside the centres.
// The nodes k_fused_build left (they run past a halo), layer by layer: each tile's last node
// of the layer, from the layer below, which is complete by then.  Run by the fused build's last
// block to finish, after every other block's nodes are visible.
template <bool QUAD>
__device__ void fused_fix(const TreeArgs& a, const uint32_t* fcnt, const uint32_t* fbase,
                          const uint32_t* ffix, uint32_t lf, uint32_t tiles, uint32_t* lcnt,
                          uint32_t* lsup) {
    for (uint32_t l = 0; l < lf; l++) {
        for (uint32_t t = threadIdx.x; t < tiles; t += kFusedThreads) {
            const uint32_t v = ffix[(uint64_t)l * tiles + t];
            if (v == 0xFFFFFFFFu) continue;
            const uint64_t lt = (uint64_t)l * tiles + t;
            const uint32_t p = fbase[lt] + fcnt[lt] - 1u;
            if (p >= a.cap[l]) continue;
            if (l == 0) {
                if (a.fp64_from == 0) emit_leaf<Agg64, QUAD>(a, v, p);
                else emit_leaf<Agg32, QUAD>(a, v, p);
            } else {
                const uint64_t cb = a.off[l - 1];
                const uint32_t capc = (uint32_t)a.cap[l - 1];
                uint32_t lv[8];
                for (uint32_t j = 0; j < 8; j++) lv[j] = v + j < capc ? a.levels[cb + v + j] : 0u;
                const TNode u0 = load_node(a, cb + v);
                const uint32_t h = emit_parent<QUAD>(a, l, v, p, lv, u0, a.counts[l - 1]);
                if (l + 1 == lf && lcnt && h >= lf) {  // a head of layer lf (as counted above)
                    atomicAdd(&lcnt[p / kLayerTile], 1u);
                    atomicAdd(&lsup[(p / kLayerTile / kSuper) * kSuperStride], 1u);
                }
            }
        }
        __syncthreads();  // layer l is complete before layer l + 1 reads it
    }
}

template <bool QUAD>
__global__ __launch_bounds__(kFusedThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_fused_build(TreeArgs a,
                                                               const uint8_t* __restrict__ plevel,
                                                               const uint32_t* __restrict__ fcnt,
                                                               const uint32_t* __restrict__ fbase,
                                                               uint32_t* ffix, uint32_t lf,
                                                               uint32_t tiles, uint32_t* lcnt,
                                                               uint32_t* lsup) {
    __shared__ float4 s_val[kFusedRange];     // the current layer's nodes (centre, mass)
    __shared__ uint32_t s_lc[8];              // heads of layer lf per candidate tile
    __shared__ float s_mom[QUAD ? 6 * kFusedRange : 1];  // and their second moments
    // per position: nodes of the layer before it; per node of the layer: its start position
    // (two of each, by layer parity, so that one layer's are written while the last layer's
    // are still read)
    __shared__ uint16_t s_rank[2][kFusedRange];
    __shared__ uint16_t s_pos[2][kFusedRange];
    __shared__ uint8_t s_h[kFusedRange];      // per position: its head level
    __shared__ uint32_t s_tmp[kFusedThreads / 64 + 1];
    constexpr uint32_t NW = kFusedThreads / 64;
    __shared__ uint32_t s_wcnt[kMaxLayers][NW + 1];  // heads per wave, then their scan
    __shared__ uint32_t s_pnext[kFusedTile];  // the next of the nodes starting at each position
    const uint32_t t = blockIdx.x, tid = threadIdx.x;
    const uint32_t a0 = t * kFusedTile;
    const uint32_t b0 = min(a0 + kFusedTile, a.nt);
    const uint32_t own = b0 - a0;  // positions [0, own) are this tile's
    if (tid < 8) s_lc[tid] = 0u;   // read after the layer loop's barriers
    // the range ends at the next head at layer lf-1 after the tile (every node of the lower
    // layers ends there), or at the halo's end
    uint32_t cand = kFusedHalo;
    {
        const uint32_t j = b0 + tid;
        if (tid < kFusedHalo && j < a.nt) {
            const uint32_t h = plevel[j];
            if (h != kNoHead && h >= lf - 1) cand = tid;
        }
    }
    const uint32_t m = block_min_u32(cand, s_tmp);
    uint32_t e0;
    bool truncated = false;
    uint32_t h_end = kNoHead;
    if (m < kFusedHalo) {
        e0 = b0 + m;
    } else if (b0 + kFusedHalo >= a.nt) {
        e0 = a.nt;
    } else {
        e0 = b0 + kFusedHalo;
        truncated = true;
        h_end = plevel[e0];
    }
    const uint32_t np = e0 - a0;
    const uint32_t W = *a.wlow;  // layers below W are not written; W's nodes end the walks
    const uint32_t q0 = 2 * tid, q1 = 2 * tid + 1;  // two adjacent positions per thread
    const uint32_t h0 = q0 < np ? plevel[a0 + q0] : kNoHead;
    const uint32_t h1 = q1 < np ? plevel[a0 + q1] : kNoHead;
    // levels of the range in LDS: the node stores of every layer read them
    s_h[q0] = (uint8_t)h0;
    s_h[q1] = (uint8_t)h1;
    // Every layer's node ranks up front: wave w holds positions 128 w .. 128 w + 127 (two per
    // lane, in order), so a position's rank among the layer's heads is the heads of the waves
    // before it (a scan over the 16 wave counts) plus those of the lanes below it (ballots).
    // Every layer below the top (nl = bits), for the next pointers: the nodes starting at a
    // head position s of level H < nl have as next the node of layer H holding s - 1, whose
    // index is off[H] plus the layer's nodes before the tile plus s's rank at layer H, minus
    // one; kNextEnd where s starts a system (H = bits).
    const uint32_t nl = a.bits;
    const int hv0 = h0 == kNoHead ? -1 : (int)h0, hv1 = h1 == kNoHead ? -1 : (int)h1;
    const uint32_t wave = tid >> 6, lane = tid & 63u;
    uint32_t pr0 = 0, pr1 = 0;  // the ranks of q0 and q1 within the wave at their own level
    for (uint32_t l = 0; l < nl; l++) {
        const uint64_t b0 = __ballot(hv0 >= (int)l), b1 = __ballot(hv1 >= (int)l);
        if (lane == 0) s_wcnt[l][wave] = (uint32_t)(__popcll(b0) + __popcll(b1));
        const uint32_t c = lanes_below(b0) + lanes_below(b1);
        if (hv0 == (int)l) pr0 = c;
        if (hv1 == (int)l) pr1 = c + (hv0 >= (int)l ? 1u : 0u);
    }
    lds_barrier();
    if (tid < nl) {  // per layer, the exclusive scan of the wave counts and the total
        uint32_t c = 0;
        for (uint32_t w = 0; w < NW; w++) {
            const uint32_t v = s_wcnt[tid][w];
            s_wcnt[tid][w] = c;
            c += v;
        }
        s_wcnt[tid][NW] = c;
    }
    lds_barrier();
    {
        // the next pointers of this tile's head positions, in LDS for the node stores below and
        // in device memory when the leftover pass will read them
        const bool keep = a.wlow[1] != 0u;
        const uint32_t qq[2] = {q0, q1};
        const int hv[2] = {hv0, hv1};
        const uint32_t pr[2] = {pr0, pr1};
#pragma unroll
        for (uint32_t i = 0; i < 2; i++) {
            if (qq[i] >= own || hv[i] < 0) continue;
            const uint32_t H = (uint32_t)hv[i];
            const uint32_t pn =
                H >= nl ? kNextEnd
                        : (uint32_t)(a.off[H] + fbase[(uint64_t)H * tiles + t] + s_wcnt[H][wave] +
                                     pr[i] - 1u);
            s_pnext[qq[i]] = pn;
            if (keep) a.pnext[a0 + qq[i]] = pn;
        }
    }
    uint32_t tot_prev = 0;
    for (uint32_t l = 0; l < lf; l++) {
        const bool f0 = hv0 >= (int)l, f1 = hv1 >= (int)l;
        const uint64_t b0 = __ballot(f0), b1 = __ballot(f1);
        const uint32_t tot = s_wcnt[l][NW];
        const uint32_t r0 = s_wcnt[l][wave] + lanes_below(b0) + lanes_below(b1);
        const uint32_t r1 = r0 + (f0 ? 1u : 0u);
        uint16_t* const sp = s_pos[l & 1u];
        const uint16_t* const srk = s_rank[(l + 1u) & 1u];  // the layer below's ranks
        if (f0) sp[r0] = (uint16_t)q0;
        if (f1) sp[r1] = (uint16_t)q1;
        // per position, the nodes of this layer before it (the next layer's child ranks)
        if (q0 < np) s_rank[l & 1u][q0] = (uint16_t)r0;
        if (q1 < np) s_rank[l & 1u][q1] = (uint16_t)r1;
        if (l == 0) {  // the leaf of every position of the tile
            const uint32_t base0 = fbase[t];
            if (q0 < own) a.leaf_of_sorted[a0 + q0] = base0 + r0 + (f0 ? 1u : 0u) - 1u;
            if (q1 < own) a.leaf_of_sorted[a0 + q1] = base0 + r1 + (f1 ? 1u : 0u) - 1u;
        }
        lds_barrier();
        // the last node ends at the range's end, which starts a node of this layer unless the
        // range was cut short
        const bool end_ok = !truncated || (h_end != kNoHead && h_end >= l);
        const bool fp64 = l >= a.fp64_from;
        float cm[2][4], S[2][6];
        uint32_t st[2], c0[2], c1[2];
        bool done[2];
#pragma unroll
        for (uint32_t i = 0; i < 2; i++) {
            const uint32_t k = tid + i * kFusedThreads;
            done[i] = false;
            if (k >= tot) continue;
            st[i] = sp[k];
            const uint32_t nx = k + 1 < tot ? sp[k + 1] : np;
            done[i] = k + 1 < tot || end_ok;
            if (l == 0) {
                c0[i] = a0 + st[i];
                c1[i] = a0 + nx;
                if (fp64) agg_particles<Agg64>(a, c0[i], c1[i], cm[i]);
                else agg_particles<Agg32>(a, c0[i], c1[i], cm[i]);
                if (QUAD) {
                    if (fp64) mom_particles<Mom64>(a, c0[i], c1[i], cm[i], S[i]);
                    else mom_particles<Mom32>(a, c0[i], c1[i], cm[i], S[i]);
                }
            } else {
                c0[i] = srk[st[i]];
                c1[i] = nx < np ? srk[nx] : tot_prev;
                if (fp64) agg_nodes<Agg64>(s_val, c0[i], c1[i], cm[i]);
                else agg_nodes<Agg32>(s_val, c0[i], c1[i], cm[i]);
                if (QUAD) {
                    if (fp64) mom_nodes<Mom64>(s_val, s_mom, c0[i], c1[i], cm[i], S[i]);
                    else mom_nodes<Mom32>(s_val, s_mom, c0[i], c1[i], cm[i], S[i]);
                }
            }
        }
        lds_barrier();  // the layer below is read; overwrite it
        const uint32_t base = fbase[(uint64_t)l * tiles + t];
        const uint32_t base_prev = l ? fbase[(uint64_t)(l - 1) * tiles + t] : 0u;
#pragma unroll
        for (uint32_t i = 0; i < 2; i++) {
            const uint32_t k = tid + i * kFusedThreads;
            if (k >= tot) continue;
            s_val[k] = make_float4(cm[i][0], cm[i][1], cm[i][2], cm[i][3]);
            if (QUAD)
                for (uint32_t e = 0; e < 6; e++) s_mom[6 * k + e] = S[i][e];
            if (st[i] >= own) continue;  // a halo node: owned by a later tile
            const uint32_t gi = base + k;
            if (gi >= a.cap[l]) continue;  // ERR_CAPACITY is raised by the scan
            if (!done[i]) {  // runs past the halo: fused_fix builds it
                ffix[l * tiles + t] = l == 0 ? c0[i] : base_prev + c0[i];
                continue;
            }
            if (l < W) continue;  // no walk reaches this layer
            const uint32_t h = s_h[st[i]];
            const uint64_t nd = a.off[l] + gi;
            TNode tn;
            tn.x = cm[i][0];
            tn.y = cm[i][1];
            tn.z = cm[i][2];
            tn.m = cm[i][3];
            tn.last = l == 0 ? ASTRO_INVALID_CHILD : (uint32_t)(a.off[l - 1] + base_prev + c1[i] - 1u);
            if (W > 0 && l == W) tn.last = kNextEnd;  // never opened by a finite target
            tn.info = tn_info(l, l == 0 ? 0u : c1[i] - c0[i], h);
            tn.next = s_pnext[st[i]];
            tn.start = a0 + st[i];
            store_node(a, nd, tn);
            a.levels[nd] = (uint8_t)h;
            if (a.sysfirst && h == a.bits)  // a system's first node of this layer
                a.sysfirst[(size_t)a.sys_of[tn.start] * kMaxLayers + l] = gi;
            if (QUAD) store_moments(a, nd, S[i], tn.last, l, tn.next);
        }
        if (l + 1 == lf && lcnt) {
            // layer lf's heads among this block's nodes of layer lf - 1 (at most 5 candidate
            // tiles of kLayerTile), added to the tile counts and supertile sums the first
            // per-layer kernel reads (in place of a counting launch)
            const uint32_t t0 = base / kLayerTile;
#pragma unroll
            for (uint32_t i = 0; i < 2; i++) {
                const uint32_t k = tid + i * kFusedThreads;
                if (k >= tot || st[i] >= own || !done[i] || base + k >= a.cap[l]) continue;
                if (s_h[st[i]] >= lf) atomicAdd(&s_lc[(base + k) / kLayerTile - t0], 1u);
            }
            lds_barrier();
            if (tid < 8 && s_lc[tid]) {
                const uint32_t tt = t0 + tid;
                atomicAdd(&lcnt[tt], s_lc[tid]);
                atomicAdd(&lsup[(tt / kSuper) * kSuperStride], s_lc[tid]);
            }
        }
        tot_prev = tot;
        // no barrier here: the next layer writes the other parity's s_pos and s_rank (last
        // read before this layer's barrier after the aggregation) and passes a barrier before
        // anything reads s_val
    }
    // Leftovers (a tile ran past its halo; the scan raised wlow[1] before this launch): the
    // last block to finish builds them.  Each block's nodes are released by its count
    // (agent-scope acq_rel after the barrier); the last block acquires them all.
    if (a.wlow[1] != 0u) {
        __shared__ uint32_t s_last;
        __syncthreads();
        if (tid == 0) {
            const uint32_t done = __hip_atomic_fetch_add(&a.wlow[2], 1u, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            s_last = done + 1u == gridDim.x ? 1u : 0u;
        }
        __syncthreads();
        if (s_last) {
            __threadfence();
            fused_fix<QUAD>(a, fcnt, fbase, ffix, lf, tiles, lcnt, lsup);
            if (tid == 0) __hip_atomic_store(&a.wlow[2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// fbuf: per-tile counts [nl][tiles], their scan [nl][tiles], the fused layers' leftovers
// [lf][tiles] (nl = bits <= kMaxLayers).
static uint32_t* fused_fcnt(uint32_t* fbuf, uint32_t) { return fbuf; }
static uint32_t* fused_fbase(uint32_t* fbuf, uint32_t tiles) {
    return fbuf + (uint64_t)kMaxLayers * tiles;
}
static uint32_t* fused_ffix(uint32_t* fbuf, uint32_t tiles) {
    return fbuf + 2ull * kMaxLayers * tiles;
}

void launch_heads_next(const TreeArgs& a, uint32_t lf, uint8_t* plevel, uint32_t* fbuf,
                       uint32_t* clr, uint32_t nclr, hipStream_t s) {
    const uint32_t tiles = fused_tiles(a.nt);
    if (!a.nt) {
        launch_fill_u32(clr, 0, nclr, s);
        return;
    }
    const uint32_t nl = a.bits;
    hipLaunchKernelGGL(k_fused_heads, dim3(tiles), dim3(kFusedThreads), 0, s, a, plevel,
                       fused_fcnt(fbuf, tiles), fused_ffix(fbuf, tiles), lf, nl, tiles, clr,
                       nclr);
    hipLaunchKernelGGL(k_fused_scan, dim3(nl), dim3(kFusedThreads), 0, s, a,
                       fused_fcnt(fbuf, tiles), fused_fbase(fbuf, tiles), lf, tiles);
    // with fused layers the fused build computes the next pointers itself
    if (lf == 0)
        hipLaunchKernelGGL(k_next_fill, dim3(tiles), dim3(kFusedThreads), 0, s, a, plevel,
                           fused_fbase(fbuf, tiles), nl, tiles);
}

void launch_fused_layers(const TreeArgs& a, uint32_t lf, const uint8_t* plevel, uint32_t* fbuf,
                         uint32_t* lcnt, uint32_t* lsup, hipStream_t s) {
    const uint32_t tiles = fused_tiles(a.nt);
    if (!a.nt || lf == 0) return;
    uint32_t* fcnt = fused_fcnt(fbuf, tiles);
    uint32_t* fbase = fused_fbase(fbuf, tiles);
    uint32_t* ffix = fused_ffix(fbuf, tiles);
    if (a.quad) {
        hipLaunchKernelGGL(k_fused_build<true>, dim3(tiles), dim3(kFusedThreads), 0, s, a, plevel,
                           fcnt, fbase, ffix, lf, tiles, lcnt, lsup);
    } else {
        hipLaunchKernelGGL(k_fused_build<false>, dim3(tiles), dim3(kFusedThreads), 0, s, a,
                           plevel, fcnt, fbase, ffix, lf, tiles, lcnt, lsup);
    }
}

uint32_t layer_tiles(uint64_t candidates) {
    return (uint32_t)((candidates + kLayerTile - 1) / kLayerTile);
}
uint64_t layer_super_words(uint64_t candidates) {
    return (uint64_t)((layer_tiles(candidates) + kSuper - 1) / kSuper) * kSuperStride;
}

void launch_layer_count(const TreeArgs& a, uint32_t layer, uint32_t* tile_cnt, uint32_t* super,
                        hipStream_t s) {
    const uint64_t cand_cap = layer == 0 ? a.nt : a.cap[layer - 1];
    if (!cand_cap) return;
    hipLaunchKernelGGL(k_layer_count, dim3(layer_tiles(cand_cap)), dim3(256), 0, s, a, layer,
                       tile_cnt, super);
}

uint32_t layers_top_first(const TreeArgs& a) {
    uint32_t l0 = a.layers;
    while (l0 > 1 && a.cap[l0 - 2] <= kTopCandidates) l0--;
    return l0;
}
void launch_layers_top(const TreeArgs& a, uint32_t l0, hipStream_t s) {
    if (l0 >= a.layers || l0 == 0) return;
    if (a.quad)
        hipLaunchKernelGGL(k_layers_top<true>, dim3(1), dim3(kTopThreads), 0, s, a, l0);
    else
        hipLaunchKernelGGL(k_layers_top<false>, dim3(1), dim3(kTopThreads), 0, s, a, l0);
}

void launch_layer_emit(const TreeArgs& a, uint32_t layer, const uint32_t* tile_cnt,
                       const uint32_t* super, uint32_t* next_cnt, uint32_t* next_super,
                       hipStream_t s) {
    const uint64_t cand_cap = layer == 0 ? a.nt : a.cap[layer - 1];
    if (!cand_cap) return;
    const dim3 grid(layer_tiles(cand_cap));
    if (layer == 0 && a.quad)
        hipLaunchKernelGGL(k_leaf_emit<true>, grid, dim3(256), 0, s, a, tile_cnt, super, next_cnt,
                           next_super);
    else if (layer == 0)
        hipLaunchKernelGGL(k_leaf_emit<false>, grid, dim3(256), 0, s, a, tile_cnt, super, next_cnt,
                           next_super);
    else if (a.quad)
        hipLaunchKernelGGL(k_layer_emit_int<true>, grid, dim3(256), 0, s, a, layer, tile_cnt,
                           super, next_cnt, next_super);
    else
        hipLaunchKernelGGL(k_layer_emit_int<false>, grid, dim3(256), 0, s, a, layer, tile_cnt,
                           super, next_cnt, next_super);
}

// The OctreeNode array (the reference layout) from the compact nodes, after all layers are
// built.  A tile is 256 slots of one layer: each thread assembles one node in registers and
// stores it into LDS field-major (rows padded by one float4, so neither the per-node stores nor
// the copy-out conflict on banks); the block then writes the tile as one contiguous run, 1 KB
// per wave store, with non-temporal stores so the 128 B per node stream past the caches instead
// of evicting the compact nodes that the traversal reads next.  Slots that held a node
// in the previous build but not in this one are emptied (Octree.metal:253-269 applied to the
// stale tail of each layer only); slots never used stay empty.
constexpr uint32_t kEmitNodes = 256;
constexpr uint32_t kEmitRow = kEmitNodes + 1;
__device__ __forceinline__ void emit_tile(const TreeArgs& a, const uint32_t* counts_prev,
                                          uint32_t l, uint32_t tile, float4* s_f) {
    const uint32_t cnt = a.counts[l], prev = counts_prev[l];
    // a count beyond the layer's capacity has raised ERR_CAPACITY; never write past it
    const uint32_t end = min(cnt > prev ? cnt : prev, (uint32_t)a.cap[l]);
    const uint32_t base = tile * kEmitNodes;
    if (base >= end) return;  // uniform over the block
    const uint32_t nslots = min(kEmitNodes, end - base);
    const uint32_t i = base + threadIdx.x;
    const float inv = __uint_as_float(ASTRO_INVALID_CHILD);
    if (threadIdx.x < nslots) {
        float4 f[8];
        if (i < cnt) {  // the fields of write_onode, from registers
            const uint64_t node = a.off[l] + i;
            const TNode t = load_node(a, node);
            const uint64_t code = a.sorted_keys[t.start] >> (3 * l);
            const uint32_t nch = l ? tn_nchild(t.info) : 0u;
            const uint32_t first = t.last - nch + 1u;  // (unused for leaves)
            f[0] = make_float4(__uint_as_float((uint32_t)code),
                               __uint_as_float((uint32_t)(code >> 32)), 0.0f, 0.0f);
            f[1] = make_float4(t.x, t.y, t.z, 0.0f);
            f[2] = make_float4(t.m, 0.0f, 0.0f, 0.0f);
            f[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            f[4] = f[3];
            if (a.quad) {
                const float4 q1 = a.qmom[2 * node + 1];
                f[3] = a.qmom[2 * node];
                f[4] = make_float4(q1.x, q1.y, 0.0f, 0.0f);
            }
            f[5] = make_float4(0u < nch ? __uint_as_float(first) : inv,
                               1u < nch ? __uint_as_float(first + 1) : inv,
                               2u < nch ? __uint_as_float(first + 2) : inv,
                               3u < nch ? __uint_as_float(first + 3) : inv);
            f[6] = make_float4(4u < nch ? __uint_as_float(first + 4) : inv,
                               5u < nch ? __uint_as_float(first + 5) : inv,
                               6u < nch ? __uint_as_float(first + 6) : inv,
                               7u < nch ? __uint_as_float(first + 7) : inv);
            f[7] = make_float4(__uint_as_float(l), 0.0f, 0.0f, 0.0f);
        } else {  // stale: an empty node
            for (int k = 0; k < 8; k++) f[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            f[0] = make_float4(inv, inv, 0.0f, 0.0f);
            f[5] = make_float4(inv, inv, inv, inv);
            f[6] = f[5];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) s_f[k * kEmitRow + threadIdx.x] = f[k];
    }
    __syncthreads();
    typedef float f4v __attribute__((ext_vector_type(4)));
    f4v* out = reinterpret_cast<f4v*>(a.onodes + a.off[l] + base);
    for (uint32_t e = threadIdx.x; e < nslots * 8u; e += kEmitNodes) {
        const float4 v = s_f[(e & 7u) * kEmitRow + (e >> 3)];
        const f4v w = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(w, out + e);
    }
    __syncthreads();  // s_f is reused by the next tile
}
// Persistent blocks: the capacity tiles of all layers, layer by layer, dealt round-robin to
// the gridDim.x blocks (tiles past a layer's live slots end at once).
__global__ __launch_bounds__(kEmitNodes) void k_emit_onodes(TreeArgs a,
                                                           const uint32_t* counts_prev) {
    __shared__ float4 s_f[8 * kEmitRow];
    const uint32_t G = gridDim.x;
    uint32_t w0 = 0;  // tiles of the layers below l
    for (uint32_t l = 0; l < a.layers; l++) {
        const uint32_t tiles = (uint32_t)((a.cap[l] + kEmitNodes - 1) / kEmitNodes);
        for (uint32_t t = (blockIdx.x + G - w0 % G) % G; t < tiles; t += G)
            emit_tile(a, counts_prev, l, t, s_f);
        w0 += tiles;
    }
}
void launch_emit_onodes(const TreeArgs& a, const uint32_t* counts_prev, uint32_t blocks,
                        hipStream_t s) {
    if (!a.onodes || !a.layers) return;
    hipLaunchKernelGGL(k_emit_onodes, dim3(blocks ? blocks : 1u), dim3(kEmitNodes), 0, s, a,
                       counts_prev);
}

// System of every position from the offsets (binary search over nsys + 1 offsets).
__global__ void k_fill_sys_of(uint32_t* sys_of, const uint32_t* off, uint32_t nsys, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = nsys;  // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    sys_of[i] = lo;
}
void launch_fill_sys_of(uint32_t* sys_of, const uint32_t* d_off, uint32_t nsys, uint32_t n,
                        hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_sys_of, dim3((n + 255) / 256), dim3(256), 0, s, sys_of, d_off, nsys, n);
}

// --------------------------------------------------- reference-shaped stage kernels
// Octree.metal:282-303 (markUniques)
__global__ void k_mark_uniques(const uint64_t* sorted, uint32_t* flags, uint32_t count,
                               uint32_t aggregate) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    uint64_t k = sorted[g];
    if (k == ASTRO_INVALID_MORTON_CODE) {
        flags[g] = 0;
        return;
    }
    bool u;
    if (g == 0) u = true;
    else if (aggregate == 1) u = (k >> 3) != (sorted[g - 1] >> 3);
    else u = k != sorted[g - 1];
    flags[g] = u ? 1u : 0u;
}
void launch_mark_uniques(const uint64_t* sorted, uint32_t* flags, uint32_t count,
                         uint32_t aggregate, hipStream_t s) {
    if (count)
        hipLaunchKernelGGL(k_mark_uniques, dim3((count + 255) / 256), dim3(256), 0, s, sorted,
                           flags, count, aggregate);
}
// Octree.metal:306-318 (scatterUniques)
__global__ void k_scatter_uniques(const uint32_t* flags, const uint32_t* scan, uint32_t* uniq,
                                  uint32_t count) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    if (flags[g]) uniq[scan[g]] = g;
}
void launch_scatter_uniques(const uint32_t* flags, const uint32_t* scan, uint32_t* uniq,
                            uint32_t count, hipStream_t s) {
    if (count)
        hipLaunchKernelGGL(k_scatter_uniques, dim3((count + 255) / 256), dim3(256), 0, s, flags,
                           scan, uniq, count);
}
// Octree.metal:321-334 (copyLastScanToParentCount)
__global__ void k_copy_last_scan(const uint32_t* scan, const uint32_t* flags, uint32_t* parent,
                                 uint32_t* child, uint32_t count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t last = count - 1;
        child[0] = parent[0];
        parent[0] = scan[last] + flags[last];
    }
}
void launch_copy_last_scan(const uint32_t* scan, const uint32_t* flags, uint32_t* parent,
                           uint32_t* child, uint32_t count, hipStream_t s) {
    if (count)
        hipLaunchKernelGGL(k_copy_last_scan, dim3(1), dim3(64), 0, s, scan, flags, parent, child,
                           count);
}

// Octree.metal:53-117 (aggregateLeafNodes): one thread per unique code.
template <typename AGG>
__device__ void agg_leaf_one(const uint64_t* sorted, const uint32_t* sorted_idx,
                             const uint32_t* uniq, const float4* pos, uint32_t num_unique,
                             OctreeNode* nodes, uint64_t* unsorted, uint32_t* unsorted_idx,
                             uint32_t g, uint32_t quirk, uint32_t nt) {
    uint32_t start = uniq[g];
    uint32_t end = g + 1 < num_unique ? uniq[g + 1] : (quirk ? num_unique : nt);
    AGG a;
    for (uint32_t i = start; i < end; i++) {
        float4 x = pos[sorted_idx[i]];
        a.add(x.x, x.y, x.z, x.w);
    }
    float cm[4];
    a.fin(cm);
    uint64_t code = sorted[start];
    write_onode(&nodes[g], code, cm, 0, 0, 0);
    unsorted[g] = code;
    unsorted_idx[g] = g;
}
__global__ void k_aggregate_leaf(const uint64_t* sorted, const uint32_t* sorted_idx,
                                 const uint32_t* uniq, const float4* pos, const uint32_t* parent,
                                 OctreeNode* nodes, uint64_t* unsorted, uint32_t* unsorted_idx,
                                 uint32_t quirk, uint32_t nt, uint32_t fp64) {
    uint32_t num_unique = parent[0];
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= num_unique) return;
    if (fp64)
        agg_leaf_one<Agg64>(sorted, sorted_idx, uniq, pos, num_unique, nodes, unsorted,
                            unsorted_idx, g, quirk, nt);
    else
        agg_leaf_one<Agg32>(sorted, sorted_idx, uniq, pos, num_unique, nodes, unsorted,
                            unsorted_idx, g, quirk, nt);
}
void launch_aggregate_leaf(const uint64_t* sorted, const uint32_t* sorted_idx,
                           const uint32_t* uniq, const float4* pos, const uint32_t* parent,
                           OctreeNode* nodes, uint64_t* unsorted, uint32_t* unsorted_idx,
                           uint32_t grid_count, uint32_t quirk, uint32_t nt, uint32_t fp64,
                           hipStream_t s) {
    if (grid_count)
        hipLaunchKernelGGL(k_aggregate_leaf, dim3((grid_count + 63) / 64), dim3(64), 0, s,
                           sorted, sorted_idx, uniq, pos, parent, nodes, unsorted, unsorted_idx,
                           quirk, nt, fp64);
}

// Octree.metal:120-202 (aggregateNodes): one thread per parent at `layer`.
template <typename AGG>
__device__ void agg_node_one(OctreeNode* nodes, const uint64_t* sorted,
                             const uint32_t* sorted_idx, const uint32_t* uniq,
                             uint32_t num_unique, uint32_t in_off, uint32_t out_off,
                             uint32_t layer, uint64_t* unsorted, uint32_t* unsorted_idx,
                             uint32_t child_count, uint32_t g) {
    uint32_t start = uniq[g];
    uint32_t end = g + 1 < num_unique ? uniq[g + 1] : child_count;
    AGG a;
    uint32_t ch[8];
    uint32_t nc = 0;
    for (int c = 0; c < 8; c++) ch[c] = ASTRO_INVALID_CHILD;
    for (uint32_t i = start; i < end; i++) {
        uint32_t ni = sorted_idx[i] + in_off;
        const OctreeNode* c = &nodes[ni];
        a.add(c->centerOfMass[0], c->centerOfMass[1], c->centerOfMass[2], c->totalMass);
        if (nc < 8) ch[nc++] = ni;
    }
    float cm[4];
    a.fin(cm);
    uint64_t code = sorted[start] >> 3;
    OctreeNode* o = &nodes[g + out_off];
    write_onode(o, code, cm, 0, 0, layer);
    for (int c = 0; c < 8; c++) o->children[c] = ch[c];
    unsorted[g] = code;
    unsorted_idx[g] = g;
}
__global__ void k_aggregate_nodes(OctreeNode* nodes, const uint64_t* sorted,
                                  const uint32_t* sorted_idx, const uint32_t* uniq,
                                  const uint32_t* parent, uint32_t in_off, uint32_t out_off,
                                  uint32_t layer, uint64_t* unsorted, uint32_t* unsorted_idx,
                                  const uint32_t* child, uint32_t fp64) {
    uint32_t num_unique = parent[0];
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= num_unique) return;
    uint32_t child_count = child[0];
    if (fp64)
        agg_node_one<Agg64>(nodes, sorted, sorted_idx, uniq, num_unique, in_off, out_off, layer,
                            unsorted, unsorted_idx, child_count, g);
    else
        agg_node_one<Agg32>(nodes, sorted, sorted_idx, uniq, num_unique, in_off, out_off, layer,
                            unsorted, unsorted_idx, child_count, g);
}
void launch_aggregate_nodes(OctreeNode* nodes, const uint64_t* sorted, const uint32_t* sorted_idx,
                            const uint32_t* uniq, const uint32_t* parent, uint32_t in_off,
                            uint32_t out_off, uint32_t layer, uint64_t* unsorted,
                            uint32_t* unsorted_idx, const uint32_t* child, uint32_t node_count,
                            uint32_t fp64, hipStream_t s) {
    if (node_count)
        hipLaunchKernelGGL(k_aggregate_nodes, dim3((node_count + 63) / 64), dim3(64), 0, s, nodes,
                           sorted, sorted_idx, uniq, parent, in_off, out_off, layer, unsorted,
                           unsorted_idx, child, fp64);
}

// Octree.metal:253-269 (clearOctreeNodeBuffer), guarded.
__global__ void k_clear_onodes(OctreeNode* nodes, uint64_t count) {
    // one thread per 16 B of the array: field k of an empty node
    float4* d = reinterpret_cast<float4*>(nodes);
    const float inv = __uint_as_float(0xFFFFFFFFu);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count * 8;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)(i & 7u);
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (k == 0) v = make_float4(inv, inv, 0.0f, 0.0f);  // mortonCode = INVALID
        if (k == 5 || k == 6) v = make_float4(inv, inv, inv, inv);  // children = INVALID
        d[i] = v;
    }
}
void launch_clear_onodes(OctreeNode* nodes, uint64_t count, hipStream_t s) {
    if (!count) return;
    uint64_t b = (count * 8 + 255) / 256;
    if (b > 16384) b = 16384;
    hipLaunchKernelGGL(k_clear_onodes, dim3((uint32_t)b), dim3(256), 0, s, nodes, count);
}

}  // namespace astro
