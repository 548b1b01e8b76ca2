// k_force.hip -- Barnes-Hut force traversal and integrator for gfx950.
//
// k_force_compact: the product traversal over the compact nodes (a 16-B centre and mass, and
// 16-B links).  Targets are processed in Morton order (sorted position j), so the 64 lanes of a
// wavefront walk nearly the same part of the tree; blocks are mapped so each XCD (and its L2)
// takes one contiguous stretch of the Morton curve.  The walk needs no stack: an opened node
// goes to its last child, any other to its `next` link.  Visit order equals the reference's
// LIFO order (children pushed 0..7, popped 7..0), so per-target sums are bit-identical to the
// oracle.
//
// k_force_reference: the reference kernel shape (BarnesHut.metal:22-101) over the 128-B
// OctreeNode array with children[8], for callers that built the tree with the stage calls.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "launch.hpp"

namespace astro {


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

// Lane masks: a condition's lanes as a wave-uniform 64-bit mask (SGPRs), and back.  Boolean
// combinations of conditions then run as scalar mask operations, where the compiler would
// otherwise materialise each condition in a vector register and select between them.
__device__ inline uint64_t lanes_of(bool b) { return __builtin_amdgcn_ballot_w64(b); }
__device__ inline bool lane_in(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// RN(1 / d) for d in [2^-45, 2^46): the hardware reciprocal refined by one Newton step.  The
// raw estimate differs from the correctly rounded reciprocal on about 11% of that range, the
// refined one on none (k_check_reciprocal checks every float of it).
__device__ inline float rcp_rn_mid(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, r, 1.0f);
    return __builtin_fmaf(e, r, r);
}

// q = RN(n / d) for d and n in [2^-45, 2^45] (n may also be 0), by Markstein's theorem: with
// y = RN(1/d), q0 = RN(n y) lies within one ulp of n / d, the residual n - d q0 is exact
// (one fma), and RN(q0 + residual y) = RN(n / d).  This replaces the compiler's IEEE division
// sequence (operand scaling, two remainder corrections, special-case fix-up), which changes
// nothing on that domain.
__device__ inline float div_rn_mid_y(float n, float d, float y) {  // y = RN(1 / d)
    const float q0 = n * y;
    const float t = __builtin_fmaf(-d, q0, n);
    return __builtin_fmaf(t, y, q0);
}
__device__ inline float div_rn_mid(float n, float d) { return div_rn_mid_y(n, d, rcp_rn_mid(d)); }

// Quadrupole term of an accepted internal node with second moments S (xx, yy, zz, xy, xz, yz)
// about its centre of mass, d = COM - target, D = r2 * r, monopole coefficient s = G M / D:
//   a = s d + u (w d - (S d + T/2 d)) = c d - u S d,
//   c = s + u (w - T/2),  u = 3G v,  w = 2.5 (d.S.d) (v D),  T = tr S,
// with v = 1/r^5 = (y y) r from y = RN(1 / D), the reciprocal the monopole term's division
// already computed (v D is 1/r2 to within a few roundings).  Folding the terms along d into
// one coefficient leaves one fma per component for them.
// The trace term is what Plummer softening adds to the expansion of  -G m / sqrt(|d|^2 + eps2).
struct QuadTerm {
    float c, u, sd0, sd1, sd2;  // the coefficient along d, u, and S d
    double phi;  // -G (1.5 d.S.d - 0.5 T r2) / (D r2), for the potential diagnostics
};
__device__ inline QuadTerm quad_term(float4 s0, float4 s1, float dx, float dy, float dz, float r2,
                                     float rr, float D, float y, float s, float G) {
    const float Sxx = s0.x, Syy = s0.y, Szz = s0.z, Sxy = s0.w, Sxz = s1.x, Syz = s1.y;
    const float Sd0 = __builtin_fmaf(Sxz, dz, __builtin_fmaf(Sxy, dy, Sxx * dx));
    const float Sd1 = __builtin_fmaf(Syz, dz, __builtin_fmaf(Syy, dy, Sxy * dx));
    const float Sd2 = __builtin_fmaf(Szz, dz, __builtin_fmaf(Syz, dy, Sxz * dx));
    const float dSd = __builtin_fmaf(dz, Sd2, __builtin_fmaf(dy, Sd1, dx * Sd0));
    const float T = (Sxx + Syy) + Szz;
    const float v = (y * y) * rr;
    const float w = (2.5f * dSd) * (v * D);
    QuadTerm t;
    t.u = (3.0f * G) * v;
    t.c = __builtin_fmaf(t.u, __builtin_fmaf(-0.5f, T, w), s);
    t.sd0 = Sd0;
    t.sd1 = Sd1;
    t.sd2 = Sd2;
    t.phi = -(double)G * (1.5 * (double)dSd - 0.5 * (double)T * (double)r2) /
            ((double)D * (double)r2);
    return t;
}

// F += c d - u S d, with the oracle's order of roundings for both accumulation modes.
template <bool KAHAN>
__device__ inline void accumulate_quad(Acc3& F, float dx, float dy, float dz, const QuadTerm& t) {
    if (KAHAN) {
        float y = __builtin_fmaf(t.c, dx, __builtin_fmaf(-t.u, t.sd0, -F.cx));
        float v = F.x + y;
        F.cx = (v - F.x) - y;
        F.x = v;
        y = __builtin_fmaf(t.c, dy, __builtin_fmaf(-t.u, t.sd1, -F.cy));
        v = F.y + y;
        F.cy = (v - F.y) - y;
        F.y = v;
        y = __builtin_fmaf(t.c, dz, __builtin_fmaf(-t.u, t.sd2, -F.cz));
        v = F.z + y;
        F.cz = (v - F.z) - y;
        F.z = v;
    } else {
        F.x = __builtin_fmaf(t.c, dx, __builtin_fmaf(-t.u, t.sd0, F.x));
        F.y = __builtin_fmaf(t.c, dy, __builtin_fmaf(-t.u, t.sd1, F.y));
        F.z = __builtin_fmaf(t.c, dz, __builtin_fmaf(-t.u, t.sd2, F.z));
    }
}

void set_accept_threshold(ForceArgs& a) {
    a.acc = make_accept_rule(a.p.theta);
    // For finite positions r2 = RN(|d|^2 + eps2) >= eps2 (rounding is monotone).  So when
    // eps2 >= r2_floor the r2 floor never binds; when RN(sqrt(eps2)) >= r_floor the r floor
    // never binds either; eps2 >= 2^-96 lets the square root skip its denormal scaling; and
    // leaf_cutoff_r2 < eps2 makes every leaf pass r2 > leaf_cutoff_r2.
    const float e2 = a.p.eps2, r2f = a.p.r2_floor, rf = a.p.r_floor, cut = a.p.leaf_cutoff_r2;
    const bool ok = std::isfinite(e2) && e2 >= 0x1p-96f && e2 >= r2f && std::sqrt(e2) >= rf &&
                    cut < e2;
    a.fast = ok ? 1u : 0u;
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

struct Visits {
    unsigned long long visits = 0, leafs = 0, monos = 0, opens = 0;
};

// The top of the tree in LDS.  A block copies the nodes of the top layers lt .. L-1 (the most
// that fit, chosen from the layer counts) into LDS before its walks start, and rewrites their
// links to name LDS slots (bit 31 set).  Every target walks the top layers, and its 64 lanes
// read scattered nodes there: from LDS those reads leave the texture address unit.
//   A top node is named with a marker bit that no global index has.  Monopoles: its slot with
//     bit 31 set (rows: centre and mass; last', info, next'), so that name << 4, the byte
//     offset of a global node's rows, is also the byte offset of its LDS rows (the marker
//     shifts out).  Quadrupoles: the byte offset of its LDS rows (slot << 4) with bit 26 set
//     (trees of fewer than 2^26 nodes; rows: centre and mass, Sxx Syy Szz Sxy, and Sxz Syz
//     (last' | layer << 27) next', the global moment rows' layout).  A visit reads either
//     source into the same registers and decodes both alike.
//   last' of a top node of layer l > lt names its last child (a top node of layer l - 1); a
//     node of layer lt keeps its child's global index (bit 31 clear); a top leaf (lt = 0) is
//     never opened.
//   next' names a top node (the next of a node of layer l lies at its head level, >= l), or
//     is kNextEnd.
// A lane leaves the top only by opening a node P of layer lt, and returns to it only from
// P's first-child chain, whose global next links all equal P's next: so a lane keeps P's
// next' (`resume`) and follows it wherever a global next points at or above off[lt].
constexpr uint32_t kTopBit = 0x80000000u;   // monopoles
constexpr uint32_t kTopQuad = 0x04000000u;  // quadrupoles (global indices below 2^26)
constexpr uint32_t kTopAddrMask = 0x0000FFFFu;  // quadrupoles: a top name's LDS byte offset
// slots per block: 80 KB of LDS per 1,024 threads (two blocks per CU), 32 B per node for
// monopoles (centre and mass; last', info, next', one free word), 48 B with the moment rows
template <int NT, bool QUAD>
struct TopCap {
    static constexpr uint32_t kSlots = (QUAD ? 1664u : 2544u) * (uint32_t)NT / 1024u;
};
template <bool QUAD>
__device__ inline uint32_t top_name(uint32_t slot) {
    return QUAD ? (kTopQuad | (slot << 4)) : (kTopBit | slot);
}
// LDS pointers keep their address space, so that the walk's LDS and global reads stay
// separate ds_read and global_load instructions (a generic pointer would turn both into flat
// loads)
#define ASTRO_LDS __attribute__((address_space(3)))
typedef float lds_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t lds_u3 __attribute__((ext_vector_type(3)));
struct TopView {
    const ASTRO_LDS lds_f4* pos;  // [slots] centre and mass
    const ASTRO_LDS uint4* lnk;   // [slots] last', info, next', 0 (monopoles)
    const ASTRO_LDS lds_f4* q0;   // [slots] Sxx, Syy, Szz, Sxy (quadrupoles)
    const ASTRO_LDS lds_f4* q1;   // [slots] Sxz, Syz, last' | layer << 27, next' (quadrupoles)
    uint32_t lt;         // first cached layer (L: none)
    uint32_t top_lo;     // off[lt]: global indices at or above it are top nodes (none:
                         // 0x7FFFFFFF for monopoles, kTopQuad for quadrupoles)
};

// One target's traversal.  MODE: 0 forces, 1 forces + visit counters, 2 forces + FP64
// potential.  QUAD: accepted internal nodes add their quadrupole term.  FAST: a.fast holds (no
// floor or cut-off can bind, square root without denormal scaling; exact for finite positions).
// FDIV: every G M / (r2 r) is in div_rn_mid's domain; with FAST and QUAD it also means every
// lane's rmin0 is in accept_r2min's domain, and nodes are then opened by r2 against
// r2min0 4^l, so that the square root, the division and the quadrupole term are computed in a
// branch that waves without an interacting lane skip.  (For monopoles the branch costs more
// than the square root and division it skips, so they stay straight-line.)
// OFF64: node byte offsets in 64 bits (trees of 2^27 nodes or more); otherwise 32-bit offsets
// from the uniform base, which the loads take as a scalar base plus a vector offset.
// G1: G = 1, so the products G M are the node masses.
//
// The walk needs no stack: every node holds `next`, the node visited after its subtree in the
// reference's order (the previous sibling, or for a first child its parent's next), so an
// opened node goes to its last child and any other node to its next, until the root's next
// (kNextEnd).  The visiting order, and with it the order of the force sums, is the
// reference's LIFO traversal (children pushed 0..7, the last popped first).
//
// A visit decides where the walk goes (distance, opening test, links) and issues the next
// node's loads before it computes its own interaction, which then runs while they are in
// flight.  Two register sets take turns (the loop visits twice per iteration), so the loads
// need no copies of the node they replace.
template <bool KAHAN, int MODE, bool QUAD, bool FAST, bool FDIV, bool OFF64, bool G1,
          bool TOPC>
__device__ __forceinline__ void traverse(const ForceArgs& a, const TopView& tv, uint32_t root,
                                         float4 p, uint32_t self, float rmin0, float r2min0,
                                         Acc3& F, double& phi, Visits& st) {
    constexpr bool R2 = FAST && FDIV && QUAD;
    // quadrupole trees below 2^27 nodes keep the links in the second moment row
    constexpr bool QLINKS = QUAD && !OFF64;
    const float eps2 = a.p.eps2, r2f = a.p.r2_floor, rf = a.p.r_floor;
    const float cut = a.p.leaf_cutoff_r2, G = a.p.G;
    const char* pos_base = reinterpret_cast<const char*>(a.tpos);
    const char* link_base = reinterpret_cast<const char*>(a.tlink);
    // Node indices are not checked per visit: the host traverses only a tree built on the
    // current layer table, and the kernel none whose build overflowed a layer, so every index
    // is the root, a written node's last child, or a written node's next.
    uint32_t resume = kNextEnd;  // TOPC: next' of the last node of layer lt this lane opened
    // a node as loaded (decoded where it is visited, so that no wait for the loads comes
    // before the previous node's interaction)
    struct Node {
        float4 cm, qm0, qm1;
        uint32_t last, info, next;  // monopoles, and quadrupole trees of 2^27 nodes or more
        uint32_t n;
    };
    auto fetch = [&](uint32_t n, Node& v) {
        v.n = n;
        if constexpr (TOPC && !QUAD) {
            // monopoles: one source or the other into the same registers; the LDS reads are
            // issued first, so that the global loads wait at most for them (not the reverse)
            const uint32_t o = n << 4;  // the rows' byte offset in either source
            if ((int32_t)n >= 0) {
                v.cm = *reinterpret_cast<const float4*>(pos_base + o);
                const TLink li = *reinterpret_cast<const TLink*>(link_base + o);
                v.last = li.last;
                v.info = li.info;
                v.next = li.next;
            } else {
                const lds_f4 c = *reinterpret_cast<const ASTRO_LDS lds_f4*>(
                    reinterpret_cast<const ASTRO_LDS char*>(tv.pos) + o);
                const lds_u3 k = *reinterpret_cast<const ASTRO_LDS lds_u3*>(
                    reinterpret_cast<const ASTRO_LDS char*>(tv.lnk) + o);
                v.cm = make_float4(c.x, c.y, c.z, c.w);
                v.last = k.x;
                v.info = k.y;
                v.next = k.z;
            }
        } else if constexpr (TOPC && QUAD) {
            // quadrupoles: the same, three rows from either source; a top node's rows carry
            // the links packed as in the global moment rows
            if ((int32_t)(n << 5) >= 0) {  // bit 26 clear: a global node
                v.cm = *reinterpret_cast<const float4*>(pos_base + (n << 4));
                const float4* qp = reinterpret_cast<const float4*>(
                    reinterpret_cast<const char*>(a.qmom) + (n << 5));
                v.qm0 = qp[0];
                v.qm1 = qp[1];
            } else {
                const uint32_t ad = n & kTopAddrMask;
                const lds_f4 c = *reinterpret_cast<const ASTRO_LDS lds_f4*>(
                    reinterpret_cast<const ASTRO_LDS char*>(tv.pos) + ad);
                const lds_f4 u = *reinterpret_cast<const ASTRO_LDS lds_f4*>(
                    reinterpret_cast<const ASTRO_LDS char*>(tv.q0) + ad);
                const lds_f4 w = *reinterpret_cast<const ASTRO_LDS lds_f4*>(
                    reinterpret_cast<const ASTRO_LDS char*>(tv.q1) + ad);
                v.cm = make_float4(c.x, c.y, c.z, c.w);
                v.qm0 = make_float4(u.x, u.y, u.z, u.w);
                v.qm1 = make_float4(w.x, w.y, w.z, w.w);
            }
        } else {
            // every node from device memory: one byte offset (16 B per node) into both
            // arrays, from uniform bases; 32-bit while the tree has fewer than 2^27 nodes
            const size_t o = OFF64 ? ((size_t)n << 4) : (size_t)(n << 4);
            v.cm = *reinterpret_cast<const float4*>(pos_base + o);
            // quadrupole mode: the moments are loaded with the node, not after the opening
            // test; below 2^27 nodes their second row also holds the links
            // (last | layer << 27, next)
            if (QUAD) {
                const float4* qp = reinterpret_cast<const float4*>(
                    reinterpret_cast<const char*>(a.qmom) +
                    (OFF64 ? ((size_t)n << 5) : (size_t)(n << 5)));
                v.qm0 = qp[0];
                v.qm1 = qp[1];
            }
            if (!QLINKS) {
                const TLink li = *reinterpret_cast<const TLink*>(link_base + o);
                v.last = li.last;
                v.info = li.info;
                v.next = li.next;
            }
        }
    };
    // Visits the node in v, loads the next one into w; returns whether the walk goes on.
    auto visit = [&](const Node& v, Node& w) -> bool {
        const float4 cm = v.cm;
        uint32_t last, next, layer;
        if constexpr (QLINKS) {
            const uint32_t lw = __float_as_uint(v.qm1.z);
            last = lw & kPackedIndexMask;
            layer = lw >> 27;
            next = __float_as_uint(v.qm1.w);
        } else {
            last = v.last;
            layer = tn_layer(v.info);
            next = v.next;
        }
        if constexpr (TOPC && !QUAD) {
            // a global node's next at or above off[lt] is a top node: back to the top through
            // the next' of the node of layer lt this lane opened last.  Top nodes' next' and
            // kNextEnd are negative as int32 and never match (a global node with next kNextEnd
            // lies on the root's first-child chain, where resume is kNextEnd too).
            if ((int32_t)next >= (int32_t)tv.top_lo) next = resume;
        } else if constexpr (TOPC && QUAD) {
            // a global node's next in [off[lt], 2^26) is a top node: back to the top through
            // the next' of the node of layer lt this lane opened last.  Top nodes' next' lie
            // at or above 2^26; kNextEnd is kept (a global node with next kNextEnd lies on the
            // root's first-child chain, where resume is kNextEnd too).
            if (next - tv.top_lo < kTopQuad - tv.top_lo) next = resume;
        }
        const float dx = cm.x - p.x, dy = cm.y - p.y, dz = cm.z - p.z;
        float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) + eps2;
        if (!FAST) r2 = fmaxf(r2, r2f);
        const bool leaf = layer == 0u;
        const float gm = G1 ? cm.w : G * cm.w;  // G1: G = 1, so RN(G M) = M
        bool open, interact, hit = false, far = false;
        float rr = 0.0f;
        if constexpr (R2) {
            // RN(sqrt(r2)) >= rmin0 2^l  <=>  r2 >= r2min0 4^l (accept.hpp): the square root
            // and the division are only computed where some lane interacts.  The empty asm
            // keeps the mass in the node's one wide load: sunk into the branch it would be a
            // second memory round trip on the loop's critical path.
            __asm__ volatile("" : : "v"(gm));
            const uint64_t mleaf = lanes_of(leaf), mfar = lanes_of(r2 >= ldexpf(r2min0, 2 * (int)layer));
            const uint64_t mself = lanes_of(v.n != self);
            interact = lane_in((mleaf & mself) | (~mleaf & mfar));
            open = lane_in(~mleaf & ~mfar);
        } else {
            // FDIV bounds r2 (every cube within +-div_lmax), as sqrt_rn_rsq needs
            rr = FAST ? (FDIV ? sqrt_rn_rsq(r2) : sqrt_rn_normal(r2)) : fmaxf(sqrtf(r2), rf);
            // (plain selects here: mask operations measured slower on this path)
            far = rr >= ldexpf(rmin0, (int)layer);
            hit = (v.n != self) & (FAST || r2 > cut);
            interact = leaf ? hit : far;
            open = !leaf & !far;
        }
        // only top nodes have layer lt (the rest of the tree lies below it)
        if (TOPC && open && layer == tv.lt) resume = next;
        const uint32_t nn = open ? last : next;
        const bool more = nn != kNextEnd;
        fetch(more ? nn : v.n, w);  // a lane about to finish reads its node again
        if constexpr (R2) {
            if (interact) {
                const float rq = sqrt_rn_rsq(r2);  // FDIV: r2 is bounded
                const float D = r2 * rq;
                const float y = rcp_rn_mid(D);
                const float s = div_rn_mid_y(gm, D, y);
                if (KAHAN) {
                    if (QUAD && !leaf) {
                        const QuadTerm qt = quad_term(v.qm0, v.qm1, dx, dy, dz, r2, rq, D, y, s, G);
                        accumulate_quad<true>(F, dx, dy, dz, qt);
                        if (MODE == 2) phi += qt.phi - (double)gm / (double)rq;
                    } else {
                        accumulate<true>(F, s, dx, dy, dz);
                        if (MODE == 2) phi -= (double)gm / (double)rq;
                    }
                } else {
                    // an internal node subtracts u S d and turns s into c; then every
                    // interacting lane adds its coefficient times d (all of a leaf's term)
                    float c = s, fx = F.x, fy = F.y, fz = F.z;
                    if (QUAD && !leaf) {
                        const QuadTerm qt = quad_term(v.qm0, v.qm1, dx, dy, dz, r2, rq, D, y, s, G);
                        c = qt.c;
                        fx = __builtin_fmaf(-qt.u, qt.sd0, fx);
                        fy = __builtin_fmaf(-qt.u, qt.sd1, fy);
                        fz = __builtin_fmaf(-qt.u, qt.sd2, fz);
                        if (MODE == 2) phi += qt.phi - (double)gm / (double)rq;
                    } else if (MODE == 2) {
                        phi -= (double)gm / (double)rq;
                    }
                    F.x = __builtin_fmaf(c, dx, fx);
                    F.y = __builtin_fmaf(c, dy, fy);
                    F.z = __builtin_fmaf(c, dz, fz);
                }
            }
        } else {
            // the term is computed for every lane and kept only where it interacts
            const float D = r2 * rr;
            const float s = FDIV ? div_rn_mid(gm, D) : gm / D;
            if (QUAD && interact && !leaf) {
                const float y = FDIV ? rcp_rn_mid(D) : 1.0f / D;
                const QuadTerm qt = quad_term(v.qm0, v.qm1, dx, dy, dz, r2, rr, D, y, s, G);
                accumulate_quad<KAHAN>(F, dx, dy, dz, qt);
                if (MODE == 2) phi += qt.phi - (double)gm / (double)rr;
            } else if (KAHAN) {
                Acc3 T = F;
                accumulate<true>(T, s, dx, dy, dz);
                F.x = interact ? T.x : F.x;
                F.y = interact ? T.y : F.y;
                F.z = interact ? T.z : F.z;
                F.cx = interact ? T.cx : F.cx;
                F.cy = interact ? T.cy : F.cy;
                F.cz = interact ? T.cz : F.cz;
            } else {
                // F never holds -0 (it starts at +0, and x + (-x) rounds to +0), so the zero
                // term fma(0, d, F) leaves it unchanged for finite d
                const float s_hit = hit ? s : 0.0f, s_acc = far ? s : 0.0f;
                const float sk = leaf ? s_hit : s_acc;
                F.x = __builtin_fmaf(sk, dx, F.x);
                F.y = __builtin_fmaf(sk, dy, F.y);
                F.z = __builtin_fmaf(sk, dz, F.z);
            }
            if (MODE == 2 && interact && !(QUAD && !leaf)) phi -= (double)gm / (double)rr;
        }
        if (MODE == 1) {
            st.visits += 1;
            st.leafs += leaf && interact;
            st.monos += !leaf && interact;
            st.opens += open;
        }
        return more;
    };
    Node A, B;
    fetch(root, A);
    while (visit(A, B) && visit(B, A)) {
    }
}

// The LDS top of the tree (TOPC): the cached layers lt .. L-1 are the most, from the root
// down, whose node counts sum to at most `slots` (in a batch, the counts of system S's tree,
// whose nodes of layer l are the index range lo[l] ..); slots are given out layer by layer
// from lt up (base[l]).  Thread 0 decides, every thread then copies its share of the nodes with their
// links rewritten.  Returns the view the walks read.
template <int NT, bool QUAD>
__device__ TopView fill_top(const ForceArgs& a, uint32_t S, float4* s_pos, uint4* s_lnk,
                            float4* s_q0, float4* s_q1, uint32_t* s_base, uint32_t* s_lo,
                            uint32_t* s_meta) {
    constexpr uint32_t slots = TopCap<NT, QUAD>::kSlots;
    const uint32_t L = a.layers, tid = threadIdx.x;
    // the cached tree's nodes of each layer: a batch's system S (its first node of the layer
    // and, through the next system's first node, its count), else the whole layer
    if (tid < L) {
        uint32_t lo = 0, c = a.counts[tid];
        if (a.sysfirst) {
            lo = a.sysfirst[(size_t)S * kMaxLayers + tid];
            if (S + 1u < a.nsys) c = a.sysfirst[(size_t)(S + 1u) * kMaxLayers + tid];
            c -= lo;
        }
        s_lo[tid] = lo;
        s_base[tid] = c;  // (the counts until thread 0 turns them into slot bases)
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t lt = L, K = 0;
        const uint32_t wl = a.wlow ? *a.wlow : 0u;  // the lowest layer the build wrote
        for (uint32_t l = L; l-- > 0;) {
            if (l < wl) break;
            const uint32_t c = s_base[l];
            if (K + c > slots) break;
            K += c;
            lt = l;
        }
        uint32_t b = 0;
        for (uint32_t l = lt; l < L; l++) {
            const uint32_t c = s_base[l];
            s_base[l] = b;
            b += c;
        }
        s_meta[0] = lt;
        s_meta[1] = K;
    }
    __syncthreads();
    const uint32_t lt = s_meta[0], K = s_meta[1];
    for (uint32_t s = tid; s < K; s += NT) {
        uint32_t l = lt;
        while (l + 1 < L && s >= s_base[l + 1]) l++;
        const uint32_t n = (uint32_t)a.off[l] + s_lo[l] + (s - s_base[l]);
        const TLink k = a.tlink[n];
        s_pos[s] = a.tpos[n];
        uint32_t last = k.last, next = k.next;
        if (l > lt)
            last = top_name<QUAD>(last - (uint32_t)a.off[l - 1] - s_lo[l - 1] + s_base[l - 1]);
        else if (l == 0u) last = QUAD ? kPackedNoChild : kNextEnd;  // a leaf: never opened
        if (next != kNextEnd) {  // (a node's next lies in its own system's tree)
            const uint32_t h = tn_level(k.info);  // the layer of the next node
            next = top_name<QUAD>(next - (uint32_t)a.off[h] - s_lo[h] + s_base[h]);
        }
        if (QUAD) {
            s_q0[s] = a.qmom[2 * (size_t)n];
            const float4 q = a.qmom[2 * (size_t)n + 1];
            s_q1[s] = make_float4(q.x, q.y, __uint_as_float(last | (l << 27)),
                                  __uint_as_float(next));
        } else {
            s_lnk[s] = make_uint4(last, k.info, next, 0u);
        }
    }
    __syncthreads();
    TopView tv;
    tv.pos = (const ASTRO_LDS lds_f4*)s_pos;
    tv.lnk = (const ASTRO_LDS uint4*)s_lnk;
    tv.q0 = (const ASTRO_LDS lds_f4*)s_q0;
    tv.q1 = (const ASTRO_LDS lds_f4*)s_q1;
    tv.lt = lt;
    tv.top_lo = lt < L ? (uint32_t)a.off[lt] : (QUAD ? kTopQuad : 0x7FFFFFFFu);
    return tv;
}

// Targets in Morton order (sorted position j), one per thread.  TOPC: the top layers are read
// from LDS (fill_top); needs 32-bit node offsets.
template <int NT, bool KAHAN, int MODE, bool QUAD, bool FAST, bool OFF64, bool G1, bool TOPC>
__global__ __launch_bounds__(NT) void k_force_compact(ForceArgs a) {
    static_assert(!(TOPC && OFF64), "the LDS top needs 32-bit node offsets");
    constexpr uint32_t kSl = TOPC ? TopCap<NT, QUAD>::kSlots : 1u;
    __shared__ float4 s_pos[kSl];
    __shared__ uint4 s_lnk[TOPC && !QUAD ? kSl : 1u];
    __shared__ float4 s_q0[TOPC && QUAD ? kSl : 1u];
    __shared__ float4 s_q1[TOPC && QUAD ? kSl : 1u];
    __shared__ uint32_t s_base[kMaxLayers];
    __shared__ uint32_t s_lo[kMaxLayers];
    __shared__ uint32_t s_meta[2];
    // XCD-aware mapping: hardware places block b on XCD b % 8.  xcd_chunk 0 gives XCD x the
    // x-th contiguous eighth of the target tiles (best L2 locality); xcd_chunk K deals runs of
    // K consecutive tiles to the XCDs in turn (better balance when work varies along the
    // curve).  Either way the mapping is a bijection on [0, gridDim.x).
    const uint32_t nb = gridDim.x, b = blockIdx.x, K = a.xcd_chunk;
    uint32_t tile;
    if (K == 0u) {
        const uint32_t xcd = b & 7u, q = nb >> 3, r = nb & 7u;
        tile = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    } else {
        const uint32_t full = (nb / (8u * K)) * 8u * K, i = b >> 3;
        tile = b < full ? ((i / K) * 8u + (b & 7u)) * K + (i % K) : b;
    }
    const uint32_t tid = threadIdx.x;
    const uint32_t j = tile * NT + tid;
    // a build that overflowed a layer left the tree incomplete: no traversal (the error word,
    // which astro_check reports and clears, already says so)
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(a.err, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) &
        ERR_CAPACITY)
        return;
    // a batch's block caches the top of the system of its middle target (nearly always the
    // system of all of them, and of most of them when the block straddles two); its other
    // targets walk from device memory
    const uint32_t jm = min(tile * NT + NT / 2u, a.ns - 1u);
    const uint32_t S = (TOPC && a.sysfirst && a.ns) ? a.sys_of[jm] : 0u;
    TopView tv{};
    if (TOPC) tv = fill_top<NT, QUAD>(a, S, s_pos, s_lnk, s_q0, s_q1, s_base, s_lo, s_meta);
    const bool active = j < a.ns;
    const uint32_t sys = (active && a.sys_of) ? a.sys_of[j] : 0u;
    // the fast division holds for the whole wave when every lane's system cube lies within
    // +-div_lmax (then |d| <= 2 div_lmax per axis; masses are bounded on the host)
    // a node at layer l is accepted iff r >= rmin0 * 2^l (accept.hpp)
    float rmin0 = 0.0f;
    bool fdiv_lane = true;
    if (active) {
        const float4 cu = a.cubes[sys];
        float side0;
        if (a.p.side_rule == ASTRO_SIDE_SHIFTED) side0 = 2097152.0f;  // 2^21
        else side0 = ldexpf(cu.w, -(int)a.p.bits_per_axis);
        rmin0 = accept_rmin(side0, a.acc);
        const float L = a.div_lmax;
        fdiv_lane = L > 0.0f && fabsf(cu.x) <= L && fabsf(cu.y) <= L && fabsf(cu.z) <= L &&
                    fabsf(cu.x + cu.w) <= L && fabsf(cu.y + cu.w) <= L && fabsf(cu.z + cu.w) <= L;
        if (FAST && QUAD) fdiv_lane = fdiv_lane && accept_r2_ok(rmin0);
    }
    const bool fdiv = __builtin_amdgcn_readfirstlane(__all(fdiv_lane ? 1 : 0)) != 0;

    Visits st;
    if (active) {
        const uint32_t pidx = j < a.nt ? a.sorted_idx[j] : j;
        // in a step, the sort left the targets' positions and types in sorted order
        const bool sorted = a.spos && j < a.nt;
        const float4 p = sorted ? a.spos[j] : a.pos[pidx];
        const uint32_t type = sorted ? a.stype[j] : a.type[pidx];
        uint32_t self =
            (type != 2u && j < a.nt) ? (uint32_t)(a.off0 + a.leaf_of_sorted[j]) : 0xFFFFFFFFu;
        uint32_t root = (uint32_t)(a.root_off + sys);
        if (TOPC && tv.lt < a.layers) {
            if (sys == S) {
                // the root is a top node; so are the leaves when the whole tree is cached
                root = top_name<QUAD>(s_base[a.layers - 1u] + sys - s_lo[a.layers - 1u]);
                if (tv.lt == 0u && self != 0xFFFFFFFFu)
                    self = top_name<QUAD>(self - (uint32_t)a.off0 - s_lo[0] + s_base[0]);
            } else {
                tv.top_lo = QUAD ? kTopQuad : 0x7FFFFFFFu;  // no global link leads to the top
            }
        }
        Acc3 F;
        double phi = 0.0;
        if (fdiv)
            traverse<KAHAN, MODE, QUAD, FAST, true, OFF64, G1, TOPC>(
                a, tv, root, p, self, rmin0, QUAD ? accept_r2min(rmin0) : 0.0f, F, phi, st);
        else
            traverse<KAHAN, MODE, QUAD, FAST, false, OFF64, G1, TOPC>(a, tv, root, p, self, rmin0,
                                                                      0.0f, F, phi, st);
        const float c = a.p.force_clamp;
        a.force[pidx] = make_float4(clampf(F.x, c), clampf(F.y, c), clampf(F.z, c), 0.0f);
        if (MODE == 2) a.phi[pidx] = phi;
    }
    if (MODE == 1) {
        const unsigned long long v = wave_sum_u64(st.visits), lf = wave_sum_u64(st.leafs),
                                 mo = wave_sum_u64(st.monos), op = wave_sum_u64(st.opens);
        if ((tid & 63u) == 0u) {
            atomicAdd(&a.stats[0], v);
            atomicAdd(&a.stats[1], lf);
            atomicAdd(&a.stats[2], mo);
            atomicAdd(&a.stats[3], op);
        }
    }
}

// Self-check: div_rn_mid against '/' on n pairs drawn log-uniformly from its domain
// (numerator and denominator in [2^-45, 2^45]), counting mismatching bit patterns.
__global__ void k_check_division(uint64_t n, uint64_t seed, unsigned long long* mismatch) {
    unsigned long long bad = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t w[2];
        for (int k = 0; k < 2; k++) {  // SplitMix64, two draws per pair
            uint64_t z = seed + (2 * i + k + 1) * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            w[k] = z ^ (z >> 31);
        }
        // exponents in [-45, 44], random 23-bit mantissas
        const uint32_t ex = (uint32_t)(w[0] >> 32) % 90u, ey = (uint32_t)(w[1] >> 32) % 90u;
        const float x = __uint_as_float(((ex + 127u - 45u) << 23) | ((uint32_t)w[0] & 0x7FFFFFu));
        const float y = __uint_as_float(((ey + 127u - 45u) << 23) | ((uint32_t)w[1] & 0x7FFFFFu));
        const float q1 = div_rn_mid(x, y), q2 = x / y;
        bad += __float_as_uint(q1) != __float_as_uint(q2);
    }
    bad = wave_sum_u64(bad);
    if ((threadIdx.x & 63u) == 0u && bad) atomicAdd(mismatch, bad);
}
int check_division(uint64_t n, uint64_t seed, unsigned long long* d_mismatch, hipStream_t s) {
    hipLaunchKernelGGL(k_check_division, dim3(2048), dim3(256), 0, s, n, seed, d_mismatch);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Self-check of the r2 form of the opening test: for T = rmin0 drawn with exponents in
// [-62, 40] (one in eight a power of two) and a layer l in [0, 21], compares
// RN(sqrt(x)) >= T 2^l  with  x >= accept_r2min(T) 4^l  on x within 8 ulps of the threshold
// (three in four draws) or drawn from the binades around it, x >= 2^-96.
__global__ void k_check_open_test(uint64_t n, uint64_t seed, unsigned long long* mismatch) {
    unsigned long long bad = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t w[2];
        for (int k = 0; k < 2; k++) {  // SplitMix64
            uint64_t z = seed + (2 * i + k + 1) * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            w[k] = z ^ (z >> 31);
        }
        const int e = (int)((w[0] >> 32) % 103u) - 62;
        const uint32_t mant = ((w[0] >> 29) & 7u) == 0u ? 0u : ((uint32_t)w[0] & 0x7FFFFFu);
        const float T = __uint_as_float(((uint32_t)(e + 127) << 23) | mant);
        const int l = (int)((w[1] >> 40) % 22u);
        const float Tl = ldexpf(T, l), Cl = ldexpf(accept_r2min(T), 2 * l);
        float x;
        if (((w[1] >> 62) & 3u) != 0u) {
            x = __uint_as_float(__float_as_uint(Cl) + (uint32_t)((w[1] >> 8) & 15u) - 8u);
        } else {
            const int ex = 2 * (e + l) + (int)((w[1] >> 24) % 5u) - 2;
            x = __uint_as_float(((uint32_t)(ex + 127) << 23) | ((uint32_t)w[1] & 0x7FFFFFu));
        }
        if (!(x >= 0x1p-96f) || !(x < INFINITY)) continue;
        const bool by_sqrt = sqrt_rn_normal(x) >= Tl, by_r2 = x >= Cl;
        bad += (by_sqrt != by_r2) || (by_sqrt != (sqrtf(x) >= Tl));
    }
    bad = wave_sum_u64(bad);
    if ((threadIdx.x & 63u) == 0u && bad) atomicAdd(mismatch, bad);
}
int check_open_test(uint64_t n, uint64_t seed, unsigned long long* d_mismatch, hipStream_t s) {
    hipLaunchKernelGGL(k_check_open_test, dim3(2048), dim3(256), 0, s, n, seed, d_mismatch);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Exhaustive self-check of both traversal square roots, sqrt_rn_normal and sqrt_rn_rsq: every
// float x in [2^-96, FLT_MAX] against the double square root rounded to float (correctly
// rounded: 53 >= 2 * 24 + 2 bits make the double rounding innocuous); a mismatch of either
// counts.  The raw hardware estimate alone differs on about 15% of them.
__global__ void k_check_sqrt(uint32_t lo, uint32_t hi, unsigned long long* mismatch) {
    unsigned long long bad = 0;
    for (uint64_t u = lo + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < hi;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const float x = __uint_as_float((uint32_t)u);
        const float ref = (float)__builtin_sqrt((double)x);
        bad += __float_as_uint(sqrt_rn_normal(x)) != __float_as_uint(ref);
        bad += __float_as_uint(sqrt_rn_rsq(x)) != __float_as_uint(ref);
    }
    bad = wave_sum_u64(bad);
    if ((threadIdx.x & 63u) == 0u && bad) atomicAdd(mismatch, bad);
}
uint64_t check_sqrt(unsigned long long* d_mismatch, hipStream_t s) {
    const uint32_t lo = (127u - 96u) << 23, hi = 0x7F800000u;  // [2^-96, +inf)
    hipLaunchKernelGGL(k_check_sqrt, dim3(8192), dim3(256), 0, s, lo, hi, d_mismatch);
    return hipGetLastError() == hipSuccess ? (uint64_t)(hi - lo) : 0u;
}

// Exhaustive self-check of rcp_rn_mid, the premise of div_rn_mid: every float d in
// [2^-45, 2^46) against the double reciprocal rounded to float (correctly rounded: the double
// division is, and 53 >= 2 * 24 + 2 bits make the second rounding innocuous).
__global__ void k_check_reciprocal(uint32_t lo, uint32_t hi, unsigned long long* mismatch) {
    unsigned long long bad = 0;
    for (uint64_t u = lo + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < hi;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const float d = __uint_as_float((uint32_t)u);
        const float ref = (float)(1.0 / (double)d);
        bad += __float_as_uint(rcp_rn_mid(d)) != __float_as_uint(ref);
    }
    bad = wave_sum_u64(bad);
    if ((threadIdx.x & 63u) == 0u && bad) atomicAdd(mismatch, bad);
}
uint64_t check_reciprocal(unsigned long long* d_mismatch, hipStream_t s) {
    const uint32_t lo = (127u - 45u) << 23, hi = (127u + 46u) << 23;  // [2^-45, 2^46)
    hipLaunchKernelGGL(k_check_reciprocal, dim3(8192), dim3(256), 0, s, lo, hi, d_mismatch);
    return hipGetLastError() == hipSuccess ? (uint64_t)(hi - lo) : 0u;
}

// One walk per wave (monopoles, one system, the fast opening test).  The wave visits the union
// of its lanes' walks in the walks' order: it opens a node when any lane that takes part there
// opens it, and a lane takes part at the nodes of its own walk only -- having accepted a node
// the wave opens for another lane, it sits out that node's subtree, until the wave reaches the
// node's next.  So every lane meets its own nodes in its own order, and its sums are the
// per-lane walk's, bit for bit.  The node index is uniform, so nodes come through scalar loads
// (the constant address space); a lane that takes no part in a visit adds fma(0, d, F), which
// leaves F as it is (F never holds -0).
typedef uint32_t cu4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) lds_f4* const_f4;
typedef const __attribute__((address_space(4))) cu4* const_u4;
template <bool FDIV, bool G1>
__device__ __forceinline__ void walk_wave(const ForceArgs& a, float4 p, uint32_t self, float rmin0,
                                          bool lane_on, Acc3& F) {
    constexpr uint32_t kOn = 0xFFFFFFFEu;  // taking part (otherwise: the node to resume at)
    const const_f4 tp = (const_f4)(uintptr_t)a.tpos;
    const const_u4 tl = (const_u4)(uintptr_t)a.tlink;
    const float eps2 = a.p.eps2, G = a.p.G;
    uint32_t resume = lane_on ? kOn : kNextEnd;
    uint32_t n = (uint32_t)a.root_off;
    while (n != kNextEnd) {
        n = __builtin_amdgcn_readfirstlane(n);
        const lds_f4 cm = tp[n];
        const cu4 lk = tl[n];  // last, info, next, start
        const uint32_t layer = lk.y & 31u;
        if (resume == n) resume = kOn;
        const bool on = resume == kOn;
        const float dx = cm.x - p.x, dy = cm.y - p.y, dz = cm.z - p.z;
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) + eps2;
        const float rr = FDIV ? sqrt_rn_rsq(r2) : sqrt_rn_normal(r2);
        const bool far = rr >= ldexpf(rmin0, (int)layer);
        const bool leaf = layer == 0u;
        const float gm = G1 ? cm.w : G * cm.w;
        const float D = r2 * rr;
        const float s = FDIV ? div_rn_mid(gm, D) : gm / D;
        const bool interact = on && (leaf ? n != self : far);
        const float sk = interact ? s : 0.0f;
        F.x = __builtin_fmaf(sk, dx, F.x);
        F.y = __builtin_fmaf(sk, dy, F.y);
        F.z = __builtin_fmaf(sk, dz, F.z);
        const bool want = on && !leaf && !far;
        if (__ballot(want) == 0ull) {
            n = lk.z;
        } else if (lk.x == kNextEnd) {
            // a node of the lowest written layer (its children were left out): a lane that
            // opens it ends its walk there, as its own walk would; the wave goes on
            if (want) resume = kNextEnd;
            n = lk.z;
        } else {
            if (on && !leaf && far) resume = lk.z;  // sit out this subtree
            n = lk.x;
        }
    }
}

template <int NT, bool G1>
__global__ __launch_bounds__(NT) void k_force_wave(ForceArgs a) {
    // the targets' tiles as in k_force_compact
    const uint32_t nb = gridDim.x, b = blockIdx.x, K = a.xcd_chunk;
    uint32_t tile;
    if (K == 0u) {
        const uint32_t xcd = b & 7u, q = nb >> 3, r = nb & 7u;
        tile = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    } else {
        const uint32_t full = (nb / (8u * K)) * 8u * K, i = b >> 3;
        tile = b < full ? ((i / K) * 8u + (b & 7u)) * K + (i % K) : b;
    }
    const uint32_t j = tile * NT + threadIdx.x;
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(a.err, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) &
        ERR_CAPACITY)
        return;
    const bool active = j < a.ns;
    const float4 cu = a.cubes[0];
    const float side0 = a.p.side_rule == ASTRO_SIDE_SHIFTED
                            ? 2097152.0f
                            : ldexpf(cu.w, -(int)a.p.bits_per_axis);
    const float rmin0 = accept_rmin(side0, a.acc);
    const float L = a.div_lmax;
    const bool fdiv = L > 0.0f && fabsf(cu.x) <= L && fabsf(cu.y) <= L && fabsf(cu.z) <= L &&
                      fabsf(cu.x + cu.w) <= L && fabsf(cu.y + cu.w) <= L &&
                      fabsf(cu.z + cu.w) <= L;
    uint32_t pidx = 0, self = 0xFFFFFFFFu;
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (active) {
        pidx = j < a.nt ? a.sorted_idx[j] : j;
        const bool sorted = a.spos && j < a.nt;
        p = sorted ? a.spos[j] : a.pos[pidx];
        const uint32_t type = sorted ? a.stype[j] : a.type[pidx];
        if (type != 2u && j < a.nt) self = (uint32_t)(a.off0 + a.leaf_of_sorted[j]);
    }
    Acc3 F;
    if (fdiv) walk_wave<true, G1>(a, p, self, rmin0, active, F);
    else walk_wave<false, G1>(a, p, self, rmin0, active, F);
    if (active) {
        const float c = a.p.force_clamp;
        a.force[pidx] = make_float4(clampf(F.x, c), clampf(F.y, c), clampf(F.z, c), 0.0f);
    }
}

// The wave walk applies: force computations only, one system, monopoles, the fast opening
// test without Kahan sums, 32-bit node offsets.
static bool wave_walk_applies(const ForceArgs& a, int mode) {
    return a.walk && mode == 0 && !a.sys_of && a.p.multipole != 2 && a.fast && !a.p.kahan &&
           !force_off64(a);
}
template <int NT>
static void launch_force_wave(const ForceArgs& a, hipStream_t s) {
    const uint32_t blocks = (a.ns + NT - 1) / NT;
    if (a.p.G == 1.0f)
        hipLaunchKernelGGL((k_force_wave<NT, true>), dim3(blocks), dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL((k_force_wave<NT, false>), dim3(blocks), dim3(NT), 0, s, a);
}

template <int NT, bool KAHAN, bool QUAD, bool FAST, bool OFF64, bool G1, bool TOPC>
static void launch_force_modes(const ForceArgs& a, int mode, hipStream_t s) {
    const uint32_t blocks = (a.ns + NT - 1) / NT;
    if (mode == 0)
        hipLaunchKernelGGL((k_force_compact<NT, KAHAN, 0, QUAD, FAST, OFF64, G1, TOPC>),
                           dim3(blocks), dim3(NT), 0, s, a);
    else if (mode == 1)
        hipLaunchKernelGGL((k_force_compact<NT, KAHAN, 1, QUAD, FAST, OFF64, G1, TOPC>),
                           dim3(blocks), dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL((k_force_compact<NT, KAHAN, 2, QUAD, FAST, OFF64, G1, TOPC>),
                           dim3(blocks), dim3(NT), 0, s, a);
}

// 64-bit offsets from 2^27 nodes (needed from 2^28), where the moment rows stop carrying
// links; `wide` asks for them on any tree (a check of that path on small trees)
bool force_off64(const ForceArgs& a) { return a.node_total >= kLinkPackNodes || a.wide; }

template <int NT, bool KAHAN, bool QUAD, bool FAST, bool G1>
static void launch_force_off(const ForceArgs& a, int mode, hipStream_t s) {
    constexpr bool kTopBlock = NT >= 512;  // the LDS top needs many waves per block
    // quadrupole top nodes are marked by bit 26, which global indices below 2^26 leave clear
    const bool top_fits = !QUAD || a.node_total < (uint64_t)kTopQuad;
    if (force_off64(a)) {
        launch_force_modes<NT, KAHAN, QUAD, FAST, true, G1, false>(a, mode, s);
    } else if (kTopBlock && a.top && top_fits) {
        launch_force_modes<NT, KAHAN, QUAD, FAST, false, G1, kTopBlock>(a, mode, s);
    } else {
        launch_force_modes<NT, KAHAN, QUAD, FAST, false, G1, false>(a, mode, s);
    }
}

template <int NT, bool KAHAN, bool QUAD, bool FAST>
static void launch_force_variant(const ForceArgs& a, int mode, hipStream_t s) {
    // G = 1 (the physics profile): the products G M are the masses themselves; compiled for
    // the plain fast paths only
    if constexpr (FAST && !KAHAN) {
        if (a.p.G == 1.0f) {
            launch_force_off<NT, KAHAN, QUAD, FAST, true>(a, mode, s);
            return;
        }
    }
    launch_force_off<NT, KAHAN, QUAD, FAST, false>(a, mode, s);
}

template <int NT>
static void launch_force_nt(const ForceArgs& a, int mode, hipStream_t s) {
    const bool k = a.p.kahan != 0, q = a.p.multipole == 2, f = a.fast != 0;
    if (k) {
        if (q) f ? launch_force_variant<NT, true, true, true>(a, mode, s)
               : launch_force_variant<NT, true, true, false>(a, mode, s);
        else f ? launch_force_variant<NT, true, false, true>(a, mode, s)
               : launch_force_variant<NT, true, false, false>(a, mode, s);
    } else {
        if (q) f ? launch_force_variant<NT, false, true, true>(a, mode, s)
               : launch_force_variant<NT, false, true, false>(a, mode, s);
        else f ? launch_force_variant<NT, false, false, true>(a, mode, s)
               : launch_force_variant<NT, false, false, false>(a, mode, s);
    }
}

bool force_block_supported(uint32_t nt) {
    return nt == 64 || nt == 128 || nt == 256 || nt == 512 || nt == 1024;
}

void launch_force_compact(const ForceArgs& a, hipStream_t s) {
    if (!a.ns) return;
    const int mode = a.phi ? 2 : (a.stats ? 1 : 0);
    if (wave_walk_applies(a, mode)) {
        switch (a.block) {
            case 64: launch_force_wave<64>(a, s); break;
            case 128: launch_force_wave<128>(a, s); break;
            case 512: launch_force_wave<512>(a, s); break;
            case 1024: launch_force_wave<1024>(a, s); break;
            default: launch_force_wave<256>(a, s); break;
        }
        return;
    }
    switch (a.block) {
        case 64: launch_force_nt<64>(a, mode, s); break;
        case 128: launch_force_nt<128>(a, mode, s); break;
        case 512: launch_force_nt<512>(a, mode, s); break;
        case 1024: launch_force_nt<1024>(a, mode, s); break;
        default: launch_force_nt<256>(a, mode, s); break;
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
        if (ni >= a.node_total) {  // empty the stack; the loop leaves at its head
            atomicOr(a.err, ERR_NODE_INDEX);
            sp = 0;
            continue;
        }
        const OctreeNode* nd = &a.onodes[ni];
        if (nd->mortonCode == ASTRO_INVALID_MORTON_CODE) continue;
        const float dx = nd->centerOfMass[0] - p.x;
        const float dy = nd->centerOfMass[1] - p.y;
        const float dz = nd->centerOfMass[2] - p.z;
        float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) + a.p.eps2;
        r2 = fmaxf(r2, a.p.r2_floor);
        const float rr = fmaxf(sqrtf(r2), a.p.r_floor);
        const uint32_t layer = nd->layer;
        bool interact;
        if (layer == 0u) {
            const bool self = has_self && mycode == nd->mortonCode;
            interact = !self && r2 > a.p.leaf_cutoff_r2;
        } else {
            const float side = ldexpf(side0, (int)layer);
            interact = accept_exact(side, rr, a.acc);
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
            const float D = r2 * rr;
            const float s = ((a.p.G * nd->totalMass) / D);
            if (a.p.multipole == 2 && layer != 0u) {
                // second moments sit in emittedColor (xx, yy, zz, xy) / emittedColorCenter
                const float4 s0 = *reinterpret_cast<const float4*>(nd->emittedColor);
                const float4 s1 = *reinterpret_cast<const float4*>(nd->emittedColorCenter);
                const QuadTerm qt = quad_term(s0, s1, dx, dy, dz, r2, rr, D, 1.0f / D, s, a.p.G);
                if (a.p.kahan) accumulate_quad<true>(F, dx, dy, dz, qt);
                else accumulate_quad<false>(F, dx, dy, dz, qt);
            } else if (a.p.kahan) {
                accumulate<true>(F, s, dx, dy, dz);
            } else {
                accumulate<false>(F, s, dx, dy, dz);
            }
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
// x = x + v*dt -- separate roundings, no contraction.  guard (optional): flags the system of
// every particle whose new position or velocity is not finite or whose position leaves
// [-extent, extent]^3 (the server evicts such systems; the arithmetic is unchanged).
__global__ __launch_bounds__(256) void k_update(float4* __restrict__ pos, float4* __restrict__ vel,
                                                const float4* __restrict__ force,
                                                const uint32_t* __restrict__ type, uint32_t n,
                                                float dt, astro_params p,
                                                const uint32_t* __restrict__ sys_of,
                                                uint32_t* __restrict__ guard, float extent) {
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
    if (guard) {
        // NaN fails every comparison, so a non-finite value also raises the flag
        const bool ok = fabsf(x.x) <= extent && fabsf(x.y) <= extent && fabsf(x.z) <= extent &&
                        fabsf(v.x) <= FLT_MAX && fabsf(v.y) <= FLT_MAX && fabsf(v.z) <= FLT_MAX;
        if (!ok) atomicOr(&guard[sys_of ? sys_of[i] : 0u], 1u);
    }
}

void launch_update(float4* pos, float4* vel, const float4* force, const uint32_t* type,
                   uint32_t n, float dt, const astro_params& p, hipStream_t s,
                   const uint32_t* sys_of, uint32_t* guard, float extent) {
    if (!n) return;
    hipLaunchKernelGGL(k_update, dim3((n + 255) / 256), dim3(256), 0, s, pos, vel, force, type, n,
                       dt, p, sys_of, guard, extent > 0.0f ? extent : FLT_MAX);
}

// FP64 diagnostics: per (system, chunk) sums of {m, kinetic, potential (1/2 m phi), m v,
// m x}; partial[(s * chunks + chunk) * 9 + k].  Summed on the host in a fixed order.
__global__ __launch_bounds__(256) void k_diag(const float4* __restrict__ pos,
                                              const float4* __restrict__ vel,
                                              const double* __restrict__ phi,
                                              const uint32_t* __restrict__ sys_off,
                                              uint32_t chunks, double* __restrict__ partial) {
    __shared__ double s_red[9][4];
    const uint32_t sys = blockIdx.y, chunk = blockIdx.x;
    const uint32_t b = sys_off[sys], e = sys_off[sys + 1];
    const uint32_t len = e - b, per = (len + chunks - 1) / chunks;
    const uint32_t lo = b + min(len, chunk * per), hi = b + min(len, (chunk + 1) * per);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) {
        const float4 x = pos[i], v = vel[i];
        const double m = x.w;
        acc[0] += m;
        acc[1] += 0.5 * m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
        acc[2] += phi ? 0.5 * m * phi[i] : 0.0;
        acc[3] += m * v.x;
        acc[4] += m * v.y;
        acc[5] += m * v.z;
        acc[6] += m * x.x;
        acc[7] += m * x.y;
        acc[8] += m * x.z;
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 9; k++) {
        double v = acc[k];
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if (lane == 0) s_red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        const int k = threadIdx.x;
        partial[((size_t)sys * chunks + chunk) * 9 + k] =
            ((s_red[k][0] + s_red[k][1]) + s_red[k][2]) + s_red[k][3];
    }
}

void launch_diag(const float4* pos, const float4* vel, const double* phi,
                 const uint32_t* d_sys_off, uint32_t nsys, uint32_t chunks, double* partial,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_diag, dim3(chunks, nsys), dim3(256), 0, s, pos, vel, phi, d_sys_off,
                       chunks, partial);
}

// BarnesHut.metal:103-167 (lightingPass): planets (type 1) and dust (type 2) take the light of
// the first min(nstars, 5) particles of their system; stars keep their colours.  Same order of
// operations as the oracle.
__global__ __launch_bounds__(256) void k_lighting(const float4* __restrict__ pos,
                                                  float4* __restrict__ color,
                                                  const uint32_t* __restrict__ type, uint32_t n,
                                                  uint32_t nstars, const uint32_t* __restrict__ sys_of,
                                                  const uint32_t* __restrict__ sys_off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = type[i];
    if (t != 1u && t != 2u) return;
    uint32_t b = 0, e = n;
    if (sys_of) {
        const uint32_t s = sys_of[i];
        b = sys_off[s];
        e = sys_off[s + 1];
    }
    const float4 p = pos[i];
    const uint32_t kmax = nstars < 5u ? nstars : 5u;
    float r = 0.0f, g = 0.0f, bl = 0.0f;
    for (uint32_t k = 0; k < kmax && b + k < e; k++) {
        const float4 sp = pos[b + k];
        const float4 sc = color[b + k];
        const float dx = sp.x - p.x, dy = sp.y - p.y, dz = sp.z - p.z;
        const float dist = sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)));
        const float q = fmaxf(dist / ASTRO_LIGHT_ATTENUATION_DISTANCE, 1e-9f);
        const float att = 1.0f / (q * q);
        r = r + sc.x * att;
        g = g + sc.y * att;
        bl = bl + sc.z * att;
    }
    const float a = t == 2u ? fminf(fmaxf(r, fmaxf(g, bl)), 0.25f) : 1.0f;
    color[i] = make_float4(r, g, bl, a);
}
void launch_lighting(const float4* pos, float4* color, const uint32_t* type, uint32_t n,
                     uint32_t nstars, const uint32_t* sys_of, const uint32_t* sys_off,
                     hipStream_t s) {
    if (!n || !nstars) return;
    hipLaunchKernelGGL(k_lighting, dim3((n + 255) / 256), dim3(256), 0, s, pos, color, type, n,
                       nstars, sys_of, sys_off);
}

// FP64 direct sum for accuracy audits: a_i = sum_j G m_j d / (|d|^2 + eps2)^(3/2), d = x_j - x_i,
// over the sources of target i's system (the self term has d = 0 and adds nothing).  Block
// (chunk, t) sums one chunk of the sources for target t; partial[(t * chunks + chunk) * 3 + k].
__global__ __launch_bounds__(256) void k_direct_fp64(const float4* __restrict__ pos,
                                                     const uint32_t* __restrict__ sys_of,
                                                     const uint32_t* __restrict__ sys_off,
                                                     uint32_t nt, const uint32_t* __restrict__ targets,
                                                     uint32_t chunks, double G, double eps2,
                                                     double* __restrict__ partial) {
    __shared__ double s_red[3][4];
    const uint32_t t = blockIdx.y, chunk = blockIdx.x;
    const uint32_t i = targets[t];
    uint32_t b = 0, e = nt;
    if (sys_of) {
        const uint32_t sys = sys_of[i];
        b = sys_off[sys];
        e = sys_off[sys + 1];
    }
    const uint32_t len = e - b, per = (len + chunks - 1) / chunks;
    const uint32_t lo = b + min(len, chunk * per), hi = b + min(len, (chunk + 1) * per);
    const float4 p = pos[i];
    const double xi = p.x, yi = p.y, zi = p.z;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (uint32_t j = lo + threadIdx.x; j < hi; j += 256) {
        const float4 q = pos[j];
        const double dx = (double)q.x - xi, dy = (double)q.y - yi, dz = (double)q.z - zi;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double sc = G * (double)q.w / (r2 * sqrt(r2));
        ax += sc * dx;
        ay += sc * dy;
        az += sc * dz;
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double v[3] = {ax, ay, az};
    for (int k = 0; k < 3; k++) {
        double x = v[k];
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        if (lane == 0) s_red[k][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        partial[((size_t)t * chunks + chunk) * 3 + k] =
            ((s_red[k][0] + s_red[k][1]) + s_red[k][2]) + s_red[k][3];
    }
}

void launch_direct_fp64(const float4* pos, const uint32_t* sys_of, const uint32_t* d_sys_off,
                        uint32_t nt, const uint32_t* targets, uint32_t n_targets, uint32_t chunks,
                        double G, double eps2, double* partial, hipStream_t s) {
    if (!n_targets) return;
    hipLaunchKernelGGL(k_direct_fp64, dim3(chunks, n_targets), dim3(256), 0, s, pos, sys_of,
                       d_sys_off, nt, targets, chunks, G, eps2, partial);
}

}  // namespace astro
