// k_resort.hip -- the step's key sort seeded by the previous step's order (gfx950).
//
// Between two steps the particles move a little, so the order the last sort produced is nearly
// the new one.  This sort gives the radix sort's result -- ascending (key, source index), the
// stable argsort of the keys -- in three launches instead of three per 8-bit digit, and computes
// the keys itself (the key kernel does not run):
//  * k_resort_sample (one block): evenly spaced entries of the previous order give samples
//    (key, index), sorted in one block; every `over`-th becomes a splitter, so that the buckets
//    between consecutive splitters hold kResortMean sources or fewer on average;
//  * k_resort_place: every source, visited in the previous order, gets its key and its bucket.
//    Bucket g's first slots are the positions that held its sources last time; a source still
//    in the bucket of its position is stored at that position's slot, with no atomic.  One that
//    moved leaves a hole there, counts itself out of the old bucket and into the new one, and
//    takes a slot past the new bucket's positions (past its kResortCap slots, in an overflow
//    list);
//  * k_resort_bucket (one block per bucket): the bucket's offset follows from the positions and
//    the counts in and out before it; its slots are sorted by (key, index), holes last, and its
//    sources written to their final positions (with their positions and types, read once here
//    for the build and the traversal), and as the next sort's previous order.  A bucket
//    beyond kResortCap gathers its sources at their final positions and sorts them there, in
//    device memory (slow, but only for orders far from the previous one).
// The block sorts are bitonic networks over rows held two per thread: steps whose partner rows
// lie in the same wave exchange them with lane shuffles, the others (10 of the 66 steps of 2,048
// rows) through LDS with one barrier.  Splitters and comparisons use (key, index), a total order
// of the sources: equal keys do not pile up in one bucket, and the result does not depend on the
// order in which the atomics handed out slots.  Any permutation as previous order gives the same
// result; a close one is fast.  The per-source key arrays the key kernel would have written are
// rebuilt from the sorted keys by k_resort_keys when a caller reads them.
#include "launch.hpp"
#include "morton.hpp"

namespace astro {

__device__ __forceinline__ bool rs_less(uint64_t ka, uint32_t ia, uint64_t kb, uint32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// The bitonic network over rows 0..m-1 padded to P = 2^ceil(log2 m), in the form whose every
// comparator (i, j), i < j, puts the smaller row at i (the first step of each merge pairs
// mirrored rows, i with i ^ (k - 1)).  Rows m..P-1 then act as +infinity and never move, so the
// comparators that reach them are skipped.  cx(i, j) compares and exchanges; bar() separates
// the steps.  m is the same for the whole block.
template <uint32_t NT, class CX, class BAR>
__device__ inline void bitonic_rows(uint32_t m, CX cx, BAR bar) {
    uint32_t P = 1;
    while (P < m) P <<= 1;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t d = k >> 1; d >= 1; d >>= 1) {
            for (uint32_t q = threadIdx.x; q < (P >> 1); q += NT) {
                const uint32_t i = ((q & ~(d - 1)) << 1) | (q & (d - 1));  // bit log2(d) clear
                const uint32_t j = d == (k >> 1) ? i ^ (k - 1) : i | d;
                if (j < m) cx(i, j);
            }
            bar();
        }
    }
}

// One row of a block sort: the key and the source index.  kRowMax sorts after every source
// (keys hold at most 63 bits).
struct SortRow {
    uint64_t k;
    uint32_t i;
};
__device__ __forceinline__ bool row_less(const SortRow& a, const SortRow& b) {
    return rs_less(a.k, a.i, b.k, b.i);
}
__device__ __forceinline__ SortRow row_max() { return SortRow{~0ull, ~0u}; }
// The row of the lane whose byte address (lane * 4) is addr.
__device__ __forceinline__ SortRow bperm_row(uint32_t addr, const SortRow& v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute((int)addr, (int)(uint32_t)v.k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute((int)addr, (int)(uint32_t)(v.k >> 32));
    const uint32_t ix = (uint32_t)__builtin_amdgcn_ds_bpermute((int)addr, (int)v.i);
    return SortRow{((uint64_t)hi << 32) | lo, ix};
}
// r = take ? o : r, as selects.
__device__ __forceinline__ void row_take(SortRow& r, const SortRow& o, bool take) {
    r.k = take ? o.k : r.k;
    r.i = take ? o.i : r.i;
}

// The same network over rows 0..P-1 (P a power of two, at most 2 NT) held two per thread:
// thread t holds rows 2t (r[0]) and 2t + 1 (r[1]).  Step (k, d) pairs row j with j ^ mask (mask
// k - 1 for the first step of a merge, else d), the lower row keeping the smaller: mask 1 pairs
// a thread's own rows; otherwise the partner thread is t ^ (mask >> 1), reached by lane shuffles
// inside the wave, or through LDS rows xk/xi (4 NT each: two alternating halves, so one barrier
// per step) beyond it.  Waves whose rows all lie at or past P sit the steps out.
// PFIX: a fixed P (0: the argument), whose steps unroll with constant masks, shuffle addresses
// and slot pairings.
template <uint32_t NT, uint32_t PFIX = 0>
__device__ __forceinline__ void bitonic_pairs(SortRow r[2], uint32_t P_arg,
                                              uint64_t* __restrict__ xk,
                                              uint32_t* __restrict__ xi) {
    const uint32_t P = PFIX ? PFIX : P_arg;
    const uint32_t t = threadIdx.x;
    const bool live = 2u * (t & ~63u) < P;  // wave-uniform
    uint32_t half = 0;
#pragma unroll
    for (uint32_t k = 2; k <= P; k <<= 1) {
#pragma unroll
        for (uint32_t d = k >> 1; d >= 1; d >>= 1) {
            const uint32_t mask = d == (k >> 1) ? k - 1 : d;
            const uint32_t tm = mask >> 1;
            if (tm == 0) {  // rows 2t and 2t + 1
                if (live) {
                    const SortRow r0 = r[0], r1 = r[1];
                    const bool sw = row_less(r1, r0);
                    row_take(r[0], r1, sw);
                    row_take(r[1], r0, sw);
                }
                continue;
            }
            SortRow q0, q1;  // the partner thread's rows
            if (tm >= 64) {
                uint64_t* hk = xk + half * 2 * NT;
                uint32_t* hi = xi + half * 2 * NT;
                if (live) {
                    hk[2 * t] = r[0].k;
                    hk[2 * t + 1] = r[1].k;
                    hi[2 * t] = r[0].i;
                    hi[2 * t + 1] = r[1].i;
                }
                lds_barrier();
                const uint32_t pt = t ^ tm;
                if (live) {
                    q0 = SortRow{hk[2 * pt], hi[2 * pt]};
                    q1 = SortRow{hk[2 * pt + 1], hi[2 * pt + 1]};
                }
                half ^= 1u;
            } else if (live) {
                const uint32_t addr = ((t & 63u) ^ tm) << 2;
                q0 = bperm_row(addr, r[0]);
                q1 = bperm_row(addr, r[1]);
            }
            if (live) {
                // an odd mask pairs row slot s with the partner's slot 1 - s; the lower row of a
                // pair (bit d clear) keeps the partner's row when that is smaller, the upper one
                // when it is not
                const bool swap_slots = mask & 1u;
                const SortRow o0 = swap_slots ? q1 : q0, o1 = swap_slots ? q0 : q1;
                const bool lower = ((2u * t) & d) == 0u;
                row_take(r[0], o0, row_less(o0, r[0]) == lower);
                row_take(r[1], o1, row_less(o1, r[1]) == lower);
            }
        }
    }
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t m) {
    uint32_t P = 1;
    while (P < m) P <<= 1;
    return P;
}

// ------------------------------------------------------------------------------ bucket map
// One system's positions [off, off + n) and buckets [b0, b0 + nb): bucket b0 + c starts at
// position off + c n / nb (rounded down).  One system: the whole sort.
struct SysBuckets {
    uint32_t s, off, n, b0, nb;
};
__device__ __forceinline__ SysBuckets sys_buckets(const ResortArgs& a, uint32_t s, uint32_t nb) {
    if (a.nsys <= 1) return SysBuckets{0u, 0u, a.n, 0u, nb};
    const uint32_t off = a.sys_off[s], b0 = a.sys_b0[s];
    return SysBuckets{s, off, a.sys_off[s + 1] - off, b0, a.sys_b0[s + 1] - b0};
}
// Position where the system's bucket c (0 .. nb) starts.
__device__ __forceinline__ uint32_t bucket_pos(const SysBuckets& m, uint32_t c) {
    return m.off + (uint32_t)((uint64_t)c * m.n / m.nb);
}
// The system's bucket holding position q (the largest c with bucket_pos(c) <= q).
__device__ __forceinline__ uint32_t bucket_of_pos(const SysBuckets& m, uint32_t q) {
    return (uint32_t)(((uint64_t)(q - m.off + 1) * m.nb + m.n - 1) / m.n - 1);
}

// --------------------------------------------------------------------------------- samples
__device__ __forceinline__ SortRow sample_row(const ResortArgs& a, const SysBuckets& m,
                                              uint32_t j, uint32_t ns) {
    const uint32_t p = a.order[m.off + (uint64_t)j * m.n / ns];
    return SortRow{morton_key(a.pos[p], a.cubes, m.s, a.params), p};
}

// The last sort's splitters still divide this sort's keys evenly: the key grid has the same cell
// size and moved by less than kResortShift cells along each axis (or does not depend on the
// cube).
constexpr float kResortShift = 1024.0f;
__device__ inline bool splitters_current(const ResortArgs& a, uint32_t s) {
    if (a.params.quantizer == ASTRO_QUANT_OFFSET_WRAP) return true;
    const float4 c = a.cubes[s], o = a.spl_cube[s];
    int e, eo;
    frexpf(c.w, &e);
    frexpf(o.w, &eo);
    if (e != eo) return false;
    const float scale = ldexpf(1.0f, 21 - (e - 1));
    return fabsf(c.x - o.x) * scale < kResortShift && fabsf(c.y - o.y) * scale < kResortShift &&
           fabsf(c.z - o.z) * scale < kResortShift;
}

// One block per system: zeroes its buckets' counts and, when its key grid changed, samples its
// previous order (over samples per bucket) and sorts the samples for its splitters.
__global__ __launch_bounds__(kResortThreads) void k_resort_sample(ResortArgs a, uint32_t nb) {
    __shared__ uint64_t s_k[kResortMaxSamples];
    __shared__ uint32_t s_i[kResortMaxSamples];
    const uint32_t tid = threadIdx.x;
    const SysBuckets m = sys_buckets(a, blockIdx.x, nb);
    // the counts and the overflow list of this sort start empty
    for (uint32_t b = tid; b < m.nb; b += kResortThreads) {
        a.moved_in[m.b0 + b] = 0;
        a.moved_out[m.b0 + b] = 0;
    }
    if (blockIdx.x == 0 && tid == 0) *a.ovf = 0;
    if (splitters_current(a, m.s)) return;  // the last sort's splitters (block-uniform)
    if (tid == 0) a.spl_cube[m.s] = a.cubes[m.s];
    uint32_t over = kResortMaxSamples / m.nb;  // samples per bucket
    if (over > kResortOversample) over = kResortOversample;
    const uint32_t ns = m.nb * over;
    uint64_t* spl_k = a.spl_k + m.b0;  // the system's splitters: nb - 1 of them
    uint32_t* spl_i = a.spl_i + m.b0;
    if (ns <= 2 * kResortThreads) {
        SortRow r[2];
        for (uint32_t s = 0; s < 2; s++) {
            const uint32_t j = 2 * tid + s;
            r[s] = j < ns ? sample_row(a, m, j, ns) : row_max();
        }
        bitonic_pairs<kResortThreads>(r, pow2_at_least(ns), s_k, s_i);
        // splitter t is sample (t + 1) * over
        for (uint32_t s = 0; s < 2; s++) {
            const uint32_t j = 2 * tid + s;
            if (j < ns && j > 0 && j % over == 0) {
                spl_k[j / over - 1] = r[s].k;
                spl_i[j / over - 1] = r[s].i;
            }
        }
        return;
    }
    for (uint32_t j = tid; j < ns; j += kResortThreads) {
        const SortRow v = sample_row(a, m, j, ns);
        s_k[j] = v.k;
        s_i[j] = v.i;
    }
    lds_barrier();
    bitonic_rows<kResortThreads>(
        ns,
        [](uint32_t i, uint32_t j) {
            const uint64_t ki = s_k[i], kj = s_k[j];
            const uint32_t ii = s_i[i], ij = s_i[j];
            if (rs_less(kj, ij, ki, ii)) {
                s_k[i] = kj;
                s_k[j] = ki;
                s_i[i] = ij;
                s_i[j] = ii;
            }
        },
        [] { lds_barrier(); });
    for (uint32_t t = tid; t + 1 < m.nb; t += kResortThreads) {
        spl_k[t] = s_k[(t + 1) * over];
        spl_i[t] = s_i[(t + 1) * over];
    }
}

// ----------------------------------------------------------------------------------- place
// Splitter t (of the system's, spl_k / spl_i) lies at or before source (key, p).
__device__ __forceinline__ bool splitter_le(const uint64_t* spl_k, const uint32_t* spl_i, int t,
                                            uint64_t key, uint32_t p) {
    const uint64_t sk = spl_k[t];
    return sk < key || (sk == key && spl_i[t] <= p);
}

// The system's bucket of (key, p): the number of its splitters (nb - 1, ascending) at or
// before it.  A galloping search from the guess g: two loads when the guess is right.
__device__ inline uint32_t find_bucket(const uint64_t* spl_k, const uint32_t* spl_i,
                                       uint64_t key, uint32_t p, uint32_t g, uint32_t nb) {
    const int ns = (int)nb - 1, gg = (int)g;
    int lo, hi;  // splitter lo at or before (p), or lo = -1; splitter hi after, or hi = ns
    if (gg > 0 && !splitter_le(spl_k, spl_i, gg - 1, key, p)) {
        hi = gg - 1;
        int step = 1;
        lo = hi - step;
        while (lo >= 0 && !splitter_le(spl_k, spl_i, lo, key, p)) {
            hi = lo;
            step <<= 1;
            lo = hi - step;
        }
        if (lo < -1) lo = -1;
    } else if (gg < ns && splitter_le(spl_k, spl_i, gg, key, p)) {
        lo = gg;
        int step = 1;
        hi = lo + step;
        while (hi < ns && splitter_le(spl_k, spl_i, hi, key, p)) {
            lo = hi;
            step <<= 1;
            hi = lo + step;
        }
        if (hi > ns) hi = ns;
    } else {
        return g;
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (splitter_le(spl_k, spl_i, mid, key, p)) lo = mid;
        else hi = mid;
    }
    return (uint32_t)hi;
}

// A source whose bucket is the one its previous position lies in (a "stayer") is stored at that
// position's slot: bucket g's slots 0 .. S_g - 1 are its positions.  A source that changed
// bucket leaves a hole (a row that sorts last) at its position's slot, counts itself out of
// the old bucket and takes a slot past the positions of the new one with one atomic.  Sources
// never change system: a batch's source looks for its bucket among its own system's.
__global__ __launch_bounds__(kResortPlaceThreads) void k_resort_place(ResortArgs a, uint32_t nb) {
    const uint32_t i = blockIdx.x * kResortPlaceThreads + threadIdx.x;
    if (i >= a.n) return;
    // the guess: the bucket whose positions hold i
    const uint32_t sy = a.nsys > 1 ? a.sys_of[i] : 0u;
    const SysBuckets m = sys_buckets(a, sy, nb);
    const uint32_t c = bucket_of_pos(m, i);
    const uint32_t g = m.b0 + c;
    const uint32_t qs = bucket_pos(m, c);
    const uint64_t* spl_k = a.spl_k + m.b0;
    const uint32_t* spl_i = a.spl_i + m.b0;
    // the splitters around the guessed bucket load beside the source's own chain
    SortRow s0 = SortRow{0ull, 0u}, s1 = row_max();
    if (c > 0) s0 = SortRow{spl_k[c - 1], spl_i[c - 1]};
    if (c + 1 < m.nb) s1 = SortRow{spl_k[c], spl_i[c]};
    const uint32_t p = a.order[i];
    const float4 xp = a.pos[p];
    const uint32_t ty = a.type[p];
    const uint64_t key = morton_key(xp, a.cubes, sy, a.params);
    const SortRow x{key, p};
    const bool at_or_after = c == 0 || !row_less(x, s0);
    const bool before = c + 1 >= m.nb || row_less(x, s1);
    const uint64_t stay = (uint64_t)g * kResortCap + (i - qs);
    if (at_or_after && before) {
        a.rk[stay] = key;
        a.ri[stay] = p;
        a.rp[stay] = xp;
        a.rt[stay] = ty;
        return;
    }
    const uint32_t bc = find_bucket(spl_k, spl_i, key, p, c, m.nb);
    const uint32_t b = m.b0 + bc;
    a.rk[stay] = ~0ull;  // a hole
    a.ri[stay] = ~0u;
    atomicAdd(&a.moved_out[g], 1u);
    const uint32_t slot =
        bucket_pos(m, bc + 1) - bucket_pos(m, bc) + atomicAdd(&a.moved_in[b], 1u);
    if (slot < kResortCap) {
        const uint64_t q = (uint64_t)b * kResortCap + slot;
        a.rk[q] = key;
        a.ri[q] = p;
        a.rp[q] = xp;
        a.rt[q] = ty;
    } else {
        const uint32_t o = atomicAdd(a.ovf, 1u);
        a.ok[o] = key;
        a.oi[o] = p;
        a.ob[o] = b;
    }
}

// ---------------------------------------------------------------------------------- bucket
// The next sort's splitter c of a system (c = 1 .. nb - 1, stored at b0 + c - 1) is the row at
// its bucket c's first position; row position q writes the one it is, if any.
__device__ __forceinline__ void emit_splitter(const ResortArgs& a, const SysBuckets& m, uint32_t q,
                                              uint64_t k, uint32_t i) {
    const uint32_t x = q - m.off;
    const uint64_t c = ((uint64_t)x * m.nb + m.n - 1) / m.n;  // the smallest c reaching x
    if (c >= 1 && c < m.nb && c * m.n / m.nb == x) {
        a.spl_k[m.b0 + c - 1] = k;
        a.spl_i[m.b0 + c - 1] = i;
    }
}

// The smallest c >= 1 whose bucket starts at or past position q0.
__device__ __forceinline__ uint32_t first_splitter(const SysBuckets& m, uint32_t q0) {
    const uint32_t c = (uint32_t)(((uint64_t)(q0 - m.off) * m.nb + m.n - 1) / m.n);
    return c ? c : 1u;
}

__global__ __launch_bounds__(kResortBucketThreads) void k_resort_bucket(ResortArgs a, uint32_t nb) {
    constexpr uint32_t NT = kResortBucketThreads;
    __shared__ uint64_t s_k[kResortCap];  // the bucket's rows (LDS sort), or exchange rows
    __shared__ uint32_t s_i[kResortCap];
    __shared__ uint32_t s_tmp[2 * (NT / 64 + 1)];
    __shared__ uint32_t s_fill;
    static_assert(kResortCap >= 4 * NT && kResortRegRows == 2 * NT, "row layout");
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const SysBuckets sb = sys_buckets(a, a.nsys > 1 ? a.bkt_sys[b] : 0u, nb);
    const uint32_t c = b - sb.b0;  // the bucket within its system
    // the bucket's positions, sources that left and sources that came: its rows are its
    // positions' slots (stayers and holes) and then those that came.  Sources stay in their
    // system, so the system's buckets before this one place it.
    const uint32_t q_b = bucket_pos(sb, c), npos = bucket_pos(sb, c + 1) - q_b;
    const uint32_t came = a.moved_in[b], left = a.moved_out[b];
    uint32_t part_in = 0, part_out = 0;
    for (uint32_t t = sb.b0 + tid; t < b; t += NT) {
        part_in += a.moved_in[t];
        part_out += a.moved_out[t];
    }
    uint32_t in_before, out_before;
    block_sum_and_scan<NT>(part_out, part_in, s_tmp, &out_before, &in_before);
    const uint32_t start = __builtin_amdgcn_readfirstlane(q_b + in_before - out_before);
    const uint32_t m = npos + came - left;  // the bucket's sources
    if (c == 0 && tid == 0) a.spl_cube[sb.s] = a.cubes[sb.s];  // the next sort's splitters' cube
    // never, on counts of a pass
    if (left > npos || start < sb.off || start > sb.off + sb.n || m > sb.off + sb.n - start) return;
    const uint64_t slots = (uint64_t)b * kResortCap;
    const uint32_t R = npos + came;  // rows in the slots when none overflowed
    if (R <= kResortRegRows) {  // holes sort after every source
        // with n <= 2^22 a row's index word is index << 10 | slot: the same order (indices are
        // distinct), and the row's position and type are read from its slot afterwards
        static_assert(kResortRegRows <= 1024u, "slot in 10 bits");
        const bool packed = a.n <= kResortPackMax;
        SortRow r[2];
        for (uint32_t s = 0; s < 2; s++) {
            const uint32_t j = 2 * tid + s;
            if (j < R) {
                const uint32_t i = a.ri[slots + j];
                r[s] = SortRow{a.rk[slots + j], packed && i != ~0u ? (i << 10) | j : i};
            } else {
                r[s] = row_max();
            }
        }
        if (R > kResortRegRows / 2) bitonic_pairs<NT, kResortRegRows>(r, 0u, s_k, s_i);
        else bitonic_pairs<NT>(r, pow2_at_least(R), s_k, s_i);
        for (uint32_t s = 0; s < 2; s++) {
            const uint32_t j = 2 * tid + s;
            if (j < m) {
                const uint32_t i = packed ? r[s].i >> 10 : r[s].i;
                a.out_k[start + j] = r[s].k;
                a.out_i[start + j] = i;
                a.order[start + j] = i;
                if (packed) {
                    const uint32_t slot = r[s].i & 1023u;
                    a.spos[start + j] = a.rp[slots + slot];
                    a.stype[start + j] = a.rt[slots + slot];
                } else {
                    a.spos[start + j] = a.pos[i];
                    a.stype[start + j] = a.type[i];
                }
                r[s].i = i;  // the splitters carry the plain index
            }
        }
        // the next sort's splitters among these rows (a few per bucket, found once per block)
        for (uint32_t cs = first_splitter(sb, start); cs < sb.nb; cs++) {
            const uint32_t q = bucket_pos(sb, cs) - start;
            if (q >= m) break;
            if ((q >> 1) == tid) {
                const SortRow v = (q & 1u) ? r[1] : r[0];
                a.spl_k[sb.b0 + cs - 1] = v.k;
                a.spl_i[sb.b0 + cs - 1] = v.i;
            }
        }
        return;
    }
    if (R <= kResortCap) {
        for (uint32_t j = tid; j < R; j += NT) {
            s_k[j] = a.rk[slots + j];
            s_i[j] = a.ri[slots + j];
        }
        lds_barrier();
        bitonic_rows<NT>(
            R,
            [](uint32_t i, uint32_t j) {
                const uint64_t ki = s_k[i], kj = s_k[j];
                const uint32_t ii = s_i[i], ij = s_i[j];
                if (rs_less(kj, ij, ki, ii)) {
                    s_k[i] = kj;
                    s_k[j] = ki;
                    s_i[i] = ij;
                    s_i[j] = ii;
                }
            },
            [] { lds_barrier(); });
        for (uint32_t j = tid; j < m; j += NT) {
            a.out_k[start + j] = s_k[j];
            a.out_i[start + j] = s_i[j];
            a.order[start + j] = s_i[j];
            a.spos[start + j] = a.pos[s_i[j]];
            a.stype[start + j] = a.type[s_i[j]];
            emit_splitter(a, sb, start + j, s_k[j], s_i[j]);
        }
        return;
    }
    // past the slots: the bucket's sources (its slots but the holes, and its overflow entries)
    // are gathered at their final positions and sorted there (__syncthreads orders the block's
    // device-memory accesses between the steps)
    uint64_t* gk = a.out_k + start;
    uint32_t* gi = a.out_i + start;
    if (tid == 0) s_fill = 0;
    __syncthreads();
    for (uint32_t j = tid; j < kResortCap; j += NT) {
        const uint32_t i = a.ri[slots + j];
        if (i == ~0u) continue;
        const uint32_t d = atomicAdd(&s_fill, 1u);
        if (d < m) {
            gk[d] = a.rk[slots + j];
            gi[d] = i;
        }
    }
    const uint32_t no = *a.ovf;
    for (uint32_t o = tid; o < no; o += NT) {
        if (a.ob[o] != b) continue;
        const uint32_t d = atomicAdd(&s_fill, 1u);
        if (d < m) {  // slots and entries of this bucket hold exactly m sources
            gk[d] = a.ok[o];
            gi[d] = a.oi[o];
        }
    }
    __syncthreads();
    bitonic_rows<NT>(
        m,
        [gk, gi](uint32_t i, uint32_t j) {
            const uint64_t ki = gk[i], kj = gk[j];
            const uint32_t ii = gi[i], ij = gi[j];
            if (rs_less(kj, ij, ki, ii)) {
                gk[i] = kj;
                gk[j] = ki;
                gi[i] = ij;
                gi[j] = ii;
            }
        },
        [] { __syncthreads(); });
    for (uint32_t j = tid; j < m; j += NT) {
        a.order[start + j] = gi[j];
        a.spos[start + j] = a.pos[gi[j]];
        a.stype[start + j] = a.type[gi[j]];
        emit_splitter(a, sb, start + j, gk[j], gi[j]);
    }
}

// After a radix sort: its order as the previous order, its rows at the splitter positions as
// the splitters, and this build's cube as theirs.
__global__ void k_resort_seed(const uint64_t* __restrict__ sk, const uint32_t* __restrict__ si,
                              ResortArgs a, uint32_t nb) {
    const uint32_t ns = a.nsys > 1 ? a.nsys : 1u;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x)
        a.spl_cube[s] = a.cubes[s];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < a.n; q += gridDim.x * blockDim.x) {
        const uint32_t i = si[q];
        a.order[q] = i;
        a.spos[q] = a.pos[i];
        a.stype[q] = a.type[i];
        emit_splitter(a, sys_buckets(a, a.nsys > 1 ? a.sys_of[q] : 0u, nb), q, sk[q], i);
    }
}

__global__ void k_sorted_gather(const uint32_t* __restrict__ si, const float4* __restrict__ pos,
                                const uint32_t* __restrict__ type, uint32_t n,
                                float4* __restrict__ spos, uint32_t* __restrict__ stype) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const uint32_t i = si[q];
        spos[q] = pos[i];
        stype[q] = type[i];
    }
}
void launch_sorted_gather(const uint32_t* si, const float4* pos, const uint32_t* type, uint32_t n,
                          float4* spos, uint32_t* stype, hipStream_t s) {
    if (!n) return;
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_sorted_gather, dim3(blocks), dim3(256), 0, s, si, pos, type, n, spos,
                       stype);
}
void launch_resort_seed(const uint64_t* sk, const uint32_t* si, const ResortArgs& a,
                        hipStream_t s) {
    const uint32_t nb = a.nb;
    if (!a.n || !nb) return;
    uint32_t blocks = (a.n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_resort_seed, dim3(blocks), dim3(256), 0, s, sk, si, a, nb);
}

// The per-source key arrays of a seeded sort, from its sorted keys and indices: what the key
// kernel writes (codes[p] = codes2[p] = key of p, idx[i] = i).
__global__ void k_resort_keys(const uint64_t* __restrict__ sk, const uint32_t* __restrict__ si,
                              uint32_t n, uint64_t* __restrict__ codes,
                              uint64_t* __restrict__ codes2, uint32_t* __restrict__ idx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = si[i];
        const uint64_t k = sk[i];
        codes[p] = k;
        codes2[p] = k;
        idx[i] = i;
    }
}
void launch_resort_keys(const uint64_t* sk, const uint32_t* si, uint32_t n, uint64_t* codes,
                        uint64_t* codes2, uint32_t* idx, hipStream_t s) {
    if (!n) return;
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_resort_keys, dim3(blocks), dim3(256), 0, s, sk, si, n, codes, codes2, idx);
}

uint32_t resort_buckets(uint32_t n) {
    const uint32_t nb = (uint32_t)(((uint64_t)n + kResortMean - 1) / kResortMean);
    return nb <= kResortMaxBuckets ? nb : 0u;
}

uint32_t resort_buckets_batch(const std::vector<uint32_t>& off, std::vector<uint32_t>* b0,
                              std::vector<uint32_t>* bkt_sys) {
    const size_t ns = off.empty() ? 0 : off.size() - 1;
    b0->assign(ns + 1, 0u);
    bkt_sys->clear();
    uint64_t nb = 0;
    for (size_t s = 0; s < ns; s++) {
        (*b0)[s] = (uint32_t)nb;
        const uint64_t k = ((uint64_t)(off[s + 1] - off[s]) + kResortMean - 1) / kResortMean;
        nb += k;
        if (nb > kResortMaxBuckets) return 0u;
        bkt_sys->insert(bkt_sys->end(), (size_t)k, (uint32_t)s);
    }
    (*b0)[ns] = (uint32_t)nb;
    return (uint32_t)nb;
}

void launch_resort(const ResortArgs& a, hipStream_t s) {
    const uint32_t nb = a.nb;
    if (!a.n || !nb) return;
    hipLaunchKernelGGL(k_resort_sample, dim3(a.nsys > 1 ? a.nsys : 1u), dim3(kResortThreads), 0,
                       s, a, nb);
    hipLaunchKernelGGL(k_resort_place, dim3((a.n + kResortPlaceThreads - 1) / kResortPlaceThreads),
                       dim3(kResortPlaceThreads), 0, s, a, nb);
    hipLaunchKernelGGL(k_resort_bucket, dim3(nb), dim3(kResortBucketThreads), 0, s, a, nb);
}

}  // namespace astro
