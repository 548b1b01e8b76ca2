// k_sort.hip -- scans and the LSD radix sort of (u64 key, u32 index) pairs for gfx950.
//
// The reference sorts with 64 one-bit split passes, each a prep + 1-3 scan dispatches + a
// scatter (RadixSort.swift:98-185, RadixSort.metal:30-131): ~3.8 KB of traffic per key.  Here
// a sort is one pass per 8-bit digit, each three launches: per-tile digit counts, a per-digit
// scan over the tiles, and a rank-and-scatter kernel that ranks keys inside a wavefront with
// ballots (no atomics), turns per-wave counts into block offsets in LDS, stages the tile in LDS
// in digit order and writes it out coalesced.  No block ever waits on another.  Every pass is
// stable, so the result is the stable ascending argsort -- the split sort's permutation.
#include "launch.hpp"

namespace astro {

// ---------------------------------------------------------------------------- small kernels
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static uint32_t fill_grid(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return b ? (uint32_t)b : 1u;
}

void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
}
void launch_fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
}

// Octree.metal:39-51 (fillIndices), guarded at the upper edge.
__global__ void k_iota(uint32_t* buf, uint32_t start, uint32_t end) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t idx = start + g;
    if (idx < end) buf[g] = idx;
}
void launch_iota(uint32_t* buf, uint32_t start, uint32_t end, hipStream_t s) {
    if (end <= start) return;
    uint32_t n = end - start;
    hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, buf, start, end);
}

// RadixSort.metal:135-153 (copyBuffer / copyIndices): copy *d_count elements.
template <typename T>
__global__ void k_copy_count(const T* in, T* out, const uint32_t* d_count) {
    uint32_t n = *d_count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[i];
}
void launch_copy_u64_count(const uint64_t* in, uint64_t* out, const uint32_t* d_count,
                           uint32_t cap, hipStream_t s) {
    if (cap)
        hipLaunchKernelGGL(k_copy_count<uint64_t>, dim3(fill_grid(cap)), dim3(256), 0, s, in,
                           out, d_count);
}
void launch_copy_u32_count(const uint32_t* in, uint32_t* out, const uint32_t* d_count,
                           uint32_t cap, hipStream_t s) {
    if (cap)
        hipLaunchKernelGGL(k_copy_count<uint32_t>, dim3(fill_grid(cap)), dim3(256), 0, s, in,
                           out, d_count);
}

__global__ void k_copy_u32(const uint32_t* in, uint32_t* out, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[i];
}
void launch_copy_u32(const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_copy_u32, dim3(fill_grid(n)), dim3(256), 0, s, in, out, n);
}

// ------------------------------------------------------------------------- exclusive scan
// Two launches over tiles of 256 threads x 8 values: per-tile sums, then each tile adds up the
// sums before it (no block waits on another), scans itself in LDS and writes.
__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t* __restrict__ in, uint32_t len,
                                                   uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_tmp[8];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * 8u;
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (base + i < len) sum += in[base + i];
    uint32_t total;
    block_exclusive_scan<256>(sum, s_tmp, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_emit(const uint32_t* __restrict__ in,
                                                   uint32_t* __restrict__ out, uint32_t len,
                                                   const uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_tmp[8];
    const uint32_t tile = blockIdx.x;
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < tile; t += 256) part += sums[t];
    uint32_t prefix;
    block_exclusive_scan<256>(part, s_tmp, &prefix);
    const uint64_t base = (uint64_t)tile * kScanTile + threadIdx.x * 8u;
    uint32_t v[8], sum = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        v[i] = base + i < len ? in[base + i] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = prefix + block_exclusive_scan<256>(sum, s_tmp, &total);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (base + i < len) out[base + i] = run;
        run += v[i];
    }
}

void launch_scan(const uint32_t* in, uint32_t* out, uint32_t len, uint32_t* sums, hipStream_t s) {
    if (!len) return;
    const uint32_t tiles = scan_tiles(len);
    hipLaunchKernelGGL(k_scan_sums, dim3(tiles), dim3(256), 0, s, in, len, sums);
    hipLaunchKernelGGL(k_scan_emit, dim3(tiles), dim3(256), 0, s, in, out, len, sums);
}

// ------------------------------------------------------------------------ 1-bit split pass
// RadixSort.metal:89-100 (split_prep): e[i] = bit clear, for i < count.
__global__ void k_split_prep(const uint64_t* kin, uint32_t* e, uint32_t bit,
                             const uint32_t* d_count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *d_count) return;
    e[i] = (kin[i] & (1ull << bit)) == 0 ? 1u : 0u;
}
// RadixSort.metal:102-131 (split_scatter).
__global__ void k_split_scatter(const uint64_t* kin, const uint32_t* iin, uint64_t* kout,
                                uint32_t* iout, uint32_t bit, const uint32_t* e,
                                const uint32_t* f, const uint32_t* d_count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t count = *d_count;
    if (i >= count) return;
    uint32_t total_falses = f[count - 1] + e[count - 1];
    uint64_t k = kin[i];
    uint32_t dst = (k & (1ull << bit)) ? total_falses + (i - f[i]) : f[i];
    kout[dst] = k;
    iout[dst] = iin[i];
}
void launch_split_prep(const uint64_t* kin, uint32_t* e, uint32_t bit, const uint32_t* d_count,
                       uint32_t cap, hipStream_t s) {
    if (cap)
        hipLaunchKernelGGL(k_split_prep, dim3((cap + 255) / 256), dim3(256), 0, s, kin, e, bit,
                           d_count);
}
void launch_split_scatter(const uint64_t* kin, const uint32_t* iin, uint64_t* kout,
                          uint32_t* iout, uint32_t bit, const uint32_t* e, const uint32_t* f,
                          const uint32_t* d_count, uint32_t cap, hipStream_t s) {
    if (cap)
        hipLaunchKernelGGL(k_split_scatter, dim3((cap + 255) / 256), dim3(256), 0, s, kin, iin,
                           kout, iout, bit, e, f, d_count);
}

// ------------------------------------------------------------------------------ radix sort
__device__ inline uint32_t radix_digit(uint64_t key, uint32_t idx, RadixPassDesc d,
                                       const uint32_t* seg_of) {
    if (d.segment) return (seg_of[idx] >> d.shift) & 0xFFu;
    return (uint32_t)(key >> d.shift) & 0xFFu;
}

// Per-tile digit counts of one pass: counts[d * tiles + tile].  Tiles past the live count
// write zeros, so the digit scan can run over the whole (host-sized) grid.
__global__ __launch_bounds__(256) void k_radix_count(const uint64_t* __restrict__ kin,
                                                     const uint32_t* __restrict__ iin,
                                                     const uint32_t* d_count, RadixPassDesc desc,
                                                     const uint32_t* __restrict__ seg_of,
                                                     uint32_t* __restrict__ counts,
                                                     uint32_t tiles) {
    __shared__ uint32_t s_h[256];
    const uint32_t tid = threadIdx.x, tile = blockIdx.x;
    s_h[tid] = 0;
    __syncthreads();
    const uint32_t count = *d_count;
    const uint64_t base = (uint64_t)tile * kSortTile;
    for (uint32_t k = tid; k < kSortTile; k += kSortThreads) {
        const uint64_t e = base + k;
        if (e < count) {
            const uint32_t id = desc.segment ? iin[e] : 0u;
            atomicAdd(&s_h[radix_digit(desc.segment ? 0ull : kin[e], id, desc, seg_of)], 1u);
        }
    }
    __syncthreads();
    counts[(uint64_t)tid * tiles + tile] = s_h[tid];
}

// Exclusive scan of each digit's row of tile counts (one block per digit) and the digit
// totals.
__global__ __launch_bounds__(256) void k_radix_digit_scan(const uint32_t* __restrict__ counts,
                                                          uint32_t tiles,
                                                          uint32_t* __restrict__ excl,
                                                          uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_tmp[8];
    const uint32_t d = blockIdx.x;
    const uint32_t* row = counts + (uint64_t)d * tiles;
    uint32_t* out = excl + (uint64_t)d * tiles;
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < tiles; t0 += 256 * 4) {
        uint32_t v[4], sum = 0;
        const uint32_t tb = t0 + threadIdx.x * 4;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[i] = tb + i < tiles ? row[tb + i] : 0u;
            sum += v[i];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan<256>(sum, s_tmp, &tot);
        uint32_t run = carry + ex;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (tb + i < tiles) out[tb + i] = run;
            run += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) totals[d] = carry;
}

// One stable 8-bit digit pass: rank inside the tile, then scatter through LDS.
__global__ __launch_bounds__(256) void k_radix_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ iin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ iout, const uint32_t* d_count,
    RadixPassDesc desc, const uint32_t* __restrict__ seg_of, const uint32_t* __restrict__ excl,
    const uint32_t* __restrict__ totals, uint32_t tiles) {
    constexpr int NW = kSortThreads / 64;
    __shared__ uint32_t s_whist[NW][256];   // per-wave digit counts, then wave offsets
    __shared__ uint32_t s_boff[256];        // block-local start of each digit
    __shared__ uint32_t s_goff[256];        // global start of each digit for this tile
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_idx[kSortTile];
    __shared__ uint8_t s_dig[kSortTile];
    __shared__ uint32_t s_tmp[8];

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const uint32_t count = *d_count;
    const uint64_t base = (uint64_t)tile * kSortTile;
    if (base >= count) return;  // uniform over the block
    for (uint32_t k = tid; k < NW * 256; k += kSortThreads) (&s_whist[0][0])[k] = 0;
    __syncthreads();

    // Load: wave w owns elements [base + w*64*ITEMS, ...); item i of lane l is element
    // base + w*64*ITEMS + i*64 + l, so (wave, item, lane) order is input order.
    uint64_t key[kSortItems];
    uint32_t id[kSortItems], dig[kSortItems], rk[kSortItems];
    const uint64_t wbase = base + (uint64_t)wave * 64 * kSortItems;
#pragma unroll
    for (uint32_t i = 0; i < kSortItems; i++) {
        const uint64_t e = wbase + i * 64 + lane;
        if (e < count) {
            key[i] = kin[e];
            id[i] = iin[e];
            dig[i] = radix_digit(key[i], id[i], desc, seg_of);
        } else {
            key[i] = 0;
            id[i] = 0;
            dig[i] = 256;  // no digit: takes part in no match and no count
        }
    }

    // Rank inside the wave: lanes with equal digits found with 8 ballots; the lowest lane of
    // each group bumps the wave's count for that digit.
#pragma unroll
    for (uint32_t i = 0; i < kSortItems; i++) {
        const uint32_t d = dig[i];
        const bool valid = d < 256;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        if (valid) {
            const uint32_t below = lanes_below(m);
            const uint32_t c = s_whist[wave][d];
            if (below == 0) s_whist[wave][d] = c + (uint32_t)__popcll(m);
            rk[i] = c + below;
        }
    }
    __syncthreads();

    // Per digit (thread = digit): wave offsets, block offsets, global offsets.
    {
        const uint32_t d = tid;
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const uint32_t c = s_whist[w][d];
            s_whist[w][d] = tot;
            tot += c;
        }
        uint32_t dummy;
        const uint32_t boff = block_exclusive_scan<kSortThreads>(tot, s_tmp, &dummy);
        const uint32_t dbase = block_exclusive_scan<kSortThreads>(totals[d], s_tmp, &dummy);
        s_boff[d] = boff;
        s_goff[d] = dbase + excl[(uint64_t)d * tiles + tile];
    }
    __syncthreads();

    // Stage the tile in LDS in digit order, then write runs out coalesced.
#pragma unroll
    for (uint32_t i = 0; i < kSortItems; i++) {
        const uint32_t d = dig[i];
        if (d < 256) {
            const uint32_t pos = s_boff[d] + s_whist[wave][d] + rk[i];
            s_keys[pos] = key[i];
            s_idx[pos] = id[i];
            s_dig[pos] = (uint8_t)d;
        }
    }
    __syncthreads();
    const uint32_t nvalid = (count - base) < kSortTile ? (uint32_t)(count - base) : kSortTile;
    for (uint32_t k = tid; k < nvalid; k += kSortThreads) {
        const uint32_t d = s_dig[k];
        const uint32_t dst = s_goff[d] + (k - s_boff[d]);
        kout[dst] = s_keys[k];
        iout[dst] = s_idx[k];
    }
}

void launch_radix_pass(const uint64_t* kin, const uint32_t* iin, uint64_t* kout,
                       uint32_t* iout, const uint32_t* d_count, uint32_t cap,
                       RadixPassDesc desc, const uint32_t* seg_of, uint32_t* work,
                       hipStream_t s) {
    if (!cap) return;
    const uint32_t tiles = sort_tiles(cap);
    uint32_t* counts = work;
    uint32_t* excl = work + (uint64_t)tiles * 256;
    uint32_t* totals = excl + (uint64_t)tiles * 256;
    hipLaunchKernelGGL(k_radix_count, dim3(tiles), dim3(kSortThreads), 0, s, kin, iin, d_count,
                       desc, seg_of, counts, tiles);
    hipLaunchKernelGGL(k_radix_digit_scan, dim3(256), dim3(256), 0, s, counts, tiles, excl,
                       totals);
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, s, kin, iin, kout,
                       iout, d_count, desc, seg_of, excl, totals, tiles);
}

}  // namespace astro
