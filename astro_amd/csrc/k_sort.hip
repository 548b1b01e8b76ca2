// k_sort.hip -- scans and the LSD radix sort of (u64 key, u32 index) pairs for gfx950.
//
// The reference sorts with 64 one-bit split passes, each a prep + 1-3 scan dispatches + a
// scatter (RadixSort.swift:98-185, RadixSort.metal:30-131): ~3.8 KB of traffic per key.  Here
// a sort is one histogram pass over the keys followed by one pass per 8-bit digit.  A digit
// pass ranks keys inside a wavefront with ballots (no atomics), turns per-wave counts into
// block offsets in LDS, gets the tile's global offset per digit by decoupled look-back, stages
// the tile in LDS in digit order and writes it out coalesced.  Every pass is stable, so the
// result is the stable ascending argsort -- the same permutation the split sort produces.
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
// One pass: tile = 256 threads x 8 consecutive values; block scan in LDS; tile prefix by
// decoupled look-back; tile ids handed out in launch order by an atomic counter so that a
// tile only ever waits on tiles that are already running.
__global__ __launch_bounds__(256) void k_scan(const uint32_t* __restrict__ in,
                                              uint32_t* __restrict__ out, uint32_t len,
                                              uint32_t* status, uint32_t* counter,
                                              uint32_t* err) {
    __shared__ uint32_t s_tile, s_prefix, s_tmp[8];
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * kScanTile + threadIdx.x * 8u;
    uint32_t v[8], sum = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t e = base + i;
        v[i] = e < len ? in[e] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t excl = block_exclusive_scan<256>(sum, s_tmp, &total);
    if (threadIdx.x == 0) s_prefix = lookback_exclusive(status, tile, total, err);
    __syncthreads();
    uint32_t run = s_prefix + excl;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t e = base + i;
        if (e < len) out[e] = run;
        run += v[i];
    }
}

void launch_scan(const uint32_t* in, uint32_t* out, uint32_t len, uint32_t* status,
                 uint32_t* err, hipStream_t s) {
    if (!len) return;
    uint32_t tiles = scan_tiles(len);
    hipLaunchKernelGGL(k_scan, dim3(tiles), dim3(256), 0, s, in, out, len, status,
                       status + tiles, err);
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

// Histograms of every digit pass in one read of the keys.
__global__ __launch_bounds__(256) void k_radix_hist(const uint64_t* __restrict__ keys,
                                                    const uint32_t* d_count,
                                                    const uint32_t* seg_of, RadixPlan plan,
                                                    uint32_t* hist) {
    __shared__ uint32_t s_h[kMaxRadixPasses][256];
    for (uint32_t k = threadIdx.x; k < kMaxRadixPasses * 256; k += 256) (&s_h[0][0])[k] = 0;
    __syncthreads();
    const uint32_t n = *d_count;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        uint64_t key = keys[i];
        for (uint32_t p = 0; p < plan.npasses; p++) {
            // the segment digit multiset does not depend on the order, so position i serves
            uint32_t d = radix_digit(key, i, plan.pass[p], seg_of);
            atomicAdd(&s_h[p][d], 1u);
        }
    }
    __syncthreads();
    for (uint32_t p = 0; p < plan.npasses; p++) {
        uint32_t c = s_h[p][threadIdx.x];
        if (c) atomicAdd(&hist[p * 256 + threadIdx.x], c);
    }
}

void launch_radix_hist(const uint64_t* keys, const uint32_t* d_count, uint32_t cap,
                       const uint32_t* seg_of, const RadixPlan& plan, uint32_t* hist,
                       hipStream_t s) {
    if (!cap || !plan.npasses) return;
    uint32_t blocks = (cap + 256 * 16 - 1) / (256 * 16);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_radix_hist, dim3(blocks), dim3(256), 0, s, keys, d_count, seg_of, plan,
                       hist);
}

// One stable 8-bit digit pass.
__global__ __launch_bounds__(256) void k_radix_pass(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ iin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ iout, const uint32_t* d_count,
    const uint32_t* __restrict__ hist, RadixPassDesc desc, const uint32_t* __restrict__ seg_of,
    uint32_t* status, uint32_t* counter, uint32_t* err) {
    constexpr int NW = kSortThreads / 64;
    __shared__ uint32_t s_whist[NW][256];   // per-wave digit counts, then wave offsets
    __shared__ uint32_t s_boff[256];        // block-local start of each digit
    __shared__ uint32_t s_goff[256];        // global start of each digit for this tile
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_idx[kSortTile];
    __shared__ uint8_t s_dig[kSortTile];
    __shared__ uint32_t s_tmp[8];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    for (uint32_t k = tid; k < NW * 256; k += kSortThreads) (&s_whist[0][0])[k] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t count = *d_count;
    const uint64_t base = (uint64_t)tile * kSortTile;
    if (base >= count) return;  // uniform over the block

    // Load: wave w owns elements [base + w*64*ITEMS, ...); item i of lane l is element
    // base + w*64*ITEMS + i*64 + l, so (wave, item, lane) order is input order.
    uint64_t key[kSortItems];
    uint32_t id[kSortItems], dig[kSortItems], rk[kSortItems];
    const uint64_t wbase = base + (uint64_t)wave * 64 * kSortItems;
#pragma unroll
    for (uint32_t i = 0; i < kSortItems; i++) {
        uint64_t e = wbase + i * 64 + lane;
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
            uint32_t c = s_whist[w][d];
            s_whist[w][d] = tot;
            tot += c;
        }
        uint32_t dummy;
        const uint32_t boff = block_exclusive_scan<kSortThreads>(tot, s_tmp, &dummy);
        const uint32_t hex = block_exclusive_scan<kSortThreads>(hist[d], s_tmp, &dummy);
        s_boff[d] = boff;
        // look-back along this digit's column of the status table
        uint32_t* col = status + d;
        uint32_t prefix = 0;
        if (tile == 0) {
            st_release(&col[0], kFlagInc | tot);
        } else {
            st_release(&col[(uint64_t)tile * 256], kFlagAgg | tot);
            int64_t t = (int64_t)tile - 1;
            uint32_t spins = 0;
            while (t >= 0) {
                uint32_t sv = ld_acquire(&col[(uint64_t)t * 256]);
                uint32_t f = sv & ~kValMask;
                if (f == 0u) {
                    if (++spins > kSpinLimit) {
                        atomicOr(err, ERR_SPIN);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += sv & kValMask;
                if (f == kFlagInc) break;
                t--;
            }
            st_release(&col[(uint64_t)tile * 256], kFlagInc | (prefix + tot));
        }
        s_goff[d] = hex + prefix;
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
                       const uint32_t* hist, RadixPassDesc desc, const uint32_t* seg_of,
                       uint32_t* status, uint32_t* err, hipStream_t s) {
    if (!cap) return;
    uint32_t tiles = sort_tiles(cap);
    hipLaunchKernelGGL(k_radix_pass, dim3(tiles), dim3(kSortThreads), 0, s, kin, iin, kout, iout,
                       d_count, hist, desc, seg_of, status, status + (uint64_t)tiles * 256, err);
}

}  // namespace astro
