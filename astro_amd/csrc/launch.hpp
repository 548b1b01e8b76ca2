// launch.hpp -- host-side launchers of the libastro kernels (defined in the k_*.hip files).
#pragma once

#include <vector>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/astro/astro.h"
#include "accept.hpp"
#include "common.hpp"

namespace astro {

// ---- k_sort.hip ---------------------------------------------------------------------------
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortItems = 4;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;  // keys per tile
constexpr uint32_t kScanTile = 2048;                       // values per scan tile

inline uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }
inline uint32_t scan_tiles(uint32_t n) { return (n + kScanTile - 1) / kScanTile; }

void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
void launch_fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);
void launch_iota(uint32_t* buf, uint32_t start, uint32_t end, hipStream_t s);
void launch_copy_u64_count(const uint64_t* in, uint64_t* out, const uint32_t* d_count,
                           uint32_t cap, hipStream_t s);
void launch_copy_u32_count(const uint32_t* in, uint32_t* out, const uint32_t* d_count,
                           uint32_t cap, hipStream_t s);
void launch_copy_u32(const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t s);

// Exclusive scan of len values; sums holds scan_tiles(len) words (no clearing needed).
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t len, uint32_t* sums, hipStream_t s);

// 1-bit split pass (RadixSort.metal:89-131 semantics).
void launch_split_prep(const uint64_t* kin, uint32_t* e, uint32_t bit, const uint32_t* d_count,
                       uint32_t cap, hipStream_t s);
void launch_split_scatter(const uint64_t* kin, const uint32_t* iin, uint64_t* kout,
                          uint32_t* iout, uint32_t bit, const uint32_t* e, const uint32_t* f,
                          const uint32_t* d_count, uint32_t cap, hipStream_t s);

// LSD radix sort building blocks.  A "digit pass" p sorts by (key >> shift[p]) & 0xFF, or,
// for a segment pass, by (seg_of[index] >> seg_shift) & 0xFF.
struct RadixPassDesc {
    uint32_t shift;      // key bit offset of the digit
    uint32_t segment;    // 1: digit from seg_of[index] instead of the key
};
constexpr uint32_t kMaxRadixPasses = 10;
struct RadixPlan {
    uint32_t npasses;
    RadixPassDesc pass[kMaxRadixPasses];
};
// One stable digit pass = per-tile digit counts, per-digit scan over tiles, rank + scatter.
// work: radix_work_words(cap) words (fully rewritten by each pass; no clearing needed).
inline uint64_t radix_work_words(uint32_t cap) { return 2ull * sort_tiles(cap) * 256 + 256; }
void launch_radix_pass(const uint64_t* kin, const uint32_t* iin, uint64_t* kout,
                       uint32_t* iout, const uint32_t* d_count, uint32_t cap,
                       RadixPassDesc desc, const uint32_t* seg_of, uint32_t* work,
                       hipStream_t s);

// ---- k_resort.hip -------------------------------------------------------------------------
// The sort seeded by the previous order: same result as the radix sort of (key, index) pairs
// (ascending key, ties by index), for any permutation `order` of 0..n-1.
constexpr uint32_t kResortThreads = 1024;      // the sample block
constexpr uint32_t kResortBucketThreads = 512; // bucket blocks
constexpr uint32_t kResortPlaceThreads = 256;
constexpr uint32_t kResortMean = 980;          // buckets = ceil(n / kResortMean)
constexpr uint32_t kResortRegRows = 2 * kResortBucketThreads;  // sorted in registers up to this
constexpr uint32_t kResortCap = 2048;          // slots per bucket (sorted in LDS up to this)
constexpr uint32_t kResortMaxBuckets = 8192;
constexpr uint32_t kResortMaxSamples = 8192;   // samples sorted in one block's LDS
constexpr uint32_t kResortOversample = 2;      // samples per bucket (at most)
// Up to this many sources the register sort's rows carry index << 10 | slot (slot < 1,024)
constexpr uint32_t kResortPackMax = 1u << 22;
struct ResortArgs {
    const float4* pos;       // [n] source positions
    const float4* cubes;     // [nsys] the systems' cubes
    astro_params params;     // quantizer and key bits
    uint32_t* order;         // [n] previous sorted order in, this sort's order out
    uint32_t n;
    uint64_t* rk;            // [buckets * kResortCap] bucket slots: key, index, position, type
    uint32_t* ri;
    float4* rp;
    uint32_t* rt;
    uint64_t* spl_k;         // [kResortMaxBuckets] splitters (the bucket kernel leaves the next
    uint32_t* spl_i;         //  sort's, from its own sorted rows)
    float4* spl_cube;        // [nsys] the cubes the splitters' keys were made with
    uint32_t* moved_in;      // [kResortMaxBuckets] sources that came into each bucket
    uint32_t* moved_out;     // [kResortMaxBuckets] sources that left each bucket's positions
    uint32_t* ovf;           // [1] overflow entries
    uint64_t* ok;            // [n] overflow list: key, index, bucket
    uint32_t* oi;
    uint32_t* ob;
    uint64_t* out_k;         // [n] sorted keys and indices
    uint32_t* out_i;
    const uint32_t* type;    // [n] source types
    float4* spos;            // [n] positions and types in sorted order (out)
    uint32_t* stype;
    // Batches (nsys > 1): system s holds positions and sources [sys_off[s], sys_off[s+1]) and
    // buckets [sys_b0[s], sys_b0[s+1]); a bucket never spans two systems, so that rows of one
    // bucket compare by (key, index) alone.  One system: nsys 1, the arrays unused.
    uint32_t nb;             // buckets in all
    uint32_t nsys;
    const uint32_t* sys_of;  // [n] system of each source (and of each position)
    const uint32_t* sys_off; // [nsys + 1]
    const uint32_t* sys_b0;  // [nsys + 1]
    const uint32_t* bkt_sys; // [buckets] the system of each bucket
};
// Buckets of a seeded sort of n sources (of one system); 0 when n is beyond it.
uint32_t resort_buckets(uint32_t n);
// Buckets of a seeded sort of the systems [off[s], off[s+1]) (each gets its own, at least one
// per source-holding system), their first buckets b0 (nsys + 1 entries) and the system of
// every bucket; 0 when beyond the sort's bucket limit.
uint32_t resort_buckets_batch(const std::vector<uint32_t>& off, std::vector<uint32_t>* b0,
                              std::vector<uint32_t>* bkt_sys);
// Bucket slots (rows of rk / ri) a seeded sort of up to nmax sources needs.
inline uint64_t resort_slots(uint32_t nmax) {
    uint64_t nb = ((uint64_t)nmax + kResortMean - 1) / kResortMean;
    if (nb > kResortMaxBuckets) nb = kResortMaxBuckets;
    return (nb ? nb : 1) * kResortCap;
}
void launch_resort(const ResortArgs& a, hipStream_t s);
// After a radix sort of the same sources: the previous order and the splitters of the next
// seeded sort from its result (sk, si), and the cube they were made with; the positions and
// types in sorted order.
void launch_resort_seed(const uint64_t* sk, const uint32_t* si, const ResortArgs& a,
                        hipStream_t s);
// Positions and types in sorted order: spos[j] = pos[si[j]], stype[j] = type[si[j]], j < n.
void launch_sorted_gather(const uint32_t* si, const float4* pos, const uint32_t* type, uint32_t n,
                          float4* spos, uint32_t* stype, hipStream_t s);
// The key kernel's per-source outputs (codes, their copy, 0..n-1) from a sort's result.
void launch_resort_keys(const uint64_t* sk, const uint32_t* si, uint32_t n, uint64_t* codes,
                        uint64_t* codes2, uint32_t* idx, hipStream_t s);

// ---- k_tree.hip ---------------------------------------------------------------------------
struct TreeArgs {
    const float4* pos;          // particle positions + mass
    const float4* spos;         // [nt] the same in sorted order (the sort writes it)
    const uint32_t* sys_of;     // system of each particle / sorted position
    const uint64_t* sorted_keys;
    const uint32_t* sorted_idx;
    float4* tpos;               // centre and mass of every node
    TLink* tlink;               // links of every node (parallel to tpos)
    uint8_t* levels;            // head level of every node (parallel to tpos)
    float4* qmom;               // second moments, 2 float4 per node (quadrupole mode)
    uint32_t quad;              // 1: compute second moments
    OctreeNode* onodes;         // may be null (no emission)
    uint32_t* counts;           // [layers] valid nodes per layer
    uint32_t* leaf_of_sorted;   // [nt]
    uint32_t* pnext;            // [nt] per head position s: the next of every node starting at s
    uint32_t* err;
    uint64_t off[kMaxLayers];
    uint64_t cap[kMaxLayers];
    uint32_t layers;            // B + 1
    uint32_t bits;              // B
    uint32_t nt;                // sources
    uint32_t fp64_from;         // first layer aggregated in FP64
    uint32_t quirk_last_leaf;
    uint32_t segmented;         // 1: heads also at system starts
    uint32_t pack;              // 1: fewer than kLinkPackNodes nodes (links in the moment rows)
    // Layers no walk reaches are not written (skip_ok: one system, monopoles, the traversal's
    // fast form): the scan of the fused layers finds the lowest layer W a finite target's walk
    // can reach from the opening test (acc, eps2, the cube) and stores it in *wlow (0: every
    // layer written); W's nodes get kNextEnd as last child, so that no walk goes below them.
    uint32_t skip_ok;
    AcceptRule acc;
    float eps2;
    uint32_t side_shifted;      // ASTRO_SIDE_SHIFTED: side0 = 2^21
    const float4* cubes;
    uint32_t* wlow;             // [3]: W, whether any tile left a node for fused_fix, and
                                //      the fused build's finished-block count (left 0)
    // Batches (else null): [nsys][kMaxLayers] the index in layer l of system s's first node
    // (the node starting at the system's first position, a head at every layer)
    uint32_t* sysfirst;
};

void launch_bbox_init(uint32_t* minmax, uint32_t nsys, hipStream_t s);
void launch_bbox(const float4* pos, uint32_t n, const uint32_t* sys_of, uint32_t* minmax,
                 hipStream_t s);
void launch_cube_finalize(const uint32_t* minmax, float4* cubes, uint32_t nsys, hipStream_t s);
// One system in one launch: part holds 8 * cube_single_blocks(n) words, *counter is zero
// before the launch and is left zero.
uint32_t cube_single_blocks(uint32_t n);
void launch_cube_single(const float4* pos, uint32_t n, uint32_t* part, uint32_t* counter,
                        float4* cubes, hipStream_t s);
void launch_morton(const float4* pos, uint32_t n, const uint32_t* sys_of, const float4* cubes,
                   const astro_params& p, uint64_t* codes, uint64_t* codes2, uint32_t* idx,
                   hipStream_t s);
// Layers: the leaf layer's heads are counted per tile by launch_layer_count; the emit of
// layer l writes its nodes and adds the heads of layer l+1 to next_cnt / next_super (zeroed;
// null for the top layer).  Per layer: layer_tiles(candidate capacity) tile counts and
// layer_super_words(...) words of supertile sums (spread one per 128-byte line).
uint32_t layer_tiles(uint64_t candidates);
uint64_t layer_super_words(uint64_t candidates);
void launch_layer_count(const TreeArgs& a, uint32_t layer, uint32_t* tile_cnt, uint32_t* super,
                        hipStream_t s);
// Before any node is written: the head level of every sorted position (plevel), the node
// count of every layer below the top per 1,024-position tile and its scan over the tiles
// (fbuf), and, without fused layers (lf = 0), from them a.pnext, the next pointer of the nodes
// starting at each position.  Layers 0 .. lf-1 (lf <= kFusedMaxLayers; 0: none) are then built
// in one launch by launch_fused_layers, which computes those next pointers itself (and stores
// them in a.pnext when its leftover pass will read them).  plevel holds nt bytes and fbuf
// fused_buf_words(nt) words, both scratch.
constexpr uint32_t kFusedMaxLayers = 16;
uint32_t fused_tiles(uint32_t nt);
inline uint64_t fused_buf_words(uint32_t nt) {
    return (2ull * kMaxLayers + kFusedMaxLayers) * fused_tiles(nt);
}
// Also zeroes clr[0 .. nclr) (the per-layer kernels' counts) before any layer kernel runs.
void launch_heads_next(const TreeArgs& a, uint32_t lf, uint8_t* plevel, uint32_t* fbuf,
                       uint32_t* clr, uint32_t nclr, hipStream_t s);
// lcnt / lsup (or null): the tile counts and supertile sums of layer lf's candidates, which the
// fused build adds to (zeroed before it), so that no counting launch precedes layer lf.
void launch_fused_layers(const TreeArgs& a, uint32_t lf, const uint8_t* plevel, uint32_t* fbuf,
                         uint32_t* lcnt, uint32_t* lsup, hipStream_t s);
// The top layers with at most kTopCandidates candidates each in one single-block launch:
// layers_top_first(a) is the first of them (a.layers if none; layer 0 never), and the layer
// before it is emitted without next-layer counts.
uint32_t layers_top_first(const TreeArgs& a);
void launch_layers_top(const TreeArgs& a, uint32_t l0, hipStream_t s);
void launch_layer_emit(const TreeArgs& a, uint32_t layer, const uint32_t* tile_cnt,
                       const uint32_t* super, uint32_t* next_cnt, uint32_t* next_super,
                       hipStream_t s);
void launch_fill_sys_of(uint32_t* sys_of, const uint32_t* d_off, uint32_t nsys, uint32_t n,
                        hipStream_t s);
// The OctreeNode array from the compact nodes (and emptied stale slots) after a fused build,
// by `blocks` persistent blocks.
void launch_emit_onodes(const TreeArgs& a, const uint32_t* counts_prev, uint32_t blocks,
                        hipStream_t s);

// Reference-shaped stage kernels (Octree.metal live kernels).
void launch_mark_uniques(const uint64_t* sorted, uint32_t* flags, uint32_t count,
                         uint32_t aggregate, hipStream_t s);
void launch_scatter_uniques(const uint32_t* flags, const uint32_t* scan, uint32_t* uniq,
                            uint32_t count, hipStream_t s);
void launch_copy_last_scan(const uint32_t* scan, const uint32_t* flags, uint32_t* parent,
                           uint32_t* child, uint32_t count, hipStream_t s);
void launch_aggregate_leaf(const uint64_t* sorted, const uint32_t* sorted_idx,
                           const uint32_t* uniq, const float4* pos, const uint32_t* parent,
                           OctreeNode* nodes, uint64_t* unsorted, uint32_t* unsorted_idx,
                           uint32_t grid_count, uint32_t quirk, uint32_t nt, uint32_t fp64,
                           hipStream_t s);
void launch_aggregate_nodes(OctreeNode* nodes, const uint64_t* sorted, const uint32_t* sorted_idx,
                            const uint32_t* uniq, const uint32_t* parent, uint32_t in_off,
                            uint32_t out_off, uint32_t layer, uint64_t* unsorted,
                            uint32_t* unsorted_idx, const uint32_t* child, uint32_t node_count,
                            uint32_t fp64, hipStream_t s);
void launch_clear_onodes(OctreeNode* nodes, uint64_t count, hipStream_t s);

// ---- k_force.hip --------------------------------------------------------------------------
struct ForceArgs {
    const float4* tpos;         // node centres and masses
    const TLink* tlink;         // node links
    const float4* qmom;         // second moments per node (quadrupole mode)
    const OctreeNode* onodes;
    const float4* pos;
    const uint32_t* type;
    const float4* spos;         // [nt] positions and types in sorted order, when the step's sort
    const uint32_t* stype;      //  wrote them since the positions last changed; else null
    const uint32_t* sorted_idx;
    const uint32_t* leaf_of_sorted;
    const uint32_t* sys_of;
    const float4* cubes;
    const uint64_t* codes;      // per-particle keys (reference-layout traversal self test)
    float4* force;
    uint32_t* err;
    unsigned long long* stats;  // 4 counters, or null
    double* phi;                // per-particle FP64 potential, or null
    const uint32_t* counts;     // [layers] valid nodes per layer (device)
    const uint32_t* wlow;       // [1] the lowest layer the build wrote (the LDS top stays above)
    const uint32_t* sysfirst;   // batches (else null): the build's first node of each system
    uint32_t nsys;              //   per layer ([nsys][kMaxLayers]), and the system count
    uint64_t off[kMaxLayers];   // layer offsets
    uint32_t layers;
    uint32_t top;               // 1: the top layers from LDS (blocks of 512 or 1,024 threads)
    uint32_t wide;              // 1: 64-bit node offsets on any tree
    uint32_t walk;              // 1: one walk per wave (k_force_wave) where it applies
    uint64_t off0;              // leaf layer offset
    uint64_t root_off;          // top layer offset (root of system s at root_off + s)
    uint64_t node_total;        // nodes in the current layer table (index bound)
    uint32_t ns, nt;
    uint32_t block;             // traversal block size (64, 128, 256, 512 or 1024)
    uint32_t xcd_chunk;         // tiles per run dealt to the XCDs in turn (0: contiguous eighths)
    AcceptRule acc;             // opening test (accept.hpp), from p.theta
    uint32_t fast;              // 1: eps2 >= max(r2_floor, 2^-96, leaf_cutoff_r2+),
                                //    sqrt(eps2) >= r_floor: no floor or cut-off can bind
    float div_lmax;             // > 0: when a system's cube lies within [-div_lmax, div_lmax]^3,
                                //    every G M / (r2 r) of its targets is in the domain of
                                //    div_rn_mid; 0: always use the full division
    astro_params p;
};
// Self-check of div_rn_mid against the IEEE division on n pseudo-random pairs of its domain.
int check_division(uint64_t n, uint64_t seed, unsigned long long* d_mismatch, hipStream_t s);
// Self-check of the r2 form of the opening test against the square-root form.
int check_open_test(uint64_t n, uint64_t seed, unsigned long long* d_mismatch, hipStream_t s);
// Exhaustive self-check of the correctly rounded square root; returns the floats checked (0 on
// a launch error).
uint64_t check_sqrt(unsigned long long* d_mismatch, hipStream_t s);
// Exhaustive self-check of the refined reciprocal behind the fast division; returns the floats
// checked (0 on a launch error).
uint64_t check_reciprocal(unsigned long long* d_mismatch, hipStream_t s);
// Fills acc and fast from p.
void set_accept_threshold(ForceArgs& a);
bool force_block_supported(uint32_t nt);
// The traversal's addressing: 64-bit node offsets (trees of 2^27 nodes or more, or `wide`).
bool force_off64(const ForceArgs& a);
void launch_force_compact(const ForceArgs& a, hipStream_t s);
void launch_force_reference(const ForceArgs& a, hipStream_t s);
void launch_diag(const float4* pos, const float4* vel, const double* phi,
                 const uint32_t* d_sys_off, uint32_t nsys, uint32_t chunks, double* partial,
                 hipStream_t s);
void launch_lighting(const float4* pos, float4* color, const uint32_t* type, uint32_t n,
                     uint32_t nstars, const uint32_t* sys_of, const uint32_t* sys_off,
                     hipStream_t s);
// guard (optional, one word per system): set to nonzero for the system of every particle
// whose new state is not finite or whose position leaves [-extent, extent]^3 (extent <= 0:
// finiteness only).
void launch_update(float4* pos, float4* vel, const float4* force, const uint32_t* type,
                   uint32_t n, float dt, const astro_params& p, hipStream_t s,
                   const uint32_t* sys_of = nullptr, uint32_t* guard = nullptr,
                   float extent = 0.0f);

void launch_direct_fp64(const float4* pos, const uint32_t* sys_of, const uint32_t* d_sys_off,
                        uint32_t nt, const uint32_t* targets, uint32_t n_targets, uint32_t chunks,
                        double G, double eps2, double* partial, hipStream_t s);

}  // namespace astro
