// ctx.hpp -- the libastro context (private to the library's host code).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/astro/astro.h"
#include "launch.hpp"

constexpr uint32_t kDiagChunks = 64;

struct astro_ctx {
    using TLink = astro::TLink;
    astro_config cfg{};
    astro_params params{};
    int device = 0;
    hipStream_t stream = nullptr;
    // the OctreeNode array is written on `beside` while the traversal runs on `stream` (both at
    // the default priority), joined before the step ends
    hipStream_t beside = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool emit_beside = true;
    uint32_t num_cus = 1;              // compute units of the device (emission grid)
    bool tree_built = false;           // the compact nodes hold a tree built on the current layer table
    std::string err;

    uint32_t ns = 0, nt = 0, nsys = 1;
    std::vector<uint32_t> sys_off;     // nsys + 1 particle offsets

    // particle state (SoA float4: 16-B rows instead of the 32-B reference structs)
    float4 *pos = nullptr, *vel = nullptr, *force = nullptr;
    float4* color = nullptr;           // ColorType.color (lit by the lighting pass)
    uint32_t* type = nullptr;
    uint32_t* sys_of = nullptr;
    // keys and sort
    uint64_t *codes = nullptr, *unsorted_codes = nullptr, *sorted_codes = nullptr;
    uint64_t *tmp_keys_a = nullptr, *tmp_keys_b = nullptr;
    uint32_t *unsorted_idx = nullptr, *sorted_idx = nullptr, *tmp_idx_a = nullptr,
             *tmp_idx_b = nullptr;
    uint32_t *unique_idx = nullptr, *flags = nullptr, *scan = nullptr, *ebuf = nullptr,
             *fbuf = nullptr;
    uint32_t *parent_count = nullptr, *child_count = nullptr;
    uint32_t* dev_consts = nullptr;    // [0] = ns, [1] = nt
    // tree
    float4* tpos = nullptr;            // centre and mass per node
    TLink* tlink = nullptr;            // links per node (parallel to tpos)
    uint8_t* levels = nullptr;         // head level per node (parallel to tpos)
    OctreeNode* onodes = nullptr;
    float4* qmom = nullptr;            // 2 per node: second moments, allocated on first use
    uint64_t node_capacity = 0;
    uint32_t *layer_counts = nullptr, *layer_counts_prev = nullptr, *leaf_of_sorted = nullptr;
    float4* cubes = nullptr;
    uint32_t* minmax = nullptr;        // per-system ordered min/max keys, or per-tile partials
    uint32_t* cube_counter = nullptr;  // tiles done in the one-system cube kernel (kept zero)
    uint8_t* plevel = nullptr;         // [max_particles] head level per sorted position
    uint32_t* pnext = nullptr;         // [max_tree] next pointer of the nodes starting there
    uint32_t* fused_buf = nullptr;     // per-tile counts, bases and leftovers of the fused layers
    uint32_t* err_word = nullptr;
    unsigned long long* stats = nullptr;
    uint32_t* d_sys_off = nullptr;     // [max_systems + 1] system offsets on the device
    double* phi = nullptr;             // [max_particles] FP64 potential (diagnostics)
    double* diag_partial = nullptr;    // [max_systems * kDiagChunks * 9]
    uint32_t force_block = 256;        // traversal block size (autotuned)
    uint32_t force_chunk = 0;          // traversal tile runs per XCD (autotuned; 0: eighths)
    uint32_t* guard = nullptr;         // per-system state flags raised by the integrator (server)
    float guard_extent = 0.0f;         // |coordinate| bound of the guard (0: finiteness only)
    bool force_top = true;             // traversal reads the top layers from LDS (blocks >= 512)
    bool force_wide = false;           // 64-bit node offsets and unpacked moment rows on any tree
    uint32_t force_walk = 0;           // 1: one walk per wave (monopoles, one system)
    // the step's sort seeded by the previous order (one system): its previous order (a
    // permutation of 0..nt-1 written by the last sort of these sources), bucket slots,
    // splitters (from the last sort's rows), bucket counts (+ the overflow count)
    uint32_t sort_mode = ASTRO_SORT_SEEDED;
    bool order_valid = false;          // `order` holds the last sort's order of the current sources
    bool last_sort_seeded = false;     // the last enqueued sort was the seeded one
    bool keys_pending = false;         // its per-source key arrays are not written yet
    uint32_t* order = nullptr;
    uint64_t* rs_keys = nullptr;
    uint32_t* rs_idx = nullptr;
    float4* rs_pos = nullptr;
    uint32_t* rs_type = nullptr;
    uint64_t* spl_keys = nullptr;
    uint32_t* spl_idx = nullptr;
    uint32_t* rs_count = nullptr;
    float4* spl_cube = nullptr;        // [max_systems] the cubes of the splitters' keys
    uint64_t rs_slots = 0;             // rows of rs_keys / rs_idx
    uint32_t rs_nb = 0;                // buckets of the current layout (0: no seeded sort)
    uint32_t* d_sys_b0 = nullptr;      // [max_systems + 1] first bucket of each system (batch)
    uint32_t* d_bkt_sys = nullptr;     // [kResortMaxBuckets] system of each bucket (batch)
    // positions and types in sorted order, written by every sort of a build (read by the fused
    // build, and by the traversal of the same step)
    float4* spos = nullptr;
    uint32_t* stype = nullptr;
    // layers no walk reaches are not written by the step's build (a switch, default on);
    // wlow: the lowest layer the last build wrote (device word); tree_partial: the last build
    // may have left layers out, so a caller's read of the tree builds it again in full
    bool skip_unreached = true;
    uint32_t* wlow = nullptr;
    uint32_t* sysfirst = nullptr;      // batches: [max_systems][kMaxLayers] each system's first
                                       // node per layer (the traversal's per-system LDS top)
    bool tree_partial = false;
    uint64_t device_bytes = 0;         // bytes allocated on the device at creation
    uint32_t* scratch = nullptr;       // zeroed per use: look-back statuses, histograms
    uint64_t scratch_words = 0;
    uint32_t emit_mode = ASTRO_EMIT_ON_READ;  // when the OctreeNode array is written
    bool onodes_pending = false;       // on read: the last build's array is not written yet
    bool onodes_dirty = true;          // stage calls wrote onodes outside the fused build
    // bounds of the uploaded masses, for the traversal's division domain: every mass finite
    // and >= 0, the smallest positive one, the largest system total (conservative)
    bool mass_ok = false;
    float mass_min_pos = 0.0f;
    double mass_total_max = 0.0;

    // layer table of the current particles
    uint32_t layers = 0;
    uint64_t cap[astro::kMaxLayers] = {}, off[astro::kMaxLayers] = {}, total = 0;

    // graph of one fused step
    hipGraphExec_t graph = nullptr;
    float graph_dt = 0.0f;
    bool graph_valid = false;
    bool use_graph = true;
};

// ------------------------------------------------------------------------------- helpers
inline int fail(astro_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
#define HIPCHK(c, call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((c), ASTRO_ERR_HIP,                                               \
                        std::string(#call) + ": " + hipGetErrorString(e_));               \
    } while (0)
#define LAUNCHCHK(c)                                                                      \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess)                                                             \
            return fail((c), ASTRO_ERR_HIP, std::string("kernel launch: ") +              \
                                                hipGetErrorString(e_));                   \
    } while (0)


// Folds masses m[0..n) (stride floats apart) into the context's mass bounds; reset first when
// the uploaded set replaces the previous one.
void astro_mass_bounds(astro_ctx* c, const float* m, uint32_t n, size_t stride, bool reset);

// host-side pieces shared by capi.hip and serve.hip
int astro_refresh_layers(astro_ctx* c);
int astro_set_counts(astro_ctx* c, uint32_t ns, uint32_t nt);
int astro_enqueue_step(astro_ctx* c, float dt);
// On read: writes the pending OctreeNode array of the last build (no-op otherwise).
int materialize_onodes(astro_ctx* c);
// On read: writes the per-source key arrays of the last seeded sort (no-op otherwise).
int materialize_keys(astro_ctx* c);
// On read: writes the tree layers the last build left out (no-op otherwise).
int materialize_tree(astro_ctx* c);
