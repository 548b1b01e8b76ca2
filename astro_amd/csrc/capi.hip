// capi.hip -- libastro's C-ABI: context, buffers, stage calls and the fused step.
//
// The context plays the role of the reference's Renderer + MetalKernelsRadixSort objects
// (Renderer.swift:56-111, RadixSort.swift:24-67): it owns every device buffer, allocated
// once at creation, and one HIP stream on which every call enqueues its kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "ctx.hpp"

using namespace astro;

template <typename T>
static hipError_t dalloc(uint64_t& tally, T** p, uint64_t count) {
    if (count == 0) count = 1;
    tally += count * sizeof(T);
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

// cap_l = sum over systems of min(N_s, 8^(B-l)); offsets are their prefix sums
// (Renderer.swift:373-386 generalised to B bits and several systems).
static void compute_layers(const std::vector<uint32_t>& sizes, uint32_t bits, uint32_t* layers,
                           uint64_t* cap, uint64_t* off, uint64_t* total) {
    *layers = bits + 1;
    uint64_t o = 0;
    for (uint32_t l = 0; l < *layers; l++) {
        uint32_t b = 3 * bits - 3 * l;
        uint64_t possible = b >= 63 ? (1ull << 63) : (1ull << b);
        uint64_t c = 0;
        for (uint32_t n : sizes) c += std::min<uint64_t>(n, possible);
        cap[l] = c;
        off[l] = o;
        o += c;
    }
    *total = o;
}

static uint64_t scratch_need(uint32_t nmax, const uint64_t* cap, uint32_t layers) {
    uint64_t sort = radix_work_words(nmax);                          // one digit pass
    uint64_t lay = layer_tiles(nmax) + layer_super_words(nmax);      // leaf layer
    for (uint32_t l = 1; l < layers; l++)
        lay += layer_tiles(cap[l - 1]) + layer_super_words(cap[l - 1]);
    uint64_t scan = scan_tiles(nmax) + 1;                            // astro_scan
    return std::max(std::max(sort, lay), scan);
}

static void invalidate_graph(astro_ctx* c) {
    if (c->graph) hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    c->graph_valid = false;
}

static int refresh_layers(astro_ctx* c) {
    c->tree_built = false;  // the layer table may change: build before traversing again
    c->order_valid = false;  // the sources may change: the next sort starts from scratch
    std::vector<uint32_t> sizes;
    if (c->nsys <= 1) sizes.push_back(c->nt);
    else
        for (uint32_t s = 0; s < c->nsys; s++) sizes.push_back(c->sys_off[s + 1] - c->sys_off[s]);
    compute_layers(sizes, c->params.bits_per_axis, &c->layers, c->cap, c->off, &c->total);
    if (c->total > c->node_capacity)
        return fail(c, ASTRO_ERR_CAPACITY, "tree needs more nodes than the context holds");
    // the seeded sort's buckets: a batch gives every system its own
    if (c->nsys <= 1) {
        c->rs_nb = resort_buckets(c->nt);
    } else {
        std::vector<uint32_t> b0, bs;
        c->rs_nb = resort_buckets_batch(c->sys_off, &b0, &bs);
        if (c->rs_nb) {
            HIPCHK(c, hipMemcpyAsync(c->d_sys_b0, b0.data(), b0.size() * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->d_bkt_sys, bs.data(), bs.size() * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));  // the host vectors go out of scope
        }
    }
    invalidate_graph(c);
    c->onodes_dirty = true;  // stale-tail clearing assumes an unchanged layer table
    return ASTRO_OK;
}

// ---------------------------------------------------------------------- lifetime/params
extern "C" void astro_params_compat(astro_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->bits_per_axis = 6;  // MortonCode.metal:43-47
    p->quantizer = ASTRO_QUANT_OFFSET_WRAP;
    p->quant_offset = 1000000.0f;
    p->side_rule = ASTRO_SIDE_SHIFTED;
    p->theta = 0.7f;        // Renderer.swift:864
    p->eps2 = 0.01f;        // BarnesHut.metal:40
    p->G = 1e11f;           // BarnesHut.metal:39
    p->r2_floor = 1e-6f;    // BarnesHut.metal:59
    p->r_floor = 1e-6f;     // BarnesHut.metal:61
    p->leaf_cutoff_r2 = 50.0f;  // BarnesHut.metal:70
    p->force_clamp = 1e5f;  // BarnesHut.metal:100
    p->damping = 0.9995f;   // BarnesHut.metal:190
    p->type_scale[0] = 5.0f;
    p->type_scale[1] = 5.0f;
    p->type_scale[2] = 1.0f;  // BarnesHut.metal:182-188
    p->quirk_last_leaf = 1;   // Octree.metal:71
    p->fp64_top_layers = 0;
    p->kahan = 0;
    p->multipole = 0;
    p->lighting_stars = 5;    // Renderer.swift:26 (numStars), BarnesHut.metal:103-167
}

extern "C" void astro_params_physics(astro_params* p, uint32_t bits_per_axis) {
    astro_params_compat(p);
    p->bits_per_axis = bits_per_axis;
    p->quantizer = ASTRO_QUANT_BBOX_POW2;
    p->quant_offset = 0.0f;
    p->side_rule = ASTRO_SIDE_EXACT;
    p->G = 1.0f;
    p->leaf_cutoff_r2 = 0.0f;
    p->force_clamp = 0.0f;
    p->damping = 1.0f;
    p->type_scale[0] = p->type_scale[1] = p->type_scale[2] = 1.0f;
    p->quirk_last_leaf = 0;
    p->fp64_top_layers = 4;
    p->lighting_stars = 0;
}

extern "C" int astro_device_count(int* count) {
    if (!count) return ASTRO_ERR_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return e == hipSuccess ? ASTRO_OK : ASTRO_ERR_HIP;
}

extern "C" int astro_destroy(astro_ctx* c) {
    if (!c) return ASTRO_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    invalidate_graph(c);
    void* ptrs[] = {c->pos, c->vel, c->color, c->force, c->type, c->sys_of, c->codes, c->unsorted_codes,
                    c->sorted_codes, c->tmp_keys_a, c->tmp_keys_b, c->unsorted_idx,
                    c->sorted_idx, c->tmp_idx_a, c->tmp_idx_b, c->unique_idx, c->flags,
                    c->scan, c->ebuf, c->fbuf, c->parent_count, c->child_count, c->dev_consts,
                    c->tpos, c->tlink, c->onodes, c->layer_counts, c->layer_counts_prev,
                    c->leaf_of_sorted, c->pnext, c->cubes, c->minmax, c->cube_counter, c->plevel, c->fused_buf,
                    c->err_word, c->stats, c->scratch,
                    c->d_sys_off, c->phi, c->diag_partial, c->levels, c->qmom, c->order,
                    c->rs_keys, c->rs_idx, c->spl_keys, c->spl_idx, c->rs_count, c->spl_cube,
                    c->spos, c->stype, c->rs_pos, c->rs_type, c->wlow, c->d_sys_b0,
                    c->d_bkt_sys, c->sysfirst};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->beside) hipStreamDestroy(c->beside);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return ASTRO_OK;
}

extern "C" int astro_create(const astro_config* cfg, astro_ctx** out) {
    if (!cfg || !out) return ASTRO_ERR_ARG;
    *out = nullptr;
    if (cfg->max_particles == 0 || cfg->bits_per_axis == 0 || cfg->bits_per_axis > 21)
        return ASTRO_ERR_ARG;
    astro_ctx* c = new astro_ctx();
    c->cfg = *cfg;
    if (c->cfg.max_tree == 0 || c->cfg.max_tree > c->cfg.max_particles)
        c->cfg.max_tree = c->cfg.max_particles;
    if (c->cfg.max_systems == 0) c->cfg.max_systems = 1;
    c->device = cfg->device;
    c->emit_mode = cfg->emit_octree_nodes <= ASTRO_EMIT_ON_READ ? cfg->emit_octree_nodes
                                                                : ASTRO_EMIT_ON_READ;
    astro_params_compat(&c->params);
    c->params.bits_per_axis = std::min<uint32_t>(c->params.bits_per_axis, cfg->bits_per_axis);
    const uint32_t NP = c->cfg.max_particles, NT = c->cfg.max_tree;

    // Node capacity: sum over layers of min(N_t, S * 8^(B-l)) bounds every split of N_t
    // into at most S systems, for every B' <= B (fewer bits only shrink the layers).
    uint64_t cap_bound = 0;
    uint64_t cap_max[kMaxLayers] = {};
    for (uint32_t bits = 1; bits <= cfg->bits_per_axis; bits++) {
        uint64_t t = 0;
        for (uint32_t l = 0; l <= bits; l++) {
            uint32_t b = 3 * bits - 3 * l;
            // S * 8^(B-l), saturated (S < 2^32, so 2^b * S fits when b < 32)
            uint64_t possible = b >= 32 ? UINT64_MAX : (1ull << b) * c->cfg.max_systems;
            uint64_t cl = std::min<uint64_t>(NT, possible);
            t += cl;
            if (bits == cfg->bits_per_axis) cap_max[l] = cl;
        }
        cap_bound = std::max(cap_bound, t);
    }
    c->node_capacity = cap_bound;
    if (cap_bound >= (1ull << 29)) {
        astro_destroy(c);
        return ASTRO_ERR_CAPACITY;  // node indices stay far below the 0xFFFFFFFF sentinels
    }

    auto bad = [&](hipError_t e) {
        if (e != hipSuccess) {
            astro_destroy(c);
            return true;
        }
        return false;
    };
    if (bad(hipSetDevice(c->device))) return ASTRO_ERR_HIP;
    int cus = 0;
    if (bad(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device)))
        return ASTRO_ERR_HIP;
    c->num_cus = cus > 0 ? (uint32_t)cus : 1u;
    if (bad(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking))) return ASTRO_ERR_HIP;
    hipError_t e = hipSuccess;
    e = e ? e : dalloc(c->device_bytes, &c->pos, NP);
    e = e ? e : dalloc(c->device_bytes, &c->vel, NP);
    e = e ? e : dalloc(c->device_bytes, &c->color, NP);
    e = e ? e : dalloc(c->device_bytes, &c->force, NP);
    e = e ? e : dalloc(c->device_bytes, &c->type, NP);
    e = e ? e : dalloc(c->device_bytes, &c->sys_of, NP);
    e = e ? e : dalloc(c->device_bytes, &c->codes, NP);
    e = e ? e : dalloc(c->device_bytes, &c->unsorted_codes, NP);
    e = e ? e : dalloc(c->device_bytes, &c->sorted_codes, NP);
    e = e ? e : dalloc(c->device_bytes, &c->tmp_keys_a, NP);
    e = e ? e : dalloc(c->device_bytes, &c->tmp_keys_b, NP);
    e = e ? e : dalloc(c->device_bytes, &c->unsorted_idx, NP);
    e = e ? e : dalloc(c->device_bytes, &c->sorted_idx, NP);
    e = e ? e : dalloc(c->device_bytes, &c->tmp_idx_a, NP);
    e = e ? e : dalloc(c->device_bytes, &c->tmp_idx_b, NP);
    e = e ? e : dalloc(c->device_bytes, &c->unique_idx, NP);
    e = e ? e : dalloc(c->device_bytes, &c->flags, NP);
    e = e ? e : dalloc(c->device_bytes, &c->scan, NP);
    e = e ? e : dalloc(c->device_bytes, &c->ebuf, NP);
    e = e ? e : dalloc(c->device_bytes, &c->fbuf, NP);
    e = e ? e : dalloc(c->device_bytes, &c->parent_count, 1);
    e = e ? e : dalloc(c->device_bytes, &c->child_count, 1);
    e = e ? e : dalloc(c->device_bytes, &c->dev_consts, 4);
    e = e ? e : dalloc(c->device_bytes, &c->tpos, cap_bound);
    e = e ? e : dalloc(c->device_bytes, &c->tlink, cap_bound);
    e = e ? e : dalloc(c->device_bytes, &c->levels, cap_bound);
    e = e ? e : dalloc(c->device_bytes, &c->onodes, cap_bound);
    e = e ? e : dalloc(c->device_bytes, &c->layer_counts, kMaxLayers);
    e = e ? e : dalloc(c->device_bytes, &c->layer_counts_prev, kMaxLayers);
    e = e ? e : dalloc(c->device_bytes, &c->leaf_of_sorted, NT);
    e = e ? e : dalloc(c->device_bytes, &c->pnext, NT);
    e = e ? e : dalloc(c->device_bytes, &c->cubes, c->cfg.max_systems);
    e = e ? e : dalloc(c->device_bytes, &c->minmax,
                       std::max(6ull * c->cfg.max_systems, 8ull * cube_single_blocks(NP)));
    e = e ? e : dalloc(c->device_bytes, &c->cube_counter, 1);
    e = e ? e : dalloc(c->device_bytes, &c->plevel, NP);
    e = e ? e : dalloc(c->device_bytes, &c->fused_buf, fused_buf_words(NP));
    e = e ? e : dalloc(c->device_bytes, &c->err_word, 1);
    e = e ? e : dalloc(c->device_bytes, &c->stats, 4);
    e = e ? e : dalloc(c->device_bytes, &c->d_sys_off, c->cfg.max_systems + 1ull);
    e = e ? e : dalloc(c->device_bytes, &c->phi, NP);
    e = e ? e : dalloc(c->device_bytes, &c->diag_partial, (uint64_t)c->cfg.max_systems * kDiagChunks * 9);
    // the seeded sort: previous order, bucket slots (kResortCap per bucket), splitters, counts
    // + overflow count
    // (batches: at most one bucket more per system)
    c->rs_slots = std::min<uint64_t>(resort_slots(NP) + (uint64_t)c->cfg.max_systems * kResortCap,
                                     (uint64_t)kResortMaxBuckets * kResortCap);
    e = e ? e : dalloc(c->device_bytes, &c->order, NP);
    e = e ? e : dalloc(c->device_bytes, &c->rs_keys, c->rs_slots);
    e = e ? e : dalloc(c->device_bytes, &c->rs_idx, c->rs_slots);
    e = e ? e : dalloc(c->device_bytes, &c->rs_pos, c->rs_slots);
    e = e ? e : dalloc(c->device_bytes, &c->rs_type, c->rs_slots);
    e = e ? e : dalloc(c->device_bytes, &c->spl_keys, kResortMaxBuckets);
    e = e ? e : dalloc(c->device_bytes, &c->spl_idx, kResortMaxBuckets);
    e = e ? e : dalloc(c->device_bytes, &c->rs_count, 2ull * kResortMaxBuckets + 1);
    e = e ? e : dalloc(c->device_bytes, &c->spl_cube, c->cfg.max_systems);
    e = e ? e : dalloc(c->device_bytes, &c->d_sys_b0, c->cfg.max_systems + 1ull);
    e = e ? e : dalloc(c->device_bytes, &c->d_bkt_sys, kResortMaxBuckets);
    e = e ? e : dalloc(c->device_bytes, &c->wlow, 3);
    if (c->cfg.max_systems > 1)
        e = e ? e : dalloc(c->device_bytes, &c->sysfirst, (uint64_t)c->cfg.max_systems * kMaxLayers);
    e = e ? e : dalloc(c->device_bytes, &c->spos, NT);  // the sort's positions and types in
    e = e ? e : dalloc(c->device_bytes, &c->stype, NT); // sorted order
    if (e != hipSuccess) {
        astro_destroy(c);
        return ASTRO_ERR_NOMEM;
    }
    c->scratch_words = scratch_need(NP, cap_max, cfg->bits_per_axis + 1);
    if (bad(dalloc(c->device_bytes, &c->scratch, c->scratch_words))) return ASTRO_ERR_NOMEM;
    hipStream_t s = c->stream;
    launch_fill_u32(c->err_word, 0, 1, s);
    launch_fill_u32(c->cube_counter, 0, 1, s);
    launch_fill_u32(c->wlow, 0, 3, s);
    launch_fill_u32(c->layer_counts, 0, kMaxLayers, s);
    launch_fill_u32(c->layer_counts_prev, 0, kMaxLayers, s);
    launch_fill_u32(c->sys_of, 0, NP, s);
    launch_fill_u32(c->parent_count, 0, 1, s);
    launch_fill_u32(c->child_count, 0, 1, s);
    launch_fill_u32(c->dev_consts, 0, 4, s);
    launch_clear_onodes(c->onodes, cap_bound, s);
    c->onodes_dirty = false;
    if (bad(hipGetLastError()) || bad(hipStreamSynchronize(s))) return ASTRO_ERR_HIP;
    *out = c;
    return ASTRO_OK;
}

extern "C" const char* astro_last_error(const astro_ctx* c) {
    return c ? c->err.c_str() : "null context";
}

extern "C" int astro_set_params(astro_ctx* c, const astro_params* p) {
    if (!c || !p) return ASTRO_ERR_ARG;
    if (p->bits_per_axis == 0 || p->bits_per_axis > c->cfg.bits_per_axis)
        return fail(c, ASTRO_ERR_ARG, "bits_per_axis outside [1, configured maximum]");
    if (p->quantizer > ASTRO_QUANT_BBOX_POW2 || p->side_rule > ASTRO_SIDE_EXACT)
        return fail(c, ASTRO_ERR_ARG, "unknown quantizer or side rule");
    if (p->multipole != 0 && p->multipole != 2)
        return fail(c, ASTRO_ERR_ARG, "multipole must be 0 or 2");
    if (p->side_rule == ASTRO_SIDE_SHIFTED && p->bits_per_axis > 10)
        return fail(c, ASTRO_ERR_ARG, "the shifted side rule overflows above 10 layers");
    if (p->multipole == 2 && !c->qmom) {  // second moments: 32 B per node, kept once made
        hipSetDevice(c->device);
        if (dalloc(c->device_bytes, &c->qmom, 2 * c->node_capacity) != hipSuccess) {
            c->qmom = nullptr;
            return fail(c, ASTRO_ERR_NOMEM, "no device memory for the second moments");
        }
    }
    c->params = *p;
    return refresh_layers(c);
}

extern "C" int astro_get_params(const astro_ctx* c, astro_params* p) {
    if (!c || !p) return ASTRO_ERR_ARG;
    *p = c->params;
    return ASTRO_OK;
}

extern "C" int astro_set_emit_octree_nodes(astro_ctx* c, int mode) {
    if (!c) return ASTRO_ERR_ARG;
    if (mode < ASTRO_EMIT_NEVER || mode > ASTRO_EMIT_ON_READ)
        return fail(c, ASTRO_ERR_ARG, "emission mode must be 0, 1 or 2");
    // a pending array of the current tree is written before the mode changes
    if (int rc = materialize_onodes(c)) return rc;
    c->emit_mode = (uint32_t)mode;
    invalidate_graph(c);
    return ASTRO_OK;
}

extern "C" int astro_set_use_graph(astro_ctx* c, int on) {
    if (!c) return ASTRO_ERR_ARG;
    c->use_graph = on != 0;
    invalidate_graph(c);
    return ASTRO_OK;
}

// ------------------------------------------------------------------------ state transfer
static int set_counts(astro_ctx* c, uint32_t ns, uint32_t nt) {
    c->ns = ns;
    c->nt = nt;
    uint32_t consts[4] = {ns, nt, 0, 0};
    HIPCHK(c, hipMemcpyAsync(c->dev_consts, consts, sizeof(consts), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ASTRO_OK;
}

void astro_mass_bounds(astro_ctx* c, const float* m, uint32_t n, size_t stride, bool reset) {
    if (reset) {
        c->mass_ok = true;
        c->mass_min_pos = INFINITY;
        c->mass_total_max = 0.0;
    }
    double total = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        const float v = m[i * stride];
        if (!(v >= 0.0f) || !std::isfinite(v)) c->mass_ok = false;
        if (v > 0.0f && v < c->mass_min_pos) c->mass_min_pos = v;
        total += v;
    }
    c->mass_total_max = std::max(c->mass_total_max, total);
}

static int upload_particles(astro_ctx* c, const PositionMass* pos, const VelocityRadius* vel,
                            const ColorType* col, uint32_t n) {
    HIPCHK(c, hipSetDevice(c->device));
    // the whole upload's total bounds every system's total
    astro_mass_bounds(c, &pos[0].mass, n, sizeof(PositionMass) / sizeof(float), true);
    std::vector<float4> p4(n), v4(n), c4(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    std::vector<uint32_t> ty(n, 0u);
    for (uint32_t i = 0; i < n; i++) {
        p4[i] = make_float4(pos[i].position[0], pos[i].position[1], pos[i].position[2],
                            pos[i].mass);
        if (vel)
            v4[i] = make_float4(vel[i].velocity[0], vel[i].velocity[1], vel[i].velocity[2],
                                vel[i].radius);
        else
            v4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (col) {
            ty[i] = col[i].type;
            c4[i] = make_float4(col[i].color[0], col[i].color[1], col[i].color[2], col[i].color[3]);
        }
    }
    if (n) {
        HIPCHK(c, hipMemcpyAsync(c->color, c4.data(), n * sizeof(float4), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->pos, p4.data(), n * sizeof(float4), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->vel, v4.data(), n * sizeof(float4), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->type, ty.data(), n * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ASTRO_OK;
}

extern "C" int astro_upload(astro_ctx* c, const PositionMass* pos, const VelocityRadius* vel,
                            const ColorType* col, uint32_t n, uint32_t n_tree) {
    if (!c || !pos) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (n > c->cfg.max_particles) return fail(c, ASTRO_ERR_CAPACITY, "too many particles");
    if (n_tree == 0 || n_tree > n) n_tree = n;
    if (n_tree > c->cfg.max_tree) return fail(c, ASTRO_ERR_CAPACITY, "too many tree sources");
    if (int rc = upload_particles(c, pos, vel, col, n)) return rc;
    launch_fill_u32(c->sys_of, 0, c->cfg.max_particles, c->stream);
    c->nsys = 1;
    c->sys_off = {0u, n};
    HIPCHK(c, hipMemcpyAsync(c->d_sys_off, c->sys_off.data(), 2 * sizeof(uint32_t),
                             hipMemcpyHostToDevice, c->stream));
    int rc = set_counts(c, n, n_tree);
    if (rc) return rc;
    return refresh_layers(c);
}

// Batched layout: system s owns particles [offsets[s], offsets[s+1]); every particle is a
// tree source; each system gets its own bounding cube, key range, tree and root.
extern "C" int astro_upload_systems(astro_ctx* c, const PositionMass* pos,
                                    const VelocityRadius* vel, const ColorType* col,
                                    const uint32_t* offsets, uint32_t n_systems) {
    if (!c || !pos || !offsets || n_systems == 0) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (n_systems > c->cfg.max_systems)
        return fail(c, ASTRO_ERR_CAPACITY, "more systems than the context was created for");
    if (offsets[0] != 0) return fail(c, ASTRO_ERR_ARG, "offsets[0] must be 0");
    for (uint32_t s = 0; s < n_systems; s++)
        if (offsets[s + 1] <= offsets[s])
            return fail(c, ASTRO_ERR_ARG, "every system needs at least one particle");
    const uint32_t n = offsets[n_systems];
    if (n > c->cfg.max_particles || n > c->cfg.max_tree)
        return fail(c, ASTRO_ERR_CAPACITY, "too many particles");
    if (int rc = upload_particles(c, pos, vel, col, n)) return rc;
    c->nsys = n_systems;
    c->sys_off.assign(offsets, offsets + n_systems + 1);
    std::vector<uint32_t> so(n);
    for (uint32_t s = 0; s < n_systems; s++)
        for (uint32_t i = offsets[s]; i < offsets[s + 1]; i++) so[i] = s;
    if (n)
        HIPCHK(c, hipMemcpyAsync(c->sys_of, so.data(), n * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_sys_off, c->sys_off.data(), (n_systems + 1) * sizeof(uint32_t),
                             hipMemcpyHostToDevice, c->stream));
    int rc = set_counts(c, n, n);
    if (rc) return rc;
    return refresh_layers(c);
}

extern "C" int astro_system_count(const astro_ctx* c, uint32_t* n_systems) {
    if (!c || !n_systems) return ASTRO_ERR_ARG;
    *n_systems = c->nsys;
    return ASTRO_OK;
}

extern "C" int astro_particle_count(const astro_ctx* c, uint32_t* n_particles, uint32_t* n_tree) {
    if (!c) return ASTRO_ERR_ARG;
    if (n_particles) *n_particles = c->ns;
    if (n_tree) *n_tree = c->nt;
    return ASTRO_OK;
}

extern "C" int astro_download(astro_ctx* c, PositionMass* pos, VelocityRadius* vel,
                              ColorType* col) {
    if (!c) return ASTRO_ERR_ARG;
    const uint32_t n = c->ns;
    std::vector<float4> p4(n), v4(n), c4(n);
    std::vector<uint32_t> ty(n);
    if (n) {
        HIPCHK(c, hipMemcpyAsync(p4.data(), c->pos, n * sizeof(float4), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c4.data(), c->color, n * sizeof(float4), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(v4.data(), c->vel, n * sizeof(float4), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(ty.data(), c->type, n * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n; i++) {
        if (pos) {
            std::memset(&pos[i], 0, sizeof(PositionMass));
            pos[i].position[0] = p4[i].x;
            pos[i].position[1] = p4[i].y;
            pos[i].position[2] = p4[i].z;
            pos[i].mass = p4[i].w;
        }
        if (vel) {
            std::memset(&vel[i], 0, sizeof(VelocityRadius));
            vel[i].velocity[0] = v4[i].x;
            vel[i].velocity[1] = v4[i].y;
            vel[i].velocity[2] = v4[i].z;
            vel[i].radius = v4[i].w;
        }
        if (col) {
            std::memset(&col[i], 0, sizeof(ColorType));
            col[i].color[0] = c4[i].x;
            col[i].color[1] = c4[i].y;
            col[i].color[2] = c4[i].z;
            col[i].color[3] = c4[i].w;
            col[i].type = ty[i];
        }
    }
    return ASTRO_OK;
}

extern "C" int astro_download_forces(astro_ctx* c, float* force4) {
    if (!c || !force4) return ASTRO_ERR_ARG;
    if (c->ns)
        HIPCHK(c, hipMemcpyAsync(force4, c->force, c->ns * sizeof(float4), hipMemcpyDeviceToHost,
                                 c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ASTRO_OK;
}

extern "C" int astro_get_buffer(astro_ctx* c, int which, void** d_ptr, uint64_t* bytes) {
    if (!c || !d_ptr) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    const uint64_t NP = c->cfg.max_particles, NT = c->cfg.max_tree;
    void* p = nullptr;
    uint64_t b = 0;
    switch (which) {
        case ASTRO_BUF_POSITION_MASS: p = c->pos; b = NP * 16; break;
        case ASTRO_BUF_VELOCITY_RADIUS: p = c->vel; b = NP * 16; break;
        case ASTRO_BUF_TYPE: p = c->type; b = NP * 4; break;
        case ASTRO_BUF_FORCE: p = c->force; b = NP * 16; break;
        case ASTRO_BUF_MORTON_CODES: p = c->codes; b = NP * 8; break;
        case ASTRO_BUF_UNSORTED_CODES: p = c->unsorted_codes; b = NP * 8; break;
        case ASTRO_BUF_UNSORTED_INDICES: p = c->unsorted_idx; b = NP * 4; break;
        case ASTRO_BUF_SORTED_CODES: p = c->sorted_codes; b = NP * 8; break;
        case ASTRO_BUF_SORTED_INDICES: p = c->sorted_idx; b = NP * 4; break;
        case ASTRO_BUF_UNIQUE_INDICES: p = c->unique_idx; b = NP * 4; break;
        case ASTRO_BUF_PARENT_COUNT: p = c->parent_count; b = 4; break;
        case ASTRO_BUF_CHILD_COUNT: p = c->child_count; b = 4; break;
        case ASTRO_BUF_FLAGS: p = c->flags; b = NP * 4; break;
        case ASTRO_BUF_SCAN: p = c->scan; b = NP * 4; break;
        case ASTRO_BUF_OCTREE_NODES:
            if (int rc = materialize_onodes(c)) return rc;
            p = c->onodes;
            b = c->node_capacity * sizeof(OctreeNode);
            break;
        case ASTRO_BUF_TREE_POS:
            if (int rc = materialize_tree(c)) return rc;
            p = c->tpos;
            b = c->node_capacity * 16ull;
            break;
        case ASTRO_BUF_TREE_LINKS:
            if (int rc = materialize_tree(c)) return rc;
            p = c->tlink;
            b = c->node_capacity * sizeof(TLink);
            break;
        case ASTRO_BUF_TREE_MOMENTS:
            if (int rc = materialize_tree(c)) return rc;
            p = c->qmom;
            b = c->qmom ? c->node_capacity * 32ull : 0;
            break;
        case ASTRO_BUF_LAYER_COUNTS: p = c->layer_counts; b = kMaxLayers * 4; break;
        case ASTRO_BUF_LEAF_OF_SORTED: p = c->leaf_of_sorted; b = NT * 4; break;
        case ASTRO_BUF_CUBES: p = c->cubes; b = c->cfg.max_systems * 16ull; break;
        default: return fail(c, ASTRO_ERR_ARG, "unknown buffer id");
    }
    *d_ptr = p;
    if (bytes) *bytes = b;
    return ASTRO_OK;
}

extern "C" int astro_memcpy_d2h(astro_ctx* c, void* host, const void* d_src, uint64_t bytes) {
    if (!c || (!host && bytes) || (!d_src && bytes)) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (bytes)
        HIPCHK(c, hipMemcpyAsync(host, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ASTRO_OK;
}

extern "C" int astro_memcpy_h2d(astro_ctx* c, void* d_dst, const void* host, uint64_t bytes) {
    if (!c || (!host && bytes) || (!d_dst && bytes)) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (bytes)
        HIPCHK(c, hipMemcpyAsync(d_dst, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ASTRO_OK;
}

extern "C" int astro_layer_table(const astro_ctx* c, uint32_t* layers, uint64_t* cap,
                                 uint64_t* off) {
    if (!c || !layers) return ASTRO_ERR_ARG;
    *layers = c->layers;
    for (uint32_t l = 0; l < c->layers; l++) {
        if (cap) cap[l] = c->cap[l];
        if (off) off[l] = c->off[l];
    }
    return ASTRO_OK;
}

extern "C" int astro_sync(astro_ctx* c) {
    if (!c) return ASTRO_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ASTRO_OK;
}

extern "C" int astro_check(astro_ctx* c) {
    if (!c) return ASTRO_ERR_ARG;
    uint32_t w = 0;
    HIPCHK(c, hipMemcpyAsync(&w, c->err_word, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (w) {
        launch_fill_u32(c->err_word, 0, 1, c->stream);
        HIPCHK(c, hipStreamSynchronize(c->stream));
        std::string m = "device error word:";
        if (w & ERR_SPIN) m += " look-back wait bound exceeded;";
        if (w & ERR_STACK) m += " traversal stack overflow;";
        if (w & ERR_CAPACITY) m += " layer capacity exceeded;";
        if (w & ERR_NODE_INDEX) m += " node index outside the tree;";
        return fail(c, ASTRO_ERR_DEVICE, m);
    }
    return ASTRO_OK;
}

// ------------------------------------------------------------------------------- sorting
// Stable LSD sort of count pairs through the context's two temporaries.  kin/iin are not
// modified unless they alias the temporaries.
static int radix_sort(astro_ctx* c, const uint64_t* kin, const uint32_t* iin, uint64_t* kout,
                      uint32_t* iout, uint32_t cap, const uint32_t* d_count,
                      const RadixPlan& plan, const uint32_t* seg_of) {
    hipStream_t s = c->stream;
    if (cap == 0 || plan.npasses == 0) return ASTRO_OK;
    if (radix_work_words(cap) > c->scratch_words)
        return fail(c, ASTRO_ERR_CAPACITY, "sort scratch too small");
    const uint64_t* ksrc = kin;
    const uint32_t* isrc = iin;
    for (uint32_t p = 0; p < plan.npasses; p++) {
        const bool last = p + 1 == plan.npasses;
        uint64_t* kdst = last ? kout : (p % 2 == 0 ? c->tmp_keys_a : c->tmp_keys_b);
        uint32_t* idst = last ? iout : (p % 2 == 0 ? c->tmp_idx_a : c->tmp_idx_b);
        launch_radix_pass(ksrc, isrc, kdst, idst, d_count, cap, plan.pass[p], seg_of,
                          c->scratch, s);
        ksrc = kdst;
        isrc = idst;
    }
    LAUNCHCHK(c);
    return ASTRO_OK;
}

static RadixPlan key_plan(uint32_t key_bits, uint32_t nsys) {
    RadixPlan plan{};
    uint32_t passes = (key_bits + 7) / 8;
    for (uint32_t p = 0; p < passes; p++) plan.pass[plan.npasses++] = {8 * p, 0};
    if (nsys > 1) {
        uint32_t sb = 0;
        while ((1ull << sb) < nsys) sb++;
        for (uint32_t p = 0; p < (sb + 7) / 8; p++) plan.pass[plan.npasses++] = {8 * p, 1};
    }
    return plan;
}

extern "C" int astro_radix_sort(astro_ctx* c, const uint64_t* kin, const uint32_t* iin,
                                uint64_t* kout, uint32_t* iout, uint32_t estimated_len,
                                const uint32_t* d_count) {
    if (!c || !kin || !iin || !kout || !iout || !d_count) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (estimated_len == 0) return ASTRO_OK;  // RadixSort.swift:106
    if (estimated_len > c->cfg.max_particles)
        return fail(c, ASTRO_ERR_CAPACITY, "estimated_len beyond the sorter's max length");
    RadixPlan plan = key_plan(64, 1);
    if ((const void*)kin == (const void*)kout || (const void*)iin == (const void*)iout) {
        // in place: stage through the second temporary first
        launch_copy_u64_count(kin, c->tmp_keys_b, d_count, estimated_len, c->stream);
        launch_copy_u32_count(iin, c->tmp_idx_b, d_count, estimated_len, c->stream);
        kin = c->tmp_keys_b;
        iin = c->tmp_idx_b;
        // 8 passes: pass 0 writes temp a, so temp b is free to hold the input
    }
    return radix_sort(c, kin, iin, kout, iout, estimated_len, d_count, plan, nullptr);
}

extern "C" int astro_scan(astro_ctx* c, const uint32_t* in, uint32_t* out, uint32_t len) {
    if (!c || (!in && len) || (!out && len)) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (len == 0) return ASTRO_OK;
    if (scan_tiles(len) > c->scratch_words)
        return fail(c, ASTRO_ERR_CAPACITY, "scan longer than the context supports");
    launch_scan(in, out, len, c->scratch, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_split_pass(astro_ctx* c, const uint64_t* kin, const uint32_t* iin,
                                uint64_t* kout, uint32_t* iout, uint32_t bit,
                                uint32_t estimated_len, const uint32_t* d_count) {
    if (!c || !kin || !iin || !kout || !iout || !d_count || bit > 63) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (estimated_len == 0) return ASTRO_OK;
    if (estimated_len > c->cfg.max_particles)
        return fail(c, ASTRO_ERR_CAPACITY, "estimated_len beyond the sorter's max length");
    launch_split_prep(kin, c->ebuf, bit, d_count, estimated_len, c->stream);
    int rc = astro_scan(c, c->ebuf, c->fbuf, estimated_len);
    if (rc) return rc;
    launch_split_scatter(kin, iin, kout, iout, bit, c->ebuf, c->fbuf, d_count, estimated_len,
                         c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

// ------------------------------------------------------------------ stage calls (Renderer)
static int need_particles(astro_ctx* c) {
    if (c->ns == 0) return fail(c, ASTRO_ERR_STATE, "no particles uploaded");
    return ASTRO_OK;
}

static void enqueue_cube(astro_ctx* c, uint32_t count) {
    hipStream_t s = c->stream;
    if (c->params.quantizer != ASTRO_QUANT_BBOX_POW2) return;
    if (c->nsys <= 1) {
        launch_cube_single(c->pos, count, c->minmax, c->cube_counter, c->cubes, s);
        return;
    }
    launch_bbox_init(c->minmax, c->nsys, s);
    launch_bbox(c->pos, count, c->sys_of, c->minmax, s);
    launch_cube_finalize(c->minmax, c->cubes, c->nsys, s);
}

extern "C" int astro_generate_morton_codes(astro_ctx* c, uint32_t count) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (int rc = need_particles(c)) return rc;
    if (count > c->nt) return fail(c, ASTRO_ERR_CAPACITY, "count beyond the tree sources");
    enqueue_cube(c, count);
    launch_morton(c->pos, count, c->nsys > 1 ? c->sys_of : nullptr, c->cubes, c->params,
                  c->codes, c->unsorted_codes, nullptr, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_fill_indices(astro_ctx* c, uint32_t* d_buf, uint32_t start, uint32_t end) {
    if (!c || !d_buf) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    launch_iota(d_buf, start, end, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_clear_buffer(astro_ctx* c, void* d_buf, uint32_t count,
                                  uint32_t elem_bytes) {
    if (!c || !d_buf) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (elem_bytes == 8)
        launch_fill_u64(static_cast<uint64_t*>(d_buf), ASTRO_INVALID_MORTON_CODE, count,
                        c->stream);
    else if (elem_bytes == 4)
        launch_fill_u32(static_cast<uint32_t*>(d_buf), 0u, count, c->stream);
    else
        return fail(c, ASTRO_ERR_ARG, "elem_bytes must be 4 or 8");
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_clear_octree_node_buffer(astro_ctx* c, OctreeNode* d_nodes,
                                              uint32_t count) {
    if (!c || !d_nodes) return ASTRO_ERR_ARG;
    launch_clear_onodes(d_nodes, count, c->stream);
    if (d_nodes == c->onodes) {
        c->onodes_dirty = true;
        c->onodes_pending = false;
    }
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_count_unique_morton_codes(astro_ctx* c, int aggregate, uint32_t count) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (count > c->cfg.max_particles) return fail(c, ASTRO_ERR_CAPACITY, "count too large");
    if (count == 0) return ASTRO_OK;
    hipStream_t s = c->stream;
    launch_fill_u32(c->flags, 0, count, s);  // Renderer.swift:572-574
    launch_fill_u32(c->scan, 0, count, s);
    launch_fill_u32(c->unique_idx, 0, count, s);
    launch_mark_uniques(c->sorted_codes, c->flags, count, aggregate ? 1u : 0u, s);
    if (int rc = astro_scan(c, c->flags, c->scan, count)) return rc;
    launch_scatter_uniques(c->flags, c->scan, c->unique_idx, count, s);
    launch_copy_last_scan(c->scan, c->flags, c->parent_count, c->child_count, count, s);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

static uint32_t fp64_from(const astro_ctx* c) {
    const uint32_t nl = c->params.bits_per_axis + 1;
    return c->params.fp64_top_layers >= nl ? 0u : nl - c->params.fp64_top_layers;
}

extern "C" int astro_aggregate_leaf_nodes(astro_ctx* c, uint32_t count) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (int rc = need_particles(c)) return rc;
    if (count > c->nt) return fail(c, ASTRO_ERR_CAPACITY, "count beyond the tree sources");
    launch_aggregate_leaf(c->sorted_codes, c->sorted_idx, c->unique_idx, c->pos, c->parent_count,
                          c->onodes + c->off[0], c->unsorted_codes, c->unsorted_idx, count,
                          c->params.quirk_last_leaf, count, fp64_from(c) == 0, c->stream);
    c->onodes_dirty = true;
    c->onodes_pending = false;
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_aggregate_internal_nodes(astro_ctx* c, uint32_t layer, uint32_t in_off,
                                              uint32_t out_off, uint32_t node_count) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (layer == 0 || layer >= c->layers) return fail(c, ASTRO_ERR_ARG, "layer out of range");
    if ((uint64_t)out_off + node_count > c->node_capacity)
        return fail(c, ASTRO_ERR_CAPACITY, "output range beyond the node buffer");
    launch_aggregate_nodes(c->onodes, c->sorted_codes, c->sorted_idx, c->unique_idx,
                           c->parent_count, in_off, out_off, layer, c->unsorted_codes,
                           c->unsorted_idx, c->child_count, node_count, layer >= fp64_from(c),
                           c->stream);
    c->onodes_dirty = true;
    c->onodes_pending = false;
    LAUNCHCHK(c);
    return ASTRO_OK;
}

// Renderer.swift:714-755.  The per-layer sort of already-ascending codes is kept, as in the
// reference; it is the identity.
extern "C" int astro_build_internal_nodes(astro_ctx* c, uint32_t count) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (int rc = need_particles(c)) return rc;
    for (uint32_t l = 1; l < c->layers; l++) {
        int rc;
        if ((rc = astro_clear_buffer(c, c->sorted_codes, count, 8))) return rc;
        if ((rc = astro_clear_buffer(c, c->sorted_idx, count, 4))) return rc;
        if ((rc = astro_radix_sort(c, c->unsorted_codes, c->unsorted_idx, c->sorted_codes,
                                   c->sorted_idx, (uint32_t)c->cap[l - 1], c->parent_count)))
            return rc;
        if ((rc = astro_count_unique_morton_codes(c, 1, count))) return rc;
        if ((rc = astro_clear_buffer(c, c->unsorted_codes, count, 8))) return rc;
        if ((rc = astro_clear_buffer(c, c->unsorted_idx, count, 4))) return rc;
        if ((rc = astro_aggregate_internal_nodes(c, l, (uint32_t)c->off[l - 1],
                                                 (uint32_t)c->off[l], (uint32_t)c->cap[l])))
            return rc;
    }
    return ASTRO_OK;
}

static ForceArgs force_args(astro_ctx* c) {
    ForceArgs a{};
    a.tpos = c->tpos;
    a.tlink = c->tlink;
    a.qmom = c->qmom;
    a.onodes = c->onodes;
    a.pos = c->pos;
    a.type = c->type;
    a.sorted_idx = c->sorted_idx;
    a.leaf_of_sorted = c->leaf_of_sorted;
    a.sys_of = c->nsys > 1 ? c->sys_of : nullptr;
    a.cubes = c->cubes;
    a.codes = c->codes;
    a.force = c->force;
    a.err = c->err_word;
    a.stats = nullptr;
    a.counts = c->layer_counts;
    a.wlow = c->wlow;
    a.sysfirst = c->nsys > 1 ? c->sysfirst : nullptr;
    a.nsys = c->nsys;
    for (uint32_t l = 0; l < c->layers; l++) a.off[l] = c->off[l];
    a.layers = c->layers;
    a.top = c->force_top ? 1u : 0u;
    a.walk = c->force_walk;
    a.wide = c->force_wide ? 1u : 0u;
    a.off0 = c->off[0];
    a.root_off = c->off[c->layers - 1];
    a.node_total = c->total;
    a.phi = nullptr;
    a.block = c->force_block;
    a.xcd_chunk = c->force_chunk;
    a.ns = c->ns;
    a.nt = c->nt;
    a.p = c->params;
    set_accept_threshold(a);  // needs a.p
    // Domain of the traversal's fast division G M / (r2 r): with every cube within +-L,
    // targets and centres of mass differ by at most 2L per axis, so r2 <= 12 L^2 + eps2 <= 2^30
    // and r2 r <= 2^45; r2 >= max(eps2, r2_floor) >= 2^-30 gives r2 r >= 2^-45; G M lies in
    // [G m_min, G M_total] or is 0.  Targets must be the tree sources (inside the cubes).
    const double L = 8192.0, lo = 0x1p-44, hi = 0x1p44;
    const astro_params& p = c->params;
    const double r2min = std::max((double)p.eps2, (double)p.r2_floor);
    const double G = p.G;
    const bool ok = p.quantizer == ASTRO_QUANT_BBOX_POW2 && c->ns == c->nt && c->mass_ok &&
                    std::isfinite(G) && G > 0.0 && r2min >= 0x1p-30 && p.eps2 <= 0x1p20 &&
                    (c->mass_min_pos == INFINITY || G * c->mass_min_pos >= lo) &&
                    G * c->mass_total_max <= hi && 12.0 * L * L + p.eps2 <= 0x1p30;
    a.div_lmax = ok ? (float)L : 0.0f;
    return a;
}

extern "C" int astro_barnes_hut(astro_ctx* c) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (int rc = materialize_tree(c)) return rc;
    if (int rc = need_particles(c)) return rc;
    if (c->nsys > 1) return fail(c, ASTRO_ERR_STATE, "stage traversal is single-system");
    if (c->params.multipole != 0)
        return fail(c, ASTRO_ERR_STATE,
                    "the stage aggregation keeps no second moments: use astro_compute_forces");
    if (int rc = materialize_onodes(c)) return rc;  // it reads the OctreeNode array
    launch_force_reference(force_args(c), c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

static void enqueue_lighting(astro_ctx* c, uint32_t stars) {
    launch_lighting(c->pos, c->color, c->type, c->ns, stars, c->nsys > 1 ? c->sys_of : nullptr,
                    c->d_sys_off, c->stream);
}

extern "C" int astro_lighting_pass(astro_ctx* c, uint32_t num_stars) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    enqueue_lighting(c, num_stars);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_update_spheres(astro_ctx* c, float dt) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    launch_update(c->pos, c->vel, c->force, c->type, c->ns, dt, c->params, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

// ------------------------------------------------------------------------ the fused path
static TreeArgs tree_args(astro_ctx* c) {
    TreeArgs a{};
    a.pos = c->pos;
    a.spos = c->spos;  // written by the sort that precedes every fused build
    a.sys_of = c->sys_of;
    a.sorted_keys = c->sorted_codes;
    a.sorted_idx = c->sorted_idx;
    a.tpos = c->tpos;
    a.tlink = c->tlink;
    a.levels = c->levels;
    a.qmom = c->qmom;
    a.quad = c->params.multipole == 2 ? 1u : 0u;
    a.onodes = c->emit_mode != ASTRO_EMIT_NEVER ? c->onodes : nullptr;
    a.counts = c->layer_counts;
    a.leaf_of_sorted = c->leaf_of_sorted;
    a.pnext = c->pnext;
    a.err = c->err_word;
    for (uint32_t l = 0; l < c->layers; l++) {
        a.off[l] = c->off[l];
        a.cap[l] = c->cap[l];
    }
    a.layers = c->layers;
    a.bits = c->params.bits_per_axis;
    a.nt = c->nt;
    a.fp64_from = fp64_from(c);
    a.quirk_last_leaf = c->params.quirk_last_leaf;
    a.segmented = c->nsys > 1;
    a.pack = c->total < kLinkPackNodes && !c->force_wide ? 1u : 0u;
    // layers no walk reaches are left unwritten (rebuilt in full when a caller reads the tree):
    // one system, monopoles, the traversal's fast form, no OctreeNode array in the same build
    ForceArgs fa{};
    fa.p = c->params;
    set_accept_threshold(fa);
    a.skip_ok = c->skip_unreached && c->nsys <= 1 && c->params.multipole != 2 &&
                        !c->params.quirk_last_leaf && fa.fast &&
                        c->emit_mode != ASTRO_EMIT_EVERY_BUILD
                    ? 1u
                    : 0u;
    a.acc = fa.acc;
    a.eps2 = c->params.eps2;
    a.side_shifted = c->params.side_rule == ASTRO_SIDE_SHIFTED ? 1u : 0u;
    a.cubes = c->cubes;
    a.wlow = c->wlow;
    a.sysfirst = c->nsys > 1 ? c->sysfirst : nullptr;
    return a;
}

// The seeded sort applies: within its bucket limit (one system, or a batch whose systems each
// get their own buckets).
static bool seeded_sort_applies(const astro_ctx* c) {
    const uint64_t nb = c->rs_nb;
    return c->sort_mode == ASTRO_SORT_SEEDED && c->nt > 0 && nb > 0 &&
           nb * kResortCap <= c->rs_slots && c->nt <= c->cfg.max_particles;
}
// This build's sort is the seeded one: it applies and the previous order is there.
static bool seeded_sort_now(const astro_ctx* c) { return seeded_sort_applies(c) && c->order_valid; }

// The seeded sort computes the keys itself: only the cube here.
static int enqueue_bbox_morton(astro_ctx* c) {
    enqueue_cube(c, c->nt);
    if (!seeded_sort_now(c))
        launch_morton(c->pos, c->nt, c->nsys > 1 ? c->sys_of : nullptr, c->cubes, c->params,
                      c->codes, c->unsorted_codes, c->unsorted_idx, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

static ResortArgs resort_args(astro_ctx* c) {
    ResortArgs r{};
    r.pos = c->pos;
    r.cubes = c->cubes;
    r.params = c->params;
    r.order = c->order;
    r.n = c->nt;
    r.rk = c->rs_keys;
    r.ri = c->rs_idx;
    r.rp = c->rs_pos;
    r.rt = c->rs_type;
    r.spl_k = c->spl_keys;
    r.spl_i = c->spl_idx;
    r.spl_cube = c->spl_cube;
    r.moved_in = c->rs_count;
    r.moved_out = c->rs_count + kResortMaxBuckets;
    r.ovf = c->rs_count + 2 * kResortMaxBuckets;
    r.ok = c->tmp_keys_a;  // the overflow list (n entries at most)
    r.oi = c->tmp_idx_a;
    r.ob = c->tmp_idx_b;
    r.out_k = c->sorted_codes;
    r.out_i = c->sorted_idx;
    r.type = c->type;
    r.spos = c->spos;
    r.stype = c->stype;
    r.nb = c->rs_nb;
    r.nsys = c->nsys > 1 ? c->nsys : 1u;
    r.sys_of = c->sys_of;
    r.sys_off = c->d_sys_off;
    r.sys_b0 = c->d_sys_b0;
    r.bkt_sys = c->d_bkt_sys;
    return r;
}

static int enqueue_sort(astro_ctx* c) {
    if (seeded_sort_now(c)) {
        launch_resort(resort_args(c), c->stream);
        c->last_sort_seeded = true;
        c->keys_pending = true;  // the per-source key arrays: written on read
        LAUNCHCHK(c);
        return ASTRO_OK;
    }
    RadixPlan plan = key_plan(3 * c->params.bits_per_axis, c->nsys);
    if (int rc = radix_sort(c, c->unsorted_codes, c->unsorted_idx, c->sorted_codes, c->sorted_idx,
                            c->nt, c->dev_consts + 1, plan, c->nsys > 1 ? c->sys_of : nullptr))
        return rc;
    c->last_sort_seeded = false;
    c->keys_pending = false;
    if (seeded_sort_applies(c)) {
        // the order and splitters the next sort starts from, and positions in sorted order
        launch_resort_seed(c->sorted_codes, c->sorted_idx, resort_args(c), c->stream);
        c->order_valid = true;
    } else {
        launch_sorted_gather(c->sorted_idx, c->pos, c->type, c->nt, c->spos, c->stype, c->stream);
    }
    LAUNCHCHK(c);
    return ASTRO_OK;
}

// The OctreeNode array of the compact tree (and the emptied stale slots).
// The layer counts become the next build's previous counts (the stale slots it empties) once
// the array is written: in line right after it, beside the traversal after the join.
static void enqueue_emit(astro_ctx* c, hipStream_t es, bool copy_counts = true) {
    TreeArgs a = tree_args(c);
    // beside the traversal: one block per CU, so that the traversal keeps most wave slots (four
    // per CU measured the same)
    launch_emit_onodes(a, c->layer_counts_prev, (es == c->stream ? 4 : 1) * c->num_cus, es);
    if (copy_counts) launch_copy_u32(c->layer_counts, c->layer_counts_prev, kMaxLayers, es);
}

// emit: write the OctreeNode array right after the layers (on the main stream); false leaves
// it to the caller (the pipelined step emits a tree while the next one is being sorted).
// full: every layer written (a caller reads the tree)
static int enqueue_layers(astro_ctx* c, bool emit = true, bool full = false) {
    hipStream_t s = c->stream;
    TreeArgs a = tree_args(c);
    if (full) a.skip_ok = 0;
    // scratch: all layers' supertile sums first (one clear), then per-layer tile counts
    uint64_t nsup = 0, ntile = 0;
    for (uint32_t l = 0; l < c->layers; l++) {
        const uint64_t cand = l == 0 ? c->nt : c->cap[l - 1];
        nsup += layer_super_words(cand);
        ntile += layer_tiles(cand);
    }
    if (nsup + ntile > c->scratch_words)
        return fail(c, ASTRO_ERR_CAPACITY, "layer scratch too small");
    const bool every = c->emit_mode == ASTRO_EMIT_EVERY_BUILD;
    if (every && c->onodes_dirty) {
        launch_clear_onodes(c->onodes, c->node_capacity, s);
        launch_fill_u32(c->layer_counts_prev, 0, kMaxLayers, s);
        c->onodes_dirty = false;
    }
    std::vector<uint32_t*> sup(c->layers), tc(c->layers);
    uint32_t* p = c->scratch;
    for (uint32_t l = 0; l < c->layers; l++) {
        sup[l] = p;
        p += layer_super_words(l == 0 ? c->nt : c->cap[l - 1]);
    }
    for (uint32_t l = 0; l < c->layers; l++) {
        tc[l] = p;
        p += layer_tiles(l == 0 ? c->nt : c->cap[l - 1]);
    }
    const uint32_t l0 = layers_top_first(a);  // layers l0.. in one single-block launch
    // layers 0 .. lf-1 (where nodes hold a few particles each) in the fused launches; the
    // reference quirk keeps the per-layer kernels
    uint32_t lf = 0;
    if (!a.quirk_last_leaf && c->layers > 10)  // 13 at B = 21 (14 runs past the halo)
        lf = std::min(std::min(l0, c->layers - 9), kFusedMaxLayers);
    if (lf < 2) lf = 0;
    // the next pointers of every layer's nodes, before any node is written; the supertile sums
    // and tile counts the layer kernels accumulate start at zero (cleared there too)
    launch_heads_next(a, lf, c->plevel, c->fused_buf, c->scratch, (uint32_t)(nsup + ntile), s);
    // the fused build counts layer lf's heads itself; without fused layers a launch does
    if (lf) launch_fused_layers(a, lf, c->plevel, c->fused_buf, lf < l0 ? tc[lf] : nullptr,
                                lf < l0 ? sup[lf] : nullptr, s);
    else if (lf < l0) launch_layer_count(a, lf, tc[lf], sup[lf], s);
    for (uint32_t l = lf; l < l0; l++) {
        const bool last = l + 1 == l0;
        launch_layer_emit(a, l, tc[l], sup[l], last ? nullptr : tc[l + 1],
                          last ? nullptr : sup[l + 1], s);
    }
    launch_layers_top(a, l0, s);
    if (every && emit) enqueue_emit(c, s);
    // on read: the array of this tree is written when a caller reads it
    c->onodes_pending = c->emit_mode == ASTRO_EMIT_ON_READ;
    c->tree_built = true;
    c->tree_partial = a.skip_ok != 0;
    LAUNCHCHK(c);
    return ASTRO_OK;
}

// On read: the layers the last build left unwritten, by building the same tree again with every
// layer (its inputs, the sorted keys and positions, are unchanged until the next sort).
int materialize_tree(astro_ctx* c) {
    if (!c->tree_partial || !c->tree_built) return ASTRO_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const bool pending = c->onodes_pending;
    if (int rc = enqueue_layers(c, false, true)) return rc;
    c->onodes_pending = pending;
    return ASTRO_OK;
}

static int enqueue_build(astro_ctx* c) {
    int rc;
    if ((rc = enqueue_bbox_morton(c))) return rc;
    if ((rc = enqueue_sort(c))) return rc;
    return enqueue_layers(c, true);
}

// sorted_pos: in a step, right after its sort and build, the targets' positions and types are
// read in sorted order; elsewhere the positions may have changed since the sort.
static int enqueue_forces(astro_ctx* c, unsigned long long* stats, bool sorted_pos = false) {
    // the traversal trusts the tree's child indices (no per-visit bounds check): only a tree
    // built on the current layer table is traversed
    if (!c->tree_built)
        return fail(c, ASTRO_ERR_STATE, "no tree for the current particles: call astro_build_tree");
    ForceArgs a = force_args(c);
    a.stats = stats;
    if (sorted_pos) {
        a.spos = c->spos;
        a.stype = c->stype;
    }
    launch_force_compact(a, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

// On read: writes the OctreeNode array of the last built tree if a build left it pending
// (stream-ordered before whatever the caller enqueues next).
int materialize_onodes(astro_ctx* c) {
    if (!c->onodes_pending) return ASTRO_OK;
    if (int rc = materialize_tree(c)) return rc;  // every layer first
    c->onodes_pending = false;
    if (!c->tree_built) return ASTRO_OK;  // the layer table changed since: nothing to write
    HIPCHK(c, hipSetDevice(c->device));
    if (c->onodes_dirty) {
        launch_clear_onodes(c->onodes, c->node_capacity, c->stream);
        launch_fill_u32(c->layer_counts_prev, 0, kMaxLayers, c->stream);
        c->onodes_dirty = false;
    }
    enqueue_emit(c, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

// On read: the per-source key arrays (MORTON_CODES, UNSORTED_CODES, UNSORTED_INDICES) of the
// last seeded sort, from its sorted keys and indices.
int materialize_keys(astro_ctx* c) {
    if (!c->keys_pending) return ASTRO_OK;
    c->keys_pending = false;
    HIPCHK(c, hipSetDevice(c->device));
    launch_resort_keys(c->sorted_codes, c->sorted_idx, c->nt, c->codes, c->unsorted_codes,
                       c->unsorted_idx, c->stream);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_build_tree(astro_ctx* c) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    return enqueue_build(c);
}

extern "C" int astro_compute_forces(astro_ctx* c) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    return enqueue_forces(c, nullptr);
}

// A stage boundary of a timed step.  While the stream is being captured the record becomes an
// external event node of the graph, which times the replay.
static void record_stage(astro_ctx* c, hipEvent_t e) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    hipStreamIsCapturing(c->stream, &st);
    hipEventRecordWithFlags(e, c->stream,
                            st == hipStreamCaptureStatusActive ? hipEventRecordExternal : 0u);
}

// One step.  ev: optional stage events (ASTRO_STAGE_COUNT + 1).
// The second stream and its events, created on the first step that writes the OctreeNode
// array (outside any capture): contexts that never write it keep one stream, so that many
// contexts still spread over the device's hardware queues.
static int ensure_beside(astro_ctx* c) {
    if (c->emit_mode != ASTRO_EMIT_EVERY_BUILD || !c->emit_beside || c->beside) return ASTRO_OK;
    HIPCHK(c, hipStreamCreateWithFlags(&c->beside, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    return ASTRO_OK;
}

// roctx ranges around the host-side enqueue of each stage of a step (Renderer.swift's debug
// groups, e.g. "Layer 0: Leaf Nodes", Renderer.swift:512,531,716,753), under the stage names
// of astro_step_timed; rocprofv3 --marker-trace shows them (for a captured step, at capture).
struct StageRange {
    explicit StageRange(const char* name) { roctxRangePushA(name); }
    ~StageRange() { roctxRangePop(); }
    StageRange(const StageRange&) = delete;
    StageRange& operator=(const StageRange&) = delete;
};

static int enqueue_step(astro_ctx* c, float dt, hipEvent_t* ev = nullptr) {
    int rc;
    const bool emit = c->emit_mode == ASTRO_EMIT_EVERY_BUILD;
    StageRange step_range("astro.step");
    if (ev) record_stage(c, ev[0]);
    {
        StageRange r("astro.bbox_morton");
        if ((rc = enqueue_bbox_morton(c))) return rc;
    }
    if (ev) record_stage(c, ev[1]);
    {
        StageRange r("astro.sort");
        if ((rc = enqueue_sort(c))) return rc;
    }
    if (ev) record_stage(c, ev[2]);
    // the OctreeNode array only reads the compact tree: written beside the traversal, which
    // is bound by instruction issue and the vector memory pipeline, not by HBM
    const bool beside = emit && c->emit_beside && c->beside;
    {
        StageRange r("astro.build");
        if ((rc = enqueue_layers(c, emit && !beside))) return rc;
    }
    if (beside) {
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->beside, c->ev_fork, 0));
        enqueue_emit(c, c->beside, false);
        HIPCHK(c, hipEventRecord(c->ev_join, c->beside));
    }
    if (ev) record_stage(c, ev[3]);
    {
        StageRange r("astro.force");
        if ((rc = enqueue_forces(c, nullptr, true))) return rc;
    }
    if (ev) record_stage(c, ev[4]);
    StageRange update_range("astro.update");
    // Renderer.swift:537-540: barnesHut, lightingPass, updateSpheres
    if (c->params.lighting_stars) enqueue_lighting(c, c->params.lighting_stars);
    launch_update(c->pos, c->vel, c->force, c->type, c->ns, dt, c->params, c->stream,
                  c->nsys > 1 ? c->sys_of : nullptr, c->guard, c->guard_extent);
    if (beside) {
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
        launch_copy_u32(c->layer_counts, c->layer_counts_prev, kMaxLayers, c->stream);
    }
    if (ev) record_stage(c, ev[5]);
    LAUNCHCHK(c);
    return ASTRO_OK;
}

extern "C" int astro_step(astro_ctx* c, float dt, uint32_t nsteps) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc = ensure_beside(c)) return rc;
    if (c->emit_mode == ASTRO_EMIT_EVERY_BUILD && c->onodes_dirty) {
        // the one-off full clear, outside any graph
        launch_clear_onodes(c->onodes, c->node_capacity, c->stream);
        launch_fill_u32(c->layer_counts_prev, 0, kMaxLayers, c->stream);
        c->onodes_dirty = false;
    }
    if (!c->use_graph) {
        for (uint32_t k = 0; k < nsteps; k++)
            if (int rc = enqueue_step(c, dt)) return rc;
        return ASTRO_OK;
    }
    if (nsteps && seeded_sort_applies(c) && !c->order_valid) {
        // no previous order yet: one step outside the graph sorts from scratch, so that the
        // captured step holds the seeded sort
        if (int rc = enqueue_step(c, dt)) return rc;
        c->onodes_pending = c->emit_mode == ASTRO_EMIT_ON_READ;
        nsteps--;
        invalidate_graph(c);
    }
    if (nsteps && (!c->graph_valid || c->graph_dt != dt)) {
        invalidate_graph(c);
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(c, dt);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(c, ASTRO_ERR_HIP, "graph capture failed");
        e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(c, ASTRO_ERR_HIP, "graph instantiate failed");
        c->graph_valid = true;
        c->graph_dt = dt;
    }
    StageRange r("astro.step (graph replays)");
    for (uint32_t k = 0; k < nsteps; k++) HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
    // each replay built a tree whose OctreeNode array (and, after a seeded sort, key arrays) is
    // left for the next read
    if (nsteps) {
        c->onodes_pending = c->emit_mode == ASTRO_EMIT_ON_READ;
        c->keys_pending = c->last_sort_seeded;
        // tree_partial keeps the captured build's value
    }
    return ASTRO_OK;
}

extern "C" int astro_step_timed(astro_ctx* c, float dt, uint32_t nsteps, float* stage_ms) {
    if (!c || !stage_ms) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (int r = ensure_beside(c)) return r;
    const uint32_t ne = ASTRO_STAGE_COUNT + 1;
    std::vector<hipEvent_t> ev((size_t)nsteps * ne);
    for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
    int rc = ASTRO_OK;
    if (c->emit_mode == ASTRO_EMIT_EVERY_BUILD && c->onodes_dirty) {
        launch_clear_onodes(c->onodes, c->node_capacity, c->stream);
        launch_fill_u32(c->layer_counts_prev, 0, kMaxLayers, c->stream);
        c->onodes_dirty = false;
    }
    hipGraphExec_t x = nullptr;
    if (c->use_graph && nsteps) {
        // the timed steps captured as one graph: astro_step's kernels, replayed the same way,
        // with timing nodes at the stage boundaries
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < nsteps && !rc; k++) rc = enqueue_step(c, dt, &ev[(size_t)k * ne]);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (!rc && e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (g) hipGraphDestroy(g);
        if (!rc && e == hipSuccess) e = hipGraphLaunch(x, c->stream);
        if (!rc && e != hipSuccess) rc = fail(c, ASTRO_ERR_HIP, "timed step graph failed");
    } else {
        for (uint32_t k = 0; k < nsteps && !rc; k++) rc = enqueue_step(c, dt, &ev[(size_t)k * ne]);
    }
    hipError_t se = hipStreamSynchronize(c->stream);
    if (x) hipGraphExecDestroy(x);
    for (uint32_t s = 0; s < ASTRO_STAGE_COUNT; s++) stage_ms[s] = 0.0f;
    if (!rc && se == hipSuccess) {
        for (uint32_t k = 0; k < nsteps; k++) {
            hipEvent_t* E = &ev[(size_t)k * ne];
            for (uint32_t s = 0; s < ASTRO_STAGE_COUNT; s++) {
                float ms = 0.0f;
                hipEventElapsedTime(&ms, E[s], E[s + 1]);
                stage_ms[s] += ms;
            }
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (rc) return rc;
    if (se != hipSuccess) return fail(c, ASTRO_ERR_HIP, hipGetErrorString(se));
    return ASTRO_OK;
}

extern "C" int astro_traversal_stats(astro_ctx* c, astro_stats* out) {
    if (!c || !out) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    launch_fill_u64(reinterpret_cast<uint64_t*>(c->stats), 0, 4, c->stream);
    // the counting traversal writes the same forces as the plain one
    if (int rc = enqueue_forces(c, c->stats)) return rc;
    unsigned long long h[4];
    HIPCHK(c, hipMemcpyAsync(h, c->stats, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->visits = h[0];
    out->leaf_terms = h[1];
    out->monopole_terms = h[2];
    out->opened = h[3];
    return ASTRO_OK;
}

extern "C" int astro_device_bytes(const astro_ctx* c, uint64_t* bytes) {
    if (!c || !bytes) return ASTRO_ERR_ARG;
    *bytes = c->device_bytes;
    return ASTRO_OK;
}

extern "C" int astro_set_force_block(astro_ctx* c, uint32_t block) {
    if (!c) return ASTRO_ERR_ARG;
    if (!force_block_supported(block))
        return fail(c, ASTRO_ERR_ARG, "block must be 64/128/256/512/1024");
    c->force_block = block;
    invalidate_graph(c);
    return ASTRO_OK;
}

extern "C" int astro_set_force_top_cache(astro_ctx* c, int on) {
    if (!c) return ASTRO_ERR_ARG;
    c->force_top = on != 0;
    invalidate_graph(c);
    return ASTRO_OK;
}

extern "C" int astro_set_sort_mode(astro_ctx* c, int mode) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_keys(c)) return rc;
    if (mode != ASTRO_SORT_RADIX && mode != ASTRO_SORT_SEEDED)
        return fail(c, ASTRO_ERR_ARG, "sort mode must be 0 (radix) or 1 (seeded)");
    c->sort_mode = (uint32_t)mode;
    c->order_valid = false;
    invalidate_graph(c);
    return ASTRO_OK;
}

extern "C" int astro_sort_info(astro_ctx* c, int* seeded, uint32_t* overflow) {
    if (!c) return ASTRO_ERR_ARG;
    uint32_t o = 0;
    if (c->last_sort_seeded) {
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipMemcpyAsync(&o, c->rs_count + 2 * kResortMaxBuckets, 4,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (seeded) *seeded = c->last_sort_seeded ? 1 : 0;
    if (overflow) *overflow = o;
    return ASTRO_OK;
}

extern "C" int astro_set_skip_unreached_layers(astro_ctx* c, int on) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = materialize_tree(c)) return rc;
    c->skip_unreached = on != 0;
    invalidate_graph(c);
    return ASTRO_OK;
}

extern "C" int astro_tree_info(astro_ctx* c, uint32_t* lowest_written, int* partial) {
    if (!c) return ASTRO_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    uint32_t w = 0;
    HIPCHK(c, hipMemcpyAsync(&w, c->wlow, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (lowest_written) *lowest_written = c->tree_partial ? w : 0u;
    if (partial) *partial = c->tree_partial && w > 0 ? 1 : 0;
    return ASTRO_OK;
}

extern "C" int astro_set_force_walk(astro_ctx* c, int mode) {
    if (!c) return ASTRO_ERR_ARG;
    if (mode != 0 && mode != 1) return fail(c, ASTRO_ERR_ARG, "walk must be 0 (per lane) or 1 (per wave)");
    c->force_walk = (uint32_t)mode;
    invalidate_graph(c);
    return ASTRO_OK;
}

extern "C" int astro_set_wide_offsets(astro_ctx* c, int on) {
    if (!c) return ASTRO_ERR_ARG;
    c->force_wide = on != 0;
    c->tree_built = false;  // the moment rows of a packed tree no longer match
    invalidate_graph(c);
    return ASTRO_OK;
}

// Times the force traversal over the current tree for every supported block size (median of
// `reps` launches each) and keeps the fastest.  Builds the tree first.
extern "C" int astro_set_force_xcd_chunk(astro_ctx* c, uint32_t chunk) {
    if (!c) return ASTRO_ERR_ARG;
    if (chunk > 4096) return fail(c, ASTRO_ERR_ARG, "xcd_chunk must be 0..4096");
    c->force_chunk = chunk;
    invalidate_graph(c);
    return ASTRO_OK;
}

extern "C" int astro_autotune(astro_ctx* c, uint32_t reps, uint32_t* chosen_block,
                              uint32_t* chosen_chunk, float* best_ms) {
    if (!c) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    if (reps == 0) reps = 3;
    if (int rc = enqueue_build(c)) return rc;
    const uint32_t blocks[5] = {64, 128, 256, 512, 1024};
    const uint32_t chunks[4] = {0, 8, 32, 128};
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    uint32_t best_b = c->force_block, best_k = c->force_chunk;
    float best_t = 1e30f;
    for (uint32_t b : blocks)
        for (uint32_t k : chunks) {
            c->force_block = b;
            c->force_chunk = k;
            if (int rc = enqueue_forces(c, nullptr)) return rc;  // warm
            std::vector<float> ts;
            for (uint32_t r = 0; r < reps; r++) {
                hipEventRecord(e0, c->stream);
                enqueue_forces(c, nullptr);
                hipEventRecord(e1, c->stream);
                HIPCHK(c, hipEventSynchronize(e1));
                float ms = 0.0f;
                hipEventElapsedTime(&ms, e0, e1);
                ts.push_back(ms);
            }
            std::sort(ts.begin(), ts.end());
            const float med = ts[ts.size() / 2];
            if (med < best_t) {
                best_t = med;
                best_b = b;
                best_k = k;
            }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c->force_block = best_b;
    c->force_chunk = best_k;
    invalidate_graph(c);
    if (chosen_block) *chosen_block = best_b;
    if (chosen_chunk) *chosen_chunk = best_k;
    if (best_ms) *best_ms = best_t;
    return ASTRO_OK;
}

// FP64 diagnostics per system: {mass, kinetic, potential, px, py, pz, mx, my, mz} (9 doubles
// per system).  The potential comes from one extra traversal of the current tree (the same
// interactions as the forces: -G M / sqrt(r^2 + eps^2), plus the quadrupole term of accepted
// nodes when the quadrupole option is on).
extern "C" int astro_diagnostics(astro_ctx* c, double* out) {
    if (!c || !out) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    if (int rc = enqueue_build(c)) return rc;
    ForceArgs a = force_args(c);
    a.phi = c->phi;
    launch_force_compact(a, c->stream);
    launch_diag(c->pos, c->vel, c->phi, c->d_sys_off, c->nsys, kDiagChunks, c->diag_partial,
                c->stream);
    LAUNCHCHK(c);
    std::vector<double> part((size_t)c->nsys * kDiagChunks * 9);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->diag_partial, part.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t s = 0; s < c->nsys; s++)
        for (int k = 0; k < 9; k++) {
            double acc = 0.0;
            for (uint32_t ch = 0; ch < kDiagChunks; ch++) acc += part[((size_t)s * kDiagChunks + ch) * 9 + k];
            out[(size_t)s * 9 + k] = acc;
        }
    return ASTRO_OK;
}

// FP64 direct-sum accelerations of the given targets (particle indices) over the sources of
// their systems, for accuracy audits of the tree forces.  Synchronous; allocates its own
// temporaries.
extern "C" int astro_direct_forces_fp64(astro_ctx* c, const uint32_t* targets, uint32_t n_targets,
                                        double* out_xyz) {
    if (!c || (n_targets && (!targets || !out_xyz))) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    if (!n_targets) return ASTRO_OK;
    for (uint32_t t = 0; t < n_targets; t++)
        if (targets[t] >= c->nt) return fail(c, ASTRO_ERR_ARG, "target beyond the tree particles");
    constexpr uint32_t kChunks = 16;
    uint32_t* d_t = nullptr;
    double* d_part = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_t), (size_t)n_targets * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void**>(&d_part), (size_t)n_targets * kChunks * 3 * sizeof(double));
    std::vector<double> part;
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_t, targets, (size_t)n_targets * sizeof(uint32_t), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess) {
        launch_direct_fp64(c->pos, c->nsys > 1 ? c->sys_of : nullptr, c->d_sys_off, c->nt, d_t,
                           n_targets, kChunks, (double)c->params.G, (double)c->params.eps2, d_part,
                           c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        part.resize((size_t)n_targets * kChunks * 3);
        e = hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (d_t) hipFree(d_t);
    if (d_part) hipFree(d_part);
    if (e != hipSuccess)
        return fail(c, ASTRO_ERR_HIP, std::string("direct sum: ") + hipGetErrorString(e));
    for (uint32_t t = 0; t < n_targets; t++)
        for (int k = 0; k < 3; k++) {
            double acc = 0.0;
            for (uint32_t ch = 0; ch < kChunks; ch++) acc += part[((size_t)t * kChunks + ch) * 3 + k];
            out_xyz[(size_t)t * 3 + k] = acc;
        }
    return ASTRO_OK;
}

// Checks the traversal's fast division against the IEEE division on n pseudo-random pairs of
// its domain; *mismatches receives the number of differing results (0 expected).  Synchronous.
extern "C" int astro_check_division(astro_ctx* c, uint64_t n, uint64_t seed, uint64_t* mismatches) {
    if (!c || !mismatches) return ASTRO_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->stats);
    launch_fill_u64(reinterpret_cast<uint64_t*>(d), 0ull, 1, c->stream);
    if (check_division(n, seed, d, c->stream)) return fail(c, ASTRO_ERR_HIP, "division check launch");
    unsigned long long h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *mismatches = h;
    return ASTRO_OK;
}

// Checks the traversal's r2 opening test against RN(sqrt(r2)) >= rmin0 2^l on n pseudo-random
// cases near the threshold; *mismatches receives the number of disagreements (0 expected).
extern "C" int astro_check_open_test(astro_ctx* c, uint64_t n, uint64_t seed,
                                     uint64_t* mismatches) {
    if (!c || !mismatches) return ASTRO_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->stats);
    launch_fill_u64(reinterpret_cast<uint64_t*>(d), 0ull, 1, c->stream);
    if (check_open_test(n, seed, d, c->stream))
        return fail(c, ASTRO_ERR_HIP, "opening-test check launch");
    unsigned long long h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *mismatches = h;
    return ASTRO_OK;
}

// Checks the traversal's correctly rounded square root on every float in [2^-96, FLT_MAX];
// *checked receives how many, *mismatches the disagreements with the double square root
// rounded to float (0 expected).  Synchronous.
extern "C" int astro_check_sqrt(astro_ctx* c, uint64_t* checked, uint64_t* mismatches) {
    if (!c || !checked || !mismatches) return ASTRO_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->stats);
    launch_fill_u64(reinterpret_cast<uint64_t*>(d), 0ull, 1, c->stream);
    const uint64_t n = check_sqrt(d, c->stream);
    if (!n) return fail(c, ASTRO_ERR_HIP, "square-root check launch");
    unsigned long long h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *checked = n;
    *mismatches = h;
    return ASTRO_OK;
}

// Checks the refined reciprocal behind the traversal's fast division on every float in
// [2^-45, 2^46); *checked receives how many, *mismatches the disagreements with the double
// reciprocal rounded to float (0 expected).  Synchronous.
extern "C" int astro_check_reciprocal(astro_ctx* c, uint64_t* checked, uint64_t* mismatches) {
    if (!c || !checked || !mismatches) return ASTRO_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->stats);
    launch_fill_u64(reinterpret_cast<uint64_t*>(d), 0ull, 1, c->stream);
    const uint64_t n = check_reciprocal(d, c->stream);
    if (!n) return fail(c, ASTRO_ERR_HIP, "reciprocal check launch");
    unsigned long long h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *checked = n;
    *mismatches = h;
    return ASTRO_OK;
}

// ------------------------------------------------------------------------- snapshots
namespace {
struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};
int read_header(FILE* f, astro_snapshot_header* h) {
    if (std::fread(h, sizeof(*h), 1, f) != 1) return ASTRO_ERR_ARG;
    if (std::memcmp(h->magic, "ASTROSNP", 8) != 0 || h->version != ASTRO_SNAPSHOT_VERSION ||
        h->n_systems == 0 || h->n_tree > h->n_particles)
        return ASTRO_ERR_ARG;
    return ASTRO_OK;
}
}  // namespace

extern "C" int astro_snapshot_info(const char* path, astro_snapshot_header* out) {
    if (!path || !out) return ASTRO_ERR_ARG;
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return ASTRO_ERR_ARG;
    return read_header(fc.f, out);
}

extern "C" int astro_snapshot_write(astro_ctx* c, const char* path, double time) {
    if (!c || !path) return ASTRO_ERR_ARG;
    if (int rc = need_particles(c)) return rc;
    const uint32_t n = c->ns;
    std::vector<PositionMass> pos(n);
    std::vector<VelocityRadius> vel(n);
    std::vector<ColorType> col(n);
    if (int rc = astro_download(c, pos.data(), vel.data(), col.data())) return rc;
    astro_snapshot_header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, "ASTROSNP", 8);
    h.version = ASTRO_SNAPSHOT_VERSION;
    h.n_particles = n;
    h.n_tree = c->nt;
    h.n_systems = c->nsys;
    h.time = time;
    h.params = c->params;
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return fail(c, ASTRO_ERR_ARG, std::string("cannot open ") + path);
    const bool ok = std::fwrite(&h, sizeof(h), 1, fc.f) == 1 &&
                    std::fwrite(c->sys_off.data(), sizeof(uint32_t), c->nsys + 1, fc.f) ==
                        c->nsys + 1 &&
                    std::fwrite(pos.data(), sizeof(PositionMass), n, fc.f) == n &&
                    std::fwrite(vel.data(), sizeof(VelocityRadius), n, fc.f) == n &&
                    std::fwrite(col.data(), sizeof(ColorType), n, fc.f) == n;
    if (!ok) return fail(c, ASTRO_ERR_ARG, std::string("short write to ") + path);
    return ASTRO_OK;
}

extern "C" int astro_snapshot_read(astro_ctx* c, const char* path, double* time) {
    if (!c || !path) return ASTRO_ERR_ARG;
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return fail(c, ASTRO_ERR_ARG, std::string("cannot open ") + path);
    astro_snapshot_header h;
    if (read_header(fc.f, &h)) return fail(c, ASTRO_ERR_ARG, "not a snapshot of this version");
    const uint32_t n = h.n_particles;
    std::vector<uint32_t> off(h.n_systems + 1ull);
    std::vector<PositionMass> pos(n);
    std::vector<VelocityRadius> vel(n);
    std::vector<ColorType> col(n);
    const bool ok = std::fread(off.data(), sizeof(uint32_t), off.size(), fc.f) == off.size() &&
                    std::fread(pos.data(), sizeof(PositionMass), n, fc.f) == n &&
                    std::fread(vel.data(), sizeof(VelocityRadius), n, fc.f) == n &&
                    std::fread(col.data(), sizeof(ColorType), n, fc.f) == n;
    if (!ok) return fail(c, ASTRO_ERR_ARG, "truncated snapshot");
    if (int rc = astro_set_params(c, &h.params)) return rc;
    int rc;
    if (h.n_systems == 1)
        rc = astro_upload(c, pos.data(), vel.data(), col.data(), n, h.n_tree);
    else
        rc = astro_upload_systems(c, pos.data(), vel.data(), col.data(), off.data(), h.n_systems);
    if (rc) return rc;
    if (time) *time = h.time;
    return ASTRO_OK;
}

// ------------------------------------------------------------- shared with serve.hip
int astro_refresh_layers(astro_ctx* c) { return refresh_layers(c); }
int astro_set_counts(astro_ctx* c, uint32_t ns, uint32_t nt) { return set_counts(c, ns, nt); }
int astro_enqueue_step(astro_ctx* c, float dt) { return enqueue_step(c, dt); }
