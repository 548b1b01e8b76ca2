/*
 * astro.h -- C-ABI of libastro, the MI355X (gfx950) Barnes-Hut step.
 *
 * Conventions
 *   - every function returns an int status: ASTRO_OK (0) or a negative ASTRO_ERR_*;
 *     astro_last_error() gives the message of the last failure on that context;
 *   - no exceptions cross this boundary; no torch or HIP types appear in it;
 *   - pointers named d_* are device pointers (from astro_get_buffer or the caller's own
 *     device allocations); all others are host pointers;
 *   - every call only enqueues work on the context's stream unless it says otherwise;
 *     variable counts stay in device memory, as in the reference (Renderer.swift:518-519,
 *     RadixSort.metal:117,140, Octree.metal:65,136,143);
 *   - a context is used from one thread at a time (the reference drives one command queue
 *     from one thread, Renderer.swift:117); separate contexts are independent.
 */
#ifndef ASTRO_ASTRO_H
#define ASTRO_ASTRO_H

#include "bridge.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ASTRO_OK 0
#define ASTRO_ERR_ARG (-1)      /* invalid argument                                    */
#define ASTRO_ERR_HIP (-2)      /* a HIP runtime call failed                           */
#define ASTRO_ERR_CAPACITY (-3) /* size beyond what the context was created for        */
#define ASTRO_ERR_STATE (-4)    /* call out of order (e.g. no particles uploaded)      */
#define ASTRO_ERR_DEVICE (-5)   /* a kernel raised its error word (see astro_check)    */
#define ASTRO_ERR_NOMEM (-6)    /* device allocation failed / HBM budget exceeded      */
#define ASTRO_ERR_BUSY (-7)     /* the server sheds load: the request would wait too long */

typedef struct astro_ctx astro_ctx;

/* Sizing of a context: the equivalent of Renderer.makeBuffers (Renderer.swift:336-391) plus
 * the sorter temporaries (RadixSort.swift:40-67), allocated once. */
typedef struct {
    uint32_t max_particles;     /* N_s: particles integrated (force targets)              */
    uint32_t max_tree;          /* N_t: tree sources (the first N_t particles); 0 = N_s   */
    uint32_t bits_per_axis;     /* largest B that will be used (tree capacity)           */
    uint32_t max_systems;       /* independent systems per batched step (>= 1)           */
    uint32_t emit_octree_nodes; /* when the 128-B OctreeNode array is written: one of    */
                                /* ASTRO_EMIT_NEVER, _EVERY_BUILD, _ON_READ              */
    int32_t device;             /* HIP device ordinal                                    */
} astro_config;

/* When the OctreeNode array (bridge.h:47-65, the reference's node buffer) is written.  The
 * traversal reads the compact nodes, so the array is an output for callers only:
 *   NEVER       not written by the fused build (the stage calls still write it);
 *   EVERY_BUILD written by every build, in a step beside the traversal (on a second stream);
 *   ON_READ     written, with the same bytes, when a caller asks for it: astro_get_buffer of
 *               ASTRO_BUF_OCTREE_NODES or astro_barnes_hut (default). */
enum { ASTRO_EMIT_NEVER = 0, ASTRO_EMIT_EVERY_BUILD = 1, ASTRO_EMIT_ON_READ = 2 };

/* How a step or astro_build_tree sorts the keys (the result is the same either way: ascending
 * key, ties by source index, RadixSort.swift:98-185; a batch by system first):
 *   RADIX   an LSD radix sort from scratch, three launches per 8-bit digit;
 *   SEEDED  buckets between splitters sampled from the previous sort's order, each sorted in
 *           LDS (default; after a first radix sort of the current sources, while the buckets
 *           number at most 8,192: one system of at most 8,028,160 sources, or a batch, whose
 *           systems get ceil(n_s / 980) buckets each; otherwise RADIX). */
enum { ASTRO_SORT_RADIX = 0, ASTRO_SORT_SEEDED = 1 };

/* Buffers reachable through astro_get_buffer (names follow Renderer.swift:56-70). */
enum {
    ASTRO_BUF_POSITION_MASS = 0,   /* float4 {x,y,z,m} per particle                       */
    ASTRO_BUF_VELOCITY_RADIUS = 1, /* float4 {vx,vy,vz,radius}                            */
    ASTRO_BUF_TYPE = 2,            /* uint32 type per particle                            */
    ASTRO_BUF_FORCE = 3,           /* float4 {fx,fy,fz,0}                                 */
    ASTRO_BUF_MORTON_CODES = 4,    /* uint64 per tree source (unsorted, per particle)     */
    ASTRO_BUF_UNSORTED_CODES = 5,  /* uint64                                              */
    ASTRO_BUF_UNSORTED_INDICES = 6,/* uint32                                              */
    ASTRO_BUF_SORTED_CODES = 7,    /* uint64                                              */
    ASTRO_BUF_SORTED_INDICES = 8,  /* uint32                                              */
    ASTRO_BUF_UNIQUE_INDICES = 9,  /* uint32                                              */
    ASTRO_BUF_PARENT_COUNT = 10,   /* uint32[1]                                           */
    ASTRO_BUF_CHILD_COUNT = 11,    /* uint32[1]                                           */
    ASTRO_BUF_FLAGS = 12,          /* uint32                                              */
    ASTRO_BUF_SCAN = 13,           /* uint32                                              */
    ASTRO_BUF_OCTREE_NODES = 14,   /* OctreeNode[sum of layer capacities]                 */
    ASTRO_BUF_TREE_POS = 15,       /* float4 {x,y,z,m} per compact traversal node         */
    ASTRO_BUF_LAYER_COUNTS = 16,   /* uint32[layers]: valid nodes per layer, last build   */
    ASTRO_BUF_LEAF_OF_SORTED = 17, /* uint32 leaf index of each sorted source             */
    ASTRO_BUF_CUBES = 18,          /* float4 {lo.x, lo.y, lo.z, side} per system          */
    ASTRO_BUF_TREE_LINKS = 19,     /* uint32 {last child, info, next, start} per compact  */
                                   /* node (info = layer | nchild << 5 | head level << 9) */
    ASTRO_BUF_TREE_MOMENTS = 20,   /* with multipole = 2: float4 {Sxx,Syy,Szz,Sxy},       */
                                   /* {Sxz,Syz, last | layer << 27, next} per compact     */
                                   /* node (the links only below 2^27 nodes; leaves'      */
                                   /* last 0x07FFFFFF)                                    */
    ASTRO_BUF_COUNT
};

/* Traversal statistics, summed over targets (for the work figures of a step). */
typedef struct {
    uint64_t visits;
    uint64_t leaf_terms;
    uint64_t monopole_terms;
    uint64_t opened;
} astro_stats;

/* ---- lifetime and parameters ---------------------------------------------------------- */
int astro_create(const astro_config* cfg, astro_ctx** out);
int astro_destroy(astro_ctx* ctx);
const char* astro_last_error(const astro_ctx* ctx);
int astro_device_count(int* count);
void astro_params_compat(astro_params* p);                    /* the reference's constants */
void astro_params_physics(astro_params* p, uint32_t bits_per_axis);
int astro_set_params(astro_ctx* ctx, const astro_params* p);
int astro_get_params(const astro_ctx* ctx, astro_params* p);
/* When the OctreeNode array is written (ASTRO_EMIT_*; default: the config value).  A pending
 * array of the current tree is written first. */
int astro_set_emit_octree_nodes(astro_ctx* ctx, int mode);
/* 1: astro_step replays one captured hipGraph per step (default). */
int astro_set_use_graph(astro_ctx* ctx, int on);

/* ---- state transfer (reference byte layouts; replaces makeSpheres' memcpy,
 *      Renderer.swift:490-492) -------------------------------------------------------- */
int astro_upload(astro_ctx* ctx, const PositionMass* pos, const VelocityRadius* vel,
                 const ColorType* col, uint32_t n_particles, uint32_t n_tree);
/* Batched layout (no reference counterpart: the reference steps one system per frame):
 * system s owns particles [offsets[s], offsets[s+1]), all of them tree sources; each system
 * gets its own bounding cube, keys, tree and root, so results do not depend on the batch. */
int astro_upload_systems(astro_ctx* ctx, const PositionMass* pos, const VelocityRadius* vel,
                         const ColorType* col, const uint32_t* offsets, uint32_t n_systems);
int astro_system_count(const astro_ctx* ctx, uint32_t* n_systems);
/* Particles of the current state (the length astro_download writes) and tree sources. */
int astro_particle_count(const astro_ctx* ctx, uint32_t* n_particles, uint32_t* n_tree);
int astro_download(astro_ctx* ctx, PositionMass* pos, VelocityRadius* vel, ColorType* col);
int astro_download_forces(astro_ctx* ctx, float* force4); /* float3 in 16-B slots */
int astro_get_buffer(astro_ctx* ctx, int which, void** d_ptr, uint64_t* bytes);
int astro_memcpy_d2h(astro_ctx* ctx, void* host, const void* d_src, uint64_t bytes);
int astro_memcpy_h2d(astro_ctx* ctx, void* d_dst, const void* host, uint64_t bytes);
int astro_layer_table(const astro_ctx* ctx, uint32_t* layers, uint64_t* cap, uint64_t* off);
int astro_sync(astro_ctx* ctx);
int astro_check(astro_ctx* ctx); /* syncs; ASTRO_ERR_DEVICE if a kernel raised an error */

/* ---- stage calls: the Renderer / RadixSort dispatch surface --------------------------- */
/* Renderer.swift:556-568 -> MortonCode.metal:26-52 (writes morton and unsorted codes) */
int astro_generate_morton_codes(astro_ctx* ctx, uint32_t count);
/* Renderer.swift:394-409 -> Octree.metal:39-51 */
int astro_fill_indices(astro_ctx* ctx, uint32_t* d_buf, uint32_t start, uint32_t end);
/* Renderer.swift:411-427 -> Octree.metal:239-251: u32 <- 0 (elem 4), u64 <- INVALID (8) */
int astro_clear_buffer(astro_ctx* ctx, void* d_buf, uint32_t count, uint32_t elem_bytes);
/* Renderer.swift:429-439 -> Octree.metal:253-269 */
int astro_clear_octree_node_buffer(astro_ctx* ctx, OctreeNode* d_nodes, uint32_t count);
/* RadixSort.swift:98-185: stable ascending sort of (u64 key, u32 index) pairs; the number
 * of pairs is read from d_count (<= estimated_len) on the device. */
int astro_radix_sort(astro_ctx* ctx, const uint64_t* d_keys_in, const uint32_t* d_idx_in,
                     uint64_t* d_keys_out, uint32_t* d_idx_out, uint32_t estimated_len,
                     const uint32_t* d_count);
/* RadixSort.swift:225-279: exclusive scan of len u32 values */
int astro_scan(astro_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint32_t len);
/* astro_scan without a context and without allocating: the same kernels and results, with
 * their per-tile sums in a fixed device array of the library (64 KiB per GPU, kept until the
 * process ends).  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses that forbid it).
 * Runs on the current device and on `stream` (a hipStream_t; NULL: the null stream);
 * asynchronous.  len <= 33,554,432 (16,384 tiles of 2,048), else ASTRO_ERR_CAPACITY.  Calls
 * that overlap on one GPU share the array and race: order them on one stream. */
int astro_scan_noalloc(const uint32_t* d_in, uint32_t* d_out, uint32_t len, void* stream);
/* RadixSort.swift:312-349: one stable 1-bit split pass */
int astro_split_pass(astro_ctx* ctx, const uint64_t* d_keys_in, const uint32_t* d_idx_in,
                     uint64_t* d_keys_out, uint32_t* d_idx_out, uint32_t bit,
                     uint32_t estimated_len, const uint32_t* d_count);
/* Renderer.swift:570-625 */
int astro_count_unique_morton_codes(astro_ctx* ctx, int aggregate, uint32_t count);
/* Renderer.swift:654-672 -> Octree.metal:53-117 */
int astro_aggregate_leaf_nodes(astro_ctx* ctx, uint32_t count);
/* Renderer.swift:674-712 -> Octree.metal:120-202 */
int astro_aggregate_internal_nodes(astro_ctx* ctx, uint32_t layer, uint32_t input_offset,
                                   uint32_t output_offset, uint32_t node_count);
/* Renderer.swift:714-755 */
int astro_build_internal_nodes(astro_ctx* ctx, uint32_t count);
/* Renderer.swift:851-869 -> BarnesHut.metal:22-101, over the OctreeNode array */
int astro_barnes_hut(astro_ctx* ctx);
/* Renderer.swift:871-883 -> BarnesHut.metal:103-167: planets and dust take the light of the
 * first min(num_stars, 5) particles (of their system); colours only */
int astro_lighting_pass(astro_ctx* ctx, uint32_t num_stars);
/* Renderer.swift:885-899 -> BarnesHut.metal:169-196 */
int astro_update_spheres(astro_ctx* ctx, float dt);

/* ---- the fused MI355X path ------------------------------------------------------------ */
/* Renderer.swift:511-534: bbox, Morton, LSD sort, layered build (lower layers fused). */
int astro_build_tree(astro_ctx* ctx);
/* Renderer.swift:536-538 over the compact tree: forces for every particle.  Needs a tree built
 * (astro_build_tree or a step) since the last change of the layer table (an upload of another
 * particle count or system layout, a new bits_per_axis): ASTRO_ERR_STATE otherwise. */
int astro_compute_forces(astro_ctx* ctx);
/* Renderer.swift:495-542 with explicit dt: build + forces (+ the lighting pass when
 * astro_params.lighting_stars > 0) + update, nsteps times. */
int astro_step(astro_ctx* ctx, float dt, uint32_t nsteps);
/* Same as astro_step with hipEvents around each stage of each step; stage_ms receives the
 * per-stage sums (ASTRO_STAGE_COUNT floats).  With graphs on (the default) the nsteps steps are
 * captured as one hipGraph whose event nodes time the stages, and replayed once. */
enum {
    ASTRO_STAGE_BBOX_MORTON = 0,
    ASTRO_STAGE_SORT = 1,
    ASTRO_STAGE_BUILD = 2,
    ASTRO_STAGE_FORCE = 3,
    ASTRO_STAGE_UPDATE = 4,
    ASTRO_STAGE_COUNT = 5
};
int astro_step_timed(astro_ctx* ctx, float dt, uint32_t nsteps, float* stage_ms);
/* Work figures of the current tree and positions (one extra counting traversal; needs a built
 * tree, as astro_compute_forces). */
int astro_traversal_stats(astro_ctx* ctx, astro_stats* out);
/* Traversal block size (64, 128, 256, 512 or 1024; default 256). */
int astro_set_force_block(astro_ctx* ctx, uint32_t block);
/* With blocks of 512 or 1,024 threads, each block first copies the top layers of the tree (as
 * many as fit in its LDS, chosen from the layer counts on the device; in a batch, of the tree of
 * its middle target's system) into LDS, and the walks read them there (default on; the results
 * are the same bits either way).  Replaces nothing in
 * the reference: BarnesHut.metal:22-101 reads every node from device memory. */
int astro_set_force_top_cache(astro_ctx* ctx, int on);
/* Check switch: build and traverse every tree as if it held 2^27 nodes or more (64-bit node
 * offsets, no links in the moment rows), so that path can be compared with the oracle on small
 * trees.  Takes effect at the next build.  No counterpart in the reference. */
int astro_set_wide_offsets(astro_ctx* ctx, int on);
/* Traversal form: 0 (default) one walk per target; 1 one walk per wave, which visits the union of
 * its 64 targets' walks in the walks' order with scalar node loads, each target taking part at
 * its own walk's nodes only (same forces, bit for bit).  The second applies to one system with
 * monopoles and the fast form of the opening test (force computations only); elsewhere the
 * first runs.  No counterpart in the reference. */
int astro_set_force_walk(astro_ctx* ctx, int mode);
/* With one system, monopoles and the traversal's fast form (eps2 >= the floors), a build does not
 * write the tree layers that no walk can reach: those whose every node passes the opening test
 * for any finite target (RN(sqrt(eps2)) >= rmin0 2^l).  The lowest written layer's nodes get
 * 0xFFFFFFFF as last child (a walk of a non-finite target ends there).  Reading the tree
 * (astro_get_buffer of ASTRO_BUF_TREE_POS, _TREE_LINKS, _TREE_MOMENTS, _OCTREE_NODES, or
 * astro_barnes_hut) first builds the same tree again with every layer.  Default on; off writes
 * every layer in every build.  No counterpart in the reference, which writes every layer. */
int astro_set_skip_unreached_layers(astro_ctx* ctx, int on);
/* The lowest tree layer the last build wrote (0: all) and whether layers are left out until the
 * tree is read; synchronises the stream. */
int astro_tree_info(astro_ctx* ctx, uint32_t* lowest_written, int* partial);
/* Key sort of steps and builds: ASTRO_SORT_RADIX or ASTRO_SORT_SEEDED (see above).  Replaces
 * nothing in the reference, whose every build sorts from scratch. */
int astro_set_sort_mode(astro_ctx* ctx, int mode);
/* The last sort's form (1: seeded, 0: radix) and, for a seeded sort, its sources that did not fit
 * in their bucket's slots (sorted on the slow path); synchronises the stream. */
int astro_sort_info(astro_ctx* ctx, int* seeded, uint32_t* overflow);
/* How target tiles are dealt to the 8 XCDs: 0 (default) gives each XCD one contiguous eighth
 * of the Morton curve; K > 0 deals runs of K consecutive tiles to the XCDs in turn. */
int astro_set_force_xcd_chunk(astro_ctx* ctx, uint32_t chunk);
/* Builds the tree, times the traversal for every block size x XCD chunk (median of reps
 * launches each) and keeps the fastest; returns the choice and its time. */
int astro_autotune(astro_ctx* ctx, uint32_t reps, uint32_t* chosen_block, uint32_t* chosen_chunk,
                   float* best_ms);
/* Builds the tree and returns, per system, 9 FP64 sums: mass, kinetic energy, potential
 * energy (1/2 sum m phi, phi from the same interactions as the forces, quadrupole terms
 * included when astro_params.multipole is 2), momentum x/y/z and mass-weighted position
 * x/y/z. */
int astro_diagnostics(astro_ctx* ctx, double* out);
/* FP64 direct-sum accelerations (3 doubles per target) of the given particles over all tree
 * particles of their system, with the context's G and eps2: the accuracy audit for the tree
 * forces (the reference has no counterpart; its README states no error bound).  Synchronous;
 * allocates temporaries. */
int astro_direct_forces_fp64(astro_ctx* ctx, const uint32_t* targets, uint32_t n_targets,
                             double* out_xyz);
/* Self-check of the traversal's fast division (used when every quotient G M / (r^2 r) lies
 * in [2^-90, 2^90]) against the IEEE division, on n pseudo-random operand pairs of its domain;
 * *mismatches receives the count of differing results (0 expected).  No reference
 * counterpart.  Synchronous. */
int astro_check_division(astro_ctx* ctx, uint64_t n, uint64_t seed, uint64_t* mismatches);
/* Self-check of the traversal's opening test in squared form (used with quadrupoles when the
 * fast division holds and eps2 >= 2^-96): a node at layer l opens unless
 * RN(sqrt(r2)) >= rmin0 2^l, which
 * is evaluated as r2 >= c0 4^l with c0 the smallest float whose rounded square root reaches
 * rmin0.  Compares both forms on n pseudo-random cases near the threshold; *mismatches
 * receives the count of disagreements (0 expected).  No reference counterpart.  Synchronous. */
int astro_check_open_test(astro_ctx* ctx, uint64_t n, uint64_t seed, uint64_t* mismatches);
/* Exhaustive self-check of the traversal's correctly rounded square roots (the hardware
 * estimate plus a one-ulp correction, and the hardware reciprocal square root plus one
 * Goldschmidt step) on every float in [2^-96, FLT_MAX]: *checked receives the count,
 * *mismatches the disagreements of either with the double-precision square root rounded to
 * float (0 expected).  No reference counterpart.  Synchronous, about 0.1 s. */
int astro_check_sqrt(astro_ctx* ctx, uint64_t* checked, uint64_t* mismatches);
/* Exhaustive self-check of the reciprocal behind the traversal's fast division (the hardware
 * estimate refined by one Newton step, which the division then uses once, by Markstein's
 * theorem) on every float in [2^-45, 2^46): *checked receives the count, *mismatches the
 * disagreements with the double-precision reciprocal rounded to float (0 expected).  No
 * reference counterpart.  Synchronous. */
int astro_check_reciprocal(astro_ctx* ctx, uint64_t* checked, uint64_t* mismatches);
/* Device memory the context allocated at creation. */
int astro_device_bytes(const astro_ctx* ctx, uint64_t* bytes);

/* ---- continuous-batching server (no reference counterpart: the reference renders one
 *      system).  Requests wait FIFO; at each tick the server retires finished systems,
 *      evicts failed ones, admits queued ones while the batch fits its HBM budget and
 *      capacities, and enqueues one batched step (a captured graph while the batch is
 *      unchanged).  The host waits for the device only when a request finishes and every
 *      16 ticks.  A request's result does not depend on what it was batched with. ---------- */
typedef struct astro_server astro_server;
typedef struct {
    uint32_t max_particles;  /* particles the batch holds at once                         */
    uint32_t max_systems;    /* systems the batch holds at once                           */
    uint64_t hbm_budget;     /* device bytes admitted requests may occupy (0: no limit)   */
    int32_t device;
    float max_extent;        /* > 0: a system with a particle outside [-max_extent,       */
                             /* max_extent]^3 fails; a non-finite state always fails       */
    double max_wait_s;       /* > 0: submit refuses (ASTRO_ERR_BUSY) a request whose       */
                             /* estimated wait before its last step exceeds this          */
} astro_server_config;
enum { ASTRO_REQUEST_DONE = 0, ASTRO_REQUEST_FAILED = 1 };
typedef struct {
    uint64_t id;
    uint32_t n, steps;
    double t_submit, t_admit, t_done; /* seconds on the server's monotonic clock */
    uint32_t status;                  /* ASTRO_REQUEST_DONE / ASTRO_REQUEST_FAILED */
    uint32_t steps_done;              /* steps run (steps when done)               */
} astro_completion;
typedef struct {
    uint32_t admitted, completed, active, queued, particles, active_after;
    double step_ms;                   /* host time of the tick                      */
    uint32_t failed;                  /* requests evicted by this tick              */
    uint32_t synced;                  /* 1: the tick waited for the device          */
} astro_tick_info;
/* p->lighting_stars must be 0: requests carry positions and velocities only. */
int astro_server_create(const astro_server_config* cfg, const astro_params* p,
                        astro_server** out);
int astro_server_destroy(astro_server* s);
const char* astro_server_last_error(const astro_server* s);
int astro_server_device_bytes(const astro_server* s, uint64_t* bytes, uint64_t* per_particle);
int astro_server_submit(astro_server* s, uint64_t id, const PositionMass* pos,
                        const VelocityRadius* vel, const ColorType* col, uint32_t n,
                        uint32_t steps);
/* Errors: a request that can never fit is refused with ASTRO_ERR_CAPACITY, one with a
 * non-finite position, mass or velocity with ASTRO_ERR_ARG, one that would wait longer than
 * max_wait_s with ASTRO_ERR_BUSY (load shedding).  A request whose state turns non-finite or
 * leaves max_extent is evicted (status FAILED, its state at eviction kept for fetch) while
 * the others continue unchanged; a device error word fails every active request and the tick
 * returns ASTRO_ERR_DEVICE, the server stays usable.  Steps are counted as enqueued. */
int astro_server_tick(astro_server* s, float dt, astro_tick_info* info);
/* Waits for the device and settles every enqueued tick (finished and failed requests). */
int astro_server_sync(astro_server* s, astro_tick_info* info);
int astro_server_poll(astro_server* s, astro_completion* out, uint32_t max, uint32_t* n_out);
int astro_server_fetch(astro_server* s, uint64_t id, PositionMass* pos, VelocityRadius* vel);
int astro_server_counters(const astro_server* s, uint64_t* ticks, uint64_t* particle_steps,
                          uint32_t* queued, uint32_t* active);

/* ---- snapshots: the state of a context in the reference byte layout, for checkpoint and
 *      resume (no reference counterpart: the reference keeps no state between runs).
 *      File: this header, then uint32 offsets[n_systems + 1], PositionMass[n],
 *      VelocityRadius[n], ColorType[n], all little-endian. ---------------------------------- */
#define ASTRO_SNAPSHOT_VERSION 1u
typedef struct {
    char magic[8];            /* "ASTROSNP"                                                  */
    uint32_t version;         /* ASTRO_SNAPSHOT_VERSION                                      */
    uint32_t n_particles;
    uint32_t n_tree;          /* tree sources (the first n_tree particles)                   */
    uint32_t n_systems;       /* 1 for astro_upload, S for astro_upload_systems              */
    double time;              /* the caller's simulation time                                */
    astro_params params;
} astro_snapshot_header;
/* Synchronises and writes the context's particles, systems and parameters. */
int astro_snapshot_write(astro_ctx* ctx, const char* path, double time);
/* Replaces the context's parameters and particles by the snapshot's; *time (optional)
 * receives the stored time.  Stepping on reproduces the run that wrote it bit for bit. */
int astro_snapshot_read(astro_ctx* ctx, const char* path, double* time);
/* Host only: reads and checks a snapshot's header. */
int astro_snapshot_info(const char* path, astro_snapshot_header* out);

/* ---- seeded initial conditions (host only; replace Renderer.makeSpheres,
 *      Renderer.swift:441-493, whose draws are unseeded) -------------------------------- */
int astro_ic_plummer(uint32_t n, uint64_t seed, PositionMass* pos, VelocityRadius* vel,
                     ColorType* col);
int astro_ic_two_disks(uint32_t n_per_disk, uint64_t seed, float separation, float pericentre,
                       PositionMass* pos, VelocityRadius* vel, ColorType* col);
int astro_ic_plummer_background(uint32_t n, uint64_t seed, float bg_fraction, float cube_side,
                                float slope, float m_lo, float m_hi, PositionMass* pos,
                                VelocityRadius* vel, ColorType* col);
int astro_ic_batch_sizes(uint32_t n_systems, uint32_t n_min, uint32_t n_max, uint64_t seed,
                         uint32_t* sizes);

#ifdef __cplusplus
}
#endif
#endif /* ASTRO_ASTRO_H */
