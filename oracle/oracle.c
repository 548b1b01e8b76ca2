/*
 * oracle.c -- CPU restatement of astro's Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, never by the product library.  Built with -ffp-contract=off, so each
 * float operation below rounds exactly once, as written.
 *
 * Where the reference leaves rounding open (it is compiled with MTL_FAST_MATH=YES,
 * Astro.xcodeproj/project.pbxproj:328,386), this file fixes one IEEE order and says so at
 * the site; those choices are the contract the HIP kernels reproduce.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INVALID_CODE ASTRO_INVALID_MORTON_CODE
#define INVALID_CHILD ASTRO_INVALID_CHILD

/* Reference constants: BarnesHut.metal:39-40,59,61,70,100,182-190; Renderer.swift:864. */
void orc_params_compat(astro_params* p) {
    memset(p, 0, sizeof(*p));
    p->bits_per_axis = 6;       /* MortonCode.metal:43-47 keeps 18 key bits */
    p->quantizer = ASTRO_QUANT_OFFSET_WRAP;
    p->quant_offset = 1000000.0f;
    p->side_rule = ASTRO_SIDE_SHIFTED;
    p->theta = 0.7f;
    p->eps2 = 0.01f;
    p->G = 1e11f;
    p->r2_floor = 1e-6f;
    p->r_floor = 1e-6f;
    p->leaf_cutoff_r2 = 50.0f;
    p->force_clamp = 1e5f;
    p->damping = 0.9995f;
    p->type_scale[0] = 5.0f;
    p->type_scale[1] = 5.0f;
    p->type_scale[2] = 1.0f;
    p->quirk_last_leaf = 1;
    p->fp64_top_layers = 0;
    p->kahan = 0;
    p->multipole = 0;
    p->lighting_stars = 5;      /* Renderer.swift:26 (numStars) */
}

/* The scaled profile for self-gravitating systems (G=1, exact cells, no cut-offs). */
void orc_params_physics(astro_params* p, uint32_t bits_per_axis) {
    orc_params_compat(p);
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

/* MortonCode.metal:7-23 -- bit i of x/y/z goes to bit 3i/3i+1/3i+2 of the key. */
uint64_t orc_interleave(uint32_t x, uint32_t y, uint32_t z) {
    x &= 0x1FFFFF;
    y &= 0x1FFFFF;
    z &= 0x1FFFFF;
    uint64_t r = 0;
    for (int i = 0; i < 21; i++) {
        r |= ((uint64_t)((x >> i) & 1u)) << (3 * i);
        r |= ((uint64_t)((y >> i) & 1u)) << (3 * i + 1);
        r |= ((uint64_t)((z >> i) & 1u)) << (3 * i + 2);
    }
    return r;
}

/* float -> u32 as the GPU's v_cvt_u32_f32 does it: NaN and negatives to 0, saturating at
 * 2^32-1, truncation otherwise.  MSL leaves out-of-range conversions open
 * (MortonCode.metal:35-37); this is the definition both sides use. */
static uint32_t cvt_u32_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

/* Power-of-two cube enclosing the points: cube = {lo.x, lo.y, lo.z, side}.  side is the
 * smallest power of two >= the largest axis extent (1 when all points coincide). */
void orc_bbox_cube(const PositionMass* pos, uint32_t n, float cube[4]) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            float v = pos[i].position[a];
            lo[a] = v < lo[a] ? v : lo[a];
            hi[a] = v > hi[a] ? v : hi[a];
        }
    if (n == 0) lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0.0f;
    float ext = 0.0f;
    for (int a = 0; a < 3; a++) {
        float e = hi[a] - lo[a];
        ext = e > ext ? e : ext;
    }
    float side = 1.0f;
    if (ext > 0.0f && isfinite(ext)) {
        int e;
        float m = frexpf(ext, &e); /* ext = m * 2^e, m in [0.5, 1) */
        side = ldexpf(1.0f, m == 0.5f ? e - 1 : e);
    }
    cube[0] = lo[0];
    cube[1] = lo[1];
    cube[2] = lo[2];
    cube[3] = side;
}

/* Exponent e of a power-of-two side (side = 2^e). */
static int pow2_exponent(float side) {
    int e;
    frexpf(side, &e);
    return e - 1;
}

/* Bbox quantizer: q = trunc((p - lo) * 2^21 / side), clamped to [0, 2^21-1].  The scale is
 * a power of two, so the product is exact. */
static uint32_t quantize_bbox(float v, float lo, float scale) {
    float t = (v - lo) * scale;
    uint32_t q = cvt_u32_sat(t);
    return q > 0x1FFFFFu ? 0x1FFFFFu : q;
}

/* MortonCode.metal:26-52 (generateMortonCodes), generalised to B bits per axis: quantize to
 * 21 bits per axis, interleave, keep the top 3B bits. */
void orc_morton_codes(const PositionMass* pos, uint32_t n, const astro_params* p,
                      const float cube[4], uint64_t* codes) {
    const int shift = 63 - 3 * (int)p->bits_per_axis;
    float scale = 1.0f;
    if (p->quantizer == ASTRO_QUANT_BBOX_POW2) scale = ldexpf(1.0f, 21 - pow2_exponent(cube[3]));
    for (uint32_t i = 0; i < n; i++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            if (p->quantizer == ASTRO_QUANT_OFFSET_WRAP)
                q[a] = cvt_u32_sat(pos[i].position[a] + p->quant_offset);
            else
                q[a] = quantize_bbox(pos[i].position[a], cube[a], scale);
        }
        codes[i] = orc_interleave(q[0], q[1], q[2]) >> shift;
    }
}

/* RadixSort.metal:30-85 + RadixSort.swift:225-279 -- exclusive scan, output[0] = 0.  The
 * reference's block scheme is exact only up to 2^20 elements (one 1024-wide aux block);
 * this is the result it computes there, for every length. */
void orc_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t len) {
    uint32_t acc = 0;
    for (uint32_t i = 0; i < len; i++) {
        uint32_t v = in[i];
        out[i] = acc;
        acc += v;
    }
}

/* RadixSort.metal:89-131 (split_prep, scan, split_scatter) -- one stable 1-bit partition of
 * count (key, index) pairs: keys with the bit clear keep their order at the front. */
void orc_split_pass(const uint64_t* kin, const uint32_t* iin, uint64_t* kout, uint32_t* iout,
                    uint32_t bit, uint32_t count) {
    if (count == 0) return;
    uint32_t* e = (uint32_t*)malloc(sizeof(uint32_t) * count);
    uint32_t* f = (uint32_t*)malloc(sizeof(uint32_t) * count);
    for (uint32_t i = 0; i < count; i++) e[i] = (kin[i] & (1ull << bit)) == 0;
    orc_exclusive_scan(e, f, count);
    uint32_t total_falses = f[count - 1] + e[count - 1];
    for (uint32_t i = 0; i < count; i++) {
        int set = (kin[i] & (1ull << bit)) != 0;
        uint32_t dst = set ? total_falses + (i - f[i]) : f[i];
        kout[dst] = kin[i];
        iout[dst] = iin[i];
    }
    free(e);
    free(f);
}

/* RadixSort.swift:98-185 -- 64 split passes, ping-pong, result in place. */
void orc_split_sort(uint64_t* keys, uint32_t* idx, uint32_t count) {
    if (count == 0) return;
    uint64_t* k2 = (uint64_t*)malloc(sizeof(uint64_t) * count);
    uint32_t* i2 = (uint32_t*)malloc(sizeof(uint32_t) * count);
    uint64_t *ka = keys, *kb = k2;
    uint32_t *ia = idx, *ib = i2;
    for (uint32_t bit = 0; bit < 64; bit++) {
        orc_split_pass(ka, ia, kb, ib, bit, count);
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* ti = ia; ia = ib; ib = ti;
    }
    /* 64 passes: the result is back in the caller's arrays */
    free(k2);
    free(i2);
}

/* Renderer.swift:570-625 = markUniques (Octree.metal:282-303) + scan + scatterUniques
 * (306-318) + copyLastScanToParentCount (321-334).  Returns the number of segment heads. */
uint32_t orc_count_unique(const uint64_t* sorted, uint32_t count, int aggregate,
                          uint32_t* flags, uint32_t* scan, uint32_t* unique_idx) {
    if (count == 0) return 0;
    for (uint32_t g = 0; g < count; g++) {
        if (sorted[g] == INVALID_CODE) {
            flags[g] = 0;
            continue;
        }
        int u;
        if (g == 0) u = 1;
        else if (aggregate) u = (sorted[g] >> 3) != (sorted[g - 1] >> 3);
        else u = sorted[g] != sorted[g - 1];
        flags[g] = u ? 1u : 0u;
    }
    orc_exclusive_scan(flags, scan, count);
    for (uint32_t g = 0; g < count; g++)
        if (flags[g]) unique_idx[scan[g]] = g;
    return scan[count - 1] + flags[count - 1];
}

/* Renderer.swift:373-386 generalised to B bits per axis: B+1 layers, cap_l = min(nt, 8^(B-l)). */
void orc_layer_table(uint32_t nt, uint32_t bits_per_axis, orc_layers* L) {
    memset(L, 0, sizeof(*L));
    L->layers = bits_per_axis + 1;
    uint64_t off = 0;
    for (uint32_t l = 0; l < L->layers; l++) {
        uint32_t bits = 3 * bits_per_axis - 3 * l;
        uint64_t possible = bits >= 63 ? (1ull << 63) : (1ull << bits);
        uint64_t cap = possible < nt ? possible : nt;
        L->cap[l] = cap;
        L->off[l] = off;
        off += cap;
    }
    L->total = off;
}

static void empty_node(OctreeNode* n) {
    /* Octree.metal:253-269 (clearOctreeNodeBuffer) */
    memset(n, 0, sizeof(*n));
    n->mortonCode = INVALID_CODE;
    for (int c = 0; c < 8; c++) n->children[c] = INVALID_CHILD;
    n->layer = 0;
}

/* Aggregate one node from a run of (position, mass) terms.  FP32: mass += m,
 * com = fma(x, m, com), COM = com / mass (0 when mass is 0) -- Octree.metal:84-94,157-172
 * with the reference's multiply-then-add fused into one rounding.  FP64 (the top layers):
 * the same sums in double, rounded to float once at the end. */
typedef struct {
    int fp64;
    float m, c[3];
    double md, cd[3];
} agg_t;

static void agg_init(agg_t* a, int fp64) {
    memset(a, 0, sizeof(*a));
    a->fp64 = fp64;
}
static void agg_add(agg_t* a, const float x[3], float m) {
    if (a->fp64) {
        a->md += (double)m;
        for (int k = 0; k < 3; k++) a->cd[k] = fma((double)x[k], (double)m, a->cd[k]);
    } else {
        a->m = a->m + m;
        for (int k = 0; k < 3; k++) a->c[k] = fmaf(x[k], m, a->c[k]);
    }
}
static void agg_final(const agg_t* a, float com[3], float* mass) {
    if (a->fp64) {
        *mass = (float)a->md;
        for (int k = 0; k < 3; k++) com[k] = a->md > 0.0 ? (float)(a->cd[k] / a->md) : 0.0f;
    } else {
        *mass = a->m;
        for (int k = 0; k < 3; k++) com[k] = a->m > 0.0f ? a->c[k] / a->m : 0.0f;
    }
}

/* Second moments S = sum m d d^T about the node's (float) centre of mass P, from terms at x
 * with mass m and their own moments Sc (zero for particles): with d = x - P,
 * S_ij = S_ij + fma(m*d_i, d_j, Sc_ij), in the order xx, yy, zz, xy, xz, yz.  FP64 layers do
 * the same in double and round once.  S is kept in the node's emittedColor (xx, yy, zz, xy)
 * and emittedColorCenter (xz, yz) slots, which the physics otherwise never reads. */
static const int SI[6] = {0, 1, 2, 0, 0, 1}, SJ[6] = {0, 1, 2, 1, 2, 2};
typedef struct {
    int fp64;
    float s[6];
    double sd[6];
} mom_t;
static void mom_init(mom_t* q, int fp64) {
    memset(q, 0, sizeof(*q));
    q->fp64 = fp64;
}
static void mom_add(mom_t* q, const float x[3], float m, const float P[3], const float Sc[6]) {
    if (q->fp64) {
        double d[3];
        for (int k = 0; k < 3; k++) d[k] = (double)x[k] - (double)P[k];
        for (int e = 0; e < 6; e++)
            q->sd[e] = q->sd[e] + fma((double)m * d[SI[e]], d[SJ[e]], Sc ? (double)Sc[e] : 0.0);
    } else {
        float d[3], md[3];
        for (int k = 0; k < 3; k++) d[k] = x[k] - P[k];
        for (int k = 0; k < 3; k++) md[k] = m * d[k];
        for (int e = 0; e < 6; e++)
            q->s[e] = q->s[e] + fmaf(md[SI[e]], d[SJ[e]], Sc ? Sc[e] : 0.0f);
    }
}
static void mom_store(const mom_t* q, OctreeNode* nd) {
    float s[6];
    for (int e = 0; e < 6; e++) s[e] = q->fp64 ? (float)q->sd[e] : q->s[e];
    nd->emittedColor[0] = s[0];
    nd->emittedColor[1] = s[1];
    nd->emittedColor[2] = s[2];
    nd->emittedColor[3] = s[3];
    nd->emittedColorCenter[0] = s[4];
    nd->emittedColorCenter[1] = s[5];
}
static void mom_load(const OctreeNode* nd, float s[6]) {
    s[0] = nd->emittedColor[0];
    s[1] = nd->emittedColor[1];
    s[2] = nd->emittedColor[2];
    s[3] = nd->emittedColor[3];
    s[4] = nd->emittedColorCenter[0];
    s[5] = nd->emittedColorCenter[1];
}

/*
 * Renderer.swift:511-534 + 714-755 -- the layered leaf-to-root build:
 *   generateMortonCodes, fillIndices, sort, countUnique, aggregateLeafNodes, then for each
 *   layer l = 1..B: sort (an identity on already-ascending codes), countUnique(aggregate),
 *   aggregateNodes.
 * nodes must hold L->total entries.  codes receives the per-particle (unsorted) keys.
 * resort_layers=1 runs the per-layer identity sorts as the reference does and returns -2 if
 * any of them is not the identity (they never are: see the aggregateNodes note below).
 * emittedColor/emittedColorCenter are left zero (the physics never reads them), or hold the
 * node's second moments when p->multipole == 2.
 */
int orc_build_tree(const PositionMass* pos, uint32_t nt, const astro_params* p,
                   const float cube[4], int resort_layers, OctreeNode* nodes, orc_layers* L,
                   uint64_t* codes, uint64_t* sorted_codes, uint32_t* sorted_idx) {
    orc_layer_table(nt, p->bits_per_axis, L);
    for (uint64_t i = 0; i < L->total; i++) empty_node(&nodes[i]);
    if (nt == 0) return 0;
    const uint32_t nl = L->layers;
    const uint32_t fp64_from = p->fp64_top_layers >= nl ? 0 : nl - p->fp64_top_layers;

    uint32_t* flags = (uint32_t*)malloc(sizeof(uint32_t) * nt);
    uint32_t* scan = (uint32_t*)malloc(sizeof(uint32_t) * nt);
    uint32_t* uniq = (uint32_t*)malloc(sizeof(uint32_t) * nt);
    uint64_t* layer_codes = (uint64_t*)malloc(sizeof(uint64_t) * nt);
    uint32_t* layer_idx = (uint32_t*)malloc(sizeof(uint32_t) * nt);
    uint64_t* sk = (uint64_t*)malloc(sizeof(uint64_t) * nt);

    /* leaf layer: MortonCode.metal:26-52, Octree.metal:39-51, RadixSort.swift:98-185 */
    orc_morton_codes(pos, nt, p, cube, codes);
    for (uint32_t i = 0; i < nt; i++) {
        sorted_codes[i] = codes[i];
        sorted_idx[i] = i;
    }
    orc_split_sort(sorted_codes, sorted_idx, nt);
    uint32_t parent = orc_count_unique(sorted_codes, nt, 0, flags, scan, uniq);

    /* Octree.metal:53-117 (aggregateLeafNodes) */
    for (uint32_t g = 0; g < parent; g++) {
        uint32_t start = uniq[g];
        uint32_t end = g + 1 < parent ? uniq[g + 1] : (p->quirk_last_leaf ? parent : nt);
        agg_t a;
        agg_init(&a, 0 >= fp64_from);
        for (uint32_t i = start; i < end; i++) {
            const PositionMass* s = &pos[sorted_idx[i]];
            agg_add(&a, s->position, s->mass);
        }
        OctreeNode* nd = &nodes[L->off[0] + g];
        agg_final(&a, nd->centerOfMass, &nd->totalMass);
        if (p->multipole == 2) {
            mom_t q;
            mom_init(&q, 0 >= fp64_from);
            for (uint32_t i = start; i < end; i++) {
                const PositionMass* s = &pos[sorted_idx[i]];
                mom_add(&q, s->position, s->mass, nd->centerOfMass, NULL);
            }
            mom_store(&q, nd);
        }
        nd->mortonCode = sorted_codes[start];
        nd->layer = 0;
        layer_codes[g] = sorted_codes[start];
        layer_idx[g] = g;
    }
    L->count[0] = parent;

    /* Renderer.swift:714-755 (buildInternalNodes) with Octree.metal:120-202 */
    int status = 0;
    for (uint32_t l = 1; l < nl; l++) {
        uint32_t child = parent;
        for (uint32_t i = 0; i < nt; i++) sk[i] = INVALID_CODE;
        for (uint32_t i = 0; i < child; i++) sk[i] = layer_codes[i];
        if (resort_layers) {
            orc_split_sort(sk, layer_idx, child);
            for (uint32_t i = 0; i < child; i++)
                if (layer_idx[i] != i) status = -2;
        }
        parent = orc_count_unique(sk, nt, 1, flags, scan, uniq);
        for (uint32_t g = 0; g < parent; g++) {
            uint32_t start = uniq[g];
            uint32_t end = g + 1 < parent ? uniq[g + 1] : child;
            agg_t a;
            agg_init(&a, l >= fp64_from);
            OctreeNode* nd = &nodes[L->off[l] + g];
            int nc = 0;
            for (uint32_t i = start; i < end; i++) {
                /* sortedIndices is the identity here, so child i sits at offset + i */
                uint64_t ci = L->off[l - 1] + i;
                const OctreeNode* c = &nodes[ci];
                agg_add(&a, c->centerOfMass, c->totalMass);
                if (nc < 8) nd->children[nc++] = (uint32_t)ci;
            }
            agg_final(&a, nd->centerOfMass, &nd->totalMass);
            if (p->multipole == 2) {
                mom_t q;
                mom_init(&q, l >= fp64_from);
                for (uint32_t i = start; i < end; i++) {
                    const OctreeNode* c = &nodes[L->off[l - 1] + i];
                    float Sc[6];
                    mom_load(c, Sc);
                    mom_add(&q, c->centerOfMass, c->totalMass, nd->centerOfMass, Sc);
                }
                mom_store(&q, nd);
            }
            nd->mortonCode = sk[start] >> 3;
            nd->layer = l;
            layer_codes[g] = nd->mortonCode;
            layer_idx[g] = g;
        }
        L->count[l] = parent;
    }
    free(flags);
    free(scan);
    free(uniq);
    free(layer_codes);
    free(layer_idx);
    free(sk);
    return status;
}

/* Node side length: BarnesHut.metal:62 (compat) or the exact cell size of the layer. */
static float node_side(const astro_params* p, uint32_t layer, float cube_side) {
    if (p->side_rule == ASTRO_SIDE_SHIFTED) return ldexpf(1.0f, 21 + (int)layer);
    return ldexpf(cube_side, (int)layer - (int)p->bits_per_axis);
}

/*
 * BarnesHut.metal:22-101 (barnesHut) for targets [t_begin, t_end).
 * Per target: LIFO DFS from the root; children pushed 0..7 (so child 7 pops first).
 *   d = com - x;  r2 = fma(dz,dz, fma(dy,dy, dx*dx)) + eps2;  r2 = max(r2, r2_floor);
 *   r = max(sqrt(r2), r_floor)
 *   leaf:      add unless it is the target's own leaf (non-dust targets) or r2 <= cutoff
 *   internal:  add if side / r < theta, else open
 *   add:       s = (G*M) / (r2*r);  F = fma(s, d, F)     (Kahan: compensated)
 * then clamp each component to +-force_clamp.  The reference's g*M*d/(r2*r) under fast
 * math has no pinned rounding; the order above is the one fixed here.
 */
void orc_barnes_hut(const OctreeNode* nodes, uint64_t root, const PositionMass* pos,
                    const ColorType* types, const uint64_t* codes, uint32_t ns, uint32_t nt,
                    const astro_params* p, float cube_side, uint32_t t_begin, uint32_t t_end,
                    float* force4, orc_stats* st_out, int threads) {
    uint64_t visits = 0, leaf_terms = 0, mono_terms = 0, opened = 0;
    uint32_t max_stack = 0;
    if (t_end > ns) t_end = ns;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : visits, leaf_terms, mono_terms, opened) reduction(max : max_stack)
#endif
    for (int64_t ii = (int64_t)t_begin; ii < (int64_t)t_end; ii++) {
        uint32_t i = (uint32_t)ii;
        uint32_t stack[1024];
        uint32_t sp = 0;
        stack[sp++] = (uint32_t)root;
        float F[3] = {0.0f, 0.0f, 0.0f}, C[3] = {0.0f, 0.0f, 0.0f};
        const float px = pos[i].position[0], py = pos[i].position[1], pz = pos[i].position[2];
        const uint32_t type = types ? types[i].type : 0u;
        uint64_t mycode = INVALID_CODE;
        int has_self = (type != 2u) && (i < nt);
        if (has_self) mycode = codes[i];
        while (sp > 0) {
            uint32_t ni = stack[--sp];
            const OctreeNode* nd = &nodes[ni];
            visits++;
            if (nd->mortonCode == INVALID_CODE) continue;
            float dx = nd->centerOfMass[0] - px;
            float dy = nd->centerOfMass[1] - py;
            float dz = nd->centerOfMass[2] - pz;
            float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) + p->eps2;
            r2 = fmaxf(r2, p->r2_floor);
            float r = fmaxf(sqrtf(r2), p->r_floor);
            int leaf = nd->layer == 0;
            int interact;
            if (leaf) {
                int self = has_self && mycode == nd->mortonCode;
                interact = !self && r2 > p->leaf_cutoff_r2;
                leaf_terms += interact;
            } else {
                float side = node_side(p, nd->layer, cube_side);
                interact = (side / r) < p->theta;
                if (interact) {
                    mono_terms++;
                } else {
                    opened++;
                    for (int c = 0; c < 8; c++) {
                        uint32_t ch = nd->children[c];
                        if (ch != INVALID_CHILD && sp < 1024) stack[sp++] = ch;
                    }
                    if (sp > max_stack) max_stack = sp;
                }
            }
            if (interact) {
                float gm = p->G * nd->totalMass;
                float D = r2 * r;
                float s = gm / D;
                float sd[3] = {0.0f, 0.0f, 0.0f}, u = 0.0f, c = s;
                int quad = p->multipole == 2 && !leaf;
                if (quad) {
                    /* a = s d + 3G v (w d - (S d + T/2 d)) = c d - u S d with
                     * c = s + u (w - T/2), u = 3G v, v = 1/r^5 = (y y) r with y = 1/D,
                     * w = 2.5 (d.S.d) v D */
                    float S[6];
                    mom_load(nd, S);
                    float Sd0 = fmaf(S[4], dz, fmaf(S[3], dy, S[0] * dx));
                    float Sd1 = fmaf(S[5], dz, fmaf(S[1], dy, S[3] * dx));
                    float Sd2 = fmaf(S[2], dz, fmaf(S[5], dy, S[4] * dx));
                    float dSd = fmaf(dz, Sd2, fmaf(dy, Sd1, dx * Sd0));
                    float T = (S[0] + S[1]) + S[2];
                    float y = 1.0f / D;
                    float v = (y * y) * r; /* v D = 1/r2 to within a few roundings */
                    float w = (2.5f * dSd) * (v * D);
                    u = (3.0f * p->G) * v;
                    c = fmaf(u, fmaf(-0.5f, T, w), s);
                    sd[0] = Sd0;
                    sd[1] = Sd1;
                    sd[2] = Sd2;
                }
                float d[3] = {dx, dy, dz};
                if (p->kahan) {
                    for (int k = 0; k < 3; k++) {
                        float y = quad ? fmaf(c, d[k], fmaf(-u, sd[k], -C[k])) : fmaf(s, d[k], -C[k]);
                        float t = F[k] + y;
                        C[k] = (t - F[k]) - y;
                        F[k] = t;
                    }
                } else {
                    for (int k = 0; k < 3; k++)
                        F[k] = quad ? fmaf(c, d[k], fmaf(-u, sd[k], F[k])) : fmaf(s, d[k], F[k]);
                }
            }
        }
        for (int k = 0; k < 3; k++) {
            float v = F[k];
            if (p->force_clamp > 0.0f) {
                float c = p->force_clamp;
                v = v < -c ? -c : (v > c ? c : v);
            }
            force4[4 * (size_t)i + k] = v;
        }
        force4[4 * (size_t)i + 3] = 0.0f;
    }
    if (st_out) {
        st_out->visits = visits;
        st_out->leaf_terms = leaf_terms;
        st_out->monopole_terms = mono_terms;
        st_out->opened = opened;
        st_out->max_stack = max_stack;
    }
}

/* BarnesHut.metal:169-196 (updateSpheres): a = F*scale[type]; v = v*damping + a*dt;
 * x = x + v*dt, each product and sum rounded separately. */
void orc_update_spheres(PositionMass* pos, VelocityRadius* vel, const float* force4,
                        const ColorType* types, uint32_t n, float dt, const astro_params* p) {
    for (uint32_t i = 0; i < n; i++) {
        uint32_t type = types ? types[i].type : 0u;
        float sc = type < 3u ? p->type_scale[type] : 1.0f;
        for (int k = 0; k < 3; k++) {
            float a = force4[4 * (size_t)i + k] * sc;
            float v = vel[i].velocity[k] * p->damping;
            v = v + a * dt;
            vel[i].velocity[k] = v;
            pos[i].position[k] = pos[i].position[k] + v * dt;
        }
    }
}

/* BarnesHut.metal:103-167 (lightingPass): every planet (type 1) and dust particle (type 2)
 * takes the light of the first min(num_stars, 5) particles (the stars) of its system:
 * rgb = sum color_k * att_k with att = 1 / n^2, n = max(|x_k - x| / 4e5, 1e-9); alpha = 1 for
 * planets, min(max(r, g, b), 0.25) for dust.  |d| = sqrt(fma(dz,dz, fma(dy,dy, dx*dx))); the
 * light sum adds each product separately.  Stars beyond the system's end are skipped (the
 * reference would read past its particle range).  sys_off: n_sys + 1 offsets, or NULL for
 * one system of n. */
void orc_lighting_pass(const PositionMass* pos, ColorType* col, uint32_t n, uint32_t num_stars,
                       const uint32_t* sys_off, uint32_t n_sys) {
    const uint32_t k_max = num_stars < 5u ? num_stars : 5u;
    for (uint32_t s = 0; s < (sys_off ? n_sys : 1u); s++) {
        const uint32_t b = sys_off ? sys_off[s] : 0u, e = sys_off ? sys_off[s + 1] : n;
        for (uint32_t i = b; i < e; i++) {
            const uint32_t type = col[i].type;
            if (type != 1u && type != 2u) continue;
            float tot[3] = {0.0f, 0.0f, 0.0f};
            for (uint32_t k = 0; k < k_max && b + k < e; k++) {
                const uint32_t si = b + k;
                const float dx = pos[si].position[0] - pos[i].position[0];
                const float dy = pos[si].position[1] - pos[i].position[1];
                const float dz = pos[si].position[2] - pos[i].position[2];
                const float dist = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
                const float q = fmaxf(dist / ASTRO_LIGHT_ATTENUATION_DISTANCE, 1e-9f);
                const float att = 1.0f / (q * q);
                for (int c = 0; c < 3; c++) tot[c] = tot[c] + col[si].color[c] * att;
            }
            for (int c = 0; c < 3; c++) col[i].color[c] = tot[c];
            col[i].color[3] = type == 2u ? fminf(fmaxf(tot[0], fmaxf(tot[1], tot[2])), 0.25f) : 1.0f;
        }
    }
}

/* Renderer.swift:495-542 (the compute part of draw(in:)) with an explicit dt: build, forces,
 * lighting pass when p->lighting_stars > 0 (col is then updated), integrator. */
int orc_step(PositionMass* pos, VelocityRadius* vel, ColorType* types, uint32_t ns,
             uint32_t nt, const astro_params* p, float dt, float* force4, orc_stats* st,
             int threads, OctreeNode* nodes_out, orc_layers* L_out) {
    orc_layers L;
    float cube[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    if (p->quantizer == ASTRO_QUANT_BBOX_POW2) orc_bbox_cube(pos, nt, cube);
    orc_layer_table(nt, p->bits_per_axis, &L);
    OctreeNode* nodes = nodes_out;
    if (!nodes) nodes = (OctreeNode*)malloc(sizeof(OctreeNode) * (L.total ? L.total : 1));
    uint64_t* codes = (uint64_t*)malloc(sizeof(uint64_t) * (nt ? nt : 1));
    uint64_t* sc = (uint64_t*)malloc(sizeof(uint64_t) * (nt ? nt : 1));
    uint32_t* si = (uint32_t*)malloc(sizeof(uint32_t) * (nt ? nt : 1));
    int rc = orc_build_tree(pos, nt, p, cube, 0, nodes, &L, codes, sc, si);
    if (ns > 0) {
        if (nt > 0)
            orc_barnes_hut(nodes, L.off[L.layers - 1], pos, types, codes, ns, nt, p, cube[3], 0,
                           ns, force4, st, threads);
        else
            memset(force4, 0, sizeof(float) * 4 * (size_t)ns);
        if (p->lighting_stars > 0 && types) orc_lighting_pass(pos, types, ns, p->lighting_stars, NULL, 1);
        orc_update_spheres(pos, vel, force4, types, ns, dt, p);
    }
    if (L_out) *L_out = L;
    if (!nodes_out) free(nodes);
    free(codes);
    free(sc);
    free(si);
    return rc;
}

/* FP64 softened direct sum for the given targets: a_i = sum_j G m_j d / (|d|^2+eps2)^1.5,
 * skipping j == i and, when skip_keys is given, every j whose key equals the target's
 * (the Barnes-Hut leaf self-skip removes the target's whole leaf). */
void orc_direct_sum_fp64(const PositionMass* pos, uint32_t n, const uint32_t* targets,
                         uint32_t n_targets, double G, double eps2, const uint64_t* skip_keys,
                         double* acc3, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int64_t t = 0; t < (int64_t)n_targets; t++) {
        uint32_t i = targets[t];
        double xi = pos[i].position[0], yi = pos[i].position[1], zi = pos[i].position[2];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (uint32_t j = 0; j < n; j++) {
            if (j == i) continue;
            if (skip_keys && skip_keys[j] == skip_keys[i]) continue;
            double dx = (double)pos[j].position[0] - xi;
            double dy = (double)pos[j].position[1] - yi;
            double dz = (double)pos[j].position[2] - zi;
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv = 1.0 / (r2 * sqrt(r2));
            double s = G * (double)pos[j].mass * inv;
            ax += s * dx;
            ay += s * dy;
            az += s * dz;
        }
        acc3[3 * t + 0] = ax;
        acc3[3 * t + 1] = ay;
        acc3[3 * t + 2] = az;
    }
}

/* FP64 diagnostics: out = {kinetic energy, px, py, pz, mass-weighted centre x, y, z}. */
void orc_energy_momentum(const PositionMass* pos, const VelocityRadius* vel, uint32_t n,
                         double out[7]) {
    double ke = 0, p[3] = {0, 0, 0}, c[3] = {0, 0, 0}, mt = 0;
    for (uint32_t i = 0; i < n; i++) {
        double m = pos[i].mass;
        double v2 = 0;
        for (int k = 0; k < 3; k++) {
            double v = vel[i].velocity[k];
            v2 += v * v;
            p[k] += m * v;
            c[k] += m * (double)pos[i].position[k];
        }
        ke += 0.5 * m * v2;
        mt += m;
    }
    out[0] = ke;
    for (int k = 0; k < 3; k++) {
        out[1 + k] = p[k];
        out[4 + k] = mt > 0 ? c[k] / mt : 0.0;
    }
}
