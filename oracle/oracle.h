/*
 * oracle.h -- CPU restatement of astro's Barnes-Hut step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker (or as the timed CPU
 * baseline).  The product path (astro_amd/libastro.so) never links or calls it.
 *
 * Every function restates one live reference kernel or host routine; the citation is on
 * the definition in oracle.c.  Arithmetic is plain IEEE-754 single precision with the
 * operation order written out (built with -ffp-contract=off; fused multiply-adds only where
 * fmaf() is spelled), so a GPU kernel that performs the same operations in the same order
 * reproduces these results bit for bit.
 */
#ifndef ASTRO_ORACLE_H
#define ASTRO_ORACLE_H

#include "../include/astro/bridge.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-target traversal statistics (summed over targets by the callers). */
typedef struct {
    uint64_t visits;          /* nodes popped from the stack                      */
    uint64_t leaf_terms;      /* leaf interactions added                          */
    uint64_t monopole_terms;  /* accepted internal nodes added                    */
    uint64_t opened;          /* internal nodes opened                            */
    uint32_t max_stack;       /* deepest stack seen                               */
} orc_stats;

/* Layer table of a tree (Renderer.swift:373-386): cap[l] = min(nt, 2^(3B-3l)). */
typedef struct {
    uint32_t layers;          /* B + 1 */
    uint64_t cap[32];
    uint64_t off[32];
    uint64_t total;           /* sum of caps: OctreeNode array length */
    uint32_t count[32];       /* valid nodes per layer after a build   */
} orc_layers;

void orc_params_compat(astro_params* p);
void orc_params_physics(astro_params* p, uint32_t bits_per_axis);

uint64_t orc_interleave(uint32_t x, uint32_t y, uint32_t z);
void orc_bbox_cube(const PositionMass* pos, uint32_t n, float cube[4]);
void orc_morton_codes(const PositionMass* pos, uint32_t n, const astro_params* p,
                      const float cube[4], uint64_t* codes);

void orc_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t len);
void orc_split_pass(const uint64_t* kin, const uint32_t* iin, uint64_t* kout, uint32_t* iout,
                    uint32_t bit, uint32_t count);
void orc_split_sort(uint64_t* keys, uint32_t* idx, uint32_t count);

uint32_t orc_count_unique(const uint64_t* sorted, uint32_t count, int aggregate,
                          uint32_t* flags, uint32_t* scan, uint32_t* unique_idx);

void orc_layer_table(uint32_t nt, uint32_t bits_per_axis, orc_layers* L);

int orc_build_tree(const PositionMass* pos, uint32_t nt, const astro_params* p,
                   const float cube[4], int resort_layers, OctreeNode* nodes,
                   orc_layers* L, uint64_t* codes, uint64_t* sorted_codes,
                   uint32_t* sorted_idx);

void orc_barnes_hut(const OctreeNode* nodes, uint64_t root, const PositionMass* pos,
                    const ColorType* types, const uint64_t* codes, uint32_t ns, uint32_t nt,
                    const astro_params* p, float cube_side, uint32_t t_begin, uint32_t t_end,
                    float* force4, orc_stats* st, int threads);

void orc_update_spheres(PositionMass* pos, VelocityRadius* vel, const float* force4,
                        const ColorType* types, uint32_t n, float dt, const astro_params* p);

void orc_lighting_pass(const PositionMass* pos, ColorType* col, uint32_t n, uint32_t num_stars,
                       const uint32_t* sys_off, uint32_t n_sys);
int orc_step(PositionMass* pos, VelocityRadius* vel, ColorType* types, uint32_t ns,
             uint32_t nt, const astro_params* p, float dt, float* force4, orc_stats* st,
             int threads, OctreeNode* nodes_out, orc_layers* L_out);

void orc_direct_sum_fp64(const PositionMass* pos, uint32_t n, const uint32_t* targets,
                         uint32_t n_targets, double G, double eps2, const uint64_t* skip_keys,
                         double* acc3, int threads);

void orc_energy_momentum(const PositionMass* pos, const VelocityRadius* vel, uint32_t n,
                         double out[7]);

#ifdef __cplusplus
}
#endif
#endif
