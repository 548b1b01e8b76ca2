/*
 * bridge.h -- structs shared by the host, the HIP kernels and the CPU oracle.
 *
 * Byte layouts follow the reference's bridging header (Astro/bridge.h:17-65), where every
 * vector member is a simd_float3/float4 occupying a 16-byte aligned slot.  HIP's own float3
 * is 12 bytes, so the vector slots here are spelled as 16-byte aligned float[4] arrays; the
 * fourth lane of a 3-vector slot is padding.  The static_asserts pin sizes and offsets.
 */
#ifndef ASTRO_BRIDGE_H
#define ASTRO_BRIDGE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define ASTRO_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define ASTRO_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

#define ASTRO_ALIGN16 __attribute__((aligned(16)))

/* Astro/bridge.h:17-24 -- position (float3 slot, 16 B) + mass @16; 32 B. */
typedef struct ASTRO_ALIGN16 {
    ASTRO_ALIGN16 float position[4]; /* x, y, z, (pad) */
    float mass;
} PositionMass;

/* Astro/bridge.h:26-33 -- velocity (float3 slot) + radius @16; 32 B. */
typedef struct ASTRO_ALIGN16 {
    ASTRO_ALIGN16 float velocity[4];
    float radius;
} VelocityRadius;

/* Astro/bridge.h:35-43 -- colour float4 + type @16 (0 star, 1 planet, 2 dust); 32 B. */
typedef struct ASTRO_ALIGN16 {
    ASTRO_ALIGN16 float color[4];
    uint32_t type;
} ColorType;

/* Astro/bridge.h:47-65 -- one octree node; 128 B.  children[] hold absolute node indices,
 * ASTRO_INVALID_CHILD marks an empty slot. */
typedef struct ASTRO_ALIGN16 {
    uint64_t mortonCode;
    ASTRO_ALIGN16 float centerOfMass[4];
    float totalMass;
    ASTRO_ALIGN16 float emittedColor[4];
    ASTRO_ALIGN16 float emittedColorCenter[4];
    uint32_t children[8];
    uint32_t layer;
} OctreeNode;

/* Astro/bridge.h:90-94 and Astro/Octree.metal:10 */
#define ASTRO_INVALID_MORTON_CODE 0xFFFFFFFFFFFFFFFFull
#define ASTRO_INVALID_CHILD 0xFFFFFFFFu

ASTRO_STATIC_ASSERT(sizeof(PositionMass) == 32, "PositionMass must be 32 B");
ASTRO_STATIC_ASSERT(offsetof(PositionMass, mass) == 16, "mass @16");
ASTRO_STATIC_ASSERT(sizeof(VelocityRadius) == 32, "VelocityRadius must be 32 B");
ASTRO_STATIC_ASSERT(offsetof(VelocityRadius, radius) == 16, "radius @16");
ASTRO_STATIC_ASSERT(sizeof(ColorType) == 32, "ColorType must be 32 B");
ASTRO_STATIC_ASSERT(offsetof(ColorType, type) == 16, "type @16");
ASTRO_STATIC_ASSERT(sizeof(OctreeNode) == 128, "OctreeNode must be 128 B");
ASTRO_STATIC_ASSERT(offsetof(OctreeNode, centerOfMass) == 16, "centerOfMass @16");
ASTRO_STATIC_ASSERT(offsetof(OctreeNode, totalMass) == 32, "totalMass @32");
ASTRO_STATIC_ASSERT(offsetof(OctreeNode, emittedColor) == 48, "emittedColor @48");
ASTRO_STATIC_ASSERT(offsetof(OctreeNode, emittedColorCenter) == 64, "emittedColorCenter @64");
ASTRO_STATIC_ASSERT(offsetof(OctreeNode, children) == 80, "children @80");
ASTRO_STATIC_ASSERT(offsetof(OctreeNode, layer) == 112, "layer @112");

/*
 * astro_params -- every constant the reference hard-codes in its kernels, as one POD.
 * astro_params_compat() reproduces the reference; astro_params_physics() is the scaled
 * profile used for self-gravitating systems (G=1, exact cell sizes, no cut-offs).
 */
enum {
    ASTRO_QUANT_OFFSET_WRAP = 0, /* (u32)(p + offset) & 0x1FFFFF   (MortonCode.metal:35-37)   */
    ASTRO_QUANT_BBOX_POW2 = 1    /* per-system power-of-two cube from the bounding box        */
};
enum {
    ASTRO_SIDE_SHIFTED = 0, /* side = 2^(21+layer)             (BarnesHut.metal:62)          */
    ASTRO_SIDE_EXACT = 1    /* side = cube side * 2^(layer - B)                              */
};

typedef struct {
    uint32_t bits_per_axis;   /* B: key bits kept per axis; keys hold 3B bits; B+1 layers   */
    uint32_t quantizer;       /* ASTRO_QUANT_*                                               */
    float quant_offset;       /* compat quantizer origin shift (1e6)                         */
    uint32_t side_rule;       /* ASTRO_SIDE_*                                                */
    float theta;              /* opening angle (Renderer.swift:864)                          */
    float eps2;               /* softening added to r^2 (BarnesHut.metal:40,58)              */
    float G;                  /* gravitational constant (BarnesHut.metal:39)                 */
    float r2_floor;           /* max(r^2, .) (BarnesHut.metal:59)                            */
    float r_floor;            /* max(r, .)   (BarnesHut.metal:61)                            */
    float leaf_cutoff_r2;     /* leaf interactions need r^2 > this (BarnesHut.metal:70)      */
    float force_clamp;        /* component clamp (BarnesHut.metal:100); <=0 disables         */
    float damping;            /* velocity factor per step (BarnesHut.metal:190)              */
    float type_scale[3];      /* acceleration scale by type (BarnesHut.metal:182-188)        */
    uint32_t quirk_last_leaf; /* 1: last leaf run ends at numUnique (Octree.metal:71)        */
    uint32_t fp64_top_layers; /* aggregate the top k layers' mass/COM in FP64                */
    uint32_t kahan;           /* 1: Kahan-compensated force accumulation                     */
    uint32_t multipole;       /* 0: monopoles (the reference); 2: + quadrupoles of accepted    *
                               * internal nodes (second moments kept per node)                  */
    uint32_t lighting_stars;  /* stars lighting the other particles in the step's lighting     *
                               * pass (BarnesHut.metal:103-167; the first particles, at most 5);*
                               * 0: no lighting pass                                            */
} astro_params;

/* BarnesHut.metal:121 (bridge.h:98-100) */
#define ASTRO_LIGHT_ATTENUATION_DISTANCE 4e5f

#ifdef __cplusplus
}
#endif
#endif /* ASTRO_BRIDGE_H */
