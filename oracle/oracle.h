/*
 * CPU oracle for the rex path-tracing core -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference renderer's algorithms for
 * the hot path so that the HIP product path can be checked against it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker: nothing in renderingengine_amd/ links or calls it.
 *
 * Reference citations are relative to the RenderingEngine source tree.
 *
 * Parity pins (what checks each function against something other than itself)
 * are listed per function in the .c files; a function whose header comment
 * says "parity unpinned" has no such independent check.
 *
 * Arithmetic: fp32 everywhere the reference shaders use fp32, compiled with
 * -ffp-contract=off and without fast-math so each +,-,*,/ and sqrt rounds once,
 * in the written order (dot = (x*x + y*y) + z*z, cross as written).
 */
#ifndef REX_ORACLE_H
#define REX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_INVALID 0xffffffffu

/* 48-byte flattened node, include/rt/software_gpu/SoftwareGpu.hpp:46-60.
 * links = leaf ? (first, count, INVALID, parent) : (left, 0, right, parent). */
typedef struct ora_node {
    float bmin[4];
    float bmax[4];
    uint32_t links[4];
} ora_node;

/* 176-byte canonical traversal triangle,
 * include/rt/gpu/CanonicalTraversalScene.hpp:15-23.
 * identity = (instance, primitive, geometry, material)
 * metadata = (geometry flags, material flags, traversal ID, instance flags) */
typedef struct ora_tri {
    float pos[3][4];
    float nrm[3][4];
    float uv[3][4];
    uint32_t identity[4];
    uint32_t metadata[4];
} ora_tri;

/* ABI-v1 queue ray, include/contracts/GpuRecordsAbiV1.hpp:76-82 (64 B). */
typedef struct ora_ray_v1 {
    float originTMin[4];
    float directionTMax[4];
    uint32_t identity[4]; /* ray ID, path ID, bounce, visibility mask */
    uint32_t rng[4];
} ora_ray_v1;

/* ABI-v1 hit record, include/contracts/GpuRecordsAbiV1.hpp:86-94 (96 B). */
typedef struct ora_hit_v1 {
    float positionT[4];
    float gnBaryU[4];
    float snBaryV[4];
    uint32_t ids[4];
    uint32_t metadata[4]; /* ray ID, HitKind, HitFlags, path ID */
    uint32_t reserved[4];
} ora_hit_v1;

/* Traversal counters, resources/shaders/traversal/software_trace_v1.hlsl:7-15. */
enum {
    ORA_CNT_NODE_TESTS = 0,
    ORA_CNT_TRIANGLE_TESTS = 1,
    ORA_CNT_STACK_OVERFLOW = 2,
    ORA_CNT_INVALID_RAY = 3,
    ORA_CNT_INVALID_HIT = 4,
    ORA_CNT_MAX_STACK = 5,
    ORA_CNT_LEAF_VISITS = 6,
    ORA_CNT_LEAF_PRIMITIVES = 7,
    ORA_CNT_MAX_LEAF_OCCUPANCY = 8,
    ORA_CNT_BOX_TESTS = 9,       /* slab tests actually evaluated (bytes model) */
    ORA_CNT_COUNT = 10
};

/* ---- BVH build + flatten (oracle_bvh.c) ---------------------------------- */

/* Binned-SAH build over triangles in canonical order, with triangle i's
 * primitive ID = tris[i].metadata[2] (the traversal ID).  Writes flattened
 * nodes (capacity 2n-1) and the leaf-order source indices (n).
 * Returns 0 on success, <0 on failure (depth > 128, invalid input). */
int ora_bvh_build(const ora_tri* tris, uint32_t n, uint32_t maxLeafSize,
                  ora_node* nodesOut, uint32_t* nodeCountOut,
                  uint32_t* primOrderOut, uint32_t* maxDepthOut);

/* ---- traversal (oracle_trace.c) ------------------------------------------ */

/* Result of the integrator's stackless walk (PbrHitL6 subset). */
typedef struct ora_walk_hit {
    uint32_t result;      /* 0 = miss, 1 = hit, 2 = fail-closed invalid */
    float t, u, v;
    uint32_t frontFace;
    uint32_t triangle;    /* leaf-order index of the winner */
    float position[3];
    float geometricNormal[3];
    float shadingNormal[3];
    uint32_t instanceId, primitiveId, materialIndex;
} ora_walk_hit;

/* Stackless parent-pointer walk of the integrator,
 * resources/shaders/integrators/pbr_software_traversal.hlsli:287-576.
 * tris are in leaf order.  counters may be NULL. */
void ora_trace_walk(const ora_node* nodes, uint32_t nodeCount,
                    const ora_tri* tris, uint32_t triCount,
                    const float origin[3], const float direction[3],
                    float tMin, float tMax,
                    uint32_t ignoredInstance, uint32_t ignoredPrimitive,
                    int shadowQuery, ora_walk_hit* hit, uint64_t* counters);

/* The same query semantics (ignored key, shadow casters, culling,
 * (instance, primitive) ties) visited in near-first order: both children of
 * an entered node are tested against the current best, the nearer one is
 * entered first and the other deferred with its clamped entry distance, as
 * the batch tracer orders its stack (software_trace_v1.hlsl:338-373).  This
 * order is the product's own production choice; it equals the left-first
 * walk wherever no box's rounded entry distance exceeds the winner's t. */
void ora_trace_walk_near_first(const ora_node* nodes, uint32_t nodeCount,
                               const ora_tri* tris, uint32_t triCount,
                               const float origin[3], const float direction[3],
                               float tMin, float tMax,
                               uint32_t ignoredInstance, uint32_t ignoredPrimitive,
                               int shadowQuery, ora_walk_hit* hit, uint64_t* counters);

/* 4-wide node of the product's production walk: the binary interior node
 * whose children are its grandchildren (a leaf child stays itself), in
 * binary pre-order (LL, LR, RL, RR).  child[k] = binary node index or
 * ORA_INVALID for an empty slot; wide[k] = the 4-wide index of an interior
 * child.  Wide nodes are numbered in pre-order from the root (index 0). */
typedef struct ora_wide4 {
    uint32_t child[4];
    uint32_t wide[4];
} ora_wide4;

/* Collapses the binary hierarchy; out needs nodeCount entries.  Returns the
 * wide node count (0 when the root is a leaf); *depth = wide levels. */
uint32_t ora_build_wide4(const ora_node* nodes, uint32_t nodeCount, ora_wide4* out, uint32_t* depth);

/* Near-first walk over the 4-wide nodes: every child box of a node is tested,
 * the hit children are ordered by (clamped entry distance, slot) and the
 * nearest is entered while the others are pushed, farthest first.  Same ray
 * validity, tie rule, shadow and culling semantics as the other walks; ties
 * between the orders are the documented rounding ties of near-first. */
void ora_trace_walk_wide(const ora_node* nodes, uint32_t nodeCount, const ora_wide4* wide,
                         const ora_tri* tris, uint32_t triCount,
                         const float origin[3], const float direction[3],
                         float tMin, float tMax,
                         uint32_t ignoredInstance, uint32_t ignoredPrimitive,
                         int shadowQuery, ora_walk_hit* hit, uint64_t* counters);

/* Batch of walk queries: rays[i] = (o.xyz, tMin, d.xyz, tMax), keys[i] =
 * (ignoredInstance, ignoredPrimitive), shadow != 0 for any-hit,
 * order 0 = left-first (reference), 1 = near-first. */
void ora_trace_walk_batch(const ora_node* nodes, uint32_t nodeCount,
                          const ora_tri* tris, uint32_t triCount,
                          const float* rays, const uint32_t* keys,
                          uint32_t count, int shadow,
                          ora_walk_hit* hits, uint64_t* counters);
/* order: 0 left-first (reference), 1 near-first binary, 2 near-first 4-wide. */
void ora_trace_walk_batch_order(const ora_node* nodes, uint32_t nodeCount,
                                const ora_tri* tris, uint32_t triCount,
                                const float* rays, const uint32_t* keys,
                                uint32_t count, int shadow, int order,
                                ora_walk_hit* hits, uint64_t* counters);

/* The same batch on `threads` host threads (identical hits; counters summed,
 * maxima taken). */
void ora_trace_walk_batch_mt(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                             const float* rays, const uint32_t* keys, uint32_t count, int shadow, int order,
                             uint32_t threads, ora_walk_hit* hits, uint64_t* counters);

/* Near-first batch tracer of the IGpuTraversalBackend path,
 * resources/shaders/traversal/software_trace_v1.hlsl:189-411.
 * queryMode 0 = closest, 1 = any.  tris are in leaf order. */
void ora_trace_batch_v1(const ora_node* nodes, uint32_t nodeCount,
                        const ora_tri* tris, uint32_t triCount,
                        const ora_ray_v1* rays, uint32_t rayCount,
                        uint32_t queryMode, uint32_t stackCapacity,
                        ora_hit_v1* hits, uint64_t* counters);

/* Brute force over the leaf-order triangles with the same per-triangle test
 * and (instance, primitive) tie rule as the walk (closest); used as the
 * order-independent check of both walks. */
void ora_trace_brute(const ora_tri* tris, uint32_t triCount,
                     const float origin[3], const float direction[3],
                     float tMin, float tMax, ora_walk_hit* hit);

/* ---- alpha-masked geometry (oracle_trace.c) ---------------------------- */

/* ABI-v0 material record (include/contracts/SceneAbiV0.hpp, 128 B). */
typedef struct ora_material_v0 {
    float baseColorFactor[4];
    float emissiveFactorStrength[4];
    float surfaceParams[4];          /* metallic, roughness, normal scale, alpha cutoff */
    float transmissionParams[4];
    float attenuationColorDistance[4];
    uint32_t textureImageIndices[4];
    uint32_t textureSamplerIndices[4];
    uint32_t metadata[4];
} ora_material_v0;

/* Private marks in a triangle's metadata[0] (geometry flags) written by
 * ora_alpha_resolve: the alpha test's resolved outcome for the walks. */
#define ORA_ALPHA_ACCEPT (1u << 30)
#define ORA_ALPHA_SKIP (1u << 31)

void ora_alpha_resolve(ora_tri* tris, uint32_t n, const ora_material_v0* materials, uint32_t materialCount,
                       uint32_t atlasLayerCount, const float* layerAlpha, uint32_t samplerId, uint8_t* outcome);

/* ---- the reference's CPU Cornell path tracer (oracle_cornell.c) -------- */

/* RenderCornellReference (src/integrators/reference_cpu/CornellReference.cpp:
 * 734-909) over rows [rowBegin, rowEnd) of a w x h film on `threads` host
 * threads; pixels: that band, row-major linear RGB.  counts: extension rays,
 * shadow rays, non-finite events.  Returns 0, or -1 for invalid options. */
int ora_cornell_render(uint32_t w, uint32_t h, uint32_t spp, uint32_t maxBounces, uint64_t seed,
                       uint32_t threads, uint32_t rowBegin, uint32_t rowEnd, float* pixels, uint64_t counts[3],
                       uint64_t* sceneHash);
/* The fixture's triangles (9 doubles each, primitive order), materials, the
 * Bvh<double> leaf order and node count; returns the triangle count. */
uint32_t ora_cornell_scene(double* verts, uint32_t* materials, uint32_t* leafOrder, uint32_t* nodeCount);
/* HashLinearRgbPixels (src/rt/cpu/ImageOutput.cpp:343-378). */
uint64_t ora_linear_rgb_hash(uint32_t w, uint32_t h, const float* rgb);
/* PCG32 (include/rt/cpu/Random.hpp:17-66): n outputs after Seed(seed, seq). */
void ora_pcg32_sequence(uint64_t seed, uint64_t seq, uint32_t n, uint32_t* out);
/* Lambert (include/bsdf/reference_cpu/Lambert.hpp): Pdf, and Sample with a
 * PCG32 stream -> (direction, pdf; 0 when invalid). */
double ora_lambert_pdf(const double n[3], const double wo[3], const double wi[3]);
void ora_lambert_sample(const double n[3], const double wo[3], uint64_t seed, uint64_t seq, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
