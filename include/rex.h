/*
 * rex -- MI355X-native wavefront path-tracing core: the C ABI.
 *
 * This header is the drop-in boundary.  Every entry point replaces one
 * operation of the reference renderer's interfaces for the Flattened-SAH BVH
 * + Wavefront + PBR (+ ReSTIR DI) path; the comment on each names the
 * reference interface it replaces (paths relative to the RenderingEngine
 * source tree).
 *
 * Conventions
 *  - Plain C types and pointers only.  Device pointers are hipDeviceptr-style
 *    `void*` addresses in the current HIP device's memory; `stream` is a
 *    hipStream_t passed as `void*` (NULL = the default stream).
 *  - Every call returns rex_status; on failure rex_last_error() returns a
 *    thread-local message.  No exception crosses the ABI
 *    (include/rt/gpu/IGpuTraversalBackend.hpp:15-33 status codes, plus
 *    REX_DEVICE_ERROR for HIP runtime failures).
 *  - One host thread drives one rex_scene / rex_renderer at a time.
 *  - Record layouts are byte-compatible with the reference's ABI structs and
 *    are static_assert-ed in the implementation.
 */
#ifndef REX_H
#define REX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REX_ABI_VERSION 1u
#define REX_INVALID_INDEX 0xffffffffu

typedef enum rex_status {
    REX_SUCCESS = 0,
    REX_INVALID_ARGUMENT = 1,
    REX_UNSUPPORTED = 2,
    REX_MISSING_SCENE = 3,
    REX_RECORDING_FAILED = 4,
    REX_DEVICE_ERROR = 5
} rex_status;

/* ---- records (byte layouts of the reference ABI) ------------------------- */

/* CanonicalTraversalTriangle, include/rt/gpu/CanonicalTraversalScene.hpp:15-23
 * (176 B): world-space positions (w = 1), normals (w = 0), texcoords;
 * identity = (instance, primitive, geometry, material IDs);
 * metadata = (geometry flags, material flags, traversal ID, instance flags). */
typedef struct rex_triangle {
    float positions[3][4];
    float normals[3][4];
    float texcoords[3][4];
    uint32_t identity[4];
    uint32_t metadata[4];
} rex_triangle;

/* GpuRayQueueRecordV1, include/contracts/GpuRecordsAbiV1.hpp:76-82 (64 B).
 * identity = (ray ID, path ID, bounce, visibility mask). */
typedef struct rex_ray_v1 {
    float originTMin[4];
    float directionTMax[4];
    uint32_t identity[4];
    uint32_t rng[4];
} rex_ray_v1;

/* GpuHitQueueRecordV1, include/contracts/GpuRecordsAbiV1.hpp:86-94 (96 B).
 * ids = (instance, primitive, geometry, material);
 * metadata = (ray ID, HitKind, HitFlags, path ID). */
typedef struct rex_hit_v1 {
    float positionT[4];
    float geometricNormalBaryU[4];
    float shadingNormalBaryV[4];
    uint32_t ids[4];
    uint32_t metadata[4];
    uint32_t reserved0[4];
} rex_hit_v1;

/* HitKind (include/contracts/RayHitAbiV0.hpp:15-21). */
#define REX_HIT_KIND_MISS 0u
#define REX_HIT_KIND_TRIANGLE 1u
#define REX_HIT_KIND_INVALID 3u

/* Traversal counter slots (resources/shaders/traversal/software_trace_v1.hlsl:7-15),
 * plus slab tests evaluated (REX_COUNTER_BOX_TESTS). */
#define REX_COUNTER_NODE_TESTS 0u
#define REX_COUNTER_TRIANGLE_TESTS 1u
#define REX_COUNTER_STACK_OVERFLOW 2u
#define REX_COUNTER_INVALID_RAY 3u
#define REX_COUNTER_INVALID_HIT 4u
#define REX_COUNTER_MAX_STACK 5u
#define REX_COUNTER_LEAF_VISITS 6u
#define REX_COUNTER_LEAF_PRIMITIVES 7u
#define REX_COUNTER_MAX_LEAF_OCCUPANCY 8u
#define REX_COUNTER_BOX_TESTS 9u
#define REX_COUNTER_COUNT 10u

/* ---- errors / device ----------------------------------------------------- */

const char* rex_last_error(void);
uint32_t rex_abi_version(void);

/* Plain device-memory helpers so a host without its own allocator can drive
 * the ABI (the reference's callers own Vulkan buffers; these stand in for
 * that ownership).  */
rex_status rex_device_alloc(size_t bytes, void** devicePtr);
rex_status rex_device_free(void* devicePtr);
rex_status rex_copy_to_device(void* dst, const void* src, size_t bytes);
rex_status rex_copy_to_host(void* dst, const void* src, size_t bytes);
rex_status rex_synchronize(void);

/* ---- scene --------------------------------------------------------------- */

typedef struct rex_scene_t* rex_scene;

struct rex_material_v0;

/* Alpha-masked geometry (pbr_software_traversal.hlsli:207-248, 430-471;
 * software_trace_v1.hlsl:127-170).  A candidate whose geometry or material
 * flags carry ALPHA_MASK is accepted when alpha >= its material's cutoff
 * (surfaceParams[3]); alpha = baseColorFactor[3] times the alpha of its
 * atlas layer textureImageIndices[0] (sampled with textureSamplerIndices[0]),
 * or baseColorFactor[3] alone when both indices are invalid (0xFFFFFFFF).
 * The atlas is described by uniform layers: layerAlpha[k] is the alpha of
 * every texel of layer k.  The renderers bind a 1x1 two-layer atlas
 * (layer 0 alpha 0, layer 1 alpha 1, sampler 0; Wave2Runtime.cpp:392-408),
 * REX_ALPHA_ATLAS_RENDERER below; textured alpha is not supported.  Each
 * masked triangle is thus accepted or skipped as a whole; a skipped one keeps
 * its place in the hierarchy (the reference's BVH) and is never hit.
 * A masked triangle whose material index is out of range, or whose mapping
 * the reference treats as invalid (an image without the atlas sampler, a
 * sampler without an image, a layer beyond the atlas), makes the reference
 * return an invalid traversal result that depends on its visiting order:
 * rex_scene_create rejects such scenes with REX_UNSUPPORTED. */
typedef struct rex_alpha_desc {
    const struct rex_material_v0* materials;  /* indexed by triangle identity[3] */
    uint32_t materialCount;
    uint32_t atlasLayerCount;
    const float* layerAlpha;                  /* atlasLayerCount values */
    uint32_t samplerId;
    uint32_t reserved;
} rex_alpha_desc;

/* Scene build options. */
typedef struct rex_scene_desc {
    uint64_t fingerprint;     /* nonzero scene identity */
    uint32_t generation;      /* nonzero, bumped on every topology change */
    uint32_t maxLeafSize;     /* 1..8; 0 selects the reference default 4 */
    uint32_t buildThreads;    /* 0 = all host cores */
    uint32_t buildMethod;     /* REX_BUILD_* */
    const rex_alpha_desc* alpha;  /* NULL: alpha-masked triangles are rejected (REX_UNSUPPORTED);
                                     copied by rex_scene_create, reused by rex_scene_refit */
} rex_scene_desc;

/* REX_BUILD_BINNED_SAH: Bvh<float>(BinnedSah) + FlattenCanonicalL3BinnedSah.
 * REX_BUILD_KARRAS_LBVH: BuildKarrasLbvh (src/rt/software_gpu/SoftwareGpu.cpp:1265-1345):
 * 30-bit Morton codes of the triangle centroids in the scene box, sorted by
 * (code, traversal ID), Karras splits at the highest differing key bit,
 * one triangle per leaf, depth <= 128, boxes by RefitBottomUp -- node for
 * node the reference's LBVH; a fast rebuild for animated scenes. */
#define REX_BUILD_BINNED_SAH 0u
#define REX_BUILD_KARRAS_LBVH 1u

/* The renderers' alpha atlas (Wave2Runtime.cpp:392-408, frame.traversal.w = 2,
 * frame.output.y = 0): two 1x1 layers, layerAlpha = {0.0f, 1.0f}, sampler 0.
 * rex_scene_create_canonical uses it with the scene's own materials. */
#define REX_ALPHA_ATLAS_RENDERER_LAYERS 2u
#define REX_ALPHA_ATLAS_RENDERER_SAMPLER 0u

/* Builds the binned-SAH BVH on the host (identical topology, bounds and leaf
 * order to Rt::Cpu::Bvh<float> + FlattenCanonicalL3BinnedSah), converts it to
 * the device layout and uploads it.  Replaces Wave2Runtime::SetScene's
 * traversal half (src/renderers/Wave2Runtime.cpp:936-969) and
 * IGpuTraversalBackend::BuildOrUpdateScene
 * (include/rt/gpu/IGpuTraversalBackend.hpp:78-79).
 * Triangles are the canonical stream in canonical order (host memory). */
rex_status rex_scene_create(const rex_triangle* triangles, uint32_t triangleCount,
                            const rex_scene_desc* desc, rex_scene* outScene);
rex_status rex_scene_destroy(rex_scene scene);

typedef struct rex_scene_info {
    uint64_t fingerprint;
    uint32_t generation;
    uint32_t triangleCount;
    uint32_t nodeCount;        /* binary (reference) node count */
    uint32_t wideNodeCount;    /* device interior records */
    uint32_t maximumDepth;
    uint32_t maxLeafSize;
    uint32_t compressedNodeCount; /* compressed 4-wide nodes of the production walk: those of the traversal
                                     hierarchy (one triangle per leaf) where the scene has one */
    double buildMilliseconds;  /* host SAH build + flatten; LBVH: the whole rex_scene_create build */
    uint64_t deviceBytes;      /* BVH + triangle streams resident in HBM */
    double deviceMilliseconds; /* wall time of the device part of the last LBVH build or refit (from the
                                  uploaded triangles to the compressed layout, synchronized), 0 otherwise */
} rex_scene_info;
rex_status rex_scene_get_info(rex_scene scene, rex_scene_info* info);

/* Host-only build (no device needed): the same binned-SAH build + flatten as
 * rex_scene_create, returning the 48-B SoftwareNodeRecord array (capacity
 * 2n-1 records), the leaf-order source indices (n) and counts.  Replaces
 * Rt::Cpu::Bvh<float>(BinnedSah) + FlattenCanonicalL3BinnedSah
 * (include/rt/cpu/Bvh.hpp:64-681, src/rt/software_gpu/SoftwareGpu.cpp:1107-1251). */
rex_status rex_bvh_build_host(const rex_triangle* triangles, uint32_t triangleCount, uint32_t maxLeafSize,
                              uint32_t buildThreads, void* nodes48, uint32_t* primitiveOrder,
                              uint32_t* nodeCount, uint32_t* maximumDepth);

/* Copies the flattened binary node array (48-B SoftwareNodeRecord layout,
 * include/rt/software_gpu/SoftwareGpu.hpp:46-60) and the leaf-order source
 * indices to host memory; either pointer may be NULL. */
rex_status rex_scene_export_flat(rex_scene scene, void* nodes48, uint32_t* primitiveOrder);

/* Animated frames without a rebuild: the scene's triangles move (same count,
 * same traversal IDs in the same canonical order, new positions and
 * normals) and the hierarchy keeps its topology while its boxes are refit
 * bottom up -- leaf boxes from their triangles, interior boxes the union of
 * their children -- as RefitBottomUp (src/rt/software_gpu/SoftwareGpu.cpp:1347-1500)
 * refits LBVH nodes; the device layouts are re-derived and uploaded.
 * generation: the scene's new generation (nonzero).  Waits for the device.
 * The shading tables depend on the emissive triangles: rendering needs
 * rex_scene_set_shading again (REX_MISSING_SCENE until then).  Batch
 * engines and servers hold their own copies of a scene added before. */
rex_status rex_scene_refit(rex_scene scene, const rex_triangle* triangles, uint32_t triangleCount,
                           uint32_t generation);
/* Host-only LBVH build (REX_BUILD_KARRAS_LBVH) into 48-B node records
 * (capacity 2n-1) and the leaf order (n). */
rex_status rex_bvh_build_lbvh_host(const rex_triangle* triangles, uint32_t triangleCount, void* nodes48,
                                   uint32_t* primitiveOrder, uint32_t* nodeCount, uint32_t* maximumDepth);
/* Host-only refit of 48-B SoftwareNodeRecords (nodeCount, pre-order as
 * rex_bvh_build_host writes them) in place, from the canonical triangles and
 * the leaf order. */
rex_status rex_bvh_refit_host(void* nodes48, uint32_t nodeCount, const rex_triangle* triangles,
                              const uint32_t* primitiveOrder, uint32_t triangleCount);

/* ---- traversal ----------------------------------------------------------- */

/* Batch closest-hit / any-hit over device-resident ABI-v1 records, with the
 * semantics of the software batch tracer
 * (resources/shaders/traversal/software_trace_v1.hlsl:189-411): near-first
 * order, closest ties to the lower traversal ID, stackCapacity in 1..64
 * (overflow -> HitKind invalid), invalid rays -> HitKind invalid, miss ->
 * t = tMax.  Offsets are record indices, never byte offsets; fingerprint and
 * generation must name the scene exactly.  Replaces
 * IGpuTraversalBackend::RecordTraceClosestBatch / RecordTraceAnyBatch
 * (include/rt/gpu/IGpuTraversalBackend.hpp:60-70, 80-83).
 * counters: optional device array of REX_COUNTER_COUNT uint64 (accumulated). */
typedef struct rex_trace_batch {
    uint64_t sceneFingerprint;
    uint32_t sceneGeneration;
    uint32_t rayOffset;
    uint32_t hitOffset;
    uint32_t rayCount;
    uint32_t stackCapacity;    /* 0 selects 64 */
    uint32_t reserved;
    const rex_ray_v1* rays;    /* device */
    rex_hit_v1* hits;          /* device */
    uint64_t* counters;        /* device, may be NULL */
} rex_trace_batch;
rex_status rex_trace_closest(rex_scene scene, const rex_trace_batch* batch, void* stream);
rex_status rex_trace_any(rex_scene scene, const rex_trace_batch* batch, void* stream);

/* No counterpart in the reference: rex_trace_closest (anyHit == 0) or
 * rex_trace_any (anyHit != 0) with identical results, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc / hipFree: the
 * persistent grid's work counter and stack overflow entries live in a static
 * device array of REX_TRACE_NO_ALLOC_SCRATCH_BYTES (29,360,384 bytes) that
 * each GPU holds until the process ends; the grid is at most 512 blocks of
 * 128 lanes.  Every call on one GPU shares that GPU's array and clears its
 * work counter with a kernel on `stream`: calls that overlap on one GPU
 * (different streams, or concurrent host threads) race and must not be
 * issued.  rayCount above REX_TRACE_NO_ALLOC_MAX_RAYS (2^31) is rejected
 * with REX_INVALID_ARGUMENT. */
#define REX_TRACE_NO_ALLOC_SCRATCH_BYTES (256u + 512u * 128u * 56u * 8u)
#define REX_TRACE_NO_ALLOC_MAX_RAYS 0x80000000u
rex_status rex_trace_no_alloc(rex_scene scene, const rex_trace_batch* batch, int anyHit, void* stream);

/* The integrator's traversal seam WfTraceClosest / WfTraceAny
 * (integrators/wavefront/shaders/include/TraversalAdapter.hlsli:4-11 ->
 * resources/shaders/integrators/pbr_software_traversal.hlsli:287-576):
 * stackless-equivalent left-first walk, closest ties to the lower
 * (instance, primitive), (instance, primitive) of `keys` ignored, shadow
 * queries skip non-casters and stop at the first accepted triangle.
 * rays: device (n, 8) floats = (origin.xyz, tMin, direction.xyz, tMax);
 * keys: device (n, 2) uint32 = ignored (instance, primitive), may be NULL;
 * hits: device records of rex_walk_hit.  walkMode selects the device walk:
 * REX_WALK_SHORT_STACK (LDS short stack) or REX_WALK_STACKLESS (parent
 * pointers) visit nodes in the reference's left-first order;
 * REX_WALK_NEAR_FIRST visits the nearer child first (the batch tracer's
 * order) and differs from the reference walk only where a box's rounded entry
 * distance exceeds the winning hit's t (documented ties); REX_WALK_WIDE
 * walks 4-wide nodes (a binary node collapsed with its children, one
 * 128-B line) nearest child first, with the same tie caveat -- the
 * production walk, about half the dependent node fetches of the binary ones.
 * Launches on one scene must be ordered on one stream (they share the
 * scene's traversal workspace). */
#define REX_WALK_SHORT_STACK 0u
#define REX_WALK_STACKLESS 1u
#define REX_WALK_NEAR_FIRST 2u
#define REX_WALK_WIDE 3u
/* REX_WALK_WIDE_COMPRESSED: the 4-wide walk over 64-B nodes whose child
 * boxes are 8-bit planes on a per-node power-of-two grid, rounded outward
 * with a margin (each contains its exact box), tested with one fma per
 * plane.  Half the node bytes of REX_WALK_WIDE; the closest hit equals the
 * exact walks' except where an exact box test's rounding culls a box that
 * holds the winning hit on its boundary (documented ties). */
#define REX_WALK_WIDE_COMPRESSED 4u
typedef struct rex_walk_hit {
    uint32_t result;          /* 0 miss, 1 hit, 2 fail-closed invalid */
    float t, u, v;
    uint32_t frontFace;
    uint32_t triangle;        /* leaf-order triangle index */
    float position[3];
    float geometricNormal[3];
    float shadingNormal[3];
    uint32_t instanceId, primitiveId, materialIndex;
} rex_walk_hit;
rex_status rex_trace_walk(rex_scene scene, const float* rays, const uint32_t* keys,
                          uint32_t count, int shadow, uint32_t walkMode,
                          rex_walk_hit* hits, uint64_t* counters, void* stream);

/* ---- shading tables ------------------------------------------------------ */

/* PbrMaterialGpu (integrators/megakernel/include/MegakernelBridge.hpp:257-338,
 * filled as src/renderers/Wave2Runtime.cpp:479-530 does), 128 B.
 * metadata = (BSDF model, allowed lobes, complex-conductor Fresnel, flags). */
typedef struct rex_material {
    float baseColorMetallic[4];
    float emissiveRoughness[4];
    float transmissionIor[4];
    float attenuationColorDistance[4];
    float f0[4];
    float conductorEta[4];
    float conductorK[4];
    uint32_t metadata[4];
} rex_material;

/* PbrLightGpu (MegakernelBridge.hpp), 96 B.  identity = (type, flags, -,
 * primitive ID); payload = (canonical triangle index, -, -, instance ID).
 * Types (MegakernelBridge.hpp:16-24): 0 point, 1 directional, 2 spot (cone
 * cosines in directionCosOuter.w (outer) and shapeParams.y (inner)), 3
 * sphere (radius in shapeParams.x), 4 emissive triangle, 5 environment. */
typedef struct rex_light {
    float positionRange[4];
    float directionCosOuter[4];
    float radianceScale[4];
    float shapeParams[4];
    uint32_t identity[4];
    uint32_t payload[4];
} rex_light;

#define REX_LIGHT_SELECTION_UNIFORM 0u
#define REX_LIGHT_SELECTION_POWER 1u

/* EnvironmentSampler (MegakernelBridge.hpp:48-52). */
#define REX_ENVIRONMENT_SAMPLER_UNIFORM_SPHERE 0u
#define REX_ENVIRONMENT_SAMPLER_IMPORTANCE_MAP 1u

typedef struct rex_shading_desc {
    const rex_material* materials;
    uint32_t materialCount;
    uint32_t lightSelection;      /* REX_LIGHT_SELECTION_* */
    const rex_light* lights;
    uint32_t lightCount;
    uint32_t environmentSampler;  /* REX_ENVIRONMENT_SAMPLER_* */
    /* Lat-long environment map: linear RGBA float texels, row-major,
     * environmentHeight rows from theta = 0 (+y) down to theta = pi, u =
     * phi / 2 pi with phi = atan2(z, x); read nearest, u repeating, v
     * clamped (the reference's Wave 2 environment image and sampler).
     * NULL: the constant map (1, 1, 1) for an enabled environment light. */
    const float* environmentTexels;
    uint32_t environmentWidth;
    uint32_t environmentHeight;
} rex_shading_desc;

/* Builds the light tables (Vose alias selection table, per-light selection
 * PMF, canonical-order emitter triangles, the (instance, primitive) -> light
 * map) and uploads materials and lights: the shading half of
 * Wave2Runtime::SetScene (src/renderers/Wave2Runtime.cpp:462-666 via
 * BuildTopLevelLightDistribution / BuildEmitterMap,
 * integrators/megakernel/MegakernelBridge.cpp:40-145, 274-454).
 * The environment map's two-level alias distribution is
 * BuildEnvironmentDistribution (MegakernelBridge.cpp:191-275) over the
 * texels' float luminance (Wave2Runtime.cpp:578-614).
 * Supported: all six BSDF models (PbrBsdf.hlsli:75-80: Lambert, GGX
 * conductor with Schlick or complex-IOR Fresnel, GGX dielectric reflection,
 * smooth glass, rough dielectric, metallic-roughness) with interior
 * attenuation of transmissive materials; every light type (one enabled
 * environment at most).
 * A map without an enabled environment light, non-finite texels or
 * negative colour -> REX_INVALID_ARGUMENT; anything else unsupported ->
 * REX_UNSUPPORTED. */
rex_status rex_scene_set_shading(rex_scene scene, const rex_shading_desc* desc);

/* Copies the environment distribution of the scene's shading tables: rows
 * (environmentHeight 16-B alias entries), columns (height x width entries)
 * and the luminance integral (texel luminance x solid angle) that weighs
 * the environment light's power.  Any output may be NULL. */
rex_status rex_scene_export_environment_tables(rex_scene scene, void* rows16, void* columns16,
                                               double* integratedLuminance);

/* Copies the light-selection alias table (q, pmf, alias, item as 16-B
 * entries; capacity lightCount) and the per-light selection PMF to host
 * memory (either may be NULL); *aliasCount receives the table size. */
rex_status rex_scene_export_light_tables(rex_scene scene, void* alias16, float* selectionPmf, uint32_t* aliasCount);

/* ---- scene ingestion: ABI-v0 canonical scenes ------------------------------ */
/* The canonical scene records of include/contracts/SceneAbiV0.hpp, byte for
 * byte: constants 96 B, vertex 64 B, geometry 64 B, instance 224 B (three
 * row-major 4x4 transforms + metadata + reserved), material 128 B, light
 * 96 B. */
typedef struct rex_scene_constants_v0 {
    uint32_t counts0[4];         /* vertices, indices, geometries, instances */
    uint32_t counts1[4];         /* materials, lights, textures, samplers */
    float sceneBoundsMin[4];
    float sceneBoundsMax[4];
    uint32_t versionFlags[4];    /* ABI version, scene generation, SceneFlags, - */
    uint32_t environment[4];     /* texture index, width, height, - */
} rex_scene_constants_v0;

typedef struct rex_vertex_v0 {
    float position[4];           /* object space, w = 1 */
    float normal[4];
    float tangent[4];
    float texcoord0[4];
} rex_vertex_v0;

typedef struct rex_geometry_v0 {
    uint32_t indexRange[4];      /* first index, index count, vertex offset, first primitive ID */
    uint32_t identity[4];        /* geometry ID, mesh ID, material index, GeometryFlags */
    float localBoundsMin[4];
    float localBoundsMax[4];
} rex_geometry_v0;

typedef struct rex_instance_v0 {
    float objectToWorld[16];     /* rows */
    float worldToObject[16];
    float previousObjectToWorld[16];
    uint32_t metadata[4];        /* first geometry, geometry count, stable instance ID, InstanceFlags */
    uint32_t reserved0[4];
} rex_instance_v0;

typedef struct rex_material_v0 {
    float baseColorFactor[4];
    float emissiveFactorStrength[4];   /* rgb factor, w = strength */
    float surfaceParams[4];            /* metallic, roughness, normal scale, alpha cutoff */
    float transmissionParams[4];       /* transmission, IOR, -, - */
    float attenuationColorDistance[4];
    uint32_t textureImageIndices[4];
    uint32_t textureSamplerIndices[4];
    uint32_t metadata[4];              /* MaterialModel, MaterialFlags, stable ID, - */
} rex_material_v0;

typedef struct rex_light_v0 {
    float positionRange[4];
    float directionCosOuter[4];
    float radianceScale[4];
    float shapeParams[4];
    uint32_t identity[4];        /* LightType, stable light ID, instance ID, primitive ID */
    uint32_t extra[4];           /* texture ID, LightFlags, distribution offset / count */
} rex_light_v0;

/* Scene::CanonicalSceneView (include/scene/CanonicalScene.hpp:70-83) without
 * the stable ID string and camera presets. */
typedef struct rex_canonical_scene {
    uint64_t fingerprint;        /* nonzero */
    uint32_t generation;         /* nonzero */
    uint32_t reserved;
    const rex_scene_constants_v0* constants;
    const rex_vertex_v0* vertices;
    const uint32_t* indices;
    const rex_geometry_v0* geometries;
    const rex_instance_v0* instances;
    const rex_material_v0* materials;
    const rex_light_v0* lights;
    uint32_t vertexCount, indexCount, geometryCount, instanceCount;
    uint32_t materialCount, lightCount;
} rex_canonical_scene;

/* Expands the visible instances' geometry into world-space canonical
 * triangles in the reference's order (instance, geometry, triangle), the
 * rex_scene_create input: BuildCanonicalTraversalScene
 * (src/rt/gpu/CanonicalTraversalScene.cpp:73-186), bit for bit (positions
 * by the rows, normals by the inverse transform's columns normalised in
 * fp64, texcoords copied, identity / metadata tuples).  With out == NULL
 * only *count is set.  REX_INVALID_ARGUMENT for the reference's rejections
 * (invalid instance, geometry, material, index or vertex ranges, non-finite
 * transforms, no visible triangle). */
rex_status rex_canonical_expand(const rex_canonical_scene* scene, rex_triangle* out, uint32_t capacity,
                                uint32_t* count);
/* The L6 material and light records of the canonical materials and lights
 * (materialCount / lightCount entries each), the host conversion of
 * Wave2Runtime::SetScene (src/renderers/Wave2Runtime.cpp:469-556):
 * emission = factor * strength, f0 0.04, model 5 / 4 / 3 for
 * metallic-roughness / transmissive / smooth dielectric with their lobes,
 * pure-emitter and thin-walled flags; a light's identity and payload
 * rearranged, an emissive triangle's payload.x = the index of the first
 * expanded triangle with its (instance, primitive). */
rex_status rex_canonical_shading(const rex_canonical_scene* scene, const rex_triangle* triangles,
                                 uint32_t triangleCount, rex_material* materials, rex_light* lights);
/* rex_canonical_expand + rex_scene_create + rex_canonical_shading +
 * rex_scene_set_shading in one call. */
rex_status rex_scene_create_canonical(const rex_canonical_scene* scene, uint32_t maxLeafSize, uint32_t lightSelection,
                                      rex_scene* outScene);
/* Moves the instances of a scene made by rex_scene_create_canonical: `scene`
 * is the same canonical scene with new instance transforms (objectToWorld,
 * worldToObject, previousObjectToWorld; the same visible instances and
 * geometry, so the canonical triangle stream keeps its order and traversal
 * IDs).  Expands it, refits the hierarchy to the new positions
 * (rex_scene_refit, which advances the scene to `generation`), converts the
 * materials and lights again (moved emissive instances move their triangle
 * lights) with the scene's light selection, and replaces the instance table
 * that motion inputs and ReSTIR's object motion read.  The ReSTIR scene
 * generation of rex_restir_desc stays the caller's: keep it when history
 * should survive the motion.  No counterpart in the reference, whose
 * runtimes take a new scene through BuildOrUpdateScene
 * (include/rt/gpu/IGpuTraversalBackend.hpp:72-75). */
rex_status rex_scene_refit_canonical(rex_scene scene, const rex_canonical_scene* canonical, uint32_t generation);

/* ---- wavefront frames ---------------------------------------------------- */

typedef struct rex_renderer_t* rex_renderer;

typedef struct rex_camera {
    float position[3];
    float tanHalfFov;
    float forward[3];
    float aspect;
    float right[3];
    float up[3];
} rex_camera;

#define REX_ESTIMATOR_BSDF_ONLY 0u
#define REX_ESTIMATOR_NEE 1u
#define REX_ESTIMATOR_MIS 2u
/* PBR_L6_ESTIMATOR_RESTIR_PRIMARY: NEE + MIS everywhere except the primary
 * vertex's non-delta direct light on surfaces ReSTIR owns (opaque
 * metallic-roughness), which ReSTIR DI supplies after the bounces
 * (integrators/wavefront/shaders/Shade.hlsl:78-106, 210-258;
 * PbrWavefrontShading.hlsli:195-199).  Requires spp 1. */
#define REX_ESTIMATOR_RESTIR 3u

/* One launch: width x height pixels, spp samples each (sample index =
 * sampleIndex + s), maxBounce path vertices.  Russian roulette, ray epsilon
 * and the stream tag follow the reference's production constants when left
 * zero: RR start 3, survival in [0.05, 0.95], epsilon 1e-4, tag 0x57a20001
 * (src/renderers/Wave2Runtime.cpp:1410, 1430-1434). */
typedef struct rex_frame_desc {
    rex_camera camera;
    uint32_t width, height;
    uint32_t spp;
    uint32_t maxBounce;
    uint32_t sampleIndex;
    uint32_t estimator;
    uint64_t seed;
    uint32_t streamTag;
    uint32_t walkMode;           /* REX_WALK_SHORT_STACK (the reference's left-first order), REX_WALK_NEAR_FIRST,
                                    REX_WALK_WIDE or REX_WALK_WIDE_COMPRESSED */
    float rrStart, rrMinimum, rrMaximum, rayEpsilon;
    uint32_t queueMode;          /* REX_QUEUE_ATOMIC_APPEND (0, production) or REX_QUEUE_PREFIX_SCAN */
    uint32_t reserved;
} rex_frame_desc;
/* Queue modes of the wavefront frame (Wavefront::QueueMode,
 * integrators/wavefront/include/WavefrontSchedule.hpp; the frame constants'
 * capacityModeSeed.y): Shade appends the next rays and shadow queries with
 * block-aggregated atomics (their order varies between runs), or writes each
 * queue item's candidates at its own index and a stable compaction
 * (ScanBlocksCS / AddOffsetsCS / ScatterCS) moves them into the queues in
 * item order.  Results are identical; the batch engine and the request loop
 * use atomic appends only. */
#define REX_QUEUE_ATOMIC_APPEND 0u
#define REX_QUEUE_PREFIX_SCAN 1u

/* Device path state, queues, Film and traversal workspace for up to
 * maxPixels * maxSpp paths.  Renderers of one scene may record frames on
 * different streams at once. */
rex_status rex_renderer_create(rex_scene scene, uint32_t maxPixels, uint32_t maxSpp, rex_renderer* out);
rex_status rex_renderer_destroy(rex_renderer renderer);
/* Points the renderer at another shaded scene, keeping its buffers; the
 * Film restarts and any ReSTIR history is dropped.  No counterpart in the
 * reference (a Wave2Runtime is rebuilt per scene); for request loops that
 * keep a pool of renderers across jobs of many resident scenes.  The
 * renderer's earlier frames must be complete or ordered before `stream`. */
rex_status rex_renderer_bind_scene(rex_renderer renderer, rex_scene scene, void* stream);
/* Clears the Film (progressive mean restarts at count 0). */
rex_status rex_renderer_reset_film(rex_renderer renderer, void* stream);

/* Renders one launch (RayGen, per bounce Intersect / Shade / TraceShadow,
 * Resolve + Film update) on `stream`, replacing Wave2Runtime::UpdateFrame +
 * RecordFrame for the Wavefront architecture
 * (src/renderers/Wave2Runtime.cpp:1346-1585, Wave3DebugRuntime.cpp:551-696).
 * keepSignals != 0 also stores the five per-sample signals. */
rex_status rex_render_frame(rex_renderer renderer, const rex_frame_desc* frame, int keepSignals, void* stream);

typedef struct rex_frame_stats {
    uint64_t closestRays;   /* closest-hit queries (all bounces) */
    uint64_t shadowRays;    /* any-hit queries */
    uint32_t raysPerBounce[16];
    uint32_t shadowsPerBounce[16];
    uint32_t pathCount;
    uint32_t bounces;
    /* With REX_PROFILE_COUNTERS: the traversal counters (REX_COUNTER_*) of
     * the frame's closest-hit and any-hit queries (the software traversal
     * telemetry of Wave2RuntimeTelemetry, include/renderers/Wave2Runtime.hpp:52-79). */
    uint64_t closestCounters[REX_COUNTER_COUNT];
    uint64_t shadowCounters[REX_COUNTER_COUNT];
    /* With REX_PROFILE_TIMING: GPU milliseconds per stage, from events
     * recorded on the frame's stream between launches: [0] reset + RayGen,
     * [1] Intersect (all bounces), [2] Shade, [3] TraceShadow, [4] Resolve.
     * Without REX_PROFILE_COUNTERS a bounce's shadow queries run in the same
     * launch as the next bounce's closest-hit queries; that launch counts
     * under [1] and only the last bounce's shadow launch under [3]. */
    float stageMs[6];       /* [5]: the ReSTIR DI passes (estimator REX_ESTIMATOR_RESTIR) */
    /* With REX_PROFILE_TIMING: GPU milliseconds of each traversal launch in
     * launch order (bounce 0's closest-hit rays first, then the launches of
     * [1] and [3] as they ran), traceLaunches of them. */
    float traceMs[17];
    uint32_t traceLaunches;
} rex_frame_stats;
/* The stable compaction of the PrefixScan queue mode on caller data: the
 * device counterpart of the reference's CPU oracle CompactStable /
 * BlellochExclusiveScan (integrators/wavefront/src/WavefrontCompactionOracle.cpp:9-81)
 * and the dispatches ScanBlocksCS / AddOffsetsCS / ScatterCS.  flags: count
 * (next, shadow) uint32 pairs; prefix (may be NULL): their exclusive prefix
 * sums (count pairs); nextIndices / shadowIndices (may be NULL): the indices
 * of the flagged elements in element order, up to the capacities; status:
 * 5 words = attempted next, attempted shadow, active next = min(attempted,
 * capacity), active shadow, fatal mask (REX_COMPACT_FATAL_*: a flag above 1
 * counts as 0 and raises SCAN_OVERFLOW; a total above a capacity raises that
 * queue's overflow).  scratch: rex_compact_scratch_bytes(count) device bytes.
 * All pointers are device memory; the call is asynchronous on `stream`. */
#define REX_COMPACT_FATAL_SCAN_OVERFLOW 1u
#define REX_COMPACT_FATAL_NEXT_OVERFLOW 2u
#define REX_COMPACT_FATAL_SHADOW_OVERFLOW 4u
uint64_t rex_compact_scratch_bytes(uint32_t count);
rex_status rex_compact_stable(const uint32_t* flags, uint32_t count, uint32_t nextCapacity, uint32_t shadowCapacity,
                              void* scratch, uint32_t* prefix, uint32_t* nextIndices, uint32_t* shadowIndices,
                              uint32_t* status, void* stream);

/* Synchronizes the stream and reads the last frame's queue counters. */
rex_status rex_renderer_stats(rex_renderer renderer, rex_frame_stats* stats);

#define REX_PROFILE_COUNTERS 1u
#define REX_PROFILE_TIMING 2u
/* Selects what later frames record for rex_renderer_stats (REX_PROFILE_*
 * flags; 0 = nothing, the default).  Counters add one atomic per wave per
 * counter to the traversal kernels. */
rex_status rex_renderer_set_profiling(rex_renderer renderer, uint32_t flags);

/* Device pointers of the Film (float4 per pixel: rgb mean, sample count) and
 * of the last frame's signals (5 float4 per path: camera emission, direct
 * diffuse, direct specular, indirect diffuse, indirect specular). */
rex_status rex_renderer_buffers(rex_renderer renderer, void** film, void** signals);

/* ---- ReSTIR DI ----------------------------------------------------------- */

#define REX_RESTIR_NEIGHBORS_RING 0u  /* src/renderers/Wave3DebugRuntime.cpp:1888-1925 */
#define REX_RESTIR_NEIGHBORS_DISK 1u  /* seeded disk of `radius` pixels (no reference counterpart) */

/* ReSTIR DI settings: RuntimeConfig::restir (include/app/RuntimeConfig.hpp:174-187)
 * plus the identity generations of ReSTIRFrameRequest
 * (src/renderers/Wave3DebugRuntime.cpp:2030-2097).  Zero generations select
 * 1 (the scene generation: the scene's own); all-zero validation selects
 * the production thresholds {0.8, 0.05, 0.02, 0.005}
 * (Wave3DebugRuntime.cpp:2085).  Only the Physical shadow mode (one any-hit
 * visibility ray per pixel) and the explicitly biased estimator mode are
 * provided. */
typedef struct rex_restir_desc {
    uint32_t candidates;          /* initial RIS candidates per pixel, 1..256 */
    uint32_t neighbors;           /* spatial neighbours, 0 (no spatial stage)..32 */
    uint32_t maxM;                /* reservoir M cap, >= 1 */
    uint32_t maxAge;              /* temporal history age cap */
    uint32_t temporal;            /* nonzero: temporal reuse stage */
    uint32_t neighborMode;        /* REX_RESTIR_NEIGHBORS_* */
    uint32_t radius;              /* disk radius in pixels, 1..1024 (disk mode) */
    uint32_t sceneGeneration;
    uint32_t lightGeneration;
    uint32_t configGeneration;    /* a change between frames is a camera cut */
    uint32_t resourceGeneration;  /* a change between frames is a resize */
    uint32_t historyGeneration;
    float validation[4];          /* normal cosine, relative depth, position, thin */
    uint32_t estimatorMode;       /* REX_RESTIR_ESTIMATOR_* */
    uint32_t objectMotion;        /* nonzero: temporal reuse reprojects each primary surface through its
                                     instance's objectToWorld and previousObjectToWorld (canonical scenes,
                                     rex_instance_v0), as MotionVectorsCS; zero or a scene without an
                                     instance table: surfaces at rest, the two cameras only */
    uint32_t reserved[2];
} rex_restir_desc;

/* ReSTIREstimatorMode (src/renderers/ReSTIRDIRuntime.cpp:112-143).  The
 * reference-correction estimator adds RestirReferenceVisibilityPrepareV3 /
 * RestirReferenceResolveV3 after spatial reuse: the selected sample's
 * visibility from the pixel and from each of its K neighbours (one any-hit
 * ray each), and W = wSum * target_src * vis_src / (target_center *
 * sum over sources of M_s * target_s * vis_s). */
#define REX_RESTIR_ESTIMATOR_BIASED 0u
#define REX_RESTIR_ESTIMATOR_REFERENCE_CORRECTION 1u

/* Sets the ReSTIR DI settings of later REX_ESTIMATOR_RESTIR frames and drops
 * any history (the next frame does no temporal reuse).  Without a call the
 * renderer uses {32 candidates, 5 ring neighbours, M cap 20, age 20,
 * temporal on}.  The ReSTIR buffers (about 1.7 KB per pixel) are allocated
 * at the first ReSTIR frame.  The reference-correction estimator adds
 * (K + 1) x 160 B per pixel of visibility queries (K = neighbours; 11 GB at
 * K = 32 and 1920 x 1080), allocated at its first frame and freed when the
 * settings return to the biased estimator. */
rex_status rex_renderer_set_restir(rex_renderer renderer, const rex_restir_desc* desc);
/* Drops the history: the next ReSTIR frame does no temporal reuse (the frame
 * index keeps counting). */
rex_status rex_renderer_reset_restir(rex_renderer renderer);

/* Device buffers of the last ReSTIR frame, in the reference's ABI records
 * (include/contracts/RestirAbiV3.hpp:109-193, ReconstructionAbiV2.hpp). */
typedef struct rex_restir_buffers {
    void* surfaces;      /* GpuPrimarySurfaceV2[pixels] (96 B) */
    void* initial;       /* GpuRestirReservoirV3[pixels] (256 B) after initial RIS */
    void* temporal;      /* after temporal reuse */
    void* spatial;       /* after spatial reuse and winner visibility */
    void* history;       /* published for the next frame */
    void* direct;        /* float4[2 * pixels]: ReSTIR direct diffuse, specular */
    void* reasons;       /* uint32[pixels]: rejection of the last reuse stage */
    void* statistics;    /* GpuRestirStatisticsV3 (64 B) */
    uint64_t frameIndex; /* the last ReSTIR frame's index (0 for the first) */
    uint32_t historyValid;  /* nonzero: the last frame reused history */
    uint32_t reserved;
} rex_restir_buffers;
rex_status rex_renderer_restir_buffers(rex_renderer renderer, rex_restir_buffers* out);

/* ---- reconstruction boundary (ABI-v2) ------------------------------------- */

/* The ABI-v2 reconstruction records (include/contracts/ReconstructionAbiV2.hpp:45-116),
 * byte for byte. */
typedef struct rex_primary_surface_v2 {
    float worldPositionLinearDepth[4];
    float geometricNormalRoughness[4];
    float shadingNormalMetallic[4];
    float diffuseAlbedo[4];
    float specularAlbedo[4];
    uint32_t identity[4];        /* material, instance, primitive, REX_PRIMARY_SURFACE_* flags */
} rex_primary_surface_v2;        /* 96 B */

typedef struct rex_motion_vector_v2 {
    float motionExpectedDepth[4];    /* previous UV - current UV, previous linear depth, 0 */
    float currentPreviousUv[4];      /* current UV, previous UV */
    uint32_t identity[4];            /* material, instance, primitive, REX_MOTION_* flags */
} rex_motion_vector_v2;          /* 48 B */

typedef struct rex_reconstruction_signal_v2 {
    float directDiffuse[4];
    float directSpecular[4];
    float indirectDiffuse[4];
    float indirectSpecular[4];
} rex_reconstruction_signal_v2;  /* 64 B */

typedef struct rex_gbuffer_record_v2 {
    rex_primary_surface_v2 primary;
    rex_motion_vector_v2 motion;
    rex_reconstruction_signal_v2 signal;
} rex_gbuffer_record_v2;         /* 208 B */

/* PbrMotionInputV2 (resources/shaders/integrators/pbr_reconstruction_export_v2.hlsli:9-15). */
typedef struct rex_motion_input_v2 {
    float objectPosition[3];
    uint32_t currentTransformIndex;
    uint32_t previousTransformIndex;
    uint32_t reserved[3];
} rex_motion_input_v2;           /* 32 B */

#define REX_PRIMARY_SURFACE_VALID 1u
#define REX_PRIMARY_SURFACE_FRONT_FACE 2u
#define REX_PRIMARY_SURFACE_HAS_DIFFUSE 4u
#define REX_PRIMARY_SURFACE_HAS_SPECULAR 8u
#define REX_MOTION_VALID 1u

/* Caller-owned device buffers that every later rex_render_frame of the
 * renderer publishes into (the reference's set-4 reconstruction bindings,
 * pbr_reconstruction_export_v2.hlsli:17-32); any pointer may be NULL.
 * Per pixel, for pixels < width * height of the frame:
 *   gbuffer         primary surface of the bounce-0 hit (PbrPublishPrimarySurfaceV2,
 *                   :44-88, valid flag set), motion cleared, signal = rawSignal;
 *                   all zero for a miss (PbrClearReconstructionInputsV2, :34-42)
 *   motionInputs    object-space position of the hit and the index of its
 *                   instance in the scene's instance table (:90-129); scenes
 *                   without one (rex_scene_create) use the world position and
 *                   index 0; requires gbuffer
 *   rawSignal       direct / indirect diffuse / specular (PbrPublishReconstructionSignalV2, :132-149),
 *                   after ReSTIR's primary-direct publication
 *   primarySurface  the ReSTIR copy of the primary surface: valid only where
 *                   ReSTIR owns the surface's direct light (:85-88)
 *   terminal        float4 camera-visible emission / background, w = 0
 * With spp > 1 the signals are the mean of the pixel's samples and the
 * surface is that of its first sample (the reference renders 1 spp per
 * launch, where both are the one path's).  pixels = the capacity of every
 * bound buffer.  NULL targets unbind.  Not used by batch engines. */
typedef struct rex_reconstruction_targets_v2 {
    void* gbuffer;          /* rex_gbuffer_record_v2[pixels] */
    void* motionInputs;     /* rex_motion_input_v2[pixels] */
    void* rawSignal;        /* rex_reconstruction_signal_v2[pixels] */
    void* primarySurface;   /* rex_primary_surface_v2[pixels] */
    void* terminal;         /* float4[pixels] */
    uint32_t pixels;
    uint32_t reserved;
} rex_reconstruction_targets_v2;
rex_status rex_renderer_set_reconstruction_v2(rex_renderer renderer, const rex_reconstruction_targets_v2* targets);

/* The renderer's camera matrices, rows: view-projection (positive clip w in
 * front of the camera, near 0.01, far 10000) and world-to-view
 * (right-handed, looking down -z; linear depth = -view.z). */
rex_status rex_camera_matrices(const rex_camera* camera, float viewProjection[16], float worldToView[16]);

/* L8MotionConstants (reconstruction/shaders/MotionVectorsCS.hlsl:16-26). */
typedef struct rex_motion_constants_v2 {
    float currentViewProjection[16];    /* rows */
    float previousViewProjection[16];
    float previousWorldToView[16];
    float currentJitterUv[2];
    float previousJitterUv[2];
    uint32_t width, height;
    uint32_t currentTransformCount, previousTransformCount;
} rex_motion_constants_v2;

/* MotionVectorsCS (reconstruction/shaders/MotionVectorsCS.hlsl:35-90) over
 * width x height pixels on stream: fills gbuffer[i].motion from
 * motionInputs[i] and the instances' current / previous object-to-world
 * rows (device float[16 * count] each); invalid surfaces, out-of-range
 * transform indices, non-finite inputs and points behind either camera
 * leave a cleared motion record. */
rex_status rex_motion_vectors_v2(void* gbuffer, const void* motionInputs, const float* currentObjectToWorld,
                                 const float* previousObjectToWorld, const rex_motion_constants_v2* constants,
                                 void* stream);

/* ComposeCS output modes (reconstruction/shaders/ReconstructionCommon.hlsli:95-110). */
#define REX_COMPOSE_RAW 0u
#define REX_COMPOSE_MOTION 4u
#define REX_COMPOSE_BASE_COLOR 10u
#define REX_COMPOSE_NORMAL 11u
#define REX_COMPOSE_ROUGHNESS 12u
#define REX_COMPOSE_METALLIC 13u
#define REX_COMPOSE_EMISSIVE 14u
#define REX_COMPOSE_PROGRESSIVE_MEAN 15u

/* ComposeCS (reconstruction/shaders/ComposeCS.hlsl:26-126) without the
 * temporal / a-trous / SVGF history modes (1-3, 5-9: REX_UNSUPPORTED).
 * Device buffers of width * height pixels; output is float4 (w = 1).
 * RAW / PROGRESSIVE_MEAN read rawSignal and terminal; PROGRESSIVE_MEAN
 * also updates film (float4: mean, sample count in w) with
 * filmSampleCount frames already in it; MOTION / NORMAL / ROUGHNESS /
 * METALLIC read gbuffer; BASE_COLOR / EMISSIVE read gbuffer and the
 * ABI-v0 materials (materialCount records). */
typedef struct rex_compose_desc_v2 {
    uint32_t width, height;
    uint32_t outputMode;
    uint32_t filmSampleCount;
    const void* gbuffer;        /* rex_gbuffer_record_v2[] */
    const void* rawSignal;      /* rex_reconstruction_signal_v2[] */
    const void* terminal;       /* float4[] */
    const void* materials;      /* rex_material_v0[materialCount] */
    uint32_t materialCount;
    uint32_t reserved;
    void* film;                 /* float4[] */
    void* output;               /* float4[] */
} rex_compose_desc_v2;
rex_status rex_compose_v2(const rex_compose_desc_v2* desc, void* stream);

/* ---- batched jobs ------------------------------------------------------- */

/* A batch engine runs many render jobs (rex_render_frame frames of any
 * resident scene, resolution, spp and max bounce) through one shared set of
 * wavefront launches: all jobs' rays share the ray and shadow queues, each
 * queue item names its scene in an arena of resident BVHs, and every engine
 * step is one launch sequence (admitted jobs' RayGen, traversal of all
 * shadow queries and rays, Resolve of the jobs that finished, Shade), so a
 * step costs the same handful of launches for 1 job or 64.  Paths live in
 * a pool: each step retires the paths that ended and starts new ones from
 * the admitted jobs' sample streams in their place (continuous batching at
 * path granularity), so a job of any size streams through a pool of any
 * size.  Jobs are admitted in submission order while a job slot and room
 * for their Film and per-sample values in the engine's output arena are
 * free.  Each job's Film is bit-identical to the same frame rendered alone
 * by rex_render_frame.  The
 * reference has no batched form (its Wave2Runtime renders one frame per
 * launch sequence); callers are render farms and request loops. */
typedef struct rex_batch_t* rex_batch;

typedef struct rex_batch_desc {
    uint32_t pathCapacity;    /* paths in flight (0 selects 16M) */
    uint32_t maxJobs;         /* jobs admitted at once (0 selects 256) */
    uint32_t useGraphs;       /* nonzero: each step replays a captured hipGraph */
    uint32_t shadeBlocksPerCu;   /* Shade grid: blocks per CU (0: autotuned) */
    uint32_t traceBlocksPerCu;   /* traversal grid: blocks per CU (0: occupancy) */
    uint32_t outputMiB;       /* device arena for the jobs' Films and per-sample values
                                 (16 B per pixel + 16 B per sample; 0 selects 16384 MiB) */
    uint32_t autotune;        /* nonzero: the first busy steps try Shade grids of 16 and 8 blocks per
                                 CU with graphs off and on at the full path budget, then path budgets
                                 of 1/2 and 3/4 of pathCapacity with the best of those (windows of 2
                                 warm-up + 4 timed steps, scored by traversal queries per second),
                                 and keep the best setting; results do not depend on it */
    uint32_t reserved;
} rex_batch_desc;

rex_status rex_batch_create(const rex_batch_desc* desc, rex_batch* out);
rex_status rex_batch_destroy(rex_batch batch);
/* Makes a shaded scene resident: copies its 4-wide hierarchy and triangles
 * into the engine's arena (the scene's shading tables are referenced and
 * must outlive the engine).  Not while jobs are in flight. */
rex_status rex_batch_add_scene(rex_batch batch, rex_scene scene, uint32_t* sceneIndex);
/* Queues one frame of a resident scene (estimator BSDF-only / NEE / MIS);
 * walkMode is ignored (the 4-wide walk).  The engine owns the job's Film,
 * allocated from the output arena at admission; REX_UNSUPPORTED when the
 * job's outputs could never fit the arena. */
rex_status rex_batch_submit(rex_batch batch, uint32_t sceneIndex, const rex_frame_desc* frame, uint64_t* jobId);
/* One engine step on `stream` (asynchronous): admits queued jobs that fit,
 * advances every path in flight by one bounce, resolves finished jobs. */
rex_status rex_batch_step(rex_batch batch, void* stream);
/* Steps until every submitted job has finished, then synchronizes. */
rex_status rex_batch_run(rex_batch batch, void* stream);

typedef struct rex_batch_job_info {
    uint32_t state;          /* 0 queued, 1 running, 2 resolve issued, 3 finished (GPU done) */
    uint32_t width, height;
    uint32_t reserved;
    uint64_t submitStep;     /* engine step counter at submit / finish */
    uint64_t finishStep;
    void* film;              /* device float4 per pixel (rgb mean, count); NULL until admitted */
} rex_batch_job_info;
rex_status rex_batch_job(rex_batch batch, uint64_t jobId, rex_batch_job_info* info);
/* Waits for a finished job's Film to be written, returns its outputs to
 * the arena and forgets the job. */
rex_status rex_batch_release(rex_batch batch, uint64_t jobId);

typedef struct rex_batch_stats {
    uint64_t steps;          /* steps that launched work */
    uint64_t jobsFinished;
    uint64_t pathsInFlight;  /* path slots held by admitted jobs */
    uint64_t queuedJobs;
    uint32_t shadeBlocksPerCu, traceBlocksPerCu;
    uint32_t graphs;         /* nonzero when steps replay hipGraphs */
    uint32_t autotune;       /* 0 off, 1 tuning, 2 tuned (the fields above and pathBudget hold the choice) */
    uint64_t outputBytesInUse;   /* output arena held by admitted and unreleased jobs */
    uint64_t outputBytesTotal;
    uint64_t raysTraced;         /* traversal queries (closest + any-hit) of all steps so far */
    uint64_t pathBudget;         /* paths the engine keeps in flight at most */
} rex_batch_stats;
/* Waits for the device (for the ray total). */
rex_status rex_batch_get_stats(rex_batch batch, rex_batch_stats* stats);

/* ---- serving ------------------------------------------------------------- */
/* A request loop over the batch engine and a pool of renderers.  Jobs arrive
 * at any time and are admitted in arrival order while their HBM footprint
 * fits the budget next to the resident scenes; admitted work overlaps on the
 * device: frames (BSDF-only / NEE / MIS) share the batch engine's launches
 * on its stream, and each ReSTIR DI sequence renders its frames back to back
 * on a renderer of its own stream, keeping reservoir and Film history on the
 * device between frames.  A job's result equals the same job run alone (the
 * frame's Film bit for bit, the sequence's last Film bit for bit).  No
 * counterpart in the reference, which renders one frame per call; for render
 * farms and interactive request streams. */
typedef struct rex_server_t* rex_server;

#define REX_JOB_FRAME 0u
#define REX_JOB_RESTIR_SEQUENCE 1u

typedef struct rex_server_desc {
    uint64_t budgetBytes;        /* resident scenes + admitted jobs (0 selects 200 GiB) */
    uint32_t pathCapacity;       /* batch engine path pool (0 selects 16M) */
    uint32_t maxJobs;            /* frames in the engine at once (0 selects 256) */
    uint32_t useGraphs;          /* nonzero: engine steps replay hipGraphs */
    uint32_t shadeBlocksPerCu;   /* engine Shade grid (0: 16 per CU) */
    uint32_t outputMiB;          /* engine output arena (0 selects 16384 MiB) */
    uint32_t sequenceSlots;      /* ReSTIR sequences in flight (0 selects 4) */
    uint32_t sequenceMaxPixels;  /* pixels of a sequence renderer (0 selects 1920 x 1080) */
    uint32_t sequenceOutputMiB;  /* arena for finished sequences' Films (0 selects 2048 MiB) */
    uint32_t reserved[2];
} rex_server_desc;

rex_status rex_server_create(const rex_server_desc* desc, rex_server* out);
rex_status rex_server_destroy(rex_server server);
/* Makes a shaded scene resident for both job kinds (the scene must outlive
 * the server).  Before the first submit. */
rex_status rex_server_add_scene(rex_server server, rex_scene scene, uint32_t* sceneIndex);

typedef struct rex_server_job_desc {
    uint32_t kind;               /* REX_JOB_* */
    uint32_t sceneIndex;
    uint32_t frames;             /* sequence length (REX_JOB_RESTIR_SEQUENCE, 1..64); frames: 1 */
    uint32_t reserved;
    rex_frame_desc frame;        /* the frame; a sequence's frame k uses sampleIndex + k */
    const rex_camera* cameras;   /* sequence: `frames` cameras, copied at submit (NULL: frame.camera) */
    const rex_restir_desc* restir;  /* sequence: ReSTIR settings, copied (NULL: the defaults) */
} rex_server_job_desc;
/* Queues a job (host only; the submit time is the job's arrival). */
rex_status rex_server_submit(rex_server server, const rex_server_job_desc* job, uint64_t* jobId);
/* One turn of the loop: notes finished jobs, admits what fits, issues one
 * engine step.  *finished (may be NULL) = jobs that finished in this call. */
rex_status rex_server_pump(rex_server server, uint32_t* finished);
/* Pumps until every submitted job has finished. */
rex_status rex_server_drain(rex_server server);

typedef struct rex_server_job_info {
    uint32_t state;              /* 0 queued, 1 admitted, 2 finished */
    uint32_t kind;
    uint32_t width, height;
    uint64_t footprintBytes;     /* charged to the budget from admission to release */
    int64_t submitNs, admitNs, finishNs;  /* host steady clock; finish = first pump that saw it done */
    uint64_t rays;               /* sequences: traversal queries of all frames; frames: 0 */
    void* film;                  /* device Film (float4 per pixel) once finished */
} rex_server_job_info;
rex_status rex_server_job(rex_server server, uint64_t jobId, rex_server_job_info* info);
/* Returns a finished job's memory to the budget and forgets it. */
rex_status rex_server_release(rex_server server, uint64_t jobId);

typedef struct rex_server_stats {
    uint64_t submitted, admitted, finished, released;
    uint64_t queuedJobs;
    uint64_t residentBytes;      /* scenes: hierarchies, triangles, shading tables, engine arena copies */
    uint64_t admittedBytes;      /* footprints of admitted, unreleased jobs */
    uint64_t peakAdmittedBytes;
    uint64_t budgetBytes;
    uint64_t engineSteps;
    uint64_t engineRays;         /* traversal queries of the engine's frames */
    uint64_t sequenceRays;       /* traversal queries of finished sequences */
    uint64_t sequenceFrames;
} rex_server_stats;
/* Waits for the device (for the engine's ray total). */
rex_status rex_server_get_stats(rex_server server, rex_server_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
