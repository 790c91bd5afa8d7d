// Internal declarations shared by the rex host library sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rex.h"
#include "bvh_build.hpp"
#include "render_layout.hpp"
#include "restir_layout.hpp"
#include "rex_layout.hpp"

static_assert(sizeof(rex_triangle) == 176, "CanonicalTraversalTriangle is 176 B");
static_assert(sizeof(rex_ray_v1) == 64, "GpuRayQueueRecordV1 is 64 B");
static_assert(sizeof(rex_hit_v1) == 96, "GpuHitQueueRecordV1 is 96 B");
static_assert(offsetof(rex_ray_v1, identity) == 32, "GpuRayQueueRecordV1 identity offset");
static_assert(offsetof(rex_hit_v1, metadata) == 64, "GpuHitQueueRecordV1 metadata offset");
static_assert(sizeof(rex_walk_hit) == 72, "walk hit record is 72 B");
static_assert(sizeof(rex_material) == 128, "PbrMaterialGpu is 128 B");
static_assert(sizeof(rex_light) == 96, "PbrLightGpu is 96 B");
static_assert(sizeof(rex_shading_desc) == 48, "shading descriptor is 48 B");
static_assert(sizeof(rex_scene_constants_v0) == 96, "GpuSceneConstantsV0 is 96 B");
static_assert(sizeof(rex_vertex_v0) == 64, "GpuVertexV0 is 64 B");
static_assert(sizeof(rex_geometry_v0) == 64, "GpuGeometryV0 is 64 B");
static_assert(sizeof(rex_instance_v0) == 224, "GpuInstanceV0 is 224 B");
static_assert(sizeof(rex_material_v0) == 128, "GpuMaterialV0 is 128 B");
static_assert(sizeof(rex_light_v0) == 96, "GpuLightV0 is 96 B");
static_assert(sizeof(rex_camera) == 56, "camera is 56 B");
static_assert(sizeof(rex_frame_desc) == 120, "frame desc is 120 B");
static_assert(sizeof(rex_frame_stats) == 408, "frame stats is 408 B");
static_assert(sizeof(rex_restir_desc) == 80, "ReSTIR desc is 80 B");
static_assert(sizeof(rex_restir_buffers) == 80, "ReSTIR buffers record is 80 B");
static_assert(sizeof(rex_batch_desc) == 32, "batch desc is 32 B");
static_assert(sizeof(rex_batch_job_info) == 40, "batch job info is 40 B");
static_assert(sizeof(rex_batch_stats) == 80, "batch stats is 80 B");

namespace rex {

// Runs body(begin, end) over [0, count) in contiguous chunks on the host's
// threads (results do not depend on the split).
template <typename F>
void parallel_ranges(size_t count, F body) {
    const size_t threads = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), 16));
    if (count < 65536 || threads == 1) {
        body(size_t{0}, count);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (count + threads - 1) / threads;
    for (size_t t = 0; t < threads; ++t) {
        const size_t b = t * chunk, e = std::min(count, b + chunk);
        if (b < e) pool.emplace_back([=] { body(b, e); });
    }
    for (std::thread& th : pool) th.join();
}

// Thread-local error text behind rex_last_error().
void set_error(const std::string& msg);
rex_status fail(rex_status code, const std::string& msg);
rex_status check_hip(hipError_t e, const char* what);

// Device scratch of the persistent traversal kernel: the work counter and the
// per-lane stack overflow region.  Owned by a scene (its trace and walk
// calls), a renderer or a batch engine; launches that share one must be
// ordered on one stream.
struct TraceWorkspace {
    uint32_t* counter = nullptr;
    uint2* overflow = nullptr;
    size_t overflowCapacity = 0;  // entries
    hipError_t reserve(size_t overflowEntries);
    void release();
};
// Sizes a workspace for every 4-wide launch over hierarchies of depth up to
// wide4Depth, so later launches never grow it (growing waits for the device).
hipError_t reserve_wide_workspace(TraceWorkspace& ws, uint32_t wide4Depth);

// Uploaded shading tables of a scene (rex_scene_set_shading).
struct ShadingData {
    bool ready = false;
    Material* dMaterials = nullptr;
    Light* dLights = nullptr;
    AliasEntry* dAlias = nullptr;
    float* dPmf = nullptr;
    Fixture* dFixtures = nullptr;
    uint32_t* dEmitterOfTri = nullptr;
    F4* dEnvTexels = nullptr;
    AliasEntry* dEnvRows = nullptr;
    AliasEntry* dEnvColumns = nullptr;
    uint32_t materialCount = 0, lightCount = 0, aliasCount = 0, fixtureCount = 0;
    uint32_t environmentLight = kInvalid;
    uint32_t envWidth = 0, envHeight = 0, envSampler = 0;
    bool extended = false;  // BSDF models 1-4 or transmission, sphere lights, or a lat-long environment
    uint32_t selection = REX_LIGHT_SELECTION_POWER;
    std::vector<AliasEntry> alias;  // host copies for export
    std::vector<float> pmf;
    std::vector<AliasEntry> envRows, envColumns;
    double envIntegrated = 0.0;
    void release();
    ShadingView view() const;
};

// One uploaded scene.
struct Scene {
    TraceWorkspace ws;
    ShadingData shade;
    uint64_t fingerprint = 0;
    uint32_t generation = 0;
    uint32_t triCount = 0;
    uint32_t nodeCount = 0;
    uint32_t wideCount = 0;
    uint32_t maxDepth = 0;
    uint32_t maxLeaf = 4;
    double buildMs = 0.0;
    uint64_t deviceBytes = 0;
    int device = 0;
    // Host copies kept for export and for the integrator's host tables.
    std::vector<BinaryNode> nodes;       // reference binary nodes, pre-order
    std::vector<uint32_t> parents;       // binary parent per node
    std::vector<uint32_t> order;         // leaf order -> canonical index
    std::vector<rex_triangle> leafTris;  // canonical records in leaf order
    // Traversal hierarchy of the compressed walk: a binned-SAH build over the
    // same triangles with one triangle per leaf, each leaf's `first` being the
    // triangle's leaf-order (reference) position.  Empty when the compressed
    // walk uses the reference hierarchy (leaves of one triangle already, a
    // device-built or refit hierarchy, or a traversal build that failed).
    std::vector<BinaryNode> tnodes;
    // The traversal hierarchy on the device (pre-order nodes, parents,
    // depths), uploaded at the first refit.
    BinaryNode* dTNodes = nullptr;
    uint32_t* dTParents = nullptr;
    uint32_t* dTDepth = nullptr;
    // Device buffers.
    WideNode* dWide = nullptr;
    TriPos* dTris = nullptr;
    TriAttr* dAttrs = nullptr;
    Wide4* dWide4 = nullptr;
    Wide4Q* dWide4Q = nullptr;
    LeafBox* dLeafBox = nullptr;
    uint32_t wide4Count = 0;   // exact 4-wide nodes (reference hierarchy)
    uint32_t wide4qCount = 0;  // compressed 4-wide nodes (traversal hierarchy)
    // The binary hierarchy on the device (pre-order nodes, parents, depths),
    // kept for device refits; host nodes are downloaded when stale.
    BinaryNode* dNodes = nullptr;
    uint32_t* dParents = nullptr;
    uint32_t* dDepth = nullptr;
    bool hostNodesStale = false;
    float deviceBuildMs = 0.0f;   // device time of the last LBVH build / refit
    // Instance table of canonical scenes (motion inputs): world-to-object
    // rows per instance, (stable ID, first index) sorted by ID.
    InstanceXform* dInstances = nullptr;
    uint2* dInstanceMap = nullptr;
    // objectToWorld and previousObjectToWorld rows (four float4 each) per
    // instance, for temporal reuse under object motion.
    float4* dObjectToWorld = nullptr;
    float4* dPrevObjectToWorld = nullptr;
    uint32_t instanceCount = 0, instanceMapCount = 0;
    // Alpha-masked geometry: per material, the outcome of the reference's
    // alpha test for a masked triangle of that material (kAlphaAccept,
    // kAlphaSkip, kAlphaInvalid); hasAlpha is false when the scene was
    // created without an alpha description.  dAlphaSkip mirrors it for the
    // device layout pass (1 = skip), allocated on first use.
    bool hasAlpha = false;
    std::vector<uint8_t> alphaTable;
    uint8_t* dAlphaSkip = nullptr;
    SceneView view{};
};

constexpr uint8_t kAlphaAccept = 0u, kAlphaSkip = 1u, kAlphaInvalid = 2u;

// The reference's alpha test per material (pbr_software_traversal.hlsli:
// 207-248) with an atlas of uniform layers; fills s.alphaTable.
rex_status set_alpha(const rex_alpha_desc* alpha, Scene& s);
// kAlphaAccept for a triangle without ALPHA_MASK flags, else its material's
// outcome (kAlphaInvalid for a material index beyond the table).
uint8_t alpha_outcome(const Scene& s, const rex_triangle& t);
// Every alpha-masked triangle has a valid outcome (REX_UNSUPPORTED otherwise).
rex_status alpha_validate(const Scene& s, const rex_triangle* tris, uint32_t n);

// Validates and builds the hierarchy on the host (fills nodes, order,
// parents, counts); defined in scene.cpp.
rex_status host_build(const rex_triangle* tris, uint32_t n, uint32_t maxLeafSize, uint32_t threads, Scene& s);
// Writes the reference's 48-B flattened node records and the leaf order.
void export_flat(const Scene& s, void* nodes48, uint32_t* primitiveOrder);
// Builds the hierarchy and device layout; defined in scene.cpp.
rex_status scene_build(const rex_triangle* tris, uint32_t n, const rex_scene_desc& desc, Scene& out);
// BuildKarrasLbvh on the host: fills nodes, order, parents, leafTris, counts.
rex_status lbvh_build(const rex_triangle* tris, uint32_t n, Scene& s);
// Device layout (binary / 4-wide / compressed nodes, triangle records, leaf
// boxes) of s.nodes and s.leafTris, uploaded.  On failure the scene keeps its
// previous layout (or none) and nothing is leaked.
rex_status scene_layout(Scene& s);
// RefitBottomUp over the binary nodes: leaf boxes from their leaf-order
// triangles, interior boxes the union of their children.
void refit_nodes(std::vector<BinaryNode>& nodes, const std::vector<rex_triangle>& leafTris);
// The refit triangles in leaf order, checked against the scene's identities.
rex_status refit_leaf_tris(const Scene& s, const rex_triangle* tris, uint32_t n, std::vector<rex_triangle>& out);
void scene_release(Scene& s);
// Device hierarchy build and layouts; defined in lbvh.hip.
// LBVH on the device from validated canonical triangles and the scene box.
rex_status lbvh_build_device(const rex_triangle* tris, uint32_t n, const float slo[3], const float shi[3],
                             Scene& s);
// Boxes of the scene's hierarchy refit on the device from the leaf-order
// triangles, then the device layout; the generation advances on success.
rex_status refit_device(Scene& s, std::vector<rex_triangle>&& leafTris, uint32_t generation);
rex_status device_layout(Scene& s, const rex_triangle* dLeafTris, uint32_t n, hipStream_t st);
rex_status upload_hierarchy(Scene& s);
void release_hierarchy(Scene& s);
// Refits the traversal hierarchy (if the scene has one) on the device from
// the leaf-order triangles and replaces the compressed nodes with its own;
// on failure the compressed walk keeps the reference hierarchy's nodes.
void refit_traversal(Scene& s, const rex_triangle* dLeafTris, hipStream_t st);
// Brings s.nodes up to date after a device refit.
rex_status ensure_host_nodes(Scene& s);
// Validation of LBVH input (finite, non-degenerate, unique traversal IDs)
// and the scene box of all vertices.
rex_status lbvh_validate(const rex_triangle* tris, uint32_t n, float slo[3], float shi[3]);

// Kernel launchers; defined in trace_kernels.hip.
hipError_t launch_walk(const SceneView& s, TraceWorkspace& ws, const float* rays, const uint32_t* keys,
                       uint32_t count, bool shadow, uint32_t mode, rex_walk_hit* hits, uint64_t* counters,
                       hipStream_t stream);
// wide: walk the exact 4-wide nodes instead of the reference's binary
// near-first order (not the compressed ones: their conservative boxes can
// admit an occluder hit within an ulp of tMax that the exact walk's box
// rounding culls, which would change a ReSTIR visibility bit).  Only for any-hit queries whose caller reads the hit kind alone, on
// a hierarchy shallow enough that the binary walk cannot overflow its stack
// (depth below stackCapacity): the kind is then the same for both walks.
hipError_t launch_trace_v1(const SceneView& s, TraceWorkspace& ws, const rex_ray_v1* rays, rex_hit_v1* hits,
                           uint32_t count, bool any, uint32_t stackCapacity, uint64_t* counters, hipStream_t stream,
                           bool wide = false);
// The batch tracer (near-first binary walk, as launch_trace_v1 without
// `wide`) with its work counter and overflow stacks in the library's static
// device scratch: no allocation, at most 512 resident blocks.  Launches that
// overlap on one device share the scratch and race.
// Any-hit visibility of ABI-v1 rays on the compressed walk (ReSTIR): writes
// only each hit record's metadata words (ray ID, kind, 0, path ID).
hipError_t launch_visibility_v1(const SceneView& s, TraceWorkspace& ws, const rex_ray_v1* rays, rex_hit_v1* hits,
                                uint32_t count, hipStream_t stream);
hipError_t launch_trace_v1_static(const SceneView& s, const rex_ray_v1* rays, rex_hit_v1* hits, uint32_t count,
                                  bool any, uint32_t stackCapacity, uint64_t* counters, hipStream_t stream);
// Stable compaction of a bounce's dense Shade output (PrefixScan queue mode);
// defined in compaction.hip.  countPtr: elements (device); the next-ray and
// shadow queue sizes are written to nextCount / shadowCount.
struct CompactQueues {
    const uint2* flags;
    const uint32_t* countPtr;
    uint32_t capacity;
    uint2* tiles;
    uint32_t* nextCount;
    uint32_t* shadowCount;
    const F4* denseRays;
    const uint32_t* densePath;
    const ShadowItem* denseShadows;
    F4* nextRays;
    uint32_t* nextPath;
    ShadowItem* shadows;
};
size_t compact_tile_bytes(uint32_t maxCount);
hipError_t launch_compact_queues(const CompactQueues& Q, hipStream_t stream);
hipError_t launch_compact_indices(const uint2* flags, uint32_t count, uint32_t capNext, uint32_t capShadow,
                                  uint2* tiles, uint2* prefix, uint32_t* nextIndices, uint32_t* shadowIndices,
                                  uint32_t* status, hipStream_t stream);
// Wavefront stages; defined in wavefront.hip.
struct ShadeLaunch {
    FrameConst f;
    ShadingView sh;
    SceneView s;
    uint32_t bounce;
    uint32_t capacity;
    const F4* rays;
    const uint32_t* rayPath;
    const uint32_t* count;
    const QueueHit* hits;
    PathState* paths;
    F4* nextRays;
    uint32_t* nextPath;
    uint32_t* nextCount;
    ShadowItem* shadows;
    uint32_t* shadowCount;
    SurfaceV2* surfaces;  // bounce 0 publishes the primary surfaces here (NULL: none)
    GBufferV2* gbuffer = nullptr;  // and the G-buffer primary records here (NULL: none)
    // Batched jobs (NULL: a single frame): job table, scenes of appended items.
    const BatchJob* jobs = nullptr;
    uint32_t* nextSceneOf = nullptr;
    uint32_t* shadowSceneOf = nullptr;
    uint32_t blocksPerCu = 0;       // Shade grid (0: 16 per CU)
    bool extended = false;          // Shade built with sphere lights and lat-long maps
    uint32_t* retire = nullptr;       // batched: slots of paths that did not continue
    uint32_t* retireCount = nullptr;
    // PrefixScan queue mode (not batched): per queue item i, the (next,
    // shadow) flags and the candidate records at index i instead of appends.
    uint2* denseFlags = nullptr;
    F4* denseRays = nullptr;
    uint32_t* densePath = nullptr;
    ShadowItem* denseShadows = nullptr;
};
// Control words of the batch engine's path pool (device memory).
enum : uint32_t {
    kCtlFree = 0,         // free path slots on the free list
    kCtlGenCount = 1,     // generation entries
    kCtlCursorLo = 2,     // next global path index to start (64-bit)
    kCtlCursorHi = 3,
    kCtlGenTotalLo = 4,   // paths of every admitted job (64-bit)
    kCtlGenTotalHi = 5,
    kCtlResolveCount = 6, // jobs resolved this step
    kCtlReserve = 7,      // free slots regen leaves unused (path budget = capacity - reserve)
    kCtlRaysLo = 8,       // traversal queries traced so far (64-bit)
    kCtlRaysHi = 9,
    kCtlWords = 16,
};
// Queue counters of the batch engine, one 128-byte line each: atomics on
// one address serialize, on separate lines they proceed in parallel.
enum : uint32_t {
    kCntRays0 = 0,        // ray queues (double-buffered)
    kCntRays1 = 32,
    kCntShadow = 64,      // shadow queue
    kCntRetire0 = 96,     // retire lists (double-buffered)
    kCntRetire1 = 128,
    kCntWords = 160,
};
// Per-job retired counts are split over this many stripes (stripe-major,
// jobStride words apart), picked by block, so that the wave-aggregated
// atomics of one big job do not all land on one address.
constexpr uint32_t kDoneStripes = 16;
struct BatchDevice {
    const BatchJob* jobs;
    PathState* paths;
    uint32_t* freeList;          // free path slots (a stack)
    uint32_t* ctl;               // kCtl* words
    const uint32_t* genJob;      // job slot of each generation entry
    const uint64_t* genStart;    // first global path index of each entry
    uint32_t* doneCount;         // [stripe * jobStride + job slot]: retired paths
    const uint32_t* resolveJob;  // jobs to resolve this step
    const uint32_t* resolveStart;  // prefix of their pixels
    uint32_t jobStride;          // job slots (maxJobs)
};
// Paths retired by the previous step's Shade (their last shadow queries
// traced): resolved sample values, slots written above the free list's top
// (positions, no atomics), per-job striped counts.
hipError_t launch_batch_finalize(const BatchDevice& d, const uint32_t* retire, const uint32_t* retireCount,
                                 uint32_t blocks, hipStream_t stream);
// After finalize: the free list's top raised by the retired count, the
// step's traced queries (tracedRays + shadow count) added to the ray total,
// the retired count and the two queue counters Shade is about to fill
// cleared.
hipError_t launch_batch_reset(uint32_t* ctl, uint32_t* retireCount, uint32_t* rayCount, uint32_t* shadowCount,
                              const uint32_t* tracedRays, hipStream_t stream);
// New paths from the admitted jobs' sample streams into free slots; their
// primary rays written after Shade's rays (positions, no atomics); then the
// free count, cursor and ray count advanced.
hipError_t launch_batch_regen(const BatchDevice& d, F4* rays, uint32_t* rayPath, uint32_t* raySceneOf,
                              uint32_t* rayCount, uint32_t blocks, hipStream_t stream);
// Admission: n staged job records copied to their slots, their count
// stripes cleared.
hipError_t launch_batch_admit(const BatchDevice& d, BatchJob* jobs, const BatchJob* staged, const uint32_t* slots,
                              uint32_t n, hipStream_t stream);
// Film running means of the finished jobs, samples in order.
hipError_t launch_batch_resolve(const BatchDevice& d, uint32_t blocks, hipStream_t stream);
hipError_t launch_frame_reset(uint32_t* counts, uint32_t n, uint32_t pathCount, hipStream_t stream);
hipError_t launch_raygen(const FrameConst& f, PathState* paths, F4* rays, uint32_t* rayPath, hipStream_t stream);
hipError_t launch_shade(const ShadeLaunch& L, hipStream_t stream);
// direct (may be NULL): ReSTIR's diffuse and specular direct light, two
// float4 per pixel, added to the path's direct signals (spp 1).
// E: the reconstruction exports (all NULL: none).
hipError_t launch_resolve_film(const FrameConst& f, PathState* paths, float4* film, float4* signals,
                               const float4* direct, const ExportV2& E, hipStream_t stream);
// MotionVectorsCS and ComposeCS; defined in reconstruction.hip.
hipError_t launch_motion_vectors(GBufferV2* gbuffer, const MotionInputV2* inputs, const float4* current,
                                 const float4* previous, const rex_motion_constants_v2& c, hipStream_t stream);
hipError_t launch_compose(const rex_compose_desc_v2& d, hipStream_t stream);
// The ReSTIR DI passes of one frame (initial, temporal, spatial, visibility
// traversal, resolve + publish); defined in restir.hip.
hipError_t launch_restir(const RestirLaunch& L, const SceneView& scene, TraceWorkspace& ws, hipStream_t stream);
// Renderer hooks of the request loop: ReSTIR buffers allocated and the
// traversal workspace sized for hierarchies up to wide4Depth ahead of time
// (neither may allocate inside the loop); the Film; the device bytes per
// pixel of a one-sample renderer (with ReSTIR buffers when restir).
rex_status renderer_prepare(rex_renderer renderer, uint32_t wide4Depth, bool restir);
void* renderer_film(rex_renderer renderer);
// Queue sizes of the renderer's last frame: words [0, 16) closest-hit rays
// per bounce, [32, 48) shadow queries per bounce.
const uint32_t* renderer_counts(rex_renderer renderer);
// The batch engine has queued or running jobs.
bool batch_busy(rex_batch batch);
size_t renderer_bytes_per_pixel(bool restir);

// Integrator queue stages (order 0 left-first, 1 near-first); the item count
// is read from device memory, capacity sizes the persistent grid.
hipError_t launch_queue_closest(const SceneView& s, TraceWorkspace& ws, const F4* rays, const uint32_t* countPtr,
                                uint32_t capacity, QueueHit* hits, uint32_t order, uint64_t* counters,
                                hipStream_t stream);
// The mixed launch over an arena of scenes (batched jobs, compressed 4-wide
// walk): every queue item names its scene, whose root the walk starts from.
hipError_t launch_queue_mixed_arena(const SceneView& s, TraceWorkspace& ws, const ShadowItem* items,
                                    const uint32_t* shadowCountPtr, const uint32_t* shadowSceneOf, const F4* rays,
                                    const uint32_t* rayCountPtr, const uint32_t* raySceneOf, const SceneRoot* roots,
                                    uint32_t capacity, QueueHit* hits, PathState* paths, hipStream_t stream);
hipError_t launch_queue_shadow(const SceneView& s, TraceWorkspace& ws, const ShadowItem* items,
                               const uint32_t* countPtr, uint32_t capacity, PathState* paths, uint32_t order,
                               uint64_t* counters, hipStream_t stream);
// One launch over a bounce's shadow queue and the next bounce's ray queue
// (capacity = both queues' capacity).
hipError_t launch_queue_mixed(const SceneView& s, TraceWorkspace& ws, const ShadowItem* items,
                              const uint32_t* shadowCountPtr, const F4* rays, const uint32_t* rayCountPtr,
                              uint32_t capacity, QueueHit* hits, PathState* paths, uint32_t order,
                              hipStream_t stream);

}  // namespace rex

struct rex_scene_t {
    rex::Scene s;
};
