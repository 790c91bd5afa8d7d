// extern "C" entry points of librex (include/rex.h).
#include <cstring>
#include <vector>
#include <exception>
#include <new>
#include <string>

#include "rex_internal.hpp"

namespace rex {

namespace {
thread_local std::string tlsError;
}

void set_error(const std::string& msg) { tlsError = msg; }

rex_status fail(rex_status code, const std::string& msg) {
    set_error(msg);
    return code;
}

rex_status check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return REX_SUCCESS;
    return fail(REX_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace rex

using rex::check_hip;
using rex::fail;

#define REX_GUARD_BEGIN try {
#define REX_GUARD_END                                                          \
    }                                                                          \
    catch (const std::bad_alloc&) {                                            \
        return fail(REX_INVALID_ARGUMENT, "host allocation failed");           \
    }                                                                          \
    catch (const std::exception& e) {                                          \
        return fail(REX_RECORDING_FAILED, std::string("internal: ") + e.what()); \
    }                                                                          \
    catch (...) {                                                              \
        return fail(REX_RECORDING_FAILED, "internal: unknown exception");      \
    }

extern "C" {

const char* rex_last_error(void) { return rex::tlsError.c_str(); }

uint32_t rex_abi_version(void) { return REX_ABI_VERSION; }

rex_status rex_device_alloc(size_t bytes, void** devicePtr) {
    if (devicePtr == nullptr) return fail(REX_INVALID_ARGUMENT, "devicePtr is NULL");
    *devicePtr = nullptr;
    return check_hip(hipMalloc(devicePtr, bytes == 0 ? 1 : bytes), "hipMalloc");
}

rex_status rex_device_free(void* devicePtr) { return check_hip(hipFree(devicePtr), "hipFree"); }

rex_status rex_copy_to_device(void* dst, const void* src, size_t bytes) {
    return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

rex_status rex_copy_to_host(void* dst, const void* src, size_t bytes) {
    return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

rex_status rex_synchronize(void) { return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

rex_status rex_scene_create(const rex_triangle* triangles, uint32_t triangleCount, const rex_scene_desc* desc,
                            rex_scene* outScene) {
    REX_GUARD_BEGIN
    if (outScene == nullptr || desc == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    *outScene = nullptr;
    rex_scene s = new rex_scene_t();
    const rex_status st = rex::scene_build(triangles, triangleCount, *desc, s->s);
    if (st != REX_SUCCESS) {
        rex::scene_release(s->s);
        delete s;
        return st;
    }
    *outScene = s;
    return REX_SUCCESS;
    REX_GUARD_END
}

rex_status rex_scene_destroy(rex_scene scene) {
    if (scene == nullptr) return REX_SUCCESS;
    rex::scene_release(scene->s);
    delete scene;
    return REX_SUCCESS;
}

rex_status rex_scene_get_info(rex_scene scene, rex_scene_info* info) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (info == nullptr) return fail(REX_INVALID_ARGUMENT, "info is NULL");
    const rex::Scene& s = scene->s;
    std::memset(info, 0, sizeof(*info));
    info->fingerprint = s.fingerprint;
    info->generation = s.generation;
    info->triangleCount = s.triCount;
    info->nodeCount = s.nodeCount;
    info->wideNodeCount = s.wideCount;
    info->maximumDepth = s.maxDepth;
    info->maxLeafSize = s.maxLeaf;
    info->compressedNodeCount = s.wide4qCount;
    info->buildMilliseconds = s.buildMs;
    info->deviceBytes = s.deviceBytes;
    info->deviceMilliseconds = s.deviceBuildMs;
    return REX_SUCCESS;
}

rex_status rex_scene_export_flat(rex_scene scene, void* nodes48, uint32_t* primitiveOrder) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (rex_status st = rex::ensure_host_nodes(scene->s)) return st;
    rex::export_flat(scene->s, nodes48, primitiveOrder);
    return REX_SUCCESS;
}

rex_status rex_bvh_build_host(const rex_triangle* triangles, uint32_t triangleCount, uint32_t maxLeafSize,
                              uint32_t buildThreads, void* nodes48, uint32_t* primitiveOrder, uint32_t* nodeCount,
                              uint32_t* maximumDepth) {
    REX_GUARD_BEGIN
    rex::Scene s;
    if (rex_status st = rex::host_build(triangles, triangleCount, maxLeafSize, buildThreads, s)) return st;
    rex::export_flat(s, nodes48, primitiveOrder);
    if (nodeCount != nullptr) *nodeCount = s.nodeCount;
    if (maximumDepth != nullptr) *maximumDepth = s.maxDepth;
    return REX_SUCCESS;
    REX_GUARD_END
}

rex_status rex_bvh_build_lbvh_host(const rex_triangle* triangles, uint32_t triangleCount, void* nodes48,
                                   uint32_t* primitiveOrder, uint32_t* nodeCount, uint32_t* maximumDepth) {
    REX_GUARD_BEGIN
    rex::Scene s;
    if (rex_status st = rex::lbvh_build(triangles, triangleCount, s)) return st;
    rex::export_flat(s, nodes48, primitiveOrder);
    if (nodeCount != nullptr) *nodeCount = s.nodeCount;
    if (maximumDepth != nullptr) *maximumDepth = s.maxDepth;
    return REX_SUCCESS;
    REX_GUARD_END
}

rex_status rex_bvh_refit_host(void* nodes48, uint32_t nodeCount, const rex_triangle* triangles,
                              const uint32_t* primitiveOrder, uint32_t triangleCount) {
    REX_GUARD_BEGIN
    if (nodes48 == nullptr || triangles == nullptr || primitiveOrder == nullptr || nodeCount == 0u ||
        triangleCount == 0u) {
        return fail(REX_INVALID_ARGUMENT, "NULL or empty argument");
    }
    uint32_t* w = static_cast<uint32_t*>(nodes48);
    std::vector<rex::BinaryNode> nodes(nodeCount);
    std::vector<uint8_t> seen(triangleCount, 0u);
    for (uint32_t i = 0; i < nodeCount; ++i) {
        const uint32_t* l = w + 12u * i + 8u;
        rex::BinaryNode& b = nodes[i];
        if (l[1] != 0u) {
            if (static_cast<uint64_t>(l[0]) + l[1] > triangleCount) return fail(REX_INVALID_ARGUMENT, "leaf range");
            b.first = l[0];
            b.count = l[1];
            b.left = b.right = rex::kInvalid;
            for (uint32_t j = l[0]; j < l[0] + l[1]; ++j) {
                if (seen[j]++) return fail(REX_INVALID_ARGUMENT, "a primitive is referenced twice");
            }
        } else {
            // Pre-order: children after their parent.
            if (l[0] <= i || l[2] <= i || l[0] >= nodeCount || l[2] >= nodeCount || l[0] == l[2]) {
                return fail(REX_INVALID_ARGUMENT, "malformed hierarchy");
            }
            b.left = l[0];
            b.right = l[2];
            b.first = 0u;
            b.count = 0u;
        }
    }
    std::vector<rex_triangle> leafTris(triangleCount);
    for (uint32_t j = 0; j < triangleCount; ++j) {
        if (!seen[j] || primitiveOrder[j] >= triangleCount) return fail(REX_INVALID_ARGUMENT, "malformed leaf order");
        leafTris[j] = triangles[primitiveOrder[j]];
    }
    rex::refit_nodes(nodes, leafTris);
    for (uint32_t i = 0; i < nodeCount; ++i) {
        const rex::BinaryNode& b = nodes[i];
        const float f[8] = {b.bmin[0], b.bmin[1], b.bmin[2], 0.0f, b.bmax[0], b.bmax[1], b.bmax[2], 0.0f};
        std::memcpy(w + 12u * i, f, sizeof(f));
    }
    return REX_SUCCESS;
    REX_GUARD_END
}

rex_status rex_scene_refit(rex_scene scene, const rex_triangle* triangles, uint32_t triangleCount,
                           uint32_t generation) {
    REX_GUARD_BEGIN
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (generation == 0u) return fail(REX_INVALID_ARGUMENT, "scene generation must be nonzero");
    rex::Scene& s = scene->s;
    std::vector<rex_triangle> leafTris;
    if (rex_status st = rex::refit_leaf_tris(s, triangles, triangleCount, leafTris)) return st;
    // Boxes refit and layouts re-derived on the device; the new layout
    // replaces the old one (and the generation advances) only when all of it
    // was built.
    if (rex_status st = rex::refit_device(s, std::move(leafTris), generation)) return st;
    // Emissive triangles moved: light fixtures, areas and power weights are
    // stale until the shading tables are set again.
    s.shade.ready = false;
    return REX_SUCCESS;
    REX_GUARD_END
}

static rex_status trace_common(rex_scene scene, const rex_trace_batch* b, void* stream, bool any) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (b == nullptr) return fail(REX_INVALID_ARGUMENT, "batch is NULL");
    const rex::Scene& s = scene->s;
    if (b->sceneFingerprint != s.fingerprint || b->sceneGeneration != s.generation) {
        return fail(REX_MISSING_SCENE, "batch does not name the built scene (fingerprint/generation)");
    }
    if (b->rayCount == 0) return REX_SUCCESS;
    if (b->rays == nullptr || b->hits == nullptr) return fail(REX_INVALID_ARGUMENT, "rays/hits are NULL");
    const uint32_t cap = b->stackCapacity == 0 ? 64u : b->stackCapacity;
    // A capacity outside 1..64 is a per-ray invalid result in the reference
    // shader (software_trace_v1.hlsl:199-206); it is passed through unchanged.
    return check_hip(rex::launch_trace_v1(s.view, scene->s.ws, b->rays + b->rayOffset, b->hits + b->hitOffset,
                                          b->rayCount, any, cap, b->counters, static_cast<hipStream_t>(stream)),
                     "trace launch");
}

rex_status rex_trace_no_alloc(rex_scene scene, const rex_trace_batch* b, int anyHit, void* stream) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (b == nullptr) return fail(REX_INVALID_ARGUMENT, "batch is NULL");
    const rex::Scene& s = scene->s;
    if (b->sceneFingerprint != s.fingerprint || b->sceneGeneration != s.generation) {
        return fail(REX_MISSING_SCENE, "batch does not name the built scene (fingerprint/generation)");
    }
    if (b->rayCount == 0) return REX_SUCCESS;
    if (b->rays == nullptr || b->hits == nullptr) return fail(REX_INVALID_ARGUMENT, "rays/hits are NULL");
    if (b->rayCount > REX_TRACE_NO_ALLOC_MAX_RAYS) {
        return fail(REX_INVALID_ARGUMENT, "rex_trace_no_alloc traces at most 2^31 rays per call");
    }
    const uint32_t cap = b->stackCapacity == 0 ? 64u : b->stackCapacity;
    return check_hip(rex::launch_trace_v1_static(s.view, b->rays + b->rayOffset, b->hits + b->hitOffset,
                                                 b->rayCount, anyHit != 0, cap, b->counters,
                                                 static_cast<hipStream_t>(stream)),
                     "trace launch");
}

rex_status rex_trace_closest(rex_scene scene, const rex_trace_batch* batch, void* stream) {
    return trace_common(scene, batch, stream, false);
}

rex_status rex_trace_any(rex_scene scene, const rex_trace_batch* batch, void* stream) {
    return trace_common(scene, batch, stream, true);
}

rex_status rex_trace_walk(rex_scene scene, const float* rays, const uint32_t* keys, uint32_t count, int shadow,
                          uint32_t walkMode, rex_walk_hit* hits, uint64_t* counters, void* stream) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (count == 0) return REX_SUCCESS;
    if (rays == nullptr || hits == nullptr) return fail(REX_INVALID_ARGUMENT, "rays/hits are NULL");
    if (walkMode != REX_WALK_SHORT_STACK && walkMode != REX_WALK_STACKLESS && walkMode != REX_WALK_NEAR_FIRST &&
        walkMode != REX_WALK_WIDE && walkMode != REX_WALK_WIDE_COMPRESSED) {
        return fail(REX_INVALID_ARGUMENT, "unknown walk mode");
    }
    return check_hip(rex::launch_walk(scene->s.view, scene->s.ws, rays, keys, count, shadow != 0, walkMode, hits,
                                      counters, static_cast<hipStream_t>(stream)),
                     "walk launch");
}

}  // extern "C"
