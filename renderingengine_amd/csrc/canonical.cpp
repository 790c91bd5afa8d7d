// Scene ingestion: ABI-v0 canonical scenes -> world-space canonical triangles
// (the traversal build's input) and the L6 material / light records.
//
// The expansion runs on the host because the SAH build that consumes it
// does; it is parallel over instances (a counting pass sizes every visible
// instance's slice of the output, then each thread fills whole instances),
// so the output order is the reference's sequential order: instances,
// their geometries, their triangles.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rex_internal.hpp"

namespace rex {
namespace {

constexpr uint32_t kInstanceVisible = 1u << 0;        // SceneAbiV0.hpp InstanceFlagVisible
constexpr uint32_t kMaterialThinWalled = 1u << 3;     // SceneAbiV0.hpp MaterialFlagThinWalled
constexpr uint32_t kModelSmoothDielectric = 1u;       // SceneAbiV0.hpp MaterialModelSmoothDielectric
constexpr uint32_t kLightEmissiveTriangle = 4u;       // SceneAbiV0.hpp LightTypeEmissiveTriangle
// L6 private material constants (MegakernelBridge.hpp:35-40; Wave2Runtime.cpp:469-476).
constexpr uint32_t kLobeDiffuse = 1u << 0, kLobeGlossy = 1u << 1, kLobeSpecular = 1u << 2,
                   kLobeReflection = 1u << 3, kLobeTransmission = 1u << 4;
constexpr uint32_t kModelMetallicRoughness = 5u, kModelSmoothGlass = 3u, kModelRoughDielectric = 4u;
constexpr uint32_t kFlagPureEmitter = 1u << 0, kFlagThinWalled = 1u << 1;

struct V3 {
    float x, y, z;
};

// TransformPosition (CanonicalTraversalScene.cpp:13-25): M * (p, 1) by rows,
// each row's products summed left to right in fp32.
V3 transform_position(const float* m, const float* p) {
    return {m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3], m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7],
            m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11]};
}

// TransformNormal (CanonicalTraversalScene.cpp:27-50): transpose(worldToObject)
// * n in fp32, normalised by an fp64 length; zero for a zero or non-finite
// length.
V3 transform_normal(const float* w, const float* n) {
    V3 r{w[0] * n[0] + w[4] * n[1] + w[8] * n[2], w[1] * n[0] + w[5] * n[1] + w[9] * n[2],
         w[2] * n[0] + w[6] * n[1] + w[10] * n[2]};
    const double ls = static_cast<double>(r.x) * r.x + static_cast<double>(r.y) * r.y + static_cast<double>(r.z) * r.z;
    if (!(ls > 0.0) || !std::isfinite(ls)) return {0.0f, 0.0f, 0.0f};
    const float inv = static_cast<float>(1.0 / std::sqrt(ls));
    r.x *= inv;
    r.y *= inv;
    r.z *= inv;
    return r;
}

bool finite3(const V3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

struct Expansion {
    const rex_canonical_scene& s;
    std::vector<uint64_t> first;  // per instance: first output triangle (visible instances)
    std::string error;

    explicit Expansion(const rex_canonical_scene& scene) : s(scene) {}

    // Triangles of instance i (0 when invisible); checks the ranges the
    // reference checks before touching indices.
    bool count(uint32_t i, uint64_t& n, std::string& err) const {
        const rex_instance_v0& inst = s.instances[i];
        n = 0;
        if ((inst.metadata[3] & kInstanceVisible) == 0u) return true;
        if (inst.metadata[0] >= s.geometryCount || inst.metadata[1] > s.geometryCount - inst.metadata[0]) {
            err = "Canonical instance geometry range is invalid during traversal expansion.";
            return false;
        }
        for (uint32_t g = 0; g < inst.metadata[1]; ++g) {
            const rex_geometry_v0& geo = s.geometries[inst.metadata[0] + g];
            if (geo.identity[2] >= s.materialCount) {
                err = "Canonical traversal geometry references an invalid material.";
                return false;
            }
            n += (static_cast<uint64_t>(geo.indexRange[1]) + 2u) / 3u;
        }
        return true;
    }

    // Writes instance i's triangles from out[first[i]] on.
    bool fill(uint32_t i, rex_triangle* out, std::string& err) const {
        const rex_instance_v0& inst = s.instances[i];
        if ((inst.metadata[3] & kInstanceVisible) == 0u) return true;
        uint64_t t = first[i];
        for (uint32_t g = 0; g < inst.metadata[1]; ++g) {
            const rex_geometry_v0& geo = s.geometries[inst.metadata[0] + g];
            const rex_material_v0& mat = s.materials[geo.identity[2]];
            for (uint32_t off = 0; off < geo.indexRange[1]; off += 3u) {
                const uint32_t firstIndex = geo.indexRange[0] + off;
                if (firstIndex > s.indexCount || 3u > s.indexCount - firstIndex) {
                    err = "Canonical traversal triangle index range is invalid.";
                    return false;
                }
                rex_triangle& tri = out[t];
                for (uint32_t c = 0; c < 3u; ++c) {
                    const uint32_t vi = s.indices[firstIndex + c] + geo.indexRange[2];
                    if (vi >= s.vertexCount) {
                        err = "Canonical traversal triangle references an invalid vertex.";
                        return false;
                    }
                    const rex_vertex_v0& v = s.vertices[vi];
                    const V3 p = transform_position(inst.objectToWorld, v.position);
                    const V3 n = transform_normal(inst.worldToObject, v.normal);
                    if (!finite3(p) || !finite3(n)) {
                        err = "Canonical traversal transform produced non-finite geometry.";
                        return false;
                    }
                    const float pw[4] = {p.x, p.y, p.z, 1.0f}, nw[4] = {n.x, n.y, n.z, 0.0f};
                    std::memcpy(tri.positions[c], pw, sizeof(pw));
                    std::memcpy(tri.normals[c], nw, sizeof(nw));
                    std::memcpy(tri.texcoords[c], v.texcoord0, sizeof(v.texcoord0));
                }
                const uint32_t traversalId = static_cast<uint32_t>(t);
                const uint32_t primitiveId = geo.indexRange[3] + off / 3u;
                const uint32_t id[4] = {inst.metadata[2], primitiveId, geo.identity[0], mat.metadata[2]};
                const uint32_t md[4] = {geo.identity[3], mat.metadata[1], traversalId, inst.metadata[3]};
                std::memcpy(tri.identity, id, sizeof(id));
                std::memcpy(tri.metadata, md, sizeof(md));
                ++t;
            }
        }
        return true;
    }
};

rex_status check_view(const rex_canonical_scene* s) {
    if (s == nullptr) return fail(REX_INVALID_ARGUMENT, "scene is NULL");
    if (s->fingerprint == 0u || s->generation == 0u || s->constants == nullptr || s->vertexCount == 0u ||
        s->indexCount == 0u || s->geometryCount == 0u || s->instanceCount == 0u || s->materialCount == 0u ||
        s->vertices == nullptr || s->indices == nullptr || s->geometries == nullptr || s->instances == nullptr ||
        s->materials == nullptr) {
        return fail(REX_INVALID_ARGUMENT, "Canonical traversal expansion requires a validated, identified scene view.");
    }
    if (static_cast<uint64_t>(s->indexCount / 3u) * s->instanceCount > 0xffffffffull) {
        return fail(REX_INVALID_ARGUMENT, "Canonical traversal expansion exceeds uint32 triangle identity.");
    }
    return REX_SUCCESS;
}

}  // namespace

// The instance table the motion inputs resolve stable instance IDs against
// (pbr_reconstruction_export_v2.hlsli:91-101 scans the instances for the
// first one with the ID): world-to-object rows per instance, and (ID, first
// index) pairs sorted by ID for a binary search.
rex_status upload_instances(const rex_canonical_scene& sc, Scene& s) {
    const uint32_t n = sc.instanceCount;
    std::vector<InstanceXform> xf(n);
    std::vector<float> cur(16u * n), prev(16u * n);
    std::vector<uint2> map;
    map.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* w = sc.instances[i].worldToObject;
        xf[i] = {{w[0], w[1], w[2], w[3]}, {w[4], w[5], w[6], w[7]}, {w[8], w[9], w[10], w[11]}};
        std::copy(sc.instances[i].objectToWorld, sc.instances[i].objectToWorld + 16, cur.begin() + 16u * i);
        std::copy(sc.instances[i].previousObjectToWorld, sc.instances[i].previousObjectToWorld + 16,
                  prev.begin() + 16u * i);
        map.push_back(make_uint2(sc.instances[i].metadata[2], i));
    }
    std::stable_sort(map.begin(), map.end(), [](const uint2& a, const uint2& b) { return a.x < b.x; });
    map.erase(std::unique(map.begin(), map.end(), [](const uint2& a, const uint2& b) { return a.x == b.x; }),
              map.end());
    auto upload = [](const void* src, size_t bytes, void** dst) -> rex_status {
        if (rex_status st = check_hip(hipMalloc(dst, bytes > 0 ? bytes : 16), "hipMalloc instance table")) return st;
        if (bytes == 0) return REX_SUCCESS;
        return check_hip(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy instance table");
    };
    if (rex_status st = upload(xf.data(), xf.size() * sizeof(InstanceXform), reinterpret_cast<void**>(&s.dInstances)))
        return st;
    if (rex_status st = upload(map.data(), map.size() * sizeof(uint2), reinterpret_cast<void**>(&s.dInstanceMap)))
        return st;
    if (rex_status st = upload(cur.data(), cur.size() * sizeof(float), reinterpret_cast<void**>(&s.dObjectToWorld)))
        return st;
    if (rex_status st =
            upload(prev.data(), prev.size() * sizeof(float), reinterpret_cast<void**>(&s.dPrevObjectToWorld)))
        return st;
    s.instanceCount = n;
    s.instanceMapCount = static_cast<uint32_t>(map.size());
    return REX_SUCCESS;
}

}  // namespace rex

using rex::fail;

extern "C" {

rex_status rex_canonical_expand(const rex_canonical_scene* scene, rex_triangle* out, uint32_t capacity,
                                uint32_t* count) {
    if (count == nullptr) return fail(REX_INVALID_ARGUMENT, "count is NULL");
    if (rex_status st = rex::check_view(scene)) return st;
    const rex_canonical_scene& s = *scene;
    rex::Expansion ex(s);
    ex.first.assign(s.instanceCount + 1u, 0u);
    // Sizes first: the reference reports range errors in instance order
    // before it expands, so they are checked sequentially here too.
    for (uint32_t i = 0; i < s.instanceCount; ++i) {
        uint64_t n = 0;
        std::string err;
        if (!ex.count(i, n, err)) return fail(REX_INVALID_ARGUMENT, err);
        ex.first[i + 1u] = ex.first[i] + n;
    }
    const uint64_t total = ex.first[s.instanceCount];
    if (total == 0u) return fail(REX_INVALID_ARGUMENT, "Canonical traversal expansion produced no visible triangles.");
    if (total > 0xffffffffull) return fail(REX_INVALID_ARGUMENT, "Canonical traversal expansion exceeds uint32.");
    *count = static_cast<uint32_t>(total);
    if (out == nullptr) return REX_SUCCESS;
    if (capacity < total) return fail(REX_INVALID_ARGUMENT, "output capacity below the triangle count");
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t threads = std::min<uint32_t>(std::min(hw, 16u), s.instanceCount);
    std::vector<std::string> errors(threads);
    std::vector<uint32_t> errorAt(threads, 0xffffffffu);
    auto work = [&](uint32_t t) {
        for (uint32_t i = t; i < s.instanceCount; i += threads) {
            if (!ex.fill(i, out, errors[t])) {
                errorAt[t] = i;
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (std::thread& th : pool) th.join();
    // The first failing instance in the reference's order wins.
    uint32_t bad = 0xffffffffu, which = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        if (errorAt[t] < bad) {
            bad = errorAt[t];
            which = t;
        }
    }
    if (bad != 0xffffffffu) return fail(REX_INVALID_ARGUMENT, errors[which]);
    return REX_SUCCESS;
}

rex_status rex_canonical_shading(const rex_canonical_scene* scene, const rex_triangle* triangles,
                                 uint32_t triangleCount, rex_material* materials, rex_light* lights) {
    if (scene == nullptr || materials == nullptr || (scene->lightCount != 0u && lights == nullptr)) {
        return fail(REX_INVALID_ARGUMENT, "NULL argument");
    }
    const rex_canonical_scene& s = *scene;
    if (s.materialCount == 0u || s.lightCount == 0u || s.materials == nullptr || s.lights == nullptr) {
        return fail(REX_INVALID_ARGUMENT, "Wave 2 production scenes require materials and at least one light.");
    }
    for (uint32_t m = 0; m < s.materialCount; ++m) {
        const rex_material_v0& src = s.materials[m];
        rex_material out{};
        const float e = src.emissiveFactorStrength[3];
        const float base[4] = {src.baseColorFactor[0], src.baseColorFactor[1], src.baseColorFactor[2],
                               src.surfaceParams[0]};
        const float emis[4] = {src.emissiveFactorStrength[0] * e, src.emissiveFactorStrength[1] * e,
                               src.emissiveFactorStrength[2] * e, src.surfaceParams[1]};
        const float trans[4] = {src.transmissionParams[0], src.transmissionParams[1], 0.0f, 0.0f};
        const float f0[4] = {0.04f, 0.04f, 0.04f, 0.0f};
        const float eta[4] = {1.0f, 1.0f, 1.0f, 0.0f};
        std::memcpy(out.baseColorMetallic, base, sizeof(base));
        std::memcpy(out.emissiveRoughness, emis, sizeof(emis));
        std::memcpy(out.transmissionIor, trans, sizeof(trans));
        std::memcpy(out.attenuationColorDistance, src.attenuationColorDistance, sizeof(out.attenuationColorDistance));
        std::memcpy(out.f0, f0, sizeof(f0));
        std::memcpy(out.conductorEta, eta, sizeof(eta));
        const bool smooth = src.metadata[0] == rex::kModelSmoothDielectric;
        const bool rough = !smooth && src.transmissionParams[0] > 0.0f;
        uint32_t flags = 0u;
        const float baseMax = std::max({src.baseColorFactor[0], src.baseColorFactor[1], src.baseColorFactor[2]});
        const float emisMax = std::max({emis[0], emis[1], emis[2]});
        if (emisMax > 0.0f && baseMax == 0.0f && src.transmissionParams[0] == 0.0f) flags |= rex::kFlagPureEmitter;
        if ((src.metadata[1] & rex::kMaterialThinWalled) != 0u) flags |= rex::kFlagThinWalled;
        out.metadata[0] = smooth ? rex::kModelSmoothGlass
                                 : (rough ? rex::kModelRoughDielectric : rex::kModelMetallicRoughness);
        out.metadata[1] = smooth ? (rex::kLobeSpecular | rex::kLobeReflection | rex::kLobeTransmission)
                                 : (rough ? (rex::kLobeGlossy | rex::kLobeReflection | rex::kLobeTransmission)
                                          : (rex::kLobeDiffuse | rex::kLobeGlossy | rex::kLobeReflection));
        out.metadata[2] = 0u;
        out.metadata[3] = flags;
        materials[m] = out;
    }
    // First expanded triangle of each (instance, primitive).
    std::unordered_map<uint64_t, uint32_t> firstOf;
    bool needMap = false;
    for (uint32_t l = 0; l < s.lightCount; ++l) needMap |= s.lights[l].identity[0] == rex::kLightEmissiveTriangle;
    if (needMap) {
        if (triangles == nullptr) return fail(REX_INVALID_ARGUMENT, "emissive lights need the expanded triangles");
        firstOf.reserve(triangleCount);
        for (uint32_t t = 0; t < triangleCount; ++t) {
            const uint64_t key = (static_cast<uint64_t>(triangles[t].identity[0]) << 32) | triangles[t].identity[1];
            firstOf.emplace(key, t);
        }
    }
    for (uint32_t l = 0; l < s.lightCount; ++l) {
        const rex_light_v0& src = s.lights[l];
        rex_light out{};
        std::memcpy(out.positionRange, src.positionRange, sizeof(out.positionRange));
        std::memcpy(out.directionCosOuter, src.directionCosOuter, sizeof(out.directionCosOuter));
        std::memcpy(out.radianceScale, src.radianceScale, sizeof(out.radianceScale));
        std::memcpy(out.shapeParams, src.shapeParams, sizeof(out.shapeParams));
        const uint32_t id[4] = {src.identity[0], src.extra[1], src.identity[1], src.identity[3]};
        uint32_t pay[4] = {src.extra[0], src.extra[2], src.extra[3], src.identity[2]};
        if (src.identity[0] == rex::kLightEmissiveTriangle) {
            auto it = firstOf.find((static_cast<uint64_t>(src.identity[2]) << 32) | src.identity[3]);
            if (it == firstOf.end()) {
                return fail(REX_INVALID_ARGUMENT, "Canonical emissive light references no traversal triangle.");
            }
            pay[0] = it->second;
        }
        std::memcpy(out.identity, id, sizeof(id));
        std::memcpy(out.payload, pay, sizeof(pay));
        lights[l] = out;
    }
    return REX_SUCCESS;
}

rex_status rex_scene_create_canonical(const rex_canonical_scene* scene, uint32_t maxLeafSize, uint32_t lightSelection,
                                      rex_scene* outScene) {
    if (outScene == nullptr) return fail(REX_INVALID_ARGUMENT, "outScene is NULL");
    *outScene = nullptr;
    uint32_t n = 0;
    if (rex_status st = rex_canonical_expand(scene, nullptr, 0u, &n)) return st;
    std::vector<rex_triangle> tris(n);
    if (rex_status st = rex_canonical_expand(scene, tris.data(), n, &n)) return st;
    std::vector<rex_material> mats(scene->materialCount);
    std::vector<rex_light> lights(scene->lightCount);
    if (rex_status st = rex_canonical_shading(scene, tris.data(), n, mats.data(), lights.data())) return st;
    // Alpha-masked geometry against the renderers' two-layer atlas
    // (Wave2Runtime.cpp:392-408) and the scene's own materials.
    static const float kRendererAtlas[REX_ALPHA_ATLAS_RENDERER_LAYERS] = {0.0f, 1.0f};
    rex_alpha_desc alpha{scene->materials, scene->materialCount, REX_ALPHA_ATLAS_RENDERER_LAYERS, kRendererAtlas,
                         REX_ALPHA_ATLAS_RENDERER_SAMPLER, 0u};
    rex_scene_desc d{};
    d.fingerprint = scene->fingerprint;
    d.generation = scene->generation;
    d.maxLeafSize = maxLeafSize;
    d.alpha = &alpha;
    rex_scene sc = nullptr;
    if (rex_status st = rex_scene_create(tris.data(), n, &d, &sc)) return st;
    rex_shading_desc sd{mats.data(), scene->materialCount, lightSelection, lights.data(), scene->lightCount, 0u,
                        nullptr, 0u, 0u};
    rex_status st = rex_scene_set_shading(sc, &sd);
    if (st == REX_SUCCESS) st = rex::upload_instances(*scene, sc->s);
    if (st != REX_SUCCESS) {
        (void)rex_scene_destroy(sc);
        return st;
    }
    *outScene = sc;
    return REX_SUCCESS;
}

rex_status rex_scene_refit_canonical(rex_scene scene, const rex_canonical_scene* canonical, uint32_t generation) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (canonical == nullptr) return fail(REX_INVALID_ARGUMENT, "canonical scene is NULL");
    rex::Scene& s = scene->s;
    if (s.dInstanceMap == nullptr) return fail(REX_INVALID_ARGUMENT, "scene was not made by rex_scene_create_canonical");
    uint32_t n = 0;
    if (rex_status st = rex_canonical_expand(canonical, nullptr, 0u, &n)) return st;
    if (n != s.triCount) return fail(REX_INVALID_ARGUMENT, "the moved scene expands to a different triangle count");
    std::vector<rex_triangle> tris(n);
    if (rex_status st = rex_canonical_expand(canonical, tris.data(), n, &n)) return st;
    std::vector<rex_material> mats(canonical->materialCount);
    std::vector<rex_light> lights(canonical->lightCount);
    if (rex_status st = rex_canonical_shading(canonical, tris.data(), n, mats.data(), lights.data())) return st;
    if (rex_status st = rex_scene_refit(scene, tris.data(), n, generation)) return st;
    rex_shading_desc sd{mats.data(), canonical->materialCount, s.shade.selection, lights.data(), canonical->lightCount,
                        0u, nullptr, 0u, 0u};
    if (rex_status st = rex_scene_set_shading(scene, &sd)) return st;
    // hipFree waits for the device, so no recorded frame still reads the old table.
    void* old[] = {s.dInstances, s.dInstanceMap, s.dObjectToWorld, s.dPrevObjectToWorld};
    for (void* p : old) {
        if (p) (void)hipFree(p);
    }
    s.dInstances = nullptr;
    s.dInstanceMap = nullptr;
    s.dObjectToWorld = s.dPrevObjectToWorld = nullptr;
    s.instanceCount = s.instanceMapCount = 0;
    return rex::upload_instances(*canonical, s);
}

}  // extern "C"
