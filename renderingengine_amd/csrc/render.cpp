// Host side of the wavefront integrator: shading tables and frame launches.
//
// Light tables follow the reference's host construction:
//   BuildVoseAliasTable            integrators/megakernel/MegakernelBridge.cpp:40-145
//   LightPowerWeight               MegakernelBridge.cpp:274-353
//   BuildEmitterMap                MegakernelBridge.cpp:355-402
//   BuildTopLevelLightDistribution MegakernelBridge.cpp:404-454
//   fixture triangles              src/renderers/Wave2Runtime.cpp:638-648
// The emitter map is resolved once per leaf-order triangle on the host with
// the shader's binary search (resources/shaders/integrators/pbr_traversal_resources.hlsli:21-69),
// so the shade stage reads one uint per hit instead of searching.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "rex_internal.hpp"

namespace rex {

void ShadingData::release() {
    if (dMaterials) (void)hipFree(dMaterials);
    if (dLights) (void)hipFree(dLights);
    if (dAlias) (void)hipFree(dAlias);
    if (dPmf) (void)hipFree(dPmf);
    if (dFixtures) (void)hipFree(dFixtures);
    if (dEmitterOfTri) (void)hipFree(dEmitterOfTri);
    if (dEnvTexels) (void)hipFree(dEnvTexels);
    if (dEnvRows) (void)hipFree(dEnvRows);
    if (dEnvColumns) (void)hipFree(dEnvColumns);
    dEnvTexels = nullptr;
    dEnvRows = nullptr;
    dEnvColumns = nullptr;
    dMaterials = nullptr;
    dLights = nullptr;
    dAlias = nullptr;
    dPmf = nullptr;
    dFixtures = nullptr;
    dEmitterOfTri = nullptr;
    ready = false;
}

ShadingView ShadingData::view() const {
    ShadingView v;
    v.materials = dMaterials;
    v.materialCount = materialCount;
    v.lights = dLights;
    v.lightCount = lightCount;
    v.alias = dAlias;
    v.aliasCount = aliasCount;
    v.selectionPmf = dPmf;
    v.fixtures = dFixtures;
    v.fixtureCount = fixtureCount;
    v.emitterOfTri = dEmitterOfTri;
    v.environmentLight = environmentLight;
    v.envTexels = dEnvTexels;
    v.envRows = dEnvRows;
    v.envColumns = dEnvColumns;
    v.envWidth = envWidth;
    v.envHeight = envHeight;
    v.envSampler = envSampler;
    v.extended = extended ? 1u : 0u;
    return v;
}

namespace {

template <typename T>
rex_status upload_vec(const std::vector<T>& host, T** dev) {
    const size_t n = std::max<size_t>(host.size(), 1) * sizeof(T);
    if (rex_status st = check_hip(hipMalloc(reinterpret_cast<void**>(dev), n), "hipMalloc shading table")) return st;
    if (!host.empty()) {
        return check_hip(hipMemcpy(*dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice),
                         "hipMemcpy shading table");
    }
    return REX_SUCCESS;
}

bool finite_nonneg(double v) { return std::isfinite(v) && v >= 0.0; }

// TexelSolidAngle (MegakernelBridge.cpp:176-189).
double texel_solid_angle(uint32_t row, uint32_t width, uint32_t height) {
    const double pi = 3.14159265358979323846;
    const double t0 = pi * static_cast<double>(row) / static_cast<double>(height);
    const double t1 = pi * static_cast<double>(row + 1u) / static_cast<double>(height);
    return 2.0 * pi / static_cast<double>(width) * (std::cos(t0) - std::cos(t1));
}

bool build_alias(const std::vector<double>& weights, const std::vector<uint32_t>& items,
                 std::vector<AliasEntry>& out);

// BuildEnvironmentDistribution (MegakernelBridge.cpp:191-275) over the
// texels' float luminance (Wave2Runtime.cpp:589-599): weights luminance x
// texel solid angle (solid angle alone for a black map), a column table per
// row and a row table over the row sums.
bool build_environment(const float* texels, uint32_t width, uint32_t height, std::vector<AliasEntry>& rows,
                       std::vector<AliasEntry>& columns, double& integrated) {
    const size_t n = static_cast<size_t>(width) * height;
    std::vector<double> w(n);
    integrated = 0.0;
    for (uint32_t r = 0; r < height; ++r) {
        const double sa = texel_solid_angle(r, width, height);
        for (uint32_t c = 0; c < width; ++c) {
            const float* t = texels + 4u * (static_cast<size_t>(r) * width + c);
            const float lumF = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
            const double weighted = static_cast<double>(lumF) * sa;
            w[static_cast<size_t>(r) * width + c] = weighted;
            integrated += weighted;
        }
    }
    if (!std::isfinite(integrated)) return false;
    if (integrated == 0.0) {
        for (uint32_t r = 0; r < height; ++r) {
            const double sa = texel_solid_angle(r, width, height);
            for (uint32_t c = 0; c < width; ++c) w[static_cast<size_t>(r) * width + c] = sa;
        }
    }
    std::vector<double> rowW(height, 0.0), span(width);
    std::vector<uint32_t> items(std::max(width, height));
    for (uint32_t k = 0; k < items.size(); ++k) items[k] = k;
    std::vector<uint32_t> colItems(items.begin(), items.begin() + width);
    columns.clear();
    columns.reserve(n);
    std::vector<AliasEntry> row;
    for (uint32_t r = 0; r < height; ++r) {
        for (uint32_t c = 0; c < width; ++c) {
            span[c] = w[static_cast<size_t>(r) * width + c];
            rowW[r] += span[c];
        }
        if (!build_alias(span, colItems, row)) return false;
        columns.insert(columns.end(), row.begin(), row.end());
    }
    std::vector<uint32_t> rowItems(items.begin(), items.begin() + height);
    return build_alias(rowW, rowItems, rows);
}

// BuildVoseAliasTable: probabilities in double, pmf stored as float.
bool build_alias(const std::vector<double>& weights, const std::vector<uint32_t>& items,
                 std::vector<AliasEntry>& out) {
    out.clear();
    const size_t count = weights.size();
    if (count == 0) return true;
    double sum = 0.0;
    for (double w : weights) {
        if (!finite_nonneg(w)) return false;
        sum += w;
        if (!std::isfinite(sum)) return false;
    }
    const bool uniform = sum == 0.0;
    const double norm = sum > 0.0 ? sum : static_cast<double>(count);
    std::vector<double> scaled(count);
    std::vector<uint32_t> small, large;
    small.reserve(count);
    large.reserve(count);
    out.resize(count);
    const double countD = static_cast<double>(count);
    for (size_t i = 0; i < count; ++i) {
        const double p = uniform ? 1.0 / countD : weights[i] / norm;
        scaled[i] = p * countD;
        const float pmf = static_cast<float>(p);
        if (p > 0.0 && pmf == 0.0f) return false;
        out[i].pmf = pmf;
        out[i].q = 0.0f;
        out[i].item = items[i];
        out[i].alias = static_cast<uint32_t>(i);
        if (scaled[i] < 1.0) small.push_back(static_cast<uint32_t>(i));
        else large.push_back(static_cast<uint32_t>(i));
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t si = small.back();
        small.pop_back();
        const uint32_t li = large.back();
        large.pop_back();
        const double q = std::clamp(scaled[si], 0.0, 1.0);
        out[si].q = static_cast<float>(q);
        out[si].alias = li;
        scaled[li] = scaled[li] + scaled[si] - 1.0;
        if (scaled[li] < 1.0) small.push_back(li);
        else large.push_back(li);
    }
    for (uint32_t i : large) {
        out[i].q = 1.0f;
        out[i].alias = i;
    }
    for (uint32_t i : small) {
        out[i].q = 1.0f;
        out[i].alias = i;
    }
    return true;
}

// LightPowerWeight (MegakernelBridge.cpp:277-359); -1 where the reference
// rejects the light (non-finite or non-positive scene radius, negative or
// non-finite radiance or scale, a directional or spot light without a
// direction, spot cone cosines outside [-1, 1] or inner < outer, a sphere
// without a radius, a triangle light without area, a negative environment
// integral).
double light_power(const rex_light& l, float sceneRadius, double envIntegratedLuminance) {
    const double pi = 3.14159265358979323846;
    if (!std::isfinite(sceneRadius) || sceneRadius <= 0.0f) return -1.0;
    const double scale = static_cast<double>(l.radianceScale[3]);
    if (!finite_nonneg(scale) || !finite_nonneg(l.radianceScale[0]) || !finite_nonneg(l.radianceScale[1]) ||
        !finite_nonneg(l.radianceScale[2])) {
        return -1.0;
    }
    const double lum = (0.2126 * static_cast<double>(l.radianceScale[0]) +
                        0.7152 * static_cast<double>(l.radianceScale[1]) +
                        0.0722 * static_cast<double>(l.radianceScale[2])) *
                       scale;
    if (!finite_nonneg(lum)) return -1.0;
    const double radius = static_cast<double>(sceneRadius);
    const double side = (l.identity[1] & kLightTwoSided) != 0u ? 2.0 : 1.0;
    switch (l.identity[0]) {
    case kLightPoint: return 4.0 * pi * lum;
    case kLightDirectional: {
        const double dx = l.directionCosOuter[0], dy = l.directionCosOuter[1], dz = l.directionCosOuter[2];
        const double len2 = dx * dx + dy * dy + dz * dz;
        if (!std::isfinite(len2) || len2 <= 0.0) return -1.0;
        return pi * radius * radius * lum;
    }
    case kLightSpot: {
        const double dx = l.directionCosOuter[0], dy = l.directionCosOuter[1], dz = l.directionCosOuter[2];
        const double len2 = dx * dx + dy * dy + dz * dz;
        if (!std::isfinite(len2) || len2 <= 0.0) return -1.0;
        const double cosOuter = l.directionCosOuter[3], cosInner = l.shapeParams[1];
        if (!std::isfinite(cosOuter) || !std::isfinite(cosInner) || cosOuter < -1.0 || cosOuter > 1.0 ||
            cosInner < -1.0 || cosInner > 1.0 || cosInner < cosOuter) {
            return -1.0;
        }
        return 2.0 * pi * lum * ((1.0 - cosInner) + 0.5 * (cosInner - cosOuter));
    }
    case kLightSphere: {
        const double r = static_cast<double>(l.shapeParams[0]);
        if (!std::isfinite(r) || r <= 0.0) return -1.0;
        return 4.0 * pi * pi * r * r * lum * side;
    }
    case kLightEmissiveTriangle: {
        const double area = static_cast<double>(l.shapeParams[2]);
        if (!std::isfinite(area) || area <= 0.0) return -1.0;
        return pi * area * lum * side;
    }
    case kLightEnvironment:
        if (!finite_nonneg(envIntegratedLuminance)) return -1.0;
        return pi * radius * radius * scale * envIntegratedLuminance;
    default: return -1.0;
    }
}

}  // namespace

rex_status set_shading(Scene& s, const rex_shading_desc& d) {
    if (rex_status st = ensure_host_nodes(s)) return st;
    if (d.materials == nullptr || d.materialCount == 0) return fail(REX_INVALID_ARGUMENT, "materials are required");
    if (d.lights == nullptr || d.lightCount == 0) return fail(REX_INVALID_ARGUMENT, "at least one light is required");
    if (d.lightSelection > REX_LIGHT_SELECTION_POWER) return fail(REX_INVALID_ARGUMENT, "unknown light selection");
    // Materials beyond Lambert / metallic-roughness without transmission go
    // to the kernels built with every BSDF model (and interior attenuation).
    bool extendedMaterials = false;
    for (uint32_t i = 0; i < d.materialCount; ++i) {
        const rex_material& m = d.materials[i];
        if (m.metadata[0] > kModelMetallicRoughness) {
            return fail(REX_UNSUPPORTED, "material " + std::to_string(i) + ": unknown BSDF model");
        }
        extendedMaterials = extendedMaterials ||
                            (m.metadata[0] != kModelLambert && m.metadata[0] != kModelMetallicRoughness) ||
                            m.transmissionIor[0] != 0.0f;
    }
    // Canonical index of each leaf-order triangle, and its inverse.
    std::vector<uint32_t> leafOf(s.triCount);
    for (uint32_t j = 0; j < s.triCount; ++j) leafOf[s.order[j]] = j;

    ShadingData sd;
    sd.materialCount = d.materialCount;
    sd.lightCount = d.lightCount;
    std::vector<rex_light> lights(d.lights, d.lights + d.lightCount);
    uint32_t env = kInvalid;
    for (uint32_t i = 0; i < d.lightCount; ++i) {
        const rex_light& l = lights[i];
        const uint32_t type = l.identity[0];
        if (type > kLightEnvironment) {
            return fail(REX_UNSUPPORTED, "light " + std::to_string(i) + ": unsupported light type");
        }
        if (type == kLightEnvironment && (l.identity[1] & kLightEnabled) != 0u) {
            if (env != kInvalid) return fail(REX_UNSUPPORTED, "one active environment light at most");
            env = i;
        }
        if (type == kLightEmissiveTriangle) {
            const uint32_t t = l.payload[0];
            if (t >= s.triCount) return fail(REX_INVALID_ARGUMENT, "emissive light references no triangle");
            const rex_triangle& tri = s.leafTris[leafOf[t]];
            if (tri.identity[0] != l.payload[3] || tri.identity[1] != l.identity[3]) {
                return fail(REX_INVALID_ARGUMENT, "emissive light's (instance, primitive) differs from its triangle");
            }
        }
    }
    sd.environmentLight = env;

    // Environment map (Wave2Runtime.cpp:578-614): a map needs an enabled
    // environment light; without a map the environment is the constant
    // (1, 1, 1) texel.
    if (d.environmentSampler > REX_ENVIRONMENT_SAMPLER_IMPORTANCE_MAP) {
        return fail(REX_INVALID_ARGUMENT, "unknown environment sampler");
    }
    std::vector<F4> envTexels;
    if (d.environmentTexels != nullptr) {
        if (env == kInvalid) return fail(REX_INVALID_ARGUMENT, "an environment map needs an enabled environment light");
        const uint64_t n = static_cast<uint64_t>(d.environmentWidth) * d.environmentHeight;
        if (n == 0 || n > 0xffffffffull) return fail(REX_INVALID_ARGUMENT, "environment dimensions are invalid");
        envTexels.resize(n);
        std::memcpy(envTexels.data(), d.environmentTexels, n * sizeof(F4));
        for (const F4& t : envTexels) {
            if (!std::isfinite(t.x) || !std::isfinite(t.y) || !std::isfinite(t.z) || !std::isfinite(t.w) ||
                t.x < 0.0f || t.y < 0.0f || t.z < 0.0f) {
                return fail(REX_INVALID_ARGUMENT, "environment contains invalid linear radiance");
            }
        }
        sd.envWidth = d.environmentWidth;
        sd.envHeight = d.environmentHeight;
    } else {
        envTexels.assign(1, F4{1.0f, 1.0f, 1.0f, 1.0f});
        sd.envWidth = sd.envHeight = 1;
    }
    sd.envSampler = d.environmentSampler;
    if (!build_environment(reinterpret_cast<const float*>(envTexels.data()), sd.envWidth, sd.envHeight, sd.envRows,
                           sd.envColumns, sd.envIntegrated)) {
        return fail(REX_INVALID_ARGUMENT, "environment distribution is invalid");
    }
    const double envIntegrated = sd.envIntegrated;
    sd.extended = extendedMaterials || sd.envWidth * sd.envHeight > 1u ||
                  sd.envSampler != REX_ENVIRONMENT_SAMPLER_UNIFORM_SPHERE;
    for (const rex_light& l : lights) sd.extended = sd.extended || l.identity[0] == kLightSphere;

    // Scene radius from the scene bounds (Wave2Runtime.cpp:649-661).
    const float* mn = s.nodes[0].bmin;
    const float* mx = s.nodes[0].bmax;
    const float hx = (mx[0] - mn[0]) * 0.5f, hy = (mx[1] - mn[1]) * 0.5f, hz = (mx[2] - mn[2]) * 0.5f;
    const float radius = std::sqrt(hx * hx + hy * hy + hz * hz);

    // Both selection strategies are built by the reference's SetScene, so an
    // enabled light whose power weight is rejected fails either way.
    std::vector<double> weights;
    std::vector<uint32_t> items;
    for (uint32_t i = 0; i < d.lightCount; ++i) {
        if ((lights[i].identity[1] & kLightEnabled) == 0u) continue;
        const double power = light_power(lights[i], radius, envIntegrated);
        if (!finite_nonneg(power)) {
            return fail(REX_INVALID_ARGUMENT, "light " + std::to_string(i) + ": invalid power (radiance, scale, "
                                              "direction or area)");
        }
        weights.push_back(d.lightSelection == REX_LIGHT_SELECTION_POWER ? power : 1.0);
        items.push_back(i);
    }
    if (!build_alias(weights, items, sd.alias)) return fail(REX_INVALID_ARGUMENT, "light alias table is invalid");
    sd.aliasCount = static_cast<uint32_t>(sd.alias.size());
    sd.pmf.assign(d.lightCount, 0.0f);
    for (const AliasEntry& e : sd.alias) sd.pmf[e.item] = e.pmf;

    // Fixtures in canonical order.
    std::vector<Fixture> fixtures(s.triCount);
    for (uint32_t i = 0; i < s.triCount; ++i) {
        const rex_triangle& t = s.leafTris[leafOf[i]];
        Fixture& f = fixtures[i];
        f.p0 = {t.positions[0][0], t.positions[0][1], t.positions[0][2], t.positions[0][3]};
        f.p1 = {t.positions[1][0], t.positions[1][1], t.positions[1][2], t.positions[1][3]};
        f.p2 = {t.positions[2][0], t.positions[2][1], t.positions[2][2], t.positions[2][3]};
        f.metadata = {t.identity[3], t.identity[1], t.identity[0], t.metadata[0]};
    }
    sd.fixtureCount = s.triCount;

    // Emitter map (sorted by instance, primitive) resolved per triangle.
    struct Entry {
        uint32_t inst, prim, light;
    };
    std::vector<Entry> map;
    for (uint32_t i = 0; i < d.lightCount; ++i) {
        if (lights[i].identity[0] != kLightEmissiveTriangle) continue;
        if (lights[i].payload[3] == kInvalid || lights[i].identity[3] == kInvalid) {
            return fail(REX_INVALID_ARGUMENT, "emissive-triangle lights need stable instance and primitive IDs");
        }
        map.push_back({lights[i].payload[3], lights[i].identity[3], i});
    }
    std::sort(map.begin(), map.end(), [](const Entry& a, const Entry& b) {
        return a.inst < b.inst || (a.inst == b.inst && a.prim < b.prim);
    });
    for (size_t k = 1; k < map.size(); ++k) {
        if (map[k].inst == map[k - 1].inst && map[k].prim == map[k - 1].prim) {
            return fail(REX_INVALID_ARGUMENT, "two emissive lights reference one (instance, primitive)");
        }
    }
    std::vector<uint32_t> emitterOfTri(s.triCount, kInvalid);
    for (uint32_t j = 0; j < s.triCount; ++j) {
        const uint32_t inst = s.leafTris[j].identity[0], prim = s.leafTris[j].identity[1];
        uint32_t first = 0, count = static_cast<uint32_t>(map.size());
        while (count > 0) {
            const uint32_t step = count >> 1;
            const uint32_t mid = first + step;
            const bool less = map[mid].inst < inst || (map[mid].inst == inst && map[mid].prim < prim);
            if (less) {
                first = mid + 1;
                count -= step + 1;
            } else {
                count = step;
            }
        }
        if (first < map.size() && map[first].inst == inst && map[first].prim == prim) {
            emitterOfTri[j] = map[first].light < d.lightCount ? map[first].light : kInvalid;
        }
    }

    std::vector<Material> mats(d.materialCount);
    std::memcpy(mats.data(), d.materials, sizeof(Material) * d.materialCount);
    std::vector<Light> lts(d.lightCount);
    std::memcpy(lts.data(), lights.data(), sizeof(Light) * d.lightCount);
    if (rex_status st = upload_vec(mats, &sd.dMaterials)) { sd.release(); return st; }
    if (rex_status st = upload_vec(lts, &sd.dLights)) { sd.release(); return st; }
    if (rex_status st = upload_vec(sd.alias, &sd.dAlias)) { sd.release(); return st; }
    if (rex_status st = upload_vec(sd.pmf, &sd.dPmf)) { sd.release(); return st; }
    if (rex_status st = upload_vec(fixtures, &sd.dFixtures)) { sd.release(); return st; }
    if (rex_status st = upload_vec(emitterOfTri, &sd.dEmitterOfTri)) { sd.release(); return st; }
    if (rex_status st = upload_vec(envTexels, &sd.dEnvTexels)) { sd.release(); return st; }
    if (rex_status st = upload_vec(sd.envRows, &sd.dEnvRows)) { sd.release(); return st; }
    if (rex_status st = upload_vec(sd.envColumns, &sd.dEnvColumns)) { sd.release(); return st; }
    sd.selection = d.lightSelection;
    sd.ready = true;
    s.shade.release();
    s.shade = std::move(sd);
    return REX_SUCCESS;
}

constexpr uint32_t kMaxBounces = 16;
constexpr uint32_t kCountWords = 64;                                   // queue sizes
constexpr uint32_t kCounterWords = 2u * REX_COUNTER_COUNT * 2u;        // 2 x 10 uint64
constexpr uint32_t kEventCount = 4u + 3u * kMaxBounces;
// Stage tags of the timing events (rex_frame_stats::stageMs indices).
constexpr uint8_t kStageRaygen = 0, kStageIntersect = 1, kStageShade = 2, kStageShadow = 3, kStageResolve = 4,
                  kStageRestir = 5;

// ReSTIR DI state of a renderer: settings, device buffers (allocated at the
// first ReSTIR frame for maxPixels), and the identity of the last published
// frame that decides history reuse (SameHistoryIdentity,
// src/renderers/ReSTIRDIRuntime.cpp:84-102).
struct RestirState {
    rex_restir_desc desc{};
    bool allocated = false;
    SurfaceV2* surfaces[2] = {nullptr, nullptr};
    ReservoirV3* initial = nullptr;
    ReservoirV3* temporal = nullptr;
    ReservoirV3* spatial = nullptr;
    ReservoirV3* history[2] = {nullptr, nullptr};
    rex_ray_v1* rays = nullptr;
    rex_hit_v1* hits = nullptr;
    float4* direct = nullptr;
    uint32_t* reasons = nullptr;
    RestirStats* stats = nullptr;
    // Reference-correction visibility queries, (K + 1) per pixel, allocated
    // at the first reference-correction frame.
    rex_ray_v1* refRays = nullptr;
    rex_hit_v1* refHits = nullptr;
    size_t refCapacity = 0;
    uint32_t cur = 0;            // buffer index the next frame writes
    uint64_t frameIndex = 0;     // index of the next ReSTIR frame
    bool hasHistory = false;     // a frame was published since the last reset
    bool lastHistoryValid = false;
    uint32_t lastWidth = 0, lastHeight = 0;
    uint32_t lastScene = 0, lastLight = 0, lastConfig = 0, lastResource = 0, lastHistoryGen = 0;
    rex_camera lastCamera{};
    RestirState() {
        desc.candidates = 32;
        desc.neighbors = 5;
        desc.maxM = 20;
        desc.maxAge = 20;
        desc.temporal = 1;
        desc.neighborMode = REX_RESTIR_NEIGHBORS_RING;
        desc.radius = 30;
    }
    void release() {
        void* ptrs[] = {surfaces[0], surfaces[1], initial, temporal, spatial, history[0], history[1],
                        rays, hits, direct, reasons, stats, refRays, refHits};
        for (void* p : ptrs) {
            if (p) (void)hipFree(p);
        }
        refRays = nullptr;
        refHits = nullptr;
        refCapacity = 0;
        surfaces[0] = surfaces[1] = nullptr;
        initial = temporal = spatial = history[0] = history[1] = nullptr;
        rays = nullptr;
        hits = nullptr;
        direct = nullptr;
        reasons = nullptr;
        stats = nullptr;
        allocated = false;
        hasHistory = false;
    }
};

// MakeWorldToView / MakeViewProjection (src/renderers/Wave3DebugRuntime.cpp:140-202)
// in float, with the camera's own tan(vfov / 2) and aspect.
void world_to_view(const rex_camera& c, float m[16]) {
    auto dot3 = [](const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    for (int i = 0; i < 16; ++i) m[i] = 0.0f;
    m[15] = 1.0f;
    m[0] = c.right[0]; m[1] = c.right[1]; m[2] = c.right[2]; m[3] = -dot3(c.right, c.position);
    m[4] = c.up[0]; m[5] = c.up[1]; m[6] = c.up[2]; m[7] = -dot3(c.up, c.position);
    m[8] = -c.forward[0]; m[9] = -c.forward[1]; m[10] = -c.forward[2]; m[11] = dot3(c.forward, c.position);
}

void view_projection(const rex_camera& c, float m[16]) {
    const float nearPlane = 0.01f, farPlane = 10000.0f;
    float proj[16] = {}, view[16];
    proj[0] = 1.0f / (c.aspect * c.tanHalfFov);
    proj[5] = 1.0f / c.tanHalfFov;
    proj[10] = farPlane / (nearPlane - farPlane);
    proj[11] = farPlane * nearPlane / (nearPlane - farPlane);
    proj[14] = -1.0f;
    world_to_view(c, view);
    for (int row = 0; row < 4; ++row) {
        for (int col = 0; col < 4; ++col) {
            float v = 0.0f;
            for (int i = 0; i < 4; ++i) v += proj[row * 4 + i] * view[i * 4 + col];
            m[row * 4 + col] = v;
        }
    }
}

struct Renderer {
    Scene* scene = nullptr;
    uint32_t maxPixels = 0, maxSpp = 0, capacity = 0;
    PathState* paths = nullptr;
    F4* rays[2] = {nullptr, nullptr};
    uint32_t* rayPath[2] = {nullptr, nullptr};
    QueueHit* hits = nullptr;
    ShadowItem* shadows = nullptr;
    // counts: [0..16] ray queue sizes per bounce, [32..47] shadow queue
    // sizes, then (8-byte aligned at word 64) the closest and shadow
    // traversal counters, REX_COUNTER_COUNT uint64 each.
    uint32_t* counts = nullptr;
    float4* film = nullptr;
    float4* signals = nullptr;
    FrameConst last{};
    hipStream_t lastStream = nullptr;
    uint32_t profile = 0;
    // Stage events of the last frame: start, after RayGen, per bounce after
    // Intersect / Shade / TraceShadow, after Resolve.
    std::vector<hipEvent_t> events;
    uint8_t eventStage[kEventCount] = {};
    RestirState restir;
    // Bound ABI-v2 reconstruction targets (caller-owned; pixels == 0: none).
    rex_reconstruction_targets_v2 recon{};
    // The renderer's own traversal work counter and overflow stacks: renderers
    // of one scene on different streams must not share them.
    TraceWorkspace ws;
    // PrefixScan queue mode: Shade's per-item flags and candidate records,
    // the compaction's tile totals (allocated at the first such frame).
    uint2* denseFlags = nullptr;
    F4* denseRays = nullptr;
    uint32_t* densePath = nullptr;
    ShadowItem* denseShadows = nullptr;
    uint2* scanTiles = nullptr;
    uint32_t eventsRecorded = 0;
    uint64_t* traversalCounters() { return reinterpret_cast<uint64_t*>(counts + kCountWords); }
    void release() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        events.clear();
        restir.release();
        ws.release();
        void* dense[] = {denseFlags, denseRays, densePath, denseShadows, scanTiles};
        for (void* p : dense) {
            if (p) (void)hipFree(p);
        }
        denseFlags = nullptr;
        denseRays = nullptr;
        densePath = nullptr;
        denseShadows = nullptr;
        scanTiles = nullptr;
        void* ptrs[] = {paths, rays[0], rays[1], rayPath[0], rayPath[1], hits, shadows, counts, film, signals};
        for (void* p : ptrs) {
            if (p) (void)hipFree(p);
        }
        paths = nullptr;
        rays[0] = rays[1] = nullptr;
        rayPath[0] = rayPath[1] = nullptr;
        hits = nullptr;
        shadows = nullptr;
        counts = nullptr;
        film = nullptr;
        signals = nullptr;
    }
};

// Device buffers of the ReSTIR passes for maxPixels pixels.
rex_status restir_allocate(Renderer& r) {
    RestirState& rs = r.restir;
    const size_t n = r.maxPixels;
    struct Alloc {
        void** p;
        size_t bytes;
    } allocs[] = {
        {reinterpret_cast<void**>(&rs.surfaces[0]), n * sizeof(SurfaceV2)},
        {reinterpret_cast<void**>(&rs.surfaces[1]), n * sizeof(SurfaceV2)},
        {reinterpret_cast<void**>(&rs.initial), n * sizeof(ReservoirV3)},
        {reinterpret_cast<void**>(&rs.temporal), n * sizeof(ReservoirV3)},
        {reinterpret_cast<void**>(&rs.spatial), n * sizeof(ReservoirV3)},
        {reinterpret_cast<void**>(&rs.history[0]), n * sizeof(ReservoirV3)},
        {reinterpret_cast<void**>(&rs.history[1]), n * sizeof(ReservoirV3)},
        {reinterpret_cast<void**>(&rs.rays), n * sizeof(rex_ray_v1)},
        {reinterpret_cast<void**>(&rs.hits), n * sizeof(rex_hit_v1)},
        {reinterpret_cast<void**>(&rs.direct), n * 2 * sizeof(float4)},
        {reinterpret_cast<void**>(&rs.reasons), n * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&rs.stats), sizeof(RestirStats)},
    };
    for (const Alloc& a : allocs) {
        if (rex_status st = check_hip(hipMalloc(a.p, a.bytes), "hipMalloc ReSTIR buffer")) {
            rs.release();
            return st;
        }
    }
    rs.allocated = true;
    rs.hasHistory = false;
    return REX_SUCCESS;
}

// Staging of the PrefixScan queue mode for the renderer's path capacity.
rex_status dense_allocate(Renderer& r) {
    const size_t cap = r.capacity;
    struct Alloc {
        void** p;
        size_t bytes;
    } allocs[] = {
        {reinterpret_cast<void**>(&r.denseFlags), cap * sizeof(uint2)},
        {reinterpret_cast<void**>(&r.denseRays), cap * 2 * sizeof(F4)},
        {reinterpret_cast<void**>(&r.densePath), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&r.denseShadows), cap * sizeof(ShadowItem)},
        {reinterpret_cast<void**>(&r.scanTiles), compact_tile_bytes(r.capacity)},
    };
    for (const Alloc& a : allocs) {
        if (rex_status st = check_hip(hipMalloc(a.p, a.bytes), "hipMalloc queue staging")) return st;
    }
    return REX_SUCCESS;
}

}  // namespace rex

struct rex_renderer_t {
    rex::Renderer r;
};

using rex::check_hip;
using rex::fail;

namespace rex {

rex_status renderer_prepare(rex_renderer renderer, uint32_t wide4Depth, bool restir) {
    Renderer& r = renderer->r;
    if (restir && !r.restir.allocated) {
        if (rex_status st = restir_allocate(r)) return st;
    }
    return check_hip(reserve_wide_workspace(r.ws, wide4Depth), "workspace reserve");
}

void* renderer_film(rex_renderer renderer) { return renderer->r.film; }

const uint32_t* renderer_counts(rex_renderer renderer) { return renderer->r.counts; }

size_t renderer_bytes_per_pixel(bool restir) {
    size_t b = sizeof(PathState) + 4 * sizeof(F4) + 2 * sizeof(uint32_t) + sizeof(QueueHit) + sizeof(ShadowItem) +
               sizeof(float4) + 5 * sizeof(float4);
    if (restir) {
        b += 2 * sizeof(SurfaceV2) + 5 * sizeof(ReservoirV3) + sizeof(rex_ray_v1) + sizeof(rex_hit_v1) +
             2 * sizeof(float4) + sizeof(uint32_t);
    }
    return b;
}

}  // namespace rex

extern "C" {

rex_status rex_scene_set_shading(rex_scene scene, const rex_shading_desc* desc) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (desc == nullptr) return fail(REX_INVALID_ARGUMENT, "desc is NULL");
    try {
        return rex::set_shading(scene->s, *desc);
    } catch (const std::exception& e) {
        return fail(REX_RECORDING_FAILED, std::string("internal: ") + e.what());
    }
}

rex_status rex_scene_export_light_tables(rex_scene scene, void* alias16, float* selectionPmf, uint32_t* aliasCount) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    const rex::ShadingData& sd = scene->s.shade;
    if (!sd.ready) return fail(REX_MISSING_SCENE, "scene has no shading tables");
    if (alias16 != nullptr) std::memcpy(alias16, sd.alias.data(), sd.alias.size() * sizeof(rex::AliasEntry));
    if (selectionPmf != nullptr) std::memcpy(selectionPmf, sd.pmf.data(), sd.pmf.size() * sizeof(float));
    if (aliasCount != nullptr) *aliasCount = sd.aliasCount;
    return REX_SUCCESS;
}

rex_status rex_scene_export_environment_tables(rex_scene scene, void* rows16, void* columns16,
                                               double* integratedLuminance) {
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    const rex::ShadingData& sd = scene->s.shade;
    if (!sd.ready) return fail(REX_MISSING_SCENE, "scene has no shading tables");
    if (rows16 != nullptr) std::memcpy(rows16, sd.envRows.data(), sd.envRows.size() * sizeof(rex::AliasEntry));
    if (columns16 != nullptr) {
        std::memcpy(columns16, sd.envColumns.data(), sd.envColumns.size() * sizeof(rex::AliasEntry));
    }
    if (integratedLuminance != nullptr) *integratedLuminance = sd.envIntegrated;
    return REX_SUCCESS;
}

rex_status rex_renderer_create(rex_scene scene, uint32_t maxPixels, uint32_t maxSpp, rex_renderer* out) {
    if (out == nullptr) return fail(REX_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    if (maxPixels == 0 || maxSpp == 0) return fail(REX_INVALID_ARGUMENT, "maxPixels and maxSpp must be nonzero");
    const uint64_t cap64 = static_cast<uint64_t>(maxPixels) * maxSpp;
    if (cap64 >= 0x7fffffffull) return fail(REX_UNSUPPORTED, "more than 2^31 paths per launch");
    rex_renderer rr = new (std::nothrow) rex_renderer_t();
    if (rr == nullptr) return fail(REX_INVALID_ARGUMENT, "host allocation failed");
    rex::Renderer& r = rr->r;
    r.scene = &scene->s;
    r.maxPixels = maxPixels;
    r.maxSpp = maxSpp;
    r.capacity = static_cast<uint32_t>(cap64);
    const size_t cap = r.capacity;
    struct Alloc {
        void** p;
        size_t bytes;
    } allocs[] = {
        {reinterpret_cast<void**>(&r.paths), cap * sizeof(rex::PathState)},
        {reinterpret_cast<void**>(&r.rays[0]), cap * 2 * sizeof(rex::F4)},
        {reinterpret_cast<void**>(&r.rays[1]), cap * 2 * sizeof(rex::F4)},
        {reinterpret_cast<void**>(&r.rayPath[0]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&r.rayPath[1]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&r.hits), cap * sizeof(rex::QueueHit)},
        {reinterpret_cast<void**>(&r.shadows), cap * sizeof(rex::ShadowItem)},
        {reinterpret_cast<void**>(&r.counts), (rex::kCountWords + rex::kCounterWords) * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&r.film), static_cast<size_t>(maxPixels) * sizeof(float4)},
        {reinterpret_cast<void**>(&r.signals), cap * 5 * sizeof(float4)},
    };
    for (const Alloc& a : allocs) {
        const rex_status st = check_hip(hipMalloc(a.p, a.bytes), "hipMalloc renderer buffer");
        if (st != REX_SUCCESS) {
            r.release();
            delete rr;
            return st;
        }
    }
    if (rex_status st = check_hip(hipMemset(r.film, 0, static_cast<size_t>(maxPixels) * sizeof(float4)), "film clear")) {
        r.release();
        delete rr;
        return st;
    }
    *out = rr;
    return REX_SUCCESS;
}

rex_status rex_renderer_destroy(rex_renderer renderer) {
    if (renderer == nullptr) return REX_SUCCESS;
    renderer->r.release();
    delete renderer;
    return REX_SUCCESS;
}


rex_status rex_renderer_bind_scene(rex_renderer renderer, rex_scene scene, void* stream) {
    if (renderer == nullptr) return fail(REX_INVALID_ARGUMENT, "renderer is NULL");
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    rex::Renderer& r = renderer->r;
    r.scene = &scene->s;
    // A new ReSTIR sequence: frame counter, buffer parity and history as
    // after creation, so the sequence renders as on a fresh renderer.
    rex::RestirState& rs = r.restir;
    rs.hasHistory = false;
    rs.lastHistoryValid = false;
    rs.frameIndex = 0;
    rs.cur = 0;
    return rex_renderer_reset_film(renderer, stream);
}

rex_status rex_renderer_reset_film(rex_renderer renderer, void* stream) {
    if (renderer == nullptr) return fail(REX_INVALID_ARGUMENT, "renderer is NULL");
    return check_hip(hipMemsetAsync(renderer->r.film, 0, static_cast<size_t>(renderer->r.maxPixels) * sizeof(float4),
                                    static_cast<hipStream_t>(stream)),
                     "film clear");
}

rex_status rex_render_frame(rex_renderer renderer, const rex_frame_desc* fd, int keepSignals, void* stream) {
    if (renderer == nullptr || fd == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    rex::Renderer& r = renderer->r;
    rex::Scene& s = *r.scene;
    if (!s.shade.ready) return fail(REX_MISSING_SCENE, "scene has no shading tables (rex_scene_set_shading)");
    const uint64_t pixels = static_cast<uint64_t>(fd->width) * fd->height;
    if (pixels == 0 || pixels > r.maxPixels || fd->spp == 0 || fd->spp > r.maxSpp) {
        return fail(REX_INVALID_ARGUMENT, "frame exceeds the renderer's pixel / spp capacity");
    }
    if (fd->width > 16384u || fd->height > 16384u) return fail(REX_INVALID_ARGUMENT, "image dimension above 16384");
    if (fd->maxBounce == 0 || fd->maxBounce > rex::kMaxBounces) {
        return fail(REX_INVALID_ARGUMENT, "maxBounce must be in 1..16");
    }
    if (fd->estimator > REX_ESTIMATOR_RESTIR) {
        return fail(REX_UNSUPPORTED, "estimator must be BSDF-only, NEE, MIS or ReSTIR");
    }
    if (fd->queueMode > REX_QUEUE_PREFIX_SCAN) return fail(REX_INVALID_ARGUMENT, "unknown queue mode");
    const bool prefixScan = fd->queueMode == REX_QUEUE_PREFIX_SCAN;
    if (prefixScan && r.denseFlags == nullptr) {
        if (rex_status st = rex::dense_allocate(r)) return st;
    }
    const bool restir = fd->estimator == REX_ESTIMATOR_RESTIR;
    if (restir && fd->spp != 1u) return fail(REX_INVALID_ARGUMENT, "ReSTIR frames render one sample per pixel");
    if (restir && !r.restir.allocated) {
        if (rex_status st = rex::restir_allocate(r)) return st;
    }
    if (fd->walkMode != REX_WALK_SHORT_STACK && fd->walkMode != REX_WALK_NEAR_FIRST &&
        fd->walkMode != REX_WALK_WIDE && fd->walkMode != REX_WALK_WIDE_COMPRESSED) {
        return fail(REX_INVALID_ARGUMENT,
                    "frames use REX_WALK_SHORT_STACK, REX_WALK_NEAR_FIRST, REX_WALK_WIDE or REX_WALK_WIDE_COMPRESSED");
    }
    const rex_reconstruction_targets_v2& rt = r.recon;
    if (rt.pixels != 0u && rt.pixels < pixels) {
        return fail(REX_INVALID_ARGUMENT, "reconstruction targets hold fewer pixels than the frame");
    }
    rex::FrameConst f{};
    for (int k = 0; k < 3; ++k) {
        f.camPos[k] = fd->camera.position[k];
        f.camFwd[k] = fd->camera.forward[k];
        f.camRight[k] = fd->camera.right[k];
        f.camUp[k] = fd->camera.up[k];
    }
    f.tanHalfFov = fd->camera.tanHalfFov;
    f.aspect = fd->camera.aspect;
    f.width = fd->width;
    f.height = fd->height;
    f.sampleIndex = fd->sampleIndex;
    f.maxBounce = fd->maxBounce;
    f.spp = fd->spp;
    f.seedLow = static_cast<uint32_t>(fd->seed);
    f.seedHigh = static_cast<uint32_t>(fd->seed >> 32);
    f.streamTag = fd->streamTag != 0u ? fd->streamTag : 0x57a20001u;
    f.estimator = fd->estimator;
    const bool defaults = fd->rrStart == 0.0f && fd->rrMinimum == 0.0f && fd->rrMaximum == 0.0f;
    f.rrStart = defaults ? 3.0f : fd->rrStart;
    f.rrMinimum = defaults ? 0.05f : fd->rrMinimum;
    f.rrMaximum = defaults ? 0.95f : fd->rrMaximum;
    f.rayEpsilon = fd->rayEpsilon != 0.0f ? fd->rayEpsilon : 1.0e-4f;
    if (!(f.rayEpsilon > 0.0f) || !std::isfinite(f.rayEpsilon)) {
        return fail(REX_INVALID_ARGUMENT, "ray epsilon must be finite and positive");
    }
    f.pathCount = static_cast<uint32_t>(pixels * fd->spp);
    const uint32_t order = fd->walkMode == REX_WALK_WIDE_COMPRESSED
                               ? 3u
                               : (fd->walkMode == REX_WALK_WIDE ? 2u : (fd->walkMode == REX_WALK_NEAR_FIRST ? 1u : 0u));
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t* rayCounts = r.counts;
    uint32_t* shadowCounts = r.counts + 32;
    const rex::ShadingView shv = s.shade.view();

#define REX_TRY(expr, what)                                           \
    do {                                                              \
        const rex_status rs_ = check_hip((expr), what);               \
        if (rs_ != REX_SUCCESS) return rs_;                           \
    } while (0)
    const bool timing = (r.profile & REX_PROFILE_TIMING) != 0u;
    uint64_t* closestCnt = (r.profile & REX_PROFILE_COUNTERS) != 0u ? r.traversalCounters() : nullptr;
    uint64_t* shadowCnt = closestCnt != nullptr ? closestCnt + REX_COUNTER_COUNT : nullptr;
    r.eventsRecorded = 0;
    // Each event closes the span of the stage named by its tag.
    auto mark = [&](uint8_t stage) -> hipError_t {
        if (!timing) return hipSuccess;
        r.eventStage[r.eventsRecorded] = stage;
        return hipEventRecord(r.events[r.eventsRecorded++], st);
    };
    // Without traversal counters, a bounce's shadow queries and the next
    // bounce's closest-hit queries share one launch (both queues are
    // complete once Shade has run), halving the launches whose duration is
    // set by their slowest ray.
    const bool merge = closestCnt == nullptr;
    REX_TRY(mark(rex::kStageRaygen), "event record");
    REX_TRY(rex::launch_frame_reset(r.counts, rex::kCountWords + rex::kCounterWords, f.pathCount, st), "frame reset");
    REX_TRY(rex::launch_raygen(f, r.paths, r.rays[0], r.rayPath[0], st), "raygen");
    REX_TRY(mark(rex::kStageRaygen), "event record");
    REX_TRY(rex::launch_queue_closest(s.view, r.ws, r.rays[0], rayCounts, f.pathCount, r.hits, order, closestCnt, st),
            "intersect");
    REX_TRY(mark(rex::kStageIntersect), "event record");
    for (uint32_t b = 0; b < f.maxBounce; ++b) {
        const uint32_t cur = b & 1u, nxt = cur ^ 1u;
        rex::ShadeLaunch L;
        L.f = f;
        L.sh = shv;
        L.extended = shv.extended != 0u;
        L.s = s.view;
        L.bounce = b;
        L.capacity = f.pathCount;
        L.rays = r.rays[cur];
        L.rayPath = r.rayPath[cur];
        L.count = rayCounts + b;
        L.hits = r.hits;
        L.paths = r.paths;
        L.nextRays = r.rays[nxt];
        L.nextPath = r.rayPath[nxt];
        L.nextCount = rayCounts + b + 1;
        L.shadows = r.shadows;
        L.shadowCount = shadowCounts + b;
        // ReSTIR frames publish into their own surface ring (copied to the
        // bound primary-surface target by Resolve), other frames straight
        // into the target.
        L.surfaces = restir ? r.restir.surfaces[r.restir.cur] : static_cast<rex::SurfaceV2*>(rt.primarySurface);
        L.gbuffer = static_cast<rex::GBufferV2*>(rt.gbuffer);
        if (prefixScan) {
            L.denseFlags = r.denseFlags;
            L.denseRays = r.denseRays;
            L.densePath = r.densePath;
            L.denseShadows = r.denseShadows;
        }
        REX_TRY(rex::launch_shade(L, st), "shade");
        if (prefixScan) {
            rex::CompactQueues Q;
            Q.flags = r.denseFlags;
            Q.countPtr = rayCounts + b;
            Q.capacity = f.pathCount;
            Q.tiles = r.scanTiles;
            Q.nextCount = rayCounts + b + 1;
            Q.shadowCount = shadowCounts + b;
            Q.denseRays = r.denseRays;
            Q.densePath = r.densePath;
            Q.denseShadows = r.denseShadows;
            Q.nextRays = r.rays[nxt];
            Q.nextPath = r.rayPath[nxt];
            Q.shadows = r.shadows;
            REX_TRY(rex::launch_compact_queues(Q, st), "stable compaction");
        }
        REX_TRY(mark(rex::kStageShade), "event record");
        const bool more = b + 1u < f.maxBounce;
        if (merge && more) {
            REX_TRY(rex::launch_queue_mixed(s.view, r.ws, r.shadows, shadowCounts + b, r.rays[nxt], rayCounts + b + 1,
                                            2u * f.pathCount, r.hits, r.paths, order, st),
                    "trace shadow + intersect");
            REX_TRY(mark(rex::kStageIntersect), "event record");
            continue;
        }
        REX_TRY(rex::launch_queue_shadow(s.view, r.ws, r.shadows, shadowCounts + b, f.pathCount, r.paths, order,
                                         shadowCnt, st),
                "trace shadow");
        REX_TRY(mark(rex::kStageShadow), "event record");
        if (more) {
            REX_TRY(rex::launch_queue_closest(s.view, r.ws, r.rays[nxt], rayCounts + b + 1, f.pathCount, r.hits, order,
                                              closestCnt, st),
                    "intersect");
            REX_TRY(mark(rex::kStageIntersect), "event record");
        }
    }
    const float4* direct = nullptr;
    bool historyValid = false;
    rex::RestirState& rs = r.restir;
    const rex_restir_desc& rd = rs.desc;
    const uint32_t sceneGen = rd.sceneGeneration != 0u ? rd.sceneGeneration : (s.generation != 0u ? s.generation : 1u);
    const uint32_t lightGen = rd.lightGeneration != 0u ? rd.lightGeneration : 1u;
    const uint32_t configGen = rd.configGeneration != 0u ? rd.configGeneration : 1u;
    const uint32_t resourceGen = rd.resourceGeneration != 0u ? rd.resourceGeneration : 1u;
    const uint32_t historyGen = rd.historyGeneration != 0u ? rd.historyGeneration : 1u;
    if (restir) {
        // SameHistoryIdentity (src/renderers/ReSTIRDIRuntime.cpp:84-102) and
        // BuildReSTIRFrameParameters (src/renderers/ReSTIRFrameParameters.cpp:73-156).
        historyValid = rd.temporal != 0u && rs.hasHistory && rs.frameIndex > 0u && rs.lastWidth == f.width &&
                       rs.lastHeight == f.height && rs.lastScene == sceneGen && rs.lastLight == lightGen &&
                       rs.lastConfig == configGen && rs.lastResource == resourceGen && rs.lastHistoryGen == historyGen;
        rex::RestirLaunch L{};
        rex::RestirParams& p = L.params;
        p.extentAndCandidates = {f.width, f.height, f.width * f.height, rd.candidates};
        p.reuseLimits = {rd.neighbors, rd.maxM, rd.maxAge, 2u};
        p.generations = {sceneGen, lightGen, historyGen, static_cast<uint32_t>(rs.frameIndex)};
        p.historyGenerations = {configGen, resourceGen, 0u, 2u};
        p.modeAndFlags = {rd.estimatorMode, s.shade.selection == REX_LIGHT_SELECTION_UNIFORM ? 1u : 2u,
                          historyValid ? 1u : 0u,
                          static_cast<uint32_t>(rs.frameIndex >> 32)};
        p.lightTableCounts = {s.shade.lightCount, historyValid ? s.shade.lightCount : 0u, 0u, 0u};
        const bool defaultValidation = rd.validation[0] == 0.0f && rd.validation[1] == 0.0f &&
                                       rd.validation[2] == 0.0f && rd.validation[3] == 0.0f;
        p.validation = defaultValidation ? rex::F4{0.8f, 0.05f, 0.02f, 0.005f}
                                         : rex::F4{rd.validation[0], rd.validation[1], rd.validation[2],
                                                   rd.validation[3]};
        p.cameraPosition = {fd->camera.position[0], fd->camera.position[1], fd->camera.position[2], 0.0f};
        L.frame = f;
        L.shading = shv;
        const uint32_t cur = rs.cur, prev = cur ^ 1u;
        L.surfaces = rs.surfaces[cur];
        L.historySurfaces = rs.surfaces[prev];
        L.initial = rs.initial;
        L.temporal = rs.temporal;
        L.spatial = rs.spatial;
        L.history = rs.history[prev];
        L.published = rs.history[cur];
        L.reasons = rs.reasons;
        L.rays = rs.rays;
        L.hits = rs.hits;
        L.direct = rs.direct;
        L.stats = rs.stats;
        const rex_camera& prevCam = rs.hasHistory ? rs.lastCamera : fd->camera;
        rex::view_projection(fd->camera, L.vpCur);
        rex::view_projection(prevCam, L.vpPrev);
        rex::world_to_view(prevCam, L.wvPrev);
        L.neighborMode = rd.neighborMode;
        L.radius = rd.radius;
        L.estimatorMode = rd.estimatorMode;
        if (rd.objectMotion != 0u && s.dInstanceMap != nullptr && s.dObjectToWorld != nullptr) {
            L.instances = s.dInstances;
            L.instanceMap = s.dInstanceMap;
            L.objectToWorld = s.dObjectToWorld;
            L.prevObjectToWorld = s.dPrevObjectToWorld;
            L.instanceCount = s.instanceCount;
            L.mapCount = s.instanceMapCount;
        }
        if (rd.estimatorMode == REX_RESTIR_ESTIMATOR_REFERENCE_CORRECTION) {
            const size_t need = static_cast<size_t>(rd.neighbors + 1u) * f.width * f.height;
            if (rs.refCapacity < need) {
                if (rs.refRays) (void)hipFree(rs.refRays);
                if (rs.refHits) (void)hipFree(rs.refHits);
                rs.refRays = nullptr;
                rs.refHits = nullptr;
                rs.refCapacity = 0;
                REX_TRY(hipMalloc(reinterpret_cast<void**>(&rs.refRays), need * sizeof(rex_ray_v1)),
                        "hipMalloc reference-correction rays");
                REX_TRY(hipMalloc(reinterpret_cast<void**>(&rs.refHits), need * sizeof(rex_hit_v1)),
                        "hipMalloc reference-correction hits");
                rs.refCapacity = need;
            }
            L.refRays = rs.refRays;
            L.refHits = rs.refHits;
        }
        // Winner visibility reads only the hit kind, which no walk order
        // changes unless the reference's binary walk could overflow its
        // 64-entry stack (a hierarchy deeper than 63 levels): the queries
        // then run on the compressed walk, whose candidates the exact leaf
        // boxes admit, so the kind is the exact walks'.
        L.wideVisibility = s.maxDepth < 63u;
        REX_TRY(rex::launch_restir(L, s.view, r.ws, st), "restir");
        REX_TRY(mark(rex::kStageRestir), "event record");
        direct = rs.direct;
    }
    rex::ExportV2 ex;
    ex.gbuffer = static_cast<rex::GBufferV2*>(rt.gbuffer);
    ex.motion = static_cast<rex::MotionInputV2*>(rt.motionInputs);
    ex.raw = static_cast<rex::SignalV2*>(rt.rawSignal);
    ex.terminal = static_cast<float4*>(rt.terminal);
    if (restir) {
        ex.primary = static_cast<rex::SurfaceV2*>(rt.primarySurface);
        ex.restirSurfaces = rs.surfaces[rs.cur];
    }
    ex.instances = s.dInstances;
    ex.instanceMap = s.dInstanceMap;
    ex.instanceCount = s.instanceCount;
    ex.mapCount = s.instanceMapCount;
    REX_TRY(rex::launch_resolve_film(f, r.paths, r.film, keepSignals ? r.signals : nullptr, direct, ex, st),
            "resolve");
    REX_TRY(mark(rex::kStageResolve), "event record");
#undef REX_TRY
    if (restir) {
        rs.lastHistoryValid = historyValid;
        rs.lastWidth = f.width;
        rs.lastHeight = f.height;
        rs.lastScene = sceneGen;
        rs.lastLight = lightGen;
        rs.lastConfig = configGen;
        rs.lastResource = resourceGen;
        rs.lastHistoryGen = historyGen;
        rs.lastCamera = fd->camera;
        rs.hasHistory = true;
        rs.frameIndex++;
        rs.cur ^= 1u;
    }
    r.last = f;
    r.lastStream = st;
    return REX_SUCCESS;
}

rex_status rex_renderer_stats(rex_renderer renderer, rex_frame_stats* stats) {
    if (renderer == nullptr || stats == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    rex::Renderer& r = renderer->r;
    if (rex_status st = check_hip(hipStreamSynchronize(r.lastStream), "stream sync")) return st;
    uint32_t counts[rex::kCountWords + rex::kCounterWords];
    if (rex_status st = check_hip(hipMemcpy(counts, r.counts, sizeof(counts), hipMemcpyDeviceToHost), "stats copy")) {
        return st;
    }
    std::memset(stats, 0, sizeof(*stats));
    stats->pathCount = r.last.pathCount;
    stats->bounces = r.last.maxBounce;
    for (uint32_t b = 0; b < r.last.maxBounce && b < rex::kMaxBounces; ++b) {
        stats->raysPerBounce[b] = counts[b];
        stats->shadowsPerBounce[b] = counts[32 + b];
        stats->closestRays += counts[b];
        stats->shadowRays += counts[32 + b];
    }
    if ((r.profile & REX_PROFILE_COUNTERS) != 0u) {
        std::memcpy(stats->closestCounters, counts + rex::kCountWords, sizeof(stats->closestCounters));
        std::memcpy(stats->shadowCounters, counts + rex::kCountWords + 2u * REX_COUNTER_COUNT,
                    sizeof(stats->shadowCounters));
    }
    if ((r.profile & REX_PROFILE_TIMING) != 0u && r.eventsRecorded >= 2u) {
        for (uint32_t e = 1; e < r.eventsRecorded; ++e) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, r.events[e - 1u], r.events[e]);
            stats->stageMs[r.eventStage[e]] += ms;
            const bool trace = r.eventStage[e] == rex::kStageIntersect || r.eventStage[e] == rex::kStageShadow;
            if (trace && stats->traceLaunches < 17u) stats->traceMs[stats->traceLaunches++] = ms;
        }
    }
    return REX_SUCCESS;
}

uint64_t rex_compact_scratch_bytes(uint32_t count) { return rex::compact_tile_bytes(count) + 256u; }

rex_status rex_compact_stable(const uint32_t* flags, uint32_t count, uint32_t nextCapacity, uint32_t shadowCapacity,
                              void* scratch, uint32_t* prefix, uint32_t* nextIndices, uint32_t* shadowIndices,
                              uint32_t* status, void* stream) {
    if (status == nullptr) return fail(REX_INVALID_ARGUMENT, "status is NULL");
    if (count != 0u && (flags == nullptr || scratch == nullptr)) return fail(REX_INVALID_ARGUMENT, "flags / scratch are NULL");
    if (count > 0x80000000u) return fail(REX_INVALID_ARGUMENT, "at most 2^31 elements");
    return check_hip(rex::launch_compact_indices(reinterpret_cast<const uint2*>(flags), count, nextCapacity,
                                                 shadowCapacity, static_cast<uint2*>(scratch),
                                                 reinterpret_cast<uint2*>(prefix), nextIndices, shadowIndices, status,
                                                 static_cast<hipStream_t>(stream)),
                     "compaction launch");
}

rex_status rex_renderer_set_profiling(rex_renderer renderer, uint32_t flags) {
    if (renderer == nullptr) return fail(REX_INVALID_ARGUMENT, "renderer is NULL");
    if ((flags & ~(REX_PROFILE_COUNTERS | REX_PROFILE_TIMING)) != 0u) {
        return fail(REX_INVALID_ARGUMENT, "unknown profiling flag");
    }
    rex::Renderer& r = renderer->r;
    if ((flags & REX_PROFILE_TIMING) != 0u && r.events.empty()) {
        r.events.resize(rex::kEventCount, nullptr);
        for (hipEvent_t& e : r.events) {
            if (rex_status st = check_hip(hipEventCreate(&e), "hipEventCreate")) return st;
        }
    }
    r.profile = flags;
    return REX_SUCCESS;
}

rex_status rex_renderer_buffers(rex_renderer renderer, void** film, void** signals) {
    if (renderer == nullptr) return fail(REX_INVALID_ARGUMENT, "renderer is NULL");
    if (film != nullptr) *film = renderer->r.film;
    if (signals != nullptr) *signals = renderer->r.signals;
    return REX_SUCCESS;
}

rex_status rex_renderer_set_restir(rex_renderer renderer, const rex_restir_desc* desc) {
    if (renderer == nullptr || desc == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    const rex_restir_desc& d = *desc;
    if (d.candidates == 0u || d.candidates > 256u) return fail(REX_INVALID_ARGUMENT, "candidates must be in 1..256");
    if (d.neighbors > 32u) return fail(REX_INVALID_ARGUMENT, "neighbors must be in 0..32");
    if (d.maxM == 0u) return fail(REX_INVALID_ARGUMENT, "maxM must be positive");
    if (d.neighborMode > REX_RESTIR_NEIGHBORS_DISK) return fail(REX_INVALID_ARGUMENT, "unknown neighbour mode");
    if (d.estimatorMode > REX_RESTIR_ESTIMATOR_REFERENCE_CORRECTION) {
        return fail(REX_INVALID_ARGUMENT, "unknown ReSTIR estimator mode");
    }
    if (d.neighborMode == REX_RESTIR_NEIGHBORS_DISK && (d.radius == 0u || d.radius > 1024u)) {
        return fail(REX_INVALID_ARGUMENT, "disk radius must be in 1..1024");
    }
    for (int k = 0; k < 4; ++k) {
        if (!std::isfinite(d.validation[k])) return fail(REX_INVALID_ARGUMENT, "validation thresholds must be finite");
    }
    if (d.validation[0] < -1.0f || d.validation[0] > 1.0f || d.validation[1] < 0.0f || d.validation[2] < 0.0f ||
        d.validation[3] < 0.0f) {
        return fail(REX_INVALID_ARGUMENT, "validation thresholds out of range");
    }
    rex::RestirState& rs = renderer->r.restir;
    if (d.estimatorMode == REX_RESTIR_ESTIMATOR_BIASED && rs.refCapacity != 0) {
        // hipFree waits for the device, so no recorded frame still reads them.
        if (rs.refRays) (void)hipFree(rs.refRays);
        if (rs.refHits) (void)hipFree(rs.refHits);
        rs.refRays = nullptr;
        rs.refHits = nullptr;
        rs.refCapacity = 0;
    }
    rs.desc = d;
    rs.hasHistory = false;
    return REX_SUCCESS;
}

rex_status rex_renderer_reset_restir(rex_renderer renderer) {
    if (renderer == nullptr) return fail(REX_INVALID_ARGUMENT, "renderer is NULL");
    renderer->r.restir.hasHistory = false;
    return REX_SUCCESS;
}

rex_status rex_camera_matrices(const rex_camera* camera, float viewProjection[16], float worldToView[16]) {
    if (camera == nullptr || viewProjection == nullptr || worldToView == nullptr) {
        return fail(REX_INVALID_ARGUMENT, "NULL argument");
    }
    rex::view_projection(*camera, viewProjection);
    rex::world_to_view(*camera, worldToView);
    return REX_SUCCESS;
}

rex_status rex_renderer_set_reconstruction_v2(rex_renderer renderer, const rex_reconstruction_targets_v2* targets) {
    if (renderer == nullptr) return fail(REX_INVALID_ARGUMENT, "renderer is NULL");
    rex::Renderer& r = renderer->r;
    if (targets == nullptr) {
        r.recon = {};
        return REX_SUCCESS;
    }
    const rex_reconstruction_targets_v2& t = *targets;
    const void* ptrs[] = {t.gbuffer, t.motionInputs, t.rawSignal, t.primarySurface, t.terminal};
    bool any = false;
    for (const void* p : ptrs) {
        any = any || p != nullptr;
        if (reinterpret_cast<uintptr_t>(p) % 16u != 0u) {
            return fail(REX_INVALID_ARGUMENT, "reconstruction targets must be 16-byte aligned");
        }
    }
    if (!any) {
        r.recon = {};
        return REX_SUCCESS;
    }
    if (t.pixels == 0u) return fail(REX_INVALID_ARGUMENT, "reconstruction targets need a pixel capacity");
    if (t.motionInputs != nullptr && t.gbuffer == nullptr) {
        return fail(REX_INVALID_ARGUMENT, "motion inputs are derived from the G-buffer target");
    }
    r.recon = t;
    return REX_SUCCESS;
}

rex_status rex_renderer_restir_buffers(rex_renderer renderer, rex_restir_buffers* out) {
    if (renderer == nullptr || out == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    const rex::RestirState& rs = renderer->r.restir;
    std::memset(out, 0, sizeof(*out));
    if (!rs.allocated || rs.frameIndex == 0u) return fail(REX_INVALID_ARGUMENT, "no ReSTIR frame rendered yet");
    const uint32_t last = rs.cur ^ 1u;
    out->surfaces = rs.surfaces[last];
    out->initial = rs.initial;
    out->temporal = rs.temporal;
    out->spatial = rs.spatial;
    out->history = rs.history[last];
    out->direct = rs.direct;
    out->reasons = rs.reasons;
    out->statistics = rs.stats;
    out->frameIndex = rs.frameIndex - 1u;
    out->historyValid = rs.lastHistoryValid ? 1u : 0u;
    return REX_SUCCESS;
}

}  // extern "C"
