// Device records of the wavefront integrator (shared by host setup and the
// HIP kernels).
//
// HBM layout of one frame launch with P = width * height * spp paths
// (path index p = pixel * spp + s, so a pixel's samples are adjacent):
//   PathState  paths[P]        192 B: throughputs, the five signal
//                              accumulators, previous vertex, flags.
//   QueueRay   rays[2][P]      32 B: origin + tMin, direction + tMax; the
//                              owning path index lives in a parallel uint
//                              array (rayPath[2][P]); ping-pong per bounce.
//   QueueHit   hits[P]         16 B: t, u, v, triangle | front-face bit (the
//                              compact closest-hit result; normals and IDs
//                              are rebuilt from the triangle streams).
//   ShadowItem shadow[P]       64 B: origin + tMin, direction + tMax, diffuse
//                              contribution + path index, specular
//                              contribution + ignored primitive.
//   counters   per bounce      queue sizes (device-side, no host readback
//                              between stages).
#pragma once

#include <stdint.h>

#include "rex_layout.hpp"

namespace rex {

// BSDF models / lobes (resources/shaders/include/bsdf/PbrBsdf.hlsli:55-80)
constexpr uint32_t kModelLambert = 0u;
constexpr uint32_t kModelConductor = 1u;
constexpr uint32_t kModelDielectricReflection = 2u;
constexpr uint32_t kModelSmoothGlass = 3u;
constexpr uint32_t kModelRoughDielectric = 4u;
constexpr uint32_t kModelMetallicRoughness = 5u;
constexpr uint32_t kLobeDiffuse = 1u;
constexpr uint32_t kLobeGlossy = 2u;
constexpr uint32_t kLobeSpecular = 4u;
constexpr uint32_t kLobeReflection = 8u;
constexpr uint32_t kLobeTransmission = 16u;
constexpr uint32_t kMaterialPureEmitter = 1u;

// Light types / flags (resources/shaders/integrators/pbr_l6_types.hlsli:10-19)
constexpr uint32_t kLightPoint = 0u;
constexpr uint32_t kLightDirectional = 1u;
constexpr uint32_t kLightSpot = 2u;
constexpr uint32_t kLightSphere = 3u;
constexpr uint32_t kLightEmissiveTriangle = 4u;
constexpr uint32_t kLightEnvironment = 5u;
constexpr uint32_t kLightEnabled = 1u;
constexpr uint32_t kLightTwoSided = 2u;

// Path flags (integrators/wavefront/include/WavefrontTypes.hpp path flags)
constexpr uint32_t kPathActive = 1u;
constexpr uint32_t kPathPrevDelta = 2u;
constexpr uint32_t kPathPrevSpecular = 4u;
constexpr uint32_t kPathPrevRestir = 8u;
constexpr uint32_t kPathTerminated = 16u;
constexpr uint32_t kPathError = 32u;

// Direct-light estimators (pbr_l6_types.hlsli:39-42)
constexpr uint32_t kEstBsdfOnly = 0u;
constexpr uint32_t kEstNee = 1u;
constexpr uint32_t kEstMis = 2u;
constexpr uint32_t kEstRestir = 3u;

struct alignas(16) Material {  // PbrMaterialGpu, 128 B
    F4 baseColorMetallic;
    F4 emissiveRoughness;
    F4 transmissionIor;
    F4 attenuationColorDistance;
    F4 f0;
    F4 conductorEta;
    F4 conductorK;
    U4 metadata;  // model, allowed lobes, complex Fresnel, private flags
};
static_assert(sizeof(Material) == 128, "PbrMaterialGpu is 128 B");

struct alignas(16) Light {  // PbrLightGpu, 96 B
    F4 positionRange;
    F4 directionCosOuter;
    F4 radianceScale;
    F4 shapeParams;
    U4 identity;  // type, flags, -, primitive
    U4 payload;   // fixture triangle, -, -, instance
};
static_assert(sizeof(Light) == 96, "PbrLightGpu is 96 B");

struct AliasEntry {
    float q, pmf;
    uint32_t alias, item;
};

struct alignas(16) Fixture {  // FixtureTriangleGpu: canonical order
    F4 p0, p1, p2;
    U4 metadata;  // material, primitive, instance, geometry flags
};

struct ShadingView {
    const Material* materials;
    uint32_t materialCount;
    const Light* lights;
    uint32_t lightCount;
    const AliasEntry* alias;
    uint32_t aliasCount;
    const float* selectionPmf;
    const Fixture* fixtures;
    uint32_t fixtureCount;
    const uint32_t* emitterOfTri;  // leaf-order triangle -> light index (or invalid)
    uint32_t environmentLight;
    // Lat-long environment map (nearest texel, u repeats, v clamps) and its
    // two-level alias distribution: envRows[height], envColumns[height x width].
    const F4* envTexels;
    const AliasEntry* envRows;
    const AliasEntry* envColumns;
    uint32_t envWidth, envHeight;
    uint32_t envSampler;  // 0 uniform sphere, 1 importance map
    uint32_t extended;    // BSDF models 1-4 / transmission, sphere lights or a lat-long map
};

struct FrameConst {
    float camPos[3];
    float tanHalfFov;
    float camFwd[3];
    float aspect;
    float camRight[3];
    float camUp[3];
    uint32_t width, height;
    uint32_t sampleIndex;
    uint32_t maxBounce;
    uint32_t spp;
    uint32_t seedLow, seedHigh, streamTag;
    uint32_t estimator;
    float rrStart, rrMinimum, rrMaximum, rayEpsilon;
    uint32_t pathCount;
};

// One job of the batch engine (an arena of resident scenes, paths of many
// jobs in one pool): its frame constants, its scene's shading tables, its
// Film and the resolved value of each of its samples.  A pooled path's state
// carries meta.y = job slot, meta.z = bounce, meta.w = the path's index
// pixel * spp + s within its job.
struct alignas(16) BatchJob {
    FrameConst f;
    ShadingView sh;
    uint32_t scene;       // arena scene index (root table entry)
    uint32_t pixelCount;  // width * height
    uint32_t pathCount;   // pixelCount * spp
    uint32_t pad;
    F4* film;             // width * height float4 running means
    F4* samples;          // pathCount resolved sample values, pixel-major
};

struct alignas(16) PathState {
    F4 betaEta;     // beta.xyz, eta scale
    F4 betaD;       // diffuse throughput, w unused
    F4 betaS;       // specular throughput
    F4 cam;         // camera emission
    F4 dd, ds;      // direct diffuse / specular
    F4 id, is;      // indirect diffuse / specular
    F4 prevPosPdf;  // previous vertex, BSDF pdf
    F4 prevGn;
    F4 prevSn;
    U4 meta;        // flags, closest rays, shadow rays, bounce
};

struct alignas(16) QueueHit {
    float t, u, v;
    uint32_t tri;  // leaf-order triangle | front << 31; kInvalid = miss; 0x7fffffff = invalid traversal
};

struct alignas(16) ShadowItem {
    F4 originTmin;
    F4 dirTmax;
    F4 diffPath;   // diffuse contribution, path index (bits)
    F4 specPrim;   // specular contribution, ignored primitive (bits)
};

constexpr uint32_t kHitMiss = 0xffffffffu;
constexpr uint32_t kHitInvalid = 0x7fffffffu;
constexpr uint32_t kHitFrontBit = 0x80000000u;

}  // namespace rex
