// ReSTIR DI V3 device library: the versioned records and the per-pixel
// arithmetic of the reservoir passes.
//
// Restates (paths relative to the RenderingEngine tree):
//   records                include/contracts/RestirAbiV3.hpp:109-193,
//                          include/contracts/ReconstructionAbiV2.hpp (GpuPrimarySurfaceV2)
//   hash / random          resources/shaders/restir/RestirProductionV3.hlsli:16-30
//   visibility ray         RestirProductionV3.hlsli:140-205 (Physical shadow mode)
//   validity / target      RestirProductionV3.hlsli:225-350
//   RIS update / finalize  RestirProductionV3.hlsli:352-452
//   surface-pair check     RestirProductionV3.hlsli:454-482
//
// Every expression keeps the operand order of the CPU oracle (shading.hpp's
// conventions: no contraction, dot = (x*x + y*y) + z*z, normalize = v * (1 /
// sqrt(dot))), so reservoirs come out bit-identical to the oracle's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "render_layout.hpp"
#include "restir_layout.hpp"
#include "rex_layout.hpp"
#include "shading.hpp"

namespace rex {

// Surface / sample / reservoir / history flags (ReconstructionAbiV2.hpp,
// RestirAbiV3.hpp:50-101).
constexpr uint32_t kSurfaceValid = 1u;
constexpr uint32_t kSurfaceFront = 2u;
constexpr uint32_t kSurfaceDiffuse = 4u;
constexpr uint32_t kSurfaceSpecular = 8u;
constexpr uint32_t kRsSampleValid = 1u;
constexpr uint32_t kRsSampleDelta = 2u;
constexpr uint32_t kRsSampleArea = 4u;
constexpr uint32_t kRsSampleSolidAngle = 8u;
constexpr uint32_t kResValid = 1u;
constexpr uint32_t kResMClamped = 2u;
constexpr uint32_t kResTemporal = 4u;
constexpr uint32_t kResSpatial = 8u;
constexpr uint32_t kResReferenceMode = 16u;
constexpr uint32_t kResVisEvaluated = 32u;
constexpr uint32_t kResVisValid = 64u;
constexpr uint32_t kHistoryValid = 1u;
constexpr uint32_t kHistoryReset = 2u;
constexpr uint32_t kReuseNone = 0u;
constexpr uint32_t kReuseTemporal = 1u;
constexpr uint32_t kReuseSpatial = 2u;
constexpr uint32_t kRejNone = 0u;
constexpr uint32_t kRejInvalid = 1u;
constexpr uint32_t kRejOutside = 2u;
constexpr uint32_t kRejMotion = 3u;
constexpr uint32_t kRejCameraCut = 4u;
constexpr uint32_t kRejResize = 5u;
constexpr uint32_t kRejDepth = 6u;
constexpr uint32_t kRejNormal = 7u;
constexpr uint32_t kRejInstance = 8u;
constexpr uint32_t kRejThin = 9u;
constexpr uint32_t kRejScene = 10u;
constexpr uint32_t kRejLight = 11u;
constexpr uint32_t kRejAge = 12u;
constexpr uint32_t kRejMaterial = 13u;
constexpr uint32_t kNeighborsRing = 0u;
constexpr uint32_t kNeighborsDisk = 1u;

__device__ __forceinline__ float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ V3 max0(V3 a) { return v3(fmaxf(a.x, 0.0f), fmaxf(a.y, 0.0f), fmaxf(a.z, 0.0f)); }
__device__ __forceinline__ bool fin4(const F4& a) { return fin(a.x) && fin(a.y) && fin(a.z) && fin(a.w); }

// RestirHashV3 / RestirRandomV3.
__device__ __forceinline__ uint32_t restir_hash(uint32_t v) {
    v ^= v >> 16u;
    v *= 0x7feb352du;
    v ^= v >> 15u;
    v *= 0x846ca68bu;
    v ^= v >> 16u;
    return v;
}
__device__ __forceinline__ float restir_random(uint32_t& state) {
    state = restir_hash(state + 0x9e3779b9u);
    return (float)(state >> 8u) * (1.0f / 16777216.0f);
}

// RestirSurfaceValidV3.
__device__ __forceinline__ bool surface_valid(const SurfaceV2& s) {
    const F4& d = s.diffuse;
    const F4& sp = s.specular;
    const V3 gn = v3(s.gnRough), sn = v3(s.snMetal);
    return (s.identity.w & kSurfaceValid) != 0u && fin4(s.posDepth) && fin4(s.gnRough) && fin4(s.snMetal) &&
           fin4(d) && fin4(sp) && d.x >= 0.0f && d.y >= 0.0f && d.z >= 0.0f && d.x <= 1.0f && d.y <= 1.0f &&
           d.z <= 1.0f && sp.x >= 0.0f && sp.y >= 0.0f && sp.z >= 0.0f && sp.x <= 1.0f && sp.y <= 1.0f &&
           sp.z <= 1.0f && s.snMetal.w >= 0.0f && s.snMetal.w <= 1.0f && roughness_finite(s.gnRough.w) &&
           dot(gn, gn) > 1.0e-12f && dot(sn, sn) > 1.0e-12f;
}

// RestirSampleValidV3.
__device__ __forceinline__ bool sample_valid(const LightSampleV3& s, const RestirParams& p) {
    const float cp = s.directionCombinedPdf.w, dp = s.radianceDiscretePdf.w;
    return (s.metadata.y & kRsSampleValid) != 0u && s.identity.x != kInvalid && s.generation.x != 0u &&
           s.generation.y == p.generations.x && fin(cp) && cp >= 0.0f && fin(dp) && dp >= 0.0f &&
           fin3(v3(s.directionCombinedPdf)) && fin3(v3(s.radianceDiscretePdf));
}

// RestirEvaluateUnshadowedV3: the surface's unshadowed diffuse and specular
// response to the sample and its luminance (the RIS target).
__device__ __forceinline__ void evaluate_unshadowed(const LightSampleV3& s, const SurfaceV2& surf,
                                                    const RestirParams& p, V3& diffuse, V3& specular,
                                                    float& target) {
    diffuse = specular = zero3();
    target = 0.0f;
    if (!sample_valid(s, p) || !surface_valid(surf)) return;
    const V3 n = normalize(v3(surf.snMetal));
    const V3 l = normalize(v3(s.directionCombinedPdf));
    const V3 v = normalize(v3(p.cameraPosition) - v3(surf.posDepth));
    const float noL = sat(dot(n, l));
    const float noV = sat(dot(n, v));
    if (!(noL > 0.0f) || !(noV > 0.0f)) return;
    const V3 h = normalize(l + v);
    const float noH = sat(dot(n, h));
    const float voH = sat(dot(v, h));
    const float roughness = surf.gnRough.w;
    const V3 diffAlb = max0(v3(surf.diffuse));
    const V3 f0 = max0(v3(surf.specular));
    const V3 f = schlick(voH, f0);
    const float alpha = roughness * roughness;
    const float d = ggx_d(v3(0.0f, 0.0f, noH), alpha);
    const float g = ggx_g2(v3(sqrtf(fmaxf(0.0f, 1.0f - noV * noV)), 0.0f, noV),
                           v3(sqrtf(fmaxf(0.0f, 1.0f - noL * noL)), 0.0f, noL), alpha);
    const V3 radiance = max0(v3(s.radianceDiscretePdf));
    const V3 dv = radiance * diffAlb * (v3(1.0f, 1.0f, 1.0f) - f) * (noL * kInvPi);
    const V3 sv = radiance * f * (d * g / fmaxf(4.0f * noV, 1.0e-6f));
    if (!fin3(dv) || !fin3(sv)) return;
    diffuse = dv;
    specular = sv;
    target = lum(max0(dv + sv));
}

// GpuRestirCandidateV3 terms of one sample at one surface
// (RestirEvaluateCandidateV3); the sample itself stays where it is.
struct CandTerms {
    float target, proposal, support, correction;
    uint32_t source, reuse, surface, sampleFlags;
};

__device__ __forceinline__ CandTerms evaluate_candidate(const LightSampleV3& s, const SurfaceV2& surf,
                                                        const RestirParams& p, uint32_t source, uint32_t reuse,
                                                        uint32_t sourceSurface) {
    V3 dv, sv;
    float target;
    evaluate_unshadowed(s, surf, p, dv, sv, target);
    CandTerms c;
    const float proposal = s.directionCombinedPdf.w;
    const bool surfaceValid = surface_valid(surf);
    const bool support = sample_valid(s, p) && surfaceValid && target > 0.0f && proposal > 0.0f;
    c.target = target;
    c.proposal = proposal;
    c.support = support ? 1.0f : 0.0f;
    c.correction = surfaceValid ? 1.0f : 0.0f;
    c.source = source;
    c.reuse = reuse;
    c.surface = sourceSurface;
    c.sampleFlags = s.metadata.y;
    return c;
}

// RestirCandidateValidV3.
__device__ __forceinline__ bool candidate_valid(const LightSampleV3& s, const CandTerms& c, const RestirParams& p) {
    const bool consistent = c.target > 0.0f ? c.support == 1.0f : c.support == 0.0f;
    return sample_valid(s, p) && fin(c.target) && c.target >= 0.0f && fin(c.proposal) && c.proposal > 0.0f &&
           (c.support == 0.0f || c.support == 1.0f) && fin(c.correction) && c.correction >= 0.0f &&
           c.correction > 0.0f && consistent;
}

// RestirCandidateWeightV3.
__device__ __forceinline__ float candidate_weight(const LightSampleV3& s, const CandTerms& c,
                                                  const RestirParams& p) {
    if (!candidate_valid(s, c, p)) return 0.0f;
    return c.target * c.support * c.correction / c.proposal;
}

// RestirEmptyReservoirV3.
__device__ __forceinline__ void reservoir_empty(ReservoirV3& r) {
    const F4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    const U4 u = {0u, 0u, 0u, 0u};
    r.selected.positionDistance = r.selected.directionCombinedPdf = r.selected.radianceDiscretePdf = z;
    r.selected.conditionalPdf = z;
    r.selected.identity = {kInvalid, kInvalid, 0u, 0u};
    r.selected.generation = r.selected.metadata = r.selected.sourceIdentity = u;
    r.weightState = r.selectedTerms = z;
    r.state = r.provenance = r.surfaceIdentity = r.sceneIdentity = r.frameIdentity = r.reprojection = u;
}

// RestirUpdateReservoirV3: streaming RIS update by one (merged) candidate.
__device__ __forceinline__ void reservoir_update(ReservoirV3& r, const LightSampleV3& s, const CandTerms& c,
                                                 const RestirParams& p, float weight, uint32_t multiplicity,
                                                 float randomValue) {
    if (!candidate_valid(s, c, p) || !fin(weight) || weight < 0.0f || multiplicity == 0u) return;
    const uint32_t oldM = r.state.x;
    const uint32_t newM = oldM > 0xffffffffu - multiplicity ? 0xffffffffu : oldM + multiplicity;
    if (weight == 0.0f) {
        r.state.x = newM;
        return;
    }
    const float newSum = r.weightState.x + weight;
    if (!(newSum > 0.0f) || !fin(newSum)) return;
    r.weightState.x = newSum;
    r.state.x = newM;
    if (randomValue * newSum < weight) {
        r.selected = s;
        r.weightState.z = c.target;
        r.weightState.w = c.proposal;
        r.selectedTerms = {c.support, c.correction, weight, 0.0f};
        r.provenance = {c.source, c.reuse, c.surface, c.sampleFlags};
    }
}

// RestirFinalizeReservoirV3: W = wSum / (M * target).
__device__ __forceinline__ void reservoir_finalize(ReservoirV3& r) {
    const float wsum = r.weightState.x, target = r.weightState.z;
    if (r.state.x == 0u || !(target > 0.0f) || !(wsum > 0.0f) || !fin(target) || !fin(wsum)) {
        r.state.z &= ~kResValid;
        r.weightState.y = 0.0f;
        return;
    }
    const float den = (float)r.state.x * target;
    if (!(den > 0.0f) || !fin(den)) {
        r.state.z &= ~kResValid;
        r.weightState.y = 0.0f;
        return;
    }
    r.weightState.y = wsum / den;
    if (!(r.weightState.y > 0.0f) || !fin(r.weightState.y)) {
        r.state.z &= ~kResValid;
        r.weightState.y = 0.0f;
        return;
    }
    r.selectedTerms.w = r.weightState.y;
    r.state.z |= kResValid;
}

// RestirClampMV3.
__device__ __forceinline__ void reservoir_clamp(ReservoirV3& r, uint32_t maxM) {
    if (r.state.x > maxM) {
        const float scale = (float)maxM / (float)r.state.x;
        r.weightState.x *= scale;
        r.state.x = maxM;
        r.state.z |= kResMClamped;
    }
}

// RestirValidateSurfacePairV3.
__device__ __forceinline__ uint32_t validate_pair(const SurfaceV2& cur, const SurfaceV2& src, const RestirParams& p) {
    if (!surface_valid(cur) || !surface_valid(src)) return kRejInvalid;
    if (cur.identity.x != src.identity.x) return kRejMaterial;
    if (cur.identity.y != src.identity.y) return kRejInstance;
    const float dc = cur.posDepth.w, ds = src.posDepth.w;
    const float scale = fmaxf(1.0f, fmaxf(fabsf(dc), fabsf(ds)));
    if (fabsf(dc - ds) / scale > p.validation.y) return kRejDepth;
    if (dot(normalize(v3(cur.snMetal)), normalize(v3(src.snMetal))) < p.validation.x) return kRejNormal;
    const float dist = length(v3(cur.posDepth) - v3(src.posDepth));
    if (dist > p.validation.z) return kRejThin;
    return kRejNone;
}

}  // namespace rex
