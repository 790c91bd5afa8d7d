// ReSTIR DI V3 reservoir passes on gfx950, one thread per pixel.
//
// A ReSTIR frame runs after the wavefront bounces (bounce 0 published the
// primary surfaces) and before Resolve + Film, on the frame's stream:
//
//   k_restir_initial     candidate export fused into initial RIS: each
//                        thread draws its C light samples in candidate
//                        order and streams them through its reservoir, so
//                        the C x 160-B candidate records of the reference
//                        (10.6 GB at 1080p x 32) are never written
//                        (pbr_restir_candidate_export_v3.hlsl:13-114 +
//                        RestirInitialV3.comp.hlsl:9-58)
//   k_restir_temporal    motion reprojection of the at-rest primary hit with
//                        the two cameras' view-projections
//                        (reconstruction/shaders/MotionVectorsCS.hlsl), then
//                        RestirTemporalV3.comp.hlsl:20-215
//   k_restir_spatial     RestirSpatialV3.comp.hlsl:12-112 (neighbour k of a
//                        pixel computed in place: the reference's host ring
//                        table, Wave3DebugRuntime.cpp:1888-1925, or a seeded
//                        disk), with the winner's visibility ray
//                        (RestirWinnerVisibilityPrepareV3.comp.hlsl:9-36)
//   any-hit traversal    ABI-v1 any-hit queries over the visibility rays
//   k_restir_resolve     RestirWinnerResolveV3.comp.hlsl:78-122 (Physical),
//                        RestirPublishDirectV3.comp.hlsl:11-42 and
//                        RestirPublishHistoryV3.comp.hlsl fused
//
// HBM per pixel and frame: surfaces 2 x 96 B read, reservoirs 256 B written
// by each of initial, temporal, spatial, resolve (+ history), read by the
// next stage (temporal also the previous frame's history, spatial K
// neighbours), 64-B ray + 96-B hit, 32 B direct light.
#include <hip/hip_runtime.h>

#include "motion.hpp"
#include "restir.hpp"
#include "rex_internal.hpp"

namespace rex {

constexpr int kRestirBlock = 256;

struct RestirArgs {
    RestirParams p;
    FrameConst f;
    ShadingView sh;
    const SurfaceV2* surfaces;
    const SurfaceV2* historySurfaces;
    ReservoirV3* initial;
    ReservoirV3* temporal;
    ReservoirV3* spatial;
    const ReservoirV3* history;
    ReservoirV3* published;
    uint32_t* reasons;
    rex_ray_v1* rays;
    const rex_hit_v1* hits;
    float4* direct;
    RestirStats* stats;
    float vpCur[16], vpPrev[16], wvPrev[16];
    uint32_t neighborMode, radius;
    uint32_t estimatorMode;
    rex_ray_v1* refRays;
    const rex_hit_v1* refHits;
    // Object motion (instanceMap NULL: surfaces at rest).
    const InstanceXform* instances;
    const uint2* instanceMap;
    const float4* objectToWorld;
    const float4* prevObjectToWorld;
    uint32_t instanceCount, mapCount;
};

// Sum of v over the wave, added to *counter by one lane.  Every lane of the
// wave must call it (the kernels keep their waves converged here).
__device__ __forceinline__ void wave_add(uint32_t* counter, uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63u) == 0u && v != 0u) atomicAdd(counter, v);
}

__global__ void k_restir_clear_stats(RestirStats* stats) {
    uint32_t* w = reinterpret_cast<uint32_t*>(stats);
    if (threadIdx.x < 16u) w[threadIdx.x] = 0u;
}

// PbrToPersistentLightSampleV3 (pbr_restir_candidate_export_v3.hlsl:35-78).
__device__ __forceinline__ LightSampleV3 to_persistent(const LightSample& ls, const ShadingView& sh, uint32_t flat,
                                                       const RestirParams& p) {
    LightSampleV3 o;
    const F4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    const U4 u = {0u, 0u, 0u, 0u};
    o.positionDistance = o.directionCombinedPdf = o.radianceDiscretePdf = o.conditionalPdf = z;
    o.identity = o.generation = o.metadata = o.sourceIdentity = u;
    if ((ls.flags & kSampleValid) == 0u || ls.lightIndex >= sh.lightCount) return o;
    o.positionDistance = {ls.position.x, ls.position.y, ls.position.z, ls.distance};
    o.directionCombinedPdf = {ls.wi.x, ls.wi.y, ls.wi.z, ls.combinedPdfW};
    o.radianceDiscretePdf = {ls.Li.x, ls.Li.y, ls.Li.z, ls.selectionPmf};
    if (ls.measure == kMeasureArea) o.conditionalPdf.x = ls.conditionalPdf;
    else if (ls.measure == kMeasureSolidAngle) o.conditionalPdf.y = ls.conditionalPdf;
    o.identity = {ls.stableId, ls.primitiveId, flat, p.generations.w};
    // The light generation stands for gCurrentToPreviousLightIndexV3[light].y
    // (Wave3DebugRuntime.cpp:1876-1886).
    o.generation = {p.generations.y, p.generations.x, p.generations.w, 0u};
    uint32_t flags = kRsSampleValid;
    if ((ls.flags & kSampleDelta) != 0u) flags |= kRsSampleDelta;
    if (ls.measure == kMeasureArea) flags |= kRsSampleArea;
    if (ls.measure == kMeasureSolidAngle) flags |= kRsSampleSolidAngle;
    // PbrPrivateLightMeasureToAbiV1L6 (pbr_l6_types.hlsli:240-254).
    const uint32_t abi = ls.measure == kMeasureDiscrete ? 1u
                         : ls.measure == kMeasureArea  ? 3u
                         : ls.measure == kMeasureSolidAngle ? 2u : 0u;
    o.metadata = {abi, flags, ls.lightFlags, ls.lightIndex};
    return o;
}

template <bool kExt>
__global__ __launch_bounds__(kRestirBlock) void k_restir_initial(RestirArgs A) {
    const RestirParams& p = A.p;
    const uint32_t pixel = blockIdx.x * kRestirBlock + threadIdx.x;
    const uint32_t pixels = p.extentAndCandidates.z, C = p.extentAndCandidates.w;
    uint32_t nValid = 0u, nRejected = 0u, nClamped = 0u;
    if (pixel < pixels) {
        const SurfaceV2 s = A.surfaces[pixel];
        ReservoirV3 r;
        reservoir_empty(r);
        uint32_t state = restir_hash(pixel ^ p.generations.w ^ 0xa511e9b3u);
        const uint32_t tag = A.f.streamTag ^ 0x52335354u;
        const V3 pos = v3(s.posDepth);
        for (uint32_t c = 0; c < C; ++c) {
            const uint32_t sampleIndex = A.f.sampleIndex * C + c;
            // Dimensions 0..4 (selection, shapes 0..3); shapes 2 and 3 are
            // drawn only by the importance-sampled environment.
            const float r0 = unit_float(philox_x(pixel, sampleIndex, 0u, tag, A.f.seedLow, A.f.seedHigh));
            const float r1 = unit_float(philox_x(pixel, sampleIndex, 1u, tag, A.f.seedLow, A.f.seedHigh));
            const float r2 = unit_float(philox_x(pixel, sampleIndex, 2u, tag, A.f.seedLow, A.f.seedHigh));
            const LightSample ls = sample_one_light<kExt>(A.sh, pos, r0, r1, r2, [&](uint32_t k) {
                return unit_float(philox_x(pixel, sampleIndex, 1u + k, tag, A.f.seedLow, A.f.seedHigh));
            });
            LightSampleV3 ps = to_persistent(ls, A.sh, pixel * C + c, p);
            ps.sourceIdentity = {pixel, s.identity.x, s.identity.y, s.identity.z};
            const CandTerms ct = evaluate_candidate(ps, s, p, p.modeAndFlags.y, kReuseNone, pixel);
            const float w = candidate_weight(ps, ct, p);
            if (candidate_valid(ps, ct, p)) ++nValid;
            else ++nRejected;
            reservoir_update(r, ps, ct, p, w, 1u, restir_random(state));
        }
        reservoir_clamp(r, p.reuseLimits.y);
        reservoir_finalize(r);
        r.state.y = 0u;
        r.surfaceIdentity = {s.identity.x, s.identity.y, s.identity.z, 0u};
        r.sceneIdentity = {p.generations.x, p.generations.y, p.historyGenerations.x, p.historyGenerations.y};
        r.frameIdentity = {p.generations.w, p.modeAndFlags.w, p.generations.z, kHistoryValid};
        r.reprojection = {pixel, p.extentAndCandidates.x, p.extentAndCandidates.y, kRejNone};
        A.initial[pixel] = r;
        nClamped = (r.state.z & kResMClamped) != 0u ? 1u : 0u;
    }
    wave_add(&A.stats->candidateCounts.x, pixel < pixels ? C : 0u);
    wave_add(&A.stats->candidateCounts.y, nValid);
    wave_add(&A.stats->candidateCounts.z, nRejected);
    wave_add(&A.stats->candidateCounts.w, nClamped);
}

// mul(row_major M, float4(p, 1)).
__device__ __forceinline__ F4 xform(const float* m, V3 q) {
    F4 o;
    o.x = ((m[0] * q.x + m[1] * q.y) + m[2] * q.z) + m[3] * 1.0f;
    o.y = ((m[4] * q.x + m[5] * q.y) + m[6] * q.z) + m[7] * 1.0f;
    o.z = ((m[8] * q.x + m[9] * q.y) + m[10] * q.z) + m[11] * 1.0f;
    o.w = ((m[12] * q.x + m[13] * q.y) + m[14] * q.z) + m[15] * 1.0f;
    return o;
}

__global__ __launch_bounds__(kRestirBlock) void k_restir_temporal(RestirArgs A) {
    const RestirParams& p = A.p;
    const uint32_t pixel = blockIdx.x * kRestirBlock + threadIdx.x;
    const uint32_t pixels = p.extentAndCandidates.z;
    const uint32_t width = p.extentAndCandidates.x, height = p.extentAndCandidates.y;
    const bool historyEnabled = p.modeAndFlags.z != 0u;
    uint32_t accepted = 0u;
    if (pixel < pixels) {
        ReservoirV3 result = A.initial[pixel];
        uint32_t rejection = kRejNone;
        if (historyEnabled) {
            const SurfaceV2 s = A.surfaces[pixel];
            bool mvalid = false;
            float pu = 0.0f, pv = 0.0f, expected = 0.0f;
            if (A.instanceMap != nullptr) {
                // MotionVectorsCS over the surface's motion input: the world
                // position taken to its instance's object space, then through
                // the current and previous object-to-world (no jitter).
                if ((s.identity.w & kSurfaceValid) != 0u) {
                    const MotionInputV2 in = surface_motion_input(A.instances, A.instanceMap, A.instanceCount,
                                                                  A.mapCount, s);
                    const float4 vc[4] = {make_float4(A.vpCur[0], A.vpCur[1], A.vpCur[2], A.vpCur[3]),
                                          make_float4(A.vpCur[4], A.vpCur[5], A.vpCur[6], A.vpCur[7]),
                                          make_float4(A.vpCur[8], A.vpCur[9], A.vpCur[10], A.vpCur[11]),
                                          make_float4(A.vpCur[12], A.vpCur[13], A.vpCur[14], A.vpCur[15])};
                    const float4 vp[4] = {make_float4(A.vpPrev[0], A.vpPrev[1], A.vpPrev[2], A.vpPrev[3]),
                                          make_float4(A.vpPrev[4], A.vpPrev[5], A.vpPrev[6], A.vpPrev[7]),
                                          make_float4(A.vpPrev[8], A.vpPrev[9], A.vpPrev[10], A.vpPrev[11]),
                                          make_float4(A.vpPrev[12], A.vpPrev[13], A.vpPrev[14], A.vpPrev[15])};
                    const float4 wv[4] = {make_float4(A.wvPrev[0], A.wvPrev[1], A.wvPrev[2], A.wvPrev[3]),
                                          make_float4(A.wvPrev[4], A.wvPrev[5], A.wvPrev[6], A.wvPrev[7]),
                                          make_float4(A.wvPrev[8], A.wvPrev[9], A.wvPrev[10], A.wvPrev[11]),
                                          make_float4(A.wvPrev[12], A.wvPrev[13], A.wvPrev[14], A.wvPrev[15])};
                    SurfaceMotion m;
                    if (surface_motion(in, A.objectToWorld, A.instanceCount, A.prevObjectToWorld, A.instanceCount, vc,
                                       vp, wv, 0.0f, 0.0f, 0.0f, 0.0f, m)) {
                        mvalid = m.valid;
                        pu = mvalid ? m.pu : m.cu;
                        pv = mvalid ? m.pv : m.cv;
                        expected = mvalid ? m.depth : 0.0f;
                    }
                }
            } else if ((s.identity.w & kSurfaceValid) != 0u && fin3(v3(s.posDepth))) {
                // MotionVectorsCS for a surface at rest (no jitter).
                const V3 q = v3(s.posDepth);
                const F4 cc = xform(A.vpCur, q), pc = xform(A.vpPrev, q), vv = xform(A.wvPrev, q);
                if (cc.w > 1.0e-8f && pc.w > 1.0e-8f && fin4(cc) && fin4(pc) && fin4(vv)) {
                    const float cnx = cc.x / cc.w, cny = cc.y / cc.w;
                    const float pnx = pc.x / pc.w, pny = pc.y / pc.w;
                    const float cu = (cnx * 0.5f + 0.5f) + 0.0f, cv = (0.5f - cny * 0.5f) + 0.0f;
                    const float qu = (pnx * 0.5f + 0.5f) + 0.0f, qv = (0.5f - pny * 0.5f) + 0.0f;
                    const float depth = -vv.z;
                    mvalid = fin(cu) && fin(cv) && fin(qu) && fin(qv) && fin(depth) && depth >= 0.0f;
                    pu = mvalid ? qu : cu;
                    pv = mvalid ? qv : cv;
                    expected = mvalid ? depth : 0.0f;
                }
            }
            uint32_t prior = kInvalid;
            if (!mvalid) {
                rejection = kRejMotion;
            } else if (pu < 0.0f || pv < 0.0f || pu >= 1.0f || pv >= 1.0f) {
                rejection = kRejOutside;
            } else {
                uint32_t qx = (uint32_t)(pu * (float)width), qy = (uint32_t)(pv * (float)height);
                if (qx > width - 1u) qx = width - 1u;
                if (qy > height - 1u) qy = height - 1u;
                prior = qy * width + qx;
            }
            ReservoirV3 h;
            if (prior < pixels) h = A.history[prior];
            else reservoir_empty(h);
            if (rejection == kRejNone &&
                ((h.state.z & kResValid) == 0u || h.state.x == 0u || (h.frameIdentity.w & kHistoryValid) == 0u))
                rejection = kRejInvalid;
            uint32_t expLow = p.generations.w, expHigh = p.modeAndFlags.w;
            if (expLow == 0u) {
                expLow = 0xffffffffu;
                expHigh -= 1u;
            } else {
                expLow -= 1u;
            }
            if (rejection == kRejNone &&
                (h.frameIdentity.x != expLow || h.frameIdentity.y != expHigh || h.frameIdentity.z != p.generations.z))
                rejection = kRejInvalid;
            else if (rejection == kRejNone && h.sceneIdentity.x != p.generations.x)
                rejection = kRejScene;
            else if (rejection == kRejNone && h.sceneIdentity.y != p.generations.y)
                rejection = kRejLight;
            else if (rejection == kRejNone && h.sceneIdentity.z != p.historyGenerations.x)
                rejection = kRejCameraCut;
            else if (rejection == kRejNone && h.sceneIdentity.w != p.historyGenerations.y)
                rejection = kRejResize;
            else if (rejection == kRejNone && h.state.y >= p.reuseLimits.z)
                rejection = kRejAge;
            else if (rejection == kRejNone && prior >= pixels)
                rejection = kRejOutside;
            else if (rejection == kRejNone)
                rejection = validate_pair(s, A.historySurfaces[prior], p);
            if (rejection == kRejNone) {
                const float hd = A.historySurfaces[prior].posDepth.w;
                const float scale = fmaxf(1.0f, fmaxf(fabsf(expected), fabsf(hd)));
                if (fabsf(expected - hd) / scale > p.validation.y) rejection = kRejDepth;
            }
            const uint32_t prevLight = h.selected.metadata.w;
            if (rejection == kRejNone) {
                // gPreviousToCurrentLightIndexV3: the identity map with the
                // light generation while history is reused.
                if (prevLight >= p.lightTableCounts.y) rejection = kRejLight;
                else if (prevLight >= p.lightTableCounts.x || p.generations.y != h.selected.generation.x)
                    rejection = kRejLight;
            }
            if (rejection == kRejNone) {
                LightSampleV3 remapped = h.selected;
                if (remapped.positionDistance.w < 1.0e29f) {
                    const V3 delta = v3(remapped.positionDistance) - v3(s.posDepth);
                    const float dist = length(delta);
                    if (fin(dist) && dist > 1.0e-6f) {
                        remapped.positionDistance.w = dist;
                        const V3 dir = delta / dist;
                        remapped.directionCombinedPdf.x = dir.x;
                        remapped.directionCombinedPdf.y = dir.y;
                        remapped.directionCombinedPdf.z = dir.z;
                    } else {
                        remapped.metadata.y &= ~kRsSampleValid;
                    }
                }
                remapped.metadata.w = prevLight;
                CandTerms ct = evaluate_candidate(remapped, s, p, h.provenance.x, kReuseTemporal, prior);
                ct.correction = h.weightState.y;
                const float mergeWeight = ct.target * ct.support * h.weightState.y * (float)h.state.x;
                uint32_t state = restir_hash(pixel ^ p.generations.w ^ 0x63d83595u);
                if (candidate_valid(remapped, ct, p) && fin(mergeWeight) && mergeWeight >= 0.0f) {
                    reservoir_update(result, remapped, ct, p, mergeWeight, h.state.x, restir_random(state));
                    result.state.z |= kResTemporal;
                    const uint32_t age = h.state.y + 1u;
                    result.state.y = age < p.reuseLimits.z ? age : p.reuseLimits.z;
                    accepted = 1u;
                } else {
                    rejection = kRejInvalid;
                }
            }
        }
        reservoir_clamp(result, p.reuseLimits.y);
        reservoir_finalize(result);
        result.state.w = rejection;
        result.reprojection.w = rejection;
        A.reasons[pixel] = rejection;
        A.temporal[pixel] = result;
    }
    wave_add(&A.stats->reuseCounts.x, pixel < pixels && historyEnabled ? 1u : 0u);
    wave_add(&A.stats->reuseCounts.y, accepted);
}

// Neighbour k of a pixel: the reference's ring table or a seeded disk.
__device__ __forceinline__ uint32_t neighbor_of(const RestirArgs& A, uint32_t pixel, uint32_t k) {
    const int32_t width = (int32_t)A.p.extentAndCandidates.x, height = (int32_t)A.p.extentAndCandidates.y;
    const int32_t x = (int32_t)(pixel % (uint32_t)width), y = (int32_t)(pixel / (uint32_t)width);
    int32_t dx, dy;
    if (A.neighborMode == kNeighborsRing) {
        const int32_t ring = (int32_t)(k / 8u) + 1;
        const uint32_t d = k % 8u;
        // (1,0) (-1,0) (0,1) (0,-1) (1,1) (-1,1) (1,-1) (-1,-1)
        const int32_t ux = d == 0u || d == 4u || d == 6u ? 1 : (d == 1u || d == 5u || d == 7u ? -1 : 0);
        const int32_t uy = d == 2u || d == 4u || d == 5u ? 1 : (d == 3u || d == 6u || d == 7u ? -1 : 0);
        dx = ux * ring;
        dy = uy * ring;
    } else {
        uint32_t state = restir_hash(pixel ^ restir_hash(A.p.generations.w + k * 0x9e3779b9u) ^ 0x2c1b3c6du);
        const float u0 = restir_random(state);
        const float u1 = restir_random(state);
        const float r = (float)A.radius * sqrtf(u0);
        float s, c;
        sincos_2pi(u1, s, c);
        dx = (int32_t)floorf(r * c + 0.5f);
        dy = (int32_t)floorf(r * s + 0.5f);
    }
    int32_t sx = x + dx, sy = y + dy;
    sx = sx < 0 ? 0 : (sx > width - 1 ? width - 1 : sx);
    sy = sy < 0 ? 0 : (sy > height - 1 ? height - 1 : sy);
    return (uint32_t)(sy * width + sx);
}

// RestirMakeVisibilityRayV3 (Physical: one ray) into an ABI-v1 record with
// identity (rayIndex, queryIndex, 0, mask); winner rays also skip reservoirs
// whose visibility was already evaluated (RestirWinnerVisibilityPrepareV3).
__device__ __forceinline__ void visibility_ray(uint32_t rayIndex, uint32_t queryIndex, const ReservoirV3& r,
                                               const SurfaceV2& s, bool winner, rex_ray_v1& out, bool& submitted) {
    F4 o = {0.0f, 0.0f, 0.0f, 0.0f}, d = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t mask = 0u;
    if ((!winner || (r.state.z & kResVisEvaluated) == 0u) && (r.state.z & kResValid) != 0u && surface_valid(s)) {
        const V3 n = normalize(v3(s.gnRough));
        V3 dir = normalize(v3(r.selected.directionCombinedPdf));
        float maxDist = 1.0e30f;
        const float pd = r.selected.positionDistance.w;
        if (fin(pd) && pd > 0.0f) {
            const V3 disp = v3(r.selected.positionDistance) - v3(s.posDepth);
            const float dist = length(disp);
            if (dist > 2.0e-4f) {
                dir = disp / dist;
                maxDist = fmaxf(dist - 1.0e-4f, 1.0e-4f);
            }
        }
        const V3 org = v3(s.posDepth) + n * 1.0e-4f;
        o = {org.x, org.y, org.z, 1.0e-4f};
        d = {dir.x, dir.y, dir.z, maxDist};
        mask = 0xffu;
    }
    submitted = mask != 0u;
    float4* f = reinterpret_cast<float4*>(&out);
    f[0] = make_float4(o.x, o.y, o.z, o.w);
    f[1] = make_float4(d.x, d.y, d.z, d.w);
    uint4* u = reinterpret_cast<uint4*>(&out);
    u[2] = make_uint4(rayIndex, queryIndex, 0u, mask);
    u[3] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ __launch_bounds__(kRestirBlock) void k_restir_spatial(RestirArgs A) {
    const RestirParams& p = A.p;
    const uint32_t pixel = blockIdx.x * kRestirBlock + threadIdx.x;
    const uint32_t pixels = p.extentAndCandidates.z, K = p.reuseLimits.x;
    uint32_t attempts = 0u, accepted = 0u, submitted = 0u;
    if (pixel < pixels) {
        const SurfaceV2 s = A.surfaces[pixel];
        ReservoirV3 result;
        if (K == 0u) {
            result = A.temporal[pixel];
        } else {
            const ReservoirV3 center = A.temporal[pixel];
            reservoir_empty(result);
            uint32_t state = restir_hash(pixel ^ p.generations.w ^ 0x9e3779b9u);
            uint32_t aggregate = A.reasons[pixel];
            CandTerms cc = evaluate_candidate(center.selected, s, p, center.provenance.x, center.provenance.y, pixel);
            cc.correction = center.weightState.y;
            reservoir_update(result, center.selected, cc, p,
                             cc.target * cc.support * center.weightState.y * (float)center.state.x, center.state.x,
                             restir_random(state));
            result.state.z |= center.state.z & (kResMClamped | kResTemporal);
            for (uint32_t k = 0; k < K; ++k) {
                ++attempts;
                const uint32_t nb = neighbor_of(A, pixel, k);
                if (nb >= pixels) {
                    aggregate = kRejOutside;
                    continue;
                }
                const uint32_t rej = validate_pair(s, A.surfaces[nb], p);
                const ReservoirV3& n = A.temporal[nb];
                const uint32_t nflags = n.state.z;
                if (rej != kRejNone || (nflags & kResValid) == 0u) {
                    aggregate = rej != kRejNone ? rej : kRejInvalid;
                    continue;
                }
                const LightSampleV3 ns = n.selected;
                const float nW = n.weightState.y;
                const uint32_t nM = n.state.x;
                CandTerms c = evaluate_candidate(ns, s, p, n.provenance.x, kReuseSpatial, nb);
                c.correction = nW;
                const float mw = c.target * c.support * nW * (float)nM;
                if (candidate_valid(ns, c, p) && fin(mw) && mw >= 0.0f) {
                    reservoir_update(result, ns, c, p, mw, nM, restir_random(state));
                    result.state.z |= kResSpatial;
                    ++accepted;
                } else {
                    aggregate = kRejInvalid;
                }
            }
            reservoir_clamp(result, p.reuseLimits.y);
            reservoir_finalize(result);
            result.state.y = center.state.y;
            result.surfaceIdentity = center.surfaceIdentity;
            result.sceneIdentity = center.sceneIdentity;
            result.frameIdentity = center.frameIdentity;
            result.reprojection = center.reprojection;
            result.reprojection.x = pixel;
            result.reprojection.w = aggregate;
            result.state.w = aggregate;
            A.reasons[pixel] = aggregate;
        }
        A.spatial[pixel] = result;
        // The reference-correction estimator prepares the winner's ray after
        // it has re-weighted the reservoir; here it writes the pixel's K + 1
        // reference queries (RestirReferenceVisibilityPrepareV3) while the
        // reservoir and the neighbours' surfaces are at hand.
        if (A.estimatorMode == REX_RESTIR_ESTIMATOR_REFERENCE_CORRECTION) {
            const uint32_t sources = K + 1u;
            for (uint32_t offset = 0u; offset < sources; ++offset) {
                const uint32_t src = offset == 0u ? pixel : neighbor_of(A, pixel, offset - 1u);
                const uint32_t q = pixel * sources + offset;
                bool sub = false;
                visibility_ray(q, q, result, A.surfaces[src], false, A.refRays[q], sub);
                submitted += sub ? 1u : 0u;
            }
        } else {
            bool sub = false;
            visibility_ray(pixel, pixel, result, s, true, A.rays[pixel], sub);
            submitted = sub ? 1u : 0u;
        }
    }
    wave_add(&A.stats->reuseCounts.z, attempts);
    wave_add(&A.stats->reuseCounts.w, accepted);
    wave_add(&A.stats->visibilityCounts.x, submitted);
}

// RestirReferenceResolveV3 (RestirReferenceResolveV3.comp.hlsl:82-144), then
// the winner's visibility ray of the re-weighted reservoir
// (RestirWinnerVisibilityPrepareV3.comp.hlsl).
__global__ __launch_bounds__(kRestirBlock) void k_restir_reference_resolve(RestirArgs A) {
    const RestirParams& p = A.p;
    const uint32_t pixels = p.extentAndCandidates.z, sources = p.reuseLimits.x + 1u;
    const uint32_t center = blockIdx.x * kRestirBlock + threadIdx.x;
    uint32_t evaluated = 0u, visibleCount = 0u, submitted = 0u;
    if (center < pixels) {
        ReservoirV3 r = A.spatial[center];
        if ((r.state.z & kResValid) != 0u) {
            float denominator = 0.0f, selectedSourceTarget = 0.0f, currentTarget = 0.0f;
            for (uint32_t offset = 0u; offset < sources; ++offset) {
                const uint32_t src = offset == 0u ? center : neighbor_of(A, center, offset - 1u);
                const uint32_t q = center * sources + offset;
                const uint4 meta = reinterpret_cast<const uint4*>(&A.refHits[q])[4];
                const bool matches = meta.x == q && meta.w == q && meta.y != 3u;
                const float visibility = matches && meta.y == 0u ? 1.0f : 0.0f;
                V3 dv, sv;
                float target;
                evaluate_unshadowed(r.selected, A.surfaces[src], p, dv, sv, target);
                const uint32_t m = A.temporal[src].state.x;
                const uint32_t sourceM = m > 1u ? m : 1u;
                denominator += (float)sourceM * target * visibility;
                if (src == center) currentTarget = target;
                if (src == r.provenance.z) selectedSourceTarget = target * visibility;
                ++evaluated;
                visibleCount += visibility > 0.0f ? 1u : 0u;
            }
            const float norm = currentTarget * denominator;
            r.weightState.y = norm > 0.0f ? r.weightState.x * selectedSourceTarget / norm : 0.0f;
            r.selectedTerms.w = r.weightState.y;
            r.state.z |= kResReferenceMode;
            if (!(r.weightState.y > 0.0f) || !fin(r.weightState.y)) r.state.z &= ~kResValid;
            A.spatial[center] = r;
        }
        bool sub = false;
        visibility_ray(center, center, r, A.surfaces[center], true, A.rays[center], sub);
        submitted = sub ? 1u : 0u;
    }
    wave_add(&A.stats->visibilityCounts.y, evaluated);
    wave_add(&A.stats->visibilityCounts.z, visibleCount);
    wave_add(&A.stats->visibilityCounts.x, submitted);
}

__global__ __launch_bounds__(kRestirBlock) void k_restir_resolve(RestirArgs A) {
    const RestirParams& p = A.p;
    const uint32_t pixel = blockIdx.x * kRestirBlock + threadIdx.x;
    const uint32_t pixels = p.extentAndCandidates.z;
    uint32_t evaluated = 0u, visibleCount = 0u, duplicates = 0u;
    if (pixel < pixels) {
        ReservoirV3 r = A.spatial[pixel];
        const bool valid = (r.state.z & kResValid) != 0u;
        const bool duplicate = (r.state.z & kResVisEvaluated) != 0u;
        if (valid && !duplicate) {
            const uint4 meta = reinterpret_cast<const uint4*>(&A.hits[pixel])[4];
            const bool matches = meta.x == pixel && meta.w == pixel && meta.y != 3u;
            const float vis = matches && meta.y == 0u ? 1.0f : 0.0f;
            const float fw = r.weightState.y * vis;
            const bool finite = fin(vis) && vis >= 0.0f && vis <= 1.0f && fin(fw) && fw >= 0.0f;
            const bool visible = finite && vis > 0.0f;
            r.state.z |= kResVisEvaluated;
            r.selectedTerms.w = finite ? fw : 0.0f;
            if (!finite) {
                r.state.z &= ~kResValid;
                r.state.w = kRejInvalid;
                r.reprojection.w = kRejInvalid;
            }
            if (visible) r.state.z |= kResVisValid;
            else r.state.z &= ~kResVisValid;
            evaluated = 1u;
            visibleCount = visible ? 1u : 0u;
        } else if (duplicate) {
            duplicates = 1u;
        }
        A.spatial[pixel] = r;
        // RestirPublishDirectV3.
        const SurfaceV2 s = A.surfaces[pixel];
        V3 dv = zero3(), sv = zero3();
        const uint32_t need = kResValid | kResVisEvaluated | kResVisValid;
        if ((r.state.z & need) == need) {
            float target;
            evaluate_unshadowed(r.selected, s, p, dv, sv, target);
            dv = dv * r.selectedTerms.w;
            sv = sv * r.selectedTerms.w;
        }
        A.direct[2u * pixel] = make_float4(dv.x, dv.y, dv.z, 1.0f);
        A.direct[2u * pixel + 1u] = make_float4(sv.x, sv.y, sv.z, 1.0f);
        // RestirPublishHistoryV3.
        r.surfaceIdentity = {s.identity.x, s.identity.y, s.identity.z, 0u};
        r.sceneIdentity.x = p.generations.x;
        r.sceneIdentity.y = p.generations.y;
        const uint32_t hneed = kResValid | kResVisEvaluated;
        r.frameIdentity = {p.generations.w, p.modeAndFlags.w, p.generations.z,
                           (r.state.z & hneed) == hneed ? kHistoryValid : kHistoryReset};
        r.reprojection = {pixel, p.extentAndCandidates.x, p.extentAndCandidates.y, r.reprojection.w};
        A.published[pixel] = r;
    }
    wave_add(&A.stats->visibilityCounts.y, evaluated);
    wave_add(&A.stats->visibilityCounts.z, visibleCount);
    wave_add(&A.stats->visibilityCounts.w, duplicates);
}

// ---- host side -----------------------------------------------------------------

hipError_t launch_restir(const RestirLaunch& L, const SceneView& scene, TraceWorkspace& ws, hipStream_t stream) {
    RestirArgs A;
    A.p = L.params;
    A.f = L.frame;
    A.sh = L.shading;
    A.surfaces = L.surfaces;
    A.historySurfaces = L.historySurfaces;
    A.initial = L.initial;
    A.temporal = L.temporal;
    A.spatial = L.spatial;
    A.history = L.history;
    A.published = L.published;
    A.reasons = L.reasons;
    A.rays = L.rays;
    A.hits = L.hits;
    A.direct = L.direct;
    A.stats = L.stats;
    for (int i = 0; i < 16; ++i) {
        A.vpCur[i] = L.vpCur[i];
        A.vpPrev[i] = L.vpPrev[i];
        A.wvPrev[i] = L.wvPrev[i];
    }
    A.neighborMode = L.neighborMode;
    A.radius = L.radius;
    A.estimatorMode = L.estimatorMode;
    A.refRays = L.refRays;
    A.refHits = L.refHits;
    A.instances = L.instances;
    A.instanceMap = L.instanceMap;
    A.objectToWorld = L.objectToWorld;
    A.prevObjectToWorld = L.prevObjectToWorld;
    A.instanceCount = L.instanceCount;
    A.mapCount = L.mapCount;
    const uint32_t pixels = L.params.extentAndCandidates.z;
    const dim3 grid((pixels + kRestirBlock - 1) / kRestirBlock), block(kRestirBlock);
    hipLaunchKernelGGL(k_restir_clear_stats, dim3(1), dim3(64), 0, stream, L.stats);
    if (A.sh.extended != 0u) hipLaunchKernelGGL(k_restir_initial<true>, grid, block, 0, stream, A);
    else hipLaunchKernelGGL(k_restir_initial<false>, grid, block, 0, stream, A);
    hipLaunchKernelGGL(k_restir_temporal, grid, block, 0, stream, A);
    hipLaunchKernelGGL(k_restir_spatial, grid, block, 0, stream, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (L.estimatorMode == REX_RESTIR_ESTIMATOR_REFERENCE_CORRECTION) {
        const uint32_t queries = pixels * (L.params.reuseLimits.x + 1u);
        e = L.wideVisibility ? launch_visibility_v1(scene, ws, L.refRays, L.refHits, queries, stream)
                             : launch_trace_v1(scene, ws, L.refRays, L.refHits, queries, true, 64u, nullptr, stream, false);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_restir_reference_resolve, grid, block, 0, stream, A);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    e = L.wideVisibility ? launch_visibility_v1(scene, ws, L.rays, L.hits, pixels, stream)
                         : launch_trace_v1(scene, ws, L.rays, L.hits, pixels, true, 64u, nullptr, stream, false);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_restir_resolve, grid, block, 0, stream, A);
    return hipGetLastError();
}

}  // namespace rex
