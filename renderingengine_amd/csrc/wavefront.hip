// Wavefront integrator stages on gfx950.
//
// One frame launch renders P = width * height * spp paths (path p = pixel *
// spp + s, sample index = base + s) as a fixed sequence of kernels on one
// stream, with queue sizes kept in device memory (no host readback between
// stages, so the sequence can be captured in a graph):
//
//   k_frame_reset                      queue counters
//   k_raygen                           RayGenCS: path state + ray queue 0
//   launch_queue_closest               IntersectCS of bounce 0: persistent closest-hit walk
//   per bounce b < maxBounce:
//     k_shade                          ShadeCS + WfPbrBuildDirectShadow + NextBounceCS
//     launch_queue_mixed               TraceShadowCS (Physical) of bounce b, accumulating
//                                      in place, and IntersectCS of bounce b + 1 in one
//                                      launch (the last bounce: launch_queue_shadow)
//   k_resolve_film                     ResolveCS + the Film running mean of ComposeCS
// (render.cpp issues the shadow and closest launches separately when
// traversal counters are recorded.)
//
// Queue appends use one atomic per wave: a ballot over the lanes that append,
// the leader reserves popcount slots, each lane takes base + its rank.  Append
// order varies from run to run, but every record carries its path index and
// every result is keyed by path, so images are deterministic.
//
// The arithmetic of each stage restates the reference shaders exactly as the
// CPU oracle does (see shading.hpp for the shared numeric conventions).
#include <hip/hip_runtime.h>

#include "motion.hpp"
#include "restir.hpp"
#include "rex_internal.hpp"
#include "shading.hpp"
#include "traverse.hpp"

namespace rex {

constexpr int kShadeBlock = 256;

struct ShadeParams {
    FrameConst f;
    ShadingView sh;
    SceneView s;
    uint32_t bounce;
    const F4* rays;            // current queue: 2 x F4 per ray
    const uint32_t* rayPath;
    const uint32_t* count;     // current queue size
    const QueueHit* hits;
    PathState* paths;
    F4* nextRays;
    uint32_t* nextPath;
    uint32_t* nextCount;
    ShadowItem* shadows;
    uint32_t* shadowCount;
    SurfaceV2* surfaces;
    GBufferV2* gbuffer;
    // Batched jobs (k_shade<true>): per-path job table and the scene of each
    // appended ray / shadow item.
    const BatchJob* jobs;
    uint32_t* nextSceneOf;
    uint32_t* shadowSceneOf;
    uint32_t* retire;
    uint32_t* retireCount;
    uint2* denseFlags;          // PrefixScan mode: flags and records at the item's index
    F4* denseRays;
    uint32_t* densePath;
    ShadowItem* denseShadows;
};

__global__ void k_frame_reset(uint32_t* counts, uint32_t n, uint32_t pathCount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counts[i] = i == 0u ? pathCount : 0u;
}

// RayGenCS (integrators/wavefront/shaders/RayGen.hlsl:10-93).
__global__ __launch_bounds__(256) void k_raygen(FrameConst f, PathState* __restrict__ paths, F4* __restrict__ rays,
                                                uint32_t* __restrict__ rayPath) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= f.pathCount) return;
    const uint32_t pixel = p / f.spp;
    const uint32_t sampleIndex = f.sampleIndex + (p - pixel * f.spp);
    const uint32_t px = pixel % f.width, py = pixel / f.width;
    const float jx = sample_dim(pixel, sampleIndex, 0u, f);
    const float jy = sample_dim(pixel, sampleIndex, 1u, f);
    float sx = ((float)px + jx) / (float)f.width;
    float sy = ((float)py + jy) / (float)f.height;
    sx = sx * 2.0f - 1.0f;
    sy = sy * 2.0f - 1.0f;
    sy = -sy;
    const V3 fwd = v3(f.camFwd[0], f.camFwd[1], f.camFwd[2]);
    const V3 right = v3(f.camRight[0], f.camRight[1], f.camRight[2]);
    const V3 up = v3(f.camUp[0], f.camUp[1], f.camUp[2]);
    const V3 d = normalize(fwd + right * (sx * f.aspect * f.tanHalfFov) + up * (sy * f.tanHalfFov));
    rays[2u * p] = {f.camPos[0], f.camPos[1], f.camPos[2], f.rayEpsilon};
    rays[2u * p + 1u] = {d.x, d.y, d.z, 1.0e30f};
    rayPath[p] = p;
    PathState st;
    st.betaEta = {1.0f, 1.0f, 1.0f, 1.0f};
    st.betaD = st.betaS = st.cam = st.dd = st.ds = st.id = st.is = {0.0f, 0.0f, 0.0f, 0.0f};
    st.prevPosPdf = st.prevGn = st.prevSn = {0.0f, 0.0f, 0.0f, 0.0f};
    st.meta = {kPathActive | kPathPrevDelta, 0u, 0u, 0u};
    paths[p] = st;
}

// Appends `want` lanes' records to a queue with one atomic per wave; returns
// the slot of this lane (valid where want).
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t* counter) {
    const uint64_t m = __ballot(want);
    if (m == 0u) return 0u;
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Appends records to kQ queues with one atomic per queue per block: the
// block's waves put their ballot counts in LDS, thread q adds queue q's total
// to its counter.  Every thread of the block calls it once per iteration of
// a block-uniform loop; `parity` alternates the LDS buffers, so a wave that
// runs ahead into the next iteration does not overwrite counts that a slower
// wave still reads.  Same-address atomics serialize (about 11 ns each on the
// MI355X), so per-wave appends would cap a queue at ~5.6 G records/s.
template <int kQ, int kBlock>
struct AppendLds {
    uint32_t cnt[2][kQ][kBlock / 64];
    uint32_t base[2][kQ];
};

template <int kQ, int kBlock>
__device__ __forceinline__ void block_append(AppendLds<kQ, kBlock>& L, uint32_t parity, const bool (&want)[kQ],
                                             uint32_t* const (&counter)[kQ], uint32_t (&slot)[kQ]) {
    constexpr int kWaves = kBlock / 64;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t m[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        m[q] = __ballot(want[q]);
        if (lane == 0u) L.cnt[parity][q][wave] = (uint32_t)__popcll(m[q]);
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kQ) {
        const uint32_t q = threadIdx.x;
        uint32_t total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) total += L.cnt[parity][q][w];
        L.base[parity][q] = total != 0u ? atomicAdd(counter[q], total) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        uint32_t off = L.base[parity][q];
        for (uint32_t w = 0; w < wave; ++w) off += L.cnt[parity][q][w];
        slot[q] = off + (uint32_t)__popcll(m[q] & below);
    }
}

// Path state held in registers by Shade.  The five signal accumulators stay
// in memory: they change only on a terminal contribution (an emitter or an
// environment miss), so they are read and written there alone instead of
// being carried through the BSDF and light sampling.
struct St {
    V3 beta;
    float eta;
    V3 betaD, betaS, prevPos;
    float prevPdf;
    V3 prevGn, prevSn;
    uint32_t flags;
};

__device__ __forceinline__ St load_state(const PathState& p) {
    St s;
    s.beta = v3(p.betaEta);
    s.eta = p.betaEta.w;
    s.betaD = v3(p.betaD);
    s.betaS = v3(p.betaS);
    s.prevPos = v3(p.prevPosPdf);
    s.prevPdf = p.prevPosPdf.w;
    s.prevGn = v3(p.prevGn);
    s.prevSn = v3(p.prevSn);
    s.flags = p.meta.x;
    return s;
}

__device__ __forceinline__ F4 f4(V3 a, float w) { return {a.x, a.y, a.z, w}; }

__device__ __forceinline__ void store_state(PathState& p, const St& s) {
    p.betaEta = f4(s.beta, s.eta);
    p.betaD = f4(s.betaD, 0.0f);
    p.betaS = f4(s.betaS, 0.0f);
    p.prevPosPdf = f4(s.prevPos, s.prevPdf);
    p.prevGn = f4(s.prevGn, 0.0f);
    p.prevSn = f4(s.prevSn, 0.0f);
    p.meta.x = s.flags;
}

// WfPbrAccumulateTerminal (PbrWavefrontShading.hlsli:105-160).
__device__ __forceinline__ bool accumulate_terminal(PathState& ps, const St& st, uint32_t bounce, V3 emitted,
                                                    float mis) {
    if (!fin_nonneg3(emitted) || !fin(mis) || mis < 0.0f || !fin_nonneg3(st.beta) || !fin_nonneg3(st.betaD) ||
        !fin_nonneg3(st.betaS))
        return false;
    const V3 w = emitted * mis;
    if (!fin_nonneg3(w)) return false;
    if (bounce == 0u) {
        const V3 a = v3(ps.cam) + st.beta * w;
        if (!fin_nonneg3(a)) return false;
        ps.cam = f4(a, 0.0f);
    } else if (bounce == 1u) {
        const V3 d = v3(ps.dd) + st.betaD * w;
        const V3 s = v3(ps.ds) + st.betaS * w;
        if (!fin_nonneg3(d) || !fin_nonneg3(s)) return false;
        ps.dd = f4(d, 0.0f);
        ps.ds = f4(s, 0.0f);
    } else {
        const V3 d = v3(ps.id) + st.betaD * w;
        const V3 s = v3(ps.is) + st.betaS * w;
        if (!fin_nonneg3(d) || !fin_nonneg3(s)) return false;
        ps.id = f4(d, 0.0f);
        ps.is = f4(s, 0.0f);
    }
    return true;
}

__device__ __forceinline__ void terminate(St& st, bool error) {
    st.flags = (st.flags & ~kPathActive) | kPathTerminated | (error ? kPathError : 0u);
}

// PbrPublishPrimarySurfaceV2 (resources/shaders/integrators/
// pbr_reconstruction_export_v2.hlsli:44-131): the ReSTIR copy of bounce 0's
// surface, valid only when ReSTIR owns its direct light.
__device__ __forceinline__ void publish_surface(const FrameConst& f, V3 pos, V3 gn, V3 sn, const TriAttr& attr,
                                                uint32_t materialIndex, bool front, const Material& mat, bool owned,
                                                SurfaceV2* out, GBufferV2* gb) {
    const float metallic = mat.baseColorMetallic.w, roughness = mat.emissiveRoughness.w;
    const float transmission = mat.transmissionIor.x;
    const V3 base = v3(mat.baseColorMetallic), f0 = v3(mat.f0);
    const float depth = dot(pos - v3(f.camPos[0], f.camPos[1], f.camPos[2]),
                            normalize(v3(f.camFwd[0], f.camFwd[1], f.camFwd[2])));
    const V3 g = normalize(gn), n = normalize(sn);
    const float sm = fminf(fmaxf(metallic, 0.0f), 1.0f);
    const float st = fminf(fmaxf(transmission, 0.0f), 1.0f);
    const V3 b0 = v3(fmaxf(base.x, 0.0f), fmaxf(base.y, 0.0f), fmaxf(base.z, 0.0f));
    const V3 diff = b0 * (1.0f - sm) * (1.0f - st);
    const V3 l = f0 + (base - f0) * sm;
    const V3 spec = v3(fmaxf(l.x, 0.0f), fmaxf(l.y, 0.0f), fmaxf(l.z, 0.0f));
    uint32_t flags = kSurfaceValid;
    if (diff.x > 0.0f || diff.y > 0.0f || diff.z > 0.0f) flags |= kSurfaceDiffuse;
    if (spec.x > 0.0f || spec.y > 0.0f || spec.z > 0.0f) flags |= kSurfaceSpecular;
    if (front) flags |= kSurfaceFront;
    const float4 r0 = make_float4(pos.x, pos.y, pos.z, fmaxf(depth, 0.0f));
    const float4 r1 = make_float4(g.x, g.y, g.z, fminf(fmaxf(roughness, 0.0f), 1.0f));
    const float4 r2 = make_float4(n.x, n.y, n.z, sm);
    const float4 r3 = make_float4(diff.x, diff.y, diff.z, 0.0f);
    const float4 r4 = make_float4(spec.x, spec.y, spec.z, 0.0f);
    // The G-buffer keeps the surface valid; the ReSTIR copy is valid only
    // where ReSTIR owns its direct light (pbr_reconstruction_export_v2.hlsli:84-88).
    SurfaceV2* dst[2] = {out, gb != nullptr ? &gb->primary : nullptr};
    const uint32_t fl[2] = {owned ? flags : flags & ~kSurfaceValid, flags};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (dst[c] == nullptr) continue;
        float4* o = reinterpret_cast<float4*>(dst[c]);
        o[0] = r0;
        o[1] = r1;
        o[2] = r2;
        o[3] = r3;
        o[4] = r4;
        reinterpret_cast<uint4*>(dst[c])[5] = make_uint4(materialIndex, attr.ids.x, attr.ids.y, fl[c]);
    }
}

// PbrClearReconstructionInputsV2 (pbr_reconstruction_export_v2.hlsli:34-42)
// for the surface records: the ReSTIR copy, and the G-buffer's primary and
// motion records (its signal is written by Resolve).
__device__ __forceinline__ void clear_surface(SurfaceV2* out, GBufferV2* gb) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (out != nullptr) {
        float4* o = reinterpret_cast<float4*>(out);
#pragma unroll
        for (int k = 0; k < 6; ++k) o[k] = z;
    }
    if (gb != nullptr) {
        float4* o = reinterpret_cast<float4*>(gb);
#pragma unroll
        for (int k = 0; k < 9; ++k) o[k] = z;
    }
}

// ShadeCS (integrators/wavefront/shaders/Shade.hlsl:6-479) with
// WfPbrBuildDirectShadow (PbrWavefrontShading.hlsli:173-292) and the
// NextBounceCS state update (NextBounce.hlsl:3-115) for one queue entry.
// kBatch: the path's job (path state meta.y) supplies the frame constants
// and shading tables, meta.z its bounce; *scene receives the job's scene.
// kExt: sphere lights and lat-long environment maps are compiled in.
template <bool kBatch, bool kExt>
__device__ __forceinline__ void shade_one(const ShadeParams& P, uint32_t i, bool& wantNext, F4& nO, F4& nD,
                                          uint32_t& path, bool& wantShadow, ShadowItem& sItem, uint32_t& scene) {
    path = P.rayPath[i];
    const F4 ro = P.rays[2u * i];
    const F4 rd = P.rays[2u * i + 1u];
    PathState& ps = P.paths[path];
    St st = load_state(ps);
    const uint4 meta = kBatch ? reinterpret_cast<const uint4*>(&ps.meta)[0] : make_uint4(0u, 0u, 0u, 0u);
    const BatchJob* J = kBatch ? P.jobs + meta.y : nullptr;
    const FrameConst f = kBatch ? J->f : P.f;
    const ShadingView sh = kBatch ? J->sh : P.sh;
    const uint32_t bounce = kBatch ? meta.z : P.bounce;
    const uint32_t lpath = kBatch ? meta.w : path;
    scene = kBatch ? J->scene : 0u;
    const uint32_t pixel = lpath / f.spp;
    const uint32_t sampleIndex = f.sampleIndex + (lpath - pixel * f.spp);
    const V3 rayO = v3(ro), rayD = v3(rd);
    const float rayTmin = ro.w, rayTmax = rd.w;
    // WfClearPrimarySurfaceV2 (Shade.hlsl:62-67): misses and errors leave
    // the pixel's surface invalid.
    // With several samples per pixel the first sample's path publishes.
    const bool primary = !kBatch && bounce == 0u && lpath == pixel * f.spp;
    SurfaceV2* surface = primary && P.surfaces != nullptr ? &P.surfaces[pixel] : nullptr;
    GBufferV2* gbuf = primary && P.gbuffer != nullptr ? &P.gbuffer[pixel] : nullptr;
    clear_surface(surface, gbuf);

    if ((st.flags & kPathError) != 0u) {
        // TraceShadowCS flagged this path after its ray was queued; the
        // reference never re-activates it (NextBounce.hlsl:76-81).
        st.flags = (st.flags & ~kPathActive) | kPathTerminated;
        ps.meta.x = st.flags;
        return;
    }
    const QueueHit qh = P.hits[i];
    if (qh.tri == kHitMiss) {
        const V3 env = environment_radiance<kExt>(sh, rayD);
        float mis = 1.0f;
        if (bounce != 0u && (st.flags & kPathPrevDelta) == 0u && f.estimator != 0u) {
            const bool owned = f.estimator == kEstRestir && bounce == 1u && (st.flags & kPathPrevRestir) != 0u;
            if (owned) {
                mis = 0.0f;
            } else {
                const float lp = selected_light_pdf<kExt>(sh, sh.environmentLight, st.prevPos, rayD, 0.0f, zero3());
                if (lp > 0.0f) mis = power_heuristic(st.prevPdf, lp);
            }
        }
        const bool ok = fin3(env) && fin(mis) && accumulate_terminal(ps, st, bounce, env, mis);
        if (!ok) st.flags |= kPathError;
        st.flags = (st.flags & ~kPathActive) | kPathTerminated;
        store_state(ps, st);
        return;
    }
    if (qh.tri == kHitInvalid) {
        terminate(st, true);
        store_state(ps, st);
        return;
    }
    const uint32_t tri = qh.tri & 0x7fffffffu;
    const bool front = (qh.tri & kHitFrontBit) != 0u;
    float gnA[3], snA[3];
    TriAttr attr;
    hit_normals(P.s, tri, qh.u, qh.v, front, gnA, snA, attr);
    const V3 pos = v3(rayO.x + rayD.x * qh.t, rayO.y + rayD.y * qh.t, rayO.z + rayD.z * qh.t);
    const V3 hgn = v3(gnA[0], gnA[1], gnA[2]);
    const V3 hsn = v3(snA[0], snA[1], snA[2]);
    const uint32_t materialIndex = attr.ids.w;
    const uint32_t emitterLight = sh.emitterOfTri[tri];
    {
        const float g2 = dot(hgn, hgn), s2 = dot(hsn, hsn);
        const bool valid = fin(qh.t) && qh.t >= rayTmin && qh.t <= rayTmax && fin3(pos) && fin(g2) && g2 > 0.0f &&
                           fin(s2) && s2 > 0.0f;
        if (!valid || materialIndex >= sh.materialCount) {
            terminate(st, true);
            store_state(ps, st);
            return;
        }
    }
    const Material mat = sh.materials[materialIndex];
    {
        const V3 em = v3(mat.emissiveRoughness);
        const F4 a = mat.attenuationColorDistance;
        bool ok = fin3(em) && !any_neg(em) && fin(a.x) && fin(a.y) && fin(a.z) && fin(a.w) && a.w >= 0.0f;
        if (ok && a.w > 0.0f && !(a.x > 0.0f && a.y > 0.0f && a.z > 0.0f && a.x <= 1.0f && a.y <= 1.0f && a.z <= 1.0f))
            ok = false;
        if (!ok) {
            terminate(st, true);
            store_state(ps, st);
            return;
        }
    }
    const BsdfParams params = material_params(mat);
    const bool currentRestirOwned = bounce == 0u && supports_restir(params);
    if (surface != nullptr || gbuf != nullptr)
        publish_surface(f, pos, hgn, hsn, attr, materialIndex, front, mat, currentRestirOwned, surface, gbuf);
    V3 beta = st.beta, betaD = st.betaD, betaS = st.betaS;
    // Interior (Beer-Lambert) attenuation of a back-face hit on a
    // transmissive material (Shade.hlsl:187-205,
    // PbrInteriorAttenuationFactorL6 pbr_material_bsdf.hlsli:77-96); plain
    // scenes hold no transmissive material.
    if (kExt && !front && mat.transmissionIor.x > 0.0f && mat.attenuationColorDistance.w > 0.0f) {
        const F4 a = mat.attenuationColorDistance;
        bool ok = a.x > 0.0f && a.y > 0.0f && a.z > 0.0f && a.x <= 1.0f && a.y <= 1.0f && a.z <= 1.0f && fin(qh.t) &&
                  qh.t >= 0.0f;
        if (ok) {
            const float e = qh.t / a.w;
            const V3 att = v3(powf(a.x, e), powf(a.y, e), powf(a.z, e));
            ok = fin3(att);
            beta = beta * att;
            betaD = betaD * att;
            betaS = betaS * att;
            ok = ok && fin3(beta) && fin3(betaD) && fin3(betaS);
        }
        if (!ok) {
            st.beta = st.betaD = st.betaS = zero3();
            terminate(st, true);
            store_state(ps, st);
            return;
        }
        st.beta = beta;
        st.betaD = betaD;
        st.betaS = betaS;
    }

    // Emitter MIS (Shade.hlsl:210-258).
    const V3 emission = v3(mat.emissiveRoughness);
    bool emitterVisible = front;
    if (!emitterVisible) {
        emitterVisible = emitterLight < sh.lightCount && (sh.lights[emitterLight].identity.y & kLightTwoSided) != 0u;
    }
    if (emitterVisible && maxc(emission) > 0.0f) {
        float mis = 1.0f;
        if (bounce != 0u && (st.flags & kPathPrevDelta) == 0u && f.estimator != 0u) {
            const bool owned = f.estimator == kEstRestir && bounce == 1u && (st.flags & kPathPrevRestir) != 0u;
            if (owned) {
                mis = 0.0f;
            } else {
                const float lp = selected_light_pdf<kExt>(sh, emitterLight, st.prevPos, rayD, qh.t, hgn);
                if (lp > 0.0f) mis = power_heuristic(st.prevPdf, lp);
            }
        }
        if (!accumulate_terminal(ps, st, bounce, emission, mis)) {
            terminate(st, true);
            store_state(ps, st);
            return;
        }
    }
    if ((mat.metadata.w & kMaterialPureEmitter) != 0u || !fin3(beta) || any_neg(beta)) {
        terminate(st, false);
        store_state(ps, st);
        return;
    }
    BsdfCtx ctx;
    ctx.gn = hgn;
    ctx.sn = hsn;
    V3 tb;
    build_basis(hsn, ctx.tangent, tb);
    ctx.etaI = front ? 1.0f : mat.transmissionIor.y;
    ctx.etaT = front ? mat.transmissionIor.y : 1.0f;
    if (!validate_params<kExt>(ctx, params)) {
        terminate(st, true);
        store_state(ps, st);
        return;
    }
    const V3 wo = -rayD;
    const uint32_t tag = f.streamTag;
    (void)tag;

    // WfPbrBuildDirectShadow.
    if (f.estimator != 0u && sh.aliasCount != 0u && !(f.estimator == kEstRestir && bounce == 0u && currentRestirOwned)) {
        const float r0 = sample_dim(pixel, sampleIndex, bounce_dim(bounce, 0u), f);
        const float r1 = sample_dim(pixel, sampleIndex, bounce_dim(bounce, 1u), f);
        const float r2 = sample_dim(pixel, sampleIndex, bounce_dim(bounce, 2u), f);
        const LightSample ls = sample_one_light<kExt>(sh, pos, r0, r1, r2, [&](uint32_t k) {
            return sample_dim(pixel, sampleIndex, bounce_dim(bounce, 1u + k), f);
        });
        if ((ls.flags & kSampleValid) != 0u) {
            const BsdfEval ev = eval_bsdf<kExt>(ctx, params, wo, ls.wi);
            if (ev.valid && fin(ev.pdf) && fin3(ev.value) && !(ev.pdf < 0.0f)) {
                const bool delta = (ls.flags & kSampleDelta) != 0u;
                const float epdf = delta ? ls.selectionPmf : ls.combinedPdfW;
                if (fin(epdf) && epdf > 0.0f) {
                    const float cosine = fabsf(dot(hgn, ls.wi));
                    if (cosine > 0.0f) {
                        const float mis = delta ? 1.0f : power_heuristic(ls.combinedPdfW, ev.pdf);
                        const V3 scale = ls.Li * (cosine * mis / epdf);
                        const V3 diff = bounce == 0u ? beta * ev.diffuse * scale : betaD * ev.value * scale;
                        const V3 spec = bounce == 0u ? beta * ev.specular * scale : betaS * ev.value * scale;
                        if (fin3(diff) && fin3(spec) && !any_neg(diff) && !any_neg(spec)) {
                            const float smax = (ls.flags & kSampleInfinite) != 0u ? 1.0e30f : ls.distance - f.rayEpsilon;
                            if (smax > f.rayEpsilon) {
                                const V3 so = offset_origin(pos, hgn, ls.wi, f.rayEpsilon);
                                wantShadow = true;
                                sItem.originTmin = f4(so, f.rayEpsilon);
                                sItem.dirTmax = f4(ls.wi, smax);
                                sItem.diffPath = f4(diff, u2f(path | (bounce == 0u ? 0x80000000u : 0u)));
                                sItem.specPrim = f4(spec, u2f(ls.primitiveId));
                            }
                        }
                    }
                }
            }
        }
    }

    // BSDF sample + Russian roulette (Shade.hlsl:307-431).
    bool nextValid = false;
    St next = st;
    V3 nextO = zero3(), nextDir = zero3();
    if (bounce + 1u < f.maxBounce) {
        const float u0 = sample_dim(pixel, sampleIndex, bounce_dim(bounce, 5u), f);
        const float u1 = sample_dim(pixel, sampleIndex, bounce_dim(bounce, 6u), f);
        const float u2 = sample_dim(pixel, sampleIndex, bounce_dim(bounce, 7u), f);
        const BsdfSample bs = sample_bsdf<kExt>(ctx, params, wo, u0, u1, u2);
        bool valid = bs.valid && fin(bs.pdf) && bs.pdf > 0.0f && fin3(bs.value) && fin3(bs.dir) && !any_neg(bs.value);
        V3 nb = zero3(), nd = zero3(), ns = zero3();
        float etaScale = st.eta;
        if (valid) {
            const float cosine = fabsf(dot(hgn, bs.dir));
            const V3 factor = bs.value * (cosine / bs.pdf);
            if (bounce == 0u) {
                nd = beta * bs.diffuse * (cosine / bs.pdf);
                ns = beta * bs.specular * (cosine / bs.pdf);
            } else {
                nd = betaD * factor;
                ns = betaS * factor;
            }
            nb = nd + ns;
            valid = fin3(nb) && fin3(nd) && fin3(ns) && !any_neg(nb);
            if (kExt && (bs.lobeFlags & kLobeTransmission) != 0u) {
                etaScale *= bs.eta * bs.eta;
                valid = valid && fin(etaScale) && etaScale > 0.0f;
            }
        }
        const uint32_t nextBounce = bounce + 1u;
        const bool rrValid = fin(f.rrStart) && fin(f.rrMinimum) && fin(f.rrMaximum) && f.rrStart >= 0.0f &&
                             floorf(f.rrStart) == f.rrStart && f.rrMinimum > 0.0f && f.rrMinimum <= f.rrMaximum &&
                             f.rrMaximum <= 1.0f;
        if (!rrValid) {
            valid = false;
        } else if (valid && (float)nextBounce >= f.rrStart) {
            const float mag = maxc(nb * etaScale);
            if (!(mag > 0.0f) || !fin(mag)) {
                valid = false;
            } else {
                const float cont = fminf(fmaxf(mag, f.rrMinimum), f.rrMaximum);
                const float rr = sample_dim(pixel, sampleIndex, bounce_dim(bounce, 9u), f);
                if (rr >= cont) {
                    valid = false;
                } else {
                    nb = nb / cont;
                    nd = nd / cont;
                    ns = ns / cont;
                }
            }
        }
        if (valid) {
            uint32_t ev = 0u;
            if (bs.delta) ev |= kPathPrevDelta;
            if ((bs.lobeFlags & kLobeDiffuse) == 0u) ev |= kPathPrevSpecular;
            if (currentRestirOwned) ev |= kPathPrevRestir;
            nextValid = true;
            nextO = offset_origin(pos, hgn, bs.dir, f.rayEpsilon);
            nextDir = normalize(bs.dir);
            next.beta = nb;
            next.eta = etaScale;
            next.betaD = nd;
            next.betaS = ns;
            next.prevPos = pos;
            next.prevPdf = bs.pdf;
            next.prevGn = hgn;
            next.prevSn = hsn;
            next.flags = ev;
        } else if (!fin3(nb)) {
            st.flags |= kPathError;
        }
    }

    if (!nextValid) {
        st.flags = (st.flags & ~kPathActive) | kPathTerminated;
        store_state(ps, st);
        return;
    }
    // NextBounceCS candidate validation.
    {
        const float dl2 = dot(nextDir, nextDir);
        const float gl2 = dot(next.prevGn, next.prevGn);
        const float sl2 = dot(next.prevSn, next.prevSn);
        const bool ok = fin3(nextO) && f.rayEpsilon > 0.0f && fin3(nextDir) && 1.0e30f > f.rayEpsilon && fin(dl2) &&
                        dl2 > 0.0f && fin3(next.beta) && !any_neg(next.beta) && fin(next.eta) && next.eta > 0.0f &&
                        fin3(next.betaD) && !any_neg(next.betaD) && fin3(next.betaS) && !any_neg(next.betaS) &&
                        fin3(next.prevPos) && next.prevPdf > 0.0f && fin(next.prevPdf) && fin3(next.prevGn) &&
                        fin(gl2) && gl2 > 0.0f && fin3(next.prevSn) && fin(sl2) && sl2 > 0.0f;
        if (!ok) {
            terminate(st, true);
            store_state(ps, st);
            return;
        }
    }
    const uint32_t carried = next.flags & (kPathPrevDelta | kPathPrevSpecular | kPathPrevRestir);
    st.beta = next.beta;
    st.eta = next.eta;
    st.betaD = next.betaD;
    st.betaS = next.betaS;
    st.prevPos = next.prevPos;
    st.prevPdf = next.prevPdf;
    st.prevGn = next.prevGn;
    st.prevSn = next.prevSn;
    st.flags = (st.flags & kPathError) | kPathActive | carried;
    store_state(ps, st);
    if (kBatch) ps.meta.z = bounce + 1u;
    wantNext = true;
    nO = f4(nextO, f.rayEpsilon);
    nD = f4(nextDir, 1.0e30f);
}

template <bool kBatch, bool kExt>
#ifndef REX_SHADE_WAVES
#define REX_SHADE_WAVES 1
#endif
__global__ __launch_bounds__(kShadeBlock, REX_SHADE_WAVES) void k_shade(ShadeParams P) {
    __shared__ AppendLds<3, kShadeBlock> lds;
    const uint32_t count = *P.count;
    const uint32_t stride = gridDim.x * kShadeBlock;
    const uint32_t rounded = (count + kShadeBlock - 1u) / kShadeBlock * kShadeBlock;
    uint32_t* const counters[3] = {P.nextCount, P.shadowCount, kBatch ? P.retireCount : P.nextCount};
    uint32_t parity = 0;
    for (uint32_t base = blockIdx.x * kShadeBlock; base < rounded; base += stride, parity ^= 1u) {
        const uint32_t i = base + threadIdx.x;
        bool wantNext = false, wantShadow = false;
        F4 nO, nD;
        uint32_t path = 0, scene = 0;
        ShadowItem sItem;
        if (i < count) shade_one<kBatch, kExt>(P, i, wantNext, nO, nD, path, wantShadow, sItem, scene);
        if (!kBatch && P.denseFlags != nullptr) {
            // PrefixScan mode (grid-uniform): the records stay at index i and
            // the stable compaction moves them into the queues.
            if (i < count) {
                P.denseFlags[i] = make_uint2(wantNext ? 1u : 0u, wantShadow ? 1u : 0u);
                if (wantNext) {
                    P.denseRays[2u * i] = nO;
                    P.denseRays[2u * i + 1u] = nD;
                    P.densePath[i] = path;
                }
                if (wantShadow) P.denseShadows[i] = sItem;
            }
            continue;
        }
        // A batched path that does not continue is finalized next step,
        // after its last shadow query.
        const bool want[3] = {wantNext, wantShadow, kBatch && i < count && !wantNext};
        uint32_t at[3];
        block_append<3, kShadeBlock>(lds, parity, want, counters, at);
        if (wantNext) {
            P.nextRays[2u * at[0]] = nO;
            P.nextRays[2u * at[0] + 1u] = nD;
            P.nextPath[at[0]] = path;
            if (kBatch) P.nextSceneOf[at[0]] = scene;
        }
        if (wantShadow) {
            P.shadows[at[1]] = sItem;
            if (kBatch) P.shadowSceneOf[at[1]] = scene;
        }
        if (kBatch && want[2]) P.retire[at[2]] = path;
    }
}

// Finalize of pooled paths retired by the previous step's Shade: the
// ResolveCS sample value (non-finite signals -> 0, error flag) stored at the
// path's place in its job's sample array, retired slot i written at
// freeList[top + i] (k_batch_reset raises the top afterwards), the job's
// retired count advanced by one atomic per job per wave on the block's
// stripe.
__global__ __launch_bounds__(256) void k_batch_finalize(BatchDevice D, const uint32_t* __restrict__ retire,
                                                       const uint32_t* __restrict__ retireCount) {
    const uint32_t n = *retireCount;
    const uint32_t top = D.ctl[kCtlFree];
    const uint32_t stride = gridDim.x * 256u;
    const uint32_t rounded = (n + 255u) / 256u * 256u;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* const done = D.doneCount + (blockIdx.x & (kDoneStripes - 1u)) * D.jobStride;
    for (uint32_t base = blockIdx.x * 256u; base < rounded; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool live = i < n;
        uint32_t job = kInvalid;
        if (live) {
            const uint32_t slot = retire[i];
            PathState& ps = D.paths[slot];
            const uint4 meta = reinterpret_cast<const uint4*>(&ps.meta)[0];
            job = meta.y;
            V3 sig[5] = {v3(ps.cam), v3(ps.dd), v3(ps.ds), v3(ps.id), v3(ps.is)};
            bool ok = true;
            for (int k = 0; k < 5; ++k) ok = ok && fin_nonneg3(sig[k]);
            if (!ok) {
                for (int k = 0; k < 5; ++k) sig[k] = zero3();
                ps.meta.x |= kPathError;
            }
            const V3 diff = sig[1] + sig[3];
            const V3 spec = sig[2] + sig[4];
            const V3 out = diff + spec + sig[0];
            D.jobs[job].samples[meta.w] = {out.x, out.y, out.z, 0.0f};
            D.freeList[top + i] = slot;
        }
        uint64_t pending = __ballot(live);
        while (pending != 0u) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const uint32_t lj = __shfl(job, leader, 64);
            const uint64_t same = __ballot(live && job == lj);
            if ((int)lane == leader) atomicAdd(done + lj, (uint32_t)__popcll(same));
            pending &= ~same;
        }
    }
}

__global__ void k_batch_reset(uint32_t* ctl, uint32_t* retireCount, uint32_t* rayCount, uint32_t* shadowCount,
                              const uint32_t* tracedRays) {
    if (threadIdx.x != 0u) return;
    const uint64_t rays = ((uint64_t)ctl[kCtlRaysLo] | ((uint64_t)ctl[kCtlRaysHi] << 32)) + *tracedRays + *shadowCount;
    ctl[kCtlRaysLo] = (uint32_t)rays;
    ctl[kCtlRaysHi] = (uint32_t)(rays >> 32);
    ctl[kCtlFree] += *retireCount;
    *retireCount = 0u;
    *rayCount = 0u;
    *shadowCount = 0u;
}

// RayGenCS for new pooled paths: thread t takes free slot
// freeList[top - 1 - t] and global path cursor + t, which belongs to
// generation entry k with genStart[k] <= cursor + t < genStart[k + 1]; its
// primary ray goes to queue position rayCount + t.
__global__ __launch_bounds__(256) void k_batch_regen(BatchDevice D, F4* __restrict__ rays,
                                                    uint32_t* __restrict__ rayPath, uint32_t* __restrict__ raySceneOf,
                                                    const uint32_t* __restrict__ rayCount) {
    const uint32_t freeSlots = D.ctl[kCtlFree];
    const uint32_t gens = D.ctl[kCtlGenCount];
    const uint64_t cursor = (uint64_t)D.ctl[kCtlCursorLo] | ((uint64_t)D.ctl[kCtlCursorHi] << 32);
    const uint64_t total = (uint64_t)D.ctl[kCtlGenTotalLo] | ((uint64_t)D.ctl[kCtlGenTotalHi] << 32);
    const uint64_t left = total > cursor ? total - cursor : 0u;
    const uint32_t reserve = D.ctl[kCtlReserve];
    const uint32_t usable = freeSlots > reserve ? freeSlots - reserve : 0u;
    const uint32_t m = left < (uint64_t)usable ? (uint32_t)left : usable;
    const uint32_t q0 = *rayCount;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < m; t += gridDim.x * 256u) {
        const uint64_t g = cursor + t;
        uint32_t lo = 0, hi = gens;  // last k with genStart[k] <= g
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (D.genStart[mid] <= g) lo = mid;
            else hi = mid;
        }
        const uint32_t j = D.genJob[lo];
        const BatchJob& J = D.jobs[j];
        const FrameConst f = J.f;
        const uint32_t lp = (uint32_t)(g - D.genStart[lo]);
        const uint32_t slot = D.freeList[freeSlots - 1u - t];
        const uint32_t pixel = lp / f.spp;
        const uint32_t sampleIndex = f.sampleIndex + (lp - pixel * f.spp);
        const uint32_t px = pixel % f.width, py = pixel / f.width;
        const float jx = sample_dim(pixel, sampleIndex, 0u, f);
        const float jy = sample_dim(pixel, sampleIndex, 1u, f);
        float sx = ((float)px + jx) / (float)f.width;
        float sy = ((float)py + jy) / (float)f.height;
        sx = sx * 2.0f - 1.0f;
        sy = sy * 2.0f - 1.0f;
        sy = -sy;
        const V3 fwd = v3(f.camFwd[0], f.camFwd[1], f.camFwd[2]);
        const V3 right = v3(f.camRight[0], f.camRight[1], f.camRight[2]);
        const V3 up = v3(f.camUp[0], f.camUp[1], f.camUp[2]);
        const V3 dir = normalize(fwd + right * (sx * f.aspect * f.tanHalfFov) + up * (sy * f.tanHalfFov));
        PathState st;
        st.betaEta = {1.0f, 1.0f, 1.0f, 1.0f};
        st.betaD = st.betaS = st.cam = st.dd = st.ds = st.id = st.is = {0.0f, 0.0f, 0.0f, 0.0f};
        st.prevPosPdf = st.prevGn = st.prevSn = {0.0f, 0.0f, 0.0f, 0.0f};
        st.meta = {kPathActive | kPathPrevDelta, j, 0u, lp};
        D.paths[slot] = st;
        const uint32_t rs = q0 + t;
        rays[2u * rs] = {f.camPos[0], f.camPos[1], f.camPos[2], f.rayEpsilon};
        rays[2u * rs + 1u] = {dir.x, dir.y, dir.z, 1.0e30f};
        rayPath[rs] = slot;
        raySceneOf[rs] = J.scene;
    }
}

// Advances the pool after k_batch_regen: the same m, from the same words.
__global__ void k_batch_commit(uint32_t* ctl, uint32_t* rayCount) {
    if (threadIdx.x != 0u) return;
    const uint32_t freeSlots = ctl[kCtlFree];
    const uint64_t cursor = (uint64_t)ctl[kCtlCursorLo] | ((uint64_t)ctl[kCtlCursorHi] << 32);
    const uint64_t total = (uint64_t)ctl[kCtlGenTotalLo] | ((uint64_t)ctl[kCtlGenTotalHi] << 32);
    const uint64_t left = total > cursor ? total - cursor : 0u;
    const uint32_t reserve = ctl[kCtlReserve];
    const uint32_t usable = freeSlots > reserve ? freeSlots - reserve : 0u;
    const uint32_t m = left < (uint64_t)usable ? (uint32_t)left : usable;
    const uint64_t next = cursor + m;
    ctl[kCtlFree] = freeSlots - m;
    ctl[kCtlCursorLo] = (uint32_t)next;
    ctl[kCtlCursorHi] = (uint32_t)(next >> 32);
    *rayCount += m;
}

// Admission of n jobs: record k into its slot, that slot's count stripes
// cleared.
__global__ void k_batch_admit(BatchDevice D, BatchJob* __restrict__ jobs, const BatchJob* __restrict__ staged,
                              const uint32_t* __restrict__ slots, uint32_t n) {
    const uint32_t k = blockIdx.x;
    if (k >= n) return;
    const uint32_t slot = slots[k];
    constexpr uint32_t kWords = sizeof(BatchJob) / 4u;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(staged + k);
    uint32_t* dst = reinterpret_cast<uint32_t*>(jobs + slot);
    for (uint32_t w = threadIdx.x; w < kWords; w += blockDim.x) dst[w] = src[w];
    if (threadIdx.x < kDoneStripes) D.doneCount[threadIdx.x * D.jobStride + slot] = 0u;
}

// ResolveCS + Film of finished jobs: thread t of the resolve list's pixel
// prefix is pixel t - resolveStart[k] of job resolveJob[k]; its samples are
// folded into the running mean in sample order.
__global__ __launch_bounds__(256) void k_batch_resolve(BatchDevice D) {
    const uint32_t n = D.ctl[kCtlResolveCount];
    const uint32_t total = n ? D.resolveStart[n] : 0u;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < total; t += gridDim.x * 256u) {
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (D.resolveStart[mid] <= t) lo = mid;
            else hi = mid;
        }
        const BatchJob& J = D.jobs[D.resolveJob[lo]];
        const uint32_t pixel = t - D.resolveStart[lo];
        const uint32_t spp = J.f.spp;
        F4& fp = J.film[pixel];
        float4 acc = make_float4(fp.x, fp.y, fp.z, fp.w);
        for (uint32_t s = 0; s < spp; ++s) {
            const F4 o = J.samples[pixel * spp + s];
            const V3 out = v3(o.x, o.y, o.z);
            const uint32_t cnt = (uint32_t)acc.w;
            const V3 prev = cnt == 0u ? zero3() : v3(acc.x, acc.y, acc.z);
            const V3 v = prev + (out - prev) / (float)(cnt + 1u);
            acc = make_float4(v.x, v.y, v.z, (float)(cnt + 1u));
        }
        fp = {acc.x, acc.y, acc.z, acc.w};
    }
}

// Motion input of a published G-buffer surface (surface_motion_input).
__device__ __forceinline__ MotionInputV2 motion_input(const ExportV2& E, const SurfaceV2& sp) {
    return surface_motion_input(E.instances, E.instanceMap, E.instanceCount, E.mapCount, sp);
}

// ResolveCS (Resolve.hlsl:4-90) and the progressive-mean Film of ComposeCS
// (reconstruction/shaders/ComposeCS.hlsl:35-47), applied to the pixel's
// samples in order.  signals (optional): five float4 per path.  E: the
// ABI-v2 exports of PbrPublishReconstructionSignalV2
// (pbr_reconstruction_export_v2.hlsli:132-149), the signals being the mean
// of the pixel's samples, plus the motion inputs and the ReSTIR surface copy.
__global__ __launch_bounds__(256) void k_resolve_film(FrameConst f, PathState* __restrict__ paths,
                                                      float4* __restrict__ film, float4* __restrict__ signals,
                                                      const float4* __restrict__ direct, ExportV2 E) {
    const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= f.width * f.height) return;
    const bool exportSignal = E.raw != nullptr || E.gbuffer != nullptr || E.terminal != nullptr;
    V3 sum[5] = {zero3(), zero3(), zero3(), zero3(), zero3()};
    float4 acc = film[pixel];
    for (uint32_t s = 0; s < f.spp; ++s) {
        const uint32_t p = pixel * f.spp + s;
        PathState& ps = paths[p];
        V3 sig[5] = {v3(ps.cam), v3(ps.dd), v3(ps.ds), v3(ps.id), v3(ps.is)};
        bool ok = true;
        for (int k = 0; k < 5; ++k) ok = ok && fin_nonneg3(sig[k]);
        if (!ok) {
            for (int k = 0; k < 5; ++k) sig[k] = zero3();
            ps.meta.x |= kPathError;
        }
        if (direct != nullptr) {
            // RestirPublishDirectV3 adds ReSTIR's primary direct light to the
            // resolved signal (spp 1: path p is pixel p).
            const float4 dd = direct[2u * p], ds = direct[2u * p + 1u];
            sig[1] = sig[1] + v3(dd.x, dd.y, dd.z);
            sig[2] = sig[2] + v3(ds.x, ds.y, ds.z);
        }
        if (signals != nullptr) {
            for (int k = 0; k < 5; ++k) signals[5u * p + k] = make_float4(sig[k].x, sig[k].y, sig[k].z, 0.0f);
        }
        if (exportSignal) {
            for (int k = 0; k < 5; ++k) sum[k] = sum[k] + sig[k];
        }
        const V3 diff = sig[1] + sig[3];
        const V3 spec = sig[2] + sig[4];
        const V3 out = diff + spec + sig[0];
        const uint32_t n = (uint32_t)acc.w;
        const V3 prev = n == 0u ? zero3() : v3(acc.x, acc.y, acc.z);
        const V3 v = prev + (out - prev) / (float)(n + 1u);
        acc = make_float4(v.x, v.y, v.z, (float)(n + 1u));
    }
    film[pixel] = acc;
    if (exportSignal) {
        const float spp = (float)f.spp;
        V3 m[5];
        for (int k = 0; k < 5; ++k) m[k] = v3(sum[k].x / spp, sum[k].y / spp, sum[k].z / spp);
        SignalV2 sg;
        sg.dd = {m[1].x, m[1].y, m[1].z, 0.0f};
        sg.ds = {m[2].x, m[2].y, m[2].z, 0.0f};
        sg.id = {m[3].x, m[3].y, m[3].z, 0.0f};
        sg.is = {m[4].x, m[4].y, m[4].z, 0.0f};
        if (E.terminal != nullptr) E.terminal[pixel] = make_float4(m[0].x, m[0].y, m[0].z, 0.0f);
        if (E.raw != nullptr) E.raw[pixel] = sg;
        if (E.gbuffer != nullptr) E.gbuffer[pixel].signal = sg;
    }
    if (E.primary != nullptr && E.restirSurfaces != nullptr) E.primary[pixel] = E.restirSurfaces[pixel];
    if (E.motion != nullptr) E.motion[pixel] = motion_input(E, E.gbuffer[pixel].primary);
}

// ---- host launchers ----------------------------------------------------------

hipError_t launch_frame_reset(uint32_t* counts, uint32_t n, uint32_t pathCount, hipStream_t stream) {
    hipLaunchKernelGGL(k_frame_reset, dim3((n + 63) / 64), dim3(64), 0, stream, counts, n, pathCount);
    return hipGetLastError();
}

hipError_t launch_raygen(const FrameConst& f, PathState* paths, F4* rays, uint32_t* rayPath, hipStream_t stream) {
    hipLaunchKernelGGL(k_raygen, dim3((f.pathCount + 255) / 256), dim3(256), 0, stream, f, paths, rays, rayPath);
    return hipGetLastError();
}

hipError_t launch_shade(const ShadeLaunch& L, hipStream_t stream) {
    ShadeParams P;
    P.f = L.f;
    P.sh = L.sh;
    P.s = L.s;
    P.bounce = L.bounce;
    P.rays = L.rays;
    P.rayPath = L.rayPath;
    P.count = L.count;
    P.hits = L.hits;
    P.paths = L.paths;
    P.nextRays = L.nextRays;
    P.nextPath = L.nextPath;
    P.nextCount = L.nextCount;
    P.shadows = L.shadows;
    P.shadowCount = L.shadowCount;
    P.surfaces = L.surfaces;
    P.gbuffer = L.gbuffer;
    P.jobs = L.jobs;
    P.nextSceneOf = L.nextSceneOf;
    P.shadowSceneOf = L.shadowSceneOf;
    P.retire = L.retire;
    P.retireCount = L.retireCount;
    P.denseFlags = L.denseFlags;
    P.denseRays = L.denseRays;
    P.densePath = L.densePath;
    P.denseShadows = L.denseShadows;
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t maxBlocks = (L.capacity + kShadeBlock - 1) / kShadeBlock;
    uint32_t blocks = (uint32_t)(cus > 0 ? cus : 1) * (L.blocksPerCu != 0u ? L.blocksPerCu : 16u);
    if (blocks > maxBlocks) blocks = maxBlocks;
    if (blocks == 0) blocks = 1;
    if (L.jobs != nullptr) {
        if (L.extended) hipLaunchKernelGGL((k_shade<true, true>), dim3(blocks), dim3(kShadeBlock), 0, stream, P);
        else hipLaunchKernelGGL((k_shade<true, false>), dim3(blocks), dim3(kShadeBlock), 0, stream, P);
    } else {
        if (L.extended) hipLaunchKernelGGL((k_shade<false, true>), dim3(blocks), dim3(kShadeBlock), 0, stream, P);
        else hipLaunchKernelGGL((k_shade<false, false>), dim3(blocks), dim3(kShadeBlock), 0, stream, P);
    }
    return hipGetLastError();
}

hipError_t launch_resolve_film(const FrameConst& f, PathState* paths, float4* film, float4* signals,
                               const float4* direct, const ExportV2& E, hipStream_t stream) {
    const uint32_t pixels = f.width * f.height;
    hipLaunchKernelGGL(k_resolve_film, dim3((pixels + 255) / 256), dim3(256), 0, stream, f, paths, film, signals,
                       direct, E);
    return hipGetLastError();
}

hipError_t launch_batch_finalize(const BatchDevice& d, const uint32_t* retire, const uint32_t* retireCount,
                                 uint32_t blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_finalize, dim3(blocks), dim3(256), 0, stream, d, retire, retireCount);
    return hipGetLastError();
}

hipError_t launch_batch_reset(uint32_t* ctl, uint32_t* retireCount, uint32_t* rayCount, uint32_t* shadowCount,
                              const uint32_t* tracedRays, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, stream, ctl, retireCount, rayCount, shadowCount,
                       tracedRays);
    return hipGetLastError();
}

hipError_t launch_batch_regen(const BatchDevice& d, F4* rays, uint32_t* rayPath, uint32_t* raySceneOf,
                              uint32_t* rayCount, uint32_t blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_regen, dim3(blocks), dim3(256), 0, stream, d, rays, rayPath, raySceneOf, rayCount);
    hipLaunchKernelGGL(k_batch_commit, dim3(1), dim3(64), 0, stream, d.ctl, rayCount);
    return hipGetLastError();
}

hipError_t launch_batch_admit(const BatchDevice& d, BatchJob* jobs, const BatchJob* staged, const uint32_t* slots,
                              uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_admit, dim3(n), dim3(64), 0, stream, d, jobs, staged, slots, n);
    return hipGetLastError();
}

hipError_t launch_batch_resolve(const BatchDevice& d, uint32_t blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_resolve, dim3(blocks), dim3(256), 0, stream, d);
    return hipGetLastError();
}

}  // namespace rex
