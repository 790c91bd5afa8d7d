/*
 * Oracle (test infrastructure only): ReSTIR DI V3 restated per pixel, with
 * the numeric conventions of oracle_shading.h (no FMA contraction,
 * dot = (x*x + y*y) + z*z, normalize = v * (1 / sqrt(dot)), saturate(x) =
 * min(max(x, 0), 1)).
 *
 * Sources restated (paths relative to the RenderingEngine tree):
 *   hash / random, validity, target, RIS update / finalize / M clamp,
 *   surface-pair validation      resources/shaders/restir/RestirProductionV3.hlsli:16-30,
 *                                140-205, 225-482
 *   candidate export             resources/shaders/integrators/pbr_restir_candidate_export_v3.hlsl:13-114
 *   initial RIS                  resources/shaders/restir/RestirInitialV3.comp.hlsl:9-58
 *   temporal reuse               resources/shaders/restir/RestirTemporalV3.comp.hlsl:20-215
 *   motion vectors               reconstruction/shaders/MotionVectorsCS.hlsl (static surfaces,
 *                                zero jitter) with the matrices of
 *                                src/renderers/Wave3DebugRuntime.cpp:140-202
 *   spatial reuse                resources/shaders/restir/RestirSpatialV3.comp.hlsl:12-112 with
 *                                the ring table of Wave3DebugRuntime.cpp:1888-1925
 *   winner visibility            RestirWinnerVisibilityPrepareV3.comp.hlsl:9-36,
 *                                RestirWinnerResolveV3.comp.hlsl:78-122 (Physical)
 *   publish                      RestirPublishDirectV3.comp.hlsl:11-42,
 *                                RestirPublishHistoryV3.comp.hlsl
 *
 * Parity pins (tests/test_oracle_restir.py): the reference's known-answer
 * reservoirs (tests/unit/L9/VulkanSmoke.cpp:1036-1043, 1834-1840), RIS
 * selection frequencies (tests/unit/L9/RestirTestsMain.cpp:109-157 restated),
 * E[W * target] = sum of weights / M, the unshadowed target against the
 * reference's double-precision BSDF, temporal / spatial merge identities,
 * and the motion of a still and of a translated camera.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_restir.h"

typedef struct { float x, y, z; } f3;
static f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f3 ld(const float* p) { return mk(p[0], p[1], p[2]); }
static f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static f3 muls(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static f3 normalize(f3 a) { return muls(a, 1.0f / sqrtf(dot(a, a))); }
static float length(f3 a) { return sqrtf(dot(a, a)); }
static int fin(float a) { return isfinite(a); }
static int fin3(f3 a) { return fin(a.x) && fin(a.y) && fin(a.z); }
static int fin4(const float* a) { return fin(a[0]) && fin(a[1]) && fin(a[2]) && fin(a[3]); }
static f3 max0(f3 a) { return mk(fmaxf(a.x, 0.0f), fmaxf(a.y, 0.0f), fmaxf(a.z, 0.0f)); }
static float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
static void st3(float* d, f3 a) { d[0] = a.x; d[1] = a.y; d[2] = a.z; }

#define INV 0xffffffffu
#define SAMPLE_FLAG_VALID 1u
#define SAMPLE_FLAG_DELTA 2u
#define SAMPLE_FLAG_AREA 4u
#define SAMPLE_FLAG_SOLID_ANGLE 8u
#define HISTORY_VALID 1u
#define HISTORY_RESET 2u
#define REUSE_NONE 0u
#define REUSE_TEMPORAL 1u
#define REUSE_SPATIAL 2u
#define REJECT_NONE 0u
#define REJECT_INVALID 1u
#define REJECT_OUTSIDE 2u
#define REJECT_MOTION 3u
#define REJECT_CAMERA_CUT 4u
#define REJECT_RESIZE 5u
#define REJECT_DEPTH 6u
#define REJECT_NORMAL 7u
#define REJECT_INSTANCE 8u
#define REJECT_THIN 9u
#define REJECT_SCENE 10u
#define REJECT_LIGHT 11u
#define REJECT_AGE 12u
#define REJECT_MATERIAL 13u
#define HIT_MISS 0u
#define HIT_INVALID 3u
#define INV_PI 0.31830988618379067154f

/* ---- primitives ------------------------------------------------------------ */

/* RestirProductionV3.hlsli:16-24 */
uint32_t ora_restir_hash(uint32_t v) {
    v ^= v >> 16u;
    v *= 0x7feb352du;
    v ^= v >> 15u;
    v *= 0x846ca68bu;
    v ^= v >> 16u;
    return v;
}

/* RestirProductionV3.hlsli:26-30 */
float ora_restir_random(uint32_t* state) {
    *state = ora_restir_hash(*state + 0x9e3779b9u);
    return (float)(*state >> 8u) * (1.0f / 16777216.0f);
}

/* RestirProductionV3.hlsli:245-266 */
int ora_restir_surface_valid(const ora_surface_v2* s) {
    const float* d = s->diffuseAlbedo;
    const float* sp = s->specularAlbedo;
    const f3 gn = ld(s->geometricNormalRoughness), sn = ld(s->shadingNormalMetallic);
    return (s->identity[3] & ORA_SURFACE_VALID) != 0u && fin4(s->worldPositionLinearDepth) &&
           fin4(s->geometricNormalRoughness) && fin4(s->shadingNormalMetallic) && fin4(d) && fin4(sp) &&
           d[0] >= 0.0f && d[1] >= 0.0f && d[2] >= 0.0f && d[0] <= 1.0f && d[1] <= 1.0f && d[2] <= 1.0f &&
           sp[0] >= 0.0f && sp[1] >= 0.0f && sp[2] >= 0.0f && sp[0] <= 1.0f && sp[1] <= 1.0f && sp[2] <= 1.0f &&
           s->shadingNormalMetallic[3] >= 0.0f && s->shadingNormalMetallic[3] <= 1.0f &&
           ora_roughness_finite(s->geometricNormalRoughness[3]) && dot(gn, gn) > 1.0e-12f && dot(sn, sn) > 1.0e-12f;
}

/* RestirProductionV3.hlsli:268-279 */
int ora_restir_sample_valid(const ora_light_sample_v3* s, const ora_restir_params* p) {
    const float cp = s->directionCombinedPdf[3], dp = s->radianceDiscretePdf[3];
    return (s->metadata[1] & SAMPLE_FLAG_VALID) != 0u && s->identity[0] != INV && s->generation[0] != 0u &&
           s->generation[1] == p->generations[0] && fin(cp) && cp >= 0.0f && fin(dp) && dp >= 0.0f &&
           fin3(ld(s->directionCombinedPdf)) && fin3(ld(s->radianceDiscretePdf));
}

/* RestirEvaluateUnshadowedV3, RestirProductionV3.hlsli:281-325 */
void ora_restir_evaluate_unshadowed(const ora_light_sample_v3* s, const ora_surface_v2* surf,
                                    const ora_restir_params* p, float diffuse[3], float specular[3], float* target) {
    st3(diffuse, mk(0.0f, 0.0f, 0.0f));
    st3(specular, mk(0.0f, 0.0f, 0.0f));
    *target = 0.0f;
    if (!ora_restir_sample_valid(s, p) || !ora_restir_surface_valid(surf)) return;
    const f3 n = normalize(ld(surf->shadingNormalMetallic));
    const f3 l = normalize(ld(s->directionCombinedPdf));
    const f3 v = normalize(sub(ld(p->cameraPosition), ld(surf->worldPositionLinearDepth)));
    const float noL = sat(dot(n, l));
    const float noV = sat(dot(n, v));
    if (!(noL > 0.0f) || !(noV > 0.0f)) return;
    const f3 h = normalize(add(l, v));
    const float noH = sat(dot(n, h));
    const float voH = sat(dot(v, h));
    const float roughness = surf->geometricNormalRoughness[3];
    const f3 diffAlb = max0(ld(surf->diffuseAlbedo));
    const f3 f0 = max0(ld(surf->specularAlbedo));
    float fa[3];
    const float f0a[3] = {f0.x, f0.y, f0.z};
    ora_schlick3(voH, f0a, fa);
    const f3 f = ld(fa);
    const float alpha = roughness * roughness;
    const float d = ora_ggx_d_cos(noH, alpha);
    const float g = ora_ggx_g2_cos(noV, noL, alpha);
    const f3 radiance = max0(ld(s->radianceDiscretePdf));
    const f3 one_f = sub(mk(1.0f, 1.0f, 1.0f), f);
    const f3 dv = muls(mul(mul(radiance, diffAlb), one_f), noL * INV_PI);
    const f3 sv = muls(mul(radiance, f), d * g / fmaxf(4.0f * noV, 1.0e-6f));
    if (!fin3(dv) || !fin3(sv)) return;
    st3(diffuse, dv);
    st3(specular, sv);
    const f3 sum = max0(add(dv, sv));
    *target = dot(sum, mk(0.2126f, 0.7152f, 0.0722f));
}

/* RestirEvaluateCandidateV3, RestirProductionV3.hlsli:327-350 */
void ora_restir_evaluate_candidate(const ora_light_sample_v3* s, const ora_surface_v2* surf,
                                   const ora_restir_params* p, uint32_t candidateSource, uint32_t reuseSource,
                                   uint32_t sourceSurface, ora_restir_candidate* out) {
    memset(out, 0, sizeof(*out));
    out->sample = *s;
    float dv[3], sv[3], target;
    ora_restir_evaluate_unshadowed(s, surf, p, dv, sv, &target);
    const float proposal = s->directionCombinedPdf[3];
    const int surfaceValid = ora_restir_surface_valid(surf);
    const int support = ora_restir_sample_valid(s, p) && surfaceValid && target > 0.0f && proposal > 0.0f;
    out->tpsc[0] = target;
    out->tpsc[1] = proposal;
    out->tpsc[2] = support ? 1.0f : 0.0f;
    out->tpsc[3] = surfaceValid ? 1.0f : 0.0f;
    out->provenance[0] = candidateSource;
    out->provenance[1] = reuseSource;
    out->provenance[2] = sourceSurface;
    out->provenance[3] = s->metadata[1];
}

/* RestirCandidateValidV3, RestirProductionV3.hlsli:352-363 */
int ora_restir_candidate_valid(const ora_restir_candidate* c, const ora_restir_params* p) {
    const float* t = c->tpsc;
    const int consistent = t[0] > 0.0f ? t[2] == 1.0f : t[2] == 0.0f;
    return ora_restir_sample_valid(&c->sample, p) && fin(t[0]) && t[0] >= 0.0f && fin(t[1]) && t[1] > 0.0f &&
           (t[2] == 0.0f || t[2] == 1.0f) && fin(t[3]) && t[3] >= 0.0f && t[3] > 0.0f && consistent;
}

/* RestirCandidateWeightV3, RestirProductionV3.hlsli:365-370 */
float ora_restir_candidate_weight(const ora_restir_candidate* c, const ora_restir_params* p) {
    if (!ora_restir_candidate_valid(c, p)) return 0.0f;
    return c->tpsc[0] * c->tpsc[2] * c->tpsc[3] / c->tpsc[1];
}

/* RestirEmptyReservoirV3, RestirProductionV3.hlsli:372-378 */
void ora_restir_empty(ora_reservoir_v3* r) {
    memset(r, 0, sizeof(*r));
    r->selected.identity[0] = INV;
    r->selected.identity[1] = INV;
}

/* RestirUpdateReservoirV3, RestirProductionV3.hlsli:380-414 */
void ora_restir_update(ora_reservoir_v3* r, const ora_restir_candidate* c, const ora_restir_params* p, float weight,
                       uint32_t multiplicity, float randomValue) {
    if (!ora_restir_candidate_valid(c, p) || !fin(weight) || weight < 0.0f || multiplicity == 0u) return;
    const uint32_t oldM = r->state[0];
    const uint32_t newM = oldM > 0xffffffffu - multiplicity ? 0xffffffffu : oldM + multiplicity;
    if (weight == 0.0f) {
        r->state[0] = newM;
        return;
    }
    const float newSum = r->weightState[0] + weight;
    if (!(newSum > 0.0f) || !fin(newSum)) return;
    r->weightState[0] = newSum;
    r->state[0] = newM;
    if (randomValue * newSum < weight) {
        r->selected = c->sample;
        r->weightState[2] = c->tpsc[0];
        r->weightState[3] = c->tpsc[1];
        r->selectedTerms[0] = c->tpsc[2];
        r->selectedTerms[1] = c->tpsc[3];
        r->selectedTerms[2] = weight;
        r->selectedTerms[3] = 0.0f;
        memcpy(r->provenance, c->provenance, sizeof(r->provenance));
    }
}

/* RestirFinalizeReservoirV3, RestirProductionV3.hlsli:416-442 */
void ora_restir_finalize(ora_reservoir_v3* r) {
    const float wsum = r->weightState[0], target = r->weightState[2];
    if (r->state[0] == 0u || !(target > 0.0f) || !(wsum > 0.0f) || !fin(target) || !fin(wsum)) {
        r->state[2] &= ~ORA_RES_VALID;
        r->weightState[1] = 0.0f;
        return;
    }
    const float den = (float)r->state[0] * target;
    if (!(den > 0.0f) || !fin(den)) {
        r->state[2] &= ~ORA_RES_VALID;
        r->weightState[1] = 0.0f;
        return;
    }
    r->weightState[1] = wsum / den;
    if (!(r->weightState[1] > 0.0f) || !fin(r->weightState[1])) {
        r->state[2] &= ~ORA_RES_VALID;
        r->weightState[1] = 0.0f;
        return;
    }
    r->selectedTerms[3] = r->weightState[1];
    r->state[2] |= ORA_RES_VALID;
}

/* RestirClampMV3, RestirProductionV3.hlsli:444-452 */
void ora_restir_clamp_m(ora_reservoir_v3* r, uint32_t maxM) {
    if (r->state[0] > maxM) {
        const float scale = (float)maxM / (float)r->state[0];
        r->weightState[0] *= scale;
        r->state[0] = maxM;
        r->state[2] |= ORA_RES_M_CLAMPED;
    }
}

/* RestirValidateSurfacePairV3, RestirProductionV3.hlsli:454-482 */
uint32_t ora_restir_validate_pair(const ora_surface_v2* cur, const ora_surface_v2* src, const ora_restir_params* p) {
    if (!ora_restir_surface_valid(cur) || !ora_restir_surface_valid(src)) return REJECT_INVALID;
    if (cur->identity[0] != src->identity[0]) return REJECT_MATERIAL;
    if (cur->identity[1] != src->identity[1]) return REJECT_INSTANCE;
    const float dc = cur->worldPositionLinearDepth[3], ds = src->worldPositionLinearDepth[3];
    const float scale = fmaxf(1.0f, fmaxf(fabsf(dc), fabsf(ds)));
    if (fabsf(dc - ds) / scale > p->validation[1]) return REJECT_DEPTH;
    if (dot(normalize(ld(cur->shadingNormalMetallic)), normalize(ld(src->shadingNormalMetallic))) < p->validation[0])
        return REJECT_NORMAL;
    const float dist = length(sub(ld(cur->worldPositionLinearDepth), ld(src->worldPositionLinearDepth)));
    if (dist > p->validation[2]) return REJECT_THIN;
    return REJECT_NONE;
}

/* ---- motion ------------------------------------------------------------------ */

/* Multiply (Wave3DebugRuntime.cpp:140-159): value += l[r][i] * r[i][c] from 0. */
static void mat_mul(const float* l, const float* r, float* out) {
    for (int row = 0; row < 4; ++row) {
        for (int col = 0; col < 4; ++col) {
            float v = 0.0f;
            for (int i = 0; i < 4; ++i) v += l[row * 4 + i] * r[i * 4 + col];
            out[row * 4 + col] = v;
        }
    }
}

/* MakeWorldToView, Wave3DebugRuntime.cpp:161-181 (RenderingEngine::Dot = x*x' + y*y' + z*z'). */
void ora_world_to_view(const ora_camera* c, float m[16]) {
    const f3 r = ld(c->right), u = ld(c->up), f = ld(c->forward), p = ld(c->position);
    memset(m, 0, 16 * sizeof(float));
    m[15] = 1.0f;
    m[0] = r.x; m[1] = r.y; m[2] = r.z; m[3] = -dot(r, p);
    m[4] = u.x; m[5] = u.y; m[6] = u.z; m[7] = -dot(u, p);
    m[8] = -f.x; m[9] = -f.y; m[10] = -f.z; m[11] = dot(f, p);
}

/* MakeViewProjection, Wave3DebugRuntime.cpp:183-203, with the camera's own
 * tan(vfov / 2) and aspect (width / height). */
void ora_view_projection(const ora_camera* c, float m[16]) {
    const float nearPlane = 0.01f, farPlane = 10000.0f;
    float proj[16], view[16];
    memset(proj, 0, sizeof(proj));
    proj[0] = 1.0f / (c->aspect * c->tanHalfFov);
    proj[5] = 1.0f / c->tanHalfFov;
    proj[10] = farPlane / (nearPlane - farPlane);
    proj[11] = farPlane * nearPlane / (nearPlane - farPlane);
    proj[14] = -1.0f;
    ora_world_to_view(c, view);
    mat_mul(proj, view, m);
}

/* mul(row_major M, float4(p, 1)): dot of each row, left to right. */
static void xform(const float* m, f3 p, float out[4]) {
    for (int r = 0; r < 4; ++r) {
        const float* row = m + 4 * r;
        out[r] = ((row[0] * p.x + row[1] * p.y) + row[2] * p.z) + row[3] * 1.0f;
    }
}

/* MotionVectorsCS main for an at-rest surface (world = object position). */
void ora_restir_motion(const float curViewProj[16], const float prevViewProj[16], const float prevWorldToView[16],
                       const ora_surface_v2* s, float out[6]) {
    memset(out, 0, 6 * sizeof(float));
    if ((s->identity[3] & ORA_SURFACE_VALID) == 0u) return;
    const f3 p = ld(s->worldPositionLinearDepth);
    if (!fin3(p)) return;
    float cc[4], pc[4], pv[4];
    xform(curViewProj, p, cc);
    xform(prevViewProj, p, pc);
    xform(prevWorldToView, p, pv);
    const int ok = cc[3] > 1.0e-8f && pc[3] > 1.0e-8f && fin4(cc) && fin4(pc) && fin4(pv);
    if (!ok) return;
    const float cnx = cc[0] / cc[3], cny = cc[1] / cc[3];
    const float pnx = pc[0] / pc[3], pny = pc[1] / pc[3];
    const float cu = (cnx * 0.5f + 0.5f) + 0.0f, cv = (0.5f - cny * 0.5f) + 0.0f;
    const float pu = (pnx * 0.5f + 0.5f) + 0.0f, pvv = (0.5f - pny * 0.5f) + 0.0f;
    const float depth = -pv[2];
    const int valid = fin(cu) && fin(cv) && fin(pu) && fin(pvv) && fin(depth) && depth >= 0.0f;
    out[0] = valid ? 1.0f : 0.0f;
    out[1] = cu;
    out[2] = cv;
    out[3] = valid ? pu : cu;
    out[4] = valid ? pvv : cv;
    out[5] = valid ? depth : 0.0f;
}

/* ---- passes ------------------------------------------------------------------ */

/* PbrToPersistentLightSampleV3, pbr_restir_candidate_export_v3.hlsl:35-78 */
static void to_persistent(const ora_shading* sh, const ora_light_sample_full* ls, uint32_t flattened,
                          const ora_restir_params* p, ora_light_sample_v3* out) {
    memset(out, 0, sizeof(*out));
    if ((ls->flags & 1u) == 0u || ls->lightIndex >= sh->lightCount) return;
    const ora_light* light = &sh->lights[ls->lightIndex];
    st3(out->positionDistance, ld(ls->position));
    out->positionDistance[3] = ls->distance;
    st3(out->directionCombinedPdf, ld(ls->wi));
    out->directionCombinedPdf[3] = ls->combinedPdfW;
    st3(out->radianceDiscretePdf, ld(ls->Li));
    out->radianceDiscretePdf[3] = ls->selectionPmf;
    if (ls->measure == ORA_MEASURE_AREA) out->conditionalPdf[0] = ls->conditionalPdf;
    else if (ls->measure == ORA_MEASURE_SOLID_ANGLE) out->conditionalPdf[1] = ls->conditionalPdf;
    out->identity[0] = light->identity[2];
    out->identity[1] = ls->primitiveId;
    out->identity[2] = flattened;
    out->identity[3] = p->generations[3];
    /* gCurrentToPreviousLightIndexV3[light].y: the light generation
     * (Wave3DebugRuntime.cpp:1876-1886). */
    out->generation[0] = p->generations[1];
    out->generation[1] = p->generations[0];
    out->generation[2] = p->generations[3];
    uint32_t flags = SAMPLE_FLAG_VALID;
    if ((ls->flags & 2u) != 0u) flags |= SAMPLE_FLAG_DELTA;
    if (ls->measure == ORA_MEASURE_AREA) flags |= SAMPLE_FLAG_AREA;
    if (ls->measure == ORA_MEASURE_SOLID_ANGLE) flags |= SAMPLE_FLAG_SOLID_ANGLE;
    /* PbrPrivateLightMeasureToAbiV1L6 (pbr_l6_types.hlsli:240-254). */
    const uint32_t abi = ls->measure == ORA_MEASURE_DISCRETE ? 1u
                         : ls->measure == ORA_MEASURE_AREA  ? 3u
                         : ls->measure == ORA_MEASURE_SOLID_ANGLE ? 2u : 0u;
    out->metadata[0] = abi;
    out->metadata[1] = flags;
    out->metadata[2] = light->identity[1];
    out->metadata[3] = ls->lightIndex;
}

void ora_restir_candidates(const ora_shading* sh, const ora_frame* fr, const ora_restir_params* p,
                           const ora_surface_v2* surfaces, ora_restir_candidate* out, ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2], C = p->extentAndCandidates[3];
    if (C == 0u) return;
    const uint32_t tag = fr->streamTag ^ 0x52335354u;
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        const ora_surface_v2* s = &surfaces[pixel];
        for (uint32_t c = 0; c < C; ++c) {
            const uint32_t flat = pixel * C + c;
            const uint32_t sampleIndex = fr->sampleIndex * C + c;
            float rnd[5];
            for (uint32_t k = 0; k < 5; ++k)
                rnd[k] = ora_sample_dimension(pixel, sampleIndex, k, tag, fr->seedLow, fr->seedHigh);
            ora_light_sample_full ls;
            ora_light_sample_full_eval(sh, s->worldPositionLinearDepth, rnd, &ls);
            ora_light_sample_v3 ps;
            to_persistent(sh, &ls, flat, p, &ps);
            ps.sourceIdentity[0] = pixel;
            ps.sourceIdentity[1] = s->identity[0];
            ps.sourceIdentity[2] = s->identity[1];
            ps.sourceIdentity[3] = s->identity[2];
            ora_restir_evaluate_candidate(&ps, s, p, p->modeAndFlags[1], REUSE_NONE, pixel, &out[flat]);
            if (stats) {
                stats->candidateCounts[0]++;
                if (ora_restir_candidate_valid(&out[flat], p)) stats->candidateCounts[1]++;
                else stats->candidateCounts[2]++;
            }
        }
    }
}

void ora_restir_initial(const ora_restir_params* p, const ora_surface_v2* surfaces,
                        const ora_restir_candidate* candidates, ora_reservoir_v3* out, ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2], C = p->extentAndCandidates[3];
    if (C == 0u) return;
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        ora_reservoir_v3 r;
        ora_restir_empty(&r);
        uint32_t state = ora_restir_hash(pixel ^ p->generations[3] ^ 0xa511e9b3u);
        for (uint32_t c = 0; c < C; ++c) {
            const ora_restir_candidate* cand = &candidates[(size_t)pixel * C + c];
            const float w = ora_restir_candidate_weight(cand, p);
            ora_restir_update(&r, cand, p, w, 1u, ora_restir_random(&state));
        }
        ora_restir_clamp_m(&r, p->reuseLimits[1]);
        ora_restir_finalize(&r);
        r.state[1] = 0u;
        r.surfaceIdentity[0] = surfaces[pixel].identity[0];
        r.surfaceIdentity[1] = surfaces[pixel].identity[1];
        r.surfaceIdentity[2] = surfaces[pixel].identity[2];
        r.surfaceIdentity[3] = 0u;
        r.sceneIdentity[0] = p->generations[0];
        r.sceneIdentity[1] = p->generations[1];
        r.sceneIdentity[2] = p->historyGenerations[0];
        r.sceneIdentity[3] = p->historyGenerations[1];
        r.frameIdentity[0] = p->generations[3];
        r.frameIdentity[1] = p->modeAndFlags[3];
        r.frameIdentity[2] = p->generations[2];
        r.frameIdentity[3] = HISTORY_VALID;
        r.reprojection[0] = pixel;
        r.reprojection[1] = p->extentAndCandidates[0];
        r.reprojection[2] = p->extentAndCandidates[1];
        r.reprojection[3] = REJECT_NONE;
        out[pixel] = r;
        if (stats && (r.state[2] & ORA_RES_M_CLAMPED) != 0u) stats->candidateCounts[3]++;
    }
}

void ora_restir_temporal_motion(const ora_restir_params* p, const float* motion, const ora_surface_v2* surfaces,
                                const ora_surface_v2* historySurfaces, const ora_reservoir_v3* initial,
                                const ora_reservoir_v3* history, ora_reservoir_v3* out, uint32_t* reasons,
                                ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2];
    const uint32_t width = p->extentAndCandidates[0], height = p->extentAndCandidates[1];
    const int historyEnabled = p->modeAndFlags[2] != 0u;
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        ora_reservoir_v3 result = initial[pixel];
        uint32_t rejection = REJECT_NONE;
        if (historyEnabled) {
            const float* mo = motion + 6u * pixel;
            uint32_t prior = INV;
            if (mo[0] == 0.0f) {
                rejection = REJECT_MOTION;
            } else if (mo[3] < 0.0f || mo[4] < 0.0f || mo[3] >= 1.0f || mo[4] >= 1.0f) {
                rejection = REJECT_OUTSIDE;
            } else {
                uint32_t qx = (uint32_t)(mo[3] * (float)width), qy = (uint32_t)(mo[4] * (float)height);
                if (qx > width - 1u) qx = width - 1u;
                if (qy > height - 1u) qy = height - 1u;
                prior = qy * width + qx;
            }
            ora_reservoir_v3 h;
            ora_restir_empty(&h);
            if (prior < pixels) h = history[prior];
            if (rejection == REJECT_NONE &&
                ((h.state[2] & ORA_RES_VALID) == 0u || h.state[0] == 0u || (h.frameIdentity[3] & HISTORY_VALID) == 0u))
                rejection = REJECT_INVALID;
            uint32_t expLow = p->generations[3], expHigh = p->modeAndFlags[3];
            if (expLow == 0u) {
                expLow = 0xffffffffu;
                expHigh -= 1u;
            } else {
                expLow -= 1u;
            }
            if (rejection == REJECT_NONE &&
                (h.frameIdentity[0] != expLow || h.frameIdentity[1] != expHigh || h.frameIdentity[2] != p->generations[2]))
                rejection = REJECT_INVALID;
            else if (rejection == REJECT_NONE && h.sceneIdentity[0] != p->generations[0])
                rejection = REJECT_SCENE;
            else if (rejection == REJECT_NONE && h.sceneIdentity[1] != p->generations[1])
                rejection = REJECT_LIGHT;
            else if (rejection == REJECT_NONE && h.sceneIdentity[2] != p->historyGenerations[0])
                rejection = REJECT_CAMERA_CUT;
            else if (rejection == REJECT_NONE && h.sceneIdentity[3] != p->historyGenerations[1])
                rejection = REJECT_RESIZE;
            else if (rejection == REJECT_NONE && h.state[1] >= p->reuseLimits[2])
                rejection = REJECT_AGE;
            else if (rejection == REJECT_NONE && prior >= pixels)
                rejection = REJECT_OUTSIDE;
            else if (rejection == REJECT_NONE)
                rejection = ora_restir_validate_pair(&surfaces[pixel], &historySurfaces[prior], p);
            if (rejection == REJECT_NONE) {
                const float expected = mo[5];
                const float hd = historySurfaces[prior].worldPositionLinearDepth[3];
                const float scale = fmaxf(1.0f, fmaxf(fabsf(expected), fabsf(hd)));
                if (fabsf(expected - hd) / scale > p->validation[1]) rejection = REJECT_DEPTH;
            }
            const uint32_t prevLight = h.selected.metadata[3];
            if (rejection == REJECT_NONE) {
                /* gPreviousToCurrentLightIndexV3: identity map with the light
                 * generation while history is reused (Wave3DebugRuntime.cpp:1873-1886). */
                if (prevLight >= p->lightTableCounts[1]) rejection = REJECT_LIGHT;
                else if (prevLight >= p->lightTableCounts[0] || p->generations[1] != h.selected.generation[0])
                    rejection = REJECT_LIGHT;
            }
            if (rejection == REJECT_NONE) {
                ora_light_sample_v3 remapped = h.selected;
                if (remapped.positionDistance[3] < 1.0e29f) {
                    const f3 delta = sub(ld(remapped.positionDistance), ld(surfaces[pixel].worldPositionLinearDepth));
                    const float dist = length(delta);
                    if (fin(dist) && dist > 1.0e-6f) {
                        remapped.positionDistance[3] = dist;
                        st3(remapped.directionCombinedPdf, divs(delta, dist));
                    } else {
                        remapped.metadata[1] &= ~SAMPLE_FLAG_VALID;
                    }
                }
                remapped.metadata[3] = prevLight;
                ora_restir_candidate cand;
                ora_restir_evaluate_candidate(&remapped, &surfaces[pixel], p, h.provenance[0], REUSE_TEMPORAL, prior,
                                              &cand);
                cand.tpsc[3] = h.weightState[1];
                const float mergeWeight = cand.tpsc[0] * cand.tpsc[2] * h.weightState[1] * (float)h.state[0];
                uint32_t state = ora_restir_hash(pixel ^ p->generations[3] ^ 0x63d83595u);
                if (ora_restir_candidate_valid(&cand, p) && fin(mergeWeight) && mergeWeight >= 0.0f) {
                    ora_restir_update(&result, &cand, p, mergeWeight, h.state[0], ora_restir_random(&state));
                    result.state[2] |= ORA_RES_TEMPORAL;
                    const uint32_t age = h.state[1] + 1u;
                    result.state[1] = age < p->reuseLimits[2] ? age : p->reuseLimits[2];
                    if (stats) stats->reuseCounts[1]++;
                } else {
                    rejection = REJECT_INVALID;
                }
            }
        }
        if (historyEnabled && stats) stats->reuseCounts[0]++;
        ora_restir_clamp_m(&result, p->reuseLimits[1]);
        ora_restir_finalize(&result);
        result.state[3] = rejection;
        result.reprojection[3] = rejection;
        reasons[pixel] = rejection;
        out[pixel] = result;
    }
}

/* Motion from the two cameras for surfaces at rest. */
void ora_restir_temporal(const ora_restir_params* p, const ora_camera* cur, const ora_camera* prev,
                         const ora_surface_v2* surfaces, const ora_surface_v2* historySurfaces,
                         const ora_reservoir_v3* initial, const ora_reservoir_v3* history, ora_reservoir_v3* out,
                         uint32_t* reasons, ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2];
    float vpCur[16], vpPrev[16], wvPrev[16];
    ora_view_projection(cur, vpCur);
    ora_view_projection(prev, vpPrev);
    ora_world_to_view(prev, wvPrev);
    float* motion = (float*)malloc(sizeof(float) * 6u * (pixels ? pixels : 1u));
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        ora_restir_motion(vpCur, vpPrev, wvPrev, &surfaces[pixel], motion + 6u * pixel);
    }
    ora_restir_temporal_motion(p, motion, surfaces, historySurfaces, initial, history, out, reasons, stats);
    free(motion);
}

uint32_t ora_restir_neighbor(const ora_restir_params* p, uint32_t mode, uint32_t radius, uint32_t pixel, uint32_t k) {
    const int32_t width = (int32_t)p->extentAndCandidates[0], height = (int32_t)p->extentAndCandidates[1];
    const int32_t x = (int32_t)(pixel % (uint32_t)width), y = (int32_t)(pixel / (uint32_t)width);
    int32_t dx, dy;
    if (mode == ORA_NEIGHBORS_RING) {
        static const int32_t dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
        const int32_t ring = (int32_t)(k / 8u) + 1;
        dx = dirs[k % 8u][0] * ring;
        dy = dirs[k % 8u][1] * ring;
    } else {
        /* Disk of radius `radius` pixels: r = radius * sqrt(u0), angle 2 pi u1,
         * rounded to the nearest pixel, u from the reservoir hash of
         * (pixel, frame, k). */
        uint32_t state = ora_restir_hash(pixel ^ ora_restir_hash(p->generations[3] + k * 0x9e3779b9u) ^ 0x2c1b3c6du);
        const float u0 = ora_restir_random(&state);
        const float u1 = ora_restir_random(&state);
        const float r = (float)radius * sqrtf(u0);
        float s, c;
        ora_sincos_2pi(u1, &s, &c);
        dx = (int32_t)floorf(r * c + 0.5f);
        dy = (int32_t)floorf(r * s + 0.5f);
    }
    int32_t sx = x + dx, sy = y + dy;
    sx = sx < 0 ? 0 : (sx > width - 1 ? width - 1 : sx);
    sy = sy < 0 ? 0 : (sy > height - 1 ? height - 1 : sy);
    return (uint32_t)(sy * width + sx);
}

void ora_restir_spatial(const ora_restir_params* p, uint32_t mode, uint32_t radius, const ora_surface_v2* surfaces,
                        const ora_reservoir_v3* temporal, uint32_t* reasons, ora_reservoir_v3* out,
                        ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2], K = p->reuseLimits[0];
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        const ora_reservoir_v3* center = &temporal[pixel];
        if (K == 0u) {
            out[pixel] = *center;
            continue;
        }
        ora_reservoir_v3 result;
        ora_restir_empty(&result);
        uint32_t state = ora_restir_hash(pixel ^ p->generations[3] ^ 0x9e3779b9u);
        uint32_t aggregate = reasons[pixel];
        ora_restir_candidate cc;
        ora_restir_evaluate_candidate(&center->selected, &surfaces[pixel], p, center->provenance[0],
                                      center->provenance[1], pixel, &cc);
        cc.tpsc[3] = center->weightState[1];
        ora_restir_update(&result, &cc, p, cc.tpsc[0] * cc.tpsc[2] * center->weightState[1] * (float)center->state[0],
                          center->state[0], ora_restir_random(&state));
        result.state[2] |= center->state[2] & (ORA_RES_M_CLAMPED | ORA_RES_TEMPORAL);
        for (uint32_t k = 0; k < K; ++k) {
            if (stats) stats->reuseCounts[2]++;
            const uint32_t nb = ora_restir_neighbor(p, mode, radius, pixel, k);
            if (nb >= pixels) {
                aggregate = REJECT_OUTSIDE;
                continue;
            }
            const uint32_t rej = ora_restir_validate_pair(&surfaces[pixel], &surfaces[nb], p);
            const ora_reservoir_v3* n = &temporal[nb];
            if (rej != REJECT_NONE || (n->state[2] & ORA_RES_VALID) == 0u) {
                aggregate = rej != REJECT_NONE ? rej : REJECT_INVALID;
                continue;
            }
            ora_restir_candidate c;
            ora_restir_evaluate_candidate(&n->selected, &surfaces[pixel], p, n->provenance[0], REUSE_SPATIAL, nb, &c);
            c.tpsc[3] = n->weightState[1];
            const float mw = c.tpsc[0] * c.tpsc[2] * n->weightState[1] * (float)n->state[0];
            if (ora_restir_candidate_valid(&c, p) && fin(mw) && mw >= 0.0f) {
                ora_restir_update(&result, &c, p, mw, n->state[0], ora_restir_random(&state));
                result.state[2] |= ORA_RES_SPATIAL;
                if (stats) stats->reuseCounts[3]++;
            } else {
                aggregate = REJECT_INVALID;
            }
        }
        ora_restir_clamp_m(&result, p->reuseLimits[1]);
        ora_restir_finalize(&result);
        result.state[1] = center->state[1];
        memcpy(result.surfaceIdentity, center->surfaceIdentity, 64);
        result.reprojection[0] = pixel;
        result.reprojection[3] = aggregate;
        result.state[3] = aggregate;
        reasons[pixel] = aggregate;
        out[pixel] = result;
    }
}

/* RestirMakeVisibilityRayV3 (Physical), RestirProductionV3.hlsli:140-205 */
static void make_visibility_ray(uint32_t rayIndex, uint32_t queryIndex, const ora_reservoir_v3* r,
                                const ora_surface_v2* s, ora_ray_v1* ray) {
    memset(ray, 0, sizeof(*ray));
    ray->identity[0] = rayIndex;
    ray->identity[1] = queryIndex;
    if ((r->state[2] & ORA_RES_VALID) == 0u || !ora_restir_surface_valid(s)) return;
    const f3 n = normalize(ld(s->geometricNormalRoughness));
    f3 dir = normalize(ld(r->selected.directionCombinedPdf));
    float maxDist = 1.0e30f;
    const float pd = r->selected.positionDistance[3];
    if (fin(pd) && pd > 0.0f) {
        const f3 disp = sub(ld(r->selected.positionDistance), ld(s->worldPositionLinearDepth));
        const float dist = length(disp);
        if (dist > 2.0e-4f) {
            dir = divs(disp, dist);
            maxDist = fmaxf(dist - 1.0e-4f, 1.0e-4f);
        }
    }
    st3(ray->originTMin, add(ld(s->worldPositionLinearDepth), muls(n, 1.0e-4f)));
    ray->originTMin[3] = 1.0e-4f;
    st3(ray->directionTMax, dir);
    ray->directionTMax[3] = maxDist;
    ray->identity[3] = 0xffu;
}

void ora_restir_visibility_rays(const ora_restir_params* p, const ora_reservoir_v3* spatial,
                                const ora_surface_v2* surfaces, ora_ray_v1* rays, ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2];
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        ora_ray_v1* ray = &rays[pixel];
        memset(ray, 0, sizeof(*ray));
        ray->identity[0] = pixel;
        ray->identity[1] = pixel;
        if ((spatial[pixel].state[2] & ORA_RES_VIS_EVALUATED) == 0u)
            make_visibility_ray(pixel, pixel, &spatial[pixel], &surfaces[pixel], ray);
        if (ray->identity[3] != 0u && stats) stats->visibilityCounts[0]++;
    }
}

/* RestirReferenceVisibilityPrepareV3 (RestirReferenceVisibilityPrepareV3.comp.hlsl:10-45),
 * Physical: query q = pixel * (K + 1) + offset, source = the pixel (offset 0)
 * or its neighbour offset - 1. */
void ora_restir_reference_rays(const ora_restir_params* p, uint32_t mode, uint32_t radius,
                               const ora_reservoir_v3* spatial, const ora_surface_v2* surfaces, ora_ray_v1* rays,
                               ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2], sources = p->reuseLimits[0] + 1u;
    for (uint32_t q = 0; q < pixels * sources; ++q) {
        const uint32_t center = q / sources, offset = q % sources;
        const uint32_t src = offset == 0u ? center : ora_restir_neighbor(p, mode, radius, center, offset - 1u);
        make_visibility_ray(q, q, &spatial[center], &surfaces[src], &rays[q]);
        if (rays[q].identity[3] != 0u && stats) stats->visibilityCounts[0]++;
    }
}

/* RestirReferenceResolveV3 (RestirReferenceResolveV3.comp.hlsl:82-144), Physical. */
void ora_restir_reference_resolve(const ora_restir_params* p, uint32_t mode, uint32_t radius,
                                  ora_reservoir_v3* spatial, const ora_reservoir_v3* temporal,
                                  const ora_surface_v2* surfaces, const ora_hit_v1* hits, ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2], sources = p->reuseLimits[0] + 1u;
    for (uint32_t center = 0; center < pixels; ++center) {
        ora_reservoir_v3* r = &spatial[center];
        if ((r->state[2] & ORA_RES_VALID) == 0u) continue;
        float denominator = 0.0f, selectedSourceTarget = 0.0f, currentTarget = 0.0f;
        for (uint32_t offset = 0; offset < sources; ++offset) {
            const uint32_t src = offset == 0u ? center : ora_restir_neighbor(p, mode, radius, center, offset - 1u);
            if (src >= pixels) continue;
            const uint32_t q = center * sources + offset;
            const ora_hit_v1* h = &hits[q];
            const int matches = h->metadata[0] == q && h->metadata[3] == q && h->metadata[1] != HIT_INVALID;
            const float visibility = matches && h->metadata[1] == HIT_MISS ? 1.0f : 0.0f;
            float dv[3], sv[3], target;
            ora_restir_evaluate_unshadowed(&r->selected, &surfaces[src], p, dv, sv, &target);
            const uint32_t m = temporal[src].state[0];
            denominator += (float)(m > 1u ? m : 1u) * target * visibility;
            if (src == center) currentTarget = target;
            if (src == r->provenance[2]) selectedSourceTarget = target * visibility;
            if (stats) {
                stats->visibilityCounts[1]++;
                if (visibility > 0.0f) stats->visibilityCounts[2]++;
            }
        }
        const float norm = currentTarget * denominator;
        r->weightState[1] = norm > 0.0f ? r->weightState[0] * selectedSourceTarget / norm : 0.0f;
        r->selectedTerms[3] = r->weightState[1];
        r->state[2] |= ORA_RES_REFERENCE_MODE;
        if (!(r->weightState[1] > 0.0f) || !fin(r->weightState[1])) r->state[2] &= ~ORA_RES_VALID;
    }
}

void ora_restir_resolve(const ora_restir_params* p, ora_reservoir_v3* spatial, const ora_hit_v1* hits,
                        ora_restir_stats* stats) {
    const uint32_t pixels = p->extentAndCandidates[2];
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        ora_reservoir_v3* r = &spatial[pixel];
        const int valid = (r->state[2] & ORA_RES_VALID) != 0u;
        const int duplicate = (r->state[2] & ORA_RES_VIS_EVALUATED) != 0u;
        if (valid && !duplicate) {
            const ora_hit_v1* h = &hits[pixel];
            const int matches = h->metadata[0] == pixel && h->metadata[3] == pixel && h->metadata[1] != HIT_INVALID;
            const float vis = matches && h->metadata[1] == HIT_MISS ? 1.0f : 0.0f;
            const float fw = r->weightState[1] * vis;
            const int finite = fin(vis) && vis >= 0.0f && vis <= 1.0f && fin(fw) && fw >= 0.0f;
            const int visible = finite && vis > 0.0f;
            r->state[2] |= ORA_RES_VIS_EVALUATED;
            r->selectedTerms[3] = finite ? fw : 0.0f;
            if (!finite) {
                r->state[2] &= ~ORA_RES_VALID;
                r->state[3] = REJECT_INVALID;
                r->reprojection[3] = REJECT_INVALID;
            }
            if (visible) r->state[2] |= ORA_RES_VIS_VALID;
            else r->state[2] &= ~ORA_RES_VIS_VALID;
            if (stats) {
                stats->visibilityCounts[1]++;
                if (visible) stats->visibilityCounts[2]++;
            }
        } else if (duplicate && stats) {
            stats->visibilityCounts[3]++;
        }
    }
}

void ora_restir_publish_direct(const ora_restir_params* p, const ora_reservoir_v3* spatial,
                               const ora_surface_v2* surfaces, float* direct) {
    const uint32_t pixels = p->extentAndCandidates[2];
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        const ora_reservoir_v3* r = &spatial[pixel];
        float dv[3] = {0.0f, 0.0f, 0.0f}, sv[3] = {0.0f, 0.0f, 0.0f};
        const uint32_t need = ORA_RES_VALID | ORA_RES_VIS_EVALUATED | ORA_RES_VIS_VALID;
        if ((r->state[2] & need) == need) {
            float target;
            ora_restir_evaluate_unshadowed(&r->selected, &surfaces[pixel], p, dv, sv, &target);
            for (int k = 0; k < 3; ++k) {
                dv[k] *= r->selectedTerms[3];
                sv[k] *= r->selectedTerms[3];
            }
        }
        float* d = direct + 8u * pixel;
        d[0] = dv[0]; d[1] = dv[1]; d[2] = dv[2]; d[3] = 1.0f;
        d[4] = sv[0]; d[5] = sv[1]; d[6] = sv[2]; d[7] = 1.0f;
    }
}

void ora_restir_publish_history(const ora_restir_params* p, const ora_reservoir_v3* spatial,
                                const ora_surface_v2* surfaces, ora_reservoir_v3* history) {
    const uint32_t pixels = p->extentAndCandidates[2];
    for (uint32_t pixel = 0; pixel < pixels; ++pixel) {
        ora_reservoir_v3 r = spatial[pixel];
        r.surfaceIdentity[0] = surfaces[pixel].identity[0];
        r.surfaceIdentity[1] = surfaces[pixel].identity[1];
        r.surfaceIdentity[2] = surfaces[pixel].identity[2];
        r.surfaceIdentity[3] = 0u;
        r.sceneIdentity[0] = p->generations[0];
        r.sceneIdentity[1] = p->generations[1];
        const uint32_t need = ORA_RES_VALID | ORA_RES_VIS_EVALUATED;
        r.frameIdentity[0] = p->generations[3];
        r.frameIdentity[1] = p->modeAndFlags[3];
        r.frameIdentity[2] = p->generations[2];
        r.frameIdentity[3] = (r.state[2] & need) == need ? HISTORY_VALID : HISTORY_RESET;
        r.reprojection[0] = pixel;
        r.reprojection[1] = p->extentAndCandidates[0];
        r.reprojection[2] = p->extentAndCandidates[1];
        history[pixel] = r;
    }
}
