/*
 * Oracle (test infrastructure only): ReSTIR DI V3, restated per pixel.
 *
 * Records mirror the reference's versioned ABI structs byte for byte:
 *   GpuPersistentLightSampleV3 / GpuRestirCandidateV3 / GpuRestirReservoirV3 /
 *   GpuRestirStatisticsV3 / GpuRestirParametersV3
 *       include/contracts/RestirAbiV3.hpp:109-193
 *   GpuPrimarySurfaceV2  include/contracts/ReconstructionAbiV2.hpp (96 B)
 *   GpuRayQueueRecordV1 / GpuHitQueueRecordV1 (ora_ray_v1 / ora_hit_v1)
 *
 * Only the Physical shadow mode (one any-hit visibility ray per pixel) and
 * the explicitly biased estimator mode are restated.
 */
#ifndef REX_ORACLE_RESTIR_H
#define REX_ORACLE_RESTIR_H

#include <stdint.h>

#include "oracle.h"
#include "oracle_shading.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_light_sample_v3 {
    float positionDistance[4];
    float directionCombinedPdf[4];
    float radianceDiscretePdf[4];
    float conditionalPdf[4];
    uint32_t identity[4];        /* stable light, primitive, sample low, sample high */
    uint32_t generation[4];      /* light, scene, sample generation, - */
    uint32_t metadata[4];        /* measure, sample flags, light flags, light-table index */
    uint32_t sourceIdentity[4];  /* source surface, material, instance, primitive */
} ora_light_sample_v3;

typedef struct ora_restir_candidate {
    ora_light_sample_v3 sample;
    float tpsc[4];               /* target, proposal pdf, support, correction */
    uint32_t provenance[4];      /* candidate source, reuse source, source surface, sample flags */
} ora_restir_candidate;

typedef struct ora_reservoir_v3 {
    ora_light_sample_v3 selected;
    float weightState[4];        /* wSum, W, selected target, selected proposal pdf */
    float selectedTerms[4];      /* support, correction, candidate weight, final weight */
    uint32_t state[4];           /* M, age, reservoir flags, rejection */
    uint32_t provenance[4];
    uint32_t surfaceIdentity[4];
    uint32_t sceneIdentity[4];
    uint32_t frameIdentity[4];
    uint32_t reprojection[4];
} ora_reservoir_v3;

typedef struct ora_surface_v2 {
    float worldPositionLinearDepth[4];
    float geometricNormalRoughness[4];
    float shadingNormalMetallic[4];
    float diffuseAlbedo[4];
    float specularAlbedo[4];
    uint32_t identity[4];        /* material, instance, primitive, flags */
} ora_surface_v2;

typedef struct ora_restir_params {
    uint32_t extentAndCandidates[4];  /* width, height, pixels, candidates */
    uint32_t reuseLimits[4];          /* neighbours, max M, max age, frames in flight */
    uint32_t generations[4];          /* scene, light, history, frame low */
    uint32_t historyGenerations[4];   /* config, resource, debug, shadow mode */
    uint32_t modeAndFlags[4];         /* estimator, candidate source, history valid, frame high */
    uint32_t lightTableCounts[4];     /* current, previous, -, - */
    float validation[4];              /* normal cosine, relative depth, position, thin */
    float cameraPosition[4];
} ora_restir_params;

typedef struct ora_restir_stats {
    uint32_t candidateCounts[4];   /* generated, valid, rejected, M-clamped */
    uint32_t reuseCounts[4];       /* temporal attempts, accepted, spatial attempts, accepted */
    uint32_t visibilityCounts[4];  /* submitted, evaluated, visible, duplicate */
    float errorMetrics[4];
} ora_restir_stats;

/* The camera the frame was rendered with (same fields as ora_frame's). */
typedef struct ora_camera {
    float position[3];
    float tanHalfFov;
    float forward[3];
    float aspect;
    float right[3];
    float up[3];
} ora_camera;

/* Primary surface flags (ReconstructionAbiV2.hpp:15-22). */
#define ORA_SURFACE_VALID 1u
#define ORA_SURFACE_FRONT 2u
#define ORA_SURFACE_DIFFUSE 4u
#define ORA_SURFACE_SPECULAR 8u

/* Reservoir flags and rejection reasons (RestirAbiV3.hpp:58-101). */
#define ORA_RES_VALID 1u
#define ORA_RES_M_CLAMPED 2u
#define ORA_RES_TEMPORAL 4u
#define ORA_RES_SPATIAL 8u
#define ORA_RES_REFERENCE_MODE 16u
#define ORA_RES_VIS_EVALUATED 32u
#define ORA_RES_VIS_VALID 64u

/* Spatial neighbour tables. */
#define ORA_NEIGHBORS_RING 0u  /* Wave3DebugRuntime.cpp:1888-1925 */
#define ORA_NEIGHBORS_DISK 1u  /* seeded disk of radius r (this library's own) */

/* ---- primitives (RestirProductionV3.hlsli) -------------------------------- */
uint32_t ora_restir_hash(uint32_t v);
float ora_restir_random(uint32_t* state);
int ora_restir_surface_valid(const ora_surface_v2* s);
int ora_restir_sample_valid(const ora_light_sample_v3* s, const ora_restir_params* p);
void ora_restir_evaluate_unshadowed(const ora_light_sample_v3* s, const ora_surface_v2* surf,
                                    const ora_restir_params* p, float diffuse[3], float specular[3], float* target);
void ora_restir_evaluate_candidate(const ora_light_sample_v3* s, const ora_surface_v2* surf,
                                   const ora_restir_params* p, uint32_t candidateSource, uint32_t reuseSource,
                                   uint32_t sourceSurface, ora_restir_candidate* out);
int ora_restir_candidate_valid(const ora_restir_candidate* c, const ora_restir_params* p);
float ora_restir_candidate_weight(const ora_restir_candidate* c, const ora_restir_params* p);
void ora_restir_empty(ora_reservoir_v3* r);
void ora_restir_update(ora_reservoir_v3* r, const ora_restir_candidate* c, const ora_restir_params* p, float weight,
                       uint32_t multiplicity, float randomValue);
void ora_restir_finalize(ora_reservoir_v3* r);
void ora_restir_clamp_m(ora_reservoir_v3* r, uint32_t maxM);
uint32_t ora_restir_validate_pair(const ora_surface_v2* cur, const ora_surface_v2* src, const ora_restir_params* p);

/* ---- camera motion of a static scene ---------------------------------------
 * MakeWorldToView / MakeViewProjection (src/renderers/Wave3DebugRuntime.cpp:161-202)
 * for width x height, row-major 4x4. */
void ora_world_to_view(const ora_camera* c, float m[16]);
void ora_view_projection(const ora_camera* c, float m[16]);
/* MotionVectorsCS (reconstruction/shaders/MotionVectorsCS.hlsl) for a
 * surface at rest with zero jitter.  out = (valid, current u, current v,
 * previous u, previous v, previous linear depth). */
void ora_restir_motion(const float curViewProj[16], const float prevViewProj[16], const float prevWorldToView[16],
                       const ora_surface_v2* s, float out[6]);

/* ---- one pass each, over all pixels ------------------------------------- */
/* Candidate export (pbr_restir_candidate_export_v3.hlsl:13-114): C per pixel. */
void ora_restir_candidates(const ora_shading* sh, const ora_frame* fr, const ora_restir_params* p,
                           const ora_surface_v2* surfaces, ora_restir_candidate* out, ora_restir_stats* stats);
/* RestirInitialV3.comp.hlsl:9-58 */
void ora_restir_initial(const ora_restir_params* p, const ora_surface_v2* surfaces,
                        const ora_restir_candidate* candidates, ora_reservoir_v3* out, ora_restir_stats* stats);
/* RestirTemporalV3.comp.hlsl:20-215 with each pixel's motion given as
 * MotionVectorsCS leaves it in gMotionGBufferV3: motion[6 * pixel] = (valid,
 * current u, current v, previous u, previous v, previous linear depth), the
 * previous UV / depth being the current UV / 0 where not valid and all zero
 * where the record was cleared. */
void ora_restir_temporal_motion(const ora_restir_params* p, const float* motion, const ora_surface_v2* surfaces,
                                const ora_surface_v2* historySurfaces, const ora_reservoir_v3* initial,
                                const ora_reservoir_v3* history, ora_reservoir_v3* out, uint32_t* reasons,
                                ora_restir_stats* stats);
/* The same with the motion of surfaces at rest from the two cameras. */
void ora_restir_temporal(const ora_restir_params* p, const ora_camera* cur, const ora_camera* prev,
                         const ora_surface_v2* surfaces, const ora_surface_v2* historySurfaces,
                         const ora_reservoir_v3* initial, const ora_reservoir_v3* history, ora_reservoir_v3* out,
                         uint32_t* reasons, ora_restir_stats* stats);
/* Neighbour k of pixel (Wave3DebugRuntime.cpp:1888-1925 for the ring). */
uint32_t ora_restir_neighbor(const ora_restir_params* p, uint32_t mode, uint32_t radius, uint32_t pixel, uint32_t k);
/* RestirSpatialV3.comp.hlsl:12-112 */
void ora_restir_spatial(const ora_restir_params* p, uint32_t mode, uint32_t radius, const ora_surface_v2* surfaces,
                        const ora_reservoir_v3* temporal, uint32_t* reasons, ora_reservoir_v3* out,
                        ora_restir_stats* stats);
/* RestirWinnerVisibilityPrepareV3.comp.hlsl:9-36 (Physical: one ray per pixel). */
void ora_restir_visibility_rays(const ora_restir_params* p, const ora_reservoir_v3* spatial,
                                const ora_surface_v2* surfaces, ora_ray_v1* rays, ora_restir_stats* stats);
/* RestirWinnerResolveV3.comp.hlsl:78-122 */
void ora_restir_reference_rays(const ora_restir_params* p, uint32_t mode, uint32_t radius,
                               const ora_reservoir_v3* spatial, const ora_surface_v2* surfaces, ora_ray_v1* rays,
                               ora_restir_stats* stats);
void ora_restir_reference_resolve(const ora_restir_params* p, uint32_t mode, uint32_t radius,
                                  ora_reservoir_v3* spatial, const ora_reservoir_v3* temporal,
                                  const ora_surface_v2* surfaces, const ora_hit_v1* hits, ora_restir_stats* stats);
void ora_restir_resolve(const ora_restir_params* p, ora_reservoir_v3* spatial, const ora_hit_v1* hits,
                        ora_restir_stats* stats);
/* RestirPublishDirectV3.comp.hlsl:11-42: direct[2 * pixel] diffuse, [2 * pixel + 1] specular (float4). */
void ora_restir_publish_direct(const ora_restir_params* p, const ora_reservoir_v3* spatial,
                               const ora_surface_v2* surfaces, float* direct);
/* RestirPublishHistoryV3.comp.hlsl */
void ora_restir_publish_history(const ora_restir_params* p, const ora_reservoir_v3* spatial,
                                const ora_surface_v2* surfaces, ora_reservoir_v3* history);

#ifdef __cplusplus
}
#endif
#endif
