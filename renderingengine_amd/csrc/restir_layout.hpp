// ReSTIR DI records in HBM: the reference's versioned ABI structs, byte for
// byte (include/contracts/RestirAbiV3.hpp:109-193 and
// include/contracts/ReconstructionAbiV2.hpp's GpuPrimarySurfaceV2), so the
// buffers can be handed to a caller of the reference's ABI unchanged.
//
// Per pixel of a ReSTIR frame (P = width * height, spp 1):
//   SurfaceV2   surfaces[2][P]     96 B  current and previous frame (ping-pong)
//   ReservoirV3 initial[P]        256 B  after initial RIS
//   ReservoirV3 temporal[P]       256 B  after temporal reuse
//   ReservoirV3 spatial[P]        256 B  after spatial reuse, then resolved
//   ReservoirV3 history[2][P]     256 B  published for the next frame (ping-pong)
//   rex_ray_v1 / rex_hit_v1       64 + 96 B  winner visibility query
//   float4 direct[2][P]           32 B   ReSTIR diffuse and specular direct light
//   uint32 reasons[P]             4 B    rejection reason of the last reuse stage
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rex.h"
#include "render_layout.hpp"
#include "rex_layout.hpp"

namespace rex {

struct alignas(16) LightSampleV3 {  // GpuPersistentLightSampleV3, 128 B
    F4 positionDistance;
    F4 directionCombinedPdf;
    F4 radianceDiscretePdf;
    F4 conditionalPdf;
    U4 identity;        // stable light, primitive, sample low, sample high
    U4 generation;      // light, scene, sample generation, -
    U4 metadata;        // ABI-v1 measure, sample flags, light flags, light-table index
    U4 sourceIdentity;  // source surface, material, instance, primitive
};
static_assert(sizeof(LightSampleV3) == 128, "GpuPersistentLightSampleV3 is 128 B");

struct alignas(16) ReservoirV3 {  // GpuRestirReservoirV3, 256 B
    LightSampleV3 selected;
    F4 weightState;    // wSum, W, selected target, selected proposal pdf
    F4 selectedTerms;  // support, correction, candidate weight, final weight
    U4 state;          // M, age, flags, rejection
    U4 provenance;     // candidate source, reuse source, source surface, sample flags
    U4 surfaceIdentity;
    U4 sceneIdentity;
    U4 frameIdentity;
    U4 reprojection;
};
static_assert(sizeof(ReservoirV3) == 256, "GpuRestirReservoirV3 is 256 B");

struct alignas(16) SurfaceV2 {  // GpuPrimarySurfaceV2, 96 B
    F4 posDepth;   // world position, linear depth
    F4 gnRough;    // geometric normal, roughness
    F4 snMetal;    // shading normal, metallic
    F4 diffuse;    // diffuse albedo
    F4 specular;   // specular albedo
    U4 identity;   // material, instance, primitive, flags
};
static_assert(sizeof(SurfaceV2) == 96, "GpuPrimarySurfaceV2 is 96 B");

// The rest of the ABI-v2 reconstruction records
// (include/contracts/ReconstructionAbiV2.hpp:58-90).
struct alignas(16) MotionV2 {  // GpuMotionVectorV2, 48 B
    F4 motionDepth;            // previous UV - current UV, previous linear depth, 0
    F4 uv;                     // current UV, previous UV
    U4 identity;               // material, instance, primitive, motion flags
};
struct alignas(16) SignalV2 {  // GpuReconstructionSignalV2, 64 B
    F4 dd, ds, id, is;         // direct / indirect diffuse / specular
};
struct alignas(16) GBufferV2 {  // GpuGBufferRecordV2, 208 B
    SurfaceV2 primary;
    MotionV2 motion;
    SignalV2 signal;
};
struct MotionInputV2 {  // PbrMotionInputV2, 32 B
    float px, py, pz;
    uint32_t current, previous, r0, r1, r2;
};
static_assert(sizeof(MotionV2) == 48 && sizeof(SignalV2) == 64 && sizeof(GBufferV2) == 208 &&
                  sizeof(MotionInputV2) == 32,
              "ABI-v2 reconstruction record sizes");
static_assert(sizeof(GBufferV2) == sizeof(rex_gbuffer_record_v2) && sizeof(MotionInputV2) == sizeof(rex_motion_input_v2),
              "public and internal ABI-v2 records agree");

// A scene instance's world-to-object rows 0-2, for the motion inputs.
struct InstanceXform {
    F4 row0, row1, row2;
};

// Reconstruction exports written by Resolve (k_resolve_film); NULL pointers
// are skipped.  instanceMap: (stable instance ID, first index) sorted by ID.
struct ExportV2 {
    GBufferV2* gbuffer = nullptr;
    MotionInputV2* motion = nullptr;
    SignalV2* raw = nullptr;
    SurfaceV2* primary = nullptr;
    float4* terminal = nullptr;
    const SurfaceV2* restirSurfaces = nullptr;  // copied to primary (ReSTIR frames)
    const InstanceXform* instances = nullptr;
    const uint2* instanceMap = nullptr;
    uint32_t instanceCount = 0, mapCount = 0;
};

struct alignas(16) RestirParams {  // GpuRestirParametersV3, 128 B
    U4 extentAndCandidates;  // width, height, pixels, candidates
    U4 reuseLimits;          // neighbours, max M, max age, frames in flight
    U4 generations;          // scene, light, history, frame low
    U4 historyGenerations;   // config, resource, debug mode, shadow mode
    U4 modeAndFlags;         // estimator mode, candidate source, history valid, frame high
    U4 lightTableCounts;     // current, previous, -, -
    F4 validation;           // normal cosine, relative depth, position, thin
    F4 cameraPosition;
};
static_assert(sizeof(RestirParams) == 128, "GpuRestirParametersV3 is 128 B");

struct alignas(16) RestirStats {  // GpuRestirStatisticsV3, 64 B
    U4 candidateCounts;   // generated, valid, rejected, M-clamped
    U4 reuseCounts;       // temporal attempts, accepted, spatial attempts, accepted
    U4 visibilityCounts;  // submitted, evaluated, visible, duplicate
    F4 errorMetrics;
};


struct RestirLaunch {
    RestirParams params;
    FrameConst frame;
    ShadingView shading;
    const SurfaceV2* surfaces;
    const SurfaceV2* historySurfaces;
    ReservoirV3* initial;
    ReservoirV3* temporal;
    ReservoirV3* spatial;
    const ReservoirV3* history;
    ReservoirV3* published;
    uint32_t* reasons;
    rex_ray_v1* rays;
    rex_hit_v1* hits;
    float4* direct;
    RestirStats* stats;
    float vpCur[16], vpPrev[16], wvPrev[16];
    uint32_t neighborMode, radius;
    bool wideVisibility;  // visibility queries may walk the compressed 4-wide nodes
    uint32_t estimatorMode = 0;       // REX_RESTIR_ESTIMATOR_*
    rex_ray_v1* refRays = nullptr;    // reference correction: (K + 1) queries per pixel
    rex_hit_v1* refHits = nullptr;
    // Object motion of the temporal reuse (NULL instanceMap: at rest).
    const InstanceXform* instances = nullptr;
    const uint2* instanceMap = nullptr;
    const float4* objectToWorld = nullptr;
    const float4* prevObjectToWorld = nullptr;
    uint32_t instanceCount = 0, mapCount = 0;
};

}  // namespace rex
