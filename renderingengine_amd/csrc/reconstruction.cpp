// C entry points of the reconstruction boundary: MotionVectorsCS and
// ComposeCS over caller-owned device buffers of ABI-v2 records
// (include/contracts/ReconstructionAbiV2.hpp:45-116).  The passes are
// stateless; the progressive-mean Film is the caller's buffer and sample
// count, as ComposeCS's gComposeFilm / gFilmSampleCount
// (reconstruction/shaders/ComposeCS.hlsl:12-24).
#include <cstdint>

#include "rex_internal.hpp"

using rex::check_hip;
using rex::fail;

namespace {

bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16u == 0u; }

}  // namespace

extern "C" {

rex_status rex_motion_vectors_v2(void* gbuffer, const void* motionInputs, const float* currentObjectToWorld,
                                 const float* previousObjectToWorld, const rex_motion_constants_v2* constants,
                                 void* stream) {
    if (gbuffer == nullptr || motionInputs == nullptr || constants == nullptr) {
        return fail(REX_INVALID_ARGUMENT, "NULL argument");
    }
    const rex_motion_constants_v2& c = *constants;
    const uint64_t pixels = static_cast<uint64_t>(c.width) * c.height;
    if (pixels == 0u || pixels > 0xffffffffull) return fail(REX_INVALID_ARGUMENT, "extent must be nonzero");
    if ((c.currentTransformCount != 0u && currentObjectToWorld == nullptr) ||
        (c.previousTransformCount != 0u && previousObjectToWorld == nullptr)) {
        return fail(REX_INVALID_ARGUMENT, "transform array is NULL but its count is not zero");
    }
    if (!aligned16(gbuffer) || !aligned16(motionInputs) || !aligned16(currentObjectToWorld) ||
        !aligned16(previousObjectToWorld)) {
        return fail(REX_INVALID_ARGUMENT, "buffers must be 16-byte aligned");
    }
    return check_hip(rex::launch_motion_vectors(static_cast<rex::GBufferV2*>(gbuffer),
                                                static_cast<const rex::MotionInputV2*>(motionInputs),
                                                reinterpret_cast<const float4*>(currentObjectToWorld),
                                                reinterpret_cast<const float4*>(previousObjectToWorld), c,
                                                static_cast<hipStream_t>(stream)),
                     "motion vectors");
}

rex_status rex_compose_v2(const rex_compose_desc_v2* desc, void* stream) {
    if (desc == nullptr) return fail(REX_INVALID_ARGUMENT, "desc is NULL");
    const rex_compose_desc_v2& d = *desc;
    const uint64_t pixels = static_cast<uint64_t>(d.width) * d.height;
    if (pixels == 0u || pixels > 0xffffffffull) return fail(REX_INVALID_ARGUMENT, "extent must be nonzero");
    if (d.output == nullptr) return fail(REX_INVALID_ARGUMENT, "output is NULL");
    switch (d.outputMode) {
        case REX_COMPOSE_RAW:
        case REX_COMPOSE_PROGRESSIVE_MEAN:
            if (d.rawSignal == nullptr || d.terminal == nullptr) {
                return fail(REX_INVALID_ARGUMENT, "raw and progressive-mean output read rawSignal and terminal");
            }
            if (d.outputMode == REX_COMPOSE_PROGRESSIVE_MEAN && d.film == nullptr) {
                return fail(REX_INVALID_ARGUMENT, "progressive-mean output needs a film");
            }
            break;
        case REX_COMPOSE_MOTION:
        case REX_COMPOSE_NORMAL:
        case REX_COMPOSE_ROUGHNESS:
        case REX_COMPOSE_METALLIC:
        case REX_COMPOSE_BASE_COLOR:
        case REX_COMPOSE_EMISSIVE:
            if (d.gbuffer == nullptr) return fail(REX_INVALID_ARGUMENT, "this output mode reads the G-buffer");
            if ((d.outputMode == REX_COMPOSE_BASE_COLOR || d.outputMode == REX_COMPOSE_EMISSIVE) &&
                d.materialCount != 0u && d.materials == nullptr) {
                return fail(REX_INVALID_ARGUMENT, "materials is NULL but materialCount is not zero");
            }
            break;
        case 1u: case 2u: case 3u: case 5u: case 6u: case 7u: case 8u: case 9u:
            return fail(REX_UNSUPPORTED, "temporal / a-trous / SVGF history outputs are not provided");
        default:
            return fail(REX_INVALID_ARGUMENT, "unknown output mode");
    }
    const void* ptrs[] = {d.gbuffer, d.rawSignal, d.terminal, d.materials, d.film, d.output};
    for (const void* p : ptrs) {
        if (!aligned16(p)) return fail(REX_INVALID_ARGUMENT, "buffers must be 16-byte aligned");
    }
    return check_hip(rex::launch_compose(d, static_cast<hipStream_t>(stream)), "compose");
}

}  // extern "C"
