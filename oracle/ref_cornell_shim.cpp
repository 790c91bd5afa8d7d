// Thin extern "C" wrapper around the reference's own CPU Cornell path tracer
// (src/integrators/reference_cpu/CornellReference.cpp), compiled from the
// reference sources where they lie into oracle/_ref/libref.so by
// oracle/Makefile.  Test infrastructure only: it pins the oracle's
// restatement and never ships.
#include "integrators/reference_cpu/CornellReference.hpp"

#include <cstdint>
#include <cstring>
#include <string>

namespace Ref = RenderingEngine::Integrators::ReferenceCpu;

extern "C" {

// Renders the whole w x h film; pixels: w*h*3 floats.  counts: rays, shadow
// rays, non-finite.  Returns 0, or 1 when the reference reports an error.
int ref_cornell_render(std::uint32_t w, std::uint32_t h, std::uint32_t spp, std::uint32_t maxBounces,
                       std::uint64_t seed, std::uint32_t threads, float* pixels, std::uint64_t counts[3],
                       std::uint64_t* sceneHash, double* elapsedMs)
{
    Ref::CornellReferenceOptions options;
    options.width = w;
    options.height = h;
    options.spp = spp;
    options.maxBounces = maxBounces;
    options.seed = seed;
    options.threads = threads;
    Ref::CornellReferenceResult result;
    std::string error;
    if (!Ref::RenderCornellReference(options, result, error)) return 1;
    std::memcpy(pixels, result.pixels.data(), result.pixels.size() * sizeof(float));
    counts[0] = result.rayCount;
    counts[1] = result.shadowRayCount;
    counts[2] = result.nonFiniteCount;
    *sceneHash = result.sceneHash;
    *elapsedMs = result.elapsedMilliseconds;
    return 0;
}
}
