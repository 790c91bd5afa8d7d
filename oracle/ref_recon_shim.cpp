// Thin extern "C" wrapper around the reference's own CPU motion vectors
// (reconstruction/src/MotionVectors.cpp), compiled from the reference
// sources where they lie into oracle/_ref/libref.so by oracle/Makefile.
// Test infrastructure only: it pins the oracle's motion vectors.
#include "reconstruction/MotionVectors.hpp"

#include <cstring>

namespace rr = rendering::reconstruction;

extern "C" {

// Matrices are 16 floats, rows.  out: current uv, previous uv, motion,
// expected previous depth, valid (as 0 / 1) -- 8 floats.
void ref_motion_vector(const float* objectPosition, const float* curO2W, const float* prevO2W, const float* vpCur,
                       const float* vpPrev, const float* wvPrev, const float* jitterCur, const float* jitterPrev,
                       float* out) {
    rr::MotionVectorInput in{};
    in.objectPosition = {objectPosition[0], objectPosition[1], objectPosition[2]};
    std::memcpy(in.currentObjectToWorld.values.data(), curO2W, 64);
    std::memcpy(in.previousObjectToWorld.values.data(), prevO2W, 64);
    std::memcpy(in.currentViewProjection.values.data(), vpCur, 64);
    std::memcpy(in.previousViewProjection.values.data(), vpPrev, 64);
    std::memcpy(in.previousWorldToView.values.data(), wvPrev, 64);
    in.currentJitterUv = {jitterCur[0], jitterCur[1]};
    in.previousJitterUv = {jitterPrev[0], jitterPrev[1]};
    const rr::MotionVectorResult r = rr::ComputeMotionVector(in);
    const float o[8] = {r.currentUv.x, r.currentUv.y, r.previousUv.x, r.previousUv.y,
                        r.motion.x,    r.motion.y,    r.expectedPreviousLinearDepth, r.valid ? 1.0f : 0.0f};
    std::memcpy(out, o, sizeof o);
}

}  // extern "C"
