// Motion of a primary surface, shared by the Resolve export of the motion
// inputs, MotionVectorsCS and ReSTIR's temporal reuse under object motion:
//
//   surface_motion_input  PbrMotionInputV2 of a published surface
//                         (resources/shaders/integrators/pbr_reconstruction_export_v2.hlsli:90-129)
//   surface_motion        MotionVectorsCS main (reconstruction/shaders/MotionVectorsCS.hlsl:35-90)
//
// Arithmetic operand for operand like the oracle's restatement (no
// contraction; a 4-term dot is ((x*x' + y*y') + z*z') + w*w').
#pragma once

#include <hip/hip_runtime.h>

#include "restir.hpp"

namespace rex {

// The world position in the object space of the first instance with the
// surface's stable instance ID (binary search over (ID, first index) pairs
// sorted by ID); an unknown ID gives the instance count as both indices.
// Scenes without an instance table (map NULL): world position, index 0.
// Invalid surfaces give a zero record.
__device__ __forceinline__ MotionInputV2 surface_motion_input(const InstanceXform* __restrict__ instances,
                                                              const uint2* __restrict__ map, uint32_t instanceCount,
                                                              uint32_t mapCount, const SurfaceV2& sp) {
    MotionInputV2 m = {0.0f, 0.0f, 0.0f, 0u, 0u, 0u, 0u, 0u};
    if ((sp.identity.w & kSurfaceValid) == 0u) return m;
    if (map == nullptr) {
        m.px = sp.posDepth.x;
        m.py = sp.posDepth.y;
        m.pz = sp.posDepth.z;
        return m;
    }
    const uint32_t id = sp.identity.y;
    uint32_t lo = 0u, hi = mapCount;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (map[mid].x < id) lo = mid + 1u;
        else hi = mid;
    }
    if (lo < mapCount && map[lo].x == id) {
        const uint32_t t = map[lo].y;
        const InstanceXform X = instances[t];
        const float wx = sp.posDepth.x, wy = sp.posDepth.y, wz = sp.posDepth.z;
        m.px = X.row0.x * wx + X.row0.y * wy + X.row0.z * wz + X.row0.w;
        m.py = X.row1.x * wx + X.row1.y * wy + X.row1.z * wz + X.row1.w;
        m.pz = X.row2.x * wx + X.row2.y * wy + X.row2.z * wz + X.row2.w;
        m.current = m.previous = t;
    } else {
        m.current = m.previous = instanceCount;
    }
    return m;
}

__device__ __forceinline__ float motion_dot4(const float4 a, const float4 b) {
    return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

__device__ __forceinline__ float4 motion_mul(const float4* __restrict__ rows, const float4 v) {
    return make_float4(motion_dot4(rows[0], v), motion_dot4(rows[1], v), motion_dot4(rows[2], v),
                       motion_dot4(rows[3], v));
}

__device__ __forceinline__ bool motion_finite4(const float4 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z) && __builtin_isfinite(v.w);
}

// The motion record MotionVectorsCS writes for a motion input: current and
// previous UV (with jitter), previous linear depth and the valid flag.
struct SurfaceMotion {
    float cu, cv, pu, pv, depth;
    bool valid;
};

// Returns false where MotionVectorsCS clears the record (out-of-range
// transform indices, a non-finite object position, a point behind either
// camera or non-finite products).  cur / prev: object-to-world rows, four
// float4 per transform; vpCur / vpPrev / wvPrev: four rows each.
__device__ __forceinline__ bool surface_motion(const MotionInputV2& in, const float4* __restrict__ cur,
                                               uint32_t curCount, const float4* __restrict__ prev, uint32_t prevCount,
                                               const float4* vpCur, const float4* vpPrev, const float4* wvPrev,
                                               float jcx, float jcy, float jpx, float jpy, SurfaceMotion& o) {
    if (in.current >= curCount || in.previous >= prevCount || !__builtin_isfinite(in.px) ||
        !__builtin_isfinite(in.py) || !__builtin_isfinite(in.pz)) {
        return false;
    }
    const float4 obj = make_float4(in.px, in.py, in.pz, 1.0f);
    const float4 curWorld = motion_mul(cur + 4u * in.current, obj);
    const float4 prevWorld = motion_mul(prev + 4u * in.previous, obj);
    const float4 curClip = motion_mul(vpCur, curWorld);
    const float4 prevClip = motion_mul(vpPrev, prevWorld);
    const float4 prevView = motion_mul(wvPrev, prevWorld);
    if (!(curClip.w > 1.0e-8f && prevClip.w > 1.0e-8f && motion_finite4(curWorld) && motion_finite4(prevWorld) &&
          motion_finite4(curClip) && motion_finite4(prevClip) && motion_finite4(prevView))) {
        return false;
    }
    const float cnx = curClip.x / curClip.w, cny = curClip.y / curClip.w;
    const float pnx = prevClip.x / prevClip.w, pny = prevClip.y / prevClip.w;
    o.cu = cnx * 0.5f + 0.5f + jcx;
    o.cv = 0.5f - cny * 0.5f + jcy;
    o.pu = pnx * 0.5f + 0.5f + jpx;
    o.pv = 0.5f - pny * 0.5f + jpy;
    o.depth = -prevView.z;
    o.valid = __builtin_isfinite(o.cu) && __builtin_isfinite(o.cv) && __builtin_isfinite(o.pu) &&
              __builtin_isfinite(o.pv) && __builtin_isfinite(o.depth) && o.depth >= 0.0f;
    return true;
}

}  // namespace rex
