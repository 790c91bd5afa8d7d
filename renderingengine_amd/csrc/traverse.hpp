// Device traversal over the wide-node scene layout (rex_layout.hpp).
//
// Two query semantics, each bit-exact with the reference shader it restates:
//
//  * walk_*  -- the integrator's closest/any walk,
//    resources/shaders/integrators/pbr_software_traversal.hlsli:287-576:
//    left-first depth-first order, a node's box tested when the walk enters
//    it (so a right child is tested only after the left subtree finished),
//    closest ties to the lower (instance, primitive), an ignored
//    (instance, primitive) pair, shadow queries skipping non-casters and
//    stopping at the first accepted triangle.
//      walk_short_stack: pending right children wait on a per-thread stack
//        whose top kLdsDepth entries live in LDS (overflow goes to a private
//        array); the deferred box test is split into its best-independent
//        part (evaluated when pushed) and the distance-limit compare
//        (evaluated when popped), which equals testing at pop time because
//        the slab arithmetic does not depend on the current best.
//      walk_stackless: the reference's parent-pointer walk on wide nodes,
//        with (node, side) state and no stack at all.
//    Both visit the same nodes in the same order as the reference walk.
//
//  * trace_v1 -- the IGpuTraversalBackend batch tracer,
//    resources/shaders/traversal/software_trace_v1.hlsl:189-411: near-first
//    order with both children tested when the parent is popped, closest ties
//    to the lower traversal ID, explicit stack capacity.
//
// Slab test (software_common.hlsli:72-123) in branch-free form: 1/d is
// computed once per ray; a zero direction component gives 1/d = +-inf, which
// makes that axis either a no-op or a miss exactly as the reference's
// containment test does; the per-axis early exits are folded into one final
// compare because fmaxf/fminf make near non-decreasing and far
// non-increasing across axes.  Triangle test: software_common.hlsli:125-172.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rex_layout.hpp"

namespace rex {

struct RayF {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;  // 1 / d
    float tmin, tmax;
};

__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

__device__ __forceinline__ RayF make_ray(float ox, float oy, float oz, float tmin, float dx, float dy,
                                         float dz, float tmax) {
    RayF r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.dx = dx; r.dy = dy; r.dz = dz;
    r.ix = 1.0f / dx; r.iy = 1.0f / dy; r.iz = 1.0f / dz;
    r.tmin = tmin; r.tmax = tmax;
    return r;
}

// Best-independent part of SoftwareIntersectBounds: slab overlap, far > tMin
// and near < tMax.  intervalNear is returned for the distance-limit compare.
__device__ __forceinline__ bool slab_indep(const RayF& r, float mnx, float mny, float mnz, float mxx,
                                           float mxy, float mxz, float& intervalNear) {
    float an0 = (mnx - r.ox) * r.ix, af0 = (mxx - r.ox) * r.ix;
    float an1 = (mny - r.oy) * r.iy, af1 = (mxy - r.oy) * r.iy;
    float an2 = (mnz - r.oz) * r.iz, af2 = (mxz - r.oz) * r.iz;
    const bool s0 = an0 > af0, s1 = an1 > af1, s2 = an2 > af2;
    const float n0 = s0 ? af0 : an0, f0 = s0 ? an0 : af0;
    const float n1 = s1 ? af1 : an1, f1 = s1 ? an1 : af1;
    const float n2 = s2 ? af2 : an2, f2 = s2 ? an2 : af2;
    const float nearV = fmaxf(fmaxf(fmaxf(-__builtin_huge_valf(), n0), n1), n2);
    const float farV = fminf(fminf(fminf(__builtin_huge_valf(), f0), f1), f2);
    intervalNear = nearV;
    return !(farV < nearV) && farV > r.tmin && nearV < r.tmax;
}

__device__ __forceinline__ bool within_limit(float nearV, float best, bool found) {
    return found ? nearV <= best : nearV < best;
}

struct TriHit {
    float t, u, v;
    bool front;
};

// SoftwareIntersectTriangle with edges precomputed (bit-identical).
__device__ __forceinline__ bool tri_test(const RayF& r, const TriPos& tp, float best, bool found,
                                         TriHit& h) {
    const float v0x = tp.p0.x, v0y = tp.p0.y, v0z = tp.p0.z;
    const float e1x = tp.p0.w, e1y = tp.p1.x, e1z = tp.p1.y;
    const float e2x = tp.p1.z, e2y = tp.p1.w, e2z = tp.p2.x;
    // p = cross(d, e2)
    const float px = r.dy * e2z - r.dz * e2y;
    const float py = r.dz * e2x - r.dx * e2z;
    const float pz = r.dx * e2y - r.dy * e2x;
    const float det = e1x * px + e1y * py + e1z * pz;
    if (det == 0.0f || !__builtin_isfinite(det)) return false;
    const float inv = 1.0f / det;
    const float tx = r.ox - v0x, ty = r.oy - v0y, tz = r.oz - v0z;
    const float u = (tx * px + ty * py + tz * pz) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    // q = cross(translated, e1)
    const float qx = ty * e1z - tz * e1y;
    const float qy = tz * e1x - tx * e1z;
    const float qz = tx * e1y - ty * e1x;
    const float v = (r.dx * qx + r.dy * qy + r.dz * qz) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    const float t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    h.t = t; h.u = u; h.v = v; h.front = det > 0.0f;
    const bool withinRay = t < r.tmax;
    const bool lim = found ? t <= best : t < best;
    return __builtin_isfinite(t) && t > r.tmin && withinRay && lim;
}

__device__ __forceinline__ TriPos load_tri(const TriPos* __restrict__ tris, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(tris + i);
    const float4 a = p[0], b = p[1], c = p[2];
    TriPos t;
    t.p0 = {a.x, a.y, a.z, a.w};
    t.p1 = {b.x, b.y, b.z, b.w};
    t.p2 = {c.x, c.y, c.z, c.w};
    return t;
}

__device__ __forceinline__ WideNode load_wide(const WideNode* __restrict__ wide, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(wide + i);
    const float4 a = p[0], b = p[1], c = p[2];
    const uint4 l = reinterpret_cast<const uint4*>(wide + i)[3];
    WideNode n;
    n.a = {a.x, a.y, a.z, a.w};
    n.b = {b.x, b.y, b.z, b.w};
    n.c = {c.x, c.y, c.z, c.w};
    n.link = {l.x, l.y, l.z, l.w};
    return n;
}

// Per-thread traversal statistics (reference counter semantics).
struct WalkCounters {
    uint32_t nodeTests;      // loop iterations of the reference walk
    uint32_t boxTests;       // slab tests evaluated
    uint32_t triTests;
    uint32_t leafVisits;
    uint32_t leafPrims;
    uint32_t maxLeaf;
    uint32_t maxStack;
};

struct WalkResult {
    uint32_t result;  // 0 miss, 1 hit, 2 invalid
    float t, u, v;
    bool front;
    uint32_t tri;
};

// Per-thread stack: top kLds entries in LDS (entry-major so a wave's lanes
// touch consecutive banks), the rest in a private array.
template <int kLds, int kTotal>
struct ThreadStack {
    uint2* lds;       // base of this block's LDS stack area
    uint32_t stride;  // threads per block
    uint32_t lane;    // thread index in block
    uint2 spill[kTotal - kLds];

    __device__ __forceinline__ void put(uint32_t sp, uint2 e) {
        if (sp < (uint32_t)kLds) {
            lds[sp * stride + lane] = e;
        } else {
            spill[sp - kLds] = e;
        }
    }
    __device__ __forceinline__ uint2 get(uint32_t sp) {
        if (sp < (uint32_t)kLds) return lds[sp * stride + lane];
        return spill[sp - kLds];
    }
};

// Tests the triangles of a leaf reference for the walk semantics.
// Returns true when a shadow query found an occluder.
template <bool kShadow, bool kCount>
__device__ __forceinline__ bool walk_leaf(const SceneView& s, const RayF& r, uint32_t ref,
                                          uint32_t ignInst, uint32_t ignPrim, bool& found,
                                          float& best, WalkResult& res, uint32_t& bestInst,
                                          uint32_t& bestPrim, WalkCounters& c) {
    const uint32_t first = ref & kLeafFirstMask;
    const uint32_t count = ((ref >> 28) & 7u) + 1u;
    if (kCount) {
        c.leafVisits++;
        c.leafPrims += count;
        c.maxLeaf = count > c.maxLeaf ? count : c.maxLeaf;
    }
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t ti = first + k;
        const TriPos tp = load_tri(s.tris, ti);
        const uint32_t inst = f2u(tp.p2.y), prim = f2u(tp.p2.z), flags = f2u(tp.p2.w);
        if (inst == ignInst && prim == ignPrim) continue;
        if (kShadow && (flags & kTriCastsShadow) == 0u) continue;
        if (kCount) c.triTests++;
        TriHit h;
        if (!tri_test(r, tp, best, found, h)) continue;
        if (!h.front && (flags & kTriDoubleSided) == 0u) continue;
        if (kShadow) {
            res.result = 1;
            return true;
        }
        const bool eqLower = h.t == best && (inst < bestInst || (inst == bestInst && prim < bestPrim));
        if (!found || h.t < best || eqLower) {
            found = true;
            best = h.t;
            res.t = h.t; res.u = h.u; res.v = h.v; res.front = h.front; res.tri = ti;
            bestInst = inst;
            bestPrim = prim;
        }
    }
    return false;
}

// Left-first walk with the LDS short stack.  The root box is tested first
// (the reference enters node 0 from its invalid parent).
template <bool kShadow, bool kCount, int kLds, int kTotal>
__device__ __forceinline__ WalkResult walk_short_stack(const SceneView& s, const RayF& r, uint32_t ignInst,
                                                       uint32_t ignPrim, ThreadStack<kLds, kTotal>& st,
                                                       WalkCounters& c) {
    WalkResult res;
    res.result = 0; res.t = r.tmax; res.u = 0.0f; res.v = 0.0f; res.front = false; res.tri = kInvalid;
    bool found = false;
    float best = r.tmax;
    uint32_t bestInst = kInvalid, bestPrim = kInvalid;
    float rn;
    if (kCount) { c.nodeTests++; c.boxTests++; }
    if (!slab_indep(r, s.rootMin[0], s.rootMin[1], s.rootMin[2], s.rootMax[0], s.rootMax[1], s.rootMax[2], rn) ||
        !within_limit(rn, best, found)) {
        return res;
    }
    uint32_t ref = s.rootRef;
    uint32_t sp = 0;
    for (;;) {
        if (is_leaf_ref(ref)) {
            if (walk_leaf<kShadow, kCount>(s, r, ref, ignInst, ignPrim, found, best, res, bestInst, bestPrim, c)) {
                return res;
            }
        } else {
            const WideNode n = load_wide(s.wide, ref);
            float ln, rnear;
            const bool li = slab_indep(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, ln);
            const bool ri = slab_indep(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, rnear);
            // The reference tests both children of an entered interior node
            // (the right one after the left subtree) and loops over the node
            // three times (enter, back from left, back from right).
            if (kCount) { c.boxTests += 2; c.nodeTests += 2 + 2; }
            const bool lp = li && within_limit(ln, best, found);
            if (lp) {
                if (ri) {
                    st.put(sp, make_uint2(n.link.y, f2u(rnear)));
                    ++sp;
                    if (kCount && sp > c.maxStack) c.maxStack = sp;
                }
                ref = n.link.x;
                continue;
            }
            // Left culled: the reference returns to this node and tests the
            // right child at once, with the same best.
            if (ri && within_limit(rnear, best, found)) {
                ref = n.link.y;
                continue;
            }
        }
        // Pop the next pending right child whose deferred compare passes.
        bool next = false;
        while (sp != 0) {
            --sp;
            const uint2 e = st.get(sp);
            if (within_limit(u2f(e.y), best, found)) {
                ref = e.x;
                next = true;
                break;
            }
        }
        if (!next) break;
    }
    if (found) res.result = 1;
    return res;
}

// The reference's parent-pointer walk on wide nodes: state (node, side) with
// side 0 = examine left child, 1 = examine right child, 2 = ascend.
template <bool kShadow, bool kCount>
__device__ __forceinline__ WalkResult walk_stackless(const SceneView& s, const RayF& r, uint32_t ignInst,
                                                     uint32_t ignPrim, WalkCounters& c) {
    WalkResult res;
    res.result = 0; res.t = r.tmax; res.u = 0.0f; res.v = 0.0f; res.front = false; res.tri = kInvalid;
    bool found = false;
    float best = r.tmax;
    uint32_t bestInst = kInvalid, bestPrim = kInvalid;
    float rn;
    if (kCount) { c.nodeTests++; c.boxTests++; }
    if (!slab_indep(r, s.rootMin[0], s.rootMin[1], s.rootMin[2], s.rootMax[0], s.rootMax[1], s.rootMax[2], rn) ||
        !within_limit(rn, best, found)) {
        return res;
    }
    if (is_leaf_ref(s.rootRef)) {
        walk_leaf<kShadow, kCount>(s, r, s.rootRef, ignInst, ignPrim, found, best, res, bestInst, bestPrim, c);
        if (kShadow) return res;
        if (found) res.result = 1;
        return res;
    }
    uint32_t w = s.rootRef;
    uint32_t side = 0;
    // Each wide node is entered once from above and returned to at most twice.
    uint32_t guard = 3u * s.nodeCount + 3u;
    while (guard-- != 0u) {
        const WideNode n = load_wide(s.wide, w);
        if (side == 0u) {
            float ln;
            if (kCount) { c.nodeTests += 3; c.boxTests++; }
            if (slab_indep(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, ln) && within_limit(ln, best, found)) {
                if (is_leaf_ref(n.link.x)) {
                    if (walk_leaf<kShadow, kCount>(s, r, n.link.x, ignInst, ignPrim, found, best, res, bestInst,
                                                   bestPrim, c)) {
                        return res;
                    }
                } else {
                    w = n.link.x;
                    continue;
                }
            }
            side = 1u;
        }
        if (side == 1u) {
            float rnear;
            if (kCount) { c.nodeTests++; c.boxTests++; }
            if (slab_indep(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, rnear) && within_limit(rnear, best, found)) {
                if (is_leaf_ref(n.link.y)) {
                    if (walk_leaf<kShadow, kCount>(s, r, n.link.y, ignInst, ignPrim, found, best, res, bestInst,
                                                   bestPrim, c)) {
                        return res;
                    }
                } else {
                    w = n.link.y;
                    side = 0u;
                    continue;
                }
            }
        }
        // side 2: ascend.
        const uint32_t parent = n.link.w & 0x7fffffffu;
        if (n.link.w == kInvalid) break;
        side = (n.link.w >> 31) != 0u ? 2u : 1u;
        w = parent;
        if (side == 2u) {
            // Came from the right child: keep ascending without re-testing.
            for (;;) {
                const uint32_t pl = s.wide[w].link.w;
                if (pl == kInvalid) { guard = 0; break; }
                side = (pl >> 31) != 0u ? 2u : 1u;
                w = pl & 0x7fffffffu;
                if (side == 1u) break;
            }
            if (guard == 0) break;
        }
    }
    if (found) res.result = 1;
    return res;
}

}  // namespace rex
