// Device traversal primitives over the wide-node scene layout (rex_layout.hpp).
//
// Slab test (SoftwareIntersectBounds, software_common.hlsli:72-123) in
// branch-free form.  1/d is computed once per ray; a zero direction
// component (either sign) gets 1/d = +inf, which turns that axis into either
// a no-op or a miss exactly as the reference's containment test does:
//   o strictly inside the slab -> (-inf, +inf); o on a slab plane -> one
//   NaN, which the swap test leaves as the "near" value and fmaxf/fminf then
//   ignore; o outside -> (+inf, +inf) or (-inf, -inf), an empty interval.
// The per-axis early exits are folded into one final compare because fmaxf /
// fminf make near non-decreasing and far non-increasing across axes.
//
// When every component of 1/d is finite (all but axis-aligned or denormal
// directions), (b - o) * (1/d) cannot be NaN, so the reference's
// compare-and-swap equals plain fminf / fmaxf; slab_fast uses that form.  A
// ray with an infinite component keeps the exact swap form (slab_exact).
//
// Triangle test: SoftwareIntersectTriangle (software_common.hlsli:125-172)
// with the edges precomputed on the host (bit-identical subtractions).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rex_layout.hpp"

namespace rex {

struct RayF {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;  // 1 / d, +inf for a zero component
    float tmin, tmax;
    bool exact;        // some component of 1/d is infinite
};

__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

__device__ __forceinline__ float safe_inv(float d) {
    return d == 0.0f ? __builtin_huge_valf() : 1.0f / d;
}

__device__ __forceinline__ RayF make_ray(float ox, float oy, float oz, float tmin, float dx, float dy,
                                         float dz, float tmax) {
    RayF r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.dx = dx; r.dy = dy; r.dz = dz;
    r.ix = safe_inv(dx); r.iy = safe_inv(dy); r.iz = safe_inv(dz);
    r.tmin = tmin; r.tmax = tmax;
    r.exact = !(__builtin_isfinite(r.ix) && __builtin_isfinite(r.iy) && __builtin_isfinite(r.iz));
    return r;
}

// Best-independent part of SoftwareIntersectBounds (slab overlap, far > tMin,
// near < tMax) in the reference's swap form; intervalNear is returned.
__device__ __forceinline__ bool slab_exact(const RayF& r, float mnx, float mny, float mnz, float mxx,
                                           float mxy, float mxz, float& intervalNear) {
    const float an0 = (mnx - r.ox) * r.ix, af0 = (mxx - r.ox) * r.ix;
    const float an1 = (mny - r.oy) * r.iy, af1 = (mxy - r.oy) * r.iy;
    const float an2 = (mnz - r.oz) * r.iz, af2 = (mxz - r.oz) * r.iz;
    const bool s0 = an0 > af0, s1 = an1 > af1, s2 = an2 > af2;
    const float n0 = s0 ? af0 : an0, f0 = s0 ? an0 : af0;
    const float n1 = s1 ? af1 : an1, f1 = s1 ? an1 : af1;
    const float n2 = s2 ? af2 : an2, f2 = s2 ? an2 : af2;
    const float nearV = fmaxf(fmaxf(fmaxf(-__builtin_huge_valf(), n0), n1), n2);
    const float farV = fminf(fminf(fminf(__builtin_huge_valf(), f0), f1), f2);
    intervalNear = nearV;
    return !(farV < nearV) && farV > r.tmin && nearV < r.tmax;
}

// Same result as slab_exact for rays whose 1/d is finite in every component.
__device__ __forceinline__ bool slab_fast(const RayF& r, float mnx, float mny, float mnz, float mxx,
                                          float mxy, float mxz, float& intervalNear) {
    const float an0 = (mnx - r.ox) * r.ix, af0 = (mxx - r.ox) * r.ix;
    const float an1 = (mny - r.oy) * r.iy, af1 = (mxy - r.oy) * r.iy;
    const float an2 = (mnz - r.oz) * r.iz, af2 = (mxz - r.oz) * r.iz;
    const float nearV = fmaxf(fmaxf(fminf(an0, af0), fminf(an1, af1)), fminf(an2, af2));
    const float farV = fminf(fminf(fmaxf(an0, af0), fmaxf(an1, af1)), fmaxf(an2, af2));
    intervalNear = nearV;
    return !(farV < nearV) && farV > r.tmin && nearV < r.tmax;
}

// slab_fast with the planes pre-ordered per ray: n* are the entry planes
// (min for a positive direction component, max for a negative one) and f*
// the exit planes.  For finite 1/d, (n - o) * inv <= (f - o) * inv on every
// axis, so fminf / fmaxf of the pair return exactly these two values and the
// result equals slab_fast's.
__device__ __forceinline__ bool slab_ordered(const RayF& r, float nx, float ny, float nz, float fx, float fy,
                                             float fz, float& intervalNear) {
    const float n0 = (nx - r.ox) * r.ix, f0 = (fx - r.ox) * r.ix;
    const float n1 = (ny - r.oy) * r.iy, f1 = (fy - r.oy) * r.iy;
    const float n2 = (nz - r.oz) * r.iz, f2 = (fz - r.oz) * r.iz;
    const float nearV = fmaxf(fmaxf(n0, n1), n2);
    const float farV = fminf(fminf(f0, f1), f2);
    intervalNear = nearV;
    return !(farV < nearV) && farV > r.tmin && nearV < r.tmax;
}

template <bool kExact>
__device__ __forceinline__ bool slab(const RayF& r, float mnx, float mny, float mnz, float mxx, float mxy,
                                     float mxz, float& intervalNear) {
    return kExact ? slab_exact(r, mnx, mny, mnz, mxx, mxy, mxz, intervalNear)
                  : slab_fast(r, mnx, mny, mnz, mxx, mxy, mxz, intervalNear);
}

// The tMin / tMax part of the reference's slab test on a triangle's exact
// leaf box, for the compressed walk's candidate hits.  The walk may have
// replaced an infinite 1/d by a large finite one; the exact test recomputes
// the reference's inverse for such rays.  Boxes are nested and the rounding
// of (b - o) * inv is monotone in b, so an ancestor box that the exact test
// rejects against tMin / tMax implies a rejected leaf box: passing the leaf
// box alone admits exactly the candidates the exact walks can reach.
__device__ __forceinline__ bool leaf_box_admits(const RayF& r, const LeafBox* __restrict__ boxes, uint32_t tri) {
    const float4* p = reinterpret_cast<const float4*>(boxes + tri);
    const float4 lo = p[0], hi = p[1];
    RayF e = r;
    if (r.exact) {
        e.ix = safe_inv(r.dx);
        e.iy = safe_inv(r.dy);
        e.iz = safe_inv(r.dz);
    }
    float nearV;
    return slab_exact(e, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, nearV);
}

// leaf_box_admits for the compressed walk, whose rays carry a large finite
// inverse where 1/d is infinite.  Bit k of infMask marks such a component;
// its exact inverse is +inf for a zero component and inf with the sign of d
// for a nonzero one (1/d overflows), so no division is needed to restore it.
__device__ __forceinline__ bool leaf_box_admits_q(const RayF& r, uint32_t infMask, const LeafBox* __restrict__ boxes,
                                                  uint32_t tri) {
    const float4* p = reinterpret_cast<const float4*>(boxes + tri);
    const float4 lo = p[0], hi = p[1];
    RayF e = r;
    const float inf = __builtin_huge_valf();
    if (infMask & 1u) e.ix = r.dx == 0.0f ? inf : __builtin_copysignf(inf, r.dx);
    if (infMask & 2u) e.iy = r.dy == 0.0f ? inf : __builtin_copysignf(inf, r.dy);
    if (infMask & 4u) e.iz = r.dz == 0.0f ? inf : __builtin_copysignf(inf, r.dz);
    float nearV;
    return slab_exact(e, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, nearV);
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

// tri_test without early exits: every lane evaluates the whole test and one
// predicate combines the reference's checks (same operations on the same
// operands, so the same bits).  A wave whose lanes disagree at an early exit
// runs every stage anyway; the flat form saves the exec-mask bookkeeping and
// the register copies of the branches.  `lim` is walk_limit(best, found).
__device__ __forceinline__ bool tri_test_flat(const RayF& r, const TriPos& tp, float lim, TriHit& h) {
    const float v0x = tp.p0.x, v0y = tp.p0.y, v0z = tp.p0.z;
    const float e1x = tp.p0.w, e1y = tp.p1.x, e1z = tp.p1.y;
    const float e2x = tp.p1.z, e2y = tp.p1.w, e2z = tp.p2.x;
    const float px = r.dy * e2z - r.dz * e2y;
    const float py = r.dz * e2x - r.dx * e2z;
    const float pz = r.dx * e2y - r.dy * e2x;
    const float det = e1x * px + e1y * py + e1z * pz;
    const float inv = 1.0f / det;
    const float tx = r.ox - v0x, ty = r.oy - v0y, tz = r.oz - v0z;
    const float u = (tx * px + ty * py + tz * pz) * inv;
    const float qx = ty * e1z - tz * e1y;
    const float qy = tz * e1x - tx * e1z;
    const float qz = tx * e1y - ty * e1x;
    const float v = (r.dx * qx + r.dy * qy + r.dz * qz) * inv;
    const float t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    h.t = t; h.u = u; h.v = v; h.front = det > 0.0f;
    // det != 0 and finite; u in [0, 1]; v >= 0 and u + v <= 1; t finite,
    // inside (tMin, tMax) and within the best hit.
    return (det != 0.0f) & __builtin_isfinite(det) & (u >= 0.0f) & (u <= 1.0f) & (v >= 0.0f) &
           (u + v <= 1.0f) & __builtin_isfinite(t) & (t > r.tmin) & (t < r.tmax) & (t < lim);
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

__device__ __forceinline__ Wide4 load_wide4(const Wide4* __restrict__ wide, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(wide + i);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5];
    const uint4 r = reinterpret_cast<const uint4*>(wide + i)[6];
    Wide4 n;
    n.lox = {a.x, a.y, a.z, a.w};
    n.loy = {b.x, b.y, b.z, b.w};
    n.loz = {c.x, c.y, c.z, c.w};
    n.hix = {d.x, d.y, d.z, d.w};
    n.hiy = {e.x, e.y, e.z, e.w};
    n.hiz = {f.x, f.y, f.z, f.w};
    n.ref = {r.x, r.y, r.z, r.w};
    n.bin = {kInvalid, kInvalid, kInvalid, kInvalid};
    return n;
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

// Hit normalization of the traversal results: each component divided by the
// length (both correctly rounded).
__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
    const float len = sqrtf(x * x + y * y + z * z);
    x = x / len;
    y = y / len;
    z = z / len;
}

// Geometric and shading normal of a hit (pbr_software_traversal.hlsli:82-118;
// software_trace_v1.hlsl:382-393): geometric = normalize(cross(e1, e2)),
// shading = normalize(w0*n0 + u*n1 + v*n2), both negated on a back face.
__device__ __forceinline__ void hit_normals(const SceneView& s, uint32_t tri, float u, float v, bool front,
                                            float gn[3], float sn[3], TriAttr& attr) {
    const TriPos tp = load_tri(s.tris, tri);
    const float e1x = tp.p0.w, e1y = tp.p1.x, e1z = tp.p1.y;
    const float e2x = tp.p1.z, e2y = tp.p1.w, e2z = tp.p2.x;
    gn[0] = e1y * e2z - e1z * e2y;
    gn[1] = e1z * e2x - e1x * e2z;
    gn[2] = e1x * e2y - e1y * e2x;
    normalize3(gn[0], gn[1], gn[2]);
    attr = s.attrs[tri];
    const float w0 = 1.0f - u - v;
    sn[0] = attr.n0.x * w0 + attr.n1.x * u + attr.n2.x * v;
    sn[1] = attr.n0.y * w0 + attr.n1.y * u + attr.n2.y * v;
    sn[2] = attr.n0.z * w0 + attr.n1.z * u + attr.n2.z * v;
    normalize3(sn[0], sn[1], sn[2]);
    if (!front) {
        for (int k = 0; k < 3; ++k) {
            gn[k] = -gn[k];
            sn[k] = -sn[k];
        }
    }
}

// Per-thread traversal statistics (reference counter semantics).
struct WalkCounters {
    uint32_t nodeTests;  // loop iterations of the reference walk / node tests of the batch tracer
    uint32_t boxTests;   // slab tests evaluated
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

__device__ __forceinline__ uint32_t leaf_first(uint32_t ref) { return ref & kLeafFirstMask; }
__device__ __forceinline__ uint32_t leaf_count(uint32_t ref) { return ((ref >> 28) & 7u) + 1u; }

// Tests the triangles of a leaf reference for the walk semantics
// (pbr_software_traversal.hlsli:385-497).  Returns true when a shadow query
// found an occluder.
template <bool kShadow, bool kCount>
__device__ __forceinline__ bool walk_leaf(const SceneView& s, const RayF& r, uint32_t ref,
                                          uint32_t ignInst, uint32_t ignPrim, bool& found,
                                          float& best, WalkResult& res, uint32_t& bestInst,
                                          uint32_t& bestPrim, WalkCounters& c) {
    const uint32_t first = leaf_first(ref);
    const uint32_t count = leaf_count(ref);
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

// The reference's parent-pointer walk (pbr_software_traversal.hlsli:287-542)
// on wide nodes: state (node, side) with side 0 = examine left child, 1 =
// examine right child, then ascend through the parent link, whose top bit
// says whether the node is its parent's right child.  No stack at all.  The
// caller has already entered the root (its box test passed).
template <bool kShadow, bool kCount>
__device__ __forceinline__ WalkResult walk_stackless(const SceneView& s, const RayF& r, uint32_t ignInst,
                                                     uint32_t ignPrim, WalkCounters& c) {
    WalkResult res;
    res.result = 0; res.t = r.tmax; res.u = 0.0f; res.v = 0.0f; res.front = false; res.tri = kInvalid;
    bool found = false;
    float best = r.tmax;
    uint32_t bestInst = kInvalid, bestPrim = kInvalid;
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
            if (slab_exact(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, ln) && within_limit(ln, best, found)) {
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
            if (slab_exact(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, rnear) && within_limit(rnear, best, found)) {
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
        // Ascend: back to side 1 of the parent when this node is its left
        // child, otherwise keep climbing.
        if (n.link.w == kInvalid) break;
        bool more = false;
        uint32_t link = n.link.w;
        for (;;) {
            const bool isRight = (link >> 31) != 0u;
            w = link & 0x7fffffffu;
            if (!isRight) {
                side = 1u;
                more = true;
                break;
            }
            link = s.wide[w].link.w;
            if (link == kInvalid) break;
        }
        if (!more) break;
    }
    if (found) res.result = 1;
    return res;
}

}  // namespace rex
