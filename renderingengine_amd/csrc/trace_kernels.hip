// Traversal kernels.
//
// k_trace: a persistent kernel for both query semantics -- the integrator's
// walk (pbr_software_traversal.hlsli, rex_trace_walk) and the batch tracer
// (software_trace_v1.hlsl, rex_trace_closest / rex_trace_any).
//
//  * Persistent lanes with refill: the grid holds exactly the resident waves;
//    every iteration, lanes without a ray take the next ones from a global
//    counter (one atomic per wave), so a wave never idles behind its longest
//    ray and the grid drains when the counter passes the batch size.
//  * One unit of work per lane per iteration: either one wide node (both
//    child boxes tested) or one triangle of the current leaf.  A wave thus
//    pays for at most one node test and one triangle test per iteration
//    instead of a whole leaf whenever any lane sits in one.
//  * Traversal stack: the first lds_entries() entries (the bottom of the
//    stack) in LDS, entry-major so a wave's lanes touch consecutive 8-byte
//    words; deeper entries go to a per-lane global overflow region (no
//    scratch memory).
//
// Orders: left-first (the reference walk: right child's box compare deferred
// until the left subtree is done, evaluated from its stored interval near)
// and near-first (the batch tracer's order; for the walk semantics this is
// the build's faster production order).  Both are bit-exact with the oracle
// walking the same order.
//
// k_walk_stackless: the reference's parent-pointer walk, one ray per thread.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rex_internal.hpp"
#include "traverse.hpp"

namespace rex {

constexpr int kBlock = 128;
constexpr uint32_t kChunk = 64;      // rays claimed per wave atomic
#ifndef REX_REFILL_MIN
#define REX_REFILL_MIN 8
#endif
constexpr uint32_t kRefillMin = REX_REFILL_MIN;  // idle lanes that trigger a refill pass

// Query semantics: the walk API, the batch tracer, and the integrator's two
// queue stages -- closest hits of the ray queue written as compact QueueHit
// records (IntersectCS), and shadow any-hits whose unoccluded contributions
// are added to the path state in place (TraceShadowCS, Physical).  The
// mixed semantic runs one bounce's shadow queue and the next bounce's ray
// queue in one launch (indices below the ray count are rays, the rest shadow
// items), so the device is not drained between the two.
// kSemVis: visibility queries on ABI-v1 rays (ReSTIR's winner and reference
// queries): any-hit over the compressed walk, writing only the hit record's
// metadata words (ray ID, hit kind, 0, path ID), which is all the ReSTIR
// passes read.
enum : int { kSemWalk = 0, kSemV1 = 1, kSemQueue = 2, kSemShadow = 3, kSemMixed = 4, kSemVis = 5 };
enum : int { kOrderLeftFirst = 0, kOrderNearFirst = 1, kOrderWide = 2, kOrderWideQ = 3 };

// The production instantiations (queue, shadow and mixed semantics over the
// compressed walk, no counters) fit in 72 VGPRs without spilling, so they are
// built for 7 waves per SIMD: more waves in flight hide more of the node and
// triangle fetch latency.  (The walk and counting instantiations would spill.)
template <int kSem, int kOrder, bool kAny, bool kCount>
constexpr bool dense_trace() {
    return !kCount && kOrder == kOrderWideQ &&
           (kSem == kSemQueue || kSem == kSemMixed || kSem == kSemShadow || kSem == kSemVis);
}

// Stack entries kept in LDS per lane (8 B each; the rest of the stack spills
// to the lane's global overflow region).  A CU has 160 KB of LDS: the dense
// kernels (7-8 waves per SIMD, 14-16 blocks of 128 lanes) fit 10 entries;
// the other compressed-walk kernels run at most 6 waves per SIMD and fit
// 12, which keeps more of their deeper stacks out of global memory.
#ifndef REX_DENSE_LDS
#define REX_DENSE_LDS 10
#endif
template <int kSem, int kOrder, bool kAny, bool kCount>
constexpr int lds_entries() {
    return dense_trace<kSem, kOrder, kAny, kCount>() ? REX_DENSE_LDS : (kOrder == kOrderWideQ ? 12 : 8);
}
constexpr bool is_wide(int order) { return order == kOrderWide || order == kOrderWideQ; }

struct TraceParams {
    SceneView s;
    const float* rays;          // walk: (n, 8) = (o, tMin, d, tMax); queue: ray queue (2 x F4 per ray)
    const uint32_t* keys;       // walk: (n, 2) ignored (instance, primitive)
    const rex_ray_v1* raysV1;   // batch tracer
    rex_walk_hit* hitsWalk;
    rex_hit_v1* hitsV1;
    QueueHit* queueHits;        // queue: one per ray-queue slot
    const ShadowItem* shadows;  // shadow queue
    PathState* paths;           // shadow: accumulation target
    const uint32_t* countPtr;   // queue / shadow: item count in device memory (mixed: ray count)
    const uint32_t* countPtr2;  // mixed: shadow item count
    uint32_t count;
    uint32_t stackCapacity;     // batch tracer: 1..64
    uint32_t* workCounter;      // zeroed before launch
    uint2* overflow;            // overflowStride entries per launched thread
    uint32_t overflowStride;
    uint64_t* counters;         // optional
    // Batched jobs over an arena of scenes (queue semantics, may be NULL):
    // the scene of each ray-queue / shadow-queue item and the scenes' roots.
    const uint32_t* raySceneOf;
    const uint32_t* shadowSceneOf;
    const SceneRoot* roots;
};

// Adds per-thread counters into the global uint64 counter array with one
// atomic per wave per slot.
__device__ __forceinline__ void flush_counters(uint64_t* __restrict__ out, const WalkCounters& c,
                                               uint32_t invalidRay, uint32_t invalidHit, uint32_t overflow) {
    uint64_t v[REX_COUNTER_COUNT] = {c.nodeTests, c.triTests, overflow, invalidRay, invalidHit,
                                     0, c.leafVisits, c.leafPrims, 0, c.boxTests};
    uint32_t mx[2] = {c.maxStack, c.maxLeaf};
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < (int)REX_COUNTER_COUNT; ++k) v[k] += __shfl_xor(v[k], off, 64);
        for (int k = 0; k < 2; ++k) {
            const uint32_t o = __shfl_xor(mx[k], off, 64);
            mx[k] = o > mx[k] ? o : mx[k];
        }
    }
    if ((threadIdx.x & 63u) == 0u) {
        for (int k = 0; k < (int)REX_COUNTER_COUNT; ++k) {
            if (k == REX_COUNTER_MAX_STACK || k == REX_COUNTER_MAX_LEAF_OCCUPANCY) continue;
            if (v[k] != 0u) atomicAdd(reinterpret_cast<unsigned long long*>(out + k), (unsigned long long)v[k]);
        }
        atomicMax(reinterpret_cast<unsigned long long*>(out + REX_COUNTER_MAX_STACK), (unsigned long long)mx[0]);
        atomicMax(reinterpret_cast<unsigned long long*>(out + REX_COUNTER_MAX_LEAF_OCCUPANCY),
                  (unsigned long long)mx[1]);
    }
}

// software_common.hlsli:63-70
__device__ __forceinline__ bool walk_ray_valid(float ox, float oy, float oz, float dx, float dy, float dz,
                                               float tmin, float tmax) {
    const float len2 = dx * dx + dy * dy + dz * dz;
    return __builtin_isfinite(ox) && __builtin_isfinite(oy) && __builtin_isfinite(oz) &&
           __builtin_isfinite(dx) && __builtin_isfinite(dy) && __builtin_isfinite(dz) &&
           __builtin_isfinite(len2) && fabsf(len2 - 1.0f) <= 1.0e-4f && !__builtin_isnan(tmin) &&
           !__builtin_isnan(tmax) && tmin < tmax;
}

// software_trace_v1.hlsl:67-75
__device__ __forceinline__ bool v1_ray_valid(const float4& a, const float4& b, uint32_t mask) {
    const float len2 = b.x * b.x + b.y * b.y + b.z * b.z;
    return __builtin_isfinite(a.x) && __builtin_isfinite(a.y) && __builtin_isfinite(a.z) &&
           __builtin_isfinite(a.w) && __builtin_isfinite(b.x) && __builtin_isfinite(b.y) &&
           __builtin_isfinite(b.z) && __builtin_isfinite(b.w) && __builtin_isfinite(len2) &&
           fabsf(len2 - 1.0f) <= 1.0e-4f && a.w >= 0.0f && a.w < b.w && (mask & ~0xffu) == 0u;
}

__device__ __forceinline__ void v1_terminal(rex_hit_v1* __restrict__ out, float tMax, uint32_t rayId,
                                            uint32_t pathId, uint32_t kind) {
    rex_hit_v1 h = {};
    h.positionT[3] = tMax;
    h.ids[0] = h.ids[1] = h.ids[2] = h.ids[3] = kInvalid;
    h.metadata[0] = rayId;
    h.metadata[1] = kind;
    h.metadata[2] = 0u;
    h.metadata[3] = pathId;
    *out = h;
}

// Per-lane traversal state of the persistent kernel.
struct Lane {
    bool active;
    uint32_t idx;
    RayF r;
    uint32_t k0, k1;  // walk: ignored (instance, primitive); batch: (ray ID, path ID)
    bool found;
    float best, bu, bv;
    bool bf;
    uint32_t bt;
    uint32_t ref;       // interior node to process when leafLeft == 0
    uint32_t leafNext;  // next triangle of the current leaf
    uint32_t leafLeft;  // triangles left in the current leaf
    uint32_t sp;
    uint32_t visited;   // batch tracer: node pops (software_trace_v1.hlsl:283)
    bool shadow;        // mixed: this lane's item is a shadow query
    uint32_t planes;    // wide walk: float4 indices of the entry planes, x | y << 2 | z << 5
    bool hold;          // compressed walk: a second leaf waits on the stack behind the pending one
};

// The LDS part of the stack is addressed through an LDS-qualified pointer so
// the compiler emits ds_read / ds_write (a generic pointer would turn every
// push and pop into a flat access through the vector memory pipeline).
typedef __attribute__((address_space(3))) unsigned long long LdsEntry;

template <int kLds>
struct Stack {
    static constexpr int kCapacity = kLds;
    LdsEntry* lds;
    uint2* over;
    uint32_t lane;
    __device__ __forceinline__ void put_lds(uint32_t sp, uint2 e) {
        lds[sp * kBlock + lane] = (unsigned long long)e.x | ((unsigned long long)e.y << 32);
    }
    __device__ __forceinline__ void put(uint32_t sp, uint2 e) {
        if (sp < (uint32_t)kLds) {
            lds[sp * kBlock + lane] = (unsigned long long)e.x | ((unsigned long long)e.y << 32);
        } else {
            over[sp - kLds] = e;
        }
    }
    __device__ __forceinline__ uint2 get(uint32_t sp) const {
        if (sp < (uint32_t)kLds) {
            const unsigned long long v = lds[sp * kBlock + lane];
            return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
        }
        return over[sp - kLds];
    }
};

template <int kSem, bool kAnyHit>
__device__ __forceinline__ void finish_lane(const TraceParams& p, Lane& L, uint32_t kind) {
    if (kSem == kSemVis) {
        reinterpret_cast<uint4*>(p.hitsV1 + L.idx)[4] = make_uint4(L.k0, kind, 0u, L.k1);
        L.active = false;
        return;
    }
    if (kSem == kSemMixed) {
        if (L.shadow) finish_lane<kSemShadow, true>(p, L, kind);
        else finish_lane<kSemQueue, false>(p, L, kind);
        return;
    }
    if (kSem == kSemQueue) {
        // The queue walk keeps only the winner's index and t in registers;
        // its barycentrics and facing come from testing the winner again
        // (the same arithmetic on the same operands gives the same bits).
        QueueHit h;
        h.t = L.r.tmax;
        h.u = 0.0f;
        h.v = 0.0f;
        h.tri = kind == 0u ? kHitMiss : kHitInvalid;
        if (kind == 1u) {
            TriHit w;
            (void)tri_test(L.r, load_tri(p.s.tris, L.bt), L.best, true, w);
            h.t = L.best;
            h.u = w.u;
            h.v = w.v;
            h.tri = L.bt | (w.front ? kHitFrontBit : 0u);
        }
        p.queueHits[L.idx] = h;
        L.active = false;
        return;
    }
    if (kSem == kSemShadow) {
        // TraceShadowCS (Physical): visibility = any hit ? 0 : 1; a
        // fail-closed invalid ray counts as occluded.  Unoccluded
        // contributions are added to the direct (primary bounce) or indirect
        // accumulators; a non-finite or negative sum flags the path.
        if (kind == 0u) {
            const ShadowItem it = p.shadows[L.idx];
            const uint32_t pathBits = f2u(it.diffPath.w);
            const uint32_t path = pathBits & 0x7fffffffu;
            const bool primary = (pathBits >> 31) != 0u;
            PathState* ps = p.paths + path;
            const F4 d0 = primary ? ps->dd : ps->id;
            const F4 s0 = primary ? ps->ds : ps->is;
            const float dx = d0.x + it.diffPath.x * 1.0f, dy = d0.y + it.diffPath.y * 1.0f,
                        dz = d0.z + it.diffPath.z * 1.0f;
            const float sx = s0.x + it.specPrim.x * 1.0f, sy = s0.y + it.specPrim.y * 1.0f,
                        sz = s0.z + it.specPrim.z * 1.0f;
            const bool ok = __builtin_isfinite(dx) && __builtin_isfinite(dy) && __builtin_isfinite(dz) &&
                            __builtin_isfinite(sx) && __builtin_isfinite(sy) && __builtin_isfinite(sz) &&
                            dx >= 0.0f && dy >= 0.0f && dz >= 0.0f && sx >= 0.0f && sy >= 0.0f && sz >= 0.0f;
            if (!ok) {
                ps->meta.x = (ps->meta.x & ~kPathActive) | kPathTerminated | kPathError;
            } else if (primary) {
                ps->dd = {dx, dy, dz, d0.w};
                ps->ds = {sx, sy, sz, s0.w};
            } else {
                ps->id = {dx, dy, dz, d0.w};
                ps->is = {sx, sy, sz, s0.w};
            }
        }
        L.active = false;
        return;
    }
    if (kSem == kSemWalk) {
        rex_walk_hit out = {};
        out.t = L.r.tmax;
        out.triangle = kInvalid;
        out.instanceId = kInvalid;
        out.primitiveId = kInvalid;
        out.materialIndex = kInvalid;
        out.result = kind;
        if (!kAnyHit && kind == 1u) {
            float gn[3], sn[3];
            TriAttr attr;
            hit_normals(p.s, L.bt, L.bu, L.bv, L.bf, gn, sn, attr);
            out.t = L.best;
            out.u = L.bu;
            out.v = L.bv;
            out.frontFace = L.bf ? 1u : 0u;
            out.triangle = L.bt;
            out.position[0] = L.r.ox + L.r.dx * L.best;
            out.position[1] = L.r.oy + L.r.dy * L.best;
            out.position[2] = L.r.oz + L.r.dz * L.best;
            for (int k = 0; k < 3; ++k) {
                out.geometricNormal[k] = gn[k];
                out.shadingNormal[k] = sn[k];
            }
            out.materialIndex = attr.ids.w;
            out.instanceId = attr.ids.x;
            out.primitiveId = attr.ids.y;
        }
        p.hitsWalk[L.idx] = out;
    } else {
        rex_hit_v1* dst = p.hitsV1 + L.idx;
        if (kind != REX_HIT_KIND_TRIANGLE) {
            v1_terminal(dst, L.r.tmax, L.k0, L.k1, kind);
        } else {
            float gn[3], sn[3];
            TriAttr attr;
            hit_normals(p.s, L.bt, L.bu, L.bv, L.bf, gn, sn, attr);
            rex_hit_v1 h = {};
            h.positionT[0] = L.r.ox + L.r.dx * L.best;
            h.positionT[1] = L.r.oy + L.r.dy * L.best;
            h.positionT[2] = L.r.oz + L.r.dz * L.best;
            h.positionT[3] = L.best;
            h.geometricNormalBaryU[0] = gn[0];
            h.geometricNormalBaryU[1] = gn[1];
            h.geometricNormalBaryU[2] = gn[2];
            h.geometricNormalBaryU[3] = L.bu;
            h.shadingNormalBaryV[0] = sn[0];
            h.shadingNormalBaryV[1] = sn[1];
            h.shadingNormalBaryV[2] = sn[2];
            h.shadingNormalBaryV[3] = L.bv;
            h.ids[0] = attr.ids.x;
            h.ids[1] = attr.ids.y;
            h.ids[2] = attr.ids.z;
            h.ids[3] = attr.ids.w;
            h.metadata[0] = L.k0;
            h.metadata[1] = REX_HIT_KIND_TRIANGLE;
            // HitFlags: front face 1, alpha tested 2 (software_trace_v1.hlsl:400-409).
            const bool alphaTested = (attr.meta.x & kGeometryFlagAlphaMask) != 0u ||
                                     (attr.meta.y & kMaterialFlagAlphaMask) != 0u;
            h.metadata[2] = (L.bf ? 1u : 0u) | (alphaTested ? 2u : 0u);
            h.metadata[3] = L.k1;
            *dst = h;
        }
    }
    L.active = false;
}

// Moves a lane onto a child reference: a leaf starts the triangle cursor, an
// interior node becomes the next node to fetch.
template <bool kCount>
__device__ __forceinline__ void enter(Lane& L, uint32_t ref, WalkCounters& c) {
    if (is_leaf_ref(ref)) {
        L.leafNext = leaf_first(ref);
        L.leafLeft = leaf_count(ref);
        if (kCount) {
            c.leafVisits++;
            c.leafPrims += L.leafLeft;
            c.maxLeaf = L.leafLeft > c.maxLeaf ? L.leafLeft : c.maxLeaf;
        }
    } else {
        L.ref = ref;
    }
}

// Starts ray `idx` on lane L.  Returns false when the ray finished at once
// (invalid ray, empty mask, empty scene, root box missed).
template <int kSem, int kOrder, bool kAnyHit, bool kCount>
__device__ __forceinline__ bool start_lane(const TraceParams& p, Lane& L, uint32_t idx, WalkCounters& c,
                                           uint32_t& invalidRay) {
    L.idx = idx;
    L.active = true;
    L.hold = false;
    L.ref = kInvalid;
    L.found = false;
    L.bu = 0.0f;
    L.bv = 0.0f;
    L.bf = false;
    L.bt = kInvalid;
    L.leafLeft = 0;
    L.sp = 0;
    L.visited = 0;
    const SceneView& s = p.s;
    float4 a, b;
    L.shadow = kSem == kSemShadow;
    if (kSem == kSemMixed) {
        // idx below the ray count is a ray of the queue, above it a shadow item.
        const uint32_t rays = *p.countPtr;
        L.shadow = idx >= rays;
        if (L.shadow) idx -= rays;
        L.idx = idx;
    }
    if (kSem == kSemQueue || kSem == kSemShadow || kSem == kSemMixed) {
        if (kSem == kSemQueue || (kSem == kSemMixed && !L.shadow)) {
            a = reinterpret_cast<const float4*>(p.rays)[2u * idx];
            b = reinterpret_cast<const float4*>(p.rays)[2u * idx + 1u];
            L.k1 = kInvalid;
        } else {
            a = reinterpret_cast<const float4*>(p.shadows + idx)[0];
            b = reinterpret_cast<const float4*>(p.shadows + idx)[1];
            L.k1 = f2u(reinterpret_cast<const float4*>(p.shadows + idx)[3].w);
        }
        L.k0 = kInvalid;
        L.r.tmax = b.w;
        if (s.nodeCount == 0u || s.triCount == 0u) {
            finish_lane<kSem, kAnyHit>(p, L, 0u);
            return false;
        }
        if (!walk_ray_valid(a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w)) {
            invalidRay++;
            finish_lane<kSem, kAnyHit>(p, L, 2u);
            return false;
        }
    } else if (kSem == kSemWalk) {
        a = reinterpret_cast<const float4*>(p.rays)[2u * idx];
        b = reinterpret_cast<const float4*>(p.rays)[2u * idx + 1u];
        if (p.keys != nullptr) {
            const uint2 k = reinterpret_cast<const uint2*>(p.keys)[idx];
            L.k0 = k.x;
            L.k1 = k.y;
        } else {
            L.k0 = kInvalid;
            L.k1 = kInvalid;
        }
        L.r.tmax = b.w;
        if (s.nodeCount == 0u || s.triCount == 0u) {
            finish_lane<kSem, kAnyHit>(p, L, 0u);
            return false;
        }
        if (!walk_ray_valid(a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w)) {
            invalidRay++;
            finish_lane<kSem, kAnyHit>(p, L, 2u);
            return false;
        }
    } else {
        a = reinterpret_cast<const float4*>(p.raysV1 + idx)[0];
        b = reinterpret_cast<const float4*>(p.raysV1 + idx)[1];
        const uint4 id = reinterpret_cast<const uint4*>(p.raysV1 + idx)[2];
        L.k0 = id.x;
        L.k1 = id.y;
        L.r.tmax = b.w;
        if (!v1_ray_valid(a, b, id.w) || p.stackCapacity == 0u || p.stackCapacity > 64u) {
            invalidRay++;
            finish_lane<kSem, kAnyHit>(p, L, REX_HIT_KIND_INVALID);
            return false;
        }
        if (id.w == 0u || s.nodeCount == 0u) {
            finish_lane<kSem, kAnyHit>(p, L, REX_HIT_KIND_MISS);
            return false;
        }
    }
    L.r = make_ray(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w);
    L.best = L.r.tmax;
    // Wide4 float4 order: lo.x, lo.y, lo.z, hi.x, hi.y, hi.z.
    L.planes = L.r.exact ? (0u | (1u << 2) | (2u << 5))
                         : ((b.x < 0.0f ? 3u : 0u) | ((b.y < 0.0f ? 4u : 1u) << 2) | ((b.z < 0.0f ? 5u : 2u) << 5));
    float rn;
    if (kCount) { c.nodeTests++; c.boxTests++; }
    uint32_t rootRef = kOrder == kOrderWideQ ? s.rootRefQ : s.rootRef;
    float r0 = s.rootMin[0], r1 = s.rootMin[1], r2 = s.rootMin[2];
    float r3 = s.rootMax[0], r4 = s.rootMax[1], r5 = s.rootMax[2];
    if ((kSem == kSemQueue || kSem == kSemShadow || kSem == kSemMixed) && p.roots != nullptr) {
        const uint32_t sid = (L.shadow ? p.shadowSceneOf : p.raySceneOf)[idx];
        const SceneRoot R = p.roots[sid];
        rootRef = R.rootRef;
        r0 = R.rootMin[0]; r1 = R.rootMin[1]; r2 = R.rootMin[2];
        r3 = R.rootMax[0]; r4 = R.rootMax[1]; r5 = R.rootMax[2];
    }
    const bool rootHit = L.r.exact ? slab_exact(L.r, r0, r1, r2, r3, r4, r5, rn)
                                   : slab_fast(L.r, r0, r1, r2, r3, r4, r5, rn);
    if (!rootHit || !(rn < L.best)) {
        finish_lane<kSem, kAnyHit>(p, L, kSem == kSemV1 || kSem == kSemVis ? REX_HIT_KIND_MISS : 0u);
        return false;
    }
    if (kSem == kSemV1) {
        L.visited = 1;
        if (kCount) c.maxStack = c.maxStack > 1u ? c.maxStack : 1u;
    }
    if (kOrder == kOrderWideQ) {
        // The compressed walk's boxes are conservative, so a zero direction
        // component may use a large finite inverse (no inf * 0).  `planes`
        // (unused by this order) keeps which components were replaced, for
        // the exact leaf-box test.
        L.planes = (__builtin_isfinite(L.r.ix) ? 0u : 1u) | (__builtin_isfinite(L.r.iy) ? 0u : 2u) |
                   (__builtin_isfinite(L.r.iz) ? 0u : 4u);
        L.r.ix = __builtin_isfinite(L.r.ix) ? L.r.ix : 1.0e30f;
        L.r.iy = __builtin_isfinite(L.r.iy) ? L.r.iy : 1.0e30f;
        L.r.iz = __builtin_isfinite(L.r.iz) ? L.r.iz : 1.0e30f;
    }
    enter<kCount>(L, rootRef, c);
    return true;
}

template <int kSem, int kOrder, bool kAnyHit, bool kCount>
__device__ __forceinline__ void trace_body(const TraceParams& p) {
    constexpr int kLds = lds_entries<kSem, kOrder, kAnyHit, kCount>();
    __shared__ uint2 ldsStack[kLds * kBlock];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack<kLds> st;
    st.lds = (LdsEntry*)ldsStack;
    st.over = p.overflow + (size_t)gtid * p.overflowStride;
    st.lane = threadIdx.x;
    const SceneView& s = p.s;
    WalkCounters c = {};
    uint32_t invalidRay = 0, invalidHit = 0, overflowed = 0;
    Lane L;
    L.active = false;
    const uint32_t count = kSem == kSemMixed ? *p.countPtr + *p.countPtr2
                                              : (p.countPtr != nullptr ? *p.countPtr : p.count);
    // Wave-private pool of ray indices [poolNext, poolEnd).  Wave w owns
    // chunk w of kChunk rays from the start; later chunks are claimed from
    // the global counter (which counts chunks handed out after the first
    // gridDim * waves-per-block), one atomic per kChunk rays.  Waves whose
    // first chunk lies beyond `count` exit without touching the counter, so
    // a launch over a short queue costs no atomic traffic.  `drained` is set
    // once the counter passed `count`.
    const uint32_t waveId = blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6);
    const uint32_t firstDynamic = gridDim.x * (kBlock / 64u) * kChunk;
    uint32_t poolNext = waveId * kChunk;
    uint32_t poolEnd = poolNext < count ? (poolNext + kChunk < count ? poolNext + kChunk : count) : poolNext;
    bool drained = poolNext >= count;
    const uint64_t below = (1ull << lane) - 1ull;

    for (;;) {
        // Refill idle lanes from the pool once at least kRefillMin of them
        // are idle (or all are): a refill pass costs the wave about the same
        // for one lane as for many.  A lane whose new ray finishes at once
        // (invalid, root box missed) takes another in the same pass.
        uint64_t m = __ballot(!L.active);
        if ((uint32_t)__popcll(m) < kRefillMin && m != ~0ull) m = 0u;
        while (m != 0u && !drained) {
            if (poolNext >= poolEnd) {
                uint32_t base = 0;
                if (lane == 0u) base = firstDynamic + atomicAdd(p.workCounter, kChunk);
                base = __shfl(base, 0, 64);
                if (base >= count) {
                    drained = true;
                    break;
                }
                poolNext = base;
                poolEnd = base + kChunk < count ? base + kChunk : count;
            }
            const uint32_t avail = poolEnd - poolNext;
            const uint32_t rank = (uint32_t)__popcll(m & below);
            if (!L.active && rank < avail) {
                start_lane<kSem, kOrder, kAnyHit, kCount>(p, L, poolNext + rank, c, invalidRay);
            }
            const uint32_t want = (uint32_t)__popcll(m);
            poolNext += want < avail ? want : avail;
            m = __ballot(!L.active);
        }
        if (!__any(L.active)) break;  // the counter is drained and every lane idle
        if (!L.active) continue;

        bool needPop = false;
        // A node step first; a lane that enters a leaf tests its first triangle
        // in the same pass.
        if (L.leafLeft == 0u) {
            if (is_wide(kOrder)) {
            // One 4-wide node: every child box; the hit children ordered by
            // (clamped entry distance, slot), the nearest entered and the
            // rest pushed farthest first.
            uint32_t ref[4];
            bool h[4];
            float nr[4];
            {
            // Each lane loads its entry planes (min for a positive direction
            // component, max for a negative one) as N and its exit planes as
            // F, so the slab test needs no per-axis min/max; rays with a zero
            // component load (min, max) and keep the reference's swap form.
            const float4* nb = reinterpret_cast<const float4*>(s.wide4 + L.ref);
            const uint32_t ox = L.planes & 3u, oy = (L.planes >> 2) & 7u, oz = (L.planes >> 5) & 7u;
            const float4 Nx = nb[ox], Ny = nb[oy], Nz = nb[oz];
            const float4 Fx = nb[3u - ox], Fy = nb[5u - oy], Fz = nb[7u - oz];
            const uint4 refs = reinterpret_cast<const uint4*>(nb)[6];
            const float nxa[4] = {Nx.x, Nx.y, Nx.z, Nx.w}, nya[4] = {Ny.x, Ny.y, Ny.z, Ny.w},
                        nza[4] = {Nz.x, Nz.y, Nz.z, Nz.w};
            const float fxa[4] = {Fx.x, Fx.y, Fx.z, Fx.w}, fya[4] = {Fy.x, Fy.y, Fy.z, Fy.w},
                        fza[4] = {Fz.x, Fz.y, Fz.z, Fz.w};
            ref[0] = refs.x; ref[1] = refs.y; ref[2] = refs.z; ref[3] = refs.w;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float nk;
                const bool box = L.r.exact ? slab_exact(L.r, nxa[k], nya[k], nza[k], fxa[k], fya[k], fza[k], nk)
                                           : slab_ordered(L.r, nxa[k], nya[k], nza[k], fxa[k], fya[k], fza[k], nk);
                h[k] = ref[k] != kInvalid && box && within_limit(nk, L.best, L.found);
                nr[k] = fmaxf(nk, L.r.tmin);
                if (kCount && ref[k] != kInvalid) { c.boxTests++; c.nodeTests++; }
            }
            }
            uint32_t rank[4];
            uint32_t m = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t r = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j != k && h[j] && (nr[j] < nr[k] || (nr[j] == nr[k] && j < k))) ++r;
                }
                rank[k] = r;
                m += h[k] ? 1u : 0u;
            }
            if (m == 0u) {
                needPop = true;
            } else {
                uint32_t first = kInvalid;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (!h[k]) continue;
                    if (rank[k] == 0u) {
                        first = ref[k];
                    } else {
                        st.put(L.sp + (m - 1u - rank[k]), make_uint2(ref[k], f2u(nr[k])));
                    }
                }
                L.sp += m - 1u;
                if (kCount && L.sp > c.maxStack) c.maxStack = L.sp;
                enter<kCount>(L, first, c);
            }
        } else {
            // One interior node: both child boxes.
            const WideNode n = load_wide(s.wide, L.ref);
            float ln, rn;
            bool li, ri;
            if (L.r.exact) {
                li = slab_exact(L.r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, ln);
                ri = slab_exact(L.r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, rn);
            } else {
                li = slab_fast(L.r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, ln);
                ri = slab_fast(L.r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, rn);
            }
            if (kCount) { c.boxTests += 2; c.nodeTests += kOrder == kOrderLeftFirst ? 4u : 2u; }
            if (kOrder == kOrderLeftFirst) {
                if (li && within_limit(ln, L.best, L.found)) {
                    if (ri) {
                        st.put(L.sp++, make_uint2(n.link.y, f2u(rn)));
                        if (kCount && L.sp > c.maxStack) c.maxStack = L.sp;
                    } else if (kCount && kAnyHit) {
                        // The reference tests the right box only on its way
                        // back from the left subtree.  The counting any-hit
                        // variant keeps a NaN-distance placeholder (never
                        // entered) so that an early exit can take back the
                        // right-box tests the reference would not have made.
                        st.put(L.sp++, make_uint2(n.link.y, 0x7fc00000u));
                    }
                    enter<kCount>(L, n.link.x, c);
                } else if (ri && within_limit(rn, L.best, L.found)) {
                    enter<kCount>(L, n.link.y, c);
                } else {
                    needPop = true;
                }
            } else {
                const bool hl = li && within_limit(ln, L.best, L.found);
                const bool hr = ri && within_limit(rn, L.best, L.found);
                if (kSem == kSemV1) {
                    const uint32_t required = (hl ? 1u : 0u) + (hr ? 1u : 0u);
                    if (L.sp + required > p.stackCapacity) {
                        overflowed++;
                        finish_lane<kSem, kAnyHit>(p, L, REX_HIT_KIND_INVALID);
                        continue;
                    }
                    if (kCount && L.sp + required > c.maxStack) c.maxStack = L.sp + required;
                }
                const float lnc = fmaxf(ln, L.r.tmin), rnc = fmaxf(rn, L.r.tmin);
                if (hl && hr) {
                    // The left child's binary index is below the right one's
                    // in pre-order, so equal distances go left first.
                    const bool leftFirst = lnc < rnc || !(rnc < lnc);
                    st.put(L.sp++, leftFirst ? make_uint2(n.link.y, f2u(rnc)) : make_uint2(n.link.x, f2u(lnc)));
                    if (kSem == kSemWalk && kCount && L.sp > c.maxStack) c.maxStack = L.sp;
                    if (kSem == kSemV1 && ++L.visited > s.nodeCount) {
                        invalidHit++;
                        finish_lane<kSem, kAnyHit>(p, L, REX_HIT_KIND_INVALID);
                        continue;
                    }
                    enter<kCount>(L, leftFirst ? n.link.x : n.link.y, c);
                } else if (hl || hr) {
                    if (kSem == kSemV1 && ++L.visited > s.nodeCount) {
                        invalidHit++;
                        finish_lane<kSem, kAnyHit>(p, L, REX_HIT_KIND_INVALID);
                        continue;
                    }
                    enter<kCount>(L, hl ? n.link.x : n.link.y, c);
                } else {
                    needPop = true;
                }
            }
        }
        }
        if (!needPop && L.leafLeft != 0u) {
            // One triangle of the current leaf.
            const uint32_t ti = L.leafNext++;
            L.leafLeft--;
            const TriPos tp = load_tri(s.tris, ti);
            const uint32_t flags = f2u(tp.p2.w);
            bool consider = true;
            if (kSem != kSemV1) {
                const uint32_t inst = f2u(tp.p2.y), prim = f2u(tp.p2.z);
                const bool anyHit = kAnyHit || (kSem == kSemMixed && L.shadow);
                // Queue items ignore (invalid instance, primitive k1) only.
                const uint32_t ignInst = kSem == kSemWalk ? L.k0 : kInvalid;
                if (inst == ignInst && prim == L.k1) consider = false;
                if (anyHit && (flags & kTriCastsShadow) == 0u) consider = false;
                if (consider) {
                    if (kCount) c.triTests++;
                    TriHit h;
                    if (tri_test(L.r, tp, L.best, L.found, h) && (h.front || (flags & kTriDoubleSided) != 0u)) {
                        if (anyHit) {
                            // Every stack entry of a left-first walk is a
                            // right box the reference has not tested yet.
                            if (kCount && kOrder == kOrderLeftFirst) c.boxTests -= L.sp;
                            finish_lane<kSem, kAnyHit>(p, L, 1u);
                            continue;
                        }
                        // A tie in t goes to the lower (instance, primitive);
                        // the winner's identity is read back only then (rare),
                        // instead of being kept in registers.
                        bool eqLower = false;
                        if (L.found && h.t == L.best) {
                            const TriPos bp = load_tri(s.tris, L.bt);
                            const uint32_t bInst = f2u(bp.p2.y), bPrim = f2u(bp.p2.z);
                            eqLower = inst < bInst || (inst == bInst && prim < bPrim);
                        }
                        if (!L.found || h.t < L.best || eqLower) {
                            L.found = true;
                            L.best = h.t;
                            L.bt = ti;
                            if (kSem == kSemWalk) {
                                L.bu = h.u;
                                L.bv = h.v;
                                L.bf = h.front;
                            }
                        }
                    }
                }
            } else {
                if (kCount) c.triTests++;
                TriHit h;
                if (tri_test(L.r, tp, L.best, L.found, h) && (h.front || (flags & kTriDoubleSided) != 0u)) {
                    bool replace = !L.found || h.t < L.best;
                    if (!replace && h.t == L.best) replace = s.attrs[ti].meta.z < s.attrs[L.bt].meta.z;
                    if (kAnyHit || replace) {
                        L.found = true;
                        L.best = h.t;
                        L.bu = h.u;
                        L.bv = h.v;
                        L.bf = h.front;
                        L.bt = ti;
                    }
                    if (kAnyHit) {
                        finish_lane<kSem, kAnyHit>(p, L, REX_HIT_KIND_TRIANGLE);
                        continue;
                    }
                }
            }
            needPop = L.leafLeft == 0u;
        }
        if (needPop) {
            bool next = false;
            bool bad = false;
            while (L.sp != 0u) {
                const uint2 e = st.get(--L.sp);
                if (kSem == kSemV1 && ++L.visited > s.nodeCount) {
                    bad = true;
                    break;
                }
                if (within_limit(u2f(e.y), L.best, L.found)) {
                    enter<kCount>(L, e.x, c);
                    next = true;
                    break;
                }
            }
            if (bad) {
                invalidHit++;
                finish_lane<kSem, kAnyHit>(p, L, REX_HIT_KIND_INVALID);
            } else if (!next) {
                if (kSem != kSemV1) {
                    finish_lane<kSem, kAnyHit>(p, L, L.found ? 1u : 0u);
                } else {
                    finish_lane<kSem, kAnyHit>(p, L, L.found ? REX_HIT_KIND_TRIANGLE : REX_HIT_KIND_MISS);
                }
            }
        }
    }
    if (kCount && p.counters != nullptr) flush_counters(p.counters, c, invalidRay, invalidHit, overflowed);
}

// ---- compressed 4-wide walk: node phase / leaf phase ------------------------
//
// The production walk runs each wave in two phases instead of giving every
// iteration a node step and a triangle test (where a wave pays for both
// whenever its lanes are split between them):
//  * node phase: lanes step through compressed nodes until at most
//    kWqLeafless (28) of the wave's lanes with work left still lack a pending
//    leaf.  A lane that already holds one keeps traversing (popping its
//    stack) so it is not idle, and stops at its next leaf, which waits on top
//    of its stack (postponed leaves, speculative traversal);
//  * leaf phase: the lanes test the triangles of their pending leaves, one
//    per iteration, until no lane has one left.
// Measured on the 1M-triangle soup (incoherent bounce rays, leaves of up to
// 4 triangles): 14-15% faster than one node step and one triangle test per
// iteration; waiting for all 64 lanes to hold a leaf, or no speculation, is
// slower than that baseline.  Over the single-triangle traversal hierarchy
// the thresholds 20-36 are within 2% of each other (28 is used); capping the
// leaf phase's iterations, leaving it early, or chaining into a sibling leaf
// on top of the stack did not help.
// The visiting order differs from the one-item-per-iteration walk only by
// testing leaves later; hits are the same (closest hit = the minimum of
// (t, instance, primitive) over the admitted candidates, up to the documented
// rounding ties of box culling; any-hit = whether any admitted candidate
// exists, which does not depend on the order).

// within_limit(t, best, found) as one comparison, t < walk_limit(best,
// found): the smallest float above best when a hit was found (best is then
// finite; -0 becomes +0 first so its successor is the smallest denormal),
// best itself otherwise.  NaN fails both forms.
__device__ __forceinline__ float walk_limit(float best, bool found) {
    const uint32_t u = f2u(best + 0.0f);
    return found ? u2f(u + (uint32_t)(((int32_t)u >> 31) | 1)) : best;
}

// Tests the child boxes of compressed node L.ref, pushes the hit children
// except the nearest (farthest deepest) and returns the nearest one's
// reference and clamped entry distance (kInvalid: no child hit).
template <bool kCount, typename StackT>
__device__ __forceinline__ uint32_t wq_node_step(const SceneView& s, Lane& L, StackT& st, WalkCounters& c,
                                                 float lim, float& nearOut) {
    uint32_t ref[4];
    bool h[4];
    float nr[4];
    // Compressed node: plane = origin + q * step per axis, so
    // t = q * (step / d) + (origin - o) / d, one fma per plane.  Entry
    // planes (lo for a non-negative inverse, hi for a negative one) and exit
    // planes are picked per axis as whole words (byte k = child k).
    const uint4* nb = reinterpret_cast<const uint4*>(s.wide4q + L.ref);
    const uint4 hdr = nb[0], pa = nb[1], pb = nb[2], rf = nb[3];
    ref[0] = rf.x; ref[1] = rf.y; ref[2] = rf.z; ref[3] = rf.w;
    const float ax = __uint_as_float((hdr.w & 0xffu) << 23) * L.r.ix;
    const float ay = __uint_as_float(((hdr.w >> 8) & 0xffu) << 23) * L.r.iy;
    const float az = __uint_as_float(((hdr.w >> 16) & 0xffu) << 23) * L.r.iz;
    const float bx = (__uint_as_float(hdr.x) - L.r.ox) * L.r.ix;
    const float by = (__uint_as_float(hdr.y) - L.r.oy) * L.r.iy;
    const float bz = (__uint_as_float(hdr.z) - L.r.oz) * L.r.iz;
    const bool nx = L.r.ix < 0.0f, ny = L.r.iy < 0.0f, nz = L.r.iz < 0.0f;
    const uint32_t enX = nx ? pa.w : pa.x, exX = nx ? pa.x : pa.w;
    const uint32_t enY = ny ? pb.x : pa.y, exY = ny ? pa.y : pb.x;
    const uint32_t enZ = nz ? pb.y : pa.z, exZ = nz ? pa.z : pb.y;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t sh = 8u * (uint32_t)k;
        const float n0 = __builtin_fmaf((float)((enX >> sh) & 0xffu), ax, bx);
        const float n1 = __builtin_fmaf((float)((enY >> sh) & 0xffu), ay, by);
        const float n2 = __builtin_fmaf((float)((enZ >> sh) & 0xffu), az, bz);
        const float f0 = __builtin_fmaf((float)((exX >> sh) & 0xffu), ax, bx);
        const float f1 = __builtin_fmaf((float)((exY >> sh) & 0xffu), ay, by);
        const float f2 = __builtin_fmaf((float)((exZ >> sh) & 0xffu), az, bz);
        const float nk = fmaxf(fmaxf(n0, n1), n2);
        const float fk = fminf(fminf(f0, f1), f2);
        // nk < lim is within_limit(nk, best, found); it also implies
        // nk < tMax (an unfound lane's best is tMax, a found one's is below).
        h[k] = ref[k] != kInvalid && !(fk < nk) && fk > L.r.tmin && nk < lim;
        nr[k] = fmaxf(nk, L.r.tmin);
        if (kCount && ref[k] != kInvalid) { c.boxTests++; c.nodeTests++; }
    }
    // far[k]: hit children farther than child k (at equal distance the
    // higher index is the farther one).  The nearest hit child is returned;
    // the others go on the stack farthest deepest.
    uint32_t far[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int j = k + 1; j < 4; ++j) {
            const bool kFarther = nr[k] > nr[j];
            far[j] += (h[k] && kFarther) ? 1u : 0u;
            far[k] += (h[j] && !kFarther) ? 1u : 0u;
        }
    }
    const uint32_t m = (h[0] ? 1u : 0u) + (h[1] ? 1u : 0u) + (h[2] ? 1u : 0u) + (h[3] ? 1u : 0u);
    uint32_t first = kInvalid;
    float firstNear = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool isFirst = h[k] && far[k] == m - 1u;
        first = isFirst ? ref[k] : first;
        firstNear = isFirst ? nr[k] : firstNear;
    }
    // The pushes go to LDS whenever all three possible entries fit there
    // (one check per node instead of one per entry).
    if (L.sp + 3u <= (uint32_t)StackT::kCapacity) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (h[k] && far[k] != m - 1u) st.put_lds(L.sp + far[k], make_uint2(ref[k], f2u(nr[k])));
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (h[k] && far[k] != m - 1u) st.put(L.sp + far[k], make_uint2(ref[k], f2u(nr[k])));
        }
    }
    if (m != 0u) L.sp += m - 1u;
    if (kCount && L.sp > c.maxStack) c.maxStack = L.sp;
    nearOut = firstNear;
    return first;
}

// One triangle of the lane's pending leaf (walk, queue, shadow and mixed
// semantics).  A candidate hit is admitted only when its exact leaf box passes
// the reference's tMin / tMax test.  An any-hit query that finds an occluder
// finishes the lane.
template <int kSem, bool kAnyHit, bool kCount>
__device__ __forceinline__ void wq_triangle(const TraceParams& p, Lane& L, WalkCounters& c) {
    const SceneView& s = p.s;
    const uint32_t ti = L.leafNext++;
    L.leafLeft--;
    const TriPos tp = load_tri(s.tris, ti);
    const uint32_t flags = f2u(tp.p2.w);
    const uint32_t inst = f2u(tp.p2.y), prim = f2u(tp.p2.z);
    const bool anyHit = kAnyHit || (kSem == kSemMixed && L.shadow);
    // Queue items ignore (invalid instance, primitive k1) only.
    const uint32_t ignInst = kSem == kSemWalk ? L.k0 : kInvalid;
    // Batch-tracer queries (visibility) test every triangle, as software_trace_v1.hlsl does.
    const bool considered = kSem == kSemVis ||
                            (!((inst == ignInst) & (prim == L.k1)) & !(anyHit & ((flags & kTriCastsShadow) == 0u)));
    if (kCount && considered) c.triTests++;
    TriHit h;
    const bool cand = tri_test_flat(L.r, tp, walk_limit(L.best, L.found), h) & considered &
                      (h.front | ((flags & kTriDoubleSided) != 0u));
    if (!cand || !leaf_box_admits_q(L.r, L.planes, s.leafBox, ti)) return;
    if (anyHit) {
        finish_lane<kSem, kAnyHit>(p, L, 1u);
        return;
    }
    // A tie in t goes to the lower (instance, primitive); the winner's
    // identity is read back only then (rare).
    bool eqLower = false;
    if (L.found && h.t == L.best) {
        const TriPos bp = load_tri(s.tris, L.bt);
        const uint32_t bInst = f2u(bp.p2.y), bPrim = f2u(bp.p2.z);
        eqLower = inst < bInst || (inst == bInst && prim < bPrim);
    }
    if (!L.found || h.t < L.best || eqLower) {
        L.found = true;
        L.best = h.t;
        L.bt = ti;
        if (kSem == kSemWalk) {
            L.bu = h.u;
            L.bv = h.v;
            L.bf = h.front;
        }
    }
}

// Node phase exit: once at most kWqLeafless lanes with work left lack a
// pending leaf (and some lane has one).  Speculation: lanes holding a
// pending leaf keep traversing.
#ifndef REX_WQ_LEAFLESS
#define REX_WQ_LEAFLESS 28
#endif
#ifndef REX_WQ_SPECULATE
#define REX_WQ_SPECULATE 1
#endif
#ifndef REX_WQ_LEAFTAIL
#define REX_WQ_LEAFTAIL 0
#endif
constexpr uint32_t kWqLeafless = REX_WQ_LEAFLESS;
// Leaf phase exit: once at most kWqLeafTail lanes still test triangles while
// others have nodes to step.
constexpr uint32_t kWqLeafTail = REX_WQ_LEAFTAIL;
// Leaf phase exit: after kWqLeafIters triangle iterations while some lane has
// nodes to step; lanes keep the rest of their leaf for the next leaf phase.
#ifndef REX_WQ_LEAFITERS
#define REX_WQ_LEAFITERS 64
#endif
constexpr uint32_t kWqLeafIters = REX_WQ_LEAFITERS;
constexpr bool kWqSpeculate = REX_WQ_SPECULATE != 0;

template <int kSem, bool kAnyHit, bool kCount>
__device__ __forceinline__ void trace_body_wq(const TraceParams& p) {
    static_assert(kSem != kSemV1, "the batch tracer walks exact nodes");
    static_assert(kSem != kSemVis || kAnyHit, "visibility queries are any-hit");
    constexpr int kOrder = kOrderWideQ;
    constexpr int kLds = lds_entries<kSem, kOrder, kAnyHit, kCount>();
    __shared__ uint2 ldsStack[kLds * kBlock];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack<kLds> st;
    st.lds = (LdsEntry*)ldsStack;
    st.over = p.overflow + (size_t)gtid * p.overflowStride;
    st.lane = threadIdx.x;
    const SceneView& s = p.s;
    WalkCounters c = {};
    uint32_t invalidRay = 0;
    Lane L;
    L.active = false;
    const uint32_t count = kSem == kSemMixed ? *p.countPtr + *p.countPtr2
                                              : (p.countPtr != nullptr ? *p.countPtr : p.count);
    // Ray pool of the wave (see trace_body).
    const uint32_t waveId = blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6);
    const uint32_t firstDynamic = gridDim.x * (kBlock / 64u) * kChunk;
    uint32_t poolNext = waveId * kChunk;
    uint32_t poolEnd = poolNext < count ? (poolNext + kChunk < count ? poolNext + kChunk : count) : poolNext;
    bool drained = poolNext >= count;
    const uint64_t below = (1ull << lane) - 1ull;

    for (;;) {
        uint64_t m = __ballot(!L.active);
        if ((uint32_t)__popcll(m) < kRefillMin && m != ~0ull) m = 0u;
        while (m != 0u && !drained) {
            if (poolNext >= poolEnd) {
                uint32_t base = 0;
                if (lane == 0u) base = firstDynamic + atomicAdd(p.workCounter, kChunk);
                base = __shfl(base, 0, 64);
                if (base >= count) {
                    drained = true;
                    break;
                }
                poolNext = base;
                poolEnd = base + kChunk < count ? base + kChunk : count;
            }
            const uint32_t avail = poolEnd - poolNext;
            const uint32_t rank = (uint32_t)__popcll(m & below);
            if (!L.active && rank < avail) {
                start_lane<kSem, kOrder, kAnyHit, kCount>(p, L, poolNext + rank, c, invalidRay);
            }
            const uint32_t want = (uint32_t)__popcll(m);
            poolNext += want < avail ? want : avail;
            m = __ballot(!L.active);
        }
        if (!__any(L.active)) break;  // the counter is drained and every lane idle

        // Node phase.  best and found change only in the leaf phase.
        const float lim = walk_limit(L.best, L.found);
        for (;;) {
            const bool work = L.active && (L.ref != kInvalid || L.sp != 0u);
            const uint32_t leafless = (uint32_t)__popcll(__ballot(work && L.leafLeft == 0u));
            // Leave early only when the leaf phase has work (otherwise the
            // leafless lanes would never advance).
            if (leafless == 0u || (leafless <= kWqLeafless && __any(L.active && L.leafLeft != 0u))) break;
            if (!work || L.hold || (!kWqSpeculate && L.leafLeft != 0u)) continue;
            if (L.ref == kInvalid) {
                // Next stack entry within the current best; a leaf waits
                // while another one is pending.
                while (L.sp != 0u) {
                    const uint2 e = st.get(L.sp - 1u);
                    if (!(u2f(e.y) < lim)) {
                        L.sp--;
                        continue;
                    }
                    if (is_leaf_ref(e.x)) {
                        if (L.leafLeft != 0u) {
                            L.hold = true;
                        } else {
                            L.sp--;
                            enter<kCount>(L, e.x, c);
                        }
                    } else {
                        L.sp--;
                        L.ref = e.x;
                    }
                    break;
                }
            }
            if (L.ref != kInvalid) {
                float nearFirst;
                const uint32_t first = wq_node_step<kCount>(s, L, st, c, lim, nearFirst);
                L.ref = kInvalid;
                if (first != kInvalid) {
                    if (!is_leaf_ref(first)) {
                        L.ref = first;
                    } else if (L.leafLeft == 0u) {
                        enter<kCount>(L, first, c);
                    } else {
                        st.put(L.sp++, make_uint2(first, f2u(nearFirst)));
                        if (kCount && L.sp > c.maxStack) c.maxStack = L.sp;
                        L.hold = true;
                    }
                }
            }
        }

        // Leaf phase.
        // At least one iteration runs whenever a lane has a pending leaf, so
        // every pass of the outer loop advances some lane.
        for (uint32_t it = 0;; ++it) {
            const uint32_t leafy = (uint32_t)__popcll(__ballot(L.active && L.leafLeft != 0u));
            if (leafy == 0u) break;
            if (it != 0u && (leafy <= kWqLeafTail || it >= kWqLeafIters) &&
                __any(L.active && L.leafLeft == 0u && (L.ref != kInvalid || L.sp != 0u))) {
                break;
            }
            if (L.active && L.leafLeft != 0u) {
                wq_triangle<kSem, kAnyHit, kCount>(p, L, c);
                if (L.leafLeft == 0u) L.hold = false;
            }
        }
        if (L.active && L.leafLeft == 0u && L.ref == kInvalid && L.sp == 0u) {
            finish_lane<kSem, kAnyHit>(p, L, L.found ? 1u : 0u);
        }
    }
    if (kCount && p.counters != nullptr) flush_counters(p.counters, c, invalidRay, 0u, 0u);
}

template <int kSem, int kOrder, bool kAnyHit, bool kCount>
__device__ __forceinline__ void trace_any_body(const TraceParams& p) {
    if constexpr (kOrder == kOrderWideQ) trace_body_wq<kSem, kAnyHit, kCount>(p);
    else trace_body<kSem, kOrder, kAnyHit, kCount>(p);
}

template <int kSem, int kOrder, bool kAnyHit, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace(TraceParams p) {
    trace_any_body<kSem, kOrder, kAnyHit, kCount>(p);
}

// Production instantiations (dense_trace): at least 7 waves per SIMD.
template <int kSem, int kOrder, bool kAnyHit, bool kCount>
#ifndef REX_DENSE_WAVES
#define REX_DENSE_WAVES 7
#endif
__global__ __launch_bounds__(kBlock, REX_DENSE_WAVES) void k_trace_dense(TraceParams p) {
    trace_any_body<kSem, kOrder, kAnyHit, kCount>(p);
}

template <int kSem, int kOrder, bool kAny, bool kCount>
auto trace_kernel() {
    if constexpr (dense_trace<kSem, kOrder, kAny, kCount>()) return &k_trace_dense<kSem, kOrder, kAny, kCount>;
    else return &k_trace<kSem, kOrder, kAny, kCount>;
}

template <bool kShadow, bool kCount>
__global__ __launch_bounds__(kBlock) void k_walk_stackless(TraceParams p) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    WalkCounters c = {};
    uint32_t invalidRay = 0;
    if (i < p.count) {
        Lane L;
        // start_lane tests the root box (uncounted here) unless the scene is
        // empty or the ray invalid.
        const bool rootTested = p.s.nodeCount != 0u && p.s.triCount != 0u;
        const uint32_t invalidBefore = invalidRay;
        const bool entered = start_lane<kSemWalk, kOrderLeftFirst, kShadow, false>(p, L, i, c, invalidRay);
        if (kCount && rootTested && invalidRay == invalidBefore) {
            c.nodeTests++;
            c.boxTests++;
        }
        if (entered) {
            const WalkResult w = walk_stackless<kShadow, kCount>(p.s, L.r, L.k0, L.k1, c);
            L.found = w.result == 1u;
            L.best = w.t;
            L.bu = w.u;
            L.bv = w.v;
            L.bf = w.front;
            L.bt = w.tri;
            finish_lane<kSemWalk, kShadow>(p, L, w.result);
        }
    }
    if (kCount && p.counters != nullptr) flush_counters(p.counters, c, invalidRay, 0u, 0u);
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

// ---- host launchers ----------------------------------------------------------

namespace {

template <typename K>
uint32_t resident_blocks(K kernel) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int perCu = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, kBlock, 0);
    if (perCu <= 0) perCu = 1;
    if (cus <= 0) cus = 1;
    return (uint32_t)(cus * perCu);
}

template <int kSem, int kOrder, bool kAny, bool kCount>
hipError_t launch_persistent(TraceParams& p, TraceWorkspace& ws, hipStream_t stream) {
    auto kernel = trace_kernel<kSem, kOrder, kAny, kCount>();
    static uint32_t resident = 0;
    if (resident == 0) resident = resident_blocks(kernel);
    const uint32_t need = (p.count + kBlock - 1) / kBlock;
    const uint32_t blocks = need < resident ? need : resident;
    // Stack bound: one entry per binary level for the binary walks, three
    // per 4-wide level for the wide walk.
    uint32_t entries = is_wide(kOrder) ? 3u * p.s.wide4Depth + 2u : 128u;
    constexpr uint32_t kLds = (uint32_t)lds_entries<kSem, kOrder, kAny, kCount>();
    p.overflowStride = entries > kLds ? entries - kLds : 1u;
    hipError_t e = ws.reserve((size_t)blocks * kBlock * p.overflowStride);
    if (e != hipSuccess) return e;
    p.workCounter = ws.counter;
    p.overflow = ws.overflow;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, ws.counter);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, stream, p);
    return hipGetLastError();
}

template <int kSem, int kOrder>
hipError_t dispatch_flags(TraceParams& p, TraceWorkspace& ws, bool any, bool count, hipStream_t stream) {
    if (any) {
        return count ? launch_persistent<kSem, kOrder, true, true>(p, ws, stream)
                     : launch_persistent<kSem, kOrder, true, false>(p, ws, stream);
    }
    return count ? launch_persistent<kSem, kOrder, false, true>(p, ws, stream)
                 : launch_persistent<kSem, kOrder, false, false>(p, ws, stream);
}

}  // namespace

hipError_t TraceWorkspace::reserve(size_t overflowEntries) {
    if (counter == nullptr) {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&counter), 256);
        if (e != hipSuccess) return e;
    }
    if (overflowEntries > overflowCapacity) {
        if (overflow != nullptr) {
            // Launches on any stream may still read the old region.
            hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) return e;
            (void)hipFree(overflow);
            overflow = nullptr;
        }
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&overflow), overflowEntries * sizeof(uint2));
        if (e != hipSuccess) {
            overflowCapacity = 0;
            return e;
        }
        overflowCapacity = overflowEntries;
    }
    return hipSuccess;
}

template <int kSem, int kOrder, bool kAny>
size_t wide_overflow_entries(uint32_t entries) {
    constexpr uint32_t kLds = (uint32_t)lds_entries<kSem, kOrder, kAny, false>();
    const size_t stride = entries > kLds ? entries - kLds : 1u;
    return (size_t)resident_blocks(trace_kernel<kSem, kOrder, kAny, false>()) * kBlock * stride;
}

hipError_t reserve_wide_workspace(TraceWorkspace& ws, uint32_t wide4Depth) {
    const uint32_t entries = 3u * wide4Depth + 2u;
    const size_t need[] = {wide_overflow_entries<kSemQueue, kOrderWide, false>(entries),
                           wide_overflow_entries<kSemMixed, kOrderWide, false>(entries),
                           wide_overflow_entries<kSemShadow, kOrderWide, true>(entries),
                           wide_overflow_entries<kSemV1, kOrderWide, true>(entries),
                           wide_overflow_entries<kSemQueue, kOrderWideQ, false>(entries),
                           wide_overflow_entries<kSemMixed, kOrderWideQ, false>(entries),
                           wide_overflow_entries<kSemShadow, kOrderWideQ, true>(entries),
                           wide_overflow_entries<kSemVis, kOrderWideQ, true>(entries)};
    size_t most = 0;
    for (size_t n : need) most = n > most ? n : most;
    return ws.reserve(most);
}

void TraceWorkspace::release() {
    if (counter) (void)hipFree(counter);
    if (overflow) (void)hipFree(overflow);
    counter = nullptr;
    overflow = nullptr;
    overflowCapacity = 0;
}

hipError_t launch_walk(const SceneView& s, TraceWorkspace& ws, const float* rays, const uint32_t* keys,
                       uint32_t count, bool shadow, uint32_t mode, rex_walk_hit* hits, uint64_t* counters,
                       hipStream_t stream) {
    if (count == 0) return hipSuccess;
    TraceParams p = {};
    p.s = s;
    p.rays = rays;
    p.keys = keys;
    p.hitsWalk = hits;
    p.count = count;
    p.counters = counters;
    const bool cnt = counters != nullptr;
    if (mode == REX_WALK_STACKLESS) {
        const dim3 grid((count + kBlock - 1) / kBlock);
        if (shadow) {
            if (cnt) hipLaunchKernelGGL((k_walk_stackless<true, true>), grid, dim3(kBlock), 0, stream, p);
            else hipLaunchKernelGGL((k_walk_stackless<true, false>), grid, dim3(kBlock), 0, stream, p);
        } else {
            if (cnt) hipLaunchKernelGGL((k_walk_stackless<false, true>), grid, dim3(kBlock), 0, stream, p);
            else hipLaunchKernelGGL((k_walk_stackless<false, false>), grid, dim3(kBlock), 0, stream, p);
        }
        return hipGetLastError();
    }
    if (mode == REX_WALK_WIDE) return dispatch_flags<kSemWalk, kOrderWide>(p, ws, shadow, cnt, stream);
    if (mode == REX_WALK_WIDE_COMPRESSED) return dispatch_flags<kSemWalk, kOrderWideQ>(p, ws, shadow, cnt, stream);
    if (mode == REX_WALK_NEAR_FIRST) return dispatch_flags<kSemWalk, kOrderNearFirst>(p, ws, shadow, cnt, stream);
    return dispatch_flags<kSemWalk, kOrderLeftFirst>(p, ws, shadow, cnt, stream);
}

// ---- allocation-free batch tracer -------------------------------------------
//
// The batch tracer's device scratch (the persistent grid's work counter and
// per-lane stack overflow entries) in a file-scope device array instead of a
// workspace allocated on first use.  The stack of the batch tracer never
// holds more than stackCapacity <= 64 entries, so each lane needs
// 64 - 8 overflow entries (8 live in LDS); the grid is bounded by kStaticBlocks blocks
// (the persistent lanes refill from the work counter, so any ray count is
// traced by that grid).  Each GPU holds its own copy of the array, found by
// hipGetSymbolAddress once per device; host code uses only that address.
constexpr uint32_t kStaticBlocks = 512;
constexpr uint32_t kStaticStride = 64u - (uint32_t)lds_entries<kSemV1, kOrderNearFirst, false, false>();
constexpr size_t kStaticCounterBytes = 256;
constexpr size_t kStaticScratchBytes =
    kStaticCounterBytes + (size_t)kStaticBlocks * kBlock * kStaticStride * sizeof(uint2);
static_assert(kStaticScratchBytes == REX_TRACE_NO_ALLOC_SCRATCH_BYTES, "scratch size stated in rex.h");
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScratchLock;
void* gScratch[kMaxDevices] = {};

hipError_t scratch_address(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchLock);
    if (gScratch[dev] == nullptr) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        gScratch[dev] = ptr;
    }
    *out = gScratch[dev];
    return hipSuccess;
}

template <bool kAny, bool kCount>
hipError_t launch_v1_static(TraceParams& p, void* base, hipStream_t stream) {
    auto kernel = trace_kernel<kSemV1, kOrderNearFirst, kAny, kCount>();
    static uint32_t resident = 0;
    if (resident == 0) resident = resident_blocks(kernel);
    const uint32_t need = (p.count + kBlock - 1) / kBlock;
    uint32_t blocks = need < resident ? need : resident;
    blocks = blocks < kStaticBlocks ? blocks : kStaticBlocks;
    p.workCounter = static_cast<uint32_t*>(base);
    p.overflow = reinterpret_cast<uint2*>(static_cast<unsigned char*>(base) + kStaticCounterBytes);
    p.overflowStride = kStaticStride;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, p.workCounter);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, stream, p);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_trace_v1_static(const SceneView& s, const rex_ray_v1* rays, rex_hit_v1* hits, uint32_t count,
                                  bool any, uint32_t stackCapacity, uint64_t* counters, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    void* base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess) return e;
    TraceParams p = {};
    p.s = s;
    p.raysV1 = rays;
    p.hitsV1 = hits;
    p.count = count;
    p.stackCapacity = stackCapacity;
    p.counters = counters;
    if (any) {
        return counters != nullptr ? launch_v1_static<true, true>(p, base, stream)
                                   : launch_v1_static<true, false>(p, base, stream);
    }
    return counters != nullptr ? launch_v1_static<false, true>(p, base, stream)
                               : launch_v1_static<false, false>(p, base, stream);
}

hipError_t launch_trace_v1(const SceneView& s, TraceWorkspace& ws, const rex_ray_v1* rays, rex_hit_v1* hits,
                           uint32_t count, bool any, uint32_t stackCapacity, uint64_t* counters, hipStream_t stream,
                           bool wide) {
    if (count == 0) return hipSuccess;
    TraceParams p = {};
    p.s = s;
    p.raysV1 = rays;
    p.hitsV1 = hits;
    p.count = count;
    p.stackCapacity = stackCapacity;
    p.counters = counters;
    if (wide) return dispatch_flags<kSemV1, kOrderWide>(p, ws, any, counters != nullptr, stream);
    return dispatch_flags<kSemV1, kOrderNearFirst>(p, ws, any, counters != nullptr, stream);
}

hipError_t launch_visibility_v1(const SceneView& s, TraceWorkspace& ws, const rex_ray_v1* rays, rex_hit_v1* hits,
                                uint32_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    TraceParams p = {};
    p.s = s;
    p.raysV1 = rays;
    p.hitsV1 = hits;
    p.count = count;
    p.stackCapacity = 64u;
    return launch_persistent<kSemVis, kOrderWideQ, true, false>(p, ws, stream);
}

}  // namespace rex

namespace rex {

hipError_t launch_queue_closest(const SceneView& s, TraceWorkspace& ws, const F4* rays, const uint32_t* countPtr,
                                uint32_t capacity, QueueHit* hits, uint32_t order, uint64_t* counters,
                                hipStream_t stream) {
    if (capacity == 0) return hipSuccess;
    TraceParams p = {};
    p.s = s;
    p.rays = reinterpret_cast<const float*>(rays);
    p.queueHits = hits;
    p.countPtr = countPtr;
    p.count = capacity;
    p.counters = counters;
    const bool cnt = counters != nullptr;
    if (order == kOrderWideQ) {
        return cnt ? launch_persistent<kSemQueue, kOrderWideQ, false, true>(p, ws, stream)
                   : launch_persistent<kSemQueue, kOrderWideQ, false, false>(p, ws, stream);
    }
    if (order == kOrderWide) {
        return cnt ? launch_persistent<kSemQueue, kOrderWide, false, true>(p, ws, stream)
                   : launch_persistent<kSemQueue, kOrderWide, false, false>(p, ws, stream);
    }
    if (order == kOrderNearFirst) {
        return cnt ? launch_persistent<kSemQueue, kOrderNearFirst, false, true>(p, ws, stream)
                   : launch_persistent<kSemQueue, kOrderNearFirst, false, false>(p, ws, stream);
    }
    return cnt ? launch_persistent<kSemQueue, kOrderLeftFirst, false, true>(p, ws, stream)
               : launch_persistent<kSemQueue, kOrderLeftFirst, false, false>(p, ws, stream);
}

hipError_t launch_queue_mixed(const SceneView& s, TraceWorkspace& ws, const ShadowItem* items,
                              const uint32_t* shadowCountPtr, const F4* rays, const uint32_t* rayCountPtr,
                              uint32_t capacity, QueueHit* hits, PathState* paths, uint32_t order,
                              hipStream_t stream) {
    if (capacity == 0) return hipSuccess;
    TraceParams p = {};
    p.s = s;
    p.rays = reinterpret_cast<const float*>(rays);
    p.queueHits = hits;
    p.shadows = items;
    p.paths = paths;
    p.countPtr = rayCountPtr;
    p.countPtr2 = shadowCountPtr;
    p.count = capacity;
    if (order == kOrderWideQ) return launch_persistent<kSemMixed, kOrderWideQ, false, false>(p, ws, stream);
    if (order == kOrderWide) return launch_persistent<kSemMixed, kOrderWide, false, false>(p, ws, stream);
    if (order == kOrderNearFirst) return launch_persistent<kSemMixed, kOrderNearFirst, false, false>(p, ws, stream);
    return launch_persistent<kSemMixed, kOrderLeftFirst, false, false>(p, ws, stream);
}

hipError_t launch_queue_mixed_arena(const SceneView& s, TraceWorkspace& ws, const ShadowItem* items,
                                    const uint32_t* shadowCountPtr, const uint32_t* shadowSceneOf, const F4* rays,
                                    const uint32_t* rayCountPtr, const uint32_t* raySceneOf, const SceneRoot* roots,
                                    uint32_t capacity, QueueHit* hits, PathState* paths, hipStream_t stream) {
    if (capacity == 0) return hipSuccess;
    TraceParams p = {};
    p.s = s;
    p.rays = reinterpret_cast<const float*>(rays);
    p.queueHits = hits;
    p.shadows = items;
    p.paths = paths;
    p.countPtr = rayCountPtr;
    p.countPtr2 = shadowCountPtr;
    p.count = capacity;
    p.raySceneOf = raySceneOf;
    p.shadowSceneOf = shadowSceneOf;
    p.roots = roots;
    return launch_persistent<kSemMixed, kOrderWideQ, false, false>(p, ws, stream);
}

hipError_t launch_queue_shadow(const SceneView& s, TraceWorkspace& ws, const ShadowItem* items,
                               const uint32_t* countPtr, uint32_t capacity, PathState* paths, uint32_t order,
                               uint64_t* counters, hipStream_t stream) {
    if (capacity == 0) return hipSuccess;
    TraceParams p = {};
    p.s = s;
    p.shadows = items;
    p.paths = paths;
    p.countPtr = countPtr;
    p.count = capacity;
    p.counters = counters;
    const bool cnt = counters != nullptr;
    if (order == kOrderWideQ) {
        return cnt ? launch_persistent<kSemShadow, kOrderWideQ, true, true>(p, ws, stream)
                   : launch_persistent<kSemShadow, kOrderWideQ, true, false>(p, ws, stream);
    }
    if (order == kOrderWide) {
        return cnt ? launch_persistent<kSemShadow, kOrderWide, true, true>(p, ws, stream)
                   : launch_persistent<kSemShadow, kOrderWide, true, false>(p, ws, stream);
    }
    if (order == kOrderNearFirst) {
        return cnt ? launch_persistent<kSemShadow, kOrderNearFirst, true, true>(p, ws, stream)
                   : launch_persistent<kSemShadow, kOrderNearFirst, true, false>(p, ws, stream);
    }
    return cnt ? launch_persistent<kSemShadow, kOrderLeftFirst, true, true>(p, ws, stream)
               : launch_persistent<kSemShadow, kOrderLeftFirst, true, false>(p, ws, stream);
}

}  // namespace rex
