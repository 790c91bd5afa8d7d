// Traversal kernels: the batch tracer behind rex_trace_closest / rex_trace_any
// (software_trace_v1.hlsl semantics) and the integrator walk behind
// rex_trace_walk (pbr_software_traversal.hlsli semantics).
//
// Launch shape: one ray per thread, 128-thread blocks (two wave64s), a 1-D
// grid over the batch.  The traversal stack's top entries live in LDS,
// entry-major ([entry][thread]) so the 64 lanes of a wave hit 64 consecutive
// 8-byte words.
#include <hip/hip_runtime.h>

#include "rex_internal.hpp"
#include "traverse.hpp"

namespace rex {

constexpr int kBlock = 128;
constexpr int kLdsEntries = 16;
constexpr int kWalkStackTotal = 128;  // depth <= 128 => at most 127 pending
constexpr int kV1StackTotal = 64;     // software_trace_v1.hlsl:6

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

__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
    const float len = sqrtf(x * x + y * y + z * z);
    x = x / len;
    y = y / len;
    z = z / len;
}

// Geometric and shading normal of a hit (pbr_software_traversal.hlsli:82-118;
// software_trace_v1.hlsl:382-393).
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

template <bool kShadow, bool kCount, uint32_t kMode>
__global__ __launch_bounds__(kBlock) void k_walk(SceneView s, const float* __restrict__ rays,
                                                 const uint32_t* __restrict__ keys, uint32_t count,
                                                 rex_walk_hit* __restrict__ hits, uint64_t* __restrict__ counters) {
    __shared__ uint2 ldsStack[kLdsEntries * kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    WalkCounters c = {};
    uint32_t invalidRay = 0;
    if (i < count) {
        const float4 a = reinterpret_cast<const float4*>(rays)[2u * i];
        const float4 b = reinterpret_cast<const float4*>(rays)[2u * i + 1u];
        uint32_t ignInst = kInvalid, ignPrim = kInvalid;
        if (keys != nullptr) {
            const uint2 k = reinterpret_cast<const uint2*>(keys)[i];
            ignInst = k.x;
            ignPrim = k.y;
        }
        rex_walk_hit out = {};
        out.t = b.w;
        out.triangle = kInvalid;
        out.instanceId = kInvalid;
        out.primitiveId = kInvalid;
        out.materialIndex = kInvalid;
        if (s.nodeCount == 0u || s.triCount == 0u) {
            out.result = 0;
        } else if (!walk_ray_valid(a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w)) {
            invalidRay = 1;
            out.result = 2;
        } else {
            const RayF r = make_ray(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w);
            WalkResult w;
            if (kMode == REX_WALK_STACKLESS) {
                w = walk_stackless<kShadow, kCount>(s, r, ignInst, ignPrim, c);
            } else {
                ThreadStack<kLdsEntries, kWalkStackTotal> st;
                st.lds = ldsStack;
                st.stride = kBlock;
                st.lane = threadIdx.x;
                w = walk_short_stack<kShadow, kCount, kLdsEntries, kWalkStackTotal>(s, r, ignInst, ignPrim, st, c);
            }
            out.result = w.result;
            if (!kShadow && w.result == 1u) {
                float gn[3], sn[3];
                TriAttr attr;
                hit_normals(s, w.tri, w.u, w.v, w.front, gn, sn, attr);
                out.t = w.t;
                out.u = w.u;
                out.v = w.v;
                out.frontFace = w.front ? 1u : 0u;
                out.triangle = w.tri;
                out.position[0] = r.ox + r.dx * w.t;
                out.position[1] = r.oy + r.dy * w.t;
                out.position[2] = r.oz + r.dz * w.t;
                for (int k = 0; k < 3; ++k) {
                    out.geometricNormal[k] = gn[k];
                    out.shadingNormal[k] = sn[k];
                }
                out.materialIndex = attr.ids.w;
                out.instanceId = attr.ids.x;
                out.primitiveId = attr.ids.y;
            }
        }
        hits[i] = out;
    }
    if (kCount && counters != nullptr) flush_counters(counters, c, invalidRay, 0u, 0u);
}

// ---- software_trace_v1.hlsl ------------------------------------------------

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

template <bool kAny, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_v1(SceneView s, const rex_ray_v1* __restrict__ rays,
                                                     rex_hit_v1* __restrict__ hits, uint32_t rayCount,
                                                     uint32_t stackCapacity, uint64_t* __restrict__ counters) {
    __shared__ uint2 ldsStack[kLdsEntries * kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    WalkCounters c = {};
    uint32_t invalidRay = 0, invalidHit = 0, overflow = 0;
    if (i < rayCount) {
        const float4 a = reinterpret_cast<const float4*>(rays + i)[0];
        const float4 b = reinterpret_cast<const float4*>(rays + i)[1];
        const uint4 id = reinterpret_cast<const uint4*>(rays + i)[2];
        rex_hit_v1* out = hits + i;
        if (!v1_ray_valid(a, b, id.w) || stackCapacity == 0u || stackCapacity > (uint32_t)kV1StackTotal) {
            invalidRay = 1;
            v1_terminal(out, b.w, id.x, id.y, REX_HIT_KIND_INVALID);
        } else if (id.w == 0u || s.nodeCount == 0u) {
            v1_terminal(out, b.w, id.x, id.y, REX_HIT_KIND_MISS);
        } else {
            const RayF r = make_ray(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w);
            ThreadStack<kLdsEntries, kV1StackTotal> st;
            st.lds = ldsStack;
            st.stride = kBlock;
            st.lane = threadIdx.x;
            bool hasHit = false;
            float best = r.tmax;
            float bu = 0.0f, bv = 0.0f;
            bool bf = false;
            uint32_t bt = kInvalid;
            uint32_t kind = REX_HIT_KIND_MISS;
            float rootNear;
            if (kCount) { c.nodeTests++; c.boxTests++; }
            if (slab_indep(r, s.rootMin[0], s.rootMin[1], s.rootMin[2], s.rootMax[0], s.rootMax[1], s.rootMax[2],
                           rootNear) &&
                rootNear < r.tmax) {
                uint32_t sp = 0;
                st.put(sp++, make_uint2(s.rootRef, f2u(fmaxf(rootNear, r.tmin))));
                if (kCount) c.maxStack = 1;
                uint32_t visited = 0;
                bool done = false;
                while (sp != 0u && !done) {
                    --sp;
                    if (++visited > s.nodeCount) {
                        invalidHit = 1;
                        kind = REX_HIT_KIND_INVALID;
                        break;
                    }
                    const uint2 e = st.get(sp);
                    const float nn = u2f(e.y);
                    if (hasHit ? nn > best : nn >= best) continue;
                    const uint32_t ref = e.x;
                    if (is_leaf_ref(ref)) {
                        const uint32_t first = ref & kLeafFirstMask;
                        const uint32_t cnt = ((ref >> 28) & 7u) + 1u;
                        if (kCount) {
                            c.leafVisits++;
                            c.leafPrims += cnt;
                            c.maxLeaf = cnt > c.maxLeaf ? cnt : c.maxLeaf;
                        }
                        for (uint32_t k = 0; k < cnt; ++k) {
                            const uint32_t ti = first + k;
                            if (kCount) c.triTests++;
                            const TriPos tp = load_tri(s.tris, ti);
                            TriHit h;
                            if (!tri_test(r, tp, best, hasHit, h)) continue;
                            const uint32_t flags = f2u(tp.p2.w);
                            if (!h.front && (flags & kTriDoubleSided) == 0u) continue;
                            bool replace = !hasHit || h.t < best;
                            if (!replace && h.t == best) {
                                replace = s.attrs[ti].meta.z < s.attrs[bt].meta.z;
                            }
                            if (kAny || replace) {
                                hasHit = true;
                                best = h.t;
                                bu = h.u;
                                bv = h.v;
                                bf = h.front;
                                bt = ti;
                            }
                            if (kAny) break;
                        }
                        if (hasHit && kAny) done = true;
                        continue;
                    }
                    const WideNode n = load_wide(s.wide, ref);
                    float ln, rn;
                    if (kCount) { c.nodeTests += 2; c.boxTests += 2; }
                    const bool hl = slab_indep(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, ln) &&
                                    within_limit(ln, best, hasHit);
                    const bool hr = slab_indep(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, rn) &&
                                    within_limit(rn, best, hasHit);
                    const uint32_t required = (hl ? 1u : 0u) + (hr ? 1u : 0u);
                    if (sp + required > stackCapacity) {
                        overflow = 1;
                        kind = REX_HIT_KIND_INVALID;
                        break;
                    }
                    const float lnc = fmaxf(ln, r.tmin), rnc = fmaxf(rn, r.tmin);
                    if (hl && hr) {
                        // The left child's binary index is always below the
                        // right one's in pre-order, so ties go left first.
                        const bool leftFirst = lnc < rnc || !(rnc < lnc);
                        st.put(sp++, leftFirst ? make_uint2(n.link.y, f2u(rnc)) : make_uint2(n.link.x, f2u(lnc)));
                        st.put(sp++, leftFirst ? make_uint2(n.link.x, f2u(lnc)) : make_uint2(n.link.y, f2u(rnc)));
                    } else if (hl) {
                        st.put(sp++, make_uint2(n.link.x, f2u(lnc)));
                    } else if (hr) {
                        st.put(sp++, make_uint2(n.link.y, f2u(rnc)));
                    }
                    if (kCount && sp > c.maxStack) c.maxStack = sp;
                }
                if (kind != REX_HIT_KIND_INVALID) kind = hasHit ? REX_HIT_KIND_TRIANGLE : REX_HIT_KIND_MISS;
            }
            if (kind != REX_HIT_KIND_TRIANGLE) {
                v1_terminal(out, r.tmax, id.x, id.y, kind);
            } else {
                float gn[3], sn[3];
                TriAttr attr;
                hit_normals(s, bt, bu, bv, bf, gn, sn, attr);
                rex_hit_v1 h = {};
                h.positionT[0] = r.ox + r.dx * best;
                h.positionT[1] = r.oy + r.dy * best;
                h.positionT[2] = r.oz + r.dz * best;
                h.positionT[3] = best;
                h.geometricNormalBaryU[0] = gn[0];
                h.geometricNormalBaryU[1] = gn[1];
                h.geometricNormalBaryU[2] = gn[2];
                h.geometricNormalBaryU[3] = bu;
                h.shadingNormalBaryV[0] = sn[0];
                h.shadingNormalBaryV[1] = sn[1];
                h.shadingNormalBaryV[2] = sn[2];
                h.shadingNormalBaryV[3] = bv;
                h.ids[0] = attr.ids.x;
                h.ids[1] = attr.ids.y;
                h.ids[2] = attr.ids.z;
                h.ids[3] = attr.ids.w;
                h.metadata[0] = id.x;
                h.metadata[1] = REX_HIT_KIND_TRIANGLE;
                h.metadata[2] = bf ? 1u : 0u;
                h.metadata[3] = id.y;
                *out = h;
            }
        }
    }
    if (kCount && counters != nullptr) flush_counters(counters, c, invalidRay, invalidHit, overflow);
}

// ---- host launchers ----------------------------------------------------------

hipError_t launch_walk(const SceneView& s, const float* rays, const uint32_t* keys, uint32_t count, bool shadow,
                       uint32_t mode, rex_walk_hit* hits, uint64_t* counters, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const dim3 grid((count + kBlock - 1) / kBlock);
    const bool cnt = counters != nullptr;
#define REX_WALK(SH, CN, MD) hipLaunchKernelGGL((k_walk<SH, CN, MD>), grid, dim3(kBlock), 0, stream, s, rays, keys, count, hits, counters)
    if (mode == REX_WALK_STACKLESS) {
        if (shadow) {
            if (cnt) REX_WALK(true, true, REX_WALK_STACKLESS); else REX_WALK(true, false, REX_WALK_STACKLESS);
        } else {
            if (cnt) REX_WALK(false, true, REX_WALK_STACKLESS); else REX_WALK(false, false, REX_WALK_STACKLESS);
        }
    } else {
        if (shadow) {
            if (cnt) REX_WALK(true, true, REX_WALK_SHORT_STACK); else REX_WALK(true, false, REX_WALK_SHORT_STACK);
        } else {
            if (cnt) REX_WALK(false, true, REX_WALK_SHORT_STACK); else REX_WALK(false, false, REX_WALK_SHORT_STACK);
        }
    }
#undef REX_WALK
    return hipGetLastError();
}

hipError_t launch_trace_v1(const SceneView& s, const rex_ray_v1* rays, rex_hit_v1* hits, uint32_t count,
                           bool any, uint32_t stackCapacity, uint64_t* counters, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const dim3 grid((count + kBlock - 1) / kBlock);
    const bool cnt = counters != nullptr;
#define REX_V1(AN, CN) hipLaunchKernelGGL((k_trace_v1<AN, CN>), grid, dim3(kBlock), 0, stream, s, rays, hits, count, stackCapacity, counters)
    if (any) {
        if (cnt) REX_V1(true, true); else REX_V1(true, false);
    } else {
        if (cnt) REX_V1(false, true); else REX_V1(false, false);
    }
#undef REX_V1
    return hipGetLastError();
}

}  // namespace rex
