/*
 * Oracle (test infrastructure only): fp32 ray traversal over the flattened
 * SAH BVH.
 *
 * Two walks, each restated from the reference shader text:
 *   ora_trace_walk      -- the integrator's stackless parent-pointer walk,
 *                          resources/shaders/integrators/pbr_software_traversal.hlsli:287-576,
 *                          with slab / Moller-Trumbore tests from
 *                          resources/shaders/traversal/software_common.hlsli:63-172;
 *   ora_trace_batch_v1  -- the near-first stack tracer behind
 *                          IGpuTraversalBackend::RecordTrace{Closest,Any}Batch,
 *                          resources/shaders/traversal/software_trace_v1.hlsl:67-411.
 * The slab test multiplies by 1/d as the shader does
 * (software_common.hlsli:96-98), not the division of the C++ mirror
 * (src/rt/software_gpu/SoftwareGpu.cpp:627-628).  max/min in the slab test
 * follow IEEE maxNum/minNum (fmaxf/fminf).
 *
 * Alpha-masked geometry is unsupported (no alpha atlas on this path): such a
 * candidate is reported through the reference's fail-closed invalid encoding.
 *
 * Parity pins (tests/test_oracle_trace.py):
 *   - FixedHitFixtureV1 on the canonical triangle: hit t = 2, (u, v) = (0.3, 0.4),
 *     front face; miss; invalid ray (src/rt/gpu/GpuTraversalContracts.cpp:252-298);
 *   - shared-edge quad -> lower primitive id wins (tests/unit/L3/BvhTests.cpp:28-50);
 *   - the L4 12-triangle grid corpus (tests/gpu/L4/TestMain.cpp:29-79, 188-218);
 *   - both walks == brute force (argmin of (t, key)) on seeded random rays,
 *     as tests/statistical/L3/TraversalParityTests.cpp:62-134 does;
 *   - the reference's own fp32 CPU mirror (SoftwareGpu::Trace / BruteForce,
 *     compiled into oracle/_ref) on hit IDs, with t within the
 *     division-vs-reciprocal rounding difference.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } f3;

static f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static f3 cross3(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* normalize(x) = x / sqrt(dot(x, x)), each component divided once. */
static f3 normalize3(f3 a) {
    const float len = sqrtf(dot3(a, a));
    return mk(a.x / len, a.y / len, a.z / len);
}
static f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
static float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

/* Flag bits, include/contracts/SceneAbiV0.hpp:16-44. */
#define GEOM_ALPHA_MASK (1u << 1)
#define GEOM_DOUBLE_SIDED (1u << 2)
#define INST_CASTS_SHADOW (1u << 1)
#define MAT_DOUBLE_SIDED (1u << 0)
#define MAT_ALPHA_MASK (1u << 1)

/* software_common.hlsli:63-70 */
static int ray_valid_walk(f3 o, f3 d, float tMin, float tMax) {
    const float len2 = dot3(d, d);
    return isfinite(o.x) && isfinite(o.y) && isfinite(o.z) && isfinite(d.x) && isfinite(d.y) &&
           isfinite(d.z) && isfinite(len2) && fabsf(len2 - 1.0f) <= 1.0e-4f && !isnan(tMin) &&
           !isnan(tMax) && tMin < tMax;
}

/* software_common.hlsli:72-123 */
static int intersect_bounds(f3 o, f3 d, float tMin, float tMax, const float* bmin,
                            const float* bmax, float limit, int inclusive, float* nearOut,
                            float* intervalNearOut) {
    float intervalNear = -INFINITY;
    float intervalFar = INFINITY;
    for (int axis = 0; axis < 3; ++axis) {
        const float direction = comp(d, axis);
        const float origin = comp(o, axis);
        if (direction == 0.0f) {
            if (origin < bmin[axis] || origin > bmax[axis]) { *nearOut = 0.0f; return 0; }
            continue;
        }
        const float inv = 1.0f / direction;
        float axisNear = (bmin[axis] - origin) * inv;
        float axisFar = (bmax[axis] - origin) * inv;
        if (axisNear > axisFar) { const float t = axisNear; axisNear = axisFar; axisFar = t; }
        intervalNear = fmaxf(intervalNear, axisNear);
        intervalFar = fminf(intervalFar, axisFar);
        if (intervalFar < intervalNear) { *nearOut = 0.0f; return 0; }
    }
    if (intervalNearOut) *intervalNearOut = intervalNear;
    const int withinRay = intervalNear < tMax;
    const int withinLimit = inclusive ? intervalNear <= limit : intervalNear < limit;
    if (!(intervalFar > tMin && withinRay && withinLimit)) { *nearOut = 0.0f; return 0; }
    *nearOut = fmaxf(intervalNear, tMin);
    return 1;
}

/* software_common.hlsli:125-172 */
static int intersect_triangle(f3 o, f3 d, float tMin, float tMax, f3 v0, f3 v1, f3 v2,
                              float limit, int inclusive, float* tOut, float* uOut,
                              float* vOut, int* frontOut) {
    const f3 edge1 = sub(v1, v0);
    const f3 edge2 = sub(v2, v0);
    const f3 p = cross3(d, edge2);
    const float det = dot3(edge1, p);
    *tOut = 0.0f; *uOut = 0.0f; *vOut = 0.0f; *frontOut = 0;
    if (det == 0.0f || !isfinite(det)) return 0;
    const float invDet = 1.0f / det;
    const f3 tr = sub(o, v0);
    const float u = dot3(tr, p) * invDet;
    *uOut = u;
    if (!(u >= 0.0f && u <= 1.0f)) { *uOut = u; return 0; }
    const f3 q = cross3(tr, edge1);
    const float v = dot3(d, q) * invDet;
    *vOut = v;
    if (!(v >= 0.0f && u + v <= 1.0f)) return 0;
    const float t = dot3(edge2, q) * invDet;
    *tOut = t;
    *frontOut = det > 0.0f;
    const int withinRay = t < tMax;
    const int withinLimit = inclusive ? t <= limit : t < limit;
    return isfinite(t) && t > tMin && withinRay && withinLimit;
}

static void cnt_add(uint64_t* c, int idx, uint64_t v) { if (c) c[idx] += v; }
static void cnt_max(uint64_t* c, int idx, uint64_t v) { if (c && v > c[idx]) c[idx] = v; }

static void walk_init(ora_walk_hit* h, float tMax) {
    memset(h, 0, sizeof(*h));
    h->t = tMax;
    h->triangle = ORA_INVALID;
    h->instanceId = ORA_INVALID;
    h->primitiveId = ORA_INVALID;
    h->materialIndex = ORA_INVALID;
}

/* pbr_software_traversal.hlsli:82-118 (emitter lookup is done by the caller) */
static void walk_populate(ora_walk_hit* h, const ora_tri* tris, f3 o, f3 d, float t, float u,
                          float v, int front, uint32_t tri) {
    const ora_tri* b = &tris[tri];
    const float w0 = 1.0f - u - v;
    f3 gn = normalize3(cross3(sub(ld3(b->pos[1]), ld3(b->pos[0])), sub(ld3(b->pos[2]), ld3(b->pos[0]))));
    f3 sn = normalize3(add(add(scl(ld3(b->nrm[0]), w0), scl(ld3(b->nrm[1]), u)), scl(ld3(b->nrm[2]), v)));
    if (!front) { gn = mk(-gn.x, -gn.y, -gn.z); sn = mk(-sn.x, -sn.y, -sn.z); }
    h->result = 1;
    h->t = t; h->u = u; h->v = v; h->frontFace = front ? 1u : 0u;
    h->triangle = tri;
    const f3 p = add(o, scl(d, t));
    h->position[0] = p.x; h->position[1] = p.y; h->position[2] = p.z;
    h->geometricNormal[0] = gn.x; h->geometricNormal[1] = gn.y; h->geometricNormal[2] = gn.z;
    h->shadingNormal[0] = sn.x; h->shadingNormal[1] = sn.y; h->shadingNormal[2] = sn.z;
    h->materialIndex = b->identity[3];
    h->instanceId = b->identity[0];
    h->primitiveId = b->identity[1];
}

static void walk_invalid(ora_walk_hit* h, float tMax, uint64_t* c) {
    cnt_add(c, ORA_CNT_INVALID_HIT, 1);
    walk_init(h, tMax);
    h->result = 2;
}

void ora_trace_walk(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris,
                    uint32_t triCount, const float origin[3], const float direction[3],
                    float tMin, float tMax, uint32_t ignoredInstance, uint32_t ignoredPrimitive,
                    int shadowQuery, ora_walk_hit* hit, uint64_t* c) {
    walk_init(hit, tMax);
    if (nodeCount == 0u || triCount == 0u) return;
    const f3 o = ld3(origin);
    const f3 d = ld3(direction);
    if (!ray_valid_walk(o, d, tMin, tMax)) {
        cnt_add(c, ORA_CNT_INVALID_RAY, 1);
        walk_init(hit, tMax);
        hit->result = 2; /* "true" with invalid material: the fail-closed encoding */
        return;
    }
    int found = 0;
    float bestDistance = tMax, bestU = 0.0f, bestV = 0.0f;
    int bestFront = 0;
    uint32_t bestTri = ORA_INVALID, bestInst = ORA_INVALID, bestPrim = ORA_INVALID;
    uint32_t previous = ORA_INVALID, nodeIndex = 0u, steps = 0u;
    const uint32_t maxSteps = nodeCount <= 0x55555555u ? nodeCount * 3u : 0xffffffffu;
    while (nodeIndex != ORA_INVALID) {
        if (nodeIndex >= nodeCount || ++steps > maxSteps) { walk_invalid(hit, tMax, c); return; }
        cnt_add(c, ORA_CNT_NODE_TESTS, 1);
        const ora_node* node = &nodes[nodeIndex];
        const uint32_t parent = node->links[3];
        uint32_t next = ORA_INVALID;
        if (previous == parent) {
            float nodeNear = 0.0f;
            cnt_add(c, ORA_CNT_BOX_TESTS, 1);
            const int hb = intersect_bounds(o, d, tMin, tMax, node->bmin, node->bmax, bestDistance,
                                            found, &nodeNear, NULL);
            if (!hb) {
                next = parent;
            } else if (node->links[1] != 0u) {
                if (node->links[0] > triCount || node->links[1] > triCount - node->links[0]) {
                    walk_invalid(hit, tMax, c);
                    return;
                }
                cnt_add(c, ORA_CNT_LEAF_VISITS, 1);
                cnt_add(c, ORA_CNT_LEAF_PRIMITIVES, node->links[1]);
                cnt_max(c, ORA_CNT_MAX_LEAF_OCCUPANCY, node->links[1]);
                for (uint32_t off = 0; off < node->links[1]; ++off) {
                    const uint32_t ti = node->links[0] + off;
                    const ora_tri* cand = &tris[ti];
                    if (cand->identity[0] == ignoredInstance && cand->identity[1] == ignoredPrimitive) continue;
                    if (shadowQuery && (cand->metadata[3] & INST_CASTS_SHADOW) == 0u) continue;
                    cnt_add(c, ORA_CNT_TRIANGLE_TESTS, 1);
                    float dist, bu, bv;
                    int front;
                    if (!intersect_triangle(o, d, tMin, tMax, ld3(cand->pos[0]), ld3(cand->pos[1]),
                                            ld3(cand->pos[2]), bestDistance, found, &dist, &bu, &bv,
                                            &front))
                        continue;
                    const int doubleSided = (cand->metadata[0] & GEOM_DOUBLE_SIDED) != 0u ||
                                            (cand->metadata[1] & MAT_DOUBLE_SIDED) != 0u;
                    if (!front && !doubleSided) continue;
                    const int alphaMasked = (cand->metadata[0] & GEOM_ALPHA_MASK) != 0u ||
                                            (cand->metadata[1] & MAT_ALPHA_MASK) != 0u;
                    if (alphaMasked) {
                        if ((cand->metadata[0] & ORA_ALPHA_SKIP) != 0u) continue;
                        if ((cand->metadata[0] & ORA_ALPHA_ACCEPT) == 0u) { walk_invalid(hit, tMax, c); return; }
                    }
                    if (shadowQuery) { hit->result = 1; return; }
                    const int eqLower = dist == bestDistance &&
                                        (cand->identity[0] < bestInst ||
                                         (cand->identity[0] == bestInst && cand->identity[1] < bestPrim));
                    if (!found || dist < bestDistance || eqLower) {
                        found = 1;
                        bestDistance = dist; bestU = bu; bestV = bv; bestFront = front;
                        bestTri = ti; bestInst = cand->identity[0]; bestPrim = cand->identity[1];
                    }
                }
                next = parent;
            } else {
                if (node->links[0] >= nodeCount || node->links[2] >= nodeCount) {
                    walk_invalid(hit, tMax, c);
                    return;
                }
                next = node->links[0];
            }
        } else if (node->links[1] == 0u && previous == node->links[0]) {
            if (node->links[2] >= nodeCount) { walk_invalid(hit, tMax, c); return; }
            next = node->links[2];
        } else if (node->links[1] == 0u && previous == node->links[2]) {
            next = parent;
        } else {
            walk_invalid(hit, tMax, c);
            return;
        }
        previous = nodeIndex;
        nodeIndex = next;
    }
    if (!found) return;
    walk_populate(hit, tris, o, d, bestDistance, bestU, bestV, bestFront, bestTri);
}

void ora_trace_walk_near_first(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris,
                               uint32_t triCount, const float origin[3], const float direction[3],
                               float tMin, float tMax, uint32_t ignoredInstance,
                               uint32_t ignoredPrimitive, int shadowQuery, ora_walk_hit* hit,
                               uint64_t* c) {
    walk_init(hit, tMax);
    if (nodeCount == 0u || triCount == 0u) return;
    const f3 o = ld3(origin);
    const f3 d = ld3(direction);
    if (!ray_valid_walk(o, d, tMin, tMax)) {
        cnt_add(c, ORA_CNT_INVALID_RAY, 1);
        hit->result = 2;
        return;
    }
    int found = 0;
    float best = tMax, bestU = 0.0f, bestV = 0.0f;
    int bestFront = 0;
    uint32_t bestTri = ORA_INVALID, bestInst = ORA_INVALID, bestPrim = ORA_INVALID;
    uint32_t stackNodes[256];
    float stackNear[256];
    uint32_t sp = 0;
    float rootNear;
    cnt_add(c, ORA_CNT_BOX_TESTS, 1);
    cnt_add(c, ORA_CNT_NODE_TESTS, 1);
    if (!intersect_bounds(o, d, tMin, tMax, nodes[0].bmin, nodes[0].bmax, best, found, &rootNear, NULL)) return;
    stackNodes[sp] = 0u;
    stackNear[sp++] = rootNear;
    while (sp != 0u) {
        --sp;
        const uint32_t ni = stackNodes[sp];
        const float nn = stackNear[sp];
        if (found ? nn > best : nn >= best) continue;
        const ora_node* node = &nodes[ni];
        if (node->links[1] != 0u) {
            cnt_add(c, ORA_CNT_LEAF_VISITS, 1);
            cnt_add(c, ORA_CNT_LEAF_PRIMITIVES, node->links[1]);
            cnt_max(c, ORA_CNT_MAX_LEAF_OCCUPANCY, node->links[1]);
            for (uint32_t off = 0; off < node->links[1]; ++off) {
                const uint32_t ti = node->links[0] + off;
                const ora_tri* cand = &tris[ti];
                if (cand->identity[0] == ignoredInstance && cand->identity[1] == ignoredPrimitive) continue;
                if (shadowQuery && (cand->metadata[3] & INST_CASTS_SHADOW) == 0u) continue;
                cnt_add(c, ORA_CNT_TRIANGLE_TESTS, 1);
                float dist, bu, bv;
                int front;
                if (!intersect_triangle(o, d, tMin, tMax, ld3(cand->pos[0]), ld3(cand->pos[1]),
                                        ld3(cand->pos[2]), best, found, &dist, &bu, &bv, &front))
                    continue;
                const int doubleSided = (cand->metadata[0] & GEOM_DOUBLE_SIDED) != 0u ||
                                        (cand->metadata[1] & MAT_DOUBLE_SIDED) != 0u;
                if (!front && !doubleSided) continue;
                if ((cand->metadata[0] & GEOM_ALPHA_MASK) != 0u || (cand->metadata[1] & MAT_ALPHA_MASK) != 0u) {
                    if ((cand->metadata[0] & ORA_ALPHA_SKIP) != 0u) continue;
                    if ((cand->metadata[0] & ORA_ALPHA_ACCEPT) == 0u) {
                        walk_invalid(hit, tMax, c);
                        return;
                    }
                }
                if (shadowQuery) { hit->result = 1; return; }
                const int eqLower = dist == best && (cand->identity[0] < bestInst ||
                                                     (cand->identity[0] == bestInst && cand->identity[1] < bestPrim));
                if (!found || dist < best || eqLower) {
                    found = 1;
                    best = dist; bestU = bu; bestV = bv; bestFront = front;
                    bestTri = ti; bestInst = cand->identity[0]; bestPrim = cand->identity[1];
                }
            }
            continue;
        }
        const ora_node* L = &nodes[node->links[0]];
        const ora_node* R = &nodes[node->links[2]];
        float ln = 0.0f, rn = 0.0f;
        cnt_add(c, ORA_CNT_BOX_TESTS, 2);
        cnt_add(c, ORA_CNT_NODE_TESTS, 2);
        const int hl = intersect_bounds(o, d, tMin, tMax, L->bmin, L->bmax, best, found, &ln, NULL);
        const int hr = intersect_bounds(o, d, tMin, tMax, R->bmin, R->bmax, best, found, &rn, NULL);
        if (hl && hr) {
            const int leftFirst = ln < rn || (!(rn < ln) && node->links[0] < node->links[2]);
            stackNodes[sp] = leftFirst ? node->links[2] : node->links[0];
            stackNear[sp++] = leftFirst ? rn : ln;
            stackNodes[sp] = leftFirst ? node->links[0] : node->links[2];
            stackNear[sp++] = leftFirst ? ln : rn;
        } else if (hl) {
            stackNodes[sp] = node->links[0];
            stackNear[sp++] = ln;
        } else if (hr) {
            stackNodes[sp] = node->links[2];
            stackNear[sp++] = rn;
        }
        cnt_max(c, ORA_CNT_MAX_STACK, sp);
    }
    if (!found) return;
    walk_populate(hit, tris, o, d, best, bestU, bestV, bestFront, bestTri);
}

/* ---- 4-wide near-first walk (the product's production order) ----------- */

static uint32_t wide_build_rec(const ora_node* nodes, uint32_t b, ora_wide4* out, uint32_t* count,
                               uint32_t level, uint32_t* depth) {
    const uint32_t w = (*count)++;
    if (level > *depth) *depth = level;
    uint32_t k = 0;
    const uint32_t kids[2] = {nodes[b].links[0], nodes[b].links[2]};
    for (int i = 0; i < 4; ++i) { out[w].child[i] = ORA_INVALID; out[w].wide[i] = ORA_INVALID; }
    for (int i = 0; i < 2; ++i) {
        const ora_node* c = &nodes[kids[i]];
        if (c->links[1] != 0u) {
            out[w].child[k++] = kids[i];
        } else {
            out[w].child[k++] = c->links[0];
            out[w].child[k++] = c->links[2];
        }
    }
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t ci = out[w].child[i];
        if (nodes[ci].links[1] == 0u) out[w].wide[i] = wide_build_rec(nodes, ci, out, count, level + 1u, depth);
    }
    return w;
}

uint32_t ora_build_wide4(const ora_node* nodes, uint32_t nodeCount, ora_wide4* out, uint32_t* depth) {
    uint32_t count = 0;
    *depth = 0;
    if (nodeCount == 0u || nodes[0].links[1] != 0u) return 0;
    wide_build_rec(nodes, 0u, out, &count, 1u, depth);
    return count;
}

void ora_trace_walk_wide(const ora_node* nodes, uint32_t nodeCount, const ora_wide4* wide,
                         const ora_tri* tris, uint32_t triCount, const float origin[3],
                         const float direction[3], float tMin, float tMax, uint32_t ignoredInstance,
                         uint32_t ignoredPrimitive, int shadowQuery, ora_walk_hit* hit, uint64_t* c) {
    walk_init(hit, tMax);
    if (nodeCount == 0u || triCount == 0u) return;
    const f3 o = ld3(origin);
    const f3 d = ld3(direction);
    if (!ray_valid_walk(o, d, tMin, tMax)) {
        cnt_add(c, ORA_CNT_INVALID_RAY, 1);
        hit->result = 2;
        return;
    }
    int found = 0;
    float best = tMax, bestU = 0.0f, bestV = 0.0f;
    int bestFront = 0;
    uint32_t bestTri = ORA_INVALID, bestInst = ORA_INVALID, bestPrim = ORA_INVALID;
    /* stack entries: binary node index + wide index (interior) + entry distance */
    uint32_t stB[256], stW[256];
    float stN[256];
    uint32_t sp = 0;
    float rootNear;
    cnt_add(c, ORA_CNT_BOX_TESTS, 1);
    cnt_add(c, ORA_CNT_NODE_TESTS, 1);
    if (!intersect_bounds(o, d, tMin, tMax, nodes[0].bmin, nodes[0].bmax, best, found, &rootNear, NULL)) return;
    uint32_t curB = 0u, curW = nodes[0].links[1] != 0u ? ORA_INVALID : 0u;
    for (;;) {
        const ora_node* node = &nodes[curB];
        if (node->links[1] != 0u) {
            cnt_add(c, ORA_CNT_LEAF_VISITS, 1);
            cnt_add(c, ORA_CNT_LEAF_PRIMITIVES, node->links[1]);
            cnt_max(c, ORA_CNT_MAX_LEAF_OCCUPANCY, node->links[1]);
            for (uint32_t off = 0; off < node->links[1]; ++off) {
                const uint32_t ti = node->links[0] + off;
                const ora_tri* cand = &tris[ti];
                if (cand->identity[0] == ignoredInstance && cand->identity[1] == ignoredPrimitive) continue;
                if (shadowQuery && (cand->metadata[3] & INST_CASTS_SHADOW) == 0u) continue;
                cnt_add(c, ORA_CNT_TRIANGLE_TESTS, 1);
                float dist, bu, bv;
                int front;
                if (!intersect_triangle(o, d, tMin, tMax, ld3(cand->pos[0]), ld3(cand->pos[1]),
                                        ld3(cand->pos[2]), best, found, &dist, &bu, &bv, &front))
                    continue;
                const int doubleSided = (cand->metadata[0] & GEOM_DOUBLE_SIDED) != 0u ||
                                        (cand->metadata[1] & MAT_DOUBLE_SIDED) != 0u;
                if (!front && !doubleSided) continue;
                if ((cand->metadata[0] & GEOM_ALPHA_MASK) != 0u || (cand->metadata[1] & MAT_ALPHA_MASK) != 0u) {
                    if ((cand->metadata[0] & ORA_ALPHA_SKIP) != 0u) continue;
                    if ((cand->metadata[0] & ORA_ALPHA_ACCEPT) == 0u) {
                        walk_invalid(hit, tMax, c);
                        return;
                    }
                }
                if (shadowQuery) { hit->result = 1; return; }
                const int eqLower = dist == best && (cand->identity[0] < bestInst ||
                                                     (cand->identity[0] == bestInst && cand->identity[1] < bestPrim));
                if (!found || dist < best || eqLower) {
                    found = 1;
                    best = dist; bestU = bu; bestV = bv; bestFront = front;
                    bestTri = ti; bestInst = cand->identity[0]; bestPrim = cand->identity[1];
                }
            }
        } else {
            const ora_wide4* w = &wide[curW];
            int hitK[4] = {0, 0, 0, 0};
            float nearK[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int k = 0; k < 4; ++k) {
                if (w->child[k] == ORA_INVALID) continue;
                cnt_add(c, ORA_CNT_BOX_TESTS, 1);
                cnt_add(c, ORA_CNT_NODE_TESTS, 1);
                const ora_node* ch = &nodes[w->child[k]];
                hitK[k] = intersect_bounds(o, d, tMin, tMax, ch->bmin, ch->bmax, best, found, &nearK[k], NULL);
            }
            int rank[4], m = 0;
            for (int k = 0; k < 4; ++k) {
                if (!hitK[k]) continue;
                ++m;
                rank[k] = 0;
                for (int j = 0; j < 4; ++j) {
                    if (j == k || !hitK[j]) continue;
                    if (nearK[j] < nearK[k] || (nearK[j] == nearK[k] && j < k)) rank[k]++;
                }
            }
            if (m != 0) {
                /* push ranks m-1 .. 1 (farthest first), enter rank 0 */
                for (int r = m - 1; r >= 0; --r) {
                    for (int k = 0; k < 4; ++k) {
                        if (!hitK[k] || rank[k] != r) continue;
                        if (r == 0) {
                            curB = w->child[k];
                            curW = w->wide[k];
                        } else {
                            stB[sp] = w->child[k];
                            stW[sp] = w->wide[k];
                            stN[sp++] = nearK[k];
                        }
                    }
                }
                cnt_max(c, ORA_CNT_MAX_STACK, sp);
                continue;
            }
        }
        /* pop the nearest pending child still within the limit */
        int next = 0;
        while (sp != 0u) {
            --sp;
            if (found ? stN[sp] <= best : stN[sp] < best) {
                curB = stB[sp];
                curW = stW[sp];
                next = 1;
                break;
            }
        }
        if (!next) break;
    }
    if (!found) return;
    walk_populate(hit, tris, o, d, best, bestU, bestV, bestFront, bestTri);
}

void ora_trace_walk_batch_order(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris,
                                uint32_t triCount, const float* rays, const uint32_t* keys,
                                uint32_t count, int shadow, int order, ora_walk_hit* hits,
                                uint64_t* counters) {
    ora_wide4* wide = NULL;
    if (order == 2) {
        uint32_t depth;
        wide = (ora_wide4*)malloc(sizeof(ora_wide4) * (nodeCount ? nodeCount : 1u));
        ora_build_wide4(nodes, nodeCount, wide, &depth);
    }
    for (uint32_t i = 0; i < count; ++i) {
        const float* r = rays + 8u * i;
        const uint32_t ki = keys ? keys[2u * i] : ORA_INVALID;
        const uint32_t kp = keys ? keys[2u * i + 1u] : ORA_INVALID;
        if (order == 2) {
            ora_trace_walk_wide(nodes, nodeCount, wide, tris, triCount, r, r + 4, r[3], r[7], ki, kp, shadow,
                                &hits[i], counters);
        } else if (order == 1) {
            ora_trace_walk_near_first(nodes, nodeCount, tris, triCount, r, r + 4, r[3], r[7], ki, kp,
                                      shadow, &hits[i], counters);
        } else {
            ora_trace_walk(nodes, nodeCount, tris, triCount, r, r + 4, r[3], r[7], ki, kp, shadow,
                           &hits[i], counters);
        }
    }
    free(wide);
}

typedef struct walk_job {
    const ora_node* nodes;
    uint32_t nodeCount;
    const ora_tri* tris;
    uint32_t triCount;
    const float* rays;
    const uint32_t* keys;
    uint32_t begin, end;
    int shadow, order;
    const ora_wide4* wide;
    ora_walk_hit* hits;
    uint64_t counters[ORA_CNT_COUNT];
} walk_job;

static void* walk_range(void* arg) {
    walk_job* j = (walk_job*)arg;
    for (uint32_t i = j->begin; i < j->end; ++i) {
        const float* r = j->rays + 8u * i;
        const uint32_t ki = j->keys ? j->keys[2u * i] : ORA_INVALID;
        const uint32_t kp = j->keys ? j->keys[2u * i + 1u] : ORA_INVALID;
        if (j->order == 2) {
            ora_trace_walk_wide(j->nodes, j->nodeCount, j->wide, j->tris, j->triCount, r, r + 4, r[3], r[7], ki, kp,
                                j->shadow, &j->hits[i], j->counters);
        } else if (j->order == 1) {
            ora_trace_walk_near_first(j->nodes, j->nodeCount, j->tris, j->triCount, r, r + 4, r[3], r[7], ki, kp,
                                      j->shadow, &j->hits[i], j->counters);
        } else {
            ora_trace_walk(j->nodes, j->nodeCount, j->tris, j->triCount, r, r + 4, r[3], r[7], ki, kp, j->shadow,
                           &j->hits[i], j->counters);
        }
    }
    return NULL;
}

/* The walk batch over `threads` host threads in contiguous ray ranges; each
 * ray is traced exactly as by ora_trace_walk_batch_order, so the hits do not
 * depend on the thread count.  Counters are summed (maxima taken) over the
 * ranges. */
void ora_trace_walk_batch_mt(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                             const float* rays, const uint32_t* keys, uint32_t count, int shadow, int order,
                             uint32_t threads, ora_walk_hit* hits, uint64_t* counters) {
    if (threads < 1u) threads = 1u;
    if (threads > 256u) threads = 256u;
    ora_wide4* wide = NULL;
    if (order == 2) {
        uint32_t depth;
        wide = (ora_wide4*)malloc(sizeof(ora_wide4) * (nodeCount ? nodeCount : 1u));
        ora_build_wide4(nodes, nodeCount, wide, &depth);
    }
    walk_job* jobs = (walk_job*)calloc(threads, sizeof(walk_job));
    pthread_t* tid = (pthread_t*)calloc(threads, sizeof(pthread_t));
    const uint32_t chunk = (count + threads - 1u) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
        walk_job* j = &jobs[t];
        j->nodes = nodes; j->nodeCount = nodeCount; j->tris = tris; j->triCount = triCount;
        j->rays = rays; j->keys = keys; j->shadow = shadow; j->order = order; j->wide = wide; j->hits = hits;
        j->begin = t * chunk < count ? t * chunk : count;
        j->end = j->begin + chunk < count ? j->begin + chunk : count;
        pthread_create(&tid[t], NULL, walk_range, j);
    }
    for (uint32_t t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    if (counters) {
        for (uint32_t t = 0; t < threads; ++t) {
            for (int k = 0; k < ORA_CNT_COUNT; ++k) {
                const uint64_t v = jobs[t].counters[k];
                if (k == ORA_CNT_MAX_STACK || k == ORA_CNT_MAX_LEAF_OCCUPANCY) {
                    if (v > counters[k]) counters[k] = v;
                } else {
                    counters[k] += v;
                }
            }
        }
    }
    free(jobs);
    free(tid);
    free(wide);
}

void ora_trace_walk_batch(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris,
                          uint32_t triCount, const float* rays, const uint32_t* keys,
                          uint32_t count, int shadow, ora_walk_hit* hits, uint64_t* counters) {
    ora_trace_walk_batch_order(nodes, nodeCount, tris, triCount, rays, keys, count, shadow, 0,
                               hits, counters);
}

void ora_trace_brute(const ora_tri* tris, uint32_t triCount, const float origin[3],
                     const float direction[3], float tMin, float tMax, ora_walk_hit* hit) {
    walk_init(hit, tMax);
    const f3 o = ld3(origin);
    const f3 d = ld3(direction);
    if (!ray_valid_walk(o, d, tMin, tMax)) { hit->result = 2; return; }
    int found = 0;
    float best = tMax, bu = 0, bv = 0;
    int bf = 0;
    uint32_t bt = ORA_INVALID, bi = ORA_INVALID, bp = ORA_INVALID;
    for (uint32_t ti = 0; ti < triCount; ++ti) {
        const ora_tri* cand = &tris[ti];
        float dist, u, v;
        int front;
        /* No distance limit: every triangle is tested against the full ray
         * interval so the result is the argmin over all of them. */
        if (!intersect_triangle(o, d, tMin, tMax, ld3(cand->pos[0]), ld3(cand->pos[1]),
                                ld3(cand->pos[2]), tMax, 0, &dist, &u, &v, &front))
            continue;
        const int doubleSided = (cand->metadata[0] & GEOM_DOUBLE_SIDED) != 0u ||
                                (cand->metadata[1] & MAT_DOUBLE_SIDED) != 0u;
        if (!front && !doubleSided) continue;
        /* Skipped alpha-masked triangles are never hit (an unresolved mask,
         * which makes the walks invalid, is ignored here). */
        if (((cand->metadata[0] & GEOM_ALPHA_MASK) != 0u || (cand->metadata[1] & MAT_ALPHA_MASK) != 0u) &&
            (cand->metadata[0] & ORA_ALPHA_ACCEPT) == 0u)
            continue;
        const int better = !found || dist < best ||
                           (dist == best && (cand->identity[0] < bi ||
                                             (cand->identity[0] == bi && cand->identity[1] < bp)));
        if (better) {
            found = 1; best = dist; bu = u; bv = v; bf = front; bt = ti;
            bi = cand->identity[0]; bp = cand->identity[1];
        }
    }
    if (found) walk_populate(hit, tris, o, d, best, bu, bv, bf, bt);
}

/* ---- software_trace_v1.hlsl ------------------------------------------------ */

#define KIND_MISS 0u
#define KIND_TRIANGLE 1u
#define KIND_INVALID 3u
#define HITFLAG_FRONT 1u
#define HITFLAG_ALPHA 2u

/* software_trace_v1.hlsl:67-75 */
static int ray_valid_v1(const ora_ray_v1* r) {
    const f3 d = ld3(r->directionTMax);
    const float len2 = dot3(d, d);
    for (int k = 0; k < 4; ++k) {
        if (!isfinite(r->originTMin[k]) || !isfinite(r->directionTMax[k])) return 0;
    }
    return isfinite(len2) && fabsf(len2 - 1.0f) <= 1.0e-4f && r->originTMin[3] >= 0.0f &&
           r->originTMin[3] < r->directionTMax[3] && (r->identity[3] & ~0xffu) == 0u;
}

/* software_trace_v1.hlsl:77-87 */
static void terminal_v1(const ora_ray_v1* r, uint32_t kind, ora_hit_v1* h) {
    memset(h, 0, sizeof(*h));
    h->positionT[3] = r->directionTMax[3];
    h->ids[0] = h->ids[1] = h->ids[2] = h->ids[3] = ORA_INVALID;
    h->metadata[0] = r->identity[0];
    h->metadata[1] = kind;
    h->metadata[2] = 0u;
    h->metadata[3] = r->identity[1];
}

typedef struct {
    int hasHit;
    float best, bu, bv;
    int bf;
    uint32_t bt;
} v1_state;

/* software_trace_v1.hlsl:92-187; returns 1 accepted, 0 rejected, -1 invalid */
static int eval_tri_v1(f3 o, f3 d, float tMin, float tMax, const ora_tri* tris, uint32_t ti,
                       uint32_t queryMode, v1_state* s, uint64_t* c) {
    cnt_add(c, ORA_CNT_TRIANGLE_TESTS, 1);
    const ora_tri* cand = &tris[ti];
    float dist, u, v;
    int front;
    if (!intersect_triangle(o, d, tMin, tMax, ld3(cand->pos[0]), ld3(cand->pos[1]),
                            ld3(cand->pos[2]), s->best, s->hasHit, &dist, &u, &v, &front))
        return 0;
    const int doubleSided = (cand->metadata[0] & GEOM_DOUBLE_SIDED) != 0u ||
                            (cand->metadata[1] & MAT_DOUBLE_SIDED) != 0u;
    if (!front && !doubleSided) return 0;
    const int alphaMasked = (cand->metadata[0] & GEOM_ALPHA_MASK) != 0u ||
                            (cand->metadata[1] & MAT_ALPHA_MASK) != 0u;
    if (alphaMasked) {
        if ((cand->metadata[0] & ORA_ALPHA_SKIP) != 0u) return 0;
        if ((cand->metadata[0] & ORA_ALPHA_ACCEPT) == 0u) return -1;
    }
    int replace = !s->hasHit || dist < s->best;
    if (!replace && dist == s->best) replace = cand->metadata[2] < tris[s->bt].metadata[2];
    if (queryMode == 1u || replace) {
        s->hasHit = 1; s->best = dist; s->bu = u; s->bv = v; s->bf = front; s->bt = ti;
    }
    return 1;
}

static void trace_one_v1(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris,
                         uint32_t triCount, const ora_ray_v1* r, uint32_t queryMode,
                         uint32_t stackCapacity, ora_hit_v1* h, uint64_t* c) {
    if (!ray_valid_v1(r) || stackCapacity == 0u || stackCapacity > 64u) {
        cnt_add(c, ORA_CNT_INVALID_RAY, 1);
        terminal_v1(r, KIND_INVALID, h);
        return;
    }
    if (r->identity[3] == 0u || nodeCount == 0u) { terminal_v1(r, KIND_MISS, h); return; }
    const f3 o = ld3(r->originTMin);
    const f3 d = ld3(r->directionTMax);
    const float tMin = r->originTMin[3];
    const float tMax = r->directionTMax[3];
    v1_state s = {0, tMax, 0.0f, 0.0f, 0, ORA_INVALID};
    uint32_t stackNodes[64];
    float stackNear[64];
    uint32_t sp = 0u;
    float rootNear = 0.0f;
    cnt_add(c, ORA_CNT_NODE_TESTS, 1);
    cnt_add(c, ORA_CNT_BOX_TESTS, 1);
    if (!intersect_bounds(o, d, tMin, tMax, nodes[0].bmin, nodes[0].bmax, tMax, 0, &rootNear, NULL)) {
        terminal_v1(r, KIND_MISS, h);
        return;
    }
    stackNodes[sp] = 0u;
    stackNear[sp++] = rootNear;
    cnt_max(c, ORA_CNT_MAX_STACK, sp);
    uint32_t visited = 0u;
    while (sp != 0u) {
        --sp;
        if (++visited > nodeCount) {
            cnt_add(c, ORA_CNT_INVALID_HIT, 1);
            terminal_v1(r, KIND_INVALID, h);
            return;
        }
        const uint32_t ni = stackNodes[sp];
        const float nn = stackNear[sp];
        if (s.hasHit ? nn > s.best : nn >= s.best) continue;
        if (ni >= nodeCount) {
            cnt_add(c, ORA_CNT_INVALID_HIT, 1);
            terminal_v1(r, KIND_INVALID, h);
            return;
        }
        const ora_node* node = &nodes[ni];
        if (node->links[1] != 0u) {
            cnt_add(c, ORA_CNT_LEAF_VISITS, 1);
            cnt_add(c, ORA_CNT_LEAF_PRIMITIVES, node->links[1]);
            cnt_max(c, ORA_CNT_MAX_LEAF_OCCUPANCY, node->links[1]);
            if (node->links[0] > triCount || node->links[1] > triCount - node->links[0]) {
                cnt_add(c, ORA_CNT_INVALID_HIT, 1);
                terminal_v1(r, KIND_INVALID, h);
                return;
            }
            for (uint32_t off = 0; off < node->links[1]; ++off) {
                const int acc = eval_tri_v1(o, d, tMin, tMax, tris, node->links[0] + off, queryMode, &s, c);
                if (acc < 0) {
                    cnt_add(c, ORA_CNT_INVALID_HIT, 1);
                    terminal_v1(r, KIND_INVALID, h);
                    return;
                }
                if (acc > 0 && queryMode == 1u) break;
            }
            if (s.hasHit && queryMode == 1u) break;
            continue;
        }
        if (node->links[0] >= nodeCount || node->links[2] >= nodeCount) {
            cnt_add(c, ORA_CNT_INVALID_HIT, 1);
            terminal_v1(r, KIND_INVALID, h);
            return;
        }
        float ln = 0.0f, rn = 0.0f;
        cnt_add(c, ORA_CNT_NODE_TESTS, 2);
        cnt_add(c, ORA_CNT_BOX_TESTS, 2);
        const ora_node* L = &nodes[node->links[0]];
        const ora_node* R = &nodes[node->links[2]];
        const int hl = intersect_bounds(o, d, tMin, tMax, L->bmin, L->bmax, s.best, s.hasHit, &ln, NULL);
        const int hr = intersect_bounds(o, d, tMin, tMax, R->bmin, R->bmax, s.best, s.hasHit, &rn, NULL);
        const uint32_t required = (uint32_t)hl + (uint32_t)hr;
        if (sp + required > stackCapacity) {
            cnt_add(c, ORA_CNT_STACK_OVERFLOW, 1);
            terminal_v1(r, KIND_INVALID, h);
            return;
        }
        if (hl && hr) {
            const int leftFirst = ln < rn || (!(rn < ln) && node->links[0] < node->links[2]);
            stackNodes[sp] = leftFirst ? node->links[2] : node->links[0];
            stackNear[sp++] = leftFirst ? rn : ln;
            stackNodes[sp] = leftFirst ? node->links[0] : node->links[2];
            stackNear[sp++] = leftFirst ? ln : rn;
        } else if (hl) {
            stackNodes[sp] = node->links[0];
            stackNear[sp++] = ln;
        } else if (hr) {
            stackNodes[sp] = node->links[2];
            stackNear[sp++] = rn;
        }
        cnt_max(c, ORA_CNT_MAX_STACK, sp);
    }
    if (!s.hasHit) { terminal_v1(r, KIND_MISS, h); return; }
    const ora_tri* b = &tris[s.bt];
    const float w0 = 1.0f - s.bu - s.bv;
    f3 gn = normalize3(cross3(sub(ld3(b->pos[1]), ld3(b->pos[0])), sub(ld3(b->pos[2]), ld3(b->pos[0]))));
    f3 sn = normalize3(add(add(scl(ld3(b->nrm[0]), w0), scl(ld3(b->nrm[1]), s.bu)), scl(ld3(b->nrm[2]), s.bv)));
    if (!s.bf) { gn = mk(-gn.x, -gn.y, -gn.z); sn = mk(-sn.x, -sn.y, -sn.z); }
    memset(h, 0, sizeof(*h));
    const f3 p = add(o, scl(d, s.best));
    h->positionT[0] = p.x; h->positionT[1] = p.y; h->positionT[2] = p.z; h->positionT[3] = s.best;
    h->gnBaryU[0] = gn.x; h->gnBaryU[1] = gn.y; h->gnBaryU[2] = gn.z; h->gnBaryU[3] = s.bu;
    h->snBaryV[0] = sn.x; h->snBaryV[1] = sn.y; h->snBaryV[2] = sn.z; h->snBaryV[3] = s.bv;
    for (int k = 0; k < 4; ++k) h->ids[k] = b->identity[k];
    uint32_t flags = s.bf ? HITFLAG_FRONT : 0u;
    if ((b->metadata[0] & GEOM_ALPHA_MASK) != 0u || (b->metadata[1] & MAT_ALPHA_MASK) != 0u) flags |= HITFLAG_ALPHA;
    h->metadata[0] = r->identity[0];
    h->metadata[1] = KIND_TRIANGLE;
    h->metadata[2] = flags;
    h->metadata[3] = r->identity[1];
}

void ora_trace_batch_v1(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris,
                        uint32_t triCount, const ora_ray_v1* rays, uint32_t rayCount,
                        uint32_t queryMode, uint32_t stackCapacity, ora_hit_v1* hits,
                        uint64_t* counters) {
    for (uint32_t i = 0; i < rayCount; ++i) {
        trace_one_v1(nodes, nodeCount, tris, triCount, &rays[i], queryMode, stackCapacity,
                     &hits[i], counters);
    }
}

/* The reference's alpha test (pbr_software_traversal.hlsli:207-248,
 * software_trace_v1.hlsl:127-170) resolved per triangle for an atlas of
 * uniform layers: masked triangles get ORA_ALPHA_ACCEPT or ORA_ALPHA_SKIP
 * in metadata[0]; those left unmarked (material index out of range, invalid
 * mapping) make the walks return an invalid result when they reach them.
 * outcome[i]: 0 accept or unmasked, 1 skip, 2 invalid. */
void ora_alpha_resolve(ora_tri* tris, uint32_t n, const ora_material_v0* materials, uint32_t materialCount,
                       uint32_t atlasLayerCount, const float* layerAlpha, uint32_t samplerId, uint8_t* outcome) {
    for (uint32_t i = 0; i < n; ++i) {
        ora_tri* t = &tris[i];
        t->metadata[0] &= ~(ORA_ALPHA_SKIP | ORA_ALPHA_ACCEPT);
        uint8_t r = 0;
        if ((t->metadata[0] & GEOM_ALPHA_MASK) != 0u || (t->metadata[1] & MAT_ALPHA_MASK) != 0u) {
            const uint32_t m = t->identity[3];
            if (m >= materialCount) {
                r = 2;
            } else {
                const ora_material_v0* mat = &materials[m];
                const uint32_t layer = mat->textureImageIndices[0], sampler = mat->textureSamplerIndices[0];
                float alpha = mat->baseColorFactor[3];
                int valid = 1;
                if (layer == 0xFFFFFFFFu) valid = sampler == 0xFFFFFFFFu;
                else if (layer >= atlasLayerCount || sampler != samplerId) valid = 0;
                else alpha *= layerAlpha[layer];
                r = !valid ? 2 : (alpha < mat->surfaceParams[3] ? 1 : 0);
            }
            if (r == 0) t->metadata[0] |= ORA_ALPHA_ACCEPT;
            if (r == 1) t->metadata[0] |= ORA_ALPHA_SKIP;
        }
        if (outcome) outcome[i] = r;
    }
}
