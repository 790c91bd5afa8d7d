/*
 * Oracle (test infrastructure only): binned-SAH BVH build + flatten.
 *
 * Restates Rt::Cpu::Bvh<float> with BvhBuildMethod::BinnedSah
 * (include/rt/cpu/Bvh.hpp:64-681) followed by FlattenCanonicalL3BinnedSah /
 * FlattenCpuSah (src/rt/software_gpu/SoftwareGpu.cpp:947-1251):
 *   - 16 bins per axis, bin index and SAH cost evaluated in double
 *     (Bvh.hpp:364-380 SurfaceAreaPromoted, :413-432 ComputeBinIndex);
 *   - strict '<' on cost so the lowest axis / bin wins ties (Bvh.hpp:487-496);
 *   - stable partition by bin (Bvh.hpp:539-560), with the stable
 *     (centroid, primitiveId, index) median fallback (Bvh.hpp:506-537);
 *   - leaves hold <= maxLeafSize triangles, depth > 128 is an error
 *     (Bvh.hpp:562-616);
 *   - nodes in build pre-order, links as in SoftwareGpu.cpp:1073-1089.
 * Triangle validity (degenerate rejection) follows SoftwareGpu.cpp:142-158.
 *
 * Parity pins: tests/test_oracle_bvh.py checks this builder node-for-node
 * against the reference's own Bvh<float> + FlattenCanonicalL3BinnedSah
 * compiled from its sources into oracle/_ref (when present), against the
 * committed golden digests in tests/golden/bvh_*.json (made from oracle/_ref),
 * the 12-triangle grid at leaf size 1 -> 23 nodes (tests/gpu/L4/TestMain.cpp:199),
 * and structural invariants (every triangle in exactly one leaf, child boxes
 * inside parent boxes, leaf boxes equal to the union of their triangles).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t tri;      /* canonical index */
    uint32_t primId;   /* traversal ID */
    float bmin[3], bmax[3];
    float c[3];
} ora_ref;

typedef struct {
    float bmin[3], bmax[3];
    uint32_t left, right, first, count;
} bnode;

typedef struct {
    ora_ref* refs;
    ora_ref* scratch;
    bnode* nodes;
    uint32_t nodeCount;
    uint32_t maxLeaf;
    uint32_t maxDepth;
    int error;
} builder;

#define NBINS 16u
#define MAX_DEPTH 128u

static float fmin_ref(float a, float b) { return a < b ? a : b; }   /* Math.hpp:178-185 */
static float fmax_ref(float a, float b) { return a > b ? a : b; }   /* Math.hpp:187-194 */

static void box_empty(float mn[3], float mx[3]) {
    for (int k = 0; k < 3; ++k) { mn[k] = INFINITY; mx[k] = -INFINITY; }
}
static void box_expand(float mn[3], float mx[3], const float bmn[3], const float bmx[3]) {
    for (int k = 0; k < 3; ++k) { mn[k] = fmin_ref(mn[k], bmn[k]); mx[k] = fmax_ref(mx[k], bmx[k]); }
}
static int box_valid(const float mn[3], const float mx[3]) {
    for (int k = 0; k < 3; ++k) {
        if (!isfinite(mn[k]) || !isfinite(mx[k]) || !(mn[k] <= mx[k])) return 0;
    }
    return 1;
}
static double sa_promoted(const float mn[3], const float mx[3]) {
    if (!box_valid(mn, mx)) return 0.0;
    const double ex = (double)mx[0] - (double)mn[0];
    const double ey = (double)mx[1] - (double)mn[1];
    const double ez = (double)mx[2] - (double)mn[2];
    return 2.0 * ((ex * ey) + (ey * ez) + (ez * ex));
}
static uint32_t bin_index(float c, float cmin, float cext) {
    const double normalized = ((double)c - (double)cmin) / (double)cext;
    const double scaled = normalized * (double)NBINS;
    if (!(scaled > 0.0)) return 0u;
    if (scaled >= (double)NBINS) return NBINS - 1u;
    return (uint32_t)scaled;
}

static int g_axis;
static int median_cmp(const void* pa, const void* pb) {
    const ora_ref* a = (const ora_ref*)pa;
    const ora_ref* b = (const ora_ref*)pb;
    if (a->c[g_axis] < b->c[g_axis]) return -1;
    if (a->c[g_axis] > b->c[g_axis]) return 1;
    if (a->primId < b->primId) return -1;
    if (a->primId > b->primId) return 1;
    if (a->tri < b->tri) return -1;
    if (a->tri > b->tri) return 1;
    return 0;
}

static uint32_t build_range(builder* b, uint32_t begin, uint32_t end, uint32_t depth) {
    if (depth > MAX_DEPTH) { b->error = -2; return 0; }
    const uint32_t nodeIndex = b->nodeCount++;
    bnode* node = &b->nodes[nodeIndex];
    memset(node, 0, sizeof(*node));
    node->left = node->right = ORA_INVALID;
    if (depth > b->maxDepth) b->maxDepth = depth;

    float mn[3], mx[3];
    box_empty(mn, mx);
    for (uint32_t i = begin; i < end; ++i) box_expand(mn, mx, b->refs[i].bmin, b->refs[i].bmax);
    const uint32_t count = end - begin;
    if (count <= b->maxLeaf) {
        memcpy(node->bmin, mn, sizeof(mn));
        memcpy(node->bmax, mx, sizeof(mx));
        node->first = begin;
        node->count = count;
        return nodeIndex;
    }

    float cmn[3], cmx[3];
    box_empty(cmn, cmx);
    for (uint32_t i = begin; i < end; ++i) box_expand(cmn, cmx, b->refs[i].c, b->refs[i].c);
    float cext[3];
    for (int k = 0; k < 3; ++k) cext[k] = cmx[k] - cmn[k];

    uint32_t middle = begin;
    /* FindSahSplit (Bvh.hpp:434-504) */
    int valid = 0;
    uint32_t bestAxis = 0, bestBin = 0;
    double bestCost = INFINITY;
    for (uint32_t axis = 0; axis < 3; ++axis) {
        if (!(cext[axis] > 0.0f)) continue;
        float bmn[NBINS][3], bmx[NBINS][3];
        uint32_t bcount[NBINS];
        for (uint32_t k = 0; k < NBINS; ++k) { box_empty(bmn[k], bmx[k]); bcount[k] = 0; }
        for (uint32_t i = begin; i < end; ++i) {
            const uint32_t bi = bin_index(b->refs[i].c[axis], cmn[axis], cext[axis]);
            bcount[bi]++;
            box_expand(bmn[bi], bmx[bi], b->refs[i].bmin, b->refs[i].bmax);
        }
        for (uint32_t split = 0; split + 1 < NBINS; ++split) {
            float lmn[3], lmx[3], rmn[3], rmx[3];
            box_empty(lmn, lmx);
            box_empty(rmn, rmx);
            uint32_t lc = 0, rc = 0;
            for (uint32_t k = 0; k <= split; ++k) { lc += bcount[k]; box_expand(lmn, lmx, bmn[k], bmx[k]); }
            for (uint32_t k = split + 1; k < NBINS; ++k) { rc += bcount[k]; box_expand(rmn, rmx, bmn[k], bmx[k]); }
            if (lc == 0 || rc == 0) continue;
            const double cost = (sa_promoted(lmn, lmx) * (double)lc) + (sa_promoted(rmn, rmx) * (double)rc);
            if (!valid || cost < bestCost) {
                valid = 1; bestAxis = axis; bestBin = split; bestCost = cost;
            }
        }
    }
    if (valid) {
        /* stable partition (Bvh.hpp:539-560) */
        uint32_t nl = 0;
        for (uint32_t i = begin; i < end; ++i) {
            if (bin_index(b->refs[i].c[bestAxis], cmn[bestAxis], cext[bestAxis]) <= bestBin) nl++;
        }
        /* two-pass stable split into scratch */
        uint32_t li = begin, ri = begin + nl;
        for (uint32_t i = begin; i < end; ++i) {
            const ora_ref* r = &b->refs[i];
            if (bin_index(r->c[bestAxis], cmn[bestAxis], cext[bestAxis]) <= bestBin) b->scratch[li++] = *r;
            else b->scratch[ri++] = *r;
        }
        memcpy(&b->refs[begin], &b->scratch[begin], (size_t)count * sizeof(ora_ref));
        middle = begin + nl;
    }
    if (middle == begin || middle == end) {
        /* MedianPartition (Bvh.hpp:506-537): LongestAxis (Math.hpp:197-209) */
        int axis = 0;
        if (cext[1] > cext[0]) axis = 1;
        if (cext[2] > cext[axis]) axis = 2;
        g_axis = axis;
        qsort(&b->refs[begin], count, sizeof(ora_ref), median_cmp); /* total order => equals stable_sort */
        middle = begin + count / 2u;
    }
    const uint32_t left = build_range(b, begin, middle, depth + 1);
    if (b->error) return 0;
    const uint32_t right = build_range(b, middle, end, depth + 1);
    if (b->error) return 0;
    node = &b->nodes[nodeIndex];
    memcpy(node->bmin, mn, sizeof(mn));
    memcpy(node->bmax, mx, sizeof(mx));
    node->left = left;
    node->right = right;
    return nodeIndex;
}

/* SoftwareGpu.cpp:142-158 */
static int primitive_valid(const ora_tri* t) {
    const float* v0 = t->pos[0];
    const float* v1 = t->pos[1];
    const float* v2 = t->pos[2];
    for (int k = 0; k < 3; ++k) {
        if (!isfinite(v0[k]) || !isfinite(v1[k]) || !isfinite(v2[k])) return 0;
    }
    if (t->metadata[2] == ORA_INVALID) return 0;
    const float e0[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const float e1[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const float cr[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                         e0[0] * e1[1] - e0[1] * e1[0]};
    const float area2 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
    const float d0 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    const float d1 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const float scale = d0 > d1 ? d0 : d1;
    return isfinite(area2) && area2 > scale * scale * 1.0e-14f;
}

int ora_bvh_build(const ora_tri* tris, uint32_t n, uint32_t maxLeafSize,
                  ora_node* nodesOut, uint32_t* nodeCountOut,
                  uint32_t* primOrderOut, uint32_t* maxDepthOut) {
    *nodeCountOut = 0;
    *maxDepthOut = 0;
    if (n == 0) return 0;
    if (maxLeafSize == 0) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        if (!primitive_valid(&tris[i])) return -3;
    }
    builder b;
    memset(&b, 0, sizeof(b));
    b.refs = (ora_ref*)malloc((size_t)n * sizeof(ora_ref));
    b.scratch = (ora_ref*)malloc((size_t)n * sizeof(ora_ref));
    b.nodes = (bnode*)malloc((size_t)(2u * n - 1u) * sizeof(bnode));
    b.maxLeaf = maxLeafSize;
    for (uint32_t i = 0; i < n; ++i) {
        ora_ref* r = &b.refs[i];
        r->tri = i;
        r->primId = tris[i].metadata[2];
        box_empty(r->bmin, r->bmax);
        for (int v = 0; v < 3; ++v) box_expand(r->bmin, r->bmax, tris[i].pos[v], tris[i].pos[v]);
        /* Triangle::Centroid, Geometry.hpp:157-175: sum in double, / 3 */
        for (int k = 0; k < 3; ++k) {
            r->c[k] = (float)(((double)tris[i].pos[0][k] + (double)tris[i].pos[1][k] +
                               (double)tris[i].pos[2][k]) / 3.0);
        }
    }
    build_range(&b, 0, n, 1);
    int rc = b.error;
    if (rc == 0) {
        /* parents: pre-order walk */
        uint32_t* parent = (uint32_t*)malloc((size_t)b.nodeCount * sizeof(uint32_t));
        parent[0] = ORA_INVALID;
        for (uint32_t i = 0; i < b.nodeCount; ++i) {
            if (b.nodes[i].count == 0) {
                parent[b.nodes[i].left] = i;
                parent[b.nodes[i].right] = i;
            }
        }
        for (uint32_t i = 0; i < b.nodeCount; ++i) {
            const bnode* s = &b.nodes[i];
            ora_node* d = &nodesOut[i];
            for (int k = 0; k < 3; ++k) { d->bmin[k] = s->bmin[k]; d->bmax[k] = s->bmax[k]; }
            d->bmin[3] = 0.0f;
            d->bmax[3] = 0.0f;
            if (s->count != 0) {
                d->links[0] = s->first; d->links[1] = s->count; d->links[2] = ORA_INVALID;
            } else {
                d->links[0] = s->left; d->links[1] = 0u; d->links[2] = s->right;
            }
            d->links[3] = parent[i];
        }
        free(parent);
        for (uint32_t i = 0; i < n; ++i) primOrderOut[i] = b.refs[i].tri;
        *nodeCountOut = b.nodeCount;
        *maxDepthOut = b.maxDepth;
    }
    free(b.refs);
    free(b.scratch);
    free(b.nodes);
    return rc;
}
