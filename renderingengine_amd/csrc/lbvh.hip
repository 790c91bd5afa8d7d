// Device hierarchy build and device layout derivation.
//
// * LBVH build (BuildKarrasLbvh, src/rt/software_gpu/SoftwareGpu.cpp:1265-1345,
//   StableRadixSortMorton :886-940, FindKarrasSplit :546-570, RefitBottomUp
//   :1347-1500): 30-bit Morton codes of the triangle centroids (midpoints of
//   the triangle boxes) in the scene box, sorted by (code, traversal ID) as
//   one 64-bit key, the binary radix tree over the sorted keys, boxes bottom
//   up.  The keys are unique, so the radix tree is unique: Karras's parallel
//   construction (one thread per interior node, splits at the highest
//   differing key bit) yields the reference's top-down tree.  The
//   reference numbers nodes in pre-order (root 0, the left subtree next);
//   a node covering sorted leaves [a, b] has pre-order index 2a + (number
//   of its ancestors whose left subtree holds it), since the nodes before it
//   are its ancestors plus the maximal subtrees left of a.
// * Device layout: from binary nodes in pre-order (parents, depths) and the
//   leaf-order triangles, the binary wide nodes, the 4-wide nodes (binary
//   interior nodes at even depth, numbered in binary pre-order, which is the
//   4-wide tree's pre-order), the compressed 4-wide nodes, triangle records
//   and leaf boxes -- the same records the host layout writes.
// * Device refit: leaf boxes from the moved triangles, interior boxes the
//   union of their children bottom up (one thread per leaf climbs; the second
//   child to arrive at a parent computes it), then the device layout.
// Box unions follow std::min / std::max order (left child first), so boxes
// are bit-identical to the host's RefitBottomUp even for signed zeros.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "rex_internal.hpp"

namespace rex {

namespace {

constexpr uint32_t kB = 256;

__host__ __device__ inline float minf_std(float a, float b) { return b < a ? b : a; }
__host__ __device__ inline float maxf_std(float a, float b) { return a < b ? b : a; }

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v &= 0x000003ffu;
    v = (v | (v << 16u)) & 0x030000ffu;
    v = (v | (v << 8u)) & 0x0300f00fu;
    v = (v | (v << 4u)) & 0x030c30c3u;
    v = (v | (v << 2u)) & 0x09249249u;
    return v;
}

// QuantizeMortonAxis (SoftwareGpu.cpp:194-207), in double.
__device__ __forceinline__ uint32_t quantize_axis(float value, float lo, float hi) {
    const double extent = (double)hi - (double)lo;
    const double t = extent > 0.0 ? ((double)value - (double)lo) / extent : 0.5;
    const double c = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    const uint32_t q = (uint32_t)(c * 1024.0);
    return q < 1023u ? q : 1023u;
}

struct Box {
    float lo[3], hi[3];
};

__global__ void k_morton(const rex_triangle* __restrict__ tris, uint32_t n, float lx, float ly, float lz, float hx,
                         float hy, float hz, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const rex_triangle& t = tris[i];
    const float slo[3] = {lx, ly, lz}, shi[3] = {hx, hy, hz};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float lo = minf_std(minf_std(t.positions[0][a], t.positions[1][a]), t.positions[2][a]);
        const float hi = maxf_std(maxf_std(t.positions[0][a], t.positions[1][a]), t.positions[2][a]);
        q[a] = quantize_axis(midpoint_f(lo, hi), slo[a], shi[a]);
    }
    const uint32_t code = (expand_bits(q[0]) << 2u) | (expand_bits(q[1]) << 1u) | expand_bits(q[2]);
    keys[i] = ((uint64_t)code << 32u) | t.metadata[2];
    vals[i] = i;
}

// Karras ids: interior i in [0, n - 1), leaf j -> (n - 1) + j.
__device__ __forceinline__ int delta(const uint64_t* keys, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    return __clzll(keys[i] ^ keys[j]);
}

__global__ void k_karras(const uint64_t* __restrict__ keys, int n, uint32_t* __restrict__ child,
                         uint32_t* __restrict__ parent, uint32_t* __restrict__ first) {
    const int i = (int)(blockIdx.x * kB + threadIdx.x);
    if (i >= n - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2) {
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    }
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0, step = l;
    do {
        step = (step + 1) >> 1;
        const int cand = s + step;
        if (cand < l && delta(keys, n, i, i + cand * d) > dnode) s = cand;
    } while (step > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const uint32_t leafBase = (uint32_t)(n - 1);
    const uint32_t left = lo == gamma ? leafBase + (uint32_t)gamma : (uint32_t)gamma;
    const uint32_t right = hi == gamma + 1 ? leafBase + (uint32_t)(gamma + 1) : (uint32_t)(gamma + 1);
    child[2 * i] = left;
    child[2 * i + 1] = right;
    parent[left] = (uint32_t)i;
    parent[right] = (uint32_t)i;
    first[i] = (uint32_t)lo;
}

// Pre-order index and depth of every Karras node; depth beyond 127 edges
// (the reference's 128 levels) raises the overflow flag.
__global__ void k_preorder(const uint32_t* __restrict__ child, const uint32_t* __restrict__ parent,
                           const uint32_t* __restrict__ first, uint32_t n, uint32_t* __restrict__ pre,
                           uint32_t* __restrict__ depth, uint32_t* flags) {
    const uint32_t id = blockIdx.x * kB + threadIdx.x;
    const uint32_t total = 2u * n - 1u;
    if (id >= total) return;
    const uint32_t leafBase = n - 1u;
    const uint32_t a = id >= leafBase ? id - leafBase : first[id];
    // Karras node 0 is the root (the single leaf when n == 1).
    uint32_t x = id, left = 0, dep = 0;
    while (x != 0u) {
        const uint32_t p = parent[x];
        if (child[2u * p] == x) ++left;
        ++dep;
        x = p;
        if (dep > 200u) break;
    }
    pre[id] = 2u * a + left;
    depth[id] = dep;
    if (dep > 127u) atomicOr(flags, 1u);
    atomicMax(flags + 1, dep + 1u);
}

// Binary nodes in pre-order: children, leaf ranges, parents, depths; the
// boxes are filled by the refit pass.
__global__ void k_emit_binary(const uint32_t* __restrict__ child, const uint32_t* __restrict__ parent,
                              const uint32_t* __restrict__ pre, const uint32_t* __restrict__ depth, uint32_t n,
                              BinaryNode* __restrict__ nodes, uint32_t* __restrict__ parentsPre,
                              uint32_t* __restrict__ depthPre) {
    const uint32_t id = blockIdx.x * kB + threadIdx.x;
    const uint32_t total = 2u * n - 1u;
    if (id >= total) return;
    const uint32_t leafBase = n - 1u;
    const uint32_t p = pre[id];
    BinaryNode b;
    for (int a = 0; a < 3; ++a) {
        b.bmin[a] = INFINITY;
        b.bmax[a] = -INFINITY;
    }
    if (id >= leafBase) {
        b.left = kInvalid;
        b.right = kInvalid;
        b.first = id - leafBase;
        b.count = 1u;
    } else {
        b.left = pre[child[2u * id]];
        b.right = pre[child[2u * id + 1u]];
        b.first = 0u;
        b.count = 0u;
    }
    nodes[p] = b;
    parentsPre[p] = (n > 1u && id != 0u) ? pre[parent[id]] : kInvalid;
    depthPre[p] = depth[id];
}

__global__ void k_gather_tris(const rex_triangle* __restrict__ src, const uint32_t* __restrict__ order, uint32_t n,
                              rex_triangle* __restrict__ dst) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j < n) dst[j] = src[order[j]];
}

__global__ void k_zero_words(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Bottom-up boxes over pre-order nodes: one thread per leaf computes the
// leaf's box and climbs; at each parent the first child to arrive stops and
// the second computes the union (left child first) and continues.
__global__ void k_refit(BinaryNode* nodes, const uint32_t* __restrict__ parents, uint32_t m,
                        const rex_triangle* __restrict__ leafTris, uint32_t* visits) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m || nodes[i].count == 0u) return;
    Box b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = INFINITY;
        b.hi[a] = -INFINITY;
    }
    const uint32_t f = nodes[i].first, c = nodes[i].count;
    // The vertices of the leaf's triangles folded in order (RefitBottomUp).
    for (uint32_t j = f; j < f + c; ++j) {
        for (int a = 0; a < 3; ++a) {
            for (int v = 0; v < 3; ++v) {
                b.lo[a] = minf_std(b.lo[a], leafTris[j].positions[v][a]);
                b.hi[a] = maxf_std(b.hi[a], leafTris[j].positions[v][a]);
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        nodes[i].bmin[a] = b.lo[a];
        nodes[i].bmax[a] = b.hi[a];
    }
    uint32_t x = i;
    for (;;) {
        const uint32_t p = parents[x];
        if (p == kInvalid) break;
        __threadfence();
        if (atomicAdd(visits + p, 1u) == 0u) break;
        __threadfence();
        volatile BinaryNode* vn = nodes;
        const uint32_t l = nodes[p].left, r = nodes[p].right;
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = minf_std(minf_std(INFINITY, vn[l].bmin[a]), vn[r].bmin[a]);
            hi[a] = maxf_std(maxf_std(-INFINITY, vn[l].bmax[a]), vn[r].bmax[a]);
        }
        for (int a = 0; a < 3; ++a) {
            vn[p].bmin[a] = lo[a];
            vn[p].bmax[a] = hi[a];
        }
        x = p;
    }
}

__device__ __forceinline__ uint32_t leaf_ref_of(const BinaryNode& c) {
    return kLeafBit | ((c.count - 1u) << 28) | c.first;
}

__global__ void k_flags(const BinaryNode* __restrict__ nodes, const uint32_t* __restrict__ depth, uint32_t m,
                        uint32_t* __restrict__ interior, uint32_t* __restrict__ wide4) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const bool in = nodes[i].count == 0u;
    interior[i] = in ? 1u : 0u;
    wide4[i] = in && (depth[i] & 1u) == 0u ? 1u : 0u;
}

struct LayoutOut {
    WideNode* wide;
    Wide4* wide4;
    Wide4Q* wide4q;
    TriPos* tris;
    TriAttr* attrs;
    LeafBox* leafBox;
};

// Conservative 8-bit quantization of a 4-wide node (the host's
// quantize_wide4, same double arithmetic).
__device__ Wide4Q quantize_dev(const Wide4& W, double sceneScale) {
    Wide4Q q;
    const float* lo[3] = {&W.lox.x, &W.loy.x, &W.loz.x};
    const float* hi[3] = {&W.hix.x, &W.hiy.x, &W.hiz.x};
    const uint32_t* ref = &W.ref.x;
    uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    uint32_t exps = 0;
    float origin[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < 3; ++a) {
        double plo = 0.0, phi = 0.0;
        bool any = false;
        for (int k = 0; k < 4; ++k) {
            if (ref[k] == kInvalid) continue;
            plo = any ? fmin(plo, (double)lo[a][k]) : (double)lo[a][k];
            phi = any ? fmax(phi, (double)hi[a][k]) : (double)hi[a][k];
            any = true;
        }
        const double pad = fmax(fmax(ldexp(fmax(fabs(plo), fabs(phi)) + (phi - plo), -16), ldexp(sceneScale, -18)),
                                1.0e-30);
        float o = (float)(plo - pad);
        if ((double)o > plo - pad) o = nextafterf(o, -INFINITY);
        const double span = (phi + pad) - (double)o;
        int e = -126;
        while (e < 127 && ldexp(255.0, e) < span) ++e;
        origin[a] = o;
        exps |= (uint32_t)(e + 127) << (8 * a);
        const double step = ldexp(1.0, e);
        for (int k = 0; k < 4; ++k) {
            uint32_t l = 255u, h = 0u;
            if (ref[k] != kInvalid) {
                const double fl = floor(((double)lo[a][k] - pad - (double)o) / step);
                const double fh = ceil(((double)hi[a][k] + pad - (double)o) / step);
                l = (uint32_t)fmin(255.0, fmax(0.0, fl));
                h = (uint32_t)fmin(255.0, fmax(0.0, fh));
            }
            qlo[a] |= l << (8 * k);
            qhi[a] |= h << (8 * k);
        }
    }
    q.origin = {origin[0], origin[1], origin[2], __uint_as_float(exps)};
    q.planeA = {qlo[0], qlo[1], qlo[2], qhi[0]};
    q.planeB = {qhi[1], qhi[2], 0u, 0u};
    q.ref = W.ref;
    return q;
}

// The 4-wide node of interior node b (at even depth): its grandchildren (or
// a leaf child itself) in pre-order slots, child boxes and references.
__device__ Wide4 wide4_of(const BinaryNode* __restrict__ nodes, const BinaryNode& b, const uint32_t* __restrict__ w4Idx) {
    uint32_t slots[4] = {kInvalid, kInvalid, kInvalid, kInvalid};
    uint32_t k = 0;
    const uint32_t kids[2] = {b.left, b.right};
    for (int c = 0; c < 2; ++c) {
        const BinaryNode cn = nodes[kids[c]];
        if (cn.count != 0u) {
            slots[k++] = kids[c];
        } else {
            slots[k++] = cn.left;
            slots[k++] = cn.right;
        }
    }
    Wide4 W;
    float* lo[3] = {&W.lox.x, &W.loy.x, &W.loz.x};
    float* hi[3] = {&W.hix.x, &W.hiy.x, &W.hiz.x};
    for (uint32_t j = 0; j < 4; ++j) {
        (&W.bin.x)[j] = slots[j];
        (&W.ref.x)[j] = kInvalid;
        if (slots[j] == kInvalid) {
            for (int a = 0; a < 3; ++a) lo[a][j] = hi[a][j] = 0.0f;
            continue;
        }
        const BinaryNode sn = nodes[slots[j]];
        for (int a = 0; a < 3; ++a) {
            lo[a][j] = sn.bmin[a];
            hi[a][j] = sn.bmax[a];
        }
        (&W.ref.x)[j] = sn.count != 0u ? leaf_ref_of(sn) : w4Idx[slots[j]];
    }
    return W;
}

__device__ double scene_scale_of(const BinaryNode* __restrict__ nodes) {
    double sceneScale = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double l = nodes[0].bmin[a], h = nodes[0].bmax[a];
        sceneScale = fmax(sceneScale, fmax(fabs(l), fabs(h)) + (h - l));
    }
    return sceneScale;
}

// Compressed 4-wide nodes only (the traversal hierarchy's).
__global__ void k_emit_wide4q(const BinaryNode* __restrict__ nodes, const uint32_t* __restrict__ depth, uint32_t m,
                              const uint32_t* __restrict__ w4Idx, Wide4Q* __restrict__ out, uint32_t* w4Depth) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const BinaryNode b = nodes[i];
    if (b.count != 0u || (depth[i] & 1u) != 0u) return;
    out[w4Idx[i]] = quantize_dev(wide4_of(nodes, b, w4Idx), scene_scale_of(nodes));
    atomicMax(w4Depth, depth[i] / 2u + 1u);
}

__global__ void k_emit_layout(const BinaryNode* __restrict__ nodes, const uint32_t* __restrict__ parents,
                              const uint32_t* __restrict__ depth, uint32_t m, const uint32_t* __restrict__ wideIdx,
                              const uint32_t* __restrict__ w4Idx, LayoutOut out, uint32_t* w4Depth) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const BinaryNode b = nodes[i];
    if (b.count != 0u) {
        const LeafBox box = {{b.bmin[0], b.bmin[1], b.bmin[2], 0.0f}, {b.bmax[0], b.bmax[1], b.bmax[2], 0.0f}};
        for (uint32_t j = b.first; j < b.first + b.count; ++j) out.leafBox[j] = box;
        return;
    }
    const BinaryNode L = nodes[b.left], R = nodes[b.right];
    auto refOf = [&](uint32_t c, const BinaryNode& cn) -> uint32_t {
        return cn.count == 0u ? wideIdx[c] : leaf_ref_of(cn);
    };
    WideNode w;
    w.a = {L.bmin[0], L.bmin[1], L.bmin[2], L.bmax[0]};
    w.b = {L.bmax[1], L.bmax[2], R.bmin[0], R.bmin[1]};
    w.c = {R.bmin[2], R.bmax[0], R.bmax[1], R.bmax[2]};
    uint32_t parentLink = kInvalid;
    if (parents[i] != kInvalid) {
        const uint32_t p = parents[i];
        parentLink = wideIdx[p] | (nodes[p].right == i ? 0x80000000u : 0u);
    }
    w.link = {refOf(b.left, L), refOf(b.right, R), b.left, parentLink};
    out.wide[wideIdx[i]] = w;
    if ((depth[i] & 1u) != 0u) return;
    // A 4-wide node: its grandchildren (or a leaf child itself) in pre-order slots.
    const Wide4 W = wide4_of(nodes, b, w4Idx);
    const uint32_t wi = w4Idx[i];
    out.wide4[wi] = W;
    out.wide4q[wi] = quantize_dev(W, scene_scale_of(nodes));
    atomicMax(w4Depth, depth[i] / 2u + 1u);
}

// alphaSkip[m] = 1: an alpha-masked triangle of material m is skipped by the
// reference's alpha test; it gets zero edges (every triangle test's
// determinant is then 0).  Scenes without masked geometry pass count 0.
__global__ void k_emit_tris(const rex_triangle* __restrict__ leafTris, uint32_t n, TriPos* __restrict__ tpos,
                            TriAttr* __restrict__ tattr, const uint8_t* __restrict__ alphaSkip, uint32_t alphaCount) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const rex_triangle& t = leafTris[j];
    const float* v0 = t.positions[0];
    const float* v1 = t.positions[1];
    const float* v2 = t.positions[2];
    float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const bool masked = (t.metadata[0] & kGeometryFlagAlphaMask) != 0u || (t.metadata[1] & kMaterialFlagAlphaMask) != 0u;
    if (masked && t.identity[3] < alphaCount && alphaSkip[t.identity[3]] != 0u) {
        for (int a = 0; a < 3; ++a) e1[a] = e2[a] = 0.0f;
    }
    uint32_t flags = 0;
    if ((t.metadata[0] & kGeometryFlagDoubleSided) != 0u || (t.metadata[1] & kMaterialFlagDoubleSided) != 0u) {
        flags |= kTriDoubleSided;
    }
    if ((t.metadata[3] & kInstanceFlagCastsShadow) != 0u) flags |= kTriCastsShadow;
    TriPos p;
    p.p0 = {v0[0], v0[1], v0[2], e1[0]};
    p.p1 = {e1[1], e1[2], e2[0], e2[1]};
    p.p2 = {e2[2], __uint_as_float(t.identity[0]), __uint_as_float(t.identity[1]), __uint_as_float(flags)};
    tpos[j] = p;
    TriAttr a;
    a.n0 = {t.normals[0][0], t.normals[0][1], t.normals[0][2], 0.0f};
    a.n1 = {t.normals[1][0], t.normals[1][1], t.normals[1][2], 0.0f};
    a.n2 = {t.normals[2][0], t.normals[2][1], t.normals[2][2], 0.0f};
    a.ids = {t.identity[0], t.identity[1], t.identity[2], t.identity[3]};
    a.meta = {t.metadata[0], t.metadata[1], t.metadata[2], t.metadata[3]};
    tattr[j] = a;
}

uint32_t grid(uint64_t n) { return (uint32_t)((n + kB - 1) / kB); }

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
    return hipMalloc(reinterpret_cast<void**>(p), (count ? count : 1) * sizeof(T));
}

struct Temp {
    std::vector<void*> ptrs;
    template <typename T>
    hipError_t alloc(T** p, size_t count) {
        hipError_t e = dmalloc(p, count);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
    ~Temp() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// Exclusive sum of m uint32 flags (hipCUB).
hipError_t exclusive_sum(const uint32_t* in, uint32_t* out, uint32_t m, Temp& tmp, hipStream_t st) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)m, st);
    if (e != hipSuccess) return e;
    void* scratch = nullptr;
    if ((e = tmp.alloc(reinterpret_cast<uint8_t**>(&scratch), bytes)) != hipSuccess) return e;
    return hipcub::DeviceScan::ExclusiveSum(scratch, bytes, in, out, (int)m, st);
}

}  // namespace

// Device layout of the scene's device hierarchy (dNodes / dParents / dDepth,
// m nodes in pre-order, boxes set) over the leaf-order triangles dLeafTris.
// The new buffers replace the scene's only when everything succeeded.
rex_status device_layout(Scene& s, const rex_triangle* dLeafTris, uint32_t n, hipStream_t st) {
    const uint32_t m = s.nodeCount;
    Temp tmp;
    uint32_t *interior = nullptr, *w4flag = nullptr, *wideIdx = nullptr, *w4Idx = nullptr, *w4Depth = nullptr;
    if (rex_status r = check_hip(tmp.alloc(&interior, m), "hipMalloc layout")) return r;
    if (rex_status r = check_hip(tmp.alloc(&w4flag, m), "hipMalloc layout")) return r;
    if (rex_status r = check_hip(tmp.alloc(&wideIdx, m), "hipMalloc layout")) return r;
    if (rex_status r = check_hip(tmp.alloc(&w4Idx, m), "hipMalloc layout")) return r;
    if (rex_status r = check_hip(tmp.alloc(&w4Depth, 4), "hipMalloc layout")) return r;
    hipLaunchKernelGGL(k_flags, dim3(grid(m)), dim3(kB), 0, st, s.dNodes, s.dDepth, m, interior, w4flag);
    if (rex_status r = check_hip(exclusive_sum(interior, wideIdx, m, tmp, st), "scan")) return r;
    if (rex_status r = check_hip(exclusive_sum(w4flag, w4Idx, m, tmp, st), "scan")) return r;
    uint32_t last[4] = {0, 0, 0, 0};
    if (rex_status r = check_hip(hipMemcpyAsync(&last[0], interior + (m - 1), 4, hipMemcpyDeviceToHost, st), "copy"))
        return r;
    if (rex_status r = check_hip(hipMemcpyAsync(&last[1], wideIdx + (m - 1), 4, hipMemcpyDeviceToHost, st), "copy"))
        return r;
    if (rex_status r = check_hip(hipMemcpyAsync(&last[2], w4flag + (m - 1), 4, hipMemcpyDeviceToHost, st), "copy"))
        return r;
    if (rex_status r = check_hip(hipMemcpyAsync(&last[3], w4Idx + (m - 1), 4, hipMemcpyDeviceToHost, st), "copy"))
        return r;
    if (rex_status r = check_hip(hipStreamSynchronize(st), "sync")) return r;
    const uint32_t wideCount = last[0] + last[1];
    const uint32_t w4Count = last[2] + last[3];
    LayoutOut out{};
    WideNode* dWide = nullptr;
    Wide4* dWide4 = nullptr;
    Wide4Q* dWide4Q = nullptr;
    TriPos* dTris = nullptr;
    TriAttr* dAttrs = nullptr;
    LeafBox* dLeafBox = nullptr;
    auto freeAll = [&]() {
        void* ptrs[] = {dWide, dWide4, dWide4Q, dTris, dAttrs, dLeafBox};
        for (void* p : ptrs) {
            if (p) (void)hipFree(p);
        }
    };
    if (dmalloc(&dWide, wideCount) != hipSuccess || dmalloc(&dWide4, w4Count) != hipSuccess ||
        dmalloc(&dWide4Q, w4Count) != hipSuccess || dmalloc(&dTris, n) != hipSuccess ||
        dmalloc(&dAttrs, n) != hipSuccess || dmalloc(&dLeafBox, n) != hipSuccess) {
        freeAll();
        return fail(REX_DEVICE_ERROR, "hipMalloc scene layout failed");
    }
    out = {dWide, dWide4, dWide4Q, dTris, dAttrs, dLeafBox};
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kB), 0, st, w4Depth, 1u);
    hipLaunchKernelGGL(k_emit_layout, dim3(grid(m)), dim3(kB), 0, st, s.dNodes, s.dParents, s.dDepth, m, wideIdx,
                       w4Idx, out, w4Depth);
    uint32_t alphaCount = 0;
    if (s.hasAlpha && !s.alphaTable.empty()) {
        alphaCount = (uint32_t)s.alphaTable.size();
        if (s.dAlphaSkip == nullptr) {
            std::vector<uint8_t> skip(alphaCount);
            for (uint32_t m = 0; m < alphaCount; ++m) skip[m] = s.alphaTable[m] == kAlphaSkip ? 1u : 0u;
            hipError_t ae = hipMalloc(&s.dAlphaSkip, alphaCount);
            if (ae == hipSuccess) ae = hipMemcpy(s.dAlphaSkip, skip.data(), alphaCount, hipMemcpyHostToDevice);
            if (ae != hipSuccess) {
                if (s.dAlphaSkip) (void)hipFree(s.dAlphaSkip);
                s.dAlphaSkip = nullptr;
                freeAll();
                return check_hip(ae, "alpha table upload");
            }
        }
    }
    hipLaunchKernelGGL(k_emit_tris, dim3(grid(n)), dim3(kB), 0, st, dLeafTris, n, dTris, dAttrs,
                       (const uint8_t*)s.dAlphaSkip, alphaCount);
    uint32_t depth4 = 0;
    BinaryNode root{};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&depth4, w4Depth, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&root, s.dNodes, sizeof(BinaryNode), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        freeAll();
        return check_hip(e, "device layout");
    }
    // In-flight work of any stream may still read the old layout.
    if (rex_status r = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize")) {
        freeAll();
        return r;
    }
    void* old[] = {s.dWide, s.dTris, s.dAttrs, s.dWide4, s.dWide4Q, s.dLeafBox};
    for (void* p : old) {
        if (p) (void)hipFree(p);
    }
    s.dWide = dWide;
    s.dWide4 = dWide4;
    s.dWide4Q = dWide4Q;
    s.dTris = dTris;
    s.dAttrs = dAttrs;
    s.dLeafBox = dLeafBox;
    s.wideCount = wideCount;
    s.wide4Count = w4Count;
    // The compressed nodes of this (device) hierarchy; a scene with a
    // traversal hierarchy replaces them (refit_traversal).
    s.wide4qCount = w4Count;
    s.deviceBytes = (uint64_t)wideCount * sizeof(WideNode) + (uint64_t)w4Count * (sizeof(Wide4) + sizeof(Wide4Q)) +
                    (uint64_t)n * (sizeof(TriPos) + sizeof(TriAttr) + sizeof(LeafBox)) +
                    (uint64_t)m * (sizeof(BinaryNode) + 8u);
    SceneView& v = s.view;
    v = SceneView{};
    v.wide = dWide;
    v.tris = dTris;
    v.attrs = dAttrs;
    v.wide4 = dWide4;
    v.wide4q = dWide4Q;
    v.leafBox = dLeafBox;
    v.wide4Depth = depth4;
    v.rootRef = root.count != 0u ? (kLeafBit | ((root.count - 1u) << 28) | root.first) : 0u;
    v.rootRefQ = v.rootRef;
    v.nodeCount = m;
    v.triCount = n;
    v.maxDepth = s.maxDepth;
    for (int k = 0; k < 3; ++k) {
        v.rootMin[k] = root.bmin[k];
        v.rootMax[k] = root.bmax[k];
    }
    return REX_SUCCESS;
}

rex_status upload_hierarchy(Scene& s) {
    const uint32_t m = s.nodeCount;
    std::vector<uint32_t> depth(m, 0u);
    for (uint32_t i = 1; i < m; ++i) depth[i] = depth[s.parents[i]] + 1u;  // parents precede children
    BinaryNode* dn = nullptr;
    uint32_t *dp = nullptr, *dd = nullptr;
    if (dmalloc(&dn, m) != hipSuccess || dmalloc(&dp, m) != hipSuccess || dmalloc(&dd, m) != hipSuccess) {
        if (dn) (void)hipFree(dn);
        if (dp) (void)hipFree(dp);
        if (dd) (void)hipFree(dd);
        return fail(REX_DEVICE_ERROR, "hipMalloc hierarchy failed");
    }
    hipError_t e = hipMemcpy(dn, s.nodes.data(), m * sizeof(BinaryNode), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, s.parents.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dd, depth.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dn);
        (void)hipFree(dp);
        (void)hipFree(dd);
        return check_hip(e, "hierarchy upload");
    }
    release_hierarchy(s);
    s.dNodes = dn;
    s.dParents = dp;
    s.dDepth = dd;
    return REX_SUCCESS;
}

void release_hierarchy(Scene& s) {
    if (s.dNodes) (void)hipFree(s.dNodes);
    if (s.dParents) (void)hipFree(s.dParents);
    if (s.dDepth) (void)hipFree(s.dDepth);
    s.dNodes = nullptr;
    s.dParents = nullptr;
    s.dDepth = nullptr;
}

rex_status ensure_host_nodes(Scene& s) {
    if (!s.hostNodesStale) return REX_SUCCESS;
    if (rex_status r = check_hip(hipMemcpy(s.nodes.data(), s.dNodes, s.nodes.size() * sizeof(BinaryNode),
                                           hipMemcpyDeviceToHost),
                                 "node download")) {
        return r;
    }
    s.hostNodesStale = false;
    return REX_SUCCESS;
}

rex_status lbvh_build_device(const rex_triangle* tris, uint32_t n, const float slo[3], const float shi[3],
                             Scene& s) {
    hipStream_t st = nullptr;
    if (rex_status r = check_hip(hipGetDevice(&s.device), "hipGetDevice")) return r;
    const auto t0 = std::chrono::steady_clock::now();
    Temp tmp;
    rex_triangle *dIn = nullptr, *dLeaf = nullptr;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    uint32_t *vals = nullptr, *vals2 = nullptr;
    if (rex_status r = check_hip(tmp.alloc(&dIn, n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&dLeaf, n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&keys, n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&keys2, n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&vals, n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&vals2, n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(hipMemcpyAsync(dIn, tris, (size_t)n * sizeof(rex_triangle), hipMemcpyHostToDevice, st),
                                 "triangle upload"))
        return r;
    if (rex_status r = check_hip(hipStreamSynchronize(st), "triangle upload")) return r;
    const auto d0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_morton, dim3(grid(n)), dim3(kB), 0, st, dIn, n, slo[0], slo[1], slo[2], shi[0], shi[1],
                       shi[2], keys, vals);
    size_t sortBytes = 0;
    if (rex_status r = check_hip(hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, keys, keys2, vals, vals2,
                                                                    (int)n, 0, 64, st),
                                 "radix sort"))
        return r;
    uint8_t* sortTemp = nullptr;
    if (rex_status r = check_hip(tmp.alloc(&sortTemp, sortBytes), "hipMalloc sort")) return r;
    if (rex_status r = check_hip(hipcub::DeviceRadixSort::SortPairs(sortTemp, sortBytes, keys, keys2, vals, vals2,
                                                                    (int)n, 0, 64, st),
                                 "radix sort"))
        return r;
    const uint32_t m = 2u * n - 1u;
    uint32_t *child = nullptr, *parent = nullptr, *first = nullptr, *pre = nullptr, *depth = nullptr, *flags = nullptr,
             *visits = nullptr;
    if (rex_status r = check_hip(tmp.alloc(&child, 2u * n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&parent, m), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&first, n), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&pre, m), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&depth, m), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&flags, 2), "hipMalloc lbvh")) return r;
    if (rex_status r = check_hip(tmp.alloc(&visits, m), "hipMalloc lbvh")) return r;
    BinaryNode* dn = nullptr;
    uint32_t *dp = nullptr, *dd = nullptr;
    if (dmalloc(&dn, m) != hipSuccess || dmalloc(&dp, m) != hipSuccess || dmalloc(&dd, m) != hipSuccess) {
        if (dn) (void)hipFree(dn);
        if (dp) (void)hipFree(dp);
        if (dd) (void)hipFree(dd);
        return fail(REX_DEVICE_ERROR, "hipMalloc hierarchy failed");
    }
    release_hierarchy(s);
    s.dNodes = dn;
    s.dParents = dp;
    s.dDepth = dd;
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kB), 0, st, flags, 2u);
    hipLaunchKernelGGL(k_zero_words, dim3(grid(m)), dim3(kB), 0, st, visits, m);
    if (n > 1u) hipLaunchKernelGGL(k_karras, dim3(grid(n - 1u)), dim3(kB), 0, st, keys2, (int)n, child, parent, first);
    hipLaunchKernelGGL(k_preorder, dim3(grid(m)), dim3(kB), 0, st, child, parent, first, n, pre, depth, flags);
    hipLaunchKernelGGL(k_emit_binary, dim3(grid(m)), dim3(kB), 0, st, child, parent, pre, depth, n, dn, dp, dd);
    hipLaunchKernelGGL(k_gather_tris, dim3(grid(n)), dim3(kB), 0, st, dIn, vals2, n, dLeaf);
    hipLaunchKernelGGL(k_refit, dim3(grid(m)), dim3(kB), 0, st, dn, dp, m, dLeaf, visits);
    uint32_t fl[2] = {0, 0};
    if (rex_status r = check_hip(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, st), "copy")) return r;
    if (rex_status r = check_hip(hipStreamSynchronize(st), "lbvh build")) return r;
    if (fl[0] != 0u) return fail(REX_UNSUPPORTED, "LBVH depth exceeds 128");
    s.triCount = n;
    s.nodeCount = m;
    s.maxLeaf = 1u;
    s.maxDepth = fl[1];
    if (rex_status r = device_layout(s, dLeaf, n, st)) return r;
    s.deviceBuildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - d0).count();
    // Host copies: nodes and parents (pre-order), leaf order, leaf-order triangles.
    s.nodes.resize(m);
    s.parents.resize(m);
    s.order.resize(n);
    hipError_t e = hipMemcpy(s.nodes.data(), dn, (size_t)m * sizeof(BinaryNode), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(s.parents.data(), dp, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(s.order.data(), vals2, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (rex_status r = check_hip(e, "hierarchy download")) return r;
    s.hostNodesStale = false;
    s.leafTris.resize(n);
    parallel_ranges(n, [&](size_t b, size_t en) {
        for (size_t j = b; j < en; ++j) s.leafTris[j] = tris[s.order[j]];
    });
    s.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return REX_SUCCESS;
}

namespace {

void release_traversal(Scene& s) {
    void* ptrs[] = {s.dTNodes, s.dTParents, s.dTDepth};
    for (void* p : ptrs) {
        if (p) (void)hipFree(p);
    }
    s.dTNodes = nullptr;
    s.dTParents = nullptr;
    s.dTDepth = nullptr;
}

}  // namespace

// The traversal hierarchy refit on the device (k_refit over its nodes: one
// triangle per leaf, at the leaf-order positions) and its compressed nodes
// emitted in place of the reference hierarchy's.  Any failure leaves the
// reference hierarchy's compressed nodes (device_layout's) in use and drops
// the traversal hierarchy.
void refit_traversal(Scene& s, const rex_triangle* dLeafTris, hipStream_t st) {
    const uint32_t m = static_cast<uint32_t>(s.tnodes.size());
    if (m == 0u) return;
    auto drop = [&]() {
        release_traversal(s);
        s.tnodes.clear();
        s.tnodes.shrink_to_fit();
        (void)hipGetLastError();
    };
    if (s.dTNodes == nullptr) {
        std::vector<uint32_t> parents(m, kInvalid), depth(m, 0u);
        for (uint32_t i = 0; i < m; ++i) {
            const BinaryNode& b = s.tnodes[i];
            if (b.count == 0u) parents[b.left] = parents[b.right] = i;
        }
        for (uint32_t i = 1; i < m; ++i) depth[i] = depth[parents[i]] + 1u;  // parents precede children
        if (dmalloc(&s.dTNodes, m) != hipSuccess || dmalloc(&s.dTParents, m) != hipSuccess ||
            dmalloc(&s.dTDepth, m) != hipSuccess ||
            hipMemcpy(s.dTNodes, s.tnodes.data(), m * sizeof(BinaryNode), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s.dTParents, parents.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s.dTDepth, depth.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
            drop();
            return;
        }
    }
    Temp tmp;
    uint32_t *visits = nullptr, *interior = nullptr, *w4flag = nullptr, *w4Idx = nullptr, *w4Depth = nullptr;
    if (tmp.alloc(&visits, m) != hipSuccess || tmp.alloc(&interior, m) != hipSuccess ||
        tmp.alloc(&w4flag, m) != hipSuccess || tmp.alloc(&w4Idx, m) != hipSuccess || tmp.alloc(&w4Depth, 4) != hipSuccess) {
        drop();
        return;
    }
    hipLaunchKernelGGL(k_zero_words, dim3(grid(m)), dim3(kB), 0, st, visits, m);
    hipLaunchKernelGGL(k_refit, dim3(grid(m)), dim3(kB), 0, st, s.dTNodes, s.dTParents, m, dLeafTris, visits);
    hipLaunchKernelGGL(k_flags, dim3(grid(m)), dim3(kB), 0, st, s.dTNodes, s.dTDepth, m, interior, w4flag);
    uint32_t last[2] = {0u, 0u};
    Wide4Q* q = nullptr;
    uint32_t depth4 = 0u;
    BinaryNode root{};
    if (exclusive_sum(w4flag, w4Idx, m, tmp, st) != hipSuccess ||
        hipMemcpyAsync(&last[0], w4flag + (m - 1), 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&last[1], w4Idx + (m - 1), 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        drop();
        return;
    }
    const uint32_t count = last[0] + last[1];
    if (count == 0u || dmalloc(&q, count) != hipSuccess) {
        drop();
        return;
    }
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kB), 0, st, w4Depth, 1u);
    hipLaunchKernelGGL(k_emit_wide4q, dim3(grid(m)), dim3(kB), 0, st, s.dTNodes, s.dTDepth, m, w4Idx, q, w4Depth);
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&depth4, w4Depth, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&root, s.dTNodes, sizeof(BinaryNode), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipFree(q);
        drop();
        return;
    }
    // device_layout already waited for in-flight work on the old layout.
    if (s.dWide4Q) (void)hipFree(s.dWide4Q);
    s.dWide4Q = q;
    s.wide4qCount = count;
    s.deviceBytes += (uint64_t)count * sizeof(Wide4Q) + (uint64_t)m * (sizeof(BinaryNode) + 8u);
    s.view.wide4q = q;
    s.view.rootRefQ = root.count != 0u ? (kLeafBit | ((root.count - 1u) << 28) | root.first) : 0u;
    s.view.wide4Depth = std::max(s.view.wide4Depth, depth4);
}

rex_status refit_device(Scene& s, std::vector<rex_triangle>&& leafTris, uint32_t generation) {
    hipStream_t st = nullptr;
    const uint32_t n = s.triCount, m = s.nodeCount;
    if (s.dNodes == nullptr) {
        if (rex_status r = upload_hierarchy(s)) return r;
    }
    Temp tmp;
    rex_triangle* dLeaf = nullptr;
    uint32_t* visits = nullptr;
    if (rex_status r = check_hip(tmp.alloc(&dLeaf, n), "hipMalloc refit")) return r;
    if (rex_status r = check_hip(tmp.alloc(&visits, m), "hipMalloc refit")) return r;
    // The node boxes are rewritten in place: wait for work reading the old
    // layout (it reads the separate wide / compressed arrays, not these).
    if (rex_status r = check_hip(hipMemcpyAsync(dLeaf, leafTris.data(), (size_t)n * sizeof(rex_triangle),
                                                hipMemcpyHostToDevice, st),
                                 "refit upload"))
        return r;
    if (rex_status r = check_hip(hipStreamSynchronize(st), "refit upload")) return r;
    const auto d0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_zero_words, dim3(grid(m)), dim3(kB), 0, st, visits, m);
    hipLaunchKernelGGL(k_refit, dim3(grid(m)), dim3(kB), 0, st, s.dNodes, s.dParents, m, dLeaf, visits);
    if (rex_status r = check_hip(hipGetLastError(), "refit")) return r;
    if (rex_status r = device_layout(s, dLeaf, n, st)) {
        // The device boxes no longer match the layout still in use; the scene
        // cannot be traced until rebuilt.
        s.view = SceneView{};
        s.generation = 0u;
        return r;
    }
    refit_traversal(s, dLeaf, st);
    s.deviceBuildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - d0).count();
    s.leafTris = std::move(leafTris);
    s.hostNodesStale = true;
    s.generation = generation;
    return REX_SUCCESS;
}

}  // namespace rex
