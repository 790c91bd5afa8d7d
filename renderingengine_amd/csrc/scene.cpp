// Scene build: validation, binned-SAH hierarchy, conversion to the device
// layout and upload.
//
// Validation follows the reference's rules: finite, non-degenerate triangles
// (src/rt/software_gpu/SoftwareGpu.cpp:142-158), unique traversal IDs
// (:171-182), depth <= 128 (include/rt/cpu/Bvh.hpp:568-572).  The hierarchy
// does not depend on alpha masks; a traceable scene resolves each masked
// triangle to accepted or skipped from its rex_alpha_desc (a skipped one gets
// zero edges in the device triangle stream) and rejects masked geometry it
// cannot resolve instead of treating it as opaque.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>

#include "rex_internal.hpp"
#include <atomic>
#include <thread>

#include "wide4q.hpp"

namespace rex {
namespace {

bool triangle_valid(const rex_triangle& t) {
    const float* v0 = t.positions[0];
    const float* v1 = t.positions[1];
    const float* v2 = t.positions[2];
    for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(v0[k]) || !std::isfinite(v1[k]) || !std::isfinite(v2[k])) return false;
    }
    if (t.metadata[2] == kInvalid) return false;
    const float e0[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const float e1[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const float cr[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
    const float area2 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
    const float d0 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    const float d1 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const float scale = std::max(d0, d1);
    return std::isfinite(area2) && area2 > scale * scale * 1.0e-14f;
}

template <typename T>
rex_status upload(const std::vector<T>& host, T** dev, uint64_t& bytes) {
    const size_t n = std::max<size_t>(host.size(), 1) * sizeof(T);
    if (rex_status st = check_hip(hipMalloc(reinterpret_cast<void**>(dev), n), "hipMalloc scene buffer")) return st;
    bytes += n;
    if (!host.empty()) {
        if (rex_status st = check_hip(hipMemcpy(*dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice),
                                      "hipMemcpy scene buffer"))
            return st;
    }
    return REX_SUCCESS;
}

}  // namespace

rex_status host_build(const rex_triangle* tris, uint32_t n, uint32_t maxLeafSize, uint32_t threads, Scene& s) {
    if (tris == nullptr || n == 0) return fail(REX_INVALID_ARGUMENT, "scene needs at least one triangle");
    if (n >= kMaxTriangles) return fail(REX_UNSUPPORTED, "scene exceeds 2^28 triangles");
    const uint32_t maxLeaf = maxLeafSize == 0 ? 4u : maxLeafSize;
    if (maxLeaf > kMaxLeafSize) return fail(REX_UNSUPPORTED, "maxLeafSize above 8");

    std::vector<float> positions(static_cast<size_t>(n) * 9u);
    std::vector<uint32_t> primIds(n);
    for (uint32_t i = 0; i < n; ++i) {
        const rex_triangle& t = tris[i];
        if (!triangle_valid(t)) {
            return fail(REX_INVALID_ARGUMENT, "triangle " + std::to_string(i) + " is non-finite or degenerate");
        }
        for (int v = 0; v < 3; ++v) {
            for (int k = 0; k < 3; ++k) positions[9u * i + 3u * v + k] = t.positions[v][k];
        }
        primIds[i] = t.metadata[2];
    }
    {
        std::vector<uint32_t> ids(primIds);
        std::sort(ids.begin(), ids.end());
        if (std::adjacent_find(ids.begin(), ids.end()) != ids.end()) {
            return fail(REX_INVALID_ARGUMENT, "traversal IDs (metadata.z) must be unique");
        }
    }

    const auto t0 = std::chrono::steady_clock::now();
    const int rc = build_binned_sah(positions.data(), primIds.data(), n, maxLeaf, threads, s.nodes, s.order,
                                    s.maxDepth);
    if (rc == -2) return fail(REX_UNSUPPORTED, "BVH depth exceeds 128");
    if (rc != 0) return fail(REX_INVALID_ARGUMENT, "BVH build failed");
    s.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s.triCount = n;
    s.maxLeaf = maxLeaf;
    s.nodeCount = static_cast<uint32_t>(s.nodes.size());
    s.parents.assign(s.nodeCount, kInvalid);
    for (uint32_t i = 0; i < s.nodeCount; ++i) {
        const BinaryNode& b = s.nodes[i];
        if (b.count == 0) {
            s.parents[b.left] = i;
            s.parents[b.right] = i;
        }
    }
    return REX_SUCCESS;
}

void export_flat(const Scene& s, void* nodes48, uint32_t* primitiveOrder) {
    if (nodes48 != nullptr) {
        // SoftwareNodeRecord (SoftwareGpu.cpp:523-531 EncodeNode, :1073-1089 links).
        uint32_t* out = static_cast<uint32_t*>(nodes48);
        for (uint32_t i = 0; i < s.nodeCount; ++i) {
            const BinaryNode& b = s.nodes[i];
            const float f[8] = {b.bmin[0], b.bmin[1], b.bmin[2], 0.0f, b.bmax[0], b.bmax[1], b.bmax[2], 0.0f};
            std::memcpy(out + 12u * i, f, sizeof(f));
            uint32_t* l = out + 12u * i + 8u;
            if (b.count != 0) {
                l[0] = b.first;
                l[1] = b.count;
                l[2] = kInvalid;
            } else {
                l[0] = b.left;
                l[1] = 0u;
                l[2] = b.right;
            }
            l[3] = s.parents[i];
        }
    }
    if (primitiveOrder != nullptr) std::memcpy(primitiveOrder, s.order.data(), s.order.size() * sizeof(uint32_t));
}

namespace {

uint32_t expand_morton_bits(uint32_t v) {
    v &= 0x000003ffu;
    v = (v | (v << 16u)) & 0x030000ffu;
    v = (v | (v << 8u)) & 0x0300f00fu;
    v = (v | (v << 4u)) & 0x030c30c3u;
    v = (v | (v << 2u)) & 0x09249249u;
    return v;
}

uint32_t quantize_morton_axis(float value, float lo, float hi) {
    const double extent = static_cast<double>(hi) - static_cast<double>(lo);
    const double t = extent > 0.0 ? (static_cast<double>(value) - static_cast<double>(lo)) / extent : 0.5;
    const double c = std::min(std::max(t, 0.0), 1.0);
    return std::min(static_cast<uint32_t>(c * 1024.0), 1023u);
}

}  // namespace

rex_status lbvh_validate(const rex_triangle* tris, uint32_t n, float slo[3], float shi[3]) {
    if (tris == nullptr || n == 0) return fail(REX_INVALID_ARGUMENT, "scene needs at least one triangle");
    if (n >= kMaxTriangles) return fail(REX_UNSUPPORTED, "scene exceeds 2^28 triangles");
    std::vector<uint32_t> ids(n);
    std::vector<uint32_t> bad(64, kInvalid);
    std::vector<std::array<float, 6>> boxes(64, {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY});
    std::vector<size_t> begins(64, ~size_t{0});
    std::atomic<uint32_t> slot{0};
    parallel_ranges(n, [&](size_t b, size_t e) {
        const uint32_t me = slot.fetch_add(1u);
        begins[me] = b;
        std::array<float, 6>& box = boxes[me];
        for (size_t i = b; i < e; ++i) {
            const rex_triangle& t = tris[i];
            if (!triangle_valid(t)) {
                bad[me] = static_cast<uint32_t>(i);
                return;
            }
            for (int v = 0; v < 3; ++v) {
                for (int a = 0; a < 3; ++a) {
                    box[a] = std::min(box[a], t.positions[v][a]);
                    box[3 + a] = std::max(box[3 + a], t.positions[v][a]);
                }
            }
            ids[i] = t.metadata[2];
        }
    });
    for (uint32_t k = 0; k < bad.size(); ++k) {
        if (bad[k] == kInvalid) continue;
        return fail(REX_INVALID_ARGUMENT, "triangle " + std::to_string(bad[k]) + " is non-finite or degenerate");
    }
    // Chunk boxes folded in chunk order: the same (first-minimum) result as
    // one sequential std::min / std::max pass over the vertices.
    std::vector<uint32_t> byBegin;
    for (uint32_t k = 0; k < begins.size(); ++k) {
        if (begins[k] != ~size_t{0}) byBegin.push_back(k);
    }
    std::sort(byBegin.begin(), byBegin.end(), [&](uint32_t x, uint32_t y) { return begins[x] < begins[y]; });
    for (int a = 0; a < 3; ++a) {
        slo[a] = INFINITY;
        shi[a] = -INFINITY;
        for (uint32_t k : byBegin) {
            slo[a] = std::min(slo[a], boxes[k][a]);
            shi[a] = std::max(shi[a], boxes[k][3 + a]);
        }
    }
    std::sort(ids.begin(), ids.end());
    if (std::adjacent_find(ids.begin(), ids.end()) != ids.end()) {
        return fail(REX_INVALID_ARGUMENT, "traversal IDs (metadata.z) must be unique");
    }
    return REX_SUCCESS;
}

rex_status lbvh_build(const rex_triangle* tris, uint32_t n, Scene& s) {
    float slo[3], shi[3];
    if (rex_status st = lbvh_validate(tris, n, slo, shi)) return st;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = tris[i].metadata[2];
    // Key = Morton code of the centroid (midpoint of the triangle's box) in
    // the high word, traversal ID in the low word: unique, so a plain sort
    // gives the reference's stable (code, ID) order.
    std::vector<std::pair<uint64_t, uint32_t>> keys(n);
    parallel_ranges(n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const rex_triangle& t = tris[i];
            uint32_t q[3];
            for (int a = 0; a < 3; ++a) {
                const float lo = std::min(std::min(t.positions[0][a], t.positions[1][a]), t.positions[2][a]);
                const float hi = std::max(std::max(t.positions[0][a], t.positions[1][a]), t.positions[2][a]);
                q[a] = quantize_morton_axis(midpoint_f(lo, hi), slo[a], shi[a]);
            }
            const uint32_t code = (expand_morton_bits(q[0]) << 2u) | (expand_morton_bits(q[1]) << 1u) |
                                  expand_morton_bits(q[2]);
            keys[i] = {(static_cast<uint64_t>(code) << 32u) | ids[i], static_cast<uint32_t>(i)};
        }
    });
    std::sort(keys.begin(), keys.end());
    s.order.resize(n);
    for (uint32_t i = 0; i < n; ++i) s.order[i] = keys[i].second;
    // Karras hierarchy over [first, last), nodes in pre-order.
    s.nodes.clear();
    s.nodes.reserve(2u * n - 1u);
    s.maxDepth = 0;
    struct Item {
        uint32_t first, last, depth, parentSlot;  // parentSlot: node * 2 + side, or kInvalid
    };
    std::vector<Item> work{{0u, n, 1u, kInvalid}};
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        if (it.depth > 128u) return fail(REX_UNSUPPORTED, "LBVH depth exceeds 128");
        s.maxDepth = std::max(s.maxDepth, it.depth);
        const uint32_t idx = static_cast<uint32_t>(s.nodes.size());
        s.nodes.push_back(BinaryNode{});
        BinaryNode& b = s.nodes.back();
        if (it.parentSlot != kInvalid) {
            BinaryNode& p = s.nodes[it.parentSlot >> 1];
            ((it.parentSlot & 1u) ? p.right : p.left) = idx;
        }
        if (it.last - it.first == 1u) {
            b.first = it.first;
            b.count = 1u;
            b.left = b.right = kInvalid;
            continue;
        }
        b.count = 0u;
        const uint64_t firstKey = keys[it.first].first, lastKey = keys[it.last - 1u].first;
        const int common = __builtin_clzll(firstKey ^ lastKey);
        uint32_t split = it.first, step = it.last - it.first;
        do {
            step = (step + 1u) >> 1u;
            const uint32_t cand = split + step;
            if (cand < it.last - 1u && __builtin_clzll(firstKey ^ keys[cand].first) > common) split = cand;
        } while (step > 1u);
        ++split;
        // Right pushed first so the left subtree is numbered next (pre-order).
        work.push_back({split, it.last, it.depth + 1u, idx * 2u + 1u});
        work.push_back({it.first, split, it.depth + 1u, idx * 2u});
    }
    std::vector<rex_triangle> leafTris(n);
    for (uint32_t j = 0; j < n; ++j) leafTris[j] = tris[s.order[j]];
    refit_nodes(s.nodes, leafTris);
    s.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s.triCount = n;
    s.maxLeaf = 1u;
    s.nodeCount = static_cast<uint32_t>(s.nodes.size());
    s.parents.assign(s.nodeCount, kInvalid);
    for (uint32_t i = 0; i < s.nodeCount; ++i) {
        if (s.nodes[i].count == 0) {
            s.parents[s.nodes[i].left] = i;
            s.parents[s.nodes[i].right] = i;
        }
    }
    s.leafTris = std::move(leafTris);
    return REX_SUCCESS;
}

rex_status set_alpha(const rex_alpha_desc* alpha, Scene& s) {
    s.hasAlpha = false;
    s.alphaTable.clear();
    if (alpha == nullptr) return REX_SUCCESS;
    if (alpha->materialCount != 0u && alpha->materials == nullptr) {
        return fail(REX_INVALID_ARGUMENT, "alpha description without materials");
    }
    if (alpha->atlasLayerCount != 0u && alpha->layerAlpha == nullptr) {
        return fail(REX_INVALID_ARGUMENT, "alpha description without layer alphas");
    }
    s.alphaTable.resize(alpha->materialCount);
    for (uint32_t m = 0; m < alpha->materialCount; ++m) {
        const rex_material_v0& mat = alpha->materials[m];
        const uint32_t layer = mat.textureImageIndices[0];
        const uint32_t sampler = mat.textureSamplerIndices[0];
        float a = mat.baseColorFactor[3];
        bool valid = true;
        if (layer == kInvalid) {
            valid = sampler == kInvalid;
        } else if (layer >= alpha->atlasLayerCount || sampler != alpha->samplerId) {
            valid = false;
        } else {
            a *= alpha->layerAlpha[layer];
        }
        s.alphaTable[m] = !valid ? kAlphaInvalid : (a < mat.surfaceParams[3] ? kAlphaSkip : kAlphaAccept);
    }
    s.hasAlpha = true;
    return REX_SUCCESS;
}

uint8_t alpha_outcome(const Scene& s, const rex_triangle& t) {
    if ((t.metadata[0] & kGeometryFlagAlphaMask) == 0u && (t.metadata[1] & kMaterialFlagAlphaMask) == 0u) {
        return kAlphaAccept;
    }
    const uint32_t m = t.identity[3];
    return m < s.alphaTable.size() ? s.alphaTable[m] : kAlphaInvalid;
}

rex_status alpha_validate(const Scene& s, const rex_triangle* tris, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        if (alpha_outcome(s, tris[i]) != kAlphaInvalid) continue;
        if (!s.hasAlpha) {
            return fail(REX_UNSUPPORTED, "alpha-masked geometry needs rex_scene_desc.alpha (triangle " +
                                             std::to_string(i) + ")");
        }
        return fail(REX_UNSUPPORTED, "alpha-masked triangle " + std::to_string(i) +
                                         " has a material index or alpha mapping the reference traverses as invalid");
    }
    return REX_SUCCESS;
}

// The compressed walk's traversal hierarchy (Scene::tnodes): a binned-SAH
// build with one triangle per leaf over the leaf-order triangles.  The
// reference hierarchy's leaves of up to 4 triangles leave a wave's lanes
// idle while the lanes with the largest leaves finish theirs; single-triangle
// leaves cull each triangle by its own box.  Hits do not depend on the
// hierarchy: a candidate is admitted by its reference leaf box (LeafBox).
void build_traversal_tree(Scene& s, uint32_t threads) {
    s.tnodes.clear();
    const uint32_t n = static_cast<uint32_t>(s.leafTris.size());
    if (s.maxLeaf <= 1u || n < 2u) return;
    std::vector<float> positions(static_cast<size_t>(n) * 9u);
    std::vector<uint32_t> primIds(n);
    parallel_ranges(n, [&](size_t b, size_t e) {
        for (size_t j = b; j < e; ++j) {
            const rex_triangle& t = s.leafTris[j];
            for (int v = 0; v < 3; ++v) {
                for (int k = 0; k < 3; ++k) positions[9u * j + 3u * v + k] = t.positions[v][k];
            }
            primIds[j] = t.metadata[2];
        }
    });
    std::vector<BinaryNode> tn;
    std::vector<uint32_t> torder;
    uint32_t depth = 0;
    if (build_binned_sah(positions.data(), primIds.data(), n, 1u, threads, tn, torder, depth) != 0) return;
    for (BinaryNode& b : tn) {
        if (b.count == 0) continue;
        if (b.count != 1u) return;  // keep the reference hierarchy
        b.first = torder[b.first];
    }
    s.tnodes = std::move(tn);
}

rex_status scene_build(const rex_triangle* tris, uint32_t n, const rex_scene_desc& desc, Scene& s) {
    if (desc.fingerprint == 0 || desc.generation == 0) {
        return fail(REX_INVALID_ARGUMENT, "scene fingerprint and generation must be nonzero");
    }
    if (desc.buildMethod > REX_BUILD_KARRAS_LBVH) return fail(REX_INVALID_ARGUMENT, "unknown build method");
    if (rex_status st = set_alpha(desc.alpha, s)) return st;
    if (tris != nullptr) {
        if (rex_status st = alpha_validate(s, tris, n)) return st;
    }
    if (desc.buildMethod == REX_BUILD_KARRAS_LBVH) {
        // BuildKarrasLbvh makes one triangle per leaf.
        if (desc.maxLeafSize > 1u) return fail(REX_UNSUPPORTED, "REX_BUILD_KARRAS_LBVH builds leaves of one triangle");
        float slo[3], shi[3];
        if (rex_status st = lbvh_validate(tris, n, slo, shi)) return st;
        s.fingerprint = desc.fingerprint;
        s.generation = desc.generation;
        return lbvh_build_device(tris, n, slo, shi, s);
    }
    if (rex_status st = host_build(tris, n, desc.maxLeafSize, desc.buildThreads, s)) return st;
    s.fingerprint = desc.fingerprint;
    s.generation = desc.generation;
    s.leafTris.resize(n);
    for (uint32_t j = 0; j < n; ++j) s.leafTris[j] = tris[s.order[j]];
    build_traversal_tree(s, desc.buildThreads);
    return scene_layout(s);
}

void refit_nodes(std::vector<BinaryNode>& nodes, const std::vector<rex_triangle>& leafTris) {
    // Pre-order numbering puts every child after its parent, so one reverse
    // pass sees both children of a node before the node itself.
    for (size_t k = nodes.size(); k-- > 0;) {
        BinaryNode& b = nodes[k];
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (b.count != 0) {
            for (uint32_t j = b.first; j < b.first + b.count; ++j) {
                for (int v = 0; v < 3; ++v) {
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min(lo[a], leafTris[j].positions[v][a]);
                        hi[a] = std::max(hi[a], leafTris[j].positions[v][a]);
                    }
                }
            }
        } else {
            for (uint32_t c : {b.left, b.right}) {
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], nodes[c].bmin[a]);
                    hi[a] = std::max(hi[a], nodes[c].bmax[a]);
                }
            }
        }
        for (int a = 0; a < 3; ++a) {
            b.bmin[a] = lo[a];
            b.bmax[a] = hi[a];
        }
    }
}

rex_status refit_leaf_tris(const Scene& s, const rex_triangle* tris, uint32_t n, std::vector<rex_triangle>& out) {
    if (tris == nullptr || n != s.triCount) return fail(REX_INVALID_ARGUMENT, "refit needs the scene's triangle count");
    out.resize(n);
    std::vector<uint32_t> firstBad(16, kInvalid), badKind(16, 0u);
    std::atomic<uint32_t> slot{0};
    parallel_ranges(n, [&](size_t b, size_t e) {
        const uint32_t me = slot.fetch_add(1u);
        for (uint32_t j = static_cast<uint32_t>(b); j < e; ++j) {
            const rex_triangle& t = tris[s.order[j]];
            uint32_t kind = t.metadata[2] != s.leafTris[j].metadata[2] ? 1u : (!triangle_valid(t) ? 2u : 0u);
            if (kind == 0u && alpha_outcome(s, t) == kAlphaInvalid) kind = 3u;
            if (kind != 0u) {
                firstBad[me] = j;
                badKind[me] = kind;
                return;
            }
            out[j] = t;
        }
    });
    for (size_t k = 0; k < firstBad.size(); ++k) {
        if (firstBad[k] == kInvalid) continue;
        if (badKind[k] == 1u) return fail(REX_INVALID_ARGUMENT, "refit triangles must keep the traversal IDs in the same order");
        if (badKind[k] == 3u) {
            return fail(REX_UNSUPPORTED, "alpha-masked triangle " + std::to_string(s.order[firstBad[k]]) +
                                             (s.hasAlpha ? " has a material index or alpha mapping the reference "
                                                           "traverses as invalid"
                                                         : " needs rex_scene_desc.alpha"));
        }
        return fail(REX_INVALID_ARGUMENT, "triangle " + std::to_string(s.order[firstBad[k]]) + " is non-finite or degenerate");
    }
    return REX_SUCCESS;
}

namespace {

// Device buffers of one layout, built before they replace a scene's.
struct Layout {
    WideNode* dWide = nullptr;
    TriPos* dTris = nullptr;
    TriAttr* dAttrs = nullptr;
    Wide4* dWide4 = nullptr;
    Wide4Q* dWide4Q = nullptr;
    LeafBox* dLeafBox = nullptr;
    uint32_t wideCount = 0, wide4Count = 0, wide4qCount = 0;
    uint64_t bytes = 0;
    SceneView view{};
    void release() {
        void* ptrs[] = {dWide, dTris, dAttrs, dWide4, dWide4Q, dLeafBox};
        for (void* p : ptrs) {
            if (p) (void)hipFree(p);
        }
        *this = Layout{};
    }
};

// 4-wide nodes: each binary interior node reached from the root through
// interior grandparent links becomes one node whose slots are its
// grandchildren (or a leaf child itself), numbered in pre-order.  depth: the
// levels of 4-wide nodes.
void collapse_wide4(const std::vector<BinaryNode>& nodes, std::vector<Wide4>& wide4, uint32_t& wide4Depth) {
    wide4.clear();
    wide4Depth = 0;
    if (nodes[0].count == 0) {
        struct Item {
            uint32_t bin, level, parentSlot, parent;
        };
        std::vector<Item> work{{0u, 1u, 0u, kInvalid}};
        while (!work.empty()) {
            const Item it = work.back();
            work.pop_back();
            const uint32_t w = static_cast<uint32_t>(wide4.size());
            wide4.push_back(Wide4{});
            if (it.parent != kInvalid) (&wide4[it.parent].ref.x)[it.parentSlot] = w;
            wide4Depth = std::max(wide4Depth, it.level);
            uint32_t slots[4] = {kInvalid, kInvalid, kInvalid, kInvalid};
            uint32_t k = 0;
            for (uint32_t child : {nodes[it.bin].left, nodes[it.bin].right}) {
                const BinaryNode& c = nodes[child];
                if (c.count != 0) {
                    slots[k++] = child;
                } else {
                    slots[k++] = c.left;
                    slots[k++] = c.right;
                }
            }
            Wide4& W = wide4[w];
            float* lo[3] = {&W.lox.x, &W.loy.x, &W.loz.x};
            float* hi[3] = {&W.hix.x, &W.hiy.x, &W.hiz.x};
            for (uint32_t j = 0; j < 4; ++j) {
                (&W.bin.x)[j] = slots[j];
                (&W.ref.x)[j] = kInvalid;
                for (int a = 0; a < 3; ++a) {
                    lo[a][j] = slots[j] == kInvalid ? 0.0f : nodes[slots[j]].bmin[a];
                    hi[a][j] = slots[j] == kInvalid ? 0.0f : nodes[slots[j]].bmax[a];
                }
                if (slots[j] != kInvalid && nodes[slots[j]].count != 0) {
                    const BinaryNode& c = nodes[slots[j]];
                    (&W.ref.x)[j] = kLeafBit | ((c.count - 1u) << 28) | c.first;
                }
            }
            // Push interior slots in reverse so the lowest slot is numbered first.
            for (int j = 3; j >= 0; --j) {
                if (slots[j] != kInvalid && nodes[slots[j]].count == 0) {
                    work.push_back({slots[j], it.level + 1u, static_cast<uint32_t>(j), w});
                }
            }
        }
    }
}

// Child reference of a hierarchy's root for the 4-wide layouts (node 0 or a leaf).
uint32_t wide4_root_ref(const std::vector<BinaryNode>& nodes) {
    const BinaryNode& r = nodes[0];
    return r.count == 0 ? 0u : (kLeafBit | ((r.count - 1u) << 28) | r.first);
}

rex_status build_layout(const Scene& s, const std::vector<BinaryNode>& nodes, const std::vector<rex_triangle>& leafTris,
                        Layout& L) {
    const uint32_t n = static_cast<uint32_t>(leafTris.size());

    // Wide indices: interior nodes in pre-order.
    const uint32_t m = static_cast<uint32_t>(nodes.size());
    std::vector<uint32_t> wideIndex(m, kInvalid);
    uint32_t wideCount = 0;
    for (uint32_t i = 0; i < m; ++i) {
        if (nodes[i].count == 0) wideIndex[i] = wideCount++;
    }
    L.wideCount = wideCount;
    auto refOf = [&](uint32_t child) -> uint32_t {
        const BinaryNode& c = nodes[child];
        if (c.count == 0) return wideIndex[child];
        return kLeafBit | ((c.count - 1u) << 28) | c.first;
    };
    std::vector<WideNode> wide(wideCount);
    for (uint32_t i = 0; i < m; ++i) {
        const BinaryNode& b = nodes[i];
        if (b.count != 0) continue;
        const BinaryNode& Lc = nodes[b.left];
        const BinaryNode& Rc = nodes[b.right];
        WideNode& w = wide[wideIndex[i]];
        w.a = {Lc.bmin[0], Lc.bmin[1], Lc.bmin[2], Lc.bmax[0]};
        w.b = {Lc.bmax[1], Lc.bmax[2], Rc.bmin[0], Rc.bmin[1]};
        w.c = {Rc.bmin[2], Rc.bmax[0], Rc.bmax[1], Rc.bmax[2]};
        uint32_t parentLink = kInvalid;
        if (s.parents[i] != kInvalid) {
            const uint32_t p = s.parents[i];
            parentLink = wideIndex[p] | (nodes[p].right == i ? 0x80000000u : 0u);
        }
        w.link = {refOf(b.left), refOf(b.right), b.left, parentLink};
    }

    std::vector<Wide4> wide4;
    uint32_t wide4Depth = 0;
    collapse_wide4(nodes, wide4, wide4Depth);
    L.wide4Count = static_cast<uint32_t>(wide4.size());

    std::vector<TriPos> tpos(n);
    std::vector<TriAttr> tattr(n);
    parallel_ranges(n, [&](size_t jb, size_t je) {
    for (uint32_t j = static_cast<uint32_t>(jb); j < je; ++j) {
        const rex_triangle& t = leafTris[j];
        const float* v0 = t.positions[0];
        const float* v1 = t.positions[1];
        const float* v2 = t.positions[2];
        float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        if (alpha_outcome(s, t) == kAlphaSkip) {
            // A skipped alpha-masked triangle: zero edges make every
            // triangle test's determinant 0, so no walk ever accepts it.
            for (int a = 0; a < 3; ++a) e1[a] = e2[a] = 0.0f;
        }
        uint32_t flags = 0;
        if ((t.metadata[0] & kGeometryFlagDoubleSided) != 0u || (t.metadata[1] & kMaterialFlagDoubleSided) != 0u) {
            flags |= kTriDoubleSided;
        }
        if ((t.metadata[3] & kInstanceFlagCastsShadow) != 0u) flags |= kTriCastsShadow;
        float instBits, primBits, flagBits;
        std::memcpy(&instBits, &t.identity[0], 4);
        std::memcpy(&primBits, &t.identity[1], 4);
        std::memcpy(&flagBits, &flags, 4);
        tpos[j].p0 = {v0[0], v0[1], v0[2], e1[0]};
        tpos[j].p1 = {e1[1], e1[2], e2[0], e2[1]};
        tpos[j].p2 = {e2[2], instBits, primBits, flagBits};
        TriAttr& a = tattr[j];
        a.n0 = {t.normals[0][0], t.normals[0][1], t.normals[0][2], 0.0f};
        a.n1 = {t.normals[1][0], t.normals[1][1], t.normals[1][2], 0.0f};
        a.n2 = {t.normals[2][0], t.normals[2][1], t.normals[2][2], 0.0f};
        a.ids = {t.identity[0], t.identity[1], t.identity[2], t.identity[3]};
        a.meta = {t.metadata[0], t.metadata[1], t.metadata[2], t.metadata[3]};
    }
    });
    // Every triangle of a leaf carries that leaf's exact box.
    std::vector<LeafBox> leafBox(n);
    for (uint32_t i = 0; i < m; ++i) {
        const BinaryNode& b = nodes[i];
        if (b.count == 0) continue;
        const LeafBox box = {{b.bmin[0], b.bmin[1], b.bmin[2], 0.0f}, {b.bmax[0], b.bmax[1], b.bmax[2], 0.0f}};
        for (uint32_t j = b.first; j < b.first + b.count; ++j) leafBox[j] = box;
    }

    auto bail = [&](rex_status st) {
        L.release();
        return st;
    };
    if (rex_status st = upload(wide, &L.dWide, L.bytes)) return bail(st);
    if (rex_status st = upload(tpos, &L.dTris, L.bytes)) return bail(st);
    if (rex_status st = upload(tattr, &L.dAttrs, L.bytes)) return bail(st);
    if (rex_status st = upload(wide4, &L.dWide4, L.bytes)) return bail(st);
    if (rex_status st = upload(leafBox, &L.dLeafBox, L.bytes)) return bail(st);
    double sceneScale = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double lo = nodes[0].bmin[k], hi = nodes[0].bmax[k];
        sceneScale = std::max(sceneScale, std::max(std::fabs(lo), std::fabs(hi)) + (hi - lo));
    }
    // Compressed nodes: of the traversal hierarchy when the scene has one
    // (same root box, leaves at leaf-order positions), else of this one.
    const bool separate = !s.tnodes.empty();
    std::vector<Wide4> travWide4;
    uint32_t travDepth = 0;
    if (separate) collapse_wide4(s.tnodes, travWide4, travDepth);
    const std::vector<Wide4>& qsrc = separate ? travWide4 : wide4;
    std::vector<Wide4Q> wide4q(qsrc.size());
    parallel_ranges(qsrc.size(), [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) wide4q[i] = quantize_wide4(qsrc[i], sceneScale);
    });
    if (rex_status st = upload(wide4q, &L.dWide4Q, L.bytes)) return bail(st);
    L.wide4qCount = static_cast<uint32_t>(wide4q.size());

    SceneView& v = L.view;
    v.wide = L.dWide;
    v.tris = L.dTris;
    v.attrs = L.dAttrs;
    v.wide4 = L.dWide4;
    v.wide4q = L.dWide4Q;
    v.leafBox = L.dLeafBox;
    v.wide4Depth = std::max(wide4Depth, travDepth);
    v.rootRef = refOf(0);
    v.rootRefQ = separate ? wide4_root_ref(s.tnodes) : wide4_root_ref(nodes);
    v.nodeCount = m;
    v.triCount = n;
    v.maxDepth = s.maxDepth;
    for (int k = 0; k < 3; ++k) {
        v.rootMin[k] = nodes[0].bmin[k];
        v.rootMax[k] = nodes[0].bmax[k];
    }
    return REX_SUCCESS;
}

void free_layout(Scene& s) {
    void* ptrs[] = {s.dWide, s.dTris, s.dAttrs, s.dWide4, s.dWide4Q, s.dLeafBox};
    for (void* p : ptrs) {
        if (p) (void)hipFree(p);
    }
    s.dWide = nullptr;
    s.dTris = nullptr;
    s.dAttrs = nullptr;
    s.dWide4 = nullptr;
    s.dWide4Q = nullptr;
    s.dLeafBox = nullptr;
    s.view = SceneView{};
}

void commit_layout(Scene& s, Layout& L) {
    free_layout(s);
    s.dWide = L.dWide;
    s.dTris = L.dTris;
    s.dAttrs = L.dAttrs;
    s.dWide4 = L.dWide4;
    s.dWide4Q = L.dWide4Q;
    s.dLeafBox = L.dLeafBox;
    s.wideCount = L.wideCount;
    s.wide4Count = L.wide4Count;
    s.wide4qCount = L.wide4qCount;
    s.deviceBytes = L.bytes;
    s.view = L.view;
    L = Layout{};
}

}  // namespace

rex_status scene_layout(Scene& s) {
    if (rex_status st = check_hip(hipGetDevice(&s.device), "hipGetDevice")) return st;
    Layout L;
    if (rex_status st = build_layout(s, s.nodes, s.leafTris, L)) return st;
    commit_layout(s, L);
    return REX_SUCCESS;
}

void scene_release(Scene& s) {
    release_hierarchy(s);
    void* traversal[] = {s.dTNodes, s.dTParents, s.dTDepth};
    for (void* p : traversal) {
        if (p) (void)hipFree(p);
    }
    s.dTNodes = nullptr;
    s.dTParents = nullptr;
    s.dTDepth = nullptr;
    s.ws.release();
    s.shade.release();
    free_layout(s);
    if (s.dInstances) (void)hipFree(s.dInstances);
    if (s.dInstanceMap) (void)hipFree(s.dInstanceMap);
    if (s.dObjectToWorld) (void)hipFree(s.dObjectToWorld);
    if (s.dPrevObjectToWorld) (void)hipFree(s.dPrevObjectToWorld);
    s.dObjectToWorld = s.dPrevObjectToWorld = nullptr;
    if (s.dAlphaSkip) (void)hipFree(s.dAlphaSkip);
    s.dInstances = nullptr;
    s.dInstanceMap = nullptr;
    s.dAlphaSkip = nullptr;
    s.instanceCount = s.instanceMapCount = 0;
}

}  // namespace rex
