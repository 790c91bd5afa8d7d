// Host binned-SAH BVH builder for the rex traversal core.
//
// Produces exactly the hierarchy of the reference's Rt::Cpu::Bvh<float> with
// BvhBuildMethod::BinnedSah (include/rt/cpu/Bvh.hpp:64-681): the same 16-bin
// SAH evaluated in double, the same strict-'<' tie rule (lowest axis, then
// lowest bin), the same stable partition and stable median fallback, the same
// leaf-size and depth rules, and nodes numbered in build pre-order as
// FlattenCpuSah keeps them (src/rt/software_gpu/SoftwareGpu.cpp:1073-1089).
//
// Parallelism: independent subtrees are built on a pool of std::threads and
// concatenated in pre-order afterwards, so the node numbering does not depend
// on the thread count or scheduling.  Inside one node the work is serial; the
// bin reductions fold elements in index order, which is what makes equal-value
// (signed zero) minima/maxima come out identical to the serial reference.
#include "bvh_build.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <vector>

namespace rex {
namespace {

constexpr uint32_t kBins = 16u;
constexpr uint32_t kMaxDepth = 128u;
// Subtrees at least this large are handed to the pool.
constexpr uint32_t kParallelGrain = 4096u;

struct Ref {
    float bmin[3];
    float bmax[3];
    float c[3];
    uint32_t tri;
    uint32_t primId;
};

inline float min_ref(float a, float b) { return a < b ? a : b; }  // Math.hpp:178-185
inline float max_ref(float a, float b) { return a > b ? a : b; }  // Math.hpp:187-194

struct Box {
    float mn[3];
    float mx[3];
    void clear() {
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::numeric_limits<float>::infinity();
            mx[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const float* bmn, const float* bmx) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = min_ref(mn[k], bmn[k]);
            mx[k] = max_ref(mx[k], bmx[k]);
        }
    }
    bool valid() const {
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(mn[k]) || !std::isfinite(mx[k]) || !(mn[k] <= mx[k])) return false;
        }
        return true;
    }
};

// Bvh.hpp:364-380
inline double surface_area(const Box& b) {
    if (!b.valid()) return 0.0;
    const double ex = static_cast<double>(b.mx[0]) - static_cast<double>(b.mn[0]);
    const double ey = static_cast<double>(b.mx[1]) - static_cast<double>(b.mn[1]);
    const double ez = static_cast<double>(b.mx[2]) - static_cast<double>(b.mn[2]);
    return 2.0 * ((ex * ey) + (ey * ez) + (ez * ex));
}

// Bvh.hpp:413-432
inline uint32_t bin_of(float c, float cmin, float cext) {
    const double normalized = (static_cast<double>(c) - static_cast<double>(cmin)) / static_cast<double>(cext);
    const double scaled = normalized * static_cast<double>(kBins);
    if (!(scaled > 0.0)) return 0u;
    if (scaled >= static_cast<double>(kBins)) return kBins - 1u;
    return static_cast<uint32_t>(scaled);
}

// A small fixed pool; tasks may enqueue further tasks.  wait_all() returns
// once every task (including those enqueued by tasks) has finished.
class Pool {
public:
    explicit Pool(unsigned threads) {
        for (unsigned i = 0; i < threads; ++i) workers_.emplace_back([this] { run(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    void submit(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(m_);
            q_.push_back(std::move(f));
            ++pending_;
        }
        cv_.notify_one();
    }
    void wait_all() {
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return pending_ == 0; });
    }

private:
    void run() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
                if (stop_ && q_.empty()) return;
                f = std::move(q_.back());
                q_.pop_back();
            }
            f();
            {
                std::lock_guard<std::mutex> lk(m_);
                if (--pending_ == 0) done_.notify_all();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::vector<std::function<void()>> q_;
    std::mutex m_;
    std::condition_variable cv_;
    std::condition_variable done_;
    size_t pending_ = 0;
    bool stop_ = false;
};

// A subtree under construction: its nodes use local pre-order indices; child
// subtrees built by other tasks are spliced in later.
struct Subtree {
    std::vector<BinaryNode> nodes;
    // For each node that has a deferred child: (local node, which child, subtree).
    struct Hole {
        uint32_t node;
        int right;
        Subtree* sub;
    };
    std::vector<Hole> holes;
    std::vector<std::unique_ptr<Subtree>> children;
    uint32_t maxDepth = 0;
};

struct Builder {
    Ref* refs;
    Ref* scratch;
    uint32_t maxLeaf;
    Pool* pool;
    std::atomic<int> error{0};

    // Builds range [begin, end) at `depth` into `st`, returning the local index
    // of the node created.
    uint32_t build(Subtree& st, uint32_t begin, uint32_t end, uint32_t depth, bool top) {
        if (depth > kMaxDepth) {
            error.store(-2);
            return 0;
        }
        const uint32_t index = static_cast<uint32_t>(st.nodes.size());
        st.nodes.push_back(BinaryNode{});
        if (depth > st.maxDepth) st.maxDepth = depth;

        Box bounds;
        bounds.clear();
        for (uint32_t i = begin; i < end; ++i) bounds.grow(refs[i].bmin, refs[i].bmax);
        const uint32_t count = end - begin;
        if (count <= maxLeaf) {
            BinaryNode& leaf = st.nodes[index];
            std::memcpy(leaf.bmin, bounds.mn, sizeof(leaf.bmin));
            std::memcpy(leaf.bmax, bounds.mx, sizeof(leaf.bmax));
            leaf.left = leaf.right = kInvalid;
            leaf.first = begin;
            leaf.count = count;
            return index;
        }

        Box cb;
        cb.clear();
        for (uint32_t i = begin; i < end; ++i) cb.grow(refs[i].c, refs[i].c);
        float cext[3];
        for (int k = 0; k < 3; ++k) cext[k] = cb.mx[k] - cb.mn[k];

        // FindSahSplit (Bvh.hpp:434-504)
        bool valid = false;
        uint32_t bestAxis = 0, bestBin = 0;
        double bestCost = std::numeric_limits<double>::infinity();
        for (uint32_t axis = 0; axis < 3; ++axis) {
            if (!(cext[axis] > 0.0f)) continue;
            Box bins[kBins];
            uint32_t counts[kBins] = {};
            for (auto& b : bins) b.clear();
            for (uint32_t i = begin; i < end; ++i) {
                const uint32_t bi = bin_of(refs[i].c[axis], cb.mn[axis], cext[axis]);
                ++counts[bi];
                bins[bi].grow(refs[i].bmin, refs[i].bmax);
            }
            // Prefix/suffix sweeps give the same boxes as the reference's
            // per-split re-accumulation: both fold bins in ascending order.
            Box left[kBins];
            uint32_t leftCount[kBins];
            Box acc;
            acc.clear();
            uint32_t n = 0;
            for (uint32_t k = 0; k < kBins; ++k) {
                acc.grow(bins[k].mn, bins[k].mx);
                n += counts[k];
                left[k] = acc;
                leftCount[k] = n;
            }
            for (uint32_t split = 0; split + 1 < kBins; ++split) {
                Box right;
                right.clear();
                uint32_t rc = 0;
                for (uint32_t k = split + 1; k < kBins; ++k) {
                    rc += counts[k];
                    right.grow(bins[k].mn, bins[k].mx);
                }
                const uint32_t lc = leftCount[split];
                if (lc == 0 || rc == 0) continue;
                const double cost = (surface_area(left[split]) * static_cast<double>(lc)) +
                                    (surface_area(right) * static_cast<double>(rc));
                if (!valid || cost < bestCost) {
                    valid = true;
                    bestAxis = axis;
                    bestBin = split;
                    bestCost = cost;
                }
            }
        }

        uint32_t middle = begin;
        if (valid) {
            // std::stable_partition (Bvh.hpp:539-560)
            uint32_t li = begin;
            uint32_t nl = 0;
            for (uint32_t i = begin; i < end; ++i) {
                if (bin_of(refs[i].c[bestAxis], cb.mn[bestAxis], cext[bestAxis]) <= bestBin) ++nl;
            }
            uint32_t ri = begin + nl;
            for (uint32_t i = begin; i < end; ++i) {
                if (bin_of(refs[i].c[bestAxis], cb.mn[bestAxis], cext[bestAxis]) <= bestBin) {
                    scratch[li++] = refs[i];
                } else {
                    scratch[ri++] = refs[i];
                }
            }
            std::memcpy(refs + begin, scratch + begin, static_cast<size_t>(count) * sizeof(Ref));
            middle = begin + nl;
        }
        if (middle == begin || middle == end) {
            // MedianPartition (Bvh.hpp:506-537) with LongestAxis (Math.hpp:197-209);
            // the key (centroid, primitive ID, index) is a total order, so any
            // sort equals std::stable_sort here.
            int axis = 0;
            if (cext[1] > cext[0]) axis = 1;
            if (cext[2] > cext[axis]) axis = 2;
            std::sort(refs + begin, refs + end, [axis](const Ref& a, const Ref& b) {
                if (a.c[axis] < b.c[axis]) return true;
                if (a.c[axis] > b.c[axis]) return false;
                if (a.primId < b.primId) return true;
                if (a.primId > b.primId) return false;
                return a.tri < b.tri;
            });
            middle = begin + count / 2u;
        }

        BinaryNode& node0 = st.nodes[index];
        std::memcpy(node0.bmin, bounds.mn, sizeof(node0.bmin));
        std::memcpy(node0.bmax, bounds.mx, sizeof(node0.bmax));
        node0.first = 0;
        node0.count = 0;

        // The left child always continues in this subtree (it follows the node
        // in pre-order); a large right child is built by another task and
        // spliced in after the left subtree.
        const uint32_t rightCount = end - middle;
        const uint32_t left = build(st, begin, middle, depth + 1, top);
        if (error.load()) return 0;
        st.nodes[index].left = left;
        if (pool != nullptr && rightCount >= kParallelGrain) {
            auto child = std::make_unique<Subtree>();
            Subtree* raw = child.get();
            st.children.push_back(std::move(child));
            st.holes.push_back({index, 1, raw});
            st.nodes[index].right = kInvalid;
            pool->submit([this, raw, middle, end, depth] {
                build(*raw, middle, end, depth + 1, false);
            });
        } else {
            const uint32_t right = build(st, middle, end, depth + 1, top);
            if (error.load()) return 0;
            st.nodes[index].right = right;
        }
        return index;
    }
};

// Splices a subtree and its deferred children into `out` in pre-order.
// Deferred right children are inserted right after the left subtree of their
// parent, which is exactly where a serial recursion would have put them.
void splice(const Subtree& st, std::vector<BinaryNode>& out, uint32_t& maxDepth) {
    maxDepth = std::max(maxDepth, st.maxDepth);
    // Map each local node to its global index.  Walk local nodes in local
    // pre-order; before emitting the node that follows a hole's left subtree,
    // the hole's subtree must be emitted.  Equivalent: recursive emission.
    std::vector<const Subtree::Hole*> holeOf(st.nodes.size(), nullptr);
    for (const auto& h : st.holes) holeOf[h.node] = &h;
    std::function<uint32_t(uint32_t)> emit = [&](uint32_t local) -> uint32_t {
        const uint32_t g = static_cast<uint32_t>(out.size());
        out.push_back(st.nodes[local]);
        const BinaryNode src = st.nodes[local];
        if (src.count != 0) return g;
        const uint32_t gl = emit(src.left);
        uint32_t gr;
        if (holeOf[local] != nullptr) {
            const Subtree* sub = holeOf[local]->sub;
            const uint32_t base = static_cast<uint32_t>(out.size());
            std::vector<BinaryNode> tmp;
            uint32_t md = 0;
            splice(*sub, tmp, md);
            maxDepth = std::max(maxDepth, md);
            for (auto& nd : tmp) {
                if (nd.count == 0) {
                    nd.left += base;
                    nd.right += base;
                }
                out.push_back(nd);
            }
            gr = base;
        } else {
            gr = emit(src.right);
        }
        out[g].left = gl;
        out[g].right = gr;
        return g;
    };
    if (!st.nodes.empty()) emit(0);
}

}  // namespace

int build_binned_sah(const float* positions, const uint32_t* primIds, uint32_t n, uint32_t maxLeaf,
                     unsigned threads, std::vector<BinaryNode>& nodes, std::vector<uint32_t>& order,
                     uint32_t& maxDepth) {
    nodes.clear();
    order.clear();
    maxDepth = 0;
    if (n == 0) return 0;
    if (maxLeaf == 0) return -1;
    std::vector<Ref> refs(n);
    std::vector<Ref> scratch(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = positions + 9u * i;
        Ref& r = refs[i];
        r.tri = i;
        r.primId = primIds[i];
        Box b;
        b.clear();
        for (int v = 0; v < 3; ++v) b.grow(p + 3 * v, p + 3 * v);
        std::memcpy(r.bmin, b.mn, sizeof(r.bmin));
        std::memcpy(r.bmax, b.mx, sizeof(r.bmax));
        // Triangle::Centroid (include/rt/cpu/Geometry.hpp:157-175)
        for (int k = 0; k < 3; ++k) {
            r.c[k] = static_cast<float>((static_cast<double>(p[k]) + static_cast<double>(p[3 + k]) +
                                         static_cast<double>(p[6 + k])) /
                                        3.0);
        }
    }
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::unique_ptr<Pool> pool;
    if (threads > 1 && n >= 2 * kParallelGrain) pool = std::make_unique<Pool>(threads);
    Builder b;
    b.refs = refs.data();
    b.scratch = scratch.data();
    b.maxLeaf = maxLeaf;
    b.pool = pool.get();
    Subtree root;
    b.build(root, 0, n, 1, true);
    if (pool) pool->wait_all();
    if (b.error.load() != 0) return b.error.load();
    nodes.reserve(2u * n - 1u);
    splice(root, nodes, maxDepth);
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = refs[i].tri;
    return 0;
}

}  // namespace rex
