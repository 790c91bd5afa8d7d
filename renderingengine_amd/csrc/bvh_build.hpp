// Host binned-SAH builder (see bvh_build.cpp).
#pragma once

#include <stdint.h>

#include <vector>

#include "rex_layout.hpp"

namespace rex {

// One node of the reference's binary hierarchy in build pre-order
// (include/rt/cpu/Bvh.hpp NodeView): leaf iff count != 0.
struct BinaryNode {
    float bmin[3];
    float bmax[3];
    uint32_t left;   // interior: left child node index
    uint32_t right;  // interior: right child node index
    uint32_t first;  // leaf: first leaf-order triangle
    uint32_t count;  // leaf: triangle count (0 for interior)
};

// positions: n * 9 floats (v0, v1, v2) in canonical order; primIds: the stable
// primitive ID of each triangle (the median fallback's tie key).  On success
// returns 0 and fills nodes (pre-order), order (leaf order -> canonical index)
// and maxDepth (root depth 1).  Returns -1 for a zero leaf size, -2 when the
// depth would exceed 128.
int build_binned_sah(const float* positions, const uint32_t* primIds, uint32_t n, uint32_t maxLeaf,
                     unsigned threads, std::vector<BinaryNode>& nodes, std::vector<uint32_t>& order,
                     uint32_t& maxDepth);

}  // namespace rex
