// Device-side scene layout of the rex traversal core (shared by the host
// converter and the HIP kernels).
//
// HBM layout (one scene):
//   WideNode  wide[interiorCount]  64 B each: the boxes of BOTH children of a
//                                  binary interior node plus their references,
//                                  so one 64-B fetch decides both children.
//   TriPos    tris[n]              48 B each, leaf order: v0, e1 = v1 - v0,
//                                  e2 = v2 - v0 (bit-identical to computing the
//                                  edges at test time), instance, primitive,
//                                  culling flags.
//   TriAttr   attrs[n]             80 B each, leaf order: shading normals,
//                                  identity and metadata, read only for the
//                                  winning hit.
//   Wide4     wide4[wide4Count] 128 B each (one cache line): a binary interior
//                                  node collapsed with its children, so its
//                                  up-to-4 children are the binary grandchildren
//                                  (a leaf child stays itself) in pre-order
//                                  slots; child boxes stored as SoA float4s.
//                                  Used by the near-first production walk.
//   Wide4Q    wide4q[wide4Count] 64 B each: the same nodes with 8-bit child
//                                  planes on a per-node grid (conservative).
//   LeafBox   leafBox[n]        32 B each, leaf order: the exact box of the
//                                  triangle's binary leaf (read only when the
//                                  compressed walk accepts a candidate hit).
// A child reference is either an interior wide-node index (bit 31 clear) or a
// leaf: bit 31 set, (count - 1) in bits 28..30, first triangle in bits 0..27.
#pragma once

#include <stdint.h>

namespace rex {

constexpr uint32_t kInvalid = 0xffffffffu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafFirstMask = 0x0fffffffu;
constexpr uint32_t kMaxLeafSize = 8u;
constexpr uint32_t kMaxTriangles = 0x10000000u;

// Triangle culling flags packed into TriPos::p2.w.
constexpr uint32_t kTriDoubleSided = 1u << 0;
constexpr uint32_t kTriCastsShadow = 1u << 1;

// include/contracts/SceneAbiV0.hpp:16-44
constexpr uint32_t kGeometryFlagAlphaMask = 1u << 1;
constexpr uint32_t kGeometryFlagDoubleSided = 1u << 2;
constexpr uint32_t kInstanceFlagCastsShadow = 1u << 1;
constexpr uint32_t kMaterialFlagDoubleSided = 1u << 0;
constexpr uint32_t kMaterialFlagAlphaMask = 1u << 1;

struct alignas(16) F4 {
    float x, y, z, w;
};
struct alignas(16) U4 {
    uint32_t x, y, z, w;
};

struct alignas(16) WideNode {
    F4 a;    // L.min.x, L.min.y, L.min.z, L.max.x
    F4 b;    // L.max.y, L.max.z, R.min.x, R.min.y
    F4 c;    // R.min.z, R.max.x, R.max.y, R.max.z
    U4 link; // left ref, right ref, binary index of left, parent wide index | isRight << 31
};
static_assert(sizeof(WideNode) == 64, "wide node is one 64-B line");

struct alignas(16) Wide4 {
    F4 lox, loy, loz;  // child k's box min (component k of each float4)
    F4 hix, hiy, hiz;  // child k's box max
    U4 ref;            // child references; kInvalid marks an empty slot
    U4 bin;            // binary node index of each child (kInvalid if empty)
};
static_assert(sizeof(Wide4) == 128, "4-wide node is one 128-B line");

// The same 4-wide node in 64 B (two per 128-B line): child planes as 8-bit
// offsets on a per-node grid, plane = origin + q * 2^e per axis, rounded
// outward with a margin so every child box contains its exact box.  Used by
// the compressed wide walk.
struct alignas(16) Wide4Q {
    F4 origin;   // x, y, z: grid origin; w (bits): biased float exponents of the
                 // per-axis steps, x | y << 8 | z << 16
    U4 planeA;   // lo.x, lo.y, lo.z, hi.x: byte k of each word is child k's plane
    U4 planeB;   // hi.y, hi.z, unused, unused
    U4 ref;      // child references (as Wide4::ref)
};
static_assert(sizeof(Wide4Q) == 64, "compressed 4-wide node is 64 B");

struct alignas(16) TriPos {
    F4 p0;   // v0.x, v0.y, v0.z, e1.x
    F4 p1;   // e1.y, e1.z, e2.x, e2.y
    F4 p2;   // e2.z, instance (bits), primitive (bits), flags (bits)
};
static_assert(sizeof(TriPos) == 48, "triangle position record is 48 B");

struct alignas(16) TriAttr {
    F4 n0, n1, n2;  // shading normals (w unused)
    U4 ids;         // instance, primitive, geometry, material
    U4 meta;        // geometry flags, material flags, traversal ID, instance flags
};
static_assert(sizeof(TriAttr) == 80, "triangle attribute record is 80 B");

// Exact box of the binary leaf that holds a triangle, one per leaf-order
// triangle (every triangle of a leaf carries its leaf's box).  The compressed
// walk's child boxes contain the exact ones, so it can reach a triangle whose
// exact leaf box the reference's slab test rejects by rounding (a ray that
// ends within an ulp of the triangle, a ray that grazes the box); a candidate
// hit of that walk is therefore accepted only when its exact leaf box passes
// the reference's tMin / tMax test as well.
struct alignas(16) LeafBox {
    F4 lo;  // x, y, z, unused
    F4 hi;
};
static_assert(sizeof(LeafBox) == 32, "leaf box is 32 B");

// Scene constants handed to every traversal kernel by value.
struct SceneView {
    const WideNode* wide;
    const TriPos* tris;
    const TriAttr* attrs;
    const Wide4* wide4;
    const Wide4Q* wide4q;    // compressed 4-wide nodes of the traversal hierarchy
    const LeafBox* leafBox;  // per leaf-order triangle: its leaf's exact box
    uint32_t rootRef;        // child reference of the root (interior 0 or a leaf)
    uint32_t rootRefQ;       // the same for the compressed walk's hierarchy
    uint32_t nodeCount;      // binary node count (reference node count)
    uint32_t triCount;
    uint32_t maxDepth;
    uint32_t wide4Depth;     // levels of 4-wide nodes of either hierarchy (stack bound: 3 per level)
    float rootMin[3];
    float rootMax[3];
};

// Root of one scene resident in a shared arena (batched jobs): the wide
// walk's root reference (relocated into the arena) and the root box.
struct alignas(16) SceneRoot {
    uint32_t rootRef;
    float rootMin[3];
    float rootMax[3];
    uint32_t pad;
};

#if defined(__HIPCC__)
#define REX_HD __host__ __device__
#else
#define REX_HD
#endif

REX_HD inline bool is_leaf_ref(uint32_t r) { return (r & kLeafBit) != 0u; }

// std::midpoint for floats (the reference's triangle centroids): (a + b) / 2
// where neither operand can overflow the sum, otherwise halves first.
REX_HD inline float midpoint_f(float a, float b) {
    const float lo = 1.17549435e-38f * 2.0f, hi = 3.40282347e+38f / 2.0f;
    const float aa = a < 0.0f ? -a : a, ab = b < 0.0f ? -b : b;
    if (aa <= hi && ab <= hi) return (a + b) / 2.0f;
    if (aa < lo) return a + b / 2.0f;
    if (ab < lo) return a / 2.0f + b;
    return a / 2.0f + b / 2.0f;
}

}  // namespace rex
