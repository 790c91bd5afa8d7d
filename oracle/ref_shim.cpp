// Thin extern "C" wrapper around the reference's own CPU BVH build, flatten and
// fp32 CPU traversal mirror, compiled from the reference sources where they
// lie (include/rt/cpu/Bvh.hpp, src/rt/software_gpu/SoftwareGpu.cpp) into
// oracle/_ref/libref.so by oracle/Makefile.  Test infrastructure only: it is
// used to pin the oracle (and, through it, the product) and never ships.
#include "rt/cpu/Bvh.hpp"
#include "rt/software_gpu/SoftwareGpu.hpp"

#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

namespace Cpu = RenderingEngine::Rt::Cpu;
namespace Sw = RenderingEngine::Rt::SoftwareGpu;

namespace
{
    std::unique_ptr<Sw::FlatBvh> gBvh;

    Sw::SoftwarePrimitiveRecord MakePrimitive(const float* p, std::uint32_t id)
    {
        Sw::SoftwarePrimitiveRecord r{};
        r.v0 = {p[0], p[1], p[2], 1.0f};
        r.v1 = {p[3], p[4], p[5], 1.0f};
        r.v2 = {p[6], p[7], p[8], 1.0f};
        r.identity = {id, 0u, 0u, 0u};
        return r;
    }
}

extern "C" {

// positions: n * 9 floats (v0, v1, v2).  Triangle i gets primitive ID i, as
// BuildCanonicalTraversalScene assigns traversal IDs
// (src/rt/gpu/CanonicalTraversalScene.cpp:205-208).  Output: nodes as 12 words
// each (SoftwareNodeRecord), primOrder[n] (leaf-order source indices).
int ref_build_flatten(const float* positions, std::uint32_t n, std::uint32_t leafSize,
                      std::uint32_t* nodesOut, std::uint32_t* nodeCount,
                      std::uint32_t* primOrder, std::uint32_t* maxDepth)
{
    try
    {
        std::vector<Cpu::Triangle<float>> tris;
        std::vector<Sw::SoftwarePrimitiveRecord> prims;
        tris.reserve(n);
        prims.reserve(n);
        for (std::uint32_t i = 0; i < n; ++i)
        {
            const float* p = positions + 9u * i;
            tris.emplace_back(Cpu::Vec3<float>{p[0], p[1], p[2]},
                              Cpu::Vec3<float>{p[3], p[4], p[5]},
                              Cpu::Vec3<float>{p[6], p[7], p[8]}, i);
            prims.push_back(MakePrimitive(p, i));
        }
        Cpu::Bvh<float> bvh(tris, Cpu::BvhBuildMethod::BinnedSah, leafSize);
        Sw::FlatBuildResult flat = Sw::FlattenCanonicalL3BinnedSah(bvh, prims);
        if (!flat.Succeeded())
        {
            return -static_cast<int>(flat.status) - 10;
        }
        *nodeCount = static_cast<std::uint32_t>(flat.bvh.nodes.size());
        std::memcpy(nodesOut, flat.bvh.nodes.data(), flat.bvh.nodes.size() * 48u);
        const auto order = bvh.PrimitiveOrder();
        for (std::uint32_t i = 0; i < n; ++i)
        {
            primOrder[i] = static_cast<std::uint32_t>(order[i]);
        }
        *maxDepth = flat.bvh.maximumDepth;
        gBvh = std::make_unique<Sw::FlatBvh>(std::move(flat.bvh));
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

// Traces rays (o.xyz, tMin, d.xyz, tMax) through the BVH of the last
// ref_build_flatten call with Rt::SoftwareGpu::Trace (mode 0 closest, 1 any).
// out: per ray (t, u, v, frontFace) floats and primitive ID (the traversal ID,
// kInvalidIndex on a miss) and the trace status.
int ref_trace(const float* rays, std::uint32_t count, std::uint32_t mode,
              float* tuvf, std::uint32_t* primitive, std::uint32_t* status,
              std::uint64_t* nodeTests, std::uint64_t* triangleTests)
{
    if (!gBvh)
    {
        return -1;
    }
    for (std::uint32_t i = 0; i < count; ++i)
    {
        const float* r = rays + 8u * i;
        const Sw::SoftwareRayRecord ray = Sw::MakeRay(
            {r[0], r[1], r[2]}, {r[4], r[5], r[6]}, r[3], r[7], i,
            mode == 0u ? Sw::QueryMode::Closest : Sw::QueryMode::Any);
        const Sw::TraceResult res = Sw::Trace(
            *gBvh, ray, mode == 0u ? Sw::QueryMode::Closest : Sw::QueryMode::Any, 64u);
        tuvf[4u * i + 0u] = res.hit.tBary.x;
        tuvf[4u * i + 1u] = res.hit.tBary.y;
        tuvf[4u * i + 2u] = res.hit.tBary.z;
        tuvf[4u * i + 3u] = res.hit.tBary.w;
        primitive[i] = res.hit.identity.x;
        status[i] = static_cast<std::uint32_t>(res.status);
        if (nodeTests)
        {
            nodeTests[i] = res.counters.nodeTests;
        }
        if (triangleTests)
        {
            triangleTests[i] = res.counters.triangleTests;
        }
    }
    return 0;
}
}

extern "C" {

// The reference's RefitBottomUp (src/rt/software_gpu/SoftwareGpu.cpp:1347)
// on 48-B node records (in place) and leaf-order triangle positions
// (n * 9 floats; primitive j gets ID j).  0, or -(BuildStatus) - 10.
int ref_refit(std::uint32_t* nodes, std::uint32_t nodeCount, const float* leafPositions, std::uint32_t n)
{
    try
    {
        Sw::FlatBvh bvh;
        bvh.nodes.resize(nodeCount);
        std::memcpy(bvh.nodes.data(), nodes, static_cast<std::size_t>(nodeCount) * 48u);
        bvh.primitives.reserve(n);
        for (std::uint32_t j = 0; j < n; ++j) bvh.primitives.push_back(MakePrimitive(leafPositions + 9u * j, j));
        const Sw::BuildStatus st = Sw::RefitBottomUp(bvh);
        if (st != Sw::BuildStatus::Success) return -static_cast<int>(st) - 10;
        std::memcpy(nodes, bvh.nodes.data(), static_cast<std::size_t>(nodeCount) * 48u);
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

}  // extern "C"

extern "C" {

// The reference's BuildKarrasLbvh (src/rt/software_gpu/SoftwareGpu.cpp:1265)
// on triangle positions (n * 9 floats) with the given stable IDs.  Output as
// ref_build_flatten; primOrder = the Morton order's source indices.
int ref_build_lbvh(const float* positions, const std::uint32_t* ids, std::uint32_t n, std::uint32_t* nodesOut,
                   std::uint32_t* nodeCount, std::uint32_t* primOrder, std::uint32_t* maxDepth)
{
    try
    {
        std::vector<Sw::SoftwarePrimitiveRecord> prims;
        prims.reserve(n);
        for (std::uint32_t i = 0; i < n; ++i) prims.push_back(MakePrimitive(positions + 9u * i, ids[i]));
        Sw::FlatBuildResult flat = Sw::BuildKarrasLbvh(prims, Sw::kMaximumTraversalStack);
        if (!flat.Succeeded()) return -static_cast<int>(flat.status) - 10;
        *nodeCount = static_cast<std::uint32_t>(flat.bvh.nodes.size());
        std::memcpy(nodesOut, flat.bvh.nodes.data(), flat.bvh.nodes.size() * 48u);
        for (std::uint32_t i = 0; i < n; ++i) primOrder[i] = flat.bvh.mortonOrder[i].sourceIndex;
        *maxDepth = flat.bvh.maximumDepth;
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

}  // extern "C"
