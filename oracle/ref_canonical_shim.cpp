// Thin extern "C" wrapper around the reference's own canonical scene
// expansion (src/rt/gpu/CanonicalTraversalScene.cpp), compiled from the
// reference sources where they lie into oracle/_ref/libref.so by
// oracle/Makefile.  Test infrastructure only: it pins the oracle's and the
// product's expansion and never ships.
#include "rt/gpu/CanonicalTraversalScene.hpp"

#include <cstdint>
#include <cstring>
#include <span>

namespace Gpu = RenderingEngine::Rt::Gpu;
namespace Abi0 = RenderingEngine::Contracts::AbiV0;
namespace Scene = RenderingEngine::Scene;

extern "C" {

// Raw ABI-v0 records in; canonical triangles (44 words each) out.  Returns 0
// and the count, or 1 with the reference's error.
int ref_canonical_expand(const void* constants, const void* vertices, std::uint32_t vertexCount,
                         const std::uint32_t* indices, std::uint32_t indexCount, const void* geometries,
                         std::uint32_t geometryCount, const void* instances, std::uint32_t instanceCount,
                         const void* materials, std::uint32_t materialCount, std::uint64_t fingerprint,
                         std::uint32_t generation, std::uint32_t* out, std::uint32_t capacity,
                         std::uint32_t* count)
{
    Scene::CanonicalSceneView view;
    view.stableId = "oracle";
    view.fingerprint = fingerprint;
    view.generation = generation;
    view.constants = static_cast<const Abi0::GpuSceneConstantsV0*>(constants);
    view.vertices = {static_cast<const Abi0::GpuVertexV0*>(vertices), vertexCount};
    view.indices = {indices, indexCount};
    view.geometries = {static_cast<const Abi0::GpuGeometryV0*>(geometries), geometryCount};
    view.instances = {static_cast<const Abi0::GpuInstanceV0*>(instances), instanceCount};
    view.materials = {static_cast<const Abi0::GpuMaterialV0*>(materials), materialCount};
    const Gpu::CanonicalTraversalSceneBuild build = Gpu::BuildCanonicalTraversalScene(view);
    if (!build) return 1;
    const auto& tris = build.scene.triangles;
    *count = static_cast<std::uint32_t>(tris.size());
    if (tris.size() > capacity) return 2;
    std::memcpy(out, tris.data(), tris.size() * sizeof(Gpu::CanonicalTraversalTriangle));
    return 0;
}

}  // extern "C"
