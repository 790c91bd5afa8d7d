// Arena relocation for batched jobs: a scene's compressed 4-wide nodes are copied into
// the shared arena with interior references moved by the arena's node base
// and leaf references' first triangle moved by its triangle base.
#include <hip/hip_runtime.h>

#include "rex_internal.hpp"

namespace rex {

__global__ void k_relocate_wide4q(Wide4Q* __restrict__ dst, const Wide4Q* __restrict__ src, uint32_t n,
                                  uint32_t nodeBase, uint32_t triBase) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Wide4Q w = src[i];
    uint32_t* r = &w.ref.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = r[k];
        if (v == kInvalid) continue;
        if (is_leaf_ref(v)) r[k] = (v & ~kLeafFirstMask) | ((v & kLeafFirstMask) + triBase);
        else r[k] = v + nodeBase;
    }
    dst[i] = w;
}

__global__ void k_iota_u32(uint32_t* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = i;
}

hipError_t launch_iota_u32(uint32_t* dst, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota_u32, dim3((n + 255) / 256), dim3(256), 0, stream, dst, n);
    return hipGetLastError();
}

hipError_t launch_relocate_wide4q(Wide4Q* dst, const Wide4Q* src, uint32_t n, uint32_t nodeBase, uint32_t triBase,
                                  hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_relocate_wide4q, dim3((n + 255) / 256), dim3(256), 0, stream, dst, src, n, nodeBase,
                       triBase);
    return hipGetLastError();
}

}  // namespace rex
