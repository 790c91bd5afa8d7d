// Reconstruction boundary kernels over the ABI-v2 records: MotionVectorsCS
// and ComposeCS.  One thread per pixel, linear pixel index y * width + x
// (L8LinearIndex), 256-thread blocks; both are streaming passes bounded by
// HBM bandwidth (a 208-B G-buffer record, a 32-B motion input and two 64-B
// transforms read per pixel for motion; at most 208 + 64 + 16 + 16 B read
// and 16 + 16 B written per pixel for compose).
//
// Restates (paths relative to the RenderingEngine tree):
//   motion vectors  reconstruction/shaders/MotionVectorsCS.hlsl:28-90
//   compose         reconstruction/shaders/ComposeCS.hlsl:26-126,
//                   ReconstructionCommon.hlsli:75-81 (raw signal), :133-137 (safe normalize)
//
// Arithmetic is written operand for operand like the C restatement in the
// oracle (no contraction; a 4-term dot is ((x*x' + y*y') + z*z') + w*w';
// 1 / sqrt for HLSL's rsqrt), so results are bit-identical to it.
#include <hip/hip_runtime.h>

#include "motion.hpp"
#include "restir.hpp"
#include "rex_internal.hpp"

namespace rex {
namespace {

struct Mat4 {
    float4 r[4];
};

__device__ __forceinline__ float dot4(const float4 a, const float4 b) {
    return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

__device__ __forceinline__ float4 mul(const Mat4& m, const float4 v) {
    return make_float4(dot4(m.r[0], v), dot4(m.r[1], v), dot4(m.r[2], v), dot4(m.r[3], v));
}

__device__ __forceinline__ bool finite4(const float4 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z) && __builtin_isfinite(v.w);
}

__device__ __forceinline__ float saturate(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

struct MotionArgs {
    Mat4 vpCur, vpPrev, wvPrev;
    float jcx, jcy, jpx, jpy;
    uint32_t pixels, curCount, prevCount;
};

__global__ __launch_bounds__(256) void k_motion_vectors(GBufferV2* __restrict__ gb,
                                                        const MotionInputV2* __restrict__ inputs,
                                                        const float4* __restrict__ cur, const float4* __restrict__ prev,
                                                        MotionArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.pixels) return;
    MotionV2& out = gb[i].motion;
    const U4 pid = gb[i].primary.identity;
    // L8ClearMotion keeps the identity's first three words.
    auto clear = [&]() {
        out.motionDepth = {0.0f, 0.0f, 0.0f, 0.0f};
        out.uv = {0.0f, 0.0f, 0.0f, 0.0f};
        out.identity.w = 0u;
    };
    if ((pid.w & kSurfaceValid) == 0u) {
        clear();
        return;
    }
    SurfaceMotion m;
    if (!surface_motion(inputs[i], cur, a.curCount, prev, a.prevCount, a.vpCur.r, a.vpPrev.r, a.wvPrev.r, a.jcx, a.jcy,
                        a.jpx, a.jpy, m)) {
        clear();
        return;
    }
    const bool valid = m.valid;
    out.motionDepth = {valid ? m.pu - m.cu : 0.0f, valid ? m.pv - m.cv : 0.0f, valid ? m.depth : 0.0f, 0.0f};
    out.uv = {m.cu, m.cv, valid ? m.pu : m.cu, valid ? m.pv : m.cv};
    out.identity = {pid.x, pid.y, pid.z, valid ? 1u : 0u};
}

struct ComposeArgs {
    const GBufferV2* gb;
    const SignalV2* raw;
    const float4* terminal;
    const rex_material_v0* materials;
    float4* film;
    float4* output;
    uint32_t pixels, mode, filmCount, materialCount;
};

__global__ __launch_bounds__(256) void k_compose(ComposeArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.pixels) return;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    if (a.mode == REX_COMPOSE_RAW || a.mode == REX_COMPOSE_PROGRESSIVE_MEAN) {
        const SignalV2 r = a.raw[i];
        const float4 t = a.terminal[i];
        // L8LoadRawSignal, then raw.diffuse + raw.specular + terminal.
        ox = (r.dd.x + r.id.x) + (r.ds.x + r.is.x) + t.x;
        oy = (r.dd.y + r.id.y) + (r.ds.y + r.is.y) + t.y;
        oz = (r.dd.z + r.id.z) + (r.ds.z + r.is.z) + t.z;
        if (a.mode == REX_COMPOSE_PROGRESSIVE_MEAN) {
            const float4 f = a.filmCount == 0u ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : a.film[i];
            const float k = (float)(a.filmCount + 1u);
            ox = f.x + (ox - f.x) / k;
            oy = f.y + (oy - f.y) / k;
            oz = f.z + (oz - f.z) / k;
            a.film[i] = make_float4(ox, oy, oz, k);
        }
    } else {
        const GBufferV2& g = a.gb[i];
        const bool valid = (g.primary.identity.w & kSurfaceValid) != 0u;
        const uint32_t mat = g.primary.identity.x;
        if (a.mode == REX_COMPOSE_MOTION) {
            ox = 0.5f + g.motion.motionDepth.x * 16.0f;
            oy = 0.5f + g.motion.motionDepth.y * 16.0f;
            oz = 0.5f;
        } else if (a.mode == REX_COMPOSE_BASE_COLOR) {
            if (valid && mat < a.materialCount) {
                const float* b = a.materials[mat].baseColorFactor;
                ox = saturate(b[0]);
                oy = saturate(b[1]);
                oz = saturate(b[2]);
            }
        } else if (a.mode == REX_COMPOSE_NORMAL) {
            const F4 n = g.primary.snMetal;
            const float l2 = n.x * n.x + n.y * n.y + n.z * n.z;
            float nx = 0.0f, ny = 0.0f, nz = 1.0f;
            if (l2 > 1.0e-20f) {
                const float s = 1.0f / sqrtf(l2);
                nx = n.x * s;
                ny = n.y * s;
                nz = n.z * s;
            }
            ox = nx * 0.5f + 0.5f;
            oy = ny * 0.5f + 0.5f;
            oz = nz * 0.5f + 0.5f;
        } else if (a.mode == REX_COMPOSE_ROUGHNESS) {
            ox = oy = oz = saturate(g.primary.gnRough.w);
        } else if (a.mode == REX_COMPOSE_METALLIC) {
            ox = oy = oz = saturate(g.primary.snMetal.w);
        } else if (a.mode == REX_COMPOSE_EMISSIVE) {
            if (valid && mat < a.materialCount) {
                const float* e = a.materials[mat].emissiveFactorStrength;
                ox = fmaxf(e[0] * e[3], 0.0f);
                oy = fmaxf(e[1] * e[3], 0.0f);
                oz = fmaxf(e[2] * e[3], 0.0f);
            }
        }
    }
    a.output[i] = make_float4(ox, oy, oz, 1.0f);
}

Mat4 load_rows(const float m[16]) {
    Mat4 r;
    for (int k = 0; k < 4; ++k) r.r[k] = make_float4(m[4 * k], m[4 * k + 1], m[4 * k + 2], m[4 * k + 3]);
    return r;
}

}  // namespace

hipError_t launch_motion_vectors(GBufferV2* gbuffer, const MotionInputV2* inputs, const float4* current,
                                 const float4* previous, const rex_motion_constants_v2& c, hipStream_t stream) {
    MotionArgs a;
    a.vpCur = load_rows(c.currentViewProjection);
    a.vpPrev = load_rows(c.previousViewProjection);
    a.wvPrev = load_rows(c.previousWorldToView);
    a.jcx = c.currentJitterUv[0];
    a.jcy = c.currentJitterUv[1];
    a.jpx = c.previousJitterUv[0];
    a.jpy = c.previousJitterUv[1];
    a.pixels = c.width * c.height;
    a.curCount = c.currentTransformCount;
    a.prevCount = c.previousTransformCount;
    hipLaunchKernelGGL(k_motion_vectors, dim3((a.pixels + 255u) / 256u), dim3(256), 0, stream, gbuffer, inputs, current,
                       previous, a);
    return hipGetLastError();
}

hipError_t launch_compose(const rex_compose_desc_v2& d, hipStream_t stream) {
    ComposeArgs a;
    a.gb = static_cast<const GBufferV2*>(d.gbuffer);
    a.raw = static_cast<const SignalV2*>(d.rawSignal);
    a.terminal = static_cast<const float4*>(d.terminal);
    a.materials = static_cast<const rex_material_v0*>(d.materials);
    a.film = static_cast<float4*>(d.film);
    a.output = static_cast<float4*>(d.output);
    a.pixels = d.width * d.height;
    a.mode = d.outputMode;
    a.filmCount = d.filmSampleCount;
    a.materialCount = d.materials != nullptr ? d.materialCount : 0u;
    hipLaunchKernelGGL(k_compose, dim3((a.pixels + 255u) / 256u), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace rex
