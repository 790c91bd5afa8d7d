/* CPU oracle (test infrastructure only; the product never links it) for the
 * reconstruction boundary over ABI-v2 records
 * (include/contracts/ReconstructionAbiV2.hpp:45-116):
 *
 *   ora_motion_vectors_v2   MotionVectorsCS (reconstruction/shaders/MotionVectorsCS.hlsl:28-90)
 *   ora_compose_v2          ComposeCS (reconstruction/shaders/ComposeCS.hlsl:26-126),
 *                           raw signal ReconstructionCommon.hlsli:75-81,
 *                           safe normalize :133-137 (rsqrt restated as 1 / sqrt)
 *   ora_motion_inputs_v2    the motion input of a published surface
 *                           (resources/shaders/integrators/pbr_reconstruction_export_v2.hlsli:90-129),
 *                           with the shader's linear scan over the instances
 *
 * Parity pins (tests/test_reconstruction.py): motion vectors against the
 * reference's own CPU ComputeMotionVector (reconstruction/src/MotionVectors.cpp,
 * compiled into oracle/_ref) and the known answers of its test
 * (reconstruction/tests/ReconstructionTests.cpp:80-127); compose against
 * closed forms (sums, the arithmetic mean of N frames, saturate, unit
 * normals); motion inputs against matrix inverses of the instances.
 *
 * fp32, no contraction: a 4-term dot is ((x*x' + y*y') + z*z') + w*w'. */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float v[4]; } f4;

typedef struct {                 /* GpuGBufferRecordV2, 208 B as 52 words */
    float primary[20];           /* position/depth, gN/rough, sN/metal, diffuse, specular */
    uint32_t identity[4];
    float motionDepth[4];
    float uv[4];
    uint32_t motionIdentity[4];
    float signal[16];
} gbuffer_v2;

typedef struct {                 /* PbrMotionInputV2, 32 B */
    float p[3];
    uint32_t current, previous, reserved[3];
} motion_input_v2;

static float dot4(const float* row, const float* v) {
    return row[0] * v[0] + row[1] * v[1] + row[2] * v[2] + row[3] * v[3];
}

static void mul(const float* m, const float* v, float* out) {
    float t[4];
    for (int r = 0; r < 4; ++r) t[r] = dot4(m + 4 * r, v);
    memcpy(out, t, sizeof t);
}

static int finite4(const float* v) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]) && isfinite(v[3]); }

static float saturate(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

static void clear_motion(gbuffer_v2* g) {
    memset(g->motionDepth, 0, sizeof g->motionDepth);
    memset(g->uv, 0, sizeof g->uv);
    g->motionIdentity[3] = 0u;
}

void ora_motion_vectors_v2(gbuffer_v2* gb, const motion_input_v2* inputs, const float* cur, const float* prev,
                           const float* vpCur, const float* vpPrev, const float* wvPrev, const float* jitterCur,
                           const float* jitterPrev, uint32_t pixels, uint32_t curCount, uint32_t prevCount) {
    for (uint32_t i = 0; i < pixels; ++i) {
        gbuffer_v2* g = &gb[i];
        if ((g->identity[3] & 1u) == 0u) { clear_motion(g); continue; }
        const motion_input_v2 in = inputs[i];
        if (in.current >= curCount || in.previous >= prevCount || !isfinite(in.p[0]) || !isfinite(in.p[1]) ||
            !isfinite(in.p[2])) {
            clear_motion(g);
            continue;
        }
        const float obj[4] = {in.p[0], in.p[1], in.p[2], 1.0f};
        float cw[4], pw[4], cc[4], pc[4], pv[4];
        mul(cur + 16u * in.current, obj, cw);
        mul(prev + 16u * in.previous, obj, pw);
        mul(vpCur, cw, cc);
        mul(vpPrev, pw, pc);
        mul(wvPrev, pw, pv);
        if (!(cc[3] > 1.0e-8f && pc[3] > 1.0e-8f && finite4(cw) && finite4(pw) && finite4(cc) && finite4(pc) &&
              finite4(pv))) {
            clear_motion(g);
            continue;
        }
        const float cnx = cc[0] / cc[3], cny = cc[1] / cc[3];
        const float pnx = pc[0] / pc[3], pny = pc[1] / pc[3];
        const float cu = cnx * 0.5f + 0.5f + jitterCur[0], cv = 0.5f - cny * 0.5f + jitterCur[1];
        const float pu = pnx * 0.5f + 0.5f + jitterPrev[0], pvv = 0.5f - pny * 0.5f + jitterPrev[1];
        const float depth = -pv[2];
        const int valid = isfinite(cu) && isfinite(cv) && isfinite(pu) && isfinite(pvv) && isfinite(depth) &&
                          depth >= 0.0f;
        g->motionDepth[0] = valid ? pu - cu : 0.0f;
        g->motionDepth[1] = valid ? pvv - cv : 0.0f;
        g->motionDepth[2] = valid ? depth : 0.0f;
        g->motionDepth[3] = 0.0f;
        g->uv[0] = cu;
        g->uv[1] = cv;
        g->uv[2] = valid ? pu : cu;
        g->uv[3] = valid ? pvv : cv;
        for (int k = 0; k < 3; ++k) g->motionIdentity[k] = g->identity[k];
        g->motionIdentity[3] = valid ? 1u : 0u;
    }
}

/* materials: ABI-v0 material records (32 words: base color at 0, emissive
 * factor / strength at 4).  film may be NULL unless mode 15. */
void ora_compose_v2(const gbuffer_v2* gb, const float* raw, const float* terminal, const float* materials,
                    uint32_t materialCount, float* film, float* output, uint32_t pixels, uint32_t mode,
                    uint32_t filmCount) {
    for (uint32_t i = 0; i < pixels; ++i) {
        float o[3] = {0.0f, 0.0f, 0.0f};
        if (mode == 0u || mode == 15u) {
            const float* r = raw + 16u * i;
            const float* t = terminal + 4u * i;
            for (int c = 0; c < 3; ++c) o[c] = (r[c] + r[8 + c]) + (r[4 + c] + r[12 + c]) + t[c];
            if (mode == 15u) {
                float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                if (filmCount != 0u) memcpy(f, film + 4u * i, sizeof f);
                const float k = (float)(filmCount + 1u);
                for (int c = 0; c < 3; ++c) o[c] = f[c] + (o[c] - f[c]) / k;
                film[4u * i] = o[0];
                film[4u * i + 1] = o[1];
                film[4u * i + 2] = o[2];
                film[4u * i + 3] = k;
            }
        } else {
            const gbuffer_v2* g = &gb[i];
            const int valid = (g->identity[3] & 1u) != 0u;
            const uint32_t mat = g->identity[0];
            if (mode == 4u) {
                o[0] = 0.5f + g->motionDepth[0] * 16.0f;
                o[1] = 0.5f + g->motionDepth[1] * 16.0f;
                o[2] = 0.5f;
            } else if (mode == 10u) {
                if (valid && mat < materialCount) {
                    for (int c = 0; c < 3; ++c) o[c] = saturate(materials[32u * mat + c]);
                }
            } else if (mode == 11u) {
                const float* n = g->primary + 8;
                const float l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
                float m[3] = {0.0f, 0.0f, 1.0f};
                if (l2 > 1.0e-20f) {
                    const float s = 1.0f / sqrtf(l2);
                    for (int c = 0; c < 3; ++c) m[c] = n[c] * s;
                }
                for (int c = 0; c < 3; ++c) o[c] = m[c] * 0.5f + 0.5f;
            } else if (mode == 12u) {
                o[0] = o[1] = o[2] = saturate(g->primary[7]);
            } else if (mode == 13u) {
                o[0] = o[1] = o[2] = saturate(g->primary[11]);
            } else if (mode == 14u) {
                if (valid && mat < materialCount) {
                    const float* e = materials + 32u * mat + 4;
                    for (int c = 0; c < 3; ++c) o[c] = fmaxf(e[c] * e[3], 0.0f);
                }
            }
        }
        output[4u * i] = o[0];
        output[4u * i + 1] = o[1];
        output[4u * i + 2] = o[2];
        output[4u * i + 3] = 1.0f;
    }
}

/* instances: ABI-v0 instance records (56 words: objectToWorld rows at 0,
 * worldToObject rows at 16, previous at 32, metadata at 48 with the stable
 * ID at 50), or NULL for a scene without an instance table. */
void ora_motion_inputs_v2(const gbuffer_v2* gb, const uint32_t* instances, uint32_t instanceCount,
                          motion_input_v2* out, uint32_t pixels) {
    for (uint32_t i = 0; i < pixels; ++i) {
        motion_input_v2 m;
        memset(&m, 0, sizeof m);
        const gbuffer_v2* g = &gb[i];
        if ((g->identity[3] & 1u) != 0u) {
            if (instances == NULL) {
                memcpy(m.p, g->primary, 3 * sizeof(float));
            } else {
                uint32_t t = 0xffffffffu;
                for (uint32_t c = 0; c < instanceCount; ++c) {
                    if (instances[56u * c + 50u] == g->identity[1]) { t = c; break; }
                }
                if (t != 0xffffffffu) {
                    const float* w2o = (const float*)(instances + 56u * t + 16u);
                    const float world[4] = {g->primary[0], g->primary[1], g->primary[2], 1.0f};
                    for (int r = 0; r < 3; ++r) m.p[r] = dot4(w2o + 4 * r, world);
                    m.current = m.previous = t;
                } else {
                    m.current = m.previous = instanceCount;
                }
            }
        }
        out[i] = m;
    }
}
