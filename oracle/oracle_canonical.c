/* CPU oracle (test infrastructure only; only tests/, the smoke test and the
 * benchmark's CPU baseline use it): ABI-v0 canonical scene ingestion.
 *
 *   ora_canonical_expand   BuildCanonicalTraversalScene
 *                          (src/rt/gpu/CanonicalTraversalScene.cpp:73-186),
 *                          TransformPosition (:13-25), TransformNormal (:27-50)
 *   ora_canonical_shading  the material / light conversion of
 *                          BuildMegakernelHostScene
 *                          (src/renderers/Wave2Runtime.cpp:469-556)
 *
 * Records are handled as raw words in the ABI-v0 layouts
 * (include/contracts/SceneAbiV0.hpp:64-121): vertex 16 words, geometry 16,
 * instance 56, material 32, light 24; canonical triangle 44 words; L6
 * material 32, L6 light 24.  Sequential, in the reference's loop order.
 * Pinned by the reference's own BuildCanonicalTraversalScene where it builds
 * (oracle/_ref) and by identities (identity transforms reproduce the object
 * space, rotations keep lengths, counts and identity tuples). */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float wf(const uint32_t* w, int k) { float f; memcpy(&f, &w[k], 4); return f; }
static void setf(uint32_t* w, int k, float f) { memcpy(&w[k], &f, 4); }

/* Returns 0 on success, else a nonzero code naming the reference's rejection:
 * 1 view, 2 instance geometry range, 3 material, 4 index range, 5 vertex,
 * 6 non-finite, 7 no visible triangle, 8 capacity. */
int ora_canonical_expand(const uint32_t* vertices, uint32_t vertexCount, const uint32_t* indices,
                         uint32_t indexCount, const uint32_t* geometries, uint32_t geometryCount,
                         const uint32_t* instances, uint32_t instanceCount, const uint32_t* materials,
                         uint32_t materialCount, uint32_t* out, uint32_t capacity, uint32_t* count) {
    if (vertexCount == 0 || indexCount == 0 || geometryCount == 0 || instanceCount == 0 || materialCount == 0) return 1;
    uint32_t t = 0;
    for (uint32_t i = 0; i < instanceCount; ++i) {
        const uint32_t* inst = instances + 56u * i;
        const float* o2w = (const float*)inst;          /* rows, 16 floats */
        const float* w2o = (const float*)(inst + 16);
        const uint32_t* meta = inst + 48;
        if ((meta[3] & 1u) == 0u) continue;             /* InstanceFlagVisible */
        if (meta[0] >= geometryCount || meta[1] > geometryCount - meta[0]) return 2;
        for (uint32_t g = 0; g < meta[1]; ++g) {
            const uint32_t* geo = geometries + 16u * (meta[0] + g);
            if (geo[6] >= materialCount) return 3;
            const uint32_t* mat = materials + 32u * geo[6];
            for (uint32_t off = 0; off < geo[1]; off += 3u) {
                const uint32_t firstIndex = geo[0] + off;
                if (firstIndex > indexCount || 3u > indexCount - firstIndex) return 4;
                uint32_t rec[44];
                memset(rec, 0, sizeof(rec));
                for (uint32_t c = 0; c < 3u; ++c) {
                    const uint32_t vi = indices[firstIndex + c] + geo[2];
                    if (vi >= vertexCount) return 5;
                    const uint32_t* v = vertices + 16u * vi;
                    const float px = wf(v, 0), py = wf(v, 1), pz = wf(v, 2);
                    const float nx = wf(v, 4), ny = wf(v, 5), nz = wf(v, 6);
                    /* TransformPosition: rows, left-to-right fp32 sums */
                    const float x = o2w[0] * px + o2w[1] * py + o2w[2] * pz + o2w[3];
                    const float y = o2w[4] * px + o2w[5] * py + o2w[6] * pz + o2w[7];
                    const float z = o2w[8] * px + o2w[9] * py + o2w[10] * pz + o2w[11];
                    /* TransformNormal: columns of worldToObject, fp64 length */
                    float mx = w2o[0] * nx + w2o[4] * ny + w2o[8] * nz;
                    float my = w2o[1] * nx + w2o[5] * ny + w2o[9] * nz;
                    float mz = w2o[2] * nx + w2o[6] * ny + w2o[10] * nz;
                    const double ls = (double)mx * mx + (double)my * my + (double)mz * mz;
                    if (!(ls > 0.0) || !isfinite(ls)) {
                        mx = 0.0f; my = 0.0f; mz = 0.0f;
                    } else {
                        const float inv = (float)(1.0 / sqrt(ls));
                        mx *= inv; my *= inv; mz *= inv;
                    }
                    if (!isfinite(x) || !isfinite(y) || !isfinite(z) || !isfinite(mx) || !isfinite(my) ||
                        !isfinite(mz)) {
                        return 6;
                    }
                    setf(rec, 4 * c + 0, x); setf(rec, 4 * c + 1, y); setf(rec, 4 * c + 2, z); setf(rec, 4 * c + 3, 1.0f);
                    setf(rec, 12 + 4 * c + 0, mx); setf(rec, 12 + 4 * c + 1, my); setf(rec, 12 + 4 * c + 2, mz);
                    setf(rec, 12 + 4 * c + 3, 0.0f);
                    memcpy(&rec[24 + 4 * c], &v[12], 16);   /* texcoord0 */
                }
                rec[36] = meta[2];                          /* instance (stable ID) */
                rec[37] = geo[3] + off / 3u;                /* primitive */
                rec[38] = geo[4];                           /* geometry ID */
                rec[39] = mat[30];                          /* material stable ID */
                rec[40] = geo[7];                           /* GeometryFlags */
                rec[41] = mat[29];                          /* MaterialFlags */
                rec[42] = t;                                /* traversal ID */
                rec[43] = meta[3];                          /* InstanceFlags */
                if (out != 0) {
                    if (t >= capacity) return 8;
                    memcpy(out + 44u * t, rec, sizeof(rec));
                }
                ++t;
            }
        }
    }
    if (t == 0) return 7;
    *count = t;
    return 0;
}

/* materials: materialCount x 32 words in, out32 x 32 words out; lights:
 * lightCount x 24 words in, out24 x 24 words out.  Returns 0, or 1 when an
 * emissive triangle light names no expanded triangle. */
int ora_canonical_shading(const uint32_t* materials, uint32_t materialCount, const uint32_t* lights,
                          uint32_t lightCount, const uint32_t* triangles, uint32_t triangleCount, uint32_t* matOut,
                          uint32_t* lightOut) {
    for (uint32_t m = 0; m < materialCount; ++m) {
        const uint32_t* s = materials + 32u * m;
        uint32_t* o = matOut + 32u * m;
        memset(o, 0, 128);
        const float e = wf(s, 7);
        setf(o, 0, wf(s, 0)); setf(o, 1, wf(s, 1)); setf(o, 2, wf(s, 2)); setf(o, 3, wf(s, 8));
        setf(o, 4, wf(s, 4) * e); setf(o, 5, wf(s, 5) * e); setf(o, 6, wf(s, 6) * e); setf(o, 7, wf(s, 9));
        setf(o, 8, wf(s, 12)); setf(o, 9, wf(s, 13));
        memcpy(&o[12], &s[16], 16);                     /* attenuation colour + distance */
        setf(o, 16, 0.04f); setf(o, 17, 0.04f); setf(o, 18, 0.04f);
        setf(o, 20, 1.0f); setf(o, 21, 1.0f); setf(o, 22, 1.0f);
        const int smooth = s[28] == 1u;
        const int rough = !smooth && wf(s, 12) > 0.0f;
        float bmax = wf(s, 0);
        if (wf(s, 1) > bmax) bmax = wf(s, 1);
        if (wf(s, 2) > bmax) bmax = wf(s, 2);
        float emax = wf(o, 4);
        if (wf(o, 5) > emax) emax = wf(o, 5);
        if (wf(o, 6) > emax) emax = wf(o, 6);
        uint32_t flags = 0;
        if (emax > 0.0f && bmax == 0.0f && wf(s, 12) == 0.0f) flags |= 1u;   /* pure emitter */
        if (s[29] & 8u) flags |= 2u;                                          /* thin walled */
        o[28] = smooth ? 3u : (rough ? 4u : 5u);
        o[29] = smooth ? (4u | 8u | 16u) : (rough ? (2u | 8u | 16u) : (1u | 2u | 8u));
        o[30] = 0u;
        o[31] = flags;
    }
    for (uint32_t l = 0; l < lightCount; ++l) {
        const uint32_t* s = lights + 24u * l;
        uint32_t* o = lightOut + 24u * l;
        memcpy(o, s, 64);
        o[16] = s[16]; o[17] = s[21]; o[18] = s[17]; o[19] = s[19];
        o[20] = s[20]; o[21] = s[22]; o[22] = s[23]; o[23] = s[18];
        if (s[16] == 4u) {                              /* emissive triangle: first match */
            uint32_t k = 0;
            for (; k < triangleCount; ++k) {
                const uint32_t* tri = triangles + 44u * k;
                if (tri[36] == s[18] && tri[37] == s[19]) break;
            }
            if (k == triangleCount) return 1;
            o[20] = k;
        }
    }
    return 0;
}
