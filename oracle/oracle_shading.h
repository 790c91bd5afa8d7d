/*
 * Oracle (test infrastructure only): PBR shading library, light sampling,
 * Philox sampling and the per-path wavefront integrator.
 *
 * Numeric conventions shared with the product kernels (both sides must
 * evaluate the same expression tree for bit-exact parity):
 *   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross         as written in HLSL
 *   rsqrt(x)      = 1.0f / sqrtf(x)           (both correctly rounded)
 *   normalize(v)  = v * rsqrt(dot(v, v))
 *   length(v)     = sqrtf(dot(v, v))
 *   lerp(a,b,t)   = a + (b - a) * t
 *   clamp(x,l,h)  = fminf(fmaxf(x, l), h);  min/max = fminf/fmaxf
 *   rcp(x)        = 1.0f / x
 *   cos(2*pi*u), sin(2*pi*u) = ora_sincos_2pi(u): quadrant reduction of 4u
 *                   (exact) and Taylor polynomials of degree 9/10 in the
 *                   reduced argument, Horner form, no FMA.
 * No other transcendental is used on the supported path.
 */
#ifndef REX_ORACLE_SHADING_H
#define REX_ORACLE_SHADING_H

#include <stdint.h>

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PbrMaterialGpu, integrators/megakernel/include/MegakernelBridge.hpp (128 B). */
typedef struct ora_material {
    float baseColorMetallic[4];
    float emissiveRoughness[4];
    float transmissionIor[4];
    float attenuationColorDistance[4];
    float f0[4];
    float conductorEta[4];
    float conductorK[4];
    uint32_t metadata[4]; /* model, allowed lobes, complex Fresnel, private flags */
} ora_material;

/* PbrLightGpu (96 B). identity = (type, flags, -, primitive),
 * payload = (fixture triangle, -, -, instance). */
typedef struct ora_light {
    float positionRange[4];
    float directionCosOuter[4];
    float radianceScale[4];
    float shapeParams[4];
    uint32_t identity[4];
    uint32_t payload[4];
} ora_light;

typedef struct ora_alias {
    float q, pmf;
    uint32_t alias, item;
} ora_alias;

/* FixtureTriangleGpu: canonical-order positions, metadata =
 * (material, primitive, instance, geometry flags). */
typedef struct ora_fixture {
    float p0[4], p1[4], p2[4];
    uint32_t metadata[4];
} ora_fixture;

typedef struct ora_emitter {
    uint32_t instanceId, primitiveId, lightIndex, reserved;
} ora_emitter;

/* Light types / flags (resources/shaders/integrators/pbr_l6_types.hlsli:10-24). */
#define ORA_LIGHT_POINT 0u
#define ORA_LIGHT_DIRECTIONAL 1u
#define ORA_LIGHT_SPOT 2u
#define ORA_LIGHT_SPHERE 3u
#define ORA_LIGHT_EMISSIVE_TRIANGLE 4u
#define ORA_LIGHT_ENVIRONMENT 5u
#define ORA_LIGHT_ENABLED 1u
#define ORA_LIGHT_TWO_SIDED 2u

/* Direct-light estimators (pbr_l6_types.hlsli:39-42). */
#define ORA_EST_BSDF_ONLY 0u
#define ORA_EST_NEE 1u
#define ORA_EST_MIS 2u
#define ORA_EST_RESTIR 3u

/* Everything the shading stages read besides the BVH. */
typedef struct ora_shading {
    const ora_material* materials;
    uint32_t materialCount;
    const ora_light* lights;
    uint32_t lightCount;
    const ora_alias* lightAlias;      /* selection alias table */
    uint32_t lightAliasCount;
    const float* selectionPmf;        /* per light */
    const ora_fixture* fixtures;      /* canonical order */
    uint32_t fixtureCount;
    const ora_emitter* emitters;      /* sorted by (instance, primitive) */
    uint32_t emitterCount;
    uint32_t environmentLight;        /* index or ORA_INVALID */
    uint32_t environmentSampler;      /* 0: uniform sphere, 1: importance map */
    const float* envTexels;           /* RGBA, envWidth x envHeight, row 0 at +y */
    const ora_alias* envRows;         /* envHeight entries */
    const ora_alias* envColumns;      /* envHeight x envWidth entries */
    uint32_t envWidth, envHeight;
} ora_shading;

/* Frame constants (the subset of PbrFrameConstantsGpuL6 the path reads). */
typedef struct ora_frame {
    float cameraPosition[3];
    float tanHalfFov;
    float cameraForward[3];
    float aspect;
    float cameraRight[3];
    float cameraUp[3];
    uint32_t width, height;
    uint32_t sampleIndex;   /* base sample index of this launch */
    uint32_t maxBounce;
    uint32_t spp;           /* samples per pixel in this launch */
    uint32_t seedLow, seedHigh, streamTag;
    uint32_t estimator;
    float rrStart, rrMinimum, rrMaximum, rayEpsilon;
} ora_frame;

/* Per-sample path result: the five signals of ResolveCS. */
typedef struct ora_path_result {
    float cameraEmission[3];
    float directDiffuse[3];
    float directSpecular[3];
    float indirectDiffuse[3];
    float indirectSpecular[3];
    uint32_t flags;          /* path flags at the end */
    uint32_t bounces;        /* closest-hit queries issued */
    uint32_t shadowRays;     /* any-hit queries issued */
} ora_path_result;

/* ---- math / RNG ---------------------------------------------------------- */
void ora_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);
uint32_t ora_bounce_dim(uint32_t bounce, uint32_t local);
void ora_bounce_dim_locals(uint32_t out[4]);
float ora_sample_dimension(uint32_t pixel, uint32_t sample, uint32_t dim, uint32_t tag,
                           uint32_t seedLow, uint32_t seedHigh);
void ora_sincos_2pi(float u, float* s, float* c);

/* ---- host tables ---------------------------------------------------------- */
/* Vose alias table (MegakernelBridge.cpp:40-145): weights in double. */
int ora_build_alias(const double* weights, const uint32_t* items, uint32_t count, ora_alias* out);
/* The environment map's float luminance per texel and its two-level alias
 * distribution (BuildEnvironmentDistribution). */
void ora_environment_luminance(const float* rgba, uint32_t count, float* out);
int ora_build_environment_distribution(uint32_t width, uint32_t height, const float* luminance, ora_alias* rows,
                                       ora_alias* columns, double* integrated, int* uniformFallback);
/* atan2 / acos from + - * / sqrt only (the product's kernels use the same). */
float ora_atan2(float y, float x);
float ora_acos(float x);
/* (u, v) of a direction on the lat-long map (PbrEnvironmentUvFromLocalDirectionL6). */
void ora_environment_uv(const float dir[3], float uv[2]);
/* LightPowerWeight (MegakernelBridge.cpp:274-353) for point, directional,
 * emissive triangle and environment (constant radiance) lights. */
double ora_light_power(const ora_light* light, float sceneRadius, double envIntegratedLuminance);

/* ---- one-sample integrator ------------------------------------------------ */
/* Runs the path of (pixel, sample) through the wavefront stage semantics
 * (RayGen, Intersect, Shade, TraceShadow, NextBounce, Resolve) with the
 * traversal walk in `order` (0 left-first, 1 near-first). */
void ora_render_path(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                     const ora_shading* sh, const ora_frame* fr, int order, uint32_t pixel,
                     uint32_t sampleIndex, ora_path_result* out, uint64_t* traversalCounters);

/* Renders every (pixel, sample) of one launch with `threads` host threads:
 * results[(pixel * spp) + s].  Returns 0. */
int ora_render_frame(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                     const ora_shading* sh, const ora_frame* fr, int order, uint32_t threads,
                     ora_path_result* results, uint64_t* traversalCounters);

/* Rows [rowBegin, rowBegin + rowCount) of one launch:
 * results[(pixel - rowBegin * width) * spp + s].  Returns 0, -1 on a bad range. */
int ora_render_rows(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                    const ora_shading* sh, const ora_frame* fr, int order, uint32_t rowBegin, uint32_t rowCount,
                    uint32_t threads, ora_path_result* results, uint64_t* traversalCounters);

/* Progressive-mean Film update of ComposeCS (reconstruction/shaders/ComposeCS.hlsl:35-47)
 * over the launch's samples in order: film[i].xyz, film[i].w = count. */
void ora_film_accumulate(const ora_path_result* results, uint32_t pixelCount, uint32_t spp, float* film);

/* As ora_render_rows, also writing bounce 0's ABI-v2 primary surface
 * (GpuPrimarySurfaceV2, 96 B) of each pixel of the rows into `surfaces`
 * (may be NULL; requires spp == 1, one path per pixel). */
int ora_render_rows_surfaces(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                             const ora_shading* sh, const ora_frame* fr, int order, uint32_t rowBegin,
                             uint32_t rowCount, uint32_t threads, ora_path_result* results,
                             uint64_t* traversalCounters, void* surfaces);

/* ---- pieces shared with the ReSTIR oracle ---------------------------------- */
/* Private light-sample measures (pbr_l6_types.hlsli:31-34). */
#define ORA_MEASURE_DISCRETE 0u
#define ORA_MEASURE_AREA 1u
#define ORA_MEASURE_SOLID_ANGLE 2u
#define ORA_MEASURE_INVALID 0xffffffffu
/* PbrLightSampleL6 (the fields the persistent sample keeps). */
typedef struct ora_light_sample_full {
    float wi[3];
    float distance;
    float Li[3];
    float selectionPmf;
    float conditionalPdf;
    float combinedPdfW;
    uint32_t lightIndex, primitiveId, flags, measure;  /* flags: 1 valid, 2 delta, 4 infinite, 8 two-sided */
    float position[3];
} ora_light_sample_full;
/* PbrSampleOneLightL6 at p with rnd = (selection, shape0..shape3). */
void ora_light_sample_full_eval(const ora_shading* sh, const float p[3], const float rnd[5], ora_light_sample_full* out);
/* BsdfRoughnessIsFiniteModelL6 (PbrBsdf.hlsli:225-237). */
int ora_roughness_finite(float perceptualRoughness);
/* BsdfFresnelSchlickL6 (PbrBsdf.hlsli:543-549). */
void ora_schlick3(float cosine, const float f0[3], float out[3]);
/* BsdfGgxDL6(float3(0, 0, cosH), alpha) and BsdfGgxG2L6 on the directions
 * (sqrt(max(0, 1 - c * c)), 0, c) (PbrBsdf.hlsli:594-632). */
float ora_ggx_d_cos(float cosH, float alpha);
float ora_ggx_g2_cos(float cosO, float cosI, float alpha);

/* ---- hooks for the oracle's own pin tests ---------------------------------- */
void ora_bsdf_eval(const ora_material* m, const float gn[3], const float sn[3], const float wo[3],
                   const float wi[3], float out[12]);
void ora_bsdf_eval_ctx(const ora_material* m, const float gn[3], const float sn[3], float etaI, float etaT,
                       uint32_t transport, const float wo[3], const float wi[3], float out[13]);
void ora_bsdf_sample_ctx(const ora_material* m, const float gn[3], const float sn[3], float etaI, float etaT,
                         uint32_t transport, const float wo[3], const float u[3], float out[12]);
void ora_material_model(const ora_material* m, uint32_t out[2]);
void ora_bsdf_sample(const ora_material* m, const float gn[3], const float sn[3], const float wo[3],
                     const float u[3], float out[9]);
void ora_light_sample(const ora_shading* sh, const float p[3], const float rnd[5], float out[11]);
float ora_light_pdf(const ora_shading* sh, uint32_t light, const float wi[3], float dist, const float en[3]);
/* The same at shading position p (sphere lights depend on it). */
float ora_light_pdf_at(const ora_shading* sh, uint32_t light, const float p[3], const float wi[3], float dist,
                       const float en[3]);
float ora_power_heuristic(float a, float b);
uint32_t ora_sample_alias(const ora_alias* table, uint32_t count, float u, float* pmf);
void ora_build_basis(const float n[3], float t[3], float b[3]);

#ifdef __cplusplus
}
#endif
#endif
