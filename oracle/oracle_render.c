/*
 * Oracle (test infrastructure only): the PBR shading library and the
 * wavefront integrator, restated per path.
 *
 * Sources restated (paths relative to the RenderingEngine tree):
 *   RayGenCS            integrators/wavefront/shaders/RayGen.hlsl:10-93
 *   IntersectCS         integrators/wavefront/shaders/Intersect.hlsl:4-39 and
 *                       PbrSoftwareTraversalImplementation.hlsli:13-78
 *   ShadeCS             integrators/wavefront/shaders/Shade.hlsl:6-479
 *   WfPbr* helpers      integrators/wavefront/shaders/include/PbrWavefrontShading.hlsli:16-292
 *   TraceShadowCS       integrators/wavefront/shaders/TraceShadow.hlsl:5-92 (Physical)
 *   NextBounceCS        integrators/wavefront/shaders/NextBounce.hlsl:3-115
 *   ResolveCS           integrators/wavefront/shaders/Resolve.hlsl:4-90
 *   BSDF                resources/shaders/include/bsdf/PbrBsdf.hlsli (models 0 Lambert and
 *                       5 metallic-roughness with transmission 0: :184-196, :239-367,
 *                       :369-496, :543-549, :594-709, :710-956, :1115-1229, :1261-1269,
 *                       :1390-1435, :1569-1677)
 *   material bridge     resources/shaders/integrators/pbr_material_bsdf.hlsli:8-89
 *   lights              resources/shaders/integrators/pbr_light_sampling.hlsli:14-72,
 *                       :99-158 (constant environment), :185-307, :422-589, :591-724
 *   emitter lookup      resources/shaders/integrators/pbr_traversal_resources.hlsli:21-69
 *   power heuristic     resources/shaders/integrators/pbr_l6_types.hlsli:423-440
 *
 * Scope: BSDF models 0 and 5 (transmission 0), point / directional /
 * emissive-triangle / constant-radiance environment lights with the
 * uniform-sphere environment sampler, estimators BSDF-only / NEE / MIS,
 * Physical shadows.  Anything else is rejected before rendering.
 *
 * Parity pins: see tests/test_oracle_shading.py -- Lambert and GGX closed
 * forms (pdf integrates to 1, white-furnace energy, reciprocity), the
 * triangle-light pdf d^2/(A cos), alias PMFs, the Philox KAT, and the Film
 * running mean equal to the arithmetic mean.  The whole-path composition
 * (Shade / NEE / MIS / RR order) has no CPU counterpart in the reference:
 * parity unpinned beyond these component checks and the white-furnace /
 * energy-conservation render checks.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_restir.h"
#include "oracle_shading.h"

/* ---- vector helpers ------------------------------------------------------ */
typedef struct { float x, y, z; } v3;
static v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v3 L3(const float* p) { return V(p[0], p[1], p[2]); }
static v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3 muls(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static v3 divs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static float rsq(float x) { return 1.0f / sqrtf(x); }
static v3 normalize(v3 a) { return muls(a, rsq(dot(a, a))); }
static float length(v3 a) { return sqrtf(dot(a, a)); }
static int fin(float a) { return isfinite(a); }
static int fin3(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static int any_neg(v3 a) { return a.x < 0.0f || a.y < 0.0f || a.z < 0.0f; }
static int fin_nonneg3(v3 a) { return fin3(a) && !any_neg(a); }
static float maxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
static float lum(v3 a) { return dot(a, V(0.2126f, 0.7152f, 0.0722f)); }
static v3 zero3(void) { return V(0.0f, 0.0f, 0.0f); }

/* ---- BSDF (PbrBsdf.hlsli) ------------------------------------------------- */
#define PI_F 3.14159265358979323846f
#define INV_PI_F 0.31830988618379067154f
#define DIR_EPS 1.0e-7f
#define MIN_ALPHA 1.0e-4f
#define LOBE_DIFFUSE 1u
#define LOBE_GLOSSY 2u
#define LOBE_SPECULAR 4u
#define LOBE_REFLECTION 8u
#define LOBE_TRANSMISSION 16u
#define MODEL_LAMBERT 0u
#define MODEL_CONDUCTOR 1u
#define MODEL_DIELECTRIC_REFLECTION 2u
#define MODEL_SMOOTH_GLASS 3u
#define MODEL_ROUGH_DIELECTRIC 4u
#define MODEL_MR 5u

typedef struct {
    v3 geometricNormal; float etaIncident;
    v3 shadingNormal; float etaTransmitted;
    v3 tangent; uint32_t transportMode;
} bsdf_ctx;

typedef struct {
    v3 baseColor; float metallic;
    v3 f0; float perceptualRoughness;
    v3 conductorEta; float transmission;
    v3 conductorK;
    uint32_t model, allowedLobes, complexFresnel;
} bsdf_params;

#define MEASURE_INVALID 0u
#define MEASURE_SOLID_ANGLE 1u
#define MEASURE_DISCRETE 2u

typedef struct {
    v3 value; float pdf;
    v3 diffuseValue, specularValue;
    uint32_t lobeFlags, isValid, measure;
} bsdf_eval;

typedef struct {
    v3 direction; float pdf;
    v3 value; float eta;
    v3 diffuseValue, specularValue;
    uint32_t lobeFlags, isDelta, isValid, measure;
} bsdf_sample;

static bsdf_eval eval_invalid(void) { bsdf_eval e; memset(&e, 0, sizeof(e)); return e; }
static bsdf_sample sample_invalid(void) { bsdf_sample s; memset(&s, 0, sizeof(s)); s.eta = 1.0f; return s; }

/* :184-196 */
static int safe_normalize(v3 v, v3* out) {
    *out = zero3();
    const float l2 = dot(v, v);
    if (!fin(l2) || l2 <= DIR_EPS * DIR_EPS) return 0;
    *out = muls(v, rsq(l2));
    return fin3(*out);
}

static int has_flags(uint32_t mask, uint32_t req) { return (mask & req) == req; }
static int color_unit(v3 c) { return fin3(c) && !any_neg(c) && c.x <= 1.0f && c.y <= 1.0f && c.z <= 1.0f; }
static int unit_range(float t) { return fin(t) && t >= 0.0f && t <= 1.0f; }

/* :218-223 */
static int positive_eta(const bsdf_ctx* c) {
    return fin(c->etaIncident) && fin(c->etaTransmitted) && c->etaIncident > 0.0f && c->etaTransmitted > 0.0f;
}

/* :225-237 */
static int roughness_finite(float r) {
    if (!fin(r) || r <= 0.0f || r > 1.0f) return 0;
    return r * r >= MIN_ALPHA;
}

/* ValidateBsdfParamsL6, :239-352 (all six models) */
static int validate_params(const bsdf_ctx* c, const bsdf_params* p) {
    if (p->allowedLobes == 0u || c->transportMode > 1u) return 0;
    switch (p->model) {
    case MODEL_LAMBERT: return color_unit(p->baseColor);
    case MODEL_CONDUCTOR:
        if (!roughness_finite(p->perceptualRoughness)) return 0;
        if (p->complexFresnel != 0u) {
            return fin3(p->conductorEta) && fin3(p->conductorK) && p->conductorEta.x > 0.0f &&
                   p->conductorEta.y > 0.0f && p->conductorEta.z > 0.0f && !any_neg(p->conductorK);
        }
        return color_unit(p->f0);
    case MODEL_DIELECTRIC_REFLECTION: return roughness_finite(p->perceptualRoughness) && positive_eta(c);
    case MODEL_SMOOTH_GLASS: return positive_eta(c) && unit_range(p->transmission);
    case MODEL_ROUGH_DIELECTRIC:
        return roughness_finite(p->perceptualRoughness) && positive_eta(c) && unit_range(p->transmission);
    case MODEL_MR:
        return color_unit(p->baseColor) && color_unit(p->f0) && roughness_finite(p->perceptualRoughness) &&
               unit_range(p->metallic) && unit_range(p->transmission);
    default: return 0;
    }
}

/* :354-367 */
static int supports_restir(const bsdf_params* p) {
    const uint32_t required = LOBE_DIFFUSE | LOBE_REFLECTION | LOBE_GLOSSY | LOBE_REFLECTION;
    const uint32_t unsupported = LOBE_TRANSMISSION | LOBE_SPECULAR;
    return p->model == MODEL_MR && p->complexFresnel == 0u && roughness_finite(p->perceptualRoughness) &&
           fin(p->transmission) && p->transmission == 0.0f && (p->allowedLobes & required) == required &&
           (p->allowedLobes & unsupported) == 0u;
}

/* :369-432 */
static int build_frame(const bsdf_ctx* c, v3 woIn, v3* gn, v3* sn, v3* t, v3* b, v3* woW, v3* woL) {
    *gn = *sn = *t = *b = *woW = *woL = zero3();
    if (!safe_normalize(c->geometricNormal, gn) || !safe_normalize(c->shadingNormal, sn) ||
        !safe_normalize(woIn, woW))
        return 0;
    if (dot(*sn, *gn) < 0.0f) *sn = neg(*sn);
    if (dot(*woW, *gn) < 0.0f) { *gn = neg(*gn); *sn = neg(*sn); }
    if (dot(*woW, *gn) <= DIR_EPS || dot(*woW, *sn) <= DIR_EPS) return 0;
    const v3 tc = sub(c->tangent, muls(*sn, dot(c->tangent, *sn)));
    if (!safe_normalize(tc, t)) {
        const v3 helper = fabsf(sn->z) < 0.999f ? V(0.0f, 0.0f, 1.0f) : V(0.0f, 1.0f, 0.0f);
        if (!safe_normalize(cross(helper, *sn), t)) return 0;
    }
    if (!safe_normalize(cross(*sn, *t), b)) return 0;
    *woL = V(dot(*woW, *t), dot(*woW, *b), dot(*woW, *sn));
    return fin3(*woL) && woL->z > DIR_EPS;
}

static v3 to_local(v3 d, v3 t, v3 b, v3 n) { return V(dot(d, t), dot(d, b), dot(d, n)); }
static v3 to_world(v3 d, v3 t, v3 b, v3 n) { return add(add(muls(t, d.x), muls(b, d.y)), muls(n, d.z)); }

/* :455-472 */
static int validate_world_event(v3 wo, v3 wi, v3 gn, uint32_t lobes) {
    const float side = dot(wo, gn) * dot(wi, gn);
    const int refl = (lobes & LOBE_REFLECTION) == LOBE_REFLECTION;
    const int trans = (lobes & LOBE_TRANSMISSION) == LOBE_TRANSMISSION;
    if (refl == trans) return 0;
    return refl ? side > 0.0f : side < 0.0f;
}

/* :474-496 */
static int shading_correction(const bsdf_ctx* c, v3 gn, v3 sn, v3 wo, v3 wi, float* corr) {
    *corr = 0.0f;
    const float num = c->transportMode == 0u ? fabsf(dot(wi, sn)) : fabsf(dot(wo, sn));
    const float den = c->transportMode == 0u ? fabsf(dot(wi, gn)) : fabsf(dot(wo, gn));
    if (!fin(num) || !fin(den) || den <= DIR_EPS) return 0;
    *corr = num / den;
    return fin(*corr) && *corr >= 0.0f;
}

/* BsdfFresnelDielectricL6, :498-541 */
static int fresnel_dielectric(float cosIn, float etaI, float etaT, float* F) {
    *F = 0.0f;
    if (!fin(cosIn) || !fin(etaI) || !fin(etaT) || etaI <= 0.0f || etaT <= 0.0f) return 0;
    const float ci = fminf(fabsf(cosIn), 1.0f);
    const float ratio = etaI / etaT;
    const float st2 = ratio * ratio * fmaxf(0.0f, 1.0f - ci * ci);
    if (st2 >= 1.0f) {
        *F = 1.0f;
        return 1;
    }
    const float ct = sqrtf(fmaxf(0.0f, 1.0f - st2));
    const float parDen = etaT * ci + etaI * ct;
    const float perDen = etaI * ci + etaT * ct;
    if (fabsf(parDen) <= DIR_EPS || fabsf(perDen) <= DIR_EPS) return 0;
    const float par = (etaT * ci - etaI * ct) / parDen;
    const float per = (etaI * ci - etaT * ct) / perDen;
    *F = 0.5f * (par * par + per * per);
    return fin(*F) && *F >= 0.0f && *F <= 1.0f;
}

/* :543-549 */
static v3 schlick(float cosine, v3 f0) {
    const float om = 1.0f - fminf(fabsf(cosine), 1.0f);
    const float f = om * om * om * om * om;
    return add(f0, muls(sub(V(1.0f, 1.0f, 1.0f), f0), f));
}

/* BsdfFresnelConductorL6, :551-592, per channel */
static int fresnel_conductor1(float ci, float c2, float s2, float eta, float k, float* F) {
    const float e2 = eta * eta, k2 = k * k;
    const float t0 = e2 - k2 - s2;
    const float a2b2 = sqrtf(fmaxf(t0 * t0 + 4.0f * e2 * k2, 0.0f));
    const float t1 = a2b2 + c2;
    const float a = sqrtf(fmaxf(0.5f * (a2b2 + t0), 0.0f));
    const float t2 = 2.0f * ci * a;
    const float rsDen = t1 + t2;
    const float t3 = c2 * a2b2 + s2 * s2;
    const float t4 = t2 * s2;
    const float rpDen = t3 + t4;
    if (fabsf(rsDen) <= DIR_EPS || fabsf(rpDen) <= DIR_EPS) return 0;
    const float rs = (t1 - t2) / rsDen;
    const float rp = rs * (t3 - t4) / rpDen;
    *F = 0.5f * (rs + rp);
    return 1;
}

static int fresnel_conductor(float cosIn, v3 eta, v3 k, v3* F) {
    *F = zero3();
    if (!fin(cosIn) || !fin3(eta) || !fin3(k) || eta.x <= 0.0f || eta.y <= 0.0f || eta.z <= 0.0f || any_neg(k))
        return 0;
    const float ci = fminf(fabsf(cosIn), 1.0f);
    const float c2 = ci * ci;
    const float s2 = fmaxf(0.0f, 1.0f - c2);
    float fx, fy, fz;
    const int okx = fresnel_conductor1(ci, c2, s2, eta.x, k.x, &fx);
    const int oky = fresnel_conductor1(ci, c2, s2, eta.y, k.y, &fy);
    const int okz = fresnel_conductor1(ci, c2, s2, eta.z, k.z, &fz);
    if (!okx || !oky || !okz) return 0;
    *F = V(fx, fy, fz);
    return fin3(*F) && !any_neg(*F) && F->x <= 1.0f && F->y <= 1.0f && F->z <= 1.0f;
}

/* :594-643 */
static float ggx_d(v3 m, float alpha) {
    if (m.z <= 0.0f || alpha < MIN_ALPHA) return 0.0f;
    const float a2 = alpha * alpha;
    const float c2 = m.z * m.z;
    const float den = c2 * (a2 - 1.0f) + 1.0f;
    if (den <= 0.0f) return 0.0f;
    return a2 / (PI_F * den * den);
}
static float ggx_lambda(v3 d, float alpha) {
    const float ac = fabsf(d.z);
    if (ac <= DIR_EPS) return 1.0e30f;
    const float s2 = fmaxf(0.0f, 1.0f - ac * ac);
    const float t2 = s2 / (ac * ac);
    return 0.5f * (sqrtf(1.0f + alpha * alpha * t2) - 1.0f);
}
static float ggx_g1(v3 d, float alpha) { return 1.0f / (1.0f + ggx_lambda(d, alpha)); }
static float ggx_g2(v3 wo, v3 wi, float alpha) {
    return 1.0f / (1.0f + ggx_lambda(wo, alpha) + ggx_lambda(wi, alpha));
}
static float ggx_vndf_pdf(v3 wo, v3 m, float alpha) {
    if (wo.z <= DIR_EPS || m.z <= 0.0f || dot(wo, m) <= 0.0f) return 0.0f;
    return ggx_d(m, alpha) * ggx_g1(wo, alpha) * fabsf(dot(wo, m)) / fabsf(wo.z);
}

/* :646-708 */
static int ggx_sample_vndf(v3 wo, float alpha, float u0, float u1, v3* m) {
    *m = zero3();
    if (wo.z <= DIR_EPS || alpha < MIN_ALPHA || !fin3(wo) || !fin(u0) || !fin(u1) || u0 < 0.0f || u1 < 0.0f ||
        u0 >= 1.0f || u1 >= 1.0f)
        return 0;
    v3 sv;
    if (!safe_normalize(V(alpha * wo.x, alpha * wo.y, wo.z), &sv)) return 0;
    if (sv.z < 0.0f) sv = neg(sv);
    v3 t1;
    const float lens2 = sv.x * sv.x + sv.y * sv.y;
    if (lens2 > DIR_EPS * DIR_EPS) t1 = muls(V(-sv.y, sv.x, 0.0f), rsq(lens2));
    else t1 = V(1.0f, 0.0f, 0.0f);
    const v3 t2 = cross(sv, t1);
    const float radius = sqrtf(u0);
    float sp, cp;
    ora_sincos_2pi(u1, &sp, &cp);
    const float dx = radius * cp;
    float dy = radius * sp;
    const float blend = 0.5f * (1.0f + sv.z);
    const float base = sqrtf(fmaxf(0.0f, 1.0f - dx * dx));
    dy = base + (dy - base) * blend;
    const float pz = sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy));
    const v3 h = add(add(muls(t1, dx), muls(t2, dy)), muls(sv, pz));
    if (!safe_normalize(V(alpha * h.x, alpha * h.y, h.z), m)) return 0;
    return m->z > 0.0f && dot(wo, *m) > 0.0f;
}

/* :710-743 */
static int ggx_terms(v3 wo, v3 wi, float alpha, v3* m, float* D, float* G, float* pdf) {
    *m = zero3(); *D = 0.0f; *G = 0.0f; *pdf = 0.0f;
    if (wo.z <= DIR_EPS || wi.z <= DIR_EPS || !safe_normalize(add(wo, wi), m)) return 0;
    if (m->z < 0.0f) *m = neg(*m);
    const float wom = fabsf(dot(wo, *m));
    if (wom <= DIR_EPS) return 0;
    *D = ggx_d(*m, alpha);
    *G = ggx_g2(wo, wi, alpha);
    *pdf = ggx_vndf_pdf(wo, *m, alpha) / (4.0f * wom);
    return *D > 0.0f && *G >= 0.0f && *pdf > 0.0f && fin(*pdf);
}

/* BsdfReflectionFresnelL6, :745-781 */
static int reflection_fresnel(const bsdf_ctx* c, const bsdf_params* p, float cosine, v3* F) {
    *F = zero3();
    if (p->model == MODEL_CONDUCTOR) {
        if (p->complexFresnel != 0u) return fresnel_conductor(cosine, p->conductorEta, p->conductorK, F);
        *F = schlick(cosine, p->f0);
        return fin3(*F);
    }
    if (p->model == MODEL_DIELECTRIC_REFLECTION) {
        float f;
        if (!fresnel_dielectric(cosine, c->etaIncident, c->etaTransmitted, &f)) return 0;
        *F = V(f, f, f);
        return 1;
    }
    *F = schlick(cosine, p->f0);
    return fin3(*F);
}

/* :783-806 */
static bsdf_eval eval_lambert_local(const bsdf_params* p, v3 wo, v3 wi) {
    bsdf_eval r = eval_invalid();
    if (!has_flags(p->allowedLobes, LOBE_DIFFUSE | LOBE_REFLECTION) || wo.z <= DIR_EPS || wi.z <= DIR_EPS)
        return r;
    r.diffuseValue = muls(p->baseColor, INV_PI_F);
    r.value = r.diffuseValue;
    r.pdf = wi.z * INV_PI_F;
    r.measure = MEASURE_SOLID_ANGLE;
    r.lobeFlags = LOBE_DIFFUSE | LOBE_REFLECTION;
    r.isValid = fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.isValid) return eval_invalid();
    return r;
}

/* BsdfEvaluateGgxReflectionLocalL6, :808-856 (models 1 and 2) */
static bsdf_eval eval_ggx_reflection_local(const bsdf_ctx* c, const bsdf_params* p, v3 wo, v3 wi) {
    bsdf_eval r = eval_invalid();
    if (!has_flags(p->allowedLobes, LOBE_GLOSSY | LOBE_REFLECTION)) return r;
    const float alpha = p->perceptualRoughness * p->perceptualRoughness;
    v3 m;
    float D, G, pdf;
    if (!ggx_terms(wo, wi, alpha, &m, &D, &G, &pdf)) return r;
    v3 F;
    if (!reflection_fresnel(c, p, fabsf(dot(wo, m)), &F)) return r;
    const float den = 4.0f * fabsf(wo.z * wi.z);
    if (den <= DIR_EPS) return r;
    r.specularValue = muls(F, D * G / den);
    r.value = r.specularValue;
    r.pdf = pdf;
    r.measure = MEASURE_SOLID_ANGLE;
    r.lobeFlags = LOBE_GLOSSY | LOBE_REFLECTION;
    r.isValid = fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.isValid) return eval_invalid();
    return r;
}

/* :858-887 */
static int mr_lobe_probabilities(const bsdf_params* p, v3 wo, float* pd, float* ps) {
    *pd = 0.0f; *ps = 0.0f;
    const int allowD = has_flags(p->allowedLobes, LOBE_DIFFUSE | LOBE_REFLECTION);
    const int allowS = has_flags(p->allowedLobes, LOBE_GLOSSY | LOBE_REFLECTION);
    const float dscale = (1.0f - p->metallic) * (1.0f - p->transmission);
    const float dw = allowD ? lum(p->baseColor) * dscale : 0.0f;
    const float sw = allowS ? lum(schlick(fabsf(wo.z), p->f0)) : 0.0f;
    const float tw = dw + sw;
    if (!fin(tw) || tw <= 0.0f) return 0;
    *pd = dw / tw;
    *ps = sw / tw;
    return fin(*pd) && fin(*ps);
}

/* :890-956 */
static bsdf_eval eval_mr_local(const bsdf_params* p, v3 wo, v3 wi) {
    bsdf_eval r = eval_invalid();
    if (wo.z <= DIR_EPS || wi.z <= DIR_EPS) return r;
    float pd, ps;
    if (!mr_lobe_probabilities(p, wo, &pd, &ps)) return r;
    const float alpha = p->perceptualRoughness * p->perceptualRoughness;
    v3 m;
    float D, G, spdf;
    const int validSpec = ggx_terms(wo, wi, alpha, &m, &D, &G, &spdf);
    if (ps > 0.0f && !validSpec) return r;
    const v3 F = validSpec ? schlick(fabsf(dot(wo, m)), p->f0) : zero3();
    const float dscale = (1.0f - p->metallic) * (1.0f - p->transmission);
    if (pd > 0.0f) {
        r.diffuseValue = muls(mul(sub(V(1.0f, 1.0f, 1.0f), F), p->baseColor), dscale * INV_PI_F);
        r.lobeFlags |= LOBE_DIFFUSE | LOBE_REFLECTION;
    }
    if (ps > 0.0f) {
        const float den = 4.0f * fabsf(wo.z * wi.z);
        if (den <= DIR_EPS) return eval_invalid();
        r.specularValue = muls(F, D * G / den);
        r.lobeFlags |= LOBE_GLOSSY | LOBE_REFLECTION;
    }
    r.value = add(r.diffuseValue, r.specularValue);
    r.pdf = pd * wi.z * INV_PI_F + ps * spdf;
    r.measure = MEASURE_SOLID_ANGLE;
    r.isValid = fin3(r.value) && fin3(r.diffuseValue) && fin3(r.specularValue) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.isValid) return eval_invalid();
    return r;
}

/* BsdfRoughDielectricHalfVectorL6, :958-980 */
static int rough_half_vector(v3 wo, v3 wi, float etaRel, int* isRefl, v3* m) {
    *isRefl = wo.z * wi.z > 0.0f;
    *m = zero3();
    const v3 h = *isRefl ? add(wo, wi) : add(wo, muls(wi, etaRel));
    if (fabsf(wo.z) <= DIR_EPS || fabsf(wi.z) <= DIR_EPS || !safe_normalize(h, m)) return 0;
    if (m->z < 0.0f) *m = neg(*m);
    return dot(*m, wi) * wi.z >= 0.0f && dot(*m, wo) * wo.z >= 0.0f;
}

/* BsdfRoughDielectricEventProbabilitiesL6, :982-1004 */
static int rough_event_probabilities(const bsdf_params* p, float F, float* pr, float* pt) {
    *pr = has_flags(p->allowedLobes, LOBE_GLOSSY | LOBE_REFLECTION) ? F : 0.0f;
    *pt = has_flags(p->allowedLobes, LOBE_GLOSSY | LOBE_TRANSMISSION) ? (1.0f - F) * p->transmission : 0.0f;
    const float total = *pr + *pt;
    if (!fin(total) || total <= 0.0f) return 0;
    *pr /= total;
    *pt /= total;
    return 1;
}

/* BsdfEvaluateRoughDielectricLocalL6, :1006-1112 */
static bsdf_eval eval_rough_dielectric_local(const bsdf_ctx* c, const bsdf_params* p, v3 wo, v3 wi) {
    bsdf_eval r = eval_invalid();
    const float etaRel = c->etaTransmitted / c->etaIncident;
    if (!fin(etaRel) || etaRel <= 0.0f) return r;
    int isRefl;
    v3 m;
    if (!rough_half_vector(wo, wi, etaRel, &isRefl, &m)) return r;
    float F;
    if (!fresnel_dielectric(dot(wo, m), c->etaIncident, c->etaTransmitted, &F)) return r;
    float pr, pt;
    if (!rough_event_probabilities(p, F, &pr, &pt)) return r;
    const float alpha = p->perceptualRoughness * p->perceptualRoughness;
    const float D = ggx_d(m, alpha);
    const float G = ggx_g2(wo, wi, alpha);
    const float npdf = ggx_vndf_pdf(wo, m, alpha);
    if (D <= 0.0f || G < 0.0f || npdf <= 0.0f) return r;
    if (isRefl) {
        if (pr <= 0.0f) return r;
        const float wom = fabsf(dot(wo, m));
        const float den = 4.0f * fabsf(wo.z * wi.z);
        if (wom <= DIR_EPS || den <= DIR_EPS) return r;
        const float v = F * (D * G / den);
        r.specularValue = V(v, v, v);
        r.pdf = npdf / (4.0f * wom) * pr;
        r.lobeFlags = LOBE_GLOSSY | LOBE_REFLECTION;
    } else {
        if (pt <= 0.0f) return r;
        const float wom = dot(wo, m);
        const float wim = dot(wi, m);
        const float term = wim + wom / etaRel;
        const float jden = term * term;
        const float cp = wo.z * wi.z;
        if (jden <= DIR_EPS || fabsf(cp) <= DIR_EPS) return r;
        const float dmdwi = fabsf(wim) / jden;
        float tv = p->transmission * (1.0f - F) * D * G * fabsf(wim * wom / (cp * jden));
        if (c->transportMode == 0u) tv /= etaRel * etaRel;
        r.specularValue = V(tv, tv, tv);
        r.pdf = npdf * dmdwi * pt;
        r.lobeFlags = LOBE_GLOSSY | LOBE_TRANSMISSION;
    }
    r.value = r.specularValue;
    r.measure = MEASURE_SOLID_ANGLE;
    r.isValid = fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.isValid) return eval_invalid();
    return r;
}

/* EvaluateBsdfL6, :1115-1229 */
static bsdf_eval eval_bsdf(const bsdf_ctx* c, const bsdf_params* p, v3 woIn, v3 wiIn) {
    bsdf_eval r = eval_invalid();
    if (!validate_params(c, p)) return r;
    v3 gn, sn, t, b, woW, woL;
    if (!build_frame(c, woIn, &gn, &sn, &t, &b, &woW, &woL)) return r;
    v3 wiW;
    if (!safe_normalize(wiIn, &wiW)) return r;
    const v3 wiL = to_local(wiW, t, b, sn);
    switch (p->model) {
    case MODEL_LAMBERT: r = eval_lambert_local(p, woL, wiL); break;
    case MODEL_CONDUCTOR:
    case MODEL_DIELECTRIC_REFLECTION: r = eval_ggx_reflection_local(c, p, woL, wiL); break;
    case MODEL_ROUGH_DIELECTRIC: r = eval_rough_dielectric_local(c, p, woL, wiL); break;
    case MODEL_MR: r = eval_mr_local(p, woL, wiL); break;
    case MODEL_SMOOTH_GLASS:
        /* a delta lobe has no finite-direction value */
        r.measure = MEASURE_DISCRETE;
        return r;
    default: break;
    }
    if (!r.isValid || !validate_world_event(woW, wiW, gn, r.lobeFlags)) return eval_invalid();
    float corr;
    if (!shading_correction(c, gn, sn, woW, wiW, &corr)) return eval_invalid();
    r.value = muls(r.value, corr);
    r.diffuseValue = muls(r.diffuseValue, corr);
    r.specularValue = muls(r.specularValue, corr);
    if (!fin3(r.value) || !fin3(r.diffuseValue) || !fin3(r.specularValue)) return eval_invalid();
    return r;
}

/* :1261-1269 */
static v3 cosine_hemisphere(float u0, float u1) {
    const float radius = sqrtf(u0);
    float sp, cp;
    ora_sincos_2pi(u1, &sp, &cp);
    return V(radius * cp, radius * sp, sqrtf(fmaxf(0.0f, 1.0f - u0)));
}

/* BsdfRefractLocalL6, :1271-1302 */
static int refract_local(v3 wo, v3 mIn, float etaRel, v3* wi) {
    *wi = zero3();
    v3 m = mIn;
    float ci = dot(m, wo);
    if (ci < 0.0f) {
        m = neg(m);
        ci = -ci;
    }
    if (ci <= DIR_EPS || !fin(etaRel) || etaRel <= 0.0f) return 0;
    const float st2 = fmaxf(0.0f, 1.0f - ci * ci) / (etaRel * etaRel);
    if (st2 >= 1.0f) return 0;
    const float ct = sqrtf(fmaxf(0.0f, 1.0f - st2));
    const v3 cand = add(divs(neg(wo), etaRel), muls(m, ci / etaRel - ct));
    return safe_normalize(cand, wi);
}

/* BsdfSampleSmoothGlassLocalL6, :1304-1388 */
static bsdf_sample sample_smooth_glass_local(const bsdf_ctx* c, const bsdf_params* p, v3 wo, float u) {
    bsdf_sample r = sample_invalid();
    float F;
    if (!fresnel_dielectric(wo.z, c->etaIncident, c->etaTransmitted, &F)) return r;
    const float rw = has_flags(p->allowedLobes, LOBE_SPECULAR | LOBE_REFLECTION) ? F : 0.0f;
    const float tw = has_flags(p->allowedLobes, LOBE_SPECULAR | LOBE_TRANSMISSION) ? (1.0f - F) * p->transmission
                                                                                    : 0.0f;
    const float total = rw + tw;
    if (!fin(total) || total <= 0.0f) return r;
    const float pr = rw / total, pt = tw / total;
    r.measure = MEASURE_DISCRETE;
    r.isDelta = 1u;
    r.eta = 1.0f;
    if (u < pr) {
        r.direction = V(-wo.x, -wo.y, wo.z);
        const float cosine = fabsf(r.direction.z);
        if (cosine <= DIR_EPS || pr <= 0.0f) return sample_invalid();
        r.pdf = pr;
        r.value = V(F / cosine, F / cosine, F / cosine);
        r.specularValue = r.value;
        r.lobeFlags = LOBE_SPECULAR | LOBE_REFLECTION;
    } else {
        const float etaRel = c->etaTransmitted / c->etaIncident;
        if (pt <= 0.0f || !refract_local(wo, V(0.0f, 0.0f, 1.0f), etaRel, &r.direction)) return sample_invalid();
        const float cosine = fabsf(r.direction.z);
        if (cosine <= DIR_EPS) return sample_invalid();
        r.pdf = pt;
        float v = (1.0f - F) * p->transmission / cosine;
        if (c->transportMode == 0u) v /= etaRel * etaRel;
        r.value = V(v, v, v);
        r.specularValue = r.value;
        r.eta = etaRel;
        r.lobeFlags = LOBE_SPECULAR | LOBE_TRANSMISSION;
    }
    r.isValid = fin3(r.direction) && fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.isValid) return sample_invalid();
    return r;
}

/* BsdfSampleFiniteLocalL6, :1390-1567 */
static bsdf_sample sample_finite_local(const bsdf_ctx* c, const bsdf_params* p, v3 wo, float u0, float u1,
                                       float u2) {
    bsdf_sample r = sample_invalid();
    r.measure = MEASURE_SOLID_ANGLE;
    r.eta = 1.0f;
    if (p->model == MODEL_LAMBERT) {
        if (!has_flags(p->allowedLobes, LOBE_DIFFUSE | LOBE_REFLECTION)) return sample_invalid();
        r.direction = cosine_hemisphere(u1, u2);
        r.lobeFlags = LOBE_DIFFUSE | LOBE_REFLECTION;
        r.isValid = 1u;
        return r;
    }
    const float alpha = p->perceptualRoughness * p->perceptualRoughness;
    if (p->model == MODEL_CONDUCTOR || p->model == MODEL_DIELECTRIC_REFLECTION) {
        if (!has_flags(p->allowedLobes, LOBE_GLOSSY | LOBE_REFLECTION)) return sample_invalid();
        v3 m;
        if (!ggx_sample_vndf(wo, alpha, u1, u2, &m)) return sample_invalid();
        r.direction = add(neg(wo), muls(m, 2.0f * dot(wo, m)));
        if (r.direction.z <= DIR_EPS) return sample_invalid();
        r.lobeFlags = LOBE_GLOSSY | LOBE_REFLECTION;
        r.isValid = 1u;
        return r;
    }
    if (p->model == MODEL_MR) {
        float pd, ps;
        if (!mr_lobe_probabilities(p, wo, &pd, &ps)) return sample_invalid();
        if (u0 < pd) {
            r.direction = cosine_hemisphere(u1, u2);
            r.lobeFlags = LOBE_DIFFUSE | LOBE_REFLECTION;
        } else {
            v3 m;
            if (ps <= 0.0f || !ggx_sample_vndf(wo, alpha, u1, u2, &m)) return sample_invalid();
            r.direction = add(neg(wo), muls(m, 2.0f * dot(wo, m)));
            if (r.direction.z <= DIR_EPS) return sample_invalid();
            r.lobeFlags = LOBE_GLOSSY | LOBE_REFLECTION;
        }
        r.isValid = 1u;
        return r;
    }
    if (p->model == MODEL_ROUGH_DIELECTRIC) {
        v3 m;
        if (!ggx_sample_vndf(wo, alpha, u1, u2, &m)) return sample_invalid();
        float F;
        if (!fresnel_dielectric(dot(wo, m), c->etaIncident, c->etaTransmitted, &F)) return sample_invalid();
        float pr, pt;
        if (!rough_event_probabilities(p, F, &pr, &pt)) return sample_invalid();
        if (u0 < pr) {
            r.direction = add(neg(wo), muls(m, 2.0f * dot(wo, m)));
            if (r.direction.z <= DIR_EPS) return sample_invalid();
            r.lobeFlags = LOBE_GLOSSY | LOBE_REFLECTION;
        } else {
            const float etaRel = c->etaTransmitted / c->etaIncident;
            if (pt <= 0.0f || !refract_local(wo, m, etaRel, &r.direction) || r.direction.z >= -DIR_EPS)
                return sample_invalid();
            r.eta = etaRel;
            r.lobeFlags = LOBE_GLOSSY | LOBE_TRANSMISSION;
        }
        r.isValid = 1u;
        return r;
    }
    return sample_invalid();
}

/* SampleBsdfL6, :1569-1677 */
static bsdf_sample sample_bsdf(const bsdf_ctx* c, const bsdf_params* p, v3 woIn, float u0, float u1, float u2) {
    bsdf_sample r = sample_invalid();
    const int unit = fin(u0) && fin(u1) && fin(u2) && u0 >= 0.0f && u1 >= 0.0f && u2 >= 0.0f && u0 < 1.0f &&
                     u1 < 1.0f && u2 < 1.0f;
    if (!validate_params(c, p) || !unit) return r;
    v3 gn, sn, t, b, woW, woL;
    if (!build_frame(c, woIn, &gn, &sn, &t, &b, &woW, &woL)) return r;
    const bsdf_sample ls = p->model == MODEL_SMOOTH_GLASS ? sample_smooth_glass_local(c, p, woL, u0)
                                                          : sample_finite_local(c, p, woL, u0, u1, u2);
    if (!ls.isValid) return r;
    v3 wiW;
    if (!safe_normalize(to_world(ls.direction, t, b, sn), &wiW) ||
        !validate_world_event(woW, wiW, gn, ls.lobeFlags))
        return sample_invalid();
    if (ls.measure == MEASURE_SOLID_ANGLE) {
        const bsdf_eval e = eval_bsdf(c, p, woW, wiW);
        if (!e.isValid) return sample_invalid();
        r.direction = wiW;
        r.pdf = e.pdf;
        r.value = e.value;
        r.eta = ls.eta;
        r.diffuseValue = e.diffuseValue;
        r.specularValue = e.specularValue;
        r.measure = MEASURE_SOLID_ANGLE;
        r.lobeFlags = ls.lobeFlags;
        r.isDelta = 0u;
        r.isValid = 1u;
        return r;
    }
    float corr;
    if (!shading_correction(c, gn, sn, woW, wiW, &corr)) return sample_invalid();
    r = ls;
    r.direction = wiW;
    r.value = muls(r.value, corr);
    r.diffuseValue = muls(r.diffuseValue, corr);
    r.specularValue = muls(r.specularValue, corr);
    r.isValid = fin3(r.value) && fin3(r.specularValue);
    if (!r.isValid) return sample_invalid();
    return r;
}

/* PbrMaterialToBsdfParamsL6, pbr_material_bsdf.hlsli:8-42: a rough
 * dielectric whose alpha is below the microfacet threshold is the smooth
 * dielectric (glossy lobes become specular). */
static bsdf_params material_params(const ora_material* m) {
    bsdf_params p;
    p.baseColor = L3(m->baseColorMetallic);
    p.metallic = m->baseColorMetallic[3];
    p.f0 = L3(m->f0);
    p.perceptualRoughness = m->emissiveRoughness[3];
    p.conductorEta = L3(m->conductorEta);
    p.transmission = m->transmissionIor[0];
    p.conductorK = L3(m->conductorK);
    p.model = m->metadata[0];
    p.allowedLobes = m->metadata[1];
    p.complexFresnel = m->metadata[2];
    if (p.model == MODEL_ROUGH_DIELECTRIC && fin(p.perceptualRoughness) && p.perceptualRoughness >= 0.0f &&
        p.perceptualRoughness * p.perceptualRoughness < MIN_ALPHA) {
        p.model = MODEL_SMOOTH_GLASS;
        if ((p.allowedLobes & LOBE_GLOSSY) != 0u) p.allowedLobes = (p.allowedLobes & ~LOBE_GLOSSY) | LOBE_SPECULAR;
    }
    return p;
}

/* pbr_light_sampling.hlsli:33-40 */
static void build_basis(v3 n, v3* t, v3* b) {
    const float sgn = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sgn + n.z);
    const float bb = n.x * n.y * a;
    *t = V(1.0f + sgn * n.x * n.x * a, sgn * bb, -sgn * n.x);
    *b = V(bb, sgn + n.y * n.y * a, -n.y);
}

/* ---- hit + lights ---------------------------------------------------------- */
typedef struct {
    float t;
    v3 position, geometricNormal, shadingNormal;
    uint32_t materialIndex, instanceId, primitiveId, emitterLightIndex, frontFace;
} pbr_hit;

typedef struct {
    v3 wi; float distance;
    v3 Li; float selectionPmf, conditionalPdf, combinedPdfW;
    uint32_t lightIndex, instanceId, primitiveId, flags;
    v3 position, normal;
    uint32_t measure;
} light_sample;

#define SAMPLE_VALID 1u
#define SAMPLE_DELTA 2u
#define SAMPLE_INFINITE 4u
#define SAMPLE_TWO_SIDED 8u

static light_sample light_invalid(void) {
    light_sample s;
    memset(&s, 0, sizeof(s));
    s.lightIndex = ORA_INVALID;
    s.instanceId = ORA_INVALID;
    s.primitiveId = ORA_INVALID;
    s.measure = ORA_MEASURE_INVALID;
    return s;
}

static v3 light_radiance_scale(const ora_light* l) { return muls(L3(l->radianceScale), l->radianceScale[3]); }

/* pbr_traversal_resources.hlsli:21-69 */
static uint32_t resolve_emitter(const ora_shading* sh, uint32_t inst, uint32_t prim) {
    uint32_t first = 0, count = sh->emitterCount;
    while (count > 0) {
        const uint32_t step = count >> 1;
        const uint32_t mid = first + step;
        const ora_emitter* e = &sh->emitters[mid];
        const int less = e->instanceId < inst || (e->instanceId == inst && e->primitiveId < prim);
        if (less) { first = mid + 1; count -= step + 1; }
        else count = step;
    }
    if (first < sh->emitterCount) {
        const ora_emitter* e = &sh->emitters[first];
        if (e->instanceId == inst && e->primitiveId == prim) {
            return e->lightIndex < sh->lightCount ? e->lightIndex : ORA_INVALID;
        }
    }
    return ORA_INVALID;
}

/* pbr_light_sampling.hlsli:47-72 */
static ora_alias sample_alias(const ora_alias* table, uint32_t count, float u) {
    ora_alias inv = {0.0f, 0.0f, 0u, ORA_INVALID};
    if (count == 0) return inv;
    const float scaled = u * (float)count;
    uint32_t bin = (uint32_t)scaled;
    if (bin > count - 1u) bin = count - 1u;
    const float within = scaled - (float)bin;
    const ora_alias* e = &table[bin];
    const uint32_t sel = within < e->q ? bin : e->alias;
    if (sel >= count) return inv;
    return table[sel];
}

#define TWO_PI_F 6.28318530717958647692f

/* PbrEnvironmentUvFromLocalDirectionL6 (pbr_light_sampling.hlsli:87-97) with
 * the shared atan2 / acos. */
static void env_uv(v3 localDirection, float* u, float* v) {
    const v3 d = normalize(localDirection);
    float phi = ora_atan2(d.z, d.x);
    if (phi < 0.0f) phi += TWO_PI_F;
    *u = phi / TWO_PI_F;
    *v = ora_acos(d.y) / PI_F;
}

void ora_environment_uv(const float dir[3], float uv[2]) { env_uv(L3(dir), &uv[0], &uv[1]); }

/* The map read as the reference's sampler does (Wave2Runtime.cpp:2571-2576):
 * nearest texel, u repeats, v clamps to the edge. */
static v3 env_texel(const ora_shading* sh, float u, float v) {
    const uint32_t w = sh->envWidth, h = sh->envHeight;
    uint32_t i = (uint32_t)fminf(fmaxf(floorf(u * (float)w), 0.0f), (float)(2u * w));
    const uint32_t j = (uint32_t)fminf(fmaxf(floorf(v * (float)h), 0.0f), (float)(h - 1u));
    i %= w;
    const float* t = sh->envTexels + 4u * ((size_t)j * w + (size_t)i);
    return V(t[0], t[1], t[2]);
}

/* cos(pi * row / height) from the shared sin / cos of 2 pi u. */
static float env_row_cos(uint32_t row, uint32_t height) {
    float sn, cs;
    ora_sincos_2pi((float)row / (float)(2u * height), &sn, &cs);
    return cs;
}

/* PbrEnvironmentTexelSolidAngleL6 (pbr_light_sampling.hlsli:74-85). */
static float env_texel_solid_angle(const ora_shading* sh, uint32_t row) {
    const uint32_t w = sh->envWidth, h = sh->envHeight;
    if (w == 0u || h == 0u || row >= h) return 0.0f;
    return (TWO_PI_F / (float)w) * (env_row_cos(row, h) - env_row_cos(row + 1u, h));
}

static int env_light_active(const ora_shading* sh, const ora_light** out) {
    const uint32_t li = sh->environmentLight;
    if (li == ORA_INVALID || li >= sh->lightCount) return 0;
    const ora_light* l = &sh->lights[li];
    if (l->identity[0] != ORA_LIGHT_ENVIRONMENT || (l->identity[1] & ORA_LIGHT_ENABLED) == 0u) return 0;
    *out = l;
    return 1;
}

/* Identity environment-to-world rows (Wave2Runtime.cpp:1416-1421). */
static v3 env_local(v3 d) {
    return normalize(V(dot(V(1.0f, 0.0f, 0.0f), d), dot(V(0.0f, 1.0f, 0.0f), d), dot(V(0.0f, 0.0f, 1.0f), d)));
}

/* PbrEnvironmentRadianceL6 (pbr_light_sampling.hlsli:99-124) */
static v3 environment_radiance(const ora_shading* sh, v3 dir) {
    const ora_light* l;
    if (!env_light_active(sh, &l) || sh->envWidth == 0u || sh->envHeight == 0u) return zero3();
    float u, v;
    env_uv(env_local(dir), &u, &v);
    return mul(env_texel(sh, u, v), light_radiance_scale(l));
}

/* PbrEnvironmentPdfL6 (pbr_light_sampling.hlsli:126-158) */
static float environment_pdf(const ora_shading* sh, v3 wi) {
    const uint32_t w = sh->envWidth, h = sh->envHeight;
    if (w == 0u || h == 0u) return 0.0f;
    if (sh->environmentSampler == 0u) return 1.0f / (4.0f * PI_F);
    float u, v;
    env_uv(env_local(wi), &u, &v);
    uint32_t row = (uint32_t)fmaxf(v * (float)h, 0.0f), col = (uint32_t)fmaxf(u * (float)w, 0.0f);
    if (row > h - 1u) row = h - 1u;
    if (col > w - 1u) col = w - 1u;
    const float pmf = sh->envRows[row].pmf * sh->envColumns[(size_t)row * w + col].pmf;
    const float sa = env_texel_solid_angle(sh, row);
    return sa > 0.0f ? pmf / sa : 0.0f;
}

/* pbr_light_sampling.hlsli:212-241 */
static light_sample sample_directional(const ora_light* l, uint32_t li, float pmf) {
    light_sample s = light_invalid();
    const v3 d = L3(l->directionCosOuter);
    const float l2 = dot(d, d);
    if (!(l2 > 0.0f) || !(pmf > 0.0f)) return s;
    s.wi = muls(neg(d), rsq(l2));
    s.distance = 1.0e30f;
    s.Li = light_radiance_scale(l);
    s.selectionPmf = pmf;
    s.conditionalPdf = 1.0f;
    s.combinedPdfW = 0.0f;
    s.lightIndex = li;
    s.flags = SAMPLE_VALID | SAMPLE_DELTA | SAMPLE_INFINITE;
    s.measure = ORA_MEASURE_DISCRETE;
    return s;
}

/* HLSL smoothstep: t = saturate((x - a) / (b - a)), t * t * (3 - 2 t). */
static float smoothstep_f(float a, float b, float x) {
    float t = (x - a) / (b - a);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

/* PbrSamplePointOrSpotLightL6, pbr_light_sampling.hlsli:160-210 */
static light_sample sample_point(v3 p, const ora_light* l, uint32_t li, float pmf) {
    light_sample s = light_invalid();
    const v3 toL = sub(L3(l->positionRange), p);
    const float d2 = dot(toL, toL);
    if (!(d2 > 0.0f) || !(pmf > 0.0f)) return s;
    const float dist = sqrtf(d2);
    if (l->positionRange[3] > 0.0f && dist > l->positionRange[3]) return s;
    const v3 wi = divs(toL, dist);
    float falloff = 1.0f;
    if (l->identity[0] == ORA_LIGHT_SPOT) {
        const v3 dir = L3(l->directionCosOuter);
        const float len2 = dot(dir, dir);
        if (!(len2 > 0.0f) || !fin(len2)) return s;
        const v3 forward = divs(dir, sqrtf(len2));
        const float cosAngle = dot(forward, neg(wi));
        const float cosOuter = l->directionCosOuter[3], cosInner = l->shapeParams[1];
        falloff = cosInner == cosOuter ? (cosAngle >= cosOuter ? 1.0f : 0.0f)
                                       : smoothstep_f(cosOuter, cosInner, cosAngle);
        if (!(falloff > 0.0f)) return s;
    }
    s.wi = wi;
    s.distance = dist;
    s.Li = muls(light_radiance_scale(l), falloff / d2);
    s.selectionPmf = pmf;
    s.conditionalPdf = 1.0f;
    s.combinedPdfW = 0.0f;
    s.lightIndex = li;
    s.position = L3(l->positionRange);
    s.flags = SAMPLE_VALID | SAMPLE_DELTA;
    s.measure = ORA_MEASURE_DISCRETE;
    return s;
}

/* PbrSampleSphereLightL6, pbr_light_sampling.hlsli:309-420 */
static light_sample sample_sphere(v3 p, const ora_light* l, uint32_t li, float pmf, float u0, float u1) {
    light_sample s = light_invalid();
    const float radius = l->shapeParams[0];
    const v3 center = L3(l->positionRange);
    if (!(radius > 0.0f) || !(pmf > 0.0f)) return s;
    const v3 toCenter = sub(center, p);
    const float d2 = dot(toCenter, toCenter);
    const float r2 = radius * radius;
    const int two = (l->identity[1] & ORA_LIGHT_TWO_SIDED) != 0u;
    v3 wi, lp, ln;
    float dist, cpdf;
    uint32_t measure;
    if (d2 > r2 * 1.000001f) {
        const float cd = sqrtf(d2);
        const v3 cdir = divs(toCenter, cd);
        const float sin2Max = r2 / d2;
        const float cosMax = sqrtf(fmaxf(0.0f, 1.0f - sin2Max));
        const float omc = sin2Max < 0.00068523f ? 0.5f * sin2Max : 1.0f - cosMax;
        const float cosTheta = 1.0f - u0 * omc;
        const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
        float sp, cp;
        ora_sincos_2pi(u1, &sp, &cp);
        v3 t, b;
        build_basis(cdir, &t, &b);
        wi = normalize(add(add(muls(cdir, cosTheta), muls(t, cp * sinTheta)), muls(b, sp * sinTheta)));
        const v3 oc = sub(p, center);
        const float halfB = dot(oc, wi);
        const float disc = halfB * halfB - (dot(oc, oc) - r2);
        if (!(disc >= 0.0f)) return s;
        dist = -halfB - sqrtf(disc);
        if (!(dist > 0.0f)) dist = -halfB + sqrtf(disc);
        if (!(dist > 0.0f) || !(omc > 0.0f)) return s;
        lp = add(p, muls(wi, dist));
        ln = normalize(sub(lp, center));
        cpdf = 1.0f / (TWO_PI_F * omc);
        measure = ORA_MEASURE_SOLID_ANGLE;
    } else {
        const float z = 1.0f - 2.0f * u0;
        const float rd = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cp;
        ora_sincos_2pi(u1, &sp, &cp);
        ln = V(rd * cp, z, rd * sp);
        lp = add(center, muls(ln, radius));
        const v3 toL = sub(lp, p);
        const float sd2 = dot(toL, toL);
        if (!(sd2 > 0.0f)) return s;
        dist = sqrtf(sd2);
        wi = divs(toL, dist);
        const float lc = dot(ln, neg(wi));
        const float ac = two ? fabsf(lc) : lc;
        if (!(ac > 0.0f)) return s;
        cpdf = 1.0f / (4.0f * PI_F * r2);
        s.combinedPdfW = pmf * (cpdf * sd2 / ac);
        measure = ORA_MEASURE_AREA;
    }
    s.wi = wi;
    s.distance = dist;
    s.Li = light_radiance_scale(l);
    s.selectionPmf = pmf;
    s.conditionalPdf = cpdf;
    if (measure == ORA_MEASURE_SOLID_ANGLE) s.combinedPdfW = pmf * cpdf;
    s.lightIndex = li;
    s.instanceId = l->payload[3];
    s.primitiveId = l->identity[3];
    s.position = lp;
    s.normal = ln;
    s.measure = measure;
    s.flags = SAMPLE_VALID | (two ? SAMPLE_TWO_SIDED : 0u);
    return s;
}

/* pbr_light_sampling.hlsli:243-307 */
static light_sample sample_triangle(const ora_shading* sh, v3 p, const ora_light* l, uint32_t li, float pmf,
                                    float u0, float u1) {
    light_sample s = light_invalid();
    if (l->payload[0] >= sh->fixtureCount || !(pmf > 0.0f)) return s;
    const ora_fixture* f = &sh->fixtures[l->payload[0]];
    if (f->metadata[2] != l->payload[3] || f->metadata[1] != l->identity[3]) return s;
    const v3 p0 = L3(f->p0), p1 = L3(f->p1), p2 = L3(f->p2);
    const v3 nu = cross(sub(p1, p0), sub(p2, p0));
    const float twiceArea = length(nu);
    if (!(twiceArea > 0.0f)) return s;
    const float root = sqrtf(u0);
    const float b0 = 1.0f - root;
    const float b1 = root * (1.0f - u1);
    const float b2 = root * u1;
    const v3 lp = add(add(muls(p0, b0), muls(p1, b1)), muls(p2, b2));
    const v3 toL = sub(lp, p);
    const float d2 = dot(toL, toL);
    if (!(d2 > 0.0f)) return s;
    const float dist = sqrtf(d2);
    const v3 wi = divs(toL, dist);
    const v3 ln = divs(nu, twiceArea);
    const float lc = dot(ln, neg(wi));
    const int two = (l->identity[1] & ORA_LIGHT_TWO_SIDED) != 0u;
    const float ac = two ? fabsf(lc) : lc;
    if (!(ac > 0.0f)) return s;
    const float area = 0.5f * twiceArea;
    const float pdfA = 1.0f / area;
    const float pdfW = pdfA * d2 / ac;
    s.wi = wi;
    s.distance = dist;
    s.Li = light_radiance_scale(l);
    s.selectionPmf = pmf;
    s.conditionalPdf = pdfA;
    s.combinedPdfW = pmf * pdfW;
    s.lightIndex = li;
    s.instanceId = l->payload[3];
    s.primitiveId = f->metadata[1];
    s.position = lp;
    s.normal = ln;
    s.flags = SAMPLE_VALID | (two ? SAMPLE_TWO_SIDED : 0u);
    s.measure = ORA_MEASURE_AREA;
    return s;
}

/* PbrSampleEnvironmentLightL6, pbr_light_sampling.hlsli:422-509 (identity
 * environment-to-world rows). */
static light_sample sample_environment(const ora_shading* sh, const ora_light* l, uint32_t li, float pmf,
                                       const float r[4]) {
    light_sample s = light_invalid();
    const uint32_t w = sh->envWidth, h = sh->envHeight;
    if (!(pmf > 0.0f) || w == 0u || h == 0u) return s;
    v3 local;
    float u, v, pdfW;
    if (sh->environmentSampler == 0u) {
        const float cosTheta = 1.0f - 2.0f * r[0];
        const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
        float sp, cp;
        ora_sincos_2pi(r[1], &sp, &cp);
        local = V(sinTheta * cp, cosTheta, sinTheta * sp);
        env_uv(local, &u, &v);
        pdfW = 1.0f / (4.0f * PI_F);
    } else {
        const ora_alias re = sample_alias(sh->envRows, h, r[0]);
        if (re.item >= h || !(re.pmf > 0.0f)) return s;
        const ora_alias ce = sample_alias(sh->envColumns + (size_t)re.item * w, w, r[1]);
        if (ce.item >= w || !(ce.pmf > 0.0f)) return s;
        const float c0 = env_row_cos(re.item, h), c1 = env_row_cos(re.item + 1u, h);
        const float cosTheta = c0 + (c1 - c0) * r[3];
        const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
        u = ((float)ce.item + r[2]) / (float)w;
        float sp, cp;
        ora_sincos_2pi(u, &sp, &cp);
        local = V(sinTheta * cp, cosTheta, sinTheta * sp);
        const float sa = env_texel_solid_angle(sh, re.item);
        if (!(sa > 0.0f)) return s;
        v = ora_acos(cosTheta) / PI_F;
        pdfW = re.pmf * ce.pmf / sa;
    }
    s.wi = env_local(local);
    s.distance = 1.0e30f;
    s.Li = mul(env_texel(sh, u, v), light_radiance_scale(l));
    s.selectionPmf = pmf;
    s.conditionalPdf = pdfW;
    s.combinedPdfW = pmf * pdfW;
    s.lightIndex = li;
    s.measure = ORA_MEASURE_SOLID_ANGLE;
    s.flags = SAMPLE_VALID | SAMPLE_INFINITE;
    return s;
}

/* pbr_light_sampling.hlsli:511-589 */
static light_sample sample_one_light(const ora_shading* sh, v3 p, const float rnd[5]) {
    if (sh->lightAliasCount == 0u) return light_invalid();
    const ora_alias sel = sample_alias(sh->lightAlias, sh->lightAliasCount, rnd[0]);
    if (sel.item >= sh->lightCount || !(sel.pmf > 0.0f)) return light_invalid();
    const ora_light* l = &sh->lights[sel.item];
    if ((l->identity[1] & ORA_LIGHT_ENABLED) == 0u) return light_invalid();
    switch (l->identity[0]) {
    case ORA_LIGHT_POINT:
    case ORA_LIGHT_SPOT: return sample_point(p, l, sel.item, sel.pmf);
    case ORA_LIGHT_DIRECTIONAL: return sample_directional(l, sel.item, sel.pmf);
    case ORA_LIGHT_SPHERE: return sample_sphere(p, l, sel.item, sel.pmf, rnd[1], rnd[2]);
    case ORA_LIGHT_EMISSIVE_TRIANGLE: return sample_triangle(sh, p, l, sel.item, sel.pmf, rnd[1], rnd[2]);
    case ORA_LIGHT_ENVIRONMENT: return sample_environment(sh, l, sel.item, sel.pmf, rnd + 1);
    default: return light_invalid();
    }
}

/* pbr_light_sampling.hlsli:591-718 */
static float light_pdf_li(const ora_shading* sh, uint32_t li, v3 p, v3 wi, float dist, v3 en) {
    if (li == ORA_INVALID || li >= sh->lightCount) return 0.0f;
    const ora_light* l = &sh->lights[li];
    if ((l->identity[1] & ORA_LIGHT_ENABLED) == 0u) return 0.0f;
    if (l->identity[0] == ORA_LIGHT_ENVIRONMENT) return environment_pdf(sh, wi);
    if (l->identity[0] == ORA_LIGHT_SPHERE) {
        const float radius = l->shapeParams[0];
        const v3 toCenter = sub(L3(l->positionRange), p);
        const float d2 = dot(toCenter, toCenter), r2 = radius * radius;
        if (!(radius > 0.0f)) return 0.0f;
        if (d2 > r2 * 1.000001f) {
            const float sin2Max = r2 / d2;
            const float cosMax = sqrtf(fmaxf(0.0f, 1.0f - sin2Max));
            const float omc = sin2Max < 0.00068523f ? 0.5f * sin2Max : 1.0f - cosMax;
            return omc > 0.0f ? 1.0f / (TWO_PI_F * omc) : 0.0f;
        }
        const float lc = dot(en, neg(wi));
        const float c = (l->identity[1] & ORA_LIGHT_TWO_SIDED) != 0u ? fabsf(lc) : lc;
        const float area = 4.0f * PI_F * r2;
        return dist > 0.0f && c > 0.0f ? dist * dist / (area * c) : 0.0f;
    }
    if (l->identity[0] == ORA_LIGHT_EMISSIVE_TRIANGLE) {
        if (l->payload[0] >= sh->fixtureCount || !(dist > 0.0f)) return 0.0f;
        const ora_fixture* f = &sh->fixtures[l->payload[0]];
        if (f->metadata[2] != l->payload[3] || f->metadata[1] != l->identity[3]) return 0.0f;
        const float area = 0.5f * length(cross(sub(L3(f->p1), L3(f->p0)), sub(L3(f->p2), L3(f->p0))));
        const float lc = dot(en, neg(wi));
        const float c = (l->identity[1] & ORA_LIGHT_TWO_SIDED) != 0u ? fabsf(lc) : lc;
        return area > 0.0f && c > 0.0f ? dist * dist / (area * c) : 0.0f;
    }
    return 0.0f;
}

static float selected_light_pdf(const ora_shading* sh, uint32_t li, v3 p, v3 wi, float dist, v3 en) {
    if (li == ORA_INVALID || li >= sh->lightCount) return 0.0f;
    return sh->selectionPmf[li] * light_pdf_li(sh, li, p, wi, dist, en);
}

/* pbr_l6_types.hlsli:423-440 */
static float power_heuristic(float a, float b) {
    if (!fin(a) || !fin(b) || a < 0.0f || b < 0.0f) return 0.0f;
    const float m = fmaxf(a, b);
    if (!(m > 0.0f) || !fin(m)) return 0.0f;
    const float na = a / m, nb = b / m;
    const float sa = na * na, sb = nb * nb;
    return sa / (sa + sb);
}

static v3 offset_origin(v3 p, v3 gn, v3 dir, float eps) {
    const float side = dot(gn, dir) >= 0.0f ? 1.0f : -1.0f;
    return add(p, muls(gn, side * eps));
}

/* ---- the path -------------------------------------------------------------- */
#define PATH_ACTIVE 1u
#define PATH_PREV_DELTA 2u
#define PATH_PREV_SPECULAR 4u
#define PATH_PREV_RESTIR 8u
#define PATH_TERMINATED 16u
#define PATH_ERROR 32u

#define DIM_LIGHT_SELECTION 0u
#define DIM_LIGHT_SHAPE0 1u
#define DIM_BSDF_LOBE 5u
#define DIM_RR 9u

typedef struct {
    v3 beta; float eta;
    v3 betaD, betaS;
    v3 cam, dd, ds, id, is;
    v3 prevPos; float prevPdf;
    v3 prevGn, prevSn;
    uint32_t flags;
} path_state;

static uint32_t bounce_dim(uint32_t bounce, uint32_t local) { return 4u + bounce * 16u + local; }

/* WavefrontRngDimensions.hpp:7-36: per-bounce dimension of a local slot, and
 * the local slots the integrator draws (light selection, light shape 0, BSDF
 * lobe, Russian roulette). */
uint32_t ora_bounce_dim(uint32_t bounce, uint32_t local) { return bounce_dim(bounce, local); }
void ora_bounce_dim_locals(uint32_t out[4]) {
    out[0] = DIM_LIGHT_SELECTION;
    out[1] = DIM_LIGHT_SHAPE0;
    out[2] = DIM_BSDF_LOBE;
    out[3] = DIM_RR;
}

/* PbrWavefrontShading.hlsli:105-160 */
static int accumulate_terminal(path_state* st, uint32_t bounce, v3 emitted, float mis) {
    if (!fin_nonneg3(emitted) || !fin(mis) || mis < 0.0f || !fin_nonneg3(st->beta) || !fin_nonneg3(st->betaD) ||
        !fin_nonneg3(st->betaS))
        return 0;
    const v3 w = muls(emitted, mis);
    if (!fin_nonneg3(w)) return 0;
    if (bounce == 0u) {
        const v3 a = add(st->cam, mul(st->beta, w));
        if (!fin_nonneg3(a)) return 0;
        st->cam = a;
    } else if (bounce == 1u) {
        const v3 d = add(st->dd, mul(st->betaD, w));
        const v3 s = add(st->ds, mul(st->betaS, w));
        if (!fin_nonneg3(d) || !fin_nonneg3(s)) return 0;
        st->dd = d;
        st->ds = s;
    } else {
        const v3 d = add(st->id, mul(st->betaD, w));
        const v3 s = add(st->is, mul(st->betaS, w));
        if (!fin_nonneg3(d) || !fin_nonneg3(s)) return 0;
        st->id = d;
        st->is = s;
    }
    return 1;
}

typedef struct {
    const ora_node* nodes; uint32_t nodeCount;
    const ora_tri* tris; uint32_t triCount;
    const ora_shading* sh;
    const ora_frame* fr;
    int order;
    uint64_t* counters;
    const ora_wide4* wide;   /* order 2 */
} path_ctx;

static int trace_closest(const path_ctx* c, v3 o, v3 d, float tmin, float tmax, pbr_hit* hit) {
    ora_walk_hit w;
    const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    if (c->order == 2) {
        ora_trace_walk_wide(c->nodes, c->nodeCount, c->wide, c->tris, c->triCount, of, df, tmin, tmax, ORA_INVALID,
                            ORA_INVALID, 0, &w, c->counters);
    } else if (c->order == 1) {
        ora_trace_walk_near_first(c->nodes, c->nodeCount, c->tris, c->triCount, of, df, tmin, tmax, ORA_INVALID,
                                  ORA_INVALID, 0, &w, c->counters);
    } else {
        ora_trace_walk(c->nodes, c->nodeCount, c->tris, c->triCount, of, df, tmin, tmax, ORA_INVALID, ORA_INVALID,
                       0, &w, c->counters);
    }
    memset(hit, 0, sizeof(*hit));
    hit->t = w.t;
    hit->position = L3(w.position);
    hit->geometricNormal = L3(w.geometricNormal);
    hit->shadingNormal = L3(w.shadingNormal);
    hit->materialIndex = w.materialIndex;
    hit->instanceId = w.instanceId;
    hit->primitiveId = w.primitiveId;
    hit->frontFace = w.frontFace;
    hit->emitterLightIndex = w.result == 1u ? resolve_emitter(c->sh, w.instanceId, w.primitiveId) : ORA_INVALID;
    return (int)w.result;  /* 0 miss, 1 hit, 2 invalid (reported as a hit with an invalid material) */
}

static int trace_any(const path_ctx* c, v3 o, v3 d, float tmin, float tmax, uint32_t ignoredPrimitive) {
    ora_walk_hit w;
    const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    if (c->order == 2) {
        ora_trace_walk_wide(c->nodes, c->nodeCount, c->wide, c->tris, c->triCount, of, df, tmin, tmax, ORA_INVALID,
                            ignoredPrimitive, 1, &w, c->counters);
    } else if (c->order == 1) {
        ora_trace_walk_near_first(c->nodes, c->nodeCount, c->tris, c->triCount, of, df, tmin, tmax, ORA_INVALID,
                                  ignoredPrimitive, 1, &w, c->counters);
    } else {
        ora_trace_walk(c->nodes, c->nodeCount, c->tris, c->triCount, of, df, tmin, tmax, ORA_INVALID,
                       ignoredPrimitive, 1, &w, c->counters);
    }
    return w.result != 0u;
}

/* WfPbrHitIsValid, PbrWavefrontShading.hlsli:95-103 */
static int hit_valid(const pbr_hit* h, float tmin, float tmax) {
    const float g2 = dot(h->geometricNormal, h->geometricNormal);
    const float s2 = dot(h->shadingNormal, h->shadingNormal);
    return fin(h->t) && h->t >= tmin && h->t <= tmax && fin3(h->position) && fin(g2) && g2 > 0.0f && fin(s2) &&
           s2 > 0.0f && h->frontFace <= 1u;
}

static void render_path(const ora_node* nodes, uint32_t nodeCount, const ora_wide4* wide, const ora_tri* tris,
                        uint32_t triCount, const ora_shading* sh, const ora_frame* fr, int order, uint32_t pixel,
                        uint32_t sampleIndex, ora_path_result* out, uint64_t* traversalCounters,
                        ora_surface_v2* surface);

/* PbrPublishPrimarySurfaceV2 (pbr_reconstruction_export_v2.hlsli:44-131): the
 * ReSTIR copy, valid only when ReSTIR owns the surface's direct light. */
static void publish_surface(const ora_frame* fr, const pbr_hit* hit, const ora_material* mat, int owned,
                            ora_surface_v2* out) {
    const float metallic = mat->baseColorMetallic[3], roughness = mat->emissiveRoughness[3];
    const float transmission = mat->transmissionIor[0];
    const v3 base = L3(mat->baseColorMetallic), f0 = L3(mat->f0);
    const float depth = dot(sub(hit->position, L3(fr->cameraPosition)), normalize(L3(fr->cameraForward)));
    const v3 gn = normalize(hit->geometricNormal), sn = normalize(hit->shadingNormal);
    const float sm = fminf(fmaxf(metallic, 0.0f), 1.0f);
    const float st = fminf(fmaxf(transmission, 0.0f), 1.0f);
    const v3 b0 = V(fmaxf(base.x, 0.0f), fmaxf(base.y, 0.0f), fmaxf(base.z, 0.0f));
    const v3 diff = muls(muls(b0, 1.0f - sm), 1.0f - st);
    const v3 l = add(f0, muls(sub(base, f0), sm));
    const v3 spec = V(fmaxf(l.x, 0.0f), fmaxf(l.y, 0.0f), fmaxf(l.z, 0.0f));
    memset(out, 0, sizeof(*out));
    out->worldPositionLinearDepth[0] = hit->position.x;
    out->worldPositionLinearDepth[1] = hit->position.y;
    out->worldPositionLinearDepth[2] = hit->position.z;
    out->worldPositionLinearDepth[3] = fmaxf(depth, 0.0f);
    out->geometricNormalRoughness[0] = gn.x;
    out->geometricNormalRoughness[1] = gn.y;
    out->geometricNormalRoughness[2] = gn.z;
    out->geometricNormalRoughness[3] = fminf(fmaxf(roughness, 0.0f), 1.0f);
    out->shadingNormalMetallic[0] = sn.x;
    out->shadingNormalMetallic[1] = sn.y;
    out->shadingNormalMetallic[2] = sn.z;
    out->shadingNormalMetallic[3] = sm;
    out->diffuseAlbedo[0] = diff.x; out->diffuseAlbedo[1] = diff.y; out->diffuseAlbedo[2] = diff.z;
    out->specularAlbedo[0] = spec.x; out->specularAlbedo[1] = spec.y; out->specularAlbedo[2] = spec.z;
    uint32_t flags = ORA_SURFACE_VALID;
    if (diff.x > 0.0f || diff.y > 0.0f || diff.z > 0.0f) flags |= ORA_SURFACE_DIFFUSE;
    if (spec.x > 0.0f || spec.y > 0.0f || spec.z > 0.0f) flags |= ORA_SURFACE_SPECULAR;
    if (hit->frontFace != 0u) flags |= ORA_SURFACE_FRONT;
    if (!owned) flags &= ~ORA_SURFACE_VALID;
    out->identity[0] = hit->materialIndex;
    out->identity[1] = hit->instanceId;
    out->identity[2] = hit->primitiveId;
    out->identity[3] = flags;
}

static ora_wide4* wide_for(const ora_node* nodes, uint32_t nodeCount, int order) {
    if (order != 2) return NULL;
    uint32_t depth;
    ora_wide4* w = (ora_wide4*)malloc(sizeof(ora_wide4) * (nodeCount ? nodeCount : 1u));
    ora_build_wide4(nodes, nodeCount, w, &depth);
    return w;
}

void ora_render_path(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                     const ora_shading* sh, const ora_frame* fr, int order, uint32_t pixel, uint32_t sampleIndex,
                     ora_path_result* out, uint64_t* traversalCounters) {
    ora_wide4* wide = wide_for(nodes, nodeCount, order);
    render_path(nodes, nodeCount, wide, tris, triCount, sh, fr, order, pixel, sampleIndex, out, traversalCounters,
                NULL);
    free(wide);
}

static void render_path(const ora_node* nodes, uint32_t nodeCount, const ora_wide4* wide, const ora_tri* tris,
                        uint32_t triCount, const ora_shading* sh, const ora_frame* fr, int order, uint32_t pixel,
                        uint32_t sampleIndex, ora_path_result* out, uint64_t* traversalCounters,
                        ora_surface_v2* surface) {
    path_ctx c = {nodes, nodeCount, tris, triCount, sh, fr, order, traversalCounters, wide};
    memset(out, 0, sizeof(*out));
    const uint32_t tag = fr->streamTag, sl = fr->seedLow, shh = fr->seedHigh;
    const uint32_t px = pixel % fr->width, py = pixel / fr->width;

    /* RayGenCS */
    const float jx = ora_sample_dimension(pixel, sampleIndex, 0u, tag, sl, shh);
    const float jy = ora_sample_dimension(pixel, sampleIndex, 1u, tag, sl, shh);
    float sx = ((float)px + jx) / (float)fr->width;
    float sy = ((float)py + jy) / (float)fr->height;
    sx = sx * 2.0f - 1.0f;
    sy = sy * 2.0f - 1.0f;
    sy = -sy;
    v3 rayO = L3(fr->cameraPosition);
    float rayTmin = fr->rayEpsilon;
    v3 rayD = normalize(add(add(L3(fr->cameraForward), muls(L3(fr->cameraRight), sx * fr->aspect * fr->tanHalfFov)),
                            muls(L3(fr->cameraUp), sy * fr->tanHalfFov)));
    float rayTmax = 1.0e30f;

    path_state st;
    memset(&st, 0, sizeof(st));
    st.beta = V(1.0f, 1.0f, 1.0f);
    st.eta = 1.0f;
    st.flags = PATH_ACTIVE | PATH_PREV_DELTA;

    const uint32_t estimator = fr->estimator;
    /* WfClearPrimarySurfaceV2 (Shade.hlsl:62-67): misses and errors leave
     * the pixel's surface invalid. */
    if (surface) memset(surface, 0, sizeof(*surface));
    for (uint32_t bounce = 0; bounce < fr->maxBounce; ++bounce) {
        /* IntersectCS */
        pbr_hit hit;
        const int kind = trace_closest(&c, rayO, rayD, rayTmin, rayTmax, &hit);
        out->bounces++;
        int nextValid = 0;
        v3 nextO = zero3(), nextD = zero3();
        path_state next = st;

        if (kind == 0) {
            /* miss (Shade.hlsl:74-125) */
            const v3 env = environment_radiance(sh, rayD);
            float mis = 1.0f;
            if (bounce != 0u && (st.flags & PATH_PREV_DELTA) == 0u && estimator != 0u) {
                const int owned = estimator == ORA_EST_RESTIR && bounce == 1u && (st.flags & PATH_PREV_RESTIR) != 0u;
                if (owned) {
                    mis = 0.0f;
                } else {
                    const float lp = selected_light_pdf(sh, sh->environmentLight, st.prevPos, rayD, 0.0f, zero3());
                    if (lp > 0.0f) mis = power_heuristic(st.prevPdf, lp);
                }
            }
            if (!(fin3(env) && fin(mis) && accumulate_terminal(&st, bounce, env, mis))) st.flags |= PATH_ERROR;
            st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED;
            break;
        }
        /* a fail-closed traversal result carries materialIndex = INVALID */
        if (!hit_valid(&hit, rayTmin, rayTmax) || hit.materialIndex >= sh->materialCount) {
            st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED | PATH_ERROR;
            break;
        }
        const ora_material* mat = &sh->materials[hit.materialIndex];
        /* PbrMaterialAuxiliaryIsValidL6 */
        {
            const v3 em = L3(mat->emissiveRoughness);
            const float* a = mat->attenuationColorDistance;
            int ok = fin3(em) && !any_neg(em) && isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]) &&
                     isfinite(a[3]) && a[3] >= 0.0f;
            if (ok && a[3] > 0.0f && !(a[0] > 0.0f && a[1] > 0.0f && a[2] > 0.0f && a[0] <= 1.0f && a[1] <= 1.0f &&
                                       a[2] <= 1.0f))
                ok = 0;
            if (!ok) {
                st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED | PATH_ERROR;
                break;
            }
        }
        const bsdf_params params = material_params(mat);
        const int currentRestirOwned = bounce == 0u && supports_restir(&params);
        if (bounce == 0u && surface) publish_surface(fr, &hit, mat, currentRestirOwned, surface);
        v3 beta = st.beta, betaD = st.betaD, betaS = st.betaS;
        /* interior (Beer-Lambert) attenuation of a back-face hit on a
         * transmissive material (Shade.hlsl:187-205,
         * PbrInteriorAttenuationFactorL6 pbr_material_bsdf.hlsli:77-96) */
        if (hit.frontFace == 0u && mat->transmissionIor[0] > 0.0f) {
            const float* a = mat->attenuationColorDistance;
            if (a[3] > 0.0f) {
                int ok = a[0] > 0.0f && a[1] > 0.0f && a[2] > 0.0f && a[0] <= 1.0f && a[1] <= 1.0f && a[2] <= 1.0f &&
                         fin(hit.t) && hit.t >= 0.0f;
                v3 att = zero3();
                if (ok) {
                    const float e = hit.t / a[3];
                    att = V(powf(a[0], e), powf(a[1], e), powf(a[2], e));
                    ok = fin3(att);
                }
                if (ok) {
                    beta = mul(beta, att);
                    betaD = mul(betaD, att);
                    betaS = mul(betaS, att);
                    ok = fin3(beta) && fin3(betaD) && fin3(betaS);
                }
                if (!ok) {
                    st.beta = st.betaD = st.betaS = zero3();
                    st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED | PATH_ERROR;
                    break;
                }
            }
        }
        st.beta = beta;
        st.betaD = betaD;
        st.betaS = betaS;

        /* emitter MIS (Shade.hlsl:210-258) */
        const v3 emission = L3(mat->emissiveRoughness);
        int emitterVisible = hit.frontFace != 0u;
        if (!emitterVisible) {
            emitterVisible = hit.emitterLightIndex < sh->lightCount &&
                             (sh->lights[hit.emitterLightIndex].identity[1] & ORA_LIGHT_TWO_SIDED) != 0u;
        }
        if (emitterVisible && maxc(emission) > 0.0f) {
            float mis = 1.0f;
            if (bounce != 0u && (st.flags & PATH_PREV_DELTA) == 0u && estimator != 0u) {
                const int owned = estimator == ORA_EST_RESTIR && bounce == 1u && (st.flags & PATH_PREV_RESTIR) != 0u;
                if (owned) {
                    mis = 0.0f;
                } else {
                    const float lp = selected_light_pdf(sh, hit.emitterLightIndex, st.prevPos, rayD, hit.t,
                                                        hit.geometricNormal);
                    if (lp > 0.0f) mis = power_heuristic(st.prevPdf, lp);
                }
            }
            if (!accumulate_terminal(&st, bounce, emission, mis)) {
                st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED | PATH_ERROR;
                break;
            }
        }
        if ((mat->metadata[3] & 1u) != 0u || !fin3(beta) || any_neg(beta)) {
            st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED;
            break;
        }
        /* PbrBuildBsdfContextL6 */
        bsdf_ctx ctx;
        ctx.geometricNormal = hit.geometricNormal;
        ctx.shadingNormal = hit.shadingNormal;
        v3 tb;
        build_basis(hit.shadingNormal, &ctx.tangent, &tb);
        const float ior = mat->transmissionIor[1];
        ctx.etaIncident = hit.frontFace != 0u ? 1.0f : ior;
        ctx.etaTransmitted = hit.frontFace != 0u ? ior : 1.0f;
        ctx.transportMode = 0u;
        if (!validate_params(&ctx, &params)) {
            st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED | PATH_ERROR;
            break;
        }
        const v3 wo = neg(rayD);

        /* WfPbrBuildDirectShadow (PbrWavefrontShading.hlsli:173-292) */
        int shadow = 0;
        v3 shO = zero3(), shD = zero3(), shDiff = zero3(), shSpec = zero3();
        float shTmin = 0.0f, shTmax = 0.0f;
        uint32_t shPrim = ORA_INVALID;
        if (estimator != 0u && sh->lightAliasCount != 0u &&
            !(estimator == ORA_EST_RESTIR && bounce == 0u && currentRestirOwned)) {
            float rnd[5];
            rnd[0] = ora_sample_dimension(pixel, sampleIndex, bounce_dim(bounce, DIM_LIGHT_SELECTION), tag, sl, shh);
            for (uint32_t k = 0; k < 4; ++k) {
                rnd[1 + k] = ora_sample_dimension(pixel, sampleIndex, bounce_dim(bounce, DIM_LIGHT_SHAPE0 + k), tag,
                                                  sl, shh);
            }
            const light_sample ls = sample_one_light(sh, hit.position, rnd);
            if ((ls.flags & SAMPLE_VALID) != 0u) {
                const bsdf_eval ev = eval_bsdf(&ctx, &params, wo, ls.wi);
                if (ev.isValid && fin(ev.pdf) && fin3(ev.value) && !(ev.pdf < 0.0f)) {
                    const int delta = (ls.flags & SAMPLE_DELTA) != 0u;
                    const float epdf = delta ? ls.selectionPmf : ls.combinedPdfW;
                    if (fin(epdf) && epdf > 0.0f) {
                        const float cosine = fabsf(dot(hit.geometricNormal, ls.wi));
                        if (cosine > 0.0f) {
                            const float mis = delta ? 1.0f : power_heuristic(ls.combinedPdfW, ev.pdf);
                            const v3 scale = muls(ls.Li, cosine * mis / epdf);
                            const v3 diff = bounce == 0u ? mul(mul(beta, ev.diffuseValue), scale)
                                                         : mul(mul(betaD, ev.value), scale);
                            const v3 spec = bounce == 0u ? mul(mul(beta, ev.specularValue), scale)
                                                         : mul(mul(betaS, ev.value), scale);
                            if (fin3(diff) && fin3(spec) && !any_neg(diff) && !any_neg(spec)) {
                                const float smax = (ls.flags & SAMPLE_INFINITE) != 0u ? 1.0e30f
                                                                                       : ls.distance - fr->rayEpsilon;
                                if (smax > fr->rayEpsilon) {
                                    shadow = 1;
                                    shO = offset_origin(hit.position, hit.geometricNormal, ls.wi, fr->rayEpsilon);
                                    shD = ls.wi;
                                    shTmin = fr->rayEpsilon;
                                    shTmax = smax;
                                    shDiff = diff;
                                    shSpec = spec;
                                    shPrim = ls.primitiveId;
                                }
                            }
                        }
                    }
                }
            }
        }

        /* BSDF sample, RR (Shade.hlsl:307-431) */
        if (bounce + 1u < fr->maxBounce) {
            const float u0 = ora_sample_dimension(pixel, sampleIndex, bounce_dim(bounce, DIM_BSDF_LOBE), tag, sl, shh);
            const float u1 = ora_sample_dimension(pixel, sampleIndex, bounce_dim(bounce, DIM_BSDF_LOBE + 1u), tag, sl, shh);
            const float u2 = ora_sample_dimension(pixel, sampleIndex, bounce_dim(bounce, DIM_BSDF_LOBE + 2u), tag, sl, shh);
            const bsdf_sample bs = sample_bsdf(&ctx, &params, wo, u0, u1, u2);
            int valid = bs.isValid && fin(bs.pdf) && bs.pdf > 0.0f && fin3(bs.value) && fin3(bs.direction) &&
                        !any_neg(bs.value);
            v3 nb = zero3(), nd = zero3(), ns = zero3();
            float etaScale = st.eta;
            if (valid) {
                const float cosine = fabsf(dot(hit.geometricNormal, bs.direction));
                const v3 factor = muls(bs.value, cosine / bs.pdf);
                if (bounce == 0u) {
                    nd = muls(mul(beta, bs.diffuseValue), cosine / bs.pdf);
                    ns = muls(mul(beta, bs.specularValue), cosine / bs.pdf);
                } else {
                    nd = mul(betaD, factor);
                    ns = mul(betaS, factor);
                }
                nb = add(nd, ns);
                valid = fin3(nb) && fin3(nd) && fin3(ns) && !any_neg(nb);
                if ((bs.lobeFlags & LOBE_TRANSMISSION) != 0u) {
                    etaScale *= bs.eta * bs.eta;
                    valid = valid && fin(etaScale) && etaScale > 0.0f;
                }
            }
            const uint32_t nextBounce = bounce + 1u;
            const float rrS = fr->rrStart, rrMin = fr->rrMinimum, rrMax = fr->rrMaximum;
            const int rrValid = fin(rrS) && fin(rrMin) && fin(rrMax) && rrS >= 0.0f && floorf(rrS) == rrS &&
                                rrMin > 0.0f && rrMin <= rrMax && rrMax <= 1.0f;
            if (!rrValid) {
                valid = 0;
            } else if (valid && (float)nextBounce >= rrS) {
                const float mag = maxc(muls(nb, etaScale));
                if (!(mag > 0.0f) || !fin(mag)) {
                    valid = 0;
                } else {
                    const float cont = fminf(fmaxf(mag, rrMin), rrMax);
                    const float rr = ora_sample_dimension(pixel, sampleIndex, bounce_dim(bounce, DIM_RR), tag, sl, shh);
                    if (rr >= cont) {
                        valid = 0;
                    } else {
                        nb = divs(nb, cont);
                        nd = divs(nd, cont);
                        ns = divs(ns, cont);
                    }
                }
            }
            if (valid) {
                uint32_t ev = 0u;
                if (bs.isDelta) ev |= PATH_PREV_DELTA;
                if ((bs.lobeFlags & LOBE_DIFFUSE) == 0u) ev |= PATH_PREV_SPECULAR;
                if (currentRestirOwned) ev |= PATH_PREV_RESTIR;
                nextValid = 1;
                nextO = offset_origin(hit.position, hit.geometricNormal, bs.direction, fr->rayEpsilon);
                nextD = normalize(bs.direction);
                next.beta = nb;
                next.eta = etaScale;
                next.betaD = nd;
                next.betaS = ns;
                next.prevPos = hit.position;
                next.prevPdf = bs.pdf;
                next.prevGn = hit.geometricNormal;
                next.prevSn = hit.shadingNormal;
                next.flags = ev;
            } else if (!fin3(nb)) {
                st.flags |= PATH_ERROR;
            }
        }

        /* TraceShadowCS (Physical) */
        if (shadow) {
            out->shadowRays++;
            const int occluded = trace_any(&c, shO, shD, shTmin, shTmax, shPrim);
            if (!occluded) {
                v3 d, s;
                if (bounce == 0u) { d = add(st.dd, shDiff); s = add(st.ds, shSpec); }
                else { d = add(st.id, shDiff); s = add(st.is, shSpec); }
                if (!fin_nonneg3(d) || !fin_nonneg3(s)) {
                    st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED | PATH_ERROR;
                } else if (bounce == 0u) {
                    st.dd = d; st.ds = s;
                } else {
                    st.id = d; st.is = s;
                }
            }
        }

        /* NextBounceCS */
        if (!nextValid) {
            st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED;
            break;
        }
        if ((st.flags & PATH_ERROR) != 0u) break;
        {
            const float dl2 = dot(nextD, nextD);
            const float gl2 = dot(next.prevGn, next.prevGn);
            const float sl2 = dot(next.prevSn, next.prevSn);
            const int ok = fin3(nextO) && fr->rayEpsilon > 0.0f && fin3(nextD) && 1.0e30f > fr->rayEpsilon &&
                           fin(dl2) && dl2 > 0.0f && fin3(next.beta) && !any_neg(next.beta) && fin(next.eta) &&
                           next.eta > 0.0f && fin3(next.betaD) && !any_neg(next.betaD) && fin3(next.betaS) &&
                           !any_neg(next.betaS) && fin3(next.prevPos) && next.prevPdf > 0.0f && fin(next.prevPdf) &&
                           fin3(next.prevGn) && fin(gl2) && gl2 > 0.0f && fin3(next.prevSn) && fin(sl2) &&
                           sl2 > 0.0f && bounce + 1u < fr->maxBounce;
            if (!ok) {
                st.flags = (st.flags & ~PATH_ACTIVE) | PATH_TERMINATED | PATH_ERROR;
                break;
            }
        }
        st.beta = next.beta;
        st.eta = next.eta;
        st.betaD = next.betaD;
        st.betaS = next.betaS;
        st.prevPos = next.prevPos;
        st.prevPdf = next.prevPdf;
        st.prevGn = next.prevGn;
        st.prevSn = next.prevSn;
        st.flags = (st.flags & PATH_ERROR) | PATH_ACTIVE |
                   (next.flags & (PATH_PREV_DELTA | PATH_PREV_SPECULAR | PATH_PREV_RESTIR));
        rayO = nextO;
        rayD = nextD;
        rayTmin = fr->rayEpsilon;
        rayTmax = 1.0e30f;
    }

    /* ResolveCS */
    const v3 sig[5] = {st.cam, st.dd, st.ds, st.id, st.is};
    int ok = 1;
    for (int k = 0; k < 5; ++k) ok = ok && fin_nonneg3(sig[k]);
    float* dst[5] = {out->cameraEmission, out->directDiffuse, out->directSpecular, out->indirectDiffuse,
                     out->indirectSpecular};
    for (int k = 0; k < 5; ++k) {
        dst[k][0] = ok ? sig[k].x : 0.0f;
        dst[k][1] = ok ? sig[k].y : 0.0f;
        dst[k][2] = ok ? sig[k].z : 0.0f;
    }
    if (!ok) st.flags |= PATH_ERROR;
    out->flags = st.flags;
}

typedef struct {
    const ora_node* nodes; uint32_t nodeCount;
    const ora_tri* tris; uint32_t triCount;
    const ora_shading* sh; const ora_frame* fr; int order;
    const ora_wide4* wide;
    ora_path_result* results;
    ora_surface_v2* surfaces;
    uint32_t next, rowEnd, firstPixel;
    pthread_mutex_t lock;
    uint64_t counters[ORA_CNT_COUNT];
} frame_job;

static void* frame_worker(void* arg) {
    frame_job* j = (frame_job*)arg;
    uint64_t local[ORA_CNT_COUNT];
    memset(local, 0, sizeof(local));
    for (;;) {
        pthread_mutex_lock(&j->lock);
        const uint32_t row = j->next++;
        pthread_mutex_unlock(&j->lock);
        if (row >= j->rowEnd) break;
        for (uint32_t x = 0; x < j->fr->width; ++x) {
            const uint32_t pixel = row * j->fr->width + x;
            for (uint32_t s = 0; s < j->fr->spp; ++s) {
                render_path(j->nodes, j->nodeCount, j->wide, j->tris, j->triCount, j->sh, j->fr, j->order, pixel,
                            j->fr->sampleIndex + s, &j->results[(size_t)(pixel - j->firstPixel) * j->fr->spp + s],
                            local, j->surfaces ? &j->surfaces[pixel - j->firstPixel] : NULL);
            }
        }
    }
    pthread_mutex_lock(&j->lock);
    for (int k = 0; k < ORA_CNT_COUNT; ++k) j->counters[k] += local[k];
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

int ora_render_rows(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                    const ora_shading* sh, const ora_frame* fr, int order, uint32_t rowBegin, uint32_t rowCount,
                    uint32_t threads, ora_path_result* results, uint64_t* traversalCounters) {
    return ora_render_rows_surfaces(nodes, nodeCount, tris, triCount, sh, fr, order, rowBegin, rowCount, threads,
                                    results, traversalCounters, NULL);
}

int ora_render_rows_surfaces(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                             const ora_shading* sh, const ora_frame* fr, int order, uint32_t rowBegin,
                             uint32_t rowCount, uint32_t threads, ora_path_result* results,
                             uint64_t* traversalCounters, void* surfaces) {
    if (rowBegin > fr->height || rowCount > fr->height - rowBegin) return -1;
    if (surfaces && fr->spp != 1u) return -1;
    frame_job j;
    memset(&j, 0, sizeof(j));
    j.nodes = nodes; j.nodeCount = nodeCount; j.tris = tris; j.triCount = triCount;
    j.sh = sh; j.fr = fr; j.order = order; j.results = results;
    j.surfaces = (ora_surface_v2*)surfaces;
    ora_wide4* wide = wide_for(nodes, nodeCount, order);
    j.wide = wide;
    j.next = rowBegin;
    j.rowEnd = rowBegin + rowCount;
    j.firstPixel = rowBegin * fr->width;
    pthread_mutex_init(&j.lock, NULL);
    if (threads == 0) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (uint32_t i = 0; i < threads; ++i) pthread_create(&th[i], NULL, frame_worker, &j);
    for (uint32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(wide);
    pthread_mutex_destroy(&j.lock);
    if (traversalCounters) {
        for (int k = 0; k < ORA_CNT_COUNT; ++k) traversalCounters[k] += j.counters[k];
    }
    return 0;
}

int ora_render_frame(const ora_node* nodes, uint32_t nodeCount, const ora_tri* tris, uint32_t triCount,
                     const ora_shading* sh, const ora_frame* fr, int order, uint32_t threads,
                     ora_path_result* results, uint64_t* traversalCounters) {
    return ora_render_rows(nodes, nodeCount, tris, triCount, sh, fr, order, 0u, fr->height, threads, results,
                           traversalCounters);
}

void ora_film_accumulate(const ora_path_result* results, uint32_t pixelCount, uint32_t spp, float* film) {
    for (uint32_t p = 0; p < pixelCount; ++p) {
        float* f = film + 4u * p;
        for (uint32_t s = 0; s < spp; ++s) {
            const ora_path_result* r = &results[(size_t)p * spp + s];
            /* ComposeCS: output = raw.diffuse + raw.specular + terminal with
             * raw.diffuse = direct + indirect diffuse, raw.specular likewise. */
            const v3 diff = add(L3(r->directDiffuse), L3(r->indirectDiffuse));
            const v3 spec = add(L3(r->directSpecular), L3(r->indirectSpecular));
            const v3 out = add(add(diff, spec), L3(r->cameraEmission));
            const uint32_t n = (uint32_t)f[3];
            const v3 prev = n == 0u ? zero3() : L3(f);
            const v3 v = add(prev, divs(sub(out, prev), (float)(n + 1u)));
            f[0] = v.x; f[1] = v.y; f[2] = v.z; f[3] = (float)(n + 1u);
        }
    }
}

/* ---- exported hooks for the oracle's own pin tests -------------------------- */

/* out[0..2] value, [3] pdf, [4..6] diffuse, [7..9] specular, [10] valid, [11] lobes */
void ora_bsdf_eval(const ora_material* m, const float gn[3], const float sn[3], const float wo[3],
                   const float wi[3], float out[12]) {
    const bsdf_params p = material_params(m);
    bsdf_ctx c;
    c.geometricNormal = L3(gn);
    c.shadingNormal = L3(sn);
    v3 b;
    build_basis(L3(sn), &c.tangent, &b);
    c.etaIncident = 1.0f;
    c.etaTransmitted = m->transmissionIor[1];
    c.transportMode = 0u;
    const bsdf_eval e = eval_bsdf(&c, &p, L3(wo), L3(wi));
    out[0] = e.value.x; out[1] = e.value.y; out[2] = e.value.z; out[3] = e.pdf;
    out[4] = e.diffuseValue.x; out[5] = e.diffuseValue.y; out[6] = e.diffuseValue.z;
    out[7] = e.specularValue.x; out[8] = e.specularValue.y; out[9] = e.specularValue.z;
    out[10] = (float)e.isValid;
    out[11] = (float)e.lobeFlags;
}

/* out[0..2] direction, [3] pdf, [4..6] value, [7] valid, [8] lobes */
void ora_bsdf_sample(const ora_material* m, const float gn[3], const float sn[3], const float wo[3],
                     const float u[3], float out[9]) {
    const bsdf_params p = material_params(m);
    bsdf_ctx c;
    c.geometricNormal = L3(gn);
    c.shadingNormal = L3(sn);
    v3 b;
    build_basis(L3(sn), &c.tangent, &b);
    c.etaIncident = 1.0f;
    c.etaTransmitted = m->transmissionIor[1];
    c.transportMode = 0u;
    const bsdf_sample s = sample_bsdf(&c, &p, L3(wo), u[0], u[1], u[2]);
    out[0] = s.direction.x; out[1] = s.direction.y; out[2] = s.direction.z; out[3] = s.pdf;
    out[4] = s.value.x; out[5] = s.value.y; out[6] = s.value.z;
    out[7] = (float)s.isValid;
    out[8] = (float)s.lobeFlags;
}

/* As ora_bsdf_eval / ora_bsdf_sample with the context's eta on both sides
 * and the transport mode; eval out[12] measure; sample out[9] eta, [10]
 * isDelta, [11] measure. */
void ora_bsdf_eval_ctx(const ora_material* m, const float gn[3], const float sn[3], float etaI, float etaT,
                       uint32_t transport, const float wo[3], const float wi[3], float out[13]) {
    const bsdf_params p = material_params(m);
    bsdf_ctx c;
    c.geometricNormal = L3(gn);
    c.shadingNormal = L3(sn);
    v3 b;
    build_basis(L3(sn), &c.tangent, &b);
    c.etaIncident = etaI;
    c.etaTransmitted = etaT;
    c.transportMode = transport;
    const bsdf_eval e = eval_bsdf(&c, &p, L3(wo), L3(wi));
    out[0] = e.value.x; out[1] = e.value.y; out[2] = e.value.z; out[3] = e.pdf;
    out[4] = e.diffuseValue.x; out[5] = e.diffuseValue.y; out[6] = e.diffuseValue.z;
    out[7] = e.specularValue.x; out[8] = e.specularValue.y; out[9] = e.specularValue.z;
    out[10] = (float)e.isValid;
    out[11] = (float)e.lobeFlags;
    out[12] = (float)e.measure;
}

void ora_bsdf_sample_ctx(const ora_material* m, const float gn[3], const float sn[3], float etaI, float etaT,
                         uint32_t transport, const float wo[3], const float u[3], float out[12]) {
    const bsdf_params p = material_params(m);
    bsdf_ctx c;
    c.geometricNormal = L3(gn);
    c.shadingNormal = L3(sn);
    v3 b;
    build_basis(L3(sn), &c.tangent, &b);
    c.etaIncident = etaI;
    c.etaTransmitted = etaT;
    c.transportMode = transport;
    const bsdf_sample s = sample_bsdf(&c, &p, L3(wo), u[0], u[1], u[2]);
    out[0] = s.direction.x; out[1] = s.direction.y; out[2] = s.direction.z; out[3] = s.pdf;
    out[4] = s.value.x; out[5] = s.value.y; out[6] = s.value.z;
    out[7] = (float)s.isValid;
    out[8] = (float)s.lobeFlags;
    out[9] = s.eta;
    out[10] = (float)s.isDelta;
    out[11] = (float)s.measure;
}

/* The material's BSDF parameters after the rough -> smooth dielectric
 * transition: out = (model, allowed lobes). */
void ora_material_model(const ora_material* m, uint32_t out[2]) {
    const bsdf_params p = material_params(m);
    out[0] = p.model;
    out[1] = p.allowedLobes;
}

/* out[0..2] wi, [3] distance, [4..6] Li, [7] selection pmf, [8] combined pdfW, [9] flags, [10] light */
void ora_light_sample(const ora_shading* sh, const float p[3], const float rnd[5], float out[11]) {
    const light_sample s = sample_one_light(sh, L3(p), rnd);
    out[0] = s.wi.x; out[1] = s.wi.y; out[2] = s.wi.z; out[3] = s.distance;
    out[4] = s.Li.x; out[5] = s.Li.y; out[6] = s.Li.z; out[7] = s.selectionPmf;
    out[8] = s.combinedPdfW; out[9] = (float)s.flags; out[10] = (float)s.lightIndex;
}

float ora_light_pdf(const ora_shading* sh, uint32_t light, const float wi[3], float dist, const float en[3]) {
    return selected_light_pdf(sh, light, zero3(), L3(wi), dist, L3(en));
}

float ora_light_pdf_at(const ora_shading* sh, uint32_t light, const float p[3], const float wi[3], float dist,
                       const float en[3]) {
    return selected_light_pdf(sh, light, L3(p), L3(wi), dist, L3(en));
}

float ora_power_heuristic(float a, float b) { return power_heuristic(a, b); }

/* Alias-table sampling (pbr_light_sampling.hlsli:47-72); returns the item or
 * ORA_INVALID and writes its pmf. */
uint32_t ora_sample_alias(const ora_alias* table, uint32_t count, float u, float* pmf) {
    const ora_alias a = sample_alias(table, count, u);
    if (pmf) *pmf = a.pmf;
    return a.item;
}

/* The tangent frame the BSDF hooks build around a shading normal. */
void ora_build_basis(const float n[3], float t[3], float b[3]) {
    v3 tt, bb;
    build_basis(L3(n), &tt, &bb);
    t[0] = tt.x; t[1] = tt.y; t[2] = tt.z;
    b[0] = bb.x; b[1] = bb.y; b[2] = bb.z;
}

void ora_light_sample_full_eval(const ora_shading* sh, const float p[3], const float rnd[5], ora_light_sample_full* out) {
    const light_sample s = sample_one_light(sh, L3(p), rnd);
    memset(out, 0, sizeof(*out));
    out->wi[0] = s.wi.x; out->wi[1] = s.wi.y; out->wi[2] = s.wi.z;
    out->distance = s.distance;
    out->Li[0] = s.Li.x; out->Li[1] = s.Li.y; out->Li[2] = s.Li.z;
    out->selectionPmf = s.selectionPmf;
    out->conditionalPdf = s.conditionalPdf;
    out->combinedPdfW = s.combinedPdfW;
    out->lightIndex = s.lightIndex;
    out->primitiveId = s.primitiveId;
    out->flags = s.flags;
    out->position[0] = s.position.x; out->position[1] = s.position.y; out->position[2] = s.position.z;
    out->measure = (s.flags & SAMPLE_VALID) != 0u && s.lightIndex < sh->lightCount ? s.measure
                                                                                  : ORA_MEASURE_INVALID;
}

int ora_roughness_finite(float perceptualRoughness) { return roughness_finite(perceptualRoughness); }

void ora_schlick3(float cosine, const float f0[3], float out[3]) {
    const v3 f = schlick(cosine, L3(f0));
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
}

float ora_ggx_d_cos(float cosH, float alpha) { return ggx_d(V(0.0f, 0.0f, cosH), alpha); }

float ora_ggx_g2_cos(float cosO, float cosI, float alpha) {
    const v3 wo = V(sqrtf(fmaxf(0.0f, 1.0f - cosO * cosO)), 0.0f, cosO);
    const v3 wi = V(sqrtf(fmaxf(0.0f, 1.0f - cosI * cosI)), 0.0f, cosI);
    return ggx_g2(wo, wi, alpha);
}
