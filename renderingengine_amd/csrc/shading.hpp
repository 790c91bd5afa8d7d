// Device PBR shading library: Philox sampling, BSDF models 0 (Lambert) and 5
// (metallic-roughness, transmission 0), light sampling and MIS.
//
// Restates (paths relative to the RenderingEngine tree):
//   Philox / dimensions  integrators/wavefront/shaders/include/WavefrontRng.hlsli:28-97
//   BSDF                 resources/shaders/include/bsdf/PbrBsdf.hlsli:184-196, 239-367,
//                        369-496, 543-549, 594-956, 1115-1229, 1261-1269, 1390-1435,
//                        1569-1677
//   material bridge      resources/shaders/integrators/pbr_material_bsdf.hlsli:8-58
//   lights               resources/shaders/integrators/pbr_light_sampling.hlsli:33-72,
//                        160-307, 422-589, 591-724
//   power heuristic      resources/shaders/integrators/pbr_l6_types.hlsli:423-440
//
// Numeric conventions (identical in the CPU oracle): dot = (x*x + y*y) + z*z,
// rsqrt(x) = 1/sqrt(x), normalize(v) = v * rsqrt(dot(v, v)), lerp(a, b, t) =
// a + (b - a) * t, sin/cos(2*pi*u) by sincos_2pi (exact quadrant reduction of
// 4u plus Taylor polynomials), no FMA contraction.  Only correctly rounded
// + - * / sqrt are used, so results are bit-identical to the oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rex_layout.hpp"
#include "render_layout.hpp"

namespace rex {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 v3(const F4& a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float rsq(float x) { return 1.0f / sqrtf(x); }
__device__ __forceinline__ V3 normalize(V3 a) { return a * rsq(dot(a, a)); }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ bool fin(float a) { return __builtin_isfinite(a); }
__device__ __forceinline__ bool fin3(V3 a) { return fin(a.x) && fin(a.y) && fin(a.z); }
__device__ __forceinline__ bool any_neg(V3 a) { return a.x < 0.0f || a.y < 0.0f || a.z < 0.0f; }
__device__ __forceinline__ bool fin_nonneg3(V3 a) { return fin3(a) && !any_neg(a); }
__device__ __forceinline__ float maxc(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ float lum(V3 a) { return dot(a, v3(0.2126f, 0.7152f, 0.0722f)); }
__device__ __forceinline__ V3 zero3() { return {0.0f, 0.0f, 0.0f}; }

// ---- Philox4x32-10 -----------------------------------------------------------
__device__ __forceinline__ uint32_t philox_x(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                             uint32_t k1) {
    // Each round's two 32x32 -> 64-bit products as one 64-bit multiply
    // (v_mad_u64_u32 gives both halves in one instruction).
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xd2511f53u * c0;
        const uint64_t p1 = (uint64_t)0xcd9e8d57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9e3779b9u;
        k1 += 0xbb67ae85u;
    }
    return c0;
}

// Top 24 bits of a Philox word as a float in [0, 1).
__device__ __forceinline__ float unit_float(uint32_t bits) { return (float)(bits >> 8) * (1.0f / 16777216.0f); }

__device__ __forceinline__ float sample_dim(uint32_t pixel, uint32_t sample, uint32_t dim, const FrameConst& f) {
    return unit_float(philox_x(pixel, sample, dim, f.streamTag, f.seedLow, f.seedHigh));
}

__device__ __forceinline__ uint32_t bounce_dim(uint32_t bounce, uint32_t local) { return 4u + bounce * 16u + local; }

// cos / sin of 2*pi*u.
__device__ __forceinline__ void sincos_2pi(float u, float& s, float& c) {
    const float x = u * 4.0f;
    const float q = floorf(x + 0.5f);
    const float r = x - q;
    const float r2 = r * r;
    const float sp = r * (1.5707963267948966f +
                          r2 * (-0.6459640975062462f +
                                r2 * (0.07969262624616703f + r2 * (-0.004681754135318687f + r2 * 0.00016044118478735982f))));
    const float cp = 1.0f + r2 * (-1.2337005501361697f +
                                  r2 * (0.25366950790104797f +
                                        r2 * (-0.020863480763352958f +
                                              r2 * (0.0009192602748394263f + r2 * -2.5202042373060605e-05f))));
    const int quadrant = ((int)q) & 3;
    if (quadrant == 0) { s = sp; c = cp; }
    else if (quadrant == 1) { s = cp; c = -sp; }
    else if (quadrant == 2) { s = -sp; c = -cp; }
    else { s = -cp; c = sp; }
}

// ---- BSDF --------------------------------------------------------------------
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kDirEps = 1.0e-7f;
constexpr float kMinAlpha = 1.0e-4f;

struct BsdfCtx {
    V3 gn, sn, tangent;
    float etaI, etaT;
};

struct BsdfParams {
    V3 base;
    float metallic;
    V3 f0;
    float rough;
    V3 conductorEta, conductorK;
    float transmission;
    uint32_t model, lobes, complexFresnel;
};

struct BsdfEval {
    V3 value;
    float pdf;
    V3 diffuse, specular;
    uint32_t lobeFlags;
    bool valid;
};

struct BsdfSample {
    V3 dir;
    float pdf;
    V3 value, diffuse, specular;
    float eta;          // relative eta of a transmission event (1 otherwise)
    uint32_t lobeFlags;
    bool delta;         // a discrete (specular) event: pdf is a probability
    bool valid;
};

__device__ __forceinline__ BsdfEval eval_invalid() {
    BsdfEval e;
    e.value = e.diffuse = e.specular = zero3();
    e.pdf = 0.0f;
    e.lobeFlags = 0u;
    e.valid = false;
    return e;
}

__device__ __forceinline__ bool safe_normalize(V3 v, V3& out) {
    out = zero3();
    const float l2 = dot(v, v);
    if (!fin(l2) || l2 <= kDirEps * kDirEps) return false;
    out = v * rsq(l2);
    return fin3(out);
}

__device__ __forceinline__ bool color_unit(V3 c) {
    return fin3(c) && !any_neg(c) && c.x <= 1.0f && c.y <= 1.0f && c.z <= 1.0f;
}

__device__ __forceinline__ bool roughness_finite(float r) {
    if (!fin(r) || r <= 0.0f || r > 1.0f) return false;
    return r * r >= kMinAlpha;
}

__device__ __forceinline__ bool has_flags(uint32_t mask, uint32_t req) { return (mask & req) == req; }
__device__ __forceinline__ bool unit_range(float t) { return fin(t) && t >= 0.0f && t <= 1.0f; }

// ValidateBsdfParamsL6 (PbrBsdf.hlsli:239-352).  Kernels built without kExt
// take models 0 and 5 only; the host routes every other material to the
// kExt build.
template <bool kExt>
__device__ __forceinline__ bool validate_params(const BsdfCtx& c, const BsdfParams& p) {
    if (p.lobes == 0u) return false;
    if (p.model == kModelLambert) return color_unit(p.base);
    if (p.model == kModelMetallicRoughness) {
        return color_unit(p.base) && color_unit(p.f0) && roughness_finite(p.rough) && unit_range(p.metallic) &&
               unit_range(p.transmission);
    }
    if (!kExt) return false;
    const bool etaOk = fin(c.etaI) && fin(c.etaT) && c.etaI > 0.0f && c.etaT > 0.0f;
    switch (p.model) {
    case kModelConductor:
        if (!roughness_finite(p.rough)) return false;
        if (p.complexFresnel != 0u) {
            return fin3(p.conductorEta) && fin3(p.conductorK) && p.conductorEta.x > 0.0f && p.conductorEta.y > 0.0f &&
                   p.conductorEta.z > 0.0f && !any_neg(p.conductorK);
        }
        return color_unit(p.f0);
    case kModelDielectricReflection: return roughness_finite(p.rough) && etaOk;
    case kModelSmoothGlass: return etaOk && unit_range(p.transmission);
    case kModelRoughDielectric: return roughness_finite(p.rough) && etaOk && unit_range(p.transmission);
    default: return false;
    }
}

__device__ __forceinline__ bool supports_restir(const BsdfParams& p) {
    const uint32_t required = kLobeDiffuse | kLobeReflection | kLobeGlossy;
    const uint32_t unsupported = kLobeTransmission | kLobeSpecular;
    return p.model == kModelMetallicRoughness && p.complexFresnel == 0u && roughness_finite(p.rough) &&
           fin(p.transmission) && p.transmission == 0.0f && (p.lobes & required) == required &&
           (p.lobes & unsupported) == 0u;
}

struct Frame {
    V3 gn, sn, t, b, woW, woL;
};

__device__ __forceinline__ bool build_frame(const BsdfCtx& c, V3 woIn, Frame& f) {
    if (!safe_normalize(c.gn, f.gn) || !safe_normalize(c.sn, f.sn) || !safe_normalize(woIn, f.woW)) return false;
    if (dot(f.sn, f.gn) < 0.0f) f.sn = -f.sn;
    if (dot(f.woW, f.gn) < 0.0f) {
        f.gn = -f.gn;
        f.sn = -f.sn;
    }
    if (dot(f.woW, f.gn) <= kDirEps || dot(f.woW, f.sn) <= kDirEps) return false;
    const V3 tc = c.tangent - f.sn * dot(c.tangent, f.sn);
    if (!safe_normalize(tc, f.t)) {
        const V3 helper = fabsf(f.sn.z) < 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(0.0f, 1.0f, 0.0f);
        if (!safe_normalize(cross(helper, f.sn), f.t)) return false;
    }
    if (!safe_normalize(cross(f.sn, f.t), f.b)) return false;
    f.woL = v3(dot(f.woW, f.t), dot(f.woW, f.b), dot(f.woW, f.sn));
    return fin3(f.woL) && f.woL.z > kDirEps;
}

__device__ __forceinline__ V3 to_local(V3 d, const Frame& f) { return v3(dot(d, f.t), dot(d, f.b), dot(d, f.sn)); }
__device__ __forceinline__ V3 to_world(V3 d, const Frame& f) { return f.t * d.x + f.b * d.y + f.sn * d.z; }

__device__ __forceinline__ bool validate_world_event(V3 wo, V3 wi, V3 gn, uint32_t lobes) {
    const float side = dot(wo, gn) * dot(wi, gn);
    const bool refl = (lobes & kLobeReflection) == kLobeReflection;
    const bool trans = (lobes & kLobeTransmission) == kLobeTransmission;
    if (refl == trans) return false;
    return refl ? side > 0.0f : side < 0.0f;
}

__device__ __forceinline__ bool shading_correction(V3 gn, V3 sn, V3 wi, float& corr) {
    corr = 0.0f;
    const float num = fabsf(dot(wi, sn));
    const float den = fabsf(dot(wi, gn));
    if (!fin(num) || !fin(den) || den <= kDirEps) return false;
    corr = num / den;
    return fin(corr) && corr >= 0.0f;
}

__device__ __forceinline__ V3 schlick(float cosine, V3 f0) {
    const float om = 1.0f - fminf(fabsf(cosine), 1.0f);
    const float f = om * om * om * om * om;
    return f0 + (v3(1.0f, 1.0f, 1.0f) - f0) * f;
}

__device__ __forceinline__ float ggx_d(V3 m, float alpha) {
    if (m.z <= 0.0f || alpha < kMinAlpha) return 0.0f;
    const float a2 = alpha * alpha;
    const float c2 = m.z * m.z;
    const float den = c2 * (a2 - 1.0f) + 1.0f;
    if (den <= 0.0f) return 0.0f;
    return a2 / (kPi * den * den);
}
__device__ __forceinline__ float ggx_lambda(V3 d, float alpha) {
    const float ac = fabsf(d.z);
    if (ac <= kDirEps) return 1.0e30f;
    const float s2 = fmaxf(0.0f, 1.0f - ac * ac);
    const float t2 = s2 / (ac * ac);
    return 0.5f * (sqrtf(1.0f + alpha * alpha * t2) - 1.0f);
}
__device__ __forceinline__ float ggx_g1(V3 d, float alpha) { return 1.0f / (1.0f + ggx_lambda(d, alpha)); }
__device__ __forceinline__ float ggx_g2(V3 wo, V3 wi, float alpha) {
    return 1.0f / (1.0f + ggx_lambda(wo, alpha) + ggx_lambda(wi, alpha));
}
__device__ __forceinline__ float ggx_vndf_pdf(V3 wo, V3 m, float alpha) {
    if (wo.z <= kDirEps || m.z <= 0.0f || dot(wo, m) <= 0.0f) return 0.0f;
    return ggx_d(m, alpha) * ggx_g1(wo, alpha) * fabsf(dot(wo, m)) / fabsf(wo.z);
}

__device__ __forceinline__ bool ggx_sample_vndf(V3 wo, float alpha, float u0, float u1, V3& m) {
    m = zero3();
    if (wo.z <= kDirEps || alpha < kMinAlpha || !fin3(wo) || !fin(u0) || !fin(u1) || u0 < 0.0f || u1 < 0.0f ||
        u0 >= 1.0f || u1 >= 1.0f)
        return false;
    V3 sv;
    if (!safe_normalize(v3(alpha * wo.x, alpha * wo.y, wo.z), sv)) return false;
    if (sv.z < 0.0f) sv = -sv;
    V3 t1;
    const float lens2 = sv.x * sv.x + sv.y * sv.y;
    if (lens2 > kDirEps * kDirEps) t1 = v3(-sv.y, sv.x, 0.0f) * rsq(lens2);
    else t1 = v3(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(sv, t1);
    const float radius = sqrtf(u0);
    float sp, cp;
    sincos_2pi(u1, sp, cp);
    const float dx = radius * cp;
    float dy = radius * sp;
    const float blend = 0.5f * (1.0f + sv.z);
    const float base = sqrtf(fmaxf(0.0f, 1.0f - dx * dx));
    dy = base + (dy - base) * blend;
    const float pz = sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy));
    const V3 h = t1 * dx + t2 * dy + sv * pz;
    if (!safe_normalize(v3(alpha * h.x, alpha * h.y, h.z), m)) return false;
    return m.z > 0.0f && dot(wo, m) > 0.0f;
}

__device__ __forceinline__ bool ggx_terms(V3 wo, V3 wi, float alpha, V3& m, float& D, float& G, float& pdf) {
    m = zero3();
    D = G = pdf = 0.0f;
    if (wo.z <= kDirEps || wi.z <= kDirEps || !safe_normalize(wo + wi, m)) return false;
    if (m.z < 0.0f) m = -m;
    const float wom = fabsf(dot(wo, m));
    if (wom <= kDirEps) return false;
    D = ggx_d(m, alpha);
    G = ggx_g2(wo, wi, alpha);
    pdf = ggx_vndf_pdf(wo, m, alpha) / (4.0f * wom);
    return D > 0.0f && G >= 0.0f && pdf > 0.0f && fin(pdf);
}

// BsdfFresnelDielectricL6 (PbrBsdf.hlsli:498-541).
__device__ __forceinline__ bool fresnel_dielectric(float cosIn, float etaI, float etaT, float& F) {
    F = 0.0f;
    if (!fin(cosIn) || !fin(etaI) || !fin(etaT) || etaI <= 0.0f || etaT <= 0.0f) return false;
    const float ci = fminf(fabsf(cosIn), 1.0f);
    const float ratio = etaI / etaT;
    const float st2 = ratio * ratio * fmaxf(0.0f, 1.0f - ci * ci);
    if (st2 >= 1.0f) {
        F = 1.0f;
        return true;
    }
    const float ct = sqrtf(fmaxf(0.0f, 1.0f - st2));
    const float parDen = etaT * ci + etaI * ct;
    const float perDen = etaI * ci + etaT * ct;
    if (fabsf(parDen) <= kDirEps || fabsf(perDen) <= kDirEps) return false;
    const float par = (etaT * ci - etaI * ct) / parDen;
    const float per = (etaI * ci - etaT * ct) / perDen;
    F = 0.5f * (par * par + per * per);
    return fin(F) && F >= 0.0f && F <= 1.0f;
}

// BsdfFresnelConductorL6 (PbrBsdf.hlsli:551-592), one channel.
__device__ __forceinline__ bool fresnel_conductor1(float ci, float c2, float s2, float eta, float k, float& F) {
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
    if (fabsf(rsDen) <= kDirEps || fabsf(rpDen) <= kDirEps) return false;
    const float rs = (t1 - t2) / rsDen;
    const float rp = rs * (t3 - t4) / rpDen;
    F = 0.5f * (rs + rp);
    return true;
}

__device__ __forceinline__ bool fresnel_conductor(float cosIn, V3 eta, V3 k, V3& F) {
    F = zero3();
    if (!fin(cosIn) || !fin3(eta) || !fin3(k) || eta.x <= 0.0f || eta.y <= 0.0f || eta.z <= 0.0f || any_neg(k))
        return false;
    const float ci = fminf(fabsf(cosIn), 1.0f);
    const float c2 = ci * ci;
    const float s2 = fmaxf(0.0f, 1.0f - c2);
    float fx, fy, fz;
    const bool okx = fresnel_conductor1(ci, c2, s2, eta.x, k.x, fx);
    const bool oky = fresnel_conductor1(ci, c2, s2, eta.y, k.y, fy);
    const bool okz = fresnel_conductor1(ci, c2, s2, eta.z, k.z, fz);
    if (!okx || !oky || !okz) return false;
    F = v3(fx, fy, fz);
    return fin3(F) && !any_neg(F) && F.x <= 1.0f && F.y <= 1.0f && F.z <= 1.0f;
}

// BsdfReflectionFresnelL6 (PbrBsdf.hlsli:745-781).
__device__ __forceinline__ bool reflection_fresnel(const BsdfCtx& c, const BsdfParams& p, float cosine, V3& F) {
    F = zero3();
    if (p.model == kModelConductor) {
        if (p.complexFresnel != 0u) return fresnel_conductor(cosine, p.conductorEta, p.conductorK, F);
        F = schlick(cosine, p.f0);
        return fin3(F);
    }
    if (p.model == kModelDielectricReflection) {
        float f;
        if (!fresnel_dielectric(cosine, c.etaI, c.etaT, f)) return false;
        F = v3(f, f, f);
        return true;
    }
    F = schlick(cosine, p.f0);
    return fin3(F);
}

// BsdfEvaluateGgxReflectionLocalL6 (PbrBsdf.hlsli:808-856), models 1 and 2.
__device__ __forceinline__ BsdfEval eval_ggx_reflection_local(const BsdfCtx& c, const BsdfParams& p, V3 wo, V3 wi) {
    if (!has_flags(p.lobes, kLobeGlossy | kLobeReflection)) return eval_invalid();
    const float alpha = p.rough * p.rough;
    V3 m;
    float D, G, pdf;
    if (!ggx_terms(wo, wi, alpha, m, D, G, pdf)) return eval_invalid();
    V3 F;
    if (!reflection_fresnel(c, p, fabsf(dot(wo, m)), F)) return eval_invalid();
    const float den = 4.0f * fabsf(wo.z * wi.z);
    if (den <= kDirEps) return eval_invalid();
    BsdfEval r = eval_invalid();
    r.specular = F * (D * G / den);
    r.value = r.specular;
    r.pdf = pdf;
    r.lobeFlags = kLobeGlossy | kLobeReflection;
    r.valid = fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.valid) return eval_invalid();
    return r;
}

// BsdfRoughDielectricHalfVectorL6 / EventProbabilitiesL6 (PbrBsdf.hlsli:958-1004).
__device__ __forceinline__ bool rough_half_vector(V3 wo, V3 wi, float etaRel, bool& isRefl, V3& m) {
    isRefl = wo.z * wi.z > 0.0f;
    m = zero3();
    const V3 h = isRefl ? wo + wi : wo + wi * etaRel;
    if (fabsf(wo.z) <= kDirEps || fabsf(wi.z) <= kDirEps || !safe_normalize(h, m)) return false;
    if (m.z < 0.0f) m = -m;
    return dot(m, wi) * wi.z >= 0.0f && dot(m, wo) * wo.z >= 0.0f;
}

__device__ __forceinline__ bool rough_event_probabilities(const BsdfParams& p, float F, float& pr, float& pt) {
    pr = has_flags(p.lobes, kLobeGlossy | kLobeReflection) ? F : 0.0f;
    pt = has_flags(p.lobes, kLobeGlossy | kLobeTransmission) ? (1.0f - F) * p.transmission : 0.0f;
    const float total = pr + pt;
    if (!fin(total) || total <= 0.0f) return false;
    pr /= total;
    pt /= total;
    return true;
}

// BsdfEvaluateRoughDielectricLocalL6 (PbrBsdf.hlsli:1006-1112), radiance transport.
__device__ __forceinline__ BsdfEval eval_rough_dielectric_local(const BsdfCtx& c, const BsdfParams& p, V3 wo,
                                                                V3 wi) {
    const float etaRel = c.etaT / c.etaI;
    if (!fin(etaRel) || etaRel <= 0.0f) return eval_invalid();
    bool isRefl;
    V3 m;
    if (!rough_half_vector(wo, wi, etaRel, isRefl, m)) return eval_invalid();
    float F;
    if (!fresnel_dielectric(dot(wo, m), c.etaI, c.etaT, F)) return eval_invalid();
    float pr, pt;
    if (!rough_event_probabilities(p, F, pr, pt)) return eval_invalid();
    const float alpha = p.rough * p.rough;
    const float D = ggx_d(m, alpha);
    const float G = ggx_g2(wo, wi, alpha);
    const float npdf = ggx_vndf_pdf(wo, m, alpha);
    if (D <= 0.0f || G < 0.0f || npdf <= 0.0f) return eval_invalid();
    BsdfEval r = eval_invalid();
    if (isRefl) {
        if (pr <= 0.0f) return eval_invalid();
        const float wom = fabsf(dot(wo, m));
        const float den = 4.0f * fabsf(wo.z * wi.z);
        if (wom <= kDirEps || den <= kDirEps) return eval_invalid();
        const float v = F * (D * G / den);
        r.specular = v3(v, v, v);
        r.pdf = npdf / (4.0f * wom) * pr;
        r.lobeFlags = kLobeGlossy | kLobeReflection;
    } else {
        if (pt <= 0.0f) return eval_invalid();
        const float wom = dot(wo, m);
        const float wim = dot(wi, m);
        const float term = wim + wom / etaRel;
        const float jden = term * term;
        const float cp = wo.z * wi.z;
        if (jden <= kDirEps || fabsf(cp) <= kDirEps) return eval_invalid();
        const float dmdwi = fabsf(wim) / jden;
        float tv = p.transmission * (1.0f - F) * D * G * fabsf(wim * wom / (cp * jden));
        tv /= etaRel * etaRel;
        r.specular = v3(tv, tv, tv);
        r.pdf = npdf * dmdwi * pt;
        r.lobeFlags = kLobeGlossy | kLobeTransmission;
    }
    r.value = r.specular;
    r.valid = fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.valid) return eval_invalid();
    return r;
}

__device__ __forceinline__ BsdfEval eval_lambert_local(const BsdfParams& p, V3 wo, V3 wi) {
    if ((p.lobes & (kLobeDiffuse | kLobeReflection)) != (kLobeDiffuse | kLobeReflection) || wo.z <= kDirEps ||
        wi.z <= kDirEps)
        return eval_invalid();
    BsdfEval r = eval_invalid();
    r.diffuse = p.base * kInvPi;
    r.value = r.diffuse;
    r.pdf = wi.z * kInvPi;
    r.lobeFlags = kLobeDiffuse | kLobeReflection;
    r.valid = fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.valid) return eval_invalid();
    return r;
}

__device__ __forceinline__ bool mr_lobe_probabilities(const BsdfParams& p, V3 wo, float& pd, float& ps) {
    pd = ps = 0.0f;
    const bool allowD = (p.lobes & (kLobeDiffuse | kLobeReflection)) == (kLobeDiffuse | kLobeReflection);
    const bool allowS = (p.lobes & (kLobeGlossy | kLobeReflection)) == (kLobeGlossy | kLobeReflection);
    const float dscale = (1.0f - p.metallic) * (1.0f - p.transmission);
    const float dw = allowD ? lum(p.base) * dscale : 0.0f;
    const float sw = allowS ? lum(schlick(fabsf(wo.z), p.f0)) : 0.0f;
    const float tw = dw + sw;
    if (!fin(tw) || tw <= 0.0f) return false;
    pd = dw / tw;
    ps = sw / tw;
    return fin(pd) && fin(ps);
}

__device__ __forceinline__ BsdfEval eval_mr_local(const BsdfParams& p, V3 wo, V3 wi) {
    if (wo.z <= kDirEps || wi.z <= kDirEps) return eval_invalid();
    float pd, ps;
    if (!mr_lobe_probabilities(p, wo, pd, ps)) return eval_invalid();
    const float alpha = p.rough * p.rough;
    V3 m;
    float D, G, spdf;
    const bool validSpec = ggx_terms(wo, wi, alpha, m, D, G, spdf);
    if (ps > 0.0f && !validSpec) return eval_invalid();
    const V3 F = validSpec ? schlick(fabsf(dot(wo, m)), p.f0) : zero3();
    const float dscale = (1.0f - p.metallic) * (1.0f - p.transmission);
    BsdfEval r = eval_invalid();
    if (pd > 0.0f) {
        r.diffuse = (v3(1.0f, 1.0f, 1.0f) - F) * p.base * (dscale * kInvPi);
        r.lobeFlags |= kLobeDiffuse | kLobeReflection;
    }
    if (ps > 0.0f) {
        const float den = 4.0f * fabsf(wo.z * wi.z);
        if (den <= kDirEps) return eval_invalid();
        r.specular = F * (D * G / den);
        r.lobeFlags |= kLobeGlossy | kLobeReflection;
    }
    r.value = r.diffuse + r.specular;
    r.pdf = pd * wi.z * kInvPi + ps * spdf;
    r.valid = fin3(r.value) && fin3(r.diffuse) && fin3(r.specular) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.valid) return eval_invalid();
    return r;
}

// EvaluateBsdfL6 for a frame already built from (context, wo).
template <bool kExt>
__device__ __forceinline__ BsdfEval eval_bsdf_frame(const BsdfCtx& c, const BsdfParams& p, const Frame& f, V3 wiIn) {
    V3 wiW;
    if (!safe_normalize(wiIn, wiW)) return eval_invalid();
    const V3 wiL = to_local(wiW, f);
    BsdfEval r;
    if (p.model == kModelLambert) r = eval_lambert_local(p, f.woL, wiL);
    else if (!kExt || p.model == kModelMetallicRoughness) r = eval_mr_local(p, f.woL, wiL);
    else if (p.model == kModelConductor || p.model == kModelDielectricReflection)
        r = eval_ggx_reflection_local(c, p, f.woL, wiL);
    else if (p.model == kModelRoughDielectric) r = eval_rough_dielectric_local(c, p, f.woL, wiL);
    else return eval_invalid();  // smooth glass: a delta lobe has no finite value
    if (!r.valid || !validate_world_event(f.woW, wiW, f.gn, r.lobeFlags)) return eval_invalid();
    float corr;
    if (!shading_correction(f.gn, f.sn, wiW, corr)) return eval_invalid();
    r.value = r.value * corr;
    r.diffuse = r.diffuse * corr;
    r.specular = r.specular * corr;
    if (!fin3(r.value) || !fin3(r.diffuse) || !fin3(r.specular)) return eval_invalid();
    return r;
}

template <bool kExt>
__device__ __forceinline__ BsdfEval eval_bsdf(const BsdfCtx& c, const BsdfParams& p, V3 woIn, V3 wiIn) {
    if (!validate_params<kExt>(c, p)) return eval_invalid();
    Frame f;
    if (!build_frame(c, woIn, f)) return eval_invalid();
    return eval_bsdf_frame<kExt>(c, p, f, wiIn);
}

__device__ __forceinline__ V3 cosine_hemisphere(float u0, float u1) {
    const float radius = sqrtf(u0);
    float sp, cp;
    sincos_2pi(u1, sp, cp);
    return v3(radius * cp, radius * sp, sqrtf(fmaxf(0.0f, 1.0f - u0)));
}

__device__ __forceinline__ BsdfSample sample_invalid() {
    BsdfSample s;
    s.dir = s.value = s.diffuse = s.specular = zero3();
    s.pdf = 0.0f;
    s.eta = 1.0f;
    s.lobeFlags = 0u;
    s.delta = false;
    s.valid = false;
    return s;
}

// BsdfRefractLocalL6 (PbrBsdf.hlsli:1271-1302).
__device__ __forceinline__ bool refract_local(V3 wo, V3 mIn, float etaRel, V3& wi) {
    wi = zero3();
    V3 m = mIn;
    float ci = dot(m, wo);
    if (ci < 0.0f) {
        m = -m;
        ci = -ci;
    }
    if (ci <= kDirEps || !fin(etaRel) || etaRel <= 0.0f) return false;
    const float st2 = fmaxf(0.0f, 1.0f - ci * ci) / (etaRel * etaRel);
    if (st2 >= 1.0f) return false;
    const float ct = sqrtf(fmaxf(0.0f, 1.0f - st2));
    const V3 cand = (-wo) / etaRel + m * (ci / etaRel - ct);
    return safe_normalize(cand, wi);
}

// BsdfSampleSmoothGlassLocalL6 (PbrBsdf.hlsli:1304-1388), radiance transport.
__device__ __forceinline__ BsdfSample sample_smooth_glass_local(const BsdfCtx& c, const BsdfParams& p, V3 wo,
                                                                float u) {
    float F;
    if (!fresnel_dielectric(wo.z, c.etaI, c.etaT, F)) return sample_invalid();
    const float rw = has_flags(p.lobes, kLobeSpecular | kLobeReflection) ? F : 0.0f;
    const float tw = has_flags(p.lobes, kLobeSpecular | kLobeTransmission) ? (1.0f - F) * p.transmission : 0.0f;
    const float total = rw + tw;
    if (!fin(total) || total <= 0.0f) return sample_invalid();
    const float pr = rw / total, pt = tw / total;
    BsdfSample r = sample_invalid();
    r.delta = true;
    if (u < pr) {
        r.dir = v3(-wo.x, -wo.y, wo.z);
        const float cosine = fabsf(r.dir.z);
        if (cosine <= kDirEps || pr <= 0.0f) return sample_invalid();
        r.pdf = pr;
        r.value = v3(F / cosine, F / cosine, F / cosine);
        r.lobeFlags = kLobeSpecular | kLobeReflection;
    } else {
        const float etaRel = c.etaT / c.etaI;
        if (pt <= 0.0f || !refract_local(wo, v3(0.0f, 0.0f, 1.0f), etaRel, r.dir)) return sample_invalid();
        const float cosine = fabsf(r.dir.z);
        if (cosine <= kDirEps) return sample_invalid();
        r.pdf = pt;
        float v = (1.0f - F) * p.transmission / cosine;
        v /= etaRel * etaRel;
        r.value = v3(v, v, v);
        r.eta = etaRel;
        r.lobeFlags = kLobeSpecular | kLobeTransmission;
    }
    r.specular = r.value;
    r.valid = fin3(r.dir) && fin3(r.value) && fin(r.pdf) && r.pdf > 0.0f;
    if (!r.valid) return sample_invalid();
    return r;
}

// SampleBsdfL6 (PbrBsdf.hlsli:1569-1677) with BsdfSampleFiniteLocalL6
// (:1390-1567); radiance transport.
template <bool kExt>
__device__ __forceinline__ BsdfSample sample_bsdf(const BsdfCtx& c, const BsdfParams& p, V3 woIn, float u0, float u1,
                                                  float u2) {
    const bool unit = fin(u0) && fin(u1) && fin(u2) && u0 >= 0.0f && u1 >= 0.0f && u2 >= 0.0f && u0 < 1.0f &&
                      u1 < 1.0f && u2 < 1.0f;
    if (!validate_params<kExt>(c, p) || !unit) return sample_invalid();
    Frame f;
    if (!build_frame(c, woIn, f)) return sample_invalid();
    V3 dirL;
    uint32_t lobes;
    float eta = 1.0f;
    const float alpha = p.rough * p.rough;
    if (kExt && p.model == kModelSmoothGlass) {
        BsdfSample ls = sample_smooth_glass_local(c, p, f.woL, u0);
        if (!ls.valid) return sample_invalid();
        V3 wiW;
        if (!safe_normalize(to_world(ls.dir, f), wiW) || !validate_world_event(f.woW, wiW, f.gn, ls.lobeFlags))
            return sample_invalid();
        float corr;
        if (!shading_correction(f.gn, f.sn, wiW, corr)) return sample_invalid();
        ls.dir = wiW;
        ls.value = ls.value * corr;
        ls.diffuse = ls.diffuse * corr;
        ls.specular = ls.specular * corr;
        ls.valid = fin3(ls.value) && fin3(ls.specular);
        return ls.valid ? ls : sample_invalid();
    }
    if (p.model == kModelLambert) {
        if (!has_flags(p.lobes, kLobeDiffuse | kLobeReflection)) return sample_invalid();
        dirL = cosine_hemisphere(u1, u2);
        lobes = kLobeDiffuse | kLobeReflection;
    } else if (!kExt || p.model == kModelMetallicRoughness) {
        float pd, ps;
        if (!mr_lobe_probabilities(p, f.woL, pd, ps)) return sample_invalid();
        if (u0 < pd) {
            dirL = cosine_hemisphere(u1, u2);
            lobes = kLobeDiffuse | kLobeReflection;
        } else {
            V3 m;
            if (ps <= 0.0f || !ggx_sample_vndf(f.woL, alpha, u1, u2, m)) return sample_invalid();
            dirL = -f.woL + m * (2.0f * dot(f.woL, m));
            if (dirL.z <= kDirEps) return sample_invalid();
            lobes = kLobeGlossy | kLobeReflection;
        }
    } else if (p.model == kModelConductor || p.model == kModelDielectricReflection) {
        if (!has_flags(p.lobes, kLobeGlossy | kLobeReflection)) return sample_invalid();
        V3 m;
        if (!ggx_sample_vndf(f.woL, alpha, u1, u2, m)) return sample_invalid();
        dirL = -f.woL + m * (2.0f * dot(f.woL, m));
        if (dirL.z <= kDirEps) return sample_invalid();
        lobes = kLobeGlossy | kLobeReflection;
    } else if (p.model == kModelRoughDielectric) {
        V3 m;
        if (!ggx_sample_vndf(f.woL, alpha, u1, u2, m)) return sample_invalid();
        float F;
        if (!fresnel_dielectric(dot(f.woL, m), c.etaI, c.etaT, F)) return sample_invalid();
        float pr, pt;
        if (!rough_event_probabilities(p, F, pr, pt)) return sample_invalid();
        if (u0 < pr) {
            dirL = -f.woL + m * (2.0f * dot(f.woL, m));
            if (dirL.z <= kDirEps) return sample_invalid();
            lobes = kLobeGlossy | kLobeReflection;
        } else {
            const float etaRel = c.etaT / c.etaI;
            if (pt <= 0.0f || !refract_local(f.woL, m, etaRel, dirL) || dirL.z >= -kDirEps) return sample_invalid();
            eta = etaRel;
            lobes = kLobeGlossy | kLobeTransmission;
        }
    } else {
        return sample_invalid();
    }
    V3 wiW;
    if (!safe_normalize(to_world(dirL, f), wiW) || !validate_world_event(f.woW, wiW, f.gn, lobes)) {
        return sample_invalid();
    }
    // SampleBsdfL6 re-evaluates from the world directions, rebuilding the frame.
    const BsdfEval e = eval_bsdf<kExt>(c, p, f.woW, wiW);
    if (!e.valid) return sample_invalid();
    BsdfSample s = sample_invalid();
    s.dir = wiW;
    s.pdf = e.pdf;
    s.value = e.value;
    s.diffuse = e.diffuse;
    s.specular = e.specular;
    s.eta = eta;
    s.lobeFlags = lobes;
    s.valid = true;
    return s;
}

__device__ __forceinline__ BsdfParams material_params(const Material& m) {
    BsdfParams p;
    p.base = v3(m.baseColorMetallic);
    p.metallic = m.baseColorMetallic.w;
    p.f0 = v3(m.f0);
    p.rough = m.emissiveRoughness.w;
    p.conductorEta = v3(m.conductorEta);
    p.conductorK = v3(m.conductorK);
    p.transmission = m.transmissionIor.x;
    p.model = m.metadata.x;
    p.lobes = m.metadata.y;
    p.complexFresnel = m.metadata.z;
    // PbrMaterialToBsdfParamsL6 (pbr_material_bsdf.hlsli:24-40): a rough
    // dielectric below the microfacet threshold is the smooth dielectric.
    if (p.model == kModelRoughDielectric && fin(p.rough) && p.rough >= 0.0f && p.rough * p.rough < kMinAlpha) {
        p.model = kModelSmoothGlass;
        if ((p.lobes & kLobeGlossy) != 0u) p.lobes = (p.lobes & ~kLobeGlossy) | kLobeSpecular;
    }
    return p;
}

__device__ __forceinline__ void build_basis(V3 n, V3& t, V3& b) {
    const float sgn = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sgn + n.z);
    const float bb = n.x * n.y * a;
    t = v3(1.0f + sgn * n.x * n.x * a, sgn * bb, -sgn * n.x);
    b = v3(bb, sgn + n.y * n.y * a, -n.y);
}

// ---- lights ----------------------------------------------------------------
struct LightSample {
    V3 wi;
    float distance;
    V3 Li;
    float selectionPmf, combinedPdfW;
    uint32_t lightIndex, primitiveId, flags;
    // Kept for ReSTIR's persistent sample: the sampled point (point lights:
    // the light position, infinite lights: 0), the conditional pdf of the
    // shape sample (area measure for triangles, solid angle for the
    // environment, 1 for delta lights), the measure, and the light's stable
    // ID and flags (identity.z / .y).
    V3 position;
    float conditionalPdf;
    uint32_t measure, stableId, lightFlags;
};

constexpr uint32_t kMeasureDiscrete = 0u;
constexpr uint32_t kMeasureArea = 1u;
constexpr uint32_t kMeasureSolidAngle = 2u;

constexpr uint32_t kSampleValid = 1u;
constexpr uint32_t kSampleDelta = 2u;
constexpr uint32_t kSampleInfinite = 4u;
constexpr uint32_t kSampleTwoSided = 8u;

__device__ __forceinline__ LightSample light_invalid() {
    LightSample s;
    s.wi = s.Li = zero3();
    s.distance = 0.0f;
    s.selectionPmf = s.combinedPdfW = 0.0f;
    s.lightIndex = kInvalid;
    s.primitiveId = kInvalid;
    s.flags = 0u;
    s.position = zero3();
    s.conditionalPdf = 0.0f;
    s.measure = kInvalid;
    s.stableId = kInvalid;
    s.lightFlags = 0u;
    return s;
}

__device__ __forceinline__ V3 light_radiance_scale(const Light& l) {
    return v3(l.radianceScale) * l.radianceScale.w;
}

__device__ __forceinline__ AliasEntry sample_alias(const AliasEntry* __restrict__ table, uint32_t count, float u) {
    AliasEntry inv = {0.0f, 0.0f, 0u, kInvalid};
    if (count == 0) return inv;
    const float scaled = u * (float)count;
    uint32_t bin = (uint32_t)scaled;
    if (bin > count - 1u) bin = count - 1u;
    const float within = scaled - (float)bin;
    const AliasEntry e = table[bin];
    const uint32_t sel = within < e.q ? bin : e.alias;
    if (sel >= count) return inv;
    return table[sel];
}

// HLSL smoothstep: t = saturate((x - a) / (b - a)), t * t * (3 - 2 t).
__device__ __forceinline__ float smoothstep_f(float a, float b, float x) {
    float t = (x - a) / (b - a);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

constexpr float kTwoPi = 6.28318530717958647692f;

// atan(t) for t in [0, 1]: the Cephes atanf reduction at tan(pi/8) and its
// odd polynomial (S. L. Moshier, Cephes Math Library), + - * / only.
__device__ __forceinline__ float atan01(float t) {
    float base = 0.0f;
    if (t > 0.41421356237309503f) {
        t = (t - 1.0f) / (t + 1.0f);
        base = 0.78539816339744831f;
    }
    const float z = t * t;
    const float poly = ((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f;
    return base + (poly * z * t + t);
}

__device__ __forceinline__ float atan2_shared(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) return 0.0f;
    const bool swap = ay > ax;
    float a = atan01(swap ? ax / ay : ay / ax);
    if (swap) a = 1.57079632679489662f - a;
    if (x < 0.0f) a = 3.14159265358979323846f - a;
    return y < 0.0f ? -a : a;
}

__device__ __forceinline__ float acos_shared(float x) {
    const float c = x < -1.0f ? -1.0f : (x > 1.0f ? 1.0f : x);
    return atan2_shared(sqrtf((1.0f - c) * (1.0f + c)), c);
}

// PbrEnvironmentUvFromLocalDirectionL6 (pbr_light_sampling.hlsli:87-97).
__device__ __forceinline__ void env_uv(V3 local, float& u, float& v) {
    const V3 d = normalize(local);
    float phi = atan2_shared(d.z, d.x);
    if (phi < 0.0f) phi += kTwoPi;
    u = phi / kTwoPi;
    v = acos_shared(d.y) / kPi;
}

// The map read as the reference's sampler does (Wave2Runtime.cpp:2571-2576):
// nearest texel, u repeats, v clamps to the edge.
__device__ __forceinline__ V3 env_texel(const ShadingView& sh, float u, float v) {
    const uint32_t w = sh.envWidth, h = sh.envHeight;
    uint32_t i = (uint32_t)fminf(fmaxf(floorf(u * (float)w), 0.0f), (float)(2u * w));
    const uint32_t j = (uint32_t)fminf(fmaxf(floorf(v * (float)h), 0.0f), (float)(h - 1u));
    i %= w;
    const F4 t = sh.envTexels[(size_t)j * w + i];
    return v3(t.x, t.y, t.z);
}

// cos(pi * row / height) from the shared sin / cos of 2 pi u.
__device__ __forceinline__ float env_row_cos(uint32_t row, uint32_t height) {
    float sn, cs;
    sincos_2pi((float)row / (float)(2u * height), sn, cs);
    return cs;
}

// PbrEnvironmentTexelSolidAngleL6 (pbr_light_sampling.hlsli:74-85).
__device__ __forceinline__ float env_texel_solid_angle(const ShadingView& sh, uint32_t row) {
    const uint32_t w = sh.envWidth, h = sh.envHeight;
    if (w == 0u || h == 0u || row >= h) return 0.0f;
    return (kTwoPi / (float)w) * (env_row_cos(row, h) - env_row_cos(row + 1u, h));
}

// Identity environment-to-world rows (Wave2Runtime.cpp:1416-1421).
__device__ __forceinline__ V3 env_local(V3 d) {
    return normalize(v3(dot(v3(1.0f, 0.0f, 0.0f), d), dot(v3(0.0f, 1.0f, 0.0f), d), dot(v3(0.0f, 0.0f, 1.0f), d)));
}

// Kernels are built twice: kExt with sphere lights and lat-long maps, and
// without, for the scenes whose environment is the constant 1 x 1 map under
// the uniform-sphere sampler and that hold no sphere light (the host picks;
// both give the same results on such scenes).

// PbrEnvironmentRadianceL6 (pbr_light_sampling.hlsli:99-124).
template <bool kExt>
__device__ __forceinline__ V3 environment_radiance(const ShadingView& sh, V3 dir) {
    const uint32_t li = sh.environmentLight;
    if (li == kInvalid || li >= sh.lightCount || sh.envWidth == 0u || sh.envHeight == 0u) return zero3();
    const Light l = sh.lights[li];
    if (l.identity.x != kLightEnvironment || (l.identity.y & kLightEnabled) == 0u) return zero3();
    if (!kExt) return v3(sh.envTexels[0].x, sh.envTexels[0].y, sh.envTexels[0].z) * light_radiance_scale(l);
    float u, v;
    env_uv(env_local(dir), u, v);
    return env_texel(sh, u, v) * light_radiance_scale(l);
}

// PbrEnvironmentPdfL6 (pbr_light_sampling.hlsli:126-158).
template <bool kExt>
__device__ __forceinline__ float environment_pdf(const ShadingView& sh, V3 wi) {
    const uint32_t w = sh.envWidth, h = sh.envHeight;
    if (w == 0u || h == 0u) return 0.0f;
    if (!kExt || sh.envSampler == 0u) return 1.0f / (4.0f * kPi);
    float u, v;
    env_uv(env_local(wi), u, v);
    uint32_t row = (uint32_t)fmaxf(v * (float)h, 0.0f), col = (uint32_t)fmaxf(u * (float)w, 0.0f);
    if (row > h - 1u) row = h - 1u;
    if (col > w - 1u) col = w - 1u;
    const float pmf = sh.envRows[row].pmf * sh.envColumns[(size_t)row * w + col].pmf;
    const float sa = env_texel_solid_angle(sh, row);
    return sa > 0.0f ? pmf / sa : 0.0f;
}

// PbrSampleSphereLightL6 (pbr_light_sampling.hlsli:309-420): the visible
// cone from outside, the sphere's area from inside.
__device__ __forceinline__ LightSample sample_sphere(const Light& l, uint32_t li, float pmf, V3 p, float u0, float u1,
                                                  LightSample s) {
    const float radius = l.shapeParams.x;
    const V3 center = v3(l.positionRange);
    if (!(radius > 0.0f) || !(pmf > 0.0f)) return s;
    const V3 toCenter = center - p;
    const float d2 = dot(toCenter, toCenter);
    const float r2 = radius * radius;
    const bool two = (l.identity.y & kLightTwoSided) != 0u;
    V3 wi, lp;
    float dist, cpdf;
    uint32_t measure;
    if (d2 > r2 * 1.000001f) {
        const float cd = sqrtf(d2);
        const V3 cdir = toCenter / cd;
        const float sin2Max = r2 / d2;
        const float cosMax = sqrtf(fmaxf(0.0f, 1.0f - sin2Max));
        const float omc = sin2Max < 0.00068523f ? 0.5f * sin2Max : 1.0f - cosMax;
        const float cosTheta = 1.0f - u0 * omc;
        const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
        float sp, cp;
        sincos_2pi(u1, sp, cp);
        V3 t, b;
        build_basis(cdir, t, b);
        wi = normalize(cdir * cosTheta + t * (cp * sinTheta) + b * (sp * sinTheta));
        const V3 oc = p - center;
        const float halfB = dot(oc, wi);
        const float disc = halfB * halfB - (dot(oc, oc) - r2);
        if (!(disc >= 0.0f)) return s;
        dist = -halfB - sqrtf(disc);
        if (!(dist > 0.0f)) dist = -halfB + sqrtf(disc);
        if (!(dist > 0.0f) || !(omc > 0.0f)) return s;
        lp = p + wi * dist;
        cpdf = 1.0f / (kTwoPi * omc);
        measure = kMeasureSolidAngle;
        s.combinedPdfW = pmf * cpdf;
    } else {
        const float z = 1.0f - 2.0f * u0;
        const float rd = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cp;
        sincos_2pi(u1, sp, cp);
        const V3 ln = v3(rd * cp, z, rd * sp);
        lp = center + ln * radius;
        const V3 toL = lp - p;
        const float sd2 = dot(toL, toL);
        if (!(sd2 > 0.0f)) return s;
        dist = sqrtf(sd2);
        wi = toL / dist;
        const float lc = dot(ln, -wi);
        const float ac = two ? fabsf(lc) : lc;
        if (!(ac > 0.0f)) return s;
        cpdf = 1.0f / (4.0f * kPi * r2);
        s.combinedPdfW = pmf * (cpdf * sd2 / ac);
        measure = kMeasureArea;
    }
    s.wi = wi;
    s.distance = dist;
    s.Li = light_radiance_scale(l);
    s.selectionPmf = pmf;
    s.lightIndex = li;
    s.primitiveId = l.identity.w;
    s.flags = kSampleValid | (two ? kSampleTwoSided : 0u);
    s.position = lp;
    s.conditionalPdf = cpdf;
    s.measure = measure;
    return s;
}

// PbrLightPdfLiL6's sphere branch (pbr_light_sampling.hlsli:676-704).
__device__ __forceinline__ float sphere_pdf(const Light& l, V3 p, V3 wi, float dist, V3 en) {
    const float radius = l.shapeParams.x;
    const V3 toCenter = v3(l.positionRange) - p;
    const float d2 = dot(toCenter, toCenter), r2 = radius * radius;
    if (!(radius > 0.0f)) return 0.0f;
    if (d2 > r2 * 1.000001f) {
        const float sin2Max = r2 / d2;
        const float cosMax = sqrtf(fmaxf(0.0f, 1.0f - sin2Max));
        const float omc = sin2Max < 0.00068523f ? 0.5f * sin2Max : 1.0f - cosMax;
        return omc > 0.0f ? 1.0f / (kTwoPi * omc) : 0.0f;
    }
    const float lc = dot(en, -wi);
    const float c = (l.identity.y & kLightTwoSided) != 0u ? fabsf(lc) : lc;
    const float area = 4.0f * kPi * r2;
    return dist > 0.0f && c > 0.0f ? dist * dist / (area * c) : 0.0f;
}

// PbrSampleEnvironmentLightL6 (pbr_light_sampling.hlsli:422-509), identity
// environment-to-world rows; out of line to keep Shade's registers down.
template <bool kExt, class Shape>
__device__ __forceinline__ LightSample sample_environment(const ShadingView& sh, const Light& l, uint32_t li, float pmf,
                                                       float r1, float r2, Shape shape, LightSample s) {
    const uint32_t w = sh.envWidth, h = sh.envHeight;
    if (!(pmf > 0.0f) || w == 0u || h == 0u) return s;
    V3 local;
    float u, v, pdfW;
    if (!kExt || sh.envSampler == 0u) {
        const float cosTheta = 1.0f - 2.0f * r1;
        const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
        float sp, cp;
        sincos_2pi(r2, sp, cp);
        local = v3(sinTheta * cp, cosTheta, sinTheta * sp);
        if (kExt) env_uv(local, u, v);
        pdfW = 1.0f / (4.0f * kPi);
    } else {
        const AliasEntry re = sample_alias(sh.envRows, h, r1);
        if (re.item >= h || !(re.pmf > 0.0f)) return s;
        const AliasEntry ce = sample_alias(sh.envColumns + (size_t)re.item * w, w, r2);
        if (ce.item >= w || !(ce.pmf > 0.0f)) return s;
        const float r3 = shape(2u), r4 = shape(3u);
        const float c0 = env_row_cos(re.item, h), c1 = env_row_cos(re.item + 1u, h);
        const float cosTheta = c0 + (c1 - c0) * r4;
        const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
        u = ((float)ce.item + r3) / (float)w;
        float sp, cp;
        sincos_2pi(u, sp, cp);
        local = v3(sinTheta * cp, cosTheta, sinTheta * sp);
        const float sa = env_texel_solid_angle(sh, re.item);
        if (!(sa > 0.0f)) return s;
        v = acos_shared(cosTheta) / kPi;
        pdfW = re.pmf * ce.pmf / sa;
    }
    s.wi = env_local(local);
    s.distance = 1.0e30f;
    s.Li = (kExt ? env_texel(sh, u, v) : v3(sh.envTexels[0].x, sh.envTexels[0].y, sh.envTexels[0].z)) *
           light_radiance_scale(l);
    s.selectionPmf = pmf;
    s.combinedPdfW = pmf * pdfW;
    s.lightIndex = li;
    s.flags = kSampleValid | kSampleInfinite;
    s.conditionalPdf = pdfW;
    s.measure = kMeasureSolidAngle;
    return s;
}

// PbrSampleOneLightL6 (pbr_light_sampling.hlsli:511-589).  shape(k) draws
// light shape dimension k (2 and 3 are read by the importance-sampled
// environment only).
template <bool kExt, class Shape>
__device__ __forceinline__ LightSample sample_one_light(const ShadingView& sh, V3 p, float r0, float r1, float r2,
                                                        Shape&& shape) {
    if (sh.aliasCount == 0u) return light_invalid();
    const AliasEntry sel = sample_alias(sh.alias, sh.aliasCount, r0);
    if (sel.item >= sh.lightCount || !(sel.pmf > 0.0f)) return light_invalid();
    const Light l = sh.lights[sel.item];
    if ((l.identity.y & kLightEnabled) == 0u) return light_invalid();
    LightSample s = light_invalid();
    const float pmf = sel.pmf;
    s.stableId = l.identity.z;
    s.lightFlags = l.identity.y;
    if (l.identity.x == kLightPoint || l.identity.x == kLightSpot) {
        const V3 toL = v3(l.positionRange) - p;
        const float d2 = dot(toL, toL);
        if (!(d2 > 0.0f) || !(pmf > 0.0f)) return s;
        const float dist = sqrtf(d2);
        if (l.positionRange.w > 0.0f && dist > l.positionRange.w) return s;
        const V3 wi = toL / dist;
        float falloff = 1.0f;
        if (l.identity.x == kLightSpot) {
            const V3 dir = v3(l.directionCosOuter);
            const float len2 = dot(dir, dir);
            if (!(len2 > 0.0f) || !fin(len2)) return s;
            const V3 forward = dir / sqrtf(len2);
            const float cosAngle = dot(forward, -wi);
            const float cosOuter = l.directionCosOuter.w, cosInner = l.shapeParams.y;
            falloff = cosInner == cosOuter ? (cosAngle >= cosOuter ? 1.0f : 0.0f)
                                           : smoothstep_f(cosOuter, cosInner, cosAngle);
            if (!(falloff > 0.0f)) return s;
        }
        s.wi = wi;
        s.distance = dist;
        s.Li = light_radiance_scale(l) * (falloff / d2);
        s.selectionPmf = pmf;
        s.combinedPdfW = 0.0f;
        s.lightIndex = sel.item;
        s.flags = kSampleValid | kSampleDelta;
        s.position = v3(l.positionRange);
        s.conditionalPdf = 1.0f;
        s.measure = kMeasureDiscrete;
        return s;
    }
    if (l.identity.x == kLightDirectional) {
        const V3 d = v3(l.directionCosOuter);
        const float l2 = dot(d, d);
        if (!(l2 > 0.0f) || !(pmf > 0.0f)) return s;
        s.wi = (-d) * rsq(l2);
        s.distance = 1.0e30f;
        s.Li = light_radiance_scale(l);
        s.selectionPmf = pmf;
        s.combinedPdfW = 0.0f;
        s.lightIndex = sel.item;
        s.flags = kSampleValid | kSampleDelta | kSampleInfinite;
        s.conditionalPdf = 1.0f;
        s.measure = kMeasureDiscrete;
        return s;
    }
    if (l.identity.x == kLightEmissiveTriangle) {
        if (l.payload.x >= sh.fixtureCount || !(pmf > 0.0f)) return s;
        const Fixture fx = sh.fixtures[l.payload.x];
        if (fx.metadata.z != l.payload.w || fx.metadata.y != l.identity.w) return s;
        const V3 p0 = v3(fx.p0), p1 = v3(fx.p1), p2 = v3(fx.p2);
        const V3 nu = cross(p1 - p0, p2 - p0);
        const float twiceArea = length(nu);
        if (!(twiceArea > 0.0f)) return s;
        const float root = sqrtf(r1);
        const float b0 = 1.0f - root;
        const float b1 = root * (1.0f - r2);
        const float b2 = root * r2;
        const V3 lp = p0 * b0 + p1 * b1 + p2 * b2;
        const V3 toL = lp - p;
        const float d2 = dot(toL, toL);
        if (!(d2 > 0.0f)) return s;
        const float dist = sqrtf(d2);
        const V3 wi = toL / dist;
        const V3 ln = nu / twiceArea;
        const float lc = dot(ln, -wi);
        const bool two = (l.identity.y & kLightTwoSided) != 0u;
        const float ac = two ? fabsf(lc) : lc;
        if (!(ac > 0.0f)) return s;
        const float area = 0.5f * twiceArea;
        const float pdfA = 1.0f / area;
        const float pdfW = pdfA * d2 / ac;
        s.wi = wi;
        s.distance = dist;
        s.Li = light_radiance_scale(l);
        s.selectionPmf = pmf;
        s.combinedPdfW = pmf * pdfW;
        s.lightIndex = sel.item;
        s.primitiveId = fx.metadata.y;
        s.flags = kSampleValid | (two ? kSampleTwoSided : 0u);
        s.position = lp;
        s.conditionalPdf = pdfA;
        s.measure = kMeasureArea;
        return s;
    }
    if (kExt && l.identity.x == kLightSphere) return sample_sphere(l, sel.item, pmf, p, r1, r2, s);
    if (l.identity.x == kLightEnvironment) return sample_environment<kExt>(sh, l, sel.item, pmf, r1, r2, shape, s);
    return s;
}

// PbrSelectedLightPdfL6 (pbr_light_sampling.hlsli:591-736) at shading
// position p (read by sphere lights only).
template <bool kExt>
__device__ __forceinline__ float selected_light_pdf(const ShadingView& sh, uint32_t li, V3 p, V3 wi, float dist,
                                                    V3 en) {
    if (li == kInvalid || li >= sh.lightCount) return 0.0f;
    const Light l = sh.lights[li];
    float pdf = 0.0f;
    if ((l.identity.y & kLightEnabled) != 0u) {
        if (l.identity.x == kLightEnvironment) {
            pdf = environment_pdf<kExt>(sh, wi);
        } else if (kExt && l.identity.x == kLightSphere) {
            pdf = sphere_pdf(l, p, wi, dist, en);
        } else if (l.identity.x == kLightEmissiveTriangle) {
            if (l.payload.x < sh.fixtureCount && dist > 0.0f) {
                const Fixture fx = sh.fixtures[l.payload.x];
                if (fx.metadata.z == l.payload.w && fx.metadata.y == l.identity.w) {
                    const float area = 0.5f * length(cross(v3(fx.p1) - v3(fx.p0), v3(fx.p2) - v3(fx.p0)));
                    const float lc = dot(en, -wi);
                    const float c = (l.identity.y & kLightTwoSided) != 0u ? fabsf(lc) : lc;
                    pdf = area > 0.0f && c > 0.0f ? dist * dist / (area * c) : 0.0f;
                }
            }
        }
    }
    return sh.selectionPmf[li] * pdf;
}

__device__ __forceinline__ float power_heuristic(float a, float b) {
    if (!fin(a) || !fin(b) || a < 0.0f || b < 0.0f) return 0.0f;
    const float m = fmaxf(a, b);
    if (!(m > 0.0f) || !fin(m)) return 0.0f;
    const float na = a / m, nb = b / m;
    const float sa = na * na, sb = nb * nb;
    return sa / (sa + sb);
}

__device__ __forceinline__ V3 offset_origin(V3 p, V3 gn, V3 dir, float eps) {
    const float side = dot(gn, dir) >= 0.0f ? 1.0f : -1.0f;
    return p + gn * (side * eps);
}

}  // namespace rex
