// Thin extern "C" wrapper around the reference's own CPU shading code,
// compiled from the reference sources where they lie
// (integrators/megakernel/MegakernelBridge.cpp: Philox, counter RNG, Vose
// alias tables, light power weights, emitter map, top-level light
// distribution; integrators/megakernel/include/ReferenceBsdf.hpp: the
// double-precision L6 BSDF) into oracle/_ref/libref.so by oracle/Makefile.
// Test infrastructure only: it pins the oracle and never ships.
#include "MegakernelBridge.hpp"
#include "include/ReferenceBsdf.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <span>
#include <vector>

namespace Mega = RenderingEngine::Integrators::Megakernel;
namespace Bsdf = RenderingEngine::Integrators::Megakernel::ReferenceBsdfL6;

namespace
{
    Bsdf::Vec3 V(const float* p) { return {p[0], p[1], p[2]}; }

    // material: 32 words of PbrMaterialGpu, converted as the reference's
    // shader-side parameter load does (base, metallic, f0, roughness, eta,
    // transmission, k, model, lobes, complex Fresnel).
    Bsdf::BsdfParamsL6 Params(const std::uint32_t* material)
    {
        const float* f = reinterpret_cast<const float*>(material);
        Bsdf::BsdfParamsL6 p;
        p.baseColor = {f[0], f[1], f[2]};
        p.metallic = f[3];
        p.perceptualRoughness = f[7];
        p.transmission = f[8];
        p.f0 = {f[16], f[17], f[18]};
        p.conductorEta = {f[20], f[21], f[22]};
        p.conductorK = {f[24], f[25], f[26]};
        p.model = material[28];
        p.allowedLobes = material[29];
        p.useComplexConductorFresnel = material[30];
        return p;
    }

    Bsdf::BsdfContextL6 Context(const float* gn, const float* sn, const float* tangent, float eta)
    {
        Bsdf::BsdfContextL6 c;
        c.geometricNormal = V(gn);
        c.shadingNormal = V(sn);
        c.tangent = V(tangent);
        c.etaIncident = 1.0;
        c.etaTransmitted = eta;
        c.transportMode = Bsdf::kTransportRadiance;
        return c;
    }
}

extern "C" {

void ref_philox(const std::uint32_t counter[4], const std::uint32_t key[2], std::uint32_t out[4])
{
    const auto r = Mega::Philox4x32TenRounds({counter[0], counter[1], counter[2], counter[3]}, {key[0], key[1]});
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

float ref_unit_float(std::uint32_t v) { return Mega::Uint32ToUnitFloat(v); }

float ref_counter_float(std::uint32_t pixel, std::uint32_t sample, std::uint32_t dim, std::uint32_t tag,
                        std::uint64_t seed)
{
    return Mega::CounterRandomFloat(pixel, sample, dim, tag, seed);
}

// out: count 16-byte AliasEntryGpu records.  Returns 0, or -1 if the
// reference rejected the weights.
int ref_build_alias(const double* weights, const std::uint32_t* items, std::uint32_t count, void* out,
                    int* uniformFallback)
{
    try
    {
        const Mega::AliasTable t = Mega::BuildVoseAliasTable(
            std::span<const double>(weights, count),
            items ? std::span<const std::uint32_t>(items, count) : std::span<const std::uint32_t>{});
        std::memcpy(out, t.entries.data(), t.entries.size() * sizeof(Mega::AliasEntryGpu));
        if (uniformFallback) *uniformFallback = t.usedUniformFallback ? 1 : 0;
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

// Returns the selected item (or 0xffffffff) and writes its pmf.
std::uint32_t ref_sample_alias(const void* table, std::uint32_t count, float u, float* pmf)
{
    const Mega::AliasSample s = Mega::SampleAliasTable(
        std::span<const Mega::AliasEntryGpu>(static_cast<const Mega::AliasEntryGpu*>(table), count), u);
    if (pmf) *pmf = s.pmf;
    return s.item;
}

double ref_environment_integrated_luminance(std::uint32_t width, std::uint32_t height, const float* luminance)
{
    const Mega::EnvironmentDistribution d = Mega::BuildEnvironmentDistribution(
        width, height, std::span<const float>(luminance, static_cast<size_t>(width) * height));
    return d.integratedLuminance;
}

// The two-level environment distribution: rows (height entries) and
// columns (height x width entries) as 16-B alias entries.  Returns 0, or -1
// where the reference throws.
int ref_environment_distribution(std::uint32_t width, std::uint32_t height, const float* luminance, void* rows,
                                 void* columns, double* integrated, int* uniformFallback)
{
    try
    {
        const Mega::EnvironmentDistribution d = Mega::BuildEnvironmentDistribution(
            width, height, std::span<const float>(luminance, static_cast<size_t>(width) * height));
        std::memcpy(rows, d.rowAlias.data(), d.rowAlias.size() * sizeof(Mega::AliasEntryGpu));
        std::memcpy(columns, d.columnAlias.data(), d.columnAlias.size() * sizeof(Mega::AliasEntryGpu));
        *integrated = d.integratedLuminance;
        *uniformFallback = d.usedUniformSphereFallback ? 1 : 0;
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

double ref_light_power(const void* light96, float sceneRadius, double envIntegratedLuminance)
{
    Mega::PbrLightGpu l;
    std::memcpy(&l, light96, sizeof(l));
    try
    {
        return Mega::LightPowerWeight(l, sceneRadius, envIntegratedLuminance);
    }
    catch (...)
    {
        return -1.0;
    }
}

// Top-level light distribution: alias entries (capacity count) and the
// per-light PMF.  Returns the alias size, or -1 on rejection.
int ref_top_level(const void* lights96, std::uint32_t count, std::uint32_t selection, float sceneRadius,
                  double envIntegratedLuminance, void* aliasOut, float* pmfOut)
{
    std::vector<Mega::PbrLightGpu> lights(count);
    std::memcpy(lights.data(), lights96, count * sizeof(Mega::PbrLightGpu));
    try
    {
        const Mega::TopLevelLightDistribution d = Mega::BuildTopLevelLightDistribution(
            lights, selection == 0u ? Mega::LightSelection::Uniform : Mega::LightSelection::PowerWeighted,
            sceneRadius, envIntegratedLuminance);
        std::memcpy(aliasOut, d.alias.data(), d.alias.size() * sizeof(Mega::AliasEntryGpu));
        std::memcpy(pmfOut, d.selectionPmfByLight.data(), d.selectionPmfByLight.size() * sizeof(float));
        return static_cast<int>(d.alias.size());
    }
    catch (...)
    {
        return -1;
    }
}

// Emitter map entries (16 B each, capacity count); returns the size or -1.
int ref_emitter_map(const void* lights96, std::uint32_t count, void* out)
{
    std::vector<Mega::PbrLightGpu> lights(count);
    std::memcpy(lights.data(), lights96, count * sizeof(Mega::PbrLightGpu));
    try
    {
        const auto m = Mega::BuildEmitterMap(lights);
        std::memcpy(out, m.data(), m.size() * sizeof(Mega::EmitterMapEntryGpu));
        return static_cast<int>(m.size());
    }
    catch (...)
    {
        return -1;
    }
}

// out[0..2] value, [3] pdf, [4..6] diffuse, [7..9] specular, [10] valid, [11] lobes
void ref_bsdf_eval(const std::uint32_t* material, const float gn[3], const float sn[3], const float tangent[3],
                   const float wo[3], const float wi[3], double out[12])
{
    const Bsdf::BsdfParamsL6 p = Params(material);
    const float eta = reinterpret_cast<const float*>(material)[9];
    const Bsdf::BsdfEvalL6 e = Bsdf::EvaluateBsdfL6(Context(gn, sn, tangent, eta), p, V(wo), V(wi));
    out[0] = e.value.x; out[1] = e.value.y; out[2] = e.value.z; out[3] = e.pdf;
    out[4] = e.diffuseValue.x; out[5] = e.diffuseValue.y; out[6] = e.diffuseValue.z;
    out[7] = e.specularValue.x; out[8] = e.specularValue.y; out[9] = e.specularValue.z;
    out[10] = e.isValid; out[11] = e.lobeFlags;
}

// out[0..2] direction, [3] pdf, [4..6] value, [7] valid, [8] lobes
void ref_bsdf_sample(const std::uint32_t* material, const float gn[3], const float sn[3], const float tangent[3],
                     const float wo[3], const float u[3], double out[9])
{
    const Bsdf::BsdfParamsL6 p = Params(material);
    const float eta = reinterpret_cast<const float*>(material)[9];
    const Bsdf::BsdfSampleL6 s = Bsdf::SampleBsdfL6(Context(gn, sn, tangent, eta), p, V(wo), V(u));
    out[0] = s.direction.x; out[1] = s.direction.y; out[2] = s.direction.z; out[3] = s.pdf;
    out[4] = s.value.x; out[5] = s.value.y; out[6] = s.value.z;
    out[7] = s.isValid; out[8] = s.lobeFlags;
}

// The same with the context's eta on both sides and the transport mode:
// eval out[0..2] value, [3] pdf, [4..6] diffuse, [7..9] specular, [10] valid,
// [11] lobes, [12] measure; sample out[0..2] direction, [3] pdf, [4..6]
// value, [7] valid, [8] lobes, [9] eta, [10] isDelta, [11] measure.
void ref_bsdf_eval_ctx(const std::uint32_t* material, const float gn[3], const float sn[3], const float tangent[3],
                       double etaI, double etaT, std::uint32_t transport, const float wo[3], const float wi[3],
                       double out[13])
{
    const Bsdf::BsdfParamsL6 p = Params(material);
    Bsdf::BsdfContextL6 c = Context(gn, sn, tangent, static_cast<float>(etaT));
    c.etaIncident = etaI;
    c.etaTransmitted = etaT;
    c.transportMode = transport;
    const Bsdf::BsdfEvalL6 e = Bsdf::EvaluateBsdfL6(c, p, V(wo), V(wi));
    out[0] = e.value.x; out[1] = e.value.y; out[2] = e.value.z; out[3] = e.pdf;
    out[4] = e.diffuseValue.x; out[5] = e.diffuseValue.y; out[6] = e.diffuseValue.z;
    out[7] = e.specularValue.x; out[8] = e.specularValue.y; out[9] = e.specularValue.z;
    out[10] = e.isValid; out[11] = e.lobeFlags; out[12] = e.measure;
}

void ref_bsdf_sample_ctx(const std::uint32_t* material, const float gn[3], const float sn[3], const float tangent[3],
                         double etaI, double etaT, std::uint32_t transport, const float wo[3], const float u[3],
                         double out[12])
{
    const Bsdf::BsdfParamsL6 p = Params(material);
    Bsdf::BsdfContextL6 c = Context(gn, sn, tangent, static_cast<float>(etaT));
    c.etaIncident = etaI;
    c.etaTransmitted = etaT;
    c.transportMode = transport;
    const Bsdf::BsdfSampleL6 s = Bsdf::SampleBsdfL6(c, p, V(wo), V(u));
    out[0] = s.direction.x; out[1] = s.direction.y; out[2] = s.direction.z; out[3] = s.pdf;
    out[4] = s.value.x; out[5] = s.value.y; out[6] = s.value.z;
    out[7] = s.isValid; out[8] = s.lobeFlags; out[9] = s.eta; out[10] = s.isDelta; out[11] = s.measure;
}

// Directional albedo of one outgoing direction estimated with the
// reference's own importance sampling: the mean over the n sample points
// u[3 k .. 3 k + 2] of value * |gn . direction| / pdf (an invalid sample
// counts zero).  out[0..2] the mean, out[3..5] its standard error.
void ref_bsdf_albedo(const std::uint32_t* material, const float gn[3], const float sn[3], const float tangent[3],
                     const float wo[3], const float* u, std::uint32_t n, double out[6])
{
    const Bsdf::BsdfParamsL6 p = Params(material);
    const float eta = reinterpret_cast<const float*>(material)[9];
    const Bsdf::BsdfContextL6 c = Context(gn, sn, tangent, eta);
    double sum[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
    for (std::uint32_t k = 0; k < n; ++k)
    {
        const Bsdf::BsdfSampleL6 s = Bsdf::SampleBsdfL6(c, p, V(wo), V(u + 3u * k));
        if (!s.isValid || !(s.pdf > 0.0))
            continue;
        const double cosine = std::abs(gn[0] * s.direction.x + gn[1] * s.direction.y + gn[2] * s.direction.z);
        const double w[3] = {s.value.x * cosine / s.pdf, s.value.y * cosine / s.pdf, s.value.z * cosine / s.pdf};
        for (int i = 0; i < 3; ++i)
        {
            sum[i] += w[i];
            sq[i] += w[i] * w[i];
        }
    }
    for (int i = 0; i < 3; ++i)
    {
        const double mean = n ? sum[i] / n : 0.0;
        out[i] = mean;
        out[3 + i] = n > 1u ? std::sqrt(std::max(0.0, sq[i] / n - mean * mean) / (n - 1u)) : 0.0;
    }
}
}
