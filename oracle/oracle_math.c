/*
 * Oracle (test infrastructure only): Philox4x32-10 sampling, the portable
 * sin/cos of 2*pi*u, the Vose alias table and light power weights.
 *
 * Philox: integrators/wavefront/shaders/include/WavefrontRng.hlsli:28-97
 * (counter = (pixel, sample, dimension, stream tag), key = 64-bit seed,
 * float = (bits >> 8) * 2^-24).  Pinned by the Random123 known-answer test
 * philox4x32_10({0,0,0,0},{0,0}) = {0x6627e8d5, 0xe169c58d, 0xbc57ac4c,
 * 0x9b00dbd8} (integrators/megakernel/tests/MegakernelCpuTests.cpp:222-228)
 * and the float endpoints 0 -> 0, 0xffffffff -> 1 - 2^-24 (:229-232).
 *
 * ora_sincos_2pi: a fixed-arithmetic replacement (shared with the kernels) for HLSL
 * cos/sin of 2*pi*u (the reference leaves them to the GPU vendor's
 * approximations).  Pinned against libm in double: at most 2 ulp over a dense
 * sweep of u (tests/test_oracle_shading.py).
 *
 * Alias table: MegakernelBridge.cpp:40-145.  Pinned by the closed-form PMF
 * (sum = 1, pmf_i = w_i / sum w) and a chi-square test of sampling frequency,
 * as MegakernelCpuTests.cpp:35-142 does.
 * Light power: MegakernelBridge.cpp:274-353 (closed forms, pinned by direct
 * formula checks on single lights).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_shading.h"

static void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void ora_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = counter[0], c1 = counter[1], c2 = counter[2], c3 = counter[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xd2511f53u, c0, &hi0, &lo0);
        mulhilo(0xcd9e8d57u, c2, &hi1, &lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9e3779b9u;
        k1 += 0xbb67ae85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float ora_sample_dimension(uint32_t pixel, uint32_t sample, uint32_t dim, uint32_t tag,
                           uint32_t seedLow, uint32_t seedHigh) {
    const uint32_t counter[4] = {pixel, sample, dim, tag};
    const uint32_t key[2] = {seedLow, seedHigh};
    uint32_t out[4];
    ora_philox4x32_10(counter, key, out);
    return (float)(out[0] >> 8) * (1.0f / 16777216.0f);
}

/* Taylor coefficients of sin(pi/2 r) and cos(pi/2 r), rounded to float. */
#define S1 1.5707963267948966f
#define S3 (-0.6459640975062462f)
#define S5 0.07969262624616703f
#define S7 (-0.004681754135318687f)
#define S9 0.00016044118478735982f
#define C2 (-1.2337005501361697f)
#define C4 0.25366950790104797f
#define C6 (-0.020863480763352958f)
#define C8 0.0009192602748394263f
#define C10 (-2.5202042373060605e-05f)

void ora_sincos_2pi(float u, float* s, float* c) {
    const float x = u * 4.0f;                 /* exact */
    const float q = floorf(x + 0.5f);
    const float r = x - q;                    /* exact, in [-0.5, 0.5] */
    const float r2 = r * r;
    const float sp = r * (S1 + r2 * (S3 + r2 * (S5 + r2 * (S7 + r2 * S9))));
    const float cp = 1.0f + r2 * (C2 + r2 * (C4 + r2 * (C6 + r2 * (C8 + r2 * C10))));
    const int quadrant = ((int)q) & 3;
    switch (quadrant) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
    }
}

int ora_build_alias(const double* weights, const uint32_t* items, uint32_t count, ora_alias* out) {
    if (count == 0) return 0;
    double sum = 0.0;
    for (uint32_t i = 0; i < count; ++i) {
        if (!isfinite(weights[i]) || weights[i] < 0.0) return -1;
        sum += weights[i];
        if (!isfinite(sum)) return -2;
    }
    const int uniform = sum == 0.0;
    const double norm = sum > 0.0 ? sum : (double)count;
    double* scaled = (double*)malloc(sizeof(double) * count);
    uint32_t* small = (uint32_t*)malloc(sizeof(uint32_t) * count);
    uint32_t* large = (uint32_t*)malloc(sizeof(uint32_t) * count);
    uint32_t ns = 0, nl = 0;
    const double countD = (double)count;
    for (uint32_t i = 0; i < count; ++i) {
        const double p = uniform ? 1.0 / countD : weights[i] / norm;
        scaled[i] = p * countD;
        out[i].pmf = (float)p;
        out[i].q = 0.0f;
        out[i].item = items ? items[i] : i;
        out[i].alias = i;
        if (scaled[i] < 1.0) small[ns++] = i; else large[nl++] = i;
    }
    while (ns > 0 && nl > 0) {
        const uint32_t si = small[--ns];
        const uint32_t li = large[--nl];
        double q = scaled[si];
        if (q < 0.0) q = 0.0;
        if (q > 1.0) q = 1.0;
        out[si].q = (float)q;
        out[si].alias = li;
        scaled[li] = scaled[li] + scaled[si] - 1.0;
        if (scaled[li] < 1.0) small[ns++] = li; else large[nl++] = li;
    }
    for (uint32_t k = 0; k < nl; ++k) { out[large[k]].q = 1.0f; out[large[k]].alias = large[k]; }
    for (uint32_t k = 0; k < ns; ++k) { out[small[k]].q = 1.0f; out[small[k]].alias = small[k]; }
    free(scaled);
    free(small);
    free(large);
    return 0;
}

static int nonneg_finite(double v) { return isfinite(v) && v >= 0.0; }

/* Returns -1 where the reference's LightPowerWeight rejects the light:
 * non-positive or non-finite scene radius, negative or non-finite radiance,
 * scale or luminance, a directional or spot light without a direction, spot
 * cone cosines outside [-1, 1] or inner < outer, a sphere light without a
 * radius, a triangle light without area, a negative environment integral, an unknown type. */
double ora_light_power(const ora_light* light, float sceneRadius, double envIntegratedLuminance) {
    const double pi = 3.14159265358979323846;
    if (!isfinite(sceneRadius) || sceneRadius <= 0.0f) return -1.0;
    const double scale = (double)light->radianceScale[3];
    if (!nonneg_finite(scale) || !nonneg_finite(light->radianceScale[0]) ||
        !nonneg_finite(light->radianceScale[1]) || !nonneg_finite(light->radianceScale[2]))
        return -1.0;
    const double lum = (0.2126 * (double)light->radianceScale[0] + 0.7152 * (double)light->radianceScale[1] +
                        0.0722 * (double)light->radianceScale[2]) * scale;
    if (!nonneg_finite(lum)) return -1.0;
    const double radius = (double)sceneRadius;
    const double side = (light->identity[1] & ORA_LIGHT_TWO_SIDED) != 0u ? 2.0 : 1.0;
    switch (light->identity[0]) {
    case ORA_LIGHT_POINT: return 4.0 * pi * lum;
    case ORA_LIGHT_DIRECTIONAL: {
        const double dx = (double)light->directionCosOuter[0], dy = (double)light->directionCosOuter[1],
                     dz = (double)light->directionCosOuter[2];
        const double len2 = dx * dx + dy * dy + dz * dz;
        if (!isfinite(len2) || len2 <= 0.0) return -1.0;
        return pi * radius * radius * lum;
    }
    case ORA_LIGHT_SPOT: {
        const double dx = (double)light->directionCosOuter[0], dy = (double)light->directionCosOuter[1],
                     dz = (double)light->directionCosOuter[2];
        const double len2 = dx * dx + dy * dy + dz * dz;
        if (!isfinite(len2) || len2 <= 0.0) return -1.0;
        const double cosOuter = (double)light->directionCosOuter[3], cosInner = (double)light->shapeParams[1];
        if (!isfinite(cosOuter) || !isfinite(cosInner) || cosOuter < -1.0 || cosOuter > 1.0 || cosInner < -1.0 ||
            cosInner > 1.0 || cosInner < cosOuter)
            return -1.0;
        return 2.0 * pi * lum * ((1.0 - cosInner) + 0.5 * (cosInner - cosOuter));
    }
    case ORA_LIGHT_SPHERE: {
        const double r = (double)light->shapeParams[0];
        if (!isfinite(r) || r <= 0.0) return -1.0;
        return 4.0 * pi * pi * r * r * lum * side;
    }
    case ORA_LIGHT_EMISSIVE_TRIANGLE: {
        const double area = (double)light->shapeParams[2];
        if (!isfinite(area) || area <= 0.0) return -1.0;
        return pi * area * lum * side;
    }
    case ORA_LIGHT_ENVIRONMENT:
        if (!nonneg_finite(envIntegratedLuminance)) return -1.0;
        return pi * radius * radius * scale * envIntegratedLuminance;
    default: return -1.0;
    }
}

/* atan(t) for t in [0, 1]: the Cephes atanf reduction at tan(pi/8) and its
 * odd polynomial (S. L. Moshier, Cephes Math Library).  Only + - * / on
 * floats, no contraction, so the product's kernels give the same bits. */
static float atan01(float t) {
    float base = 0.0f;
    if (t > 0.41421356237309503f) {
        t = (t - 1.0f) / (t + 1.0f);
        base = 0.78539816339744831f;
    }
    const float z = t * t;
    const float poly = ((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f;
    return base + (poly * z * t + t);
}

float ora_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) return 0.0f;
    const int swap = ay > ax;
    float a = atan01(swap ? ax / ay : ay / ax);
    if (swap) a = 1.57079632679489662f - a;
    if (x < 0.0f) a = 3.14159265358979323846f - a;
    return y < 0.0f ? -a : a;
}

float ora_acos(float x) {
    const float c = x < -1.0f ? -1.0f : (x > 1.0f ? 1.0f : x);
    return ora_atan2(sqrtf((1.0f - c) * (1.0f + c)), c);
}

/* TexelSolidAngle, integrators/megakernel/MegakernelBridge.cpp:176-189. */
static double texel_solid_angle_d(uint32_t row, uint32_t width, uint32_t height) {
    const double pi = 3.14159265358979323846;
    const double t0 = pi * (double)row / (double)height;
    const double t1 = pi * (double)(row + 1u) / (double)height;
    return 2.0 * pi / (double)width * (cos(t0) - cos(t1));
}

/* Texel luminance of the environment map, Wave2Runtime.cpp:589-599 (float). */
void ora_environment_luminance(const float* rgba, uint32_t count, float* out) {
    for (uint32_t i = 0; i < count; ++i) {
        const float* t = rgba + 4u * (size_t)i;
        out[i] = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
    }
}

/* BuildEnvironmentDistribution, MegakernelBridge.cpp:191-275: texel weights
 * luminance x solid angle (solid angle alone when the map is black), one
 * column alias table per row (width entries each) and a row alias table
 * over the row sums.  Returns 0, or -1 where the reference throws. */
int ora_build_environment_distribution(uint32_t width, uint32_t height, const float* luminance, ora_alias* rows,
                                       ora_alias* columns, double* integrated, int* uniformFallback) {
    if (width == 0u || height == 0u) return -1;
    const size_t n = (size_t)width * height;
    double* w = (double*)malloc(sizeof(double) * n);
    double* rowW = (double*)calloc(height, sizeof(double));
    double total = 0.0;
    int rc = 0;
    for (uint32_t r = 0; r < height && rc == 0; ++r) {
        const double sa = texel_solid_angle_d(r, width, height);
        for (uint32_t c = 0; c < width; ++c) {
            const double l = (double)luminance[(size_t)r * width + c];
            if (!isfinite(l) || l < 0.0) { rc = -1; break; }
            w[(size_t)r * width + c] = l * sa;
            total += l * sa;
        }
    }
    if (rc == 0 && !isfinite(total)) rc = -1;
    if (rc == 0) {
        *uniformFallback = total == 0.0;
        if (total == 0.0) {
            for (uint32_t r = 0; r < height; ++r) {
                const double sa = texel_solid_angle_d(r, width, height);
                for (uint32_t c = 0; c < width; ++c) w[(size_t)r * width + c] = sa;
            }
        }
        for (uint32_t r = 0; r < height && rc == 0; ++r) {
            for (uint32_t c = 0; c < width; ++c) rowW[r] += w[(size_t)r * width + c];
            if (ora_build_alias(w + (size_t)r * width, NULL, width, columns + (size_t)r * width) != 0) rc = -1;
        }
        if (rc == 0 && ora_build_alias(rowW, NULL, height, rows) != 0) rc = -1;
        *integrated = total;
    }
    free(w);
    free(rowW);
    return rc;
}
