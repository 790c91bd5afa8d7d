// Compressed 4-wide node construction (host only; no HIP dependency, so
// host tests can build it alone).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "rex_layout.hpp"

namespace rex {

// Conservative 8-bit quantization of a 4-wide node's child boxes.  Per axis
// the grid starts a margin below the node's box and steps by the smallest
// power of two that spans the box plus the margin in 255 steps; child planes
// round outward after moving out by the margin, so each quantized box
// contains the exact box with room for the rounding of the slab arithmetic.
// The margin is the larger of 2^-16 of the axis' magnitude plus extent and
// 2^-18 of sceneScale (the scene's largest coordinate magnitude plus extent):
// the slab distances of rays starting within a few scene sizes round by far
// less.  Empty slots get lo 255 > hi 0.
inline Wide4Q quantize_wide4(const Wide4& W, double sceneScale) {
    Wide4Q q{};
    const float* lo[3] = {&W.lox.x, &W.loy.x, &W.loz.x};
    const float* hi[3] = {&W.hix.x, &W.hiy.x, &W.hiz.x};
    const uint32_t* ref = &W.ref.x;
    uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    uint32_t exps = 0;
    float origin[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < 3; ++a) {
        double plo = 0.0, phi = 0.0;
        bool any = false;
        for (int k = 0; k < 4; ++k) {
            if (ref[k] == kInvalid) continue;
            plo = any ? std::min(plo, (double)lo[a][k]) : (double)lo[a][k];
            phi = any ? std::max(phi, (double)hi[a][k]) : (double)hi[a][k];
            any = true;
        }
        const double pad = std::max(std::max(std::ldexp(std::max(std::fabs(plo), std::fabs(phi)) + (phi - plo), -16),
                                             std::ldexp(sceneScale, -18)),
                                    1.0e-30);
        float o = static_cast<float>(plo - pad);
        if ((double)o > plo - pad) o = std::nextafter(o, -HUGE_VALF);
        const double span = (phi + pad) - (double)o;
        int e = -126;
        while (e < 127 && std::ldexp(255.0, e) < span) ++e;
        origin[a] = o;
        exps |= static_cast<uint32_t>(e + 127) << (8 * a);
        const double step = std::ldexp(1.0, e);
        for (int k = 0; k < 4; ++k) {
            uint32_t l = 255u, h = 0u;
            if (ref[k] != kInvalid) {
                const double fl = std::floor(((double)lo[a][k] - pad - (double)o) / step);
                const double fh = std::ceil(((double)hi[a][k] + pad - (double)o) / step);
                l = static_cast<uint32_t>(std::min(255.0, std::max(0.0, fl)));
                h = static_cast<uint32_t>(std::min(255.0, std::max(0.0, fh)));
            }
            qlo[a] |= l << (8 * k);
            qhi[a] |= h << (8 * k);
        }
    }
    float expBits;
    std::memcpy(&expBits, &exps, 4);
    q.origin = {origin[0], origin[1], origin[2], expBits};
    q.planeA = {qlo[0], qlo[1], qlo[2], qhi[0]};
    q.planeB = {qhi[1], qhi[2], 0u, 0u};
    q.ref = W.ref;
    return q;
}

}  // namespace rex
