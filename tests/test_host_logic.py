"""Host-side logic of the library built alone with g++ (no GPU): the output
arenas' range allocator and the compressed walk's node quantization.

* Range allocator: a seeded stress of allocations and releases checked
  against a unit bitmap (no overlap, in-use bytes, first fit, full
  coalescing once everything is released).
* Quantization (renderingengine_amd/csrc/wide4q.hpp): on random 4-wide
  nodes of extents 1e-3..1e3, every dequantized child box contains its exact
  box (exact arithmetic), empty slots quantize to an empty box, and for rays
  from up to 30 units away the compressed slab test (one fp32 fma per plane,
  as in the kernel) accepts every ray the exact slab test accepts and never
  enters later -- the property the compressed walk's hit equality rests on.
"""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "renderingengine_amd", "csrc")

DRIVER = r'''
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <random>
#include <vector>
#include "range_allocator.hpp"
#include "wide4q.hpp"
using namespace rex;

static int allocator_stress() {
    RangeAllocator ra;
    const size_t units = 4096;
    ra.reset(units * RangeAllocator::kAlign + 100);          // rounds down to whole units
    if (ra.total() != units * RangeAllocator::kAlign) return 1;
    std::vector<int> owner(units, -1);
    std::vector<std::pair<size_t, size_t>> live;             // (offset, bytes)
    std::mt19937 rng(7);
    for (int it = 0; it < 200000; ++it) {
        if (live.empty() || rng() % 3 != 0) {
            const size_t bytes = 1 + rng() % (20 * RangeAllocator::kAlign);
            const size_t need = RangeAllocator::round(bytes) / RangeAllocator::kAlign;
            // first fit by the bitmap
            size_t expect = RangeAllocator::kNone, run = 0;
            for (size_t u = 0; u < units; ++u) {
                run = owner[u] < 0 ? run + 1 : 0;
                if (run == need) { expect = (u + 1 - need) * RangeAllocator::kAlign; break; }
            }
            const size_t off = ra.alloc(bytes);
            if (off != expect) { printf("first fit: got %zu expected %zu\n", off, expect); return 2; }
            if (off == RangeAllocator::kNone) continue;
            for (size_t u = off / RangeAllocator::kAlign; u < off / RangeAllocator::kAlign + need; ++u) owner[u] = it;
            live.push_back({off, need * RangeAllocator::kAlign});
        } else {
            const size_t k = rng() % live.size();
            if (!ra.release(live[k].first)) return 3;
            for (size_t u = live[k].first / RangeAllocator::kAlign;
                 u < (live[k].first + live[k].second) / RangeAllocator::kAlign; ++u) owner[u] = -1;
            live.erase(live.begin() + k);
        }
        size_t used = 0;
        for (auto& l : live) used += l.second;
        if (used != ra.in_use()) return 4;
    }
    if (ra.release(12345)) return 5;                          // not an allocation
    for (auto& l : live) ra.release(l.first);
    if (ra.in_use() != 0 || ra.largest_free() != ra.total()) return 6;
    return 0;
}

static float byte_of(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); }

static int quantize_check(long* rays, long* hits) {
    std::mt19937 rng(1);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    for (int it = 0; it < 100000; ++it) {
        Wide4 W{};
        const float scale = std::pow(10.f, U(rng) * 6 - 3), center = (U(rng) - 0.5f) * 20;
        const int nvalid = 1 + (int)(U(rng) * 4);
        float* lo[3] = {&W.lox.x, &W.loy.x, &W.loz.x};
        float* hi[3] = {&W.hix.x, &W.hiy.x, &W.hiz.x};
        for (int k = 0; k < 4; ++k) {
            (&W.ref.x)[k] = k < nvalid ? (uint32_t)k : kInvalid;
            for (int a = 0; a < 3; ++a) {
                float x0 = center + (U(rng) - 0.5f) * scale, x1 = center + (U(rng) - 0.5f) * scale;
                if (U(rng) < 0.1f) x1 = x0;
                lo[a][k] = std::min(x0, x1);
                hi[a][k] = std::max(x0, x1);
            }
        }
        const Wide4Q q = quantize_wide4(W, 30.0);
        uint32_t ebits;
        std::memcpy(&ebits, &q.origin.w, 4);
        const float org[3] = {q.origin.x, q.origin.y, q.origin.z};
        const uint32_t qlo[3] = {q.planeA.x, q.planeA.y, q.planeA.z}, qhi[3] = {q.planeA.w, q.planeB.x, q.planeB.y};
        for (int k = 0; k < 4; ++k) {
            for (int a = 0; a < 3; ++a) {
                const double step = std::ldexp(1.0, (int)((ebits >> (8 * a)) & 0xffu) - 127);
                const double dl = (double)org[a] + byte_of(qlo[a], k) * step;
                const double dh = (double)org[a] + byte_of(qhi[a], k) * step;
                if (k >= nvalid) {
                    if (!(dl > dh)) return 10;                    // empty slot -> empty box
                } else if (!(dl < lo[a][k]) || !(dh > hi[a][k])) {
                    return 11;                                    // containment
                }
            }
        }
        if (q.ref.x != W.ref.x || q.ref.w != W.ref.w) return 12;
        for (int r = 0; r < 4; ++r) {
            float p[3], o[3], d[3], inv[3], len = 0.f;
            for (int a = 0; a < 3; ++a) p[a] = lo[a][0] + (hi[a][0] - lo[a][0]) * U(rng);
            for (int a = 0; a < 3; ++a) o[a] = p[a] + (U(rng) - 0.5f) * 60.f;
            for (int a = 0; a < 3; ++a) { d[a] = p[a] - o[a]; len += d[a] * d[a]; }
            len = std::sqrt(len);
            for (int a = 0; a < 3; ++a) { d[a] /= len; inv[a] = 1.0f / d[a]; }
            float ne = -INFINITY, fe = INFINITY, nq = -INFINITY, fq = INFINITY;
            for (int a = 0; a < 3; ++a) {
                const float t0 = (lo[a][0] - o[a]) * inv[a], t1 = (hi[a][0] - o[a]) * inv[a];
                ne = std::max(ne, std::min(t0, t1));
                fe = std::min(fe, std::max(t0, t1));
                float s;
                const uint32_t sb = ((ebits >> (8 * a)) & 0xffu) << 23;
                std::memcpy(&s, &sb, 4);
                const float A = s * inv[a], B = (org[a] - o[a]) * inv[a];
                const bool neg = inv[a] < 0.f;
                nq = std::max(nq, std::fma(byte_of(neg ? qhi[a] : qlo[a], 0), A, B));
                fq = std::min(fq, std::fma(byte_of(neg ? qlo[a] : qhi[a], 0), A, B));
            }
            const bool ex = !(fe < ne) && fe > 0.f;
            const bool qq = !(fq < nq) && fq > 0.f;
            ++*rays;
            *hits += ex;
            if (ex && (!qq || nq > ne)) return 13;              // not conservative
        }
    }
    return 0;
}

int main() {
    const int a = allocator_stress();
    long rays = 0, hits = 0;
    const int q = quantize_check(&rays, &hits);
    printf("allocator %d quantize %d rays %ld hits %ld\n", a, q, rays, hits);
    return a || q;
}
'''


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("hostlogic")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "driver"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    return exe


def test_range_allocator_and_quantization(driver):
    res = subprocess.run([str(driver)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    fields = res.stdout.split()
    assert fields[1] == "0" and fields[3] == "0"
    rays, hits = int(fields[5]), int(fields[7])
    assert rays == 400000 and hits > 390000   # rays aimed at child 0's box nearly all hit it
