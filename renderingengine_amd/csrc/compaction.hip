// Stable queue compaction (the PrefixScan queue mode of the wavefront
// integrator): ScanBlocksCS / AddOffsetsCS / ScatterCS
// (integrators/wavefront/shaders/Scan.hlsl:5-113, AddOffsets.hlsl:3-25,
// Scatter.hlsl:3-59), whose result is the reference's CompactStable
// (integrators/wavefront/src/WavefrontCompactionOracle.cpp:56-81): an
// exclusive prefix of the (next, shadow) flag pairs, and every flagged
// element moved to its prefix position, in element order.
//
// MI355X design: the flags are bits, so a wave's prefix is a ballot and a
// popcount per component instead of the reference's 128-wide LDS Blelloch
// tree.  Three kernels:
//   k_scan_reduce   one 1024-element tile per block (4 passes of 256
//                   threads): tile totals from per-wave popcounts;
//   k_scan_tiles    one block: exclusive scan of the tile totals (wave
//                   shuffles + an LDS pass over the 16 wave totals), queue
//                   sizes min(total, capacity), overflow status;
//   k_scan_scatter  per tile again: position = tile offset + the wave
//                   offsets before this wave's pass (LDS) + popcount of the
//                   lanes below; records (or indices) written there.
// A non-binary flag is a fatal scan overflow and counts as (0, 0), as in
// ScanBlocksCS; positions at or beyond a queue's capacity are not written
// and set the queue's overflow bit, as in ScatterCS.
#include <hip/hip_runtime.h>

#include "rex_internal.hpp"

namespace rex {

namespace {

constexpr uint32_t kScanBlock = 256;
constexpr uint32_t kScanPasses = 4;
constexpr uint32_t kScanTile = kScanBlock * kScanPasses;
constexpr uint32_t kScanWaves = kScanBlock / 64;

struct CompactParams {
    const uint2* flags;
    const uint32_t* countPtr;   // elements (device)
    uint32_t count;             // elements when countPtr is NULL
    uint32_t capNext, capShadow;
    uint2* tiles;               // tile totals, then exclusive tile offsets
    uint32_t* nextCount;        // queue sizes written (may be NULL)
    uint32_t* shadowCount;
    uint32_t* status;           // attemptedNext, attemptedShadow, activeNext, activeShadow, fatal (may be NULL)
    uint2* prefix;              // exclusive prefix per element (may be NULL)
    // record mode
    const F4* denseRays;
    const uint32_t* densePath;
    const ShadowItem* denseShadows;
    F4* nextRays;
    uint32_t* nextPath;
    ShadowItem* shadows;
    // index mode (records NULL)
    uint32_t* nextIndices;
    uint32_t* shadowIndices;
};

__device__ __forceinline__ uint32_t element_count(const CompactParams& P) {
    return P.countPtr != nullptr ? *P.countPtr : P.count;
}

// The masked flag of element e: (0, 0) past the count and for a non-binary
// flag (which also raises the fatal scan-overflow bit).
__device__ __forceinline__ uint2 load_flag(const CompactParams& P, uint32_t e, uint32_t n) {
    if (e >= n) return make_uint2(0u, 0u);
    const uint2 f = P.flags[e];
    if (f.x > 1u || f.y > 1u) {
        if (P.status != nullptr) atomicOr(P.status + 4, REX_COMPACT_FATAL_SCAN_OVERFLOW);
        return make_uint2(0u, 0u);
    }
    return f;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(CompactParams P) {
    __shared__ uint32_t sums[2][kScanWaves];
    const uint32_t n = element_count(P);
    const uint32_t tile = blockIdx.x;
    if ((uint64_t)tile * kScanTile >= n) return;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t cx = 0, cy = 0;
#pragma unroll
    for (uint32_t p = 0; p < kScanPasses; ++p) {
        const uint32_t e = tile * kScanTile + p * kScanBlock + threadIdx.x;
        const uint2 f = load_flag(P, e, n);
        cx += (uint32_t)__popcll(__ballot(f.x != 0u));
        cy += (uint32_t)__popcll(__ballot(f.y != 0u));
    }
    if (lane == 0u) {
        sums[0][wave] = cx;
        sums[1][wave] = cy;
    }
    __syncthreads();
    if (threadIdx.x == 0u) {
        uint32_t tx = 0, ty = 0;
        for (uint32_t w = 0; w < kScanWaves; ++w) {
            tx += sums[0][w];
            ty += sums[1][w];
        }
        P.tiles[tile] = make_uint2(tx, ty);
    }
}

// Inclusive scan of a uint2 across the 64 lanes of a wave.
__device__ __forceinline__ uint2 wave_inclusive(uint2 v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t x = __shfl_up(v.x, d, 64), y = __shfl_up(v.y, d, 64);
        if (lane >= d) {
            v.x += x;
            v.y += y;
        }
    }
    return v;
}

constexpr uint32_t kTileScanBlock = 1024;

__global__ __launch_bounds__(kTileScanBlock) void k_scan_tiles(CompactParams P) {
    __shared__ uint2 waveTotals[kTileScanBlock / 64];
    __shared__ uint2 carryLds;
    const uint32_t n = element_count(P);
    const uint32_t tiles = (uint32_t)(((uint64_t)n + kScanTile - 1u) / kScanTile);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0u) carryLds = make_uint2(0u, 0u);
    __syncthreads();
    for (uint32_t base = 0; base < tiles; base += kTileScanBlock) {
        const uint32_t t = base + threadIdx.x;
        const uint2 v = t < tiles ? P.tiles[t] : make_uint2(0u, 0u);
        const uint2 inc = wave_inclusive(v);
        if (lane == 63u) waveTotals[wave] = inc;
        __syncthreads();
        if (wave == 0u) {
            const uint2 w = lane < kTileScanBlock / 64 ? waveTotals[lane] : make_uint2(0u, 0u);
            const uint2 wi = wave_inclusive(w);
            if (lane < kTileScanBlock / 64) waveTotals[lane] = make_uint2(wi.x - w.x, wi.y - w.y);
        }
        __syncthreads();
        const uint2 carry = carryLds;
        const uint2 wo = waveTotals[wave];
        if (t < tiles) P.tiles[t] = make_uint2(carry.x + wo.x + inc.x - v.x, carry.y + wo.y + inc.y - v.y);
        __syncthreads();
        if (threadIdx.x == kTileScanBlock - 1u) carryLds = make_uint2(carry.x + wo.x + inc.x, carry.y + wo.y + inc.y);
        __syncthreads();
    }
    if (threadIdx.x == 0u) {
        const uint2 total = carryLds;
        const uint32_t an = total.x < P.capNext ? total.x : P.capNext;
        const uint32_t as = total.y < P.capShadow ? total.y : P.capShadow;
        if (P.nextCount != nullptr) *P.nextCount = an;
        if (P.shadowCount != nullptr) *P.shadowCount = as;
        if (P.status != nullptr) {
            P.status[0] = total.x;
            P.status[1] = total.y;
            P.status[2] = an;
            P.status[3] = as;
            uint32_t fatal = 0;
            if (total.x > P.capNext) fatal |= REX_COMPACT_FATAL_NEXT_OVERFLOW;
            if (total.y > P.capShadow) fatal |= REX_COMPACT_FATAL_SHADOW_OVERFLOW;
            if (fatal != 0u) atomicOr(P.status + 4, fatal);
        }
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_scatter(CompactParams P) {
    __shared__ uint2 cnt[kScanPasses * kScanWaves];
    const uint32_t n = element_count(P);
    const uint32_t tile = blockIdx.x;
    if ((uint64_t)tile * kScanTile >= n) return;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint2 f[kScanPasses];
    uint64_t mx[kScanPasses], my[kScanPasses];
#pragma unroll
    for (uint32_t p = 0; p < kScanPasses; ++p) {
        const uint32_t e = tile * kScanTile + p * kScanBlock + threadIdx.x;
        // The flags were validated by k_scan_reduce; mask without reporting.
        uint2 v = e < n ? P.flags[e] : make_uint2(0u, 0u);
        if (v.x > 1u || v.y > 1u) v = make_uint2(0u, 0u);
        f[p] = v;
        mx[p] = __ballot(v.x != 0u);
        my[p] = __ballot(v.y != 0u);
        if (lane == 0u) cnt[p * kScanWaves + wave] = make_uint2((uint32_t)__popcll(mx[p]), (uint32_t)__popcll(my[p]));
    }
    __syncthreads();
    if (threadIdx.x == 0u) {
        // Exclusive prefix of the (pass, wave) counts in element order.
        uint2 run = P.tiles[tile];
        for (uint32_t k = 0; k < kScanPasses * kScanWaves; ++k) {
            const uint2 c = cnt[k];
            cnt[k] = run;
            run.x += c.x;
            run.y += c.y;
        }
    }
    __syncthreads();
    uint32_t overflow = 0;
#pragma unroll
    for (uint32_t p = 0; p < kScanPasses; ++p) {
        const uint32_t e = tile * kScanTile + p * kScanBlock + threadIdx.x;
        if (e >= n) continue;
        const uint2 o = cnt[p * kScanWaves + wave];
        const uint32_t px = o.x + (uint32_t)__popcll(mx[p] & below);
        const uint32_t py = o.y + (uint32_t)__popcll(my[p] & below);
        if (P.prefix != nullptr) P.prefix[e] = make_uint2(px, py);
        if (f[p].x != 0u) {
            if (px < P.capNext) {
                if (P.nextRays != nullptr) {
                    P.nextRays[2u * px] = P.denseRays[2u * e];
                    P.nextRays[2u * px + 1u] = P.denseRays[2u * e + 1u];
                    P.nextPath[px] = P.densePath[e];
                } else if (P.nextIndices != nullptr) {
                    P.nextIndices[px] = e;
                }
            } else {
                overflow |= REX_COMPACT_FATAL_NEXT_OVERFLOW;
            }
        }
        if (f[p].y != 0u) {
            if (py < P.capShadow) {
                if (P.shadows != nullptr) P.shadows[py] = P.denseShadows[e];
                else if (P.shadowIndices != nullptr) P.shadowIndices[py] = e;
            } else {
                overflow |= REX_COMPACT_FATAL_SHADOW_OVERFLOW;
            }
        }
    }
    if (overflow != 0u && P.status != nullptr) atomicOr(P.status + 4, overflow);
}

// Clears the status words with a kernel (no memset node in a captured graph).
__global__ void k_status_clear(uint32_t* status) {
    if (threadIdx.x < 5u) status[threadIdx.x] = 0u;
}

hipError_t launch_compact(CompactParams& P, uint32_t maxCount, hipStream_t stream) {
    if (P.status != nullptr) hipLaunchKernelGGL(k_status_clear, dim3(1), dim3(64), 0, stream, P.status);
    const uint32_t tiles = (uint32_t)(((uint64_t)maxCount + kScanTile - 1u) / kScanTile);
    if (tiles == 0u) {
        // Empty input: sizes 0 (the single-block pass writes them).
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kTileScanBlock), 0, stream, P);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kScanBlock), 0, stream, P);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kTileScanBlock), 0, stream, P);
    hipLaunchKernelGGL(k_scan_scatter, dim3(tiles), dim3(kScanBlock), 0, stream, P);
    return hipGetLastError();
}

}  // namespace

size_t compact_tile_bytes(uint32_t maxCount) {
    return (size_t)(((uint64_t)maxCount + kScanTile - 1u) / kScanTile) * sizeof(uint2);
}

hipError_t launch_compact_queues(const CompactQueues& Q, hipStream_t stream) {
    CompactParams P = {};
    P.flags = Q.flags;
    P.countPtr = Q.countPtr;
    P.capNext = Q.capacity;
    P.capShadow = Q.capacity;
    P.tiles = Q.tiles;
    P.nextCount = Q.nextCount;
    P.shadowCount = Q.shadowCount;
    P.denseRays = Q.denseRays;
    P.densePath = Q.densePath;
    P.denseShadows = Q.denseShadows;
    P.nextRays = Q.nextRays;
    P.nextPath = Q.nextPath;
    P.shadows = Q.shadows;
    return launch_compact(P, Q.capacity, stream);
}

hipError_t launch_compact_indices(const uint2* flags, uint32_t count, uint32_t capNext, uint32_t capShadow,
                                  uint2* tiles, uint2* prefix, uint32_t* nextIndices, uint32_t* shadowIndices,
                                  uint32_t* status, hipStream_t stream) {
    CompactParams P = {};
    P.flags = flags;
    P.count = count;
    P.capNext = capNext;
    P.capShadow = capShadow;
    P.tiles = tiles;
    P.status = status;
    P.prefix = prefix;
    P.nextIndices = nextIndices;
    P.shadowIndices = shadowIndices;
    return launch_compact(P, count, stream);
}

}  // namespace rex
