/*
 * Oracle (test infrastructure only): stable queue compaction of the
 * wavefront integrator's PrefixScan mode.
 *
 * ora_blelloch_exclusive_scan restates BlellochExclusiveScan
 * (integrators/wavefront/src/WavefrontCompactionOracle.cpp:9-54): the input
 * padded with zero pairs to the next power of two, the up-sweep
 * (stride 1, 2, 4, ...: a[i] += a[i - stride] at i = 2*stride*k - 1), the
 * root cleared, the down-sweep (swap-and-add); non-binary flags are rejected.
 * ora_compact_stable restates CompactStable (:56-81): totals = last prefix +
 * last flag, every flagged index written at its prefix.
 *
 * Pins (tests/test_oracle_compaction.py): the reference's own vectors
 * (integrators/wavefront/tests/WavefrontCpuOracleTests.cpp:30-89: the
 * 5-element example, the stale-capacity tail, boundary counts 127 / 128 /
 * 129 / 4097 with index % 3 / index % 5 flags), the empty input, and the
 * closed form: an exclusive scan equals the serial prefix sum.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Returns 0, or -1 for a non-binary flag (std::invalid_argument in the
 * reference), -2 on allocation failure. flags / out: count (next, shadow)
 * uint32 pairs. */
int ora_blelloch_exclusive_scan(const uint32_t* flags, uint32_t count, uint32_t* out) {
    if (count == 0u) return 0;
    uint64_t padded = 1u;
    while (padded < count) padded <<= 1u;
    uint32_t* a = (uint32_t*)calloc((size_t)padded * 2u, sizeof(uint32_t));
    if (a == NULL) return -2;
    for (uint32_t i = 0; i < count; ++i) {
        if (flags[2u * i] > 1u || flags[2u * i + 1u] > 1u) {
            free(a);
            return -1;
        }
        a[2u * i] = flags[2u * i];
        a[2u * i + 1u] = flags[2u * i + 1u];
    }
    for (uint64_t stride = 1u; stride < padded; stride <<= 1u) {
        for (uint64_t i = stride * 2u - 1u; i < padded; i += stride * 2u) {
            a[2u * i] += a[2u * (i - stride)];
            a[2u * i + 1u] += a[2u * (i - stride) + 1u];
        }
    }
    a[2u * (padded - 1u)] = 0u;
    a[2u * (padded - 1u) + 1u] = 0u;
    for (uint64_t stride = padded >> 1u; stride != 0u; stride >>= 1u) {
        for (uint64_t i = stride * 2u - 1u; i < padded; i += stride * 2u) {
            const uint32_t lx = a[2u * (i - stride)], ly = a[2u * (i - stride) + 1u];
            a[2u * (i - stride)] = a[2u * i];
            a[2u * (i - stride) + 1u] = a[2u * i + 1u];
            a[2u * i] += lx;
            a[2u * i + 1u] += ly;
        }
    }
    memcpy(out, a, (size_t)count * 2u * sizeof(uint32_t));
    free(a);
    return 0;
}

/* totals: (next, shadow). nextIndices / shadowIndices: room for the totals
 * (count each is enough). Returns as ora_blelloch_exclusive_scan. */
int ora_compact_stable(const uint32_t* flags, uint32_t count, uint32_t* prefix, uint32_t* nextIndices,
                       uint32_t* shadowIndices, uint32_t* totals) {
    totals[0] = totals[1] = 0u;
    if (count == 0u) return 0;
    const int rc = ora_blelloch_exclusive_scan(flags, count, prefix);
    if (rc != 0) return rc;
    totals[0] = prefix[2u * (count - 1u)] + flags[2u * (count - 1u)];
    totals[1] = prefix[2u * (count - 1u) + 1u] + flags[2u * (count - 1u) + 1u];
    for (uint32_t i = 0; i < count; ++i) {
        if (flags[2u * i] != 0u) nextIndices[prefix[2u * i]] = i;
        if (flags[2u * i + 1u] != 0u) shadowIndices[prefix[2u * i + 1u]] = i;
    }
    return 0;
}
