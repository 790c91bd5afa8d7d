// Device output arena: one hipMalloc'd slab carved into ranges by a host-side
// first-fit allocator with coalescing.  Jobs of the batch engine and the
// server take their Film / sample / history buffers from it, so admitting
// and retiring a job never calls hipMalloc or hipFree (hipFree waits for the
// whole device, which would drain every stream of a serving loop).
#pragma once

#include <hip/hip_runtime.h>

#include "range_allocator.hpp"

namespace rex {

struct DeviceArena {
    char* base = nullptr;
    RangeAllocator ranges;

    hipError_t create(size_t bytes) {
        bytes = RangeAllocator::round(bytes);
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), bytes);
        if (e != hipSuccess) {
            base = nullptr;
            return e;
        }
        ranges.reset(bytes);
        return hipSuccess;
    }
    void destroy() {
        if (base) (void)hipFree(base);
        base = nullptr;
        ranges.reset(0);
    }
    // nullptr when no free range fits.
    void* alloc(size_t bytes) {
        const size_t off = ranges.alloc(bytes);
        return off == RangeAllocator::kNone ? nullptr : base + off;
    }
    bool release(void* p) {
        if (p == nullptr || base == nullptr) return false;
        return ranges.release(static_cast<size_t>(static_cast<char*>(p) - base));
    }
};

}  // namespace rex
