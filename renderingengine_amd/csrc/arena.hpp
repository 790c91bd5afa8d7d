// Device output arena: one hipMalloc'd slab carved into ranges by a host-side
// first-fit allocator with coalescing.  Jobs of the batch engine and the
// server take their Film / sample / history buffers from it, so admitting
// and retiring a job never calls hipMalloc or hipFree (hipFree waits for the
// whole device, which would drain every stream of a serving loop).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace rex {

// Offsets of free ranges -> their sizes; all sizes and offsets are multiples
// of kAlign.
class RangeAllocator {
public:
    static constexpr size_t kAlign = 256;
    static constexpr size_t kNone = ~size_t(0);

    void reset(size_t bytes) {
        free_.clear();
        used_.clear();
        total_ = bytes / kAlign * kAlign;
        inUse_ = 0;
        if (total_ != 0) free_[0] = total_;
    }
    size_t total() const { return total_; }
    size_t in_use() const { return inUse_; }
    static size_t round(size_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

    // Lowest-offset free range that fits; kNone when none does.
    size_t alloc(size_t bytes) {
        const size_t need = round(bytes == 0 ? 1 : bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, left = it->second - need;
            free_.erase(it);
            if (left != 0) free_[off + need] = left;
            used_[off] = need;
            inUse_ += need;
            return off;
        }
        return kNone;
    }
    // Returns false for an offset that is not an allocation.
    bool release(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t start = off, size = u->second;
        inUse_ -= size;
        used_.erase(u);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && next->first == start + size) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                size += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = size;
        return true;
    }
    size_t largest_free() const {
        size_t m = 0;
        for (const auto& kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }

private:
    std::map<size_t, size_t> free_, used_;
    size_t total_ = 0, inUse_ = 0;
};

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
