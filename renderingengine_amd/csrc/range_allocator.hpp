// First-fit range allocator with coalescing (host-side bookkeeping of a
// device slab; no HIP dependency, so host tests can build it alone).
#pragma once

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

}  // namespace rex
