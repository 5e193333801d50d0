// Host-side first-fit allocator over one device allocation (the serving
// loop's slot arena).  Plain C++: no HIP calls, so it is unit-tested on the
// CPU.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <map>

namespace mlmq {

constexpr uint64_t kArenaAlign = 256;

inline uint64_t round_up(uint64_t x) { return (x + kArenaAlign - 1) & ~(kArenaAlign - 1); }

// First-fit allocator over one device allocation, with coalescing of
// neighbouring free ranges.
struct Arena {
    uint8_t *base = nullptr;
    uint64_t size = 0, used = 0, peak = 0;
    std::map<uint64_t, uint64_t> free_;   // offset -> bytes

    void reset(uint8_t *b, uint64_t n) {
        base = b;
        size = n;
        used = peak = 0;
        free_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(uint64_t bytes, uint64_t *off) {
        bytes = round_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const uint64_t o = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[o + bytes] = rest;
            *off = o;
            used += bytes;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = round_up(bytes);
        used -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                bytes += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            free_.erase(next);
        }
        free_[off] = bytes;
    }
};

}  // namespace mlmq
