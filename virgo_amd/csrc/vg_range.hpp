// First-fit allocator over the offsets [0, size) of a region reserved once
// (the proof server's HBM and pinned host regions): ranges are 256-byte
// aligned, freed ranges coalesce with their free neighbours.  Not
// thread-safe (the server holds its lock).  Host code only.
#pragma once
#include <stdint.h>
#include <iterator>
#include <map>

namespace vg {

struct range_alloc {
    static constexpr int64_t ALIGN = 256;
    static int64_t up(int64_t n) { return (n + ALIGN - 1) & ~(ALIGN - 1); }
    std::map<int64_t, int64_t> free_;      // offset -> length of each free range
    void init(int64_t n) { free_.clear(); if (n >= ALIGN) free_[0] = n & ~(ALIGN - 1); }
    // offset of a range of n bytes (rounded up to ALIGN), or -1
    int64_t take(int64_t n)
    {
        n = up(n);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                int64_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[off + n] = len - n;
                return off;
            }
        return -1;
    }
    // returns a range that take(n) gave
    void give(int64_t off, int64_t n)
    {
        n = up(n);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
    int64_t largest() const
    {
        int64_t m = 0;
        for (auto &kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }
};

}  // namespace vg
