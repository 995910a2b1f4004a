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

namespace vg {

// Admission order of the proof server's requests that wait for memory:
// arrival order, except that a later request whose lease fits now may pass
// the oldest waiter while that one does not fit, at most max_bypass times;
// after that the oldest is admitted next.  Not thread-safe (the server holds
// its lock).  Host code only.
struct admit_queue {
    int max_bypass = 8;
    std::map<int64_t, int> waiting;        // id -> times passed while it was the oldest waiter
    void wait(int64_t id) { waiting.emplace(id, 0); }
    // may request id try to take its lease now?
    bool may_try(int64_t id) const
    {
        if (waiting.empty()) return false;
        auto head = waiting.begin();
        return head->first == id || head->second < max_bypass;
    }
    // request id took its lease
    void admitted(int64_t id)
    {
        auto head = waiting.begin();
        if (head != waiting.end() && head->first != id) ++head->second;
        waiting.erase(id);
    }
    void cancel(int64_t id) { waiting.erase(id); }
};

}  // namespace vg
