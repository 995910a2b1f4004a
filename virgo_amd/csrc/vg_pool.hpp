// Host-side work sharing and memory reuse for the parts of a proof that run
// on the CPU (loading and planning a circuit, packing its uploads).  Every
// context of the process shares one pool, so a proof server's workers do
// not each start threads of their own.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <functional>

namespace vg {

// Runs fn(k) for every k in [0, n), on the calling thread and on up to
// pool_width() - 1 pool threads; returns when every call has returned.
// Calls may run in any order and concurrently.
void parallel_for(int64_t n, const std::function<void(int64_t)> &fn);
// threads parallel_for may use at once (the calling thread included):
// min(16, hardware threads), or VIRGO_HOST_THREADS
int pool_width();

// Large host blocks are kept for reuse when freed (up to a fixed total), so
// a context built after another of similar size writes into pages that are
// already mapped instead of faulting every page in again.
void *host_take(size_t bytes);
void host_give(void *p, size_t bytes);

}  // namespace vg
