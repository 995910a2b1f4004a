// Host work pool and recycled host blocks (see vg_pool.hpp).
#include "vg_pool.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace vg {
namespace {

// The pool lives for the whole process: it is never destroyed, so no thread
// waits on a mutex or condition variable that static destruction has torn
// down at exit.
struct pool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> tasks;
    int width = 1;
    int started = 0;
    void run()
    {
        for (;;) {
            std::function<void()> t;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !tasks.empty(); });
                t = std::move(tasks.front());
                tasks.pop_front();
            }
            t();
        }
    }
};

pool *the_pool()
{
    static pool *p = [] {
        pool *q = new pool();
        int hw = (int)std::thread::hardware_concurrency();
        int w = std::max(1, std::min(16, hw));
        if (const char *e = getenv("VIRGO_HOST_THREADS")) w = std::max(1, std::min(64, atoi(e)));
        q->width = w;
        return q;
    }();
    return p;
}

// one parallel_for: items are claimed from `next`; the caller waits until
// `done` reaches n
struct job {
    std::atomic<int64_t> next{0}, done{0};
    int64_t n = 0;
    const std::function<void(int64_t)> *fn = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    void work()
    {
        int64_t k;
        while ((k = next.fetch_add(1)) < n) {
            (*fn)(k);
            if (done.fetch_add(1) + 1 == n) {
                std::lock_guard<std::mutex> lk(mu);
                cv.notify_all();
            }
        }
    }
};

}  // namespace

int pool_width() { return the_pool()->width; }

void parallel_for(int64_t n, const std::function<void(int64_t)> &fn)
{
    if (n <= 0) return;
    pool *p = the_pool();
    int helpers = (int)std::min<int64_t>(p->width - 1, n - 1);
    if (helpers <= 0) {
        for (int64_t k = 0; k < n; ++k) fn(k);
        return;
    }
    auto j = std::make_shared<job>();
    j->n = n;
    j->fn = &fn;       // alive until done == n: a helper only calls it for a claimed k < n
    {
        std::lock_guard<std::mutex> lk(p->mu);
        while (p->started < p->width - 1) {
            std::thread([p] { p->run(); }).detach();
            ++p->started;
        }
        for (int h = 0; h < helpers; ++h) p->tasks.push_back([j] { j->work(); });
    }
    p->cv.notify_all();
    j->work();
    std::unique_lock<std::mutex> lk(j->mu);
    j->cv.wait(lk, [&] { return j->done.load() == n; });
}

namespace {
struct block_cache {
    std::mutex mu;
    std::multimap<size_t, void *> free;
    size_t held = 0;
    size_t cap = (size_t)4 << 30;
};
block_cache *cache()
{
    static block_cache *c = new block_cache();   // lives until the process ends, like the pool
    return c;
}
constexpr size_t kMinCached = (size_t)1 << 20;
}  // namespace

void *host_take(size_t bytes)
{
    if (bytes >= kMinCached) {
        block_cache *c = cache();
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->free.find(bytes);
        if (it != c->free.end()) {
            void *p = it->second;
            c->free.erase(it);
            c->held -= bytes;
            return p;
        }
    }
    return malloc(bytes ? bytes : 1);
}

void host_give(void *p, size_t bytes)
{
    if (!p) return;
    if (bytes >= kMinCached) {
        block_cache *c = cache();
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->held + bytes <= c->cap) {
            c->free.emplace(bytes, p);
            c->held += bytes;
            return;
        }
    }
    free(p);
}

}  // namespace vg
