// Host-side helpers shared by the prover drivers: the reference's challenge
// generator and a byte buffer for transcripts.
#pragma once
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <algorithm>
#include <new>
#include <vector>
#include <hip/hip_runtime.h>
#include "vg_field.hpp"
#include "vg_pool.hpp"

namespace vg {

// ---------------------------------------------------------------------------
// glibc rand() TYPE_3 (additive feedback, r[i] = r[i-3] + r[i-31]), seed 1;
// prime_field::random (prime_field.cpp:35-41) draws real, then img.
// ---------------------------------------------------------------------------
struct glibc_rand {
    uint32_t s[34];
    uint64_t k;
    explicit glibc_rand(uint32_t seed)
    {
        int32_t r[34];
        r[0] = (int32_t)(seed ? seed : 1);
        for (int i = 1; i < 31; ++i) {
            int64_t t = (16807LL * r[i - 1]) % 2147483647LL;
            if (t < 0) t += 2147483647LL;
            r[i] = (int32_t)t;
        }
        for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
        for (int i = 0; i < 34; ++i) s[i] = (uint32_t)r[i];
        k = 34;
        for (int i = 0; i < 310; ++i) next();
    }
    uint32_t next()
    {
        uint32_t v = s[(k - 31) % 34] + s[(k - 3) % 34];
        s[k % 34] = v;
        ++k;
        return v >> 1;
    }
    fe field()
    {
        fe r;
        r.real = next() % P;
        r.img = next() % P;
        return r;
    }
};

// ---------------------------------------------------------------------------
// Where a proof context takes its device arena and pinned host staging from:
// hipMalloc / hipHostMalloc for a context of its own (hip_mem()), or a lease
// carved from a region a proof server reserved once (no allocation, and no
// hipFree / hipHostFree, which would wait for the whole device, while other
// proofs run).  A context asks for its memory in a few large blocks at
// creation and frees them at destruction.
// ---------------------------------------------------------------------------
struct mem_source {
    virtual ~mem_source() {}
    virtual void *device(int64_t bytes) = 0;     // nullptr: no memory
    virtual void device_free(void *p) = 0;
    virtual void *pinned(int64_t bytes) = 0;     // page-locked host memory
    virtual void pinned_free(void *p) = 0;
    // true: a context stages its set-up uploads through page-locked memory
    // of this source and copies them asynchronously on its stream (see
    // uploader); the planned pinned byte count then includes the staging
    virtual bool async_uploads() const { return false; }
};
mem_source *hip_mem();

// ---------------------------------------------------------------------------
// Host-to-device uploads of a context's static data (gates, descriptors,
// tables) at creation.  A context of its own copies each array with a
// synchronous hipMemcpy.  A context in a proof server's lease packs the
// bytes into page-locked staging taken from the lease and enqueues
// hipMemcpyAsync on its own stream: a synchronous copy runs on the null
// stream, whose hardware queue other proofs share, so each one would wait
// for whatever work of other proofs is queued ahead of it.  The staging
// stays part of the lease until the context is destroyed (after its stream
// is drained).  stage_bytes(n) is what put() of n bytes takes from it.
// Staged puts whose destinations lie at the same distances apart as their
// staging (256-byte aligned blocks carved back to back, like the staging)
// become one copy; flush() enqueues the pending copy and must come before
// any launch that reads what was put.  A staged put of 1 MB or more (a gate
// layer's arrays) is packed at the next flush, all such puts together on the
// host pool: its source must stay valid until then.
// ---------------------------------------------------------------------------
struct uploader {
    uint8_t *stage = nullptr;
    int64_t cap = 0, off = 0;
    hipStream_t st = nullptr;
    uint8_t *run_dst = nullptr;                  // pending copy: staging [run_off, run_end) to run_dst
    int64_t run_off = 0, run_end = 0;
    struct piece { int64_t off; const void *src; int64_t n; };
    std::vector<piece> deferred;                 // large puts not yet packed
    static constexpr int64_t kDefer = (int64_t)1 << 20;
    static int64_t stage_bytes(int64_t n) { return (n + 255) & ~(int64_t)255; }
    // takes `bytes` of staging from m when it stages; false: no memory
    bool init(mem_source *m, int64_t bytes, hipStream_t s)
    {
        st = s;
        if (!m->async_uploads() || bytes <= 0) return true;
        stage = (uint8_t *)m->pinned(bytes);
        cap = bytes;
        return stage != nullptr;
    }
    hipError_t put(void *dst, const void *src, int64_t n)
    {
        if (n <= 0) return hipSuccess;
        if (!stage) return hipMemcpy(dst, src, (size_t)n, hipMemcpyHostToDevice);
        if (off + n > cap) return hipErrorInvalidValue;      // more than the plan counted
        hipError_t e = hipSuccess;
        if (run_dst && (uint8_t *)dst != run_dst + (off - run_off)) e = flush();
        if (!run_dst) { run_dst = (uint8_t *)dst; run_off = off; }
        if (n >= kDefer) deferred.push_back({off, src, n});
        else memcpy(stage + off, src, (size_t)n);
        run_end = off + n;                       // the padding after the last block is not copied
        off += stage_bytes(n);
        return e;
    }
    // packs every deferred put, in 1 MB pieces over the host pool
    void pack()
    {
        if (deferred.empty()) return;
        std::vector<piece> pc;
        for (const piece &d : deferred)
            for (int64_t o = 0; o < d.n; o += kDefer)
                pc.push_back({d.off + o, (const uint8_t *)d.src + o, std::min(kDefer, d.n - o)});
        deferred.clear();
        parallel_for((int64_t)pc.size(), [&](int64_t k) { memcpy(stage + pc[k].off, pc[k].src, (size_t)pc[k].n); });
    }
    hipError_t flush()
    {
        pack();
        if (!run_dst) return hipSuccess;
        hipError_t e = hipMemcpyAsync(run_dst, stage + run_off, (size_t)(run_end - run_off), hipMemcpyHostToDevice, st);
        run_dst = nullptr;
        return e;
    }
};

// transcript bytes: F = 16 bytes {img, real}, little-endian u32.  The bytes
// live in malloc'd memory that release() hands to the caller as is (freed
// with free(), i.e. vg_free), so a finished transcript is never copied.
struct tbuf {
    uint8_t *p = nullptr;
    size_t n = 0, cap = 0;
    tbuf() = default;
    tbuf(const tbuf &) = delete;
    tbuf &operator=(const tbuf &) = delete;
    ~tbuf() { free(p); }
    void reserve(size_t c)
    {
        if (c <= cap) return;
        uint8_t *q = (uint8_t *)realloc(p, c);
        if (!q) throw std::bad_alloc();
        p = q;
        cap = c;
    }
    void put(const void *src, size_t k)
    {
        if (n + k > cap) reserve(std::max(2 * cap, n + k + 4096));
        memcpy(p + n, src, k);
        n += k;
    }
    void f(const fe &x) { put(&x, 16); }
    void u32(uint32_t x) { put(&x, 4); }
    void append(const tbuf &o) { if (o.n) put(o.p, o.n); }
    size_t size() const { return n; }
    uint8_t *release()
    {
        uint8_t *q = p;
        p = nullptr;
        n = cap = 0;
        return q;
    }
};

}  // namespace vg
