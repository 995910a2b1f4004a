// Proof server: continuous batching of proof requests on one device.
//
// The reference proves one circuit per process (main_zk.cpp:11-25).  Here
// requests of any size arrive at any time and are admitted while others run:
//
// * One device region of the caller's HBM budget and one page-locked host
//   region are reserved when the server starts.  Every admitted proof takes
//   its whole working set (circuit, layer values, sumcheck tables, codewords,
//   Merkle layers, fft_gkr buffers) as one lease from each region, keeps it
//   resident until its transcript is on the host, and gives it back.  No
//   hipMalloc / hipFree runs while proofs are in flight (hipFree waits for
//   the whole device).
// * `slots` worker threads, each with its own non-blocking stream, take the
//   requests in arrival order.  A worker plans its request on the host
//   (circuit build, CSR transposes, layout: the exact byte count), then waits
//   for memory: requests waiting for memory are admitted in arrival order,
//   each as soon as its lease fits (first fit over the free ranges), except
//   that a later one whose lease fits may pass the oldest waiter while that
//   one does not fit, at most VIRGO_SERVE_BYPASS (default 8) times per
//   oldest waiter (vg::admit_queue), so small leases do not queue behind a
//   large one and the large one is not starved.
//   The proof then runs in order on the worker's stream, overlapping on the
//   device with every other admitted proof.
// * Each result is the transcript vg_prove makes for the same circuit alone:
//   a proof's kernels read and write only its own lease, and its challenges
//   depend only on its circuit.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>
#include "vg_internal.hpp"
#include "vg_range.hpp"

using clk = std::chrono::steady_clock;

namespace {

using vg::range_alloc;
int64_t up(int64_t n) { return vg::range_alloc::up(n); }
double ms(clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

// A proof's lease: bump allocation inside its reserved ranges; frees are
// no-ops (the whole lease goes back at once).
struct lease : vg::mem_source {
    uint8_t *d = nullptr, *h = nullptr;
    int64_t dn = 0, hn = 0, doff = 0, hoff = 0;
    void *device(int64_t b) override
    {
        doff = up(doff);
        if (doff + b > dn) return nullptr;
        void *p = d + doff;
        doff += b;
        return p;
    }
    void device_free(void *) override {}
    void *pinned(int64_t b) override
    {
        hoff = up(hoff);
        if (hoff + b > hn) return nullptr;
        void *p = h + hoff;
        hoff += b;
        return p;
    }
    void pinned_free(void *) override {}
    bool async_uploads() const override { return true; }
};

struct job {
    int64_t id = 0;
    vg_circuit_desc desc = {};
    int state = 0;                         // 0 queued, 1 planning / waiting for memory / proving, 2 done
    int rc = 0;
    std::string err;
    uint8_t *tr = nullptr;
    size_t len = 0;
    int64_t dev_bytes = 0;
    clk::time_point t_submit, t_start, t_admit, t_done;
};

}  // namespace

struct vg_server {
    int device = 0;
    uint8_t *dbase = nullptr, *hbase = nullptr;
    int64_t dsize = 0, hsize = 0;
    range_alloc dpool, hpool;
    std::mutex mu;
    std::condition_variable cv_work, cv_mem, cv_done;
    std::deque<job *> queue;
    std::map<int64_t, job *> jobs;
    vg::admit_queue waiting;               // requests planned and waiting for memory
    int64_t next_id = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, running = 0, in_use = 0, peak_bytes = 0;
    int admitted = 0, peak_admitted = 0;
};

namespace {

void worker(vg_server *s, hipStream_t st)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            if (s->queue.empty()) return;
            j = s->queue.front();
            s->queue.pop_front();
            j->state = 1;
            j->t_start = clk::now();
            ++s->running;
        }
        vg_ctx *c = nullptr;
        int rc = ctx_create_planned(&c, &j->desc, st);
        if (!rc) {
            int64_t dn = 0, hn = 0;
            ctx_needs(c, &dn, &hn);
            j->dev_bytes = dn;
            if (dn > s->dsize || hn > s->hsize) {
                vg_ctx_destroy(c);
                char b[256];
                snprintf(b, sizeof b, "proof needs %lld device / %lld pinned bytes; the server holds %lld / %lld",
                         (long long)dn, (long long)hn, (long long)s->dsize, (long long)s->hsize);
                vg_set_error(b);
                rc = -21;
            } else {
                int64_t doff = -1, hoff = -1;
                {
                    std::unique_lock<std::mutex> lk(s->mu);
                    s->waiting.wait(j->id);
                    s->cv_mem.wait(lk, [&] {
                        if (!s->waiting.may_try(j->id)) return false;
                        doff = s->dpool.take(dn);
                        if (doff < 0) return false;
                        hoff = s->hpool.take(hn);
                        if (hoff < 0) { s->dpool.give(doff, dn); doff = -1; return false; }
                        return true;
                    });
                    s->waiting.admitted(j->id);
                    s->in_use += up(dn);
                    s->peak_bytes = std::max(s->peak_bytes, s->in_use);
                    ++s->admitted;
                    s->peak_admitted = std::max(s->peak_admitted, s->admitted);
                }
                s->cv_mem.notify_all();            // the next waiter may fit as well
                j->t_admit = clk::now();
                lease L;
                L.d = s->dbase + doff; L.dn = up(dn);
                L.h = s->hbase + hoff; L.hn = up(hn);
                rc = ctx_materialize(c, &L, &c);   // destroys c on failure
                if (!rc) {
                    rc = vg_prove(c, &j->tr, &j->len, nullptr);
                    vg_ctx_destroy(c);             // waits for the stream before the lease goes back
                }
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->dpool.give(doff, dn);
                    s->hpool.give(hoff, hn);
                    s->in_use -= up(dn);
                    --s->admitted;
                }
                s->cv_mem.notify_all();
            }
        }
        if (rc) {
            j->err = vg_last_error();
            (void)hipGetLastError();
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            j->rc = rc;
            j->t_done = clk::now();
            if (!j->t_admit.time_since_epoch().count()) j->t_admit = j->t_done;
            j->state = 2;
            --s->running;
            ++s->completed;
            if (rc) ++s->failed;
        }
        s->cv_done.notify_all();
    }
}

}  // namespace

#define SFAIL(code, ...) do { char b_[512]; snprintf(b_, sizeof b_, __VA_ARGS__); vg_set_error(b_); return (code); } while (0)

int vg_server_create(vg_server **out, int64_t device_budget, int64_t pinned_budget, int slots)
{
    if (!out) SFAIL(-1, "vg_server_create: null argument");
    *out = nullptr;
    if (device_budget <= 0 || slots < 1 || slots > 64) SFAIL(-1, "vg_server_create: bad budget or slot count");
    if (pinned_budget <= 0) pinned_budget = (int64_t)2 << 30;   // staging of uploads and transcripts
    vg_server *s = new vg_server();
    if (const char *e = getenv("VIRGO_SERVE_BYPASS")) s->waiting.max_bypass = std::max(0, atoi(e));
    s->dsize = up(device_budget);
    s->hsize = up(pinned_budget);
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipMalloc(&s->dbase, s->dsize);
    if (e == hipSuccess) e = hipHostMalloc(&s->hbase, s->hsize, hipHostMallocDefault);
    for (int k = 0; k < slots && e == hipSuccess; ++k) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) s->streams.push_back(st);
    }
    if (e == hipSuccess) {
        // touch every page of the page-locked region once, on a few threads:
        // its pages are mapped on first touch, and a proof packing its
        // uploads into an untouched range wrote at about 2 GB/s instead of
        // about 20
        constexpr int64_t pg = 4096;
        constexpr int nt = 8;
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([s, t] {
                for (int64_t o = (int64_t)t * pg; o < s->hsize; o += (int64_t)nt * pg) s->hbase[o] = 0;
            });
        for (auto &x : th) x.join();
    }
    if (e != hipSuccess) {
        std::string m = std::string("vg_server_create: ") + hipGetErrorString(e);
        for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
        if (s->dbase) (void)hipFree(s->dbase);
        if (s->hbase) (void)hipHostFree(s->hbase);
        delete s;
        vg_set_error(m);
        return -20;
    }
    s->dpool.init(s->dsize);
    s->hpool.init(s->hsize);
    for (hipStream_t st : s->streams) s->workers.emplace_back(worker, s, st);
    *out = s;
    return 0;
}

int vg_server_submit(vg_server *s, const vg_circuit_desc *c, int64_t *id)
{
    if (!s || !c || !id) SFAIL(-1, "vg_server_submit: null argument");
    job *j = new job();
    j->desc = *c;
    j->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) { delete j; SFAIL(-1, "vg_server_submit: server is shutting down"); }
        j->id = s->next_id++;
        s->jobs[j->id] = j;
        s->queue.push_back(j);
        ++s->submitted;
        *id = j->id;
    }
    s->cv_work.notify_one();
    return 0;
}

int vg_server_wait(vg_server *s, int64_t id, uint8_t **transcript, size_t *len, double *times_ms)
{
    if (!s || !transcript || !len) SFAIL(-1, "vg_server_wait: null argument");
    *transcript = nullptr;
    *len = 0;
    job *j = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) SFAIL(-1, "vg_server_wait: unknown or already collected request %lld", (long long)id);
        j = it->second;
        s->cv_done.wait(lk, [&] { return j->state == 2; });
        s->jobs.erase(it);
    }
    if (times_ms) {
        times_ms[0] = ms(j->t_submit, j->t_done);     // latency
        times_ms[1] = ms(j->t_submit, j->t_start);    // queued for a slot
        times_ms[2] = ms(j->t_start, j->t_admit);     // host planning + waiting for memory
        times_ms[3] = ms(j->t_admit, j->t_done);      // upload + proof on the device
        times_ms[4] = (double)j->dev_bytes;
    }
    int rc = j->rc;
    if (rc) vg_set_error(j->err);
    else { *transcript = j->tr; *len = j->len; }
    delete j;
    return rc;
}

int vg_server_stats(vg_server *s, vg_server_stats_t *out)
{
    if (!s || !out) SFAIL(-1, "vg_server_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->queued = (int64_t)s->queue.size();
    out->running = s->running;
    out->device_in_use = s->in_use;
    out->device_peak = s->peak_bytes;
    out->peak_admitted = s->peak_admitted;
    out->device_budget = s->dsize;
    return 0;
}

int vg_server_destroy(vg_server *s)
{
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &kv : s->jobs) {
        free(kv.second->tr);
        delete kv.second;
    }
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    (void)hipFree(s->dbase);
    (void)hipHostFree(s->hbase);
    delete s;
    return 0;
}
