// Checks the proof server's first-fit range allocator (virgo_amd/csrc/
// vg_range.hpp) against a brute-force model: a bitmap of 256-byte units
// where first fit is the lowest-offset run of free units long enough.
// Random take / give sequences; every offset must equal the model's, live
// ranges never overlap, and once everything is given back the region is one
// free range again.  Prints "ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "../../virgo_amd/csrc/vg_range.hpp"

int main()
{
    const int64_t U = vg::range_alloc::ALIGN, units = 4096, size = units * U;
    std::mt19937_64 rng(7);
    for (int trial = 0; trial < 20; ++trial) {
        vg::range_alloc A;
        A.init(size);
        std::vector<char> used(units, 0);
        std::vector<std::pair<int64_t, int64_t>> live;   // offset, bytes
        for (int step = 0; step < 4000; ++step) {
            bool take = live.empty() || rng() % 100 < 55;
            if (take) {
                int64_t n = 1 + (int64_t)(rng() % (trial % 2 ? 200000 : 20000));
                int64_t nu = (n + U - 1) / U;
                int64_t want = -1;
                for (int64_t s = 0, run = 0; s < units; ++s) {
                    run = used[s] ? 0 : run + 1;
                    if (run == nu) { want = s - nu + 1; break; }
                }
                int64_t got = A.take(n);
                if (got != (want < 0 ? -1 : want * U)) {
                    printf("trial %d step %d: take(%lld) = %lld, model %lld\n", trial, step, (long long)n, (long long)got,
                           (long long)(want * U));
                    return 1;
                }
                if (got >= 0) {
                    for (int64_t k = 0; k < nu; ++k) {
                        if (used[got / U + k]) { printf("overlap\n"); return 1; }
                        used[got / U + k] = 1;
                    }
                    live.push_back({got, n});
                }
            } else {
                size_t i = rng() % live.size();
                int64_t off = live[i].first, n = live[i].second;
                A.give(off, n);
                for (int64_t k = 0; k < (n + U - 1) / U; ++k) used[off / U + k] = 0;
                live[i] = live.back();
                live.pop_back();
            }
        }
        for (auto &l : live) A.give(l.first, l.second);
        if (A.free_.size() != 1 || A.free_.begin()->first != 0 || A.free_.begin()->second != size || A.largest() != size) {
            printf("trial %d: region not whole after giving everything back\n", trial);
            return 1;
        }
    }
    printf("ok\n");
    return 0;
}
