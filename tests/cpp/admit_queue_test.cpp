// Checks the proof server's admission order (virgo_amd/csrc/vg_range.hpp,
// admit_queue) against a model: requests arrive in id order with random
// lease sizes into a region of fixed size; at each step every waiter the
// policy lets try, oldest first, takes its lease if it fits; running leases
// end at random.  Properties: (1) a waiter whose lease fits is never held
// back while the oldest waiter has been passed fewer than max_bypass times;
// (2) the oldest waiter is passed at most max_bypass times; (3) with
// max_bypass = 0 admission is strictly in arrival order.  Prints "ok".
#include <cstdio>
#include <random>
#include <vector>
#include "../../virgo_amd/csrc/vg_range.hpp"

int main()
{
    std::mt19937_64 rng(11);
    for (int mb : {0, 1, 3, 8}) {
        long bypasses = 0;
        for (int trial = 0; trial < 40; ++trial) {
            vg::admit_queue Q;
            Q.max_bypass = mb;
            const int64_t cap = 1000;
            int64_t used = 0;
            std::vector<int64_t> need(400);
            for (auto &x : need) x = 1 + (int64_t)(rng() % (trial % 2 ? 900 : 300));
            std::vector<std::pair<int64_t, int64_t>> run;   // id, bytes
            std::vector<int64_t> order;                     // admission order
            int64_t next = 0;
            int passed_head = 0;
            int64_t head_id = -1;
            for (int step = 0; step < 4000 && (next < (int64_t)need.size() || !Q.waiting.empty() || !run.empty()); ++step) {
                if (next < (int64_t)need.size() && rng() % 3 == 0) Q.wait(next++);
                if (!run.empty() && rng() % 4 == 0) {
                    size_t k = rng() % run.size();
                    used -= run[k].second;
                    run.erase(run.begin() + k);
                }
                // one admission pass, oldest first (the server wakes every waiter)
                std::vector<int64_t> ids;
                for (auto &kv : Q.waiting) ids.push_back(kv.first);
                for (int64_t id : ids) {
                    bool fits = used + need[id] <= cap;
                    int64_t head = Q.waiting.begin()->first;
                    if (head != head_id) { head_id = head; passed_head = 0; }
                    bool may = Q.may_try(id);
                    if (id != head && fits && passed_head < mb && !may) { printf("held back id %lld\n", (long long)id); return 1; }
                    if (!may || !fits) continue;
                    if (id != head) { ++passed_head; ++bypasses; }
                    if (passed_head > mb) { printf("oldest passed %d > %d times\n", passed_head, mb); return 1; }
                    Q.admitted(id);
                    used += need[id];
                    run.push_back({id, need[id]});
                    order.push_back(id);
                }
            }
            if (mb == 0)
                for (size_t k = 1; k < order.size(); ++k)
                    if (order[k] < order[k - 1]) { printf("out of arrival order with max_bypass 0\n"); return 1; }
        }
        if ((mb > 0) != (bypasses > 0)) { printf("max_bypass %d: %ld bypasses\n", mb, bypasses); return 1; }
    }
    printf("ok\n");
    return 0;
}
