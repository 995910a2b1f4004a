// Host driver of the MI355X Virgo prover and the C ABI of libvirgo_hip.so.
//
// One proof is one device-resident schedule on one stream.  The reference's
// challenge stream (glibc rand(), default seed 1) is drawn in the order of
// zk_verifier::verify (src/linear_gkr/zk_verifier.cpp:910-1253), which depends
// only on the circuit shape, so every challenge is uploaded before the first
// kernel and no round waits for the host.  The prover's mask bookkeeping
// (zk_prover.cpp:162-333 and the mask terms of every sumcheck round) is host
// scalar work on challenges only; it is computed beside the device work and
// added to the device round sums when the transcript is assembled.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vg_kernels.hpp"
#include "vg_host.hpp"
#include "vg_fftgkr.hpp"
#include "vg_internal.hpp"
#include "vg_pool.hpp"
#include "../../include/virgo.h"

using namespace vg;

static thread_local std::string g_err;
void vg_set_error(const std::string &msg) { g_err = msg; }
#define FAILV(code, ...) do { char b_[512]; snprintf(b_, sizeof b_, __VA_ARGS__); g_err = b_; return (code); } while (0)
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) FAILV(-20, "%s: %s", #x, hipGetErrorString(e_)); } while (0)

namespace {

// ---------------------------------------------------------------------------
// circuit, as zk_verifier::read_circuit builds it (zk_verifier.cpp:15-217)
// ---------------------------------------------------------------------------
// std::allocator whose resize() leaves new elements uninitialised (the
// loader writes every gate it keeps, so a layer's arrays are not zeroed
// first: hundreds of MB for the largest circuits) and whose large blocks
// come from the process's recycled host blocks (no page faults for a
// circuit loaded after another of similar size)
template <class T> struct noinit_alloc : std::allocator<T> {
    using value_type = T;
    template <class U> struct rebind { using other = noinit_alloc<U>; };
    noinit_alloc() = default;
    template <class U> noinit_alloc(const noinit_alloc<U> &) {}
    T *allocate(size_t n) { return (T *)host_take(n * sizeof(T)); }
    void deallocate(T *p, size_t n) { host_give(p, n * sizeof(T)); }
    template <class U> void construct(U *p) noexcept { ::new ((void *)p) U; }
    template <class U, class... A> void construct(U *p, A &&...a) { ::new ((void *)p) U(std::forward<A>(a)...); }
};
template <class T> bool operator==(const noinit_alloc<T> &, const noinit_alloc<T> &) { return true; }
template <class T> bool operator!=(const noinit_alloc<T> &, const noinit_alloc<T> &) { return false; }
template <class T> using hvec = std::vector<T, noinit_alloc<T>>;
// The CSR transposes of a gate layer, as counted while the layer is loaded:
// records by u (phase 1) and by v (phase 2), the densest gather tile's
// records, and the gates phase 2 adds into entry 0 (types 5, 12, 6, 10), in
// gate order.
struct wiring_plan {
    int64_t e1 = 0, e2 = 0, m1 = 0, m2 = 0;
    std::vector<int32_t> z2;
};
struct host_circuit {
    int depth = 0;                         // total_depth (layer 0 = inputs)
    std::vector<int> bl;
    std::vector<hvec<uint8_t>> ty;
    std::vector<hvec<int32_t>> u, v;
    std::vector<wiring_plan> W;            // gate layers 1 .. depth-1
};

int bitlen(int64_t x) { int c = 0; while (x) { ++c; x >>= 1; } return c; }
int ilog2(int64_t x) { int l = 0; while (((int64_t)1 << l) < x) ++l; return l; }

// Loads a circuit as zk_verifier::read_circuit builds it (padding of the
// input layer, the relay layer above it, bit lengths), checks every gate,
// and counts each gate layer's wiring records for the CSR plan, in one pass
// over the gates.  The pass is split into chunks of 2^16 gates run on the
// host pool; the first error in file order (layer, then gate) is reported,
// a layer's relay check only when its other checks pass, as before.
int build_circuit(const vg_circuit_desc *d, host_circuit &C)
{
    int D = d->d;
    if (D < 2) FAILV(-1, "circuit needs an input layer and at least one gate layer");
    C.depth = D + 1;
    C.bl.assign(D + 1, 0);
    C.ty.assign(D + 1, {});
    C.u.assign(D + 1, {});
    C.v.assign(D + 1, {});
    C.W.assign(D + 1, {});
    std::vector<int64_t> n(D + 1), n_pad(D + 1), goff(D + 1);
    int64_t off = 0;
    for (int i = 1; i <= D; ++i) {
        n[i] = d->n[i - 1];
        int pad_req = D > 3 ? 17 : 15;
        n_pad[i] = (i == 1 && n[i] < (1LL << pad_req)) ? (1LL << pad_req) : n[i];
        if (n[i] < 1 || (n[i] != 1 && (n[i] & (n[i] - 1)))) FAILV(-2, "layer %d: gate count %lld is not a power of two", i, (long long)n[i]);
        if (i == 1 && n[i] == 1) FAILV(-2, "single-input circuits are not supported");
        if (n_pad[i] > (1LL << 30)) FAILV(-2, "layer %d too wide", i);
        goff[i] = off;
        off += n[i];
        C.bl[i] = n_pad[i] == 1 ? 1 : bitlen(n_pad[i] - 1);
    }
    C.bl[0] = C.bl[1];
    for (int i = 1; i <= D; ++i) {
        int64_t alloc = n[i] == 1 ? 2 : n_pad[i];
        C.ty[i].resize(alloc); C.u[i].resize(alloc); C.v[i].resize(alloc);
        if (i == 1) { C.ty[0].resize(alloc); C.u[0].resize(alloc); C.v[0].resize(alloc); }
    }
    // work items: (layer, chunk) over each layer's loaded range (the input
    // layer's padding included)
    constexpr int64_t CH = (int64_t)1 << 16;
    struct item {
        int layer;
        int64_t lo, hi;
        int code = 0;
        int64_t bad = -1;             // gate of the first failed check
        std::string err;
        int64_t relay_bad = -1;
        int64_t e1 = 0, e2 = 0;
        std::vector<int64_t> t1, t2;  // records per gather tile
        std::vector<int32_t> z2;
    };
    std::vector<item> items;
    for (int i = 1; i <= D; ++i) {
        int64_t len = i == 1 ? n_pad[1] : n[i];
        for (int64_t lo = 0; lo < len; lo += CH) {
            item it;
            it.layer = i; it.lo = lo; it.hi = std::min(len, lo + CH);
            items.push_back(std::move(it));
        }
    }
    auto chunk = [&](int64_t k) {
        item &it = items[k];
        const int i = it.layer;
        const int tsh = ilog2(launch::gf_tile(C.bl[i - 1]));
        const int64_t ntile = std::max<int64_t>(1, ((int64_t)1 << C.bl[i - 1]) >> tsh);
        it.t1.assign(ntile, 0); it.t2.assign(ntile, 0);
        int64_t *t1 = it.t1.data(), *t2 = it.t2.data();
        int64_t e1 = 0, e2 = 0;
        // the CSR records of one stored gate (zk_prover.cpp:375-592, 789-983)
        auto count = [&](int64_t g, int t, int64_t uu, int64_t vv) {
            int64_t lo = 0, hi = 0;
            switch (t) {
            case 0: case 1: case 4: case 6: case 7: case 8: case 9: case 10: case 13: lo = uu; hi = uu + 1; break;
            case 5: lo = uu; hi = vv; break;
            case 12: lo = uu; hi = vv + 1; break;
            default: break;
            }
            if (hi == lo + 1) t1[lo >> tsh]++;
            else
                for (int64_t x = lo; x < hi;) {
                    int64_t q = x >> tsh, end = std::min((q + 1) << tsh, hi);
                    t1[q] += end - x;
                    x = end;
                }
            e1 += hi > lo ? hi - lo : 0;
            if (t == 0 || t == 1 || t == 7 || t == 8 || t == 9 || t == 13) { t2[vv >> tsh]++; ++e2; }
            if (t == 5 || t == 12 || t == 6 || t == 10) it.z2.push_back((int32_t)g);
        };
        char b[256];
        auto fail = [&](int64_t j) { it.code = -3; it.bad = j; it.err = b; };
        const int prev_bl = i == 1 ? 0 : C.bl[i - 1];   // the input layer's gates are checked against 2^0
        const int64_t lim = (int64_t)1 << prev_bl;
        const int64_t own = (int64_t)1 << C.bl[i];
        const int32_t *dty = d->ty + goff[i];
        const int64_t *du = d->u + goff[i], *dv = d->v + goff[i];
        uint8_t *ty = C.ty[i].data();
        int32_t *u = C.u[i].data(), *v = C.v[i].data();
        const int64_t end = std::min(it.hi, n[i]);
        for (int64_t j = it.lo; j < end; ++j) {
            int t = dty[j];
            int64_t uu = du[j], vv = dv[j];
            if (t != 3) {
                if (t == 5) {
                    if (!(uu >= 0 && uu < lim) || !(vv > uu && vv <= lim)) {
                        snprintf(b, sizeof b, "layer %d gate %lld: sum range out of bounds", i, (long long)j);
                        return fail(j);
                    }
                } else if (!(uu >= 0 && uu < lim) || !(vv >= 0 && vv < lim)) {
                    snprintf(b, sizeof b, "layer %d gate %lld: input index out of bounds", i, (long long)j);
                    return fail(j);
                }
            }
            if (t == 6 || t == 10) vv = 0;
            if (t == 13 && uu != vv) { snprintf(b, sizeof b, "bit-test gate needs u == v"); return fail(j); }
            if (t == 12 && vv - uu + 1 > 60) {                                   // zk_prover.cpp:129
                snprintf(b, sizeof b, "power-sum gate spans more than 60 entries");
                return fail(j);
            }
            if (t < 0 || t > 13 || t == 11) { snprintf(b, sizeof b, "unknown gate type %d", t); return fail(j); }
            if (i != 1) {
                ty[j] = (uint8_t)t; u[j] = (int32_t)uu; v[j] = (int32_t)vv;
                if (t == 4 && uu >= own && it.relay_bad < 0) it.relay_bad = j;
                count(j, t, uu, vv);
            } else {
                if (t != 2 && t != 3) { snprintf(b, sizeof b, "input layer gates must be type 2 or 3"); return fail(j); }
                ty[j] = 4; u[j] = (int32_t)j; v[j] = 0;
                C.ty[0][j] = (uint8_t)t; C.u[0][j] = (int32_t)uu; C.v[0][j] = (int32_t)vv;   // (int) truncation, zk_prover.cpp:60
                count(j, 4, j, 0);
            }
        }
        if (i == 1)                                      // padding: zero inputs relayed
            for (int64_t g = std::max(it.lo, n[1]); g < it.hi; ++g) {
                ty[g] = 4; u[g] = (int32_t)g; v[g] = 0;
                C.ty[0][g] = 3; C.u[0][g] = 0; C.v[0][g] = 0;
                count(g, 4, g, 0);
            }
        if (n[i] == 1 && it.lo == 0) { ty[1] = 2; u[1] = 0; v[1] = 0; }   // single-gate layer: a dummy second gate
        it.e1 = e1; it.e2 = e2;
    };
    parallel_for((int64_t)items.size(), chunk);
    // errors: the first failing layer in file order; in it, the first failed
    // check by gate, else its first relay gate reading beyond its own layer
    size_t k = 0;
    for (int i = 1; i <= D; ++i) {
        size_t k0 = k;
        int64_t relay = -1;
        for (; k < items.size() && items[k].layer == i; ++k) {
            if (items[k].code) FAILV(items[k].code, "%s", items[k].err.c_str());
            if (relay < 0) relay = items[k].relay_bad;
        }
        if (relay >= 0)
            FAILV(-3, "layer %d: relay gate %zu reads u=%d beyond its own layer size", i, (size_t)relay, C.u[i][relay]);
        // the layer's wiring counts
        wiring_plan &W = C.W[i];
        std::vector<int64_t> t1, t2;
        for (size_t q = k0; q < k; ++q) {
            item &it = items[q];
            if (t1.empty()) { t1.swap(it.t1); t2.swap(it.t2); }
            else
                for (size_t x = 0; x < t1.size(); ++x) { t1[x] += it.t1[x]; t2[x] += it.t2[x]; }
            W.e1 += it.e1;
            W.e2 += it.e2;
            W.z2.insert(W.z2.end(), it.z2.begin(), it.z2.end());
        }
        for (size_t x = 0; x < t1.size(); ++x) { W.m1 = std::max(W.m1, t1[x]); W.m2 = std::max(W.m2, t2[x]); }
    }
    if (C.bl[0] < 6 + 1) FAILV(-2, "input layer too small for the commitment");
    return 0;
}

// ---------------------------------------------------------------------------
// challenge / mask schedule (draw order of zk_verifier::verify,
// zk_prover::generate_maskR / generate_maskpoly_pre_rho, commit_phase and
// verify_poly_commitment)
// ---------------------------------------------------------------------------
struct layer_sched {
    int bl = 0;
    fe rho, maskR[6], maskr[2], r_c, random_combine, linear_coeff, linear_combine, alpha, beta;
    std::vector<fe> maskpoly, r_u, r_v;
};
struct schedule {
    std::vector<fe> r0, r1, fri_r;
    std::vector<layer_sched> L;
    int64_t pow[33];
};

void build_schedule(const host_circuit &C, schedule &S)
{
    glibc_rand g(1);
    int bl_out = C.bl[C.depth - 1];
    S.r0.resize(bl_out); S.r1.resize(bl_out);
    for (auto &x : S.r0) x = g.field();
    for (auto &x : S.r1) x = g.field();
    S.L.assign(C.depth, layer_sched());
    for (int i = C.depth - 1; i >= 1; --i) {
        layer_sched &l = S.L[i];
        int bl = l.bl = C.bl[i - 1];
        l.rho = g.field();
        if (i > 1) for (auto &x : l.maskR) x = g.field();
        l.maskr[0] = g.field(); l.maskr[1] = g.field();
        l.maskpoly.resize((2 * bl + 1) * 2 + 1 + 6);
        for (auto &x : l.maskpoly) x = g.field();
        l.r_u.resize(bl); l.r_v.resize(bl);
        for (auto &x : l.r_u) x = g.field();
        for (auto &x : l.r_v) x = g.field();
        l.r_c = g.field();
        l.random_combine = g.field();
        l.linear_coeff = g.field();
        l.linear_combine = g.field();
        l.alpha = g.field();
        l.beta = g.field();
    }
    int n = C.bl[0];
    int cw = 1 << (n - 1);
    S.fri_r.clear();
    while (cw > 32) { S.fri_r.push_back(g.field()); cw /= 2; }
    for (int rep = 0; rep < 33; ++rep) {
        int64_t p;
        do { p = (int64_t)(g.next() % (uint32_t)(1u << (n - 1))); } while (p < (1LL << (n - 6)) || p % 2 == 1);
        S.pow[rep] = p;
    }
}

struct qpoly { fe a, b, c; };
fe q_eval(const qpoly &p, const fe &x) { return add(mul(add(mul(p.a, x), p.b), x), p.c); }
fe dbl_n(fe x, int n) { for (int i = 0; i < n; ++i) x = add(x, x); return x; }

// Host part of every prover message of one layer.  Device round sums are
// added to these (field addition, so the split is exact).
struct layer_host {
    int bl = 0;
    fe maskpoly_sumc;
    std::vector<fe> p1;           // 3 per round; the last round has 6 (a..f before reordering)
    std::vector<fe> p2;
    fe lb1[4], lb2[4];            // sumRc.a, .b, .c, Z for the last-bit terms
    fe vu_const, vv_const;
    fe fin[3];                    // final round a, b (without general value), c
    fe query, rg1, rg2, Iuv, preZu, preZv;
};

// The scalar state of zk_prover (zk_prover.h:283-348) that does not touch
// the bookkeeping tables.
struct mask_machine {
    fe alpha, beta, rho, maskpoly_sumc, maskpoly_sumr;
    std::vector<fe> maskpoly;
    fe maskR[6], maskR_sumcu, maskR_sumcv, preZu, preZv, Zu, Zv, preu1, prev1, Iuv, prepreu1, preprev1;
    qpoly Rg1, Rg2, sumRc;
    fe maskr[2];
    int lu = 0, lv = 0;
    fe inv2;
    std::vector<fe> pri;
    mask_machine()
    {
        alpha = beta = rho = maskpoly_sumc = maskpoly_sumr = fzero();
        for (auto &x : maskR) x = fzero();
        maskR_sumcu = maskR_sumcv = preZu = preZv = Zu = Zv = preu1 = prev1 = Iuv = prepreu1 = preprev1 = fzero();
        Rg1 = Rg2 = sumRc = {fzero(), fzero(), fzero()};
        maskr[0] = maskr[1] = fzero();
        inv2 = finv(mk(2, 0));
    }
    void gen_maskR(const layer_sched &l, int layer_id)       // zk_prover.cpp:192-255
    {
        for (auto &x : maskR) pri.push_back(x);
        prepreu1 = preu1; preprev1 = prev1;
        fe *m = maskR;
        Rg1.a = m[4]; Rg1.b = add(m[3], mul(m[5], preu1));
        Rg1.c = add(add(m[0], mul(m[1], preu1)), mul(mul(m[2], preu1), preu1));
        Rg2.a = m[4]; Rg2.b = add(m[3], mul(m[5], prev1));
        Rg2.c = add(add(m[0], mul(m[1], prev1)), mul(mul(m[2], prev1), prev1));
        fe sRu = add(add(add(Rg1.a, Rg1.b), Rg1.c), Rg1.c);
        fe sRv = add(add(add(Rg2.a, Rg2.b), Rg2.c), Rg2.c);
        maskR_sumcu = mul(mul(alpha, Zu), sRu);
        maskR_sumcv = mul(mul(beta, Zv), sRv);
        preZu = Zu; preZv = Zv;
        Zu = fone(); Zv = fone(); Iuv = fone();
        if (layer_id > 1) {
            for (int i = 0; i < 6; ++i) m[i] = l.maskR[i];
            sumRc.a = add(m[2], m[2]);
            sumRc.b = add(add(m[1], m[1]), m[5]);
            sumRc.c = add(add(add(m[0], m[0]), m[3]), m[4]);
        }
        maskr[0] = l.maskr[0]; maskr[1] = l.maskr[1];
        if (layer_id == 1) {
            m[0] = maskr[0]; m[1] = maskr[1];
            m[2] = m[3] = m[4] = m[5] = fzero();
            sumRc.a = fzero(); sumRc.b = m[1]; sumRc.c = m[0];
        }
    }
    void maskpolys(const layer_sched &l)                      // zk_prover.cpp:162-188
    {
        int length = 2 * l.bl + 1, degree = 2;
        maskpoly = l.maskpoly;
        for (auto &x : maskpoly) pri.push_back(x);
        rho = l.rho;
        for (auto &x : maskpoly) x = mul(x, rho);
        fe s = maskpoly[length * degree];
        for (int i = length * degree; i >= 0; --i) s = add(s, maskpoly[i]);
        for (int i = 1; i <= 6; ++i) s = add(s, maskpoly[length * degree + i]);
        for (int i = 1; i < length; ++i) s = add(s, s);
        maskpoly_sumc = s;
        maskpoly_sumr = fzero();
    }
    // phase-1 round (zk_prover.cpp:596-787), mask part only
    void p1_round(fe r, int cb, bool last, std::vector<fe> &out, fe *lb)
    {
        fe one = fone();
        Iuv = mul(Iuv, sub(one, r));
        if (cb > 0) {
            maskR_sumcu = mul(maskR_sumcu, sub(one, r));
            maskR_sumcv = mul(maskR_sumcv, sub(one, r));
            Zu = mul(mul(Zu, sub(one, r)), r);
        }
        fe db = neg(add(maskR_sumcu, maskR_sumcv)), dc = add(maskR_sumcu, maskR_sumcv);
        int k = lu + lv - cb, base = (lu + lv + 1) * 2;
        fe t1 = dbl_n(maskpoly[cb << 1], k), t2 = dbl_n(maskpoly[(cb << 1) + 1], k);
        fe t4 = fzero(), t5 = fzero(), t6 = fzero();
        if (last) {
            lb[0] = sumRc.a; lb[1] = sumRc.b; lb[2] = sumRc.c; lb[3] = Zu;
            t4 = dbl_n(maskpoly[base + 1], k); t5 = dbl_n(maskpoly[base + 2], k); t6 = dbl_n(maskpoly[base + 3], k);
            maskpoly_sumc = mul(sub(sub(sub(sub(sub(maskpoly_sumc, t1), t2), t4), t5), t6), inv2);
        } else {
            maskpoly_sumc = mul(sub(sub(maskpoly_sumc, t1), t2), inv2);
        }
        fe t3 = fzero();
        if (cb > 0) {
            maskpoly_sumr = add(add(maskpoly_sumr, mul(mul(maskpoly[(cb << 1) - 2], r), r)), mul(maskpoly[(cb << 1) - 1], r));
            t3 = dbl_n(maskpoly_sumr, k);
        }
        out.push_back(t1);
        out.push_back(add(db, t2));
        out.push_back(add(add(dc, maskpoly_sumc), t3));
        if (last) { out.push_back(t4); out.push_back(t5); out.push_back(t6); }
    }
    fe p2_init(fe r)                                          // zk_prover.cpp:789-983, mask part
    {
        fe one = fone();
        preu1 = r;
        maskR_sumcu = mul(maskR_sumcu, sub(one, r));
        maskR_sumcv = mul(maskR_sumcv, sub(one, r));
        Iuv = mul(Iuv, sub(one, r));
        Zu = mul(mul(Zu, sub(one, r)), r);
        fe vconst = mul(Zu, q_eval(sumRc, r));
        int base = (lu + lv + 1) * 2;
        fe r2 = mul(r, r), r3 = mul(r2, r), r4 = mul(r3, r), r5 = mul(r4, r);
        maskpoly_sumr = add(add(maskpoly_sumr, mul(maskpoly[lu * 2 - 2], r2)), mul(maskpoly[lu * 2 - 1], r));
        maskpoly_sumr = add(maskpoly_sumr, mul(maskpoly[base + 1], r5));
        maskpoly_sumr = add(maskpoly_sumr, mul(maskpoly[base + 2], r4));
        maskpoly_sumr = add(maskpoly_sumr, mul(maskpoly[base + 3], r3));
        return vconst;
    }
    // phase-2 round (zk_prover.cpp:988-1171), mask part only
    void p2_round(fe r, int cb, bool last, std::vector<fe> &out, fe *lb)
    {
        fe one = fone();
        if (cb > 0 || last) Iuv = mul(Iuv, sub(one, r));
        if (cb > 0) {
            maskR_sumcu = mul(maskR_sumcu, sub(one, r));
            maskR_sumcv = mul(maskR_sumcv, sub(one, r));
            Zv = mul(mul(Zv, sub(one, r)), r);
        }
        fe db = neg(add(maskR_sumcu, maskR_sumcv)), dc = add(maskR_sumcu, maskR_sumcv);
        int cur = cb + lu, k = lu + lv - cur, base = (lu + lv + 1) * 2;
        fe t1 = dbl_n(maskpoly[cur << 1], k), t2 = dbl_n(maskpoly[(cur << 1) + 1], k);
        fe t4 = fzero(), t5 = fzero(), t6 = fzero();
        if (last) {
            lb[0] = sumRc.a; lb[1] = sumRc.b; lb[2] = sumRc.c; lb[3] = Zv;
            t4 = dbl_n(maskpoly[base + 4], k); t5 = dbl_n(maskpoly[base + 5], k); t6 = dbl_n(maskpoly[base + 6], k);
            maskpoly_sumc = mul(sub(sub(sub(sub(sub(maskpoly_sumc, t1), t2), t4), t5), t6), inv2);
        } else {
            maskpoly_sumc = mul(sub(sub(maskpoly_sumc, t1), t2), inv2);
        }
        maskpoly_sumr = add(add(maskpoly_sumr, mul(mul(maskpoly[(cur << 1) - 2], r), r)), mul(maskpoly[(cur << 1) - 1], r));
        fe t3 = dbl_n(maskpoly_sumr, k);
        out.push_back(t1);
        out.push_back(add(db, t2));
        out.push_back(add(add(dc, maskpoly_sumc), t3));
        if (last) { out.push_back(t4); out.push_back(t5); out.push_back(t6); }
    }
    fe finalize(fe r)                                          // zk_prover.cpp:1206-1215
    {
        fe one = fone();
        prev1 = r;
        Iuv = mul(Iuv, sub(one, r));
        Zv = mul(mul(Zv, sub(one, r)), r);
        return mul(Zv, q_eval(sumRc, r));
    }
    void finalround(fe r, int current, fe *out)                // zk_prover.cpp:1173-1203
    {
        fe zu = mul(Iuv, preZu), zv = mul(Iuv, preZv);
        fe a = add(mul(mul(zu, Rg1.a), alpha), mul(mul(zv, Rg2.a), beta));
        fe b = add(mul(mul(zu, Rg1.b), alpha), mul(mul(zv, Rg2.b), beta));
        fe c = add(mul(mul(zu, Rg1.c), alpha), mul(mul(zv, Rg2.c), beta));
        int k = lu + lv - current, base = (lu + lv + 1) * 2;
        fe t1 = dbl_n(maskpoly[current << 1], k), t2 = dbl_n(maskpoly[(current << 1) + 1], k);
        maskpoly_sumc = mul(sub(sub(maskpoly_sumc, t1), t2), inv2);
        fe r2 = mul(r, r), r3 = mul(r2, r), r4 = mul(r3, r), r5 = mul(r4, r);
        maskpoly_sumr = add(add(maskpoly_sumr, mul(maskpoly[(current << 1) - 2], r2)), mul(maskpoly[(current << 1) - 1], r));
        maskpoly_sumr = add(maskpoly_sumr, mul(maskpoly[base + 4], r5));
        maskpoly_sumr = add(maskpoly_sumr, mul(maskpoly[base + 5], r4));
        maskpoly_sumr = add(maskpoly_sumr, mul(maskpoly[base + 6], r3));
        fe t3 = dbl_n(maskpoly_sumr, k);
        out[0] = add(a, t1);
        out[1] = add(b, t2);
        out[2] = add(add(c, maskpoly_sumc), t3);
    }
    fe query(const fe *u, const fe *v, fe rc) const            // zk_prover.cpp:284-306
    {
        fe res = fzero();
        int base = 2 * (lu + lv + 1);
        const fe *m = maskpoly.data();
        for (int i = 0; i < lu; ++i) {
            res = add(add(res, mul(mul(m[2 * i], u[i]), u[i])), mul(m[2 * i + 1], u[i]));
            if (i == lu - 1) {
                fe x = u[i], x2 = mul(x, x), x3 = mul(x2, x), x4 = mul(x3, x), x5 = mul(x4, x);
                res = add(add(add(res, mul(m[base + 1], x5)), mul(m[base + 2], x4)), mul(m[base + 3], x3));
            }
        }
        for (int i = 0; i < lv; ++i) {
            res = add(add(res, mul(mul(m[2 * (i + lu)], v[i]), v[i])), mul(m[2 * (i + lu) + 1], v[i]));
            if (i == lv - 1) {
                fe x = v[i], x2 = mul(x, x), x3 = mul(x2, x), x4 = mul(x3, x), x5 = mul(x4, x);
                res = add(add(add(res, mul(m[base + 4], x5)), mul(m[base + 5], x4)), mul(m[base + 6], x3));
            }
        }
        res = add(add(res, mul(mul(m[2 * (lu + lv)], rc), rc)), mul(m[2 * (lu + lv) + 1], rc));
        return add(res, m[2 * (lu + lv + 1)]);
    }
};

// verifier's public mask coefficients (zk_verifier.cpp:983-1025)
void push_pub_masks(std::vector<fe> &pub, const layer_sched &l, int i, fe preu1, fe prev1)
{
    int ll = l.bl;
    fe one = fone();
    fe rc = i == 1 ? fzero() : l.r_c;
    fe rcb = l.random_combine, lc = l.linear_coeff, lcm = l.linear_combine, rho = l.rho;
    std::vector<fe> rv(ll);
    for (int k = 0; k < ll; ++k) rv[k] = i == 1 ? fzero() : l.r_v[k];
    pub.push_back(add(mul(one, rcb), mul(mul(lc, one), rcb)));
    pub.push_back(add(mul(preu1, rcb), mul(mul(lc, prev1), rcb)));
    pub.push_back(add(mul(mul(preu1, preu1), rcb), mul(mul(mul(lc, prev1), prev1), rcb)));
    pub.push_back(add(mul(rc, rcb), mul(mul(lc, rc), rcb)));
    pub.push_back(add(mul(mul(rc, rc), rcb), mul(mul(mul(lc, rc), rc), rcb)));
    pub.push_back(add(mul(mul(preu1, rc), rcb), mul(mul(mul(lc, prev1), rc), rcb)));
    for (int k = 0; k < ll; ++k) {
        pub.push_back(mul(mul(mul(lcm, l.r_u[k]), l.r_u[k]), rho));
        pub.push_back(mul(mul(lcm, l.r_u[k]), rho));
    }
    for (int k = 0; k < ll; ++k) {
        pub.push_back(mul(mul(mul(lcm, rv[k]), rv[k]), rho));
        pub.push_back(mul(mul(lcm, rv[k]), rho));
    }
    pub.push_back(mul(mul(mul(lcm, rc), rc), rho));
    pub.push_back(mul(mul(lcm, rc), rho));
    pub.push_back(mul(mul(lcm, one), rho));
    for (int w = 0; w < 2; ++w) {
        fe x = w == 0 ? l.r_u[ll - 1] : rv[ll - 1];
        fe x2 = mul(x, x), x3 = mul(x2, x), x4 = mul(x3, x), x5 = mul(x4, x);
        pub.push_back(mul(mul(lcm, x5), rho));
        pub.push_back(mul(mul(lcm, x4), rho));
        pub.push_back(mul(mul(lcm, x3), rho));
    }
}

// last-bit terms of sumcheck_phase{1,2}_updatelastbit (zk_prover.cpp:736-747, 1122-1133)
void lastbit_terms(const fe *lb, fe d, fe e, fe &A, fe &B, fe &Dd, fe &E, fe &Fv)
{
    fe a = lb[0], b = lb[1], c = lb[2], Z = lb[3];
    Dd = mul(mul(neg(a), d), Z);
    E = mul(sub(mul(a, sub(d, e)), mul(b, d)), Z);
    Fv = mul(sub(add(mul(a, e), mul(b, sub(d, e))), mul(c, d)), Z);
    A = add(A, mul(add(mul(c, sub(d, e)), mul(b, e)), Z));
    B = add(B, mul(mul(c, e), Z));
}


}  // namespace

// ---------------------------------------------------------------------------
// proof context: device arena + per-proof host state
// ---------------------------------------------------------------------------
struct vg_ctx {
    host_circuit C;
    hipStream_t st = nullptr;   // evaluation, sumcheck, queries
    hipStream_t st2 = nullptr;  // commitment (runs beside the sumcheck)
    bool own_stream = false;
    uint8_t *arena = nullptr;
    int64_t arena_bytes = 0;
    int64_t pin_bytes = 0, gout_len = 0;     // pinned staging: chal, gsrc, gout, msg
    int64_t up_bytes = 0;                    // page-locked upload staging when the source stages (upload_bytes)
    uint8_t *h_up = nullptr;
    int64_t fg_bytes = 0, fg_pin = 0;        // the fft_gkr context's
    mem_source *mem = nullptr;               // where arena and staging come from
    // device views
    std::vector<uint8_t *> d_ty;
    std::vector<int32_t *> d_u, d_v;
    std::vector<fe *> cv;
    fe *chal = nullptr;         // challenges
    static constexpr int NGS = 4;             // sumcheck streams (layers run concurrently)
    // per layer i: flat tables, fold chunk results (6), block partials
    std::vector<fe *> amb_l, aVb_l, part_l;
    std::vector<fe *> ws_l;
    // layers grouped by table size; one launch per step covers a group
    struct batch {
        int bl = 0;
        std::vector<int> layers;
        int nz_max = 0, ng_max = 0;
        int d0 = 0;                       // first entry in d_layers / d_inst
        bool fused[2] = {true, true};     // phase tables built by gather_fold
        int kl[2] = {0, 0};               // log2 lanes per entry of a flat table build
    };
    std::vector<batch> batches;
    launch::layer_dev *d_layers = nullptr;    // in batch order
    launch::fold_inst *d_inst = nullptr;      // batch order: phase 1 of its layers, then phase 2
    std::vector<launch::layer_dev> h_layers;  // host copies (merged launches of many proofs)
    std::vector<launch::fold_inst> h_inst;
    std::vector<int> slot0;                   // first slot round of (layer entry, phase)
    hipStream_t gst[NGS] = {};
    hipEvent_t gev[NGS] = {};
    // CSR transposes of every gate layer (circuit structure only, built once)
    // (built on the device at creation; the host plans sizes and the z2 lists)
    std::vector<int64_t> n_ent1, n_ent2;      // records by u / by v per layer
    std::vector<char> crowd1, crowd2;         // a gather_fold tile holds more than GF_MAX_FANIN records per entry
    std::vector<int> kl1, kl2;                // log2 lanes per entry for a flat build (densest tile's records per entry)
    std::vector<std::vector<int32_t>> h_z2;
    launch::csr_layer *d_csr = nullptr;                 // descriptors of the device CSR build, one per gate layer
    std::vector<uint32_t *> tops1, tops2;               // its scan scratch per layer
    std::vector<uint32_t *> off1, off2;
    std::vector<launch::gate_rec *> ent1, ent2;
    std::vector<int32_t *> z2;
    // eq half tables of every layer: 8 per layer (g0f g0s g1f g1s uf us vf vs) + 2 for q
    fe *ht = nullptr;
    int64_t ht_stride = 0;
    launch::half_desc *hdesc = nullptr;
    int n_hdesc = 0;
    int64_t ht_total = 0;
    // round-sum slots of the block folds (see launch::sumcheck_multi)
    fe *slots = nullptr;
    int n_slots = 0;
    int *slot_nb = nullptr;
    int64_t *slot_dst = nullptr;
    int64_t *slot_rof = nullptr;               // offset of r[j-1] in chal; -1 for round 0
    std::vector<int> h_slot_nb;
    std::vector<int64_t> h_slot_dst, h_slot_rof;
    fe *bt[8] = {};             // beta half tables: g0f g0s g1f g1s uf us vf vs
    fe *msg = nullptr;          // device results of the GKR part
    fe *mle_t[2] = {};
    // commitment
    int n = 0, log_SS = 0, log_C = 0, log_R = 0;
    int64_t SS = 0, real = 0, m = 0, C_ = 0, LQ = 0;
    fe *ntW = nullptr, *ntWi = nullptr, *ntTA = nullptr, *ntTB = nullptr;   // NTT stage and coset tables
    launch::ntt_tabs NT = {};
    fe *tw = nullptr, *T12 = nullptr, *coefP = nullptr, *coefQ = nullptr, *l = nullptr, *q = nullptr, *h = nullptr,
       *qv = nullptr, *cq = nullptr, *eqs = nullptr, *lqc = nullptr, *c0 = nullptr, *all_sum = nullptr, *pri = nullptr,
       *pubm = nullptr, *cw = nullptr, *gout = nullptr;
    int *lens = nullptr;
    uint8_t *treeL = nullptr, *treeH = nullptr, *ctree = nullptr;
    uint64_t *gsrc = nullptr;
    int64_t n_gather = 0, gather_cap = 0;
    std::vector<int64_t> cw_off, ctree_off;
    // host staging (pinned)
    fe *h_chal = nullptr;
    int64_t n_chal = 0;
    uint64_t *h_gsrc = nullptr;
    fe *h_gout = nullptr;
    fe *h_msg = nullptr;
    int64_t msg_len = 0;
    // per-proof host state
    schedule S;
    std::vector<layer_host> LH;
    std::vector<fe> pub_mask, pri_mask;       // all_pub_mask / all_pri_mask of the last proof (unpadded)
    bool proved = false;                      // a finished proof's values are on the host
    std::vector<int64_t> chal_ru, chal_rv, chal_ab;   // offsets into chal per layer
    int64_t chal_r0 = 0, chal_r1 = 0, chal_fri = 0, chal_one = 0, chal_vc = 0;
    std::vector<int64_t> msg_layer;                   // offset of each layer's block in msg
    int64_t msg_vres = 0, msg_allsum = 0;
    fe post[5];
    hipEvent_t ev[10] = {};     // stage boundaries; ev[9]: inputs of the side streams ready
    hipEvent_t ev_leaf[2] = {}; // around the leaf-chain kernel
    std::chrono::steady_clock::time_point t_start, t_enq;
    double host_ms[3] = {};     // prepare (challenges, masks, gather table), enqueue, assemble
    bool pending = false;
    bool serial = false;        // the whole proof in order on st (caller's stream, large input or VIRGO_SERIAL=1)
    bool own_aux = false;       // st2 created by the context
    // a proof in order on its own streams (large input) still runs the
    // extension of its input layer (commit_private_array) on st2 beside the
    // evaluation, whose gathers leave the VALU idle: ev_in (input layer
    // evaluated, on st), ev_chal (challenges uploaded, on st), ev_l0 (the
    // extension's start, on st2); ev[3] marks its end
    bool early_l = false;
    hipEvent_t ev_in = nullptr, ev_chal = nullptr, ev_l0 = nullptr;
    int own_gst = 0;            // gst[0..own_gst) created by the context (else gst[] = st)
    fg_ctx *fg = nullptr;       // fft_gkr of the commitment (lg = n - 6), on its own stream
};

namespace {

template <class T>
T *carve(uint8_t *base, int64_t &off, int64_t count)
{
    off = (off + 255) & ~(int64_t)255;
    T *p = base ? (T *)(base + off) : nullptr;
    off += (int64_t)sizeof(T) * count;
    return p;
}

// lays out (base == nullptr: measures) every device buffer of the context
int64_t layout(vg_ctx *c, uint8_t *base)
{
    int64_t off = 0;
    const host_circuit &C = c->C;
    int D = C.depth;
    c->d_ty.assign(D, nullptr); c->d_u.assign(D, nullptr); c->d_v.assign(D, nullptr); c->cv.assign(D, nullptr);
    int maxbl = 0;
    // the gate arrays back to back in the order ctx_alloc uploads them, so
    // that a staged upload copies them all at once
    for (int i = 0; i < D; ++i) {
        int64_t n = (int64_t)1 << C.bl[i];
        if (C.bl[i] > maxbl) maxbl = C.bl[i];
        c->d_u[i] = carve<int32_t>(base, off, n);
        if (i > 0) {
            c->d_ty[i] = carve<uint8_t>(base, off, n);
            c->d_v[i] = carve<int32_t>(base, off, n);
        }
    }
    for (int i = 0; i < D; ++i) c->cv[i] = carve<fe>(base, off, (int64_t)1 << C.bl[i]);
    c->chal = carve<fe>(base, off, c->n_chal);
    c->amb_l.assign(D, nullptr); c->aVb_l.assign(D, nullptr); c->part_l.assign(D, nullptr);
    c->ws_l.assign(6 * D, nullptr);
    for (int i = 1; i < D; ++i) {
        // flat tables only where gather_fold does not build them on the fly
        bool flat = launch::gf_rounds(C.bl[i - 1]) == 0;
        for (auto &b : c->batches)
            if (b.bl == C.bl[i - 1] && (!b.fused[0] || !b.fused[1])) flat = true;
        if (flat) {
            int64_t T = (int64_t)1 << C.bl[i - 1];
            c->amb_l[i] = carve<fe>(base, off, T);
            c->aVb_l[i] = carve<fe>(base, off, T);
        }
        c->part_l[i] = carve<fe>(base, off, launch::MAX_PARTS + 1);
        for (int k = 0; k < 6; ++k) c->ws_l[6 * i + k] = carve<fe>(base, off, launch::fold_ws_len(C.bl[i - 1], launch::GF_B));
    }
    c->d_layers = carve<launch::layer_dev>(base, off, D);
    c->d_inst = carve<launch::fold_inst>(base, off, 2 * D);
    c->off1.assign(D, nullptr); c->off2.assign(D, nullptr);
    c->ent1.assign(D, nullptr); c->ent2.assign(D, nullptr); c->z2.assign(D, nullptr);
    c->tops1.assign(D, nullptr); c->tops2.assign(D, nullptr);
    for (int i = 1; i < D; ++i) {
        int64_t nt = (int64_t)1 << C.bl[i - 1];
        c->off1[i] = carve<uint32_t>(base, off, nt + 1);
        c->off2[i] = carve<uint32_t>(base, off, nt + 1);
        c->ent1[i] = carve<launch::gate_rec>(base, off, c->n_ent1[i] + 1);
        c->ent2[i] = carve<launch::gate_rec>(base, off, c->n_ent2[i] + 1);
        c->z2[i] = carve<int32_t>(base, off, (int64_t)c->h_z2[i].size() + 1);
        c->tops1[i] = carve<uint32_t>(base, off, launch::csr_tops_len(nt));
        c->tops2[i] = carve<uint32_t>(base, off, launch::csr_tops_len(nt));
    }
    c->d_csr = carve<launch::csr_layer>(base, off, D);
    c->ht_stride = (int64_t)1 << ((maxbl >> 1) + 1);
    c->ht = carve<fe>(base, off, (int64_t)D * 8 * c->ht_stride + 2 * ((int64_t)1 << (C.bl[0] / 2 + 1)));
    c->hdesc = carve<launch::half_desc>(base, off, (int64_t)D * 8 + 2);
    c->slots = carve<fe>(base, off, (int64_t)(c->n_slots > 0 ? c->n_slots : 1) * 3 * launch::FW_BLOCKS);
    c->slot_nb = carve<int>(base, off, c->n_slots + 1);
    c->slot_dst = carve<int64_t>(base, off, c->n_slots + 1);
    c->slot_rof = carve<int64_t>(base, off, c->n_slots + 1);
    int half = (maxbl >> 1) + 1;
    for (int k = 0; k < 8; ++k) c->bt[k] = carve<fe>(base, off, (int64_t)1 << half);
    c->msg = carve<fe>(base, off, c->msg_len);
    int64_t no = (int64_t)1 << C.bl[D - 1];
    c->mle_t[0] = carve<fe>(base, off, no);
    c->mle_t[1] = carve<fe>(base, off, no);
    // commitment
    int64_t SS = c->SS;
    c->T12 = carve<fe>(base, off, 1024 + (SS >> 10) + 1);
    c->tw = carve<fe>(base, off, SS);
    {
        int maxlog = c->log_C + 1, th = (c->log_C + 1) / 2, tb = c->log_C - th;
        int64_t R = (int64_t)1 << c->log_R;
        c->ntW = carve<fe>(base, off, (int64_t)1 << maxlog);
        c->ntWi = carve<fe>(base, off, (int64_t)1 << maxlog);
        c->ntTA = carve<fe>(base, off, R << th);
        c->ntTB = carve<fe>(base, off, R << tb);
        c->NT = {c->ntW, c->ntWi, c->ntTA, c->ntTB, th, tb};
    }
    c->coefP = carve<fe>(base, off, 65 * c->C_);
    // the public array's slices 0..63 are cq[j] Q (eq factors over the top 6
    // and the low n-6 input bits), so q holds two extensions: Q's and the
    // public mask's
    c->coefQ = carve<fe>(base, off, 2 * c->C_);
    c->l = carve<fe>(base, off, 65 * SS);
    c->q = carve<fe>(base, off, 2 * SS);
    c->h = carve<fe>(base, off, 65 * SS);
    c->qv = carve<fe>(base, off, c->real);
    c->cq = carve<fe>(base, off, 64);
    c->eqs = carve<fe>(base, off, 2 * ((int64_t)1 << (c->n / 2 + 1)));
    c->lqc = carve<fe>(base, off, 65 * c->LQ);
    c->c0 = carve<fe>(base, off, 65);
    c->all_sum = carve<fe>(base, off, 65);
    c->pri = carve<fe>(base, off, c->m);
    c->pubm = carve<fe>(base, off, c->m);
    c->lens = carve<int>(base, off, 65);
    c->treeL = carve<uint8_t>(base, off, SS * 32);
    c->treeH = carve<uint8_t>(base, off, SS * 32);
    int nf = c->n - 6;
    c->cw_off.assign(nf, 0); c->ctree_off.assign(nf, 0);
    int64_t cwt = 0, ctt = 0;
    for (int k = 0, S = (int)SS; k < nf; ++k, S /= 2) {
        c->cw_off[k] = cwt; cwt += 65LL * (S / 2);
        c->ctree_off[k] = ctt; ctt += (int64_t)(S / 2) * 32;
    }
    c->cw = carve<fe>(base, off, cwt);
    c->ctree = carve<uint8_t>(base, off, ctt);
    c->gsrc = carve<uint64_t>(base, off, c->gather_cap);
    c->gout = carve<fe>(base, off, c->gather_cap);
    return off + 256;
}

// number of gathered 16-byte items of the commitment (roots, final codeword, queries)
int64_t gather_count(int n)
{
    int64_t k = 2 * 2 + 2 * (n - 6) + 2048 + 32;
    int64_t per = 2 * (130 + 2 * (n - 1));
    for (int lvl = 0; lvl < n - 6; ++lvl) per += 130 + 2 * (n - 2 - lvl);
    return k + 33 * per;
}

// host work of a proof that depends only on the shape: schedule, masks,
// challenge buffer, gather addresses
int prepare(vg_ctx *c)
{
    host_circuit &C = c->C;
    build_schedule(C, c->S);
    schedule &S = c->S;
    int D = C.depth;
    // challenge buffer (offsets fixed by chal_layout)
    fe *ch = c->h_chal;
    ch[c->chal_one] = fone(); ch[c->chal_one + 1] = fzero();
    for (size_t k = 0; k < S.r0.size(); ++k) { ch[c->chal_r0 + k] = S.r0[k]; ch[c->chal_r1 + k] = S.r1[k]; }
    for (int i = D - 1; i >= 1; --i) {
        layer_sched &l = S.L[i];
        for (int k = 0; k < l.bl; ++k) {
            ch[c->chal_ru[i] + k] = l.r_u[k];
            ch[c->chal_rv[i] + k] = i == 1 ? fzero() : l.r_v[k];
        }
        ch[c->chal_ab[i]] = l.alpha; ch[c->chal_ab[i] + 1] = l.beta;
    }
    for (size_t k = 0; k < S.fri_r.size(); ++k) ch[c->chal_fri + k] = S.fri_r[k];
    // host parts of the prover messages (the mask bookkeeping)
    mask_machine M;
    std::vector<fe> pub;
    c->LH.assign(D, layer_host());
    fe alpha = fone(), beta = fzero();
    for (int i = D - 1; i >= 1; --i) {
        layer_sched &l = S.L[i];
        layer_host &H = c->LH[i];
        int bl = H.bl = l.bl;
        M.alpha = alpha; M.beta = beta; M.lu = bl; M.lv = bl;
        M.gen_maskR(l, i);
        M.maskpolys(l);
        H.maskpoly_sumc = M.maskpoly_sumc;
        push_pub_masks(pub, l, i, M.prepreu1, M.preprev1);
        std::vector<fe> rv(bl);
        for (int k = 0; k < bl; ++k) rv[k] = i == 1 ? fzero() : l.r_v[k];
        for (int j = 0; j < bl; ++j) M.p1_round(j ? l.r_u[j - 1] : fzero(), j, j == bl - 1, H.p1, H.lb1);
        H.vu_const = M.p2_init(l.r_u[bl - 1]);
        for (int j = 0; j < bl; ++j) M.p2_round(j ? rv[j - 1] : fzero(), j, j == bl - 1, H.p2, H.lb2);
        H.vv_const = M.finalize(rv[bl - 1]);
        ch[c->chal_vc + 2 * i] = H.vu_const;
        ch[c->chal_vc + 2 * i + 1] = H.vv_const;
        M.finalround(rv[bl - 1], bl << 1, H.fin);
        fe rc = i == 1 ? fzero() : l.r_c;
        H.query = M.query(l.r_u.data(), rv.data(), rc);
        H.rg1 = q_eval(M.Rg1, rc);
        H.rg2 = q_eval(M.Rg2, rc);
        H.Iuv = M.Iuv; H.preZu = M.preZu; H.preZv = M.preZv;
        alpha = l.alpha; beta = l.beta;
    }
    pub.push_back(M.Zu);
    pub.push_back(mul(M.Zu, M.preu1));
    M.pri.push_back(M.maskr[0]);
    M.pri.push_back(M.maskr[1]);
    c->pub_mask = pub;
    c->pri_mask = M.pri;
    c->post[0] = M.Zu; c->post[1] = M.preu1; c->post[2] = M.sumRc.a; c->post[3] = M.sumRc.b; c->post[4] = M.sumRc.c;
    // masks padded to m (poly_commit.h:56-69, 136-139)
    std::vector<fe> prim(c->m, fzero()), pubm(c->m, fzero());
    for (size_t k = 0; k < M.pri.size() && (int64_t)k < c->m; ++k) prim[k] = M.pri[k];
    for (size_t k = 0; k < pub.size() && (int64_t)k < c->m; ++k) pubm[k] = pub[k];
    // pri and pub masks ride along with the challenges
    memcpy(c->h_chal + c->n_chal - 2 * c->m, prim.data(), sizeof(fe) * c->m);
    memcpy(c->h_chal + c->n_chal - c->m, pubm.data(), sizeof(fe) * c->m);
    return 0;
}

// addresses of every gathered item, in transcript order
int build_gather(vg_ctx *c)
{
    int n = c->n;
    int64_t SS = c->SS, half = SS / 2;
    int log_R = c->log_R;
    uint64_t *g = c->h_gsrc;
    int64_t k = 0;
    auto dig = [&](const uint8_t *tree, int64_t node) {
        g[k++] = (uint64_t)(uintptr_t)(tree + node * 32);
        g[k++] = (uint64_t)(uintptr_t)(tree + node * 32 + 16);
    };
    auto val = [&](const fe *p) { g[k++] = (uint64_t)(uintptr_t)p; };
    dig(c->treeL, 1);
    dig(c->treeH, 1);
    int nf = n - 6;
    for (int lv = 0; lv < nf; ++lv) dig(c->ctree + c->ctree_off[lv], 1);
    // final codeword (cpd.rs_codeword[last]: [(i<<7)|(j<<1)|h] = slice j at i + 16h)
    const fe *fin = c->cw + c->cw_off[nf - 1];
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 64; ++j)
            for (int h = 0; h < 2; ++h) val(fin + j * 32 + cw_idx(i + 16 * h, 32, log_R));
    for (int i = 0; i < 16; ++i)
        for (int h = 0; h < 2; ++h) val(fin + 64 * 32 + cw_idx(i + 16 * h, 32, log_R));
    for (int rep = 0; rep < 33; ++rep) {
        int64_t pow = c->S.pow[rep];
        int64_t s0 = pow / 2;
        for (int o = 0; o < 2; ++o) {
            const fe *src = o == 0 ? c->l : c->h;
            const uint8_t *tree = o == 0 ? c->treeL : c->treeH;
            for (int j = 0; j <= 64; ++j) {
                val(src + j * SS + cw_idx(s0, SS, log_R));
                val(src + j * SS + cw_idx(s0 + half, SS, log_R));
            }
            int64_t pos = s0 + half;
            for (int d = 0; d < n - 2; ++d) { dig(tree, pos ^ 1); pos /= 2; }
            dig(tree, s0 + half);
        }
        for (int lv = 0; lv < nf; ++lv) {
            int64_t S = SS >> (lv + 1);              // codeword size after step lv
            int64_t pw = pow % ((int64_t)1 << (n - 1 - lv));
            int64_t q = (pw / 2) % (S / 2);
            const fe *cwp = c->cw + c->cw_off[lv];
            for (int j = 0; j <= 64; ++j) {
                val(cwp + j * S + cw_idx(q, S, log_R));
                val(cwp + j * S + cw_idx(q + S / 2, S, log_R));
            }
            const uint8_t *tree = c->ctree + c->ctree_off[lv];
            int64_t p0 = q + S / 2;
            int64_t leaf = p0;
            while (p0 != 1) { dig(tree, p0 ^ 1); p0 /= 2; }
            dig(tree, leaf);
        }
    }
    c->n_gather = k;
    if (k > c->gather_cap) FAILV(-30, "gather list overflow (%lld > %lld)", (long long)k, (long long)c->gather_cap);
    return 0;
}

// ---- evaluate (zk_prover.cpp:54-157): needs no challenge, so it is
// enqueued before the host draws them ----
void enqueue_lde_l(vg_ctx *c, hipStream_t st, int part);
int enqueue_eval(vg_ctx *c)
{
    hipStream_t st = c->st;
    host_circuit &C = c->C;
    HIPCHK(hipEventRecord(c->ev[0], st));
    launch::eval_input(c->d_u[0], c->cv[0], 1 << C.bl[0], st);
    if (c->early_l) {
        // the extension of the 64 input slices needs the input layer only:
        // st2 runs it beside the evaluation
        HIPCHK(hipEventRecord(c->ev_in, st));
        HIPCHK(hipStreamWaitEvent(c->st2, c->ev_in, 0));
        HIPCHK(hipEventRecord(c->ev_l0, c->st2));
        enqueue_lde_l(c, c->st2, 1);
    }
    for (int i = 1; i < C.depth; ++i)
        launch::eval_layer({c->d_ty[i], c->d_u[i], c->d_v[i], C.bl[i]}, c->cv[i - 1], c->cv[i], st);
    HIPCHK(hipEventRecord(c->ev[1], st));
    return 0;
}

// challenges, half tables, V_res and the sumcheck layers of a proof (ev[9]:
// what the other streams need); gkr = false leaves the sumcheck layers and
// their slot reduction to a batch's merged launches (batch_gkr) and
// enqueue_post
int enqueue_gkr(vg_ctx *c, bool gkr = true)
{
    hipStream_t st = c->st, st2 = c->st2;
    host_circuit &C = c->C;
    int D = C.depth;
    HIPCHK(hipMemcpyAsync(c->chal, c->h_chal, sizeof(fe) * c->n_chal, hipMemcpyHostToDevice, st));
    if (c->early_l) {
        // the private mask's slice of l, once the challenges are uploaded
        HIPCHK(hipEventRecord(c->ev_chal, st));
        HIPCHK(hipStreamWaitEvent(st2, c->ev_chal, 0));
        enqueue_lde_l(c, st2, 2);
        HIPCHK(hipEventRecord(c->ev[3], st2));
    }
    HIPCHK(hipMemcpyAsync(c->gsrc, c->h_gsrc, sizeof(uint64_t) * c->n_gather, hipMemcpyHostToDevice, st));
    // every eq half table of the proof, from the uploaded challenges
    launch::half_multi(c->hdesc, c->n_hdesc, c->ht_total, st);
    // evaluated circuit + challenges + half tables: what the other streams need
    HIPCHK(hipEventRecord(c->ev[9], st));
    // The commitment of the input layer needs only the input values, the
    // host-known masks and challenges, never a sumcheck result: it runs on the
    // second stream beside the latency-bound sumcheck rounds.
    if (!c->serial) HIPCHK(hipStreamWaitEvent(st2, c->ev[9], 0));
    // ---- V_res (zk_prover.cpp:30-52) ----
    launch::mle_eval(c->cv[D - 1], C.bl[D - 1], c->chal + c->chal_r0, c->mle_t[0], c->mle_t[1], c->msg + c->msg_vres, st);
    // ---- GKR layers ----
    // Every layer's sumcheck depends only on the evaluated circuit and on
    // challenges known in advance, so layer batches may run on several
    // streams (gst[], the main stream unless VIRGO_GKR_STREAMS asks for more).
    if (!c->serial && gkr)
        for (int w = 0; w < vg_ctx::NGS; ++w) HIPCHK(hipStreamWaitEvent(c->gst[w], c->ev[9], 0));
    int slot = 0;
    for (size_t bi = 0; gkr && bi < c->batches.size(); ++bi) {
        const vg_ctx::batch &b = c->batches[bi];
        hipStream_t gs = c->gst[bi % vg_ctx::NGS];
        int n = (int)b.layers.size();
        const launch::layer_dev *Dl = c->d_layers + b.d0;
        const launch::fold_inst *I = c->d_inst + 2 * b.d0;
        for (int ph = 0; ph < 2; ++ph) {
            const launch::fold_inst *Ip = I + ph * n;
            int nz = ph ? b.nz_max : 0;
            int f = b.fused[ph] ? launch::gather_fold(ph, Dl, Ip, n, b.bl, nz, gs) : 0;
            if (!f) launch::layer_tables(ph, Dl, n, 1 << b.bl, nz, gs, b.kl[ph]);
            slot += n * launch::sumcheck_multi(Ip, n, b.bl, gs, launch::GF_B, f > 0);
        }
        launch::general_multi(Dl, n, b.ng_max, gs);
    }
    if (gkr) {
        if (!c->serial)
            for (int w = 0; w < vg_ctx::NGS; ++w) {
                HIPCHK(hipEventRecord(c->gev[w], c->gst[w]));
                HIPCHK(hipStreamWaitEvent(st, c->gev[w], 0));
            }
        if (slot != c->n_slots) FAILV(-30, "fold slot plan mismatch");
        launch::reduce_slots(c->slots, c->n_slots, c->slot_nb, c->slot_dst, c->slot_rof, c->chal, c->msg, st);
        HIPCHK(hipEventRecord(c->ev[2], st));
    }
    return 0;
}

// the commitment up to its leaf chains: LDEs, public array, FRI folds (ends
// on ev[5]); on the second stream when the context overlaps its stages
// ---- commit_private_array (poly_commit.h:42-124): the 64 input slices and
// the private mask interpolated and extended into l (slice-major, SS values
// per slice; every row is independent, so the slices may be extended in
// parts): part 0 all 65 slices, part 1 the 64 input slices (they need the
// input layer only), part 2 the mask slice (it needs the challenge upload)
void enqueue_lde_l(vg_ctx *c, hipStream_t st, int part = 0)
{
    int64_t real = c->real, m = c->m, Cc = c->C_;
    int lr = ilog2(real), lm = ilog2(m);
    fe *pri_d = c->chal + (c->n_chal - 2 * m);
    fe inv_real = finv(mk((uint64_t)real, 0)), inv_m = finv(mk((uint64_t)m, 0));
    if (part != 2) launch::intt({0, c->cv[0], real, nullptr, nullptr, 0, 0, 0, 0}, c->coefP, Cc, 64, lr, inv_real, c->NT, st);
    if (part == 1) launch::lde(c->coefP, Cc, c->lens, 64, c->l, c->log_C, c->log_R, c->NT, st);
    if (part != 1) launch::intt({0, pri_d, m, nullptr, nullptr, 0, 0, 0, 0}, c->coefP + 64 * Cc, Cc, 1, lm, inv_m, c->NT, st);
    if (part == 0) launch::lde(c->coefP, Cc, c->lens, 65, c->l, c->log_C, c->log_R, c->NT, st);
    if (part == 2)
        launch::lde(c->coefP + 64 * Cc, Cc, c->lens + 64, 1, c->l + 64 * c->SS, c->log_C, c->log_R, c->NT, st);
}
int enqueue_commit(vg_ctx *c)
{
    hipStream_t st = c->serial ? c->st : c->st2;
    host_circuit &C = c->C;
    int D = C.depth;
    int n = c->n;
    int64_t SS = c->SS, real = c->real, m = c->m, Cc = c->C_;
    int lr = ilog2(real), lm = ilog2(m);
    fe *pub_d = c->chal + (c->n_chal - m);
    fe inv_real = finv(mk((uint64_t)real, 0)), inv_m = finv(mk((uint64_t)m, 0));
    fe inv_2real = finv(mk((uint64_t)(2 * real), 0)), inv_2m = finv(mk((uint64_t)(2 * m), 0));
    if (c->early_l) {
        HIPCHK(hipStreamWaitEvent(st, c->ev[3], 0));       // l, extended on st2 beside the evaluation
    } else {
        enqueue_lde_l(c, st, 0);
        HIPCHK(hipEventRecord(c->ev[3], st));
    }
    // ---- commit_public_array (poly_commit.h:125-351) ----
    // Slice j < 64 of the public array is cq[j] Q exactly (a product of the
    // same eq factors), so its interpolation and extension are cq[j] times
    // Q's: one extension of Q and one of the public mask replace the 65 of
    // poly_commit.h:136-170; every consumer scales Q's values by cq[j].
    {
        fe *q = c->ht + (int64_t)D * 8 * c->ht_stride;
        int fq = (n - 6) >> 1;
        launch::eq_expand(c->qv, q, q + ((int64_t)1 << fq), fq, real, st);
    }
    launch::intt({0, c->qv, real, nullptr, nullptr, 0, 0, 0, 0}, c->coefQ, Cc, 1, lr, inv_real, c->NT, st);
    launch::intt({0, pub_d, m, nullptr, nullptr, 0, 0, 0, 0}, c->coefQ + Cc, Cc, 1, lm, inv_m, c->NT, st);
    launch::lde(c->coefQ, Cc, c->lens + 63, 2, c->q, c->log_C, c->log_R, c->NT, st);     // lens[63] = real, lens[64] = m
    {
        launch::intt_src lq = {1, nullptr, 0, c->l, c->q, SS, SS, SS / (2 * real), c->log_R};
        lq.qscale = c->cq;
        launch::intt(lq, c->lqc, c->LQ, 64, lr + 1, inv_2real, c->NT, st);
    }
    launch::intt({1, nullptr, 0, c->l + 64 * SS, c->q + SS, SS, SS, SS / (2 * m), c->log_R}, c->lqc + 64 * c->LQ, c->LQ, 1,
                 lm + 1, inv_2m, c->NT, st);
    launch::slice_sums(c->lqc, c->LQ, 65, real, m, c->c0, c->msg + c->msg_allsum, st);
    launch::lde(c->lqc, c->LQ, c->lens, 65, c->h, c->log_C, c->log_R, c->NT, st, 1);
    HIPCHK(hipEventRecord(c->ev[4], st));
    // ---- commit_phase folds (vpd_verifier.cpp:46-78, fri.cpp:321-390).  The
    // challenges are known in advance, so every fold runs before any hashing.
    {
        fe *next[24];
        for (int k = 0; k < n - 6; ++k) next[k] = c->cw + c->cw_off[k];
        launch::fri_folds_vo(c->l, c->q, c->cq, c->h, c->c0, real, m, SS, n - 6, c->log_R, c->chal + c->chal_fri, c->tw, next,
                             st);
    }
    HIPCHK(hipEventRecord(c->ev[5], st));
    return 0;
}

// The sumcheck layers of many proofs in merged launches on stream gs: the
// layer batches of equal table width and table-build path of every context
// form one instance list (each instance keeps its own tables, slots and
// message block), so a launch covers them all.  wait: gs waits for every
// context's challenges (ev[9]) on the device (false: the caller has waited
// on the host); records gkr_done.  *buf: device descriptor arrays, *host_buf:
// their pinned copy for the asynchronous upload; the caller frees both once
// gs has drained.
int batch_gkr(vg_ctx *const *ctxs, int n, hipStream_t gs, hipEvent_t gkr_done, void **buf, void **host_buf,
              bool wait = true)
{
    struct group {
        int bl;
        bool f0, f1;
        int nz_max = 0, ng_max = 0, kl[2] = {0, 0};
        std::vector<launch::layer_dev> L;
        std::vector<launch::fold_inst> I[2];
    };
    std::vector<group> G;
    for (int i = 0; i < n; ++i) {
        vg_ctx *c = ctxs[i];
        if (wait) HIPCHK(hipStreamWaitEvent(gs, c->ev[9], 0));
        for (auto &b : c->batches) {
            group *g = nullptr;
            for (auto &x : G)
                if (x.bl == b.bl && x.f0 == b.fused[0] && x.f1 == b.fused[1]) g = &x;
            if (!g) {
                G.push_back(group());
                g = &G.back();
                g->bl = b.bl; g->f0 = b.fused[0]; g->f1 = b.fused[1];
            }
            int m = (int)b.layers.size();
            g->nz_max = b.nz_max > g->nz_max ? b.nz_max : g->nz_max;
            g->ng_max = b.ng_max > g->ng_max ? b.ng_max : g->ng_max;
            g->kl[0] = std::max(g->kl[0], b.kl[0]);
            g->kl[1] = std::max(g->kl[1], b.kl[1]);
            for (int q = 0; q < m; ++q) {
                g->L.push_back(c->h_layers[b.d0 + q]);
                g->I[0].push_back(c->h_inst[2 * b.d0 + q]);
                g->I[1].push_back(c->h_inst[2 * b.d0 + m + q]);
            }
        }
    }
    size_t bytes = 0;
    for (auto &g : G) bytes += sizeof(launch::layer_dev) * g.L.size() + 2 * sizeof(launch::fold_inst) * g.L.size();
    HIPCHK(hipHostMalloc(host_buf, bytes ? bytes : 1, hipHostMallocDefault));
    uint8_t *hk = (uint8_t *)*host_buf;
    std::vector<size_t> offL(G.size()), offI(G.size());
    size_t o = 0;
    for (size_t k = 0; k < G.size(); ++k) {
        offL[k] = o;
        memcpy(hk + o, G[k].L.data(), sizeof(launch::layer_dev) * G[k].L.size());
        o += sizeof(launch::layer_dev) * G[k].L.size();
        offI[k] = o;
        for (int ph = 0; ph < 2; ++ph) {
            memcpy(hk + o, G[k].I[ph].data(), sizeof(launch::fold_inst) * G[k].I[ph].size());
            o += sizeof(launch::fold_inst) * G[k].I[ph].size();
        }
    }
    HIPCHK(hipMalloc(buf, bytes ? bytes : 1));
    uint8_t *d = (uint8_t *)*buf;
    HIPCHK(hipMemcpyAsync(d, hk, bytes, hipMemcpyHostToDevice, gs));
    for (size_t k = 0; k < G.size(); ++k) {
        const group &g = G[k];
        int m = (int)g.L.size();
        const launch::layer_dev *Dl = (const launch::layer_dev *)(d + offL[k]);
        for (int ph = 0; ph < 2; ++ph) {
            const launch::fold_inst *Ip = (const launch::fold_inst *)(d + offI[k]) + ph * m;
            int nz = ph ? g.nz_max : 0;
            int f = (ph ? g.f1 : g.f0) ? launch::gather_fold(ph, Dl, Ip, m, g.bl, nz, gs) : 0;
            if (!f) launch::layer_tables(ph, Dl, m, 1 << g.bl, nz, gs, g.kl[ph]);
            launch::sumcheck_multi(Ip, m, g.bl, gs, launch::GF_B, f > 0);
        }
        launch::general_multi(Dl, m, g.ng_max, gs);
    }
    HIPCHK(hipEventRecord(gkr_done, gs));
    HIPCHK(hipGetLastError());
    return 0;
}

// the committed codewords, in leaf-launch order: L, H, then every FRI level
int commit_segments(const vg_ctx *c, launch::leaf_seg *g)
{
    int ns = 0;
    g[ns++] = {c->l, c->SS, c->treeL, c->log_R};
    g[ns++] = {c->h, c->SS, c->treeH, c->log_R};
    int64_t Sk = c->SS / 2;
    for (int k = 0; k < c->n - 6; ++k, Sk /= 2) g[ns++] = {c->cw + c->cw_off[k], Sk, c->ctree + c->ctree_off[k], c->log_R};
    return ns;
}

// ---- every SHA3 leaf chain (fri.cpp:97-126, 397-420) in one launch, then
// every Merkle tree (merkle_tree.cpp:46-56) level by level, and the query
// answers (a batch does all of this in merged launches instead).
int enqueue_post(vg_ctx *c)
{
    hipStream_t st = c->serial ? c->st : c->st2;
    launch::leaf_seg g[24];
    int ns = commit_segments(c, g);
    const fe *cws[24];
    int64_t Ss[24], nl[24];
    uint8_t *trees[24];
    for (int k = 0; k < ns; ++k) { cws[k] = g[k].cw; Ss[k] = g[k].S; trees[k] = g[k].tree; nl[k] = g[k].S / 2; }
    // the leaf chains keep every SIMD busy on their own: beside them the
    // latency-bound sumcheck and fft_gkr launches only slow both down, so
    // they start once those are done (the LDEs and folds before them do
    // overlap the sumcheck)
    if (!c->serial)
        for (int w = 0; w < vg_ctx::NGS; ++w) HIPCHK(hipStreamWaitEvent(st, c->gev[w], 0));
    if (c->fg && c->own_stream) HIPCHK(hipStreamWaitEvent(st, fg_done_event(c->fg), 0));   // fft_gkr on its own stream
    HIPCHK(hipEventRecord(c->ev_leaf[0], st));
    launch::leaf_hash_multi(ns, cws, Ss, trees, c->log_R, st);
    HIPCHK(hipEventRecord(c->ev_leaf[1], st));
    launch::merkle_multi(ns, trees, nl, st);
    HIPCHK(hipEventRecord(c->ev[6], st));
    st = c->st;
    if (!c->serial) HIPCHK(hipStreamWaitEvent(st, c->ev[6], 0));
    HIPCHK(hipEventRecord(c->ev[8], st));
    // ---- query answers (fri.cpp:148-289) ----
    launch::gather16(c->gsrc, c->n_gather, c->gout, st);
    HIPCHK(hipMemcpyAsync(c->h_gout, c->gout, sizeof(fe) * c->n_gather, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(c->h_msg, c->msg, sizeof(fe) * c->msg_len, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev[7], st));
    return 0;
}

void assemble(vg_ctx *c, tbuf &T)
{
    host_circuit &C = c->C;
    int D = C.depth;
    T.put("VGT1", 4);
    T.u32((uint32_t)D);
    T.u32((uint32_t)C.bl[0]);
    for (int i = 0; i < D; ++i) T.u32((uint32_t)C.bl[i]);
    T.f(c->h_msg[c->msg_vres]);
    for (int i = D - 1; i >= 1; --i) {
        layer_host &H = c->LH[i];
        int bl = H.bl;
        const fe *M = c->h_msg + c->msg_layer[i];
        const fe *sums1 = M, *end1 = M + 3 * bl, *sums2 = M + 3 * bl + 5, *end2 = M + 6 * bl + 5, *gv = M + 6 * bl + 10;
        T.f(H.maskpoly_sumc);
        for (int ph = 0; ph < 2; ++ph) {
            const fe *sums = ph == 0 ? sums1 : sums2;
            const fe *end = ph == 0 ? end1 : end2;
            const std::vector<fe> &hp = ph == 0 ? H.p1 : H.p2;
            const fe *lb = ph == 0 ? H.lb1 : H.lb2;
            for (int j = 0; j < bl - 1; ++j) {
                T.f(add(sums[3 * j], hp[3 * j]));
                T.f(add(sums[3 * j + 1], hp[3 * j + 1]));
                T.f(add(sums[3 * j + 2], hp[3 * j + 2]));
            }
            int j = bl - 1;
            fe A = sums[3 * j], B = sums[3 * j + 1], Cc = sums[3 * j + 2], Dd, E, Fv;
            lastbit_terms(lb, end[2], end[3], A, B, Dd, E, Fv);
            const fe *h6 = &hp[3 * j];
            A = add(A, h6[0]); B = add(B, h6[1]); Cc = add(Cc, h6[2]);
            Dd = add(Dd, h6[3]); E = add(E, h6[4]); Fv = add(Fv, h6[5]);
            T.f(Dd); T.f(E); T.f(Fv); T.f(A); T.f(B); T.f(Cc);      // quintuple (d,e,f,a,b,c), zk_prover.cpp:781
            T.f(end[4]);                                            // v_u, then v_v
        }
        fe g = i > 1 ? *gv : fzero();
        T.f(H.fin[0]); T.f(add(H.fin[1], g)); T.f(H.fin[2]);
        T.f(H.query); T.f(H.rg1); T.f(H.rg2);
        T.f(H.Iuv); T.f(H.preZu); T.f(H.preZv);
    }
    for (int k = 0; k < 5; ++k) T.f(c->post[k]);
    const fe *go = c->h_gout;
    int64_t k = 0;
    T.put(go + k, 64); k += 4;                                     // roots L, H
    for (int s = 0; s < 65; ++s) T.f(c->h_msg[c->msg_allsum + s]);
    T.put(go + k, (size_t)16 * (c->n_gather - k));
    c->proved = true;
}

// Plan of the CSR transposes of one gate layer (the structure of
// zk_prover.cpp:420-587 and :845-972; the transposes themselves are built on
// the device, launch::csr_build_multi): phase 1 lists, for every entry u of the
// previous layer, the gates that add into it through u (a sum gate into
// every u of its range); phase 2 lists by v the gates that add into entry v,
// and separately the gates that always add into entry 0 (z2, kept here).
// Takes the layer's record counts (counted while loading, build_circuit)
// and decides whether any gather_fold tile of the layer's table holds more
// than GF_MAX_FANIN records per entry, and the lanes per entry of a flat
// build.
void csr_plan(vg_ctx *c, int i)
{
    wiring_plan &W = c->C.W[i];
    int64_t nt = (int64_t)1 << c->C.bl[i - 1];
    int64_t tile = launch::gf_tile(c->C.bl[i - 1]);
    c->n_ent1[i] = W.e1;
    c->n_ent2[i] = W.e2;
    c->h_z2[i].swap(W.z2);
    c->crowd1[i] = W.m1 > launch::GF_MAX_FANIN * tile;
    c->crowd2[i] = W.m2 > launch::GF_MAX_FANIN * tile;
    // lanes per entry of a flat build: the densest tile's records per entry,
    // rounded up to a power of two, at most a wave
    auto lanes = [&](int64_t m) {
        int64_t per = (m + std::min(tile, nt) - 1) / std::min(tile, nt);
        int k = 0;
        while (k < 6 && ((int64_t)1 << k) < per) ++k;
        return k;
    };
    c->kl1[i] = lanes(W.m1);
    c->kl2[i] = lanes(W.m2);
}

// Host planning of a proof: sizes, challenge and message layout, layer
// batches, CSR transposes and the arena layout (nothing allocated).
// staging bytes of ctx_alloc's uploads, put by put in the same order
int64_t upload_bytes(const vg_ctx *c)
{
    const host_circuit &C = c->C;
    auto sb = uploader::stage_bytes;
    int64_t t = 0;
    for (int i = 0; i < C.depth; ++i) {
        int64_t nn = (int64_t)1 << C.bl[i];
        t += sb(4 * nn);
        if (i > 0) t += sb(nn) + sb(4 * nn);
    }
    if (c->n_slots) t += sb((int64_t)sizeof(int) * c->n_slots) + 2 * sb((int64_t)sizeof(int64_t) * c->n_slots);
    for (int i = 1; i < C.depth; ++i) t += sb(4 * (int64_t)c->h_z2[i].size());
    t += sb((int64_t)sizeof(launch::csr_layer) * (C.depth - 1));
    t += sb((int64_t)sizeof(launch::half_desc) * (8 * (C.depth - 1) + 3));
    int nd = C.depth - 1;
    t += sb((int64_t)sizeof(launch::layer_dev) * (nd > 0 ? nd : 1));
    t += sb((int64_t)sizeof(launch::fold_inst) * (nd > 0 ? 2 * nd : 1));
    t += sb((int64_t)sizeof(int) * 65);
    t += sb((int64_t)sizeof(fe) * (1024 + (c->SS >> 10) + 1));
    return t;
}
int ctx_plan(vg_ctx *c)
{
    host_circuit &C = c->C;
    int n = c->n = C.bl[0];
    if (n > 26) FAILV(-2, "input layer of 2^%d exceeds this build's commitment limit", n);
    c->SS = (int64_t)1 << (n - 1);
    c->log_SS = n - 1;
    c->real = c->SS >> 5;
    // mask count (4 bl + 15 per layer, + 2) and the padded mask length m (poly_commit.h:56-66)
    int64_t nm = 2;
    for (int i = C.depth - 1; i >= 1; --i) nm += 4 * C.bl[i - 1] + 15;
    int64_t gap = c->SS / nm;
    for (int j = 0; j < 64; ++j)
        if ((1LL << j) <= gap && (1LL << (j + 1)) > gap) { gap = 1LL << j; break; }
    if (gap < 2) FAILV(-2, "too many mask coefficients for the input size");
    c->m = c->SS / gap;
    c->C_ = c->real > c->m ? c->real : c->m;
    c->log_C = ilog2(c->C_);
    c->log_R = c->log_SS - c->log_C;
    c->LQ = 2 * c->C_;
    // challenge buffer layout: 1, 0 | r_0 | r_1 | per layer r_u, r_v, alpha, beta | fri | pri masks | pub masks
    {
        int64_t k = 0;
        int bl_out = C.bl[C.depth - 1];
        c->chal_one = k; k += 2;
        c->chal_r0 = k; k += bl_out;
        c->chal_r1 = k; k += bl_out;
        c->chal_ru.assign(C.depth, 0); c->chal_rv.assign(C.depth, 0); c->chal_ab.assign(C.depth, 0);
        for (int i = C.depth - 1; i >= 1; --i) {
            c->chal_ru[i] = k; k += C.bl[i - 1];
            c->chal_rv[i] = k; k += C.bl[i - 1];
            c->chal_ab[i] = k; k += 2;
        }
        c->chal_fri = k; k += n - 6;
        c->chal_vc = k; k += 2 * C.depth;      // per layer: the host parts of v_u, v_v
        c->n_chal = k + 2 * c->m;
    }
    // device message block
    c->msg_layer.assign(C.depth, 0);
    int64_t ml = 0;
    c->msg_vres = ml++;
    for (int i = C.depth - 1; i >= 1; --i) { c->msg_layer[i] = ml; ml += 6 * C.bl[i - 1] + 11; }
    c->msg_allsum = ml; ml += 65;
    c->msg_len = ml;
    // layer batches (equal table size 2^bl(i-1)), from the top layer down
    c->batches.clear();
    for (int i = C.depth - 1; i >= 1; --i) {
        int bl = C.bl[i - 1];
        vg_ctx::batch *b = nullptr;
        for (auto &x : c->batches) if (x.bl == bl) b = &x;
        if (!b) { c->batches.push_back(vg_ctx::batch()); b = &c->batches.back(); b->bl = bl; }
        b->layers.push_back(i);
        if ((1 << C.bl[i]) > b->ng_max) b->ng_max = 1 << C.bl[i];
    }
    {
        int d0 = 0;
        for (auto &b : c->batches) { b.d0 = d0; d0 += (int)b.layers.size(); }
    }
    // slot rounds of the block folds: batch by batch, phase 1 of its layers,
    // then phase 2; each instance's rounds contiguous
    c->h_slot_nb.clear(); c->h_slot_dst.clear(); c->h_slot_rof.clear();
    c->slot0.assign(2 * C.depth, 0);
    for (auto &b : c->batches) {
        int js[32], ks[32];
        // the main prover's phases build their tables with gather_fold
        int nl = launch::fold_plan(b.bl, js, ks, launch::GF_B);
        for (int ph = 0; ph < 2; ++ph)
            for (size_t q = 0; q < b.layers.size(); ++q) {
                int i = b.layers[q];
                c->slot0[2 * (b.d0 + q) + ph] = (int)c->h_slot_nb.size();
                int64_t sums = c->msg_layer[i] + (ph == 0 ? 0 : 3 * b.bl + 5);
                int64_t r0 = ph == 0 ? c->chal_ru[i] : c->chal_rv[i];
                for (int k = 0; k < nl; ++k)
                    for (int t = 0; t < ks[k]; ++t) {
                        int j = js[k] + t;
                        c->h_slot_nb.push_back(launch::fold_blocks(b.bl, js[k], launch::GF_B));
                        c->h_slot_dst.push_back(sums + 3 * j);
                        c->h_slot_rof.push_back(j == 0 ? -1 : r0 + j - 1);
                    }
            }
    }
    c->n_slots = (int)c->h_slot_nb.size();
    c->gather_cap = gather_count(n);
    c->n_ent1.assign(C.depth, 0); c->n_ent2.assign(C.depth, 0);
    c->crowd1.assign(C.depth, 0); c->crowd2.assign(C.depth, 0); c->h_z2.assign(C.depth, {});
    c->kl1.assign(C.depth, 0); c->kl2.assign(C.depth, 0);
    for (int i = 1; i < C.depth; ++i) csr_plan(c, i);
    for (int i = 1; i < C.depth; ++i)
        if (c->n_ent1[i] >= (1LL << 31) || c->n_ent2[i] >= (1LL << 31)) FAILV(-2, "layer %d: too many wiring records", i);
    // gather_fold walks a block's records in sequence: phases whose records
    // crowd onto few entries (high fan-out wiring) build flat tables with
    // one thread per entry instead
    for (auto &b : c->batches)
        for (int i : b.layers) {
            if (c->crowd1[i]) b.fused[0] = false;
            if (c->crowd2[i]) b.fused[1] = false;
            b.kl[0] = std::max(b.kl[0], c->kl1[i]);
            b.kl[1] = std::max(b.kl[1], c->kl2[i]);
        }
    c->arena_bytes = layout(c, nullptr);
    c->gout_len = c->gather_cap > (1024 + (c->SS >> 10) + 1) ? c->gather_cap : (1024 + (c->SS >> 10) + 1);
    c->pin_bytes = (int64_t)sizeof(fe) * (c->n_chal + c->gout_len + c->msg_len) + (int64_t)sizeof(uint64_t) * c->gather_cap + 1024;
    c->up_bytes = upload_bytes(c);
    return 0;
}

// Takes the planned arena and pinned staging from c->mem and uploads the
// circuit and the static tables.  Work runs on the context's stream only
// (other proofs may be running on the device).
int ctx_alloc(vg_ctx *c)
{
    host_circuit &C = c->C;
    int n = c->n;
    c->arena = (uint8_t *)c->mem->device(c->arena_bytes);
    if (!c->arena) FAILV(-21, "cannot allocate %lld device bytes", (long long)c->arena_bytes);
    layout(c, c->arena);
    c->h_chal = (fe *)c->mem->pinned(sizeof(fe) * c->n_chal);
    c->h_gsrc = (uint64_t *)c->mem->pinned(sizeof(uint64_t) * c->gather_cap);
    c->h_gout = (fe *)c->mem->pinned(sizeof(fe) * c->gout_len);
    c->h_msg = (fe *)c->mem->pinned(sizeof(fe) * c->msg_len);
    if (!c->h_chal || !c->h_gsrc || !c->h_gout || !c->h_msg) FAILV(-21, "pinned host allocation failed");
    for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
    for (auto &e : c->gev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto &e : c->ev_leaf) HIPCHK(hipEventCreate(&e));
    if (c->early_l) {
        HIPCHK(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ev_chal, hipEventDisableTiming));
        HIPCHK(hipEventCreate(&c->ev_l0));
    }
    // every upload in the order upload_bytes counts them
    uploader U;
    if (!U.init(c->mem, c->up_bytes, c->st)) FAILV(-21, "pinned upload staging failed");
    c->h_up = U.stage;
    // gates and inputs
    for (int i = 0; i < C.depth; ++i) {
        int64_t nn = (int64_t)1 << C.bl[i];
        HIPCHK(U.put(c->d_u[i], C.u[i].data(), 4 * nn));
        if (i > 0) {
            HIPCHK(U.put(c->d_ty[i], C.ty[i].data(), nn));
            HIPCHK(U.put(c->d_v[i], C.v[i].data(), 4 * nn));
        }
    }
    if (c->n_slots) {
        HIPCHK(U.put(c->slot_nb, c->h_slot_nb.data(), (int64_t)sizeof(int) * c->n_slots));
        HIPCHK(U.put(c->slot_dst, c->h_slot_dst.data(), (int64_t)sizeof(int64_t) * c->n_slots));
        HIPCHK(U.put(c->slot_rof, c->h_slot_rof.data(), (int64_t)sizeof(int64_t) * c->n_slots));
    }
    {
        // the CSR transposes of every gate layer, in six launches
        std::vector<launch::csr_layer> CL;
        int64_t max_ng = 0, max_nt = 0;
        for (int i = 1; i < C.depth; ++i) {
            HIPCHK(U.put(c->z2[i], c->h_z2[i].data(), 4 * (int64_t)c->h_z2[i].size()));
            launch::csr_layer d;
            d.ty = c->d_ty[i]; d.u = c->d_u[i]; d.v = c->d_v[i];
            d.ng = (int64_t)C.ty[i].size(); d.nt = (int64_t)1 << C.bl[i - 1];
            d.off1 = c->off1[i]; d.off2 = c->off2[i]; d.ent1 = c->ent1[i]; d.ent2 = c->ent2[i];
            d.tops1 = c->tops1[i]; d.tops2 = c->tops2[i];
            max_ng = std::max(max_ng, d.ng);
            max_nt = std::max(max_nt, d.nt);
            CL.push_back(d);
        }
        HIPCHK(U.put(c->d_csr, CL.data(), (int64_t)sizeof(launch::csr_layer) * (int64_t)CL.size()));
        HIPCHK(U.flush());
        launch::csr_build_multi(c->d_csr, (int)CL.size(), max_ng, max_nt, c->st);
    }
    {
        // eq half tables of every layer, computed in one launch per proof
        // from the uploaded challenges (zk_prover.cpp:391-418, 802-822)
        std::vector<launch::half_desc> hd;
        int64_t tot = 0;
        auto addd = [&](fe *out, const fe *f0, const fe *r, int nbits) {
            launch::half_desc d; d.out = out; d.f0 = f0; d.r = r; d.nbits = nbits; d.start = tot;
            hd.push_back(d); tot += (int64_t)1 << nbits;
        };
        const fe *r0 = c->chal + c->chal_r0, *r1 = c->chal + c->chal_r1;
        const fe *al = c->chal + c->chal_one, *be = al + 1;
        for (int i = C.depth - 1; i >= 1; --i) {
            int bg = C.bl[i], bu = C.bl[i - 1];
            fe *t = c->ht + (int64_t)i * 8 * c->ht_stride;
            const fe *ru = c->chal + c->chal_ru[i], *rv = c->chal + c->chal_rv[i];
            addd(t + 0 * c->ht_stride, al, r0, bg >> 1);
            addd(t + 1 * c->ht_stride, nullptr, r0 + (bg >> 1), bg - (bg >> 1));
            addd(t + 2 * c->ht_stride, be, r1, bg >> 1);
            addd(t + 3 * c->ht_stride, nullptr, r1 + (bg >> 1), bg - (bg >> 1));
            addd(t + 4 * c->ht_stride, nullptr, ru, bu >> 1);
            addd(t + 5 * c->ht_stride, nullptr, ru + (bu >> 1), bu - (bu >> 1));
            addd(t + 6 * c->ht_stride, nullptr, rv, bu >> 1);
            addd(t + 7 * c->ht_stride, nullptr, rv + (bu >> 1), bu - (bu >> 1));
            r0 = ru; r1 = rv;
            al = c->chal + c->chal_ab[i]; be = al + 1;
        }
        // the public array eq(rl, x) over the n input bits (zk_verifier.cpp:838-847)
        // is cq[x >> (n-6)] Q[x mod 2^(n-6)]: half tables of Q over the low
        // n-6 bits, and the 64 slice factors over the top 6
        fe *q = c->ht + (int64_t)C.depth * 8 * c->ht_stride;
        const fe *rl = c->chal + c->chal_ru[1];
        int nq = n - 6;
        addd(q, nullptr, rl, nq >> 1);
        addd(q + ((int64_t)1 << (nq >> 1)), nullptr, rl + (nq >> 1), nq - (nq >> 1));
        addd(c->cq, nullptr, rl + nq, 6);
        c->n_hdesc = (int)hd.size();
        c->ht_total = tot;
        HIPCHK(U.put(c->hdesc, hd.data(), (int64_t)sizeof(launch::half_desc) * (int64_t)hd.size()));
    }
    {
        // static per-layer descriptors of the batched sumcheck launches
        int nd = C.depth - 1;
        std::vector<launch::layer_dev> LD(nd > 0 ? nd : 1);
        std::vector<launch::fold_inst> FI(nd > 0 ? 2 * nd : 1);
        for (auto &b : c->batches)
            for (size_t q = 0; q < b.layers.size(); ++q) {
                int i = b.layers[q], bl = C.bl[i - 1], bg = C.bl[i];
                int e = b.d0 + (int)q;
                fe *M = c->msg + c->msg_layer[i];
                fe *t = c->ht + (int64_t)i * 8 * c->ht_stride;
                int64_t hs = c->ht_stride;
                launch::layer_dev &d = LD[e];
                d.ty = c->d_ty[i]; d.u = c->d_u[i]; d.v = c->d_v[i];
                d.nt = 1 << bl; d.ng = 1 << bg;
                d.off1 = c->off1[i]; d.off2 = c->off2[i]; d.ent1 = c->ent1[i]; d.ent2 = c->ent2[i];
                d.z2 = c->z2[i]; d.nz = (int)c->h_z2[i].size();
                if (d.nz > b.nz_max) b.nz_max = d.nz;
                d.pv = c->cv[i - 1];
                d.bg = {t, t + hs, t + 2 * hs, t + 3 * hs, bg >> 1};
                d.bu = {t + 4 * hs, t + 5 * hs, bl >> 1};
                d.bv = {t + 6 * hs, t + 7 * hs, bl >> 1};
                d.amb = c->amb_l[i]; d.aVb = c->aVb_l[i];
                d.vu = M + 3 * bl + 4; d.vv = M + 6 * bl + 9;
                d.part = c->part_l[i];
                d.gv = i > 1 ? M + 6 * bl + 10 : nullptr;
                for (int ph = 0; ph < 2; ++ph) {
                    launch::fold_inst &x = FI[2 * b.d0 + ph * (int)b.layers.size() + (int)q];
                    x.Vb = c->cv[i - 1]; x.aVb = d.aVb; x.amb = d.amb;
                    for (int k = 0; k < 2; ++k) {
                        x.V[k] = c->ws_l[6 * i + 3 * k]; x.aV[k] = c->ws_l[6 * i + 3 * k + 1]; x.am[k] = c->ws_l[6 * i + 3 * k + 2];
                    }
                    x.r = c->chal + (ph == 0 ? c->chal_ru[i] : c->chal_rv[i]);
                    x.slots = c->slots + (int64_t)c->slot0[2 * e + ph] * 3 * launch::FW_BLOCKS;
                    x.sums = M + (ph == 0 ? 0 : 3 * bl + 5);
                    x.end = M + (ph == 0 ? 3 * bl : 6 * bl + 5);
                    x.vconst = c->chal + c->chal_vc + 2 * i + ph;
                }
            }
        HIPCHK(U.put(c->d_layers, LD.data(), (int64_t)sizeof(launch::layer_dev) * (int64_t)LD.size()));
        HIPCHK(U.put(c->d_inst, FI.data(), (int64_t)sizeof(launch::fold_inst) * (int64_t)FI.size()));
        c->h_layers = LD;
        c->h_inst = FI;
    }
    std::vector<int> lens(65, (int)c->real);
    lens[64] = (int)c->m;
    HIPCHK(U.put(c->lens, lens.data(), (int64_t)sizeof(int) * 65));
    {
        // twiddle table w_SS^e, e < SS (depends only on the input size)
        int64_t SS = c->SS;
        std::vector<fe> T12(1024 + (SS >> 10) + 1);
        fe w = root_of_unity(c->log_SS), x = fone();
        for (int j = 0; j < 1024; ++j) { T12[j] = x; x = mul(x, w); }
        fe w10 = x, y = fone();
        for (int64_t j = 0; j < (SS >> 10) + 1; ++j) { T12[1024 + j] = y; y = mul(y, w10); }
        HIPCHK(U.put(c->T12, T12.data(), (int64_t)sizeof(fe) * (int64_t)T12.size()));
        HIPCHK(U.flush());                       // the last upload
        launch::twiddles(c->tw, SS, c->T12, c->T12 + 1024, c->st);
        launch::ntt_tables(c->ntW, c->ntWi, c->log_C + 1, c->ntTA, c->ntTB, c->log_R, c->NT.th, c->NT.tb, c->tw, SS, c->st);
    }
    // the LDS tile of this proof's transform sizes (the extensions: 65 x 2^R
    // rows of 2^log_C; the l q products: 65 rows of 2^(log_C+1)), tuned once
    // per process on the context's codeword buffers l and h before they hold
    // anything
    if (launch::ntt_tune(c->log_C, (int64_t)65 << c->log_R, c->l, c->h, c->NT, c->st) ||
        launch::ntt_tune(c->log_C + 1, 65, c->l, c->h, c->NT, c->st))
        FAILV(-20, "NTT tile tuning failed");
    // a context of its own reports set-up errors here; a staged one at its
    // proof's end (the host does not wait behind other proofs' work)
    if (!U.stage) HIPCHK(hipStreamSynchronize(c->st));
    return 0;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
// host work and device work of a proof up to its sumcheck layers (gkr: as
// enqueue_gkr)
int prove_start_gkr(vg_ctx *c, bool gkr)
{
    if (!c) FAILV(-1, "null context");
    if (c->pending) FAILV(-1, "a proof is already in flight on this context");
    c->proved = false;
    c->t_start = std::chrono::steady_clock::now();
    // the evaluation runs while the host draws the challenges
    int rc = enqueue_eval(c);
    if (rc) return rc;
    auto tp = std::chrono::steady_clock::now();
    rc = prepare(c);
    if (rc) return rc;
    rc = build_gather(c);
    if (rc) return rc;
    c->host_ms[0] = ms_since(tp);
    c->t_enq = std::chrono::steady_clock::now();
    return enqueue_gkr(c, gkr);
}
// fft_gkr (its own stream unless the proof runs in order: its small
// latency-bound launches overlap the main proof) and the commitment up to
// the leaf chains
int prove_start_commit(vg_ctx *c)
{
    int rc = fg_enqueue(c->fg);
    if (rc) return rc;
    return enqueue_commit(c);
}
// everything of a proof up to its leaf chains
int prove_start(vg_ctx *c, bool gkr = true)
{
    int rc = prove_start_gkr(c, gkr);
    if (rc) return rc;
    return prove_start_commit(c);
}

// ---------------------------------------------------------------------------
// Verifier (zk_verifier::verify, zk_verifier.cpp:910-1253, with
// poly_commit_verifier::verify_poly_commitment, vpd_verifier.cpp:80-335, and
// the fft_gkr program's checks) reading the transcript assemble() writes.
// The round checks are a few field operations each and run on the host; the
// hot loops run on the device: the wiring predicates of every layer
// (zk_verifier.cpp:219-606) in one launch, public_array_prepare (:884-908)
// as an eq table and an iNTT, the first-round evaluations of the public
// polynomial at the query points (vpd_verifier.cpp:215-226) and every query
// path's leaf chain and Merkle climb (vpd_verifier.cpp:8-41).  Slice j of
// the public array is Q[k] q_r[j] (public_array_prepare), so its value at x
// is q_r[j] Q(x): one evaluation per point serves all 64 slices, with the
// exact values of the reference's per-slice loop.
// ---------------------------------------------------------------------------
struct treader {
    const uint8_t *p;
    size_t n, off;
    bool bad;
    void get(void *x, size_t k)
    {
        if (off + k > n) { bad = true; memset(x, 0, k); off = n; return; }
        memcpy(x, p + off, k);
        off += k;
    }
    // a non-canonical element (a component >= p) makes the transcript
    // malformed: the field operations assume canonical inputs, as the
    // reference's in-process messages always are
    fe f()
    {
        fe x;
        get(&x, 16);
        if (x.real >= P || x.img >= P) bad = true;
        return x;
    }
    uint32_t u32() { uint32_t x; get(&x, 4); return x; }
    qpoly q3() { qpoly q; q.a = f(); q.b = f(); q.c = f(); return q; }
};
// quintuple_poly in member order, evaluated as polynomial.cpp:19-22
fe q5_eval(const fe *q, const fe &x)
{
    fe r = q[0];
    for (int k = 1; k < 6; ++k) r = add(mul(r, x), q[k]);
    return r;
}
// relay term of layer 1 (zk_verifier.cpp:1015-1044)
fe direct_relay(const host_circuit &C, int i, const std::vector<fe> &rg, const std::vector<fe> &ru)
{
    if (i != 1) return fzero();
    fe r = fone(), two = from_u64(2);
    for (int k = 0; k < C.bl[1]; ++k) r = mul(r, add(sub(sub(fone(), rg[k]), ru[k]), mul(mul(two, rg[k]), ru[k])));
    return r;
}
// the layer's predicates combined with v_u, v_v (zk_verifier.cpp:1149)
fe general_value(const fe *pv, const fe &vu, const fe &vv)
{
    fe one = fone(), uv = mul(vu, vv);
    fe g = add(mul(pv[0], add(vu, vv)), mul(pv[1], uv));
    g = add(g, mul(pv[6], sub(one, vu)));
    g = add(g, mul(pv[7], sub(vu, vv)));
    g = add(g, mul(pv[8], sub(add(vu, vv), dbl(uv))));
    g = add(g, mul(pv[9], sub(vv, uv)));
    g = add(g, mul(add(add(pv[5], pv[10]), pv[12]), vu));
    g = add(g, mul(pv[13], mul(vu, sub(one, vv))));
    return g;
}
struct vscratch {                 // device buffers of one verification
    std::vector<void *> p;
    ~vscratch() { for (void *x : p) (void)hipFree(x); }
    template <class T>
    T *get(size_t count)
    {
        void *x = nullptr;
        if (hipMalloc(&x, sizeof(T) * (count ? count : 1)) != hipSuccess) return nullptr;
        p.push_back(x);
        return (T *)x;
    }
};
#define VREJ(...) do { char b_[512]; snprintf(b_, sizeof b_, __VA_ARGS__); g_err = b_; *accepted = 0; return 0; } while (0)
#define VNULL(x) do { if (!(x)) FAILV(-21, "verifier: device allocation failed"); } while (0)

int verify(vg_ctx *c, const uint8_t *tr, size_t len, int *accepted)
{
    *accepted = 0;
    const host_circuit &C = c->C;
    int D = C.depth, n = c->n, nf = n - 6;
    hipStream_t st = c->st;
    int rc = prepare(c);                          // the challenge schedule (public randomness)
    if (rc) return rc;
    const schedule &S = c->S;
    vscratch V;
    // ---- wiring predicates of every layer, on the device ----
    int nl = D - 1, ngm = 1;
    for (auto &b : c->batches) ngm = b.ng_max > ngm ? b.ng_max : ngm;
    int pblocks = (ngm + 255) / 256;
    if (pblocks > 256) pblocks = 256;
    fe *ppart = V.get<fe>((size_t)nl * pblocks * 14), *pvd = V.get<fe>((size_t)nl * 14);
    VNULL(ppart); VNULL(pvd);
    HIPCHK(hipMemcpyAsync(c->chal, c->h_chal, sizeof(fe) * c->n_chal, hipMemcpyHostToDevice, st));
    launch::half_multi(c->hdesc, c->n_hdesc, c->ht_total, st);
    launch::predicates(c->d_layers, nl, ngm, ppart, pblocks, pvd, st);
    std::vector<fe> pvh((size_t)nl * 14);
    HIPCHK(hipMemcpyAsync(pvh.data(), pvd, sizeof(fe) * pvh.size(), hipMemcpyDeviceToHost, st));
    std::vector<const fe *> pv_of(D, nullptr);
    for (auto &b : c->batches)
        for (size_t q = 0; q < b.layers.size(); ++q) pv_of[b.layers[q]] = pvh.data() + (size_t)(b.d0 + q) * 14;
    // ---- header ----
    treader R = {tr, len, 0, false};
    char magic[4];
    R.get(magic, 4);
    uint32_t hd = R.u32(), hn = R.u32();
    if (R.bad || memcmp(magic, "VGT1", 4) || hd != (uint32_t)D || hn != (uint32_t)C.bl[0]) VREJ("bad transcript header");
    for (int i = 0; i < D; ++i)
        if (R.u32() != (uint32_t)C.bl[i]) VREJ("bit length mismatch at layer %d", i);
    HIPCHK(hipStreamSynchronize(st));
    // ---- GKR layers ----
    fe one = fone(), alpha = one, beta = fzero();
    std::vector<fe> r0 = S.r0, r1 = S.r1;
    fe claim = R.f();                             // alpha * V_res
    fe mask_sum = fzero(), preu1_prev = fzero(), prev1_prev = fzero();
    std::vector<fe> pub;
    // one sumcheck phase: bl rounds, the last a quintuple; `shift` is added
    // to every round's p(0) + p(1) and to the next claim
    auto phase = [&](const std::vector<fe> &rr, int bl, const fe &shift, int i, int ph) -> bool {
        for (int j = 0; j < bl; ++j) {
            fe p0, p1, pr;
            if (j == bl - 1) {
                fe q[6];
                for (auto &x : q) x = R.f();
                p0 = q5_eval(q, fzero()); p1 = q5_eval(q, one); pr = q5_eval(q, rr[j]);
            } else {
                qpoly q = R.q3();
                p0 = q_eval(q, fzero()); p1 = q_eval(q, one); pr = q_eval(q, rr[j]);
            }
            if (R.bad) { g_err = "transcript truncated or not canonical"; return false; }
            if (!feq(add(add(p0, p1), shift), claim)) {
                char b_[128];
                snprintf(b_, sizeof b_, "phase-%d check failed, layer %d bit %d", ph, i, j);
                g_err = b_;
                return false;
            }
            claim = add(pr, shift);
        }
        return true;
    };
    for (int i = D - 1; i >= 1; --i) {
        const layer_sched &l = S.L[i];
        int bl = l.bl;
        claim = add(claim, R.f());                // maskpoly_sumc
        std::vector<fe> ru = l.r_u, rv(bl, fzero());
        if (i != 1) rv = l.r_v;
        fe drv = add(mul(alpha, direct_relay(C, i, r0, ru)), mul(beta, direct_relay(C, i, r1, ru)));
        fe rcv = i == 1 ? fzero() : l.r_c;
        push_pub_masks(pub, l, i, preu1_prev, prev1_prev);
        if (!phase(ru, bl, fzero(), i, 1)) { *accepted = 0; return 0; }
        fe vu = R.f();
        fe dv = mul(drv, vu);
        if (!phase(rv, bl, dv, i, 2)) { *accepted = 0; return 0; }
        fe vv = R.f();
        fe gv = general_value(pv_of[i], vu, vv);
        qpoly fp = R.q3();
        if (R.bad) VREJ("transcript truncated or not canonical");
        if (!feq(add(add(q_eval(fp, fzero()), q_eval(fp, one)), dv), claim)) VREJ("final round check failed, layer %d", i);
        claim = add(q_eval(fp, rcv), dv);
        fe mpv = R.f(), rg1 = R.f(), rg2 = R.f();
        mask_sum = add(mask_sum, add(add(mul(rg1, l.random_combine), mul(mul(l.linear_coeff, rg2), l.random_combine)),
                                     mul(l.linear_combine, mpv)));
        fe Iuv = R.f(), pZu = R.f(), pZv = R.f();
        fe rhs = add(add(add(add(mul(rcv, gv), mul(mul(mul(alpha, Iuv), pZu), rg1)), mul(mul(mul(beta, Iuv), pZv), rg2)), mpv), dv);
        if (R.bad) VREJ("transcript truncated or not canonical");
        if (!feq(claim, rhs)) VREJ("semi-final check failed, layer %d", i);
        alpha = l.alpha; beta = l.beta;
        claim = i != 1 ? add(mul(alpha, vu), mul(beta, vv)) : vu;
        r0 = ru; r1 = rv;
        preu1_prev = ru[bl - 1];
        prev1_prev = rv[bl - 1];
    }
    fe Zu = R.f(), preu1 = R.f();
    qpoly sumRc = R.q3();
    pub.push_back(Zu);
    pub.push_back(mul(Zu, preu1));
    fe zs = mul(Zu, q_eval(sumRc, preu1));
    fe target = add(sub(claim, zs), add(mask_sum, zs));
    // ---- polynomial commitment (vpd_verifier.cpp:80-335) ----
    int64_t SS = c->SS, cs = c->real, gap = SS / (int64_t)pub.size();
    for (int j = 0; j < 64; ++j)
        if ((1LL << j) <= gap && (1LL << (j + 1)) > gap) { gap = 1LL << j; break; }
    int64_t m = SS / gap;
    if (m != c->m) FAILV(-30, "verifier: public mask length %lld differs from the context's %lld", (long long)m, (long long)c->m);
    pub.resize(m, fzero());
    uint8_t rootL[32], rootH[32];
    R.get(rootL, 32);
    R.get(rootH, 32);
    fe all_sum[65], tot = fzero();
    for (auto &x : all_sum) { x = R.f(); tot = add(tot, x); }
    if (R.bad) VREJ("transcript truncated or not canonical");
    // the prover's own assertion (poly_commit.h:193): the slices add up to the target
    if (!feq(tot, target)) VREJ("sum of all_sum != inner-product target");
    std::vector<uint8_t> froot((size_t)32 * nf);
    R.get(froot.data(), froot.size());
    std::vector<fe> fin(2048 + 32);
    for (auto &x : fin) x = R.f();
    // final codewords constant (vpd_verifier.cpp:315-332)
    for (int i = 0; i < 64; ++i)
        for (int j = 1; j < 16; ++j)
            if (!feq(fin[(j << 7) | (i << 1)], fin[i << 1])) VREJ("final rs code not constant");
    for (int j = 1; j < 32; ++j)
        if (!feq(fin[2048 + j], fin[2048 + j - 1])) VREJ("final mask rs code not constant");
    // query answers: every value pair and digest, and one path job per
    // authentication path; digests 0, 1 = roots L, H, 2 + k = FRI root k
    std::vector<fe> vals;
    std::vector<uint8_t> dg(32 * 2 + froot.size());
    memcpy(dg.data(), rootL, 32);
    memcpy(dg.data() + 32, rootH, 32);
    memcpy(dg.data() + 64, froot.data(), froot.size());
    std::vector<launch::path_job> jobs;
    std::vector<int64_t> vpos((size_t)33 * nf), alpos(33), ahpos(33);
    auto take_vals = [&]() { int64_t o = (int64_t)vals.size(); for (int k = 0; k < 130; ++k) vals.push_back(R.f()); return o; };
    auto take_path = [&](int plen) {
        int64_t o = (int64_t)(dg.size() / 32);
        size_t at = dg.size();
        dg.resize(at + (size_t)32 * plen);
        R.get(dg.data() + at, (size_t)32 * plen);
        return o;
    };
    std::vector<fe> pts(66);
    fe wS = root_of_unity(c->log_SS);
    for (int rep = 0; rep < 33; ++rep) {
        int64_t pw = S.pow[rep];
        for (int i = 0; i < nf; ++i) {
            int64_t mod = (int64_t)1 << (n - 1 - i);
            if (i > 0) pw %= mod;
            int64_t s0p = pw / 2, s1p = (pw + mod) / 2;
            if (i == 0) {
                int64_t mn = s0p < s1p ? s0p : s1p;
                alpos[rep] = take_vals();
                int64_t pl = take_path(n - 1);
                ahpos[rep] = take_vals();
                int64_t ph = take_path(n - 1);
                jobs.push_back({mn, 0, pl, alpos[rep], n - 1});
                jobs.push_back({mn, 1, ph, ahpos[rep], n - 1});
                pts[2 * rep] = fpow(wS, (unsigned __int128)s0p);
                pts[2 * rep + 1] = fpow(wS, (unsigned __int128)s1p);
            }
            vpos[(size_t)rep * nf + i] = take_vals();
            int plen = n - 2 - i;
            int64_t pp = take_path(plen);
            jobs.push_back({pw / 2, 2 + i, pp, vpos[(size_t)rep * nf + i], plen});
        }
    }
    if (R.bad) VREJ("transcript truncated or not canonical");
    // ---- device: public polynomial Q, mask polynomial, their values at the
    // query points, every path ----
    int lcs = ilog2(cs), lm = ilog2(m);
    fe *qv = V.get<fe>(cs), *qc = V.get<fe>(cs), *pubd = V.get<fe>(m), *pubc = V.get<fe>(m), *ptsd = V.get<fe>(66),
       *evq = V.get<fe>(66), *evm = V.get<fe>(66), *vd = V.get<fe>(vals.size());
    uint64_t *dgd = V.get<uint64_t>(dg.size() / 8);
    launch::path_job *jd = V.get<launch::path_job>(jobs.size());
    int *okd = V.get<int>(jobs.size());
    launch::half_desc *hq = V.get<launch::half_desc>(1);
    VNULL(qv); VNULL(qc); VNULL(pubd); VNULL(pubc); VNULL(ptsd); VNULL(evq); VNULL(evm); VNULL(vd); VNULL(dgd); VNULL(jd);
    VNULL(okd); VNULL(hq);
    // q_eval_real of the input point (zk_verifier.cpp:884-908): eq over its
    // low n-6 bits, then the 64 slice weights over the high 6
    launch::half_desc hdq = {qv, nullptr, c->chal + c->chal_ru[1], lcs, 0};
    HIPCHK(hipMemcpyAsync(hq, &hdq, sizeof hdq, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(pubd, pub.data(), sizeof(fe) * m, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ptsd, pts.data(), sizeof(fe) * 66, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(vd, vals.data(), sizeof(fe) * vals.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dgd, dg.data(), dg.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(jd, jobs.data(), sizeof(launch::path_job) * jobs.size(), hipMemcpyHostToDevice, st));
    launch::half_multi(hq, 1, cs, st);
    launch::intt({0, qv, cs, nullptr, nullptr, 0, 0, 0, 0}, qc, cs, 1, lcs, finv(from_u64((uint64_t)cs)), c->NT, st);
    launch::intt({0, pubd, m, nullptr, nullptr, 0, 0, 0, 0}, pubc, m, 1, lm, finv(from_u64((uint64_t)m)), c->NT, st);
    launch::poly_eval(qc, cs, ptsd, 66, evq, st);
    launch::poly_eval(pubc, m, ptsd, 66, evm, st);
    launch::verify_paths(jd, (int)jobs.size(), dgd, vd, okd, st);
    std::vector<fe> hq_ev(66), hm_ev(66);
    std::vector<int> ok(jobs.size());
    HIPCHK(hipMemcpyAsync(hq_ev.data(), evq, sizeof(fe) * 66, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hm_ev.data(), evm, sizeof(fe) * 66, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ok.data(), okd, sizeof(int) * ok.size(), hipMemcpyDeviceToHost, st));
    // meanwhile: the slice weights q_r over the top 6 bits of the input point
    fe qr[65];
    {
        const std::vector<fe> &rin = S.L[1].r_u;
        for (int j = 0; j < 64; ++j) {
            fe w = one;
            for (int b = 0; b < 6; ++b) w = mul(w, ((j >> b) & 1) ? rin[n - 6 + b] : sub(one, rin[n - 6 + b]));
            qr[j] = w;
        }
        qr[64] = one;
    }
    HIPCHK(hipStreamSynchronize(st));
    for (size_t k = 0; k < jobs.size(); ++k)
        if (!ok[k]) VREJ("merkle path %zu failed (query %zu)", k, k / (size_t)(nf + 2));
    // ---- FRI consistency (vpd_verifier.cpp:160-313) ----
    fe inv2 = finv(from_u64(2)), sg = from_u64((uint64_t)(SS / gap)), s32 = from_u64((uint64_t)(SS >> 5));
    for (int rep = 0; rep < 33; ++rep) {
        int64_t pw = S.pow[rep];
        fe rou = root_of_unity(n - 1);
        for (int i = 0; i < nf; ++i) {
            int64_t mod = (int64_t)1 << (n - 1 - i);
            if (i > 0) { rou = mul(rou, rou); pw %= mod; }
            int64_t s0p = pw / 2, s1p = (pw + mod) / 2;
            fe inv_mu = fpow(rou, (unsigned __int128)(mod - pw / 2));      // rou has order mod
            fe rmu = mul(mul(inv2, S.fri_r[i]), inv_mu);
            const fe *v = vals.data() + vpos[(size_t)rep * nf + i];
            if (i == 0) {
                const fe *al = vals.data() + alpos[rep], *ah = vals.data() + ahpos[rep];
                fe x0 = pts[2 * rep], x1 = pts[2 * rep + 1];
                fe ix0 = fpow(wS, (unsigned __int128)(SS - s0p)), ix1 = fpow(wS, (unsigned __int128)(SS - s1p));
                fe mr0 = fpow(x0, (unsigned __int128)(SS / gap)), mr1 = fpow(x1, (unsigned __int128)(SS / gap));
                fe ro0 = fpow(x0, (unsigned __int128)(SS >> 5)), ro1 = fpow(x1, (unsigned __int128)(SS >> 5));
                for (int j = 0; j < 65; ++j) {
                    bool msk = j == 64;
                    fe t0 = msk ? hm_ev[2 * rep] : mul(qr[j], hq_ev[2 * rep]);
                    fe t1 = msk ? hm_ev[2 * rep + 1] : mul(qr[j], hq_ev[2 * rep + 1]);
                    fe z0 = msk ? mr0 : ro0, z1 = msk ? mr1 : ro1, sc = msk ? sg : s32;
                    fe a = sub(mul(al[2 * j], t0), mul(sub(z0, one), ah[2 * j]));
                    a = mul(sub(mul(a, sc), all_sum[j]), ix0);
                    fe b = sub(mul(al[2 * j + 1], t1), mul(sub(z1, one), ah[2 * j + 1]));
                    b = mul(sub(mul(b, sc), all_sum[j]), ix1);
                    if (s0p > s1p) std::swap(a, b);
                    fe pval = add(mul(add(a, b), inv2), mul(sub(a, b), rmu));
                    if (!feq(pval, v[2 * j]) && !feq(pval, v[2 * j + 1]))
                        VREJ("FRI first-round consistency failed, query %d slice %d", rep, j);
                }
            } else {
                const fe *pv = vals.data() + vpos[(size_t)rep * nf + i - 1];
                for (int j = 0; j < 65; ++j) {
                    fe a = pv[2 * j], b = pv[2 * j + 1];
                    fe h0 = mul(add(a, b), inv2);
                    fe p0 = add(h0, mul(sub(a, b), rmu)), p1 = add(h0, mul(sub(b, a), rmu));
                    if (!feq(p0, v[2 * j]) && !feq(p0, v[2 * j + 1]) && !feq(p1, v[2 * j]) && !feq(p1, v[2 * j + 1]))
                        VREJ("FRI consistency failed, query %d round %d", rep, i);
                }
            }
        }
    }
    // ---- the fft_gkr section ----
    int fok = 0;
    rc = fg_verify(c->fg, tr + R.off, len - R.off, &fok);
    if (rc) return rc;
    if (!fok) VREJ("fft_gkr section rejected");
    *accepted = 1;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace vg {
namespace {
struct hip_source : mem_source {
    void *device(int64_t bytes) override
    {
        void *p = nullptr;
        return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? p : nullptr;
    }
    void device_free(void *p) override { (void)hipFree(p); }
    void *pinned(int64_t bytes) override
    {
        void *p = nullptr;
        return hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? p : nullptr;
    }
    void pinned_free(void *p) override { (void)hipHostFree(p); }
};
}  // namespace
mem_source *hip_mem()
{
    static hip_source s;
    return &s;
}
}  // namespace vg

extern "C" {

int vg_init(int device)
{
    HIPCHK(hipSetDevice(device));
    return 0;
}
int vg_stream_create(void **stream)
{
    if (!stream) FAILV(-1, "null argument");
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *stream = st;
    return 0;
}
int vg_stream_destroy(void *stream)
{
    if (stream) HIPCHK(hipStreamDestroy((hipStream_t)stream));
    return 0;
}
const char *vg_last_error(void) { return g_err.c_str(); }
void vg_free(void *p) { free(p); }

}  // extern "C"

// Context creation in two steps, so that a proof server can reserve the
// memory before anything is allocated: ctx_create_planned builds the circuit
// and plans the proof (host work only); ctx_materialize takes the device
// arena and pinned staging from `mem` (nullptr: hipMalloc / hipHostMalloc)
// and uploads.  Both destroy the context on failure.
int ctx_create_planned(vg_ctx **out, const vg_circuit_desc *d, void *stream, bool host_only)
{
    vg_ctx *c = new vg_ctx();
    int rc = build_circuit(d, c->C);
    if (rc) { delete c; return rc; }
    if (host_only) {
        // sizes only: no stream, the proof's layout as on a caller's stream
        c->serial = true;
        rc = ctx_plan(c);
        if (!rc) rc = fg_plan(&c->fg, c->n - 6, nullptr, &c->fg_bytes, &c->fg_pin, true);
        if (rc) { vg_ctx_destroy(c); return rc; }
        *out = c;
        return 0;
    }
    if (stream) c->st = (hipStream_t)stream;
    else {
        hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
        if (e != hipSuccess) { delete c; FAILV(-20, "hipStreamCreate: %s", hipGetErrorString(e)); }
        c->own_stream = true;
    }
    {
        // A caller's stream runs the whole proof in order on one stream:
        // several contexts then overlap on the device without sharing
        // hardware queues.  Otherwise the context owns three streams (main:
        // evaluation and sumcheck; commitment; fft_gkr), each on its own
        // hardware queue, so that a small proof's latency-bound chains
        // overlap (median ms, serial vs overlapped: 16x16 matmul 1.63 vs
        // 1.36, 2^16 x 12 layers 2.66 vs 2.31, 2^18 x 16 3.81 vs 3.55); from
        // 2^20 inputs on, the sumcheck and the extensions each keep the VALU
        // busy and overlapping them gains nothing (2^20 x 16: 12.24 vs
        // 12.50), so such proofs run in order, fft_gkr still on its own
        // stream.  VIRGO_SERIAL=1 / =0 forces one choice.
        const char *e = getenv("VIRGO_SERIAL");
        c->serial = stream != nullptr || (e && e[0] == '1') || (!(e && e[0] == '0') && c->C.bl[0] >= 20);
    }
    {
        hipError_t e = hipSuccess;
        if (c->serial) {
            c->st2 = c->st;
            for (int w = 0; w < vg_ctx::NGS; ++w) c->gst[w] = c->st;
            // in order on its own streams because of its size: the input
            // layer's extension still runs beside the evaluation (VIRGO_EARLY_L=0
            // keeps it in order too)
            const char *es = getenv("VIRGO_SERIAL"), *el = getenv("VIRGO_EARLY_L");
            if (c->own_stream && !(es && es[0] == '1') && !(el && el[0] == '0')) {
                e = hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking);
                if (e == hipSuccess) {
                    c->own_aux = true;
                    c->early_l = true;
                }
            }
        } else {
            e = hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking);
            c->own_aux = true;
            // The sumcheck layers run on the main stream, or on k streams of
            // their own (VIRGO_GKR_STREAMS=k).  Every stream takes one of the
            // process's hardware queues (4 by default) in turn, and kernels of
            // streams that share a queue run one after another: the main
            // stream, the commitment's and fft_gkr's then hold three distinct
            // queues and truly overlap.
            int k = 1;
            if (const char *g = getenv("VIRGO_GKR_STREAMS")) k = std::max(1, std::min(vg_ctx::NGS, atoi(g)));
            for (int w = 0; w < vg_ctx::NGS; ++w) c->gst[w] = c->st;
            for (int w = 0; k > 1 && w < k && e == hipSuccess; ++w)
                if ((e = hipStreamCreateWithFlags(&c->gst[w], hipStreamNonBlocking)) == hipSuccess) c->own_gst = w + 1;
            for (int w = k; k > 1 && w < vg_ctx::NGS; ++w) c->gst[w] = c->gst[w % k];
        }
        if (e != hipSuccess) { vg_ctx_destroy(c); FAILV(-20, "hipStreamCreate: %s", hipGetErrorString(e)); }
    }
    rc = ctx_plan(c);
    if (rc) { vg_ctx_destroy(c); return rc; }
    // fft_gkr depends on nothing of the main proof: a context that owns its
    // streams gives it one of its own even when the main proof runs in
    // order (its latency-bound launches then hide beside the main proof's
    // large ones: 12.4 vs 13.0 ms at 2^20 gates x 16 layers); on a caller's
    // stream everything stays in order
    rc = fg_plan(&c->fg, c->n - 6, c->own_stream ? nullptr : c->st, &c->fg_bytes, &c->fg_pin);
    if (rc) { vg_ctx_destroy(c); return rc; }
    *out = c;
    return 0;
}

// device and pinned host bytes a planned context will take from its source
void ctx_needs(const vg_ctx *c, int64_t *dev, int64_t *pin)
{
    *dev = ((c->arena_bytes + 255) & ~(int64_t)255) + ((c->fg_bytes + 255) & ~(int64_t)255);
    *pin = c->pin_bytes + uploader::stage_bytes(c->up_bytes) + c->fg_pin;
}

int ctx_materialize(vg_ctx *c, mem_source *mem, vg_ctx **out)
{
    c->mem = mem ? mem : hip_mem();
    int rc = ctx_alloc(c);
    if (!rc) rc = fg_alloc(c->fg, c->mem);
    if (rc) { vg_ctx_destroy(c); return rc; }
    *out = c;
    return 0;
}

extern "C" {

int vg_ctx_create(vg_ctx **out, const vg_circuit_desc *d, void *stream)
{
    if (!out || !d) FAILV(-1, "null argument");
    *out = nullptr;
    vg_ctx *c = nullptr;
    int rc = ctx_create_planned(&c, d, stream);
    if (rc) return rc;
    return ctx_materialize(c, nullptr, out);
}

int vg_plan(const vg_circuit_desc *d, int64_t *device_bytes, int64_t *pinned_bytes)
{
    if (!d || !device_bytes || !pinned_bytes) FAILV(-1, "null argument");
    *device_bytes = *pinned_bytes = 0;
    vg_ctx *c = nullptr;
    int rc = ctx_create_planned(&c, d, nullptr, true);
    if (rc) return rc;
    ctx_needs(c, device_bytes, pinned_bytes);
    return vg_ctx_destroy(c);     // host state only: no stream, event or device memory was made
}

int vg_ctx_destroy(vg_ctx *c)
{
    if (!c) return 0;
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->own_aux) {
        if (c->st2) (void)hipStreamSynchronize(c->st2);
        for (int w = 0; w < vg_ctx::NGS; ++w)
            if (c->gst[w]) (void)hipStreamSynchronize(c->gst[w]);
    }
    fg_destroy(c->fg);
    if (c->mem) {
        if (c->arena) c->mem->device_free(c->arena);
        if (c->h_chal) c->mem->pinned_free(c->h_chal);
        if (c->h_gsrc) c->mem->pinned_free(c->h_gsrc);
        if (c->h_gout) c->mem->pinned_free(c->h_gout);
        if (c->h_msg) c->mem->pinned_free(c->h_msg);
        if (c->h_up) c->mem->pinned_free(c->h_up);
    }
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    for (auto &e : c->ev_leaf) if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->ev_in, c->ev_chal, c->ev_l0}) if (e) (void)hipEventDestroy(e);
    if (c->own_stream && c->st) (void)hipStreamDestroy(c->st);
    if (c->own_aux && c->st2) (void)hipStreamDestroy(c->st2);
    for (int w = 0; w < c->own_gst; ++w) (void)hipStreamDestroy(c->gst[w]);
    for (int w = 0; w < vg_ctx::NGS; ++w)
        if (c->gev[w]) (void)hipEventDestroy(c->gev[w]);
    delete c;
    return 0;
}

int vg_ntt_plan(int log_n, int *tile_log, float *us)
{
    if (!tile_log) FAILV(-1, "vg_ntt_plan: null argument");
    float t[4] = {0, 0, 0, 0};
    *tile_log = launch::ntt_plan(log_n, t);
    if (us) for (int k = 0; k < 4; ++k) us[k] = t[k];
    return 0;
}

int64_t vg_ctx_device_bytes(const vg_ctx *c) { return c ? c->arena_bytes + fg_device_bytes(c->fg) : 0; }

int vg_prove_async(vg_ctx *c)
{
    int rc = prove_start(c);
    if (rc) return rc;
    rc = enqueue_post(c);
    if (rc) return rc;
    c->host_ms[1] = ms_since(c->t_enq);
    c->pending = true;
    return 0;
}

int vg_finish(vg_ctx *c, uint8_t **transcript, size_t *len, double *stage_ms)
{
    if (!c || !c->pending) FAILV(-1, "no proof in flight");
    c->pending = false;
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipGetLastError());
    auto t2 = std::chrono::steady_clock::now();
    tbuf T;
    // the transcript is the messages and query answers plus framing: one
    // allocation instead of a doubling vector's twenty-odd copies
    T.reserve((size_t)(16 * (c->msg_len + c->n_gather)) + ((size_t)1 << 16));
    assemble(c, T);
    double fg_ms = 0;
    int rc = fg_finish(c->fg, T, nullptr, nullptr, &fg_ms);
    if (rc) return rc;
    c->host_ms[2] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
    *len = T.size();
    *transcript = T.release();            // the caller frees it with vg_free
    if (stage_ms) {
        // evaluate, gkr | lde_l, lde_qh, fri_fold, hash (second stream) | queries, leaf kernel, total
        // (in order on one stream, lde_l starts where gkr ends)
        static const int pairs[9][2] = {{0, 1}, {1, 2}, {1, 3}, {3, 4}, {4, 5}, {5, 6}, {8, 7}, {-1, -1}, {0, 7}};
        for (int k = 0; k < 9; ++k) {
            float ms = 0;
            int a = k == 2 && c->serial ? 2 : pairs[k][0];
            if (c->early_l && k == 3) a = 2;                   // lde_qh starts once the sumcheck is done
            if (pairs[k][0] < 0) (void)hipEventElapsedTime(&ms, c->ev_leaf[0], c->ev_leaf[1]);
            else if (c->early_l && k == 2) (void)hipEventElapsedTime(&ms, c->ev_l0, c->ev[3]);
            else (void)hipEventElapsedTime(&ms, c->ev[a], c->ev[pairs[k][1]]);
            stage_ms[k] = ms;
        }
        stage_ms[9] = fg_ms;
        for (int k = 0; k < 3; ++k) stage_ms[10 + k] = c->host_ms[k];
    }
    return 0;
}

int vg_prove(vg_ctx *c, uint8_t **transcript, size_t *len, double *stage_ms)
{
    int rc = vg_prove_async(c);
    if (rc) return rc;
    return vg_finish(c, transcript, len, stage_ms);
}

int vg_prove_batch(vg_ctx *const *ctxs, int n, int workers, uint8_t **transcripts, size_t *lens)
{
    if (n < 0 || (n > 0 && (!ctxs || !transcripts || !lens))) FAILV(-1, "vg_prove_batch: bad arguments");
    for (int i = 0; i < n; ++i) { transcripts[i] = nullptr; lens[i] = 0; }
    if (n == 0) return 0;
    if (workers < 1) workers = 1;
    if (workers > n) workers = n;
    // fn(i) for every i in idx, worker k taking idx[k], idx[k + workers],
    // ...; returns the first error (its thread-local message moved here)
    auto run = [&](const std::vector<int> &idx, const std::function<int(int)> &fn) {
        int m = (int)idx.size(), nw = workers < m ? workers : m;
        std::vector<int> rc(nw, 0);
        std::vector<std::string> msg(nw);
        std::vector<std::thread> th;
        for (int k = 0; k < nw; ++k)
            th.emplace_back([&, k]() {
                for (int j = k; j < m && rc[k] == 0; j += nw) {
                    rc[k] = fn(idx[j]);
                    if (rc[k]) msg[k] = g_err;
                }
            });
        for (auto &t : th) t.join();
        for (int k = 0; k < nw; ++k)
            if (rc[k]) { g_err = msg[k]; return rc[k]; }
        return 0;
    };
    // The proofs run in groups, pipelined: (1a) every proof of a group up to
    // its sumcheck layers (evaluation, challenges, half tables) and (1b) its
    // commitment up to the leaf chains, on the proofs' streams; (2) the
    // group's sumcheck layers in merged launches (one per step and table
    // width across its proofs) once every (1a) is done, and its leaf chains
    // and Merkle trees in merged launches (60 % of a mixed batch's VALU work)
    // once every (1b) is done, on two streams of the batch; (3) the rest of
    // each proof.  The dependencies of (2) and (3) are resolved on the host
    // (helper threads wait for the events, then enqueue): a stream waiting on
    // the device would stall the whole hardware queue it shares with other
    // streams.  Group g+1's (1a) and (1b) are enqueued before group g's (3),
    // so the proofs' streams run group g+1's latency-bound stages beside
    // group g's full-width merged launches.
    int ngroups = n >= 96 ? 4 : n >= 32 ? 2 : 1;
    if (const char *e = getenv("VIRGO_BATCH_GROUPS")) ngroups = std::max(1, std::min(n, atoi(e)));
    struct group {
        std::vector<int> idx;
        std::vector<vg_ctx *> S;
        hipEvent_t done = nullptr, gdone = nullptr;
        launch::leaf_seg *dG = nullptr;
        int64_t *dstart = nullptr;
        launch::mk_item *dI = nullptr;
        void *gbuf = nullptr, *gpin = nullptr, *pin = nullptr;   // device / pinned launch tables
        void *pin3 = nullptr, *d3 = nullptr;                       // the same for (3)
        bool ok = false, gok = false;
        std::string err;
        std::thread tg, tl;                                        // helper threads of (2)
    };
    std::vector<group> grp(ngroups);
    for (int g = 0; g < ngroups; ++g)
        for (int i = (int)((int64_t)n * g / ngroups); i < (int)((int64_t)n * (g + 1) / ngroups); ++i) grp[g].idx.push_back(i);
    std::vector<char> started(n, 0);
    hipStream_t ls = nullptr, gs = nullptr;
    bool streams_ok = hipStreamCreateWithFlags(&ls, hipStreamNonBlocking) == hipSuccess &&
                      hipStreamCreateWithFlags(&gs, hipStreamNonBlocking) == hipSuccess;
    int rc = 0;
    auto merged_gkr = [&](group &G) {
        bool ok = streams_ok && hipEventCreateWithFlags(&G.gdone, hipEventDisableTiming) == hipSuccess;
        for (vg_ctx *c : G.S) ok = ok && hipEventSynchronize(c->ev[9]) == hipSuccess;
        if (ok && batch_gkr(G.S.data(), (int)G.S.size(), gs, G.gdone, &G.gbuf, &G.gpin, false) != 0) { ok = false; G.err = g_err; }
        G.gok = ok;
    };
    auto merged = [&](group &G) {
        std::vector<launch::leaf_seg> segs;
        std::vector<int64_t> start(1, 0);
        for (vg_ctx *c : G.S) {
            launch::leaf_seg sg[24];
            int ns = commit_segments(c, sg);
            for (int k = 0; k < ns; ++k) { segs.push_back(sg[k]); start.push_back(start.back() + sg[k].S / 2); }
        }
        bool ok = streams_ok && hipEventCreateWithFlags(&G.done, hipEventDisableTiming) == hipSuccess;
        // every Merkle tree of the group: the large levels from item tables,
        // then the tails one block per tree
        std::vector<launch::mk_item> items;
        std::vector<size_t> lvl_off;
        std::vector<int64_t> lvl_tot;
        for (int h = 1;; ++h) {
            int64_t tot = 0;
            size_t o = items.size();
            for (auto &sg : segs) {
                int64_t cnt = (sg.S / 2) >> h;
                if (cnt > launch::MERKLE_TAIL) { items.push_back({sg.tree, cnt, tot}); tot += cnt; }
            }
            if (!tot) break;
            lvl_off.push_back(o);
            lvl_tot.push_back(tot);
        }
        // segment, start and item tables: one pinned staging block, one
        // device block
        size_t bG = sizeof(launch::leaf_seg) * segs.size(), bS = sizeof(int64_t) * start.size(),
               bI = sizeof(launch::mk_item) * items.size(), oS = (bG + 255) & ~(size_t)255,
               oI = oS + ((bS + 255) & ~(size_t)255), tot = oI + bI + 256;
        uint8_t *dblk = nullptr;
        ok = ok && hipHostMalloc(&G.pin, tot, hipHostMallocDefault) == hipSuccess &&
             hipMalloc((void **)&dblk, tot) == hipSuccess;
        if (ok) {
            uint8_t *h = (uint8_t *)G.pin;
            memcpy(h, segs.data(), bG);
            memcpy(h + oS, start.data(), bS);
            if (bI) memcpy(h + oI, items.data(), bI);
            G.dG = (launch::leaf_seg *)dblk;
            G.dstart = (int64_t *)(dblk + oS);
            G.dI = (launch::mk_item *)(dblk + oI);
        }
        for (vg_ctx *c : G.S) ok = ok && hipEventSynchronize(c->ev[5]) == hipSuccess;
        ok = ok && hipMemcpyAsync(dblk, G.pin, tot, hipMemcpyHostToDevice, ls) == hipSuccess;
        if (ok) {
            launch::leaf_hash_table(G.dG, G.dstart, (int)segs.size(), start.back(), ls);
            for (size_t l = 0; l < lvl_off.size(); ++l) {
                size_t end = l + 1 < lvl_off.size() ? lvl_off[l + 1] : items.size();
                launch::merkle_level_table(G.dI + lvl_off[l], (int)(end - lvl_off[l]), lvl_tot[l], ls);
            }
            launch::merkle_tails_table(G.dG, (int)segs.size(), ls);
            ok = hipEventRecord(G.done, ls) == hipSuccess && hipGetLastError() == hipSuccess;
        }
        G.ok = ok;
    };
    // (3) of a group in two launches on the leaf stream, after its Merkle
    // trees: every proof's slot reduction (once the merged sumcheck layers
    // are done), then every proof's query answers and message block copied
    // straight into its page-locked host staging
    auto post = [&](group &G) {
        G.tg.join();
        G.tl.join();
        bool ok = G.ok && G.gok && hipEventSynchronize(G.gdone) == hipSuccess;
        std::vector<launch::slot_job> sj;
        std::vector<launch::copy_job> cj;
        std::vector<int64_t> ss(1, 0), cs(1, 0);
        for (vg_ctx *c : G.S) {
            if (c->n_slots) {
                sj.push_back({c->slots, c->slot_nb, c->slot_dst, c->msg, c->slot_rof, c->chal});
                ss.push_back(ss.back() + c->n_slots);
            }
            if (c->n_gather) { cj.push_back({c->gsrc, nullptr, c->h_gout}); cs.push_back(cs.back() + c->n_gather); }
            cj.push_back({nullptr, c->msg, c->h_msg});
            cs.push_back(cs.back() + c->msg_len);
        }
        size_t b1 = sizeof(launch::slot_job) * sj.size(), b2 = sizeof(int64_t) * ss.size(),
               b3 = sizeof(launch::copy_job) * cj.size(), b4 = sizeof(int64_t) * cs.size();
        size_t o2 = (b1 + 255) & ~(size_t)255, o3 = o2 + ((b2 + 255) & ~(size_t)255), o4 = o3 + ((b3 + 255) & ~(size_t)255),
               tot = o4 + b4 + 256;
        uint8_t *d = nullptr;
        ok = ok && hipHostMalloc(&G.pin3, tot, hipHostMallocDefault) == hipSuccess && hipMalloc((void **)&d, tot) == hipSuccess;
        if (ok) {
            G.d3 = d;
            uint8_t *h = (uint8_t *)G.pin3;
            if (b1) memcpy(h, sj.data(), b1);
            memcpy(h + o2, ss.data(), b2);
            if (b3) memcpy(h + o3, cj.data(), b3);
            memcpy(h + o4, cs.data(), b4);
            ok = hipMemcpyAsync(d, h, tot, hipMemcpyHostToDevice, ls) == hipSuccess;
        }
        if (ok) {
            launch::reduce_slots_table((const launch::slot_job *)d, (const int64_t *)(d + o2), (int)sj.size(), ss.back(), ls);
            launch::copy_table((const launch::copy_job *)(d + o3), (const int64_t *)(d + o4), (int)cj.size(), cs.back(), ls);
            ok = hipGetLastError() == hipSuccess;
        }
        if (!ok && rc == 0) { g_err = G.err.empty() ? std::string("vg_prove_batch: merged launches failed") : G.err; rc = -20; }
        for (int i : G.idx)
            if (started[i]) ctxs[i]->pending = true;
    };
    for (int g = 0; g < ngroups; ++g) {
        group &G = grp[g];
        // the evaluations, challenges and half tables of every proof of the
        // group first, so that the merged sumcheck layers start early, beside
        // the commitments enqueued after them on the same streams
        int r = run(G.idx, [&](int i) { int q = prove_start_gkr(ctxs[i], false); started[i] = q == 0; return q; });
        if (r && rc == 0) rc = r;
        for (int i : G.idx)
            if (started[i]) G.S.push_back(ctxs[i]);
        G.tg = std::thread([&, g] { merged_gkr(grp[g]); });
        r = run(G.idx, [&](int i) { return started[i] ? prove_start_commit(ctxs[i]) : 0; });
        if (r && rc == 0) rc = r;
        G.tl = std::thread([&, g] { merged(grp[g]); });
        if (g > 0) post(grp[g - 1]);
    }
    post(grp[ngroups - 1]);
    // every (3) done: the proofs' answers are in their host staging
    if (ls && hipStreamSynchronize(ls) != hipSuccess && rc == 0) { g_err = "vg_prove_batch: leaf stream failed"; rc = -20; }
    // transcripts (a failed merge still drains every proof in flight; its
    // transcripts are dropped)
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    int rc3 = run(all, [&](int i) {
        if (!started[i] || !ctxs[i]->pending) return 0;
        return vg_finish(ctxs[i], &transcripts[i], &lens[i], nullptr);
    });
    if (ls) (void)hipStreamSynchronize(ls);
    if (gs) (void)hipStreamSynchronize(gs);
    for (auto &G : grp) {
        if (G.dG) (void)hipFree(G.dG);                 // the block holding dstart and dI too
        if (G.gbuf) (void)hipFree(G.gbuf);
        if (G.gpin) (void)hipHostFree(G.gpin);
        if (G.pin) (void)hipHostFree(G.pin);
        if (G.pin3) (void)hipHostFree(G.pin3);
        if (G.d3) (void)hipFree(G.d3);
        if (G.done) (void)hipEventDestroy(G.done);
        if (G.gdone) (void)hipEventDestroy(G.gdone);
    }
    if (ls) (void)hipStreamDestroy(ls);
    if (gs) (void)hipStreamDestroy(gs);
    int first = rc ? rc : rc3;
    if (first)
        for (int i = 0; i < n; ++i) { free(transcripts[i]); transcripts[i] = nullptr; lens[i] = 0; }
    return first;
}

int vg_verify(vg_ctx *c, const uint8_t *transcript, size_t len, int *accepted, double *ms)
{
    if (!c || !transcript || !accepted) FAILV(-1, "vg_verify: null argument");
    if (c->pending) FAILV(-1, "vg_verify: a proof is in flight on this context");
    auto t0 = std::chrono::steady_clock::now();
    for (hipStream_t s : {c->st2, c->gst[0], c->gst[1], c->gst[2], c->gst[3]})
        if (s && s != c->st) HIPCHK(hipStreamSynchronize(s));
    int rc = verify(c, transcript, len, accepted);
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int vg_layer_values(vg_ctx *c, int layer, vg_fe *out, size_t cap, int *bits)
{
    if (!c || !bits) FAILV(-1, "vg_layer_values: null argument");
    if (layer < 0 || layer >= c->C.depth) FAILV(-1, "vg_layer_values: layer %d outside 0..%d", layer, c->C.depth - 1);
    if (c->pending) FAILV(-1, "vg_layer_values: a proof is in flight on this context");
    int bl = c->C.bl[layer];
    *bits = bl;
    if (!out) return 0;
    if (((size_t)1 << bl) > cap) FAILV(-1, "vg_layer_values: buffer of %zu values holds less than 2^%d", cap, bl);
    HIPCHK(hipMemcpy(out, c->cv[layer], sizeof(fe) << bl, hipMemcpyDeviceToHost));
    return 0;
}

int vg_output(vg_ctx *c, vg_fe *out, size_t cap, int *bits)
{
    if (!c) FAILV(-1, "null context");
    return vg_layer_values(c, c->C.depth - 1, out, cap, bits);
}

int vg_proof_values(vg_ctx *c, int what, vg_fe *out, size_t cap, size_t *count)
{
    if (!c || !count) FAILV(-1, "vg_proof_values: null argument");
    if (c->pending || !c->proved) FAILV(-1, "vg_proof_values: no finished proof on this context");
    const schedule &S = c->S;
    const host_circuit &C = c->C;
    std::vector<fe> v;
    switch (what) {
    case VG_VALUES_DRAWS:
        v.insert(v.end(), S.r0.begin(), S.r0.end());
        v.insert(v.end(), S.r1.begin(), S.r1.end());
        for (int i = C.depth - 1; i >= 1; --i) {
            const layer_sched &l = S.L[i];
            v.push_back(l.rho);
            if (i > 1) v.insert(v.end(), l.maskR, l.maskR + 6);
            v.insert(v.end(), l.maskr, l.maskr + 2);
            v.insert(v.end(), l.maskpoly.begin(), l.maskpoly.end());
            v.insert(v.end(), l.r_u.begin(), l.r_u.end());
            v.insert(v.end(), l.r_v.begin(), l.r_v.end());
            for (const fe &x : {l.r_c, l.random_combine, l.linear_coeff, l.linear_combine, l.alpha, l.beta}) v.push_back(x);
        }
        v.insert(v.end(), S.fri_r.begin(), S.fri_r.end());
        break;
    case VG_VALUES_GENERAL:
        for (int i = C.depth - 1; i >= 1; --i)
            v.push_back(i > 1 ? c->h_msg[c->msg_layer[i] + 6 * c->LH[i].bl + 10] : fzero());
        break;
    case VG_VALUES_PUB_MASK: v = c->pub_mask; break;
    case VG_VALUES_PRI_MASK: v = c->pri_mask; break;
    default: FAILV(-1, "vg_proof_values: unknown selector %d", what);
    }
    *count = v.size();
    if (out) memcpy(out, v.data(), sizeof(fe) * (v.size() < cap ? v.size() : cap));
    return 0;
}

int vg_query_positions(vg_ctx *c, int64_t *pos)
{
    if (!c || !pos) FAILV(-1, "vg_query_positions: null argument");
    if (c->pending || !c->proved) FAILV(-1, "vg_query_positions: no finished proof on this context");
    for (int r = 0; r < 33; ++r) pos[r] = c->S.pow[r];
    return 0;
}

}  // extern "C"
