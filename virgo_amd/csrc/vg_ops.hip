// Single-operation entry points of the C ABI on host arrays: the drop-in
// replacements of fast_fourier_transform / inverse_fast_fourier_transform
// (RS_polynomial.cpp:27-240), my_hhash (my_hhash.h:39-46),
// merkle_tree_prover::create_tree (merkle_tree.cpp:4-57), one FRI fold of a
// codeword (fri.cpp:321-340) and the field product (prime_field.h:125-151).  Each call copies in, runs the HIP kernels
// on the null stream and copies out.  Also the kernel micro-set (transform,
// inverse, extension, Merkle tree, FRI commit of one polynomial, timed per
// stage on the device) and the SHA-256 variants of the hash and tree.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vg_kernels.hpp"
#include "../../include/virgo.h"

using namespace vg;

void vg_set_error(const std::string &msg);   // vg_prover.hip

namespace {
struct err_proxy {
    void operator=(const std::string &m) { vg_set_error(m); }
    void operator=(const char *m) { vg_set_error(m); }
} op_err;
int fail(const char *what, hipError_t e)
{
    vg_set_error(std::string(what) + ": " + hipGetErrorString(e));
    return -20;
}
#define OPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return fail(#x, e_); } while (0)

struct dbuf {
    void *p = nullptr;
    ~dbuf() { if (p) (void)hipFree(p); }
};
int ilog2(int64_t x) { int l = 0; while (((int64_t)1 << l) < x) ++l; return l; }
bool pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }

// tw[e] = w^e for e < n on the device (from two small tables)
int make_powers(const fe &w, int64_t n, dbuf &tw)
{
    std::vector<fe> t(1024 + (n >> 10) + 1);
    fe x = fone();
    for (int j = 0; j < 1024; ++j) { t[j] = x; x = mul(x, w); }
    fe y = fone();
    for (int64_t j = 0; j < (n >> 10) + 1; ++j) { t[1024 + j] = y; y = mul(y, x); }
    dbuf T12;
    OPCHK(hipMalloc(&T12.p, sizeof(fe) * t.size()));
    OPCHK(hipMemcpy(T12.p, t.data(), sizeof(fe) * t.size(), hipMemcpyHostToDevice));
    OPCHK(hipMalloc(&tw.p, sizeof(fe) * (n ? n : 1)));
    launch::twiddles((fe *)tw.p, n, (fe *)T12.p, (fe *)T12.p + 1024, 0);
    OPCHK(hipDeviceSynchronize());
    return 0;
}
// tw[e] = w^e for e < n on the device, and the stage-major tables of an
// NTT of size n built from them
int make_twiddles(const fe &w, int64_t n, dbuf &tw, dbuf &W, launch::ntt_tabs &T)
{
    int rc = make_powers(w, n, tw);
    if (rc) return rc;
    int lg = ilog2(n);
    OPCHK(hipMalloc(&W.p, sizeof(fe) * 2 * n));
    fe *w0 = (fe *)W.p, *w1 = w0 + n;
    launch::ntt_tables(w0, w1, lg, nullptr, nullptr, 0, 0, 0, (const fe *)tw.p, n, 0);
    T = {w0, w1, nullptr, nullptr, 0, 0};
    OPCHK(hipDeviceSynchronize());
    return 0;
}
}  // namespace

// ---- working memory of the allocation-free transform (vg_fft_noalloc) ----
// One static device array per GPU, held until the process ends: the powers
// tw[e] = w^e (e < order), the forward and inverse stage-major tables
// W[order], Wi[order] (k_ntt_tables), the zero-padded input pad[order], and
// the two-level power table T12 (w^j for j < 1024, then w^(1024 k) for
// k <= order / 1024) that k_twiddles multiplies.  Sized for order <= 2^20:
// 4 * 2^20 + 2049 values of 16 bytes, 64.03 MiB.
constexpr int NA_MAX_LOG = 20;
constexpr int64_t NA_MAX = (int64_t)1 << NA_MAX_LOG;
constexpr int64_t NA_T12 = 1024 + (NA_MAX >> 10) + 1;
constexpr int64_t NA_VALUES = 4 * NA_MAX + NA_T12;
static __device__ __align__(256) uint64_t scratch[2 * NA_VALUES];

// T[j] = w^j for j < 1024 and T[1024 + k] = w^(1024 k) for k < m, each by
// square-and-multiply (exact, so the same values as repeated products)
__global__ void k_pow_tables(fe *T, fe w, int64_t m)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= 1024 + m) return;
    uint64_t e = j < 1024 ? (uint64_t)j : (uint64_t)(j - 1024) << 10;
    fe x = w, acc = fone();
    while (e) {
        if (e & 1) acc = mul(acc, x);
        x = mul(x, x);
        e >>= 1;
    }
    T[j] = acc;
}
// pad[i] = src[i] for i < len, 0 for len <= i < n (a kernel, not a memset, so
// that the clear is an ordinary node when a caller captures the stream)
__global__ void k_pad_copy(const fe *src, int64_t len, fe *pad, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) pad[i] = i < len ? src[i] : fzero();
}

// VALU issue-rate kernels for vg_valu_rates: 16 independent 32-bit chains
// per lane, iters steps of one instruction class (CLS 0: three-input
// bitwise op, v_bitop3_b32; CLS 1: funnel shift, v_alignbit_b32); thread 0
// of block 0 records its shader clock and 100 MHz constant clock spans.
template <int CLS>
__global__ __launch_bounds__(256) void k_issue(uint32_t *out, int iters, unsigned long long *clk)
{
    uint32_t a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = threadIdx.x * (k + 3) + blockIdx.x;
    unsigned long long t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 16; ++k)
            a[k] = CLS == 0 ? a[k] ^ (a[(k + 5) & 15] & ~a[(k + 9) & 15])
                            : __builtin_amdgcn_alignbit(a[k], a[(k + 5) & 15], 13);
    unsigned long long t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) x ^= a[k];
    out[blockIdx.x * (int64_t)blockDim.x + threadIdx.x] = x;
    if (threadIdx.x == 0 && blockIdx.x == 0) { clk[0] = t1 - t0; clk[1] = r1 - r0; }
}

namespace {
std::mutex g_na_mu;
constexpr int NA_DEVICES = 64;
fe *g_na_ptr[NA_DEVICES];        // device address of scratch, per device
// The device address of this GPU's copy of scratch, looked up once per
// device with that device current.  The host-side name of the array is not a
// device address and is never handed to a kernel or a copy.
int na_scratch(fe **out)
{
    int dev = 0;
    OPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_DEVICES) { op_err = "vg_fft_noalloc: device index out of range"; return -1; }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        OPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) { op_err = "vg_fft_noalloc: no device address for the scratch array"; return -20; }
        g_na_ptr[dev] = (fe *)p;
    }
    *out = g_na_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int vg_fft(const vg_fe *coef, int coef_len, int order, const vg_fe *rou, vg_fe *result)
{
    if (!pow2(coef_len) || !pow2(order) || coef_len > order) { op_err = "vg_fft: sizes must be powers of two, coef_len <= order"; return -1; }
    fe w; memcpy(&w, rou, sizeof w);
    dbuf tw, buf, W;
    launch::ntt_tabs T;
    int rc = make_twiddles(w, order, tw, W, T);
    if (rc) return rc;
    OPCHK(hipMalloc(&buf.p, sizeof(fe) * order));
    OPCHK(hipMemset(buf.p, 0, sizeof(fe) * order));
    OPCHK(hipMemcpy(buf.p, coef, sizeof(fe) * coef_len, hipMemcpyHostToDevice));
    dbuf out;
    OPCHK(hipMalloc(&out.p, sizeof(fe) * order));
    launch::ntt({0, (const fe *)buf.p, order, nullptr, nullptr, 0, 0, 0, 0}, (fe *)out.p, order, 1, ilog2(order), 0, fone(), 0,
                T, 0);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(result, out.p, sizeof(fe) * order, hipMemcpyDeviceToHost));
    return 0;
}

int vg_ifft(const vg_fe *evals, int coef_len, int order, const vg_fe *rou, vg_fe *result)
{
    if (!pow2(coef_len) || !pow2(order)) { op_err = "vg_ifft: sizes must be powers of two"; return -1; }
    if (coef_len > order) coef_len = order;            // RS_polynomial.cpp:171-177
    int64_t step = order / coef_len;
    std::vector<fe> sub(coef_len);
    for (int i = 0; i < coef_len; ++i) memcpy(&sub[i], &evals[(int64_t)i * step], sizeof(fe));
    fe w; memcpy(&w, rou, sizeof w);
    fe nw = fone();
    for (int64_t i = 0; i < step; ++i) nw = mul(nw, w);
    dbuf tw, in, out, W;
    launch::ntt_tabs T;
    int rc = make_twiddles(nw, coef_len, tw, W, T);
    if (rc) return rc;
    OPCHK(hipMalloc(&in.p, sizeof(fe) * coef_len));
    OPCHK(hipMalloc(&out.p, sizeof(fe) * coef_len));
    OPCHK(hipMemcpy(in.p, sub.data(), sizeof(fe) * coef_len, hipMemcpyHostToDevice));
    fe inv_n = finv(mk((uint64_t)coef_len, 0));
    launch::ntt({0, (const fe *)in.p, coef_len, nullptr, nullptr, 0, 0, 0, 0}, (fe *)out.p, coef_len, 1, ilog2(coef_len), 1,
                inv_n, 1, T, 0);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(result, out.p, sizeof(fe) * coef_len, hipMemcpyDeviceToHost));
    return 0;
}

int vg_sha3_64(const uint8_t *in, uint8_t *out, size_t n)
{
    if (!n) return 0;
    dbuf a, b;
    OPCHK(hipMalloc(&a.p, 64 * n));
    OPCHK(hipMalloc(&b.p, 32 * n));
    OPCHK(hipMemcpy(a.p, in, 64 * n, hipMemcpyHostToDevice));
    launch::sha3_batch((const uint8_t *)a.p, (uint8_t *)b.p, (int64_t)n, 0);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(out, b.p, 32 * n, hipMemcpyDeviceToHost));
    return 0;
}

int vg_sha256_64(const uint8_t *in, uint8_t *out, size_t n)
{
    if (!n) return 0;
    dbuf a, b;
    OPCHK(hipMalloc(&a.p, 64 * n));
    OPCHK(hipMalloc(&b.p, 32 * n));
    OPCHK(hipMemcpy(a.p, in, 64 * n, hipMemcpyHostToDevice));
    launch::sha256_batch((const uint8_t *)a.p, (uint8_t *)b.p, (int64_t)n, 0);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(out, b.p, 32 * n, hipMemcpyDeviceToHost));
    return 0;
}

static int merkle_tree_with(const uint8_t *leaves, int n_leaf, uint8_t *tree, int hash, const char *name)
{
    if (n_leaf < 1) { op_err = std::string(name) + ": no leaves"; return -1; }
    int64_t pad = 1;
    while (pad < n_leaf) pad *= 2;
    dbuf t;
    OPCHK(hipMalloc(&t.p, 64 * pad));
    OPCHK(hipMemset(t.p, 0, 64 * pad));
    OPCHK(hipMemcpy((uint8_t *)t.p + 32 * pad, leaves, 32 * (size_t)n_leaf, hipMemcpyHostToDevice));
    launch::merkle_pad((uint8_t *)t.p, pad, n_leaf, 0, hash);
    uint8_t *tp = (uint8_t *)t.p;
    int64_t npad = pad;
    launch::merkle_multi(1, &tp, &npad, 0, hash);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(tree, t.p, 64 * pad, hipMemcpyDeviceToHost));
    return 0;
}

int vg_merkle_tree(const uint8_t *leaves, int n_leaf, uint8_t *tree)
{
    return merkle_tree_with(leaves, n_leaf, tree, launch::HASH_SHA3, "vg_merkle_tree");
}

int vg_merkle_tree_sha256(const uint8_t *leaves, int n_leaf, uint8_t *tree)
{
    return merkle_tree_with(leaves, n_leaf, tree, launch::HASH_SHA256, "vg_merkle_tree_sha256");
}

int vg_microset(int log_n, int rate_log, int hash, const vg_fe *coef, const vg_fe *fri_r, int reps, uint8_t *roots,
                vg_fe *final_cw, int *roundtrip_ok, double *stage_ms)
{
    if (log_n < 5 || log_n > 26 || rate_log < 1 || rate_log > 5 || log_n + rate_log > 30 ||
        (hash != VG_HASH_SHA3_256 && hash != VG_HASH_SHA256) || reps < 1 || !coef || !fri_r || !roots || !final_cw) {
        op_err = "vg_microset: bad arguments (5 <= log_n <= 26, 1 <= rate_log <= 5, log_n + rate_log <= 30, "
                 "hash VG_HASH_SHA3_256 or VG_HASH_SHA256, reps >= 1, no null pointers)";
        return -1;
    }
    const int L = log_n + rate_log, nfold = L - 5;
    const int64_t N = (int64_t)1 << log_n, M = (int64_t)1 << L, R = (int64_t)1 << rate_log;
    const int th = (log_n + 1) / 2, tb = log_n - th;
    dbuf T12, tw, W, TA, TB, lens, cf, F, G, E, tree, ftree, cw, rr, rts, cnt, pbuf;
    // powers tw[e] = w_M^e, stage tables for transforms up to N, coset twists of R rows
    std::vector<fe> t(1024 + (M >> 10) + 1);
    fe w = root_of_unity(L), x = fone();
    for (int j = 0; j < 1024; ++j) { t[j] = x; x = mul(x, w); }
    fe y = fone();
    for (int64_t j = 0; j < (M >> 10) + 1; ++j) { t[1024 + j] = y; y = mul(y, x); }
    OPCHK(hipMalloc(&T12.p, sizeof(fe) * t.size()));
    OPCHK(hipMemcpy(T12.p, t.data(), sizeof(fe) * t.size(), hipMemcpyHostToDevice));
    OPCHK(hipMalloc(&tw.p, sizeof(fe) * M));
    OPCHK(hipMalloc(&W.p, sizeof(fe) * 2 * N));
    OPCHK(hipMalloc(&TA.p, sizeof(fe) * (R << th)));
    OPCHK(hipMalloc(&TB.p, sizeof(fe) * (R << tb)));
    fe *tw_ = (fe *)tw.p, *W0 = (fe *)W.p;
    launch::twiddles(tw_, M, (const fe *)T12.p, (const fe *)T12.p + 1024, 0);
    launch::ntt_tables(W0, W0 + N, log_n, (fe *)TA.p, (fe *)TB.p, rate_log, th, tb, tw_, M, 0);
    launch::ntt_tabs T = {W0, W0 + N, (const fe *)TA.p, (const fe *)TB.p, th, tb};
    int len = (int)N;
    OPCHK(hipMalloc(&lens.p, sizeof(int)));
    OPCHK(hipMemcpy(lens.p, &len, sizeof(int), hipMemcpyHostToDevice));
    // data: coefficients, transform, inverse, extension (coset-major), its
    // tree, the folded codewords of every level and their trees
    OPCHK(hipMalloc(&cf.p, sizeof(fe) * N));
    OPCHK(hipMalloc(&pbuf.p, sizeof(fe) * N));
    OPCHK(hipMalloc(&F.p, sizeof(fe) * N));
    OPCHK(hipMalloc(&G.p, sizeof(fe) * N));
    OPCHK(hipMalloc(&E.p, sizeof(fe) * M));
    OPCHK(hipMalloc(&tree.p, (size_t)32 * M));
    OPCHK(hipMalloc(&ftree.p, (size_t)32 * M));
    OPCHK(hipMalloc(&cw.p, sizeof(fe) * M));
    OPCHK(hipMalloc(&rr.p, sizeof(fe) * nfold));
    OPCHK(hipMalloc(&rts.p, (size_t)32 * nfold));
    OPCHK(hipMalloc(&cnt.p, sizeof(unsigned int)));
    OPCHK(hipMemcpy(cf.p, coef, sizeof(fe) * N, hipMemcpyHostToDevice));
    OPCHK(hipMemcpy(rr.p, fri_r, sizeof(fe) * nfold, hipMemcpyHostToDevice));
    const fe *cf_ = (const fe *)cf.p, *r_ = (const fe *)rr.p;
    fe *F_ = (fe *)F.p, *G_ = (fe *)G.p, *E_ = (fe *)E.p;
    uint8_t *tr_ = (uint8_t *)tree.p, *rt_ = (uint8_t *)rts.p;
    // codeword of level k >= 1 (M >> k values) at lv[k]; its tree (levels
    // with more than 32 values, (M >> k) / 2 leaves) at lt[k]
    std::vector<fe *> lv(nfold + 1, nullptr);
    std::vector<uint8_t *> lt(nfold + 1, nullptr);
    std::vector<int64_t> ln(nfold + 1, 0);
    {
        int64_t off = 0, toff = 0;
        for (int k = 1; k <= nfold; ++k) {
            lv[k] = (fe *)cw.p + off;
            off += M >> k;
            if (k < nfold) {
                lt[k] = (uint8_t *)ftree.p + toff;
                ln[k] = (M >> k) / 2;
                toff += 64 * ln[k];
            }
        }
    }
    lt[0] = tr_;
    if (launch::ntt_tune(log_n, R, E_, (fe *)tree.p, T, 0)) { op_err = "vg_microset: NTT tile tuning failed"; return -20; }
    const fe inv_n = finv(mk((uint64_t)N, 0));
    const bool big = log_n >= 22;
    fe *P_ = (fe *)pbuf.p;
    hipEvent_t ev[6] = {};
    struct evs { hipEvent_t *e; ~evs() { for (int i = 0; i < 6; ++i) if (e[i]) (void)hipEventDestroy(e[i]); } } ev_guard{ev};
    for (auto &e : ev) OPCHK(hipEventCreate(&e));
    double acc[5] = {0, 0, 0, 0, 0};
    for (int rep = 0; rep <= reps; ++rep) {          // rep 0 warms up, untimed
        // transforms of 2^22 values and more read a bit-reversed copy (one
        // tiled permutation) instead of gathering at stride 2^(log_n-12):
        // the extension of 2^24 values takes 7.8 instead of 8.6 ms; below
        // 2^22 the permutation costs more than it saves
        OPCHK(hipEventRecord(ev[0], 0));
        if (big) {
            launch::brev_permute(cf_, N, P_, log_n, 0);
            launch::ntt({0, P_, N, nullptr, nullptr, 0, 0, 0, 0, 1}, F_, N, 1, log_n, 0, fone(), 0, T, 0);
        } else {
            launch::ntt({0, cf_, N, nullptr, nullptr, 0, 0, 0, 0}, F_, N, 1, log_n, 0, fone(), 0, T, 0);
        }
        OPCHK(hipEventRecord(ev[1], 0));
        if (big) {
            launch::brev_permute(F_, N, P_, log_n, 0);
            launch::ntt({0, P_, N, nullptr, nullptr, 0, 0, 0, 0, 1}, G_, N, 1, log_n, 1, inv_n, 1, T, 0);
        } else {
            launch::ntt({0, F_, N, nullptr, nullptr, 0, 0, 0, 0}, G_, N, 1, log_n, 1, inv_n, 1, T, 0);
        }
        OPCHK(hipEventRecord(ev[2], 0));
        if (big) {
            launch::brev_permute(cf_, N, P_, log_n, 0);
            launch::lde(P_, N, (const int *)lens.p, 1, E_, log_n, rate_log, T, 0, 0, 1);
        } else {
            launch::lde(cf_, N, (const int *)lens.p, 1, E_, log_n, rate_log, T, 0);
        }
        OPCHK(hipEventRecord(ev[3], 0));
        // level 0: the tree of the extension's pairs (E[j], E[j + M/2])
        int64_t half = M / 2;
        launch::pair_leaves(E_, half, rate_log, N, tr_, 0);
        launch::merkle_multi(1, &tr_, &half, 0, hash);
        OPCHK(hipEventRecord(ev[4], 0));
        // FRI: every fold first (each writes the next codeword and its tree's
        // leaves), then the trees of all levels together
        for (int k = 0; k < nfold; ++k) {
            half = (M >> k) / 2;
            if (k == 0) launch::fri_fold_any(E_, half, rate_log, N, r_, tw_, M, 0, lv[1], lt[1], 0);
            else launch::fri_fold_any(lv[k], half, 0, 0, r_ + k, tw_, M, k, lv[k + 1], lt[k + 1], 0);
        }
        if (nfold > 1) launch::merkle_multi(nfold - 1, lt.data() + 1, ln.data() + 1, 0, hash);
        if (launch::copy_roots(lt.data(), nfold, rt_, 0)) { op_err = "vg_microset: too many levels"; return -1; }
        OPCHK(hipEventRecord(ev[5], 0));
        OPCHK(hipGetLastError());
        OPCHK(hipEventSynchronize(ev[5]));
        if (rep == 0) continue;
        for (int sgi = 0; sgi < 5; ++sgi) {
            float ms = 0;
            OPCHK(hipEventElapsedTime(&ms, ev[sgi], ev[sgi + 1]));
            acc[sgi] += ms;
        }
    }
    OPCHK(hipMemset(cnt.p, 0, sizeof(unsigned int)));
    launch::count_diff(G_, cf_, N, (unsigned int *)cnt.p, 0);
    unsigned int nd = 0;
    OPCHK(hipMemcpy(&nd, cnt.p, sizeof nd, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(roots, rts.p, (size_t)32 * nfold, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(final_cw, lv[nfold], sizeof(fe) * 32, hipMemcpyDeviceToHost));
    if (roundtrip_ok) *roundtrip_ok = nd == 0;
    if (stage_ms)
        for (int sgi = 0; sgi < 5; ++sgi) stage_ms[sgi] = acc[sgi] / reps;
    return 0;
}

int vg_fri_fold(const vg_fe *cw, int log_n, const vg_fe *r, vg_fe *out)
{
    if (log_n < 1 || log_n > 30 || !cw || !r || !out) { op_err = "vg_fri_fold: bad arguments (1 <= log_n <= 30)"; return -1; }
    int64_t n = (int64_t)1 << log_n, half = n / 2;
    dbuf iw, in, o, rr;
    int rc = make_powers(finv(root_of_unity(log_n)), half, iw);
    if (rc) return rc;
    OPCHK(hipMalloc(&in.p, sizeof(fe) * n));
    OPCHK(hipMalloc(&o.p, sizeof(fe) * half));
    OPCHK(hipMalloc(&rr.p, sizeof(fe)));
    OPCHK(hipMemcpy(in.p, cw, sizeof(fe) * n, hipMemcpyHostToDevice));
    OPCHK(hipMemcpy(rr.p, r, sizeof(fe), hipMemcpyHostToDevice));
    launch::fri_fold_nat((const fe *)in.p, half, (const fe *)rr.p, (const fe *)iw.p, (fe *)o.p, 0);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(out, o.p, sizeof(fe) * half, hipMemcpyDeviceToHost));
    return 0;
}

int vg_valu_rates(double *rates)
{
    if (!rates) { op_err = "vg_valu_rates: null argument"; return -1; }
    int dev = 0, ncu = 0;
    OPCHK(hipGetDevice(&dev));
    OPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // 8 blocks of 256 threads per CU: 8 waves on each SIMD
    const int blocks = 8 * ncu, iters = 20000;
    dbuf out, clk;
    OPCHK(hipMalloc(&out.p, (size_t)blocks * 256 * 4));
    OPCHK(hipMalloc(&clk.p, 16));
    double mhz = 0;
    for (int cls = 0; cls < 2; ++cls) {
        for (int rep = 0; rep < 2; ++rep) {              // the first run warms up and is discarded
            auto kfn = cls == 0 ? k_issue<0> : k_issue<1>;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            OPCHK(hipEventCreate(&e0));
            OPCHK(hipEventCreate(&e1));
            OPCHK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(kfn, dim3(blocks), dim3(256), 0, 0, (uint32_t *)out.p, iters, (unsigned long long *)clk.p);
            OPCHK(hipGetLastError());
            OPCHK(hipEventRecord(e1, 0));
            OPCHK(hipEventSynchronize(e1));
            float ms = 0;
            OPCHK(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            unsigned long long h[2] = {0, 0};
            OPCHK(hipMemcpy(h, clk.p, 16, hipMemcpyDeviceToHost));
            if (rep == 0) continue;
            double f = h[1] ? h[0] * 100.0 / h[1] : 0.0;   // MHz: shader clocks per 100 MHz tick
            double per_simd = (double)blocks * 4 * iters * 16 / (ncu * 4.0);
            rates[cls] = f > 0 ? per_simd / (ms * 1e-3 * f * 1e6) : 0.0;
            mhz += f / 2;
        }
    }
    rates[2] = mhz;
    return 0;
}

int vg_fe_mul(const vg_fe *a, const vg_fe *b, vg_fe *out, size_t n)
{
    if (!n) return 0;
    dbuf x, y, z;
    OPCHK(hipMalloc(&x.p, 16 * n));
    OPCHK(hipMalloc(&y.p, 16 * n));
    OPCHK(hipMalloc(&z.p, 16 * n));
    OPCHK(hipMemcpy(x.p, a, 16 * n, hipMemcpyHostToDevice));
    OPCHK(hipMemcpy(y.p, b, 16 * n, hipMemcpyHostToDevice));
    launch::fe_mul((const fe *)x.p, (const fe *)y.p, (fe *)z.p, (int64_t)n, 0);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(out, z.p, 16 * n, hipMemcpyDeviceToHost));
    return 0;
}

int vg_fft_noalloc(const vg_fe *coef, int coef_len, int order, const vg_fe *rou, vg_fe *result, void *stream)
{
    if (!coef || !rou || !result) { op_err = "vg_fft_noalloc: null pointer"; return -1; }
    if (!pow2(coef_len) || !pow2(order) || coef_len > order || order < 2) {
        op_err = "vg_fft_noalloc: sizes must be powers of two, coef_len <= order, order >= 2";
        return -1;
    }
    if (order > NA_MAX) {
        op_err = "vg_fft_noalloc: order " + std::to_string(order) + " exceeds the scratch limit of 2^" +
                 std::to_string(NA_MAX_LOG);
        return -2;
    }
    fe *S = nullptr;
    int rc = na_scratch(&S);
    if (rc) return rc;
    fe *tw = S, *W = S + NA_MAX, *Wi = S + 2 * NA_MAX, *pad = S + 3 * NA_MAX, *T12 = S + 4 * NA_MAX;
    hipStream_t st = (hipStream_t)stream;
    fe w;
    memcpy(&w, rou, sizeof w);
    int64_t n = order, m = (n >> 10) + 1;
    int lg = ilog2(n);
    hipLaunchKernelGGL(k_pow_tables, dim3((unsigned)((1024 + m + 255) / 256)), dim3(256), 0, st, T12, w, m);
    launch::twiddles(tw, n, T12, T12 + 1024, st);
    launch::ntt_tables(W, Wi, lg, nullptr, nullptr, 0, 0, 0, tw, n, st);
    hipLaunchKernelGGL(k_pad_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const fe *)coef,
                       (int64_t)coef_len, pad, n);
    launch::ntt_tabs T = {W, Wi, nullptr, nullptr, 0, 0};
    launch::ntt({0, pad, n, nullptr, nullptr, 0, 0, 0, 0}, (fe *)result, n, 1, lg, 0, fone(), 0, T, st);
    OPCHK(hipGetLastError());
    return 0;
}

}  // extern "C"
