// CDNA4 (gfx950) kernels of the Virgo prover hot path.  All arithmetic is
// exact F_{p^2} integer work on the VALU (no MFMA: there is no dense float
// contraction here).  Reductions and atomics may run in any order because
// every field operation returns the canonical representative.
#include "vg_kernels.hpp"
#include "vg_keccak.hpp"
#include <cstring>

namespace vg {
namespace {

constexpr int TPB = 256;
constexpr int NTT_LM = 11;          // LDS-resident NTT block: 2^11 elements = 32 KB
constexpr int FOLD_TAIL = 512;      // fold rounds with <= this many outputs run in one block
constexpr int TAIL_T = 256;         // threads of that block
constexpr int MAX_PARTS = 1024;     // max blocks of a reduction with per-block partials

__host__ __device__ inline unsigned nblk(int64_t n, int t = TPB) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ uint64_t red64(uint64_t x)       // x mod p for any 64-bit x
{
    uint64_t r = (x & P) + (x >> 61);
    return r >= P ? r - P : r;
}

// ---- exact modular atomics (CAS on each 64-bit component) ----
__device__ __forceinline__ void atomic_add_p(uint64_t *a, uint64_t x)
{
    if (x == 0) return;
    unsigned long long *p = (unsigned long long *)a;
    unsigned long long old = *p, assumed;
    do {
        assumed = old;
        old = atomicCAS(p, assumed, (unsigned long long)addp(assumed, x));
    } while (old != assumed);
}
// Table builds accumulate into split limbs: each canonical component x < 2^61
// is added as (x mod 2^32) and (x >> 32) into two 64-bit counters with
// non-returning atomicAdd, which cannot overflow before 2^32 terms.  The first
// fold round folds the limbs back into canonical field elements.
__device__ __forceinline__ void atomic_add_fe(acc4 *a, const fe &x)
{
    if (x.real) {
        atomicAdd((unsigned long long *)&a->rl, (unsigned long long)(x.real & 0xFFFFFFFFull));
        atomicAdd((unsigned long long *)&a->rh, (unsigned long long)(x.real >> 32));
    }
    if (x.img) {
        atomicAdd((unsigned long long *)&a->il, (unsigned long long)(x.img & 0xFFFFFFFFull));
        atomicAdd((unsigned long long *)&a->ih, (unsigned long long)(x.img >> 32));
    }
}
// (hi * 2^32 + lo) mod p, hi < 2^61, lo < 2^64
__device__ __forceinline__ uint64_t limbs_to_p(uint64_t lo, uint64_t hi)
{
    uint64_t s = ((hi & 0x1FFFFFFFull) << 32) + (hi >> 29) + (lo & P) + (lo >> 61);
    s = (s & P) + (s >> 61);
    return s >= P ? s - P : s;
}
__device__ __forceinline__ fe acc_value(const acc4 &a) { return mk(limbs_to_p(a.rl, a.rh), limbs_to_p(a.il, a.ih)); }

// ---- reductions ----
__device__ __forceinline__ fe wave_sum(fe x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        fe y;
        y.real = __shfl_xor((unsigned long long)x.real, off);
        y.img = __shfl_xor((unsigned long long)x.img, off);
        x = add(x, y);
    }
    return x;
}
// block-wide sum of K field elements; result valid in thread 0
template <int K>
__device__ __forceinline__ void block_sum(fe *v)
{
    __shared__ fe red[K][16];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < K; ++k) red[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) {
            fe s = red[k][0];
            for (int i = 1; i < nw; ++i) s = add(s, red[k][i]);
            v[k] = s;
        }
    __syncthreads();
}

// ---------------------------------------------------------------------
// circuit evaluation (zk_prover.cpp:54-157)
// ---------------------------------------------------------------------
__global__ void k_zero(uint32_t *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}
__global__ void k_eval_input(const int32_t *u, fe *out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = mk(red64((uint64_t)(int64_t)u[i]), 0);   // field_element((int)u), :60-65
}
__global__ void k_eval_layer(const uint8_t *ty, const int32_t *gu, const int32_t *gv, const fe *p, fe *o, int n)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int u = gu[g], v = gv[g];
    fe r;
    switch (ty[g]) {
    case 0: r = add(p[u], p[v]); break;
    case 1: r = mul(p[u], p[v]); break;
    case 3: r = mk(red64((uint64_t)(int64_t)u), 0); break;
    case 4: case 10: r = p[u]; break;
    case 5: { r = fzero(); for (int k = u; k < v; ++k) r = add(r, p[k]); break; }
    case 6: r = sub(fone(), p[u]); break;
    case 7: r = sub(p[u], p[v]); break;
    case 8: { fe x = p[u], y = p[v]; r = sub(add(x, y), mul(mul(mk(2, 0), x), y)); break; }
    case 9: { fe x = p[u], y = p[v]; r = sub(y, mul(x, y)); break; }
    case 12: { r = fzero(); for (int k = u; k <= v; ++k) r = add(r, mul_base(p[k], 1ull << (k - u))); break; }
    case 13: r = mul(p[u], sub(fone(), p[v])); break;
    default: r = fzero(); break;
    }
    o[g] = r;
}

// ---------------------------------------------------------------------
// multilinear evaluation by halving (V_res, zk_prover.cpp:30-52)
// ---------------------------------------------------------------------
__global__ void k_mle_fold(const fe *in, fe *out, int64_t n_out, const fe *rp)
{
    fe r = *rp, om = sub(fone(), r);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_out; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = add(mul(in[2 * i], om), mul(in[2 * i + 1], r));
}
__global__ __launch_bounds__(1024) void k_mle_tail(fe *a, fe *b, int64_t n, const fe *r, int nr, fe *dst)
{
    for (int k = 0; k < nr; ++k) {
        fe rr = r[k], om = sub(fone(), rr);
        int64_t no = n >> 1;
        for (int64_t i = threadIdx.x; i < no; i += blockDim.x) b[i] = add(mul(a[2 * i], om), mul(a[2 * i + 1], rr));
        __syncthreads();
        fe *t = a; a = b; b = t;
        n = no;
    }
    if (threadIdx.x == 0) *dst = a[0];
}

// ---------------------------------------------------------------------
// eq / beta half tables (zk_prover.cpp:391-418, 802-822)
// ---------------------------------------------------------------------
__global__ void k_half_tables(fe *f, fe *s, const fe *f0, const fe *r, int fh, int sh)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    fe one = fone();
    if (j < (1 << fh)) {
        fe acc = f0 ? *f0 : one;
        for (int i = 0; i < fh; ++i) acc = mul(acc, ((j >> i) & 1) ? r[i] : sub(one, r[i]));
        f[j] = acc;
    } else {
        j -= 1 << fh;
        if (j < (1 << sh)) {
            fe acc = one;
            for (int i = 0; i < sh; ++i) acc = mul(acc, ((j >> i) & 1) ? r[fh + i] : sub(one, r[fh + i]));
            s[j] = acc;
        }
    }
}
__global__ void k_eq_expand(fe *out, const fe *lo, const fe *hi, int fh, int64_t n)
{
    int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (x < n) out[x] = mul(lo[x & ((1 << fh) - 1)], hi[x >> fh]);
}

struct bg_t { const fe *g0f, *g0s, *g1f, *g1s; int fh; };
struct bu_t { const fe *f, *s; int fh; };
__device__ __forceinline__ fe beta_g(const bg_t &b, int i)
{
    int lo = i & ((1 << b.fh) - 1), hi = i >> b.fh;
    return add(mul(b.g0f[lo], b.g0s[hi]), mul(b.g1f[lo], b.g1s[hi]));
}
__device__ __forceinline__ fe beta_u(const bu_t &b, int u)
{
    return mul(b.f[u & ((1 << b.fh) - 1)], b.s[u >> b.fh]);
}

// ---------------------------------------------------------------------
// sumcheck phase-1 table build: scatter by u (zk_prover.cpp:420-587)
// ---------------------------------------------------------------------
__global__ void k_p1_scatter(const uint8_t *gty, const int32_t *gu, const int32_t *gv, int n, const fe *pv, bg_t bg,
                             acc4 *amb, acc4 *aVb)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ty = gty[i], u = gu[i], v = gv[i];
    if (ty == 2 || ty == 3) return;
    fe tmp = beta_g(bg, i);
    switch (ty) {
    case 0: atomic_add_fe(&aVb[u], mul(pv[v], tmp)); atomic_add_fe(&amb[u], tmp); break;
    case 1: atomic_add_fe(&amb[u], mul(pv[v], tmp)); break;
    case 5: for (int j = u; j < v; ++j) atomic_add_fe(&amb[j], tmp); break;
    case 12: for (int j = u; j <= v; ++j) { atomic_add_fe(&amb[j], tmp); tmp = dbl(tmp); } break;
    case 4: atomic_add_fe(&amb[u], beta_g(bg, u)); break;
    case 6: atomic_add_fe(&amb[u], neg(tmp)); atomic_add_fe(&aVb[u], tmp); break;
    case 7: atomic_add_fe(&aVb[u], neg(mul(pv[v], tmp))); atomic_add_fe(&amb[u], tmp); break;
    case 8: { fe tv = mul(tmp, pv[v]); atomic_add_fe(&aVb[u], tv); atomic_add_fe(&amb[u], sub(tmp, dbl(tv))); break; }
    case 13: { fe tv = mul(tmp, pv[v]); atomic_add_fe(&amb[u], sub(tmp, tv)); break; }
    case 9: { fe tv = mul(tmp, pv[v]); atomic_add_fe(&aVb[u], tv); atomic_add_fe(&amb[u], neg(tv)); break; }
    case 10: atomic_add_fe(&amb[u], tmp); break;
    default: break;
    }
}

// ---------------------------------------------------------------------
// sumcheck phase-2 table build: scatter by v (zk_prover.cpp:845-972).
// Gates whose target is index 0 (sum, exp-sum, NOT, relay) are reduced in
// the block first.
// ---------------------------------------------------------------------
__global__ void k_p2_scatter(const uint8_t *gty, const int32_t *gu, const int32_t *gv, int n, bg_t bg, bu_t bu,
                             const fe *vup, acc4 *amb, acc4 *aVb)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    fe acc0 = fzero();
    if (i < n) {
        int ty = gty[i], u = gu[i], v = gv[i];
        if (ty != 2 && ty != 3 && ty != 4) {
            fe vu = *vup;
            fe tg = beta_g(bg, i);
            switch (ty) {
            case 1: atomic_add_fe(&amb[v], mul(mul(tg, beta_u(bu, u)), vu)); break;
            case 0: { fe t = mul(tg, beta_u(bu, u)); atomic_add_fe(&amb[v], t); atomic_add_fe(&aVb[v], mul(t, vu)); break; }
            case 5: { fe t = mul(tg, vu); for (int j = u; j < v; ++j) acc0 = add(acc0, mul(t, beta_u(bu, j))); break; }
            case 12: { fe t = mul(tg, vu); for (int j = u; j <= v; ++j) { acc0 = add(acc0, mul(t, beta_u(bu, j))); t = dbl(t); } break; }
            case 6: { fe t = mul(tg, beta_u(bu, u)); acc0 = add(acc0, sub(t, mul(t, vu))); break; }
            case 7: { fe t = mul(tg, beta_u(bu, u)); atomic_add_fe(&amb[v], neg(t)); atomic_add_fe(&aVb[v], mul(t, vu)); break; }
            case 8: { fe t = mul(tg, beta_u(bu, u)), tv = mul(t, vu); atomic_add_fe(&amb[v], sub(t, dbl(tv))); atomic_add_fe(&aVb[v], tv); break; }
            case 13: { fe t = mul(tg, beta_u(bu, u)), tv = mul(t, vu); atomic_add_fe(&amb[v], neg(tv)); atomic_add_fe(&aVb[v], tv); break; }
            case 9: { fe t = mul(tg, beta_u(bu, u)); atomic_add_fe(&amb[v], sub(t, mul(vu, t))); break; }
            case 10: { fe t = mul(tg, beta_u(bu, u)); acc0 = add(acc0, mul(t, vu)); break; }
            default: break;
            }
        }
    }
    int any = __syncthreads_or(!is_zero(acc0));
    if (!any) return;
    block_sum<1>(&acc0);
    if (threadIdx.x == 0) atomic_add_fe(&aVb[0], acc0);
}

// ---------------------------------------------------------------------
// bookkeeping fold (zk_prover.cpp:600-633, 992-1025)
// ---------------------------------------------------------------------
struct tabs { lpoly *V, *aV, *am; };
struct ctabs { const lpoly *V, *aV, *am; };

__device__ __forceinline__ void fold_one(bool first, const fe *Vb, const acc4 *aVb, const acc4 *amb, ctabs in, const fe &r,
                                         int64_t i, lpoly &V, lpoly &aV, lpoly &am)
{
    if (first) {
        fe x0 = Vb[2 * i], x1 = Vb[2 * i + 1];
        V.b = x0; V.a = sub(x1, x0);
        x0 = acc_value(aVb[2 * i]); x1 = acc_value(aVb[2 * i + 1]);
        aV.b = x0; aV.a = sub(x1, x0);
        x0 = acc_value(amb[2 * i]); x1 = acc_value(amb[2 * i + 1]);
        am.b = x0; am.a = sub(x1, x0);
    } else {
        lpoly y0 = in.V[2 * i], y1 = in.V[2 * i + 1];
        V.b = add(mul(y0.a, r), y0.b); V.a = sub(add(mul(y1.a, r), y1.b), V.b);
        y0 = in.aV[2 * i]; y1 = in.aV[2 * i + 1];
        aV.b = add(mul(y0.a, r), y0.b); aV.a = sub(add(mul(y1.a, r), y1.b), aV.b);
        y0 = in.am[2 * i]; y1 = in.am[2 * i + 1];
        am.b = add(mul(y0.a, r), y0.b); am.a = sub(add(mul(y1.a, r), y1.b), am.b);
    }
}
__device__ __forceinline__ void fold_acc(fe *acc, const lpoly &V, const lpoly &aV, const lpoly &am)
{
    acc[0] = add(acc[0], mul(am.a, V.a));
    acc[1] = add(add(acc[1], add(mul(am.a, V.b), mul(am.b, V.a))), aV.a);
    acc[2] = add(add(acc[2], mul(am.b, V.b)), aV.b);
}
// block 0 first reduces the previous round's per-block partials
__device__ void reduce_prev(const fe *prev, int nprev, fe *dst)
{
    if (blockIdx.x != 0 || nprev == 0) return;
    fe acc[3] = {fzero(), fzero(), fzero()};
    for (int k = threadIdx.x; k < nprev; k += blockDim.x)
        for (int c = 0; c < 3; ++c) acc[c] = add(acc[c], prev[3 * k + c]);
    block_sum<3>(acc);
    if (threadIdx.x == 0) { dst[0] = acc[0]; dst[1] = acc[1]; dst[2] = acc[2]; }
}
__global__ __launch_bounds__(TPB) void k_fold(int first, const fe *Vb, const acc4 *aVb, const acc4 *amb, ctabs in, const fe *rp,
                                             tabs out, int64_t n_out, fe *part, const fe *prev_part, int nprev, fe *prev_dst)
{
    reduce_prev(prev_part, nprev, prev_dst);
    fe r = first ? fzero() : *rp;
    fe acc[3] = {fzero(), fzero(), fzero()};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_out; i += (int64_t)gridDim.x * blockDim.x) {
        lpoly V, aV, am;
        fold_one(first, Vb, aVb, amb, in, r, i, V, aV, am);
        out.V[i] = V; out.aV[i] = aV; out.am[i] = am;
        fold_acc(acc, V, aV, am);
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) { part[3 * blockIdx.x] = acc[0]; part[3 * blockIdx.x + 1] = acc[1]; part[3 * blockIdx.x + 2] = acc[2]; }
}
// remaining rounds j0..L-1 in one block; X holds round j0's input (unless j0 == 0)
__global__ __launch_bounds__(TAIL_T) void k_fold_tail(int j0, int L, const fe *Vb, const acc4 *aVb, const acc4 *amb, tabs X, tabs Y,
                                                    const fe *r, fe vconst, const fe *prev_part, int nprev, fe *sums, fe *end)
{
    if (nprev) {
        fe acc[3] = {fzero(), fzero(), fzero()};
        for (int k = threadIdx.x; k < nprev; k += blockDim.x)
            for (int c = 0; c < 3; ++c) acc[c] = add(acc[c], prev_part[3 * k + c]);
        block_sum<3>(acc);
        if (threadIdx.x == 0) { sums[3 * (j0 - 1)] = acc[0]; sums[3 * (j0 - 1) + 1] = acc[1]; sums[3 * (j0 - 1) + 2] = acc[2]; }
    }
    for (int j = j0; j < L; ++j) {
        int64_t n_out = (int64_t)1 << (L - 1 - j);
        fe rr = j == 0 ? fzero() : r[j - 1];
        ctabs in = {X.V, X.aV, X.am};
        fe acc[3] = {fzero(), fzero(), fzero()};
        for (int64_t i = threadIdx.x; i < n_out; i += blockDim.x) {
            lpoly V, aV, am;
            fold_one(j == 0, Vb, aVb, amb, in, rr, i, V, aV, am);
            Y.V[i] = V; Y.aV[i] = aV; Y.am[i] = am;
            fold_acc(acc, V, aV, am);
        }
        block_sum<3>(acc);      // contains __syncthreads: Y complete before the swap
        if (threadIdx.x == 0) { sums[3 * j] = acc[0]; sums[3 * j + 1] = acc[1]; sums[3 * j + 2] = acc[2]; }
        tabs t = X; X = Y; Y = t;
    }
    if (threadIdx.x == 0) {
        lpoly V0 = X.V[0], am0 = X.am[0];
        end[0] = V0.a; end[1] = V0.b; end[2] = am0.a; end[3] = am0.b;
        end[4] = add(add(mul(V0.a, r[L - 1]), V0.b), vconst);
    }
}


// ---------------------------------------------------------------------
// Table builds by gather over CSR transposes (owner computes, no atomics).
// Phase 1 (zk_prover.cpp:420-587): every table entry t sums the gates that
// feed it through u (a sum gate feeds every t in its range).  Phase 2
// (:845-972): entry t sums the gates whose v is t; the gates that feed entry
// 0 unconditionally (sum, exp-sum, NOT, relay) are reduced separately.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_p1_gather(const uint32_t *off, const int32_t *ent, const uint8_t *gty,
                                                  const int32_t *gu, const int32_t *gv, int nt, const fe *pv, bg_t bg,
                                                  fe *amb, fe *aVb)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    fe am = fzero(), aV = fzero();
    for (uint32_t e = off[t]; e < off[t + 1]; ++e) {
        int g = ent[e];
        int ty = gty[g], u = gu[g], v = gv[g];
        if (ty == 4) { am = add(am, beta_g(bg, u)); continue; }
        fe tmp = beta_g(bg, g);
        switch (ty) {
        case 0: aV = add(aV, mul(pv[v], tmp)); am = add(am, tmp); break;
        case 1: am = add(am, mul(pv[v], tmp)); break;
        case 5: am = add(am, tmp); break;
        case 12: am = add(am, mul_base(tmp, 1ull << (t - u))); break;
        case 6: am = sub(am, tmp); aV = add(aV, tmp); break;
        case 7: aV = sub(aV, mul(pv[v], tmp)); am = add(am, tmp); break;
        case 8: { fe tv = mul(tmp, pv[v]); aV = add(aV, tv); am = sub(add(am, tmp), dbl(tv)); break; }
        case 13: { fe tv = mul(tmp, pv[v]); am = add(sub(am, tv), tmp); break; }
        case 9: { fe tv = mul(tmp, pv[v]); aV = add(aV, tv); am = sub(am, tv); break; }
        case 10: am = add(am, tmp); break;
        default: break;
        }
    }
    amb[t] = am;
    aVb[t] = aV;
}
__global__ __launch_bounds__(TPB) void k_p2_gather(const uint32_t *off, const int32_t *ent, const uint8_t *gty,
                                                  const int32_t *gu, int nt, bg_t bg, bu_t bu, const fe *vup, fe *amb,
                                                  fe *aVb)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    fe vu = *vup;
    fe am = fzero(), aV = fzero();
    for (uint32_t e = off[t]; e < off[t + 1]; ++e) {
        int g = ent[e];
        int ty = gty[g], u = gu[g];
        fe tgu = mul(beta_g(bg, g), beta_u(bu, u));
        switch (ty) {
        case 1: am = add(am, mul(tgu, vu)); break;
        case 0: am = add(am, tgu); aV = add(aV, mul(tgu, vu)); break;
        case 7: am = sub(am, tgu); aV = add(aV, mul(tgu, vu)); break;
        case 8: { fe tv = mul(tgu, vu); am = sub(add(am, tgu), dbl(tv)); aV = add(aV, tv); break; }
        case 13: { fe tv = mul(tgu, vu); am = sub(am, tv); aV = add(aV, tv); break; }
        case 9: am = add(am, sub(tgu, mul(vu, tgu))); break;
        default: break;
        }
    }
    amb[t] = am;
    aVb[t] = aV;
}
// gates adding into addV[0] in phase 2; per-block partials
__global__ __launch_bounds__(TPB) void k_p2_zero(const int32_t *list, int n, const uint8_t *gty, const int32_t *gu,
                                                const int32_t *gv, bg_t bg, bu_t bu, const fe *vup, fe *part)
{
    fe acc = fzero();
    fe vu = *vup;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int g = list[k];
        int ty = gty[g], u = gu[g], v = gv[g];
        fe tg = beta_g(bg, g);
        switch (ty) {
        case 5: { fe t = mul(tg, vu); for (int j = u; j < v; ++j) acc = add(acc, mul(t, beta_u(bu, j))); break; }
        case 12: { fe t = mul(tg, vu); for (int j = u; j <= v; ++j) { acc = add(acc, mul(t, beta_u(bu, j))); t = dbl(t); } break; }
        case 6: { fe t = mul(tg, beta_u(bu, u)); acc = add(acc, sub(t, mul(t, vu))); break; }
        case 10: { fe t = mul(tg, beta_u(bu, u)); acc = add(acc, mul(t, vu)); break; }
        default: break;
        }
    }
    block_sum<1>(&acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}
__global__ void k_add_parts(const fe *part, int n, fe *dst)
{
    fe acc = fzero();
    for (int k = threadIdx.x; k < n; k += blockDim.x) acc = add(acc, part[k]);
    block_sum<1>(&acc);
    if (threadIdx.x == 0) *dst = add(*dst, acc);
}

// ---------------------------------------------------------------------
// K fold rounds per launch: a thread reads 2^K consecutive entries of the
// three tables, folds K times in registers and writes one entry; each round's
// (a, b, c) sums are block-reduced into per-block partials [block][K][3].
// ---------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(TPB) void k_fold_k(int j0, const fe *Vb, const fe *aVb, const fe *amb, ctabs in,
                                               const fe *r, tabs out, int64_t n_out, fe *part, const fe *prev_part,
                                               int nprev, int kprev, fe *prev_dst)
{
    if (blockIdx.x == 0 && nprev) {           // finish the previous launch's round sums
        for (int q = 0; q < kprev; ++q) {
            fe acc[3] = {fzero(), fzero(), fzero()};
            for (int k = threadIdx.x; k < nprev; k += blockDim.x)
                for (int c = 0; c < 3; ++c) acc[c] = add(acc[c], prev_part[(k * kprev + q) * 3 + c]);
            block_sum<3>(acc);
            if (threadIdx.x == 0) { prev_dst[3 * q] = acc[0]; prev_dst[3 * q + 1] = acc[1]; prev_dst[3 * q + 2] = acc[2]; }
        }
    }
    const bool first = j0 == 0;
    fe acc[K][3];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q][0] = acc[q][1] = acc[q][2] = fzero();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_out; i += (int64_t)gridDim.x * blockDim.x) {
        lpoly V[1 << K], aV[1 << K], am[1 << K];
        int64_t base = i << K;
        if (first) {
#pragma unroll
            for (int m = 0; m < (1 << K); ++m) {
                V[m].a = fzero(); V[m].b = Vb[base + m];
                aV[m].a = fzero(); aV[m].b = aVb[base + m];
                am[m].a = fzero(); am[m].b = amb[base + m];
            }
        } else {
#pragma unroll
            for (int m = 0; m < (1 << K); ++m) { V[m] = in.V[base + m]; aV[m] = in.aV[base + m]; am[m] = in.am[base + m]; }
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
            bool flat = first && q == 0;
            fe rr = flat ? fzero() : r[j0 + q - 1];      // round j0+q folds with r[j0+q-1]
#pragma unroll
            for (int m = 0; m < (1 << (K - 1 - q)); ++m) {
                lpoly y0, y1, z;
#define VG_FOLD(T) \
                y0 = T[2 * m]; y1 = T[2 * m + 1]; \
                if (flat) { z.b = y0.b; z.a = sub(y1.b, y0.b); } \
                else { z.b = add(mul(y0.a, rr), y0.b); z.a = sub(add(mul(y1.a, rr), y1.b), z.b); } \
                T[m] = z;
                VG_FOLD(V) VG_FOLD(aV) VG_FOLD(am)
#undef VG_FOLD
                fold_acc(acc[q], V[m], aV[m], am[m]);
            }
        }
        out.V[i] = V[0]; out.aV[i] = aV[0]; out.am[i] = am[0];
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        block_sum<3>(acc[q]);
        if (threadIdx.x == 0)
            for (int c = 0; c < 3; ++c) part[(blockIdx.x * K + q) * 3 + c] = acc[q][c];
    }
}

// ---------------------------------------------------------------------
// Wave-cooperative fold of K rounds (K <= 6): lane gi folds the input pair
// (2gi, 2gi+1) of round j0; for each later round the lanes combine their
// neighbours' results with __shfl_down, so only round j0+K-1's table is
// written.  Per-round (a, b, c) sums go to per-block partials [block][K][3].
// ---------------------------------------------------------------------
__device__ __forceinline__ fe shfl_down_fe(const fe &x, int d)
{
    fe y;
    y.real = __shfl_down((unsigned long long)x.real, d);
    y.img = __shfl_down((unsigned long long)x.img, d);
    return y;
}
__device__ __forceinline__ lpoly shfl_down_lp(const lpoly &x, int d)
{
    lpoly y;
    y.a = shfl_down_fe(x.a, d);
    y.b = shfl_down_fe(x.b, d);
    return y;
}
__device__ __forceinline__ lpoly fold_pair(const lpoly &y0, const lpoly &y1, const fe &r)
{
    lpoly z;
    z.b = add(mul(y0.a, r), y0.b);
    z.a = sub(add(mul(y1.a, r), y1.b), z.b);
    return z;
}
template <int K>
__global__ __launch_bounds__(TPB) void k_fold_wave(int j0, const fe *Vb, const fe *aVb, const fe *amb, ctabs in,
                                                  const fe *r, tabs out, int64_t n0, fe *part)
{
    int lane = threadIdx.x & 63;
    int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    int64_t nchunks = n0 >> 6;
    fe acc[K][3];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q][0] = acc[q][1] = acc[q][2] = fzero();
    fe r0 = j0 ? r[j0 - 1] : fzero();
    for (int64_t chunk = wave; chunk < nchunks; chunk += nwaves) {     // uniform per wave
        int64_t gi = (chunk << 6) + lane;
        lpoly V, aV, am;
        if (j0 == 0) {
            fe x0 = Vb[2 * gi], x1 = Vb[2 * gi + 1];
            V.b = x0; V.a = sub(x1, x0);
            x0 = aVb[2 * gi]; x1 = aVb[2 * gi + 1];
            aV.b = x0; aV.a = sub(x1, x0);
            x0 = amb[2 * gi]; x1 = amb[2 * gi + 1];
            am.b = x0; am.a = sub(x1, x0);
        } else {
            V = fold_pair(in.V[2 * gi], in.V[2 * gi + 1], r0);
            aV = fold_pair(in.aV[2 * gi], in.aV[2 * gi + 1], r0);
            am = fold_pair(in.am[2 * gi], in.am[2 * gi + 1], r0);
        }
        fold_acc(acc[0], V, aV, am);
#pragma unroll
        for (int q = 1; q < K; ++q) {
            lpoly V2 = shfl_down_lp(V, 1 << (q - 1));
            lpoly aV2 = shfl_down_lp(aV, 1 << (q - 1));
            lpoly am2 = shfl_down_lp(am, 1 << (q - 1));
            if ((lane & ((1 << q) - 1)) == 0) {
                fe rr = r[j0 + q - 1];
                V = fold_pair(V, V2, rr);
                aV = fold_pair(aV, aV2, rr);
                am = fold_pair(am, am2, rr);
                fold_acc(acc[q], V, aV, am);
            }
        }
        if ((lane & ((1 << (K - 1)) - 1)) == 0) {
            int64_t o = gi >> (K - 1);
            out.V[o] = V; out.aV[o] = aV; out.am[o] = am;
        }
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        block_sum<3>(acc[q]);
        if (threadIdx.x == 0)
            for (int c = 0; c < 3; ++c) part[(int64_t)(q * 3 + c) * launch::FW_BLOCKS + blockIdx.x] = acc[q][c];
    }
}
// round sums of every wave-fold launch of the proof: slot R (a round) holds
// 3 x FW_BLOCKS block partials; nb[R] of them are live; dst[R] indexes msg.
__global__ void k_reduce_slots(const fe *part, const int *nb, const int64_t *dst, fe *msg)
{
    int R = blockIdx.x;
    fe acc[3] = {fzero(), fzero(), fzero()};
    for (int b = threadIdx.x; b < nb[R]; b += blockDim.x)
        for (int c = 0; c < 3; ++c) acc[c] = add(acc[c], part[((int64_t)R * 3 + c) * launch::FW_BLOCKS + b]);
    block_sum<3>(acc);
    if (threadIdx.x == 0) { msg[dst[R]] = acc[0]; msg[dst[R] + 1] = acc[1]; msg[dst[R] + 2] = acc[2]; }
}
// Remaining rounds of a phase in one block, tables in LDS (input of round j0:
// 2^(L-j0) <= TAIL_IN entries, or the flat tables if j0 == 0).  Rounds with
// more than 64 outputs use the block; the last ones run in wave 0 alone,
// folding through shuffles with no barriers.
constexpr int TAIL_IN = 512;
__global__ __launch_bounds__(256) void k_fold_tail_lds(int j0, int L, const fe *Vb, const fe *aVb, const fe *amb, ctabs X,
                                                      const fe *r, fe vconst, fe *sums, fe *end)
{
    __shared__ lpoly sV[TAIL_IN], saV[TAIL_IN], sam[TAIL_IN];
    int n_in = 1 << (L - j0);
    for (int i = threadIdx.x; i < n_in; i += blockDim.x) {
        if (j0 == 0) {
            sV[i].a = fzero(); sV[i].b = Vb[i];
            saV[i].a = fzero(); saV[i].b = aVb[i];
            sam[i].a = fzero(); sam[i].b = amb[i];
        } else {
            sV[i] = X.V[i]; saV[i] = X.aV[i]; sam[i] = X.am[i];
        }
    }
    __syncthreads();
    int j = j0;
    for (; j < L && (1 << (L - 1 - j)) > 64; ++j) {
        int n_out = 1 << (L - 1 - j);
        bool flat = j == 0;
        fe rr = flat ? fzero() : r[j - 1];
        fe acc[3] = {fzero(), fzero(), fzero()};
        lpoly nV, naV, nam;
        bool act = threadIdx.x < n_out;
        if (act) {
            int i = threadIdx.x;
#define VG_TFOLD(S, N) { lpoly y0 = S[2 * i], y1 = S[2 * i + 1]; \
                if (flat) { N.b = y0.b; N.a = sub(y1.b, y0.b); } \
                else N = fold_pair(y0, y1, rr); }
            VG_TFOLD(sV, nV) VG_TFOLD(saV, naV) VG_TFOLD(sam, nam)
#undef VG_TFOLD
            fold_acc(acc, nV, naV, nam);
        }
        __syncthreads();
        if (act) { sV[threadIdx.x] = nV; saV[threadIdx.x] = naV; sam[threadIdx.x] = nam; }
        block_sum<3>(acc);
        if (threadIdx.x == 0) { sums[3 * j] = acc[0]; sums[3 * j + 1] = acc[1]; sums[3 * j + 2] = acc[2]; }
    }
    if (threadIdx.x >= 64) return;
    // wave rounds: after local round q lane i holds entry i >> q (valid if i % 2^q == 0)
    int lane = threadIdx.x;
    lpoly V, aV, am;
    for (int q = 0; j < L; ++j, ++q) {
        int n_out = 1 << (L - 1 - j);
        bool flat = j == 0;
        fe rr = flat ? fzero() : r[j - 1];
        fe acc[3] = {fzero(), fzero(), fzero()};
        if (q == 0) {
            if (lane < n_out) {
                lpoly y0 = sV[2 * lane], y1 = sV[2 * lane + 1];
                if (flat) { V.b = y0.b; V.a = sub(y1.b, y0.b); } else V = fold_pair(y0, y1, rr);
                y0 = saV[2 * lane]; y1 = saV[2 * lane + 1];
                if (flat) { aV.b = y0.b; aV.a = sub(y1.b, y0.b); } else aV = fold_pair(y0, y1, rr);
                y0 = sam[2 * lane]; y1 = sam[2 * lane + 1];
                if (flat) { am.b = y0.b; am.a = sub(y1.b, y0.b); } else am = fold_pair(y0, y1, rr);
                fold_acc(acc, V, aV, am);
            }
        } else {
            lpoly V2 = shfl_down_lp(V, 1 << (q - 1));
            lpoly aV2 = shfl_down_lp(aV, 1 << (q - 1));
            lpoly am2 = shfl_down_lp(am, 1 << (q - 1));
            if ((lane & ((1 << q) - 1)) == 0 && (lane >> q) < n_out) {
                V = fold_pair(V, V2, rr);
                aV = fold_pair(aV, aV2, rr);
                am = fold_pair(am, am2, rr);
                fold_acc(acc, V, aV, am);
            }
        }
        for (int c = 0; c < 3; ++c) acc[c] = wave_sum(acc[c]);
        if (lane == 0) { sums[3 * j] = acc[0]; sums[3 * j + 1] = acc[1]; sums[3 * j + 2] = acc[2]; }
    }
    if (lane == 0) {
        end[0] = V.a; end[1] = V.b; end[2] = am.a; end[3] = am.b;
        end[4] = add(add(mul(V.a, r[L - 1]), V.b), vconst);
    }
}

// every eq half table of the proof in one launch: out[j] = f0 * prod_{i<nb} eq(bit_i(j), r[i])
__global__ void k_half_multi(const launch::half_desc *d, int nd, int64_t total)
{
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= total) return;
    int lo = 0, hi = nd - 1;
    while (lo < hi) {                         // last descriptor with start <= tid
        int mid = (lo + hi + 1) >> 1;
        if (d[mid].start <= tid) lo = mid; else hi = mid - 1;
    }
    const launch::half_desc &h = d[lo];
    int64_t j = tid - h.start;
    fe one = fone();
    fe acc = h.f0 ? *h.f0 : one;
    for (int i = 0; i < h.nbits; ++i) acc = mul(acc, ((j >> i) & 1) ? h.r[i] : sub(one, h.r[i]));
    h.out[j] = acc;
}

// ---------------------------------------------------------------------
// wiring predicates combined with v_u, v_v (zk_verifier.cpp:219-606, :1149)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_general(const uint8_t *gty, const int32_t *gu, const int32_t *gv, int n, bg_t bg,
                                                bu_t bu, bu_t bv, const fe *vup, const fe *vvp, fe *part)
{
    fe vu = *vup, vv = *vvp, one = fone();
    fe uv = mul(vu, vv);
    fe acc = fzero();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int ty = gty[i], u = gu[i], v = gv[i];
        if (ty == 2 || ty == 3 || ty == 4) continue;
        fe tg = beta_g(bg, i);
        fe f;
        switch (ty) {
        case 0: f = add(vu, vv); break;
        case 1: f = uv; break;
        case 6: f = sub(one, vu); break;
        case 7: f = sub(vu, vv); break;
        case 8: f = sub(add(vu, vv), dbl(uv)); break;
        case 9: f = sub(vv, uv); break;
        case 13: f = mul(vu, sub(one, vv)); break;
        default: f = vu; break;                          /* 5, 10, 12 */
        }
        if (ty == 5 || ty == 12) {
            fe b0 = mul(tg, beta_u(bv, 0));
            fe s = fzero();
            for (int j = u; (ty == 5) ? (j < v) : (j <= v); ++j) {
                s = add(s, mul(b0, beta_u(bu, j)));
                if (ty == 12) b0 = dbl(b0);
            }
            acc = add(acc, mul(s, f));
        } else {
            int vv_idx = (ty == 10) ? 0 : v;
            acc = add(acc, mul(mul(mul(tg, beta_u(bu, u)), beta_u(bv, vv_idx)), f));
        }
    }
    block_sum<1>(&acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}
__global__ void k_reduce(const fe *part, int n, fe *dst)
{
    fe acc = fzero();
    for (int k = threadIdx.x; k < n; k += blockDim.x) acc = add(acc, part[k]);
    block_sum<1>(&acc);
    if (threadIdx.x == 0) *dst = acc;
}

// ---------------------------------------------------------------------
// NTT.  DIT radix-2 with bit-reversed loads and natural-order output:
// stages [0, log_m) in LDS on blocks of 2^log_m, later stages in registers
// (radix 2^g passes).  Twiddles w_N^e = tw[e * SS/N] from one table of
// w_SS powers; inverse uses tw[SS - e].
// ---------------------------------------------------------------------
struct ntt_in {
    int kind;                 // 0 plain, 1 l*q product at codeword positions, 2 coset twist
    const fe *src; int64_t stride;
    const fe *l, *q; int64_t cw_S, cw_stride, step; int log_R;
    const int *lens; int log_Rl; int hmode;
};
__device__ __forceinline__ fe ntt_fetch(const ntt_in &in, int64_t b, int64_t j, const fe *tw, int64_t SS)
{
    if (in.kind == 0) return in.src[b * in.stride + j];
    if (in.kind == 1) {
        int64_t k = cw_idx(j * in.step, in.cw_S, in.log_R);
        return mul(in.l[b * in.cw_stride + k], in.q[b * in.cw_stride + k]);
    }
    int64_t slice = b >> in.log_Rl, s = b & ((1 << in.log_Rl) - 1);
    int len = in.lens[slice];
    if (j >= len) return fzero();
    fe c = in.src[slice * in.stride + (in.hmode ? len : 0) + j];
    return mul(c, tw[(s * j) & (SS - 1)]);
}
__device__ __forceinline__ fe twid(const fe *tw, int64_t e, int64_t SS, int inverse)
{
    return inverse ? tw[(SS - e) & (SS - 1)] : tw[e];
}
__global__ __launch_bounds__(TPB) void k_ntt_lds(ntt_in in, fe *out, int64_t out_stride, int log_n, int log_m, int inverse,
                                                fe scale, int do_scale, const fe *tw, int64_t SS, int log_SS)
{
    __shared__ fe sh[1 << NTT_LM];
    int64_t b = blockIdx.y;
    int64_t q = blockIdx.x;
    int M = 1 << log_m;
    for (int i = threadIdx.x; i < M; i += blockDim.x) {
        int64_t pos = (q << log_m) + i;
        int64_t j = log_n ? (int64_t)(__brevll((unsigned long long)pos) >> (64 - log_n)) : 0;
        sh[i] = ntt_fetch(in, b, j, tw, SS);
    }
    __syncthreads();
    for (int s = 0; s < log_m; ++s) {
        int half = 1 << s;
        for (int t = threadIdx.x; t < M / 2; t += blockDim.x) {
            int k = t & (half - 1);
            int i0 = ((t >> s) << (s + 1)) | k, i1 = i0 + half;
            fe w = twid(tw, (int64_t)k << (log_SS - s - 1), SS, inverse);
            fe a = sh[i0], c = mul(w, sh[i1]);
            sh[i0] = add(a, c);
            sh[i1] = sub(a, c);
        }
        __syncthreads();
    }
    fe *o = out + b * out_stride + (q << log_m);
    for (int i = threadIdx.x; i < M; i += blockDim.x) o[i] = do_scale ? mul(sh[i], scale) : sh[i];
}
// stages [s0, s0+G) in registers; thread = (batch, base with bits [s0,s0+G) clear)
template <int G>
__global__ __launch_bounds__(TPB) void k_ntt_regs(fe *data, int64_t stride, int64_t nb, int log_n, int s0, int inverse,
                                                 fe scale, int do_scale, const fe *tw, int64_t SS, int log_SS)
{
    int64_t per = (int64_t)1 << (log_n - G);
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= nb * per) return;
    int64_t b = tid / per, r = tid % per;
    int64_t lo = r & (((int64_t)1 << s0) - 1), hi = r >> s0;
    int64_t base = (hi << (s0 + G)) | lo;
    fe *d = data + b * stride;
    fe x[1 << G];
#pragma unroll
    for (int m = 0; m < (1 << G); ++m) x[m] = d[base + ((int64_t)m << s0)];
#pragma unroll
    for (int t = 0; t < G; ++t) {
        int s = s0 + t;
#pragma unroll
        for (int m = 0; m < (1 << G); ++m) {
            if (m & (1 << t)) continue;
            int m1 = m | (1 << t);
            int64_t e = lo + ((int64_t)(m & ((1 << t) - 1)) << s0);
            fe w = twid(tw, e << (log_SS - s - 1), SS, inverse);
            fe a = x[m], c = mul(w, x[m1]);
            x[m] = add(a, c);
            x[m1] = sub(a, c);
        }
    }
#pragma unroll
    for (int m = 0; m < (1 << G); ++m) d[base + ((int64_t)m << s0)] = do_scale ? mul(x[m], scale) : x[m];
}

// tw[e] = T1[e & 1023] * T2[e >> 10]
__global__ void k_twiddles(fe *tw, int64_t n, const fe *T1, const fe *T2)
{
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < n) tw[e] = mul(T1[e & 1023], T2[e >> 10]);
}

// ---------------------------------------------------------------------
// commitment helpers (poly_commit.h:203-330)
// ---------------------------------------------------------------------
__global__ void k_slice_sums(const fe *lq, int64_t stride, int nslice, int64_t real, int64_t m, fe *c0, fe *all_sum)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslice) return;
    int64_t len = (i == nslice - 1) ? m : real;
    fe c = add(lq[i * stride], lq[i * stride + len]);
    c0[i] = c;
    all_sum[i] = mul_base(c, (uint64_t)len);
}
__global__ void k_vo(const fe *l, const fe *q, const fe *h, const fe *c0, int nslice, int64_t SS, int log_R, int64_t real,
                     int64_t m, const fe *tw, fe *vo)
{
    int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= (int64_t)nslice * SS) return;
    int slice = (int)(id / SS);
    int64_t st = id - slice * SS;
    int64_t cols = SS >> log_R;
    int64_t s = st / cols, t = st - s * cols;
    int64_t p = (t << log_R) + s;
    int64_t e = (slice == nslice - 1) ? m : real;
    fe xe = tw[(p * e) & (SS - 1)], xinv = tw[(SS - p) & (SS - 1)];
    fe g = sub(mul(l[id], q[id]), mul(sub(xe, fone()), h[id]));
    vo[id] = mul_base(mul(sub(g, c0[slice]), xinv), (uint64_t)e);
}

// ---------------------------------------------------------------------
// hashing: leaf chains and Merkle levels
// ---------------------------------------------------------------------
__device__ __forceinline__ void chain_step(uint64_t *h, const fe &a, const fe &b)
{
    uint64_t m[8] = {a.img, a.real, b.img, b.real, h[0], h[1], h[2], h[3]};
    sha3_64(m, h);
}
__device__ __forceinline__ void store_digest(uint8_t *tree, int64_t node, const uint64_t *h)
{
    ulonglong2 *d = (ulonglong2 *)(tree + node * 32);
    d[0] = make_ulonglong2(h[0], h[1]);
    d[1] = make_ulonglong2(h[2], h[3]);
}
// leaf (s-major thread order) -> natural leaf index i and the row/column of its pair
__device__ __forceinline__ int64_t leaf_of(int64_t tid, int64_t nleaf, int log_R)
{
    int64_t cl = nleaf >> log_R;
    if (cl == 0) return tid;
    int64_t s = tid / cl, t = tid - s * cl;
    return (t << log_R) + s;
}
__global__ __launch_bounds__(TPB) void k_merkle_level(uint8_t *tree, int64_t start, int64_t n)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    k += start;
    const uint64_t *c = (const uint64_t *)(tree + 2 * k * 32);
    uint64_t m[8], h[4];
    for (int i = 0; i < 8; ++i) m[i] = c[i];
    sha3_64(m, h);
    store_digest(tree, k, h);
}
__global__ __launch_bounds__(1024) void k_merkle_tail(uint8_t *tree, int64_t n)
{
    for (; n >= 1; n >>= 1) {
        for (int64_t k = n + threadIdx.x; k < 2 * n; k += blockDim.x) {
            const uint64_t *c = (const uint64_t *)(tree + 2 * k * 32);
            uint64_t m[8], h[4];
            for (int i = 0; i < 8; ++i) m[i] = c[i];
            sha3_64(m, h);
            store_digest(tree, k, h);
        }
        __syncthreads();
    }
}
// FRI fold of 65 slices (fri.cpp:321-340, 367-379): next[p] = ((a+b) + w_S^-p r (a-b)) / 2,
// a = prev[p], b = prev[p + S/2]; one thread per (slice, position), row-contiguous.
__global__ __launch_bounds__(TPB) void k_fri_fold(const fe *prev, int64_t S, int log_R, const fe *rp, const fe *tw, int64_t SS,
                                                 fe *next)
{
    int64_t Sn = S / 2;
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= 65 * Sn) return;
    int64_t j = tid / Sn, st = tid - j * Sn;
    int64_t cols = Sn >> log_R;
    int64_t srow = st / cols, t = st - srow * cols;
    int64_t p = (t << log_R) + srow;
    fe a = prev[j * S + cw_idx(p, S, log_R)], b = prev[j * S + cw_idx(p + Sn, S, log_R)];
    fe imr = mul(tw[(SS - p * (SS / S)) & (SS - 1)], *rp);
    const fe inv2 = mk(1ull << 60, 0);
    next[tid] = mul(inv2, add(add(a, b), mul(imr, sub(a, b))));
}
struct leaf_segs {
    int n;
    int log_R;
    const fe *cw[24];
    int64_t S[24];
    uint8_t *tree[24];
    int64_t start[25];          // prefix sums of leaf counts
};
// Merkle leaves of every commitment at once: a leaf i of a codeword with S
// positions per slice chains SHA3 over the 64 slice pairs (i, i+S/2) and then
// the mask pair (fri.cpp:97-126, 397-420); the digest goes to tree[S/2 + i].
__global__ __launch_bounds__(TPB) void k_leaf_multi(leaf_segs L)
{
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= L.start[L.n]) return;
    int sg = 0;
    while (tid >= L.start[sg + 1]) ++sg;
    tid -= L.start[sg];
    int64_t S = L.S[sg], half = S / 2;
    const fe *cw = L.cw[sg];
    int64_t i = leaf_of(tid, half, L.log_R);
    int64_t k0 = cw_idx(i, S, L.log_R), k1 = cw_idx(i + half, S, L.log_R);
    uint64_t h[4] = {0, 0, 0, 0};
    for (int j = 0; j <= 64; ++j) chain_step(h, cw[j * S + k0], cw[j * S + k1]);
    store_digest(L.tree[sg], half + i, h);
}
struct tree_set {
    int n;
    uint8_t *tree[24];
    int64_t nleaf[24];
};
// level h (h >= 1) of every tree that has it: nodes [nleaf>>h, 2*(nleaf>>h))
__global__ __launch_bounds__(TPB) void k_merkle_multi(tree_set T, int h, int64_t total)
{
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= total) return;
    int sg = 0;
    for (; sg < T.n; ++sg) {
        int64_t c = T.nleaf[sg] >> h;
        if (tid < c) break;
        tid -= c;
    }
    if (sg >= T.n) return;
    int64_t k = (T.nleaf[sg] >> h) + tid;
    const uint64_t *c = (const uint64_t *)(T.tree[sg] + 2 * k * 32);
    uint64_t m[8], hh[4];
    for (int q = 0; q < 8; ++q) m[q] = c[q];
    sha3_64(m, hh);
    store_digest(T.tree[sg], k, hh);
}
__global__ void k_gather16(const uint64_t *src, int64_t n, fe *out)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) out[k] = *(const fe *)(uintptr_t)src[k];
}
__global__ void k_sha3_batch(const uint64_t *in, uint64_t *out, int64_t n)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint64_t m[8], h[4];
    for (int i = 0; i < 8; ++i) m[i] = in[8 * k + i];
    sha3_64(m, h);
    for (int i = 0; i < 4; ++i) out[4 * k + i] = h[i];
}
// pad leaves of a Merkle tree: H(0^64) (merkle_tree.cpp:105-117)
__global__ void k_merkle_pad(uint8_t *tree, int64_t pad, int64_t n_leaf)
{
    int64_t k = n_leaf + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= pad) return;
    uint64_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0}, h[4];
    sha3_64(m, h);
    store_digest(tree, pad + k, h);
}
__global__ void k_fe_mul(const fe *a, const fe *b, fe *o, int64_t n)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) o[k] = mul(a[k], b[k]);
}

int ilog2(int64_t x) { int l = 0; while (((int64_t)1 << l) < x) ++l; return l; }

void ntt_run(const ntt_in &in, fe *out, int64_t out_stride, int64_t nb, int log_n, int inverse, fe scale, int do_scale,
             const fe *tw, int64_t SS, hipStream_t st)
{
    int log_SS = ilog2(SS);
    int log_m = log_n < NTT_LM ? log_n : NTT_LM;
    int rest = log_n - log_m;
    dim3 grid((unsigned)((int64_t)1 << (log_n - log_m)), (unsigned)nb);
    hipLaunchKernelGGL(k_ntt_lds, grid, dim3(TPB), 0, st, in, out, out_stride, log_n, log_m, inverse, scale,
                       rest == 0 ? do_scale : 0, tw, SS, log_SS);
    int s0 = log_m;
    while (rest > 0) {
        int g = rest > 4 ? 4 : rest;
        rest -= g;
        int64_t threads = nb * ((int64_t)1 << (log_n - g));
        int ds = rest == 0 ? do_scale : 0;
        dim3 gr(nblk(threads)), bk(TPB);
        switch (g) {
        case 1: hipLaunchKernelGGL(k_ntt_regs<1>, gr, bk, 0, st, out, out_stride, nb, log_n, s0, inverse, scale, ds, tw, SS, log_SS); break;
        case 2: hipLaunchKernelGGL(k_ntt_regs<2>, gr, bk, 0, st, out, out_stride, nb, log_n, s0, inverse, scale, ds, tw, SS, log_SS); break;
        case 3: hipLaunchKernelGGL(k_ntt_regs<3>, gr, bk, 0, st, out, out_stride, nb, log_n, s0, inverse, scale, ds, tw, SS, log_SS); break;
        default: hipLaunchKernelGGL(k_ntt_regs<4>, gr, bk, 0, st, out, out_stride, nb, log_n, s0, inverse, scale, ds, tw, SS, log_SS); break;
        }
        s0 += g;
    }
}

}  // namespace

namespace launch {

void zero(void *p, size_t bytes, hipStream_t s)
{
    int64_t n = (int64_t)(bytes / 4);
    if (!n) return;
    unsigned g = nblk(n);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_zero, dim3(g), dim3(TPB), 0, s, (uint32_t *)p, n);
}
void eval_input(const int32_t *u, fe *out, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_eval_input, dim3(nblk(n)), dim3(TPB), 0, s, u, out, n);
}
void eval_layer(gates_dev g, const fe *prev, fe *out, hipStream_t s)
{
    int n = 1 << g.bl;
    hipLaunchKernelGGL(k_eval_layer, dim3(nblk(n)), dim3(TPB), 0, s, g.ty, g.u, g.v, prev, out, n);
}
void mle_eval(const fe *in, int log_n, const fe *r, fe *t0, fe *t1, fe *dst, hipStream_t s)
{
    int64_t n = (int64_t)1 << log_n;
    int k = 0;
    const fe *src = in;
    fe *bufs[2] = {t0, t1};
    int cur = 0;
    while (n > FOLD_TAIL) {
        int64_t no = n / 2;
        unsigned g = nblk(no);
        if (g > 4096) g = 4096;
        hipLaunchKernelGGL(k_mle_fold, dim3(g), dim3(TPB), 0, s, src, bufs[cur], no, r + k);
        src = bufs[cur];
        cur ^= 1;
        n = no;
        ++k;
    }
    // tail: copy into the ping-pong buffers if still reading the input
    if (src == in) {
        (void)hipMemcpyAsync(t0, in, sizeof(fe) * n, hipMemcpyDeviceToDevice, s);
        src = t0; cur = 1;
    }
    hipLaunchKernelGGL(k_mle_tail, dim3(1), dim3(1024), 0, s, (fe *)src, bufs[cur], n, r + k, log_n - k, dst);
}
void half_tables(fe *f, fe *s, const fe *f0, const fe *r, int len, hipStream_t st)
{
    int fh = len >> 1, sh = len - fh;
    int n = (1 << fh) + (1 << sh);
    hipLaunchKernelGGL(k_half_tables, dim3(nblk(n)), dim3(TPB), 0, st, f, s, f0, r, fh, sh);
}
void eq_table(fe *out, const fe *r, int len, fe *scratch, hipStream_t st)
{
    int fh = len >> 1;
    fe *lo = scratch, *hi = scratch + (1 << fh);
    half_tables(lo, hi, nullptr, r, len, st);
    int64_t n = (int64_t)1 << len;
    hipLaunchKernelGGL(k_eq_expand, dim3(nblk(n)), dim3(TPB), 0, st, out, lo, hi, fh, n);
}
void eq_expand(fe *out, const fe *lo, const fe *hi, int fh, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_eq_expand, dim3(nblk(n)), dim3(TPB), 0, st, out, lo, hi, fh, n);
}
void p1_scatter(gates_dev g, const fe *prevV, beta_g bg, acc4 *amb, acc4 *aVb, hipStream_t s)
{
    int n = 1 << g.bl;
    bg_t b = {bg.g0f, bg.g0s, bg.g1f, bg.g1s, bg.fh};
    hipLaunchKernelGGL(k_p1_scatter, dim3(nblk(n)), dim3(TPB), 0, s, g.ty, g.u, g.v, n, prevV, b, amb, aVb);
}
void p2_scatter(gates_dev g, beta_g bg, beta_u bu, const fe *vu, acc4 *amb, acc4 *aVb, hipStream_t s)
{
    int n = 1 << g.bl;
    bg_t b = {bg.g0f, bg.g0s, bg.g1f, bg.g1s, bg.fh};
    bu_t u = {bu.f, bu.s, bu.fh};
    hipLaunchKernelGGL(k_p2_scatter, dim3(nblk(n)), dim3(TPB), 0, s, g.ty, g.u, g.v, n, b, u, vu, amb, aVb);
}
void sumcheck_phase(const fe *Vb, const acc4 *aVb, const acc4 *amb, int L, const fe *r, fe vconst, fold_ws ws, fe *sums,
                    fe *end, hipStream_t s)
{
    int j = 0;
    int cur = 0;
    const fe *prev_part = nullptr;
    int nprev = 0;
    fe *part[2] = {ws.part, ws.part + 3 * MAX_PARTS};
    while (j < L && ((int64_t)1 << (L - 1 - j)) > FOLD_TAIL) {
        int64_t n_out = (int64_t)1 << (L - 1 - j);
        unsigned g = nblk(n_out);
        if (g > MAX_PARTS) g = MAX_PARTS;
        ctabs in = {ws.V[cur ^ 1], ws.aV[cur ^ 1], ws.am[cur ^ 1]};
        tabs out = {ws.V[cur], ws.aV[cur], ws.am[cur]};
        hipLaunchKernelGGL(k_fold, dim3(g), dim3(TPB), 0, s, j == 0 ? 1 : 0, Vb, aVb, amb, in, j == 0 ? r : r + j - 1, out,
                           n_out, part[j & 1], prev_part, nprev, j > 0 ? sums + 3 * (j - 1) : sums);
        prev_part = part[j & 1];
        nprev = (int)g;
        cur ^= 1;
        ++j;
    }
    tabs X = {ws.V[cur ^ 1], ws.aV[cur ^ 1], ws.am[cur ^ 1]};
    tabs Y = {ws.V[cur], ws.aV[cur], ws.am[cur]};
    hipLaunchKernelGGL(k_fold_tail, dim3(1), dim3(TAIL_T), 0, s, j, L, Vb, aVb, amb, X, Y, r, vconst, prev_part, nprev, sums, end);
}
void p1_gather(const uint32_t *off, const int32_t *ent, gates_dev g, int nt, const fe *prevV, beta_g bg, fe *amb, fe *aVb,
               hipStream_t s)
{
    bg_t b = {bg.g0f, bg.g0s, bg.g1f, bg.g1s, bg.fh};
    hipLaunchKernelGGL(k_p1_gather, dim3(nblk(nt)), dim3(TPB), 0, s, off, ent, g.ty, g.u, g.v, nt, prevV, b, amb, aVb);
}
void p2_gather(const uint32_t *off, const int32_t *ent, const int32_t *zlist, int nz, gates_dev g, int nt, beta_g bg,
               beta_u bu, const fe *vu, fe *amb, fe *aVb, fe *part, hipStream_t s)
{
    bg_t b = {bg.g0f, bg.g0s, bg.g1f, bg.g1s, bg.fh};
    bu_t u = {bu.f, bu.s, bu.fh};
    hipLaunchKernelGGL(k_p2_gather, dim3(nblk(nt)), dim3(TPB), 0, s, off, ent, g.ty, g.u, nt, b, u, vu, amb, aVb);
    if (nz > 0) {
        unsigned blocks = nblk(nz);
        if (blocks > MAX_PARTS) blocks = MAX_PARTS;
        hipLaunchKernelGGL(k_p2_zero, dim3(blocks), dim3(TPB), 0, s, zlist, nz, g.ty, g.u, g.v, b, u, vu, part);
        hipLaunchKernelGGL(k_add_parts, dim3(1), dim3(1024), 0, s, part, (int)blocks, aVb);
    }
}
int fold_plan(int L, int *js, int *ks)
{
    int n = 0, j = 0;
    while ((1 << (L - j)) > TAIL_IN) {
        int K = L - j - 9;
        if (K > 6) K = 6;
        js[n] = j; ks[n] = K; ++n;
        j += K;
    }
    return n;
}
int sumcheck_phase2(const fe *Vb, const fe *aVb, const fe *amb, int L, const fe *r, fe vconst, fold_ws ws, fe *slots,
                    fe *sums, fe *end, hipStream_t s)
{
    int js[32], ks[32];
    int nl = fold_plan(L, js, ks);
    int cur = 0, rounds = 0;
    for (int k = 0; k < nl; ++k) {
        int j = js[k], K = ks[k];
        int64_t n0 = (int64_t)1 << (L - j - 1);
        int64_t g64 = n0 / TPB;
        unsigned g = (unsigned)(g64 > FW_BLOCKS ? FW_BLOCKS : g64);
        ctabs in = {ws.V[cur ^ 1], ws.aV[cur ^ 1], ws.am[cur ^ 1]};
        tabs out = {ws.V[cur], ws.aV[cur], ws.am[cur]};
        fe *pp = slots + (int64_t)rounds * 3 * FW_BLOCKS;
        switch (K) {
#define VG_FW(k) case k: hipLaunchKernelGGL(k_fold_wave<k>, dim3(g), dim3(TPB), 0, s, j, Vb, aVb, amb, in, r, out, n0, pp); break;
        VG_FW(1) VG_FW(2) VG_FW(3) VG_FW(4) VG_FW(5) VG_FW(6)
#undef VG_FW
        }
        cur ^= 1;
        rounds += K;
    }
    int j = nl ? js[nl - 1] + ks[nl - 1] : 0;
    ctabs X = {ws.V[cur ^ 1], ws.aV[cur ^ 1], ws.am[cur ^ 1]};
    hipLaunchKernelGGL(k_fold_tail_lds, dim3(1), dim3(256), 0, s, j, L, Vb, aVb, amb, X, r, vconst, sums, end);
    return rounds;
}
int fold_blocks(int L, int j)
{
    int64_t g = ((int64_t)1 << (L - j - 1)) / TPB;
    return (int)(g > FW_BLOCKS ? FW_BLOCKS : g);
}
void reduce_slots(const fe *slots, int nslots, const int *nb, const int64_t *dst, fe *msg, hipStream_t s)
{
    if (nslots) hipLaunchKernelGGL(k_reduce_slots, dim3(nslots), dim3(256), 0, s, slots, nb, dst, msg);
}
void half_multi(const half_desc *d, int nd, int64_t total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(k_half_multi, dim3(nblk(total)), dim3(TPB), 0, s, d, nd, total);
}
void general_value(gates_dev g, beta_g bg, beta_u bu, beta_u bv, const fe *vu, const fe *vv, fe *part, unsigned *counter,
                   fe *dst, hipStream_t s)
{
    (void)counter;
    int n = 1 << g.bl;
    unsigned blocks = nblk(n);
    if (blocks > MAX_PARTS) blocks = MAX_PARTS;
    bg_t b = {bg.g0f, bg.g0s, bg.g1f, bg.g1s, bg.fh};
    bu_t u = {bu.f, bu.s, bu.fh}, v = {bv.f, bv.s, bv.fh};
    hipLaunchKernelGGL(k_general, dim3(blocks), dim3(TPB), 0, s, g.ty, g.u, g.v, n, b, u, v, vu, vv, part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, part, (int)blocks, dst);
}
void twiddles(fe *tw, int64_t n, const fe *T1, const fe *T2, hipStream_t s)
{
    hipLaunchKernelGGL(k_twiddles, dim3(nblk(n)), dim3(TPB), 0, s, tw, n, T1, T2);
}
void intt(intt_src in, fe *out, int64_t out_stride, int nb, int log_n, fe inv_n, const fe *tw, int64_t SS, hipStream_t s)
{
    ntt_in ni;
    ni.kind = in.kind; ni.src = in.src; ni.stride = in.src_stride;
    ni.l = in.l; ni.q = in.q; ni.cw_S = in.cw_S; ni.cw_stride = in.cw_stride; ni.step = in.step; ni.log_R = in.log_R;
    ni.lens = nullptr; ni.log_Rl = 0; ni.hmode = 0;
    ntt_run(ni, out, out_stride, nb, log_n, 1, inv_n, 1, tw, SS, s);
}
void lde(const fe *coef, int64_t coef_stride, const int *lens, int nb, fe *out, int log_C, int log_R, const fe *tw,
         int64_t SS, hipStream_t s, int hmode)
{
    ntt_in ni;
    memset(&ni, 0, sizeof ni);
    ni.kind = 2; ni.src = coef; ni.stride = coef_stride; ni.lens = lens; ni.log_Rl = log_R; ni.hmode = hmode;
    ntt_run(ni, out, (int64_t)1 << log_C, (int64_t)nb << log_R, log_C, 0, fone(), 0, tw, SS, s);
}
void virtual_oracle(const fe *l, const fe *q, const fe *h, const fe *c0, int nslice, int64_t SS, int log_R, int64_t real,
                    int64_t m, const fe *tw, fe *vo, hipStream_t s)
{
    int64_t n = (int64_t)nslice * SS;
    hipLaunchKernelGGL(k_vo, dim3(nblk(n)), dim3(TPB), 0, s, l, q, h, c0, nslice, SS, log_R, real, m, tw, vo);
}
void slice_sums(const fe *lq_coef, int64_t stride, int nslice, int64_t real, int64_t m, fe *c0, fe *all_sum, hipStream_t s)
{
    hipLaunchKernelGGL(k_slice_sums, dim3(1), dim3(128), 0, s, lq_coef, stride, nslice, real, m, c0, all_sum);
}
void merkle(uint8_t *tree, int64_t n_leaf, hipStream_t s)
{
    int64_t n = n_leaf / 2;
    while (n > 1024) {
        hipLaunchKernelGGL(k_merkle_level, dim3(nblk(n)), dim3(TPB), 0, s, tree, n, n);
        n /= 2;
    }
    if (n >= 1) hipLaunchKernelGGL(k_merkle_tail, dim3(1), dim3(1024), 0, s, tree, n);
}
void ntt(intt_src in, fe *out, int64_t out_stride, int nb, int log_n, int inverse, fe scale, int do_scale, const fe *tw,
         int64_t SS, hipStream_t s)
{
    ntt_in ni;
    memset(&ni, 0, sizeof ni);
    ni.kind = in.kind; ni.src = in.src; ni.stride = in.src_stride;
    ni.l = in.l; ni.q = in.q; ni.cw_S = in.cw_S; ni.cw_stride = in.cw_stride; ni.step = in.step; ni.log_R = in.log_R;
    ntt_run(ni, out, out_stride, nb, log_n, inverse, scale, do_scale, tw, SS, s);
}
void sha3_batch(const uint8_t *in, uint8_t *out, int64_t n, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_sha3_batch, dim3(nblk(n)), dim3(TPB), 0, s, (const uint64_t *)in, (uint64_t *)out, n);
}
void merkle_pad(uint8_t *tree, int64_t pad, int64_t n_leaf, hipStream_t s)
{
    if (pad > n_leaf) hipLaunchKernelGGL(k_merkle_pad, dim3(nblk(pad - n_leaf)), dim3(TPB), 0, s, tree, pad, n_leaf);
}
void fe_mul(const fe *a, const fe *b, fe *o, int64_t n, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_fe_mul, dim3(nblk(n)), dim3(TPB), 0, s, a, b, o, n);
}

void fri_fold(const fe *prev, int64_t S, int log_R, const fe *r, const fe *tw, int64_t SS, fe *next, hipStream_t s)
{
    hipLaunchKernelGGL(k_fri_fold, dim3(nblk(65 * (S / 2))), dim3(TPB), 0, s, prev, S, log_R, r, tw, SS, next);
}
void leaf_hash_multi(int nseg, const fe *const *cw, const int64_t *S, uint8_t *const *tree, int log_R, hipStream_t s)
{
    leaf_segs L;
    memset(&L, 0, sizeof L);
    L.n = nseg; L.log_R = log_R;
    L.start[0] = 0;
    for (int k = 0; k < nseg; ++k) {
        L.cw[k] = cw[k]; L.S[k] = S[k]; L.tree[k] = tree[k];
        L.start[k + 1] = L.start[k] + S[k] / 2;
    }
    hipLaunchKernelGGL(k_leaf_multi, dim3(nblk(L.start[nseg])), dim3(TPB), 0, s, L);
}
void merkle_multi(int ntree, uint8_t *const *tree, const int64_t *nleaf, hipStream_t s)
{
    tree_set T;
    memset(&T, 0, sizeof T);
    T.n = ntree;
    int64_t mx = 0;
    for (int k = 0; k < ntree; ++k) { T.tree[k] = tree[k]; T.nleaf[k] = nleaf[k]; if (nleaf[k] > mx) mx = nleaf[k]; }
    for (int h = 1; ((int64_t)1 << h) <= mx; ++h) {
        int64_t total = 0;
        for (int k = 0; k < ntree; ++k) total += nleaf[k] >> h;
        hipLaunchKernelGGL(k_merkle_multi, dim3(nblk(total)), dim3(TPB), 0, s, T, h, total);
    }
}
void gather16(const uint64_t *src, int64_t n, fe *out, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_gather16, dim3(nblk(n)), dim3(TPB), 0, s, src, n, out);
}

}  // namespace launch
}  // namespace vg
