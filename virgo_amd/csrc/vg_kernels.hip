// CDNA4 (gfx950) kernels of the Virgo prover hot path.  All arithmetic is
// exact F_{p^2} integer work on the VALU (no MFMA: there is no dense float
// contraction here).  Reductions and atomics may run in any order because
// every field operation returns the canonical representative.
#include "vg_kernels.hpp"
#include "vg_keccak.hpp"
#include "vg_sha256.hpp"
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace vg {
namespace {

constexpr int TPB = 256;
constexpr int NTT_LM = 11;          // LDS-resident NTT block: 2^11 elements = 32 KB
constexpr int FOLD_TAIL = 512;      // MLE rounds with <= this many outputs run in one block

__host__ __device__ inline unsigned nblk(int64_t n, int t = TPB) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ uint64_t red64(uint64_t x)       // x mod p for any 64-bit x
{
    uint64_t r = (x & P) + (x >> 61);
    return r >= P ? r - P : r;
}


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
    if (threadIdx.x < 64) {
        for (int k = 0; k < K; ++k) {
            fe s = lane < nw ? red[k][lane] : fzero();
            s = wave_sum(s);
            if (threadIdx.x == 0) v[k] = s;
        }
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
// Both operands are gathered before the gate type is looked at: a wave holds
// add and mult gates at random, and loads inside the type's branch would be
// issued once per branch the wave takes, each batch waiting for its own
// round trip.  Every index read here is in range (the loader checks u and v
// of every gate but the input type 3, whose u is a value, and a sum's v may
// equal the layer size, so neither of those is read as an operand).
__global__ void k_eval_layer(const uint8_t *ty, const int32_t *gu, const int32_t *gv, const fe *p, fe *o, int n)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int t = ty[g], u = gu[g], v = gv[g];
    fe x = fzero(), y = fzero();
    if (t != 3) x = p[u];
    if (t != 3 && t != 5) y = p[v];
    fe r;
    switch (t) {
    case 0: r = add(x, y); break;
    case 1: r = mul(x, y); break;
    case 3: r = mk(red64((uint64_t)(int64_t)u), 0); break;
    case 4: case 10: r = x; break;
    case 5: { r = x; for (int k = u + 1; k < v; ++k) r = add(r, p[k]); break; }
    case 6: r = sub(fone(), x); break;
    case 7: r = sub(x, y); break;
    case 8: r = sub(add(x, y), mul(mul(mk(2, 0), x), y)); break;
    case 9: r = sub(y, mul(x, y)); break;
    case 12: { r = fzero(); for (int k = u; k <= v; ++k) r = add(r, mul_base(p[k], 1ull << ((k - u) % 61))); break; }
    case 13: r = mul(x, sub(fone(), y)); break;
    default: r = fzero(); break;
    }
    o[g] = r;
}

// ---------------------------------------------------------------------
// multilinear evaluation by halving (V_res, zk_prover.cpp:30-52)
// ---------------------------------------------------------------------
// h rounds of an MLE evaluation in one block of 2^(h-1) threads: block b
// folds in[b 2^h, (b+1) 2^h) to out[b], round k taking each pair (x0, x1)
// to x0 (1 - r[k]) + x1 r[k] = x0 + r[k] (x1 - x0); the first round in
// registers, the rest in LDS
__global__ __launch_bounds__(1024) void k_mle_block(const fe *in, fe *out, const fe *r, int h)
{
    __shared__ fe buf[1024];
    const int t = threadIdx.x;
    const fe *src = in + ((int64_t)blockIdx.x << h);
    int no = 1 << (h - 1);
    fe v = fzero();
    if (t < no) { fe a = src[2 * t], b = src[2 * t + 1]; v = add(a, mul(sub(b, a), r[0])); }
    for (int k = 1; k < h; ++k) {
        if (t < no) buf[t] = v;
        __syncthreads();
        no >>= 1;
        if (t < no) { fe a = buf[2 * t], b = buf[2 * t + 1]; v = add(a, mul(sub(b, a), r[k])); }
        __syncthreads();
    }
    if (t == 0) out[blockIdx.x] = v;
}
// the last rounds of an MLE evaluation (n <= 2 FOLD_TAIL entries) in LDS
__global__ __launch_bounds__(FOLD_TAIL) void k_mle_tail(const fe *in, int n, const fe *r, int nr, fe *dst)
{
    __shared__ fe buf[2 * FOLD_TAIL];
    int t = threadIdx.x;
    for (int i = t; i < n; i += blockDim.x) buf[i] = in[i];
    __syncthreads();
    for (int k = 0; k < nr; ++k) {
        fe rr = r[k], om = sub(fone(), rr);
        int no = n >> 1;
        fe v = fzero();
        if (t < no) v = add(mul(buf[2 * t], om), mul(buf[2 * t + 1], rr));
        __syncthreads();
        if (t < no) buf[t] = v;
        __syncthreads();
        n = no;
    }
    if (t == 0) *dst = buf[0];
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
    return dot2(b.g0f[lo], b.g0s[hi], b.g1f[lo], b.g1s[hi]);
}
__device__ __forceinline__ fe beta_u(const bu_t &b, int u)
{
    return mul(b.f[u & ((1 << b.fh) - 1)], b.s[u >> b.fh]);
}

// ---------------------------------------------------------------------
// bookkeeping fold (zk_prover.cpp:600-633, 992-1025)
// ---------------------------------------------------------------------
struct tabs { lpoly *V, *aV, *am; };
struct ctabs { const lpoly *V, *aV, *am; };

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

// ---------------------------------------------------------------------
// Table builds by gather over CSR transposes (owner computes, no atomics).
// Phase 1 (zk_prover.cpp:375-592): table entry t sums the contributions of
// the gates that reach it through u (a sum gate reaches every t of its
// range; its record carries t in place of v).  Phase 2 (:789-983): entry t
// sums the gates whose v is t; the gates that reach entry 0 unconditionally
// (sum, exp-sum, NOT, relay) are reduced separately into part[MAX_PARTS].
// A record's contribution (am, aV) depends on the record alone; add and
// mult, the common types, take one branch-free path.
// ---------------------------------------------------------------------
// types whose phase-1 contribution reads V(v)
__device__ __forceinline__ bool p1_reads_v(int ty) { return ty <= 1 || ty == 7 || ty == 8 || ty == 9 || ty == 13; }
// pv = V(R.v) for the types that read it (p1_reads_v), anything otherwise
__device__ __forceinline__ void p1_rec(const launch::gate_rec &R, const bg_t &bg, const fe &pv, fe &am, fe &aV)
{
    int ty = R.ty;
    if (ty <= 1) {
        fe tmp = beta_g(bg, R.g), x = mul(pv, tmp);
        am = ty == 0 ? tmp : x;
        aV = ty == 0 ? x : fzero();
        return;
    }
    am = fzero(); aV = fzero();
    if (ty == 4) { am = beta_g(bg, R.u); return; }
    fe tmp = beta_g(bg, R.g);
    switch (ty) {
    case 5: case 10: am = tmp; break;
    case 12: am = mul_base(tmp, 1ull << ((R.v - R.u) % 61)); break;   // R.v = t; 2^61 == 1
    case 6: am = neg(tmp); aV = tmp; break;
    case 7: aV = neg(mul(pv, tmp)); am = tmp; break;
    case 8: { fe tv = mul(tmp, pv); aV = tv; am = sub(tmp, dbl(tv)); break; }
    case 13: { fe tv = mul(tmp, pv); am = sub(tmp, tv); break; }
    case 9: { fe tv = mul(tmp, pv); aV = tv; am = neg(tv); break; }
    default: break;
    }
}
__device__ __forceinline__ void p2_rec(const launch::gate_rec &R, const bg_t &bg, const bu_t &bu, const fe &vu, fe &am,
                                       fe &aV)
{
    int ty = R.ty;
    fe tgu = mul(beta_g(bg, R.g), beta_u(bu, R.u)), x = mul(tgu, vu);
    if (ty <= 1) {
        am = ty == 0 ? tgu : x;
        aV = ty == 0 ? x : fzero();
        return;
    }
    am = fzero(); aV = fzero();
    switch (ty) {
    case 7: am = neg(tgu); aV = x; break;
    case 8: am = sub(tgu, dbl(x)); aV = x; break;
    case 13: am = neg(x); aV = x; break;
    case 9: am = sub(tgu, x); break;
    default: break;
    }
}
// 2^kl lanes per entry (kl <= 6, the lanes of an entry inside one wave):
// lane `sub` takes records sub, sub + 2^kl, ... of its entry and the lanes
// add their sums with butterfly shuffles; crowded tables (many records per
// entry, e.g. the inputs of a matrix product) get many lanes per entry
template <int C>
__device__ __forceinline__ void lanes_sum(fe *x, int kl)
{
    for (int o = 1; o < (1 << kl); o <<= 1)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            fe y;
            y.real = __shfl_xor((unsigned long long)x[c].real, o);
            y.img = __shfl_xor((unsigned long long)x[c].img, o);
            x[c] = add(x[c], y);
        }
}
__global__ __launch_bounds__(TPB) void k_p1_gather(const launch::layer_dev *D, int kl)
{
    const launch::layer_dev &d = D[blockIdx.y];
    int64_t gt = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t t = gt >> kl;
    int sub = (int)(gt & ((1 << kl) - 1));
    bool live = t < d.nt;
    bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
    fe acc[2] = {fzero(), fzero()};     // am, aV
    if (live)
        for (uint32_t e = d.off1[t] + sub; e < d.off1[t + 1]; e += 1u << kl) {
            launch::gate_rec R = d.ent1[e];
            fe m, a;
            p1_rec(R, bg, d.pv[p1_reads_v(R.ty) ? R.v : 0], m, a);
            acc[0] = add(acc[0], m); acc[1] = add(acc[1], a);
        }
    lanes_sum<2>(acc, kl);
    if (live && sub == 0) {
        d.amb[t] = acc[0];
        d.aVb[t] = acc[1];
    }
}
__global__ __launch_bounds__(TPB) void k_p2_gather(const launch::layer_dev *D, int kl)
{
    const launch::layer_dev &d = D[blockIdx.y];
    int64_t gt = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t t = gt >> kl;
    int sub = (int)(gt & ((1 << kl) - 1));
    bool live = t < d.nt;
    bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
    bu_t bu = {d.bu.f, d.bu.s, d.bu.fh};
    fe vu = *d.vu;
    fe acc[2] = {fzero(), fzero()};     // am, aV
    if (live && t == 0 && sub == 0 && d.nz > 0) acc[1] = d.part[launch::MAX_PARTS];
    if (live)
        for (uint32_t e = d.off2[t] + sub; e < d.off2[t + 1]; e += 1u << kl) {
            fe m, a;
            p2_rec(d.ent2[e], bg, bu, vu, m, a);
            acc[0] = add(acc[0], m); acc[1] = add(acc[1], a);
        }
    lanes_sum<2>(acc, kl);
    if (live && sub == 0) {
        d.amb[t] = acc[0];
        d.aVb[t] = acc[1];
    }
}
// gates adding into addV[0] in phase 2 (sum, exp-sum, NOT, relay); per-block partials
__global__ __launch_bounds__(TPB) void k_p2_zero(const launch::layer_dev *D)
{
    const launch::layer_dev &d = D[blockIdx.y];
    bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
    bu_t bu = {d.bu.f, d.bu.s, d.bu.fh};
    fe acc = fzero();
    fe vu = *d.vu;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < d.nz; k += gridDim.x * blockDim.x) {
        int g = d.z2[k];
        int ty = d.ty[g], u = d.u[g], v = d.v[g];
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
    if (threadIdx.x == 0) d.part[blockIdx.x] = acc;
}
__global__ void k_zero_sum(const launch::layer_dev *D, int n)
{
    const launch::layer_dev &d = D[blockIdx.y];
    fe acc = fzero();
    for (int k = threadIdx.x; k < n; k += blockDim.x) acc = add(acc, d.part[k]);
    block_sum<1>(&acc);
    if (threadIdx.x == 0) d.part[launch::MAX_PARTS] = acc;
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
__device__ __forceinline__ lpoly fold_pair(const lpoly &y0, const lpoly &y1, const fe &r)
{
    lpoly z;
    z.b = add(mul(y0.a, r), y0.b);
    z.a = sub(add(mul(y1.a, r), y1.b), z.b);
    return z;
}
// round contributions a = am.a V.a, b = am.a V.b + am.b V.a + aV.a, c = am.b V.b + aV.b
__device__ __forceinline__ void fold_terms(const lpoly &V, const lpoly &aV, const lpoly &am, fe *t)
{
    t[0] = mul(am.a, V.a);
    t[1] = add(add(mul(am.a, V.b), mul(am.b, V.a)), aV.a);
    t[2] = add(mul(am.b, V.b), aV.b);
}
// ---------------------------------------------------------------------
// Sumcheck fold (zk_prover.cpp:600-633, 992-1025).  A phase is a sequence of
// rounds on three value tables V, addV, amsum: round j pairs (x0, x1) =
// (T[2i], T[2i+1]) of each table into the linear polynomial x0 + X (x1 - x0),
// adds the pair's terms to the round polynomial, and the next round's table
// is that polynomial at r[j].  Per pair, with d = x1 - x0:
//   a += d_am d_V,  c += am0 V0 + aV0,  p(1) += am1 V1 + aV1,  b = p(1) - a - c
// (three products; the reference's (a, b, c) follow exactly).
//
// Streaming launches (k_fold_reg) run up to 3 rounds: each wave folds 64 2^b
// consecutive values of every table down to 64, loading them with wave-wide
// contiguous 1 KB loads and pairing values across lanes (fold_tree), keeping
// its round sums in registers; the block reduces them once at the end into
// per-block partials.  Rounds j >= 1 sum a and p(0) only; k_derive_b sets
// their b after the reduction.  When at most 64 values remain, one wave per instance
// (k_fold_tail) runs the last rounds in LDS and writes the round sums and the
// final pair directly.
// ---------------------------------------------------------------------
using launch::FB_B;
using launch::FB_TAIL;
constexpr int FB_MAXB = launch::FW_BLOCKS;   // block partials per slot round
struct ftab { const fe *V, *aV, *am; };
struct fsum { fe a, c, p1; };
__device__ __forceinline__ void fsum_add(fsum &s, const fe &V0, const fe &V1, const fe &aV0, const fe &aV1,
                                         const fe &am0, const fe &am1)
{
    s.a = add(s.a, mul(sub(am1, am0), sub(V1, V0)));
    s.c = add(s.c, add(mul(am0, V0), aV0));
    s.p1 = add(s.p1, add(mul(am1, V1), aV1));
}
// a and p(0) only.  For every round j >= 1 of a phase the tables are those of
// round j-1 folded at r[j-1], so p_j(0) + p_j(1) = p_{j-1}(r[j-1]) exactly;
// k_derive_b sets b = p_{j-1}(r[j-1]) - a - 2 c once the sums are reduced,
// and the fold skips the product of p(1).
__device__ __forceinline__ void fsum_add_ac(fsum &s, const fe &V0, const fe &V1, const fe &aV0, const fe &am0,
                                            const fe &am1)
{
    s.a = add(s.a, mul(sub(am1, am0), sub(V1, V0)));
    s.c = add(s.c, add(mul(am0, V0), aV0));
}
__device__ __forceinline__ fe lerp(const fe &x0, const fe &x1, const fe &r) { return add(x0, mul(sub(x1, x0), r)); }
__device__ __forceinline__ ftab fold_input(const launch::fold_inst &t, int j0, int cur)
{
    if (j0 == 0) return {t.Vb, t.aVb, t.amb};
    return {t.V[cur ^ 1], t.aV[cur ^ 1], t.am[cur ^ 1]};
}
// B rounds (the first of a phase and its successors) on the 2^B consecutive
// values one lane holds of each table; ld(k, V, aV, am) supplies value k of
// the three tables (called in order, one pair at a time, so at most half of
// them are live at once).  Round q's sums go into acc[q] (p(1) for round 0
// only: later rounds' b comes from k_derive_b); the folded values into oV,
// oaV, oam.
template <int B, class Load>
__device__ __forceinline__ void fold_lane(Load ld, const fe *rq, fsum *acc, fe &oV, fe &oaV, fe &oam)
{
    constexpr int H = 1 << (B - 1);
    fe V[H], aV[H], am[H];
#pragma unroll
    for (int p2 = 0; p2 < H; ++p2) {
        fe V0, V1, aV0, aV1, am0, am1;
        ld(2 * p2, V0, aV0, am0);
        ld(2 * p2 + 1, V1, aV1, am1);
        fsum_add(acc[0], V0, V1, aV0, aV1, am0, am1);
        V[p2] = lerp(V0, V1, rq[0]); aV[p2] = lerp(aV0, aV1, rq[0]); am[p2] = lerp(am0, am1, rq[0]);
    }
#pragma unroll
    for (int q = 1; q < B; ++q) {
#pragma unroll
        for (int p2 = 0; p2 < (H >> q); ++p2) {
            fsum_add_ac(acc[q], V[2 * p2], V[2 * p2 + 1], aV[2 * p2], am[2 * p2], am[2 * p2 + 1]);
            V[p2] = lerp(V[2 * p2], V[2 * p2 + 1], rq[q]);
            aV[p2] = lerp(aV[2 * p2], aV[2 * p2 + 1], rq[q]);
            am[p2] = lerp(am[2 * p2], am[2 * p2 + 1], rq[q]);
        }
    }
    oV = V[0]; oaV = aV[0]; oam = am[0];
}
// per-block partials (a, b, c) of B rounds, b = p(1) - a - c; without P1
// (a and p(0) summed only) b = 0, set after the reduction by k_derive_b
template <int B, bool P1 = true>
__device__ __forceinline__ void fold_partials(fsum *acc, const launch::fold_inst &t, int slot_round)
{
#pragma unroll
    for (int q = 0; q < B; ++q) {
        fe v3[3] = {acc[q].a, acc[q].c, P1 ? acc[q].p1 : fzero()};
        block_sum<P1 ? 3 : 2>(v3);
        if (threadIdx.x == 0) {
            fe o3[3] = {v3[0], P1 ? sub(sub(v3[2], v3[0]), v3[1]) : fzero(), v3[1]};
            for (int c = 0; c < 3; ++c) t.slots[((int64_t)(slot_round + q) * 3 + c) * FB_MAXB + blockIdx.x] = o3[c];
        }
    }
}
// lane l's copy of lane (l ^ m)'s v, m a power of two below 64 (a constant
// once the callers' loops are unrolled): DPP quad permutes for 1 and 2,
// ds_swizzle's xor mode (within 32 lanes) up to 16, ds_bpermute for 32.
// Every lane of the wave must be active.
__device__ __forceinline__ uint32_t lane_xor32(uint32_t v, int m)
{
    switch (m) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (4 << 10) | 0x1F);
    case 8: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (8 << 10) | 0x1F);
    case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (16 << 10) | 0x1F);
    default: return (uint32_t)__shfl_xor((int)v, m);
    }
}
__device__ __forceinline__ uint64_t lane_xor64(uint64_t v, int m)
{
    return (uint64_t)lane_xor32((uint32_t)v, m) | (uint64_t)lane_xor32((uint32_t)(v >> 32), m) << 32;
}
__device__ __forceinline__ fe lane_xor(const fe &x, int m)
{
    fe y;
    y.real = lane_xor64(x.real, m);
    y.img = lane_xor64(x.img, m);
    return y;
}
// B rounds on the 64 2^B consecutive values E[0 .. 64 2^B) of each table one
// wave holds, lane l holding E[64k + l] as value k (so each wave-wide load
// reads 1 KB contiguous).  Before round q, lane l holds F_q[64k + rot_q(l)]
// as value k < 2^(B-q), F_q being the tables after q rounds and rot_q
// rotating the 6 lane bits right by q; so the pair (F_q[2i], F_q[2i+1]) lies
// in lanes l and l ^ 2^q as the same value k.  The lane with bit q clear
// folds the pairs of even k, its partner those of odd k, each receiving the
// other's element; both keep the result as value k / 2.  Afterwards lane l
// holds F_B[rot_B(l)] as value 0.  Round q's sums go into acc[q].
//
// The values are folded as a post-order walk: values [k0, k0 + 2^Q) become
// value k0 / 2^Q of round Q once both halves are folded, so they are loaded
// two at a time and at most Q + 2 of each table are live (4 waves per SIMD
// instead of 2 with all 2^B loaded first; the round sums are field sums, so
// their order is free).  P1: round 0 (q = 0 of a launch at j0 = 0) sums
// p(1) too; every other round a and p(0) only.
template <int Q, bool P1>
__device__ __forceinline__ void fold_tree(const ftab &X, int64_t o, int k0, const fe *rq, fsum *acc, int lane, fe &V,
                                          fe &aV, fe &am)
{
    if constexpr (Q == 0) {
        V = X.V[o + 64 * k0]; aV = X.aV[o + 64 * k0]; am = X.am[o + 64 * k0];
    } else {
        constexpr int q = Q - 1, m = 1 << q;
        fe Va, aVa, ama, Vb, aVb, amb;
        fold_tree<Q - 1, P1>(X, o, k0, rq, acc, lane, Va, aVa, ama);
        fold_tree<Q - 1, P1>(X, o, k0 + m, rq, acc, lane, Vb, aVb, amb);
        const bool hi = (lane >> q) & 1;
        fe rV = lane_xor(hi ? Va : Vb, m), raV = lane_xor(hi ? aVa : aVb, m), ram = lane_xor(hi ? ama : amb, m);
        fe V0 = hi ? rV : Va, V1 = hi ? Vb : rV;
        fe aV0 = hi ? raV : aVa, aV1 = hi ? aVb : raV;
        fe am0 = hi ? ram : ama, am1 = hi ? amb : ram;
        if constexpr (P1 && q == 0) fsum_add(acc[q], V0, V1, aV0, aV1, am0, am1);
        else fsum_add_ac(acc[q], V0, V1, aV0, am0, am1);
        V = lerp(V0, V1, rq[q]); aV = lerp(aV0, aV1, rq[q]); am = lerp(am0, am1, rq[q]);
    }
}
// rounds j0 .. j0+B-1: wave unit u (64 outputs; nout is a multiple of 64)
// folds values [u 64 2^B, (u+1) 64 2^B) to out[u 64 .. u 64 + 64).  P1 for
// the launch that starts a phase (j0 = 0): its round 0, which has no
// predecessor, sums p(1) too.
template <int B, bool P1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(P1 ? 3 : 4))) void k_fold_reg(const launch::fold_inst *I, int cur, int j0, int64_t nout, int slot_round)
{
    const launch::fold_inst &t = I[blockIdx.y];
    ftab X = fold_input(t, j0, cur);
    fe rq[B];
#pragma unroll
    for (int q = 0; q < B; ++q) rq[q] = t.r[j0 + q];
    fsum acc[B];
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] = {fzero(), fzero(), fzero()};
    const int lane = threadIdx.x & 63;
    const int64_t nu = nout >> 6, ustep = ((int64_t)gridDim.x * blockDim.x) >> 6;
    // u is uniform over the wave, so all its lanes take part in every exchange
    for (int64_t u = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; u < nu; u += ustep) {
        const int64_t o = (u << (B + 6)) + lane;
        fe V, aV, am;
        fold_tree<B, P1>(X, o, 0, rq, acc, lane, V, aV, am);
        const int64_t g = (u << 6) + (((lane >> B) | (lane << (6 - B))) & 63);
        t.V[cur][g] = V; t.aV[cur][g] = aV; t.am[cur][g] = am;
    }
    fold_partials<B, P1>(acc, t, slot_round);
}
// Table build of one phase fused with its first B fold rounds: the full
// tables are never stored.  A block takes tiles of 256 2^B consecutive
// entries (thread: 2^B of them); the tile's records, contiguous in the CSR
// order, are evaluated in chunks of GF_CH, one record per thread (uniform
// work whatever the per-entry counts), into LDS; each thread then sums the
// chunk's records of its own entries and, after the last chunk, folds its
// entries with the V values (the previous layer) as k_fold_reg does.
constexpr int GF_CH = 1152;                  // records per chunk (a tile holds ~1024 at one per entry)
constexpr int GF_NR = (GF_CH + 255) / 256;   // records per thread and chunk
template <int PH, int B>
__global__ __launch_bounds__(256) void k_gather_fold(const launch::layer_dev *D, const launch::fold_inst *I)
{
    constexpr int E = 1 << B;
    const launch::layer_dev &d = D[blockIdx.y];
    const launch::fold_inst &t = I[blockIdx.y];
    __shared__ fe cam[2 * GF_CH];
    __shared__ fe tot[3 * B];             // the block's round sums (a, c, p(1)) over its tiles
    fe *caV = cam + GF_CH;
    const uint32_t *off = PH == 0 ? d.off1 : d.off2;
    const launch::gate_rec *ent = PH == 0 ? d.ent1 : d.ent2;
    bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
    bu_t bu = {d.bu.f, d.bu.s, d.bu.fh};
    fe vu = PH ? *d.vu : fzero();
    if (threadIdx.x < 3 * B) tot[threadIdx.x] = fzero();
    const int64_t nt = d.nt, tile = (int64_t)blockDim.x * E;
    for (int64_t T0 = blockIdx.x * tile; T0 < nt; T0 += (int64_t)gridDim.x * tile) {
        int64_t tl = T0 + (int64_t)threadIdx.x * E;
        bool live = tl < nt;
        uint32_t E0 = off[T0], E1 = off[T0 + tile < nt ? T0 + tile : nt];
        fe am[E], aV[E];
#pragma unroll
        for (int k = 0; k < E; ++k) { am[k] = fzero(); aV[k] = fzero(); }
        for (uint32_t S = E0; S < E1; S += GF_CH) {
            uint32_t n = E1 - S < (uint32_t)GF_CH ? E1 - S : (uint32_t)GF_CH;
            __syncthreads();
            // all loads of the thread's records first, then their contributions
            launch::gate_rec R[GF_NR];
            fe pvv[PH == 0 ? GF_NR : 1];
#pragma unroll
            for (int i = 0; i < GF_NR; ++i) {
                uint32_t q = threadIdx.x + 256 * i;
                R[i] = q < n ? ent[S + q] : launch::gate_rec{0, 0, 0, -1};
            }
            if (PH == 0) {
#pragma unroll
                for (int i = 0; i < GF_NR; ++i) pvv[i] = d.pv[R[i].ty >= 0 && p1_reads_v(R[i].ty) ? R[i].v : 0];
            }
#pragma unroll
            for (int i = 0; i < GF_NR; ++i) {
                uint32_t q = threadIdx.x + 256 * i;
                if (q < n) {
                    fe m, a;
                    if (PH == 0) p1_rec(R[i], bg, pvv[i], m, a);
                    else p2_rec(R[i], bg, bu, vu, m, a);
                    cam[q] = m; caV[q] = a;
                }
            }
            __syncthreads();
            if (live) {
                uint32_t ok = off[tl];
#pragma unroll
                for (int k = 0; k < E; ++k) {
                    uint32_t ok1 = off[tl + k + 1];
                    uint32_t lo = ok > S ? ok : S, hi = ok1 < S + n ? ok1 : S + n;
#pragma unroll 1
                    for (uint32_t j = lo; j < hi; ++j) { am[k] = add(am[k], cam[j - S]); aV[k] = add(aV[k], caV[j - S]); }
                    ok = ok1;
                }
            }
        }
        fsum acc[B];
#pragma unroll
        for (int q = 0; q < B; ++q) acc[q] = {fzero(), fzero(), fzero()};
        if (live) {
            if (PH == 1 && tl == 0 && d.nz > 0) aV[0] = add(aV[0], d.part[launch::MAX_PARTS]);
            fe oV, oaV, oam;
            // the entries' own V values straight from memory (64 contiguous
            // bytes per thread; staging them through LDS cost two barriers)
            const fe *myv = d.pv + tl;
            fold_lane<B>([&](int k, fe &V, fe &a, fe &m) { V = myv[k]; a = aV[k]; m = am[k]; }, t.r, acc, oV, oaV,
                         oam);
            int64_t g = tl >> B;
            t.V[0][g] = oV; t.aV[0][g] = oaV; t.am[0][g] = oam;
        }
        fe v[3 * B];
#pragma unroll
        for (int q = 0; q < B; ++q) { v[3 * q] = acc[q].a; v[3 * q + 1] = acc[q].c; v[3 * q + 2] = acc[q].p1; }
        block_sum<3 * B>(v);
        if (threadIdx.x == 0)
            for (int k = 0; k < 3 * B; ++k) tot[k] = add(tot[k], v[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < B; ++q) {
            fe o3[3] = {tot[3 * q], sub(sub(tot[3 * q + 2], tot[3 * q]), tot[3 * q + 1]), tot[3 * q + 1]};
            for (int c = 0; c < 3; ++c) t.slots[((int64_t)q * 3 + c) * FB_MAXB + blockIdx.x] = o3[c];
        }
}
// the last b <= 6 rounds of each instance, one wave each: round sums to
// sums[3j..], the final pair's V.a, V.b, am.a, am.b and v = V(r[L-1]) + vconst to end[]
__global__ __launch_bounds__(64) void k_fold_tail(const launch::fold_inst *I, int cur, int j0, int b)
{
    const launch::fold_inst &t = I[blockIdx.y];
    ftab X = fold_input(t, j0, cur);
    __shared__ fe sV[64], saV[64], sam[64];
    __shared__ fsum ls[32];
    int lane = threadIdx.x, n = 1 << b;
    if (lane < n) { sV[lane] = X.V[lane]; saV[lane] = X.aV[lane]; sam[lane] = X.am[lane]; }
    __syncthreads();
    for (int q = 0; q < b; ++q) {
        int no = n >> 1;
        fe nv0, nv1, nv2;
        if (lane < no) {
            fe V0 = sV[2 * lane], V1 = sV[2 * lane + 1], aV0 = saV[2 * lane], aV1 = saV[2 * lane + 1];
            fe am0 = sam[2 * lane], am1 = sam[2 * lane + 1];
            fsum f = {fzero(), fzero(), fzero()};
            fsum_add(f, V0, V1, aV0, aV1, am0, am1);
            ls[lane] = f;
            fe rr = t.r[j0 + q];
            nv0 = lerp(V0, V1, rr); nv1 = lerp(aV0, aV1, rr); nv2 = lerp(am0, am1, rr);
            if (no == 1) {
                t.end[0] = sub(V1, V0); t.end[1] = V0; t.end[2] = sub(am1, am0); t.end[3] = am0;
                fe vc = t.vconst ? *t.vconst : fzero();
                t.end[4] = add(nv0, vc);
            }
        }
        __syncthreads();
        if (lane == 0) {
            fsum sm = {fzero(), fzero(), fzero()};
            for (int k = 0; k < no; ++k) { sm.a = add(sm.a, ls[k].a); sm.c = add(sm.c, ls[k].c); sm.p1 = add(sm.p1, ls[k].p1); }
            fe *o = t.sums + 3 * (j0 + q);
            o[0] = sm.a; o[1] = sub(sub(sm.p1, sm.a), sm.c); o[2] = sm.c;
        }
        if (lane < no) { sV[lane] = nv0; saV[lane] = nv1; sam[lane] = nv2; }
        __syncthreads();
        n = no;
    }
}

// round sums of every wave-fold launch of the proof: slot R (a round) holds
// 3 x FW_BLOCKS block partials; nb[R] of them are live; dst[R] indexes msg.
__global__ void k_reduce_slots(const fe *part, const int *nb, const int64_t *dst, fe *msg)
{
    int R = blockIdx.x;
    fe acc[3] = {fzero(), fzero(), fzero()};
    for (int b = threadIdx.x; b < nb[R]; b += blockDim.x)
        for (int c = 0; c < 3; ++c) acc[c] = add(acc[c], part[((int64_t)R * 3 + c) * FB_MAXB + b]);
    block_sum<3>(acc);
    if (threadIdx.x == 0) { msg[dst[R]] = acc[0]; msg[dst[R] + 1] = acc[1]; msg[dst[R] + 2] = acc[2]; }
}
// last job j with start[j] <= x
__device__ __forceinline__ int job_of(const int64_t *start, int n, int64_t x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}
// k_reduce_slots over the slots of many proofs: block = one slot of one job
__global__ void k_reduce_slots_table(const launch::slot_job *J, const int64_t *start, int n)
{
    int j = job_of(start, n, blockIdx.x);
    const launch::slot_job &q = J[j];
    int R = (int)(blockIdx.x - start[j]);
    fe acc[3] = {fzero(), fzero(), fzero()};
    for (int b = threadIdx.x; b < q.nb[R]; b += blockDim.x)
        for (int c = 0; c < 3; ++c) acc[c] = add(acc[c], q.part[((int64_t)R * 3 + c) * FB_MAXB + b]);
    block_sum<3>(acc);
    if (threadIdx.x == 0) { q.msg[q.dst[R]] = acc[0]; q.msg[q.dst[R] + 1] = acc[1]; q.msg[q.dst[R] + 2] = acc[2]; }
}
// b of the slot rounds j >= 1 of the phase whose round 0 is slot R (rof[R] <
// 0), in round order: p_j(0) + p_j(1) = p_{j-1}(r[j-1]) exactly, because
// round j's tables are round j-1's folded at r[j-1], so
// b_j = p_j(1) - a_j - c_j = (c' + r (b' + r a')) - a_j - 2 c_j with
// (a', b', c') round j-1's message (msg order a, b, c: a x^2 + b x + c).
__device__ __forceinline__ void derive_chain(const int64_t *dst, const int64_t *rof, int n, int R, const fe *chal, fe *msg)
{
    if (rof[R] >= 0) return;
    for (int k = R + 1; k < n && rof[k] >= 0; ++k) {
        const fe *pv = msg + dst[k - 1];
        fe *cv = msg + dst[k];
        fe r = chal[rof[k]];
        fe s = add(pv[2], mul(r, add(pv[1], mul(r, pv[0]))));
        cv[1] = sub(sub(s, cv[0]), dbl(cv[2]));
    }
}
__global__ void k_derive_b(const int64_t *dst, const int64_t *rof, int n, const fe *chal, fe *msg)
{
    int R = blockIdx.x * blockDim.x + threadIdx.x;
    if (R < n) derive_chain(dst, rof, n, R, chal, msg);
}
__global__ void k_derive_b_table(const launch::slot_job *J, const int64_t *start, int n, int64_t total)
{
    int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (x >= total) return;
    int j = job_of(start, n, x);
    const launch::slot_job &q = J[j];
    derive_chain(q.dst, q.rof, (int)(start[j + 1] - start[j]), (int)(x - start[j]), q.chal, q.msg);
}
__global__ void k_copy_table(const launch::copy_job *J, const int64_t *start, int n, int64_t total)
{
    int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (x >= total) return;
    int j = job_of(start, n, x);
    const launch::copy_job &q = J[j];
    int64_t k = x - start[j];
    q.out[k] = q.addr ? *(const fe *)(uintptr_t)q.addr[k] : q.base[k];
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
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(6))) void k_general(const launch::layer_dev *D)
{
    const launch::layer_dev &d = D[blockIdx.y];
    if (!d.gv) return;                                   // layer 1: no general value
    bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
    bu_t bu = {d.bu.f, d.bu.s, d.bu.fh}, bv = {d.bv.f, d.bv.s, d.bv.fh};
    fe vu = *d.vu, vv = *d.vv, one = fone();
    fe uv = mul(vu, vv);
    fe acc = fzero();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += gridDim.x * blockDim.x) {
        int ty = d.ty[i], u = d.u[i], v = d.v[i];
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
    if (threadIdx.x == 0) d.part[blockIdx.x] = acc;
}
__global__ void k_reduce(const launch::layer_dev *D, int n)
{
    const launch::layer_dev &d = D[blockIdx.y];
    if (!d.gv) return;
    fe acc = fzero();
    for (int k = threadIdx.x; k < n; k += blockDim.x) acc = add(acc, d.part[k]);
    block_sum<1>(&acc);
    if (threadIdx.x == 0) *d.gv = acc;
}

// ---------------------------------------------------------------------
// verifier: wiring predicates per gate type (zk_verifier.cpp:219-606).
// For layer d (blockIdx.y) and each type t, sum over its gates g of type t
// of beta_g(g) beta_u(u) beta_v(v); a sum gate (5) contributes beta_v(0)
// times beta_u over its range [u, v), an exp-sum gate (12) beta_v(0) 2^k
// beta_u(u + k) over [u, v]; type 10 sums beta_g beta_u(u) and is scaled by
// beta_v(0) at the end.  Layer 1 (gv null) has no predicates.  Each thread
// keeps its per-type sums in its own LDS column; block partials go to
// part[(layer * blocks + block) * 14 + type].
// ---------------------------------------------------------------------
constexpr int PRED_T = 14;
__global__ __launch_bounds__(TPB) void k_predicates(const launch::layer_dev *D, fe *part)
{
    const launch::layer_dev &d = D[blockIdx.y];
    __shared__ fe acc[PRED_T][TPB];
    for (int t = 0; t < PRED_T; ++t) acc[t][threadIdx.x] = fzero();
    if (d.gv) {
        bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
        bu_t bu = {d.bu.f, d.bu.s, d.bu.fh}, bv = {d.bv.f, d.bv.s, d.bv.fh};
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += gridDim.x * blockDim.x) {
            int ty = d.ty[i], u = d.u[i], v = d.v[i];
            fe c;
            switch (ty) {
            case 0: case 1: case 6: case 7: case 8: case 9: case 13:
                c = mul(mul(beta_g(bg, i), beta_u(bu, u)), beta_u(bv, v));
                break;
            case 5: case 12: {
                fe tg = beta_g(bg, i), b0 = beta_u(bv, 0);
                c = fzero();
                for (int j = u; ty == 5 ? j < v : j <= v; ++j) {
                    c = add(c, mul(mul(tg, b0), beta_u(bu, j)));
                    if (ty == 12) b0 = dbl(b0);
                }
                break;
            }
            case 10: c = mul(beta_g(bg, i), beta_u(bu, u)); break;
            default: continue;
            }
            acc[ty][threadIdx.x] = add(acc[ty][threadIdx.x], c);
        }
    }
    __syncthreads();
    for (int t = 0; t < PRED_T; ++t) {
        fe x = acc[t][threadIdx.x];
        block_sum<1>(&x);
        if (threadIdx.x == 0) {
            if (t == 10 && d.gv) x = mul(x, beta_u({d.bv.f, d.bv.s, d.bv.fh}, 0));
            part[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * PRED_T + t] = x;
        }
    }
}
__global__ void k_predicates_sum(const fe *part, int blocks, fe *out)
{
    int l = blockIdx.x, t = threadIdx.x;
    if (t >= PRED_T) return;
    fe s = fzero();
    for (int b = 0; b < blocks; ++b) s = add(s, part[((int64_t)l * blocks + b) * PRED_T + t]);
    out[l * PRED_T + t] = s;
}
// out[p] = sum_k coef[k] x_p^k (one block per point)
__global__ __launch_bounds__(TPB) void k_poly_eval(const fe *coef, int64_t n, const fe *pts, fe *out)
{
    fe x = pts[blockIdx.x];
    fe xt = fone(), base = x;
    for (int e = threadIdx.x; e; e >>= 1) {               // x^threadIdx.x
        if (e & 1) xt = mul(xt, base);
        base = mul(base, base);
    }
    fe step = fone();
    base = x;
    for (int e = blockDim.x; e; e >>= 1) {
        if (e & 1) step = mul(step, base);
        base = mul(base, base);
    }
    fe acc = fzero();
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
        acc = add(acc, mul(coef[k], xt));
        xt = mul(xt, step);
    }
    block_sum<1>(&acc);
    if (threadIdx.x == 0) out[blockIdx.x] = acc;
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
    int brev_src;             // kinds 0 and 2 (hmode 0): src already in bit-reversed order
    const fe *qscale;         // kind 1: q of row b = qscale[b] times the one row at q
};
// input value j of row b; jm: where src holds it (j, or brev(j) for a
// bit-reversed source)
__device__ __forceinline__ fe ntt_fetch(const ntt_in &in, int64_t b, int64_t j, int64_t jm, const launch::ntt_tabs &T)
{
    if (in.kind == 0) return in.src[b * in.stride + jm];
    if (in.kind == 1) {
        int64_t k = cw_idx(j * in.step, in.cw_S, in.log_R);
        fe qv = in.qscale ? mul(in.qscale[b], in.q[k]) : in.q[b * in.cw_stride + k];
        return mul(in.l[b * in.cw_stride + k], qv);
    }
    int64_t slice = b >> in.log_Rl, s = b & ((1 << in.log_Rl) - 1);
    int len = in.lens[slice];
    if (j >= len) return fzero();
    fe c = in.src[slice * in.stride + (in.hmode ? len : 0) + jm];
    // coset twist w_SS^(s j) from two small tables
    fe x = mul(T.TA[(s << T.th) + (j & ((1 << T.th) - 1))], T.TB[(s << T.tb) + (j >> T.th)]);
    return mul(c, x);
}
// x w_4 (forward w_4 = -i) or x w_4^-1 (= i)
__device__ __forceinline__ fe mul_w4(const fe &x, int inverse)
{
    return inverse ? mk(negp(x.img), x.real) : mk(x.img, negp(x.real));
}
// x 2^30 mod p for canonical x: a rotation within 61 bits
__device__ __forceinline__ uint64_t rot30(uint64_t x) { return ((x << 30) & P) | (x >> 31); }
// x w_8 (forward w_8 = 2^30 (1 - i)) or x w_8^-1 (= 2^30 (1 + i))
__device__ __forceinline__ fe mul_w8(const fe &x, int inverse)
{
    uint64_t s = addp(x.real, x.img);
    uint64_t d = inverse ? subp(x.real, x.img) : subp(x.img, x.real);
    return inverse ? mk(rot30(d), rot30(s)) : mk(rot30(s), rot30(d));
}
__device__ __forceinline__ void bfly(fe &a, fe &b)
{
    fe t = b;
    b = sub(a, t);
    a = add(a, t);
}
// radix-2^G DIT on v[0..2^G) whose inputs are already premultiplied: stage
// t pairs (m, m + 2^t) with twiddle w_{2^(t+1)}^(m mod 2^t) (1, w_4, w_8)
template <int G>
__device__ __forceinline__ void radix_core(fe *v, int inverse)
{
    if (G >= 1) {
#pragma unroll
        for (int m = 0; m < (1 << G); m += 2) bfly(v[m], v[m + 1]);
    }
    if (G >= 2) {
#pragma unroll
        for (int m = 0; m < (1 << G); m += 4) {
            bfly(v[m], v[m + 2]);
            v[m + 3] = mul_w4(v[m + 3], inverse);
            bfly(v[m + 1], v[m + 3]);
        }
    }
    if (G >= 3) {
        bfly(v[0], v[4]);
        v[5] = mul_w8(v[5], inverse);
        bfly(v[1], v[5]);
        v[6] = mul_w4(v[6], inverse);
        bfly(v[2], v[6]);
        v[7] = mul_w4(mul_w8(v[7], inverse), inverse);
        bfly(v[3], v[7]);
    }
}
// premultiply v[m] by w_{2^(s0+G)}^(lo brev_G(m)), stages [s0, s0+G) of a
// DIT NTT then reduce to radix_core (w^(e + 2^S) = -w^e, S = s0+G-1)
template <int G>
__device__ __forceinline__ void radix_premul(fe *v, int s0, int64_t lo, const launch::ntt_tabs &T, int inverse)
{
    if (s0 == 0) return;
    int S = s0 + G - 1;
    const fe *W = (inverse ? T.Wi : T.W) + ((int64_t)1 << S);
    int64_t half = (int64_t)1 << S;
#pragma unroll
    for (int m = 1; m < (1 << G); ++m) {
        int br = 0;
#pragma unroll
        for (int i = 0; i < G; ++i) br |= ((m >> i) & 1) << (G - 1 - i);
        int64_t e = lo * br;
        fe w = W[e & (half - 1)];
        fe x = mul(v[m], w);
        v[m] = (e & half) ? neg(x) : x;
    }
}
// One block: positions [q 2^log_m, (q+1) 2^log_m) of batch row b, read in
// bit-reversed order, stages [0, log_m) as radix-8/4/2 passes (the first
// reading global memory, the last writing it, the rest in LDS).
constexpr int NTT_PAD = 3;          // one pad element per 8 breaks power-of-two LDS strides
__device__ __forceinline__ int lpad(int i) { return i + (i >> NTT_PAD); }
// group g (2^G values) of a radix-2^G pass over stages [s0, s0+G) of the
// tile of positions [q 2^log_m, (q+1) 2^log_m) held in sh
template <int G>
__device__ __forceinline__ void ntt_group(fe *sh, int g, int s0, int first, int last, const ntt_in &in, int64_t b,
                                          int64_t q, int log_m, int log_n, fe *out, int64_t out_stride, fe scale,
                                          int do_scale, const launch::ntt_tabs &T, int inverse)
{
    int lo = g & ((1 << s0) - 1), hi = g >> s0;
    int base = (hi << (s0 + G)) | lo;
    fe v[1 << G];
#pragma unroll
    for (int m = 0; m < (1 << G); ++m) {
        int i = base + (m << s0);
        if (first) {
            int64_t pos = (q << log_m) + i;
            int64_t j = log_n ? (int64_t)(__brevll((unsigned long long)pos) >> (64 - log_n)) : 0;
            v[m] = ntt_fetch(in, b, j, in.brev_src ? pos : j, T);
        } else {
            v[m] = sh[lpad(i)];
        }
    }
    radix_premul<G>(v, s0, lo, T, inverse);
    radix_core<G>(v, inverse);
#pragma unroll
    for (int m = 0; m < (1 << G); ++m) {
        int i = base + (m << s0);
        if (last) out[b * out_stride + (q << log_m) + i] = do_scale ? mul(v[m], scale) : v[m];
        else sh[lpad(i)] = v[m];
    }
}
template <int G>
__device__ __forceinline__ void ntt_pass(fe *sh, int M, int s0, int first, int last, const ntt_in &in, int64_t b, int64_t q,
                                         int log_m, int log_n, fe *out, int64_t out_stride, fe scale, int do_scale,
                                         const launch::ntt_tabs &T, int inverse)
{
    int ng = M >> G;
    for (int g = threadIdx.x; g < ng; g += blockDim.x)
        ntt_group<G>(sh, g, s0, first, last, in, b, q, log_m, log_n, out, out_stride, scale, do_scale, T, inverse);
}
// A whole transform of 2^(LM+3) values per block of NTT_FUSED_TPB threads
// (row b = blockIdx.x), with one global write instead of an LDS pass and a
// register pass: the 8 tiles of 2^LM bit-reversed positions go through
// stages [0, LM) in LDS four at a time; each thread keeps the values of its
// positions i < 2^LM in all 8 tiles in registers and finishes stages
// [LM, LM+3) as one radix-8 step across the tiles (the k_ntt_regs<3> step
// with s0 = LM, lo = i).
constexpr int NTT_FUSED_TPB = 1024;
// tiles 4 ROUND .. 4 ROUND + 3 of k_ntt_fused through stages [0, LM) in LDS,
// then into v[p][4 ROUND + k] (constant indices, so v stays in registers)
template <int LM, int ROUND, int PP>
__device__ __forceinline__ void ntt_fused_round(fe (&v)[PP][8], fe *sh, int TL, const ntt_in &in, int64_t b, fe *out,
                                                int64_t out_stride, fe scale, const launch::ntt_tabs &T, int inverse)
{
    constexpr int M = 1 << LM;
    for (int s0 = 0; s0 < LM;) {
        int G = LM - s0 >= 3 ? 3 : LM - s0;
        int ng = M >> G;
        for (int gg = threadIdx.x; gg < 4 * ng; gg += NTT_FUSED_TPB) {
            int k = gg / ng, g = gg - k * ng;
            fe *tile = sh + k * TL;
            int64_t q = 4 * ROUND + k;
            if (G == 3) ntt_group<3>(tile, g, s0, s0 == 0, 0, in, b, q, LM, LM + 3, out, out_stride, scale, 0, T, inverse);
            else if (G == 2) ntt_group<2>(tile, g, s0, s0 == 0, 0, in, b, q, LM, LM + 3, out, out_stride, scale, 0, T, inverse);
            else ntt_group<1>(tile, g, s0, s0 == 0, 0, in, b, q, LM, LM + 3, out, out_stride, scale, 0, T, inverse);
        }
        s0 += G;
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < PP; ++p)
#pragma unroll
        for (int k = 0; k < 4; ++k) v[p][4 * ROUND + k] = sh[k * TL + lpad(threadIdx.x + p * NTT_FUSED_TPB)];
    __syncthreads();
}
template <int LM>
__global__ __launch_bounds__(NTT_FUSED_TPB) void k_ntt_fused(ntt_in in, fe *out, int64_t out_stride, int inverse, fe scale,
                                                            int do_scale, launch::ntt_tabs T)
{
    constexpr int M = 1 << LM, TL = M + (M >> NTT_PAD), PP = M / NTT_FUSED_TPB;
    static_assert(PP >= 1, "tile smaller than the block");
    __shared__ fe sh[4 * TL];
    const int64_t b = blockIdx.x;
    fe v[PP][8];
    ntt_fused_round<LM, 0, PP>(v, sh, TL, in, b, out, out_stride, scale, T, inverse);
    ntt_fused_round<LM, 1, PP>(v, sh, TL, in, b, out, out_stride, scale, T, inverse);
#pragma unroll
    for (int p = 0; p < PP; ++p) {
        int i = threadIdx.x + p * NTT_FUSED_TPB;
        radix_premul<3>(v[p], LM, i, T, inverse);
        radix_core<3>(v[p], inverse);
        fe *o = out + b * out_stride + i;
#pragma unroll
        for (int m = 0; m < 8; ++m) o[(int64_t)m << LM] = do_scale ? mul(v[p][m], scale) : v[p][m];
    }
}
// LM: log2 of the LDS tile (the block's transform size when log_n >= LM)
template <int LM>
__global__ __launch_bounds__(TPB) void k_ntt_lds(ntt_in in, fe *out, int64_t out_stride, int log_n, int log_m, int inverse,
                                                fe scale, int do_scale, launch::ntt_tabs T)
{
    __shared__ fe sh[(1 << LM) + (1 << (LM - NTT_PAD))];
    int64_t b = blockIdx.y;
    int64_t q = blockIdx.x;
    int M = 1 << log_m;
    if (log_m == 0) {                                  // a transform of size 1
        if (threadIdx.x == 0) {
            fe x = ntt_fetch(in, b, 0, 0, T);
            out[b * out_stride + q] = do_scale ? mul(x, scale) : x;
        }
        return;
    }
    int s0 = 0;
    while (s0 < log_m) {
        int G = log_m - s0 >= 3 ? 3 : log_m - s0;
        int first = s0 == 0, last = s0 + G == log_m;
        if (G == 3) ntt_pass<3>(sh, M, s0, first, last, in, b, q, log_m, log_n, out, out_stride, scale, do_scale, T, inverse);
        else if (G == 2) ntt_pass<2>(sh, M, s0, first, last, in, b, q, log_m, log_n, out, out_stride, scale, do_scale, T, inverse);
        else ntt_pass<1>(sh, M, s0, first, last, in, b, q, log_m, log_n, out, out_stride, scale, do_scale, T, inverse);
        s0 += G;
        __syncthreads();
    }
}
// stages [s0, s0+G) of rows in global memory; thread = (row, group)
template <int G>
__global__ __launch_bounds__(TPB) void k_ntt_regs(fe *data, int64_t stride, int64_t nb, int log_n, int s0, int inverse,
                                                 fe scale, int do_scale, launch::ntt_tabs T)
{
    int64_t per = (int64_t)1 << (log_n - G);
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= nb * per) return;
    int64_t b = tid / per, r = tid % per;
    int64_t lo = r & (((int64_t)1 << s0) - 1), hi = r >> s0;
    int64_t base = (hi << (s0 + G)) | lo;
    fe *d = data + b * stride;
    fe v[1 << G];
#pragma unroll
    for (int m = 0; m < (1 << G); ++m) v[m] = d[base + ((int64_t)m << s0)];
    radix_premul<G>(v, s0, lo, T, inverse);
    radix_core<G>(v, inverse);
#pragma unroll
    for (int m = 0; m < (1 << G); ++m) d[base + ((int64_t)m << s0)] = do_scale ? mul(v[m], scale) : v[m];
}

// tw[e] = T1[e & 1023] * T2[e >> 10]
// one FRI fold of a natural-order codeword over the 2n-th roots of unity
// (fri.cpp:321-340 for one slice): out[i] = ((a + b) + r w^-i (a - b)) / 2
// with a = in[i], b = in[i + n] = f(-w^i); iw[i] = w^-i
__global__ void k_fri_fold_nat(const fe *in, int64_t n, const fe *r, const fe *iw, fe *out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    fe a = in[i], b = in[i + n];
    out[i] = add(halve(add(a, b)), mul(mul(halve(*r), iw[i]), sub(a, b)));
}
// value k of a codeword held in natural order (log_rows = 0) or coset-major
// (the extension's layout: value k in row k mod 2^log_rows, column
// k >> log_rows of rows of length cols)
__device__ __forceinline__ fe cw_at(const fe *cw, int64_t k, int log_rows, int64_t cols)
{
    return cw[(k & (((int64_t)1 << log_rows) - 1)) * cols + (k >> log_rows)];
}
// Merkle leaves of a codeword of 2 half values: leaf j = value j || value j + half
__global__ void k_pair_leaves(const fe *cw, int64_t half, int log_rows, int64_t cols, uint8_t *tree)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= half) return;
    fe *leaf = (fe *)(tree + 32 * (half + j));
    leaf[0] = cw_at(cw, j, log_rows, cols);
    leaf[1] = cw_at(cw, j + half, log_rows, cols);
}
// the FRI fold of k_fri_fold_nat on a natural or coset-major codeword of
// 2 half values with w^-i read from the powers tw[e] = w_SS^e:
// w_{2 half}^-i = tw[(SS - (i << sh)) mod SS], 2 half << sh = SS.  With
// leaves set, value i of the result is also stored in the leaf level of
// the result's own tree (half / 2 leaves: leaf i mod half/2, second value
// when i >= half/2).
__global__ void k_fri_fold_any(const fe *in, int64_t half, int log_rows, int64_t cols, const fe *r, const fe *tw,
                               int64_t SS, int sh, fe *out, uint8_t *leaves)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= half) return;
    fe a = cw_at(in, i, log_rows, cols), b = cw_at(in, i + half, log_rows, cols);
    fe iw = tw[(SS - (i << sh)) & (SS - 1)];
    fe o = add(halve(add(a, b)), mul(mul(halve(*r), iw), sub(a, b)));
    out[i] = o;
    if (leaves) {
        int64_t h2 = half >> 1;
        *(fe *)(leaves + 32 * (h2 + (i & (h2 - 1))) + (i >= h2 ? 16 : 0)) = o;
    }
}
// dst[brev(j)] = j < len ? src[j] : 0 over log_n >= 12 bits, in 64 x 64
// tiles: the tile of middle bits mid holds j = a 2^(log_n-6) + mid 2^6 + c
// (a, c < 64), read along c and written along brev6(a), both contiguous
__global__ __launch_bounds__(256) void k_brev_tile(const fe *src, int64_t len, fe *dst, int log_n)
{
    __shared__ fe t[64][65];
    int lm = log_n - 12, sh = log_n - 6;
    int64_t mid = blockIdx.x;
    int64_t rmid = lm ? (int64_t)(__brevll((unsigned long long)mid) >> (64 - lm)) : 0;
    for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
        int a = e >> 6, c = e & 63;
        int64_t j = ((int64_t)a << sh) | (mid << 6) | c;
        t[a][c] = j < len ? src[j] : fzero();
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
        int r = e >> 6, u = e & 63;                  // r = brev6(c), u = brev6(a)
        int c = (int)(__brev((unsigned)r) >> 26), a = (int)(__brev((unsigned)u) >> 26);
        dst[((int64_t)r << sh) | (rmid << 6) | u] = t[a][c];
    }
}
struct root_src {
    const uint8_t *p[32];
};
// out[32 k ..] = the root (node 1) of tree k, k < n
__global__ void k_copy_roots(root_src S, int n, uint8_t *out)
{
    int t = threadIdx.x;
    if (t >= 4 * n) return;
    ((uint64_t *)out)[t] = ((const uint64_t *)(S.p[t >> 2] + 32))[t & 3];
}
__global__ void k_count_diff(const fe *a, const fe *b, int64_t n, unsigned int *cnt)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && (a[i].img != b[i].img || a[i].real != b[i].real)) atomicAdd(cnt, 1u);
}
__global__ void k_twiddles(fe *tw, int64_t n, const fe *T1, const fe *T2)
{
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < n) tw[e] = mul(T1[e & 1023], T2[e >> 10]);
}
// stage-major twiddles and coset twist tables, read from tw[e] = w^e (e < SS)
__global__ void k_ntt_tables(fe *W, fe *Wi, int maxlog, fe *TA, fe *TB, int R, int th, int tb, const fe *tw, int64_t SS)
{
    int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t nw = (int64_t)1 << maxlog, na = (int64_t)R << th, nbb = (int64_t)R << tb;
    if (id < nw) {
        if (id == 0) { W[0] = fone(); Wi[0] = fone(); return; }
        int s = 63 - __clzll((unsigned long long)id);
        int64_t k = id - ((int64_t)1 << s);
        int64_t e = k * (SS >> (s + 1));
        W[id] = tw[e];
        Wi[id] = tw[(SS - e) & (SS - 1)];
        return;
    }
    id -= nw;
    if (id < na) {
        int64_t s = id >> th, i = id & (((int64_t)1 << th) - 1);
        TA[id] = tw[(s * i) & (SS - 1)];
        return;
    }
    id -= na;
    if (id < nbb) {
        int64_t s = id >> tb, i = id & (((int64_t)1 << tb) - 1);
        TB[id] = tw[((s * i) << th) & (SS - 1)];
    }
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
// The virtual oracle of slice j (poly_commit.h:246-263, 305-322) at x = w_SS^p
// is vo(p) = (G(p) - c0) x^-1 e with G = l q - (x^e - 1) h, e the slice
// length; q of slice j < nslice - 1 is cq[j] times Q (q's first SS values),
// the last slice's q its own (q's second SS values).  It is fused into the
// first FRI fold (fri.cpp:337-338) and never stored: with a = vo(p),
// b = vo(p + SS/2) and x(p + SS/2) = -x(p),
//   next = ((a + b) + r x^-1 (a - b)) / 2
//        = (e / 2) x^-1 ((G_a - G_b) + r x^-1 (G_a + G_b - 2 c0)),
// the same field element with two general products fewer.
__device__ __forceinline__ fe vo_numer(const fe *l, const fe *q, const fe *cq, const fe *h, int slice, bool last,
                                       int64_t p, int64_t e, int64_t SS, int log_R, const fe *tw)
{
    int64_t k = cw_idx(p, SS, log_R), id = (int64_t)slice * SS + k;
    fe qv = last ? q[SS + k] : mul(cq[slice], q[k]);
    fe xe = tw[(p * e) & (SS - 1)];
    return dot2(l[id], qv, sub(fone(), xe), h[id]);               // l q - (x^e - 1) h
}
// the first FRI fold's value at p < SS/2 of slice j from the virtual oracle
// at p and p + SS/2
__device__ __forceinline__ fe vo_value(const fe *l, const fe *q, const fe *cq, const fe *h, const fe *c0, int j, bool last,
                                       int64_t p, int64_t e, int64_t SS, int log_R, const fe *tw, const fe &r)
{
    fe ga = vo_numer(l, q, cq, h, j, last, p, e, SS, log_R, tw);
    fe gb = vo_numer(l, q, cq, h, j, last, p + SS / 2, e, SS, log_R, tw);
    fe xinv = tw[(SS - p) & (SS - 1)];
    fe sum = sub(add(ga, gb), add(c0[j], c0[j]));
    fe v = mul(xinv, add(sub(ga, gb), mul(mul(r, xinv), sum)));
    return mul_pow2(v, (63 - __builtin_clzll((unsigned long long)e) + 60) % 61);     // e / 2 = 2^(log2 e - 1)
}
// next (SS/2 per slice, codeword layout) from the virtual oracle at p and p + SS/2
__global__ __launch_bounds__(TPB) void k_vo_fold(const fe *l, const fe *q, const fe *cq, const fe *h, const fe *c0, int nslice,
                                                int64_t SS, int log_R, int64_t real, int64_t m, const fe *tw, const fe *rp,
                                                fe *next)
{
    int64_t Sn = SS / 2;
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nslice * Sn) return;
    int j = (int)(tid / Sn);
    int64_t st = tid - (int64_t)j * Sn;
    int64_t cols = Sn >> log_R;
    int64_t srow = st / cols, t = st - srow * cols;
    int64_t p = (t << log_R) + srow;
    bool last = j == nslice - 1;
    next[tid] = vo_value(l, q, cq, h, c0, j, last, p, last ? m : real, SS, log_R, tw, *rp);
}
// The first two FRI folds in one pass: thread = a position p < SS/4 of
// slice j of the second fold's output (row-contiguous, as k_fri_fold); it
// forms the first fold's values a at p and b at p + SS/4 from the virtual
// oracle, writes both (they are committed) and folds them at r[1] as
// k_fri_fold does, instead of reading them back.
__global__ __launch_bounds__(TPB) void k_vo_fold2(const fe *l, const fe *q, const fe *cq, const fe *h, const fe *c0,
                                                 int nslice, int64_t SS, int log_R, int64_t real, int64_t m, const fe *tw,
                                                 const fe *rp, fe *next0, fe *next1)
{
    int64_t S0 = SS / 2, S1 = SS / 4;
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nslice * S1) return;
    int j = (int)(tid / S1);
    int64_t st = tid - (int64_t)j * S1;
    int64_t cols = S1 >> log_R;
    int64_t srow = st / cols, t = st - srow * cols;
    int64_t p = (t << log_R) + srow;
    bool last = j == nslice - 1;
    int64_t e = last ? m : real;
    fe r0 = rp[0];
    fe a = vo_value(l, q, cq, h, c0, j, last, p, e, SS, log_R, tw, r0);
    fe b = vo_value(l, q, cq, h, c0, j, last, p + S1, e, SS, log_R, tw, r0);
    next0[(int64_t)j * S0 + cw_idx(p, S0, log_R)] = a;
    next0[(int64_t)j * S0 + cw_idx(p + S1, S0, log_R)] = b;
    fe imr = mul(tw[(SS - p * (SS / S0)) & (SS - 1)], halve(rp[1]));
    next1[tid] = add(halve(add(a, b)), mul(imr, sub(a, b)));
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
// FRI fold of 65 slices (fri.cpp:321-340, 367-379): next[p] = ((a+b) + w_S^-p r (a-b)) / 2,
// a = prev[p], b = prev[p + S/2], as (a+b)/2 + w_S^-p (r/2) (a-b) (halving is a
// rotation: two general products per position); one thread per (slice,
// position), row-contiguous.
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
    fe imr = mul(tw[(SS - p * (SS / S)) & (SS - 1)], halve(*rp));
    next[tid] = add(halve(add(a, b)), mul(imr, sub(a, b)));
}
// FRI levels with at most FRI_TAIL positions per slice run in k_fri_tail
constexpr int FRI_TAIL = 4096;
struct fri_levels {
    fe *next[24];               // codeword after each tail level
    int n;                      // tail levels
};
// the remaining folds of slice blockIdx.x in LDS: prev holds S positions per
// slice (S <= FRI_TAIL), level k halves it with challenge r[k]
__global__ __launch_bounds__(256) void k_fri_tail(const fe *prev, int64_t S, int log_R, const fe *r, const fe *tw,
                                                int64_t SS, fri_levels F)
{
    __shared__ fe v[FRI_TAIL];
    int64_t j = blockIdx.x;
    for (int64_t p = threadIdx.x; p < S; p += blockDim.x) v[p] = prev[j * S + cw_idx(p, S, log_R)];
    __syncthreads();
    for (int k = 0; k < F.n; ++k) {
        int64_t Sn = S / 2;
        fe rr = halve(r[k]);
        fe nv[FRI_TAIL / 2 / 256];
        int c = 0;
        for (int64_t p = threadIdx.x; p < Sn; p += blockDim.x, ++c) {
            fe a = v[p], b = v[p + Sn];
            fe imr = mul(tw[(SS - p * (SS / S)) & (SS - 1)], rr);
            nv[c] = add(halve(add(a, b)), mul(imr, sub(a, b)));
            F.next[k][j * Sn + cw_idx(p, Sn, log_R)] = nv[c];
        }
        __syncthreads();
        c = 0;
        for (int64_t p = threadIdx.x; p < Sn; p += blockDim.x, ++c) v[p] = nv[c];
        __syncthreads();
        S = Sn;
    }
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
// leaf t (in s-major thread order) of a codeword with S positions per slice
__device__ __forceinline__ void leaf_chain(const fe *cw, int64_t S, uint8_t *tree, int log_R, int64_t t)
{
    int64_t half = S / 2;
    int64_t i = leaf_of(t, half, log_R);
    int64_t k0 = cw_idx(i, S, log_R), k1 = cw_idx(i + half, S, log_R);
    uint64_t h[4] = {0, 0, 0, 0};
    for (int j = 0; j <= 64; ++j) chain_step(h, cw[j * S + k0], cw[j * S + k1]);
    store_digest(tree, half + i, h);
}
// The chain is written as the body of a loop over the thread's chains (one
// per thread at the launch's grid): the compiler then schedules the 65
// permutations so that the SIMDs issue 0.286 instead of 0.248 VALU
// instructions per cycle (rocprofv3 SQ_INSTS_VALU over GRBM_GUI_ACTIVE, the
// same instruction count; 5.06 against 5.70 ms for a 2^20-input proof's
// 786416 chains), the straight-line form being scheduled worse.
__global__ __launch_bounds__(TPB) void k_leaf_multi(leaf_segs L)
{
    const int64_t total = L.start[L.n], step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; tid < total; tid += step) {
        int sg = 0;
        while (tid >= L.start[sg + 1]) ++sg;
        leaf_chain(L.cw[sg], L.S[sg], L.tree[sg], L.log_R, tid - L.start[sg]);
    }
}
// the same over a segment table in device memory (the codewords of many
// proofs in one launch): start[] = prefix sums of leaf counts
// (in the loop form of k_leaf_multi, for the same schedule)
__global__ __launch_bounds__(TPB) void k_leaf_table(const launch::leaf_seg *G, const int64_t *start, int nseg)
{
    const int64_t total = start[nseg], step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; tid < total; tid += step) {
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {                     // last segment with start <= tid
            int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= tid) lo = mid; else hi = mid - 1;
        }
        const launch::leaf_seg &g = G[lo];
        leaf_chain(g.cw, g.S, g.tree, g.log_R, tid - start[lo]);
    }
}
// one query path (vpd_verifier.cpp:8-41): the leaf chain over the 65 value
// pairs must give the last path entry, and hashing up the path (sibling
// order by the position's bits) must give the root
__global__ void k_verify_paths(const launch::path_job *J, int njobs, const uint64_t *dg, const fe *vals, int *ok)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= njobs) return;
    launch::path_job j = J[k];
    const uint64_t *path = dg + 4 * j.path;
    uint64_t vh[4] = {0, 0, 0, 0};
    for (int i = 0; i < 65; ++i) chain_step(vh, vals[j.vals + 2 * i], vals[j.vals + 2 * i + 1]);
    bool good = true;
    const uint64_t *last = path + 4 * (j.len - 1);
    for (int q = 0; q < 4; ++q) good &= vh[q] == last[q];
    uint64_t cur[4] = {last[0], last[1], last[2], last[3]};
    int64_t pw = j.pow;
    for (int i = 0; i < j.len - 1; ++i) {
        const uint64_t *sib = path + 4 * i;
        uint64_t m[8];
        for (int q = 0; q < 4; ++q) {
            m[q] = (pw & 1) ? sib[q] : cur[q];
            m[4 + q] = (pw & 1) ? cur[q] : sib[q];
        }
        pw >>= 1;
        sha3_64(m, cur);
    }
    const uint64_t *root = dg + 4 * j.root;
    for (int q = 0; q < 4; ++q) good &= cur[q] == root[q];
    ok[k] = good ? 1 : 0;
}

struct tree_set {
    int n;
    uint8_t *tree[24];
    int64_t nleaf[24];
};
// Merkle levels with at most MK_TAIL nodes run in k_merkle_tails
constexpr int MK_TAIL = launch::MERKLE_TAIL;
// level h (h >= 1) of every tree whose level h has more than MK_TAIL nodes:
// nodes [nleaf>>h, 2*(nleaf>>h))
// the node hash of a tree: SHA3-256 (the proofs' trees) or SHA-256
template <int HASH>
__device__ __forceinline__ void node_hash(const uint64_t *m, uint64_t *h)
{
    if constexpr (HASH == launch::HASH_SHA256) sha256_64(m, h); else sha3_64(m, h);
}
// SHA3 nodes in the loop form of k_leaf_multi (one node per thread at this
// grid; 2^24 values' tree 9.03 -> 7.84 ms), SHA-256 nodes straight (the loop
// form measured 3 % slower for them)
template <int HASH>
__global__ __launch_bounds__(TPB) void k_merkle_multi(tree_set T, int h, int64_t total)
{
    const int64_t step = HASH == launch::HASH_SHA256 ? total : (int64_t)gridDim.x * blockDim.x;
    for (int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t0 < total; t0 += step) {
        int64_t tid = t0;
        int sg = 0;
        for (; sg < T.n; ++sg) {
            int64_t c = T.nleaf[sg] >> h;
            if (c <= MK_TAIL) continue;
            if (tid < c) break;
            tid -= c;
        }
        if (sg >= T.n) break;
        int64_t k = (T.nleaf[sg] >> h) + tid;
        const uint64_t *c = (const uint64_t *)(T.tree[sg] + 2 * k * 32);
        uint64_t m[8], hh[4];
        for (int q = 0; q < 8; ++q) m[q] = c[q];
        node_hash<HASH>(m, hh);
        store_digest(T.tree[sg], k, hh);
    }
}
// level h of many trees (a batch of proofs): item k covers the nodes
// [first, 2 first) of its tree, start = prefix sum of the node counts
__global__ __launch_bounds__(TPB) void k_merkle_level_table(const launch::mk_item *L, int n, int64_t total)
{
    for (int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; tid < total; tid += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (L[mid].start <= tid) lo = mid; else hi = mid - 1;
        }
        int64_t k = L[lo].first + (tid - L[lo].start);
        uint8_t *tree = L[lo].tree;
        const uint64_t *c = (const uint64_t *)(tree + 2 * k * 32);
        uint64_t m[8], hh[4];
        for (int q = 0; q < 8; ++q) m[q] = c[q];
        sha3_64(m, hh);
        store_digest(tree, k, hh);
    }
}
// the remaining levels of one tree (every level with <= MK_TAIL nodes) in
// one block of 256 threads: the first from the tree in memory, the rest
// from LDS
template <int HASH>
__device__ __forceinline__ void merkle_tail(uint8_t *tree, int64_t nl)
{
    __shared__ uint64_t dg[MK_TAIL][4];
    int h = 1;
    while ((nl >> h) > MK_TAIL) ++h;
    int cnt = (int)(nl >> h);
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        int64_t k = cnt + i;
        const uint64_t *c = (const uint64_t *)(tree + 2 * k * 32);
        uint64_t m[8], hh[4];
        for (int q = 0; q < 8; ++q) m[q] = c[q];
        node_hash<HASH>(m, hh);
        store_digest(tree, k, hh);
        for (int q = 0; q < 4; ++q) dg[i][q] = hh[q];
    }
    __syncthreads();
    while (cnt > 1) {
        cnt >>= 1;                                    // <= MK_TAIL / 2 nodes: at most one per thread
        uint64_t hh[(MK_TAIL + 511) / 512][4];
        int c = 0;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x, ++c) {
            uint64_t m[8];
            for (int q = 0; q < 4; ++q) { m[q] = dg[2 * i][q]; m[4 + q] = dg[2 * i + 1][q]; }
            node_hash<HASH>(m, hh[c]);
            store_digest(tree, cnt + i, hh[c]);
        }
        __syncthreads();
        c = 0;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x, ++c)
            for (int q = 0; q < 4; ++q) dg[i][q] = hh[c][q];
        __syncthreads();
    }
}
template <int HASH>
__global__ __launch_bounds__(256) void k_merkle_tails(tree_set T)
{
    merkle_tail<HASH>(T.tree[blockIdx.x], T.nleaf[blockIdx.x]);
}
__global__ __launch_bounds__(256) void k_merkle_tails_table(const launch::leaf_seg *G)
{
    merkle_tail<launch::HASH_SHA3>(G[blockIdx.x].tree, G[blockIdx.x].S / 2);
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
__global__ void k_sha256_batch(const uint64_t *in, uint64_t *out, int64_t n)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint64_t m[8], h[4];
    for (int i = 0; i < 8; ++i) m[i] = in[8 * k + i];
    sha256_64(m, h);
    for (int i = 0; i < 4; ++i) out[4 * k + i] = h[i];
}
// pad leaves of a Merkle tree: H(0^64) (merkle_tree.cpp:105-117)
template <int HASH>
__global__ void k_merkle_pad(uint8_t *tree, int64_t pad, int64_t n_leaf)
{
    int64_t k = n_leaf + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= pad) return;
    uint64_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0}, h[4];
    node_hash<HASH>(m, h);
    store_digest(tree, pad + k, h);
}
__global__ void k_fe_mul(const fe *a, const fe *b, fe *o, int64_t n)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) o[k] = mul(a[k], b[k]);
}

int ilog2(int64_t x) { int l = 0; while (((int64_t)1 << l) < x) ++l; return l; }

// ---- NTT tile autotuner: the LDS tile (2^10, 2^11 or 2^12 elements) of
// each transform size, timed once per process on the first context that
// needs that size; any tile gives the same bits (exact arithmetic) ----
constexpr int NTT_LM_MIN = 10, NTT_LM_MAX = 12;
// the one-block transform k_ntt_fused (sizes 2^13 and 2^14), a candidate
// beside the tiles
constexpr int NTT_FUSED = 1;
bool fused_fits(int log_n) { return log_n == 13 || log_n == 14; }
std::mutex g_tune_mu;
int g_ntt_lm[64];                     // chosen tile per log_n (0: not tuned)
float g_ntt_us[64][NTT_LM_MAX + 1];   // measured us per candidate (0: not run)
int forced_lm()
{
    static int v = [] {
        const char *e = getenv("VIRGO_NTT_LM");
        int x = e ? atoi(e) : 0;
        return (x >= NTT_LM_MIN && x <= NTT_LM_MAX) || x == NTT_FUSED ? x : 0;
    }();
    return v;
}
int ntt_lm_for(int log_n)
{
    if (int f = forced_lm()) return f == NTT_FUSED && !fused_fits(log_n) ? NTT_LM : f;
    int t = (log_n >= 0 && log_n < 64) ? __atomic_load_n(&g_ntt_lm[log_n], __ATOMIC_ACQUIRE) : 0;
    return t ? t : NTT_LM;
}

void ntt_run_lm(const ntt_in &in, fe *out, int64_t out_stride, int64_t nb, int log_n, int inverse, fe scale, int do_scale,
                const launch::ntt_tabs &T, hipStream_t st, int lm)
{
    if (lm == NTT_FUSED) {
        if (log_n == 14) {
            hipLaunchKernelGGL(k_ntt_fused<11>, dim3((unsigned)nb), dim3(NTT_FUSED_TPB), 0, st, in, out, out_stride, inverse,
                               scale, do_scale, T);
            return;
        }
        if (log_n == 13) {
            hipLaunchKernelGGL(k_ntt_fused<10>, dim3((unsigned)nb), dim3(NTT_FUSED_TPB), 0, st, in, out, out_stride, inverse,
                               scale, do_scale, T);
            return;
        }
        lm = NTT_LM;
    }
    int log_m = log_n < lm ? log_n : lm;
    int rest = log_n - log_m;
    dim3 grid((unsigned)((int64_t)1 << (log_n - log_m)), (unsigned)nb);
    int ds0 = rest == 0 ? do_scale : 0;
    switch (lm) {
    case 10: hipLaunchKernelGGL(k_ntt_lds<10>, grid, dim3(TPB), 0, st, in, out, out_stride, log_n, log_m, inverse, scale, ds0, T); break;
    case 12: hipLaunchKernelGGL(k_ntt_lds<12>, grid, dim3(TPB), 0, st, in, out, out_stride, log_n, log_m, inverse, scale, ds0, T); break;
    default: hipLaunchKernelGGL(k_ntt_lds<NTT_LM>, grid, dim3(TPB), 0, st, in, out, out_stride, log_n, log_m, inverse, scale, ds0, T); break;
    }
    int s0 = log_m;
    while (rest > 0) {
        int g = rest > 3 ? 3 : rest;
        rest -= g;
        int64_t threads = nb * ((int64_t)1 << (log_n - g));
        int ds = rest == 0 ? do_scale : 0;
        dim3 gr(nblk(threads)), bk(TPB);
        switch (g) {
        case 1: hipLaunchKernelGGL(k_ntt_regs<1>, gr, bk, 0, st, out, out_stride, nb, log_n, s0, inverse, scale, ds, T); break;
        case 2: hipLaunchKernelGGL(k_ntt_regs<2>, gr, bk, 0, st, out, out_stride, nb, log_n, s0, inverse, scale, ds, T); break;
        default: hipLaunchKernelGGL(k_ntt_regs<3>, gr, bk, 0, st, out, out_stride, nb, log_n, s0, inverse, scale, ds, T); break;
        }
        s0 += g;
    }
}
void ntt_run(const ntt_in &in, fe *out, int64_t out_stride, int64_t nb, int log_n, int inverse, fe scale, int do_scale,
             const launch::ntt_tabs &T, hipStream_t st)
{
    ntt_run_lm(in, out, out_stride, nb, log_n, inverse, scale, do_scale, T, st, ntt_lm_for(log_n));
}


// ---------------------------------------------------------------------
// fft_gkr: the FFT/evaluation circuit of the commitment and the table builds
// of its GKR layers (fft_circuit_GKR.cpp:22-743).  N = 2^lg coefficients.
// ---------------------------------------------------------------------
// layer lg of the circuit: out[x] = prod_i (bit (lg-1-i) of x ? 1 - r_i : r_i)
__global__ void k_fg_eq(fe *out, const fe *r, int lg)
{
    int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (x >= ((int64_t)1 << lg)) return;
    fe one = fone(), v = one;
    for (int i = 0; i < lg; ++i) v = mul(v, ((x >> (lg - 1 - i)) & 1) ? sub(one, r[i]) : r[i]);
    out[x] = v;
}
// power tables of npts points: lo[i][t] = x_i^t (t < 2^lo), hi[i][s] = x_i^(s 2^lo)
__global__ void k_fg_pow(const fe *pts, int npts, int lo, int hi, fe *plo, fe *phi)
{
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t nlo = (int64_t)npts << lo, nhi = (int64_t)npts << hi;
    if (tid >= nlo + nhi) return;
    bool isl = tid < nlo;
    int64_t k = isl ? tid : tid - nlo;
    int bits = isl ? lo : hi;
    int i = (int)(k >> bits);
    uint64_t e = (uint64_t)(k & (((int64_t)1 << bits) - 1));
    if (!isl) e <<= lo;
    fe x = pts[i], acc = fone();
    while (e) { if (e & 1) acc = mul(acc, x); x = mul(x, x); e >>= 1; }
    (isl ? plo : phi)[k] = acc;
}
__device__ __forceinline__ fe fg_pw(const fe *plo, const fe *phi, int lo, int hi, int i, int64_t e)
{
    return mul(plo[((int64_t)i << lo) + (e & (((int64_t)1 << lo) - 1))], phi[((int64_t)i << hi) + (e >> lo)]);
}
// one butterfly layer of the circuit's inverse DFT (:44-75); the x_arr[k]
// twiddle is inv_rou^(k 2^dep); scale (the 1/N of :76-82) when do_scale
__global__ void k_fg_butterfly(const fe *pre, fe *cur, int dep, int lg, const fe *plo, const fe *phi, int lo, int hi,
                               int tw_pt, fe scale, int do_scale)
{
    int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t half = (int64_t)1 << (lg - 1);
    if (idx >= half) return;
    int64_t k = idx >> dep, j = idx & (((int64_t)1 << dep) - 1);
    fe x = fg_pw(plo, phi, lo, hi, tw_pt, k << dep);
    fe l = pre[k << (dep + 1) | j], r = mul(x, pre[k << (dep + 1) | ((int64_t)1 << dep) | j]);
    fe a = add(l, r), b = sub(l, r);
    int64_t h2 = (int64_t)1 << (lg - dep - 1);
    if (do_scale) { a = mul(a, scale); b = mul(b, scale); }
    cur[k << dep | j] = a;
    cur[(k + h2) << dep | j] = b;
}
// the 64 x N evaluation layer (:84-102): out[j + (i << lg)] = coef[j] x_i^j
__global__ void k_fg_eval(const fe *coef, const fe *plo, const fe *phi, int lo, int hi, int lg, fe *out)
{
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= ((int64_t)64 << lg)) return;
    int i = (int)(tid >> lg);
    int64_t j = tid & (((int64_t)1 << lg) - 1);
    out[tid] = mul(coef[j], fg_pw(plo, phi, lo, hi, i, j));
}
// the 64 outputs (:103-111): one block per row
__global__ __launch_bounds__(TPB) void k_fg_rowsum(const fe *ev, int lg, fe *outs)
{
    int64_t N = (int64_t)1 << lg;
    const fe *row = ev + blockIdx.x * N;
    fe acc = fzero();
    for (int64_t j = threadIdx.x; j < N; j += blockDim.x) acc = add(acc, row[j]);
    block_sum<1>(&acc);
    if (threadIdx.x == 0) outs[blockIdx.x] = acc;
}
// addition layer tables (:289-298): amb[j] = beta_g(j >> lg), aVb = 0
__global__ void k_fg_add_tables(bg_t bg, int lg, fe *amb, fe *aVb)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= ((int64_t)64 << lg)) return;
    amb[j] = beta_g(bg, (int)(j >> lg));
    aVb[j] = fzero();
}
// mult layer tables (:411-421): amb[i] = sum_j beta_g(j N + i) x_j^i, aVb = 0
__global__ void k_fg_mult_tables(bg_t bg, const fe *plo, const fe *phi, int lo, int hi, int lg, fe *amb, fe *aVb)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ((int64_t)1 << lg)) return;
    fe acc = fzero();
    for (int j = 0; j < 64; ++j)
        acc = add(acc, mul(beta_g(bg, (int)(((int64_t)j << lg) + i)), fg_pw(plo, phi, lo, hi, j, i)));
    amb[i] = acc;
    aVb[i] = fzero();
}
// ifft layers, phase 1 (:583-605), blockIdx.y = layer: gate g = k 2^dep + j
// and g + N/2 both read u = k 2^(dep+1) + j and v = u + 2^dep
__global__ void k_fg_ifft_p1m(const launch::fg_layer_dev *D, const fe *plo, const fe *phi, int lo, int hi, int tw_pt,
                              int lg)
{
    const launch::fg_layer_dev &d = D[blockIdx.y];
    int dep = d.dep;
    int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (u >= ((int64_t)1 << lg)) return;
    if ((u >> dep) & 1) { d.amb[u] = fzero(); d.aVb[u] = fzero(); return; }
    bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
    int64_t k = u >> (dep + 1), j = u & (((int64_t)1 << dep) - 1);
    int g = (int)(k << dep | j);
    fe b1 = beta_g(bg, g), b2 = beta_g(bg, g + (1 << (lg - 1)));
    fe xv = mul(fg_pw(plo, phi, lo, hi, tw_pt, k << dep), d.pre[u | ((int64_t)1 << dep)]);
    d.amb[u] = add(b1, b2);
    d.aVb[u] = sub(mul(b1, xv), mul(b2, xv));
}
// ifft layers, phase 2 (:638-665), blockIdx.y = layer: entries v = u + 2^dep
__global__ void k_fg_ifft_p2m(const launch::fg_layer_dev *D, const fe *plo, const fe *phi, int lo, int hi, int tw_pt,
                              int lg)
{
    const launch::fg_layer_dev &d = D[blockIdx.y];
    int dep = d.dep;
    int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= ((int64_t)1 << lg)) return;
    if (!((v >> dep) & 1)) { d.amb[v] = fzero(); d.aVb[v] = fzero(); return; }
    bg_t bg = {d.bg.g0f, d.bg.g0s, d.bg.g1f, d.bg.g1s, d.bg.fh};
    bu_t bu = {d.bu.f, d.bu.s, d.bu.fh};
    int64_t u = v ^ ((int64_t)1 << dep);
    int64_t k = u >> (dep + 1), j = u & (((int64_t)1 << dep) - 1);
    int g = (int)(k << dep | j);
    fe tu = beta_u(bu, (int)u);
    fe t1 = mul(beta_g(bg, g), tu), t2 = mul(beta_g(bg, g + (1 << (lg - 1))), tu);
    fe x = fg_pw(plo, phi, lo, hi, tw_pt, k << dep), vu = *d.vu;
    d.amb[v] = sub(mul(t1, x), mul(t2, x));
    d.aVb[v] = add(mul(t1, vu), mul(t2, vu));
}

// ---------------------------------------------------------------------
// CSR transposes of every gate layer, built on the device when a context is
// created (the structure of zk_prover.cpp:420-587 and :845-972): phase 1
// lists, for every entry x of the previous layer, the gates that add into
// it through u (a sum gate into every x of its range, its record carrying x
// in place of v); phase 2 lists by v the gates that add into entry v.
// All layers in six launches (blockIdx.y = layer, or layer and list): the
// offsets A (nt + 1 words) are zeroed, entry x's records counted into
// A[x + 2] (the last entry's count is never needed), A scanned inclusively
// so that A[x + 1] = start(x), and each record placed at atomicAdd(&A[x +
// 1], 1), which leaves A[x + 1] = end(x) = start(x + 1): A ends as the
// offsets, A[0] = 0 and A[nt] = the total, with no cursor array.  The order
// of records within one entry is arbitrary, which no result depends on (the
// table builds only add field elements).
// ---------------------------------------------------------------------
__device__ __forceinline__ void csr_p1_range(int t, int32_t u, int32_t v, int64_t &lo, int64_t &hi)
{
    switch (t) {
    case 0: case 1: case 4: case 6: case 7: case 8: case 9: case 10: case 13: lo = u; hi = (int64_t)u + 1; break;
    case 5: lo = u; hi = v; break;
    case 12: lo = u; hi = (int64_t)v + 1; break;
    default: lo = hi = 0; break;
    }
}
__device__ __forceinline__ bool csr_p2_by_v(int t) { return t == 0 || t == 1 || t == 7 || t == 8 || t == 9 || t == 13; }

__global__ void k_csr_zero(const launch::csr_layer *L)
{
    const launch::csr_layer &d = L[blockIdx.y];
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x <= d.nt; x += (int64_t)gridDim.x * blockDim.x) {
        d.off1[x] = 0u;
        d.off2[x] = 0u;
    }
}
__global__ void k_csr_count(const launch::csr_layer *L)
{
    const launch::csr_layer &d = L[blockIdx.y];
    int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= d.ng) return;
    int t = d.ty[g];
    int32_t uu = d.u[g], vv = d.v[g];
    int64_t lo, hi;
    csr_p1_range(t, uu, vv, lo, hi);
    for (int64_t x = lo; x < hi; ++x)
        if (x + 2 <= d.nt) atomicAdd(&d.off1[x + 2], 1u);
    if (csr_p2_by_v(t) && (int64_t)vv + 2 <= d.nt) atomicAdd(&d.off2[vv + 2], 1u);
}
__global__ void k_csr_fill(const launch::csr_layer *L)
{
    const launch::csr_layer &d = L[blockIdx.y];
    int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= d.ng) return;
    int t = d.ty[g];
    int32_t uu = d.u[g], vv = d.v[g];
    launch::gate_rec rec = {(int32_t)g, uu, vv, t};
    int64_t lo, hi;
    csr_p1_range(t, uu, vv, lo, hi);
    bool range = t == 5 || t == 12;
    for (int64_t x = lo; x < hi; ++x) {
        launch::gate_rec r1 = rec;
        if (range) r1.v = (int32_t)x;
        d.ent1[atomicAdd(&d.off1[x + 1], 1u)] = r1;
    }
    if (csr_p2_by_v(t)) d.ent2[atomicAdd(&d.off2[vv + 1], 1u)] = rec;
}

// inclusive scan of the 2048 consecutive words of block blockIdx.x of list
// blockIdx.y (layer y / 2, off1 or off2); tops[b] = block b's total
constexpr int SCAN_T = 256, SCAN_E = 8, SCAN_B = SCAN_T * SCAN_E;
__device__ __forceinline__ void scan_list(const launch::csr_layer &d, int which, uint32_t *&a, int64_t &n, uint32_t *&tops)
{
    a = which ? d.off2 : d.off1;
    tops = which ? d.tops2 : d.tops1;
    n = d.nt + 1;
}
// block-wide inclusive scan of one value per thread; returns the thread's
// inclusive prefix, *total the block's sum
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t *sh, uint32_t *total)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        uint32_t y = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    uint32_t r = sh[threadIdx.x];
    *total = sh[SCAN_T - 1];
    __syncthreads();
    return r;
}
__global__ void __launch_bounds__(SCAN_T) k_scan_block(const launch::csr_layer *L)
{
    __shared__ uint32_t sh[SCAN_T];
    uint32_t *a, *tops;
    int64_t n;
    scan_list(L[blockIdx.y >> 1], blockIdx.y & 1, a, n, tops);
    int64_t b0 = blockIdx.x * (int64_t)SCAN_B;
    if (b0 >= n) return;                       // the same for the whole block
    int64_t base = b0 + threadIdx.x * SCAN_E;
    uint32_t x[SCAN_E], sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_E; ++k) {
        sum += base + k < n ? a[base + k] : 0u;
        x[k] = sum;
    }
    uint32_t tot;
    uint32_t pre = block_scan_u32(sum, sh, &tot) - sum;
#pragma unroll
    for (int k = 0; k < SCAN_E; ++k)
        if (base + k < n) a[base + k] = x[k] + pre;
    if (threadIdx.x == 0) tops[blockIdx.x] = tot;
}
// inclusive scan of one list's block totals, one block per list, 2048 at a
// time with the running total carried
__global__ void __launch_bounds__(SCAN_T) k_scan_tops(const launch::csr_layer *L)
{
    __shared__ uint32_t sh[SCAN_T];
    uint32_t *a, *tops;
    int64_t n;
    scan_list(L[blockIdx.x >> 1], blockIdx.x & 1, a, n, tops);
    int64_t nb = (n + SCAN_B - 1) / SCAN_B;
    uint32_t carry = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += SCAN_B) {
        int64_t base = c0 + threadIdx.x * SCAN_E;
        uint32_t x[SCAN_E], sum = 0;
#pragma unroll
        for (int k = 0; k < SCAN_E; ++k) {
            sum += base + k < nb ? tops[base + k] : 0u;
            x[k] = sum;
        }
        uint32_t tot;
        uint32_t pre = block_scan_u32(sum, sh, &tot) - sum + carry;
#pragma unroll
        for (int k = 0; k < SCAN_E; ++k)
            if (base + k < nb) tops[base + k] = x[k] + pre;
        carry += tot;
    }
}
// adds the inclusive prefix of the previous blocks' totals
__global__ void k_scan_add(const launch::csr_layer *L)
{
    uint32_t *a, *tops;
    int64_t n;
    scan_list(L[blockIdx.y >> 1], blockIdx.y & 1, a, n, tops);
    int64_t base = blockIdx.x * (int64_t)SCAN_B;
    if (blockIdx.x == 0 || base >= n) return;
    uint32_t add = tops[blockIdx.x - 1];
    for (int k = threadIdx.x; k < SCAN_B; k += blockDim.x)
        if (base + k < n) a[base + k] += add;
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
// ---- CSR transposes ----
int64_t csr_tops_len(int64_t nt) { return (nt + 1 + SCAN_B - 1) / SCAN_B; }
void csr_build_multi(const csr_layer *L, int n, int64_t max_ng, int64_t max_nt, hipStream_t s)
{
    if (n <= 0) return;
    unsigned zb = nblk(max_nt + 1);
    if (zb > 4096) zb = 4096;
    hipLaunchKernelGGL(k_csr_zero, dim3(zb, n), dim3(TPB), 0, s, L);
    unsigned gb = nblk(max_ng > 0 ? max_ng : 1);
    hipLaunchKernelGGL(k_csr_count, dim3(gb, n), dim3(TPB), 0, s, L);
    unsigned sb = (unsigned)csr_tops_len(max_nt);
    hipLaunchKernelGGL(k_scan_block, dim3(sb, 2 * n), dim3(SCAN_T), 0, s, L);
    hipLaunchKernelGGL(k_scan_tops, dim3(2 * n), dim3(SCAN_T), 0, s, L);
    hipLaunchKernelGGL(k_scan_add, dim3(sb, 2 * n), dim3(SCAN_T), 0, s, L);
    hipLaunchKernelGGL(k_csr_fill, dim3(gb, n), dim3(TPB), 0, s, L);
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
    // up to 11 rounds per launch (a 2^20-entry layer: one block launch, then the tail)
    while (n > FOLD_TAIL) {
        int h = log_n - k < 11 ? log_n - k : 11;
        int64_t no = n >> h;
        hipLaunchKernelGGL(k_mle_block, dim3((unsigned)no), dim3(1 << (h - 1)), 0, s, src, bufs[cur], r + k, h);
        src = bufs[cur];
        cur ^= 1;
        n = no;
        k += h;
    }
    hipLaunchKernelGGL(k_mle_tail, dim3(1), dim3(FOLD_TAIL), 0, s, src, (int)n, r + k, log_n - k, dst);
}
void eq_expand(fe *out, const fe *lo, const fe *hi, int fh, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_eq_expand, dim3(nblk(n)), dim3(TPB), 0, st, out, lo, hi, fh, n);
}
// phase-2 gates that reach entry 0 unconditionally -> part[MAX_PARTS]
static void zero_list(const layer_dev *D, int n, int nz_max, hipStream_t s)
{
    if (nz_max <= 0) return;
    unsigned blocks = nblk(nz_max);
    if (blocks > MAX_PARTS) blocks = MAX_PARTS;
    hipLaunchKernelGGL(k_p2_zero, dim3(blocks, n), dim3(TPB), 0, s, D);
    hipLaunchKernelGGL(k_zero_sum, dim3(1, n), dim3(1024), 0, s, D, (int)blocks);
}
void layer_tables(int phase, const layer_dev *D, int n, int nt, int nz_max, hipStream_t s, int kl)
{
    if (kl < 0) kl = 0;
    if (kl > 6) kl = 6;
    dim3 grid(nblk((int64_t)nt << kl), n);
    if (phase == 0) {
        hipLaunchKernelGGL(k_p1_gather, grid, dim3(TPB), 0, s, D, kl);
        return;
    }
    zero_list(D, n, nz_max, s);
    hipLaunchKernelGGL(k_p2_gather, grid, dim3(TPB), 0, s, D, kl);
}
int gather_fold(int phase, const layer_dev *D, const fold_inst *I, int n, int L, int nz_max, hipStream_t s)
{
    int js[32], ks[32];
    if (fold_plan(L, js, ks, GF_B) == 0) return 0;
    if (phase == 1) zero_list(D, n, nz_max, s);
    int b = ks[0];
    dim3 g((unsigned)fold_blocks(L, 0, GF_B), n);
    static_assert(GF_B == 2, "gather_fold instantiates B = 1, 2");
    if (phase == 0) {
        if (b == 2) hipLaunchKernelGGL((k_gather_fold<0, 2>), g, dim3(256), 0, s, D, I);
        else hipLaunchKernelGGL((k_gather_fold<0, 1>), g, dim3(256), 0, s, D, I);
    } else {
        if (b == 2) hipLaunchKernelGGL((k_gather_fold<1, 2>), g, dim3(256), 0, s, D, I);
        else hipLaunchKernelGGL((k_gather_fold<1, 1>), g, dim3(256), 0, s, D, I);
    }
    return b;
}
int fold_plan(int L, int *js, int *ks, int first)
{
    // streaming launches of <= FB_B rounds (the first <= first) until at
    // most FB_TAIL rounds remain
    int n = 0, j = 0;
    while (L - j > FB_TAIL) {
        int cap = n == 0 ? first : FB_B;
        int b = L - j - FB_TAIL < cap ? L - j - FB_TAIL : cap;
        js[n] = j; ks[n] = b; ++n;
        j += b;
    }
    return n;
}
int fold_blocks(int L, int j, int first)
{
    int js[32], ks[32];
    int n = fold_plan(L, js, ks, first);
    for (int k = 0; k < n; ++k)
        if (js[k] == j) {
            int64_t lanes = (int64_t)1 << (L - j - ks[k]);
            int64_t g = (lanes + 255) / 256;
            return (int)(g > FB_MAXB ? FB_MAXB : g);
        }
    return 0;
}
int sumcheck_multi(const fold_inst *I, int ninst, int L, hipStream_t s, int first, int skip_first)
{
    int js[32], ks[32];
    int nl = fold_plan(L, js, ks, first);
    int cur = 0, rounds = 0;
    for (int k = 0; k < nl; ++k) {
        int j = js[k], b = ks[k];
        if (k > 0 || !skip_first) {
            int64_t nout = (int64_t)1 << (L - j - b);
            dim3 g((unsigned)fold_blocks(L, j, first), ninst);
            if (j == 0) {
                if (b == 3) hipLaunchKernelGGL((k_fold_reg<3, true>), g, dim3(256), 0, s, I, cur, j, nout, rounds);
                else if (b == 2) hipLaunchKernelGGL((k_fold_reg<2, true>), g, dim3(256), 0, s, I, cur, j, nout, rounds);
                else hipLaunchKernelGGL((k_fold_reg<1, true>), g, dim3(256), 0, s, I, cur, j, nout, rounds);
            } else {
                if (b == 3) hipLaunchKernelGGL((k_fold_reg<3, false>), g, dim3(256), 0, s, I, cur, j, nout, rounds);
                else if (b == 2) hipLaunchKernelGGL((k_fold_reg<2, false>), g, dim3(256), 0, s, I, cur, j, nout, rounds);
                else hipLaunchKernelGGL((k_fold_reg<1, false>), g, dim3(256), 0, s, I, cur, j, nout, rounds);
            }
        }
        cur ^= 1;
        rounds += b;
    }
    int j = nl ? js[nl - 1] + ks[nl - 1] : 0;
    hipLaunchKernelGGL(k_fold_tail, dim3(1, ninst), dim3(64), 0, s, I, cur, j, L - j);
    return rounds;
}
void reduce_slots(const fe *slots, int nslots, const int *nb, const int64_t *dst, const int64_t *rof, const fe *chal,
                  fe *msg, hipStream_t s)
{
    if (!nslots) return;
    hipLaunchKernelGGL(k_reduce_slots, dim3(nslots), dim3(256), 0, s, slots, nb, dst, msg);
    hipLaunchKernelGGL(k_derive_b, dim3((nslots + 63) / 64), dim3(64), 0, s, dst, rof, nslots, chal, msg);
}
void half_multi(const half_desc *d, int nd, int64_t total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(k_half_multi, dim3(nblk(total)), dim3(TPB), 0, s, d, nd, total);
}
void general_multi(const layer_dev *D, int n, int ng_max, hipStream_t s)
{
    unsigned blocks = nblk(ng_max);
    if (blocks > MAX_PARTS) blocks = MAX_PARTS;
    hipLaunchKernelGGL(k_general, dim3(blocks, n), dim3(TPB), 0, s, D);
    hipLaunchKernelGGL(k_reduce, dim3(1, n), dim3(1024), 0, s, D, (int)blocks);
}
void ntt_tables(fe *W, fe *Wi, int maxlog, fe *TA, fe *TB, int log_R, int th, int tb, const fe *tw, int64_t SS,
                hipStream_t s)
{
    int R = TA ? 1 << log_R : 0;
    int64_t n = ((int64_t)1 << maxlog) + ((int64_t)R << th) + ((int64_t)R << tb);
    hipLaunchKernelGGL(k_ntt_tables, dim3(nblk(n)), dim3(TPB), 0, s, W, Wi, maxlog, TA, TB, R, th, tb, tw, SS);
}
int ntt_tune(int log_n, int64_t nb, const fe *in, fe *out, const ntt_tabs &T, hipStream_t s)
{
    if (log_n <= NTT_LM_MIN || log_n >= 64 || forced_lm()) return 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_ntt_lm[log_n]) return 0;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return -1; }
    ntt_in ni;
    memset(&ni, 0, sizeof ni);
    ni.kind = 0; ni.src = in; ni.stride = (int64_t)1 << log_n;
    int best = NTT_LM;
    float best_us = 0;
    bool ok = true;
    for (int lm = fused_fits(log_n) ? NTT_FUSED : NTT_LM_MIN; lm <= NTT_LM_MAX && lm <= log_n && ok;
         lm = lm == NTT_FUSED ? NTT_LM_MIN : lm + 1) {
        float us = 0;
        for (int rep = 0; rep < 4 && ok; ++rep) {
            ok = hipEventRecord(e0, s) == hipSuccess;
            ntt_run_lm(ni, out, (int64_t)1 << log_n, nb, log_n, 0, fone(), 0, T, s, lm);
            ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0;
            ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (rep > 0 && (us == 0 || ms * 1000 < us)) us = ms * 1000;   // best of 3 after a warm-up
        }
        g_ntt_us[log_n][lm] = us;
        if (ok && (best_us == 0 || us < best_us)) { best_us = us; best = lm; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) return -1;
    __atomic_store_n(&g_ntt_lm[log_n], best, __ATOMIC_RELEASE);
    return 0;
}
int ntt_plan(int log_n, float *us)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (log_n < 0 || log_n >= 64) return 0;
    for (int lm = NTT_LM_MIN; lm <= NTT_LM_MAX; ++lm) us[lm - NTT_LM_MIN] = g_ntt_us[log_n][lm];
    us[NTT_LM_MAX - NTT_LM_MIN + 1] = g_ntt_us[log_n][NTT_FUSED];
    return ntt_lm_for(log_n);
}
void fri_fold_nat(const fe *in, int64_t n, const fe *r, const fe *iw, fe *out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fri_fold_nat, dim3(nblk(n)), dim3(TPB), 0, s, in, n, r, iw, out);
}
void pair_leaves(const fe *cw, int64_t half, int log_rows, int64_t cols, uint8_t *tree, hipStream_t s)
{
    if (half > 0) hipLaunchKernelGGL(k_pair_leaves, dim3(nblk(half)), dim3(TPB), 0, s, cw, half, log_rows, cols, tree);
}
void fri_fold_any(const fe *in, int64_t half, int log_rows, int64_t cols, const fe *r, const fe *tw, int64_t SS, int sh,
                  fe *out, uint8_t *leaves, hipStream_t s)
{
    if (half > 0)
        hipLaunchKernelGGL(k_fri_fold_any, dim3(nblk(half)), dim3(TPB), 0, s, in, half, log_rows, cols, r, tw, SS, sh, out,
                           leaves);
}
int brev_permute(const fe *src, int64_t len, fe *dst, int log_n, hipStream_t s)
{
    if (log_n < 12 || log_n > 40 || len > ((int64_t)1 << log_n)) return -1;
    hipLaunchKernelGGL(k_brev_tile, dim3((unsigned)((int64_t)1 << (log_n - 12))), dim3(256), 0, s, src, len, dst, log_n);
    return 0;
}
int copy_roots(const uint8_t *const *trees, int n, uint8_t *out, hipStream_t s)
{
    if (n < 1 || n > 32) return -1;
    root_src S;
    memset(&S, 0, sizeof S);
    for (int k = 0; k < n; ++k) S.p[k] = trees[k];
    hipLaunchKernelGGL(k_copy_roots, dim3(1), dim3(128), 0, s, S, n, out);
    return 0;
}
void count_diff(const fe *a, const fe *b, int64_t n, unsigned int *cnt, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_count_diff, dim3(nblk(n)), dim3(TPB), 0, s, a, b, n, cnt);
}
void twiddles(fe *tw, int64_t n, const fe *T1, const fe *T2, hipStream_t s)
{
    hipLaunchKernelGGL(k_twiddles, dim3(nblk(n)), dim3(TPB), 0, s, tw, n, T1, T2);
}
void intt(intt_src in, fe *out, int64_t out_stride, int nb, int log_n, fe inv_n, const ntt_tabs &T, hipStream_t s)
{
    ntt_in ni;
    ni.kind = in.kind; ni.src = in.src; ni.stride = in.src_stride;
    ni.l = in.l; ni.q = in.q; ni.cw_S = in.cw_S; ni.cw_stride = in.cw_stride; ni.step = in.step; ni.log_R = in.log_R;
    ni.lens = nullptr; ni.log_Rl = 0; ni.hmode = 0; ni.brev_src = in.brev_src; ni.qscale = in.qscale;
    ntt_run(ni, out, out_stride, nb, log_n, 1, inv_n, 1, T, s);
}
void lde(const fe *coef, int64_t coef_stride, const int *lens, int nb, fe *out, int log_C, int log_R, const ntt_tabs &T,
         hipStream_t s, int hmode, int brev_src)
{
    ntt_in ni;
    memset(&ni, 0, sizeof ni);
    ni.kind = 2; ni.src = coef; ni.stride = coef_stride; ni.lens = lens; ni.log_Rl = log_R; ni.hmode = hmode;
    ni.brev_src = brev_src;
    ntt_run(ni, out, (int64_t)1 << log_C, (int64_t)nb << log_R, log_C, 0, fone(), 0, T, s);
}
void slice_sums(const fe *lq_coef, int64_t stride, int nslice, int64_t real, int64_t m, fe *c0, fe *all_sum, hipStream_t s)
{
    hipLaunchKernelGGL(k_slice_sums, dim3(1), dim3(128), 0, s, lq_coef, stride, nslice, real, m, c0, all_sum);
}
void ntt(intt_src in, fe *out, int64_t out_stride, int nb, int log_n, int inverse, fe scale, int do_scale,
         const ntt_tabs &T, hipStream_t s)
{
    ntt_in ni;
    memset(&ni, 0, sizeof ni);
    ni.kind = in.kind; ni.src = in.src; ni.stride = in.src_stride;
    ni.l = in.l; ni.q = in.q; ni.cw_S = in.cw_S; ni.cw_stride = in.cw_stride; ni.step = in.step; ni.log_R = in.log_R;
    ni.brev_src = in.brev_src; ni.qscale = in.qscale;
    ntt_run(ni, out, out_stride, nb, log_n, inverse, scale, do_scale, T, s);
}
void sha3_batch(const uint8_t *in, uint8_t *out, int64_t n, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_sha3_batch, dim3(nblk(n)), dim3(TPB), 0, s, (const uint64_t *)in, (uint64_t *)out, n);
}
void sha256_batch(const uint8_t *in, uint8_t *out, int64_t n, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_sha256_batch, dim3(nblk(n)), dim3(TPB), 0, s, (const uint64_t *)in, (uint64_t *)out, n);
}
void merkle_pad(uint8_t *tree, int64_t pad, int64_t n_leaf, hipStream_t s, int hash)
{
    if (pad <= n_leaf) return;
    if (hash == HASH_SHA256)
        hipLaunchKernelGGL(k_merkle_pad<HASH_SHA256>, dim3(nblk(pad - n_leaf)), dim3(TPB), 0, s, tree, pad, n_leaf);
    else
        hipLaunchKernelGGL(k_merkle_pad<HASH_SHA3>, dim3(nblk(pad - n_leaf)), dim3(TPB), 0, s, tree, pad, n_leaf);
}
void fe_mul(const fe *a, const fe *b, fe *o, int64_t n, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_fe_mul, dim3(nblk(n)), dim3(TPB), 0, s, a, b, o, n);
}

void fri_fold(const fe *prev, int64_t S, int log_R, const fe *r, const fe *tw, int64_t SS, fe *next, hipStream_t s)
{
    hipLaunchKernelGGL(k_fri_fold, dim3(nblk(65 * (S / 2))), dim3(TPB), 0, s, prev, S, log_R, r, tw, SS, next);
}
void fri_folds_vo(const fe *l, const fe *q, const fe *cq, const fe *h, const fe *c0, int64_t real, int64_t m, int64_t SS,
                  int nlev, int log_R, const fe *r, const fe *tw, fe *const *next, hipStream_t s)
{
    // levels 0 and 1 straight from l, q, h (level 0 alone for a single
    // level); then as fri_folds
    if (nlev > 1) {
        hipLaunchKernelGGL(k_vo_fold2, dim3(nblk(65 * (SS / 4))), dim3(TPB), 0, s, l, q, cq, h, c0, 65, SS, log_R, real, m, tw,
                           r, next[0], next[1]);
        if (nlev > 2) fri_folds(next[1], SS / 4, SS, nlev - 2, log_R, r + 2, tw, next + 2, s);
        return;
    }
    hipLaunchKernelGGL(k_vo_fold, dim3(nblk(65 * (SS / 2))), dim3(TPB), 0, s, l, q, cq, h, c0, 65, SS, log_R, real, m, tw, r,
                       next[0]);
}
void fri_folds(const fe *vo, int64_t S0, int64_t SS, int nlev, int log_R, const fe *r, const fe *tw, fe *const *next,
               hipStream_t s)
{
    const fe *prev = vo;
    int64_t S = S0;
    int k = 0;
    for (; k < nlev && S > FRI_TAIL; ++k, S /= 2) {
        fri_fold(prev, S, log_R, r + k, tw, SS, next[k], s);
        prev = next[k];
    }
    if (k < nlev) {
        fri_levels F;
        memset(&F, 0, sizeof F);
        F.n = nlev - k;
        for (int q = 0; q < F.n; ++q) F.next[q] = next[k + q];
        hipLaunchKernelGGL(k_fri_tail, dim3(65), dim3(256), 0, s, prev, S, log_R, r + k, tw, SS, F);
    }
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
void leaf_hash_table(const leaf_seg *G, const int64_t *start, int nseg, int64_t total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(k_leaf_table, dim3(nblk(total)), dim3(TPB), 0, s, G, start, nseg);
}
void merkle_level_table(const mk_item *items, int n, int64_t total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(k_merkle_level_table, dim3(nblk(total)), dim3(TPB), 0, s, items, n, total);
}
void merkle_tails_table(const leaf_seg *G, int ntree, hipStream_t s)
{
    if (ntree) hipLaunchKernelGGL(k_merkle_tails_table, dim3(ntree), dim3(256), 0, s, G);
}
void merkle_multi(int ntree, uint8_t *const *tree, const int64_t *nleaf, hipStream_t s, int hash)
{
    tree_set T;
    memset(&T, 0, sizeof T);
    T.n = ntree;
    for (int k = 0; k < ntree; ++k) { T.tree[k] = tree[k]; T.nleaf[k] = nleaf[k]; }
    for (int h = 1;; ++h) {
        int64_t total = 0;
        for (int k = 0; k < ntree; ++k)
            if ((nleaf[k] >> h) > MK_TAIL) total += nleaf[k] >> h;
        if (!total) break;
        if (hash == HASH_SHA256)
            hipLaunchKernelGGL(k_merkle_multi<HASH_SHA256>, dim3(nblk(total)), dim3(TPB), 0, s, T, h, total);
        else
            hipLaunchKernelGGL(k_merkle_multi<HASH_SHA3>, dim3(nblk(total)), dim3(TPB), 0, s, T, h, total);
    }
    if (hash == HASH_SHA256)
        hipLaunchKernelGGL(k_merkle_tails<HASH_SHA256>, dim3(ntree), dim3(256), 0, s, T);
    else
        hipLaunchKernelGGL(k_merkle_tails<HASH_SHA3>, dim3(ntree), dim3(256), 0, s, T);
}
void reduce_slots_table(const slot_job *J, const int64_t *start, int n, int64_t total, hipStream_t s)
{
    if (!total) return;
    hipLaunchKernelGGL(k_reduce_slots_table, dim3((unsigned)total), dim3(256), 0, s, J, start, n);
    hipLaunchKernelGGL(k_derive_b_table, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, s, J, start, n, total);
}
void copy_table(const copy_job *J, const int64_t *start, int n, int64_t total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(k_copy_table, dim3(nblk(total)), dim3(TPB), 0, s, J, start, n, total);
}
void gather16(const uint64_t *src, int64_t n, fe *out, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_gather16, dim3(nblk(n)), dim3(TPB), 0, s, src, n, out);
}


void predicates(const layer_dev *D, int n, int ng_max, fe *part, int blocks, fe *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_predicates, dim3(blocks, n), dim3(TPB), 0, s, D, part);
    hipLaunchKernelGGL(k_predicates_sum, dim3(n), dim3(64), 0, s, part, blocks, out);
}
void poly_eval(const fe *coef, int64_t n, const fe *pts, int npts, fe *out, hipStream_t s)
{
    if (npts) hipLaunchKernelGGL(k_poly_eval, dim3(npts), dim3(TPB), 0, s, coef, n, pts, out);
}
void verify_paths(const path_job *J, int njobs, const uint64_t *dg, const fe *vals, int *ok, hipStream_t s)
{
    if (njobs) hipLaunchKernelGGL(k_verify_paths, dim3(nblk(njobs, 64)), dim3(64), 0, s, J, njobs, dg, vals, ok);
}

// ---- fft_gkr ----
void fg_eq(fe *out, const fe *r, int lg, hipStream_t s)
{
    hipLaunchKernelGGL(k_fg_eq, dim3(nblk((int64_t)1 << lg)), dim3(TPB), 0, s, out, r, lg);
}
void fg_pow(const fe *pts, int npts, int lo, int hi, fe *plo, fe *phi, hipStream_t s)
{
    int64_t n = ((int64_t)npts << lo) + ((int64_t)npts << hi);
    hipLaunchKernelGGL(k_fg_pow, dim3(nblk(n)), dim3(TPB), 0, s, pts, npts, lo, hi, plo, phi);
}
void fg_butterfly(const fe *pre, fe *cur, int dep, int lg, fg_pw_t pw, fe scale, int do_scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_fg_butterfly, dim3(nblk((int64_t)1 << (lg - 1))), dim3(TPB), 0, s, pre, cur, dep, lg, pw.plo,
                       pw.phi, pw.lo, pw.hi, pw.tw_pt, scale, do_scale);
}
void fg_eval(const fe *coef, fg_pw_t pw, int lg, fe *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_fg_eval, dim3(nblk((int64_t)64 << lg)), dim3(TPB), 0, s, coef, pw.plo, pw.phi, pw.lo, pw.hi,
                       lg, out);
}
void fg_rowsum(const fe *ev, int lg, fe *outs, hipStream_t s)
{
    hipLaunchKernelGGL(k_fg_rowsum, dim3(64), dim3(TPB), 0, s, ev, lg, outs);
}
void fg_add_tables(beta_g bg, int lg, fe *amb, fe *aVb, hipStream_t s)
{
    bg_t b = {bg.g0f, bg.g0s, bg.g1f, bg.g1s, bg.fh};
    hipLaunchKernelGGL(k_fg_add_tables, dim3(nblk((int64_t)64 << lg)), dim3(TPB), 0, s, b, lg, amb, aVb);
}
void fg_mult_tables(beta_g bg, fg_pw_t pw, int lg, fe *amb, fe *aVb, hipStream_t s)
{
    bg_t b = {bg.g0f, bg.g0s, bg.g1f, bg.g1s, bg.fh};
    hipLaunchKernelGGL(k_fg_mult_tables, dim3(nblk((int64_t)1 << lg)), dim3(TPB), 0, s, b, pw.plo, pw.phi, pw.lo,
                       pw.hi, lg, amb, aVb);
}
void fg_ifft_tables(int phase, const fg_layer_dev *D, int nlayer, fg_pw_t pw, int lg, hipStream_t s)
{
    dim3 grid(nblk((int64_t)1 << lg), nlayer);
    if (phase == 0)
        hipLaunchKernelGGL(k_fg_ifft_p1m, grid, dim3(TPB), 0, s, D, pw.plo, pw.phi, pw.lo, pw.hi, pw.tw_pt, lg);
    else
        hipLaunchKernelGGL(k_fg_ifft_p2m, grid, dim3(TPB), 0, s, D, pw.plo, pw.phi, pw.lo, pw.hi, pw.tw_pt, lg);
}
}  // namespace launch
}  // namespace vg
