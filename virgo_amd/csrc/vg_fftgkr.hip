// Device schedule of fft_gkr (src/VPD/fft_circuit_GKR.cpp:22-845).
//
// Everything the program computes is fixed by lg and its own rand() stream
// (seed 1), so all challenges are drawn on the host first and the whole
// proof -- circuit, every sumcheck of the addition, mult and lg inverse-DFT
// layers -- is one stream of kernels with a single upload and a single copy
// back.  The round polynomials come from the same block-fold kernels as the
// main prover (these layers have no zero-knowledge masks).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vg_fftgkr.hpp"
#include "vg_host.hpp"
#include "vg_kernels.hpp"

void vg_set_error(const std::string &msg);

namespace vg {

#define FG_FAIL(code, ...) do { char b_[512]; snprintf(b_, sizeof b_, __VA_ARGS__); vg_set_error(b_); return (code); } while (0)
#define FG_CHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) FG_FAIL(-20, "%s: %s", #x, hipGetErrorString(e_)); } while (0)

namespace {

// challenge offsets of one GKR layer of the program
struct fg_layer {
    int64_t ru, rv;         // this layer's r_u, r_v
    int64_t r0, r1;         // beta_g points (previous layer's r_u, r_v)
    int64_t alpha, beta;    // beta_g weights
    int64_t ab_next;        // alpha, beta drawn after an inverse-DFT layer (-1: none)
    int len_g;              // beta_g bits
    int bits;               // sumcheck rounds per phase
    int64_t msg1, msg2;     // message offsets: 3 bits sums + 5 end values per phase
    int hd;                 // first half-table descriptor (g0f g0s g1f g1s [uf us])
};

}  // namespace

struct fg_ctx {
    int lg = 0;
    int64_t N = 0;
    hipStream_t st = nullptr;
    bool own_stream = false;
    mem_source *mem = nullptr;       // where arena, h_chal and h_msg came from
    uint8_t *arena = nullptr;
    int64_t bytes = 0;
    // challenges: [0] one, [1] zero, r[lg], pts[64] + inv_rou, r0, r1, layers
    int64_t n_chal = 0, c_r = 0, c_pts = 0, c_r0 = 0, c_r1 = 0;
    std::vector<fg_layer> L;        // 0: addition, 1: mult, 2 + dep: inverse DFT
    fe *chal = nullptr, *h_chal = nullptr;
    // circuit: lay[t] = layer lg + t of the program (t < lg), lay[lg] = its
    // layer 2lg+1 (the scaled coefficients); ev = the 64 x N layer
    std::vector<fe *> lay;
    fe *ev = nullptr, *outs = nullptr, *mle_t[2] = {};
    fe *plo = nullptr, *phi = nullptr;
    int lo = 0, hi = 0;
    fe inv_n;
    // sumcheck: instance 0 = addition, 1 = mult, 2 + dep = inverse-DFT phase
    // 1, 2 + lg + dep = its phase 2 (the order of the slot rounds too)
    fe *amb = nullptr, *aVb = nullptr;      // addition / mult tables (64 N)
    std::vector<fe *> amb_d, aVb_d;         // per inverse-DFT layer (N)
    std::vector<fe *> ws;                   // 6 chunk-result tables per instance
    std::vector<int> slot0;                 // first slot round per instance
    launch::fold_inst *inst = nullptr;
    launch::fg_layer_dev *ldev = nullptr;
    fe *slots = nullptr;
    int *slot_nb = nullptr;
    int64_t *slot_dst = nullptr;
    int64_t *slot_rof = nullptr;               // offset of r[j-1] in chal; -1 for round 0
    std::vector<int> h_slot_nb;
    std::vector<int64_t> h_slot_dst, h_slot_rof;
    fe *ht = nullptr;
    launch::half_desc *hdesc = nullptr;
    std::vector<launch::half_desc> h_hdesc;
    int64_t ht_total = 0;
    fe *msg = nullptr, *h_msg = nullptr;
    uint8_t *h_up = nullptr;          // page-locked upload staging (when the source stages)
    int64_t up_bytes = 0;
    int64_t msg_len = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool pending = false;
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

// challenge and message layout; half-table descriptors (relative offsets)
void plan(fg_ctx *c)
{
    int lg = c->lg;
    int64_t k = 2;
    c->c_r = k; k += lg;
    c->c_pts = k; k += 65;
    c->c_r0 = k; k += lg + 10;
    c->c_r1 = k; k += lg + 10;
    int64_t m = 1;                       // msg[0] = claim on the outputs
    c->L.clear();
    auto add_layer = [&](int len_g, int bits, int64_t r0, int64_t r1, int64_t a, int64_t b, bool ifft) {
        fg_layer l;
        l.ru = k; k += bits;
        l.rv = k; k += bits;
        l.r0 = r0; l.r1 = r1; l.alpha = a; l.beta = b;
        l.ab_next = -1;
        if (ifft) { l.ab_next = k; k += 2; }
        l.len_g = len_g; l.bits = bits;
        l.msg1 = m; m += 3 * bits + 5;
        l.msg2 = -1;
        if (ifft) { l.msg2 = m; m += 3 * bits + 5; }
        l.hd = 0;
        c->L.push_back(l);
    };
    // alpha = 1, beta = 0 until the first inverse-DFT layer draws new ones
    add_layer(6, lg + 6, c->c_r0, c->c_r1, 0, 1, false);
    add_layer(lg + 6, lg, c->L[0].ru, c->L[0].rv, 0, 1, false);
    for (int dep = 0; dep < lg; ++dep) {
        const fg_layer &p = c->L.back();
        int64_t a = dep == 0 ? 0 : p.ab_next, b = dep == 0 ? 1 : p.ab_next + 1;
        add_layer(lg, lg, p.ru, p.rv, a, b, true);
    }
    c->n_chal = k;
    c->msg_len = m;
    // slot rounds of the multi-block fold launches, instance by instance
    c->h_slot_nb.clear(); c->h_slot_dst.clear(); c->h_slot_rof.clear(); c->slot0.clear();
    int ninst = 2 + 2 * lg;
    for (int t = 0; t < ninst; ++t) {
        const fg_layer &l = c->L[t < 2 + lg ? t : t - lg];
        int ph = t >= 2 + lg;
        int js[32], ks[32];
        int nl = launch::fold_plan(l.bits, js, ks);
        int64_t sums = ph == 0 ? l.msg1 : l.msg2;
        c->slot0.push_back((int)c->h_slot_nb.size());
        for (int q = 0; q < nl; ++q)
            for (int r = 0; r < ks[q]; ++r) {
                c->h_slot_nb.push_back(launch::fold_blocks(l.bits, js[q]));
                int j = js[q] + r;
                c->h_slot_dst.push_back(sums + 3 * j);
                c->h_slot_rof.push_back(j == 0 ? -1 : (ph == 0 ? l.ru : l.rv) + j - 1);
            }
    }
    // half tables: beta_g halves (weight alpha / beta) and, for the
    // inverse-DFT layers, beta_u halves
    c->h_hdesc.clear();
    int64_t tot = 0;
    auto half = [&](int64_t r, int nbits, int64_t f0) {
        launch::half_desc d;
        d.out = (fe *)(intptr_t)tot;     // relative until the arena exists
        d.f0 = (const fe *)(intptr_t)f0; // challenge index or -1
        d.r = (const fe *)(intptr_t)r;
        d.nbits = nbits;
        d.start = tot;
        tot += (int64_t)1 << nbits;
        c->h_hdesc.push_back(d);
    };
    for (auto &l : c->L) {
        l.hd = (int)c->h_hdesc.size();
        int fh = l.len_g >> 1, sh = l.len_g - fh;
        half(l.r0, fh, l.alpha); half(l.r0 + fh, sh, -1);
        half(l.r1, fh, l.beta); half(l.r1 + fh, sh, -1);
        if (l.msg2 >= 0) { half(l.ru, fh, -1); half(l.ru + fh, sh, -1); }
    }
    c->ht_total = tot;
}

int64_t layout(fg_ctx *c, uint8_t *base)
{
    int64_t off = 0, N = c->N;
    c->chal = carve<fe>(base, off, c->n_chal);
    c->lay.assign(c->lg + 1, nullptr);
    for (int t = 0; t <= c->lg; ++t) c->lay[t] = carve<fe>(base, off, N);
    c->ev = carve<fe>(base, off, 64 * N);
    c->outs = carve<fe>(base, off, 64);
    c->mle_t[0] = carve<fe>(base, off, 64);
    c->mle_t[1] = carve<fe>(base, off, 64);
    c->plo = carve<fe>(base, off, (int64_t)65 << c->lo);
    c->phi = carve<fe>(base, off, (int64_t)65 << c->hi);
    c->amb = carve<fe>(base, off, 64 * N);
    c->aVb = carve<fe>(base, off, 64 * N);
    c->amb_d.assign(c->lg, nullptr); c->aVb_d.assign(c->lg, nullptr);
    for (int d = 0; d < c->lg; ++d) { c->amb_d[d] = carve<fe>(base, off, N); c->aVb_d[d] = carve<fe>(base, off, N); }
    int ninst = 2 + 2 * c->lg;
    c->ws.assign(6 * ninst, nullptr);
    for (int t = 0; t < ninst; ++t)
        for (int k = 0; k < 6; ++k)
            c->ws[6 * t + k] = carve<fe>(base, off, launch::fold_ws_len(t == 0 ? c->lg + 6 : c->lg));
    c->inst = carve<launch::fold_inst>(base, off, ninst);
    c->ldev = carve<launch::fg_layer_dev>(base, off, c->lg);
    int ns = (int)c->h_slot_nb.size();
    c->slots = carve<fe>(base, off, (int64_t)(ns > 0 ? ns : 1) * 3 * launch::FW_BLOCKS);
    c->slot_nb = carve<int>(base, off, ns + 1);
    c->slot_dst = carve<int64_t>(base, off, ns + 1);
    c->slot_rof = carve<int64_t>(base, off, ns + 1);
    c->ht = carve<fe>(base, off, c->ht_total);
    c->hdesc = carve<launch::half_desc>(base, off, (int64_t)c->h_hdesc.size());
    c->msg = carve<fe>(base, off, c->msg_len);
    return off;
}

void draw(fg_ctx *c)
{
    fe *h = c->h_chal;
    int lg = c->lg;
    glibc_rand g(1);
    h[0] = fone(); h[1] = fzero();
    for (int i = 0; i < lg; ++i) h[c->c_r + i] = g.field();             // main :834-836
    for (int i = 0; i < 64; ++i) h[c->c_pts + i] = g.field();           // build_circuit :88-91
    h[c->c_pts + 64] = finv(root_of_unity(lg));
    for (int i = 0; i < lg + 10; ++i) h[c->c_r0 + i] = g.field();       // engage_gkr :771-772
    for (int i = 0; i < lg + 10; ++i) h[c->c_r1 + i] = g.field();
    for (auto &l : c->L) {                                               // refresh_randomness per layer
        for (int i = 0; i < l.bits; ++i) h[l.ru + i] = g.field();
        for (int i = 0; i < l.bits; ++i) h[l.rv + i] = g.field();
        if (l.ab_next >= 0) { h[l.ab_next] = g.field(); h[l.ab_next + 1] = g.field(); }   // :736-737
    }
}

launch::beta_g bg_of(const fg_ctx *c, const fg_layer &l)
{
    const fe *t[4];
    for (int k = 0; k < 4; ++k) t[k] = c->ht + c->h_hdesc[l.hd + k].start;
    return {t[0], t[1], t[2], t[3], l.len_g >> 1};
}

// the program's inline checks (fft_circuit_GKR.cpp:327-352, :456-484,
// :677-731 and the round sums) on the message block M
bool check(const fg_ctx *c, const fe *M)
{
    const fe *h = c->h_chal;
    int lg = c->lg;
    fe one = fone();
    auto q_eval = [](const fe *q, const fe &x) { return add(mul(add(mul(q[0], x), q[1]), x), q[2]); };
    auto rounds = [&](int64_t off, const fe *rr, int bits, fe &claim) {
        bool ok = true;
        for (int i = 0; i < bits; ++i) {
            const fe *q = M + off + 3 * i;
            if (!feq(add(q_eval(q, fzero()), q_eval(q, one)), claim)) ok = false;
            claim = q_eval(q, rr[i]);
        }
        return ok;
    };
    auto eqb = [&](const fe *r, int nbits, int64_t x, fe w) {
        for (int i = 0; i < nbits; ++i) w = mul(w, ((x >> i) & 1) ? r[i] : sub(one, r[i]));
        return w;
    };
    bool ok = true;
    fe claim = M[0];
    fe alpha = one, beta = fzero();
    std::vector<fe> r0(h + c->c_r0, h + c->c_r0 + lg + 10), r1(h + c->c_r1, h + c->c_r1 + lg + 10);
    {   // addition layer
        const fg_layer &l = c->L[0];
        int log_uv = lg + 6;
        ok &= rounds(l.msg1, h + l.ru, log_uv, claim);
        fe v_u = M[l.msg1 + 3 * log_uv + 4], sv = fzero();
        for (int i = 0; i < 64; ++i) {
            fe tg = add(eqb(r0.data(), 6, i, alpha), eqb(r1.data(), 6, i, beta));
            sv = add(sv, mul(tg, eqb(h + l.ru + log_uv - 6, 6, i, one)));
        }
        for (int i = 0; i < log_uv; ++i) { r0[i] = h[l.ru + i]; r1[i] = h[l.rv + i]; }
        if (!feq(claim, mul(sv, v_u))) ok = false;
        claim = mul(alpha, v_u);
    }
    {   // mult layer
        const fg_layer &l = c->L[1];
        const fe *ru = h + l.ru;
        ok &= rounds(l.msg1, ru, lg, claim);
        fe v_u = M[l.msg1 + 3 * lg + 4], sm = fzero();
        for (int i = 0; i < 64; ++i) {
            fe tg0 = eqb(r0.data() + lg, 6, i, alpha), tg1 = eqb(r1.data() + lg, 6, i, beta);
            fe tu0 = one, tu1 = one, x = h[c->c_pts + i];
            for (int j = 0; j < lg; ++j) {
                tu0 = mul(tu0, add(mul(mul(r0[j], ru[j]), x), mul(sub(one, r0[j]), sub(one, ru[j]))));
                tu1 = mul(tu1, add(mul(mul(r1[j], ru[j]), x), mul(sub(one, r1[j]), sub(one, ru[j]))));
                x = mul(x, x);
            }
            sm = add(add(sm, mul(tg0, tu0)), mul(tg1, tu1));
        }
        for (int i = 0; i < lg; ++i) { r0[i] = ru[i]; r1[i] = h[l.rv + i]; }
        if (!feq(claim, mul(sm, v_u))) ok = false;
        claim = mul(alpha, v_u);
    }
    claim = mul(claim, from_u64((uint64_t)1 << lg));     // intermediate_layer :497-502
    fe rot = h[c->c_pts + 64];
    for (int dep = 0; dep < lg; ++dep) {                 // inverse-DFT layers
        const fg_layer &l = c->L[2 + dep];
        const fe *ru = h + l.ru, *rv = h + l.rv;
        ok &= rounds(l.msg1, ru, lg, claim);
        fe v_u = M[l.msg1 + 3 * lg + 4];
        ok &= rounds(l.msg2, rv, lg, claim);
        fe v_v = M[l.msg2 + 3 * lg + 4];
        fe xr = rot;
        for (int i = 0; i < dep; ++i) xr = mul(xr, xr);
        int log_k = lg - dep - 1, log_j = dep;
        fe uv = mul(sub(one, ru[log_j]), rv[log_j]);
        fe u0A = mul(mul(sub(one, r0[lg - 1]), uv), alpha), u1A = mul(mul(sub(one, r1[lg - 1]), uv), beta);
        fe v0A = u0A, v1A = u1A;
        fe u0B = mul(mul(r0[lg - 1], uv), alpha), u1B = mul(mul(r1[lg - 1], uv), beta);
        fe v0B = u0B, v1B = u1B;
        for (int i = 0; i < log_k; ++i) {
            fe a0 = r0[log_j + i], a1 = r1[log_j + i], bu = ru[log_j + 1 + i], bv = rv[log_j + 1 + i];
            fe nn = mul(sub(one, bu), sub(one, bv)), pp = mul(bu, bv);
            fe t0 = add(mul(a0, pp), mul(sub(one, a0), nn)), t1 = add(mul(a1, pp), mul(sub(one, a1), nn));
            fe t0x = add(mul(mul(a0, pp), xr), mul(sub(one, a0), nn));
            fe t1x = add(mul(mul(a1, pp), xr), mul(sub(one, a1), nn));
            u0A = mul(u0A, t0); u1A = mul(u1A, t1); v0A = mul(v0A, t0x); v1A = mul(v1A, t1x);
            u0B = mul(u0B, t0); u1B = mul(u1B, t1); v0B = mul(v0B, t0x); v1B = mul(v1B, t1x);
            xr = mul(xr, xr);
        }
        for (int i = 0; i < log_j; ++i) {
            fe nn = mul(sub(one, ru[i]), sub(one, rv[i])), pp = mul(ru[i], rv[i]);
            fe t0 = add(mul(r0[i], pp), mul(sub(one, r0[i]), nn)), t1 = add(mul(r1[i], pp), mul(sub(one, r1[i]), nn));
            u0A = mul(u0A, t0); u1A = mul(u1A, t1); v0A = mul(v0A, t0); v1A = mul(v1A, t1);
            u0B = mul(u0B, t0); u1B = mul(u1B, t1); v0B = mul(v0B, t0); v1B = mul(v1B, t1);
        }
        fe want = add(mul(add(add(add(u0A, u1A), u0B), u1B), v_u), mul(sub(sub(add(v0A, v1A), v0B), v1B), v_v));
        if (!feq(claim, want)) ok = false;
        for (int i = 0; i < lg; ++i) { r0[i] = ru[i]; r1[i] = rv[i]; }
        alpha = h[l.ab_next];
        beta = h[l.ab_next + 1];
        claim = add(mul(alpha, v_u), mul(beta, v_v));
    }
    return ok;
}

}  // namespace

int fg_plan(fg_ctx **out, int lg, hipStream_t stream, int64_t *dev_bytes, int64_t *pin_bytes, bool host_only)
{
    *out = nullptr;
    if (lg < 1 || lg > 20) FG_FAIL(-1, "fft_gkr: lg = %d out of range [1, 20]", lg);
    fg_ctx *c = new fg_ctx();
    c->lg = lg;
    c->N = (int64_t)1 << lg;
    c->lo = (lg + 1) / 2;
    c->hi = lg - c->lo + 1;            // exponents up to 2^lg - 1 (and 2^(lg-1) for twiddles)
    c->inv_n = finv(from_u64((uint64_t)c->N));
    plan(c);
    if (stream || host_only) c->st = stream;
    else {
        if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; FG_FAIL(-20, "stream"); }
        c->own_stream = true;
    }
    c->bytes = layout(c, nullptr);
    {
        // fg_alloc's uploads, put by put
        auto sb = uploader::stage_bytes;
        int lg2 = 2 + 2 * lg;
        c->up_bytes = sb((int64_t)sizeof(launch::fold_inst) * lg2) + sb((int64_t)sizeof(launch::fg_layer_dev) * lg) +
                      sb((int64_t)sizeof(launch::half_desc) * (int64_t)c->h_hdesc.size()) +
                      sb((int64_t)sizeof(int) * (int64_t)c->h_slot_nb.size()) +
                      2 * sb((int64_t)sizeof(int64_t) * (int64_t)c->h_slot_dst.size());
    }
    if (dev_bytes) *dev_bytes = c->bytes;
    if (pin_bytes) *pin_bytes = (int64_t)sizeof(fe) * (c->n_chal + c->msg_len) + 512 + c->up_bytes;
    *out = c;
    return 0;
}

int fg_alloc(fg_ctx *c, mem_source *mem)
{
    int lg = c->lg;
    c->mem = mem;
    c->arena = (uint8_t *)mem->device(c->bytes);
    if (!c->arena) FG_FAIL(-21, "fft_gkr: cannot allocate %lld bytes", (long long)c->bytes);
    layout(c, c->arena);
    c->h_chal = (fe *)mem->pinned(sizeof(fe) * c->n_chal);
    c->h_msg = (fe *)mem->pinned(sizeof(fe) * c->msg_len);
    if (!c->h_chal || !c->h_msg) FG_FAIL(-21, "fft_gkr: pinned host allocation failed");
    // resolve the half-table descriptors and upload the static tables
    std::vector<launch::half_desc> d = c->h_hdesc;
    for (auto &x : d) {
        x.out = c->ht + (intptr_t)x.out;
        int64_t f0 = (int64_t)(intptr_t)x.f0;
        x.f0 = f0 < 0 ? nullptr : c->chal + f0;
        x.r = c->chal + (intptr_t)x.r;
    }
    c->h_hdesc = d;
    // static per-instance descriptors (every pointer is fixed by the arena)
    std::vector<launch::fold_inst> I(2 + 2 * lg);
    for (int t = 0; t < 2 + 2 * lg; ++t) {
        const fg_layer &l = c->L[t < 2 + lg ? t : t - lg];
        int ph = t >= 2 + lg, dep = t < 2 + lg ? t - 2 : t - 2 - lg;
        launch::fold_inst &x = I[t];
        if (t == 0) { x.Vb = c->ev; x.aVb = c->aVb; x.amb = c->amb; }
        else if (t == 1) { x.Vb = c->lay[lg]; x.aVb = c->aVb; x.amb = c->amb; }
        else { x.Vb = c->lay[lg - dep - 1]; x.aVb = c->aVb_d[dep]; x.amb = c->amb_d[dep]; }
        for (int k = 0; k < 2; ++k) { x.V[k] = c->ws[6 * t + 3 * k]; x.aV[k] = c->ws[6 * t + 3 * k + 1]; x.am[k] = c->ws[6 * t + 3 * k + 2]; }
        int64_t m = ph == 0 ? l.msg1 : l.msg2;
        x.r = c->chal + (ph == 0 ? l.ru : l.rv);
        x.slots = c->slots + (int64_t)c->slot0[t] * 3 * launch::FW_BLOCKS;
        x.sums = c->msg + m;
        x.end = c->msg + m + 3 * l.bits;
        x.vconst = nullptr;
    }
    std::vector<launch::fg_layer_dev> LD(lg);
    for (int dep = 0; dep < lg; ++dep) {
        const fg_layer &l = c->L[2 + dep];
        LD[dep].bg = bg_of(c, l);
        LD[dep].bu = {c->ht + c->h_hdesc[l.hd + 4].start, c->ht + c->h_hdesc[l.hd + 5].start, lg >> 1};
        LD[dep].pre = c->lay[lg - dep - 1];
        LD[dep].amb = c->amb_d[dep];
        LD[dep].aVb = c->aVb_d[dep];
        LD[dep].vu = c->msg + l.msg1 + 3 * lg + 4;
        LD[dep].dep = dep;
    }
    uploader U;
    if (!U.init(mem, c->up_bytes, c->st)) FG_FAIL(-21, "fft_gkr: pinned upload staging failed");
    c->h_up = U.stage;
    bool ok = U.put(c->inst, I.data(), (int64_t)sizeof(launch::fold_inst) * (int64_t)I.size()) == hipSuccess &&
              U.put(c->ldev, LD.data(), (int64_t)sizeof(launch::fg_layer_dev) * (int64_t)LD.size()) == hipSuccess;
    ok = ok && U.put(c->hdesc, d.data(), (int64_t)sizeof(launch::half_desc) * (int64_t)d.size()) == hipSuccess;
    if (ok && !c->h_slot_nb.empty())
        ok = U.put(c->slot_nb, c->h_slot_nb.data(), (int64_t)sizeof(int) * (int64_t)c->h_slot_nb.size()) == hipSuccess &&
             U.put(c->slot_dst, c->h_slot_dst.data(), (int64_t)sizeof(int64_t) * (int64_t)c->h_slot_dst.size()) == hipSuccess &&
             U.put(c->slot_rof, c->h_slot_rof.data(), (int64_t)sizeof(int64_t) * (int64_t)c->h_slot_rof.size()) == hipSuccess;
    ok = ok && U.flush() == hipSuccess;
    if (!ok || hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess)
        FG_FAIL(-20, "fft_gkr: setup copies failed");
    return 0;
}

int fg_create(fg_ctx **out, int lg, hipStream_t stream)
{
    int rc = fg_plan(out, lg, stream, nullptr, nullptr);
    if (rc) return rc;
    rc = fg_alloc(*out, hip_mem());
    if (rc) { fg_destroy(*out); *out = nullptr; }
    return rc;
}

void fg_destroy(fg_ctx *c)
{
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->mem) {
        if (c->arena) c->mem->device_free(c->arena);
        if (c->h_chal) c->mem->pinned_free(c->h_chal);
        if (c->h_msg) c->mem->pinned_free(c->h_msg);
        if (c->h_up) c->mem->pinned_free(c->h_up);
    }
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->own_stream && c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int64_t fg_device_bytes(const fg_ctx *c) { return c ? c->bytes : 0; }

int fg_enqueue(fg_ctx *c)
{
    if (c->pending) FG_FAIL(-3, "fft_gkr: a proof is already in flight on this context");
    draw(c);
    hipStream_t s = c->st;
    int lg = c->lg;
    FG_CHK(hipEventRecord(c->e0, s));
    FG_CHK(hipMemcpyAsync(c->chal, c->h_chal, sizeof(fe) * c->n_chal, hipMemcpyHostToDevice, s));
    launch::half_multi(c->hdesc, (int)c->h_hdesc.size(), c->ht_total, s);
    // circuit (build_circuit :22-112)
    launch::fg_pw_t pw = {c->plo, c->phi, c->lo, c->hi, 64};
    launch::fg_pow(c->chal + c->c_pts, 65, c->lo, c->hi, c->plo, c->phi, s);
    launch::fg_eq(c->lay[0], c->chal + c->c_r, lg, s);
    for (int dep = lg - 1; dep >= 0; --dep)
        launch::fg_butterfly(c->lay[lg - dep - 1], c->lay[lg - dep], dep, lg, pw, c->inv_n, dep == 0, s);
    launch::fg_eval(c->lay[lg], pw, lg, c->ev, s);
    launch::fg_rowsum(c->ev, lg, c->outs, s);
    launch::mle_eval(c->outs, 6, c->chal + c->c_r0, c->mle_t[0], c->mle_t[1], c->msg, s);   // V_output :134-150
    // sumchecks: addition, mult, then all inverse-DFT layers at once
    int slot = 0;
    launch::fg_add_tables(bg_of(c, c->L[0]), lg, c->amb, c->aVb, s);
    slot += launch::sumcheck_multi(c->inst, 1, lg + 6, s);
    launch::fg_mult_tables(bg_of(c, c->L[1]), pw, lg, c->amb, c->aVb, s);
    slot += launch::sumcheck_multi(c->inst + 1, 1, lg, s);
    launch::fg_ifft_tables(0, c->ldev, lg, pw, lg, s);
    slot += lg * launch::sumcheck_multi(c->inst + 2, lg, lg, s);
    launch::fg_ifft_tables(1, c->ldev, lg, pw, lg, s);
    slot += lg * launch::sumcheck_multi(c->inst + 2 + lg, lg, lg, s);
    if (slot != (int)c->h_slot_nb.size()) FG_FAIL(-30, "fft_gkr: fold slot plan mismatch");
    launch::reduce_slots(c->slots, slot, c->slot_nb, c->slot_dst, c->slot_rof, c->chal, c->msg, s);
    FG_CHK(hipMemcpyAsync(c->h_msg, c->msg, sizeof(fe) * c->msg_len, hipMemcpyDeviceToHost, s));
    FG_CHK(hipEventRecord(c->e1, s));
    FG_CHK(hipGetLastError());
    c->pending = true;
    return 0;
}

hipEvent_t fg_done_event(const fg_ctx *c) { return c->e1; }

int fg_finish(fg_ctx *c, tbuf &out, int *accepted, int *proof_size, double *p_ms)
{
    if (!c->pending) FG_FAIL(-3, "fft_gkr: no proof in flight");
    c->pending = false;
    FG_CHK(hipEventSynchronize(c->e1));
    int lg = c->lg;
    tbuf t;
    t.reserve((size_t)(16 * c->msg_len) + 4096);
    t.put("VGF1", 4);
    t.u32((uint32_t)lg);
    t.f(c->h_msg[0]);
    for (auto &l : c->L) {
        for (int ph = 0; ph < (l.msg2 >= 0 ? 2 : 1); ++ph) {
            int64_t m = ph == 0 ? l.msg1 : l.msg2;
            for (int i = 0; i < 3 * l.bits; ++i) t.f(c->h_msg[m + i]);
            t.f(c->h_msg[m + 3 * l.bits + 4]);           // v_u or v_v
        }
    }
    out.append(t);
    if (accepted) *accepted = check(c, c->h_msg) ? 1 : 0;
    if (proof_size) {                                    // sizeof(quadratic_poly) per round (:342 and like)
        int q = (lg + 6) + lg + 2 * lg * lg;
        for (int i = 1; i <= lg; ++i) q += i;            // extension_gkr :745-754
        *proof_size = 48 * q;
    }
    if (p_ms) {
        float ms = 0.f;
        FG_CHK(hipEventElapsedTime(&ms, c->e0, c->e1));
        *p_ms = ms;
    }
    return 0;
}

int fg_verify(fg_ctx *c, const uint8_t *sec, size_t len, int *accepted)
{
    *accepted = 0;
    if (c->pending) FG_FAIL(-3, "fft_gkr: a proof is in flight");
    draw(c);
    std::vector<fe> M(c->msg_len, fzero());
    size_t off = 0;
    bool bad = false;
    auto get = [&](void *x, size_t k) {
        if (off + k > len) { bad = true; memset(x, 0, k); off = len; return; }
        memcpy(x, sec + off, k);
        off += k;
    };
    auto f = [&]() {                              // canonical elements only
        fe x;
        get(&x, 16);
        if (x.real >= P || x.img >= P) bad = true;
        return x;
    };
    char magic[4];
    uint32_t lg = 0;
    get(magic, 4);
    get(&lg, 4);
    if (bad || memcmp(magic, "VGF1", 4) || lg != (uint32_t)c->lg) return 0;
    M[0] = f();
    for (auto &l : c->L)
        for (int ph = 0; ph < (l.msg2 >= 0 ? 2 : 1); ++ph) {       // fg_finish's order
            int64_t m = ph == 0 ? l.msg1 : l.msg2;
            for (int i = 0; i < 3 * l.bits; ++i) M[m + i] = f();
            M[m + 3 * l.bits + 4] = f();
        }
    if (bad || off != len) return 0;
    *accepted = check(c, M.data()) ? 1 : 0;
    return 0;
}

}  // namespace vg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int vg_fft_gkr(int lg, uint8_t **section, size_t *len, int *accepted, int *proof_size, double *p_ms)
{
    if (!section || !len) { vg_set_error("vg_fft_gkr: null output"); return -1; }
    *section = nullptr; *len = 0;
    vg::fg_ctx *c = nullptr;
    int rc = vg::fg_create(&c, lg, nullptr);
    if (rc) return rc;
    vg::tbuf out;
    rc = vg::fg_enqueue(c);
    if (!rc) rc = vg::fg_finish(c, out, accepted, proof_size, p_ms);
    vg::fg_destroy(c);
    if (rc) return rc;
    *len = out.size();
    *section = out.release();
    if (!*section) { vg_set_error("vg_fft_gkr: out of host memory"); return -21; }
    return 0;
}
