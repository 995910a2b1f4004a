/*
 * vg_oracle.c -- CPU oracle (test infrastructure only) for the Virgo prover.
 *
 * Plain single-threaded C restatement of the reference prover and verifier.
 * It is the checker for the HIP product and the CPU baseline timer; nothing
 * in virgo_amd/ links or calls it.  Citations are relative to the reference
 * tree (sunblaze-ucb/Virgo).  See vg_oracle.h for how each part is pinned.
 *
 * Transcript layout (the reference has no serializer; this byte order follows
 * the order in which zk_verifier.cpp:910-1253 and vpd_verifier.cpp:80-335
 * consume prover messages).  F = 16 bytes {u64 img, u64 real} little-endian,
 * D = 32-byte digest.
 *   "VGT1", u32 total_depth, u32 n_input_bits, u32 bit_length[total_depth]
 *   F V_res(output, r_0)                                  zk_verifier.cpp:942
 *   per layer i = total_depth-1 .. 1, bl = bit_length[i-1]:
 *     F maskpoly_sumc                                     zk_verifier.cpp:963
 *     phase 1: (bl-1) x {F a,b,c} then {F a,b,c,d,e,f}    :1037-1074
 *     F v_u                                               :1088
 *     phase 2: (bl-1) x {F a,b,c} then {F a,b,c,d,e,f}    :1079-1117
 *     F v_v                                               :1123-1127
 *     F finalround a,b,c                                  :1149
 *     F query, queryRg1, queryRg2                         :1167-1169
 *     F Iuv, preZu, preZv                                 :1175
 *   F Zu, F preu1, F sumRc.a, sumRc.b, sumRc.c            :1203-1207
 *   D root_L, D root_H, F all_sum[65]                     :1212-1216
 *   D fri_root[n-6], F final_codeword[2048], F final_msk[32]  vpd_verifier.cpp:46-78,315-332
 *   33 x { 65 x (F,F) + (n-1) D   (L init query, fri.cpp:148-201)
 *          65 x (F,F) + (n-1) D   (H init query)
 *          for lvl < n-6: 65 x (F,F) + (n-2-lvl) D  (fri.cpp:227-289) }
 *   fft_gkr section for lg = n-6 ("VGF1", see vgo_fft_gkr below;
 *   vpd_verifier.cpp:90-99 runs that program during the commitment)
 */
#include "vg_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>

#define P61 0x1FFFFFFFFFFFFFFFULL
typedef unsigned __int128 u128;
typedef vgo_fe fe;
typedef vgo_digest dig;

static char g_err[512];
const char *vgo_last_error(void) { return g_err; }
#define FAIL(code, ...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); return (code); } while (0)

static double now_s(void)
{
    struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* ===================================================================== */
/* F_{p^2} (prime_field.h:28-176, prime_field.cpp:60-92).  Every operator */
/* returns the canonical representative in [0,p), so any exact algorithm  */
/* gives the same bits as the reference's mulx/AVX2 code.                 */
/* ===================================================================== */
static inline uint64_t red(u128 x)            /* x < 2^125 */
{
    uint64_t lo = (uint64_t)x & P61;
    uint64_t hi = (uint64_t)(x >> 61);
    uint64_t r = lo + (hi & P61) + (hi >> 61);
    r = (r & P61) + (r >> 61);
    if (r >= P61) r -= P61;
    return r;
}
static inline fe F(uint64_t real, uint64_t img) { fe r; r.real = real; r.img = img; return r; }
static inline fe fe_u(uint64_t x) { return F(x % P61, 0); }          /* prime_field.h:108 */
static inline uint64_t addp(uint64_t a, uint64_t b) { uint64_t s = a + b; return s >= P61 ? s - P61 : s; }
static inline uint64_t subp(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P61 - b; }
static inline fe add(fe a, fe b) { return F(addp(a.real, b.real), addp(a.img, b.img)); }
static inline fe sub(fe a, fe b) { return F(subp(a.real, b.real), subp(a.img, b.img)); }
static inline fe mul(fe a, fe b)                /* prime_field.h:125-151 */
{
    uint64_t ac = red((u128)a.real * b.real);
    uint64_t bd = red((u128)a.img * b.img);
    uint64_t t = red((u128)(a.real + a.img) * (b.real + b.img));
    return F(subp(ac, bd), subp(t, addp(ac, bd)));
}
static inline int feq(fe a, fe b) { return a.real == b.real && a.img == b.img; }
static const fe ZERO = {0, 0};
static inline fe ONE(void) { return F(1, 0); }

static fe fpow(fe x, u128 e)                     /* prime_field.cpp:74-87 */
{
    fe r = ONE(), t = x;
    while (e) { if (e & 1) r = mul(r, t); t = mul(t, t); e >>= 1; }
    return r;
}
static fe finv(fe x)                             /* prime_field.cpp:88-92 */
{
    u128 p = P61;
    return fpow(x, p * p - 2);
}
static fe root_of_unity(int log_order)           /* prime_field.cpp:60-73 */
{
    fe r = F(2147483648ULL, 1033321771269002680ULL);
    for (int i = 0; i < 62 - log_order; ++i) r = mul(r, r);
    return r;
}

/* ===================================================================== */
/* glibc rand() TYPE_3 (the reference never calls srand: seed 1).  Used   */
/* by prime_field::random (prime_field.cpp:35-41): real first, then img.  */
/* ===================================================================== */
typedef struct { uint32_t s[34]; uint64_t k; } grand_t;
static uint32_t grand_next(grand_t *g)
{
    uint32_t v = g->s[(g->k - 31) % 34] + g->s[(g->k - 3) % 34];
    g->s[g->k % 34] = v;
    g->k++;
    return v >> 1;
}
static void grand_init(grand_t *g, uint32_t seed)
{
    int32_t r[34];
    r[0] = (int32_t)(seed ? seed : 1);
    for (int i = 1; i < 31; ++i) {
        int64_t t = (16807LL * r[i - 1]) % 2147483647LL;
        if (t < 0) t += 2147483647LL;
        r[i] = (int32_t)t;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 34; ++i) g->s[i] = (uint32_t)r[i];
    g->k = 34;
    for (int i = 0; i < 310; ++i) grand_next(g);
}
static fe grand_fe(grand_t *g)
{
    fe r;
    r.real = (uint64_t)grand_next(g) % P61;
    r.img = (uint64_t)grand_next(g) % P61;
    return r;
}

/* ===================================================================== */
/* FFT (RS_polynomial.cpp:27-240).  result[k] = sum_{j<coef_len} c_j w^jk */
/* for k < order, computed with the reference's replicated-leaf decimation */
/* (scalar path of :62-160).                                              */
/* ===================================================================== */
static fe *g_fft_buf[2];
static size_t g_fft_cap;
static void fft(const fe *coef, int coef_len, int order, fe rou, fe *result)
{
    if ((size_t)order > g_fft_cap) {
        free(g_fft_buf[0]); free(g_fft_buf[1]);
        g_fft_buf[0] = (fe *)malloc(sizeof(fe) * order);
        g_fft_buf[1] = (fe *)malloc(sizeof(fe) * order);
        g_fft_cap = order;
    }
    fe rot[62];
    int lg_order = -1, lg_coef = -1;
    rot[0] = rou;
    for (int i = 0; i < 62; ++i) {
        if (i > 0) rot[i] = mul(rot[i - 1], rot[i - 1]);
        if ((1LL << i) == order) lg_order = i;
        if ((1LL << i) == coef_len) lg_coef = i;
    }
    fe **dst = g_fft_buf;
    int blk_sz = order / coef_len;
    for (int j = 0; j < blk_sz; ++j)
        for (int i = 0; i < coef_len; ++i) dst[lg_coef & 1][(j << lg_coef) | i] = coef[i];
    for (int dep = lg_coef - 1; dep >= 0; --dep) {
        int blk = 1 << (lg_order - dep), half = blk >> 1;
        int cur = dep & 1, pre = cur ^ 1;
        fe x = ONE();
        for (int k = 0; k < blk / 2; ++k) {
            int dk = k & (half - 1);
            for (int j = 0; j < (1 << dep); ++j) {
                fe l = dst[pre][(dk << (dep + 1)) | j];
                fe r = mul(x, dst[pre][(dk << (dep + 1)) | (1 << dep) | j]);
                dst[cur][(k << dep) | j] = add(l, r);
                dst[cur][((k + blk / 2) << dep) | j] = sub(l, r);
            }
            x = mul(x, rot[dep]);
        }
    }
    memcpy(result, dst[0], sizeof(fe) * order);
}
static void ifft(const fe *evals, int coef_len, int order, fe rou, fe *dst)
{                                                /* RS_polynomial.cpp:169-240 */
    if (coef_len > order) coef_len = order;
    if (coef_len < 1) return;
    fe *sub_eval = (fe *)calloc((size_t)coef_len, sizeof(fe));
    for (int i = 0; i < coef_len; ++i) sub_eval[i] = evals[(size_t)i * (order / coef_len)];
    fe new_rou = ONE();
    for (int i = 0; i < order / coef_len; ++i) new_rou = mul(new_rou, rou);
    order = coef_len;
    fe inv_rou = ONE(), tmp = new_rou;
    int lg_order = 0;
    while ((1 << lg_order) < order) lg_order++;
    for (int i = 0; i < lg_order; ++i) { inv_rou = mul(inv_rou, tmp); tmp = mul(tmp, tmp); }
    fft(sub_eval, order, coef_len, inv_rou, dst);
    free(sub_eval);
    fe inv_n = fpow(fe_u((uint64_t)order), P61 - 2);
    for (int i = 0; i < coef_len; ++i) dst[i] = mul(dst[i], inv_n);
}
static int mylog(long long x)                    /* utility.cpp:3-10 */
{
    for (int i = 0; i < 64; ++i) if ((1LL << i) == x) return i;
    return -1;
}

/* ===================================================================== */
/* SHA3-256 of exactly 64 bytes (my_hhash.h:39-46 with USESHA3), FIPS 202 */
/* ===================================================================== */
static const uint64_t KRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808AULL, 0x8000000080008000ULL,
    0x000000000000808BULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008AULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000AULL,
    0x000000008000808BULL, 0x800000000000008BULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800AULL, 0x800000008000000AULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

static inline uint64_t rol(uint64_t x, int s) { return (x << s) | (x >> (64 - s)); }
/* Keccak-f[1600]: theta, rho+pi (b[y + 5*((2x+3y)%5)] = rot(a[x+5y] ^ d[x], r[x+5y])), chi, iota */
static void keccakf(uint64_t *a)
{
    for (int rd = 0; rd < 24; ++rd) {
        uint64_t c0 = a[0] ^ a[5] ^ a[10] ^ a[15] ^ a[20], c1 = a[1] ^ a[6] ^ a[11] ^ a[16] ^ a[21];
        uint64_t c2 = a[2] ^ a[7] ^ a[12] ^ a[17] ^ a[22], c3 = a[3] ^ a[8] ^ a[13] ^ a[18] ^ a[23];
        uint64_t c4 = a[4] ^ a[9] ^ a[14] ^ a[19] ^ a[24];
        uint64_t d0 = c4 ^ rol(c1, 1), d1 = c0 ^ rol(c2, 1), d2 = c1 ^ rol(c3, 1), d3 = c2 ^ rol(c4, 1), d4 = c3 ^ rol(c0, 1);
        uint64_t b[25];
        b[0] = a[0] ^ d0;
        b[16] = rol(a[5] ^ d0, 36);
        b[7] = rol(a[10] ^ d0, 3);
        b[23] = rol(a[15] ^ d0, 41);
        b[14] = rol(a[20] ^ d0, 18);
        b[10] = rol(a[1] ^ d1, 1);
        b[1] = rol(a[6] ^ d1, 44);
        b[17] = rol(a[11] ^ d1, 10);
        b[8] = rol(a[16] ^ d1, 45);
        b[24] = rol(a[21] ^ d1, 2);
        b[20] = rol(a[2] ^ d2, 62);
        b[11] = rol(a[7] ^ d2, 6);
        b[2] = rol(a[12] ^ d2, 43);
        b[18] = rol(a[17] ^ d2, 15);
        b[9] = rol(a[22] ^ d2, 61);
        b[5] = rol(a[3] ^ d3, 28);
        b[21] = rol(a[8] ^ d3, 55);
        b[12] = rol(a[13] ^ d3, 25);
        b[3] = rol(a[18] ^ d3, 21);
        b[19] = rol(a[23] ^ d3, 56);
        b[15] = rol(a[4] ^ d4, 27);
        b[6] = rol(a[9] ^ d4, 20);
        b[22] = rol(a[14] ^ d4, 39);
        b[13] = rol(a[19] ^ d4, 8);
        b[4] = rol(a[24] ^ d4, 14);
        a[0] = b[0] ^ (~b[1] & b[2]);
        a[1] = b[1] ^ (~b[2] & b[3]);
        a[2] = b[2] ^ (~b[3] & b[4]);
        a[3] = b[3] ^ (~b[4] & b[0]);
        a[4] = b[4] ^ (~b[0] & b[1]);
        a[5] = b[5] ^ (~b[6] & b[7]);
        a[6] = b[6] ^ (~b[7] & b[8]);
        a[7] = b[7] ^ (~b[8] & b[9]);
        a[8] = b[8] ^ (~b[9] & b[5]);
        a[9] = b[9] ^ (~b[5] & b[6]);
        a[10] = b[10] ^ (~b[11] & b[12]);
        a[11] = b[11] ^ (~b[12] & b[13]);
        a[12] = b[12] ^ (~b[13] & b[14]);
        a[13] = b[13] ^ (~b[14] & b[10]);
        a[14] = b[14] ^ (~b[10] & b[11]);
        a[15] = b[15] ^ (~b[16] & b[17]);
        a[16] = b[16] ^ (~b[17] & b[18]);
        a[17] = b[17] ^ (~b[18] & b[19]);
        a[18] = b[18] ^ (~b[19] & b[15]);
        a[19] = b[19] ^ (~b[15] & b[16]);
        a[20] = b[20] ^ (~b[21] & b[22]);
        a[21] = b[21] ^ (~b[22] & b[23]);
        a[22] = b[22] ^ (~b[23] & b[24]);
        a[23] = b[23] ^ (~b[24] & b[20]);
        a[24] = b[24] ^ (~b[20] & b[21]);
        a[0] ^= KRC[rd];
    }
}

static void hhash(const void *src64, void *dst32)
{
    uint64_t a[25];
    memset(a, 0, sizeof a);
    memcpy(a, src64, 64);               /* little-endian host */
    a[8] ^= 0x06ULL;
    a[16] ^= 0x8000000000000000ULL;
    keccakf(a);
    memcpy(dst32, a, 32);
}

/* Merkle tree, heap layout, raw leaves, pad leaves = H(0^64)
 * (merkle_tree.cpp:4-57).  Returns padded leaf count. */
static int merkle_build(const dig *leaves, int ele_num, dig *dst)
{
    int sz = 1;
    while (sz < ele_num) sz *= 2;
    memset(dst, 0, sizeof(dig) * 2 * sz);
    dig data[2];
    memset(data, 0, sizeof data);
    for (int i = sz - 1; i >= 0; --i) {
        if (i < ele_num) dst[i + sz] = leaves[i];
        else { memset(data, 0, sizeof data); hhash(data, &dst[i + sz]); }
    }
    int start = sz, lvl = sz;
    lvl /= 2; start -= lvl;
    while (lvl >= 1) {
        for (int i = 0; i < lvl; ++i) {
            data[0] = dst[start + lvl + i * 2];
            data[1] = dst[start + lvl + i * 2 + 1];
            hhash(data, &dst[start + i]);
        }
        lvl /= 2; start -= lvl;
    }
    return sz;
}

/* ===================================================================== */
/* Circuit (zk_verifier.cpp:15-217 read_circuit, circuit_fast_track.h)    */
/* ===================================================================== */
typedef struct { int32_t ty; int64_t u, v; } gate;
typedef struct { int bl; int n_alloc; gate *g; } layer;
typedef struct { int depth; layer *L; int max_bl; } circuit;

static void circ_free(circuit *c)
{
    if (!c->L) return;
    for (int i = 0; i < c->depth; ++i) free(c->L[i].g);
    free(c->L);
    c->L = NULL;
}
static int bitlen(long long x) { int c = 0; while (x) { c++; x >>= 1; } return c; }

static int circ_build(const vgo_circuit_desc *dsc, circuit *c)
{
    int d = dsc->d;
    memset(c, 0, sizeof *c);
    if (d < 2) FAIL(-1, "circuit needs an input layer and at least one gate layer");
    c->depth = d + 1;
    c->L = (layer *)calloc(d + 1, sizeof(layer));
    c->max_bl = -1;
    size_t off = 0;
    for (int i = 1; i <= d; ++i) {
        int64_t n = dsc->n[i - 1];
        int pad_req = d > 3 ? 17 : 15;                         /* :34-37 */
        int64_t n_pad = (i == 1 && n < (1LL << pad_req)) ? (1LL << pad_req) : n;
        if (n < 1 || (n != 1 && (n & (n - 1))))
            FAIL(-2, "layer %d: gate count %lld is not a power of two", i, (long long)n);
        if (i == 1 && n == 1) FAIL(-2, "single-input circuits are not supported");
        int64_t alloc = (n == 1) ? 2 : n_pad;
        if (i != 1) {
            c->L[i].g = (gate *)calloc(alloc, sizeof(gate));
            c->L[i].n_alloc = (int)alloc;
        } else {
            c->L[0].g = (gate *)calloc(alloc, sizeof(gate));
            c->L[1].g = (gate *)calloc(alloc, sizeof(gate));
            c->L[0].n_alloc = c->L[1].n_alloc = (int)alloc;
        }
        int prev_bl = c->L[i - 1].bl;
        for (int64_t j = 0; j < n; ++j, ++off) {
            int ty = dsc->ty[off];
            int64_t u = dsc->u[off], v = dsc->v[off];
            if (ty != 3) {                                      /* :70-86 */
                if (ty == 5) {
                    if (!(u >= 0 && u < (1LL << prev_bl)) || !(v > u && v <= (1LL << prev_bl)))
                        FAIL(-3, "layer %d gate %lld: sum range out of bounds", i, (long long)j);
                } else if (!(u >= 0 && u < (1LL << prev_bl)) || !(v >= 0 && v < (1LL << prev_bl)))
                    FAIL(-3, "layer %d gate %lld: input index out of bounds", i, (long long)j);
            }
            if (ty == 6 || ty == 10) v = 0;                     /* :87-98 */
            if (ty == 13 && u != v) FAIL(-3, "bit-test gate needs u == v");
            if (ty == 12 && v - u + 1 > 60) FAIL(-3, "power-sum gate spans more than 60 entries");   /* zk_prover.cpp:129 */
            if (ty < 0 || ty > 13 || ty == 11) FAIL(-3, "unknown gate type %d", ty);                /* zk_prover.cpp:145 */
            if (i != 1) {
                c->L[i].g[j].ty = ty; c->L[i].g[j].u = u; c->L[i].g[j].v = v;
            } else {
                if (ty != 2 && ty != 3) FAIL(-3, "input layer gates must be type 2 or 3");
                c->L[1].g[j].ty = 4; c->L[1].g[j].u = j; c->L[1].g[j].v = 0;
                c->L[0].g[j].ty = ty; c->L[0].g[j].u = u; c->L[0].g[j].v = v;
            }
        }
        if (i == 1) {                                           /* :119-128 */
            for (int64_t g = n; g < n_pad; ++g) {
                c->L[1].g[g].ty = 4; c->L[1].g[g].u = g; c->L[1].g[g].v = 0;
                c->L[0].g[g].ty = 3; c->L[0].g[g].u = 0; c->L[0].g[g].v = 0;
            }
            n = n_pad;
        }
        if (n == 1) {                                           /* :148-163 */
            c->L[i].g[1].ty = 2; c->L[i].g[1].u = 0; c->L[i].g[1].v = 0;
            c->L[i].bl = 1;
        } else {
            c->L[i].bl = bitlen(n - 1);
            if (i == 1) c->L[0].bl = c->L[i].bl;
        }
        if (c->L[i].bl > c->max_bl) c->max_bl = c->L[i].bl;
    }
    return 0;
}

/* layer evaluation (zk_prover.cpp:54-157) */
static fe **circ_evaluate(const circuit *C)
{
    fe **cv = (fe **)calloc(C->depth, sizeof(fe *));
    int n0 = 1 << C->L[0].bl;
    cv[0] = (fe *)malloc(sizeof(fe) * n0);
    for (int g = 0; g < n0; ++g) {
        int u = (int)C->L[0].g[g].u;
        cv[0][g] = fe_u((uint64_t)(int64_t)u);
    }
    for (int i = 1; i < C->depth; ++i) {
        int n = 1 << C->L[i].bl;
        fe *o = cv[i] = (fe *)malloc(sizeof(fe) * n);
        const fe *p = cv[i - 1];
        for (int g = 0; g < n; ++g) {
            int ty = C->L[i].g[g].ty, u = (int)C->L[i].g[g].u, v = (int)C->L[i].g[g].v;
            switch (ty) {
            case 0: o[g] = add(p[u], p[v]); break;
            case 1: o[g] = mul(p[u], p[v]); break;
            case 2: o[g] = ZERO; break;
            case 3: o[g] = fe_u((uint64_t)(int64_t)u); break;
            case 4: o[g] = p[u]; break;
            case 5: { fe s = ZERO; for (int k = u; k < v; ++k) s = add(s, p[k]); o[g] = s; break; }
            case 6: o[g] = sub(ONE(), p[u]); break;
            case 7: o[g] = sub(p[u], p[v]); break;
            case 8: { fe x = p[u], y = p[v]; o[g] = sub(add(x, y), mul(mul(fe_u(2), x), y)); break; }
            case 9: { fe x = p[u], y = p[v]; o[g] = sub(y, mul(x, y)); break; }
            case 10: o[g] = p[u]; break;
            case 12: { fe s = ZERO; for (int k = u; k <= v; ++k) s = add(s, mul(p[k], fe_u(1ULL << (k - u)))); o[g] = s; break; }
            case 13: o[g] = mul(p[u], sub(ONE(), p[v])); break;
            default: o[g] = ZERO; break;
            }
        }
    }
    return cv;
}

/* ===================================================================== */
/* Challenge / mask schedule.  Every rand() draw of zk_verifier::verify,  */
/* zk_prover::generate_maskR / generate_maskpoly_pre_rho and              */
/* poly_commit_prover::commit_phase / verify_poly_commitment, in order.   */
/* The order depends only on the circuit shape.                           */
/* ===================================================================== */
typedef struct {
    int bl;
    fe rho, maskR[6], maskr[2], *maskpoly;          /* maskpoly: 4bl+9 raw */
    fe *r_u, *r_v, r_c, random_combine, linear_coeff, linear_combine, alpha, beta;
} lsched;
typedef struct {
    int depth, bl_out, n_in;
    fe *r0, *r1;
    lsched *L;
    fe fri_r[64];
    int n_fri;
    int64_t pow[33];
} sched;

static void sched_free(sched *s)
{
    free(s->r0); free(s->r1);
    if (s->L) for (int i = 1; i < s->depth; ++i) { free(s->L[i].maskpoly); free(s->L[i].r_u); free(s->L[i].r_v); }
    free(s->L);
}
static void sched_build(const circuit *C, sched *s)
{
    grand_t g;
    grand_init(&g, 1);
    memset(s, 0, sizeof *s);
    s->depth = C->depth;
    s->bl_out = C->L[C->depth - 1].bl;
    s->n_in = C->L[0].bl;
    s->r0 = (fe *)malloc(sizeof(fe) * s->bl_out);
    s->r1 = (fe *)malloc(sizeof(fe) * s->bl_out);
    for (int k = 0; k < s->bl_out; ++k) s->r0[k] = grand_fe(&g);   /* zk_verifier.cpp:929 */
    for (int k = 0; k < s->bl_out; ++k) s->r1[k] = grand_fe(&g);
    s->L = (lsched *)calloc(C->depth, sizeof(lsched));
    for (int i = C->depth - 1; i >= 1; --i) {
        lsched *l = &s->L[i];
        int bl = l->bl = C->L[i - 1].bl;
        l->rho = grand_fe(&g);                                      /* :955 */
        if (i > 1) for (int k = 0; k < 6; ++k) l->maskR[k] = grand_fe(&g);   /* zk_prover.cpp:231-233 */
        l->maskr[0] = grand_fe(&g); l->maskr[1] = grand_fe(&g);     /* zk_prover.cpp:239-241 */
        int mlen = (2 * bl + 1) * 2 + 1 + 6;                        /* zk_prover.cpp:167 */
        l->maskpoly = (fe *)malloc(sizeof(fe) * mlen);
        for (int k = 0; k < mlen; ++k) l->maskpoly[k] = grand_fe(&g);
        l->r_u = (fe *)malloc(sizeof(fe) * bl);
        l->r_v = (fe *)malloc(sizeof(fe) * bl);
        for (int k = 0; k < bl; ++k) l->r_u[k] = grand_fe(&g);      /* :968 */
        for (int k = 0; k < bl; ++k) l->r_v[k] = grand_fe(&g);      /* :969 */
        l->r_c = grand_fe(&g);                                      /* :971 */
        l->random_combine = grand_fe(&g);                           /* :978 */
        l->linear_coeff = grand_fe(&g);                             /* :980 */
        l->linear_combine = grand_fe(&g);                           /* :1002 */
        l->alpha = grand_fe(&g);                                    /* :1183 */
        l->beta = grand_fe(&g);
    }
    int n = s->n_in;
    int cw = 1 << (n + 5 - 6);                                      /* vpd_verifier.cpp:52-64 */
    s->n_fri = 0;
    while (cw > 32) { s->fri_r[s->n_fri++] = grand_fe(&g); cw /= 2; }
    for (int rep = 0; rep < 33; ++rep) {                            /* vpd_verifier.cpp:127-130 */
        int64_t p;
        do { p = (int64_t)(grand_next(&g) % (uint32_t)(1u << (n - 1))); }
        while (p < (1LL << (n - 6)) || p % 2 == 1);
        s->pow[rep] = p;
    }
}

/* ===================================================================== */
/* Transcript buffer                                                      */
/* ===================================================================== */
typedef struct { uint8_t *p; size_t n, cap; } tbuf;
static void tb_put(tbuf *t, const void *x, size_t k)
{
    if (t->n + k > t->cap) {
        size_t c = t->cap ? t->cap * 2 : 4096;
        while (c < t->n + k) c *= 2;
        t->p = (uint8_t *)realloc(t->p, c);
        t->cap = c;
    }
    memcpy(t->p + t->n, x, k);
    t->n += k;
}
static void tb_fe(tbuf *t, fe x) { tb_put(t, &x, 16); }
static void tb_u32(tbuf *t, uint32_t x) { tb_put(t, &x, 4); }
static void tb_dig(tbuf *t, const dig *d) { tb_put(t, d, 32); }
typedef struct { const uint8_t *p; size_t n, off; int bad; } trd;
static void tr_get(trd *r, void *x, size_t k)
{
    if (r->off + k > r->n) { r->bad = 1; memset(x, 0, k); return; }
    memcpy(x, r->p + r->off, k);
    r->off += k;
}
/* a field element of the transcript; a non-canonical one (a component
 * >= p) marks the transcript malformed: every operation assumes canonical
 * inputs, as the reference's in-process messages always are */
static fe tr_fe(trd *r)
{
    fe x;
    tr_get(r, &x, 16);
    if (x.real >= P61 || x.img >= P61) r->bad = 1;
    return x;
}
static dig tr_dig(trd *r) { dig d; tr_get(r, &d, 32); return d; }

/* polynomial helpers (polynomial.cpp) */
typedef struct { fe a, b; } lpoly;
typedef struct { fe a, b, c; } qpoly;
typedef struct { fe a, b, c, d, e, f; } q5poly;
static fe q_eval(qpoly p, fe x) { return add(mul(add(mul(p.a, x), p.b), x), p.c); }
static fe q5_eval(q5poly p, fe x)                 /* polynomial.cpp:19-22 */
{
    return add(mul(add(mul(add(mul(add(mul(add(mul(p.a, x), p.b), x), p.c), x), p.d), x), p.e), x), p.f);
}

/* ===================================================================== */
/* zk_prover (zk_prover.cpp)                                              */
/* ===================================================================== */
typedef struct {
    const circuit *C;
    fe **cv;
    lpoly *V, *addV, *ams;
    fe *bg0f, *bg0s, *bg1f, *bg1s, *buf, *bus;
    int total_uv, layer_id, length_g, length_u, length_v;
    fe alpha, beta, v_u, v_v;
    const fe *r0, *r1, *omr0, *omr1;
    fe *maskpoly; int maskpoly_len;
    fe maskpoly_sumc, maskpoly_sumr, rho;
    fe maskR[6], preR[6];
    fe maskR_sumcu, maskR_sumcv, preZu, preZv, Zu, Zv, preu1, prev1, Iuv, prepreu1, preprev1;
    qpoly Rg1, Rg2, sumRc;
    fe maskr[2];
    fe *pri; int n_pri, cap_pri;         /* all_pri_mask */
    fe inv_2;
} zkp;

static void pri_push(zkp *p, fe x)
{
    if (p->n_pri == p->cap_pri) {
        p->cap_pri = p->cap_pri ? p->cap_pri * 2 : 256;
        p->pri = (fe *)realloc(p->pri, sizeof(fe) * p->cap_pri);
    }
    p->pri[p->n_pri++] = x;
}

static void zkp_init(zkp *p, const circuit *C)    /* zk_prover.cpp:24-28, 337-350 */
{
    memset(p, 0, sizeof *p);
    p->C = C;
    size_t n = (size_t)1 << C->max_bl;
    p->V = (lpoly *)calloc(n, sizeof(lpoly));
    p->addV = (lpoly *)calloc(n, sizeof(lpoly));
    p->ams = (lpoly *)calloc(n, sizeof(lpoly));
    int half = (C->max_bl >> 1) + 1;
    p->bg0f = (fe *)calloc((size_t)1 << half, sizeof(fe));
    p->bg0s = (fe *)calloc((size_t)1 << half, sizeof(fe));
    p->bg1f = (fe *)calloc((size_t)1 << half, sizeof(fe));
    p->bg1s = (fe *)calloc((size_t)1 << half, sizeof(fe));
    p->buf = (fe *)calloc((size_t)1 << half, sizeof(fe));
    p->bus = (fe *)calloc((size_t)1 << half, sizeof(fe));
    p->inv_2 = finv(fe_u(2));
}
static void zkp_free(zkp *p)
{
    free(p->V); free(p->addV); free(p->ams);
    free(p->bg0f); free(p->bg0s); free(p->bg1f); free(p->bg1s); free(p->buf); free(p->bus);
    if (p->cv) { for (int i = 0; i < p->C->depth; ++i) free(p->cv[i]); free(p->cv); }
    free(p->pri);
}

static fe V_res(const fe *omr0, const fe *r0, const fe *out_raw, int r0_size, int out_size)
{                                                  /* zk_prover.cpp:30-52 */
    fe *o = (fe *)malloc(sizeof(fe) * out_size);
    memcpy(o, out_raw, sizeof(fe) * out_size);
    for (int i = 0; i < r0_size; ++i) {
        for (int j = 0; j < (out_size >> 1); ++j)
            o[j] = add(mul(o[j << 1], omr0[i]), mul(o[j << 1 | 1], r0[i]));
        out_size >>= 1;
    }
    fe r = o[0];
    free(o);
    return r;
}

static void gen_maskpoly_pre_rho(zkp *p, const lsched *l, int length, int degree)
{                                                  /* zk_prover.cpp:162-172 */
    free(p->maskpoly);
    p->maskpoly_len = length * degree + 1 + 6;
    p->maskpoly = (fe *)malloc(sizeof(fe) * p->maskpoly_len);
    for (int i = 0; i < p->maskpoly_len; ++i) { p->maskpoly[i] = l->maskpoly[i]; pri_push(p, p->maskpoly[i]); }
}
static void gen_maskpoly_after_rho(zkp *p, int length, int degree)
{                                                  /* zk_prover.cpp:174-188 */
    for (int i = 0; i < length * degree + 1 + 6; ++i) p->maskpoly[i] = mul(p->maskpoly[i], p->rho);
    fe s = p->maskpoly[length * degree];
    for (int i = length * degree; i >= 0; --i) s = add(s, p->maskpoly[i]);
    for (int i = 1; i <= 6; ++i) s = add(s, p->maskpoly[length * degree + i]);
    for (int i = 1; i < length; ++i) s = add(s, s);
    p->maskpoly_sumc = s;
    p->maskpoly_sumr = ZERO;
}
static void gen_maskR(zkp *p, const lsched *l, int layer_id)
{                                                  /* zk_prover.cpp:192-255 */
    for (int i = 0; i < 6; ++i) pri_push(p, p->maskR[i]);
    p->prepreu1 = p->preu1;
    p->preprev1 = p->prev1;
    for (int i = 0; i < 6; ++i) p->preR[i] = p->maskR[i];
    fe *m = p->maskR;
    p->Rg1.a = m[4];
    p->Rg1.b = add(m[3], mul(m[5], p->preu1));
    p->Rg1.c = add(add(m[0], mul(m[1], p->preu1)), mul(mul(m[2], p->preu1), p->preu1));
    p->Rg2.a = m[4];
    p->Rg2.b = add(m[3], mul(m[5], p->prev1));
    p->Rg2.c = add(add(m[0], mul(m[1], p->prev1)), mul(mul(m[2], p->prev1), p->prev1));
    fe sumRu = add(add(add(p->Rg1.a, p->Rg1.b), p->Rg1.c), p->Rg1.c);
    fe sumRv = add(add(add(p->Rg2.a, p->Rg2.b), p->Rg2.c), p->Rg2.c);
    p->maskR_sumcu = mul(mul(p->alpha, p->Zu), sumRu);
    p->maskR_sumcv = mul(mul(p->beta, p->Zv), sumRv);
    p->preZu = p->Zu; p->preZv = p->Zv;
    p->Zu = ONE(); p->Zv = ONE(); p->Iuv = ONE();
    if (layer_id > 1) {
        for (int i = 0; i < 6; ++i) m[i] = l->maskR[i];
        p->sumRc.a = add(m[2], m[2]);
        p->sumRc.b = add(add(m[1], m[1]), m[5]);
        p->sumRc.c = add(add(add(m[0], m[0]), m[3]), m[4]);
    }
    p->maskr[0] = l->maskr[0];
    p->maskr[1] = l->maskr[1];
    if (layer_id == 1) {
        m[0] = p->maskr[0]; m[1] = p->maskr[1];
        m[2] = m[3] = m[4] = m[5] = ZERO;
        p->sumRc.a = ZERO; p->sumRc.b = m[1]; p->sumRc.c = m[0];
    }
}
static void sumcheck_init(zkp *p, const lsched *l, int layer_id, int bg, int bu, int bv, fe a, fe b,
                          const fe *R0, const fe *R1, const fe *o0, const fe *o1)
{                                                  /* zk_prover.cpp:317-333 */
    p->r0 = R0; p->r1 = R1; p->alpha = a; p->beta = b;
    p->layer_id = layer_id; p->length_g = bg; p->length_u = bu; p->length_v = bv;
    p->omr0 = o0; p->omr1 = o1;
    gen_maskR(p, l, layer_id);
}

/* half-size eq tables (zk_prover.cpp:391-418) */
static void half_tables(fe *f, fe *s, fe f0, const fe *r, const fe *omr, int len)
{
    int fh = len >> 1, sh = len - fh;
    f[0] = f0; s[0] = ONE();
    for (int i = 0; i < fh; ++i)
        for (int j = 0; j < (1 << i); ++j) {
            f[j | (1 << i)] = mul(f[j], r[i]);
            f[j] = mul(f[j], omr[i]);
        }
    for (int i = 0; i < sh; ++i)
        for (int j = 0; j < (1 << i); ++j) {
            s[j | (1 << i)] = mul(s[j], r[i + fh]);
            s[j] = mul(s[j], omr[i + fh]);
        }
}

static void phase1_init(zkp *p)                    /* zk_prover.cpp:375-592 */
{
    const circuit *C = p->C;
    int lid = p->layer_id;
    p->total_uv = 1 << C->L[lid - 1].bl;
    const fe *pv = p->cv[lid - 1];
    for (int i = 0; i < p->total_uv; ++i) {
        p->V[i].a = ZERO; p->V[i].b = pv[i];
        p->addV[i].a = p->addV[i].b = ZERO;
        p->ams[i].a = p->ams[i].b = ZERO;
    }
    half_tables(p->bg0f, p->bg0s, p->alpha, p->r0, p->omr0, p->length_g);
    half_tables(p->bg1f, p->bg1s, p->beta, p->r1, p->omr1, p->length_g);
    int fh = p->length_g >> 1, mf = (1 << fh) - 1;
    const gate *G = C->L[lid].g;
    for (int i = 0; i < (1 << p->length_g); ++i) {
        int u = (int)G[i].u, v = (int)G[i].v;
        fe tmp = add(mul(p->bg0f[i & mf], p->bg0s[i >> fh]), mul(p->bg1f[i & mf], p->bg1s[i >> fh]));
        switch (G[i].ty) {
        case 0: p->addV[u].b = add(p->addV[u].b, mul(pv[v], tmp)); p->ams[u].b = add(p->ams[u].b, tmp); break;
        case 2: break;
        case 1: p->ams[u].b = add(p->ams[u].b, mul(pv[v], tmp)); break;
        case 5: for (int j = u; j < v; ++j) p->ams[j].b = add(p->ams[j].b, tmp); break;
        case 12: for (int j = u; j <= v; ++j) { p->ams[j].b = add(p->ams[j].b, tmp); tmp = add(tmp, tmp); } break;
        case 4: {
            fe t4 = add(mul(p->bg0f[u & mf], p->bg0s[u >> fh]), mul(p->bg1f[u & mf], p->bg1s[u >> fh]));
            p->ams[u].b = add(p->ams[u].b, t4);
            break;
        }
        case 6: p->ams[u].b = sub(p->ams[u].b, tmp); p->addV[u].b = add(p->addV[u].b, tmp); break;
        case 7: p->addV[u].b = sub(p->addV[u].b, mul(pv[v], tmp)); p->ams[u].b = add(p->ams[u].b, tmp); break;
        case 8: {
            fe tv = mul(tmp, pv[v]), t2 = add(tv, tv);
            p->addV[u].b = add(p->addV[u].b, tv);
            p->ams[u].b = sub(add(p->ams[u].b, tmp), t2);
            break;
        }
        case 13: { fe tv = mul(tmp, pv[v]); p->ams[u].b = add(sub(p->ams[u].b, tv), tmp); break; }
        case 9: { fe tv = mul(tmp, pv[v]); p->addV[u].b = add(p->addV[u].b, tv); p->ams[u].b = sub(p->ams[u].b, tv); break; }
        case 10: p->ams[u].b = add(p->ams[u].b, tmp); break;
        default: break;                            /* "Warning Unknown gate" */
        }
    }
}

/* bookkeeping fold + round sums shared by the four update routines
 * (zk_prover.cpp:600-633, 684-717, 992-1025, 1075-1108) */
static void fold3(zkp *p, fe r, int current_bit, fe *ra, fe *rb, fe *rc)
{
    fe A = ZERO, B = ZERO, Cc = ZERO;
    lpoly *V = p->V, *aV = p->addV, *am = p->ams;
    for (int i = 0; i < (p->total_uv >> 1); ++i) {
        int g0 = i << 1, g1 = g0 | 1;
        if (current_bit == 0) {
            V[i].b = V[g0].b; V[i].a = sub(V[g1].b, V[i].b);
            aV[i].b = aV[g0].b; aV[i].a = sub(aV[g1].b, aV[i].b);
            am[i].b = am[g0].b; am[i].a = sub(am[g1].b, am[i].b);
        } else {
            V[i].b = add(mul(V[g0].a, r), V[g0].b);
            V[i].a = sub(add(mul(V[g1].a, r), V[g1].b), V[i].b);
            aV[i].b = add(mul(aV[g0].a, r), aV[g0].b);
            aV[i].a = sub(add(mul(aV[g1].a, r), aV[g1].b), aV[i].b);
            am[i].b = add(mul(am[g0].a, r), am[g0].b);
            am[i].a = sub(add(mul(am[g1].a, r), am[g1].b), am[i].b);
        }
        A = add(A, mul(am[i].a, V[i].a));
        B = add(add(add(B, mul(am[i].a, V[i].b)), mul(am[i].b, V[i].a)), aV[i].a);
        Cc = add(add(Cc, mul(am[i].b, V[i].b)), aV[i].b);
    }
    p->total_uv >>= 1;
    *ra = A; *rb = B; *rc = Cc;
}
static fe dbl_n(fe x, int n) { for (int i = 0; i < n; ++i) x = add(x, x); return x; }

static qpoly phase1_update(zkp *p, fe r, int cb)   /* zk_prover.cpp:596-678 */
{
    qpoly ret;
    fold3(p, r, cb, &ret.a, &ret.b, &ret.c);
    fe one = ONE();
    p->Iuv = mul(p->Iuv, sub(one, r));
    if (cb > 0) {
        p->maskR_sumcu = mul(p->maskR_sumcu, sub(one, r));
        p->maskR_sumcv = mul(p->maskR_sumcv, sub(one, r));
        p->Zu = mul(mul(p->Zu, sub(one, r)), r);
    }
    ret.b = sub(sub(ret.b, p->maskR_sumcu), p->maskR_sumcv);
    ret.c = add(add(ret.c, p->maskR_sumcu), p->maskR_sumcv);
    int k = p->length_u + p->length_v - cb;
    fe t1 = dbl_n(p->maskpoly[cb << 1], k), t2 = dbl_n(p->maskpoly[(cb << 1) + 1], k);
    p->maskpoly_sumc = mul(sub(sub(p->maskpoly_sumc, t1), t2), p->inv_2);
    fe t3 = ZERO;
    if (cb > 0) {
        p->maskpoly_sumr = add(add(p->maskpoly_sumr, mul(mul(p->maskpoly[(cb << 1) - 2], r), r)),
                               mul(p->maskpoly[(cb << 1) - 1], r));
        t3 = dbl_n(p->maskpoly_sumr, k);
    }
    ret.a = add(ret.a, t1);
    ret.b = add(ret.b, t2);
    ret.c = add(add(ret.c, p->maskpoly_sumc), t3);
    return ret;
}

/* last-bit extra terms (zk_prover.cpp:736-747, 1122-1133) */
static void lastbit_terms(zkp *p, fe Z, fe *ra, fe *rb, fe *rd, fe *re, fe *rf)
{
    fe a = p->sumRc.a, b = p->sumRc.b, c = p->sumRc.c;
    fe d = p->ams[0].a, e = p->ams[0].b;
    *rd = mul(mul(sub(ZERO, a), d), Z);
    *re = mul(sub(mul(a, sub(d, e)), mul(b, d)), Z);
    *rf = mul(sub(add(mul(a, e), mul(b, sub(d, e))), mul(c, d)), Z);
    *ra = add(*ra, mul(add(mul(c, sub(d, e)), mul(b, e)), Z));
    *rb = add(*rb, mul(mul(c, e), Z));
}

static q5poly phase1_updatelastbit(zkp *p, fe r, int cb)   /* zk_prover.cpp:680-787 */
{
    q5poly ret = {ZERO, ZERO, ZERO, ZERO, ZERO, ZERO};
    fold3(p, r, cb, &ret.a, &ret.b, &ret.c);
    fe one = ONE();
    p->Iuv = mul(p->Iuv, sub(one, r));
    if (cb > 0) {
        p->maskR_sumcv = mul(p->maskR_sumcv, sub(one, r));
        p->maskR_sumcu = mul(p->maskR_sumcu, sub(one, r));
        p->Zu = mul(mul(p->Zu, sub(one, r)), r);
    }
    ret.b = sub(sub(ret.b, p->maskR_sumcu), p->maskR_sumcv);
    ret.c = add(add(ret.c, p->maskR_sumcu), p->maskR_sumcv);
    if (cb == p->length_u - 1) lastbit_terms(p, p->Zu, &ret.a, &ret.b, &ret.d, &ret.e, &ret.f);
    int k = p->length_u + p->length_v - cb, base = (p->length_u + p->length_v + 1) << 1;
    fe t1 = dbl_n(p->maskpoly[cb << 1], k), t2 = dbl_n(p->maskpoly[(cb << 1) + 1], k);
    fe t4 = dbl_n(p->maskpoly[base + 1], k), t5 = dbl_n(p->maskpoly[base + 2], k), t6 = dbl_n(p->maskpoly[base + 3], k);
    p->maskpoly_sumc = mul(sub(sub(sub(sub(sub(p->maskpoly_sumc, t1), t2), t4), t5), t6), p->inv_2);
    fe t3 = ZERO;
    if (cb > 0) {
        p->maskpoly_sumr = add(add(p->maskpoly_sumr, mul(mul(p->maskpoly[(cb << 1) - 2], r), r)),
                               mul(p->maskpoly[(cb << 1) - 1], r));
        t3 = dbl_n(p->maskpoly_sumr, k);
    }
    ret.a = add(ret.a, t1); ret.b = add(ret.b, t2);
    ret.c = add(add(ret.c, p->maskpoly_sumc), t3);
    ret.d = add(ret.d, t4); ret.e = add(ret.e, t5); ret.f = add(ret.f, t6);
    q5poly r5 = {ret.d, ret.e, ret.f, ret.a, ret.b, ret.c};   /* :781 */
    return r5;
}

static void phase2_init(zkp *p, fe prev_r, const fe *r_u, const fe *omr_u)   /* :789-983 */
{
    const circuit *C = p->C;
    fe one = ONE();
    p->preu1 = prev_r;
    p->maskR_sumcu = mul(p->maskR_sumcu, sub(one, prev_r));
    p->maskR_sumcv = mul(p->maskR_sumcv, sub(one, prev_r));
    p->Iuv = mul(p->Iuv, sub(one, prev_r));
    p->v_u = add(mul(p->V[0].a, prev_r), p->V[0].b);
    p->Zu = mul(mul(p->Zu, sub(one, prev_r)), prev_r);
    p->v_u = add(p->v_u, mul(p->Zu, q_eval(p->sumRc, prev_r)));
    half_tables(p->buf, p->bus, ONE(), r_u, omr_u, p->length_u);
    int fh = p->length_u >> 1, mf = (1 << fh) - 1;
    int fgh = p->length_g >> 1, mgf = (1 << fgh) - 1;
    int lid = p->layer_id;
    p->total_uv = 1 << C->L[lid - 1].bl;
    int total_g = 1 << C->L[lid].bl;
    const fe *pv = p->cv[lid - 1];
    for (int i = 0; i < p->total_uv; ++i) {
        p->ams[i].a = p->ams[i].b = ZERO;
        p->addV[i].a = p->addV[i].b = ZERO;
        p->V[i].a = ZERO; p->V[i].b = pv[i];
    }
    const gate *G = C->L[lid].g;
    fe vu = p->v_u;
    for (int i = 0; i < total_g; ++i) {
        int ty = G[i].ty, u = (int)G[i].u, v = (int)G[i].v;
        fe tg = add(mul(p->bg0f[i & mgf], p->bg0s[i >> fgh]), mul(p->bg1f[i & mgf], p->bg1s[i >> fgh]));
        switch (ty) {
        case 1: { fe tu = mul(p->buf[u & mf], p->bus[u >> fh]); p->ams[v].b = add(p->ams[v].b, mul(mul(tg, tu), vu)); break; }
        case 0: {
            fe tu = mul(p->buf[u & mf], p->bus[u >> fh]), tgu = mul(tg, tu);
            p->ams[v].b = add(p->ams[v].b, tgu);
            p->addV[v].b = add(mul(tgu, vu), p->addV[v].b);
            break;
        }
        case 2: case 4: break;
        case 5: {
            fe tgv = mul(tg, vu);
            for (int j = u; j < v; ++j) { fe tu = mul(p->buf[j & mf], p->bus[j >> fh]); p->addV[0].b = add(p->addV[0].b, mul(tgv, tu)); }
            break;
        }
        case 12: {
            fe tgv = mul(tg, vu);
            for (int j = u; j <= v; ++j) {
                fe tu = mul(p->buf[j & mf], p->bus[j >> fh]);
                p->addV[0].b = add(p->addV[0].b, mul(tgv, tu));
                tgv = add(tgv, tgv);
            }
            break;
        }
        case 6: { fe tu = mul(p->buf[u & mf], p->bus[u >> fh]), tgu = mul(tg, tu); p->addV[v].b = sub(add(p->addV[v].b, tgu), mul(tgu, vu)); break; }
        case 7: { fe tu = mul(p->buf[u & mf], p->bus[u >> fh]), t = mul(tg, tu); p->ams[v].b = sub(p->ams[v].b, t); p->addV[v].b = add(mul(t, vu), p->addV[v].b); break; }
        case 8: {
            fe tu = mul(p->buf[u & mf], p->bus[u >> fh]), t = mul(tg, tu), tv = mul(t, vu);
            p->ams[v].b = sub(sub(add(p->ams[v].b, t), tv), tv);
            p->addV[v].b = add(p->addV[v].b, tv);
            break;
        }
        case 13: { fe tu = mul(p->buf[u & mf], p->bus[u >> fh]), t = mul(tg, tu), tv = mul(t, vu); p->ams[v].b = sub(p->ams[v].b, tv); p->addV[v].b = add(p->addV[v].b, tv); break; }
        case 9: { fe tu = mul(p->buf[u & mf], p->bus[u >> fh]), t = mul(tg, tu); p->ams[v].b = sub(add(p->ams[v].b, t), mul(vu, t)); break; }
        case 10: { fe tu = mul(p->buf[u & mf], p->bus[u >> fh]), t = mul(tg, tu); p->addV[v].b = add(p->addV[v].b, mul(t, vu)); break; }
        default: break;
        }
    }
    fe r = prev_r, r2 = mul(r, r), r3 = mul(r2, r);
    int base = (p->length_u + p->length_v + 1) * 2;
    p->maskpoly_sumr = add(add(p->maskpoly_sumr, mul(mul(p->maskpoly[p->length_u * 2 - 2], r), r)), mul(p->maskpoly[p->length_u * 2 - 1], r));
    p->maskpoly_sumr = add(p->maskpoly_sumr, mul(mul(mul(mul(mul(p->maskpoly[base + 1], r), r), r), r), r));
    p->maskpoly_sumr = add(p->maskpoly_sumr, mul(mul(mul(mul(p->maskpoly[base + 2], r), r), r), r));
    p->maskpoly_sumr = add(p->maskpoly_sumr, mul(mul(mul(p->maskpoly[base + 3], r), r), r));
    (void)r3;
}

static qpoly phase2_update(zkp *p, fe r, int cb)   /* zk_prover.cpp:988-1069 */
{
    qpoly ret;
    fold3(p, r, cb, &ret.a, &ret.b, &ret.c);
    fe one = ONE();
    if (cb > 0) p->Iuv = mul(p->Iuv, sub(one, r));
    if (cb > 0) {
        p->maskR_sumcu = mul(p->maskR_sumcu, sub(one, r));
        p->maskR_sumcv = mul(p->maskR_sumcv, sub(one, r));
        p->Zv = mul(mul(p->Zv, sub(one, r)), r);
    }
    ret.b = sub(sub(ret.b, p->maskR_sumcu), p->maskR_sumcv);
    ret.c = add(add(ret.c, p->maskR_sumcu), p->maskR_sumcv);
    int cur = cb + p->length_u, k = p->length_u + p->length_v - cur;
    fe t1 = dbl_n(p->maskpoly[cur << 1], k), t2 = dbl_n(p->maskpoly[(cur << 1) + 1], k);
    p->maskpoly_sumc = mul(sub(sub(p->maskpoly_sumc, t1), t2), p->inv_2);
    p->maskpoly_sumr = add(add(p->maskpoly_sumr, mul(mul(p->maskpoly[(cur << 1) - 2], r), r)), mul(p->maskpoly[(cur << 1) - 1], r));
    fe t3 = dbl_n(p->maskpoly_sumr, k);
    ret.a = add(ret.a, t1); ret.b = add(ret.b, t2);
    ret.c = add(add(ret.c, p->maskpoly_sumc), t3);
    return ret;
}

static q5poly phase2_updatelastbit(zkp *p, fe r, int cb)   /* zk_prover.cpp:1071-1171 */
{
    q5poly ret = {ZERO, ZERO, ZERO, ZERO, ZERO, ZERO};
    fold3(p, r, cb, &ret.a, &ret.b, &ret.c);
    fe one = ONE();
    p->Iuv = mul(p->Iuv, sub(one, r));
    if (cb > 0) {
        p->maskR_sumcu = mul(p->maskR_sumcu, sub(one, r));
        p->maskR_sumcv = mul(p->maskR_sumcv, sub(one, r));
        p->Zv = mul(mul(p->Zv, sub(one, r)), r);
    }
    ret.b = sub(sub(ret.b, p->maskR_sumcu), p->maskR_sumcv);
    ret.c = add(add(ret.c, p->maskR_sumcu), p->maskR_sumcv);
    if (cb == p->length_v - 1) lastbit_terms(p, p->Zv, &ret.a, &ret.b, &ret.d, &ret.e, &ret.f);
    int cur = cb + p->length_u, k = p->length_u + p->length_v - cur;
    int base = (p->length_u + p->length_v + 1) * 2;
    fe t1 = dbl_n(p->maskpoly[cur << 1], k), t2 = dbl_n(p->maskpoly[(cur << 1) + 1], k);
    fe t4 = dbl_n(p->maskpoly[base + 4], k), t5 = dbl_n(p->maskpoly[base + 5], k), t6 = dbl_n(p->maskpoly[base + 6], k);
    p->maskpoly_sumc = mul(sub(sub(sub(sub(sub(p->maskpoly_sumc, t1), t2), t4), t5), t6), p->inv_2);
    p->maskpoly_sumr = add(add(p->maskpoly_sumr, mul(mul(p->maskpoly[(cur << 1) - 2], r), r)), mul(p->maskpoly[(cur << 1) - 1], r));
    fe t3 = dbl_n(p->maskpoly_sumr, k);
    ret.d = add(ret.d, t4); ret.e = add(ret.e, t5); ret.f = add(ret.f, t6);
    ret.a = add(ret.a, t1); ret.b = add(ret.b, t2);
    ret.c = add(add(ret.c, p->maskpoly_sumc), t3);
    q5poly r5 = {ret.d, ret.e, ret.f, ret.a, ret.b, ret.c};
    return r5;
}

static qpoly finalround(zkp *p, fe r, int current, fe general)   /* :1173-1203 */
{
    qpoly ret;
    fe zu = mul(p->Iuv, p->preZu), zv = mul(p->Iuv, p->preZv);
    ret.a = add(mul(mul(zu, p->Rg1.a), p->alpha), mul(mul(zv, p->Rg2.a), p->beta));
    ret.b = add(add(mul(mul(zu, p->Rg1.b), p->alpha), mul(mul(zv, p->Rg2.b), p->beta)), general);
    ret.c = add(mul(mul(zu, p->Rg1.c), p->alpha), mul(mul(zv, p->Rg2.c), p->beta));
    int k = p->length_u + p->length_v - current, base = (p->length_u + p->length_v + 1) * 2;
    fe t1 = dbl_n(p->maskpoly[current << 1], k), t2 = dbl_n(p->maskpoly[(current << 1) + 1], k);
    p->maskpoly_sumc = mul(sub(sub(p->maskpoly_sumc, t1), t2), p->inv_2);
    fe r2 = mul(r, r), r3 = mul(r2, r), r4 = mul(r3, r), r5 = mul(r4, r);
    p->maskpoly_sumr = add(add(p->maskpoly_sumr, mul(p->maskpoly[(current << 1) - 2], r2)), mul(p->maskpoly[(current << 1) - 1], r));
    p->maskpoly_sumr = add(p->maskpoly_sumr, mul(p->maskpoly[base + 4], r5));
    p->maskpoly_sumr = add(p->maskpoly_sumr, mul(p->maskpoly[base + 5], r4));
    p->maskpoly_sumr = add(p->maskpoly_sumr, mul(p->maskpoly[base + 6], r3));
    fe t3 = dbl_n(p->maskpoly_sumr, k);
    ret.a = add(ret.a, t1); ret.b = add(ret.b, t2);
    ret.c = add(add(ret.c, p->maskpoly_sumc), t3);
    return ret;
}
static void finalize(zkp *p, fe r)                 /* zk_prover.cpp:1206-1215 */
{
    fe one = ONE();
    p->prev1 = r;
    p->Iuv = mul(p->Iuv, sub(one, r));
    p->v_v = add(mul(p->V[0].a, r), p->V[0].b);
    p->Zv = mul(mul(p->Zv, sub(one, r)), r);
    p->v_v = add(p->v_v, mul(p->Zv, q_eval(p->sumRc, r)));
}
static fe query(zkp *p, const fe *u, const fe *v, fe rc)   /* zk_prover.cpp:284-306 */
{
    fe res = ZERO;
    int lu = p->length_u, lv = p->length_v, base = 2 * (lu + lv + 1);
    const fe *m = p->maskpoly;
    for (int i = 0; i < lu; ++i) {
        res = add(add(res, mul(mul(m[2 * i], u[i]), u[i])), mul(m[2 * i + 1], u[i]));
        if (i == lu - 1) {
            fe x = u[i], x2 = mul(x, x), x3 = mul(x2, x), x4 = mul(x3, x), x5 = mul(x4, x);
            res = add(res, mul(m[base + 1], x5));
            res = add(res, mul(m[base + 2], x4));
            res = add(res, mul(m[base + 3], x3));
        }
    }
    for (int i = 0; i < lv; ++i) {
        res = add(add(res, mul(mul(m[2 * (i + lu)], v[i]), v[i])), mul(m[2 * (i + lu) + 1], v[i]));
        if (i == lv - 1) {
            fe x = v[i], x2 = mul(x, x), x3 = mul(x2, x), x4 = mul(x3, x), x5 = mul(x4, x);
            res = add(res, mul(m[base + 4], x5));
            res = add(res, mul(m[base + 5], x4));
            res = add(res, mul(m[base + 6], x3));
        }
    }
    res = add(add(res, mul(mul(m[2 * (lu + lv)], rc), rc)), mul(m[2 * (lu + lv) + 1], rc));
    res = add(res, m[2 * (lu + lv + 1)]);
    return res;
}

/* ===================================================================== */
/* Verifier-side wiring predicates (zk_verifier.cpp:219-606, non-parallel */
/* branch; the parallel branch is asserted equal to it at :600-604),      */
/* beta_init (:621-707) and direct_relay (:608-619).  The prover needs    */
/* their combination for sumcheck_finalround's general_value (:1149).     */
/* ===================================================================== */
static void predicates(const circuit *C, int depth, const fe *r0, const fe *r1, const fe *omr0, const fe *omr1,
                       const fe *r_u, const fe *r_v, const fe *omr_u, const fe *omr_v, fe alpha, fe beta, fe ret[14])
{
    for (int i = 0; i < 14; ++i) ret[i] = ZERO;
    if (depth == 1) return;
    int bg = C->L[depth].bl, bu = C->L[depth - 1].bl;
    int hg = ((bg >> 1) > (bg - (bg >> 1)) ? (bg >> 1) : (bg - (bg >> 1))) + 1;
    int hu = ((bu >> 1) > (bu - (bu >> 1)) ? (bu >> 1) : (bu - (bu >> 1))) + 1;
    fe *g0f = (fe *)calloc((size_t)1 << hg, sizeof(fe)), *g0s = (fe *)calloc((size_t)1 << hg, sizeof(fe));
    fe *g1f = (fe *)calloc((size_t)1 << hg, sizeof(fe)), *g1s = (fe *)calloc((size_t)1 << hg, sizeof(fe));
    fe *uf = (fe *)calloc((size_t)1 << hu, sizeof(fe)), *us = (fe *)calloc((size_t)1 << hu, sizeof(fe));
    fe *vf = (fe *)calloc((size_t)1 << hu, sizeof(fe)), *vs = (fe *)calloc((size_t)1 << hu, sizeof(fe));
    half_tables(g0f, g0s, alpha, r0, omr0, bg);
    half_tables(g1f, g1s, beta, r1, omr1, bg);
    half_tables(uf, us, ONE(), r_u, omr_u, bu);
    half_tables(vf, vs, ONE(), r_v, omr_v, bu);
    int fg = bg / 2, fu = bu / 2;
    fe *tmp_u = NULL;
    const gate *G = C->L[depth].g;
    for (int i = 0; i < (1 << bg); ++i) {
        int g = i, u = (int)G[i].u, v = (int)G[i].v, ty = G[i].ty;
        int gf = g & ((1 << fg) - 1), gs = g >> fg;
        fe bgv = add(mul(g0f[gf], g0s[gs]), mul(g1f[gf], g1s[gs]));
        switch (ty) {
        case 0: case 1: case 6: case 7: case 8: case 9: case 13: {
            fe bu_ = mul(uf[u & ((1 << fu) - 1)], us[u >> fu]);
            fe bv_ = mul(vf[v & ((1 << fu) - 1)], vs[v >> fu]);
            ret[ty] = add(ret[ty], mul(mul(bgv, bu_), bv_));
            break;
        }
        case 5: case 12: {
            fe bv0 = mul(vf[0], vs[0]);
            for (int j = u; (ty == 5) ? (j < v) : (j <= v); ++j) {
                fe bj = mul(uf[j & ((1 << fu) - 1)], us[j >> fu]);
                ret[ty] = add(ret[ty], mul(mul(bgv, bv0), bj));
                if (ty == 12) bv0 = add(bv0, bv0);
            }
            break;
        }
        case 10: {
            if (!tmp_u) {
                tmp_u = (fe *)malloc(sizeof(fe) << bu);
                for (int k = 0; k < (1 << bu); ++k) tmp_u[k] = mul(uf[k & ((1 << fu) - 1)], us[k >> fu]);
            }
            ret[10] = add(ret[10], mul(bgv, tmp_u[u]));
            break;
        }
        default: break;
        }
    }
    ret[10] = mul(ret[10], mul(vf[0], vs[0]));
    free(tmp_u);
    free(g0f); free(g0s); free(g1f); free(g1s); free(uf); free(us); free(vf); free(vs);
}
static fe direct_relay(const circuit *C, int depth, const fe *rg, const fe *ru)
{
    if (depth != 1) return ZERO;
    fe r = ONE();
    for (int i = 0; i < C->L[depth].bl; ++i)
        r = mul(r, add(sub(sub(ONE(), rg[i]), ru[i]), mul(mul(fe_u(2), rg[i]), ru[i])));
    return r;
}
static fe general_value(const fe pv[14], fe vu, fe vv)       /* zk_verifier.cpp:1149 */
{
    fe one = ONE(), two = fe_u(2);
    fe g = add(mul(pv[0], add(vu, vv)), mul(mul(pv[1], vu), vv));
    g = add(g, mul(pv[6], sub(one, vu)));
    g = add(g, mul(pv[7], sub(vu, vv)));
    g = add(g, mul(pv[8], sub(add(vu, vv), mul(mul(two, vu), vv))));
    g = add(g, mul(pv[9], sub(vv, mul(vu, vv))));
    g = add(g, mul(pv[5], vu));
    g = add(g, mul(pv[10], vu));
    g = add(g, mul(pv[12], vu));
    g = add(g, mul(pv[13], mul(vu, sub(one, vv))));
    return g;
}

/* ===================================================================== */
/* Poly commitment + FRI prover state (poly_commit.h, fri.cpp)            */
/* ===================================================================== */
#define NSL 64
typedef struct {
    int n;                    /* log input length */
    int slice_size, slice_count, real, gap, msk_len;
    fe *l_eval, *q_eval, *h_eval_arr, *vo, *vo_msk;
    fe *W[2];                 /* interleaved witness codewords */
    dig *tree[2];             /* Merkle trees (heap) of the init commits */
    fe *L_group;
    int log_cur, n_steps;
    fe *rs[64], *rs_msk[64];  /* cpd.rs_codeword / rs_codeword_msk, interleaved */
    dig *ctree[64];
    int csize[64];
    fe all_sum[65];
} pcs;
static void pcs_free(pcs *s)
{
    free(s->l_eval); free(s->q_eval); free(s->h_eval_arr); free(s->vo); free(s->vo_msk);
    free(s->W[0]); free(s->W[1]); free(s->tree[0]); free(s->tree[1]); free(s->L_group);
    for (int i = 0; i < 64; ++i) { free(s->rs[i]); free(s->rs_msk[i]); free(s->ctree[i]); }
}

/* fri::request_init_commit (fri.cpp:36-139) */
static dig request_init_commit(pcs *s, int bit_len, int o)
{
    s->log_cur = bit_len + 5 - 6;
    int S = 1 << s->log_cur, half = S / 2;
    if (o == 0) {
        s->L_group = (fe *)malloc(sizeof(fe) * S);
        fe w = root_of_unity(s->log_cur);
        s->L_group[0] = ONE();
        for (int i = 1; i < S; ++i) s->L_group[i] = mul(s->L_group[i - 1], w);
    }
    const fe *src = o == 0 ? s->l_eval : s->h_eval_arr;
    fe *W = s->W[o] = (fe *)malloc(sizeof(fe) * ((size_t)1 << (bit_len + 5)));
    for (int i = 0; i < NSL; ++i)
        for (int j = 0; j < half; ++j) {
            W[((size_t)j << 7) | (i << 1) | 0] = src[(size_t)i * s->slice_size + j];
            W[((size_t)j << 7) | (i << 1) | 1] = src[(size_t)i * s->slice_size + j + half];
        }
    dig *leaf = (dig *)malloc(sizeof(dig) * half);
    const fe *msk = src + (size_t)(s->slice_count - 1) * s->slice_size;
    for (int i = 0; i < half; ++i) {
        dig t; memset(&t, 0, sizeof t);
        uint8_t data[64];
        for (int j = 0; j < 128; j += 2) {
            memcpy(data, &W[((size_t)i << 7) | j], 32);
            memcpy(data + 32, &t, 32);
            hhash(data, &t);
        }
        fe e[2] = {msk[i], msk[i + half]};
        memcpy(data, e, 32);
        memcpy(data + 32, &t, 32);
        hhash(data, &t);
        leaf[i] = t;
    }
    s->tree[o] = (dig *)malloc(sizeof(dig) * 2 * half);
    merkle_build(leaf, half, s->tree[o]);
    free(leaf);
    return s->tree[o][1];
}

/* poly_commit_prover::commit_private_array (poly_commit.h:42-124) */
static dig commit_private_array(pcs *s, const fe *priv, int n, const fe *pri_mask, int n_mask)
{
    s->n = n;
    s->slice_count = NSL + 1;
    s->slice_size = 1 << (n + 5 - 6);
    s->real = s->slice_size >> 5;
    s->l_eval = (fe *)malloc(sizeof(fe) * (size_t)s->slice_count * s->slice_size);
    int gap = s->slice_size / n_mask;
    for (int j = 0; j < 64; ++j)
        if ((1LL << j) <= gap && (1LL << (j + 1)) > gap) { gap = 1 << j; break; }
    s->gap = gap;
    int m = s->msk_len = s->slice_size / gap;
    fe *mk = (fe *)calloc(m, sizeof(fe));
    for (int i = 0; i < n_mask && i < m; ++i) mk[i] = pri_mask[i];
    fe *tmp = (fe *)malloc(sizeof(fe) * (m > s->real ? m : s->real));
    fe w_slice = root_of_unity(mylog(s->slice_size));
    for (int i = 0; i < s->slice_count; ++i) {
        fe *dst = s->l_eval + (size_t)i * s->slice_size;
        if (i == s->slice_count - 1) {
            ifft(mk, m, m, root_of_unity(mylog(m)), tmp);
            fft(tmp, m, s->slice_size, w_slice, dst);
        } else {
            const fe *src = priv + (size_t)i * s->real;
            int all_zero = 1;
            for (int j = 0; j < s->real; ++j) if (!feq(src[j], ZERO)) { all_zero = 0; break; }
            if (all_zero) memset(dst, 0, sizeof(fe) * s->slice_size);
            else {
                ifft(src, s->real, s->real, root_of_unity(mylog(s->real)), tmp);
                fft(tmp, s->real, s->slice_size, w_slice, dst);
            }
        }
    }
    free(tmp); free(mk);
    return request_init_commit(s, n, 0);
}

/* poly_commit_prover::commit_public_array (poly_commit.h:125-351).
 * pub_mask must hold msk_len entries (zero padded). */
static int commit_public_array(pcs *s, const fe *pub_mask, const fe *pub_arr, int n, fe target, dig *root_h)
{
    int SS = s->slice_size, m = s->msk_len, real = s->real, gap = s->gap;
    size_t tot = (size_t)SS * s->slice_count;
    s->vo = (fe *)malloc(sizeof(fe) * tot);
    s->vo_msk = (fe *)malloc(sizeof(fe) * SS);
    s->q_eval = (fe *)malloc(sizeof(fe) * tot);
    fe w_slice = root_of_unity(mylog(SS));
    int tmpn = m > SS ? m : SS;
    fe *tmp = (fe *)malloc(sizeof(fe) * tmpn);
    ifft(pub_mask, m, m, root_of_unity(mylog(m)), tmp);
    fft(tmp, m, SS, w_slice, s->q_eval + (size_t)(s->slice_count - 1) * SS);
    for (int i = 0; i < s->slice_count - 1; ++i) {
        ifft(pub_arr + (size_t)i * real, real, real, root_of_unity(mylog(real)), tmp);
        fft(tmp, real, SS, w_slice, s->q_eval + (size_t)i * SS);
    }
    /* inner-product assertion (:179-193) */
    fe sum = ZERO;
    for (size_t i = 0; i < (size_t)(s->slice_count - 1) * real; ++i)
        sum = add(sum, mul(s->q_eval[i << 5], s->l_eval[i << 5]));
    const fe *lmsk = s->l_eval + (size_t)(s->slice_count - 1) * SS;
    for (int i = 0; i < SS; i += gap) sum = add(sum, mul(lmsk[i], pub_mask[i / gap]));
    if (!feq(sum, target)) FAIL(-10, "commit_public_array: inner product != target sum");
    int lqn = (2 * m > 2 * real) ? 2 * m : 2 * real;
    fe *lq_eval = (fe *)malloc(sizeof(fe) * lqn), *lq_coef = (fe *)malloc(sizeof(fe) * lqn);
    fe *h_coef = (fe *)malloc(sizeof(fe) * (m > real ? m : real));
    fe *h_eval = (fe *)malloc(sizeof(fe) * SS);
    s->h_eval_arr = (fe *)malloc(sizeof(fe) * tot);
    fe inv_rou = finv(w_slice);
    for (int i = 0; i < s->slice_count; ++i) {
        const fe *l = s->l_eval + (size_t)i * SS, *q = s->q_eval + (size_t)i * SS;
        if (i == s->slice_count - 1) {
            int id = 0, all_zero = 1;
            for (int j = 0; j < SS; j += gap / 2, id++) {
                lq_eval[id] = mul(l[j], q[j]);
                if (!feq(lq_eval[id], ZERO)) all_zero = 0;
            }
            if (all_zero) {
                memset(lq_coef, 0, sizeof(fe) * 2 * m); memset(h_coef, 0, sizeof(fe) * m); memset(h_eval, 0, sizeof(fe) * SS);
            } else {
                ifft(lq_eval, 2 * m, 2 * m, root_of_unity(mylog(2 * m)), lq_coef);
                for (int j = 0; j < m; ++j) h_coef[j] = lq_coef[j + m];
                fft(h_coef, m, SS, w_slice, h_eval);
            }
            fe inv_x = ONE(), rou_n = fpow(w_slice, (u128)m), x_n = ONE();
            fe c0 = add(lq_coef[0], h_coef[0]);
            for (int j = 0; j < SS; ++j) {
                fe g = sub(mul(l[j], q[j]), mul(sub(x_n, ONE()), h_eval[j]));
                fe val = mul(mul(sub(g, c0), fe_u((uint64_t)m)), inv_x);
                if (j < SS / 2) s->vo_msk[j << 1] = val;
                else s->vo_msk[((j - SS / 2) << 1) | 1] = val;
                inv_x = mul(inv_x, inv_rou);
                x_n = mul(x_n, rou_n);
            }
            s->all_sum[i] = mul(c0, fe_u((uint64_t)m));
        } else {
            int all_zero = 1, stride = SS / (2 * real);
            for (int j = 0; j < 2 * real; ++j) {
                lq_eval[j] = mul(l[(size_t)j * stride], q[(size_t)j * stride]);
                if (!feq(lq_eval[j], ZERO)) all_zero = 0;
            }
            if (all_zero) {
                memset(lq_coef, 0, sizeof(fe) * 2 * real); memset(h_coef, 0, sizeof(fe) * real); memset(h_eval, 0, sizeof(fe) * SS);
            } else {
                ifft(lq_eval, 2 * real, 2 * real, root_of_unity(mylog(2 * real)), lq_coef);
                for (int j = 0; j < real; ++j) h_coef[j] = lq_coef[j + real];
                fft(h_coef, real, SS, w_slice, h_eval);
            }
            fe rou_n = fpow(w_slice, (u128)real), x_n = ONE(), inv_x = fe_u((uint64_t)real);
            fe cs = sub(ZERO, add(lq_coef[0], h_coef[0]));
            for (int j = 0; j < SS; ++j) {
                fe g = sub(mul(l[j], q[j]), mul(sub(x_n, ONE()), h_eval[j]));
                fe val = mul(add(g, cs), inv_x);
                if (j < SS / 2) s->vo[((size_t)j << 7) | (i << 1) | 0] = val;
                else s->vo[((size_t)(j - SS / 2) << 7) | (i << 1) | 1] = val;
                inv_x = mul(inv_x, inv_rou);
                x_n = mul(x_n, rou_n);
            }
            s->all_sum[i] = mul(add(lq_coef[0], h_coef[0]), fe_u((uint64_t)real));
        }
        memcpy(s->h_eval_arr + (size_t)i * SS, h_eval, sizeof(fe) * SS);
    }
    free(tmp); free(lq_eval); free(lq_coef); free(h_coef); free(h_eval);
    *root_h = request_init_commit(s, n, 1);
    return 0;
}

/* fri::commit_phase_step (fri.cpp:291-430).  Fold pairs (i, i+S/2) of the
 * previous codeword of every slice; the interleaved index maps of the
 * reference reduce to mapping[(p<<6)|j] = ((p mod S/2)<<7)|(j<<1). */
static dig commit_phase_step(pcs *s, fe r)
{
    int k = s->n_steps;
    int S = 1 << s->log_cur, nxt = S / 2;
    const fe *prev = k == 0 ? s->vo : s->rs[k - 1];
    const fe *prev_m = k == 0 ? s->vo_msk : s->rs_msk[k - 1];
    fe inv_2 = finv(fe_u(2));
    fe *folded = (fe *)malloc(sizeof(fe) * (size_t)nxt * s->slice_count);
    for (int i = 0; i < nxt; ++i) {
        fe inv_mu = s->L_group[(S - i) & (S - 1)];
        fe imr = mul(inv_mu, r);
        for (int j = 0; j < NSL; ++j) {
            size_t rp = ((size_t)i << 7) | (j << 1);
            fe a = prev[rp], b = prev[rp | 1];
            folded[((size_t)i << 6) | j] = mul(inv_2, add(add(a, b), mul(imr, sub(a, b))));
        }
    }
    for (int i = 0; i < nxt; ++i) s->L_group[i] = s->L_group[i * 2];
    fe *t = (fe *)malloc(sizeof(fe) * (size_t)nxt * s->slice_count);
    for (int i = 0; i < nxt / 2; ++i)
        for (int j = 0; j < NSL; ++j) {
            t[((size_t)i << 7) | (j << 1) | 0] = folded[((size_t)i << 6) | j];
            t[((size_t)i << 7) | (j << 1) | 1] = folded[((size_t)(i + nxt / 2) << 6) | j];
        }
    free(folded);
    s->rs[k] = t;
    fe *fm = (fe *)malloc(sizeof(fe) * nxt);
    for (int i = 0; i < nxt; ++i) {
        fe inv_mu = s->L_group[(S - i) & (S - 1)];
        fe a = prev_m[i << 1], b = prev_m[(i << 1) + 1];
        fm[i] = mul(inv_2, add(add(a, b), mul(mul(inv_mu, r), sub(a, b))));
    }
    fe *tm = (fe *)malloc(sizeof(fe) * nxt);
    for (int i = 0; i < nxt / 2; ++i) { tm[i << 1] = fm[i]; tm[i << 1 | 1] = fm[i + nxt / 2]; }
    free(fm);
    s->rs_msk[k] = tm;
    dig *hv = (dig *)malloc(sizeof(dig) * (nxt / 2));
    for (int i = 0; i < nxt / 2; ++i) {
        dig h; memset(&h, 0, sizeof h);
        uint8_t data[64];
        for (int j = 0; j < NSL; ++j) {
            memcpy(data, &t[((size_t)i << 7) | (j << 1)], 32);
            memcpy(data + 32, &h, 32);
            hhash(data, &h);
        }
        memcpy(data, &tm[i << 1], 32);
        memcpy(data + 32, &h, 32);
        hhash(data, &h);
        hv[i] = h;
    }
    s->ctree[k] = (dig *)malloc(sizeof(dig) * nxt);
    merkle_build(hv, nxt / 2, s->ctree[k]);
    free(hv);
    s->csize[k] = nxt / 2;
    s->log_cur--;
    s->n_steps++;
    return s->ctree[k][1];
}

/* query answers: fri.cpp:148-201 and :227-289 */
static void tb_init_query(tbuf *t, pcs *s, int64_t pow0, int o)
{
    int half = 1 << (s->log_cur - 1);
    const fe *W = s->W[o];
    for (int i = 0; i < NSL; ++i) {
        tb_fe(t, W[((size_t)pow0 << 7) | (i << 1) | 0]);
        tb_fe(t, W[((size_t)pow0 << 7) | (i << 1) | 1]);
    }
    const fe *src = o == 0 ? s->l_eval : s->h_eval_arr;
    const fe *msk = src + (size_t)(s->slice_count - 1) * s->slice_size;
    tb_fe(t, msk[pow0]);
    tb_fe(t, msk[pow0 + half]);
    int depth = s->log_cur - 1;
    int64_t pos = pow0 + half;
    for (int i = 0; i < depth; ++i) { tb_dig(t, &s->tree[o][pos ^ 1]); pos /= 2; }
    tb_dig(t, &s->tree[o][pow0 + half]);
}
static void tb_step_query(tbuf *t, pcs *s, int lvl, int64_t pow)
{
    int nxt2 = s->csize[lvl];                   /* nxt/2 */
    int64_t q = pow % nxt2;
    const fe *rs = s->rs[lvl];
    for (int j = 0; j < NSL; ++j) {
        tb_fe(t, rs[((size_t)q << 7) | (j << 1)]);
        tb_fe(t, rs[((size_t)q << 7) | (j << 1) | 1]);
    }
    tb_fe(t, s->rs_msk[lvl][q << 1]);
    tb_fe(t, s->rs_msk[lvl][(q << 1) | 1]);
    int64_t p0 = q + nxt2;
    dig leaf = s->ctree[lvl][p0];
    while (p0 != 1) { tb_dig(t, &s->ctree[lvl][p0 ^ 1]); p0 /= 2; }
    tb_dig(t, &leaf);
}

/* eq(r, x) expansion, dfs_for_public_eval order (zk_verifier.cpp:838-847) */
static void eq_table(const fe *r, int len, fe *out)
{
    out[0] = ONE();
    for (int i = len - 1, sz = 1; i >= 0; --i, sz <<= 1) {
        /* dfs: bit for r[len-1-dep] is appended as the next lower bit */
        for (int x = sz - 1; x >= 0; --x) {
            fe v = out[x];
            out[2 * x] = mul(v, sub(ONE(), r[i]));
            out[2 * x + 1] = mul(v, r[i]);
        }
    }
}

/* ===================================================================== */
/* The protocol: zk_verifier::verify restated on the prover side, i.e.    */
/* the sequence of prover calls with the challenges of the schedule.      */
/* ===================================================================== */
typedef struct { fe *v; int n, cap; } fvec;
static void fv_push(fvec *f, fe x)
{
    if (f->n == f->cap) { f->cap = f->cap ? 2 * f->cap : 256; f->v = (fe *)realloc(f->v, sizeof(fe) * f->cap); }
    f->v[f->n++] = x;
}

/* public mask coefficients pushed by the verifier for layer i
 * (zk_verifier.cpp:983-1025) */
static void push_pub_masks(fvec *pub, const lsched *l, int i, fe preu1, fe prev1)
{
    int ll = l->bl;
    fe one = ONE();
    fe rc = i == 1 ? ZERO : l->r_c;
    fe rcb = l->random_combine, lc = l->linear_coeff, lcm = l->linear_combine, rho = l->rho;
    const fe *ru = l->r_u;
    fe rv[64];
    for (int k = 0; k < ll; ++k) rv[k] = i == 1 ? ZERO : l->r_v[k];
    fv_push(pub, add(mul(one, rcb), mul(mul(lc, one), rcb)));
    fv_push(pub, add(mul(preu1, rcb), mul(mul(lc, prev1), rcb)));
    fv_push(pub, add(mul(mul(preu1, preu1), rcb), mul(mul(mul(lc, prev1), prev1), rcb)));
    fv_push(pub, add(mul(rc, rcb), mul(mul(lc, rc), rcb)));
    fv_push(pub, add(mul(mul(rc, rc), rcb), mul(mul(mul(lc, rc), rc), rcb)));
    fv_push(pub, add(mul(mul(preu1, rc), rcb), mul(mul(mul(lc, prev1), rc), rcb)));
    for (int k = 0; k < ll; ++k) {
        fv_push(pub, mul(mul(mul(lcm, ru[k]), ru[k]), rho));
        fv_push(pub, mul(mul(lcm, ru[k]), rho));
    }
    for (int k = 0; k < ll; ++k) {
        fv_push(pub, mul(mul(mul(lcm, rv[k]), rv[k]), rho));
        fv_push(pub, mul(mul(lcm, rv[k]), rho));
    }
    fv_push(pub, mul(mul(mul(lcm, rc), rc), rho));
    fv_push(pub, mul(mul(lcm, rc), rho));
    fv_push(pub, mul(mul(lcm, one), rho));
    fe x = ru[ll - 1], x2 = mul(x, x), x3 = mul(x2, x), x4 = mul(x3, x), x5 = mul(x4, x);
    fv_push(pub, mul(mul(lcm, x5), rho));
    fv_push(pub, mul(mul(lcm, x4), rho));
    fv_push(pub, mul(mul(lcm, x3), rho));
    x = rv[ll - 1]; x2 = mul(x, x); x3 = mul(x2, x); x4 = mul(x3, x); x5 = mul(x4, x);
    fv_push(pub, mul(mul(lcm, x5), rho));
    fv_push(pub, mul(mul(lcm, x4), rho));
    fv_push(pub, mul(mul(lcm, x3), rho));
}

static void tb_q(tbuf *t, qpoly p) { tb_fe(t, p.a); tb_fe(t, p.b); tb_fe(t, p.c); }
static void tb_q5(tbuf *t, q5poly p) { tb_fe(t, p.a); tb_fe(t, p.b); tb_fe(t, p.c); tb_fe(t, p.d); tb_fe(t, p.e); tb_fe(t, p.f); }

static void fg_prove(int lg, tbuf *t, double *p_sec);

int vgo_prove(const vgo_circuit_desc *dsc, uint8_t **out, size_t *out_len, double *times)
{
    circuit C;
    int rc = circ_build(dsc, &C);
    if (rc) { circ_free(&C); return rc; }
    sched S;
    sched_build(&C, &S);
    zkp P;
    zkp_init(&P, &C);
    tbuf T = {0};
    double t0 = now_s();
    P.cv = circ_evaluate(&C);
    double t1 = now_s();
    tb_put(&T, "VGT1", 4);
    tb_u32(&T, C.depth);
    tb_u32(&T, C.L[0].bl);
    for (int i = 0; i < C.depth; ++i) tb_u32(&T, C.L[i].bl);

    int bl_out = C.L[C.depth - 1].bl;
    fe alpha = ONE(), beta = ZERO;
    fe *r0 = (fe *)malloc(sizeof(fe) * 64), *r1 = (fe *)malloc(sizeof(fe) * 64);
    fe *o0 = (fe *)malloc(sizeof(fe) * 64), *o1 = (fe *)malloc(sizeof(fe) * 64);
    for (int k = 0; k < bl_out; ++k) { r0[k] = S.r0[k]; r1[k] = S.r1[k]; o0[k] = sub(ONE(), r0[k]); o1[k] = sub(ONE(), r1[k]); }
    tb_fe(&T, V_res(o0, r0, P.cv[C.depth - 1], bl_out, 1 << bl_out));
    fvec pub = {0};
    for (int i = C.depth - 1; i >= 1; --i) {
        const lsched *l = &S.L[i];
        int bl = l->bl;
        sumcheck_init(&P, l, i, C.L[i].bl, bl, bl, alpha, beta, r0, r1, o0, o1);
        gen_maskpoly_pre_rho(&P, l, bl * 2 + 1, 2);
        P.rho = l->rho;
        gen_maskpoly_after_rho(&P, bl * 2 + 1, 2);
        tb_fe(&T, P.maskpoly_sumc);
        phase1_init(&P);
        fe ru[64], rv[64], oru[64], orv[64];
        for (int k = 0; k < bl; ++k) {
            ru[k] = l->r_u[k]; rv[k] = i == 1 ? ZERO : l->r_v[k];
            oru[k] = sub(ONE(), ru[k]); orv[k] = sub(ONE(), rv[k]);
        }
        fe rcv = i == 1 ? ZERO : l->r_c;
        push_pub_masks(&pub, l, i, P.prepreu1, P.preprev1);
        fe prev = ZERO;
        for (int j = 0; j < bl; ++j) {
            if (j == bl - 1) tb_q5(&T, phase1_updatelastbit(&P, prev, j));
            else tb_q(&T, phase1_update(&P, prev, j));
            prev = ru[j];
        }
        phase2_init(&P, prev, ru, oru);
        tb_fe(&T, P.v_u);
        prev = ZERO;
        for (int j = 0; j < bl; ++j) {
            if (j == bl - 1) tb_q5(&T, phase2_updatelastbit(&P, prev, j));
            else tb_q(&T, phase2_update(&P, prev, j));
            prev = rv[j];
        }
        finalize(&P, prev);
        tb_fe(&T, P.v_v);
        fe pv[14];
        predicates(&C, i, r0, r1, o0, o1, ru, rv, oru, orv, alpha, beta, pv);
        tb_q(&T, finalround(&P, prev, bl << 1, general_value(pv, P.v_u, P.v_v)));
        tb_fe(&T, query(&P, ru, rv, rcv));
        tb_fe(&T, q_eval(P.Rg1, rcv));
        tb_fe(&T, q_eval(P.Rg2, rcv));
        tb_fe(&T, P.Iuv); tb_fe(&T, P.preZu); tb_fe(&T, P.preZv);
        alpha = l->alpha; beta = l->beta;
        for (int k = 0; k < bl; ++k) { r0[k] = ru[k]; r1[k] = rv[k]; o0[k] = oru[k]; o1[k] = orv[k]; }
    }
    /* post sumcheck (zk_verifier.cpp:1202-1207) */
    fv_push(&pub, P.Zu);
    fv_push(&pub, mul(P.Zu, P.preu1));
    pri_push(&P, P.maskr[0]);
    pri_push(&P, P.maskr[1]);
    tb_fe(&T, P.Zu); tb_fe(&T, P.preu1); tb_q(&T, P.sumRc);
    double t2 = now_s();

    /* commitment */
    int n = C.L[0].bl;
    pcs PC;
    memset(&PC, 0, sizeof PC);
    dig rootL = commit_private_array(&PC, P.cv[0], n, P.pri, P.n_pri);
    fe *q = (fe *)malloc(sizeof(fe) << n);
    eq_table(r0, n, q);
    /* target_sum (zk_verifier.cpp:1216) is recomputed by the prover's own
     * assertion; it needs the verifier's running sums, which this side
     * does not track, so it is checked in vgo_verify instead. */
    fe *pubm = (fe *)calloc(PC.msk_len, sizeof(fe));
    for (int k = 0; k < pub.n && k < PC.msk_len; ++k) pubm[k] = pub.v[k];
    /* inner product = sum l*q + pri.pub is what the assertion compares */
    fe target = ZERO;
    {
        for (size_t k = 0; k < ((size_t)1 << n); ++k) target = add(target, mul(P.cv[0][k], q[k]));
        for (int k = 0; k < P.n_pri && k < PC.msk_len; ++k) target = add(target, mul(P.pri[k], pubm[k]));
    }
    dig rootH;
    rc = commit_public_array(&PC, pubm, q, n, target, &rootH);
    if (rc) goto done;
    tb_dig(&T, &rootL); tb_dig(&T, &rootH);
    for (int k = 0; k < 65; ++k) tb_fe(&T, PC.all_sum[k]);
    /* commit_phase (vpd_verifier.cpp:46-78) */
    {
        int save = PC.log_cur;
        for (int k = 0; k < S.n_fri; ++k) {
            dig rt = commit_phase_step(&PC, S.fri_r[k]);
            tb_dig(&T, &rt);
        }
        const fe *fin = PC.rs[PC.n_steps - 1];
        for (int k = 0; k < 2048; ++k) tb_fe(&T, fin[k]);
        for (int k = 0; k < 32; ++k) tb_fe(&T, PC.rs_msk[PC.n_steps - 1][k]);
        PC.log_cur = save;
    }
    double t3 = now_s();
    for (int rep = 0; rep < 33; ++rep) {
        int64_t pow = S.pow[rep];
        int64_t s0 = pow / 2;
        tb_init_query(&T, &PC, s0, 0);
        tb_init_query(&T, &PC, s0, 1);
        for (int lvl = 0; lvl < n - 6; ++lvl) {
            int64_t pw = pow % (1LL << (n - 1 - lvl));
            tb_step_query(&T, &PC, lvl, pw / 2);
        }
    }
    double t4 = now_s();
    /* fft_gkr section (vpd_verifier.cpp:90-99: "./fft_gkr n-6", own rand stream) */
    fg_prove(n - 6, &T, NULL);
    double t5 = now_s();
    if (times) { times[0] = t1 - t0; times[1] = t2 - t1; times[2] = t3 - t2; times[3] = t4 - t3; times[4] = t5 - t4; }
    *out = T.p; *out_len = T.n; T.p = NULL;
done:
    free(q); free(pubm); free(pub.v);
    free(r0); free(r1); free(o0); free(o1);
    pcs_free(&PC);
    zkp_free(&P);
    sched_free(&S);
    circ_free(&C);
    free(T.p);
    return rc;
}

/* ===================================================================== */
/* Verifier on a transcript (zk_verifier.cpp:910-1253,                    */
/* vpd_verifier.cpp:8-41, 80-335).                                        */
/* ===================================================================== */
#define REJ(...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); res = 0; goto vdone; } while (0)
static qpoly tr_q(trd *r) { qpoly p; p.a = tr_fe(r); p.b = tr_fe(r); p.c = tr_fe(r); return p; }
static q5poly tr_q5(trd *r) { q5poly p; p.a = tr_fe(r); p.b = tr_fe(r); p.c = tr_fe(r); p.d = tr_fe(r); p.e = tr_fe(r); p.f = tr_fe(r); return p; }

static int verify_merkle(dig root, const dig *path, int len, int64_t pow, const fe *vals /* 65 pairs */)
{                                                  /* vpd_verifier.cpp:8-41 */
    dig cur = path[len - 1];
    uint8_t data[64];
    for (int i = 0; i < len - 1; ++i) {
        if (pow & 1) { memcpy(data, &path[i], 32); memcpy(data + 32, &cur, 32); }
        else { memcpy(data, &cur, 32); memcpy(data + 32, &path[i], 32); }
        pow /= 2;
        hhash(data, &cur);
    }
    dig vh; memset(&vh, 0, sizeof vh);
    for (int i = 0; i < 65; ++i) {
        memcpy(data, &vals[2 * i], 32);
        memcpy(data + 32, &vh, 32);
        hhash(data, &vh);
    }
    return memcmp(&root, &cur, 32) == 0 && memcmp(&vh, &path[len - 1], 32) == 0;
}

int vgo_verify(const vgo_circuit_desc *dsc, const uint8_t *tr, size_t len)
{
    circuit C;
    int res = 1;
    if (circ_build(dsc, &C)) { circ_free(&C); return -1; }
    sched S;
    sched_build(&C, &S);
    trd R = {tr, len, 0, 0};
    fvec pub = {0};
    fe *r0 = (fe *)malloc(sizeof(fe) * 64), *r1 = (fe *)malloc(sizeof(fe) * 64);
    fe *o0 = (fe *)malloc(sizeof(fe) * 64), *o1 = (fe *)malloc(sizeof(fe) * 64);
    fe *qcoef = NULL, *pubcoef = NULL, *allpath = NULL, *vals = NULL;
    dig *path = NULL;
    char magic[4];
    tr_get(&R, magic, 4);
    uint32_t hd, hn;
    tr_get(&R, &hd, 4); tr_get(&R, &hn, 4);
    if (memcmp(magic, "VGT1", 4) || hd != (uint32_t)C.depth || hn != (uint32_t)C.L[0].bl) REJ("bad transcript header");
    for (int i = 0; i < C.depth; ++i) { uint32_t b; tr_get(&R, &b, 4); if (b != (uint32_t)C.L[i].bl) REJ("bit length mismatch"); }
    int bl_out = C.L[C.depth - 1].bl;
    fe alpha = ONE(), beta = ZERO;
    for (int k = 0; k < bl_out; ++k) { r0[k] = S.r0[k]; r1[k] = S.r1[k]; o0[k] = sub(ONE(), r0[k]); o1[k] = sub(ONE(), r1[k]); }
    fe a0 = mul(alpha, tr_fe(&R));
    fe abs_ = a0, all_mask_sum = ZERO;
    fe preu1_prev = ZERO, prev1_prev = ZERO;   /* prover's preu1/prev1 before the layer */
    for (int i = C.depth - 1; i >= 1; --i) {
        const lsched *l = &S.L[i];
        int bl = l->bl;
        abs_ = add(abs_, tr_fe(&R));           /* maskpoly_sumc */
        fe ru[64], rv[64], oru[64], orv[64];
        for (int k = 0; k < bl; ++k) {
            ru[k] = l->r_u[k]; rv[k] = i == 1 ? ZERO : l->r_v[k];
            oru[k] = sub(ONE(), ru[k]); orv[k] = sub(ONE(), rv[k]);
        }
        fe drv = add(mul(alpha, direct_relay(&C, i, r0, ru)), mul(beta, direct_relay(&C, i, r1, ru)));
        fe rc = i == 1 ? ZERO : l->r_c;
        push_pub_masks(&pub, l, i, preu1_prev, prev1_prev);
        for (int j = 0; j < bl; ++j) {
            if (j == bl - 1) {
                q5poly p = tr_q5(&R);
                if (!feq(add(q5_eval(p, ZERO), q5_eval(p, ONE())), abs_)) REJ("phase1 check failed, layer %d bit %d", i, j);
                abs_ = q5_eval(p, ru[j]);
            } else {
                qpoly p = tr_q(&R);
                if (!feq(add(q_eval(p, ZERO), q_eval(p, ONE())), abs_)) REJ("phase1 check failed, layer %d bit %d", i, j);
                abs_ = q_eval(p, ru[j]);
            }
        }
        fe vu = tr_fe(&R);
        for (int j = 0; j < bl; ++j) {
            if (j == bl - 1) {
                q5poly p = tr_q5(&R);
                if (!feq(add(add(q5_eval(p, ZERO), q5_eval(p, ONE())), mul(drv, vu)), abs_)) REJ("phase2 check failed, layer %d bit %d", i, j);
                abs_ = add(q5_eval(p, rv[j]), mul(drv, vu));
            } else {
                qpoly p = tr_q(&R);
                if (!feq(add(add(q_eval(p, ZERO), q_eval(p, ONE())), mul(drv, vu)), abs_)) REJ("phase2 check failed, layer %d bit %d", i, j);
                abs_ = add(q_eval(p, rv[j]), mul(drv, vu));
            }
        }
        fe vv = tr_fe(&R);
        fe pv[14];
        predicates(&C, i, r0, r1, o0, o1, ru, rv, oru, orv, alpha, beta, pv);
        fe gv = general_value(pv, vu, vv);
        qpoly fp = tr_q(&R);
        if (!feq(add(add(q_eval(fp, ZERO), q_eval(fp, ONE())), mul(drv, vu)), abs_)) REJ("final round check failed, layer %d", i);
        abs_ = add(q_eval(fp, rc), mul(drv, vu));
        fe mpv = tr_fe(&R), rg1 = tr_fe(&R), rg2 = tr_fe(&R);
        all_mask_sum = add(add(add(all_mask_sum, mul(rg1, l->random_combine)), mul(mul(l->linear_coeff, rg2), l->random_combine)),
                           mul(l->linear_combine, mpv));
        fe Iuv = tr_fe(&R), pZu = tr_fe(&R), pZv = tr_fe(&R);
        fe rhs = add(add(add(add(mul(rc, gv), mul(mul(mul(alpha, Iuv), pZu), rg1)), mul(mul(mul(beta, Iuv), pZv), rg2)), mpv), mul(drv, vu));
        if (!feq(abs_, rhs)) REJ("semi-final check failed, layer %d", i);
        alpha = l->alpha; beta = l->beta;
        abs_ = i != 1 ? add(mul(alpha, vu), mul(beta, vv)) : vu;
        for (int k = 0; k < bl; ++k) { r0[k] = ru[k]; r1[k] = rv[k]; o0[k] = oru[k]; o1[k] = orv[k]; }
        preu1_prev = ru[bl - 1];
        prev1_prev = rv[bl - 1];
    }
    fe Zu = tr_fe(&R), preu1 = tr_fe(&R);
    qpoly sumRc = tr_q(&R);
    fv_push(&pub, Zu);
    fv_push(&pub, mul(Zu, preu1));
    all_mask_sum = add(all_mask_sum, mul(Zu, q_eval(sumRc, preu1)));
    fe target = add(sub(abs_, mul(Zu, q_eval(sumRc, preu1))), all_mask_sum);

    /* ---- VPD ---- */
    int n = C.L[0].bl;
    int SS = 1 << (n - 1), real = 1 << (n - 6), cs = real;
    int gap = SS / pub.n;
    for (int j = 0; j < 64; ++j) if ((1LL << j) <= gap && (1LL << (j + 1)) > gap) { gap = 1 << j; break; }
    int m = SS / gap;
    while (pub.n < m) fv_push(&pub, ZERO);
    dig rootL = tr_dig(&R), rootH = tr_dig(&R);
    fe all_sum[65];
    fe tot = ZERO;
    for (int k = 0; k < 65; ++k) { all_sum[k] = tr_fe(&R); tot = add(tot, all_sum[k]); }
    /* Extra check, not in the reference verifier: the prover asserts that
     * the slice sums add up to the target (poly_commit.h:193). */
    if (!feq(tot, target)) REJ("sum of all_sum != inner-product target");
    int nf = n - 6;
    dig froot[64];
    for (int k = 0; k < nf; ++k) froot[k] = tr_dig(&R);
    fe *fin = (fe *)malloc(sizeof(fe) * (2048 + 32));
    for (int k = 0; k < 2080; ++k) fin[k] = tr_fe(&R);
    /* public_array_prepare (zk_verifier.cpp:884-908) */
    qcoef = (fe *)malloc(sizeof(fe) << n);
    {
        fe *qv = (fe *)malloc(sizeof(fe) * cs), *qr = (fe *)malloc(sizeof(fe) * 65), *qc = (fe *)malloc(sizeof(fe) * cs);
        eq_table(r0 + n - 6, 6, qr);
        qr[64] = ONE();
        eq_table(r0, n - 6, qv);
        ifft(qv, cs, cs, root_of_unity(n - 6), qc);
        for (int i = 0; i < 64; ++i)
            for (int j = 0; j < cs; ++j) qcoef[(size_t)i * cs + j] = mul(qc[j], qr[i]);
        free(qv); free(qr); free(qc);
    }
    pubcoef = (fe *)malloc(sizeof(fe) * m);
    ifft(pub.v, m, m, root_of_unity(mylog(m)), pubcoef);
    fe inv_2 = finv(fe_u(2));
    vals = (fe *)malloc(sizeof(fe) * 130 * 2);
    path = (dig *)malloc(sizeof(dig) * 64);
    for (int rep = 0; rep < 33; ++rep) {
        int64_t pow = S.pow[rep];
        fe rou = root_of_unity(n - 1);
        fe alpha_v[65][2];
        fe beta_v[65][2];
        for (int i = 0; i < nf; ++i) {
            int64_t mod = 1LL << (n - 1 - i);
            if (i > 0) { rou = mul(rou, rou); pow = pow % mod; }
            int64_t s0p = pow / 2, s1p = (pow + mod) / 2;
            if (i == 0) {
                fe al[130], ah[130];
                dig pl[64], ph[64];
                for (int k = 0; k < 130; ++k) al[k] = tr_fe(&R);
                for (int k = 0; k < n - 1; ++k) pl[k] = tr_dig(&R);
                for (int k = 0; k < 130; ++k) ah[k] = tr_fe(&R);
                for (int k = 0; k < n - 1; ++k) ph[k] = tr_dig(&R);
                int64_t mn = s0p < s1p ? s0p : s1p;
                if (!verify_merkle(rootL, pl, n - 1, mn, al)) REJ("merkle L failed rep %d", rep);
                if (!verify_merkle(rootH, ph, n - 1, mn, ah)) REJ("merkle H failed rep %d", rep);
                for (int k = 0; k < 130; ++k) vals[k] = tr_fe(&R);
                int plen = (n - 3) + 1;
                for (int k = 0; k < plen; ++k) path[k] = tr_dig(&R);
                if (!verify_merkle(froot[0], path, plen, pow / 2, vals)) REJ("merkle step 0 failed rep %d", rep);
                fe inv_mu = finv(fpow(rou, (u128)(pow / 2)));
                fe w = root_of_unity(mylog(SS));
                fe x0 = fpow(w, (u128)s0p), x1 = fpow(w, (u128)s1p);
                fe mr0 = fpow(x0, (u128)(SS / gap)), mr1 = fpow(x1, (u128)(SS / gap));
                fe ro0 = fpow(x0, (u128)(SS >> 5)), ro1 = fpow(x1, (u128)(SS >> 5));
                fe ix0 = finv(x0), ix1 = finv(x1);
                fe qm0 = ZERO, qm1 = ZERO, xx0 = ONE(), xx1 = ONE();
                for (int k = 0; k < m; ++k) {
                    qm0 = add(qm0, mul(xx0, pubcoef[k])); xx0 = mul(xx0, x0);
                    qm1 = add(qm1, mul(xx1, pubcoef[k])); xx1 = mul(xx1, x1);
                }
                for (int j = 0; j < 65; ++j) {
                    fe a, b;
                    if (j == 64) {
                        a = sub(mul(al[2 * j], qm0), mul(sub(mr0, ONE()), ah[2 * j]));
                        a = mul(sub(mul(a, fe_u((uint64_t)(SS / gap))), all_sum[j]), ix0);
                        b = sub(mul(al[2 * j + 1], qm1), mul(sub(mr1, ONE()), ah[2 * j + 1]));
                        b = mul(sub(mul(b, fe_u((uint64_t)(SS / gap))), all_sum[j]), ix1);
                    } else {
                        fe t0 = ZERO, t1 = ZERO;
                        xx0 = ONE(); xx1 = ONE();
                        for (int k = 0; k < cs; ++k) {
                            t0 = add(t0, mul(xx0, qcoef[(size_t)k + (size_t)j * cs])); xx0 = mul(xx0, x0);
                            t1 = add(t1, mul(xx1, qcoef[(size_t)k + (size_t)j * cs])); xx1 = mul(xx1, x1);
                        }
                        a = sub(mul(al[2 * j], t0), mul(sub(ro0, ONE()), ah[2 * j]));
                        a = mul(sub(mul(a, fe_u((uint64_t)(SS >> 5))), all_sum[j]), ix0);
                        b = sub(mul(al[2 * j + 1], t1), mul(sub(ro1, ONE()), ah[2 * j + 1]));
                        b = mul(sub(mul(b, fe_u((uint64_t)(SS >> 5))), all_sum[j]), ix1);
                    }
                    if (s0p > s1p) { fe t = a; a = b; b = t; }
                    fe pval = add(mul(add(a, b), inv_2), mul(mul(mul(mul(sub(a, b), inv_2), S.fri_r[i]), inv_mu), ONE()));
                    if (!feq(pval, vals[2 * j]) && !feq(pval, vals[2 * j + 1])) REJ("FRI first-round consistency failed rep %d slice %d", rep, j);
                }
            } else {
                for (int j = 0; j < 65; ++j) { alpha_v[j][0] = beta_v[j][0]; alpha_v[j][1] = beta_v[j][1]; }
                for (int k = 0; k < 130; ++k) vals[k] = tr_fe(&R);
                int plen = (n - 3 - i) + 1;
                for (int k = 0; k < plen; ++k) path[k] = tr_dig(&R);
                if (!verify_merkle(froot[i], path, plen, pow / 2, vals)) REJ("merkle step %d failed rep %d", i, rep);
                fe inv_mu = finv(fpow(rou, (u128)(pow / 2)));
                for (int j = 0; j < 65; ++j) {
                    fe a = alpha_v[j][0], b = alpha_v[j][1];
                    fe h0 = mul(add(a, b), inv_2);
                    fe p0 = add(h0, mul(mul(mul(sub(a, b), inv_2), S.fri_r[i]), inv_mu));
                    fe p1 = add(h0, mul(mul(mul(sub(b, a), inv_2), S.fri_r[i]), inv_mu));
                    fe v0 = vals[2 * j], v1 = vals[2 * j + 1];
                    if (!feq(p0, v0) && !feq(p0, v1) && !feq(p1, v0) && !feq(p1, v1)) REJ("FRI consistency failed rep %d round %d", rep, i);
                }
            }
            for (int j = 0; j < 65; ++j) { beta_v[j][0] = vals[2 * j]; beta_v[j][1] = vals[2 * j + 1]; }
        }
        /* last rs code check (vpd_verifier.cpp:315-332) */
        for (int i = 0; i < 64; ++i) {
            fe tpl = fin[(0 << 7) | (i << 1)];
            for (int j = 0; j < 16; ++j) if (!feq(fin[(j << 7) | (i << 1)], tpl)) REJ("final rs code not constant");
        }
        for (int j = 1; j < 32; ++j) if (!feq(fin[2048 + j], fin[2048 + j - 1])) REJ("final mask rs code not constant");
    }
    free(fin);
    fin = NULL;
    if (R.bad || R.off > R.n) REJ("transcript truncated or not canonical");
    if (!vgo_fft_gkr_verify(n - 6, R.p + R.off, R.n - R.off)) REJ("fft_gkr section rejected");
    R.off = R.n;
vdone:
    free(qcoef); free(pubcoef); free(vals); free(path); free(allpath);
    free(pub.v); free(r0); free(r1); free(o0); free(o1);
    sched_free(&S);
    circ_free(&C);
    return res;
}

/* ===================================================================== */
/* exported primitives                                                    */
/* ===================================================================== */
void vgo_fe_mul(const fe *a, const fe *b, fe *o, size_t n) { for (size_t i = 0; i < n; ++i) o[i] = mul(a[i], b[i]); }
void vgo_fe_add(const fe *a, const fe *b, fe *o, size_t n) { for (size_t i = 0; i < n; ++i) o[i] = add(a[i], b[i]); }
void vgo_fe_sub(const fe *a, const fe *b, fe *o, size_t n) { for (size_t i = 0; i < n; ++i) o[i] = sub(a[i], b[i]); }
void vgo_fe_inv(const fe *a, fe *o, size_t n) { for (size_t i = 0; i < n; ++i) o[i] = finv(a[i]); }
void vgo_root_of_unity(int lo, fe *o) { *o = root_of_unity(lo); }
void vgo_rand_seq(uint32_t seed, int32_t *out, size_t n)
{
    grand_t g; grand_init(&g, seed);
    for (size_t i = 0; i < n; ++i) out[i] = (int32_t)grand_next(&g);
}
void vgo_fft(const fe *c, int cl, int order, const fe *rou, fe *out) { fft(c, cl, order, *rou, out); }
void vgo_ifft(const fe *e, int cl, int order, const fe *rou, fe *out) { ifft(e, cl, order, *rou, out); }
void vgo_sha3_64(const uint8_t *in, uint8_t *out, size_t n) { for (size_t i = 0; i < n; ++i) hhash(in + 64 * i, out + 32 * i); }
int vgo_merkle_tree(const dig *leaves, int n_leaf, dig *tree) { return merkle_build(leaves, n_leaf, tree); }
void vgo_lde_slice(const fe *vals, int log_real, int log_slice, fe *out)
{
    int real = 1 << log_real;
    fe *tmp = (fe *)malloc(sizeof(fe) * real);
    ifft(vals, real, real, root_of_unity(log_real), tmp);
    fft(tmp, real, 1 << log_slice, root_of_unity(log_slice), out);
    free(tmp);
}
void vgo_leaf_hashes(const fe *sl, const fe *msk, int log_S, dig *leaves)
{
    int S = 1 << log_S, half = S / 2;
    uint8_t data[64];
    for (int i = 0; i < half; ++i) {
        dig t; memset(&t, 0, sizeof t);
        for (int j = 0; j < NSL; ++j) {
            fe e[2] = {sl[(size_t)j * S + i], sl[(size_t)j * S + i + half]};
            memcpy(data, e, 32); memcpy(data + 32, &t, 32);
            hhash(data, &t);
        }
        fe e[2] = {msk[i], msk[i + half]};
        memcpy(data, e, 32); memcpy(data + 32, &t, 32);
        hhash(data, &t);
        leaves[i] = t;
    }
}
void vgo_fri_fold(const fe *sl, int log_S, const fe *r, fe *out)
{
    int S = 1 << log_S, half = S / 2;
    fe w = root_of_unity(log_S), inv_2 = finv(fe_u(2));
    fe iw = finv(w), imu = ONE();
    for (int i = 0; i < half; ++i) {
        fe imr = mul(imu, *r);
        for (int j = 0; j < 65; ++j) {
            fe a = sl[(size_t)j * S + i], b = sl[(size_t)j * S + i + half];
            out[(size_t)j * half + i] = mul(inv_2, add(add(a, b), mul(imr, sub(a, b))));
        }
        imu = mul(imu, iw);
    }
}
/* one slice of the fold above (fri.cpp:321-340): a natural-order codeword of
 * 2^log_S values over the 2^log_S-th roots of unity */
void vgo_fri_fold1(const fe *cw, int log_S, const fe *r, fe *out)
{
    int64_t S = (int64_t)1 << log_S, half = S / 2;
    fe iw = finv(root_of_unity(log_S)), inv_2 = finv(fe_u(2)), imu = ONE();
    for (int64_t i = 0; i < half; ++i) {
        fe a = cw[i], b = cw[i + half];
        out[i] = mul(inv_2, add(add(a, b), mul(mul(imu, *r), sub(a, b))));
        imu = mul(imu, iw);
    }
}
void vgo_eq_table(const fe *r, int len, fe *out) { eq_table(r, len, out); }
int vgo_evaluate(const vgo_circuit_desc *dsc, fe *out, size_t cap, int *bits)
{
    circuit C;
    int rc = circ_build(dsc, &C);
    if (rc) { circ_free(&C); return rc; }
    fe **cv = circ_evaluate(&C);
    int bl = C.L[C.depth - 1].bl;
    *bits = bl;
    if (((size_t)1 << bl) <= cap) memcpy(out, cv[C.depth - 1], sizeof(fe) << bl);
    for (int i = 0; i < C.depth; ++i) free(cv[i]);
    free(cv);
    circ_free(&C);
    return 0;
}
void vgo_free(void *p) { free(p); }

/* ===================================================================== */
/* fft_gkr: the GKR proof of the commitment's FFT/evaluation circuit      */
/* (fft_circuit_GKR.cpp:22-825).  The reference runs it as its own        */
/* process (vpd_verifier.cpp:90-99) with lg = n - 6, so it draws from a   */
/* fresh rand() stream (seed 1).  Prover messages plus the program's own  */
/* inline checks (its `result`).                                          */
/*                                                                        */
/* Transcript section: "VGF1", u32 lg, F a_0 (claim on the 64 outputs),   */
/*   addition layer: (lg+6) x {F a,b,c}, F v_u                            */
/*   mult layer:      lg x {F a,b,c}, F v_u                               */
/*   per dep = 0..lg-1: lg x {F a,b,c}, F v_u, lg x {F a,b,c}, F v_v      */
/* ===================================================================== */
typedef struct {
    int lg, nl;
    fe *val[64];
    int64_t size[64];
    fe eval_points[64];
    fe rou, inv_rou, inv_n;
} fgc;

static void fg_free(fgc *c)
{
    for (int i = 0; i < c->nl; ++i) free(c->val[i]);
    c->nl = 0;
}
static fe *fg_push(fgc *c, int64_t n)
{
    fe *p = (fe *)calloc((size_t)n, sizeof(fe));
    c->val[c->nl] = p;
    c->size[c->nl] = n;
    c->nl++;
    return p;
}

/* build_circuit (fft_circuit_GKR.cpp:22-112) */
static void fg_build(fgc *c, int lg, const fe *r, grand_t *g)
{
    c->lg = lg; c->nl = 0;
    fe *l0 = fg_push(c, 1);
    l0[0] = ONE();
    for (int i = 0; i < lg; ++i) {
        fe *prev = c->val[i], *cur = fg_push(c, (int64_t)2 << i);
        for (int j = 0; j < (1 << i); ++j) {
            cur[j << 1] = mul(prev[j], r[i]);
            cur[j << 1 | 1] = mul(prev[j], sub(ONE(), r[i]));
        }
    }
    c->rou = root_of_unity(lg);
    c->inv_rou = finv(c->rou);
    c->inv_n = fpow(fe_u((uint64_t)1 << lg), P61 - 2);
    fe rot[62];
    rot[0] = c->inv_rou;
    for (int i = 1; i < 62; ++i) rot[i] = mul(rot[i - 1], rot[i - 1]);
    int64_t N = (int64_t)1 << lg;
    fe *x_arr = (fe *)malloc(sizeof(fe) * N);
    for (int dep = lg - 1; dep >= 0; --dep) {          /* :44-75 */
        int blk = 1 << (lg - dep), half = blk >> 1;
        fe *pre = c->val[c->nl - 1], *cur = fg_push(c, N);
        x_arr[0] = ONE();
        for (int j = 1; j < blk; ++j) x_arr[j] = mul(x_arr[j - 1], rot[dep]);
        for (int k = 0; k < half; ++k) {
            int dk = k & (half - 1);
            for (int j = 0; j < (1 << dep); ++j) {
                fe lv = pre[(int64_t)dk << (dep + 1) | j];
                fe rv = mul(x_arr[k], pre[(int64_t)dk << (dep + 1) | (1 << dep) | j]);
                cur[(int64_t)k << dep | j] = add(lv, rv);
                cur[(int64_t)(k + half) << dep | j] = sub(lv, rv);
            }
        }
    }
    free(x_arr);
    {                                                  /* :76-82 */
        fe *pre = c->val[c->nl - 1], *cur = fg_push(c, N);
        for (int64_t i = 0; i < N; ++i) cur[i] = mul(pre[i], c->inv_n);
    }
    {                                                  /* :84-102 */
        fe *pre = c->val[c->nl - 1], *cur = fg_push(c, 64 * N);
        for (int i = 0; i < 64; ++i) {
            fe ep = grand_fe(g), x = ONE();
            c->eval_points[i] = ep;
            for (int64_t j = 0; j < N; ++j) {
                cur[j + ((int64_t)i << lg)] = mul(pre[j], x);
                x = mul(x, ep);
            }
        }
    }
    {                                                  /* :103-111 */
        fe *pre = c->val[c->nl - 1], *cur = fg_push(c, 64);
        for (int i = 0; i < 64; ++i) {
            fe s = ZERO;
            for (int64_t j = 0; j < N; ++j) s = add(s, pre[j + (int64_t)i * N]);
            cur[i] = s;
        }
    }
}

typedef struct {
    lpoly *V, *aV, *am;
    fe *b0f, *b0s, *b1f, *b1s, *buf, *bus;
    fe alpha, beta;
} fgp;

/* sumcheck_phase1_update / phase2_update (fft_circuit_GKR.cpp:162-246):
 * the same fold and round sums on the three tables */
static qpoly fg_update(fgp *p, fe r, int cb, int64_t total)
{
    qpoly q = {ZERO, ZERO, ZERO};
    for (int64_t i = 0; i < (total >> 1); ++i) {
        int64_t g0 = i << 1, g1 = g0 | 1;
        lpoly *T[3] = {p->V, p->aV, p->am};
        for (int t = 0; t < 3; ++t) {
            lpoly *X = T[t];
            if (cb == 0) {
                X[i].b = X[g0].b; X[i].a = sub(X[g1].b, X[i].b);
            } else {
                X[i].b = add(mul(X[g0].a, r), X[g0].b);
                X[i].a = sub(add(mul(X[g1].a, r), X[g1].b), X[i].b);
            }
        }
        q.a = add(q.a, mul(p->am[i].a, p->V[i].a));
        q.b = add(add(add(q.b, mul(p->am[i].a, p->V[i].b)), mul(p->am[i].b, p->V[i].a)), p->aV[i].a);
        q.c = add(add(q.c, mul(p->am[i].b, p->V[i].b)), p->aV[i].b);
    }
    return q;
}
static void fg_refresh(grand_t *g, fe *r, fe *omr, int n)        /* :114-121 */
{
    for (int i = 0; i < n; ++i) { r[i] = grand_fe(g); omr[i] = sub(ONE(), r[i]); }
}
static inline fe fg_bg(const fgp *p, int64_t g, int fh)
{
    int64_t m = ((int64_t)1 << fh) - 1;
    return add(mul(p->b0f[g & m], p->b0s[g >> fh]), mul(p->b1f[g & m], p->b1s[g >> fh]));
}
static void fg_tables_init(fgp *p, const fe *V, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        p->V[i].a = ZERO; p->V[i].b = V[i];
        p->aV[i].a = p->aV[i].b = ZERO;
        p->am[i].a = p->am[i].b = ZERO;
    }
}
/* a phase of `bits` rounds (:330-340, :426-436, :604-614, :660-670);
 * returns V[0] at the last challenge (v_u or v_v) */
static fe fg_phase(fgp *p, tbuf *t, const fe *rr, int bits)
{
    fe prev = ZERO;
    for (int i = 0; i < bits; ++i) {
        qpoly q = fg_update(p, prev, i, (int64_t)1 << (bits - i));
        tb_q(t, q);
        prev = rr[i];
    }
    return add(mul(p->V[0].a, prev), p->V[0].b);
}
static void fg_alloc(fgp *p, int lg)
{
    int64_t maxn = (int64_t)64 << lg;
    p->V = (lpoly *)malloc(sizeof(lpoly) * maxn);
    p->aV = (lpoly *)malloc(sizeof(lpoly) * maxn);
    p->am = (lpoly *)malloc(sizeof(lpoly) * maxn);
    int hl = ((lg + 6) >> 1) + 2;
    p->b0f = (fe *)malloc(sizeof(fe) << hl); p->b0s = (fe *)malloc(sizeof(fe) << hl);
    p->b1f = (fe *)malloc(sizeof(fe) << hl); p->b1s = (fe *)malloc(sizeof(fe) << hl);
    p->buf = (fe *)malloc(sizeof(fe) << hl); p->bus = (fe *)malloc(sizeof(fe) << hl);
}
static void fg_dealloc(fgp *p)
{
    free(p->V); free(p->aV); free(p->am);
    free(p->b0f); free(p->b0s); free(p->b1f); free(p->b1s); free(p->buf); free(p->bus);
}
static void fg_rot(fe inv_rou, fe *rot)
{
    rot[0] = inv_rou;
    for (int i = 1; i < 62; ++i) rot[i] = mul(rot[i - 1], rot[i - 1]);
}
/* proof_size as the program counts it: sizeof(quadratic_poly) per round,
 * plus the extension part it does not prove (extension_gkr :745-754) */
static int fg_proof_size(int lg)
{
    int q = (lg + 6) + lg + 2 * lg * lg;
    for (int i = 1; i <= lg; ++i) q += i;
    return 48 * q;
}

/* engage_gkr (:756-806) and main (:827-845): the prover side */
static void fg_prove(int lg, tbuf *t, double *p_sec)
{
    grand_t g;
    grand_init(&g, 1);
    fe r[64];
    for (int i = 0; i < lg; ++i) r[i] = grand_fe(&g);
    fgc C;
    fg_build(&C, lg, r, &g);
    double t0 = now_s();
    int64_t N = (int64_t)1 << lg;
    fgp p;
    fg_alloc(&p, lg);
    fe r0[80], omr0[80], r1[80], omr1[80], ru[80], omru[80], rv[80], omrv[80];
    p.alpha = ONE(); p.beta = ZERO;
    fg_refresh(&g, r0, omr0, lg + 10);
    fg_refresh(&g, r1, omr1, lg + 10);
    tb_put(t, "VGF1", 4);
    tb_u32(t, (uint32_t)lg);
    tb_fe(t, V_res(omr0, r0, C.val[C.nl - 1], 6, 64));          /* V_output :134-150 */
    {                                                            /* addition_layer :248-369 */
        half_tables(p.b0f, p.b0s, p.alpha, r0, omr0, 6);
        half_tables(p.b1f, p.b1s, p.beta, r1, omr1, 6);
        fg_tables_init(&p, C.val[C.nl - 2], 64 * N);
        for (int i = 0; i < 64; ++i) {
            fe tmp = fg_bg(&p, i, 3);
            for (int64_t j = (int64_t)i * N; j < (int64_t)(i + 1) * N; ++j) p.am[j].b = add(p.am[j].b, tmp);
        }
        fg_refresh(&g, ru, omru, lg + 6);
        fg_refresh(&g, rv, omrv, lg + 6);
        tb_fe(t, fg_phase(&p, t, ru, lg + 6));
        for (int i = 0; i < lg + 6; ++i) { r0[i] = ru[i]; r1[i] = rv[i]; omr0[i] = omru[i]; omr1[i] = omrv[i]; }
    }
    {                                                            /* mult_layer :371-495 */
        int length_g = lg + 6, fh = length_g >> 1;
        half_tables(p.b0f, p.b0s, p.alpha, r0, omr0, length_g);
        half_tables(p.b1f, p.b1s, p.beta, r1, omr1, length_g);
        fg_tables_init(&p, C.val[C.nl - 3], N);
        fe x_eval[64];
        for (int j = 0; j < 64; ++j) x_eval[j] = ONE();
        for (int64_t i = 0; i < N; ++i)
            for (int j = 0; j < 64; ++j) {
                fe tmp = fg_bg(&p, (int64_t)j * N + i, fh);
                p.am[i].b = add(p.am[i].b, mul(tmp, x_eval[j]));
                x_eval[j] = mul(x_eval[j], C.eval_points[j]);
            }
        fg_refresh(&g, ru, omru, lg);
        fg_refresh(&g, rv, omrv, lg);
        tb_fe(t, fg_phase(&p, t, ru, lg));
        for (int i = 0; i < lg; ++i) { r0[i] = ru[i]; r1[i] = rv[i]; omr0[i] = omru[i]; omr1[i] = omrv[i]; }
    }
    fe rot[62];                                                  /* ifft_gkr :504-743 */
    fg_rot(C.inv_rou, rot);
    int fh = lg >> 1;
    int64_t mf = ((int64_t)1 << fh) - 1;
    for (int dep = 0; dep < lg; ++dep) {
        int half = 1 << (lg - dep - 1);
        const fe *pre = C.val[2 * lg - dep - 1];
        half_tables(p.b0f, p.b0s, p.alpha, r0, omr0, lg);
        half_tables(p.b1f, p.b1s, p.beta, r1, omr1, lg);
        fg_tables_init(&p, pre, N);
        fe x = ONE();
        for (int k = 0; k < half; ++k) {
            for (int j = 0; j < (1 << dep); ++j) {
                int64_t gid = (int64_t)k << dep | j, uid = (int64_t)k << (dep + 1) | j, vid = uid | (1 << dep);
                fe tmp = fg_bg(&p, gid, fh);
                p.am[uid].b = add(p.am[uid].b, tmp);
                p.aV[uid].b = add(p.aV[uid].b, mul(mul(tmp, x), pre[vid]));
                gid = (int64_t)(k + half) << dep | j;
                tmp = fg_bg(&p, gid, fh);
                p.am[uid].b = add(p.am[uid].b, tmp);
                p.aV[uid].b = sub(p.aV[uid].b, mul(mul(tmp, x), pre[vid]));
            }
            x = mul(x, rot[dep]);
        }
        fg_refresh(&g, ru, omru, lg);
        fg_refresh(&g, rv, omrv, lg);
        fe v_u = fg_phase(&p, t, ru, lg);
        tb_fe(t, v_u);
        half_tables(p.buf, p.bus, ONE(), ru, omru, lg);          /* phase 2 init :615-656 */
        fg_tables_init(&p, pre, N);
        x = ONE();
        for (int k = 0; k < half; ++k) {
            for (int j = 0; j < (1 << dep); ++j) {
                int64_t gid = (int64_t)k << dep | j, uid = (int64_t)k << (dep + 1) | j, vid = uid | (1 << dep);
                fe tu = mul(p.buf[uid & mf], p.bus[uid >> fh]);
                fe tgu = mul(fg_bg(&p, gid, fh), tu);
                p.am[vid].b = add(p.am[vid].b, mul(tgu, x));
                p.aV[vid].b = add(p.aV[vid].b, mul(tgu, v_u));
                gid = (int64_t)(k + half) << dep | j;
                tgu = mul(fg_bg(&p, gid, fh), tu);
                p.am[vid].b = sub(p.am[vid].b, mul(tgu, x));
                p.aV[vid].b = add(p.aV[vid].b, mul(tgu, v_u));
            }
            x = mul(x, rot[dep]);
        }
        tb_fe(t, fg_phase(&p, t, rv, lg));
        for (int i = 0; i < lg; ++i) { r0[i] = ru[i]; r1[i] = rv[i]; omr0[i] = omru[i]; omr1[i] = omrv[i]; }
        p.alpha = grand_fe(&g);
        p.beta = grand_fe(&g);
    }
    if (p_sec) *p_sec = now_s() - t0;
    fg_dealloc(&p);
    fg_free(&C);
}

/* the program's inline verifier checks, replayed on a transcript section:
 * round sums (:342-352 and like), the closed forms that end each layer
 * (:327-352, :456-484, :677-731) and the claim chaining between layers */
static int fg_rounds(trd *rd, const fe *rr, int bits, fe *claim)
{
    int ok = 1;
    for (int i = 0; i < bits; ++i) {
        qpoly q = tr_q(rd);
        if (!feq(add(q_eval(q, ZERO), q_eval(q, ONE())), *claim)) ok = 0;
        *claim = q_eval(q, rr[i]);
    }
    return ok;
}
int vgo_fft_gkr_verify(int lg, const uint8_t *tr, size_t len)
{
    if (lg < 1 || lg > 24) return 0;
    trd rd = {tr, len, 0, 0};
    char magic[4];
    tr_get(&rd, magic, 4);
    uint32_t l32;
    tr_get(&rd, &l32, 4);
    if (memcmp(magic, "VGF1", 4) || (int)l32 != lg) return 0;
    grand_t g;
    grand_init(&g, 1);
    fe r[64];
    for (int i = 0; i < lg; ++i) r[i] = grand_fe(&g);
    fgc C;
    fg_build(&C, lg, r, &g);
    fgp p;
    int hl = ((lg + 6) >> 1) + 2;
    p.b0f = (fe *)malloc(sizeof(fe) << hl); p.b0s = (fe *)malloc(sizeof(fe) << hl);
    p.b1f = (fe *)malloc(sizeof(fe) << hl); p.b1s = (fe *)malloc(sizeof(fe) << hl);
    fe r0[80], omr0[80], r1[80], omr1[80], ru[80], omru[80], rv[80], omrv[80];
    fe one = ONE();
    int ok = 1;
    p.alpha = one; p.beta = ZERO;
    fg_refresh(&g, r0, omr0, lg + 10);
    fg_refresh(&g, r1, omr1, lg + 10);
    fe claim = tr_fe(&rd);
    if (!feq(claim, V_res(omr0, r0, C.val[C.nl - 1], 6, 64))) ok = 0;
    {                                                            /* addition layer */
        half_tables(p.b0f, p.b0s, p.alpha, r0, omr0, 6);
        half_tables(p.b1f, p.b1s, p.beta, r1, omr1, 6);
        int log_uv = lg + 6;
        fg_refresh(&g, ru, omru, log_uv);
        fg_refresh(&g, rv, omrv, log_uv);
        ok &= fg_rounds(&rd, ru, log_uv, &claim);
        fe v_u = tr_fe(&rd);
        fe sv = ZERO;
        for (int i = 0; i < 64; ++i) {
            fe tg = fg_bg(&p, i, 3), tu = one;
            for (int j = 0; j < 6; ++j)
                tu = mul(tu, ((i >> j) & 1) ? ru[log_uv - 6 + j] : sub(one, ru[log_uv - 6 + j]));
            sv = add(sv, mul(tg, tu));
        }
        for (int i = 0; i < log_uv; ++i) { r0[i] = ru[i]; r1[i] = rv[i]; omr0[i] = omru[i]; omr1[i] = omrv[i]; }
        if (!feq(claim, mul(sv, v_u))) ok = 0;
        claim = mul(p.alpha, v_u);
    }
    {                                                            /* mult layer */
        int length_g = lg + 6;
        fg_refresh(&g, ru, omru, lg);
        fg_refresh(&g, rv, omrv, lg);
        ok &= fg_rounds(&rd, ru, lg, &claim);
        fe v_u = tr_fe(&rd);
        fe sm = ZERO;
        for (int i = 0; i < 64; ++i) {
            fe tg0 = p.alpha, tg1 = p.beta;
            for (int j = 0; j < 6; ++j) {
                int b = (i >> j) & 1;
                tg0 = mul(tg0, b ? r0[length_g - 6 + j] : sub(one, r0[length_g - 6 + j]));
                tg1 = mul(tg1, b ? r1[length_g - 6 + j] : sub(one, r1[length_g - 6 + j]));
            }
            fe tu0 = one, tu1 = one, x = C.eval_points[i];
            for (int j = 0; j < lg; ++j) {
                tu0 = mul(tu0, add(mul(mul(r0[j], ru[j]), x), mul(sub(one, r0[j]), sub(one, ru[j]))));
                tu1 = mul(tu1, add(mul(mul(r1[j], ru[j]), x), mul(sub(one, r1[j]), sub(one, ru[j]))));
                x = mul(x, x);
            }
            sm = add(add(sm, mul(tg0, tu0)), mul(tg1, tu1));
        }
        for (int i = 0; i < lg; ++i) { r0[i] = ru[i]; r1[i] = rv[i]; omr0[i] = omru[i]; omr1[i] = omrv[i]; }
        if (!feq(claim, mul(sm, v_u))) ok = 0;
        claim = mul(p.alpha, v_u);
    }
    claim = mul(claim, fe_u((uint64_t)1 << lg));                 /* intermediate_layer */
    fe rot[62];
    fg_rot(C.inv_rou, rot);
    for (int dep = 0; dep < lg; ++dep) {                          /* ifft layers */
        fg_refresh(&g, ru, omru, lg);
        fg_refresh(&g, rv, omrv, lg);
        ok &= fg_rounds(&rd, ru, lg, &claim);
        fe v_u = tr_fe(&rd);
        ok &= fg_rounds(&rd, rv, lg, &claim);
        fe v_v = tr_fe(&rd);
        fe xr = rot[dep];
        int log_k = lg - dep - 1, log_j = dep;
        fe uv = mul(sub(one, ru[log_j]), rv[log_j]);
        fe u0A = mul(mul(sub(one, r0[lg - 1]), uv), p.alpha), u1A = mul(mul(sub(one, r1[lg - 1]), uv), p.beta);
        fe v0A = u0A, v1A = u1A;
        fe u0B = mul(mul(r0[lg - 1], uv), p.alpha), u1B = mul(mul(r1[lg - 1], uv), p.beta);
        fe v0B = u0B, v1B = u1B;
        for (int i = 0; i < log_k; ++i) {
            fe a0 = r0[log_j + i], a1 = r1[log_j + i], bu = ru[log_j + 1 + i], bv = rv[log_j + 1 + i];
            fe nn = mul(sub(one, bu), sub(one, bv)), pp = mul(bu, bv);
            fe t0 = add(mul(a0, pp), mul(sub(one, a0), nn));
            fe t1 = add(mul(a1, pp), mul(sub(one, a1), nn));
            fe t0x = add(mul(mul(a0, pp), xr), mul(sub(one, a0), nn));
            fe t1x = add(mul(mul(a1, pp), xr), mul(sub(one, a1), nn));
            u0A = mul(u0A, t0); u1A = mul(u1A, t1); v0A = mul(v0A, t0x); v1A = mul(v1A, t1x);
            u0B = mul(u0B, t0); u1B = mul(u1B, t1); v0B = mul(v0B, t0x); v1B = mul(v1B, t1x);
            xr = mul(xr, xr);
        }
        for (int i = 0; i < log_j; ++i) {
            fe nn = mul(sub(one, ru[i]), sub(one, rv[i])), pp = mul(ru[i], rv[i]);
            fe t0 = add(mul(r0[i], pp), mul(sub(one, r0[i]), nn));
            fe t1 = add(mul(r1[i], pp), mul(sub(one, r1[i]), nn));
            u0A = mul(u0A, t0); u1A = mul(u1A, t1); v0A = mul(v0A, t0); v1A = mul(v1A, t1);
            u0B = mul(u0B, t0); u1B = mul(u1B, t1); v0B = mul(v0B, t0); v1B = mul(v1B, t1);
        }
        fe want = add(mul(add(add(add(u0A, u1A), u0B), u1B), v_u), mul(sub(sub(add(v0A, v1A), v0B), v1B), v_v));
        if (!feq(claim, want)) ok = 0;
        for (int i = 0; i < lg; ++i) { r0[i] = ru[i]; r1[i] = rv[i]; omr0[i] = omru[i]; omr1[i] = omrv[i]; }
        p.alpha = grand_fe(&g);
        p.beta = grand_fe(&g);
        claim = add(mul(p.alpha, v_u), mul(p.beta, v_v));
    }
    if (rd.bad || rd.off != rd.n) ok = 0;
    free(p.b0f); free(p.b0s); free(p.b1f); free(p.b1s);
    fg_free(&C);
    return ok;
}

int vgo_fft_gkr(int lg, uint8_t **out, size_t *out_len, int *accepted, int *proof_size, double *p_sec)
{
    if (lg < 1 || lg > 24) FAIL(-1, "fft_gkr: lg %d out of range", lg);
    tbuf t = {0};
    fg_prove(lg, &t, p_sec);
    if (accepted) *accepted = vgo_fft_gkr_verify(lg, t.p, t.n);
    if (proof_size) *proof_size = fg_proof_size(lg);
    *out = t.p;
    *out_len = t.n;
    return 0;
}

/* one layer of the FFT circuit (for closed-form tests): index as in
 * C.circuit_val (0 = the constant 1 ... 2lg+3 = the 64 outputs) */
int vgo_fft_gkr_layer(int lg, int idx, fe *out, size_t cap)
{
    if (lg < 1 || lg > 20 || idx < 0 || idx > 2 * lg + 3) FAIL(-1, "bad layer");
    grand_t g;
    grand_init(&g, 1);
    fe r[64];
    for (int i = 0; i < lg; ++i) r[i] = grand_fe(&g);
    fgc C;
    fg_build(&C, lg, r, &g);
    int64_t n = C.size[idx];
    if ((size_t)n > cap) { fg_free(&C); FAIL(-2, "capacity"); }
    memcpy(out, C.val[idx], sizeof(fe) * n);
    fg_free(&C);
    return (int)n;
}
