// Host-side launchers of the CDNA4 kernels of the Virgo prover.  Every
// launcher enqueues on the given stream and returns immediately; all pointers
// are device pointers.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>
#include "vg_field.hpp"

namespace vg {

// Codeword storage used by the commitment: each slice of S natural positions
// is kept as R rows ("cosets" s) by S/R columns t, natural position
// p = t*R + s stored at s*(S/R) + t.  Every FRI pair (p, p+S/2) and every
// Merkle leaf pair then lies in one row, so all passes read contiguously.
struct cw_layout {
    int R;          // rows (cosets), R <= 32
    int log_R;
};
__host__ __device__ inline int64_t cw_idx(int64_t p, int64_t S, int log_R)
{
    int64_t R = (int64_t)1 << log_R;
    return (p & (R - 1)) * (S >> log_R) + (p >> log_R);
}

struct gates_dev {            // one layer's gates, SoA
    const uint8_t *ty;
    const int32_t *u, *v;
    int bl;
};

namespace launch {

void zero(void *p, size_t bytes, hipStream_t s);
void eval_input(const int32_t *u, fe *out, int n, hipStream_t s);
void eval_layer(gates_dev g, const fe *prev, fe *out, hipStream_t s);

// out[j] = in[2j](1-r[k]) + in[2j+1] r[k] for all rounds k < nr, result -> dst
void mle_eval(const fe *in, int log_n, const fe *r, fe *tmp0, fe *tmp1, fe *dst, hipStream_t s);

// out[x] = lo[x mod 2^fh] * hi[x >> fh]
void eq_expand(fe *out, const fe *lo, const fe *hi, int fh, int64_t n, hipStream_t st);

struct beta_g { const fe *g0f, *g0s, *g1f, *g1s; int fh; };
struct beta_u { const fe *f, *s; int fh; };


// a gate as listed in the CSR transposes: everything a table build reads of
// it, in one 16-byte load
struct __align__(16) gate_rec {
    int32_t g, u, v, ty;
};
// CSR transposes of gate layers (ng gates over nt = 2^bl(i-1) table
// entries each): off1 / off2 (nt + 1 offsets each) and ent1 / ent2 by u and
// by v (record counts as the host's plan counted them); tops1 / tops2:
// csr_tops_len(nt) words of scan scratch each.  csr_build_multi builds the
// n layers of the device array L in six launches on s (max_ng, max_nt: the
// largest layer's).
struct csr_layer {
    const uint8_t *ty;
    const int32_t *u, *v;
    int64_t ng, nt;
    uint32_t *off1, *off2;
    gate_rec *ent1, *ent2;
    uint32_t *tops1, *tops2;
};
int64_t csr_tops_len(int64_t nt);
void csr_build_multi(const csr_layer *L, int n, int64_t max_ng, int64_t max_nt, hipStream_t s);
// One GKR layer of the main prover on the device (all pointers fixed per
// context).  Batches of layers with equal table size run in one launch.
struct layer_dev {
    const uint8_t *ty;
    const int32_t *u, *v;                 // the layer's gates
    int nt, ng;                           // table entries 2^bl(i-1), gates 2^bl(i)
    const uint32_t *off1, *off2;          // CSR by u (phase 1), by v (phase 2)
    const gate_rec *ent1, *ent2;
    const int32_t *z2;                    // phase-2 gates adding into addV[0]
    int nz;
    const fe *pv;                         // values of layer i-1
    beta_g bg;
    beta_u bu, bv;
    fe *amb, *aVb;                        // flat tables
    const fe *vu, *vv;                    // end[4] of the two phases
    fe *part;                             // MAX_PARTS block partials
    fe *gv;                               // general value (null for layer 1)
};
constexpr int MAX_PARTS = 1024;        // block partials; part[MAX_PARTS] holds the zero-list sum
// phase 0: tables by u (zk_prover.cpp:375-592); phase 1: by v (:789-983),
// including the gates that add into addV[0] (nz_max: largest z2 list), into
// the flat tables amb, aVb (used when the phase has no streaming launch),
// 2^kl lanes per table entry (0..6: more for tables with many records per
// entry)
void layer_tables(int phase, const layer_dev *D, int n, int nt, int nz_max, hipStream_t s, int kl = 0);
// the same fused with the phase's first streaming fold launch, of at most
// GF_B rounds (the flat tables are never written): returns its rounds, 0 when
// the phase has no streaming launch (then layer_tables and the plain fold
// apply)
struct fold_inst;
int gather_fold(int phase, const layer_dev *D, const fold_inst *I, int n, int L, int nz_max, hipStream_t s);
// Sumcheck phases (zk_prover.cpp:600-633, 992-1025) of a batch of ninst
// instances with flat tables of 2^L entries each, blockIdx.y = instance:
// streaming launches of up to FB_B rounds (2^FB_B consecutive values per
// lane, no communication), then the last <= FB_TAIL rounds in one wave per
// instance.  Round j's pair polynomials are evaluated at r[j].  Each
// streaming launch writes its block partials of round q to slots[(q*3 + c) *
// FW_BLOCKS + block] (q counted over the instance's launches; rounds j >= 1
// with b = 0); the caller reduces every slot of the proof once with
// reduce_slots, which also sets those rounds' b.  Tail round sums go
// straight to sums[3j..].  After the last round: end[0..3] = the final pair's
// V.a, V.b, am.a, am.b and end[4] = V.a * r[L-1] + V.b + *vconst (= v_u or
// v_v).  Returns the slot rounds per instance.
constexpr int FB_B = 3;                // rounds per streaming launch
constexpr int FB_TAIL = 6;             // rounds of the tail (<= 64 values)
constexpr int GF_B = 2;                // rounds of a fused table build + fold (gather_fold)
constexpr int FW_BLOCKS = 1024;
// launch plan of a phase on 2^L entries whose first launch folds at most
// `first` rounds: launch k starts at round js[k] and folds ks[k] rounds
int fold_plan(int L, int *js, int *ks, int first = FB_B);
int fold_blocks(int L, int j, int first = FB_B);
// values of each table a phase hands from one launch to the next (the
// output of its first streaming launch)
// rounds and table entries per block of gather_fold on 2^L entries (0: none)
inline int gf_rounds(int L) { return L > FB_TAIL ? (L - FB_TAIL < GF_B ? L - FB_TAIL : GF_B) : 0; }
inline int64_t gf_tile(int L) { return (int64_t)256 << gf_rounds(L); }
// gather_fold pays off when no block's entries hold more than this many
// records per entry on average (a block walks its records sequentially)
constexpr int GF_MAX_FANIN = 4;
inline int64_t fold_ws_len(int L, int first = FB_B)
{
    int b = L - FB_TAIL < first ? L - FB_TAIL : first;
    return L > FB_TAIL ? ((int64_t)1 << (L - b)) : 64;
}
struct fold_inst {
    const fe *Vb, *aVb, *amb;       // flat tables
    fe *V[2], *aV[2], *am[2];       // chunk results, fold_ws_len(L) values each
    const fe *r;                    // the phase's challenges
    fe *slots;                      // the instance's slot rounds
    fe *sums, *end;                 // tail round sums, end values
    const fe *vconst;               // added to end[4]; null = 0
};
// plan with first-launch cap `first`; skip_first: the first launch already
// ran (gather_fold)
int sumcheck_multi(const fold_inst *I, int ninst, int L, hipStream_t s, int first = FB_B, int skip_first = 0);
// msg[dst[R]..+3] = sum over the nb[R] block partials of slot round R; then
// for every slot round j >= 1 of a phase (rof[R] >= 0, the offset of r[j-1]
// in chal) b = p_{j-1}(r[j-1]) - a - 2c from the previous round's message
// (the folds of rounds j >= 1 sum a and p(0) only).  A phase's slot rounds
// are consecutive slots from its round 0 (rof = -1).
void reduce_slots(const fe *slots, int nslots, const int *nb, const int64_t *dst, const int64_t *rof, const fe *chal,
                  fe *msg, hipStream_t s);
struct half_desc {            // one eq half table: out[j] = f0 * prod_{i<nbits} eq(bit_i(j), r[i])
    fe *out;
    const fe *f0;
    const fe *r;
    int nbits;
    int64_t start;            // prefix sum of table sizes
};
void half_multi(const half_desc *d, int nd, int64_t total, hipStream_t s);

// per layer: sum over gates of beta_g(g) beta_u(u) beta_v(v) f_type(v_u, v_v)
// (zk_verifier.cpp:219-606 and the combination at :1149) -> *gv
void general_multi(const layer_dev *D, int n, int ng_max, hipStream_t s);

// tw[e] = w^e, e < n, from T1[j] = w^j (j < 1024) and T2[j] = w^(1024 j)
void twiddles(fe *tw, int64_t n, const fe *T1, const fe *T2, hipStream_t s);
// one FRI fold of a natural-order codeword of 2n values (fri.cpp:321-340):
// out[i] = ((in[i] + in[i+n]) + r iw[i] (in[i] - in[i+n])) / 2, iw[i] = w^-i
void fri_fold_nat(const fe *in, int64_t n, const fe *r, const fe *iw, fe *out, hipStream_t s);
// kernel micro-set pieces.  A codeword of 2 half values is natural
// (log_rows = 0) or coset-major (value k in row k mod 2^log_rows, column
// k >> log_rows, rows of cols values).  pair_leaves: the leaf level of its
// Merkle tree (leaf j = value j || value j + half, at tree + 32 (half + j)).
// fri_fold_any: the fold of fri_fold_nat with w^-i = tw[(SS - (i << sh))
// mod SS] from the powers tw[e] = w_SS^e (2 half << sh == SS); with leaves
// set it also writes the leaf level of the result's tree (half/2 leaves,
// half >= 2).  copy_roots: out[32 k ..] = root of tree k (k < n <= 32).
// count_diff: *cnt += number of positions where a and b differ.
void pair_leaves(const fe *cw, int64_t half, int log_rows, int64_t cols, uint8_t *tree, hipStream_t s);
void fri_fold_any(const fe *in, int64_t half, int log_rows, int64_t cols, const fe *r, const fe *tw, int64_t SS, int sh,
                  fe *out, uint8_t *leaves, hipStream_t s);
int copy_roots(const uint8_t *const *trees, int n, uint8_t *out, hipStream_t s);
void count_diff(const fe *a, const fe *b, int64_t n, unsigned int *cnt, hipStream_t s);
// Twiddles of the NTTs: stage-major W[2^s + k] = w_{2^(s+1)}^k and Wi (the
// inverse powers) for s < maxlog, so each stage reads a contiguous run; the
// LDE's coset twist w^(s j) = TA[s][j mod 2^th] TB[s][j >> th] for rows
// s < 2^log_R (TA, TB null when there is no LDE).  All read from the table
// tw[e] = w^e (e < SS, w of order SS).
struct ntt_tabs {
    const fe *W, *Wi;
    const fe *TA, *TB;
    int th, tb;
};
void ntt_tables(fe *W, fe *Wi, int maxlog, fe *TA, fe *TB, int log_R, int th, int tb, const fe *tw, int64_t SS,
                hipStream_t s);

// Batched inverse NTT: out[b][k] = N^-1 sum_j in(b, j) w_N^{-jk}, natural order.
//   kind 0: in(b, j) = src[b*src_stride + j]
//   kind 1: in(b, j) = l[b][idx(j*step)] * q[b][idx(j*step)]  (codeword layout)
struct intt_src {
    int kind;
    const fe *src; int64_t src_stride;
    const fe *l, *q; int64_t cw_S, cw_stride; int64_t step; int log_R;
    int brev_src = 0;      // kind 0: src holds the row in bit-reversed order (brev_permute)
    const fe *qscale = nullptr;   // kind 1: q of row b is qscale[b] times the one row at q
};
void intt(intt_src in, fe *out, int64_t out_stride, int nb, int log_n, fe inv_n, const ntt_tabs &T, hipStream_t s);
// Coset LDE: for each slice b and row s < R, out[b][s][t] = sum_{j<len_b} c_b[j] w_SS^{(tR+s) j}.
// coef[b] has lens[b] (<= C) entries at coef + b*coef_stride.
// hmode: coefficients of slice b start at coef + b*coef_stride + lens[b] (upper half of lq)
void lde(const fe *coef, int64_t coef_stride, const int *lens, int nb, fe *out, int log_C, int log_R,
         const ntt_tabs &T, hipStream_t s, int hmode = 0, int brev_src = 0);
// dst[brev(j)] = j < len ? src[j] : 0 for j < 2^log_n (log_n >= 12): the
// bit-reversed copy a transform reads with brev_src set, so that its first
// pass reads contiguous tiles instead of one value per 2^(log_n-12) stride
int brev_permute(const fe *src, int64_t len, fe *dst, int log_n, hipStream_t s);
// all_sum[i] = (lq_coef[0] + lq_coef[len_i]) * len_i; c0[i] = lq_coef[0] + lq_coef[len_i]
void slice_sums(const fe *lq_coef, int64_t stride, int nslice, int64_t real, int64_t m, fe *c0, fe *all_sum, hipStream_t s);

// generic batched NTT (inverse: w^-1), natural order in and out
void ntt(intt_src in, fe *out, int64_t out_stride, int nb, int log_n, int inverse, fe scale, int do_scale,
         const ntt_tabs &T, hipStream_t s);
// NTT tile autotuner: times the LDS tile candidates (2^10..2^12 elements)
// and, for 2^13 and 2^14, the one-block transform (code 1) of a forward
// transform of nb rows of 2^log_n on stream s (in: any nb * 2^log_n
// values, out: as many, overwritten) and keeps the fastest for every later
// transform of that size in the process; once per size (later calls return
// at once).  ntt_plan: the schedule in use for log_n (the default 2^11
// until tuned; VIRGO_NTT_LM=k forces one) and the measured us of the four
// candidates in us[0..3] (tiles 2^10..2^12, one-block; 0: not run).
int ntt_tune(int log_n, int64_t nb, const fe *in, fe *out, const ntt_tabs &T, hipStream_t s);
int ntt_plan(int log_n, float *us);
void sha3_batch(const uint8_t *in, uint8_t *out, int64_t n, hipStream_t s);
// node hashes of a Merkle tree: the proofs use SHA3-256; SHA-256 is the
// alternative the kernel micro-set may build its trees with
constexpr int HASH_SHA3 = 0, HASH_SHA256 = 1;
void sha256_batch(const uint8_t *in, uint8_t *out, int64_t n, hipStream_t s);
// fills leaves [n_leaf, pad) of a tree with pad leaves with H(0^64)
void merkle_pad(uint8_t *tree, int64_t pad, int64_t n_leaf, hipStream_t s, int hash = HASH_SHA3);
void fe_mul(const fe *a, const fe *b, fe *o, int64_t n, hipStream_t s);
// FRI fold only (no hashing): next (S/2 per slice) from prev (S per slice)
void fri_fold(const fe *prev, int64_t S, int log_R, const fe *r, const fe *tw, int64_t SS, fe *next, hipStream_t s);
// FRI folds: nlev levels from a codeword of S0 positions per slice (twiddles
// w_SS), level k with challenge r[k] into next[k]; the small levels in one launch
void fri_folds(const fe *cw, int64_t S0, int64_t SS, int nlev, int log_R, const fe *r, const fe *tw, fe *const *next,
               hipStream_t s);
// the same from the virtual oracle of l, q, h (poly_commit.h:246-263,
// 305-322), computed inside the first fold and never stored; q holds two
// extensions (codeword layout, SS each): Q's, which slice j < 64 scales by
// cq[j], and the last slice's own
void fri_folds_vo(const fe *l, const fe *q, const fe *cq, const fe *h, const fe *c0, int64_t real, int64_t m, int64_t SS,
                  int nlev, int log_R, const fe *r, const fe *tw, fe *const *next, hipStream_t s);
// leaf chains of nseg codewords in one launch (codeword k: S[k] positions per
// slice, leaves S[k]/2 written to tree[k] + S[k]/2 nodes)
void leaf_hash_multi(int nseg, const fe *const *cw, const int64_t *S, uint8_t *const *tree, int log_R, hipStream_t s);
// leaf chains of many codewords (of many proofs) in one launch, from a
// segment table in device memory; start[k] = sum of S/2 over segments < k,
// start[nseg] = total
struct leaf_seg {
    const fe *cw;
    int64_t S;
    uint8_t *tree;
    int log_R;
};
void leaf_hash_table(const leaf_seg *G, const int64_t *start, int nseg, int64_t total, hipStream_t s);
// Merkle trees of many proofs: levels with more than MERKLE_TAIL nodes level
// by level from item tables (item: a tree's nodes [first, 2 first), start =
// prefix sum of the node counts of the level), the rest one block per tree
// G[k] (nleaf = S / 2)
constexpr int MERKLE_TAIL = 256;
// phase 3 of many proofs in two launches (batched proving): every proof's
// slot reduction (as reduce_slots), then every proof's query answers and
// message block copied straight into its page-locked host staging
struct slot_job {
    const fe *part;
    const int *nb;
    const int64_t *dst;
    fe *msg;
    const int64_t *rof;     // as reduce_slots
    const fe *chal;
};
struct copy_job {
    const uint64_t *addr;     // gather: out[k] = *(fe *)addr[k]; null: out[k] = base[k]
    const fe *base;
    fe *out;                  // may be page-locked host memory
};
// start[j]: first slot (reduce) / element (copy) of job j; start[n] = total
void reduce_slots_table(const slot_job *J, const int64_t *start, int n, int64_t total, hipStream_t s);
void copy_table(const copy_job *J, const int64_t *start, int n, int64_t total, hipStream_t s);
struct mk_item {
    uint8_t *tree;
    int64_t first, start;
};
void merkle_level_table(const mk_item *items, int n, int64_t total, hipStream_t s);
void merkle_tails_table(const leaf_seg *G, int ntree, hipStream_t s);
// every Merkle tree of the set, level by level across trees
void merkle_multi(int ntree, uint8_t *const *tree, const int64_t *nleaf, hipStream_t s, int hash = HASH_SHA3);
// out[k] = 16 bytes at address src[k]
void gather16(const uint64_t *src, int64_t n, fe *out, hipStream_t s);

// ---- verifier (zk_verifier.cpp:219-606, 884-908; vpd_verifier.cpp:8-41, 80-335) ----
// out[layer * 14 + type] = the layer's wiring predicate of that gate type
// (zero for the layer without a general value); part: n * blocks * 14
void predicates(const layer_dev *D, int n, int ng_max, fe *part, int blocks, fe *out, hipStream_t s);
// out[p] = sum_{k < n} coef[k] pts[p]^k
void poly_eval(const fe *coef, int64_t n, const fe *pts, int npts, fe *out, hipStream_t s);
struct path_job {              // one authentication path of a query answer
    int64_t pow;               // position whose bits order each level's pair
    int64_t root, path;        // digest indices: root, first path entry
    int64_t vals;              // first of the 65 value pairs
    int len;                   // path entries (the last is the leaf digest)
};
// ok[k] = 1 when job k's leaf chain and path reach its root
void verify_paths(const path_job *J, int njobs, const uint64_t *dg, const fe *vals, int *ok, hipStream_t s);

// ---- fft_gkr (fft_circuit_GKR.cpp:22-743), N = 2^lg ----
struct fg_pw_t {              // power tables of 65 points: 64 evaluation points, then inv_rou
    const fe *plo, *phi;      // x_i^t (t < 2^lo) and x_i^(s 2^lo) (s < 2^hi), row i
    int lo, hi;
    int tw_pt;                // row of inv_rou
};
// layer lg of the circuit: out[x] = prod_i (bit (lg-1-i) of x ? 1 - r_i : r_i)
void fg_eq(fe *out, const fe *r, int lg, hipStream_t s);
void fg_pow(const fe *pts, int npts, int lo, int hi, fe *plo, fe *phi, hipStream_t s);
// one butterfly layer of the circuit's inverse DFT, times scale if do_scale
void fg_butterfly(const fe *pre, fe *cur, int dep, int lg, fg_pw_t pw, fe scale, int do_scale, hipStream_t s);
// out[j + (i << lg)] = coef[j] x_i^j for the 64 evaluation points
void fg_eval(const fe *coef, fg_pw_t pw, int lg, fe *out, hipStream_t s);
void fg_rowsum(const fe *ev, int lg, fe *outs, hipStream_t s);
// sumcheck tables of the addition, mult and inverse-DFT layers
void fg_add_tables(beta_g bg, int lg, fe *amb, fe *aVb, hipStream_t s);
void fg_mult_tables(beta_g bg, fg_pw_t pw, int lg, fe *amb, fe *aVb, hipStream_t s);
struct fg_layer_dev {          // one inverse-DFT layer of the circuit
    beta_g bg;                // beta_g half tables
    beta_u bu;                // beta_u half tables (phase 2)
    const fe *pre;            // the layer's input values
    fe *amb, *aVb;            // its flat tables
    const fe *vu;             // v_u (end[4] of its phase 1)
    int dep;
};
// phase (0 or 1) tables of every inverse-DFT layer in one launch
void fg_ifft_tables(int phase, const fg_layer_dev *D, int nlayer, fg_pw_t pw, int lg, hipStream_t s);

}  // namespace launch
}  // namespace vg
