/*
 * vg_oracle.h -- CPU oracle for the Virgo prover path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded, plain-C restatement
 * of the reference Virgo prover/verifier (sunblaze-ucb/Virgo), used as the
 * bit-exact checker for the HIP product and as the CPU baseline timer.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (virgo_amd/) never links, loads or calls anything in oracle/.
 *
 * Every function cites the reference file:line it restates.
 *
 * Parity pinning (what checks each part of this restatement):
 *   field ops ............ Python big-int closed form + the reference's own
 *                          prime_field.cpp compiled from source (oracle/_ref)
 *   glibc rand ........... libc rand() through ctypes (same recurrence)
 *   FFT / iFFT ........... brute-force DFT on small sizes, round trips
 *   SHA3-256 ............. FIPS 202 known answers and hashlib.sha3_256
 *   Merkle / leaf chains . recomputation with hashlib from the closed form
 *   sumcheck / GKR ....... the restated verifier's round checks (p(0)+p(1)),
 *                          brute-force multilinear sums on tiny circuits,
 *                          mutation tests (a flipped message is rejected)
 *   FRI / commitment ..... restated verifier (query + fold consistency,
 *                          constant final codeword), closed-form fold checks
 * The exact transcript bytes of a full proof are NOT pinned against a run of
 * the reference: the reference needs a header (flo-shani.h) and a prebuilt
 * SHA3 library that this image lacks, so it is unbuildable here.
 */
#ifndef VG_ORACLE_H
#define VG_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* F_{p^2}, p = 2^61-1, i^2 = -1.  Memory order {img, real} as in
 * include/linear_gkr/prime_field.h:93 (this order is what gets hashed). */
typedef struct { uint64_t img, real; } vgo_fe;
typedef struct { uint8_t b[32]; } vgo_digest;

/* Circuit in the reference's file layout (zk_verifier.cpp:15-217): d file
 * layers; file layer 1 holds the inputs (type 3 value u, or type 2 zero);
 * gates of each layer are listed in gate order g = 0..n-1 and concatenated. */
typedef struct {
    int d;
    const int64_t *n;      /* d entries */
    const int32_t *ty;     /* sum(n) entries */
    const int64_t *u;
    const int64_t *v;
} vgo_circuit_desc;

/* ---- primitives (exported for known-answer tests) ---- */
void vgo_fe_mul(const vgo_fe *a, const vgo_fe *b, vgo_fe *out, size_t n);
void vgo_fe_add(const vgo_fe *a, const vgo_fe *b, vgo_fe *out, size_t n);
void vgo_fe_sub(const vgo_fe *a, const vgo_fe *b, vgo_fe *out, size_t n);
void vgo_fe_inv(const vgo_fe *a, vgo_fe *out, size_t n);
void vgo_root_of_unity(int log_order, vgo_fe *out);
void vgo_rand_seq(uint32_t seed, int32_t *out, size_t n);
void vgo_fft(const vgo_fe *coef, int coef_len, int order, const vgo_fe *rou, vgo_fe *out);
void vgo_ifft(const vgo_fe *evals, int coef_len, int order, const vgo_fe *rou, vgo_fe *out);
void vgo_sha3_64(const uint8_t *in64, uint8_t *out32, size_t n);
/* heap-layout Merkle tree over n_leaf 32-byte leaves; tree has 2*pad entries */
int  vgo_merkle_tree(const vgo_digest *leaves, int n_leaf, vgo_digest *tree_out);

/* Component-level restatements used as per-kernel oracles. */
/* Reed-Solomon LDE of one slice: iFFT(real) then FFT(real -> slice) */
void vgo_lde_slice(const vgo_fe *vals, int log_real, int log_slice, vgo_fe *out);
/* leaf hash chain over 64 interleaved slices + mask pair (fri.cpp:97-126) */
void vgo_leaf_hashes(const vgo_fe *slices /*64 x S*/, const vgo_fe *mask /*S*/,
                     int log_S, vgo_digest *leaves /*S/2*/);
/* one FRI fold of 65 slices of size S (fri.cpp:321-340, 367-379) */
void vgo_fri_fold(const vgo_fe *slices /*65 x S*/, int log_S, const vgo_fe *r,
                  vgo_fe *out /*65 x S/2*/);
/* eq-table q[x] = prod_i (x_i ? r_i : 1-r_i)  (zk_verifier.cpp:838-847) */
/* fri.cpp:321-340 for one natural-order codeword of 2^log_S values */
void vgo_fri_fold1(const vgo_fe *cw, int log_S, const vgo_fe *r, vgo_fe *out);
void vgo_eq_table(const vgo_fe *r, int len, vgo_fe *out);

/* ---- whole protocol ---- */
/* Proves the circuit with the reference's challenge stream (glibc rand, seed
 * 1).  Writes a malloc'd transcript.  Returns 0 on success, else an error
 * code (the message is in vgo_last_error()).  times[0..4] (seconds):
 * evaluate, gkr sumcheck, commitment (LDE + Merkle + FRI), queries, fft_gkr
 * (whose section ends the transcript). */
int  vgo_prove(const vgo_circuit_desc *c, uint8_t **transcript, size_t *len, double *times);
/* Restated verifier (zk_verifier.cpp:910-1253, vpd_verifier.cpp:80-335) run
 * on a transcript.  Returns 1 = accept, 0 = reject, <0 = malformed input. */
int  vgo_verify(const vgo_circuit_desc *c, const uint8_t *transcript, size_t len);
/* Evaluates the circuit and returns the output layer (for cross-checks). */
int  vgo_evaluate(const vgo_circuit_desc *c, vgo_fe *out, size_t out_cap, int *out_bits);

/* fft_gkr (fft_circuit_GKR.cpp:22-845): prover of the FFT/evaluation circuit
 * with lg = log2(coefficient count); fresh rand() stream (seed 1).  Writes
 * the "VGF1" transcript section, the program's verdict (its inline checks,
 * replayed by vgo_fft_gkr_verify), its proof_size and prover seconds. */
int  vgo_fft_gkr(int lg, uint8_t **out, size_t *out_len, int *accepted, int *proof_size, double *p_sec);
int  vgo_fft_gkr_verify(int lg, const uint8_t *tr, size_t len);
/* layer idx of the circuit (0 .. 2lg+3); returns its length */
int  vgo_fft_gkr_layer(int lg, int idx, vgo_fe *out, size_t cap);
void vgo_free(void *p);
const char *vgo_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
