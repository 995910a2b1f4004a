/*
 * virgo.h -- C ABI of the MI355X Virgo prover (libvirgo_hip.so).
 *
 * Plain pointers and sizes only.  Field elements are 16 bytes
 * {uint64 img, uint64 real}, the memory layout of the reference's
 * prime_field::field_element (include/linear_gkr/prime_field.h:93).
 * Every entry point returns 0 on success and a negative code on error;
 * vg_last_error() then describes the failure.  Nothing here aborts.
 *
 * The reference's prover is a C++ class driven in-process by its verifier
 * (src/linear_gkr/zk_verifier.cpp:910-1253).  Its challenge stream is glibc
 * rand() with the default seed 1, and the order of draws depends only on the
 * circuit shape, so the whole proof is run here as one device-resident
 * schedule; vg_prove returns the prover's messages as a transcript in the
 * order the reference verifier consumes them.
 */
#ifndef VIRGO_H
#define VIRGO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t img, real; } vg_fe;

/* A layered circuit in the reference's file layout (zk_verifier.cpp:15-217):
 * d file layers; layer 1 holds the inputs (type 3: value u, type 2: zero);
 * gates of each layer are in gate order, concatenated over layers. */
typedef struct {
    int d;
    const int64_t *n;
    const int32_t *ty;
    const int64_t *u;
    const int64_t *v;
} vg_circuit_desc;

typedef struct vg_ctx vg_ctx;

/* Selects the HIP device for this thread's later calls. */
int vg_init(int device);
const char *vg_last_error(void);
void vg_free(void *p);

/* Builds the layered circuit exactly as zk_verifier::read_circuit does
 * (input padding to 2^15 or 2^17, relay layer 1, bit lengths), uploads gates
 * and inputs, and allocates every device buffer of one proof.
 * stream: NULL -- the context owns its streams and overlaps the commitment,
 * the sumcheck layers and fft_gkr where that lowers one proof's latency
 * (inputs below 2^20; larger proofs run in order on the context's stream,
 * VIRGO_SERIAL=1 / =0 forces either); or a hipStream_t -- the whole proof
 * runs in order on that stream, so many contexts on different streams
 * overlap on the device (batched proving). */
int vg_ctx_create(vg_ctx **ctx, const vg_circuit_desc *c, void *stream);
int vg_ctx_destroy(vg_ctx *ctx);
/* Creates / destroys a non-blocking HIP stream for vg_ctx_create (no
 * counterpart in the reference: a convenience for callers without a HIP
 * runtime binding of their own). */
int vg_stream_create(void **stream);
int vg_stream_destroy(void *stream);
/* Bytes of device memory held by the context. */
int64_t vg_ctx_device_bytes(const vg_ctx *ctx);
/* Loads and checks circuit c as vg_ctx_create does (the checks and padding
 * of zk_verifier::read_circuit, zk_verifier.cpp:15-217; the same error codes
 * and messages) and plans one proof of it on the host alone -- no HIP call,
 * no device memory: *device_bytes and *pinned_bytes are what a context of it
 * on a caller's stream takes (a proof server's lease).  No counterpart in
 * the reference, which has no separate planning step; for admission control
 * and capacity planning by callers. */
int vg_plan(const vg_circuit_desc *c, int64_t *device_bytes, int64_t *pinned_bytes);
/* NTT schedule in use for transforms of 2^log_n values (no counterpart in
 * the reference, whose FFT has one fixed loop, RS_polynomial.cpp:27-240):
 * the LDS tile (log2 of its elements, 10..12), or 1 for the one-block
 * transform of sizes 2^13 and 2^14 (all stages in one launch), picked by
 * the library's autotuner -- the first context that needs a transform size
 * times the candidates on its own buffers and keeps the fastest for the
 * process -- and, in us[0..3] (may be NULL), the measured times of the
 * 2^10, 2^11 and 2^12 tiles and of the one-block transform (0: not
 * measured; the default 2^11 until tuned, VIRGO_NTT_LM=10|11|12|1 forces
 * one).  Every schedule gives the same results. */
int vg_ntt_plan(int log_n, int *tile_log, float *us);

/* Replaces zk_verifier::verify's prover calls (zk_verifier.cpp:922-1226) and
 * poly_commit_prover::commit_private_array / commit_public_array /
 * commit_phase (poly_commit.h:42-351, vpd_verifier.cpp:46-78) plus the
 * query answers of fri::request_init_value_with_merkle /
 * request_step_commit (fri.cpp:148-289).
 * The transcript ends with the fft_gkr section of the commitment (as
 * vg_fft_gkr with lg = n - 6; vpd_verifier.cpp:90-99 runs that program and
 * adds its prover time to the proof's).
 * Writes a malloc'd transcript (free with vg_free).  stage_ms (may be NULL)
 * receives 13 timings in ms.  Device: evaluate, gkr (incl. V_res), then the
 * commitment, which runs on a second stream beside gkr, or after it when the
 * proof runs in order: lde_l, lde_qh (the public array's and h's
 * extensions), fri_fold (virtual oracle + folds), hash (all leaf chains
 * and Merkle trees); then queries (gathers + copies to the host), the SHA3
 * leaf-chain kernel alone (part of hash), the whole device schedule, and
 * fft_gkr (on its own stream).  Host: prepare (challenge stream, mask
 * bookkeeping, query gather table), enqueue, assemble (transcript). */
int vg_prove(vg_ctx *ctx, uint8_t **transcript, size_t *len, double *stage_ms);
/* Same, but only enqueues the device work of the proof on the context's
 * stream and returns; vg_finish waits and assembles the transcript.  Lets
 * several contexts on different streams overlap on one device. */
int vg_prove_async(vg_ctx *ctx);
int vg_finish(vg_ctx *ctx, uint8_t **transcript, size_t *len, double *stage_ms);

/* Batched proving (no single counterpart in the reference, which proves one
 * circuit per process): proves ctxs[0..n) with `workers` host threads,
 * thread k taking i = k, k + workers, ...; transcripts[i] / lens[i] receive
 * each proof (free with vg_free), byte-identical to a single vg_prove of the
 * same context.  Contexts created on distinct caller streams overlap on the
 * device; give the contexts of one worker the same stream.  Returns the
 * first error (its message in vg_last_error); proofs not made are NULL. */
int vg_prove_batch(vg_ctx *const *ctxs, int n, int workers, uint8_t **transcripts, size_t *lens);

/* ---- proof server: continuous batching (no counterpart in the reference,
 * which proves one circuit per process, main_zk.cpp:11-25) ----
 * Requests of any size are submitted at any time and admitted while others
 * run.  The server reserves device_budget bytes of HBM and pinned_budget
 * bytes of page-locked host memory (<= 0: 2 GiB) once; every admitted proof
 * takes its whole working set from them as one lease (the pinned part holds
 * its set-up uploads, about 9 bytes per gate, and its transcript staging),
 * keeps it resident on the device until its transcript is on the host, and
 * gives it back, so no device memory is allocated or freed while proofs run.
 * `slots` worker threads (1..64), each on its own stream, take requests in
 * arrival order; requests waiting for memory are admitted in arrival order
 * as soon as their lease fits, except that a later request whose lease fits
 * may pass the oldest waiter up to VIRGO_SERVE_BYPASS (default 8) times.  A
 * proof's uploads and launches all go to its worker's stream, and its worker
 * waits only for its own proof's end.  Every transcript is byte-identical to vg_prove of the same
 * circuit alone.  A request larger than the budgets fails (its wait returns
 * the error); the others are unaffected. */
typedef struct vg_server vg_server;
typedef struct {
    int64_t submitted, completed, failed;   /* requests so far (failed: among completed) */
    int64_t queued, running;                /* waiting for a slot; in a slot (planning, waiting for memory or proving) */
    int64_t device_in_use, device_peak;     /* bytes of the device budget leased now / at most */
    int64_t peak_admitted;                  /* most proofs holding a lease at once */
    int64_t device_budget;
} vg_server_stats_t;
int vg_server_create(vg_server **server, int64_t device_budget, int64_t pinned_budget, int slots);
/* Queues a proof of circuit c; *id names it.  The arrays c points to must
 * stay valid until vg_server_wait returns for this id. */
int vg_server_submit(vg_server *server, const vg_circuit_desc *c, int64_t *id);
/* Waits for request id and hands over its transcript (free with vg_free).
 * times_ms (may be NULL) receives 5 values: latency (submit to transcript),
 * time queued for a slot, host planning + waiting for memory, upload + proof,
 * and the device bytes of its lease.  Each id is collected once. */
int vg_server_wait(vg_server *server, int64_t id, uint8_t **transcript, size_t *len, double *times_ms);
int vg_server_stats(vg_server *server, vg_server_stats_t *out);
/* Finishes every queued request, then releases the server (results not
 * collected are dropped). */
int vg_server_destroy(vg_server *server);

/* Output layer of the last proof (2^bits elements); out may be NULL to ask
 * for *bits alone.  Fails when cap < 2^bits. */
int vg_output(vg_ctx *ctx, vg_fe *out, size_t cap, int *bits);
/* Values of layer `layer` (0: the padded input layer, circuit_value[0] of
 * zk_prover.cpp:54-66; depth-1: the output layer) of the padded,
 * relay-extended circuit, as last evaluated; out may be NULL to ask for
 * *bits alone.  No proof may be in flight on ctx. */
int vg_layer_values(vg_ctx *ctx, int layer, vg_fe *out, size_t cap, int *bits);

/* Values a caller that drives the prover one message at a time checks its
 * own arguments against (include/virgo_prover.hpp does, for each argument
 * the reference verifier passes).  Available once a proof of ctx has
 * finished.  Writes min(cap, *count) values to out (may be NULL) and the
 * total to *count.  what:
 *  VG_VALUES_DRAWS: every field element the proof takes from glibc rand()
 *    (seed 1; each value = rand() % p as real, then rand() % p as img,
 *    prime_field.cpp:35-41), in draw order: r_0[b], r_1[b] (b = bits of the
 *    output layer, zk_verifier.cpp:929); then for each layer i = depth-1
 *    .. 1 with bl = bits of layer i-1: rho (:955), maskR[6] when i > 1 and
 *    maskr[2] (zk_prover.cpp:233-241, drawn in sumcheck_init), the mask
 *    polynomial's (2 bl + 1) * 2 + 7 coefficients (zk_prover.cpp:162-172,
 *    generate_maskpoly_pre_rho), r_u[bl], r_v[bl], r_c, random_combine,
 *    linear_coeff, linear_combine (zk_verifier.cpp:968-1002), alpha, beta
 *    (:1183); then the n-6 FRI challenges (vpd_verifier.cpp:60,
 *    commit_phase).  Values as drawn: the verifier zeroes r_v and r_c of
 *    layer 1 afterwards (zk_verifier.cpp:973-977).
 *  VG_VALUES_GENERAL: per layer i = depth-1 .. 1, the value of the wiring
 *    predicates at the layer's challenges that sumcheck_finalround adds
 *    (zk_prover.cpp:1173-1178; the verifier computes it at
 *    zk_verifier.cpp:1126-1147); 0 for layer 1.
 *  VG_VALUES_PUB_MASK / VG_VALUES_PRI_MASK: all_pub_mask as the verifier
 *    builds it (zk_verifier.cpp:983-1025, 1201-1202) and all_pri_mask as
 *    the prover and verifier build it (zk_prover.cpp:195-201, 170,
 *    zk_verifier.cpp:1203-1204), the lists the commitment used. */
#define VG_VALUES_DRAWS 0
#define VG_VALUES_GENERAL 1
#define VG_VALUES_PUB_MASK 2
#define VG_VALUES_PRI_MASK 3
int vg_proof_values(vg_ctx *ctx, int what, vg_fe *out, size_t cap, size_t *count);
/* The 33 query positions of ctx's last proof: each the first raw rand()
 * draw r % 2^(n-1) that is even and at least 2^(n-6), after the draws
 * above (vpd_verifier.cpp:127-130). */
int vg_query_positions(vg_ctx *ctx, int64_t *pos);

/* Verifies a transcript of this context's circuit, as vg_prove writes it,
 * with the checks of zk_verifier::verify (src/linear_gkr/zk_verifier.cpp:
 * 910-1253), poly_commit_verifier::verify_poly_commitment
 * (src/VPD/vpd_verifier.cpp:80-335) and the fft_gkr program.  The round
 * checks run on the host; the wiring predicates (zk_verifier.cpp:219-606),
 * public_array_prepare (:884-908), the first-round evaluations at the query
 * points (vpd_verifier.cpp:215-226) and every authentication path
 * (vpd_verifier.cpp:8-41) run on the device.  Returns 0 when the check ran:
 * *accepted = 1, or 0 with the failed check in vg_last_error().  ms: wall
 * time.  No proof may be in flight on ctx. */
int vg_verify(vg_ctx *ctx, const uint8_t *transcript, size_t len, int *accepted, double *ms);

/* ---- single-operation entry points on host arrays ---- */

/* fast_fourier_transform (RS_polynomial.cpp:27-167): result[k] =
 * sum_{j<coef_len} c_j w^{jk}, k < order, w = root of unity of order `order`. */
int vg_fft(const vg_fe *coef, int coef_len, int order, const vg_fe *rou, vg_fe *result);
/* Allocation-free variant of vg_fft.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  coef (coef_len values) and result (order
 * values) are DEVICE pointers; rou points to one host value.  Same result
 * as vg_fft.  Enqueues its kernels on `stream` (NULL: the null stream) and
 * returns; result is ready once that stream's work is done.  All working
 * memory is one static device array per GPU (64 MiB, held until the process
 * ends), so 2 <= order <= 2^20 (a larger order fails with an error) and no
 * hipMalloc or hipFree is called.  Calls that overlap on one GPU share that
 * array and race: serialise them (one stream, or wait between calls). */
int vg_fft_noalloc(const vg_fe *coef, int coef_len, int order, const vg_fe *rou, vg_fe *result, void *stream);
/* Measures this device's VALU issue rate for the two instruction classes of
 * the SHA3-256 leaf chains, 8 waves per SIMD, 16 independent chains per
 * lane: rates[0] = three-input bitwise operations (v_bitop3_b32) and
 * rates[1] = funnel shifts (v_alignbit_b32), in wave instructions per SIMD
 * per shader cycle; rates[2] = the shader clock during the runs, MHz (shader
 * cycle counter over the 100 MHz constant counter).  No counterpart in the
 * reference: the measured side of the leaf kernel's roofline. */
int vg_valu_rates(double *rates);
/* inverse_fast_fourier_transform (RS_polynomial.cpp:169-240) */
int vg_ifft(const vg_fe *evals, int coef_len, int order, const vg_fe *rou, vg_fe *result);
/* my_hhash (my_hhash.h:39-46): n independent SHA3-256 digests of 64 bytes */
int vg_sha3_64(const uint8_t *in, uint8_t *out, size_t n);
/* merkle_tree_prover::create_tree (merkle_tree.cpp:4-57): heap tree of
 * 2*pad digests, pad = next power of two >= n_leaf. */
int vg_merkle_tree(const uint8_t *leaves, int n_leaf, uint8_t *tree);
/* One FRI fold of a natural-order codeword of 2^log_n values over the
 * 2^log_n-th roots of unity w = root of order 2^log_n, as
 * fri::commit_phase_step folds each slice (fri.cpp:321-340): out[i] =
 * ((a + b) + r w^-i (a - b)) / 2 with a = cw[i], b = cw[i + 2^(log_n-1)],
 * i < 2^(log_n-1); the result is the codeword over the squared domain. */
int vg_fri_fold(const vg_fe *cw, int log_n, const vg_fe *r, vg_fe *out);
/* element-wise F_{p^2} product (prime_field.h:125-151) */
int vg_fe_mul(const vg_fe *a, const vg_fe *b, vg_fe *out, size_t n);

/* SHA-256 (FIPS 180-4) of n independent 64-byte messages, and the Merkle
 * tree of vg_merkle_tree with SHA-256 node hashes (pad leaves SHA-256 of 64
 * zero bytes).  No counterpart in the reference, whose Merkle hash is
 * SHA3-256 (my_hhash.h:39-46): an alternative hash for the kernel micro-set,
 * never used by proofs. */
int vg_sha256_64(const uint8_t *in, uint8_t *out, size_t n);
int vg_merkle_tree_sha256(const uint8_t *leaves, int n_leaf, uint8_t *tree);

/* Kernel micro-set on one polynomial of N = 2^log_n coefficients, every
 * stage device-resident and timed on the device: (0) forward transform of
 * size N (fast_fourier_transform, RS_polynomial.cpp:27-167), (1) its
 * inverse (RS_polynomial.cpp:169-240), (2) the Reed-Solomon extension to
 * M = N 2^rate_log points (the coset extension of commit_private_array,
 * poly_commit.h:91-104, at rate 2^-rate_log), (3) the Merkle tree over the
 * M/2 32-byte leaves (E[j], E[j + M/2]) of the extension E in natural order
 * (merkle_tree.cpp:4-57), (4) FRI commit: L - 5 folds of vg_fri_fold
 * (L = log_n + rate_log) with challenges fri_r[0..L-6], each folded
 * codeword of more than 32 values committed by the same tree.  Writes
 * roots[32 k] = root of level k's tree (level 0: E), k < L - 5; final_cw =
 * the 32 values left; *roundtrip_ok = 1 when the inverse returned coef
 * exactly; stage_ms[5] = mean device ms of each stage over reps runs after
 * one untimed run (either may be NULL).  hash: VG_HASH_SHA3_256 (the
 * reference's) or VG_HASH_SHA256.  No counterpart in the reference as one
 * call: it strings together the drop-ins above on device buffers. */
#define VG_HASH_SHA3_256 0
#define VG_HASH_SHA256 1
int vg_microset(int log_n, int rate_log, int hash, const vg_fe *coef, const vg_fe *fri_r, int reps, uint8_t *roots,
                vg_fe *final_cw, int *roundtrip_ok, double *stage_ms);

/* fft_gkr (src/VPD/fft_circuit_GKR.cpp:827-845; the program that
 * vpd_verifier.cpp:90-99 runs as "./fft_gkr lg log_fftgkr.txt"): the GKR
 * proof of the FFT/evaluation circuit of a commitment with 2^lg
 * coefficients, 1 <= lg <= 20, with its own challenge stream (rand() seed 1).
 * Writes the malloc'd "VGF1" transcript section (free with vg_free): F claim
 * on the 64 outputs; addition layer (lg+6) x {a,b,c} then v_u; mult layer
 * lg x {a,b,c} then v_u; per inverse-DFT layer lg x {a,b,c}, v_u,
 * lg x {a,b,c}, v_v.  accepted: the program's inline verdict (its `result`);
 * proof_size: bytes as the program counts them; p_ms: device time.  Any of
 * the last three may be NULL. */
int vg_fft_gkr(int lg, uint8_t **section, size_t *len, int *accepted, int *proof_size, double *p_ms);

#ifdef __cplusplus
}
#endif
#endif
