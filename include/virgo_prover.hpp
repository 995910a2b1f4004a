/*
 * virgo_prover.hpp -- header-only C++ mirror of the reference's prover
 * interface over the C ABI of virgo.h, for a verifier written against the
 * reference's classes (zk_verifier::verify, zk_verifier.cpp:910-1253, and
 * poly_commit_verifier::verify_poly_commitment, vpd_verifier.cpp:80-335).
 *
 * vg::zk_prover keeps the method names, argument meaning, public members and
 * call order of the reference's class zk_prover (zk_prover.h:16-122);
 * vg::poly_commit_prover (its member poly_prover) those of
 * poly_commit::poly_commit_prover (poly_commit.h:34-361,
 * vpd_verifier.cpp:46-78); namespace vg::fri those of the fri:: functions
 * and state the verifier calls and reads (fri.h:59-100).
 *
 * evaluate() runs the whole proof on the device: the reference draws every
 * challenge from glibc rand() with its default seed 1 and the order of draws
 * depends only on the circuit's shape, so all of them are known in advance.
 * Each later method hands out the next prover message.  Every argument that
 * carries a challenge, a query position, a verifier-computed value or a mask
 * list is compared with the one the device used; on a mismatch the method
 * throws std::runtime_error naming the argument, so a caller never receives
 * a message made for other challenges.  Errors of the library also throw
 * (the reference asserts).
 *
 * The rand() stream is shared, as in the reference: the verifier draws the
 * challenges, and the prover draws its masks inside sumcheck_init (maskR,
 * maskr; zk_prover.cpp:233-241) and generate_maskpoly_pre_rho
 * (zk_prover.cpp:162-172), and commit_phase draws the FRI challenges
 * (vpd_verifier.cpp:60).  The mirror makes those draws with rand() at the
 * same methods and checks each against the device's, so a verifier that
 * draws its own challenges with rand() stays in step, and one whose stream
 * is out of step is reported at the first prover draw.
 *
 * Also as in the reference, the fri:: functions are process-global: they
 * answer for the prover whose commit_private_array ran last.  Not reentrant.
 *
 * circuit_value holds the input layer ([0], the padded inputs) and the
 * output layer ([depth-1]) on the host; the layers between stay on the
 * device (nullptr here).  fri::cpd holds the final FRI codeword and its mask
 * codeword (rs_codeword[mx_depth-1], rs_codeword_msk[mx_depth-1], in the
 * reference's interleaved layout); the other levels stay on the device.
 * new_size of the query functions follows the reference's accounting
 * (fri.cpp:148-289): first visits of values and authentication-path nodes.
 */
#ifndef VIRGO_PROVER_HPP
#define VIRGO_PROVER_HPP

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "virgo.h"

namespace vg {

using hhash_digest = std::array<uint8_t, 32>;        // __hhash_digest (my_hhash.h)
struct quadratic_poly { vg_fe a, b, c; };            // polynomial.h: ax^2 + bx + c
struct quintuple_poly { vg_fe a, b, c, d, e, f; };   // polynomial.h: degree-5, member order
using query_answer = std::pair<std::vector<std::pair<vg_fe, vg_fe>>, std::vector<hhash_digest>>;
// poly_commit.h:24-32
struct ldt_commitment {
    std::vector<hhash_digest> commitment_hhash;
    std::vector<vg_fe> randomness;
    const vg_fe *final_rs_code = nullptr;             // cpd.rs_codeword[mx_depth - 1]
    int mx_depth = 0;
    int repeat_no = 33;
};

namespace field {
// F_{p^2}, p = 2^61 - 1, canonical representatives (prime_field.h:28-176)
constexpr uint64_t P = (1ULL << 61) - 1;
inline uint64_t addp(uint64_t a, uint64_t b) { uint64_t s = a + b; return s >= P ? s - P : s; }
inline uint64_t subp(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P - b; }
inline uint64_t mulp(uint64_t a, uint64_t b)
{
    unsigned __int128 x = (unsigned __int128)a * b;
    uint64_t r = ((uint64_t)x & P) + (uint64_t)(x >> 61);
    r = (r & P) + (r >> 61);
    return r >= P ? r - P : r;
}
inline vg_fe mk(uint64_t real, uint64_t img) { vg_fe x; x.real = real; x.img = img; return x; }
inline vg_fe add(vg_fe x, vg_fe y) { return mk(addp(x.real, y.real), addp(x.img, y.img)); }
inline vg_fe sub(vg_fe x, vg_fe y) { return mk(subp(x.real, y.real), subp(x.img, y.img)); }
inline vg_fe mul(vg_fe x, vg_fe y)
{
    return mk(subp(mulp(x.real, y.real), mulp(x.img, y.img)), addp(mulp(x.real, y.img), mulp(x.img, y.real)));
}
inline bool eq(vg_fe x, vg_fe y) { return x.real == y.real && x.img == y.img; }
inline vg_fe zero() { return mk(0, 0); }
inline vg_fe one() { return mk(1, 0); }
// prime_field::random (prime_field.cpp:35-41): real, then img
inline vg_fe random() { vg_fe x; x.real = (unsigned long long)rand() % P; x.img = (unsigned long long)rand() % P; return x; }
}  // namespace field

class zk_prover;

namespace fri {
// fri::commit_phase_data (fri.h:45-58), the parts the verifier reads
struct commit_phase_data {
    const vg_fe *rs_codeword[64] = {};
    const vg_fe *rs_codeword_msk[64] = {};
    int merkle_size[64] = {};
};
inline commit_phase_data cpd;
inline int log_current_witness_size_per_slice = 0, current_step_no = 0, witness_bit_length_per_slice = 0;
inline zk_prover *active = nullptr;                  // set by commit_private_array
query_answer request_init_value_with_merkle(long long pow_0, long long pow_1, int &new_size, int oracle_indicator);
query_answer request_step_commit(int lvl, long long pow, int &new_size);
hhash_digest commit_phase_step(const vg_fe &r);
const vg_fe *commit_phase_final();
}  // namespace fri

// poly_commit::poly_commit_prover (poly_commit.h:34-361)
class poly_commit_prover {
public:
    explicit poly_commit_prover(zk_prover *p) : p_(p) {}
    // poly_commit.h:42-124: the root of the private array's commitment (L)
    hhash_digest commit_private_array(const vg_fe *private_array, int log_array_length,
                                      const std::vector<vg_fe> &private_mask_array);
    // poly_commit.h:125-351: writes all_sum[65] (64 slices + mask), returns root H
    hhash_digest commit_public_array(const std::vector<vg_fe> &all_pub_msk, const vg_fe *public_array, int r_0_len,
                                     const vg_fe &target_sum, vg_fe *all_sum);
    // vpd_verifier.cpp:46-78: draws the n-6 FRI challenges, commits each level
    ldt_commitment commit_phase(int log_length);
    double total_time = 0;

private:
    zk_prover *p_;
};

class zk_prover {
public:
    explicit zk_prover(const vg_circuit_desc &c, int device = 0) : poly_prover(this)
    {
        check(vg_init(device));
        get_circuit(c);
    }
    ~zk_prover()
    {
        if (fri::active == this) fri::active = nullptr;
        release();
    }
    zk_prover(const zk_prover &) = delete;
    zk_prover &operator=(const zk_prover &) = delete;

    // ---- the reference's public members the verifier reads or writes ----
    vg_fe rho{}, maskpoly_sumc{}, v_u{}, v_v{}, Iuv{}, preZu{}, preZv{}, Zu{}, preu1{}, prev1{}, prepreu1{}, preprev1{};
    quadratic_poly sumRc{};
    std::vector<vg_fe> maskr = std::vector<vg_fe>(2);
    std::vector<vg_fe> all_pri_mask;
    std::vector<vg_fe *> circuit_value;               // [0] input layer, [depth-1] output layer
    double total_time = 0;                              // device seconds of the proof
    poly_commit_prover poly_prover;

    // zk_prover::get_circuit (zk_prover.cpp:24-29): the padded,
    // relay-extended circuit resident on the device
    void get_circuit(const vg_circuit_desc &c)
    {
        release();
        check(vg_ctx_create(&ctx_, &c, nullptr));
        bl_of_.clear();
        max_bl_ = 0;
        for (int i = 0;; ++i) {                           // bit lengths of every layer
            int b = 0;
            if (vg_layer_values(ctx_, i, nullptr, 0, &b)) break;
            bl_of_.push_back(b);
            if (b > max_bl_) max_bl_ = b;
        }
        depth_ = (int)bl_of_.size();
        if (depth_ < 2) fail("get_circuit: the context reports fewer than two layers");
        circuit_value.assign(depth_, nullptr);
    }
    // zk_prover::init_array (zk_prover.cpp:337-350); the device context
    // already holds every table, sized for the circuit
    void init_array(int max_bit_length)
    {
        if (max_bit_length < max_bl_) fail("init_array: max_bit_length below the circuit's widest layer");
    }
    void proof_init() {}

    // zk_prover::evaluate (zk_prover.cpp:54-157): runs the whole proof on the
    // device and returns the output layer
    vg_fe *evaluate()
    {
        uint8_t *tr = nullptr;
        size_t len = 0;
        double ms[13];
        check(vg_prove(ctx_, &tr, &len, ms));
        tr_.assign(tr, tr + len);
        vg_free(tr);
        total_time = ms[8] * 1e-3;
        int bits = 0;
        check(vg_output(ctx_, nullptr, 0, &bits));
        output_.assign((size_t)1 << bits, field::zero());
        check(vg_output(ctx_, output_.data(), output_.size(), &bits));
        check(vg_layer_values(ctx_, 0, nullptr, 0, &bits));
        input_.assign((size_t)1 << bits, field::zero());
        check(vg_layer_values(ctx_, 0, input_.data(), input_.size(), &bits));
        circuit_value.assign(depth_, nullptr);
        circuit_value[0] = input_.data();
        circuit_value[depth_ - 1] = output_.data();
        fetch(VG_VALUES_DRAWS, draws_);
        fetch(VG_VALUES_GENERAL, general_);
        fetch(VG_VALUES_PUB_MASK, pub_);
        fetch(VG_VALUES_PRI_MASK, pri_);
        check(vg_query_positions(ctx_, pos_));
        parse_header();
        index_draws();
        reset_state();
        return output_.data();
    }
    // zk_prover::V_res (zk_prover.cpp:30-52)
    vg_fe V_res(const vg_fe *one_minus_r_0, const vg_fe *r_0, const vg_fe *output_raw, int r_0_size, int output_size)
    {
        need_proof("V_res");
        int b = bl_of_[depth_ - 1];
        if (r_0_size != b || output_size != (1 << b)) fail("V_res: sizes differ from the output layer's");
        expect_vec("V_res: r_0", r_0, draws_.data() + r0_at_, b);
        expect_one_minus("V_res: one_minus_r_0", one_minus_r_0, draws_.data() + r0_at_, b);
        if (output_raw != output_.data() &&
            (!output_raw || std::memcmp(output_raw, output_.data(), sizeof(vg_fe) * output_.size())))
            fail("V_res: output values differ from the evaluated circuit's");
        return vres_;
    }

    // zk_prover::sumcheck_init (zk_prover.cpp:317-335) with generate_maskR
    // (:192-255): layers run from depth-1 down to 1; alpha, beta, r_0, r_1 are
    // the previous layer's (1, 0 and the initial draws at the top)
    void sumcheck_init(int layer_id, int bit_length_g, int bit_length_u, int bit_length_v, const vg_fe &alpha,
                       const vg_fe &beta, const vg_fe *r_0, const vg_fe *r_1, const vg_fe *one_minus_r_0,
                       const vg_fe *one_minus_r_1)
    {
        need_proof("sumcheck_init");
        if (layer_id != next_layer_ || layer_id < 1) fail("sumcheck_init: layers must run from depth-1 down to 1");
        cur_ = layer_id;
        const layer_view &l = L_[cur_];
        bl_ = l.bl;
        if (bit_length_g != bl_of_[cur_] || bit_length_u != bl_ || bit_length_v != bl_)
            fail("sumcheck_init: bit lengths differ from the circuit's");
        int bg = bl_of_[cur_];
        if (cur_ == depth_ - 1) {
            expect("sumcheck_init: alpha", alpha, field::one());
            expect("sumcheck_init: beta", beta, field::zero());
            expect_vec("sumcheck_init: r_0", r_0, draws_.data() + r0_at_, bg);
            expect_vec("sumcheck_init: r_1", r_1, draws_.data() + r0_at_ + bg, bg);
            expect_one_minus("sumcheck_init: one_minus_r_0", one_minus_r_0, draws_.data() + r0_at_, bg);
            expect_one_minus("sumcheck_init: one_minus_r_1", one_minus_r_1, draws_.data() + r0_at_ + bg, bg);
        } else {
            const layer_view &up = L_[cur_ + 1];
            expect("sumcheck_init: alpha", alpha, up.d[up.alpha]);
            expect("sumcheck_init: beta", beta, up.d[up.alpha + 1]);
            expect_vec("sumcheck_init: r_0", r_0, up.d + up.r_u, bg);
            expect_vec("sumcheck_init: r_1", r_1, up.d + up.r_v, bg);
            expect_one_minus("sumcheck_init: one_minus_r_0", one_minus_r_0, up.d + up.r_u, bg);
            expect_one_minus("sumcheck_init: one_minus_r_1", one_minus_r_1, up.d + up.r_v, bg);
        }
        // generate_maskR: the last layer's masks join all_pri_mask, then fresh ones
        for (int i = 0; i < 6; ++i) all_pri_mask.push_back(maskR_[i]);
        prepreu1 = preu1;
        preprev1 = prev1;
        if (cur_ > 1)
            for (int i = 0; i < 6; ++i) maskR_[i] = draw("sumcheck_init: maskR", l.d[l.maskR + i]);
        maskr[0] = draw("sumcheck_init: maskr", l.d[l.maskr]);
        maskr[1] = draw("sumcheck_init: maskr", l.d[l.maskr + 1]);
        if (cur_ == 1) {
            maskR_[0] = maskr[0];
            maskR_[1] = maskr[1];
            for (int i = 2; i < 6; ++i) maskR_[i] = field::zero();
        }
        maskpoly_sumc = field::zero();
        stage_ = 1;
        --next_layer_;
    }
    // zk_prover::generate_maskpoly_pre_rho (zk_prover.cpp:162-172)
    void generate_maskpoly_pre_rho(int length, int degree)
    {
        if (stage_ != 1) fail("generate_maskpoly_pre_rho out of the verifier's order");
        const layer_view &l = L_[cur_];
        if (length != 2 * bl_ + 1 || degree != 2) fail("generate_maskpoly_pre_rho: length or degree differ from the layer's");
        for (int i = 0; i < l.n_maskpoly; ++i) all_pri_mask.push_back(draw("generate_maskpoly_pre_rho", l.d[l.maskpoly + i]));
        stage_ = 2;
    }
    // zk_prover::generate_maskpoly_after_rho (zk_prover.cpp:174-188); rho is
    // the member the verifier sets before the call
    void generate_maskpoly_after_rho(int length, int degree)
    {
        if (stage_ != 2) fail("generate_maskpoly_after_rho out of the verifier's order");
        if (length != 2 * bl_ + 1 || degree != 2) fail("generate_maskpoly_after_rho: length or degree differ from the layer's");
        expect("generate_maskpoly_after_rho: rho", rho, L_[cur_].d[L_[cur_].rho]);
        maskpoly_sumc = fe();
        stage_ = 3;
    }
    void sumcheck_phase1_init()
    {
        if (stage_ != 3) fail("sumcheck_phase1_init out of the verifier's order");
        round_ = 0;
        stage_ = 4;
    }
    // zk_prover.cpp:596-787: previous_random is 0, then r_u[j-1]
    quadratic_poly sumcheck_phase1_update(const vg_fe &previous_random, int current_bit)
    {
        round_check(4, "sumcheck_phase1_update", previous_random, current_bit, false);
        return q3();
    }
    quintuple_poly sumcheck_phase1_updatelastbit(const vg_fe &previous_random, int current_bit)
    {
        round_check(4, "sumcheck_phase1_updatelastbit", previous_random, current_bit, true);
        stage_ = 5;
        return q6();
    }
    // zk_prover.cpp:789-983: previous_random = r_u[bl-1]
    void sumcheck_phase2_init(const vg_fe &previous_random, const vg_fe *r_u, const vg_fe *one_minus_r_u)
    {
        if (stage_ != 5) fail("sumcheck_phase2_init out of the verifier's order");
        const layer_view &l = L_[cur_];
        expect("sumcheck_phase2_init: previous_random", previous_random, l.d[l.r_u + bl_ - 1]);
        expect_vec("sumcheck_phase2_init: r_u", r_u, l.d + l.r_u, bl_);
        expect_one_minus("sumcheck_phase2_init: one_minus_r_u", one_minus_r_u, l.d + l.r_u, bl_);
        preu1 = previous_random;
        v_u = fe();
        round_ = 0;
        stage_ = 6;
    }
    // zk_prover.cpp:988-1171: previous_random is 0, then r_v[j-1] (0 on layer 1)
    quadratic_poly sumcheck_phase2_update(const vg_fe &previous_random, int current_bit)
    {
        round_check(6, "sumcheck_phase2_update", previous_random, current_bit, false);
        return q3();
    }
    quintuple_poly sumcheck_phase2_updatelastbit(const vg_fe &previous_random, int current_bit)
    {
        round_check(6, "sumcheck_phase2_updatelastbit", previous_random, current_bit, true);
        stage_ = 7;
        return q6();
    }
    // zk_prover.cpp:1205-1214: previous_random = r_v[bl-1]
    std::pair<vg_fe, vg_fe> sumcheck_finalize(const vg_fe &previous_random)
    {
        if (stage_ != 7) fail("sumcheck_finalize out of the verifier's order");
        expect("sumcheck_finalize: previous_random", previous_random, rv(bl_ - 1));
        prev1 = previous_random;
        v_v = fe();
        stage_ = 8;
        return {v_u, v_v};
    }
    // zk_prover.cpp:1173-1203: general_value is the verifier's value of the
    // wiring predicates at the layer's challenges
    quadratic_poly sumcheck_finalround(const vg_fe &previous_random, int current, const vg_fe &general_value)
    {
        if (stage_ != 8) fail("sumcheck_finalround out of the verifier's order");
        expect("sumcheck_finalround: previous_random", previous_random, rv(bl_ - 1));
        if (current != bl_ << 1) fail("sumcheck_finalround: current differs from 2 * bit length");
        expect("sumcheck_finalround: general_value", general_value, general_[depth_ - 1 - cur_]);
        quadratic_poly p = q3();
        query_ = fe(); rg1_ = fe(); rg2_ = fe();
        Iuv = fe(); preZu = fe(); preZv = fe();
        stage_ = 9;
        if (cur_ == 1) parse_commitment();            // Zu, preu1, sumRc, then the commitment sections
        return p;
    }
    // zk_prover::query / queryRg1 / queryRg2 (zk_prover.cpp:257-315): r_v and
    // r_c are 0 on layer 1
    vg_fe query(const vg_fe *r_u, const vg_fe *r_v, const vg_fe &r_c)
    {
        if (stage_ != 9) fail("query out of the verifier's order");
        const layer_view &l = L_[cur_];
        expect_vec("query: r_u", r_u, l.d + l.r_u, bl_);
        for (int k = 0; k < bl_; ++k) expect("query: r_v", r_v[k], rv(k));
        expect("query: r_c", r_c, rc());
        return query_;
    }
    vg_fe queryRg1(const vg_fe &r_c)
    {
        if (stage_ != 9) fail("queryRg1 out of the verifier's order");
        expect("queryRg1: r_c", r_c, rc());
        return rg1_;
    }
    vg_fe queryRg2(const vg_fe &r_c)
    {
        if (stage_ != 9) fail("queryRg2 out of the verifier's order");
        expect("queryRg2: r_c", r_c, rc());
        return rg2_;
    }
    void delete_self() { release(); }

    // ---- raw transcript sections (a caller without the reference's classes) ----
    const std::vector<uint8_t> &transcript() const { return tr_; }
    std::pair<const uint8_t *, const uint8_t *> commitment_roots() const { return {root_l_, root_h_}; }
    const vg_fe *all_sum() const { return all_sum_; }
    std::pair<const uint8_t *, size_t> fft_gkr_section() const { return {fg_sec_, fg_len_}; }
    // the general value the device used for layer i (what the verifier's
    // predicates must give, zk_verifier.cpp:1126-1147)
    vg_fe general_value(int layer_id) const { return general_[depth_ - 1 - layer_id]; }
    int depth() const { return depth_; }
    int bit_length(int layer) const
    {
        if (layer < 0 || layer >= depth_) fail("bit_length: no layer " + std::to_string(layer));
        return bl_of_[layer];
    }

private:
    friend class poly_commit_prover;
    friend query_answer fri::request_init_value_with_merkle(long long, long long, int &, int);
    friend query_answer fri::request_step_commit(int, long long, int &);
    friend hhash_digest fri::commit_phase_step(const vg_fe &);
    friend const vg_fe *fri::commit_phase_final();

    // offsets of one layer's values in the draw list (VG_VALUES_DRAWS)
    struct layer_view {
        const vg_fe *d = nullptr;
        int bl = 0, rho = 0, maskR = 0, maskr = 0, maskpoly = 0, n_maskpoly = 0, r_u = 0, r_v = 0, r_c = 0, alpha = 0;
    };

    vg_ctx *ctx_ = nullptr;
    std::vector<uint8_t> tr_;
    std::vector<vg_fe> output_, input_, draws_, general_, pub_, pri_;
    int64_t pos_[33] = {};
    std::vector<int> bl_of_;
    std::vector<layer_view> L_;
    size_t off_ = 0, draw_at_ = 0, r0_at_ = 0, fri_at_ = 0;
    int depth_ = 0, max_bl_ = 0, n_ = 0, next_layer_ = 0, cur_ = 0, bl_ = 0, stage_ = 0, round_ = 0;
    vg_fe vres_{}, query_{}, rg1_{}, rg2_{}, maskR_[6] = {};
    const uint8_t *root_l_ = nullptr, *root_h_ = nullptr, *fri_roots_ = nullptr, *query_sec_ = nullptr,
                  *fg_sec_ = nullptr;
    const vg_fe *all_sum_ = nullptr, *final_cw_ = nullptr;
    size_t query_len_ = 0, fg_len_ = 0, qoff_ = 0;
    // commitment progress: 0 nothing, 1 private array, 2 public array, 3 FRI levels, 4 queries
    int com_stage_ = 0, fri_step_ = 0, rep_ = 0, qlvl_ = 0, init_next_ = 0;
    std::vector<std::vector<bool>> seen_w_, seen_i_, seen_s_;   // visited_witness, visited_init, visited

    static void check(int rc)
    {
        if (rc) throw std::runtime_error(std::string("virgo: ") + vg_last_error());
    }
    [[noreturn]] static void fail(const std::string &m) { throw std::runtime_error("virgo: " + m); }
    void release()
    {
        if (ctx_) { vg_ctx_destroy(ctx_); ctx_ = nullptr; }
        tr_.clear();
    }
    void need_proof(const char *what) const
    {
        if (tr_.empty()) fail(std::string(what) + " before evaluate()");
    }
    void fetch(int what, std::vector<vg_fe> &v)
    {
        size_t n = 0;
        check(vg_proof_values(ctx_, what, nullptr, 0, &n));
        v.assign(n, field::zero());
        check(vg_proof_values(ctx_, what, v.data(), n, &n));
    }
    static void expect(const char *what, const vg_fe &got, const vg_fe &want)
    {
        if (!field::eq(got, want)) fail(std::string(what) + " differs from the value the device used");
    }
    static void expect_vec(const char *what, const vg_fe *got, const vg_fe *want, int n)
    {
        if (n > 0 && !got) fail(std::string(what) + " is null");
        for (int k = 0; k < n; ++k)
            if (!field::eq(got[k], want[k])) fail(std::string(what) + "[" + std::to_string(k) + "] differs from the value the device used");
    }
    static void expect_one_minus(const char *what, const vg_fe *got, const vg_fe *r, int n)
    {
        if (n > 0 && !got) fail(std::string(what) + " is null");
        for (int k = 0; k < n; ++k)
            if (!field::eq(got[k], field::sub(field::one(), r[k])))
                fail(std::string(what) + "[" + std::to_string(k) + "] is not 1 - r");
    }
    // a draw the prover makes from the shared rand() stream, which must be
    // the value the device used
    vg_fe draw(const char *what, const vg_fe &want)
    {
        vg_fe x = field::random();
        if (!field::eq(x, want))
            fail(std::string(what) + ": the rand() stream is out of step with the proof's schedule");
        return x;
    }
    // r_v and r_c as the verifier passes them (zero on layer 1, zk_verifier.cpp:973-977, 1078-1079)
    vg_fe rv(int k) const { return cur_ == 1 ? field::zero() : L_[cur_].d[L_[cur_].r_v + k]; }
    vg_fe rc() const { return cur_ == 1 ? field::zero() : L_[cur_].d[L_[cur_].r_c]; }
    void round_check(int stage, const char *what, const vg_fe &previous_random, int current_bit, bool last)
    {
        if (stage_ != stage) fail(std::string(what) + " out of the verifier's order");
        if (current_bit != round_ || (last != (round_ == bl_ - 1))) fail(std::string(what) + ": current_bit out of order");
        const layer_view &l = L_[cur_];
        vg_fe want = round_ == 0 ? field::zero() : stage == 4 ? l.d[l.r_u + round_ - 1] : rv(round_ - 1);
        expect((std::string(what) + ": previous_random").c_str(), previous_random, want);
        ++round_;
    }
    void reset_state()
    {
        next_layer_ = depth_ - 1;
        stage_ = 0;
        for (auto &x : maskR_) x = field::zero();
        preu1 = prev1 = prepreu1 = preprev1 = field::zero();
        all_pri_mask.clear();
        com_stage_ = 0;
        fri_step_ = rep_ = qlvl_ = init_next_ = 0;
        qoff_ = 0;
        seen_w_.clear(); seen_i_.clear(); seen_s_.clear();
    }
    void index_draws()
    {
        size_t k = 0;
        int bo = bl_of_[depth_ - 1];
        r0_at_ = k;
        k += 2 * (size_t)bo;
        L_.assign(depth_, layer_view());
        std::vector<size_t> base(depth_, 0);
        for (int i = depth_ - 1; i >= 1; --i) {
            layer_view &l = L_[i];
            int bl = l.bl = bl_of_[i - 1];
            base[i] = k;
            l.rho = 0;
            int o = 1;
            if (i > 1) { l.maskR = o; o += 6; }
            l.maskr = o; o += 2;
            l.maskpoly = o; l.n_maskpoly = (2 * bl + 1) * 2 + 7; o += l.n_maskpoly;
            l.r_u = o; o += bl;
            l.r_v = o; o += bl;
            l.r_c = o; o += 4;                            // r_c, random_combine, linear_coeff, linear_combine
            l.alpha = o; o += 2;
            k += o;
        }
        fri_at_ = k;
        if (k + (size_t)(n_ - 6) != draws_.size()) fail("the proof's draw list does not match the circuit's shape");
        for (int i = depth_ - 1; i >= 1; --i) L_[i].d = draws_.data() + base[i];
        if (general_.size() != (size_t)(depth_ - 1)) fail("the proof's general values do not match the circuit's depth");
    }
    const uint8_t *take(size_t k)
    {
        if (off_ + k > tr_.size()) fail("transcript shorter than its layout");
        const uint8_t *p = tr_.data() + off_;
        off_ += k;
        return p;
    }
    vg_fe fe() { vg_fe x; std::memcpy(&x, take(16), 16); return x; }
    uint32_t u32() { uint32_t x; std::memcpy(&x, take(4), 4); return x; }
    quadratic_poly q3() { quadratic_poly p; p.a = fe(); p.b = fe(); p.c = fe(); return p; }
    quintuple_poly q6() { quintuple_poly p; p.a = fe(); p.b = fe(); p.c = fe(); p.d = fe(); p.e = fe(); p.f = fe(); return p; }
    static hhash_digest digest(const uint8_t *p)
    {
        hhash_digest d;
        std::memcpy(d.data(), p, 32);
        return d;
    }
    // "VGT1", depth, input bits, bit lengths, V_res
    void parse_header()
    {
        off_ = 0;
        if (std::memcmp(take(4), "VGT1", 4)) fail("bad transcript header");
        int depth = (int)u32();
        if (depth != depth_) fail("transcript depth differs from the circuit's");
        n_ = (int)u32();
        if (n_ != bl_of_[0]) fail("transcript input bits differ from the circuit's");
        for (int i = 0; i < depth; ++i)
            if ((int)u32() != bl_of_[i]) fail("transcript bit lengths differ from the circuit's");
        vres_ = fe();
    }
    // after the last layer's query values: Zu, preu1, sumRc (read by the
    // verifier after its layer loop, zk_verifier.cpp:1201-1207), then the
    // commitment sections
    void parse_commitment()
    {
        Zu = fe();
        vg_fe pu = fe();
        if (!field::eq(pu, preu1)) fail("preu1 of the transcript differs from the phase-2 challenge");
        sumRc.a = fe(); sumRc.b = fe(); sumRc.c = fe();
        root_l_ = take(32);
        root_h_ = take(32);
        all_sum_ = (const vg_fe *)take(65 * 16);
        fri_roots_ = take((size_t)32 * (n_ - 6));
        final_cw_ = (const vg_fe *)take((2048 + 32) * 16);
        size_t q = 0;
        for (int rep = 0; rep < 33; ++rep) {
            q += 2 * (130 * 16 + (size_t)(n_ - 1) * 32);
            for (int lvl = 0; lvl < n_ - 6; ++lvl) q += 130 * 16 + (size_t)(n_ - 2 - lvl) * 32;
        }
        query_sec_ = take(q);
        query_len_ = q;
        fg_sec_ = tr_.data() + off_;
        fg_len_ = tr_.size() - off_;
    }
    // the next 65 value pairs and `path` digests of the query section
    query_answer answer(int path)
    {
        size_t need = 65 * 32 + (size_t)path * 32;
        if (!query_sec_ || qoff_ + need > query_len_) fail("no query answer left");
        query_answer a;
        const uint8_t *q = query_sec_ + qoff_;
        for (int i = 0; i < 65; ++i) {
            vg_fe x, y;
            std::memcpy(&x, q + 32 * i, 16);
            std::memcpy(&y, q + 32 * i + 16, 16);
            a.first.push_back({x, y});
        }
        for (int k = 0; k < path; ++k) a.second.push_back(digest(q + 65 * 32 + 32 * k));
        qoff_ += need;
        return a;
    }

    // ---- commitment (called through poly_prover and fri::) ----
    hhash_digest commit_private(const vg_fe *arr, int log_len, const std::vector<vg_fe> &pri)
    {
        need_proof("commit_private_array");
        if (next_layer_ != 0 || stage_ != 9 || com_stage_ != 0) fail("commit_private_array out of the verifier's order");
        if (log_len != n_) fail("commit_private_array: log_array_length differs from the input layer's");
        if (arr != input_.data() && (!arr || std::memcmp(arr, input_.data(), sizeof(vg_fe) * input_.size())))
            fail("commit_private_array: private_array differs from the circuit's input layer");
        if (pri.size() != pri_.size()) fail("commit_private_array: all_pri_mask has a different length");
        for (size_t k = 0; k < pri.size(); ++k)
            if (!field::eq(pri[k], pri_[k]))
                fail("commit_private_array: all_pri_mask[" + std::to_string(k) + "] differs from the device's");
        fri::active = this;
        fri::log_current_witness_size_per_slice = n_ - 1;   // fri.cpp:46, rate 1/32, 64 slices
        fri::witness_bit_length_per_slice = n_ - 6;
        fri::current_step_no = 0;
        fri::cpd = fri::commit_phase_data();
        com_stage_ = 1;
        return digest(root_l_);
    }
    hhash_digest commit_public(const std::vector<vg_fe> &pub, const vg_fe *q, int r0_len, const vg_fe &target, vg_fe *all_sum)
    {
        if (com_stage_ != 1) fail("commit_public_array out of the verifier's order");
        if (r0_len != n_) fail("commit_public_array: r_0_len differs from the input layer's");
        if (pub.size() != pub_.size()) fail("commit_public_array: all_pub_msk has a different length");
        for (size_t k = 0; k < pub.size(); ++k)
            if (!field::eq(pub[k], pub_[k]))
                fail("commit_public_array: all_pub_msk[" + std::to_string(k) + "] differs from the device's");
        // public_array: eq(r, x), bit b of x taking r[b] (dfs_for_public_eval,
        // zk_verifier.cpp:838-847), r = the last layer's r_u
        const layer_view &l1 = L_[1];
        std::vector<vg_fe> e(1, field::one());
        for (int b = 0; b < n_; ++b) {
            vg_fe r = l1.d[l1.r_u + b], om = field::sub(field::one(), r);
            std::vector<vg_fe> nx(e.size() * 2);
            for (size_t x = 0; x < e.size(); ++x) { nx[x] = field::mul(e[x], om); nx[x + e.size()] = field::mul(e[x], r); }
            e.swap(nx);
        }
        if (!q) fail("commit_public_array: public_array is null");
        for (size_t x = 0; x < e.size(); ++x)
            if (!field::eq(q[x], e[x])) fail("commit_public_array: public_array[" + std::to_string(x) + "] is not eq(r_0, x)");
        vg_fe tot = field::zero();
        for (int k = 0; k < 65; ++k) tot = field::add(tot, all_sum_[k]);
        expect("commit_public_array: target_sum", target, tot);
        if (all_sum) std::memcpy(all_sum, all_sum_, 65 * sizeof(vg_fe));
        com_stage_ = 2;
        return digest(root_h_);
    }
    hhash_digest fri_step(const vg_fe &r)
    {
        if (com_stage_ != 2 && com_stage_ != 3) fail("commit_phase_step before commit_public_array");
        if (fri_step_ >= n_ - 6) fail("commit_phase_step: every FRI level is committed");
        expect("commit_phase_step: r", r, draws_[fri_at_ + fri_step_]);
        com_stage_ = 3;
        fri::cpd.merkle_size[fri_step_] = 1 << (n_ - 3 - fri_step_);
        hhash_digest d = digest(fri_roots_ + 32 * fri_step_);
        ++fri_step_;
        fri::current_step_no = fri_step_;
        --fri::log_current_witness_size_per_slice;
        if (fri_step_ == n_ - 6) {
            fri::cpd.rs_codeword[fri_step_ - 1] = final_cw_;
            fri::cpd.rs_codeword_msk[fri_step_ - 1] = final_cw_ + 2048;
        }
        return d;
    }
    ldt_commitment commit_phase(int log_length)
    {
        if (log_length != n_) fail("commit_phase: log_length differs from the input layer's");
        if (com_stage_ != 2 || fri_step_ != 0) fail("commit_phase out of the verifier's order");
        int saved = fri::log_current_witness_size_per_slice;
        ldt_commitment c;
        for (int k = 0; k < n_ - 6; ++k) {
            vg_fe r = draw("commit_phase", draws_[fri_at_ + k]);
            c.randomness.push_back(r);
            c.commitment_hhash.push_back(fri_step(r));
        }
        c.final_rs_code = final_cw_;
        c.mx_depth = n_ - 6;
        fri::log_current_witness_size_per_slice = saved;   // vpd_verifier.cpp:50,70
        com_stage_ = 4;
        return c;
    }
    // fri.cpp:148-201: s0 = pow/2 and s1 = s0 + 2^(n-2) of the repetition's
    // position, L then H
    query_answer init_query(long long pow_0, long long pow_1, int &new_size, int oracle)
    {
        if (com_stage_ != 4 || qlvl_ != 0 || rep_ >= 33 || init_done_ || oracle != init_next_)
            fail("request_init_value_with_merkle out of the verifier's order");
        if (pow_0 > pow_1) std::swap(pow_0, pow_1);
        long long s0 = pos_[rep_] / 2, half = 1LL << (n_ - 2);
        if (pow_0 != s0 || pow_1 != s0 + half) fail("request_init_value_with_merkle: query position differs from the device's");
        if (seen_w_.empty()) {
            seen_w_.assign(2, std::vector<bool>((size_t)half << 7));
            seen_i_.assign(2, std::vector<bool>((size_t)half << 1));
        }
        new_size = 0;
        for (int i = 0; i < 64; ++i)
            for (int h = 0; h < 2; ++h) {
                size_t idx = (size_t)pow_0 << 7 | (size_t)i << 1 | h;
                if (!seen_w_[oracle][idx]) { seen_w_[oracle][idx] = true; new_size += 16; }
            }
        long long pos = pow_0 + half;
        for (int d = 0; d < n_ - 2; ++d) {
            if (!seen_i_[oracle][pos ^ 1]) new_size += 32;
            seen_i_[oracle][pos] = true;
            seen_i_[oracle][pos ^ 1] = true;
            pos /= 2;
        }
        init_next_ ^= 1;
        return answer(n_ - 1);
    }
    // fri.cpp:227-289: pow is (position mod 2^(n-1-lvl)) / 2
    query_answer step_query(int lvl, long long pow, int &new_size)
    {
        if (com_stage_ != 4 || rep_ >= 33 || lvl != qlvl_ || (lvl == 0 && !init_done_))
            fail("request_step_commit out of the verifier's order");
        long long want = (pos_[rep_] % (1LL << (n_ - 1 - lvl))) / 2;
        if (pow != want) fail("request_step_commit: query position differs from the device's");
        if (seen_s_.empty()) seen_s_.resize(n_ - 6);
        long long half = 1LL << (n_ - 3 - lvl);
        std::vector<bool> &vis = seen_s_[lvl];
        if (vis.empty()) vis.assign((size_t)half << 7, false);     // codeword indices and tree nodes share it
        long long q = pow % half;
        bool seen_value = false;
        for (int i = 0; i < 64; ++i)
            if (vis[(size_t)q << 7 | (size_t)i << 1]) seen_value = true;
        new_size = seen_value ? 0 : 16;
        for (long long node = q + half; node != 1; node /= 2)
            if (!vis[node ^ 1]) { new_size += 32; vis[node ^ 1] = true; vis[node] = true; }
        query_answer a = answer(n_ - 2 - lvl);
        ++qlvl_;
        if (qlvl_ == n_ - 6) { qlvl_ = 0; ++rep_; init_done_ = false; }
        return a;
    }
    bool init_done_ = false;
    query_answer init_query_counted(long long pow_0, long long pow_1, int &new_size, int oracle)
    {
        query_answer a = init_query(pow_0, pow_1, new_size, oracle);
        if (oracle == 1) init_done_ = true;
        return a;
    }
};

inline hhash_digest poly_commit_prover::commit_private_array(const vg_fe *private_array, int log_array_length,
                                                             const std::vector<vg_fe> &private_mask_array)
{
    return p_->commit_private(private_array, log_array_length, private_mask_array);
}
inline hhash_digest poly_commit_prover::commit_public_array(const std::vector<vg_fe> &all_pub_msk, const vg_fe *public_array,
                                                            int r_0_len, const vg_fe &target_sum, vg_fe *all_sum)
{
    return p_->commit_public(all_pub_msk, public_array, r_0_len, target_sum, all_sum);
}
inline ldt_commitment poly_commit_prover::commit_phase(int log_length)
{
    ldt_commitment c = p_->commit_phase(log_length);
    total_time = 0;
    return c;
}

namespace fri {
inline zk_prover &need_active()
{
    if (!active) throw std::runtime_error("virgo: fri:: called before commit_private_array");
    return *active;
}
inline query_answer request_init_value_with_merkle(long long pow_0, long long pow_1, int &new_size, int oracle_indicator)
{
    return need_active().init_query_counted(pow_0, pow_1, new_size, oracle_indicator);
}
inline query_answer request_step_commit(int lvl, long long pow, int &new_size)
{
    return need_active().step_query(lvl, pow, new_size);
}
inline hhash_digest commit_phase_step(const vg_fe &r) { return need_active().fri_step(r); }
inline const vg_fe *commit_phase_final()
{
    zk_prover &p = need_active();
    if (p.fri_step_ == 0) throw std::runtime_error("virgo: commit_phase_final before any FRI level");
    return cpd.rs_codeword[p.fri_step_ - 1] ? cpd.rs_codeword[p.fri_step_ - 1] : p.final_cw_;
}
}  // namespace fri

}  // namespace vg
#endif
