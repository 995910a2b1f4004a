/*
 * virgo_prover.hpp -- header-only C++ mirror of the reference's prover
 * interface, over the C ABI of virgo.h.
 *
 * vg::zk_prover keeps the method names, argument meaning and call order of
 * the reference's class zk_prover (include/linear_gkr/zk_prover.h:16-122) as
 * zk_verifier::verify drives it (src/linear_gkr/zk_verifier.cpp:910-1253),
 * so a verifier written against that interface runs unchanged on top of the
 * GPU prover.  evaluate() runs the whole proof on the device at once (the
 * challenge stream is the reference's glibc rand() with seed 1, so every
 * challenge is known in advance); each later method returns the next prover
 * message in the verifier's order.  The challenges a caller passes are the
 * ones the device already used; they are accepted for signature
 * compatibility and not re-read.  Errors throw std::runtime_error with the
 * library's message (the reference asserts).
 *
 * The commitment side keeps the names of poly_commit::poly_commit_prover
 * (poly_commit.h:42-351) and of the fri:: query functions (fri.cpp:148-289)
 * as verify_poly_commitment calls them (vpd_verifier.cpp:80-335):
 * commit_private_array() returns root L, commit_public_array() fills all_sum
 * and returns root H, commit_phase() returns the FRI roots and the final
 * codeword, and request_init_value_with_merkle() / request_step_commit()
 * return the 65 value pairs and the authentication path of each query in the
 * verifier's order (33 repetitions: L, H, then one step per FRI level).
 * new_size receives the bytes of that answer (the reference's proof-size
 * statistic also drops what earlier queries already sent; that accounting is
 * not reproduced).  The raw sections stay available: commitment_roots(),
 * all_sum(), fri_roots(), final_codeword(), query_section(),
 * fft_gkr_section().
 */
#ifndef VIRGO_PROVER_HPP
#define VIRGO_PROVER_HPP

#include <array>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "virgo.h"

namespace vg {

using hhash_digest = std::array<uint8_t, 32>;        // __hhash_digest (my_hhash.h)
// poly_commit.h:24-32: the FRI roots of commit_phase and its final codeword
struct ldt_commitment {
    std::vector<hhash_digest> commitment_hhash;
    const vg_fe *final_rs_code = nullptr;             // 2048 values (+ 32 of the mask)
    int mx_depth = 0;
    int repeat_no = 33;
};
using query_answer = std::pair<std::vector<std::pair<vg_fe, vg_fe>>, std::vector<hhash_digest>>;
struct quadratic_poly { vg_fe a, b, c; };            // polynomial.h: ax^2 + bx + c
struct quintuple_poly { vg_fe a, b, c, d, e, f; };   // polynomial.h: degree-5 member order

class zk_prover {
public:
    explicit zk_prover(const vg_circuit_desc &c, int device = 0) { check(vg_init(device)); get_circuit(c); }
    ~zk_prover() { release(); }
    zk_prover(const zk_prover &) = delete;
    zk_prover &operator=(const zk_prover &) = delete;

    // zk_prover::get_circuit (zk_prover.cpp:337-350) + init_array: the
    // padded, relay-extended circuit resident on the device
    void get_circuit(const vg_circuit_desc &c)
    {
        release();
        check(vg_ctx_create(&ctx_, &c, nullptr));
    }
    void proof_init() {}

    // zk_prover::evaluate (zk_prover.cpp:54-157): runs the whole proof and
    // returns the output layer
    const std::vector<vg_fe> &evaluate()
    {
        uint8_t *tr = nullptr;
        size_t len = 0;
        double ms[13];
        check(vg_prove(ctx_, &tr, &len, ms));
        tr_.assign(tr, tr + len);
        vg_free(tr);
        total_time = ms[8] * 1e-3;
        int bits = 0;
        output_.resize((size_t)1 << 26);
        check(vg_output(ctx_, output_.data(), output_.size(), &bits));
        output_.resize((size_t)1 << bits);
        parse_header();
        return output_;
    }
    // zk_prover::V_res (zk_prover.cpp:30-52)
    vg_fe V_res(const vg_fe *, const vg_fe *, const vg_fe *, int, int) { return vres_; }

    // zk_prover::sumcheck_init (zk_prover.cpp:162-200): layer_id counts down
    // from total_depth - 1 to 1, as the verifier calls it
    void sumcheck_init(int layer_id, int, int, int, const vg_fe &, const vg_fe &, const vg_fe *, const vg_fe *,
                       vg_fe *, vg_fe *)
    {
        if (layer_id != next_layer_) fail("sumcheck_init: layers must run from the top down");
        bl_ = bl_of_[layer_id - 1];
        maskpoly_sumc = fe();
        --next_layer_;
    }
    void generate_maskpoly_pre_rho(int, int) {}
    void generate_maskpoly_after_rho(int, int) {}
    void sumcheck_phase1_init() { phase_round_ = 0; }
    quadratic_poly sumcheck_phase1_update(const vg_fe &, int) { return q3(); }
    quintuple_poly sumcheck_phase1_updatelastbit(const vg_fe &, int)
    {
        quintuple_poly p = q6();
        v_u = fe();
        return p;
    }
    void sumcheck_phase2_init(const vg_fe &, const vg_fe *, const vg_fe *) {}
    quadratic_poly sumcheck_phase2_update(const vg_fe &, int) { return q3(); }
    quintuple_poly sumcheck_phase2_updatelastbit(const vg_fe &, int)
    {
        quintuple_poly p = q6();
        v_v = fe();
        return p;
    }
    std::pair<vg_fe, vg_fe> sumcheck_finalize(const vg_fe &) { return {v_u, v_v}; }
    quadratic_poly sumcheck_finalround(const vg_fe &, int, const vg_fe &)
    {
        quadratic_poly p = q3();
        query_ = fe(); rg1_ = fe(); rg2_ = fe();
        Iuv = fe(); preZu = fe(); preZv = fe();
        if (next_layer_ == 0) { Zu = fe(); preu1 = fe(); sumRc = q3(); parse_commitment(); }
        return p;
    }
    vg_fe query(const vg_fe *, const vg_fe *, const vg_fe &) { return query_; }
    vg_fe queryRg1(const vg_fe &) { return rg1_; }
    vg_fe queryRg2(const vg_fe &) { return rg2_; }
    void delete_self() { release(); }

    // values the reference verifier reads from the prover object
    vg_fe maskpoly_sumc{}, v_u{}, v_v{}, Iuv{}, preZu{}, preZv{}, Zu{}, preu1{};
    quadratic_poly sumRc{};
    double total_time = 0;                              // device seconds of the proof

    // commitment sections (poly_commit.h:42-351, vpd_verifier.cpp:46-78, fri.cpp:148-289)
    std::pair<const uint8_t *, const uint8_t *> commitment_roots() const { return {root_l_, root_h_}; }
    const vg_fe *all_sum() const { return all_sum_; }
    const uint8_t *fri_roots() const { return fri_roots_; }
    const vg_fe *final_codeword() const { return final_cw_; }
    std::pair<const uint8_t *, size_t> query_section() const { return {query_sec_, query_len_}; }
    std::pair<const uint8_t *, size_t> fft_gkr_section() const { return {fg_sec_, fg_len_}; }
    const std::vector<uint8_t> &transcript() const { return tr_; }

    // poly_commit_prover::commit_private_array (poly_commit.h:42-124): root L
    hhash_digest commit_private_array(const vg_fe *, int, const std::vector<vg_fe> &) const { return digest(root_l_); }
    // poly_commit_prover::commit_public_array (poly_commit.h:125-351): writes
    // all_sum[65] (64 slices + mask; may be null), returns root H
    hhash_digest commit_public_array(const std::vector<vg_fe> &, const vg_fe *, int, const vg_fe &, vg_fe *all_sum) const
    {
        if (all_sum) std::memcpy(all_sum, all_sum_, 65 * sizeof(vg_fe));
        return digest(root_h_);
    }
    // poly_commit_prover::commit_phase (vpd_verifier.cpp:46-78)
    ldt_commitment commit_phase(int) const
    {
        ldt_commitment c;
        for (int k = 0; k < n_ - 6; ++k) c.commitment_hhash.push_back(digest(fri_roots_ + 32 * k));
        c.final_rs_code = final_cw_;
        c.mx_depth = n_ - 6;
        return c;
    }
    // fri::request_init_value_with_merkle (fri.cpp:148-201): oracle 0 = L, 1 = H
    query_answer request_init_value_with_merkle(long long, long long, int &new_size, int oracle_indicator)
    {
        if (step_ != 0 || oracle_indicator != init_next_) fail("request_init_value_with_merkle out of the verifier's order");
        init_next_ ^= 1;
        if (init_next_ == 0) step_ = 1;
        return answer(n_ - 1, new_size);
    }
    // fri::request_step_commit (fri.cpp:227-289)
    query_answer request_step_commit(int lvl, long long, int &new_size)
    {
        if (step_ == 0 || lvl != step_ - 1) fail("request_step_commit out of the verifier's order");
        step_ = lvl + 2 > n_ - 6 ? 0 : lvl + 2;
        return answer(n_ - 2 - lvl, new_size);
    }

private:
    vg_ctx *ctx_ = nullptr;
    std::vector<uint8_t> tr_;
    std::vector<vg_fe> output_;
    std::vector<int> bl_of_;
    size_t off_ = 0;
    int next_layer_ = 0, bl_ = 0, phase_round_ = 0, n_ = 0;
    vg_fe vres_{}, query_{}, rg1_{}, rg2_{};
    const uint8_t *root_l_ = nullptr, *root_h_ = nullptr, *fri_roots_ = nullptr, *query_sec_ = nullptr,
                  *fg_sec_ = nullptr;
    const vg_fe *all_sum_ = nullptr, *final_cw_ = nullptr;
    size_t query_len_ = 0, fg_len_ = 0, qoff_ = 0;
    int init_next_ = 0, step_ = 0;                     // position in the query order

    static hhash_digest digest(const uint8_t *p)
    {
        hhash_digest d;
        std::memcpy(d.data(), p, 32);
        return d;
    }
    // the next 65 value pairs and `path` digests of the query section
    query_answer answer(int path, int &new_size)
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
        new_size = (int)need;
        return a;
    }

    static void check(int rc)
    {
        if (rc) throw std::runtime_error(std::string("virgo: ") + vg_last_error());
    }
    [[noreturn]] static void fail(const char *m) { throw std::runtime_error(std::string("virgo: ") + m); }
    void release()
    {
        if (ctx_) { vg_ctx_destroy(ctx_); ctx_ = nullptr; }
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
    // "VGT1", depth, input bits, bit lengths, V_res
    void parse_header()
    {
        off_ = 0;
        if (std::memcmp(take(4), "VGT1", 4)) fail("bad transcript header");
        int depth = (int)u32();
        n_ = (int)u32();
        bl_of_.resize(depth);
        for (int i = 0; i < depth; ++i) bl_of_[i] = (int)u32();
        vres_ = fe();
        next_layer_ = depth - 1;
    }
    // after the last layer: roots, all_sum, FRI roots, final codeword, queries, fft_gkr
    void parse_commitment()
    {
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
};

}  // namespace vg
#endif
