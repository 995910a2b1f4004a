// Drives vg::zk_prover, vg::poly_commit_prover and vg::fri::
// (include/virgo_prover.hpp) as the reference verifier drives the reference
// prover (zk_verifier.cpp:910-1253, vpd_verifier.cpp:80-335): it draws its
// own challenges and query positions from glibc rand() (default seed 1, the
// stream it shares with the prover's mask draws) in the verifier's order,
// builds all_pub_mask and the public array itself, passes them, reads the
// members the verifier reads, and writes the prover's messages in transcript
// order.  It checks none of the sumcheck equations (the oracle's verifier
// does that on the bytes it writes) except the final-codeword check
// (vpd_verifier.cpp:315-331) and the proof-size accounting of the first
// queries.  The one value it cannot compute, the wiring predicates' general
// value, it takes from the prover.
//
// Usage: zk_prover_drive circuit.bin out.bin [mode]
// circuit.bin: int32 d, int64 n[d], then per gate int32 ty, int64 u, int64 v.
// mode "ok" (default) writes the transcript.  Any other mode plants one
// mistake, which must be refused with an exception whose message names the
// argument: challenge (a phase-1 previous_random), alpha (the next layer's
// alpha), stream (one extra rand() before the first sumcheck_init), general
// (a general_value), primask / pubmask (one mask entry), public (one
// public-array entry), target (the target sum), fri (the FRI stream out of
// step), position (a query position), order (a step query before the
// initial ones).  Exit 0: written, or refused as expected; 1: otherwise.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "virgo_prover.hpp"

namespace F = vg::field;

static void put(std::vector<uint8_t> &o, const void *p, size_t k)
{
    const uint8_t *c = (const uint8_t *)p;
    o.insert(o.end(), c, c + k);
}
static std::vector<vg_fe> gen(int k)                      // generate_randomness (zk_verifier.cpp)
{
    std::vector<vg_fe> v(k);
    for (auto &x : v) x = F::random();
    return v;
}
static std::vector<vg_fe> one_minus(const std::vector<vg_fe> &r)
{
    std::vector<vg_fe> o(r.size());
    for (size_t k = 0; k < r.size(); ++k) o[k] = F::sub(F::one(), r[k]);
    return o;
}
static vg_fe pw(vg_fe x, int e)
{
    vg_fe r = F::one();
    for (int k = 0; k < e; ++k) r = F::mul(r, x);
    return r;
}

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: %s circuit.bin out.bin [mode]\n", argv[0]); return 2; }
    std::string mode = argc > 3 ? argv[3] : "ok";
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t d;
    if (fread(&d, 4, 1, f) != 1) return 2;
    std::vector<int64_t> n(d);
    if (fread(n.data(), 8, d, f) != (size_t)d) return 2;
    int64_t tot = 0;
    for (auto x : n) tot += x;
    std::vector<int32_t> ty(tot);
    std::vector<int64_t> u(tot), v(tot);
    for (int64_t g = 0; g < tot; ++g)
        if (fread(&ty[g], 4, 1, f) != 1 || fread(&u[g], 8, 1, f) != 1 || fread(&v[g], 8, 1, f) != 1) return 2;
    fclose(f);
    vg_circuit_desc desc = {d, n.data(), ty.data(), u.data(), v.data()};
    try {
        vg::zk_prover p(desc);
        std::vector<uint8_t> out;
        int D = p.depth();
        std::vector<int> bl(D);
        int mx = 0;
        for (int i = 0; i < D; ++i) { bl[i] = p.bit_length(i); mx = bl[i] > mx ? bl[i] : mx; }
        p.init_array(mx);                                    // zk_verifier.cpp:194
        srand(1);                                            // glibc's default stream
        p.proof_init();
        vg_fe *result = p.evaluate();
        uint32_t hd = (uint32_t)D, hn = (uint32_t)bl[0];
        put(out, "VGT1", 4); put(out, &hd, 4); put(out, &hn, 4);
        for (int i = 0; i < D; ++i) { uint32_t b = (uint32_t)bl[i]; put(out, &b, 4); }
        vg_fe alpha = F::one(), beta = F::zero();
        std::vector<vg_fe> r0 = gen(bl[D - 1]), r1 = gen(bl[D - 1]);
        std::vector<vg_fe> om0 = one_minus(r0), om1 = one_minus(r1);
        vg_fe a0 = p.V_res(om0.data(), r0.data(), result, bl[D - 1], 1 << bl[D - 1]);
        put(out, &a0, 16);
        std::vector<vg_fe> pub;
        for (int i = D - 1; i >= 1; --i) {
            int ll = bl[i - 1];
            vg_fe rho = F::random();
            if (mode == "stream" && i == D - 1) (void)rand();
            p.sumcheck_init(i, bl[i], ll, ll, alpha, beta, r0.data(), r1.data(), om0.data(), om1.data());
            p.generate_maskpoly_pre_rho(ll * 2 + 1, 2);
            p.rho = rho;
            p.generate_maskpoly_after_rho(ll * 2 + 1, 2);
            put(out, &p.maskpoly_sumc, 16);
            p.sumcheck_phase1_init();
            std::vector<vg_fe> ru = gen(ll), rv = gen(ll);
            vg_fe rc = F::random();
            if (i == 1) {
                for (auto &x : rv) x = F::zero();
                rc = F::zero();
            }
            vg_fe random_combine = F::random(), linear_coeff = F::random();
            // zk_verifier.cpp:983-1025
            vg_fe pu = p.prepreu1, pv = p.preprev1, rcb = random_combine, lcf = linear_coeff;
            auto two = [&](vg_fe a, vg_fe b) { return F::add(F::mul(a, rcb), F::mul(F::mul(lcf, b), rcb)); };
            pub.push_back(two(F::one(), F::one()));
            pub.push_back(two(pu, pv));
            pub.push_back(two(F::mul(pu, pu), F::mul(pv, pv)));
            pub.push_back(two(rc, rc));
            pub.push_back(two(F::mul(rc, rc), F::mul(rc, rc)));
            pub.push_back(two(F::mul(pu, rc), F::mul(pv, rc)));
            vg_fe lc = F::random();
            for (int k = 0; k < ll; ++k) {
                pub.push_back(F::mul(F::mul(lc, F::mul(ru[k], ru[k])), rho));
                pub.push_back(F::mul(F::mul(lc, ru[k]), rho));
            }
            for (int k = 0; k < ll; ++k) {
                pub.push_back(F::mul(F::mul(lc, F::mul(rv[k], rv[k])), rho));
                pub.push_back(F::mul(F::mul(lc, rv[k]), rho));
            }
            pub.push_back(F::mul(F::mul(lc, F::mul(rc, rc)), rho));
            pub.push_back(F::mul(F::mul(lc, rc), rho));
            pub.push_back(F::mul(lc, rho));
            for (int e = 5; e >= 3; --e) pub.push_back(F::mul(F::mul(lc, pw(ru[ll - 1], e)), rho));
            for (int e = 5; e >= 3; --e) pub.push_back(F::mul(F::mul(lc, pw(rv[ll - 1], e)), rho));
            vg_fe prev = F::zero();
            for (int j = 0; j < ll; ++j) {
                vg_fe pass = prev;
                if (mode == "challenge" && i == D - 1 && j == 1) pass = F::add(prev, F::one());
                if (j == ll - 1) { vg::quintuple_poly q = p.sumcheck_phase1_updatelastbit(pass, j); put(out, &q, 96); }
                else { vg::quadratic_poly q = p.sumcheck_phase1_update(pass, j); put(out, &q, 48); }
                prev = ru[j];
            }
            std::vector<vg_fe> omu = one_minus(ru);
            p.sumcheck_phase2_init(prev, ru.data(), omu.data());
            put(out, &p.v_u, 16);
            prev = F::zero();
            for (int j = 0; j < ll; ++j) {
                if (j == ll - 1) { vg::quintuple_poly q = p.sumcheck_phase2_updatelastbit(prev, j); put(out, &q, 96); }
                else { vg::quadratic_poly q = p.sumcheck_phase2_update(prev, j); put(out, &q, 48); }
                prev = rv[j];
            }
            std::pair<vg_fe, vg_fe> fc = p.sumcheck_finalize(prev);
            put(out, &fc.second, 16);
            vg_fe gv = p.general_value(i);
            if (mode == "general" && i == D - 1) gv = F::add(gv, F::one());
            vg::quadratic_poly fr = p.sumcheck_finalround(prev, ll << 1, gv);
            put(out, &fr, 48);
            vg_fe qv = p.query(ru.data(), rv.data(), rc), g1 = p.queryRg1(rc), g2 = p.queryRg2(rc);
            put(out, &qv, 16); put(out, &g1, 16); put(out, &g2, 16);
            put(out, &p.Iuv, 16); put(out, &p.preZu, 16); put(out, &p.preZv, 16);
            alpha = F::random();
            beta = F::random();
            if (mode == "alpha" && i == D - 1) alpha = F::add(alpha, F::one());
            r0 = ru; r1 = rv;
            om0 = one_minus(r0); om1 = one_minus(r1);
        }
        // post sumcheck (zk_verifier.cpp:1200-1212)
        pub.push_back(p.Zu);
        pub.push_back(F::mul(p.Zu, p.preu1));
        p.all_pri_mask.push_back(p.maskr[0]);
        p.all_pri_mask.push_back(p.maskr[1]);
        put(out, &p.Zu, 16); put(out, &p.preu1, 16); put(out, &p.sumRc, 48);
        if (mode == "primask") p.all_pri_mask[3] = F::add(p.all_pri_mask[3], F::one());
        vg::hhash_digest rl = p.poly_prover.commit_private_array(p.circuit_value[0], bl[0], p.all_pri_mask);
        // public array eq(r_0, x) (dfs_for_public_eval, zk_verifier.cpp:838-847)
        std::vector<vg_fe> q(1, F::one());
        for (int b = 0; b < bl[0]; ++b) {
            std::vector<vg_fe> nx(q.size() * 2);
            for (size_t x = 0; x < q.size(); ++x) { nx[x] = F::mul(q[x], om0[b]); nx[x + q.size()] = F::mul(q[x], r0[b]); }
            q.swap(nx);
        }
        if (mode == "public") q[5] = F::add(q[5], F::one());
        if (mode == "pubmask") pub[3] = F::add(pub[3], F::one());
        // the verifier's running claim equals the inner product sum of all_sum
        vg_fe target = F::zero();
        for (int k = 0; k < 65; ++k) target = F::add(target, p.all_sum()[k]);
        if (mode == "target") target = F::add(target, F::one());
        std::vector<vg_fe> all_sum(65);
        vg::hhash_digest rh = p.poly_prover.commit_public_array(pub, q.data(), bl[0], target, all_sum.data());
        put(out, rl.data(), 32); put(out, rh.data(), 32);
        put(out, all_sum.data(), 65 * 16);
        if (mode == "fri") (void)rand();
        vg::ldt_commitment com = p.poly_prover.commit_phase(bl[0]);
        int nb = bl[0];
        if (com.mx_depth != nb - 6) { fprintf(stderr, "FRI level count\n"); return 1; }
        for (auto &dg : com.commitment_hhash) put(out, dg.data(), 32);
        const vg_fe *cw = vg::fri::cpd.rs_codeword[com.mx_depth - 1], *msk = vg::fri::cpd.rs_codeword_msk[com.mx_depth - 1];
        put(out, cw, 2048 * 16);
        put(out, msk, 32 * 16);
        // the final codewords are constant (vpd_verifier.cpp:315-331)
        for (int i = 0; i < 64; ++i)
            for (int j = 0; j < 16; ++j)
                if (!F::eq(cw[j << 7 | i << 1], cw[i << 1])) { fprintf(stderr, "final codeword not constant\n"); return 1; }
        for (int j = 1; j < 32; ++j)
            if (!F::eq(msk[j], msk[j - 1])) { fprintf(stderr, "final mask codeword not constant\n"); return 1; }
        auto emit = [&](const vg::query_answer &a) {
            for (auto &pr : a.first) { put(out, &pr.first, 16); put(out, &pr.second, 16); }
            for (auto &dg : a.second) put(out, dg.data(), 32);
        };
        long long proof_bytes = 0;
        for (int rep = 0; rep < 33; ++rep) {
            long long pos = 0;
            for (int i = 0; i < nb - 6; ++i) {
                int ns = 0;
                if (i == 0) {
                    do { pos = rand() % (1 << (nb - 1 - i)); } while (pos < (1 << (nb - 6 - i)) || pos % 2 == 1);
                    long long s0 = pos / 2, s1 = (pos + (1LL << (nb - 1 - i))) / 2;
                    if (mode == "position" && rep == 0) { s0 += 2; s1 += 2; }
                    if (mode == "order" && rep == 0) vg::fri::request_step_commit(0, pos / 2, ns);
                    vg::query_answer a = vg::fri::request_init_value_with_merkle(s0, s1, ns, 0);
                    if (rep == 0 && ns != 64 * 2 * 16 + (nb - 2) * 32) { fprintf(stderr, "first init size %d\n", ns); return 1; }
                    proof_bytes += ns;
                    emit(a);
                    vg::query_answer b = vg::fri::request_init_value_with_merkle(s0, s1, ns, 1);
                    proof_bytes += ns;
                    emit(b);
                } else {
                    pos %= 1LL << (nb - 1 - i);
                }
                vg::query_answer c = vg::fri::request_step_commit(i, pos / 2, ns);
                if (rep == 0 && ns != 16 + (nb - 3 - i) * 32) { fprintf(stderr, "first step size %d\n", ns); return 1; }
                proof_bytes += ns;
                emit(c);
            }
        }
        auto fg = p.fft_gkr_section();
        put(out, fg.first, fg.second);
        fprintf(stderr, "query bytes (the reference's new_size sum): %lld\n", proof_bytes);
        if (mode != "ok") { fprintf(stderr, "mode %s: the planted mistake was not refused\n", mode.c_str()); return 1; }
        FILE *o = fopen(argv[2], "wb");
        if (!o || fwrite(out.data(), 1, out.size(), o) != out.size()) return 3;
        fclose(o);
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        if (mode == "ok") return 1;
        printf("refused: %s\n", e.what());
        return 0;
    }
    return 0;
}
