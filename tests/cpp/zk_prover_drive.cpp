// Drives vg::zk_prover (include/virgo_prover.hpp) in the call order of the
// reference verifier (zk_verifier.cpp:910-1253; the commitment's
// poly_commit_prover and fri:: calls in the order of vpd_verifier.cpp:46-335)
// and rebuilds the transcript from the values the methods return.  Usage: zk_prover_drive circuit.bin out.bin
// circuit.bin: int32 d, int64 n[d], then per gate int32 ty, int64 u, int64 v.
#include <cstdio>
#include <cstring>
#include <vector>
#include "virgo_prover.hpp"

static void put(std::vector<uint8_t> &o, const void *p, size_t k)
{
    const uint8_t *c = (const uint8_t *)p;
    o.insert(o.end(), c, c + k);
}

int main(int argc, char **argv)
{
    if (argc != 3) { fprintf(stderr, "usage: %s circuit.bin out.bin\n", argv[0]); return 2; }
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
        p.proof_init();
        p.evaluate();
        const std::vector<uint8_t> &tr = p.transcript();
        // header as the prover reports it: "VGT1", depth, input bits, bit lengths
        uint32_t depth, nbits;
        memcpy(&depth, tr.data() + 4, 4);
        memcpy(&nbits, tr.data() + 8, 4);
        std::vector<uint32_t> bl(depth);
        memcpy(bl.data(), tr.data() + 12, 4 * depth);
        put(out, "VGT1", 4); put(out, &depth, 4); put(out, &nbits, 4); put(out, bl.data(), 4 * depth);
        vg_fe zero = {0, 0};
        std::vector<vg_fe> r(64, zero);
        vg_fe vres = p.V_res(r.data(), r.data(), nullptr, 0, 0);
        put(out, &vres, 16);
        for (int i = (int)depth - 1; i >= 1; --i) {
            p.sumcheck_init(i, bl[i], bl[i - 1], bl[i - 1], zero, zero, r.data(), r.data(), r.data(), r.data());
            p.generate_maskpoly_pre_rho(bl[i - 1] * 2 + 1, 2);
            p.generate_maskpoly_after_rho(bl[i - 1] * 2 + 1, 2);
            put(out, &p.maskpoly_sumc, 16);
            p.sumcheck_phase1_init();
            for (uint32_t j = 0; j < bl[i - 1]; ++j) {
                if (j == bl[i - 1] - 1) { vg::quintuple_poly q = p.sumcheck_phase1_updatelastbit(zero, j); put(out, &q, 96); }
                else { vg::quadratic_poly q = p.sumcheck_phase1_update(zero, j); put(out, &q, 48); }
            }
            put(out, &p.v_u, 16);
            p.sumcheck_phase2_init(zero, r.data(), r.data());
            for (uint32_t j = 0; j < bl[i - 1]; ++j) {
                if (j == bl[i - 1] - 1) { vg::quintuple_poly q = p.sumcheck_phase2_updatelastbit(zero, j); put(out, &q, 96); }
                else { vg::quadratic_poly q = p.sumcheck_phase2_update(zero, j); put(out, &q, 48); }
            }
            std::pair<vg_fe, vg_fe> fc = p.sumcheck_finalize(zero);
            put(out, &fc.second, 16);
            vg::quadratic_poly fr = p.sumcheck_finalround(zero, bl[i - 1] << 1, zero);
            put(out, &fr, 48);
            vg_fe qv = p.query(r.data(), r.data(), zero), g1 = p.queryRg1(zero), g2 = p.queryRg2(zero);
            put(out, &qv, 16); put(out, &g1, 16); put(out, &g2, 16);
            put(out, &p.Iuv, 16); put(out, &p.preZu, 16); put(out, &p.preZv, 16);
        }
        put(out, &p.Zu, 16); put(out, &p.preu1, 16); put(out, &p.sumRc, 48);
        // the commitment, as verify_poly_commitment drives poly_commit_prover
        // and fri:: (vpd_verifier.cpp:46-335)
        std::vector<vg_fe> msk;
        vg::hhash_digest rl = p.commit_private_array(nullptr, (int)nbits, msk);
        std::vector<vg_fe> all_sum(65);
        vg::hhash_digest rh = p.commit_public_array(msk, nullptr, 0, zero, all_sum.data());
        put(out, rl.data(), 32); put(out, rh.data(), 32);
        put(out, all_sum.data(), 65 * 16);
        vg::ldt_commitment com = p.commit_phase((int)nbits);
        if ((int)com.commitment_hhash.size() != (int)nbits - 6) { fprintf(stderr, "FRI root count\n"); return 1; }
        for (auto &dg : com.commitment_hhash) put(out, dg.data(), 32);
        put(out, com.final_rs_code, (2048 + 32) * 16);
        auto emit = [&](const vg::query_answer &a, int new_size) {
            size_t before = out.size();
            for (auto &pr : a.first) { put(out, &pr.first, 16); put(out, &pr.second, 16); }
            for (auto &dg : a.second) put(out, dg.data(), 32);
            return (int)(out.size() - before) == new_size;
        };
        for (int rep = 0; rep < 33; ++rep)
            for (int i = 0; i < (int)nbits - 6; ++i) {
                int ns = 0;
                bool ok = true;
                if (i == 0) {
                    vg::query_answer a = p.request_init_value_with_merkle(0, 0, ns, 0);
                    ok &= emit(a, ns);
                    vg::query_answer b = p.request_init_value_with_merkle(0, 0, ns, 1);
                    ok &= emit(b, ns);
                }
                vg::query_answer c = p.request_step_commit(i, 0, ns);
                ok &= emit(c, ns);
                if (!ok) { fprintf(stderr, "query answer size\n"); return 1; }
            }
        // the verifier asks in its own order; any other order is refused
        bool refused = false;
        try { int ns; p.request_step_commit(0, 0, ns); } catch (const std::exception &) { refused = true; }
        if (!refused) { fprintf(stderr, "out-of-order query accepted\n"); return 1; }
        auto fg = p.fft_gkr_section();
        put(out, fg.first, fg.second);
        FILE *o = fopen(argv[2], "wb");
        if (!o || fwrite(out.data(), 1, out.size(), o) != out.size()) return 3;
        fclose(o);
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
