// Test-only harness around the reference's own fft_gkr program
// (src/VPD/fft_circuit_GKR.cpp), compiled from /root/reference by
// oracle/ref_field.mk with its main renamed; never shipped.  It exposes, via
// a C ABI:
//  * ref_fg_run: the program's main body (fft_circuit_GKR.cpp:827-845:
//    init_array, lg draws, build_circuit, engage_gkr) on a fresh glibc
//    rand() stream (srand(1) is the default seed the program relies on),
//    returning engage_gkr's result and the global proof_size;
//  * every round polynomial the program's prover produces and every v_u /
//    v_v it computes, recorded as they are evaluated: the link wraps
//    quadratic_poly::eval and linear_poly::eval (polynomial.cpp, out of
//    line), which the program calls only in its round checks
//    (p(0) + p(1), p(r); :317-321, :440-444, :624-628, :699-703) and in
//    V_mult_add[0].eval(previous_random) (:324, :446, :630, :706);
//  * the circuit values C.circuit_val of every layer.
#include <cstdlib>
#include <vector>

#define main vg_ref_fft_gkr_program_main
#include "VPD/fft_circuit_GKR.cpp"
#undef main

using prime_field::field_element;

namespace {
struct rfe { unsigned long long img, real; };
rfe from(const field_element &x) { rfe r; r.img = x.img; r.real = x.real; return r; }
struct quad_rec { rfe a, b, c, x; };
struct lin_rec { rfe a, b, x, y; };
std::vector<quad_rec> g_quad;
std::vector<lin_rec> g_lin;
}  // namespace

extern "C" {
field_element __real__ZNK14quadratic_poly4evalERKN11prime_field13field_elementE(const quadratic_poly *, const field_element &);
field_element __real__ZNK11linear_poly4evalERKN11prime_field13field_elementE(const linear_poly *, const field_element &);

field_element __wrap__ZNK14quadratic_poly4evalERKN11prime_field13field_elementE(const quadratic_poly *p, const field_element &x)
{
    g_quad.push_back({from(p->a), from(p->b), from(p->c), from(x)});
    return __real__ZNK14quadratic_poly4evalERKN11prime_field13field_elementE(p, x);
}
field_element __wrap__ZNK11linear_poly4evalERKN11prime_field13field_elementE(const linear_poly *p, const field_element &x)
{
    field_element y = __real__ZNK11linear_poly4evalERKN11prime_field13field_elementE(p, x);
    g_lin.push_back({from(p->a), from(p->b), from(x), from(y)});
    return y;
}

// main of fft_circuit_GKR.cpp:827-845 without its argv and output file
int ref_fg_run(int lg, int *result, int *psize)
{
    for (auto *v : C.circuit_val) delete[] v;
    C.circuit_val.clear();
    C.size.clear();
    proof_size = 0;
    v_time = 0;
    p_time = 0;
    g_quad.clear();
    g_lin.clear();
    srand(1);
    prime_field::init();
    prime_field::init_random();
    init_array(lg + 6);
    field_element *r = new field_element[lg];
    for (int i = 0; i < lg; ++i) r[i] = prime_field::random();
    build_circuit(lg, r);
    *result = engage_gkr(lg) ? 1 : 0;
    *psize = proof_size;
    delete[] r;
    return (int)C.circuit_val.size();
}
// recorded quadratic_poly::eval calls: 4 values each (a, b, c, x)
long ref_fg_quad(rfe *out, long cap)
{
    long n = (long)g_quad.size();
    for (long i = 0; i < n && i < cap; ++i) {
        out[4 * i] = g_quad[i].a; out[4 * i + 1] = g_quad[i].b;
        out[4 * i + 2] = g_quad[i].c; out[4 * i + 3] = g_quad[i].x;
    }
    return n;
}
// recorded linear_poly::eval calls: 4 values each (a, b, x, result)
long ref_fg_lin(rfe *out, long cap)
{
    long n = (long)g_lin.size();
    for (long i = 0; i < n && i < cap; ++i) {
        out[4 * i] = g_lin[i].a; out[4 * i + 1] = g_lin[i].b;
        out[4 * i + 2] = g_lin[i].x; out[4 * i + 3] = g_lin[i].y;
    }
    return n;
}
// C.circuit_val[idx] (C.size[idx] values) of the last run
long ref_fg_layer(int idx, rfe *out, long cap)
{
    if (idx < 0 || idx >= (int)C.circuit_val.size()) return -1;
    long n = C.size[idx];
    for (long i = 0; i < n && i < cap; ++i) out[i] = from(C.circuit_val[idx][i]);
    return n;
}
}
