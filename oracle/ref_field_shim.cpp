// Test-only harness: exposes the reference's own field arithmetic
// (src/linear_gkr/prime_field.cpp, src/linear_gkr/polynomial.cpp, compiled
// from /root/reference by oracle/ref_field.mk) through a C ABI so the tests can
// compare the oracle's F_{p^2} restatement against it.  Never shipped.
#include "linear_gkr/prime_field.h"
#include "linear_gkr/polynomial.h"
using prime_field::field_element;
using prime_field::field_element_packed;
extern "C" {
struct rfe { unsigned long long img, real; };
static field_element to(const rfe &x) { field_element r; r.img = x.img; r.real = x.real; return r; }
static rfe from(const field_element &x) { rfe r; r.img = x.img; r.real = x.real; return r; }
void ref_init() { prime_field::init(); }
void ref_mul(const rfe *a, const rfe *b, rfe *o, long n) { for (long i = 0; i < n; ++i) o[i] = from(to(a[i]) * to(b[i])); }
void ref_add(const rfe *a, const rfe *b, rfe *o, long n) { for (long i = 0; i < n; ++i) o[i] = from(to(a[i]) + to(b[i])); }
void ref_sub(const rfe *a, const rfe *b, rfe *o, long n) { for (long i = 0; i < n; ++i) o[i] = from(to(a[i]) - to(b[i])); }
void ref_inv(const rfe *a, rfe *o, long n) { for (long i = 0; i < n; ++i) o[i] = from(prime_field::inv(to(a[i]))); }
void ref_root(int k, rfe *o) { *o = from(prime_field::get_root_of_unity(k)); }
// AVX2 packed multiply / add / sub (prime_field.h:178-258), four lanes at a time
void ref_packed_mul(const rfe *a, const rfe *b, rfe *o, long n)
{
    for (long i = 0; i + 4 <= n; i += 4) {
        field_element_packed x(to(a[i]), to(a[i + 1]), to(a[i + 2]), to(a[i + 3]));
        field_element_packed y(to(b[i]), to(b[i + 1]), to(b[i + 2]), to(b[i + 3]));
        field_element r[4];
        (x * y).get_field_element(r);
        for (int k = 0; k < 4; ++k) o[i + k] = from(r[k]);
    }
}
void ref_quintuple_eval(const rfe *c6, const rfe *x, rfe *o)
{
    quintuple_poly p(to(c6[0]), to(c6[1]), to(c6[2]), to(c6[3]), to(c6[4]), to(c6[5]));
    *o = from(p.eval(to(*x)));
}
}
