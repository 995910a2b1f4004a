// F_{p^2}, p = 2^61 - 1, i^2 = -1: the field of the Virgo prover
// (include/linear_gkr/prime_field.h:89-176).  Memory order {img, real}
// (prime_field.h:93) is kept because Merkle leaves hash these bytes.
//
// Every operation returns the canonical representative in [0, p), exactly as
// the reference does, so any exact evaluation order (parallel reductions,
// atomics, different FFT factorisations) produces the reference's bits.
//
// Multiplication is built for the CDNA integer VALU: a 64x64 product is
// split into four 32x32->64 partial products (v_mad_u64_u32) and folded with
// 2^61 == 1 and 2^64 == 8 (mod p); there is no 128-bit intermediate.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

#define VG_HD __host__ __device__ __forceinline__

namespace vg {

constexpr uint64_t P = 0x1FFFFFFFFFFFFFFFull;

struct __align__(16) fe {
    uint64_t img, real;
};

VG_HD fe mk(uint64_t real, uint64_t img) { fe r; r.real = real; r.img = img; return r; }
VG_HD fe fzero() { return mk(0, 0); }
VG_HD fe fone() { return mk(1, 0); }
VG_HD fe from_u64(uint64_t x) { return mk(x % P, 0); }     // prime_field.h:108-112
VG_HD bool feq(const fe &a, const fe &b) { return a.real == b.real && a.img == b.img; }
VG_HD bool is_zero(const fe &a) { return (a.real | a.img) == 0; }

VG_HD uint64_t addp(uint64_t a, uint64_t b) { uint64_t s = a + b; return s >= P ? s - P : s; }
VG_HD uint64_t subp(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P - b; }
VG_HD uint64_t negp(uint64_t a) { return a ? P - a : 0; }

// a, b < 2^62 (so Karatsuba sums of canonical values are allowed).
VG_HD uint64_t mulp(uint64_t a, uint64_t b)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t t0 = (uint64_t)a0 * b0;                          // < 2^64
    uint64_t t1 = (uint64_t)a1 * b0 + (uint64_t)a0 * b1;      // < 2^63
    uint64_t t2 = (uint64_t)a1 * b1;                          // < 2^60
    // a*b = t0 + t1*2^32 + t2*2^64
    //    == (t0 mod 2^61) + (t0 >> 61) + (t1 mod 2^29)*2^32 + (t1 >> 29) + 8*t2
    uint64_t s = (t0 & P) + (t0 >> 61) + ((t1 & 0x1FFFFFFFull) << 32) + (t1 >> 29) + (t2 << 3);
    s = (s & P) + (s >> 61);
    return s >= P ? s - P : s;
}

VG_HD fe add(const fe &a, const fe &b) { return mk(addp(a.real, b.real), addp(a.img, b.img)); }
VG_HD fe sub(const fe &a, const fe &b) { return mk(subp(a.real, b.real), subp(a.img, b.img)); }
VG_HD fe neg(const fe &a) { return mk(negp(a.real), negp(a.img)); }
VG_HD fe dbl(const fe &a) { return add(a, a); }

// (a + b i)(c + d i) with three base products (prime_field.h:125-151)
VG_HD fe mul(const fe &x, const fe &y)
{
    uint64_t ac = mulp(x.real, y.real);
    uint64_t bd = mulp(x.img, y.img);
    uint64_t t = mulp(x.real + x.img, y.real + y.img);
    return mk(subp(ac, bd), subp(t, addp(ac, bd)));
}
// multiply by an element of the base field F_p (img == 0)
VG_HD fe mul_base(const fe &x, uint64_t c) { return mk(mulp(x.real, c), mulp(x.img, c)); }

inline fe fpow(fe x, unsigned __int128 e)
{
    fe r = fone(), t = x;
    while (e) { if (e & 1) r = mul(r, t); t = mul(t, t); e >>= 1; }
    return r;
}
inline fe finv(const fe &x)                    // x^(p^2-2), prime_field.cpp:88-92
{
    unsigned __int128 p = P;
    return fpow(x, p * p - 2);
}
// generator of the 2-adic subgroup: order exactly 2^62 (prime_field.cpp:60-73)
inline fe root_of_unity(int log_order)
{
    fe r = mk(2147483648ull, 1033321771269002680ull);
    for (int i = 0; i < 62 - log_order; ++i) r = mul(r, r);
    return r;
}

struct lpoly { fe a, b; };                      // a*x + b   (polynomial.h:36-46)

} // namespace vg
