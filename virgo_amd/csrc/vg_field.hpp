// F_{p^2}, p = 2^61 - 1, i^2 = -1: the field of the Virgo prover
// (include/linear_gkr/prime_field.h:89-176).  Memory order {img, real}
// (prime_field.h:93) is kept because Merkle leaves hash these bytes.
//
// Every operation returns the canonical representative in [0, p), exactly as
// the reference does, so any exact evaluation order (parallel reductions,
// atomics, different FFT factorisations) produces the reference's bits.
//
// Multiplication is built for the CDNA integer VALU: operands (< 2^61) are
// split into 31- and 30-bit limbs, so the 32x32->64 partial products
// (v_mad_u64_u32, which accumulates a 64-bit addend for free) of up to two
// products sum in 64 bits without overflow, and are folded once with
// 2^61 == 1 (mod p).  An F_{p^2} product is two such two-term dot products:
// re = xr yr + (p - xi) yi, im = xr yi + xi yr (16 v_mad_u64_u32, one
// reduction per component, no 128-bit intermediate).
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

// s mod p for any 64-bit s
VG_HD uint64_t fold61(uint64_t s)
{
    uint64_t r = (s & P) + (s >> 61);
    return r >= P ? r - P : r;
}
// a*b mod p for a, b < 2^61
VG_HD uint64_t mulp(uint64_t a, uint64_t b)
{
    uint32_t a0 = (uint32_t)a & 0x7FFFFFFFu, a1 = (uint32_t)(a >> 31);
    uint32_t b0 = (uint32_t)b & 0x7FFFFFFFu, b1 = (uint32_t)(b >> 31);
    uint64_t t0 = (uint64_t)a0 * b0;                          // < 2^62
    uint64_t t1 = (uint64_t)a1 * b0 + (uint64_t)a0 * b1;      // < 2^62
    uint64_t t2 = (uint64_t)a1 * b1;                          // < 2^60
    // a*b = t0 + t1 2^31 + t2 2^62 == t0 + (t1 mod 2^30) 2^31 + (t1 >> 30) + 2 t2
    return fold61(t0 + (t2 << 1) + ((t1 & 0x3FFFFFFFull) << 31) + (t1 >> 30));
}
// a*b + c*d mod p for a, b, c, d < 2^61 (each partial sum < 2^63, total < 2^64)
VG_HD uint64_t dot2p(uint64_t a, uint64_t b, uint64_t c, uint64_t d)
{
    uint32_t a0 = (uint32_t)a & 0x7FFFFFFFu, a1 = (uint32_t)(a >> 31);
    uint32_t b0 = (uint32_t)b & 0x7FFFFFFFu, b1 = (uint32_t)(b >> 31);
    uint32_t c0 = (uint32_t)c & 0x7FFFFFFFu, c1 = (uint32_t)(c >> 31);
    uint32_t d0 = (uint32_t)d & 0x7FFFFFFFu, d1 = (uint32_t)(d >> 31);
    uint64_t t0 = (uint64_t)a0 * b0 + (uint64_t)c0 * d0;
    uint64_t t1 = (uint64_t)a1 * b0 + (uint64_t)a0 * b1 + (uint64_t)c1 * d0 + (uint64_t)c0 * d1;
    uint64_t t2 = (uint64_t)a1 * b1 + (uint64_t)c1 * d1;
    return fold61(t0 + (t2 << 1) + ((t1 & 0x3FFFFFFFull) << 31) + (t1 >> 30));
}

// a*b + c*d + e*f + g*h mod p for operands < 2^61, one reduction: t0 sums
// four 31x31-bit products (< 2^64), t1 eight 30x31-bit ones (< 2^64), t2
// four 30x30-bit ones (< 2^62); t0 is folded once first so that the final
// sum stays below 2^64
VG_HD uint64_t dot4p(uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e, uint64_t f, uint64_t g, uint64_t h)
{
    const uint64_t x[4] = {a, c, e, g}, y[4] = {b, d, f, h};
    uint64_t t0 = 0, t1 = 0, t2 = 0;
    for (int k = 0; k < 4; ++k) {
        uint32_t x0 = (uint32_t)x[k] & 0x7FFFFFFFu, x1 = (uint32_t)(x[k] >> 31);
        uint32_t y0 = (uint32_t)y[k] & 0x7FFFFFFFu, y1 = (uint32_t)(y[k] >> 31);
        t0 += (uint64_t)x0 * y0;
        t1 += (uint64_t)x1 * y0 + (uint64_t)x0 * y1;
        t2 += (uint64_t)x1 * y1;
    }
    uint64_t s0 = (t0 & P) + (t0 >> 61);
    return fold61(s0 + (t2 << 1) + ((t1 & 0x3FFFFFFFull) << 31) + (t1 >> 30));
}

VG_HD fe add(const fe &a, const fe &b) { return mk(addp(a.real, b.real), addp(a.img, b.img)); }
VG_HD fe sub(const fe &a, const fe &b) { return mk(subp(a.real, b.real), subp(a.img, b.img)); }
VG_HD fe neg(const fe &a) { return mk(negp(a.real), negp(a.img)); }
VG_HD fe dbl(const fe &a) { return add(a, a); }

// (a + b i)(c + d i) = (ac - bd) + (ad + bc) i (prime_field.h:125-151),
// as two dot products with -b = p - b (p itself is a valid limb input)
VG_HD fe mul(const fe &x, const fe &y)
{
    return mk(dot2p(x.real, y.real, P - x.img, y.img), dot2p(x.real, y.img, x.img, y.real));
}
// x y + z w with one reduction per component
VG_HD fe dot2(const fe &x, const fe &y, const fe &z, const fe &w)
{
    return mk(dot4p(x.real, y.real, P - x.img, y.img, z.real, w.real, P - z.img, w.img),
              dot4p(x.real, y.img, x.img, y.real, z.real, w.img, z.img, w.real));
}
// multiply by an element of the base field F_p (img == 0)
VG_HD fe mul_base(const fe &x, uint64_t c) { return mk(mulp(x.real, c), mulp(x.img, c)); }
// a 2^k mod p for a < p, 0 <= k < 61: since 2^61 == 1, a rotation of the 61
// bits (the result is canonical: only p's own pattern is all ones)
VG_HD uint64_t rotp(uint64_t a, int k) { return ((a << k) & P) | (a >> (61 - k)); }
VG_HD fe mul_pow2(const fe &x, int k) { return mk(rotp(x.real, k), rotp(x.img, k)); }
// x / 2 = x 2^60
VG_HD fe halve(const fe &x) { return mul_pow2(x, 60); }

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
