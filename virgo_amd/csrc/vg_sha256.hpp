// SHA-256 of exactly one 64-byte message, written from FIPS 180-4: two
// compressions, the first over the message, the second over the constant
// padding block (0x80, zeros, bit length 512).  The second block's message
// schedule is fixed, so its K[t] + W[t] sums are one table (SHA256_KW_PAD)
// and that compression runs without message expansion.  One thread hashes
// one message.  Words are big-endian: message word j is byte-swapped from
// the little-endian 64-bit lane j/2, and the digest is swapped back.
//
// The tables follow from FIPS 180-4 section 4.2.2 (K: first 32 bits of the
// fractional parts of the cube roots of the first 64 primes) and 5.3.3 (H0:
// square roots of the first 8 primes); the pad sums add the schedule of the
// padding block to K.  SHA-256 has no counterpart in the reference prover
// (its Merkle hash is SHA3-256): this is the hash the kernel micro-set may
// build its trees with instead.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace vg {

__constant__ static const uint32_t SHA256_K[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

__constant__ static const uint32_t SHA256_KW_PAD[64] = {
    0xc28a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf374u,
    0x649b69c1u, 0xf0fe4786u, 0x0fe1edc6u, 0x240cf254u, 0x4fe9346fu, 0x6cc984beu, 0x61b9411eu, 0x16f988fau,
    0xf2c65152u, 0xa88e5a6du, 0xb019fc65u, 0xb9d99ec7u, 0x9a1231c3u, 0xe70eeaa0u, 0xfdb1232bu, 0xc7353eb0u,
    0x3069bad5u, 0xcb976d5fu, 0x5a0f118fu, 0xdc1eeefdu, 0x0a35b689u, 0xde0b7a04u, 0x58f4ca9du, 0xe15d5b16u,
    0x007f3e86u, 0x37088980u, 0xa507ea32u, 0x6fab9537u, 0x17406110u, 0x0d8cd6f1u, 0xcdaa3b6du, 0xc0bbbe37u,
    0x83613bdau, 0xdb48a363u, 0x0b02e931u, 0x6fd15ca7u, 0x521afacau, 0x31338431u, 0x6ed41a95u, 0x6d437890u,
    0xc39c91f2u, 0x9eccabbdu, 0xb5c9a0e6u, 0x532fb63cu, 0xd2c741c6u, 0x07237ea3u, 0xa4954b68u, 0x4c191d76u};

__device__ __forceinline__ uint32_t sh_rotr(uint32_t x, int n)
{
    return __builtin_amdgcn_alignbit(x, x, n);
}
__device__ __forceinline__ uint32_t sh_x3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// one round over the working variables a..h, kwt = K[t] + W[t]
#define VG_SHA256_ROUND(kwt)                                                              \
    {                                                                                     \
        uint32_t t1 = h + sh_x3(sh_rotr(e, 6), sh_rotr(e, 11), sh_rotr(e, 25)) +          \
                      ((e & f) ^ (~e & g)) + (kwt);                                       \
        uint32_t t2 = sh_x3(sh_rotr(a, 2), sh_rotr(a, 13), sh_rotr(a, 22)) +              \
                      ((a & b) | (c & (a | b)));                                          \
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;               \
    }

__device__ __forceinline__ void sha256_64(const uint64_t *m, uint64_t *out)
{
    uint32_t W[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        W[2 * j] = __builtin_bswap32((uint32_t)m[j]);
        W[2 * j + 1] = __builtin_bswap32((uint32_t)(m[j] >> 32));
    }
    const uint32_t H0[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au, 0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    uint32_t a = H0[0], b = H0[1], c = H0[2], d = H0[3], e = H0[4], f = H0[5], g = H0[6], h = H0[7];
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        if (t >= 16) {
            uint32_t w15 = W[(t - 15) & 15], w2 = W[(t - 2) & 15];
            uint32_t s0 = sh_x3(sh_rotr(w15, 7), sh_rotr(w15, 18), w15 >> 3);
            uint32_t s1 = sh_x3(sh_rotr(w2, 17), sh_rotr(w2, 19), w2 >> 10);
            W[t & 15] = s1 + W[(t - 7) & 15] + s0 + W[t & 15];
        }
        VG_SHA256_ROUND(SHA256_K[t] + W[t & 15]);
    }
    uint32_t A[8] = {a + H0[0], b + H0[1], c + H0[2], d + H0[3], e + H0[4], f + H0[5], g + H0[6], h + H0[7]};
    a = A[0]; b = A[1]; c = A[2]; d = A[3]; e = A[4]; f = A[5]; g = A[6]; h = A[7];
#pragma unroll
    for (int t = 0; t < 64; ++t) VG_SHA256_ROUND(SHA256_KW_PAD[t]);
    A[0] += a; A[1] += b; A[2] += c; A[3] += d; A[4] += e; A[5] += f; A[6] += g; A[7] += h;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        out[j] = (uint64_t)__builtin_bswap32(A[2 * j]) | ((uint64_t)__builtin_bswap32(A[2 * j + 1]) << 32);
}
#undef VG_SHA256_ROUND

}  // namespace vg
