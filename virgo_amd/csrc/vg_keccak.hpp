// SHA3-256 of exactly one 64-byte message (my_hhash.h:39-46 with USESHA3),
// written from FIPS 202: Keccak-f[1600], rate 136 bytes, domain byte 0x06,
// final bit 0x80 in byte 135.  One thread hashes one message; the 25 lanes
// live in VGPR pairs and 64-bit rotations lower to v_alignbit funnel shifts.
// The round body is generated (theta, rho+pi as b[y+5((2x+3y)%5)] =
// rot(a[x+5y]^d[x], r[x+5y]), chi, iota).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace vg {

__constant__ static const uint64_t KECCAK_RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

__device__ __forceinline__ uint64_t krol(uint64_t x, int s) { return (x << s) | (x >> (64 - s)); }

__device__ __forceinline__ void keccak_f1600(uint64_t *a)
{
#pragma unroll 1
    for (int rd = 0; rd < 24; ++rd) {
        uint64_t c0 = a[0] ^ a[5] ^ a[10] ^ a[15] ^ a[20];
        uint64_t c1 = a[1] ^ a[6] ^ a[11] ^ a[16] ^ a[21];
        uint64_t c2 = a[2] ^ a[7] ^ a[12] ^ a[17] ^ a[22];
        uint64_t c3 = a[3] ^ a[8] ^ a[13] ^ a[18] ^ a[23];
        uint64_t c4 = a[4] ^ a[9] ^ a[14] ^ a[19] ^ a[24];
        uint64_t d0 = c4 ^ krol(c1, 1), d1 = c0 ^ krol(c2, 1), d2 = c1 ^ krol(c3, 1);
        uint64_t d3 = c2 ^ krol(c4, 1), d4 = c3 ^ krol(c0, 1);
        uint64_t b[25];
        b[0] = a[0] ^ d0;
        b[16] = krol(a[5] ^ d0, 36);
        b[7] = krol(a[10] ^ d0, 3);
        b[23] = krol(a[15] ^ d0, 41);
        b[14] = krol(a[20] ^ d0, 18);
        b[10] = krol(a[1] ^ d1, 1);
        b[1] = krol(a[6] ^ d1, 44);
        b[17] = krol(a[11] ^ d1, 10);
        b[8] = krol(a[16] ^ d1, 45);
        b[24] = krol(a[21] ^ d1, 2);
        b[20] = krol(a[2] ^ d2, 62);
        b[11] = krol(a[7] ^ d2, 6);
        b[2] = krol(a[12] ^ d2, 43);
        b[18] = krol(a[17] ^ d2, 15);
        b[9] = krol(a[22] ^ d2, 61);
        b[5] = krol(a[3] ^ d3, 28);
        b[21] = krol(a[8] ^ d3, 55);
        b[12] = krol(a[13] ^ d3, 25);
        b[3] = krol(a[18] ^ d3, 21);
        b[19] = krol(a[23] ^ d3, 56);
        b[15] = krol(a[4] ^ d4, 27);
        b[6] = krol(a[9] ^ d4, 20);
        b[22] = krol(a[14] ^ d4, 39);
        b[13] = krol(a[19] ^ d4, 8);
        b[4] = krol(a[24] ^ d4, 14);
        a[0] = b[0] ^ (~b[1] & b[2]);
        a[1] = b[1] ^ (~b[2] & b[3]);
        a[2] = b[2] ^ (~b[3] & b[4]);
        a[3] = b[3] ^ (~b[4] & b[0]);
        a[4] = b[4] ^ (~b[0] & b[1]);
        a[5] = b[5] ^ (~b[6] & b[7]);
        a[6] = b[6] ^ (~b[7] & b[8]);
        a[7] = b[7] ^ (~b[8] & b[9]);
        a[8] = b[8] ^ (~b[9] & b[5]);
        a[9] = b[9] ^ (~b[5] & b[6]);
        a[10] = b[10] ^ (~b[11] & b[12]);
        a[11] = b[11] ^ (~b[12] & b[13]);
        a[12] = b[12] ^ (~b[13] & b[14]);
        a[13] = b[13] ^ (~b[14] & b[10]);
        a[14] = b[14] ^ (~b[10] & b[11]);
        a[15] = b[15] ^ (~b[16] & b[17]);
        a[16] = b[16] ^ (~b[17] & b[18]);
        a[17] = b[17] ^ (~b[18] & b[19]);
        a[18] = b[18] ^ (~b[19] & b[15]);
        a[19] = b[19] ^ (~b[15] & b[16]);
        a[20] = b[20] ^ (~b[21] & b[22]);
        a[21] = b[21] ^ (~b[22] & b[23]);
        a[22] = b[22] ^ (~b[23] & b[24]);
        a[23] = b[23] ^ (~b[24] & b[20]);
        a[24] = b[24] ^ (~b[20] & b[21]);
        a[0] ^= KECCAK_RC[rd];
    }
}

// out = SHA3-256(m[0..7] as 64 little-endian bytes)
__device__ __forceinline__ void sha3_64(const uint64_t m[8], uint64_t out[4])
{
    uint64_t a[25];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = m[i];
    a[8] = 0x06ull;
#pragma unroll
    for (int i = 9; i < 25; ++i) a[i] = 0;
    a[16] = 0x8000000000000000ull;
    keccak_f1600(a);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = a[i];
}

} // namespace vg
