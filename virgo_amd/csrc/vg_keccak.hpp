// SHA3-256 of exactly one 64-byte message (my_hhash.h:39-46 with USESHA3),
// written from FIPS 202: Keccak-f[1600], rate 136 bytes, domain byte 0x06,
// final bit 0x80 in byte 135.  One thread hashes one message.
//
// CDNA4 mapping: each 64-bit lane is two 32-bit VGPRs (lo, hi).  A 64-bit
// rotation is two v_alignbit_b32 funnel shifts, the 5-way XOR of theta and
// the "A ^ C[x-1] ^ rot(C[x+1])" update are v_bitop3_b32 XOR3s, and chi
// (a ^ (~b & c)) is one v_bitop3_b32 per half: about 180 VALU instructions
// per round.  The round body is generated (rho+pi as
// b[y + 5((2x+3y)%5)] = rot(a[x+5y], r[x+5y])).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace vg {

__constant__ static const uint32_t KECCAK_RC_LO[24] = {
    0x00000001u, 0x00008082u, 0x0000808Au, 0x80008000u, 0x0000808Bu, 0x80000001u, 0x80008081u, 0x00008009u,
    0x0000008Au, 0x00000088u, 0x80008009u, 0x8000000Au, 0x8000808Bu, 0x0000008Bu, 0x00008089u, 0x00008003u,
    0x00008002u, 0x00000080u, 0x0000800Au, 0x8000000Au, 0x80008081u, 0x00008080u, 0x80000001u, 0x80008008u};
__constant__ static const uint32_t KECCAK_RC_HI[24] = {
    0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u,
    0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u,
    0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u};

__device__ __forceinline__ uint32_t kx3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint32_t kal(uint32_t hi, uint32_t lo, uint32_t s)
{
    return __builtin_amdgcn_alignbit(hi, lo, s);
}

__device__ __forceinline__ void keccak_f1600(uint32_t *L, uint32_t *H)
{
#pragma unroll 24
    for (int rd = 0; rd < 24; ++rd) {
        uint32_t cl0 = kx3(kx3(L[0], L[5], L[10]), L[15], L[20]);
        uint32_t ch0 = kx3(kx3(H[0], H[5], H[10]), H[15], H[20]);
        uint32_t cl1 = kx3(kx3(L[1], L[6], L[11]), L[16], L[21]);
        uint32_t ch1 = kx3(kx3(H[1], H[6], H[11]), H[16], H[21]);
        uint32_t cl2 = kx3(kx3(L[2], L[7], L[12]), L[17], L[22]);
        uint32_t ch2 = kx3(kx3(H[2], H[7], H[12]), H[17], H[22]);
        uint32_t cl3 = kx3(kx3(L[3], L[8], L[13]), L[18], L[23]);
        uint32_t ch3 = kx3(kx3(H[3], H[8], H[13]), H[18], H[23]);
        uint32_t cl4 = kx3(kx3(L[4], L[9], L[14]), L[19], L[24]);
        uint32_t ch4 = kx3(kx3(H[4], H[9], H[14]), H[19], H[24]);
        uint32_t rl0 = kal(cl0, ch0, 31), rh0 = kal(ch0, cl0, 31);
        uint32_t rl1 = kal(cl1, ch1, 31), rh1 = kal(ch1, cl1, 31);
        uint32_t rl2 = kal(cl2, ch2, 31), rh2 = kal(ch2, cl2, 31);
        uint32_t rl3 = kal(cl3, ch3, 31), rh3 = kal(ch3, cl3, 31);
        uint32_t rl4 = kal(cl4, ch4, 31), rh4 = kal(ch4, cl4, 31);
        uint32_t bl[25], bh[25];
        { uint32_t l = kx3(L[0], cl4, rl1), h = kx3(H[0], ch4, rh1);
          bl[0] = l; bh[0] = h; }
        { uint32_t l = kx3(L[5], cl4, rl1), h = kx3(H[5], ch4, rh1);
          bh[16] = kal(l, h, 28); bl[16] = kal(h, l, 28); }
        { uint32_t l = kx3(L[10], cl4, rl1), h = kx3(H[10], ch4, rh1);
          bh[7] = kal(h, l, 29); bl[7] = kal(l, h, 29); }
        { uint32_t l = kx3(L[15], cl4, rl1), h = kx3(H[15], ch4, rh1);
          bh[23] = kal(l, h, 23); bl[23] = kal(h, l, 23); }
        { uint32_t l = kx3(L[20], cl4, rl1), h = kx3(H[20], ch4, rh1);
          bh[14] = kal(h, l, 14); bl[14] = kal(l, h, 14); }
        { uint32_t l = kx3(L[1], cl0, rl2), h = kx3(H[1], ch0, rh2);
          bh[10] = kal(h, l, 31); bl[10] = kal(l, h, 31); }
        { uint32_t l = kx3(L[6], cl0, rl2), h = kx3(H[6], ch0, rh2);
          bh[1] = kal(l, h, 20); bl[1] = kal(h, l, 20); }
        { uint32_t l = kx3(L[11], cl0, rl2), h = kx3(H[11], ch0, rh2);
          bh[17] = kal(h, l, 22); bl[17] = kal(l, h, 22); }
        { uint32_t l = kx3(L[16], cl0, rl2), h = kx3(H[16], ch0, rh2);
          bh[8] = kal(l, h, 19); bl[8] = kal(h, l, 19); }
        { uint32_t l = kx3(L[21], cl0, rl2), h = kx3(H[21], ch0, rh2);
          bh[24] = kal(h, l, 30); bl[24] = kal(l, h, 30); }
        { uint32_t l = kx3(L[2], cl1, rl3), h = kx3(H[2], ch1, rh3);
          bh[20] = kal(l, h, 2); bl[20] = kal(h, l, 2); }
        { uint32_t l = kx3(L[7], cl1, rl3), h = kx3(H[7], ch1, rh3);
          bh[11] = kal(h, l, 26); bl[11] = kal(l, h, 26); }
        { uint32_t l = kx3(L[12], cl1, rl3), h = kx3(H[12], ch1, rh3);
          bh[2] = kal(l, h, 21); bl[2] = kal(h, l, 21); }
        { uint32_t l = kx3(L[17], cl1, rl3), h = kx3(H[17], ch1, rh3);
          bh[18] = kal(h, l, 17); bl[18] = kal(l, h, 17); }
        { uint32_t l = kx3(L[22], cl1, rl3), h = kx3(H[22], ch1, rh3);
          bh[9] = kal(l, h, 3); bl[9] = kal(h, l, 3); }
        { uint32_t l = kx3(L[3], cl2, rl4), h = kx3(H[3], ch2, rh4);
          bh[5] = kal(h, l, 4); bl[5] = kal(l, h, 4); }
        { uint32_t l = kx3(L[8], cl2, rl4), h = kx3(H[8], ch2, rh4);
          bh[21] = kal(l, h, 9); bl[21] = kal(h, l, 9); }
        { uint32_t l = kx3(L[13], cl2, rl4), h = kx3(H[13], ch2, rh4);
          bh[12] = kal(h, l, 7); bl[12] = kal(l, h, 7); }
        { uint32_t l = kx3(L[18], cl2, rl4), h = kx3(H[18], ch2, rh4);
          bh[3] = kal(h, l, 11); bl[3] = kal(l, h, 11); }
        { uint32_t l = kx3(L[23], cl2, rl4), h = kx3(H[23], ch2, rh4);
          bh[19] = kal(l, h, 8); bl[19] = kal(h, l, 8); }
        { uint32_t l = kx3(L[4], cl3, rl0), h = kx3(H[4], ch3, rh0);
          bh[15] = kal(h, l, 5); bl[15] = kal(l, h, 5); }
        { uint32_t l = kx3(L[9], cl3, rl0), h = kx3(H[9], ch3, rh0);
          bh[6] = kal(h, l, 12); bl[6] = kal(l, h, 12); }
        { uint32_t l = kx3(L[14], cl3, rl0), h = kx3(H[14], ch3, rh0);
          bh[22] = kal(l, h, 25); bl[22] = kal(h, l, 25); }
        { uint32_t l = kx3(L[19], cl3, rl0), h = kx3(H[19], ch3, rh0);
          bh[13] = kal(h, l, 24); bl[13] = kal(l, h, 24); }
        { uint32_t l = kx3(L[24], cl3, rl0), h = kx3(H[24], ch3, rh0);
          bh[4] = kal(h, l, 18); bl[4] = kal(l, h, 18); }
        L[0] = bl[0] ^ (~bl[1] & bl[2]); H[0] = bh[0] ^ (~bh[1] & bh[2]);
        L[1] = bl[1] ^ (~bl[2] & bl[3]); H[1] = bh[1] ^ (~bh[2] & bh[3]);
        L[2] = bl[2] ^ (~bl[3] & bl[4]); H[2] = bh[2] ^ (~bh[3] & bh[4]);
        L[3] = bl[3] ^ (~bl[4] & bl[0]); H[3] = bh[3] ^ (~bh[4] & bh[0]);
        L[4] = bl[4] ^ (~bl[0] & bl[1]); H[4] = bh[4] ^ (~bh[0] & bh[1]);
        L[5] = bl[5] ^ (~bl[6] & bl[7]); H[5] = bh[5] ^ (~bh[6] & bh[7]);
        L[6] = bl[6] ^ (~bl[7] & bl[8]); H[6] = bh[6] ^ (~bh[7] & bh[8]);
        L[7] = bl[7] ^ (~bl[8] & bl[9]); H[7] = bh[7] ^ (~bh[8] & bh[9]);
        L[8] = bl[8] ^ (~bl[9] & bl[5]); H[8] = bh[8] ^ (~bh[9] & bh[5]);
        L[9] = bl[9] ^ (~bl[5] & bl[6]); H[9] = bh[9] ^ (~bh[5] & bh[6]);
        L[10] = bl[10] ^ (~bl[11] & bl[12]); H[10] = bh[10] ^ (~bh[11] & bh[12]);
        L[11] = bl[11] ^ (~bl[12] & bl[13]); H[11] = bh[11] ^ (~bh[12] & bh[13]);
        L[12] = bl[12] ^ (~bl[13] & bl[14]); H[12] = bh[12] ^ (~bh[13] & bh[14]);
        L[13] = bl[13] ^ (~bl[14] & bl[10]); H[13] = bh[13] ^ (~bh[14] & bh[10]);
        L[14] = bl[14] ^ (~bl[10] & bl[11]); H[14] = bh[14] ^ (~bh[10] & bh[11]);
        L[15] = bl[15] ^ (~bl[16] & bl[17]); H[15] = bh[15] ^ (~bh[16] & bh[17]);
        L[16] = bl[16] ^ (~bl[17] & bl[18]); H[16] = bh[16] ^ (~bh[17] & bh[18]);
        L[17] = bl[17] ^ (~bl[18] & bl[19]); H[17] = bh[17] ^ (~bh[18] & bh[19]);
        L[18] = bl[18] ^ (~bl[19] & bl[15]); H[18] = bh[18] ^ (~bh[19] & bh[15]);
        L[19] = bl[19] ^ (~bl[15] & bl[16]); H[19] = bh[19] ^ (~bh[15] & bh[16]);
        L[20] = bl[20] ^ (~bl[21] & bl[22]); H[20] = bh[20] ^ (~bh[21] & bh[22]);
        L[21] = bl[21] ^ (~bl[22] & bl[23]); H[21] = bh[21] ^ (~bh[22] & bh[23]);
        L[22] = bl[22] ^ (~bl[23] & bl[24]); H[22] = bh[22] ^ (~bh[23] & bh[24]);
        L[23] = bl[23] ^ (~bl[24] & bl[20]); H[23] = bh[23] ^ (~bh[24] & bh[20]);
        L[24] = bl[24] ^ (~bl[20] & bl[21]); H[24] = bh[24] ^ (~bh[20] & bh[21]);
        L[0] ^= KECCAK_RC_LO[rd];
        H[0] ^= KECCAK_RC_HI[rd];
    }
}

// out = SHA3-256(m[0..7] as 64 little-endian bytes)
__device__ __forceinline__ void sha3_64(const uint64_t m[8], uint64_t out[4])
{
    uint32_t L[25], H[25];
#pragma unroll
    for (int i = 0; i < 8; ++i) { L[i] = (uint32_t)m[i]; H[i] = (uint32_t)(m[i] >> 32); }
#pragma unroll
    for (int i = 8; i < 25; ++i) { L[i] = 0; H[i] = 0; }
    L[8] = 0x06u;
    H[16] = 0x80000000u;
    keccak_f1600(L, H);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = (uint64_t)L[i] | ((uint64_t)H[i] << 32);
}

} // namespace vg
