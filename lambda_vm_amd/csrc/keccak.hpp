// Keccak-f[1600] (FIPS 202 section 3: theta, rho, pi, chi, iota; 24 rounds)
// for gfx950. Each 64-bit lane is held as two 32-bit VGPRs (lo, hi):
//  - column parities and theta use v_bitop3_b32 as a 3-input XOR (LUT 0x96),
//    and the theta term D[x] is folded into the per-lane XOR (A ^ C[x-1] ^
//    rot(C[x+1], 1) is one 3-input XOR per half);
//  - a 64-bit rotate is two v_alignbit_b32 (funnel shifts of the halves);
//  - chi, a ^ (~b & c), is one v_bitop3_b32 per half (LUT 0xD2 for
//    operands (a, b, c), i.e. f(0xF0, 0xCC, 0xAA)).
// ~180 VALU instructions per round. Keccak-256 framing (rate 136 bytes, pad
// 0x01 ... 0x80, little-endian lanes) is what the reference's Merkle leaves,
// nodes and transcript use (crypto/crypto/src/hash/platform_keccak.rs:
// 124-125, sha3 0.10.8; crypto/math-cuda/kernels/keccak.cu:114-138 states the
// same framing).
#pragma once
#include "field.hpp"

namespace lvm {

__constant__ static const uint32_t KECCAK_RC_LO[24] = {
    0x00000001u, 0x00008082u, 0x0000808au, 0x80008000u, 0x0000808bu, 0x80000001u, 0x80008081u, 0x00008009u,
    0x0000008au, 0x00000088u, 0x80008009u, 0x8000000au, 0x8000808bu, 0x0000008bu, 0x00008089u, 0x00008003u,
    0x00008002u, 0x00000080u, 0x0000800au, 0x8000000au, 0x80008081u, 0x00008080u, 0x80000001u, 0x80008008u};
__constant__ static const uint32_t KECCAK_RC_HI[24] = {
    0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u,
    0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u,
    0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u};

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
// a ^ (~b & c)
__device__ __forceinline__ uint32_t chi(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);
}

// rotate the 64-bit value (hi, lo) left by N, 0 <= N < 64
template <int N>
__device__ __forceinline__ void rotl(uint32_t hi, uint32_t lo, uint32_t &oh, uint32_t &ol) {
    if constexpr (N == 0) {
        oh = hi;
        ol = lo;
    } else if constexpr (N == 32) {
        oh = lo;
        ol = hi;
    } else if constexpr (N < 32) {
        oh = __builtin_amdgcn_alignbit(hi, lo, 32 - N);
        ol = __builtin_amdgcn_alignbit(lo, hi, 32 - N);
    } else {
        oh = __builtin_amdgcn_alignbit(lo, hi, 64 - N);
        ol = __builtin_amdgcn_alignbit(hi, lo, 64 - N);
    }
}

// B[dst] = rot(A[src] ^ C[x-1] ^ rot(C[x+1], 1), R) where x = src % 5
#define LVM_KB(dst, src, R)                                                                      \
    rotl<R>(xor3(h[src], Ch[((src) + 4) % 5], Rh[((src) + 1) % 5]),                             \
            xor3(l[src], Cl[((src) + 4) % 5], Rl[((src) + 1) % 5]), bh[dst], bl[dst])

// one round (theta, rho, pi, chi, iota) on the 32-bit halves
__device__ __forceinline__ void keccak_round_32(uint32_t l[25], uint32_t h[25], int round) {
    uint32_t Cl[5], Ch[5], Rl[5], Rh[5];
#pragma unroll
    for (int x = 0; x < 5; x++) {
        Cl[x] = xor3(xor3(l[x], l[x + 5], l[x + 10]), l[x + 15], l[x + 20]);
        Ch[x] = xor3(xor3(h[x], h[x + 5], h[x + 10]), h[x + 15], h[x + 20]);
    }
#pragma unroll
    for (int x = 0; x < 5; x++) rotl<1>(Ch[x], Cl[x], Rh[x], Rl[x]);
    uint32_t bl[25], bh[25];
    // theta + rho + pi: B[y, 2x+3y] = rot(A[x,y] ^ D[x], r[x,y]), lane index x + 5y
    LVM_KB(0, 0, 0);
    LVM_KB(10, 1, 1);
    LVM_KB(20, 2, 62);
    LVM_KB(5, 3, 28);
    LVM_KB(15, 4, 27);
    LVM_KB(16, 5, 36);
    LVM_KB(1, 6, 44);
    LVM_KB(11, 7, 6);
    LVM_KB(21, 8, 55);
    LVM_KB(6, 9, 20);
    LVM_KB(7, 10, 3);
    LVM_KB(17, 11, 10);
    LVM_KB(2, 12, 43);
    LVM_KB(12, 13, 25);
    LVM_KB(22, 14, 39);
    LVM_KB(23, 15, 41);
    LVM_KB(8, 16, 45);
    LVM_KB(18, 17, 15);
    LVM_KB(3, 18, 21);
    LVM_KB(13, 19, 8);
    LVM_KB(14, 20, 18);
    LVM_KB(24, 21, 2);
    LVM_KB(9, 22, 61);
    LVM_KB(19, 23, 56);
    LVM_KB(4, 24, 14);
    // chi
#pragma unroll
    for (int y = 0; y < 25; y += 5) {
#pragma unroll
        for (int x = 0; x < 5; x++) {
            l[y + x] = chi(bl[y + x], bl[y + (x + 1) % 5], bl[y + (x + 2) % 5]);
            h[y + x] = chi(bh[y + x], bh[y + (x + 1) % 5], bh[y + (x + 2) % 5]);
        }
    }
    // iota
    l[0] ^= KECCAK_RC_LO[round];
    h[0] ^= KECCAK_RC_HI[round];
}

// The 24 rounds as a loop (throughput kernels: one copy of the round body
// keeps the code small), or fully unrolled (latency-bound tree tops: no loop
// overhead, round constants as literals).
template <bool UNROLL = false>
__device__ __forceinline__ void keccak_f1600_32(uint32_t l[25], uint32_t h[25]) {
    if constexpr (UNROLL) {
#pragma unroll
        for (int round = 0; round < 24; round++) keccak_round_32(l, h, round);
    } else {
#pragma unroll 1
        for (int round = 0; round < 24; round++) keccak_round_32(l, h, round);
    }
}
#undef LVM_KB

// 64-bit-lane interface over the 32-bit-half permutation
__device__ __forceinline__ void keccak_f1600(uint64_t s[25]) {
    uint32_t l[25], h[25];
#pragma unroll
    for (int i = 0; i < 25; i++) {
        l[i] = (uint32_t)s[i];
        h[i] = (uint32_t)(s[i] >> 32);
    }
    keccak_f1600_32(l, h);
#pragma unroll
    for (int i = 0; i < 25; i++) s[i] = ((uint64_t)h[i] << 32) | l[i];
}

// Keccak-256 of exactly two 32-byte nodes (one permutation): lanes 0..7 hold
// left || right, pad byte 0x01 at lane 8, 0x80 at the top of lane 16.
template <bool UNROLL = false>
__device__ __forceinline__ void keccak_pair_node(const uint64_t lft[4], const uint64_t rgt[4], uint64_t out[4]) {
    uint32_t l[25], h[25];
#pragma unroll
    for (int i = 0; i < 25; i++) l[i] = h[i] = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        l[i] = (uint32_t)lft[i];
        h[i] = (uint32_t)(lft[i] >> 32);
        l[4 + i] = (uint32_t)rgt[i];
        h[4 + i] = (uint32_t)(rgt[i] >> 32);
    }
    l[8] = 0x01;
    h[16] = 0x80000000u;
    keccak_f1600_32<UNROLL>(l, h);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = ((uint64_t)h[i] << 32) | l[i];
}

}  // namespace lvm
