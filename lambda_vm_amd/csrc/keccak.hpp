// Keccak-f[1600] (FIPS 202 section 3: theta, rho, pi, chi, iota; 24 rounds)
// for gfx950. The 25 lanes live in VGPRs; each 64-bit lane is a pair of
// 32-bit registers, so a rotate is two v_alignbit_b32 and chi maps to
// bit-select ops. Keccak-256 framing (rate 136 bytes, pad 0x01 ... 0x80,
// little-endian lanes) is what the reference's Merkle leaves, nodes and
// transcript use (crypto/crypto/src/hash/platform_keccak.rs:124-125, sha3
// 0.10.8; crypto/math-cuda/kernels/keccak.cu:114-138 states the same framing).
#pragma once
#include "field.hpp"

namespace lvm {

__constant__ static const uint64_t KECCAK_RC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

template <int N>
__device__ __forceinline__ uint64_t rotl64(uint64_t v) {
    if constexpr (N == 0) {
        return v;
    } else if constexpr (N == 32) {
        return (v << 32) | (v >> 32);
    } else {
        return (v << N) | (v >> (64 - N));
    }
}

__device__ __forceinline__ void keccak_f1600(uint64_t s[25]) {
#pragma unroll 1
    for (int round = 0; round < 24; round++) {
        uint64_t C0 = s[0] ^ s[5] ^ s[10] ^ s[15] ^ s[20];
        uint64_t C1 = s[1] ^ s[6] ^ s[11] ^ s[16] ^ s[21];
        uint64_t C2 = s[2] ^ s[7] ^ s[12] ^ s[17] ^ s[22];
        uint64_t C3 = s[3] ^ s[8] ^ s[13] ^ s[18] ^ s[23];
        uint64_t C4 = s[4] ^ s[9] ^ s[14] ^ s[19] ^ s[24];
        uint64_t D0 = C4 ^ rotl64<1>(C1);
        uint64_t D1 = C0 ^ rotl64<1>(C2);
        uint64_t D2 = C1 ^ rotl64<1>(C3);
        uint64_t D3 = C2 ^ rotl64<1>(C4);
        uint64_t D4 = C3 ^ rotl64<1>(C0);
        // theta + rho + pi: B[y, 2x+3y] = rot(A[x,y] ^ D[x], r[x,y])
        uint64_t B0 = s[0] ^ D0;
        uint64_t B10 = rotl64<1>(s[1] ^ D1);
        uint64_t B20 = rotl64<62>(s[2] ^ D2);
        uint64_t B5 = rotl64<28>(s[3] ^ D3);
        uint64_t B15 = rotl64<27>(s[4] ^ D4);
        uint64_t B16 = rotl64<36>(s[5] ^ D0);
        uint64_t B1 = rotl64<44>(s[6] ^ D1);
        uint64_t B11 = rotl64<6>(s[7] ^ D2);
        uint64_t B21 = rotl64<55>(s[8] ^ D3);
        uint64_t B6 = rotl64<20>(s[9] ^ D4);
        uint64_t B7 = rotl64<3>(s[10] ^ D0);
        uint64_t B17 = rotl64<10>(s[11] ^ D1);
        uint64_t B2 = rotl64<43>(s[12] ^ D2);
        uint64_t B12 = rotl64<25>(s[13] ^ D3);
        uint64_t B22 = rotl64<39>(s[14] ^ D4);
        uint64_t B23 = rotl64<41>(s[15] ^ D0);
        uint64_t B8 = rotl64<45>(s[16] ^ D1);
        uint64_t B18 = rotl64<15>(s[17] ^ D2);
        uint64_t B3 = rotl64<21>(s[18] ^ D3);
        uint64_t B13 = rotl64<8>(s[19] ^ D4);
        uint64_t B14 = rotl64<18>(s[20] ^ D0);
        uint64_t B24 = rotl64<2>(s[21] ^ D1);
        uint64_t B9 = rotl64<61>(s[22] ^ D2);
        uint64_t B19 = rotl64<56>(s[23] ^ D3);
        uint64_t B4 = rotl64<14>(s[24] ^ D4);
        // chi
        s[0] = B0 ^ (~B1 & B2);
        s[1] = B1 ^ (~B2 & B3);
        s[2] = B2 ^ (~B3 & B4);
        s[3] = B3 ^ (~B4 & B0);
        s[4] = B4 ^ (~B0 & B1);
        s[5] = B5 ^ (~B6 & B7);
        s[6] = B6 ^ (~B7 & B8);
        s[7] = B7 ^ (~B8 & B9);
        s[8] = B8 ^ (~B9 & B5);
        s[9] = B9 ^ (~B5 & B6);
        s[10] = B10 ^ (~B11 & B12);
        s[11] = B11 ^ (~B12 & B13);
        s[12] = B12 ^ (~B13 & B14);
        s[13] = B13 ^ (~B14 & B10);
        s[14] = B14 ^ (~B10 & B11);
        s[15] = B15 ^ (~B16 & B17);
        s[16] = B16 ^ (~B17 & B18);
        s[17] = B17 ^ (~B18 & B19);
        s[18] = B18 ^ (~B19 & B15);
        s[19] = B19 ^ (~B15 & B16);
        s[20] = B20 ^ (~B21 & B22);
        s[21] = B21 ^ (~B22 & B23);
        s[22] = B22 ^ (~B23 & B24);
        s[23] = B23 ^ (~B24 & B20);
        s[24] = B24 ^ (~B20 & B21);
        s[0] ^= KECCAK_RC[round];
    }
}

// Keccak-256 of exactly two 32-byte nodes (one permutation): lanes 0..7 hold
// left || right, pad byte 0x01 at lane 8, 0x80 at the top of lane 16.
__device__ __forceinline__ void keccak_pair_node(const uint64_t l[4], const uint64_t r[4], uint64_t out[4]) {
    uint64_t s[25];
#pragma unroll
    for (int i = 0; i < 25; i++) s[i] = 0;
    s[0] = l[0]; s[1] = l[1]; s[2] = l[2]; s[3] = l[3];
    s[4] = r[0]; s[5] = r[1]; s[6] = r[2]; s[7] = r[3];
    s[8] = 0x01;
    s[16] = 0x8000000000000000ULL;
    keccak_f1600(s);
    out[0] = s[0]; out[1] = s[1]; out[2] = s[2]; out[3] = s[3];
}

}  // namespace lvm
