/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * Keccak-256 as the reference uses it: the host prover hashes with the `sha3`
 * crate 0.10.8 (`sha3::Keccak256`, crypto/crypto/src/hash/platform_keccak.rs:
 * 124-125; Cargo.lock pins sha3 0.10.8), i.e. the original Keccak submission
 * padding (0x01 ... 0x80), rate 136 bytes, little-endian lanes, 32-byte output
 * (crypto/math-cuda/kernels/keccak.cu:114-138 states the same framing).
 * The permutation is Keccak-f[1600] as published in FIPS 202 section 3
 * (theta, rho, pi, chi, iota; 24 rounds). A `pad` argument lets the tests run
 * the same permutation with SHA3's 0x06 domain byte and compare against
 * Python's hashlib.sha3_256, which pins the permutation independently of the
 * Keccak-256 known-answer vectors.
 */
#include <stdint.h>
#include <string.h>
#include "lvm_oracle.h"

static const uint64_t RC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL,
    0x8000000080008000ULL, 0x000000000000808bULL, 0x0000000080000001ULL,
    0x8000000080008081ULL, 0x8000000000008009ULL, 0x000000000000008aULL,
    0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL,
    0x8000000000008003ULL, 0x8000000000008002ULL, 0x8000000000000080ULL,
    0x000000000000800aULL, 0x800000008000000aULL, 0x8000000080008081ULL,
    0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

/* rho offsets r[x][y] indexed as lane x + 5y (FIPS 202 Table 2) */
static const unsigned RHO[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                                 25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

static inline uint64_t rotl(uint64_t v, unsigned r) { return r ? (v << r) | (v >> (64 - r)) : v; }

void ora_keccak_f1600(uint64_t st[25]) {
    for (int round = 0; round < 24; round++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = st[x] ^ st[x + 5] ^ st[x + 10] ^ st[x + 15] ^ st[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) st[i] ^= D[i % 5];
        /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) {
                int nx = y, ny = (2 * x + 3 * y) % 5;
                B[nx + 5 * ny] = rotl(st[x + 5 * y], RHO[x + 5 * y]);
            }
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                st[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
        st[0] ^= RC[round];
    }
}

/* Streaming sponge so callers can absorb in pieces (the transcript and the
 * leaf builder both feed bytes incrementally, like `Digest::update`). */
void ora_keccak_init(ora_keccak_ctx *c) {
    memset(c, 0, sizeof(*c));
}

void ora_keccak_update(ora_keccak_ctx *c, const uint8_t *data, size_t len) {
    uint8_t *s = (uint8_t *)c->st; /* little-endian host: lane bytes in order */
    while (len > 0) {
        size_t take = 136 - c->pos;
        if (take > len) take = len;
        for (size_t i = 0; i < take; i++) s[c->pos + i] ^= data[i];
        c->pos += take;
        data += take;
        len -= take;
        if (c->pos == 136) {
            ora_keccak_f1600(c->st);
            c->pos = 0;
        }
    }
}

void ora_keccak_final_pad(ora_keccak_ctx *c, uint8_t out[32], uint8_t pad) {
    uint8_t *s = (uint8_t *)c->st;
    s[c->pos] ^= pad;
    s[135] ^= 0x80;
    ora_keccak_f1600(c->st);
    memcpy(out, s, 32);
}

void ora_keccak_final(ora_keccak_ctx *c, uint8_t out[32]) { ora_keccak_final_pad(c, out, 0x01); }

void ora_keccak256(const uint8_t *data, size_t len, uint8_t out[32]) {
    ora_keccak_ctx c;
    ora_keccak_init(&c);
    ora_keccak_update(&c, data, len);
    ora_keccak_final(&c, out);
}

void ora_keccak_pad(const uint8_t *data, size_t len, uint8_t pad, uint8_t out[32]) {
    ora_keccak_ctx c;
    ora_keccak_init(&c);
    ora_keccak_update(&c, data, len);
    ora_keccak_final_pad(&c, out, pad);
}
