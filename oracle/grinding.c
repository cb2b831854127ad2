/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * Proof-of-work nonce, restating crypto/stark/src/grinding.rs:
 *   :6      PREFIX = 01 23 45 67 89 ab cd ed
 *   :84-93  inner = Keccak(PREFIX || seed || grinding_factor)
 *   :66-80  valid iff BE u64 of Keccak(inner || BE(nonce))[0..8] < 2^(64 - g)
 *   :45-62  search; the sequential `find` returns the smallest valid nonce,
 *           which is the deterministic choice the parity tests need.
 */
#include <string.h>
#include "lvm_oracle.h"

static const uint8_t PREFIX[8] = {0x01, 0x23, 0x45, 0x67, 0x89, 0xab, 0xcd, 0xed};

static void inner_hash(const uint8_t seed[32], uint8_t g, uint8_t out[32]) {
    uint8_t d[41];
    memcpy(d, PREFIX, 8);
    memcpy(d + 8, seed, 32);
    d[40] = g;
    ora_keccak256(d, 41, out);
}

static int valid_for_inner(const uint8_t inner[32], uint64_t nonce, uint64_t limit) {
    uint8_t d[40], h[32];
    memcpy(d, inner, 32);
    for (int i = 0; i < 8; i++) d[32 + i] = (uint8_t)(nonce >> (56 - 8 * i));
    ora_keccak256(d, 40, h);
    uint64_t head = 0;
    for (int i = 0; i < 8; i++) head = (head << 8) | h[i];
    return head < limit;
}

int ora_is_valid_nonce(const uint8_t seed[32], uint64_t nonce, uint8_t g) {
    if (g < 1 || g > 64) return 0;
    uint8_t inner[32];
    inner_hash(seed, g, inner);
    uint64_t limit = g == 64 ? 1 : ((uint64_t)1 << (64 - g));
    return valid_for_inner(inner, nonce, limit);
}

uint64_t ora_generate_nonce(const uint8_t seed[32], uint8_t g, uint64_t limit_search) {
    uint8_t inner[32];
    inner_hash(seed, g, inner);
    uint64_t limit = g == 64 ? 1 : ((uint64_t)1 << (64 - g));
    for (uint64_t n = 0; n < limit_search; n++)
        if (valid_for_inner(inner, n, limit)) return n;
    return UINT64_MAX;
}
