/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * Keccak duplex Fiat-Shamir transcript, restating
 * crypto/crypto/src/fiat_shamir/default_transcript.rs:
 *   :17      SQUEEZE_LEN = 32
 *   :75-81   sample(): finalize_reset, reverse the 32 bytes, absorb them back,
 *            invalidate the output buffer
 *   :86-95   next_sample_u64(): 8 buffered bytes as a big-endian u64
 *   :113-126 append_bytes / append_field_element (canonical BE bytes)
 *   :129-131 state(): digest of a clone
 *   :136-145 sample_u64(ub): reject candidates below (2^64 mod ub)
 * Field sampling: crypto/math/src/field/goldilocks.rs:548-555 (reject >= p),
 * extensions_goldilocks.rs:574-582 (three base samples, components 0,1,2).
 */
#include <stdlib.h>
#include <string.h>
#include "gl.h"
#include "lvm_oracle.h"

struct ora_transcript {
    ora_keccak_ctx h;
    uint8_t out_buf[32];
    size_t out_pos;
};

ora_transcript *ora_transcript_new(const uint8_t *data, size_t len) {
    ora_transcript *t = (ora_transcript *)calloc(1, sizeof(*t));
    ora_keccak_init(&t->h);
    t->out_pos = 32;
    ora_transcript_append_bytes(t, data, len);
    return t;
}

ora_transcript *ora_transcript_clone(const ora_transcript *t) {
    ora_transcript *c = (ora_transcript *)malloc(sizeof(*c));
    memcpy(c, t, sizeof(*c));
    return c;
}

void ora_transcript_free(ora_transcript *t) { free(t); }

void ora_transcript_append_bytes(ora_transcript *t, const uint8_t *data, size_t len) {
    t->out_pos = 32;
    if (len) ora_keccak_update(&t->h, data, len);
}

static void be64(uint64_t v, uint8_t *o) {
    for (int i = 0; i < 8; i++) o[i] = (uint8_t)(v >> (56 - 8 * i));
}

void ora_transcript_append_base(ora_transcript *t, uint64_t v) {
    uint8_t b[8];
    be64(gl_canon(v), b);
    ora_transcript_append_bytes(t, b, 8);
}

void ora_transcript_append_fp3(ora_transcript *t, const uint64_t v[3]) {
    uint8_t b[24];
    for (int i = 0; i < 3; i++) be64(gl_canon(v[i]), b + 8 * i);
    ora_transcript_append_bytes(t, b, 24);
}

void ora_transcript_sample(ora_transcript *t, uint8_t out[32]) {
    uint8_t d[32];
    ora_keccak_final(&t->h, d); /* finalize_reset */
    ora_keccak_init(&t->h);
    for (int i = 0; i < 32; i++) out[i] = d[31 - i];
    ora_keccak_update(&t->h, out, 32);
    t->out_pos = 32;
}

void ora_transcript_state(const ora_transcript *t, uint8_t out[32]) {
    ora_keccak_ctx c = t->h;
    ora_keccak_final(&c, out);
}

static uint64_t next_u64(ora_transcript *t) {
    if (t->out_pos + 8 > 32) {
        uint8_t s[32];
        ora_transcript_sample(t, s);
        memcpy(t->out_buf, s, 32);
        t->out_pos = 0;
    }
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v = (v << 8) | t->out_buf[t->out_pos + i];
    t->out_pos += 8;
    return v;
}

uint64_t ora_transcript_sample_base(ora_transcript *t) {
    for (;;) {
        uint64_t c = next_u64(t);
        if (c < GL_P) return c;
    }
}

void ora_transcript_sample_fp3(ora_transcript *t, uint64_t out[3]) {
    for (int i = 0; i < 3; i++) out[i] = ora_transcript_sample_base(t);
}

uint64_t ora_transcript_sample_u64(ora_transcript *t, uint64_t ub) {
    if (ub == 0) return 0;
    uint64_t threshold = (0 - ub) % ub;
    for (;;) {
        uint64_t c = next_u64(t);
        if (c >= threshold) return c % ub;
    }
}
