/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * FRI commit phase over Fp3 codewords in bit-reversed order, restating
 *   crypto/stark/src/fri/fri_functions.rs:15-29  fold: out[j] = (lo+hi) + inv_tw[j]*(zeta*(lo-hi))
 *   crypto/stark/src/fri/fri_functions.rs:37-49  inverse coset twiddles: g*w^i for
 *        i < N/2, bit-reversed over log(N)-1 bits, batch-inverted
 *   crypto/stark/src/fri/fri_functions.rs:55-61  tw'[j] = tw[2j]^2
 *   crypto/stark/src/fri/terminal.rs:45-55       FriFoldLayout
 *   crypto/stark/src/fri/terminal.rs:72-103      terminal coefficients: gather the
 *        2^k sub-coset from the bit-reversed codeword, coset iFFT
 *   crypto/math/src/polynomial.rs:332-370        interpolate_fft / interpolate_offset_fft
 *   crypto/stark/src/fri/mod.rs:27-151           commit loop: zeta <- transcript, fold,
 *        pair leaves Keccak(BE24(e[2j]) || BE24(e[2j+1])) (field_element_vector.rs:129-134),
 *        tree, root -> transcript, update twiddles; final fold; coefficients -> transcript
 */
#include <stdlib.h>
#include <string.h>
#include "gl.h"
#include "lvm_oracle.h"
#include "par.h"

static inline fp3 ld3(const uint64_t *p) { return fp3_make(p[0], p[1], p[2]); }
static inline void st3(uint64_t *p, fp3 v) {
    p[0] = v.c[0];
    p[1] = v.c[1];
    p[2] = v.c[2];
}

void ora_fri_inv_twiddles(uint64_t offset, size_t domain_size, uint64_t *out) {
    size_t half = domain_size / 2;
    if (half == 0) return;
    unsigned order = (unsigned)__builtin_ctzll(domain_size);
    uint64_t root = gl_root_of_unity(order);
    uint64_t cur = offset;
    unsigned lh = (unsigned)__builtin_ctzll(half);
    for (size_t i = 0; i < half; i++) {
        out[bitrev(i, lh)] = cur;
        cur = gl_mul(cur, root);
    }
    /* sequential batch inversion, element.rs:95-111 */
    ora_batch_inverse_base(out, half);
}

void ora_fri_fold(const uint64_t *evals, size_t len, const uint64_t zeta[3], const uint64_t *inv_tw,
                  uint64_t *out) {
    fp3 z = ld3(zeta);
    size_t half = len / 2;
    for (size_t j = 0; j < half; j++) {
        fp3 lo = ld3(evals + 6 * j), hi = ld3(evals + 6 * j + 3);
        fp3 sum = fp3_add(lo, hi);
        fp3 diff = fp3_sub(lo, hi);
        fp3 t = fp3_mul_base(fp3_mul(z, diff), inv_tw[j]);
        st3(out + 3 * j, fp3_add(sum, t));
    }
}

void ora_fri_update_twiddles(const uint64_t *tw, size_t len, uint64_t *out) {
    for (size_t j = 0; j < len / 2; j++) out[j] = gl_sqr(tw[2 * j]);
}

void ora_fri_layout(unsigned lde_log, unsigned blowup_log, unsigned k, unsigned *total_folds,
                    size_t *num_committed, size_t *terminal_len, unsigned *effective_k) {
    unsigned terminal_log = blowup_log + k;
    if (terminal_log > lde_log) terminal_log = lde_log;
    unsigned tf = lde_log - terminal_log;
    *total_folds = tf;
    *num_committed = tf > 0 ? tf - 1 : 0;
    *terminal_len = (size_t)1 << terminal_log;
    *effective_k = terminal_log - blowup_log;
}

static inline void put_be64(uint8_t *o, uint64_t v) {
    for (int i = 0; i < 8; i++) o[i] = (uint8_t)(v >> (56 - 8 * i));
}

typedef struct {
    const uint64_t *e;
    uint8_t *leaves;
} fleafctx;

static void fleaf_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    fleafctx *c = (fleafctx *)vctx;
    uint8_t buf[48];
    for (size_t j = b; j < e; j++) {
        for (int i = 0; i < 6; i++) put_be64(buf + 8 * i, gl_canon(c->e[6 * j + i]));
        ora_keccak256(buf, 48, c->leaves + 32 * j);
    }
}

void ora_fri_layer_leaves(const uint64_t *evals, size_t len, uint8_t *leaves_out) {
    fleafctx c = {evals, leaves_out};
    par_for(len / 2, 1024, fleaf_body, &c);
}

/* Size-n inverse DFT of Fp3 values (natural in, natural out), no 1/n.
 * O(n log n) radix-2 DIT with a bit-reversal first. */
static void fp3_idft(uint64_t *v, size_t n) {
    if (n <= 1) return;
    unsigned lg = (unsigned)__builtin_ctzll(n);
    for (size_t i = 0; i < n; i++) {
        size_t j = (size_t)bitrev(i, lg);
        if (j > i)
            for (int c = 0; c < 3; c++) {
                uint64_t t = v[3 * i + c];
                v[3 * i + c] = v[3 * j + c];
                v[3 * j + c] = t;
            }
    }
    for (unsigned s = 1; s <= lg; s++) {
        size_t len = (size_t)1 << s, h = len / 2;
        uint64_t wl = gl_inv(gl_root_of_unity(s));
        for (size_t b = 0; b < n; b += len) {
            uint64_t w = 1;
            for (size_t j = 0; j < h; j++) {
                fp3 a = ld3(v + 3 * (b + j)), x = ld3(v + 3 * (b + j + h));
                fp3 t = fp3_mul_base(x, w);
                st3(v + 3 * (b + j), fp3_add(a, t));
                st3(v + 3 * (b + j + h), fp3_sub(a, t));
                w = gl_mul(w, wl);
            }
        }
    }
}

void ora_fri_terminal_coeffs(const uint64_t *cw, size_t len, uint64_t terminal_offset, unsigned k,
                             uint64_t *coeffs_out) {
    size_t keep = (size_t)1 << k;
    size_t blowup = len / keep;
    unsigned ll = (unsigned)__builtin_ctzll(len);
    for (size_t m = 0; m < keep; m++) {
        size_t src = (size_t)bitrev(m * blowup, ll);
        st3(coeffs_out + 3 * m, ld3(cw + 3 * src));
    }
    fp3_idft(coeffs_out, keep);
    uint64_t inv_n = gl_inv((uint64_t)keep);
    uint64_t off_inv = gl_inv(terminal_offset);
    uint64_t pw = inv_n;
    for (size_t i = 0; i < keep; i++) {
        st3(coeffs_out + 3 * i, fp3_mul_base(ld3(coeffs_out + 3 * i), pw));
        pw = gl_mul(pw, off_inv);
    }
}

int ora_fri_commit(uint64_t *evals, size_t len, ora_transcript *t, uint64_t offset,
                   unsigned blowup_log, unsigned k, uint8_t *roots_out, uint64_t *coeffs_out,
                   uint64_t *layer_evals_out) {
    if (len < 2 || (len & (len - 1))) return -1;
    unsigned total_folds, eff_k;
    size_t num_committed, terminal_len;
    ora_fri_layout((unsigned)__builtin_ctzll(len), blowup_log, k, &total_folds, &num_committed,
                   &terminal_len, &eff_k);
    uint64_t *tw = (uint64_t *)malloc((len / 2) * sizeof(uint64_t));
    uint64_t *tmp = (uint64_t *)malloc(3 * (len / 2) * sizeof(uint64_t));
    ora_fri_inv_twiddles(offset, len, tw);
    size_t cur = len;
    uint64_t *layer_dst = layer_evals_out;
    for (size_t l = 0; l < num_committed; l++) {
        uint64_t zeta[3];
        ora_transcript_sample_fp3(t, zeta);
        ora_fri_fold(evals, cur, zeta, tw, tmp);
        cur /= 2;
        memcpy(evals, tmp, 3 * cur * sizeof(uint64_t));
        if (layer_dst) {
            memcpy(layer_dst, evals, 3 * cur * sizeof(uint64_t));
            layer_dst += 3 * cur;
        }
        size_t nl = cur / 2;
        size_t L = ora_merkle_padded_leaves(nl);
        uint8_t *leaves = (uint8_t *)malloc(nl * 32);
        uint8_t *nodes = (uint8_t *)malloc((2 * L - 1) * 32);
        ora_fri_layer_leaves(evals, cur, leaves);
        ora_merkle_build(leaves, nl, nodes);
        memcpy(roots_out + 32 * l, nodes, 32);
        ora_transcript_append_bytes(t, nodes, 32);
        free(leaves);
        free(nodes);
        ora_fri_update_twiddles(tw, cur, tw);
    }
    if (total_folds > 0) {
        uint64_t zeta[3];
        ora_transcript_sample_fp3(t, zeta);
        ora_fri_fold(evals, cur, zeta, tw, tmp);
        cur /= 2;
        memcpy(evals, tmp, 3 * cur * sizeof(uint64_t));
    }
    uint64_t toff = gl_pow(offset, (uint64_t)1 << total_folds);
    ora_fri_terminal_coeffs(evals, cur, toff, eff_k, coeffs_out);
    for (size_t i = 0; i < ((size_t)1 << eff_k); i++) ora_transcript_append_fp3(t, coeffs_out + 3 * i);
    free(tw);
    free(tmp);
    return 0;
}
