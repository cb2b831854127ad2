/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * Batched row-major FFT and coset LDE, restating
 *   crypto/math/src/fft/two_half_fft.rs:51-81    TwoHalfTwiddles::new (flat + bit-reversed)
 *   crypto/math/src/fft/two_half_fft.rs:86-161   DIT butterflies (first-half / second-half layers)
 *   crypto/math/src/fft/two_half_fft.rs:167-236  fft_batch_two_half: bit-reverse rows,
 *        layers [0, mid) in 2^mid-row chunks, bit-reverse, layers [mid, log n) in
 *        2^(log n - mid)-row chunks, bit-reverse (natural -> natural, no 1/n)
 *   crypto/math/src/fft/bit_reversing.rs:15-21, :43-110  row-major bit reversal
 *   crypto/math/src/polynomial.rs:527-597  coset_lde_full_expand_row_major:
 *        iFFT(n) -> row r scaled by weights[r] -> zero-pad to n*blowup -> FFT(n*blowup)
 *   crypto/stark/src/prover.rs:451-467     coset weights w_i = n^-1 * offset^i
 * Row chunks run in parallel like the reference's rayon par_chunks_mut.
 */
#include <stdlib.h>
#include <string.h>
#include "gl.h"
#include "lvm_oracle.h"
#include "par.h"

typedef struct {
    uint64_t *buf;
    size_t n, m, ew;
    unsigned log_n;
} brctx;

static void br_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    brctx *c = (brctx *)vctx;
    size_t w = c->m * c->ew;
    for (size_t i = b; i < e; i++) {
        size_t j = (size_t)bitrev(i, c->log_n);
        if (j > i) {
            uint64_t *x = c->buf + i * w, *y = c->buf + j * w;
            for (size_t k = 0; k < w; k++) {
                uint64_t t = x[k];
                x[k] = y[k];
                y[k] = t;
            }
        }
    }
}

void ora_bitrev_permute_row_major(uint64_t *buf, size_t n, size_t m, size_t elem_words) {
    if (n <= 1 || m == 0) return;
    brctx c = {buf, n, m, elem_words, (unsigned)__builtin_ctzll(n)};
    par_for(n, 1024, br_body, &c);
}

typedef struct {
    uint64_t *buf;
    size_t m;
    unsigned log_n, mid;
    const uint64_t *tw, *bitrev_tw;
} fftctx;

static inline void dit_rows(uint64_t *lo, uint64_t *hi, size_t m, uint64_t tw) {
    for (size_t k = 0; k < m; k++) {
        uint64_t t = gl_mul(tw, hi[k]);
        uint64_t a = lo[k];
        lo[k] = gl_add(a, t);
        hi[k] = gl_sub(a, t);
    }
}

/* chunk index range [b, e) of 2^mid-row chunks */
static void first_half_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    fftctx *c = (fftctx *)vctx;
    size_t rows = (size_t)1 << c->mid, m = c->m;
    for (size_t ch = b; ch < e; ch++) {
        uint64_t *chunk = c->buf + ch * rows * m;
        for (unsigned layer = 0; layer < c->mid; layer++) {
            size_t half = (size_t)1 << layer, block_rows = half * 2;
            size_t step = (size_t)1 << (c->log_n - 1 - layer);
            for (size_t blk = 0; blk < rows; blk += block_rows) {
                uint64_t *lows = chunk + blk * m, *highs = lows + half * m;
                for (size_t j = 0; j < half; j++) dit_rows(lows + j * m, highs + j * m, m, c->tw[j * step]);
            }
        }
    }
}

static void second_half_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    fftctx *c = (fftctx *)vctx;
    size_t rows = (size_t)1 << (c->log_n - c->mid), m = c->m;
    for (size_t thread = b; thread < e; thread++) {
        uint64_t *chunk = c->buf + thread * rows * m;
        for (unsigned layer = c->mid; layer < c->log_n; layer++) {
            size_t half_block = (size_t)1 << (c->log_n - 1 - layer), block_rows = half_block * 2;
            size_t first_block = thread << (layer - c->mid);
            for (size_t bi = 0; bi * block_rows < rows; bi++) {
                uint64_t tw = c->bitrev_tw[first_block + bi];
                uint64_t *lows = chunk + bi * block_rows * m;
                dit_rows(lows, lows + half_block * m, half_block * m, tw);
            }
        }
    }
}

int ora_fft_two_half(uint64_t *buf, size_t n, size_t m, int inverse) {
    if (m == 0 || n == 0) return 0;
    if (n & (n - 1)) return -1;
    unsigned log_n = (unsigned)__builtin_ctzll(n);
    if (log_n == 0) return 0;
    size_t half = n / 2;
    uint64_t omega = gl_root_of_unity(log_n);
    if (inverse) omega = gl_inv(omega);
    uint64_t *tw = (uint64_t *)malloc(half * sizeof(uint64_t));
    uint64_t *btw = (uint64_t *)malloc(half * sizeof(uint64_t));
    uint64_t cur = 1;
    for (size_t i = 0; i < half; i++) {
        tw[i] = cur;
        cur = gl_mul(cur, omega);
    }
    unsigned lh = log_n - 1;
    for (size_t i = 0; i < half; i++) btw[bitrev(i, lh)] = tw[i];
    unsigned mid = (log_n + 1) / 2;
    fftctx c = {buf, m, log_n, mid, tw, btw};
    ora_bitrev_permute_row_major(buf, n, m, 1);
    par_for(n >> mid, 1, first_half_body, &c);
    ora_bitrev_permute_row_major(buf, n, m, 1);
    par_for((size_t)1 << mid, 1, second_half_body, &c);
    ora_bitrev_permute_row_major(buf, n, m, 1);
    free(tw);
    free(btw);
    return 0;
}

void ora_lde_weights(size_t n, uint64_t offset, uint64_t *out) {
    uint64_t inv_n = gl_inv((uint64_t)n);
    uint64_t cur = inv_n;
    for (size_t i = 0; i < n; i++) {
        out[i] = cur;
        cur = gl_mul(offset, cur);
    }
}

typedef struct {
    uint64_t *buf;
    size_t m;
    const uint64_t *w;
} scalectx;

static void scale_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    scalectx *c = (scalectx *)vctx;
    for (size_t r = b; r < e; r++) {
        uint64_t w = c->w[r];
        uint64_t *row = c->buf + r * c->m;
        for (size_t k = 0; k < c->m; k++) row[k] = gl_mul(w, row[k]);
    }
}

int ora_coset_lde_row_major(const uint64_t *trace, size_t n, size_t m, size_t blowup,
                            const uint64_t *weights, uint64_t *out) {
    if (m == 0 || n == 0) return 0;
    if ((n & (n - 1)) || (blowup & (blowup - 1))) return -1;
    size_t N = n * blowup;
    if (__builtin_ctzll(N) > GL_TWO_ADICITY) return -2;
    memcpy(out, trace, n * m * sizeof(uint64_t));
    ora_fft_two_half(out, n, m, 1);
    scalectx sc = {out, m, weights};
    par_for(n, 256, scale_body, &sc);
    memset(out + n * m, 0, (N - n) * m * sizeof(uint64_t));
    return ora_fft_two_half(out, N, m, 0);
}
