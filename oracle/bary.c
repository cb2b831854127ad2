/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * Barycentric out-of-domain evaluation of base columns from their LDE,
 * restating
 *   crypto/stark/src/domain.rs:17-47          DomainConstants: points x_i = lde_coset[i*B]
 *        = g*w_N^(i*B), g^n, 1/n, g^-n
 *   crypto/math/src/polynomial.rs:190-204     inv_denoms = batch_inverse(z - x_i)
 *   crypto/stark/src/trace.rs:831-1110        col_scale[i] = x_i * inv_denom_i,
 *        sum_col = sum_i LDE[i*B][col] * col_scale[i],
 *        value = (1/n * g^-n * (z^n - g^n)) * sum_col
 *   crypto/stark/src/trace.rs:1112-1132       frame evaluation points z*w^(off*step+s)
 *   crypto/math/src/field/element.rs:64-111   Montgomery batch inverse (sequential and
 *        chunk-parallel above 2^16 elements)
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

static int batch_inv_fp3_seq(uint64_t *v, size_t n) {
    if (n == 0) return 0;
    fp3 *pre = (fp3 *)malloc(n * sizeof(fp3));
    pre[0] = ld3(v);
    for (size_t i = 1; i < n; i++) pre[i] = fp3_mul(pre[i - 1], ld3(v + 3 * i));
    fp3 last = fp3_canon(pre[n - 1]);
    if (!last.c[0] && !last.c[1] && !last.c[2]) {
        free(pre);
        return -1;
    }
    fp3 bi = fp3_inv(pre[n - 1]);
    for (size_t i = n - 1; i >= 1; i--) {
        fp3 ai = fp3_mul(bi, pre[i - 1]);
        bi = fp3_mul(bi, ld3(v + 3 * i));
        st3(v + 3 * i, ai);
    }
    st3(v, bi);
    free(pre);
    return 0;
}

static int batch_inv_base_seq(uint64_t *v, size_t n) {
    if (n == 0) return 0;
    uint64_t *pre = (uint64_t *)malloc(n * sizeof(uint64_t));
    pre[0] = v[0];
    for (size_t i = 1; i < n; i++) pre[i] = gl_mul(pre[i - 1], v[i]);
    if (gl_canon(pre[n - 1]) == 0) {
        free(pre);
        return -1;
    }
    uint64_t bi = gl_inv(pre[n - 1]);
    for (size_t i = n - 1; i >= 1; i--) {
        uint64_t ai = gl_mul(bi, pre[i - 1]);
        bi = gl_mul(bi, v[i]);
        v[i] = ai;
    }
    v[0] = bi;
    free(pre);
    return 0;
}

typedef struct {
    uint64_t *v;
    int fp3;
    int rc;
} binvctx;

static void binv_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    binvctx *c = (binvctx *)vctx;
    int rc = c->fp3 ? batch_inv_fp3_seq(c->v + 3 * b, e - b) : batch_inv_base_seq(c->v + b, e - b);
    if (rc) c->rc = rc;
}

int ora_batch_inverse_base(uint64_t *buf, size_t n) { return batch_inv_base_seq(buf, n); }

int ora_batch_inverse_fp3(uint64_t *buf, size_t n) {
    if (n >= ((size_t)1 << 16) && par_threads() > 1) {
        for (size_t i = 0; i < n; i++)
            if (!gl_canon(buf[3 * i]) && !gl_canon(buf[3 * i + 1]) && !gl_canon(buf[3 * i + 2])) return -1;
        size_t chunk = (n + par_threads() - 1) / par_threads();
        binvctx c = {buf, 1, 0};
        par_for(n, chunk, binv_body, &c);
        return c.rc;
    }
    return batch_inv_fp3_seq(buf, n);
}

typedef struct {
    const uint64_t *lde;
    size_t m, blowup, n;
    const uint64_t *scale; /* n Fp3 */
    uint64_t *sums;        /* m Fp3 */
} barcctx;

static void bary_col_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    barcctx *c = (barcctx *)vctx;
    for (size_t col = b; col < e; col++) {
        fp3 acc = fp3_zero();
        for (size_t i = 0; i < c->n; i++) {
            uint64_t f = c->lde[i * c->blowup * c->m + col];
            acc = fp3_add(acc, fp3_mul_base(ld3(c->scale + 3 * i), f));
        }
        st3(c->sums + 3 * col, acc);
    }
}

void ora_barycentric_base(const uint64_t *lde, size_t n, size_t m, size_t blowup, uint64_t offset,
                          const uint64_t zz[3], uint64_t *out) {
    size_t N = n * blowup;
    uint64_t wN = gl_root_of_unity((unsigned)__builtin_ctzll(N));
    uint64_t step = gl_pow(wN, blowup);
    uint64_t *points = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t cur = offset;
    for (size_t i = 0; i < n; i++) {
        points[i] = cur;
        cur = gl_mul(cur, step);
    }
    fp3 z = ld3(zz);
    uint64_t *den = (uint64_t *)malloc(3 * n * sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) st3(den + 3 * i, fp3_neg(fp3_base_sub(points[i], z)));
    ora_batch_inverse_fp3(den, n);
    for (size_t i = 0; i < n; i++) st3(den + 3 * i, fp3_mul_base(ld3(den + 3 * i), points[i]));
    uint64_t *sums = (uint64_t *)malloc(3 * m * sizeof(uint64_t));
    barcctx c = {lde, m, blowup, n, den, sums};
    par_for(m, 1, bary_col_body, &c);
    uint64_t g_n = gl_pow(offset, n);
    uint64_t n_inv = gl_inv((uint64_t)n);
    uint64_t g_n_inv = gl_inv(g_n);
    uint64_t nig = gl_mul(n_inv, g_n_inv);
    fp3 zn = fp3_pow(z, n);
    fp3 vf = fp3_mul_base(fp3_sub_base(zn, g_n), nig);
    for (size_t col = 0; col < m; col++) st3(out + 3 * col, fp3_mul(vf, ld3(sums + 3 * col)));
    free(points);
    free(den);
    free(sums);
}

void ora_frame_points(const uint64_t zz[3], const size_t *frame_offsets, size_t nf, uint64_t trace_root,
                      size_t step_size, uint64_t *out) {
    fp3 z = ld3(zz);
    size_t k = 0;
    for (size_t f = 0; f < nf; f++) {
        fp3 cur = fp3_mul_base(z, gl_pow(trace_root, frame_offsets[f] * step_size));
        for (size_t s = 0; s < step_size; s++) {
            st3(out + 3 * k++, cur);
            cur = fp3_mul_base(cur, trace_root);
        }
    }
}
