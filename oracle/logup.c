/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (see lvm_oracle.h).
 *
 * LogUp auxiliary trace of one table, restated from the reference CPU prover:
 *   crypto/stark/src/lookup.rs:117-125    split_interactions (committed pairs + 1-2 absorbed)
 *   crypto/stark/src/lookup.rs:1141-1313  build_auxiliary_trace (committed columns, then the
 *                                         accumulated column after them)
 *   crypto/stark/src/lookup.rs:1660-1780  compute_logup_term_column: fingerprints, batch
 *                                         inverse, term = sum of +-m * fp^-1 per row
 *   crypto/stark/src/lookup.rs:1790-1834  build_accumulated_column_from_terms: L = sum of all
 *                                         terms, acc[0] = 0, acc[i+1] = acc[i] + rowsum[i] - L/N
 * over the flat CSR descriptor that crypto/stark/src/logup_gpu.rs:100-290
 * (build_fingerprint_descriptor) compiles a table's bus interactions into, with
 * the per-row semantics of its CPU evaluators eval_fingerprint (:497-520) and
 * eval_term (:527-550):
 *   fp_k(i)   = z - (bus_k + sum_e alpha^{aidx_e} * (const_e + sum_t coef_t * main[col_t][i]))
 *   m_k(i)    = mconst_k + sum_t mcoef_t * main[mcol_t][i]   (receiver sign folded in)
 *   term_c(i) = sum_{k in column c} m_k(i) * fp_k(i)^-1
 * Output columns 0..C-2 are committed pairs, column C-1 the absorbed
 * (virtual) one; only committed columns and the accumulated column are stored.
 */
#include <stdlib.h>
#include <string.h>

#include "gl.h"
#include "lvm_oracle.h"

static uint64_t main_at(const uint64_t *main, size_t row_stride, size_t col_stride, size_t row, size_t col) {
    return main[row * row_stride + col * col_stride];
}

static fp3 fingerprint(const ora_logup_desc *d, size_t k, const uint64_t *main, size_t rs, size_t cs, size_t row,
                       const uint64_t *alpha, const uint64_t z[3]) {
    fp3 lc = fp3_from_base(d->bus_ids[k]);
    for (uint32_t e = d->elem_offsets[k]; e < d->elem_offsets[k + 1]; e++) {
        uint64_t base = d->elem_const[e];
        for (uint32_t t = d->term_offsets[e]; t < d->term_offsets[e + 1]; t++)
            base = gl_add(base, gl_mul(d->term_coef[t], main_at(main, rs, cs, row, d->term_col[t])));
        const uint64_t *a = alpha + 3 * d->elem_alpha_idx[e];
        lc = fp3_add(lc, fp3_mul_base(fp3_make(a[0], a[1], a[2]), base));
    }
    return fp3_sub(fp3_make(z[0], z[1], z[2]), lc);
}

static uint64_t multiplicity(const ora_logup_desc *d, size_t k, const uint64_t *main, size_t rs, size_t cs,
                             size_t row) {
    uint64_t m = d->mult_const[k];
    for (uint32_t t = d->mult_term_offsets[k]; t < d->mult_term_offsets[k + 1]; t++)
        m = gl_add(m, gl_mul(d->mult_term_coef[t], main_at(main, rs, cs, row, d->mult_term_col[t])));
    return m;
}

/* Montgomery batch inverse of v[0..n) in place (crypto/math/src/field/element.rs:95-111). */
static void batch_inverse(fp3 *v, size_t n, fp3 *scratch) {
    if (n == 0) return;
    scratch[0] = v[0];
    for (size_t i = 1; i < n; i++) scratch[i] = fp3_mul(scratch[i - 1], v[i]);
    fp3 inv = fp3_inv(scratch[n - 1]);
    for (size_t i = n - 1; i > 0; i--) {
        fp3 vi = v[i];
        v[i] = fp3_mul(inv, scratch[i - 1]);
        inv = fp3_mul(inv, vi);
    }
    v[0] = inv;
}

int ora_logup_fingerprints(const ora_logup_desc *d, const uint64_t *main, size_t row_stride, size_t col_stride,
                           size_t n, const uint64_t *alpha, const uint64_t z[3], uint64_t *fp_out) {
    for (size_t k = 0; k < d->num_interactions; k++)
        for (size_t i = 0; i < n; i++) {
            fp3 f = fingerprint(d, k, main, row_stride, col_stride, i, alpha, z);
            memcpy(fp_out + 3 * (k * n + i), f.c, 24);
        }
    return 0;
}

int ora_logup_aux(const ora_logup_desc *d, const uint64_t *main, size_t row_stride, size_t col_stride, size_t n,
                  const uint64_t *alpha, const uint64_t z[3], uint64_t *terms_out, uint64_t *aux_out,
                  uint64_t L_out[3]) {
    if (n == 0 || d->num_out_cols == 0) return -1;
    const size_t K = d->num_interactions, C = d->num_out_cols;
    fp3 *recip = malloc(sizeof(fp3) * K * n);
    fp3 *scratch = malloc(sizeof(fp3) * n);
    fp3 *terms = malloc(sizeof(fp3) * C * n);
    if (!recip || !scratch || !terms) {
        free(recip);
        free(scratch);
        free(terms);
        return -2;
    }
    for (size_t k = 0; k < K; k++) {
        for (size_t i = 0; i < n; i++) recip[k * n + i] = fingerprint(d, k, main, row_stride, col_stride, i, alpha, z);
        batch_inverse(recip + k * n, n, scratch);
    }
    for (size_t c = 0; c < C; c++)
        for (size_t i = 0; i < n; i++) {
            fp3 t = fp3_zero();
            for (uint32_t ki = d->out_col_offsets[c]; ki < d->out_col_offsets[c + 1]; ki++) {
                const uint32_t k = d->out_col_interactions[ki];
                t = fp3_add(t, fp3_mul_base(recip[k * n + i], multiplicity(d, k, main, row_stride, col_stride, i)));
            }
            terms[c * n + i] = t;
        }
    /* L = sum over rows of the sum over columns, in the reference's order */
    fp3 L = fp3_zero();
    for (size_t i = 0; i < n; i++)
        for (size_t c = 0; c < C; c++) L = fp3_add(L, terms[c * n + i]);
    const fp3 offset = fp3_mul(L, fp3_from_base(gl_inv((uint64_t)n)));
    /* aux: committed columns 0..C-2 then the accumulated column, row-major */
    fp3 acc = fp3_zero();
    for (size_t i = 0; i < n; i++) {
        uint64_t *row = aux_out + 3 * C * i;
        for (size_t c = 0; c + 1 < C; c++) memcpy(row + 3 * c, terms[c * n + i].c, 24);
        memcpy(row + 3 * (C - 1), acc.c, 24);
        fp3 rs = fp3_zero();
        for (size_t c = 0; c < C; c++) rs = fp3_add(rs, terms[c * n + i]);
        acc = fp3_sub(fp3_add(acc, rs), offset);
    }
    if (terms_out)
        for (size_t c = 0; c < C; c++)
            for (size_t i = 0; i < n; i++) memcpy(terms_out + 3 * (c * n + i), terms[c * n + i].c, 24);
    memcpy(L_out, L.c, 24);
    free(recip);
    free(scratch);
    free(terms);
    return 0;
}
