/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * DEEP composition codeword, restating crypto/stark/src/prover.rs:2400-2604
 * (`compute_deep_composition_poly_evaluations`) with the inverse
 * denominators of crypto/stark/src/r4_denoms.rs:25-45 (1/(x_i - pole),
 * poles [z^P, z_0, ..., z_{K-1}]):
 *   result_i = sum_j gamma_h[j] * (H_j(x_i) - H_j(z^P)) * inv_h[i]
 *            + sum_k inv_t[k][i] * (sum_j t_j(x_i) * gamma_tr[j][k] - ood_compressed[k])
 *   ood_compressed[k] = sum_j gamma_tr[j][k] * t_j(z_k)
 * over the natural-order LDE coset x_i = g * w_N^i. Inputs are row-major
 * natural-order tables (main base words, aux and parts Fp3).
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

typedef struct {
    size_t N, m, ma, np, K;
    const uint64_t *main, *aux, *parts, *gtr, *gh, *parts_ood, *inv;
    const fp3 *oodc;
    uint64_t *out;
} deepctx;

static void deep_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    deepctx *c = (deepctx *)vctx;
    for (size_t i = b; i < e; i++) {
        fp3 r = fp3_zero();
        for (size_t j = 0; j < c->np; j++) {
            fp3 hv = ld3(c->parts + 3 * (i * c->np + j));
            fp3 t = fp3_mul(fp3_mul(ld3(c->gh + 3 * j), fp3_sub(hv, ld3(c->parts_ood + 3 * j))), ld3(c->inv + 3 * i));
            r = fp3_add(r, t);
        }
        size_t off = c->np ? 1 : 0;
        for (size_t k = 0; k < c->K; k++) {
            fp3 cs = fp3_zero();
            for (size_t j = 0; j < c->m; j++)
                cs = fp3_add(cs, fp3_mul_base(ld3(c->gtr + 3 * (j * c->K + k)), c->main[i * c->m + j]));
            for (size_t j = 0; j < c->ma; j++)
                cs = fp3_add(cs, fp3_mul(ld3(c->aux + 3 * (i * c->ma + j)), ld3(c->gtr + 3 * ((c->m + j) * c->K + k))));
            r = fp3_add(r, fp3_mul(ld3(c->inv + 3 * ((off + k) * c->N + i)), fp3_sub(cs, c->oodc[k])));
        }
        st3(c->out + 3 * i, r);
    }
}

int ora_deep_composition(const uint64_t *main, size_t m, const uint64_t *aux, size_t ma, const uint64_t *parts,
                         size_t np, size_t N, uint64_t offset, const uint64_t *z_points, size_t K,
                         const uint64_t *z_parts, const uint64_t *trace_ood, const uint64_t *parts_ood,
                         const uint64_t *gammas_tr, const uint64_t *gammas_h, uint64_t *out) {
    size_t poles = K + (np ? 1 : 0);
    uint64_t *inv = (uint64_t *)malloc(3 * poles * N * sizeof(uint64_t));
    uint64_t wN = gl_root_of_unity((unsigned)__builtin_ctzll(N));
    uint64_t x = offset;
    for (size_t i = 0; i < N; i++) {
        size_t q = 0;
        if (np) st3(inv + 3 * (q++ * N + i), fp3_base_sub(x, ld3(z_parts)));
        for (size_t k = 0; k < K; k++) st3(inv + 3 * ((q + k) * N + i), fp3_base_sub(x, ld3(z_points + 3 * k)));
        x = gl_mul(x, wN);
    }
    int rc = ora_batch_inverse_fp3(inv, poles * N);
    if (rc) {
        free(inv);
        return rc;
    }
    fp3 *oodc = (fp3 *)calloc(K ? K : 1, sizeof(fp3));
    for (size_t j = 0; j < m + ma; j++)
        for (size_t k = 0; k < K; k++)
            oodc[k] = fp3_add(oodc[k], fp3_mul(ld3(gammas_tr + 3 * (j * K + k)), ld3(trace_ood + 3 * (j * K + k))));
    deepctx c = {N, m, ma, np, K, main, aux, parts, gammas_tr, gammas_h, parts_ood, inv, oodc, out};
    par_for(N, 1024, deep_body, &c);
    free(oodc);
    free(inv);
    return 0;
}
