/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (see lvm_oracle.h).
 *
 * Constraint evaluation and the composition polynomial on the LDE, restated
 * from the reference CPU prover:
 *   crypto/stark/src/constraint_ir/device.rs:60-125, :330-470  flat program
 *     wire format and eval_device_program (the per-row forward walk: slots in
 *     two classes, uniform operands read from the tables, roots read last)
 *   crypto/stark/src/constraints/evaluator.rs:36-140  evaluate_transitions:
 *     H(i) = sum_c zinv_c(i) * C_c(i) * beta_c  (+ the boundary part)
 *   crypto/stark/src/constraints/evaluator.rs:244-270  boundary part:
 *     sum_b zinv_b(i) * beta_b * (trace_b(i) - value_b), zinv_b = 1/(x_i - w^step_b)
 *   crypto/stark/src/constraints/zerofier.rs:8-110  transition zerofier inverse:
 *     prod_{j=1..e}(x - w^(n-j)) / (x^n - 1), e = the constraint's end exemptions
 * x_i = offset * w_N^i is LDE point i (natural order); a frame offset k reads
 * LDE row (i + k*next_step) mod N. Tables here are natural-order row-major.
 */
#include <stdlib.h>
#include <string.h>

#include "gl.h"
#include "lvm_oracle.h"
#include "par.h"

enum { OP_CONST_BASE = 0, OP_CONST_EXT, OP_VAR, OP_RAP, OP_ALPHA, OP_OFFSET, OP_ADD, OP_SUB, OP_MUL, OP_NEG, OP_EMBED };
enum { K_BASE_SLOT = 0, K_EXT_SLOT, K_BASE_CONST, K_EXT_CONST, K_RAP, K_ALPHA, K_OFFSET };

typedef struct {
    int ext;
    uint64_t b;
    fp3 e;
} val;

static fp3 ld3(const uint64_t *p) { return fp3_make(p[0], p[1], p[2]); }

static val operand(const ora_program *p, uint32_t enc, const uint64_t *vb, const fp3 *ve) {
    const uint32_t kind = enc >> 29, x = enc & 0x1FFFFFFFu;
    val v = {0, 0, {{0, 0, 0}}};
    switch (kind) {
    case K_BASE_SLOT: v.b = vb[x]; break;
    case K_EXT_SLOT: v.ext = 1; v.e = ve[x]; break;
    case K_BASE_CONST: v.b = p->base_consts[x]; break;
    case K_EXT_CONST: v.ext = 1; v.e = ld3(p->ext_consts + 3 * x); break;
    case K_RAP: v.ext = 1; v.e = ld3(p->rap + 3 * x); break;
    case K_ALPHA: v.ext = 1; v.e = ld3(p->alpha + 3 * x); break;
    default: v.ext = 1; v.e = ld3(p->table_offset); break;
    }
    return v;
}

static fp3 as_ext(val v) { return v.ext ? v.e : fp3_from_base(v.b); }

/* Node computing each constraint: given (root_nodes), or for pinned root
 * slots the last node writing the slot (device.rs pins roots). */
static int root_nodes_of(const ora_program *p, uint32_t *out) {
    for (size_t c = 0; c < p->num_roots; c++) {
        if (p->root_nodes) {
            out[c] = p->root_nodes[c];
            continue;
        }
        size_t k = p->num_nodes;
        while (k > 0 && (uint32_t)(p->nodes[2 * (k - 1) + 1] >> 32) != (uint32_t)p->roots[c]) k--;
        if (k == 0) return -1;
        out[c] = (uint32_t)(k - 1);
    }
    return 0;
}

/* Constraints computed by each node, as chains: head[k] is the first
 * constraint whose value node k computes (or -1), nxt[c] the next one. */
static void root_chains(const ora_program *p, const uint32_t *root_node, int32_t *head, int32_t *nxt) {
    for (size_t k = 0; k < p->num_nodes; k++) head[k] = -1;
    for (size_t c = p->num_roots; c-- > 0;) {
        nxt[c] = head[root_node[c]];
        head[root_node[c]] = (int32_t)c;
    }
}

/* one row's walk; slots vb[num_base_slots], ve[num_ext_slots]; the value of
 * constraint c is captured into rv[c] when node root_node[c] has run (the
 * chains of root_chains list them per node) */
static void walk(const ora_program *p, const uint64_t *main, size_t m, const uint64_t *aux, size_t ma, size_t N,
                 size_t next_step, size_t i, uint64_t *vb, fp3 *ve, const int32_t *head, const int32_t *nxt,
                 fp3 *rv) {
    for (size_t k = 0; k < p->num_nodes; k++) {
        const uint64_t w0 = p->nodes[2 * k], w1 = p->nodes[2 * k + 1];
        const uint32_t op = (uint32_t)w0, a = (uint32_t)(w0 >> 32), b = (uint32_t)w1, res = (uint32_t)(w1 >> 32);
        const int rext = (res >> 31) & 1;
        const uint32_t slot = res & 0x7FFFFFFFu;
        val r = {0, 0, {{0, 0, 0}}};
        switch (op) {
        case OP_CONST_BASE: r.b = p->base_consts[a]; break;
        case OP_CONST_EXT: r.ext = 1; r.e = ld3(p->ext_consts + 3 * a); break;
        case OP_VAR: {
            const uint32_t col = a & 0xFFFF, is_main = (b >> 16) & 1, off = (b >> 8) & 0xFF;
            const size_t row = (i + (size_t)off * next_step) % N;
            if (is_main) {
                r.b = main[row * m + col];
            } else {
                r.ext = 1;
                r.e = ld3(aux + 3 * (row * ma + col));
            }
            break;
        }
        case OP_RAP: r.ext = 1; r.e = ld3(p->rap + 3 * a); break;
        case OP_ALPHA: r.ext = 1; r.e = ld3(p->alpha + 3 * a); break;
        case OP_OFFSET: r.ext = 1; r.e = ld3(p->table_offset); break;
        case OP_ADD: case OP_SUB: case OP_MUL: {
            const val x = operand(p, a, vb, ve), y = operand(p, b, vb, ve);
            if (!x.ext && !y.ext) {
                r.b = op == OP_ADD ? gl_add(x.b, y.b) : op == OP_SUB ? gl_sub(x.b, y.b) : gl_mul(x.b, y.b);
            } else {
                r.ext = 1;
                if (op == OP_MUL && !x.ext) r.e = fp3_mul_base(y.e, x.b);
                else if (op == OP_MUL && !y.ext) r.e = fp3_mul_base(x.e, y.b);
                else {
                    const fp3 xe = as_ext(x), ye = as_ext(y);
                    r.e = op == OP_ADD ? fp3_add(xe, ye) : op == OP_SUB ? fp3_sub(xe, ye) : fp3_mul(xe, ye);
                }
            }
            break;
        }
        case OP_NEG: {
            const val x = operand(p, a, vb, ve);
            if (x.ext) { r.ext = 1; r.e = fp3_neg(x.e); } else r.b = gl_neg(x.b);
            break;
        }
        default: { /* OP_EMBED */
            const val x = operand(p, a, vb, ve);
            r.ext = 1;
            r.e = as_ext(x);
            break;
        }
        }
        if (rext) ve[slot] = as_ext(r);
        else vb[slot] = r.b;
        for (int32_t c = head[k]; c >= 0; c = nxt[c]) rv[c] = as_ext(r);
    }
}

int ora_eval_constraints(const ora_program *p, const uint64_t *main, size_t m, const uint64_t *aux, size_t ma,
                         size_t N, size_t next_step, uint64_t *evals_out) {
    uint64_t *vb = calloc(p->num_base_slots + 1, 8);
    fp3 *ve = calloc(p->num_ext_slots + 1, sizeof(fp3));
    uint32_t *rn = calloc(p->num_roots + 1, 4);
    fp3 *rv = calloc(p->num_roots + 1, sizeof(fp3));
    int32_t *head = calloc(p->num_nodes + 1, 4), *nxt = calloc(p->num_roots + 1, 4);
    int rc = 0;
    if (!vb || !ve || !rn || !rv || !head || !nxt || root_nodes_of(p, rn)) {
        rc = -1;
    } else {
        root_chains(p, rn, head, nxt);
        for (size_t i = 0; i < N; i++) {
            walk(p, main, m, aux, ma, N, next_step, i, vb, ve, head, nxt, rv);
            for (size_t c = 0; c < p->num_roots; c++) memcpy(evals_out + 3 * (c * N + i), rv[c].c, 24);
        }
    }
    free(vb);
    free(ve);
    free(rn);
    free(rv);
    free(head);
    free(nxt);
    return rc;
}

typedef struct {
    const ora_program *p;
    const uint64_t *main, *aux;
    size_t m, ma, N, B, next_step;
    const ora_composition *cp;
    const int32_t *head, *nxt;
    const uint64_t *zcyc;
    uint64_t wN, w_last, wn;
    uint64_t *h_out;
    int failed;
} compctx;

/* rows [b, e): the walk and the composition sum per row, rayon-style chunks
 * (crypto/stark/src/constraints/evaluator.rs:117-137 per-thread buffers) */
static void comp_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    compctx *c = (compctx *)vctx;
    const ora_program *p = c->p;
    const ora_composition *cp = c->cp;
    uint64_t *vb = calloc(p->num_base_slots + 1, 8);
    fp3 *ve = calloc(p->num_ext_slots + 1, sizeof(fp3));
    fp3 *rv = calloc(p->num_roots + 1, sizeof(fp3));
    if (!vb || !ve || !rv) {
        c->failed = 1;
        free(vb);
        free(ve);
        free(rv);
        return;
    }
    uint64_t zg[64];
    uint64_t x = gl_mul(cp->coset_offset, gl_pow(c->wN, b));
    for (size_t i = b; i < e; i++) {
        walk(p, c->main, c->m, c->aux, c->ma, c->N, c->next_step, i, vb, ve, c->head, c->nxt, rv);
        /* each zerofier group's inverse zerofier at this row, once */
        for (size_t g = 0; g < cp->num_groups && g < 64; g++) {
            uint64_t z = c->zcyc[i % c->B];
            uint64_t r = c->w_last;
            for (uint32_t j = 0; j < cp->group_exemptions[g]; j++) {
                z = gl_mul(z, gl_sub(x, r));
                r = gl_mul(r, c->w_last);
            }
            zg[g] = z;
        }
        fp3 h = fp3_zero();
        for (size_t k = 0; k < p->num_roots; k++) {
            const fp3 beta = ld3(cp->beta + 3 * k);
            h = fp3_add(h, fp3_mul(fp3_mul_base(rv[k], zg[cp->root_group[k]]), beta));
        }
        uint64_t last_step = 0, zb = 0;
        int have = 0;
        for (size_t bi = 0; bi < cp->num_boundary; bi++) {
            /* 1/(x - w^step); consecutive constraints on one step share it */
            if (!have || cp->b_step[bi] != last_step) {
                last_step = cp->b_step[bi];
                zb = gl_inv(gl_sub(x, gl_pow(c->wn, last_step)));
                have = 1;
            }
            fp3 t;
            if (cp->b_is_aux[bi])
                t = ld3(c->aux + 3 * (i * c->ma + cp->b_col[bi]));
            else
                t = fp3_from_base(c->main[i * c->m + cp->b_col[bi]]);
            const fp3 bp = fp3_sub(t, ld3(cp->b_value + 3 * bi));
            h = fp3_add(h, fp3_mul(fp3_mul_base(ld3(cp->b_beta + 3 * bi), zb), bp));
        }
        memcpy(c->h_out + 3 * i, h.c, 24);
        x = gl_mul(x, c->wN);
    }
    free(vb);
    free(ve);
    free(rv);
}

int ora_eval_composition(const ora_program *p, const uint64_t *main, size_t m, const uint64_t *aux, size_t ma,
                         size_t N, size_t next_step, const ora_composition *cp, uint64_t *h_out) {
    const size_t n = cp->trace_rows;
    if (n == 0 || N % n) return -2;
    const size_t B = N / n;
    unsigned log_n = 0, log_N = 0;
    while (((size_t)1 << log_n) < n) log_n++;
    while (((size_t)1 << log_N) < N) log_N++;
    const uint64_t wn = gl_root_of_unity(log_n), wN = gl_root_of_unity(log_N);
    /* 1/(x^n - 1) repeats with period B: x_i^n = offset^n * (w_N^n)^i */
    uint64_t *zcyc = malloc(8 * B);
    uint32_t *rn = calloc(p->num_roots + 1, 4);
    int32_t *head = calloc(p->num_nodes + 1, 4), *nxt = calloc(p->num_roots + 1, 4);
    if (!zcyc || !rn || !head || !nxt || root_nodes_of(p, rn)) {
        free(zcyc);
        free(rn);
        free(head);
        free(nxt);
        return -1;
    }
    root_chains(p, rn, head, nxt);
    {
        const uint64_t step = gl_pow(wN, n);
        uint64_t d = gl_pow(cp->coset_offset, n);
        for (size_t j = 0; j < B; j++) {
            zcyc[j] = gl_inv(gl_sub(d, 1));
            d = gl_mul(d, step);
        }
    }
    int bad = cp->num_groups > 64;
    for (size_t k = 0; k < p->num_roots; k++)
        if (cp->root_group[k] >= cp->num_groups) bad = 1;
    compctx c = {p, main, aux, m, ma, N, B, next_step, cp, head, nxt, zcyc, wN, gl_pow(wn, n - 1), wn, h_out, 0};
    if (!bad) par_for(N, 1024, comp_body, &c);
    free(zcyc);
    free(rn);
    free(head);
    free(nxt);
    return bad ? -3 : c.failed ? -1 : 0;
}

/* Degree-2 split (crypto/stark/src/prover.rs:1605-1654): h natural N Fp3 ->
 * halves natural (N/2) x [H0 | H1], H0 = (h[i] + h[i+N/2])/2,
 * H1 = (h[i] - h[i+N/2]) / (2 offset w_N^i). */
int ora_decompose_d2(const uint64_t *h, size_t N, uint64_t offset, uint64_t *halves) {
    if (N < 2 || (N & (N - 1))) return -1;
    unsigned log_N = 0;
    while (((size_t)1 << log_N) < N) log_N++;
    const uint64_t wN = gl_root_of_unity(log_N), inv2 = gl_inv(2);
    const size_t half = N / 2;
    uint64_t x = offset;
    for (size_t i = 0; i < half; i++) {
        const fp3 a = ld3(h + 3 * i), b = ld3(h + 3 * (i + half));
        const fp3 h0 = fp3_mul_base(fp3_add(a, b), inv2);
        const fp3 h1 = fp3_mul_base(fp3_sub(a, b), gl_inv(gl_double(x)));
        memcpy(halves + 6 * i, h0.c, 24);
        memcpy(halves + 6 * i + 3, h1.c, 24);
        x = gl_mul(x, wN);
    }
    return 0;
}
