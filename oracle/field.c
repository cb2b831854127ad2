/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 * Vector entry points over the scalar field arithmetic of gl.h (restating
 * crypto/math/src/field/goldilocks.rs and extensions_goldilocks.rs), plus the
 * thread-count knob.
 */
#include "gl.h"
#include "lvm_oracle.h"
#include "par.h"

void ora_set_threads(int t) { par_set_threads(t); }
int ora_get_threads(void) { return par_threads(); }

uint64_t ora_gl_root_of_unity(unsigned order) { return gl_root_of_unity(order); }
uint64_t ora_gl_pow(uint64_t a, uint64_t e) { return gl_pow(a, e); }

void ora_gl_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        switch (op) {
        case ORA_ADD: out[i] = gl_add(a[i], b[i]); break;
        case ORA_SUB: out[i] = gl_sub(a[i], b[i]); break;
        case ORA_MUL: out[i] = gl_mul(a[i], b[i]); break;
        case ORA_NEG: out[i] = gl_neg(a[i]); break;
        case ORA_INV: out[i] = gl_inv(a[i]); break;
        case ORA_SQR: out[i] = gl_sqr(a[i]); break;
        default: out[i] = gl_canon(a[i]); break;
        }
    }
}

void ora_fp3_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        fp3 x = fp3_make(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
        fp3 y = fp3_zero();
        if (b) y = fp3_make(b[3 * i], b[3 * i + 1], b[3 * i + 2]);
        fp3 r;
        switch (op) {
        case ORA_ADD: r = fp3_add(x, y); break;
        case ORA_SUB: r = fp3_sub(x, y); break;
        case ORA_MUL: r = fp3_mul(x, y); break;
        case ORA_NEG: r = fp3_neg(x); break;
        case ORA_INV: r = fp3_inv(x); break;
        case ORA_SQR: r = fp3_sqr(x); break;
        default: r = fp3_canon(x); break;
        }
        out[3 * i] = r.c[0];
        out[3 * i + 1] = r.c[1];
        out[3 * i + 2] = r.c[2];
    }
}
