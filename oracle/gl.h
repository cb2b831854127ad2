/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker or the timed CPU baseline; the
 * product path (lambda_vm_amd) never links or calls it.
 *
 * Goldilocks field p = 2^64 - 2^32 + 1 and its cubic extension
 * Fp3 = Fp[w]/(w^3 - 2), restated from the reference CPU prover:
 *   crypto/math/src/field/goldilocks.rs:79-121   add/sub/mul (non-canonical u64 repr)
 *   crypto/math/src/field/goldilocks.rs:197-248  reduce128, add_no_canonicalize
 *   crypto/math/src/field/goldilocks.rs:263-315  dot_product_2/3 (overflow count, EPS^2 fix-up)
 *   crypto/math/src/field/goldilocks.rs:335-393  inverse by the p-2 addition chain
 *   crypto/math/src/field/goldilocks.rs:506-512  canonical
 *   crypto/math/src/field/goldilocks.rs:536-540  two-adicity 32, root 1753635133440165772
 *   crypto/math/src/field/traits.rs:83-95        primitive root of order 2^k by squaring
 *   crypto/math/src/field/extensions_goldilocks.rs:305-388  Fp3 mul/square/inv
 *   crypto/math/src/field/extensions_goldilocks.rs:420-476  base x ext ops
 * Raw (non-canonical) limbs follow the reference's operation sequence, so the
 * oracle's raw outputs equal the reference CPU's raw outputs for the same
 * operation order; parity checks compare canonical values.
 */
#ifndef LVM_ORACLE_GL_H
#define LVM_ORACLE_GL_H

#include <stdint.h>
#include <stddef.h>

#define GL_P 0xFFFFFFFF00000001ULL
#define GL_EPS 0xFFFFFFFFULL
#define GL_TWO_ADICITY 32
#define GL_ROOT_2_32 1753635133440165772ULL

typedef unsigned __int128 u128;

static inline uint64_t gl_canon(uint64_t a) { return a >= GL_P ? a - GL_P : a; }

/* goldilocks.rs:79-94 */
static inline uint64_t gl_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    uint64_t over = s < a;
    uint64_t s2 = s + over * GL_EPS;
    if (s2 < s) s2 += GL_EPS; /* double overflow: both inputs > p */
    return s2;
}

/* goldilocks.rs:98-113 */
static inline uint64_t gl_sub(uint64_t a, uint64_t b) {
    uint64_t d = a - b;
    uint64_t under = a < b;
    uint64_t corr = under * GL_EPS;
    uint64_t d2 = d - corr;
    if (d < corr) d2 -= GL_EPS;
    return d2;
}

/* goldilocks.rs:226-248 (portable form): x + y < 2^64 + p */
static inline uint64_t gl_add_nc(uint64_t x, uint64_t y) {
    uint64_t r = x + y;
    return r + (r < x ? GL_EPS : 0);
}

/* goldilocks.rs:197-216 */
static inline uint64_t gl_reduce128(u128 x) {
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t hi_hi = hi >> 32, hi_lo = hi & GL_EPS;
    uint64_t t0 = lo - hi_hi;
    if (lo < hi_hi) t0 -= GL_EPS;
    uint64_t t1 = (hi_lo << 32) - hi_lo;
    return gl_add_nc(t0, t1);
}

static inline uint64_t gl_mul(uint64_t a, uint64_t b) { return gl_reduce128((u128)a * b); }
static inline uint64_t gl_sqr(uint64_t a) { return gl_reduce128((u128)a * a); }
static inline uint64_t gl_double(uint64_t a) { return gl_add(a, a); }

/* goldilocks.rs:126-134 */
static inline uint64_t gl_neg(uint64_t a) {
    uint64_t c = gl_canon(a);
    return c == 0 ? 0 : GL_P - c;
}

/* goldilocks.rs:164-171 from_u64 */
static inline uint64_t gl_from_u64(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

#define GL_EPS_SQ ((uint64_t)(GL_EPS * GL_EPS))

/* goldilocks.rs:263-281 */
static inline uint64_t gl_dot2(uint64_t a0, uint64_t b0, uint64_t a1, uint64_t b1) {
    u128 p0 = (u128)a0 * b0, p1 = (u128)a1 * b1;
    u128 s = p0 + p1;
    int over = s < p0;
    uint64_t r = gl_reduce128(s);
    if (over) r = gl_add_nc(r, GL_EPS_SQ);
    return r;
}

/* goldilocks.rs:289-315 */
static inline uint64_t gl_dot3(uint64_t a0, uint64_t b0, uint64_t a1, uint64_t b1,
                               uint64_t a2, uint64_t b2) {
    u128 p0 = (u128)a0 * b0, p1 = (u128)a1 * b1, p2 = (u128)a2 * b2;
    u128 s01 = p0 + p1;
    int o1 = s01 < p0;
    u128 s = s01 + p2;
    int o2 = s < s01;
    uint64_t r = gl_reduce128(s);
    if (o1 + o2 > 0) {
        r = gl_add_nc(r, GL_EPS_SQ);
        if (o1 + o2 > 1) r = gl_add_nc(r, GL_EPS_SQ);
    }
    return r;
}

/* goldilocks.rs:335-393: x^(p-2) */
static inline uint64_t gl_exp_acc(uint64_t base, uint64_t tail, int n) {
    uint64_t r = base;
    for (int i = 0; i < n; i++) r = gl_sqr(r);
    return gl_mul(r, tail);
}

static inline uint64_t gl_inv(uint64_t a) {
    uint64_t x = gl_canon(a);
    if (x == 0) return 0; /* caller must not invert zero; reference returns Err */
    uint64_t x2 = gl_sqr(x);
    uint64_t x3 = gl_mul(x2, x);
    uint64_t x7 = gl_exp_acc(x3, x, 1);
    uint64_t x63 = gl_exp_acc(x7, x7, 3);
    uint64_t x12m1 = gl_exp_acc(x63, x63, 6);
    uint64_t x24m1 = gl_exp_acc(x12m1, x12m1, 12);
    uint64_t x30m1 = gl_exp_acc(x24m1, x63, 6);
    uint64_t x31m1 = gl_exp_acc(x30m1, x, 1);
    uint64_t x32m1 = gl_exp_acc(x31m1, x, 1);
    uint64_t t = x31m1;
    for (int i = 0; i < 33; i++) t = gl_sqr(t);
    return gl_mul(t, x32m1);
}

static inline uint64_t gl_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1, b = a;
    while (e) {
        if (e & 1) r = gl_mul(r, b);
        b = gl_sqr(b);
        e >>= 1;
    }
    return r;
}

/* traits.rs:83-95: primitive 2^order-th root of unity = root^(2^(32-order)) */
static inline uint64_t gl_root_of_unity(unsigned order) {
    if (order == 0) return 1;
    uint64_t r = GL_ROOT_2_32;
    for (unsigned i = 0; i < GL_TWO_ADICITY - order; i++) r = gl_sqr(r);
    return r;
}

/* ---------------- Fp3 = Fp[w]/(w^3 - 2) ---------------- */
typedef struct { uint64_t c[3]; } fp3;

static inline fp3 fp3_make(uint64_t a, uint64_t b, uint64_t c) { fp3 r = {{a, b, c}}; return r; }
static inline fp3 fp3_zero(void) { return fp3_make(0, 0, 0); }
static inline fp3 fp3_one(void) { return fp3_make(1, 0, 0); }
static inline fp3 fp3_add(fp3 a, fp3 b) {
    return fp3_make(gl_add(a.c[0], b.c[0]), gl_add(a.c[1], b.c[1]), gl_add(a.c[2], b.c[2]));
}
static inline fp3 fp3_sub(fp3 a, fp3 b) {
    return fp3_make(gl_sub(a.c[0], b.c[0]), gl_sub(a.c[1], b.c[1]), gl_sub(a.c[2], b.c[2]));
}
static inline fp3 fp3_neg(fp3 a) { return fp3_make(gl_neg(a.c[0]), gl_neg(a.c[1]), gl_neg(a.c[2])); }

/* extensions_goldilocks.rs:305-321 */
static inline fp3 fp3_mul(fp3 a, fp3 b) {
    uint64_t b1_2 = gl_double(b.c[1]), b2_2 = gl_double(b.c[2]);
    uint64_t c0 = gl_dot3(a.c[0], b.c[0], a.c[1], b2_2, a.c[2], b1_2);
    uint64_t c1 = gl_dot3(a.c[0], b.c[1], a.c[1], b.c[0], a.c[2], b2_2);
    uint64_t c2 = gl_dot3(a.c[0], b.c[2], a.c[1], b.c[1], a.c[2], b.c[0]);
    return fp3_make(c0, c1, c2);
}

/* extensions_goldilocks.rs:329-344 */
static inline fp3 fp3_sqr(fp3 a) {
    uint64_t a0_2 = gl_double(a.c[0]);
    uint64_t a2_4 = gl_double(gl_double(a.c[2]));
    uint64_t c0 = gl_dot2(a.c[0], a.c[0], a.c[1], a2_4);
    uint64_t a2_2 = gl_double(a.c[2]);
    uint64_t c1 = gl_dot2(a0_2, a.c[1], a2_2, a.c[2]);
    uint64_t c2 = gl_dot2(a.c[1], a.c[1], a0_2, a.c[2]);
    return fp3_make(c0, c1, c2);
}

/* extensions_goldilocks.rs:359-388 (norm-based inverse) */
static inline fp3 fp3_inv(fp3 a) {
    uint64_t a0_sq = gl_sqr(a.c[0]), a1_sq = gl_sqr(a.c[1]), a2_sq = gl_sqr(a.c[2]);
    uint64_t a0_cu = gl_mul(a0_sq, a.c[0]);
    uint64_t a1_cu = gl_mul(a1_sq, a.c[1]);
    uint64_t a2_cu = gl_mul(a2_sq, a.c[2]);
    uint64_t a012 = gl_mul(gl_mul(a.c[0], a.c[1]), a.c[2]);
    uint64_t norm = gl_add(gl_add(a0_cu, gl_double(a1_cu)), gl_double(gl_double(a2_cu)));
    norm = gl_sub(norm, gl_double(gl_add(gl_double(a012), a012)));
    uint64_t ni = gl_inv(norm);
    uint64_t a1a2 = gl_mul(a.c[1], a.c[2]);
    uint64_t a0a1 = gl_mul(a.c[0], a.c[1]);
    uint64_t a0a2 = gl_mul(a.c[0], a.c[2]);
    return fp3_make(gl_mul(gl_sub(a0_sq, gl_double(a1a2)), ni),
                    gl_mul(gl_sub(gl_double(a2_sq), a0a1), ni),
                    gl_mul(gl_sub(a1_sq, a0a2), ni));
}

/* extensions_goldilocks.rs:424-433: base * ext, component-wise */
static inline fp3 fp3_mul_base(fp3 a, uint64_t b) {
    return fp3_make(gl_mul(b, a.c[0]), gl_mul(b, a.c[1]), gl_mul(b, a.c[2]));
}
/* extensions_goldilocks.rs:436-443: base + ext */
static inline fp3 fp3_add_base(fp3 a, uint64_t b) { return fp3_make(gl_add(b, a.c[0]), a.c[1], a.c[2]); }
/* ext - base (E - F via sub_subfield): component 0 only */
static inline fp3 fp3_sub_base(fp3 a, uint64_t b) { return fp3_make(gl_sub(a.c[0], b), a.c[1], a.c[2]); }
/* extensions_goldilocks.rs:455-466: base - ext */
static inline fp3 fp3_base_sub(uint64_t a, fp3 b) {
    return fp3_make(gl_sub(a, b.c[0]), gl_neg(b.c[1]), gl_neg(b.c[2]));
}
static inline fp3 fp3_from_base(uint64_t a) { return fp3_make(a, 0, 0); }
static inline fp3 fp3_canon(fp3 a) { return fp3_make(gl_canon(a.c[0]), gl_canon(a.c[1]), gl_canon(a.c[2])); }

static inline fp3 fp3_pow(fp3 a, uint64_t e) {
    fp3 r = fp3_one(), b = a;
    while (e) {
        if (e & 1) r = fp3_mul(r, b);
        b = fp3_sqr(b);
        e >>= 1;
    }
    return r;
}

/* bit_reversing.rs:15-21 */
static inline uint64_t bitrev64(uint64_t x) {
    x = ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
    x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
    return __builtin_bswap64(x);
}
static inline uint64_t bitrev(uint64_t i, unsigned log_n) {
    if (log_n == 0) return i;
    return bitrev64(i) >> (64 - log_n);
}

#endif
