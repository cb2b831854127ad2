// Goldilocks (p = 2^64 - 2^32 + 1) and Fp3 = Fp[w]/(w^3 - 2) arithmetic for
// gfx950 device code.
//
// Semantics follow the reference CPU field (crypto/math/src/field/goldilocks.rs
// :79-121 add/sub/mul, :197-216 reduce128, :506-512 canonical;
// extensions_goldilocks.rs:305-388 Fp3): values are u64 limbs in [0, 2^64),
// not necessarily canonical, and every operation returns the same residue
// class as the reference. Results are compared canonically.
//
// The 64x64->128 product is built from four 32x32->64 partial products
// (v_mad_u64_u32 on CDNA) and reduced with 2^64 = 2^32 - 1, 2^96 = -1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LVM_HD __host__ __device__ __forceinline__

// Device-memory views in the global address space: loads and stores through
// them are global_* instructions, which count only against vmcnt, where
// generic (flat) ones also count against lgkmcnt and so make every LDS wait
// wait for the memory loads in flight as well.
typedef __attribute__((address_space(1))) uint64_t gmem_u64;
#define LVM_GLOBAL(p) ((gmem_u64 *)(p))
#define LVM_GLOBAL_CONST(p) ((const gmem_u64 *)(p))

namespace lvm {

constexpr uint64_t P = 0xFFFFFFFF00000001ULL;
constexpr uint64_t EPS = 0xFFFFFFFFULL;

LVM_HD uint64_t canon(uint64_t a) { return a >= P ? a - P : a; }

LVM_HD uint64_t add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    uint64_t s2 = s + (s < a ? EPS : 0);
    return s2 + (s2 < s ? EPS : 0);
}

LVM_HD uint64_t sub(uint64_t a, uint64_t b) {
    uint64_t d = a - b;
    uint64_t corr = a < b ? EPS : 0;
    uint64_t d2 = d - corr;
    return d2 - (d < corr ? EPS : 0);
}

// x + y with x + y < 2^64 + p (goldilocks.rs:226-248)
LVM_HD uint64_t add_nc(uint64_t x, uint64_t y) {
    uint64_t r = x + y;
    return r + (r < x ? EPS : 0);
}

LVM_HD uint64_t reduce128(uint64_t lo, uint64_t hi) {
    uint64_t hi_hi = hi >> 32, hi_lo = hi & EPS;
    uint64_t t0 = lo - hi_hi;
    t0 -= (lo < hi_hi) ? EPS : 0;
    uint64_t t1 = (hi_lo << 32) - hi_lo;
    return add_nc(t0, t1);
}

LVM_HD void mul_wide(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = (uint64_t)a0 * b0;
    uint64_t mid = (uint64_t)a0 * b1 + (p00 >> 32);
    uint64_t mid2 = (uint64_t)a1 * b0 + (uint32_t)mid;
    lo = (mid2 << 32) | (uint32_t)p00;
    hi = (uint64_t)a1 * b1 + (mid >> 32) + (mid2 >> 32);
}

LVM_HD uint64_t mul(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    mul_wide(a, b, lo, hi);
    return reduce128(lo, hi);
}

LVM_HD uint64_t sqr(uint64_t a) { return mul(a, a); }
LVM_HD uint64_t dbl(uint64_t a) { return add(a, a); }

LVM_HD uint64_t neg(uint64_t a) {
    uint64_t c = canon(a);
    return c == 0 ? 0 : P - c;
}

LVM_HD uint64_t pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mul(r, a);
        a = sqr(a);
        e >>= 1;
    }
    return r;
}

// a^(p-2), the addition chain of goldilocks.rs:335-393
LVM_HD uint64_t exp_acc(uint64_t base, uint64_t tail, int n) {
    for (int i = 0; i < n; i++) base = sqr(base);
    return mul(base, tail);
}

LVM_HD uint64_t inv(uint64_t a) {
    uint64_t x = canon(a);
    uint64_t x2 = sqr(x);
    uint64_t x3 = mul(x2, x);
    uint64_t x7 = exp_acc(x3, x, 1);
    uint64_t x63 = exp_acc(x7, x7, 3);
    uint64_t x12 = exp_acc(x63, x63, 6);
    uint64_t x24 = exp_acc(x12, x12, 12);
    uint64_t x30 = exp_acc(x24, x63, 6);
    uint64_t x31 = exp_acc(x30, x, 1);
    uint64_t x32 = exp_acc(x31, x, 1);
    uint64_t t = x31;
    for (int i = 0; i < 33; i++) t = sqr(t);
    return mul(t, x32);
}

// primitive 2^k-th root of unity: the 2^32-th root of goldilocks.rs:540
// squared 32 - k times (crypto/math/src/field/traits.rs:83-95)
LVM_HD uint64_t root_of_unity(unsigned k) {
    uint64_t r = 1753635133440165772ULL;
    for (unsigned i = k; i < 32; i++) r = sqr(r);
    return r;
}

// ---------------- canonical arithmetic ----------------
// Operations on canonical values (< p) that return canonical values. Every
// input range is stated; the NTT passes canonicalise what they load and then
// stay in this set, which needs one conditional correction per add or sub
// instead of two.

LVM_HD uint64_t ccanon(uint64_t r) { return r >= P ? r + EPS : r; }  // any r -> [0, p)

// a, b < p. A carry means a + b - p = s + EPS (< p, no second wrap); no
// carry and s >= p means s - p = s + EPS mod 2^64.
LVM_HD uint64_t cadd(uint64_t a, uint64_t b) {
    const uint64_t s = a + b;
    const bool c = (s < a) | (s >= P);
    return s + (c ? EPS : 0);
}

// a, b < p. A borrow means a - b + p = d - EPS with d = a - b + 2^64 > EPS.
LVM_HD uint64_t csub(uint64_t a, uint64_t b) {
    const uint64_t d = a - b;
    return a < b ? d - EPS : d;
}

// canonical reduction of any 128-bit lo + 2^64 hi (2^64 = EPS, 2^96 = -1):
// t0 = lo - hi_hi (mod p, in [0, 2^64)), t1 = hi_lo * EPS <= (2^32 - 1)^2,
// t0 + t1 wraps at most once and its correction cannot wrap again.
LVM_HD uint64_t cred(uint64_t lo, uint64_t hi) {
    const uint64_t hi_hi = hi >> 32, hi_lo = hi & EPS;
    uint64_t t0, r;
    // the borrow and carry as the operations' own flags (no 64-bit compare
    // against a zero-extended word)
    const bool borrow = __builtin_sub_overflow(lo, hi_hi, &t0);
    t0 -= borrow ? EPS : 0;
    const bool carry = __builtin_add_overflow(t0, (hi_lo << 32) - hi_lo, &r);
    r += carry ? EPS : 0;
    return ccanon(r);
}

// any a, b; canonical product
LVM_HD uint64_t cmul(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    mul_wide(a, b, lo, hi);
    return cred(lo, hi);
}

// 2^s mod p for s < 96: 2^s = 2^(s-64) * 2^64 = 2^(s-32) - 2^(s-64) for s >= 64
constexpr uint64_t pow2_mod(int s) {
    return s < 64 ? (1ull << s) : (1ull << (s - 32)) - (1ull << (s - 64));
}

// x * 2^S for canonical x and 0 <= S < 96: the shifted 128-bit value
// reduced (for S <= 32 the high word is below 2^32, so hi_hi = 0); for
// S >= 64 a product with the constant.
template <int S>
LVM_HD uint64_t cmul_pow2(uint64_t x) {
    static_assert(S >= 0 && S < 96, "shift out of range");
    if constexpr (S == 0) {
        return x;
    } else if constexpr (S <= 32) {
        const uint64_t lo = x << S, hi = x >> (64 - S);
        const uint64_t t1 = (hi << 32) - hi;
        uint64_t r = lo + t1;
        r += (r < lo) ? EPS : 0;
        return ccanon(r);
    } else if constexpr (S < 64) {
        return cred(x << S, x >> (64 - S));
    } else {
        return cmul(x, pow2_mod(S));
    }
}

// x * 2^-K for canonical x and 0 < K <= 32. p = 1 mod 2^32, so m = -x mod 2^K
// makes x + m*p a multiple of 2^K, and
//   (x + m*p) / 2^K = ((x + m) >> K) + (m * 2^(32-K)) * (2^32 - 1)
// where x + m <= p - 1 + 2^32 - 1 < 2^64. The first term is at most
// 2^(64-K) - 2^(32-K) and the second at most 2^64 - 2^32 - 2^(64-K) + 2^(32-K),
// so the result is at most p - 1: canonical with no correction (a shift, an
// add and one v_mad_u64_u32 where a general product needs four and a reduction).
template <int K>
LVM_HD uint64_t cmul_inv_pow2_32(uint64_t x) {
    static_assert(K > 0 && K <= 32, "divisor out of range");
    constexpr uint32_t MASK = K == 32 ? 0xFFFFFFFFu : ((1u << (K & 31)) - 1u);
    const uint32_t m = (0u - (uint32_t)x) & MASK;
    const uint64_t q = (x + m) >> K;
    return q + (uint64_t)(m << (32 - K)) * EPS;
}

// x * 2^-K for canonical x and 0 < K < 96, as divisions by at most 2^32
template <int K>
LVM_HD uint64_t cmul_inv_pow2(uint64_t x) {
    static_assert(K > 0 && K < 96, "divisor out of range");
    if constexpr (K <= 32)
        return cmul_inv_pow2_32<K>(x);
    else
        return cmul_inv_pow2<K - 32>(cmul_inv_pow2_32<32>(x));
}

// A multiplication by 2^T (any T >= 0; 2^192 = 1, 2^96 = -1) as the cheapest
// of the two forms: with S = T mod 96, 2^T = +-2^S; S <= 32 is a left shift
// and reduction (cmul_pow2), S > 32 the division 2^S = -2^-(96 - S).
// pow2_negates<T>() says whether the form computes -x * 2^T, so that a
// caller folds the sign into an operand order.
template <int T>
constexpr bool pow2_negates() {
    constexpr int S = (T % 192) % 96;
    return ((T % 192) >= 96) != (S > 32);
}
template <int T>
LVM_HD uint64_t cmul_pow2_form(uint64_t x) {
    constexpr int S = (T % 192) % 96;
    if constexpr (S <= 32)
        return cmul_pow2<S>(x);
    else
        return cmul_inv_pow2<96 - S>(x);
}

// ---------------- Fp3 ----------------
struct fp3 {
    uint64_t c0, c1, c2;
};

LVM_HD fp3 f3(uint64_t a, uint64_t b, uint64_t c) { return fp3{a, b, c}; }
LVM_HD fp3 f3_add(fp3 a, fp3 b) { return fp3{add(a.c0, b.c0), add(a.c1, b.c1), add(a.c2, b.c2)}; }
LVM_HD fp3 f3_sub(fp3 a, fp3 b) { return fp3{sub(a.c0, b.c0), sub(a.c1, b.c1), sub(a.c2, b.c2)}; }
LVM_HD fp3 f3_neg(fp3 a) { return fp3{neg(a.c0), neg(a.c1), neg(a.c2)}; }
LVM_HD fp3 f3_mul_base(fp3 a, uint64_t b) { return fp3{mul(b, a.c0), mul(b, a.c1), mul(b, a.c2)}; }
LVM_HD fp3 f3_canon(fp3 a) { return fp3{canon(a.c0), canon(a.c1), canon(a.c2)}; }

// sum of three 128-bit products reduced once, with 2^128 = EPS^2 for each
// carry out of 128 bits (goldilocks.rs:289-315)
LVM_HD uint64_t dot3(uint64_t a0, uint64_t b0, uint64_t a1, uint64_t b1, uint64_t a2, uint64_t b2) {
    uint64_t l0, h0, l1, h1, l2, h2;
    mul_wide(a0, b0, l0, h0);
    mul_wide(a1, b1, l1, h1);
    mul_wide(a2, b2, l2, h2);
    uint64_t lo = l0 + l1;
    uint64_t c = lo < l0;
    uint64_t hi = h0 + h1;
    uint32_t over = hi < h0;
    uint64_t hi2 = hi + c;
    over += hi2 < hi;
    uint64_t lo2 = lo + l2;
    c = lo2 < lo;
    uint64_t hi3 = hi2 + h2;
    over += hi3 < hi2;
    uint64_t hi4 = hi3 + c;
    over += hi4 < hi3;
    uint64_t r = reduce128(lo2, hi4);
    constexpr uint64_t EPS_SQ = EPS * EPS;
    if (over) r = add_nc(r, EPS_SQ);
    if (over > 1) r = add_nc(r, EPS_SQ);
    return r;
}

LVM_HD fp3 f3_mul(fp3 a, fp3 b) {
    uint64_t b1_2 = dbl(b.c1), b2_2 = dbl(b.c2);
    return fp3{dot3(a.c0, b.c0, a.c1, b2_2, a.c2, b1_2), dot3(a.c0, b.c1, a.c1, b.c0, a.c2, b2_2),
               dot3(a.c0, b.c2, a.c1, b.c1, a.c2, b.c0)};
}

LVM_HD fp3 f3_sqr(fp3 a) { return f3_mul(a, a); }

LVM_HD fp3 f3_inv(fp3 a) {
    uint64_t a0s = sqr(a.c0), a1s = sqr(a.c1), a2s = sqr(a.c2);
    uint64_t a0c = mul(a0s, a.c0), a1c = mul(a1s, a.c1), a2c = mul(a2s, a.c2);
    uint64_t a012 = mul(mul(a.c0, a.c1), a.c2);
    uint64_t norm = add(add(a0c, dbl(a1c)), dbl(dbl(a2c)));
    norm = sub(norm, dbl(add(dbl(a012), a012)));
    uint64_t ni = inv(norm);
    uint64_t a1a2 = mul(a.c1, a.c2), a0a1 = mul(a.c0, a.c1), a0a2 = mul(a.c0, a.c2);
    return fp3{mul(sub(a0s, dbl(a1a2)), ni), mul(sub(dbl(a2s), a0a1), ni), mul(sub(a1s, a0a2), ni)};
}

LVM_HD fp3 f3_pow(fp3 a, uint64_t e) {
    fp3 r = fp3{1, 0, 0};
    while (e) {
        if (e & 1) r = f3_mul(r, a);
        a = f3_sqr(a);
        e >>= 1;
    }
    return r;
}

// Delayed-reduction accumulator for dot products (the idea of the
// reference's dot_product_3, goldilocks.rs:289-315, stretched over a whole
// row of columns): the four 32x32-bit partial products of every 64x64-bit
// product go into three 64-bit running sums (weights 2^0, 2^32, 2^64), each
// with a 32-bit count of its wrap-arounds; one reduction at the end. A
// product costs four v_mad_u64_u32 plus carry bookkeeping, against a full
// modular multiply-add per product otherwise.
struct acc128 {
    uint64_t s0, s1, s2;
    uint32_t c0, c1, c2;
};

LVM_HD void acc_zero(acc128 &a) {
    a.s0 = a.s1 = a.s2 = 0;
    a.c0 = a.c1 = a.c2 = 0;
}

LVM_HD void acc_mad(acc128 &a, uint64_t x, uint64_t y) {
    const uint32_t a0 = (uint32_t)x, a1 = (uint32_t)(x >> 32), b0 = (uint32_t)y, b1 = (uint32_t)(y >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    // each partial product: one v_mad_u64_u32 into its running sum with the
    // carry-out mask, one v_addc_co_u32 adding that carry to the sum's count
    uint64_t cy;
    asm("v_mad_u64_u32 %0, %6, %7, %9, %0\n\t"
        "v_addc_co_u32 %3, %6, 0, %3, %6\n\t"
        "v_mad_u64_u32 %1, %6, %7, %10, %1\n\t"
        "v_addc_co_u32 %4, %6, 0, %4, %6\n\t"
        "v_mad_u64_u32 %1, %6, %8, %9, %1\n\t"
        "v_addc_co_u32 %4, %6, 0, %4, %6\n\t"
        "v_mad_u64_u32 %2, %6, %8, %10, %2\n\t"
        "v_addc_co_u32 %5, %6, 0, %5, %6"
        : "+v"(a.s0), "+v"(a.s1), "+v"(a.s2), "+v"(a.c0), "+v"(a.c1), "+v"(a.c2), "=&s"(cy)
        : "v"(a0), "v"(a1), "v"(b0), "v"(b1));
#else
    uint64_t t;
    t = a.s0 + (uint64_t)a0 * b0;
    a.c0 += t < a.s0;
    a.s0 = t;
    t = a.s1 + (uint64_t)a0 * b1;
    a.c1 += t < a.s1;
    a.s1 = t;
    t = a.s1 + (uint64_t)a1 * b0;
    a.c1 += t < a.s1;
    a.s1 = t;
    t = a.s2 + (uint64_t)a1 * b1;
    a.c2 += t < a.s2;
    a.s2 = t;
#endif
}

// value = s0 + c0*2^64 + s1*2^32 + c1*2^96 + s2*2^64 + c2*2^128 (mod p), with
// 2^64 = EPS, 2^96 = -1, 2^128 = EPS^2
LVM_HD uint64_t acc_reduce(const acc128 &a) {
    uint64_t r = a.s0;
    r = add(r, mul((uint64_t)a.c0, EPS));
    r = add(r, reduce128(a.s1 << 32, a.s1 >> 32));
    r = sub(r, (uint64_t)a.c1);
    r = add(r, reduce128(0, a.s2));
    r = add(r, mul((uint64_t)a.c2, EPS * EPS));
    return r;
}

LVM_HD uint64_t bswap64(uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    return ((uint64_t)__builtin_bswap32(lo) << 32) | __builtin_bswap32(hi);
#else
    return __builtin_bswap64(v);
#endif
}

LVM_HD uint32_t brev32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse32(x);
#else
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    return __builtin_bswap32(x);
#endif
}

// reverse the low `bits` bits of i (bits <= 32)
LVM_HD uint32_t bitrev(uint32_t i, unsigned bits) { return bits ? brev32(i) >> (32 - bits) : 0u; }

}  // namespace lvm
