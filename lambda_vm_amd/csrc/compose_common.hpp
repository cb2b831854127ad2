// Shared by the constraint-program interpreter and the run-time compiled
// constraint kernels (compose.hip): launch arguments, table access helpers
// and the per-row composition tail (zerofiers and boundary terms). The
// build embeds this text for the run-time compiler, so it includes nothing
// beyond field.hpp and uses only plain C++ and HIP device code.
#pragma once
#include "field.hpp"

namespace lvm {

constexpr uint32_t OP_CONST_BASE = 0, OP_CONST_EXT = 1, OP_VAR = 2, OP_RAP = 3, OP_ALPHA = 4, OP_OFFSET = 5,
                   OP_ADD = 6, OP_SUB = 7, OP_MUL = 8, OP_NEG = 9, OP_EMBED = 10;
constexpr uint32_t K_BASE_SLOT = 0, K_EXT_SLOT = 1, K_BASE_CONST = 2, K_EXT_CONST = 3, K_RAP = 4, K_ALPHA = 5;
constexpr int COMP_MAX_GROUPS = 4;
constexpr int COMP_MAX_EXEMPT = 8;
constexpr int COMP_MAX_STEPS = 16;
constexpr int COMP_BLOCK = 128;

struct CompArgs {
    // program
    const uint64_t *nodes;
    uint32_t num_nodes, num_roots;
    const uint64_t *base_consts, *ext_consts, *rap, *alpha, *toff;
    const uint32_t *roots;
    const uint32_t *emit_off, *emit_idx;  // per node: the constraints whose value it computes (CSR)
    uint32_t nbs, nes;  // base / ext slot counts
    // column-major copies of the tables in natural row order (natural LDE
    // index r): main (col, r) at main[col * N + r]; aux Fp3 at aux[3 * (col * N + r)]
    const uint64_t *main;
    const uint64_t *aux;
    uint32_t log_N;
    uint64_t next_step;
    // composition (mode 0) or per-constraint evals (mode 1)
    const uint64_t *beta;        // num_roots Fp3
    const uint32_t *root_group;  // num_roots
    uint32_t num_groups;
    uint32_t group_ex[COMP_MAX_GROUPS];
    uint64_t ex_root[COMP_MAX_EXEMPT];  // w^(n-1), w^(n-2), ...
    uint32_t max_ex;
    const uint64_t *zcyc;  // 1/(x^n - 1), period B
    uint32_t log_B;
    uint64_t offset;
    const uint64_t *tw_lo, *tw_hi;
    uint32_t num_steps;
    uint64_t step_point[COMP_MAX_STEPS];
    uint32_t num_boundary;
    uint32_t step_off[COMP_MAX_STEPS + 1];  // boundaries of step s: [step_off[s], step_off[s+1])
    uint64_t step_const[3 * COMP_MAX_STEPS];  // sum of beta_b * v_b over the boundaries of step s
    const uint32_t *b_col, *b_is_aux;      // boundary constraints, grouped by step
    const uint64_t *b_value, *b_beta;
    const uint64_t *binv;  // num_steps x N: 1 / (x_r - w^step_s) at natural index r
    // main-column boundaries per pair of steps (2k, 2k+1): entries
    // [bp_off[k], bp_off[k+1]), one per column, with the column, a presence
    // flag per step (bit j for step 2k+j) and the summed betas of the
    // column's boundaries at each step (6 words: step 2k, then 2k+1)
    const uint32_t *bp_col, *bp_flags;
    const uint64_t *bp_beta;
    uint32_t bp_off[COMP_MAX_STEPS / 2 + 2];
    uint64_t *out;       // mode 0: N Fp3 (order below); mode 1: [(c*N + r)*3]
    uint32_t natural_out;
    uint64_t *scratch;   // global slots when not in LDS
    uint64_t total_threads;
};

__device__ __forceinline__ fp3 ld3(const uint64_t *p) { return fp3{p[0], p[1], p[2]}; }
// w_{2^32}^e from the two-level twiddle tables
__device__ __forceinline__ uint64_t tw32_table(const uint64_t *lo, const uint64_t *hi, uint32_t e) {
    return mul(lo[e & 0xFFFFu], hi[e >> 16]);
}
__device__ __forceinline__ fp3 embed(uint64_t b) { return fp3{b, 0, 0}; }

// row of the column-major copies holding natural index (r + off*next_step) mod N
__device__ __forceinline__ uint64_t frame_row(const CompArgs &a, uint32_t r, uint32_t off) {
    const uint64_t N = 1ull << a.log_N;
    return ((uint64_t)r + off * a.next_step) & (N - 1);
}

// H(x_r) from the per-group transition sums: each group's sum times its
// inverse zerofier prod_j (x - w^(n-j)) / (x^n - 1), plus the boundary terms
// beta_b (t_b(x) - v_b) / (x - w^step_b), boundaries grouped by step, with the
// step denominators inverted beforehand (binv). A step's main-column terms
// sum_b beta_b t_b are three delayed-reduction dot products.
// Transition part: sum over groups of the group's sum times its inverse zerofier.
__device__ __forceinline__ fp3 composition_groups(const CompArgs &a, uint32_t r, const fp3 *gs) {
    // x = offset * w_N^r
    const uint64_t x = mul(a.offset, tw32_table(a.tw_lo, a.tw_hi, r << (32 - a.log_N)));
    const uint64_t zc = a.zcyc[r & ((1u << a.log_B) - 1)];
    fp3 h = fp3{0, 0, 0};
    for (uint32_t g = 0; g < a.num_groups; g++) {
        uint64_t z = zc;
        for (uint32_t j = 0; j < a.group_ex[g]; j++) z = mul(z, sub(x, a.ex_root[j]));
        fp3 gv = gs[0];
#pragma unroll
        for (int q = 1; q < COMP_MAX_GROUPS; q++)
            if ((uint32_t)q == g) gv = gs[q];
        h = f3_add(h, f3_mul_base(gv, z));
    }
    return h;
}

__device__ __forceinline__ fp3 composition_tail(const CompArgs &a, uint32_t r, const fp3 *gs) {
    fp3 h = composition_groups(a, r, gs);
    // steps in pairs: each main column of a pair's boundaries is read once and
    // feeds its boundaries' sums sum_b beta_b t_b, kept as delayed-reduction
    // dot products (three per step)
    for (uint32_t s0 = 0, k = 0; s0 < a.num_steps; s0 += 2, k++) {
        acc128 m00, m01, m02, m10, m11, m12;
        acc_zero(m00);
        acc_zero(m01);
        acc_zero(m02);
        acc_zero(m10);
        acc_zero(m11);
        acc_zero(m12);
        // unrolled so that several columns' loads are in flight together
#pragma unroll 8
        for (uint32_t e = a.bp_off[k]; e < a.bp_off[k + 1]; e++) {
            const uint32_t fl = a.bp_flags[e];
            const uint64_t v = a.main[((uint64_t)a.bp_col[e] << a.log_N) + r];
            const uint64_t *bb = a.bp_beta + 6 * (uint64_t)e;
            if (fl & 1) {
                acc_mad(m00, bb[0], v);
                acc_mad(m01, bb[1], v);
                acc_mad(m02, bb[2], v);
            }
            if (fl & 2) {
                acc_mad(m10, bb[3], v);
                acc_mad(m11, bb[4], v);
                acc_mad(m12, bb[5], v);
            }
        }
        for (uint32_t j = 0; j < 2 && s0 + j < a.num_steps; j++) {
            const uint32_t s = s0 + j;
            // sum_b beta_b (t_b - v_b) = sum_b beta_b t_b - step_const[s]
            fp3 acc = f3_neg(ld3(a.step_const + 3 * s));
            if (j == 0)
                acc = f3_add(acc, fp3{acc_reduce(m00), acc_reduce(m01), acc_reduce(m02)});
            else
                acc = f3_add(acc, fp3{acc_reduce(m10), acc_reduce(m11), acc_reduce(m12)});
            for (uint32_t bi = a.step_off[s]; bi < a.step_off[s + 1]; bi++)
                if (a.b_is_aux[bi])
                    acc = f3_add(acc, f3_mul(ld3(a.b_beta + 3 * (uint64_t)bi),
                                             ld3(a.aux + 3 * (((uint64_t)a.b_col[bi] << a.log_N) + r))));
            h = f3_add(h, f3_mul_base(acc, a.binv[((uint64_t)s << a.log_N) + r]));
        }
    }
    return h;
}

}  // namespace lvm
