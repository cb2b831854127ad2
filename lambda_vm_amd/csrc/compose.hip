// Constraint evaluation and the composition polynomial over a resident LDE,
// for gfx950.
//
// Program format: the flat constraint program of
// crypto/stark/src/constraint_ir/device.rs:60-160 (16-byte nodes op/a/b/res,
// operands kind << 29 | payload, dim-split base/ext value slots, uniform
// operands read from small tables). Semantics of the walk: device.rs
// eval_device_program; of the composition: crypto/stark/src/constraints/
// evaluator.rs:36-140 (transitions) and :244-270 (boundary), with the
// zerofiers of crypto/stark/src/constraints/zerofier.rs:8-110:
//   H(x) = sum_c beta_c C_c(x) prod_{j<=e_c}(x - w^(n-j)) / (x^n - 1)
//        + sum_b beta_b (t_b(x) - v_b) / (x - w^(step_b)).
//
// Design:
//  - one thread per LDE row in storage order: the tables keep rows
//    bit-reversed, so thread p reads physical row p and the frame row of
//    offset k at bitrev(bitrev(p) + k*next_step); for a wave of consecutive
//    p both are runs of consecutive rows, so the row reads stay coalesced;
//  - value slots live in LDS (slot s of lane t at [s][t]) when the program's
//    slot set fits, otherwise in a strided global scratch;
//  - zerofiers are computed per row instead of read: x = offset * w_N^r from
//    the two-level twiddle tables, 1/(x^n - 1) from a blowup-length table,
//    end-exemption factors as products, and the boundary denominators of
//    all distinct steps inverted together by Montgomery's trick (one base
//    inversion per row);
//  - transition sums are kept per zerofier group and multiplied by the
//    group's zerofier once.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "lvm_internal.hpp"
#include "prof.hpp"

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
    uint32_t nbs, nes;  // base / ext slot counts
    // tables: element (physical row, word) at lde[row * w + word]
    const uint64_t *main;
    uint64_t wm;
    const uint64_t *aux;
    uint64_t wa;
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
    const uint32_t *b_col, *b_is_aux, *b_step_idx;  // boundary constraints, step index into step_point
    const uint64_t *b_value, *b_beta;
    uint64_t *out;       // mode 0: N Fp3 (order below); mode 1: [(c*N + r)*3]
    uint32_t natural_out;
    uint64_t *scratch;   // global slots when not in LDS
    uint64_t total_threads;
};

struct Slots {
    uint64_t *base;  // [s * stride]
    uint64_t *ext;   // [(s*3 + k) * stride]
    uint64_t stride;
    __device__ __forceinline__ uint64_t &b(uint32_t s) const { return base[s * stride]; }
    __device__ __forceinline__ fp3 e(uint32_t s) const {
        return fp3{ext[(3 * s) * stride], ext[(3 * s + 1) * stride], ext[(3 * s + 2) * stride]};
    }
    __device__ __forceinline__ void set_e(uint32_t s, const fp3 &v) const {
        ext[(3 * s) * stride] = v.c0;
        ext[(3 * s + 1) * stride] = v.c1;
        ext[(3 * s + 2) * stride] = v.c2;
    }
};

struct Val {
    bool ext;
    uint64_t b;
    fp3 e;
};

__device__ __forceinline__ fp3 ld3(const uint64_t *p) { return fp3{p[0], p[1], p[2]}; }
// w_{2^32}^e from the two-level twiddle tables
__device__ __forceinline__ uint64_t tw32_table(const uint64_t *lo, const uint64_t *hi, uint32_t e) {
    return mul(lo[e & 0xFFFFu], hi[e >> 16]);
}
__device__ __forceinline__ fp3 embed(uint64_t b) { return fp3{b, 0, 0}; }

__device__ __forceinline__ Val operand(const CompArgs &a, uint32_t enc, const Slots &sl) {
    const uint32_t kind = enc >> 29, x = enc & 0x1FFFFFFFu;
    Val v;
    v.ext = kind != K_BASE_SLOT && kind != K_BASE_CONST;
    v.b = 0;
    v.e = fp3{0, 0, 0};
    switch (kind) {
    case K_BASE_SLOT: v.b = sl.b(x); break;
    case K_EXT_SLOT: v.e = sl.e(x); break;
    case K_BASE_CONST: v.b = a.base_consts[x]; break;
    case K_EXT_CONST: v.e = ld3(a.ext_consts + 3 * (uint64_t)x); break;
    case K_RAP: v.e = ld3(a.rap + 3 * (uint64_t)x); break;
    case K_ALPHA: v.e = ld3(a.alpha + 3 * (uint64_t)x); break;
    default: v.e = ld3(a.toff); break;
    }
    return v;
}

// physical row of natural index (r + off*next_step) mod N
__device__ __forceinline__ uint64_t frame_row(const CompArgs &a, uint32_t r, uint32_t off) {
    const uint64_t N = 1ull << a.log_N;
    const uint64_t nat = ((uint64_t)r + off * a.next_step) & (N - 1);
    return bitrev((uint32_t)nat, a.log_N);
}

__device__ __forceinline__ void walk(const CompArgs &a, uint32_t r, const Slots &sl) {
    for (uint32_t k = 0; k < a.num_nodes; k++) {
        const uint64_t w0 = a.nodes[2 * k], w1 = a.nodes[2 * k + 1];
        const uint32_t op = (uint32_t)w0, x = (uint32_t)(w0 >> 32), y = (uint32_t)w1, res = (uint32_t)(w1 >> 32);
        const bool rext = res >> 31;
        const uint32_t slot = res & 0x7FFFFFFFu;
        Val v;
        v.ext = false;
        v.b = 0;
        v.e = fp3{0, 0, 0};
        switch (op) {
        case OP_CONST_BASE: v.b = a.base_consts[x]; break;
        case OP_CONST_EXT: v.ext = true; v.e = ld3(a.ext_consts + 3 * (uint64_t)x); break;
        case OP_VAR: {
            const uint32_t col = x & 0xFFFF, is_main = (y >> 16) & 1, off = (y >> 8) & 0xFF;
            const uint64_t row = frame_row(a, r, off);
            if (is_main) {
                v.b = a.main[row * a.wm + col];
            } else {
                v.ext = true;
                v.e = ld3(a.aux + row * a.wa + 3 * (uint64_t)col);
            }
            break;
        }
        case OP_RAP: v.ext = true; v.e = ld3(a.rap + 3 * (uint64_t)x); break;
        case OP_ALPHA: v.ext = true; v.e = ld3(a.alpha + 3 * (uint64_t)x); break;
        case OP_OFFSET: v.ext = true; v.e = ld3(a.toff); break;
        case OP_ADD:
        case OP_SUB:
        case OP_MUL: {
            const Val p = operand(a, x, sl), q = operand(a, y, sl);
            if (!p.ext && !q.ext) {
                v.b = op == OP_ADD ? add(p.b, q.b) : op == OP_SUB ? sub(p.b, q.b) : mul(p.b, q.b);
            } else {
                v.ext = true;
                if (op == OP_MUL && !p.ext)
                    v.e = f3_mul_base(q.e, p.b);
                else if (op == OP_MUL && !q.ext)
                    v.e = f3_mul_base(p.e, q.b);
                else {
                    const fp3 pe = p.ext ? p.e : embed(p.b), qe = q.ext ? q.e : embed(q.b);
                    v.e = op == OP_ADD ? f3_add(pe, qe) : op == OP_SUB ? f3_sub(pe, qe) : f3_mul(pe, qe);
                }
            }
            break;
        }
        case OP_NEG: {
            const Val p = operand(a, x, sl);
            v.ext = p.ext;
            if (p.ext)
                v.e = f3_neg(p.e);
            else
                v.b = neg(p.b);
            break;
        }
        default: {  // OP_EMBED
            const Val p = operand(a, x, sl);
            v.ext = true;
            v.e = p.ext ? p.e : embed(p.b);
            break;
        }
        }
        if (rext)
            sl.set_e(slot, v.ext ? v.e : embed(v.b));
        else
            sl.b(slot) = v.b;
    }
}

__device__ __forceinline__ fp3 root_value(const CompArgs &a, uint32_t c, const Slots &sl) {
    const uint32_t rs = a.roots[c];
    return (rs >> 31) ? sl.e(rs & 0x7FFFFFFFu) : embed(sl.b(rs));
}

template <bool IN_LDS, int MODE>
__global__ __launch_bounds__(COMP_BLOCK) void compose_kernel(CompArgs a) {
    extern __shared__ uint64_t s_slots[];
    Slots sl;
    if (IN_LDS) {
        sl.base = s_slots + threadIdx.x;
        sl.ext = s_slots + (uint64_t)a.nbs * COMP_BLOCK + threadIdx.x;
        sl.stride = COMP_BLOCK;
    } else {
        const uint64_t gt = (uint64_t)blockIdx.x * COMP_BLOCK + threadIdx.x;
        sl.base = a.scratch + gt;
        sl.ext = a.scratch + (uint64_t)a.nbs * a.total_threads + gt;
        sl.stride = a.total_threads;
    }
    const uint64_t N = 1ull << a.log_N;
    for (uint64_t p = (uint64_t)blockIdx.x * COMP_BLOCK + threadIdx.x; p < N; p += a.total_threads) {
        const uint32_t r = bitrev((uint32_t)p, a.log_N);  // natural index of physical row p
        walk(a, r, sl);
        if (MODE == 1) {
            for (uint32_t c = 0; c < a.num_roots; c++) {
                const fp3 v = root_value(a, c, sl);
                uint64_t *o = a.out + 3 * ((uint64_t)c * N + r);
                o[0] = v.c0;
                o[1] = v.c1;
                o[2] = v.c2;
            }
            continue;
        }
        fp3 gsum[COMP_MAX_GROUPS];
#pragma unroll
        for (int g = 0; g < COMP_MAX_GROUPS; g++) gsum[g] = fp3{0, 0, 0};
        for (uint32_t c = 0; c < a.num_roots; c++) {
            const uint32_t rs = a.roots[c];
            const fp3 beta = ld3(a.beta + 3 * (uint64_t)c);
            const fp3 t = (rs >> 31) ? f3_mul(sl.e(rs & 0x7FFFFFFFu), beta) : f3_mul_base(beta, sl.b(rs));
            const uint32_t g = a.root_group[c];
#pragma unroll
            for (int q = 0; q < COMP_MAX_GROUPS; q++)
                if ((uint32_t)q == g) gsum[q] = f3_add(gsum[q], t);
        }
        // x = offset * w_N^r
        const uint64_t x = mul(a.offset, tw32_table(a.tw_lo, a.tw_hi, r << (32 - a.log_N)));
        const uint64_t zc = a.zcyc[r & ((1u << a.log_B) - 1)];
        uint64_t exf[COMP_MAX_EXEMPT + 1];
        exf[0] = zc;
#pragma unroll
        for (int j = 0; j < COMP_MAX_EXEMPT; j++)
            exf[j + 1] = (uint32_t)j < a.max_ex ? mul(exf[j], sub(x, a.ex_root[j])) : exf[j];
        fp3 h = fp3{0, 0, 0};
#pragma unroll
        for (int g = 0; g < COMP_MAX_GROUPS; g++) {
            if ((uint32_t)g < a.num_groups) {
                uint64_t z = exf[0];
#pragma unroll
                for (int j = 1; j <= COMP_MAX_EXEMPT; j++)
                    if ((uint32_t)j == a.group_ex[g]) z = exf[j];
                h = f3_add(h, f3_mul_base(gsum[g], z));
            }
        }
        if (a.num_boundary) {
            // 1/(x - w^s) for every distinct step s by one inversion
            uint64_t pre[COMP_MAX_STEPS];
            uint64_t run = 1;
#pragma unroll
            for (int s = 0; s < COMP_MAX_STEPS; s++) {
                if ((uint32_t)s < a.num_steps) run = mul(run, sub(x, a.step_point[s]));
                pre[s] = run;
            }
            uint64_t inv_all = inv(run);
            uint64_t dinv[COMP_MAX_STEPS];
#pragma unroll
            for (int s = COMP_MAX_STEPS - 1; s >= 0; s--) {
                if ((uint32_t)s < a.num_steps) {
                    dinv[s] = s ? mul(inv_all, pre[s - 1]) : inv_all;
                    inv_all = mul(inv_all, sub(x, a.step_point[s]));
                } else {
                    dinv[s] = 0;
                }
            }
            for (uint32_t bi = 0; bi < a.num_boundary; bi++) {
                const uint32_t col = a.b_col[bi];
                fp3 t;
                if (a.b_is_aux[bi])
                    t = ld3(a.aux + p * a.wa + 3 * (uint64_t)col);
                else
                    t = embed(a.main[p * a.wm + col]);
                const fp3 bp = f3_sub(t, ld3(a.b_value + 3 * (uint64_t)bi));
                const uint32_t si = a.b_step_idx[bi];
                uint64_t zb = 0;
#pragma unroll
                for (int s = 0; s < COMP_MAX_STEPS; s++)
                    if ((uint32_t)s == si) zb = dinv[s];
                h = f3_add(h, f3_mul(f3_mul_base(ld3(a.b_beta + 3 * (uint64_t)bi), zb), bp));
            }
        }
        const uint64_t orow = a.natural_out ? r : p;
        uint64_t *o = a.out + 3 * orow;
        o[0] = h.c0;
        o[1] = h.c1;
        o[2] = h.c2;
    }
}

// Degree-2 split of the composition polynomial (crypto/stark/src/prover.rs
// :1605-1654 decompose_and_extend_d2): on the LDE coset, -x_r = x_{r+N/2},
// so with rows stored bit-reversed H(x) and H(-x) are the physical pair
// (2q, 2q+1) with r = bitrev_{N/2}(q). Writes natural-order half rows
// [H0(x_r^2) | H1(x_r^2)] (6 words):
//   H0 = (H(x) + H(-x)) / 2,  H1 = (H(x) - H(-x)) / (2x),
// with 1/(2x_r) = (2g)^-1 * w_N^-r from the inverse twiddle tables.
__global__ __launch_bounds__(256) void d2_split_kernel(const uint64_t *__restrict__ hb, uint32_t log_N,
                                                        uint64_t inv2, uint64_t inv_2g, const uint64_t *inv_lo,
                                                        const uint64_t *inv_hi, uint64_t *__restrict__ halves) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (1ull << (log_N - 1))) return;
    const uint32_t r = bitrev((uint32_t)q, log_N - 1);
    const fp3 a = ld3(hb + 6 * q), b = ld3(hb + 6 * q + 3);
    const fp3 h0 = f3_mul_base(f3_add(a, b), inv2);
    const uint64_t ix = mul(inv_2g, tw32_table(inv_lo, inv_hi, r << (32 - log_N)));
    const fp3 h1 = f3_mul_base(f3_sub(a, b), ix);
    uint64_t *o = halves + 6 * (uint64_t)r;
    o[0] = h0.c0;
    o[1] = h0.c1;
    o[2] = h0.c2;
    o[3] = h1.c0;
    o[4] = h1.c1;
    o[5] = h1.c2;
}

void launch_d2_split(const uint64_t *hb, unsigned log_N, uint64_t coset_offset, const Tables &tb, uint64_t *halves,
                     hipStream_t st) {
    const uint64_t half = 1ull << (log_N - 1);
    const uint64_t inv2 = inv(2), inv_2g = inv(dbl(coset_offset));
    hipLaunchKernelGGL(d2_split_kernel, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, st, hb, log_N, inv2,
                       inv_2g, tb.inv_lo, tb.inv_hi, halves);
}

namespace {

template <class T>
size_t put(std::vector<uint8_t> &blob, const T *v, size_t n) {
    size_t at = (blob.size() + 15) & ~(size_t)15;
    blob.resize(at + (n ? n : 1) * sizeof(T));
    if (n) std::memcpy(blob.data() + at, v, n * sizeof(T));
    return at;
}

bool is_pow2(size_t x) { return x && !(x & (x - 1)); }
unsigned log2u(size_t x) { return (unsigned)__builtin_ctzll(x); }

}  // namespace

// Shared by the composition and the per-constraint entry points.
int compose_impl(const lvm_program *pg, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                 const lvm_composition *cp, int mode, int natural_out, uint64_t *out) {
    if (!pg || !main || !out) return fail(LVM_ERR_ARG, "null argument");
    if (mode == 0 && !cp) return fail(LVM_ERR_ARG, "null composition");
    const size_t N = main->N;
    if (!is_pow2(N) || log2u(N) > 32) return fail(LVM_ERR_ARG, "bad LDE size");
    if (aux && aux->N != N) return fail(LVM_ERR_ARG, "main and aux LDE sizes differ");
    if (next_step >= N) return fail(LVM_ERR_ARG, "frame step must be smaller than the LDE");
    if (pg->num_nodes == 0 || pg->num_roots == 0 || !pg->nodes || !pg->roots)
        return fail(LVM_ERR_ARG, "empty program");
    // validate the program: every operand, slot, column and table index in range
    const size_t ma = aux ? aux->w / 3 : 0;
    uint32_t max_off = 0;
    for (size_t k = 0; k < pg->num_nodes; k++) {
        const uint64_t w0 = pg->nodes[2 * k], w1 = pg->nodes[2 * k + 1];
        const uint32_t op = (uint32_t)w0, x = (uint32_t)(w0 >> 32), y = (uint32_t)w1, res = (uint32_t)(w1 >> 32);
        const uint32_t slot = res & 0x7FFFFFFFu;
        if (op > OP_EMBED) return fail(LVM_ERR_ARG, "unknown program op");
        if ((res >> 31) ? slot >= pg->num_ext_slots : slot >= pg->num_base_slots)
            return fail(LVM_ERR_ARG, "result slot out of range");
        auto check_operand = [&](uint32_t enc) -> bool {
            const uint32_t kind = enc >> 29, i = enc & 0x1FFFFFFFu;
            switch (kind) {
            case K_BASE_SLOT: return i < pg->num_base_slots;
            case K_EXT_SLOT: return i < pg->num_ext_slots;
            case K_BASE_CONST: return i < pg->num_base_consts;
            case K_EXT_CONST: return i < pg->num_ext_consts;
            case K_RAP: return i < pg->num_rap;
            case K_ALPHA: return i < pg->num_alpha;
            case 6: return pg->table_offset != nullptr;
            default: return false;
            }
        };
        switch (op) {
        case OP_CONST_BASE:
            if (x >= pg->num_base_consts) return fail(LVM_ERR_ARG, "constant index out of range");
            break;
        case OP_CONST_EXT:
            if (x >= pg->num_ext_consts) return fail(LVM_ERR_ARG, "constant index out of range");
            break;
        case OP_VAR: {
            const uint32_t col = x & 0xFFFF, is_main = (y >> 16) & 1, off = (y >> 8) & 0xFF;
            if (is_main ? col >= main->w : (!aux || col >= ma)) return fail(LVM_ERR_ARG, "column out of range");
            max_off = off > max_off ? off : max_off;
            break;
        }
        case OP_RAP:
            if (x >= pg->num_rap) return fail(LVM_ERR_ARG, "challenge index out of range");
            break;
        case OP_ALPHA:
            if (x >= pg->num_alpha) return fail(LVM_ERR_ARG, "alpha index out of range");
            break;
        case OP_OFFSET:
            if (!pg->table_offset) return fail(LVM_ERR_ARG, "table offset missing");
            break;
        case OP_ADD:
        case OP_SUB:
        case OP_MUL:
            if (!check_operand(x) || !check_operand(y)) return fail(LVM_ERR_ARG, "operand out of range");
            break;
        default:
            if (!check_operand(x)) return fail(LVM_ERR_ARG, "operand out of range");
            break;
        }
    }
    (void)max_off;
    for (size_t c = 0; c < pg->num_roots; c++) {
        const uint32_t rs = (uint32_t)pg->roots[c];
        if ((rs >> 31) ? (rs & 0x7FFFFFFFu) >= pg->num_ext_slots : rs >= pg->num_base_slots)
            return fail(LVM_ERR_ARG, "root slot out of range");
    }
    const unsigned logN = log2u(N);
    CompArgs a{};
    std::vector<uint8_t> blob;
    std::vector<uint32_t> roots32(pg->num_roots);
    for (size_t c = 0; c < pg->num_roots; c++) roots32[c] = (uint32_t)pg->roots[c];
    const size_t o_nodes = put(blob, pg->nodes, 2 * pg->num_nodes), o_bc = put(blob, pg->base_consts, pg->num_base_consts),
                 o_ec = put(blob, pg->ext_consts, 3 * pg->num_ext_consts), o_rap = put(blob, pg->rap_challenges, 3 * pg->num_rap),
                 o_al = put(blob, pg->alpha_powers, 3 * pg->num_alpha),
                 o_to = put(blob, pg->table_offset, pg->table_offset ? 3 : 0), o_roots = put(blob, roots32.data(), roots32.size());
    size_t o_beta = 0, o_rg = 0, o_z = 0, o_bcol = 0, o_baux = 0, o_bsi = 0, o_bval = 0, o_bbeta = 0;
    if (mode == 0) {
        const size_t n = cp->trace_rows;
        if (!is_pow2(n) || N % n || n < 2) return fail(LVM_ERR_ARG, "trace rows must be a power of two dividing the LDE size");
        if (!cp->beta || !cp->root_group || !cp->group_exemptions) return fail(LVM_ERR_ARG, "composition arrays missing");
        if (cp->num_groups == 0 || cp->num_groups > COMP_MAX_GROUPS) return fail(LVM_ERR_ARG, "1 to 4 zerofier groups");
        for (size_t c = 0; c < pg->num_roots; c++)
            if (cp->root_group[c] >= cp->num_groups) return fail(LVM_ERR_ARG, "zerofier group out of range");
        uint32_t max_ex = 0;
        for (size_t g = 0; g < cp->num_groups; g++) {
            if (cp->group_exemptions[g] > COMP_MAX_EXEMPT) return fail(LVM_ERR_ARG, "at most 8 end exemptions");
            a.group_ex[g] = cp->group_exemptions[g];
            max_ex = cp->group_exemptions[g] > max_ex ? cp->group_exemptions[g] : max_ex;
        }
        if (max_ex >= n) return fail(LVM_ERR_ARG, "exemptions exceed the trace");
        a.num_groups = (uint32_t)cp->num_groups;
        a.max_ex = max_ex;
        const unsigned logn = log2u(n), logB = logN - logn;
        const uint64_t wn = root_of_unity(logn), wN = root_of_unity(logN);
        const uint64_t w_last = pow(wn, n - 1);
        uint64_t rr = w_last;
        for (uint32_t j = 0; j < max_ex; j++) {
            a.ex_root[j] = rr;
            rr = mul(rr, w_last);
        }
        const size_t B = N / n;
        std::vector<uint64_t> zc(B);
        {
            const uint64_t step = pow(wN, n);
            uint64_t d = pow(cp->coset_offset, n);
            for (size_t j = 0; j < B; j++) {
                const uint64_t den = sub(d, 1);
                if (canon(den) == 0) return fail(LVM_ERR_ARG, "coset offset makes x^n - 1 vanish on the LDE");
                zc[j] = inv(den);
                d = mul(d, step);
            }
        }
        a.log_B = logB;
        a.offset = cp->coset_offset;
        // boundary steps
        std::vector<uint32_t> step_idx(cp->num_boundary);
        std::vector<uint64_t> steps;
        for (size_t b = 0; b < cp->num_boundary; b++) {
            if (cp->b_is_aux[b] ? (!aux || cp->b_col[b] >= ma) : cp->b_col[b] >= main->w)
                return fail(LVM_ERR_ARG, "boundary column out of range");
            const uint64_t s = cp->b_step[b] % n;
            size_t k = 0;
            while (k < steps.size() && steps[k] != s) k++;
            if (k == steps.size()) {
                if (steps.size() == COMP_MAX_STEPS) return fail(LVM_ERR_ARG, "at most 16 distinct boundary steps");
                steps.push_back(s);
            }
            step_idx[b] = (uint32_t)k;
        }
        a.num_steps = (uint32_t)steps.size();
        for (size_t k = 0; k < steps.size(); k++) a.step_point[k] = pow(wn, steps[k]);
        a.num_boundary = (uint32_t)cp->num_boundary;
        o_beta = put(blob, cp->beta, 3 * pg->num_roots);
        o_rg = put(blob, cp->root_group, pg->num_roots);
        o_z = put(blob, zc.data(), B);
        o_bcol = put(blob, cp->b_col, cp->num_boundary);
        o_baux = put(blob, cp->b_is_aux, cp->num_boundary);
        o_bsi = put(blob, step_idx.data(), cp->num_boundary);
        o_bval = put(blob, cp->b_value, 3 * cp->num_boundary);
        o_bbeta = put(blob, cp->b_beta, 3 * cp->num_boundary);
    }
    Device *dv = device_state();
    if (!dv) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = lib_stream();
    const size_t slot_bytes = 8 * (size_t)pg->num_base_slots + 24 * (size_t)pg->num_ext_slots;
    const bool in_lds = slot_bytes * COMP_BLOCK <= 64 * 1024;
    const uint64_t max_blocks = 256 * 16;
    const uint64_t blocks = std::min<uint64_t>((N + COMP_BLOCK - 1) / COMP_BLOCK, max_blocks);
    a.total_threads = blocks * COMP_BLOCK;
    uint8_t *db = nullptr;
    uint64_t *scratch = nullptr;
    if (lvm_malloc(&db, blob.size()) != hipSuccess ||
        (!in_lds && lvm_malloc(&scratch, slot_bytes * a.total_threads) != hipSuccess)) {
        lvm_free(db);
        return fail(LVM_ERR_NOMEM, "device allocation for the constraint program failed");
    }
    hipError_t e = hipMemcpyAsync(db, blob.data(), blob.size(), hipMemcpyHostToDevice, st);
    a.nodes = (const uint64_t *)(db + o_nodes);
    a.num_nodes = (uint32_t)pg->num_nodes;
    a.num_roots = (uint32_t)pg->num_roots;
    a.base_consts = (const uint64_t *)(db + o_bc);
    a.ext_consts = (const uint64_t *)(db + o_ec);
    a.rap = (const uint64_t *)(db + o_rap);
    a.alpha = (const uint64_t *)(db + o_al);
    a.toff = (const uint64_t *)(db + o_to);
    a.roots = (const uint32_t *)(db + o_roots);
    a.nbs = (uint32_t)pg->num_base_slots;
    a.nes = (uint32_t)pg->num_ext_slots;
    a.main = main->lde;
    a.wm = main->w;
    a.aux = aux ? aux->lde : nullptr;
    a.wa = aux ? aux->w : 0;
    a.log_N = logN;
    a.next_step = next_step;
    a.tw_lo = dv->tb.fwd_lo;
    a.tw_hi = dv->tb.fwd_hi;
    if (mode == 0) {
        a.beta = (const uint64_t *)(db + o_beta);
        a.root_group = (const uint32_t *)(db + o_rg);
        a.zcyc = (const uint64_t *)(db + o_z);
        a.b_col = (const uint32_t *)(db + o_bcol);
        a.b_is_aux = (const uint32_t *)(db + o_baux);
        a.b_step_idx = (const uint32_t *)(db + o_bsi);
        a.b_value = (const uint64_t *)(db + o_bval);
        a.b_beta = (const uint64_t *)(db + o_bbeta);
    }
    a.out = out;
    a.natural_out = natural_out ? 1u : 0u;
    a.scratch = scratch;
    const size_t lds = in_lds ? slot_bytes * COMP_BLOCK : 0;
    {
        ProfScope ps(mode == 0 ? "composition" : "constraint_evals", st);
        if (in_lds) {
            if (mode == 0)
                hipLaunchKernelGGL((compose_kernel<true, 0>), dim3((unsigned)blocks), dim3(COMP_BLOCK), lds, st, a);
            else
                hipLaunchKernelGGL((compose_kernel<true, 1>), dim3((unsigned)blocks), dim3(COMP_BLOCK), lds, st, a);
        } else {
            if (mode == 0)
                hipLaunchKernelGGL((compose_kernel<false, 0>), dim3((unsigned)blocks), dim3(COMP_BLOCK), 0, st, a);
            else
                hipLaunchKernelGGL((compose_kernel<false, 1>), dim3((unsigned)blocks), dim3(COMP_BLOCK), 0, st, a);
        }
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    lvm_free(db);
    lvm_free(scratch);
    if (e != hipSuccess) return fail(LVM_ERR_DEVICE, std::string("constraint evaluation failed: ") + hipGetErrorString(e));
    return LVM_OK;
}

}  // namespace lvm

extern "C" {

int lvm_eval_composition(const lvm_program *program, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                         const lvm_composition *comp, int natural_order, uint64_t *dev_h_out) {
    return lvm::compose_impl(program, main, aux, next_step, comp, 0, natural_order, dev_h_out);
}

int lvm_eval_constraints(const lvm_program *program, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                         uint64_t *dev_evals_out) {
    return lvm::compose_impl(program, main, aux, next_step, nullptr, 1, 1, dev_evals_out);
}

}  // extern "C"
