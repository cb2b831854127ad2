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

#include "compose_common.hpp"
#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

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

// Consumes the value of constraint c produced at row r: the composition adds
// beta_c * value into its zerofier group's sum, the per-constraint mode
// stores it.
template <int MODE>
__device__ __forceinline__ void emit_value(const CompArgs &a, uint32_t c, bool ext, uint64_t vb, const fp3 &ve,
                                           uint32_t r, uint64_t *gsum) {
    if (MODE == 1) {
        const uint64_t N = 1ull << a.log_N;
        uint64_t *o = a.out + 3 * ((uint64_t)c * N + r);
        o[0] = ext ? ve.c0 : vb;
        o[1] = ext ? ve.c1 : 0;
        o[2] = ext ? ve.c2 : 0;
        return;
    }
    const fp3 beta = ld3(a.beta + 3 * (uint64_t)c);
    const fp3 t = ext ? f3_mul(ve, beta) : f3_mul_base(beta, vb);
    uint64_t *gs = gsum + 3 * a.root_group[c] * COMP_BLOCK;  // this lane's group sum in LDS
    const fp3 cur = fp3{gs[0], gs[COMP_BLOCK], gs[2 * COMP_BLOCK]};
    const fp3 nxt = f3_add(cur, t);
    gs[0] = nxt.c0;
    gs[COMP_BLOCK] = nxt.c1;
    gs[2 * COMP_BLOCK] = nxt.c2;
}

constexpr int VAR_BATCH = 8;

// Forward walk of the program at natural row r. Runs of consecutive trace
// reads are issued together (up to VAR_BATCH loads in flight per lane)
// before their slots are written. Every constraint value is handed to
// emit_value when its node has run.
template <int MODE>
__device__ __forceinline__ void walk(const CompArgs &a, const uint64_t *__restrict__ nodes, uint32_t num_nodes,
                                     const uint32_t *__restrict__ emit_off, const uint32_t *__restrict__ emit_idx,
                                     uint32_t r, const Slots &sl, uint64_t *gsum) {
    for (uint32_t k = 0; k < num_nodes; k++) {
        const uint64_t w0 = nodes[2 * k], w1 = nodes[2 * k + 1];
        const uint32_t op = (uint32_t)w0;
        if (op == OP_VAR && ((uint32_t)nodes[2 * k + 1] >> 16 & 1)) {
            // a run of consecutive main-trace reads, loads issued together
            uint32_t cnt = 1;
            while (cnt < VAR_BATCH && k + cnt < num_nodes && (uint32_t)nodes[2 * (k + cnt)] == OP_VAR &&
                   ((uint32_t)nodes[2 * (k + cnt) + 1] >> 16 & 1))
                cnt++;
            uint64_t lv[VAR_BATCH];
#pragma unroll
            for (int j = 0; j < VAR_BATCH; j++) {
                if ((uint32_t)j < cnt) {
                    const uint64_t u0 = nodes[2 * (k + j)], u1 = nodes[2 * (k + j) + 1];
                    const uint32_t col = (uint32_t)(u0 >> 32) & 0xFFFF, off = ((uint32_t)u1 >> 8) & 0xFF;
                    lv[j] = a.main[((uint64_t)col << a.log_N) + frame_row(a, r, off)];
                }
            }
#pragma unroll
            for (int j = 0; j < VAR_BATCH; j++) {
                if ((uint32_t)j < cnt) {
                    const uint32_t res = (uint32_t)(nodes[2 * (k + j) + 1] >> 32);
                    const uint32_t slot = res & 0x7FFFFFFFu;
                    if (res >> 31)
                        sl.set_e(slot, embed(lv[j]));
                    else
                        sl.b(slot) = lv[j];
                    for (uint32_t q = emit_off[k + j]; q < emit_off[k + j + 1]; q++)
                        emit_value<MODE>(a, emit_idx[q], false, lv[j], embed(lv[j]), r, gsum);
                }
            }
            k += cnt - 1;
            continue;
        }
        const uint32_t x = (uint32_t)(w0 >> 32), y = (uint32_t)w1, res = (uint32_t)(w1 >> 32);
        const bool rext = res >> 31;
        const uint32_t slot = res & 0x7FFFFFFFu;
        Val v;
        v.ext = false;
        v.b = 0;
        v.e = fp3{0, 0, 0};
        switch (op) {
        case OP_CONST_BASE: v.b = a.base_consts[x]; break;
        case OP_CONST_EXT: v.ext = true; v.e = ld3(a.ext_consts + 3 * (uint64_t)x); break;
        case OP_VAR: {  // aux (Fp3) read
            const uint32_t col = x & 0xFFFF, off = (y >> 8) & 0xFF;
            v.ext = true;
            v.e = ld3(a.aux + 3 * (((uint64_t)col << a.log_N) + frame_row(a, r, off)));
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
        if (v.ext && !rext) v.ext = false;  // a base slot keeps component 0 (not produced by valid programs)
        if (rext)
            sl.set_e(slot, v.ext ? v.e : embed(v.b));
        else
            sl.b(slot) = v.b;
        for (uint32_t q = emit_off[k]; q < emit_off[k + 1]; q++)
            emit_value<MODE>(a, emit_idx[q], v.ext, v.b, v.ext ? v.e : embed(v.b), r, gsum);
    }
}

// ptab (optional): the program split into parts, block row y taking part y
// (its node list at nodes + 2*ptab[4y], ptab[4y+1] nodes, its emission CSR at
// emit_off + ptab[4y+2] and emit_idx + ptab[4y+3]); each part's zerofier-group
// sums go to gpart[(y*3G + k) * N + r] for compose_parts_tail_kernel.
template <bool IN_LDS, int MODE>
__global__ __launch_bounds__(COMP_BLOCK) void compose_kernel(CompArgs a, const uint32_t *__restrict__ ptab,
                                                             uint64_t *__restrict__ gpart) {
    extern __shared__ uint64_t s_slots[];
    __shared__ uint64_t s_gsum[3 * COMP_MAX_GROUPS][COMP_BLOCK];  // per-lane zerofier-group sums
    const uint64_t *nodes = a.nodes;
    uint32_t num_nodes = a.num_nodes;
    const uint32_t *emit_off = a.emit_off, *emit_idx = a.emit_idx;
    if (ptab) {
        const uint32_t *pt = ptab + 4 * blockIdx.y;
        nodes = a.nodes + 2 * (uint64_t)pt[0];
        num_nodes = pt[1];
        emit_off = a.emit_off + pt[2];
        emit_idx = a.emit_idx + pt[3];
    }
    Slots sl;
    if (IN_LDS) {
        sl.base = s_slots + threadIdx.x;
        sl.ext = s_slots + (uint64_t)a.nbs * COMP_BLOCK + threadIdx.x;
        sl.stride = COMP_BLOCK;
    } else {
        // every (part, lane) has its own slots
        const uint64_t stride = a.total_threads * gridDim.y;
        const uint64_t gt = (uint64_t)blockIdx.y * a.total_threads + (uint64_t)blockIdx.x * COMP_BLOCK + threadIdx.x;
        sl.base = a.scratch + gt;
        sl.ext = a.scratch + (uint64_t)a.nbs * stride + gt;
        sl.stride = stride;
    }
    const uint64_t N = 1ull << a.log_N;
    for (uint64_t q = (uint64_t)blockIdx.x * COMP_BLOCK + threadIdx.x; q < N; q += a.total_threads) {
        const uint32_t r = (uint32_t)q;  // natural index; its LDE row is bitrev(r)
        uint64_t *gsum = &s_gsum[0][threadIdx.x];
        for (uint32_t k = 0; k < 3 * a.num_groups; k++) gsum[k * COMP_BLOCK] = 0;
        walk<MODE>(a, nodes, num_nodes, emit_off, emit_idx, r, sl, gsum);
        if (MODE == 1) continue;
        if (ptab) {
            for (uint32_t k = 0; k < 3 * a.num_groups; k++)
                gpart[((uint64_t)blockIdx.y * 3 * a.num_groups + k) * N + r] = gsum[k * COMP_BLOCK];
            continue;
        }
        fp3 gs[COMP_MAX_GROUPS];
#pragma unroll
        for (int g = 0; g < COMP_MAX_GROUPS; g++)
            gs[g] = (uint32_t)g < a.num_groups ? fp3{gsum[3 * g * COMP_BLOCK], gsum[(3 * g + 1) * COMP_BLOCK],
                                                    gsum[(3 * g + 2) * COMP_BLOCK]}
                                               : fp3{0, 0, 0};
        const fp3 h = composition_tail(a, r, gs);
        const uint64_t orow = a.natural_out ? r : bitrev(r, a.log_N);
        uint64_t *o = a.out + 3 * orow;
        o[0] = h.c0;
        o[1] = h.c1;
        o[2] = h.c2;
    }
}

// The group sums of the parts added up, then the composition tail.
__global__ __launch_bounds__(256) void compose_parts_tail_kernel(CompArgs a, const uint64_t *__restrict__ gpart,
                                                                 uint32_t parts) {
    const uint64_t N = 1ull << a.log_N;
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const uint32_t r = (uint32_t)q;
    const uint32_t G3 = 3 * a.num_groups;
    fp3 gs[COMP_MAX_GROUPS];
#pragma unroll
    for (int g = 0; g < COMP_MAX_GROUPS; g++) {
        gs[g] = fp3{0, 0, 0};
        if ((uint32_t)g < a.num_groups) {
            for (uint32_t p = 0; p < parts; p++) {
                const uint64_t *b = gpart + ((uint64_t)p * G3 + 3 * g) * N + r;
                gs[g] = f3_add(gs[g], fp3{b[0], b[N], b[2 * N]});
            }
        }
    }
    const fp3 h = composition_tail(a, r, gs);
    const uint64_t orow = a.natural_out ? r : bitrev(r, a.log_N);
    uint64_t *o = a.out + 3 * orow;
    o[0] = h.c0;
    o[1] = h.c1;
    o[2] = h.c2;
}

// binv[s * N + r] = 1 / (x_r - w^step_s) at every natural LDE index r
// (x_r = offset * w_N^r). Each thread takes STEP_INV_E indices r = t,
// t + T, ... and inverts the product of their D(r) = prod_s (x_r - w^step_s)
// with one base-field inversion (Montgomery's trick, prefixes in registers);
// then 1 / (x - c_s) = D^-1 * prod_{u != s} (x - c_u).
constexpr int STEP_INV_E = 16;
// thread t of T threads covering one table
__device__ __forceinline__ void step_inverse_body(const CompArgs &a, uint64_t *__restrict__ binv, uint64_t t,
                                                  uint64_t T) {
    const uint64_t N = 1ull << a.log_N;
    if (t >= N) return;
    const uint32_t S = a.num_steps;
    auto point = [&](uint64_t p) -> uint64_t {
        return mul(a.offset, tw32_table(a.tw_lo, a.tw_hi, (uint32_t)p << (32 - a.log_N)));
    };
    uint64_t pre[STEP_INV_E];
    uint64_t acc = 1;
#pragma unroll
    for (int j = 0; j < STEP_INV_E; j++) {
        const uint64_t p = t + (uint64_t)j * T;
        if (p < N) {
            const uint64_t x = point(p);
            uint64_t d = 1;
            for (uint32_t s = 0; s < S; s++) d = mul(d, sub(x, a.step_point[s]));
            acc = mul(acc, d);
        }
        pre[j] = acc;
    }
    uint64_t inv_acc = inv(acc);
#pragma unroll
    for (int j = STEP_INV_E - 1; j >= 0; j--) {
        const uint64_t p = t + (uint64_t)j * T;
        if (p >= N) continue;
        const uint64_t x = point(p);
        uint64_t d = 1;
        for (uint32_t s = 0; s < S; s++) d = mul(d, sub(x, a.step_point[s]));
        const uint64_t dinv = j ? mul(inv_acc, pre[j - 1]) : inv_acc;
        inv_acc = mul(inv_acc, d);
        for (uint32_t s = 0; s < S; s++) {
            uint64_t v = dinv;
            for (uint32_t u = 0; u < S; u++)
                if (u != s) v = mul(v, sub(x, a.step_point[u]));
            binv[((uint64_t)s << a.log_N) + p] = v;
        }
    }
}

__global__ __launch_bounds__(256) void step_inverse_kernel(const CompArgs a, uint64_t *__restrict__ binv) {
    step_inverse_body(a, binv, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, (uint64_t)gridDim.x * blockDim.x);
}

// blocks per table of the step-inverse pass (STEP_INV_E points per thread)
static uint64_t step_inverse_blocks(uint64_t N) { return ((N + STEP_INV_E - 1) / STEP_INV_E + 255) / 256; }

// The step inverses of several tables: blocks [prefix[t], prefix[t+1]) take
// table t (args[t].binv is its output).
__global__ __launch_bounds__(256) void step_inverse_batched_kernel(const CompArgs *__restrict__ args,
                                                                   const uint32_t *__restrict__ prefix,
                                                                   uint32_t count) {
    const uint32_t b = blockIdx.x;
    const uint32_t t = find_slot(prefix, count, b);
    const CompArgs &a = args[t];
    const uint64_t T = (uint64_t)(prefix[t + 1] - prefix[t]) * blockDim.x;
    step_inverse_body(a, const_cast<uint64_t *>(a.binv), (uint64_t)(b - prefix[t]) * blockDim.x + threadIdx.x, T);
}


// Column-major copy of a row-major LDE table (rows in bit-reversed order) in
// natural row order for the constraint walk: element (natural row r, c) of E
// words, stored at input row bitrev(r), moves to out[(c * N + r) * E + k].
// A wave's reads then cover consecutive natural rows of one column, and the
// frame rows r + k*blowup of neighbouring lanes overlap. 32 x 32-element tiles
// through LDS: a tile's 32 natural rows are 32 whole input row segments, its
// stores 32-row column segments.
template <int E>
__global__ __launch_bounds__(256) void transpose_kernel(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                        uint64_t N, uint32_t ncols) {
    constexpr int T = 32;
    __shared__ uint64_t tile[T][T * E + 1];
    const uint64_t r0 = (uint64_t)blockIdx.x * T;
    const uint32_t c0 = blockIdx.y * T;
    const uint64_t wpr = (uint64_t)ncols * E;  // words per input row
    const unsigned logN = (unsigned)__builtin_ctzll(N);
    for (int i = threadIdx.x; i < T * T * E; i += 256) {
        const int rr = i / (T * E), ww = i % (T * E);
        const uint64_t r = r0 + rr, w = (uint64_t)c0 * E + ww;
        if (r < N && w < wpr) tile[rr][ww] = in[(uint64_t)bitrev((uint32_t)r, logN) * wpr + w];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < T * T * E; i += 256) {
        const int k = i % E, rr = (i / E) % T, cc = i / (E * T);
        const uint64_t r = r0 + rr;
        const uint32_t c = c0 + cc;
        if (r < N && c < ncols) out[((uint64_t)c * N + r) * E + k] = tile[rr][cc * E + k];
    }
}

// The same copy in tiles that follow the bit reversal (N = 2^L, L >= 8): the
// natural rows r = (h << (L-3)) | (m << 5) | l (h < 8, l < 32, one middle m
// per block) sit at the input rows (bitrev5(l) << (L-5)) | (bitrev(m) << 3) |
// bitrev3(h), i.e. for each l a run of 8 consecutive input rows. A block moves
// one such 256-row tile for 16 words of the row: its reads are runs of 8
// row segments, its stores 32-row column runs.
template <int E>
__device__ __forceinline__ void transpose_bitrev_tile_body(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                           uint64_t N, uint32_t wpr, unsigned logN, uint32_t blk) {
    constexpr int WC = 16;
    __shared__ uint64_t tile[256][WC + 1];
    const uint32_t nch = (wpr + WC - 1) / WC;  // a tile's word chunks are adjacent blocks
    const uint32_t m = blk / nch;
    const uint32_t w0 = (blk % nch) * WC;
    const uint32_t wn = wpr - w0 < (uint32_t)WC ? wpr - w0 : (uint32_t)WC;
    const unsigned mid = logN - 8;
    const uint64_t mrev = mid ? (uint64_t)(__brev(m) >> (32 - mid)) : 0;
    // all 16 loads of a thread in flight before the LDS stores
    uint64_t v[WC];
#pragma unroll
    for (int q = 0; q < WC; q++) {
        const uint32_t i = threadIdx.x + 256 * q;
        const uint32_t ww = i % WC, j = (i / WC) % 8, l = i / (WC * 8);
        const uint64_t s = ((uint64_t)(__brev(l) >> 27) << (logN - 5)) | (mrev << 3) | j;
        v[q] = ww < wn ? in[s * wpr + w0 + ww] : 0;
    }
#pragma unroll
    for (int q = 0; q < WC; q++) {
        const uint32_t i = threadIdx.x + 256 * q;
        const uint32_t ww = i % WC, j = (i / WC) % 8, l = i / (WC * 8);
        tile[(__brev(j) >> 29) * 32 + l][ww] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < WC; q++) {
        const uint32_t i = threadIdx.x + 256 * q;
        const uint32_t l = i % 32, ww = (i / 32) % WC, h = i / (32 * WC);
        if (ww < wn) {
            const uint32_t w = w0 + ww;
            const uint64_t r = ((uint64_t)h << (logN - 3)) | ((uint64_t)m << 5) | l;
            out[((uint64_t)(w / E) * N + r) * E + (w % E)] = tile[h * 32 + l][ww];
        }
    }
}

template <int E>
__global__ __launch_bounds__(256) void transpose_bitrev_tile_kernel(const uint64_t *__restrict__ in,
                                                                    uint64_t *__restrict__ out, uint64_t N,
                                                                    uint32_t wpr, unsigned logN) {
    transpose_bitrev_tile_body<E>(in, out, N, wpr, logN, blockIdx.x);
}

// One table's copy for the batched transposes (N >= 256).
struct TransposeJob {
    const uint64_t *in;
    uint64_t *out;
    uint64_t N;
    uint32_t wpr, logN;
};
static uint32_t transpose_tile_blocks(uint64_t N, uint32_t wpr) { return (uint32_t)((N >> 8) * ((wpr + 15) / 16)); }

template <int E>
__global__ __launch_bounds__(256) void transpose_bitrev_tile_batched_kernel(const TransposeJob *__restrict__ jobs,
                                                                            const uint32_t *__restrict__ prefix,
                                                                            uint32_t count) {
    const uint32_t b = blockIdx.x;
    const uint32_t t = find_slot(prefix, count, b);
    const TransposeJob j = jobs[t];
    transpose_bitrev_tile_body<E>(j.in, j.out, j.N, j.wpr, j.logN, b - prefix[t]);
}

static void launch_transpose(const uint64_t *in, uint64_t *out, uint64_t N, uint32_t ncols, int E, hipStream_t st) {
    const unsigned logN = (unsigned)__builtin_ctzll(N);
    if (ncols == 0) return;
    if (logN >= 8) {
        const uint32_t wpr = ncols * (uint32_t)E;
        dim3 g((unsigned)((N >> 8) * ((wpr + 15) / 16)));
        if (E == 1)
            hipLaunchKernelGGL(transpose_bitrev_tile_kernel<1>, g, dim3(256), 0, st, in, out, N, wpr, logN);
        else
            hipLaunchKernelGGL(transpose_bitrev_tile_kernel<3>, g, dim3(256), 0, st, in, out, N, wpr, logN);
        return;
    }
    dim3 grid((unsigned)((N + 31) / 32), (ncols + 31) / 32);
    if (E == 1)
        hipLaunchKernelGGL(transpose_kernel<1>, grid, dim3(256), 0, st, in, out, N, ncols);
    else
        hipLaunchKernelGGL(transpose_kernel<3>, grid, dim3(256), 0, st, in, out, N, ncols);
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

// Composition parts for three or more parts (crypto/stark/src/prover.rs
// :1858-1900: interpolate_offset_fft, break_in_parts, each part evaluated on
// the LDE coset). coef holds the unnormalised inverse DFT of the N natural
// composition values on offset*<w_N>, rows in bit-reversed order, so row
// bitrev(k) = N * c_k * offset^k for the interpolant's coefficients c_k.
// Part i has coefficients c_{i + P t}; its values on the coset are the DFT of
// c_{i+Pt} * offset^t = coef[bitrev(i+Pt)] * (N^-1 offset^(t(1-P))) * offset^-i,
// written here as natural row t, words [3i, 3i+3) of a 3P-word row (zero
// where i + P t >= N). wt[t] = N^-1 * offset^(t(1-P)).
struct PartShifts {
    uint64_t v[8];  // offset^-i, i < P
};
__global__ __launch_bounds__(256) void parts_scatter_kernel(const uint64_t *__restrict__ coef, uint32_t log_N,
                                                            uint32_t P, const uint64_t *__restrict__ wt,
                                                            PartShifts gi, uint64_t *__restrict__ out) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t N = 1ull << log_N;
    if (idx >= N * P) return;
    const uint64_t t = idx / P, i = idx - t * P;
    const uint64_t k = i + (uint64_t)P * t;
    uint64_t *o = out + 3 * idx;
    if (k >= N) {
        o[0] = o[1] = o[2] = 0;
        return;
    }
    const fp3 c = ld3(coef + 3 * (uint64_t)bitrev((uint32_t)k, log_N));
    const fp3 v = f3_mul_base(c, mul(wt[t], gi.v[i]));
    o[0] = v.c0;
    o[1] = v.c1;
    o[2] = v.c2;
}

void launch_parts_scatter(const uint64_t *coef, unsigned log_N, uint32_t P, const uint64_t *wt,
                          const uint64_t *offset_inv_pows, uint64_t *out, hipStream_t st) {
    PartShifts gi{};
    for (uint32_t i = 0; i < P && i < 8; i++) gi.v[i] = offset_inv_pows[i];
    const uint64_t total = (1ull << log_N) * P;
    hipLaunchKernelGGL(parts_scatter_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, coef, log_N,
                       P, wt, gi, out);
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

// One table's composition (or per-constraint evaluation) ready to launch:
// its host-side arguments (device pointers filled by compose_bind), the blob
// of its small tables and their offsets in it.
struct ComposeJob {
    const lvm_program *pg = nullptr;
    const lvm_lde *main = nullptr, *aux = nullptr;
    int mode = 0;
    uint64_t *out = nullptr;
    size_t N = 0;
    unsigned logN = 0;
    CompArgs a{};
    std::vector<uint8_t> blob;
    std::vector<uint32_t> root_node;
    const uint32_t *root_group = nullptr;
    size_t o_nodes = 0, o_bc = 0, o_ec = 0, o_rap = 0, o_al = 0, o_to = 0, o_roots = 0, o_eoff = 0, o_eidx = 0;
    size_t o_beta = 0, o_rg = 0, o_z = 0, o_bcol = 0, o_baux = 0, o_bval = 0, o_bbeta = 0, o_bpcol = 0, o_bpflag = 0,
           o_bpbeta = 0;
};

// Validation and host-side arguments of one table's composition (mode 0)
// or per-constraint evaluation (mode 1); no device work.
int compose_prepare(const lvm_program *pg, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                    const lvm_composition *cp, int mode, int natural_out, uint64_t *out, ComposeJob &j) {
    if (!pg || !main || !out) return fail(LVM_ERR_ARG, "null argument");
    if (mode == 0 && !cp) return fail(LVM_ERR_ARG, "null composition");
    const size_t N = main->N;
    if (!is_pow2(N) || log2u(N) > 32) return fail(LVM_ERR_ARG, "bad LDE size");
    if (aux && aux->N != N) return fail(LVM_ERR_ARG, "main and aux LDE sizes differ");
    if (aux && aux->w % 3) return fail(LVM_ERR_ARG, "aux table rows must hold whole Fp3 elements");
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
    (void)max_off;  // the frame reads stay inside the table for every offset (rows wrap mod N)
    for (size_t c = 0; c < pg->num_roots; c++) {
        const uint32_t rs = (uint32_t)pg->roots[c];
        if ((rs >> 31) ? (rs & 0x7FFFFFFFu) >= pg->num_ext_slots : rs >= pg->num_base_slots)
            return fail(LVM_ERR_ARG, "root slot out of range");
    }
    const unsigned logN = log2u(N);
    j = ComposeJob{};
    j.pg = pg;
    j.main = main;
    j.aux = aux;
    j.mode = mode;
    j.out = out;
    j.N = N;
    j.logN = logN;
    j.root_group = mode == 0 ? cp->root_group : nullptr;
    CompArgs &a = j.a;
    BoundarySteps jb;
    std::vector<uint8_t> &blob = j.blob;
    std::vector<uint32_t> roots32(pg->num_roots);
    for (size_t c = 0; c < pg->num_roots; c++) roots32[c] = (uint32_t)pg->roots[c];
    // node computing each constraint: given, or (pinned root slots) the last
    // node writing the root's slot
    std::vector<uint32_t> &root_node = j.root_node;
    if (!root_nodes_of(pg, root_node)) return fail(LVM_ERR_ARG, "root node index out of range or no node writes a root slot");
    std::vector<uint32_t> emit_off(pg->num_nodes + 1, 0), emit_idx(pg->num_roots);
    for (size_t c = 0; c < pg->num_roots; c++) emit_off[root_node[c] + 1]++;
    for (size_t k = 0; k < pg->num_nodes; k++) emit_off[k + 1] += emit_off[k];
    {
        std::vector<uint32_t> fill(emit_off.begin(), emit_off.end() - 1);
        for (size_t c = 0; c < pg->num_roots; c++) emit_idx[fill[root_node[c]]++] = (uint32_t)c;
    }
    j.o_nodes = put(blob, pg->nodes, 2 * pg->num_nodes);
    j.o_bc = put(blob, pg->base_consts, pg->num_base_consts);
    j.o_ec = put(blob, pg->ext_consts, 3 * pg->num_ext_consts);
    j.o_rap = put(blob, pg->rap_challenges, 3 * pg->num_rap);
    j.o_al = put(blob, pg->alpha_powers, 3 * pg->num_alpha);
    j.o_to = put(blob, pg->table_offset, pg->table_offset ? 3 : 0);
    j.o_roots = put(blob, roots32.data(), roots32.size());
    j.o_eoff = put(blob, emit_off.data(), emit_off.size());
    j.o_eidx = put(blob, emit_idx.data(), emit_idx.size());
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
        for (size_t b = 0; b < cp->num_boundary; b++)
            if (cp->b_is_aux[b] ? (!aux || cp->b_col[b] >= ma) : cp->b_col[b] >= main->w)
                return fail(LVM_ERR_ARG, "boundary column out of range");
        std::vector<uint32_t> order;
        if (!boundary_steps_of(cp->b_col, cp->b_is_aux, cp->b_step, cp->num_boundary, n, jb, &order))
            return fail(LVM_ERR_ARG, "at most 16 distinct boundary steps");
        a.num_steps = jb.num_steps;
        for (uint32_t k = 0; k < jb.num_steps; k++) a.step_point[k] = pow(wn, jb.step_row[k]);
        a.num_boundary = (uint32_t)cp->num_boundary;
        for (uint32_t k = 0; k <= jb.num_steps; k++) a.step_off[k] = jb.step_off[k];
        std::vector<uint32_t> bcol(order.size()), baux(order.size());
        std::vector<uint64_t> bval(3 * order.size()), bbeta(3 * order.size());
        for (size_t i = 0; i < order.size(); i++) {
            const uint32_t b = order[i];
            bcol[i] = cp->b_col[b];
            baux[i] = cp->b_is_aux[b] ? 1u : 0u;
            for (int k = 0; k < 3; k++) {
                bval[3 * i + k] = cp->b_value[3 * b + k];
                bbeta[3 * i + k] = cp->b_beta[3 * b + k];
            }
        }
        for (uint32_t k = 0; k < a.num_steps; k++) {
            fp3 c{0, 0, 0};
            for (uint32_t i = a.step_off[k]; i < a.step_off[k + 1]; i++)
                c = f3_add(c, f3_mul(fp3{bbeta[3 * i], bbeta[3 * i + 1], bbeta[3 * i + 2]},
                                     fp3{bval[3 * i], bval[3 * i + 1], bval[3 * i + 2]}));
            a.step_const[3 * k] = c.c0;
            a.step_const[3 * k + 1] = c.c1;
            a.step_const[3 * k + 2] = c.c2;
        }
        j.o_beta = put(blob, cp->beta, 3 * pg->num_roots);
        j.o_rg = put(blob, cp->root_group, pg->num_roots);
        j.o_z = put(blob, zc.data(), B);
        j.o_bcol = put(blob, bcol.data(), bcol.size());
        j.o_baux = put(blob, baux.data(), baux.size());
        j.o_bval = put(blob, bval.data(), bval.size());
        j.o_bbeta = put(blob, bbeta.data(), bbeta.size());
        // main-column boundaries per pair of steps: one entry per column, the
        // betas of its boundaries at each step of the pair summed
        std::vector<uint32_t> pcol, pflag;
        std::vector<uint64_t> pbeta;
        for (uint32_t k = 0; 2 * k < a.num_steps; k++) {
            a.bp_off[k] = (uint32_t)pcol.size();
            const uint32_t s0 = 2 * k, s1 = std::min(a.num_steps, s0 + 2);
            std::vector<uint32_t> cols;
            for (uint32_t i = a.step_off[s0]; i < a.step_off[s1]; i++)
                if (!baux[i]) cols.push_back(bcol[i]);
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            for (uint32_t c : cols) {
                uint32_t fl = 0;
                fp3 bsum[2] = {fp3{0, 0, 0}, fp3{0, 0, 0}};
                for (uint32_t st = s0; st < s1; st++)
                    for (uint32_t i = a.step_off[st]; i < a.step_off[st + 1]; i++)
                        if (!baux[i] && bcol[i] == c) {
                            fl |= 1u << (st - s0);
                            bsum[st - s0] = f3_add(bsum[st - s0], fp3{bbeta[3 * i], bbeta[3 * i + 1], bbeta[3 * i + 2]});
                        }
                pcol.push_back(c);
                pflag.push_back(fl);
                for (int j = 0; j < 2; j++) {
                    pbeta.push_back(bsum[j].c0);
                    pbeta.push_back(bsum[j].c1);
                    pbeta.push_back(bsum[j].c2);
                }
            }
            a.bp_off[k + 1] = (uint32_t)pcol.size();
        }
        if (pcol.empty()) {
            pcol.push_back(0);
            pflag.push_back(0);
            pbeta.resize(6, 0);
        }
        j.o_bpcol = put(blob, pcol.data(), pcol.size());
        j.o_bpflag = put(blob, pflag.data(), pflag.size());
        j.o_bpbeta = put(blob, pbeta.data(), pbeta.size());
    }
    a.log_N = logN;
    a.next_step = next_step;
    a.nbs = (uint32_t)pg->num_base_slots;
    a.nes = (uint32_t)pg->num_ext_slots;
    a.num_nodes = (uint32_t)pg->num_nodes;
    a.num_roots = (uint32_t)pg->num_roots;
    a.out = out;
    a.natural_out = natural_out ? 1u : 0u;
    return LVM_OK;
}

// Device pointers of a prepared job: its blob uploaded at db, the column-major
// copies of its tables, its step inverses.
void compose_bind(ComposeJob &j, const uint8_t *db, const uint64_t *cm_main, const uint64_t *cm_aux,
                  const uint64_t *binv, const Tables &tb) {
    CompArgs &a = j.a;
    a.nodes = (const uint64_t *)(db + j.o_nodes);
    a.base_consts = (const uint64_t *)(db + j.o_bc);
    a.ext_consts = (const uint64_t *)(db + j.o_ec);
    a.rap = (const uint64_t *)(db + j.o_rap);
    a.alpha = (const uint64_t *)(db + j.o_al);
    a.toff = (const uint64_t *)(db + j.o_to);
    a.roots = (const uint32_t *)(db + j.o_roots);
    a.emit_off = (const uint32_t *)(db + j.o_eoff);
    a.emit_idx = (const uint32_t *)(db + j.o_eidx);
    a.main = cm_main;
    a.aux = cm_aux;
    a.tw_lo = tb.fwd_lo;
    a.tw_hi = tb.fwd_hi;
    if (j.mode == 0) {
        a.beta = (const uint64_t *)(db + j.o_beta);
        a.root_group = (const uint32_t *)(db + j.o_rg);
        a.zcyc = (const uint64_t *)(db + j.o_z);
        a.b_col = (const uint32_t *)(db + j.o_bcol);
        a.b_is_aux = (const uint32_t *)(db + j.o_baux);
        a.b_value = (const uint64_t *)(db + j.o_bval);
        a.b_beta = (const uint64_t *)(db + j.o_bbeta);
        a.bp_col = (const uint32_t *)(db + j.o_bpcol);
        a.bp_flags = (const uint32_t *)(db + j.o_bpflag);
        a.bp_beta = (const uint64_t *)(db + j.o_bpbeta);
    }
    a.binv = const_cast<uint64_t *>(binv);
}

// A program's constraints split into `parts` runs, in the order of the nodes
// computing them. A part's node list is the subsequence of nodes its
// constraints depend on, each slot read traced to the last node writing that
// slot before it: the subsequence keeps every write a kept read depends on
// and drops only writes, so the program's slot assignment stays valid. Its
// emission CSR over that list holds only its own constraints. ptab: per part
// its node offset, node count, emit_off offset and emit_idx offset.
static void split_program(const lvm_program *pg, const std::vector<uint32_t> &root_node, uint32_t parts,
                          std::vector<uint64_t> &nodes, std::vector<uint32_t> &eoff, std::vector<uint32_t> &eidx,
                          std::vector<uint32_t> &ptab) {
    const uint32_t K = (uint32_t)pg->num_nodes, R = (uint32_t)pg->num_roots;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    std::vector<uint32_t> dep(2 * (size_t)K, NONE);
    std::vector<uint32_t> wb(pg->num_base_slots, NONE), we(pg->num_ext_slots, NONE);
    auto src = [&](uint32_t enc) -> uint32_t {
        const uint32_t kind = enc >> 29, x = enc & 0x1FFFFFFFu;
        if (kind == K_BASE_SLOT) return x < wb.size() ? wb[x] : NONE;
        if (kind == K_EXT_SLOT) return x < we.size() ? we[x] : NONE;
        return NONE;
    };
    for (uint32_t k = 0; k < K; k++) {
        const uint64_t w0 = pg->nodes[2 * k], w1 = pg->nodes[2 * k + 1];
        const uint32_t op = (uint32_t)w0, x = (uint32_t)(w0 >> 32), y = (uint32_t)w1;
        if (op == OP_ADD || op == OP_SUB || op == OP_MUL) {
            dep[2 * k] = src(x);
            dep[2 * k + 1] = src(y);
        } else if (op > OP_MUL) {  // OP_NEG, OP_EMBED: one operand
            dep[2 * k] = src(x);
        }
        const uint32_t res = (uint32_t)(w1 >> 32), slot = res & 0x7FFFFFFFu;
        if (res >> 31) {
            if (slot < we.size()) we[slot] = k;
        } else if (slot < wb.size()) {
            wb[slot] = k;
        }
    }
    std::vector<uint32_t> order(R);
    for (uint32_t c = 0; c < R; c++) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t u, uint32_t v) { return root_node[u] < root_node[v]; });
    std::vector<uint32_t> part_of(R), by_node_off(K + 1, 0), by_node(R);
    for (uint32_t i = 0; i < R; i++) part_of[order[i]] = (uint32_t)((uint64_t)i * parts / R);
    for (uint32_t c = 0; c < R; c++) by_node_off[root_node[c] + 1]++;
    for (uint32_t k = 0; k < K; k++) by_node_off[k + 1] += by_node_off[k];
    {
        std::vector<uint32_t> fill(by_node_off.begin(), by_node_off.end() - 1);
        for (uint32_t i = 0; i < R; i++) by_node[fill[root_node[order[i]]]++] = order[i];
    }
    std::vector<uint32_t> stamp(K, NONE), stack;
    for (uint32_t p = 0; p < parts; p++) {
        for (uint32_t c = 0; c < R; c++)
            if (part_of[c] == p) stack.push_back(root_node[c]);
        while (!stack.empty()) {
            const uint32_t k = stack.back();
            stack.pop_back();
            if (stamp[k] == p) continue;
            stamp[k] = p;
            for (int i = 0; i < 2; i++)
                if (dep[2 * k + i] != NONE && stamp[dep[2 * k + i]] != p) stack.push_back(dep[2 * k + i]);
        }
        ptab.push_back((uint32_t)(nodes.size() / 2));
        const size_t count_at = ptab.size();
        ptab.push_back(0);
        ptab.push_back((uint32_t)eoff.size());
        ptab.push_back((uint32_t)eidx.size());
        const size_t e0 = eidx.size();
        uint32_t count = 0;
        for (uint32_t k = 0; k < K; k++) {
            if (stamp[k] != p) continue;
            nodes.push_back(pg->nodes[2 * k]);
            nodes.push_back(pg->nodes[2 * k + 1]);
            eoff.push_back((uint32_t)(eidx.size() - e0));
            for (uint32_t q = by_node_off[k]; q < by_node_off[k + 1]; q++)
                if (part_of[by_node[q]] == p) eidx.push_back(by_node[q]);
            count++;
        }
        eoff.push_back((uint32_t)(eidx.size() - e0));
        ptab[count_at] = count;
    }
}

// Shared by the composition and the per-constraint entry points.
int compose_impl(const lvm_program *pg, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                 const lvm_composition *cp, int mode, int natural_out, uint64_t *out) {
    ComposeJob j;
    int rc = compose_prepare(pg, main, aux, next_step, cp, mode, natural_out, out, j);
    if (rc != LVM_OK) return rc;
    const size_t N = j.N;
    CompArgs &a = j.a;
    std::vector<uint8_t> &blob = j.blob;
    const std::vector<uint32_t> &root_node = j.root_node;

    Device *dv = device_state();
    if (!dv) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = lib_stream();
    const size_t slot_bytes = 8 * (size_t)pg->num_base_slots + 24 * (size_t)pg->num_ext_slots;
    const bool in_lds = slot_bytes * COMP_BLOCK <= 64 * 1024;
    const uint64_t max_blocks = 256 * 16;
    const uint64_t blocks = std::min<uint64_t>((N + COMP_BLOCK - 1) / COMP_BLOCK, max_blocks);
    a.total_threads = blocks * COMP_BLOCK;
    // the compiled form of the program when there is one (compiled on first use)
    const bool jit = jit_enabled() && jit_prepare(pg, root_node, mode == 0 ? cp->root_group : nullptr, mode);
    // An interpreted composition over few points (a short table of a wide
    // chip) walks its program in parts side by side, about 2^17 lanes in all.
    uint32_t parts = 1;
    if (mode == 0 && !jit && pg->num_roots >= 2 && pg->num_nodes >= 256) {
        const uint64_t want = (1ull << 17) / a.total_threads;
        parts = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(want, pg->num_roots), 64);
        if (parts < 2) parts = 1;
    }
    size_t o_pnodes = 0, o_peoff = 0, o_peidx = 0, o_ptab = 0;
    if (parts > 1) {
        std::vector<uint64_t> pnodes;
        std::vector<uint32_t> peoff, peidx, ptab;
        split_program(pg, root_node, parts, pnodes, peoff, peidx, ptab);
        o_pnodes = put(blob, pnodes.data(), pnodes.size());
        o_peoff = put(blob, peoff.data(), peoff.size());
        o_peidx = put(blob, peidx.data(), peidx.size());
        o_ptab = put(blob, ptab.data(), ptab.size());
    }
    uint8_t *db = nullptr;
    uint64_t *scratch = nullptr, *cm_main = nullptr, *cm_aux = nullptr, *binv = nullptr, *gpart = nullptr;
    auto release = [&]() {
        lvm_free(db);
        lvm_free(scratch);
        lvm_free(cm_main);
        lvm_free(cm_aux);
        lvm_free(binv);
        lvm_free(gpart);
    };
    if (lvm_malloc(&db, blob.size()) != hipSuccess ||
        (!in_lds && lvm_malloc(&scratch, slot_bytes * a.total_threads * parts) != hipSuccess) ||
        (parts > 1 && lvm_malloc(&gpart, (size_t)parts * 3 * a.num_groups * N * 8) != hipSuccess) ||
        lvm_malloc(&cm_main, N * main->w * 8) != hipSuccess ||
        (aux && lvm_malloc(&cm_aux, N * aux->w * 8) != hipSuccess) ||
        (mode == 0 && a.num_steps && lvm_malloc(&binv, N * a.num_steps * 8) != hipSuccess)) {
        release();
        return fail(LVM_ERR_NOMEM, "device allocation for the constraint program failed");
    }
    {
        ProfScope ps("composition_transpose", st);
        launch_transpose(main->lde, cm_main, N, (uint32_t)main->w, 1, st);
        if (aux) launch_transpose(aux->lde, cm_aux, N, (uint32_t)(aux->w / 3), 3, st);
    }
    StageScope ss(st);
    hipError_t e = stage_upload(db, blob.data(), blob.size(), st);
    compose_bind(j, db, cm_main, cm_aux, binv, dv->tb);
    if (mode == 0 && a.num_steps) {
        ProfScope ps("composition_step_inverse", st);
        const uint64_t threads = (N + STEP_INV_E - 1) / STEP_INV_E;
        hipLaunchKernelGGL(step_inverse_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, a, binv);
    }
    a.scratch = scratch;
    const size_t lds = in_lds ? slot_bytes * COMP_BLOCK : 0;
    if (jit) {
        ProfScope ps(mode == 0 ? "composition_compiled" : "constraint_evals_compiled", st);
        if (e == hipSuccess && !jit_launch(pg, root_node, mode == 0 ? cp->root_group : nullptr, mode, a, blocks, st))
            e = hipErrorLaunchFailure;
    } else {
        ProfScope ps(mode == 0 ? "composition" : "constraint_evals", st);
        const uint32_t *ptab = nullptr;
        if (parts > 1) {
            a.nodes = (const uint64_t *)(db + o_pnodes);
            a.emit_off = (const uint32_t *)(db + o_peoff);
            a.emit_idx = (const uint32_t *)(db + o_peidx);
            ptab = (const uint32_t *)(db + o_ptab);
        }
        const dim3 grid((unsigned)blocks, parts);
        if (in_lds) {
            if (mode == 0)
                hipLaunchKernelGGL((compose_kernel<true, 0>), grid, dim3(COMP_BLOCK), lds, st, a, ptab, gpart);
            else
                hipLaunchKernelGGL((compose_kernel<true, 1>), grid, dim3(COMP_BLOCK), lds, st, a, ptab, gpart);
        } else {
            if (mode == 0)
                hipLaunchKernelGGL((compose_kernel<false, 0>), grid, dim3(COMP_BLOCK), 0, st, a, ptab, gpart);
            else
                hipLaunchKernelGGL((compose_kernel<false, 1>), grid, dim3(COMP_BLOCK), 0, st, a, ptab, gpart);
        }
        if (parts > 1)
            hipLaunchKernelGGL(compose_parts_tail_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, a,
                               (const uint64_t *)gpart, parts);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = stage_sync(st);
    release();
    if (e != hipSuccess) return fail(LVM_ERR_DEVICE, std::string("constraint evaluation failed: ") + hipGetErrorString(e));
    return LVM_OK;
}

// The compositions of several tables in shared launches: the column-major
// copies in one launch per element width (tables of at least 256 LDE rows),
// the boundary step inverses in one launch, and every table's program
// compiled into one kernel launched once over all tables' rows. Tables are
// evaluated one by one (compose_impl) when the run-time compiler is off or
// does not take one of the programs.
int compose_batch_impl(const lvm_program *const *pgs, const lvm_lde *const *mains, const lvm_lde *const *auxs,
                       const size_t *next_steps, const lvm_composition *const *cps, size_t count, int natural_out,
                       uint64_t *const *outs) {
    if (count == 0) return LVM_OK;
    if (!pgs || !mains || !auxs || !next_steps || !cps || !outs) return fail(LVM_ERR_ARG, "null argument");
    std::vector<ComposeJob> jobs(count);
    for (size_t t = 0; t < count; t++) {
        if (!cps[t]) return fail(LVM_ERR_ARG, "null composition");
        int rc = compose_prepare(pgs[t], mains[t], auxs[t], next_steps[t], cps[t], 0, natural_out, outs[t], jobs[t]);
        if (rc != LVM_OK) return rc;
    }
    std::vector<const lvm_program *> vp(count);
    std::vector<std::vector<uint32_t>> vr(count);
    std::vector<const uint32_t *> vg(count);
    for (size_t t = 0; t < count; t++) {
        vp[t] = jobs[t].pg;
        vr[t] = jobs[t].root_node;
        vg[t] = jobs[t].root_group;
    }
    if (!jit_enabled() || !jit_prepare_batch(vp, vr, vg)) {
        for (size_t t = 0; t < count; t++) {
            int rc = compose_impl(pgs[t], mains[t], auxs[t], next_steps[t], cps[t], 0, natural_out, outs[t]);
            if (rc != LVM_OK) return rc;
        }
        return LVM_OK;
    }
    Device *dv = device_state();
    if (!dv) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = lib_stream();
    std::vector<uint64_t *> cm_main(count, nullptr), cm_aux(count, nullptr), binv(count, nullptr);
    uint8_t *arena = nullptr;
    auto release = [&]() {
        for (size_t t = 0; t < count; t++) {
            lvm_free(cm_main[t]);
            lvm_free(cm_aux[t]);
            lvm_free(binv[t]);
        }
        lvm_free(arena);
    };
    for (size_t t = 0; t < count; t++) {
        const ComposeJob &j = jobs[t];
        if (lvm_malloc(&cm_main[t], j.N * j.main->w * 8) != hipSuccess ||
            (j.aux && lvm_malloc(&cm_aux[t], j.N * j.aux->w * 8) != hipSuccess) ||
            (j.a.num_steps && lvm_malloc(&binv[t], j.N * j.a.num_steps * 8) != hipSuccess)) {
            release();
            return fail(LVM_ERR_NOMEM, "device allocation for the constraint programs failed");
        }
    }
    // arena: every table's blob, then the launch descriptors
    std::vector<uint8_t> host;
    auto place = [&](const void *p, size_t bytes) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        if (bytes) std::memcpy(host.data() + off, p, bytes);
        return off;
    };
    std::vector<size_t> blob_off(count);
    for (size_t t = 0; t < count; t++) blob_off[t] = place(jobs[t].blob.data(), jobs[t].blob.size());
    std::vector<TransposeJob> tj[2];
    std::vector<uint32_t> tp[2] = {std::vector<uint32_t>(1, 0), std::vector<uint32_t>(1, 0)};
    std::vector<uint32_t> sp(1, 0), cpfx(1, 0);
    uint64_t s_total = 0, c_total = 0;
    for (size_t t = 0; t < count; t++) {
        const ComposeJob &j = jobs[t];
        if (j.logN >= 8) {
            tj[0].push_back({j.main->lde, cm_main[t], j.N, (uint32_t)j.main->w, j.logN});
            tp[0].push_back(tp[0].back() + transpose_tile_blocks(j.N, (uint32_t)j.main->w));
            if (j.aux) {
                tj[1].push_back({j.aux->lde, cm_aux[t], j.N, (uint32_t)j.aux->w, j.logN});
                tp[1].push_back(tp[1].back() + transpose_tile_blocks(j.N, (uint32_t)j.aux->w));
            }
        }
        s_total += j.a.num_steps ? step_inverse_blocks(j.N) : 0;
        c_total += (j.N + COMP_BLOCK - 1) / COMP_BLOCK;
        if (s_total >= (1ull << 31) || c_total >= (1ull << 31) || tp[0].back() >= (1u << 31) || tp[1].back() >= (1u << 31)) {
            release();
            return fail(LVM_ERR_ARG, "batched composition exceeds 2^31 blocks");
        }
        sp.push_back((uint32_t)s_total);
        cpfx.push_back((uint32_t)c_total);
    }
    // the CompArgs array needs the arena's address: allocate, bind, then place
    const size_t args_bytes = count * sizeof(CompArgs);
    size_t o_args = (host.size() + 255) & ~(size_t)255;
    const size_t total = o_args + args_bytes + 4 * (sp.size() + cpfx.size() + tp[0].size() + tp[1].size()) +
                         sizeof(TransposeJob) * (tj[0].size() + tj[1].size()) + 8 * 256;
    if (lvm_malloc(&arena, total) != hipSuccess) {
        release();
        return fail(LVM_ERR_NOMEM, "device allocation for the composition descriptors failed");
    }
    std::vector<CompArgs> args(count);
    for (size_t t = 0; t < count; t++) {
        compose_bind(jobs[t], arena + blob_off[t], cm_main[t], cm_aux[t], binv[t], dv->tb);
        args[t] = jobs[t].a;
    }
    o_args = place(args.data(), args_bytes);
    const size_t o_sp = place(sp.data(), 4 * sp.size()), o_cp = place(cpfx.data(), 4 * cpfx.size());
    size_t o_tj[2], o_tp[2];
    for (int e = 0; e < 2; e++) {
        o_tj[e] = place(tj[e].data(), sizeof(TransposeJob) * tj[e].size());
        o_tp[e] = place(tp[e].data(), 4 * tp[e].size());
    }
    if (host.size() > total) {
        release();
        return fail(LVM_ERR_STATE, "composition descriptor layout overflow");
    }
    StageScope ss(st);
    hipError_t e = stage_upload(arena, host.data(), host.size(), st);
    if (e == hipSuccess) {
        ProfScope ps("composition_transpose", st);
        if (!tj[0].empty())
            hipLaunchKernelGGL(transpose_bitrev_tile_batched_kernel<1>, dim3(tp[0].back()), dim3(256), 0, st,
                               (const TransposeJob *)(arena + o_tj[0]), (const uint32_t *)(arena + o_tp[0]),
                               (uint32_t)tj[0].size());
        if (!tj[1].empty())
            hipLaunchKernelGGL(transpose_bitrev_tile_batched_kernel<3>, dim3(tp[1].back()), dim3(256), 0, st,
                               (const TransposeJob *)(arena + o_tj[1]), (const uint32_t *)(arena + o_tp[1]),
                               (uint32_t)tj[1].size());
        for (size_t t = 0; t < count; t++) {
            const ComposeJob &j = jobs[t];
            if (j.logN >= 8) continue;
            launch_transpose(j.main->lde, cm_main[t], j.N, (uint32_t)j.main->w, 1, st);
            if (j.aux) launch_transpose(j.aux->lde, cm_aux[t], j.N, (uint32_t)(j.aux->w / 3), 3, st);
        }
    }
    if (e == hipSuccess && s_total) {
        ProfScope ps("composition_step_inverse", st);
        hipLaunchKernelGGL(step_inverse_batched_kernel, dim3((unsigned)s_total), dim3(256), 0, st,
                           (const CompArgs *)(arena + o_args), (const uint32_t *)(arena + o_sp), (uint32_t)count);
    }
    if (e == hipSuccess) {
        ProfScope ps("composition_compiled", st);
        if (!jit_launch_batch(vp, vr, vg, (const CompArgs *)(arena + o_args), (const uint32_t *)(arena + o_cp),
                              (uint32_t)count, c_total, st))
            e = hipErrorLaunchFailure;
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = stage_sync(st);
    release();
    if (e != hipSuccess)
        return fail(LVM_ERR_DEVICE, std::string("batched composition failed: ") + hipGetErrorString(e));
    return LVM_OK;
}

}  // namespace lvm

extern "C" {

int lvm_eval_composition_batch(const lvm_program *const *programs, const lvm_lde *const *mains,
                               const lvm_lde *const *auxs, const size_t *next_steps,
                               const lvm_composition *const *comps, size_t count, int natural_order,
                               uint64_t *const *dev_h_out) {
    return lvm::compose_batch_impl(programs, mains, auxs, next_steps, comps, count, natural_order, dev_h_out);
}

int lvm_eval_composition(const lvm_program *program, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                         const lvm_composition *comp, int natural_order, uint64_t *dev_h_out) {
    return lvm::compose_impl(program, main, aux, next_step, comp, 0, natural_order, dev_h_out);
}

int lvm_eval_constraints(const lvm_program *program, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                         uint64_t *dev_evals_out) {
    return lvm::compose_impl(program, main, aux, next_step, nullptr, 1, 1, dev_evals_out);
}

}  // extern "C"
