// Batched row-major Goldilocks NTT passes for gfx950.
//
// Layout: a table is row-major (row r holds the `m` columns of trace row r,
// `pitch` u64 apart), exactly the reference's host layout
// (crypto/math/src/polynomial.rs:527-597, crypto/math-cuda/src/lde.rs:674).
// Every column is transformed independently; a workgroup tile covers 16
// adjacent columns, so each row touch is one 128-byte segment and all
// twiddles of a tile are computed once and shared by its 16 columns.
//
// A size-2^L DIF transform is split into passes. A pass with sub-transform
// size R = R1*R2 and row stride S works on blocks of M = R*S rows:
//   Y[k1] = sum_t x[blk*M + t*S + j] * w_R^(t*k1)          (size-R DFT over t)
//   out[blk*M + bitrev_R(k1)*S + j] = Y[k1] * w_M^(j*k1)   (post twiddle)
// after which every M/R... block of S contiguous rows is an independent
// size-S DFT (the four-step identity). Chaining passes down to S = 1 leaves
// the full DFT in bit-reversed row order. Inside a pass the first log2(R1)
// radix-2 stages run in registers on elements spaced R2 apart, the tile is
// exchanged through LDS, and the last log2(R2) stages run in registers on
// R2 contiguous elements.
//
// Inside a pass the size-R DFT uses the four-step form: round 1 is a size-R1
// DFT over t = t0 + R2*i followed by the inner twiddle w_R^(t0*k_a), round 2
// a size-R2 DFT over t0. Both small DFTs (R1, R2 <= 16) are radix-2 DIF
// whose twiddles are roots of unity of order <= 16, all of the form +-2^s
// in this field (forward w_16 = 2^156, inverse w_16 = 2^36), so they are
// shifts and a reduction, chosen at compile time; a negative sign swaps the
// butterfly's subtraction. Values are kept canonical from the load on, which
// makes each add and sub one conditional correction (field.hpp cadd/csub).
//
// Optional input remaps make the same kernel serve the coset LDE:
//  - input rows may be read at bitrev(logical row) (the coefficient buffer
//    produced by the inverse transform is in bit-reversed order);
//  - element j may be pre-multiplied by weights[j] * w_N^(coset*j)
//    (LDE weights n^-1*g^j, crypto/stark/src/prover.rs:451-467, times the
//    coset shift that turns one size-N transform into B size-n ones);
//  - output rows may be offset per coset by bitrev_B(coset)*n.
#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

// w_{2^32}^e via the two-level table (canonical)
__device__ __forceinline__ uint64_t tw32(const uint64_t *lo, const uint64_t *hi, uint32_t e) {
    return cmul(LVM_GLOBAL_CONST(lo)[e & 0xFFFFu], LVM_GLOBAL_CONST(hi)[e >> 16]);
}

// bit reversal of i over `bits` bits, for compile-time indices
constexpr uint32_t bitrev_small(uint32_t i, int bits) {
    uint32_t r = 0;
    for (int b = 0; b < bits; b++) r |= ((i >> b) & 1u) << (bits - 1 - b);
    return r;
}

// log2 of w_R as a power of two (w_R = w_16^(16/R)), forward and inverse
constexpr int log2_root(int R, bool inverse) { return ((inverse ? 36 : 156) * (16 / R)) % 192; }

constexpr uint64_t cmod_mul(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) % P); }
constexpr uint64_t cmod_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = cmod_mul(r, a);
        a = cmod_mul(a, a);
        e >>= 1;
    }
    return r;
}
// w_16 = w_{2^32}^(2^28) of the field's root of unity, and its inverse
static_assert(cmod_pow(1753635133440165772ull, 1ull << 28) == cmod_pow(2, 156), "forward w_16 is 2^156");
static_assert(cmod_mul(cmod_pow(2, 156), cmod_pow(2, 36)) == 1, "inverse w_16 is 2^36");

// DIF butterfly (u, v) -> (u + v, (u - v) * 2^T); where the cheaper form of
// the power of two computes -2^T, the subtraction is swapped instead
template <int T>
__device__ __forceinline__ void bfly_pow2(uint64_t &u, uint64_t &v) {
    const uint64_t a = u, b = v;
    u = cadd(a, b);
    if constexpr (T % 192 == 0)
        v = csub(a, b);
    else if constexpr (pow2_negates<T>())
        v = cmul_pow2_form<T>(csub(b, a));
    else
        v = cmul_pow2_form<T>(csub(a, b));
}

template <int R, int LOGW, int H, int I>
__device__ __forceinline__ void dif_stage(uint64_t *x) {
    if constexpr (I < R) {
        if constexpr ((I & H) == 0) bfly_pow2<(LOGW * ((I & (H - 1)) * (R / (2 * H)))) % 192>(x[I], x[I + H]);
        dif_stage<R, LOGW, H, I + 1>(x);
    }
}

// In-register radix-2 DIF DFT of size R <= 16 with root 2^LOGW: natural
// input order, bit-reversed output order, canonical values.
template <int R, int LOGW, int H = R / 2>
__device__ __forceinline__ void dif_small(uint64_t *x) {
    if constexpr (H >= 1) {
        dif_stage<R, LOGW, H, 0>(x);
        dif_small<R, LOGW, H / 2>(x);
    }
}

template <int R>
__device__ __forceinline__ void dft_small(uint64_t *x, bool inverse) {
    if (inverse)
        dif_small<R, log2_root(R, true)>(x);
    else
        dif_small<R, log2_root(R, false)>(x);
}

// Body of one pass for block (bx, by, bz) of the single-table grid
// (bx: row-tile index, by: 16-column group, bz: coset slice).
template <int LOG_R1, int LOG_R2, bool MULTI, int C = 16, int NT = 256>
__device__ __forceinline__ void ntt_pass_body(const PassArgs &a, uint64_t bx, uint32_t by, uint32_t bz) {
    constexpr int R1 = 1 << LOG_R1, R2 = 1 << LOG_R2, R = R1 * R2, LOG_R = LOG_R1 + LOG_R2;
    // C columns per tile: 16 (128-byte row segments), or 8 for narrow tables
    // whose last 16-column group would leave most lanes idle
    // TPC = min(R1, R2) threads per column; each holds R / TPC elements in
    // both rounds: D1 = R2 / TPC size-R1 DFTs in round 1, D2 = R1 / TPC
    // size-R2 DFTs in round 2 (for R1 != R2 no thread idles in either round)
    constexpr int TPC = R1 < R2 ? R1 : R2;
    constexpr int D1 = R2 / TPC, D2 = R1 / TPC;
    constexpr int TILES = NT / (C * TPC);
    static_assert(TILES >= 1 && D1 * R1 == D2 * R2, "tile shape");

    __shared__ uint64_t s_data[TILES][R * C];
    __shared__ uint64_t s_pre[TILES][R];
    __shared__ uint64_t s_post[TILES][R];
    __shared__ uint64_t s_tw[R];

    const int tid = threadIdx.x;
    const int c = tid % C;
    const int tc = (tid / C) % TPC;
    const int tl = tid / (C * TPC);
    const uint64_t g = bx * TILES + tl;
    const bool g_ok = g < a.total_groups;
    const uint32_t col = by * C + c;
    const bool c_ok = col < a.m && g_ok;

    const uint32_t log_s = a.log_s;
    const uint64_t S = 1ull << log_s;
    const uint64_t blk = g >> log_s;
    const uint64_t jlo = g & (S - 1);
    const uint64_t row0 = (blk << a.log_msub) + jlo;  // logical row of t = 0

    // inner twiddles w_R^e = w_256^(e*256/R), e < R
    for (int e = tid; e < R; e += NT) s_tw[e] = LVM_GLOBAL_CONST(a.tw_r)[e << (8 - LOG_R)];
    // per-tile post-twiddle factors, shared by the C columns
    const int tin = tid % (C * TPC);
    for (int t = tin; t < R; t += C * TPC) {
        if (g_ok) {
            uint64_t post = 1;
            if (log_s) {
                // position t holds frequency bitrev_R(t); w_M^(jlo*k1)
                uint64_t k1 = bitrev((uint32_t)t, LOG_R);
                uint32_t e = (uint32_t)((jlo * k1) << (32 - a.log_msub));
                post = tw32(a.tw_lo, a.tw_hi, e);
            }
            s_post[tl][t] = post;
        }
    }

    uint64_t x[R1 > R2 ? R1 : R2];
    uint64_t xin[D1][R1];
    const bool pre_needed = a.weights != nullptr || a.shift_log != 0;
    // Input element (d, i) of a thread is t = t0 + R2*i with t0 = tc + TPC*d,
    // logical row rowA + ((TPC*d + R2*i) << log_s) with rowA = row0 + (tc << log_s):
    // disjoint bit fields, so with bit-reversed input rows (over in_bitrev_bits
    // = log_msub bits) the physical row is bitrev(rowA) + bitrev_R(TPC*d + R2*i)
    // << (bits - log_s - LOG_R). Either way a per-thread base plus a uniform
    // stride per element, and the loads are issued back to back. A lane outside
    // the table reads row 0, column 0 (its results are never stored).
    const uint32_t bits = a.in_bitrev_bits;
    const uint64_t rowA = row0 + ((uint64_t)tc << log_s);
    const uint64_t physA = bits ? (uint64_t)bitrev((uint32_t)rowA, bits) : rowA;
    const gmem_u64 *src = LVM_GLOBAL_CONST(a.in) + (c_ok ? physA * a.in_pitch + col : 0);
    const uint64_t in_step = (uint64_t)a.in_pitch << (bits ? bits - log_s - LOG_R : log_s);
#pragma unroll
    for (int d = 0; d < D1; d++) {
#pragma unroll
        for (int i = 0; i < R1; i++) {
            const uint32_t u = (uint32_t)(TPC * d + R2 * i);
            const uint64_t sel = bits ? bitrev_small(u, LOG_R) : (uint64_t)u;
            xin[d][i] = src[sel * in_step];
        }
    }

    const uint32_t ncos = MULTI ? a.cosets_per_block : 1u;
    for (uint32_t ci = 0; ci < ncos; ci++) {
    const uint32_t coset = bz * a.cosets_per_block + ci;
    if (pre_needed) {
        // pre-multiply factors weights[j] * w_N^(coset*j) of this coset
        for (int t = tin; t < R; t += C * TPC) {
            if (g_ok) {
                uint64_t pre = 1;
                uint64_t lrow = row0 + ((uint64_t)t << log_s);
                if (a.weights) pre = LVM_GLOBAL_CONST(a.weights)[lrow];
                if (a.shift_log) {
                    uint32_t e = (uint32_t)(((uint64_t)coset * lrow) << (32 - a.shift_log));
                    pre = cmul(pre, tw32(a.shift_lo, a.shift_hi, e));
                }
                s_pre[tl][t] = pre;
            }
        }
    }
    __syncthreads();

    // ---- round 1: size-R1 DFTs over t = t0 + R2*i, then w_R^(t0*k_a) ----
#pragma unroll
    for (int d = 0; d < D1; d++) {
        const int t0 = tc + TPC * d;
#pragma unroll
        for (int i = 0; i < R1; i++)
            x[i] = pre_needed ? cmul(xin[d][i], s_pre[tl][t0 + R2 * i])
                              : a.in_canonical ? xin[d][i] : ccanon(xin[d][i]);
        dft_small<R1>(x, a.inverse);
        // position i holds frequency k_a = bitrev_R1(i)
        if constexpr (R2 > 1) {
#pragma unroll
            for (int i = 1; i < R1; i++) x[i] = cmul(x[i], s_tw[t0 * (int)bitrev_small(i, LOG_R1)]);
        }
#pragma unroll
        for (int i = 0; i < R1; i++) s_data[tl][(t0 + R2 * i) * C + c] = x[i];
    }
    __syncthreads();

    // ---- round 2: size-R2 DFTs over t0 at t = R2*t1 + t0 ----
    uint64_t obase = 0;
    if (a.out_coset_rows) obase = (uint64_t)bitrev(coset, a.coset_bits) * a.out_coset_rows;
    // output q = R2*t1 + i is row obase + row0 + (q << log_s): a base and a uniform stride
    gmem_u64 *dst = LVM_GLOBAL(a.out) + (obase + row0 + ((uint64_t)(R2 * tc) << log_s)) * a.out_pitch + col;
    const uint64_t out_step = (uint64_t)a.out_pitch << log_s;
#pragma unroll
    for (int d = 0; d < D2; d++) {
        const int t1 = tc + TPC * d;
#pragma unroll
        for (int i = 0; i < R2; i++) x[i] = s_data[tl][(R2 * t1 + i) * C + c];
        dft_small<R2>(x, a.inverse);
        if (c_ok) {
#pragma unroll
            for (int i = 0; i < R2; i++)
                dst[(R2 * TPC * d + i) * out_step] = log_s ? cmul(x[i], s_post[tl][R2 * t1 + i]) : x[i];
        }
    }
    __syncthreads();  // the LDS tile and factors are rewritten by the next coset
    }
}

template <int LOG_R1, int LOG_R2, bool MULTI, int C>
__global__ __launch_bounds__(256) void ntt_pass_kernel(PassArgs a, uint32_t groups) {
    // one-dimensional grid, column groups fastest, XCD-contiguous (the column
    // groups of a row tile share cache lines)
    const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
    ntt_pass_body<LOG_R1, LOG_R2, MULTI, C>(a, b / groups, b % groups, 0);
}

// One launch for the same pass shape over several tables: `args[t]` is
// table t's pass, blocks [prefix[t], prefix[t+1]) are its tiles, column
// groups fastest so that neighbouring blocks cover whole rows.
// columns and threads per tile of the batched passes: 8-column tiles measured
// slower (28.4 vs 30.8 Gcells/s), and so did 32-column tiles of 512 threads
// (NTT stage 15.4 -> 16.8 ms on the ragged 14-table step)
constexpr int BATCH_C = 16, BATCH_T = 256;

template <int LOG_R1, int LOG_R2, bool MULTI>
__global__ __launch_bounds__(BATCH_T) void ntt_pass_batched_kernel(const PassArgs *__restrict__ args,
                                                               const uint32_t *__restrict__ prefix, uint32_t count) {
    const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t t = find_slot(prefix, count, b);
    const PassArgs a = args[t];
    const uint32_t local = b - prefix[t];
    const uint32_t groups = (a.m + BATCH_C - 1) / BATCH_C;
    ntt_pass_body<LOG_R1, LOG_R2, MULTI, BATCH_C, BATCH_T>(a, local / groups, local % groups, 0);
}

template <int A, int B>
static constexpr uint32_t pass_tiles() {
    constexpr int TPC = (1 << A) < (1 << B) ? (1 << A) : (1 << B);
    return BATCH_T / (BATCH_C * TPC);
}

// pass shapes: (LOG_R1, LOG_R2) for r = 1..8
template <int A, int B, int C>
static void launch_one_c(const PassArgs &a0, uint32_t cosets, hipStream_t st) {
    constexpr int TPC = (1 << A) < (1 << B) ? (1 << A) : (1 << B);
    constexpr int TILES = 256 / (C * TPC);
    PassArgs a = a0;
    // all cosets of a tile come from one block, which reads the tile once
    a.cosets_per_block = cosets;
    const uint32_t groups = (a.m + C - 1) / C;
    const dim3 grid((unsigned)((a.total_groups + TILES - 1) / TILES * groups));
    if (cosets > 1)
        hipLaunchKernelGGL((ntt_pass_kernel<A, B, true, C>), grid, dim3(256), 0, st, a, groups);
    else
        hipLaunchKernelGGL((ntt_pass_kernel<A, B, false, C>), grid, dim3(256), 0, st, a, groups);
}

// 8-column tiles for tables of at most 8 columns (a 16-column tile would
// leave half its lanes or more idle); 16 otherwise: on wide tables 8-column
// tiles measured slower even where they idle fewer lanes (64-byte row segments)
template <int A, int B>
static void launch_one(const PassArgs &a, uint32_t cosets, hipStream_t st) {
    if (a.m <= 8)
        launch_one_c<A, B, 8>(a, cosets, st);
    else
        launch_one_c<A, B, 16>(a, cosets, st);
}

static void launch_pass(int r, const PassArgs &a, uint32_t cosets, hipStream_t st, const char *name) {
    ProfScope ps(name, st);
    switch (r) {
    case 1: launch_one<1, 0>(a, cosets, st); break;
    case 2: launch_one<1, 1>(a, cosets, st); break;
    case 3: launch_one<2, 1>(a, cosets, st); break;
    case 4: launch_one<2, 2>(a, cosets, st); break;
    case 5: launch_one<3, 2>(a, cosets, st); break;
    case 6: launch_one<3, 3>(a, cosets, st); break;
    case 7: launch_one<4, 3>(a, cosets, st); break;
    case 8: launch_one<4, 4>(a, cosets, st); break;
    default: break;
    }
}

// Pass plans per log2 size: default splits log n into passes of at most 8
// bits, largest first; set_ntt_plan installs a measured split.
static int g_plan[33][8];
static int g_plan_len[33];

int plan_passes(unsigned log_n, int *r) {
    if (log_n <= 32 && g_plan_len[log_n] > 0) {
        for (int i = 0; i < g_plan_len[log_n]; i++) r[i] = g_plan[log_n][i];
        return g_plan_len[log_n];
    }
    int np = 0;
    unsigned left = log_n;
    while (left > 0) {
        unsigned take = left > 8 ? 8 : left;
        r[np++] = (int)take;
        left -= take;
    }
    return np;
}

bool set_ntt_plan(unsigned log_n, const int *r, int np) {
    if (log_n > 32 || np < 0 || np > 8) return false;
    int sum = 0;
    for (int i = 0; i < np; i++) {
        if (r[i] < 1 || r[i] > 8) return false;
        sum += r[i];
    }
    if (np > 0 && sum != (int)log_n) return false;
    for (int i = 0; i < np; i++) g_plan[log_n][i] = r[i];
    g_plan_len[log_n] = np;  // 0 restores the default split
    return true;
}

// ---- work-list launches over several tables ----

uint32_t pass_blocks(int r, const PassArgs &a) {
    uint32_t tiles = 1;
    switch (r) {
    case 1: tiles = pass_tiles<1, 0>(); break;
    case 2: tiles = pass_tiles<1, 1>(); break;
    case 3: tiles = pass_tiles<2, 1>(); break;
    case 4: tiles = pass_tiles<2, 2>(); break;
    case 5: tiles = pass_tiles<3, 2>(); break;
    case 6: tiles = pass_tiles<3, 3>(); break;
    case 7: tiles = pass_tiles<4, 3>(); break;
    case 8: tiles = pass_tiles<4, 4>(); break;
    default: break;
    }
    return (uint32_t)((a.total_groups + tiles - 1) / tiles) * ((a.m + BATCH_C - 1) / BATCH_C);
}

template <int A, int B>
static void launch_batched_one(const PassArgs *args, const uint32_t *prefix, uint32_t count, uint32_t blocks,
                               bool multi, hipStream_t st) {
    if (multi)
        hipLaunchKernelGGL((ntt_pass_batched_kernel<A, B, true>), dim3(blocks), dim3(BATCH_T), 0, st, args, prefix,
                           count);
    else
        hipLaunchKernelGGL((ntt_pass_batched_kernel<A, B, false>), dim3(blocks), dim3(BATCH_T), 0, st, args, prefix,
                           count);
}

void launch_pass_batched(int r, const PassArgs *d_args, const uint32_t *d_prefix, uint32_t count, uint32_t blocks,
                         bool multi, hipStream_t st) {
    if (blocks == 0) return;
    switch (r) {
    case 1: launch_batched_one<1, 0>(d_args, d_prefix, count, blocks, multi, st); break;
    case 2: launch_batched_one<1, 1>(d_args, d_prefix, count, blocks, multi, st); break;
    case 3: launch_batched_one<2, 1>(d_args, d_prefix, count, blocks, multi, st); break;
    case 4: launch_batched_one<2, 2>(d_args, d_prefix, count, blocks, multi, st); break;
    case 5: launch_batched_one<3, 2>(d_args, d_prefix, count, blocks, multi, st); break;
    case 6: launch_batched_one<3, 3>(d_args, d_prefix, count, blocks, multi, st); break;
    case 7: launch_batched_one<4, 3>(d_args, d_prefix, count, blocks, multi, st); break;
    case 8: launch_batched_one<4, 4>(d_args, d_prefix, count, blocks, multi, st); break;
    default: break;
    }
}

// The pass sequence of one table's coset LDE (inverse transform of the
// trace, then the coset expansion), as the argument blocks the single-table
// path launches: radix in r_out, arguments in a_out, coset count in cos_out.
int lde_pass_list(const uint64_t *trace, uint32_t m, unsigned log_n, unsigned log_blowup, const uint64_t *weights,
                  uint64_t *coef, uint64_t *lde, const Tables &tb, int *r_out, PassArgs *a_out, uint32_t *cos_out) {
    int r[8];
    const int np = plan_passes(log_n, r);
    if (np == 0) return -1;
    int k = 0;
    unsigned done = 0;
    for (int p = 0; p < np; p++) {
        PassArgs a{};
        a.in = p == 0 ? trace : coef;
        a.out = coef;
        a.in_pitch = a.out_pitch = m;
        a.m = m;
        a.log_msub = log_n - done;
        a.log_s = log_n - done - r[p];
        a.total_groups = (1ull << log_n) >> r[p];
        a.tw_lo = tb.inv_lo;
        a.tw_hi = tb.inv_hi;
        a.tw_r = tb.inv_r;
        a.inverse = 1;
        a.in_canonical = p > 0;
        a.cosets_per_block = 1;
        r_out[k] = r[p];
        cos_out[k] = 1;
        a_out[k++] = a;
        done += r[p];
    }
    const uint32_t B = 1u << log_blowup;
    done = 0;
    for (int p = 0; p < np; p++) {
        PassArgs a{};
        a.in = p == 0 ? coef : lde;
        a.out = lde;
        a.in_pitch = a.out_pitch = m;
        a.m = m;
        a.log_msub = log_n - done;
        a.log_s = log_n - done - r[p];
        a.tw_lo = tb.fwd_lo;
        a.tw_hi = tb.fwd_hi;
        a.tw_r = tb.fwd_r;
        a.in_canonical = 1;
        if (p == 0) {
            a.in_bitrev_bits = log_n;
            a.total_groups = (1ull << log_n) >> r[0];
            a.out_coset_rows = 1ull << log_n;
            a.coset_bits = log_blowup;
            a.shift_log = log_n + log_blowup;
            a.weights = weights;
            a.shift_lo = tb.fwd_lo;
            a.shift_hi = tb.fwd_hi;
            a.cosets_per_block = B;
            cos_out[k] = B;
        } else {
            a.total_groups = ((uint64_t)B << log_n) >> r[p];
            a.cosets_per_block = 1;
            cos_out[k] = 1;
        }
        r_out[k] = r[p];
        a_out[k++] = a;
        done += r[p];
    }
    return k;
}

// In-place DIF passes over `rows` rows of `buf` (independent size-2^log_n
// problems stacked), starting at pass index `first`.
static const char *PASS_NAMES[2][4] = {{"ntt_fwd_pass0", "ntt_fwd_pass1", "ntt_fwd_pass2", "ntt_fwd_pass3"},
                                       {"ntt_inv_pass0", "ntt_inv_pass1", "ntt_inv_pass2", "ntt_inv_pass3"}};

static void run_passes(uint64_t *buf, uint32_t pitch, uint32_t m, size_t rows, unsigned log_n, int first,
                       const int *r, int np, const Tables &tb, bool inverse, hipStream_t st) {
    unsigned done = 0;
    for (int p = 0; p < first; p++) done += r[p];
    for (int p = first; p < np; p++) {
        PassArgs a{};
        a.in = buf;
        a.out = buf;
        a.in_pitch = a.out_pitch = pitch;
        a.m = m;
        a.log_msub = log_n - done;
        a.log_s = log_n - done - r[p];
        a.total_groups = rows >> r[p];
        a.tw_lo = inverse ? tb.inv_lo : tb.fwd_lo;
        a.tw_hi = inverse ? tb.inv_hi : tb.fwd_hi;
        a.tw_r = inverse ? tb.inv_r : tb.fwd_r;
        a.inverse = inverse ? 1u : 0u;
        a.in_canonical = 1;
        launch_pass(r[p], a, 1, st, PASS_NAMES[inverse][p < 3 ? p : 3]);
        done += r[p];
    }
}

// Unnormalised inverse DFT of each column of a natural-order row-major
// table; output rows are in bit-reversed order.
void ntt_inverse_rows(const uint64_t *in, uint32_t in_pitch, uint64_t *out, uint32_t out_pitch, uint32_t m,
                      unsigned log_n, const Tables &tb, hipStream_t st) {
    int r[8];
    int np = plan_passes(log_n, r);
    if (np == 0) {
        (void)hipMemcpy2DAsync(out, out_pitch * 8, in, in_pitch * 8, m * 8, 1, hipMemcpyDeviceToDevice, st);
        return;
    }
    PassArgs a{};
    a.in = in;
    a.out = out;
    a.in_pitch = in_pitch;
    a.out_pitch = out_pitch;
    a.m = m;
    a.log_msub = log_n;
    a.log_s = log_n - r[0];
    a.total_groups = (1ull << log_n) >> r[0];
    a.tw_lo = tb.inv_lo;
    a.tw_hi = tb.inv_hi;
    a.tw_r = tb.inv_r;
    a.inverse = 1;
    launch_pass(r[0], a, 1, st, PASS_NAMES[1][0]);
    run_passes(out, out_pitch, m, 1ull << log_n, log_n, 1, r, np, tb, true, st);
}

// Forward DFT of each column (natural in, bit-reversed out).
void ntt_forward_rows(const uint64_t *in, uint32_t in_pitch, uint64_t *out, uint32_t out_pitch, uint32_t m,
                      unsigned log_n, const Tables &tb, hipStream_t st) {
    int r[8];
    int np = plan_passes(log_n, r);
    if (np == 0) {
        (void)hipMemcpy2DAsync(out, out_pitch * 8, in, in_pitch * 8, m * 8, 1, hipMemcpyDeviceToDevice, st);
        return;
    }
    PassArgs a{};
    a.in = in;
    a.out = out;
    a.in_pitch = in_pitch;
    a.out_pitch = out_pitch;
    a.m = m;
    a.log_msub = log_n;
    a.log_s = log_n - r[0];
    a.total_groups = (1ull << log_n) >> r[0];
    a.tw_lo = tb.fwd_lo;
    a.tw_hi = tb.fwd_hi;
    a.tw_r = tb.fwd_r;
    launch_pass(r[0], a, 1, st, PASS_NAMES[0][0]);
    run_passes(out, out_pitch, m, 1ull << log_n, log_n, 1, r, np, tb, false, st);
}

// Coset LDE expansion: coefficient table `coef` (n rows, bit-reversed order,
// unnormalised iDFT of the trace) -> `lde` (n*B rows, bit-reversed order over
// log2(n*B) bits) of the evaluations on offset*<w_N> (offset folded into weights).
// log_n >= 1 (every entry point rejects tables of fewer than two rows); returns
// LVM_ERR_ARG without writing `lde` for a size that has no pass plan.
int ntt_coset_expand(const uint64_t *coef, uint32_t coef_pitch, uint64_t *lde, uint32_t lde_pitch, uint32_t m,
                      unsigned log_n, unsigned log_blowup, const uint64_t *weights, const Tables &tb,
                      hipStream_t st) {
    int r[8];
    int np = plan_passes(log_n, r);
    if (np == 0) return LVM_ERR_ARG;
    const uint32_t B = 1u << log_blowup;
    PassArgs a{};
    a.in = coef;
    a.out = lde;
    a.in_pitch = coef_pitch;
    a.out_pitch = lde_pitch;
    a.m = m;
    a.log_msub = log_n;
    a.log_s = log_n - r[0];
    a.in_bitrev_bits = log_n;
    a.total_groups = (1ull << log_n) >> r[0];
    a.out_coset_rows = 1ull << log_n;
    a.coset_bits = log_blowup;
    a.shift_log = log_n + log_blowup;
    a.weights = weights;
    a.in_canonical = 1;
    a.tw_lo = tb.fwd_lo;
    a.tw_hi = tb.fwd_hi;
    a.shift_lo = tb.fwd_lo;
    a.shift_hi = tb.fwd_hi;
    a.tw_r = tb.fwd_r;
    launch_pass(r[0], a, B, st, "ntt_coset_pass0");
    run_passes(lde, lde_pitch, m, (size_t)B << log_n, log_n, 1, r, np, tb, false, st);
    return LVM_OK;
}

}  // namespace lvm
