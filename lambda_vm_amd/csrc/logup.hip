// LogUp auxiliary trace on gfx950: fingerprints, batch inversion, term
// columns and the accumulated running-sum column of one table, built on the
// device from its resident main trace.
//
// Semantics (crypto/stark/src/lookup.rs:1141-1313 build_auxiliary_trace,
// :1660-1780 compute_logup_term_column, :1790-1834
// build_accumulated_column_from_terms) over the flat descriptor of
// crypto/stark/src/logup_gpu.rs:100-290:
//   fp_k(i)   = z - (bus_k + sum_e alpha^{aidx_e} (const_e + sum_t coef_t main[i][col_t]))
//   m_k(i)    = mconst_k + sum_t mcoef_t main[i][mcol_t]
//   term_c(i) = sum_{k in column c} m_k(i) / fp_k(i)
//   L = sum_i sum_c term_c(i),  acc(i) = sum_{j<i} sum_c term_c(j) - i * L/N
// Aux layout: row-major, C = committed + 1 Fp3 columns per row (committed
// term columns, then the accumulated column).
//
// Kernel plan: each thread owns G consecutive rows. For every interaction it
// builds the running products of its G fingerprints, inverts the last one
// (one Fp3 inversion per G elements, Montgomery's trick within the thread),
// and walks back recomputing each fingerprint instead of keeping it, so only
// the running products and the term accumulators live in registers. Row sums
// go to LDS, the block scans them, and a single-block pass scans the block
// totals; a last pass writes the accumulated column.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

constexpr int LOGUP_G = 4;        // rows per thread
constexpr int LOGUP_BLOCK = 256;  // threads per block
constexpr uint32_t LOGUP_ROWS_PER_BLOCK = LOGUP_G * LOGUP_BLOCK;

struct LogupDev {
    const uint64_t *bus_ids;
    const uint32_t *elem_offsets, *elem_alpha_idx;
    const uint64_t *elem_const;
    const uint32_t *term_offsets;
    const uint64_t *term_coef;
    const uint32_t *term_col;
    const uint32_t *out_col_offsets, *out_col_interactions;
    const uint64_t *mult_const;
    const uint32_t *mult_term_offsets;
    const uint64_t *mult_term_coef;
    const uint32_t *mult_term_col;
    const uint64_t *alpha;  // Fp3 powers
    uint32_t num_out_cols;
    uint64_t z0, z1, z2;
};

struct MainView {
    const uint64_t *p;
    uint64_t rs, cs;  // element (row, col) at p[row*rs + col*cs]
    __device__ __forceinline__ uint64_t at(uint64_t row, uint32_t col) const { return p[row * rs + col * cs]; }
};

__device__ __forceinline__ fp3 fingerprint(const LogupDev &d, uint32_t k, const MainView &mv, uint64_t row) {
    fp3 lc = fp3{d.bus_ids[k], 0, 0};
    const uint32_t e1 = d.elem_offsets[k + 1];
    for (uint32_t e = d.elem_offsets[k]; e < e1; e++) {
        uint64_t base = d.elem_const[e];
        const uint32_t t1 = d.term_offsets[e + 1];
        for (uint32_t t = d.term_offsets[e]; t < t1; t++) base = add(base, mul(d.term_coef[t], mv.at(row, d.term_col[t])));
        const uint64_t *a = d.alpha + 3 * (uint64_t)d.elem_alpha_idx[e];
        lc = f3_add(lc, f3_mul_base(fp3{a[0], a[1], a[2]}, base));
    }
    return f3_sub(fp3{d.z0, d.z1, d.z2}, lc);
}

__device__ __forceinline__ uint64_t multiplicity(const LogupDev &d, uint32_t k, const MainView &mv, uint64_t row) {
    uint64_t m = d.mult_const[k];
    const uint32_t t1 = d.mult_term_offsets[k + 1];
    for (uint32_t t = d.mult_term_offsets[k]; t < t1; t++) m = add(m, mul(d.mult_term_coef[t], mv.at(row, d.mult_term_col[t])));
    return m;
}

// The terms of output columns [c0, c1) at the thread's rows r0 .. r0+cnt-1
// (one batch inversion per interaction over those rows): committed columns
// into aux, every column into terms_out, their sum added to rowsum.
__device__ __forceinline__ void logup_terms(const LogupDev &d, const MainView &mv, uint64_t n, uint64_t r0, int cnt,
                                            uint32_t c0, uint32_t c1, uint64_t *__restrict__ aux,
                                            uint64_t *__restrict__ terms_out, fp3 rowsum[LOGUP_G]) {
    const uint32_t C = d.num_out_cols;
    for (uint32_t c = c0; c < c1; c++) {
        fp3 term[LOGUP_G];
#pragma unroll
        for (int j = 0; j < LOGUP_G; j++) term[j] = fp3{0, 0, 0};
        const uint32_t k1 = d.out_col_offsets[c + 1];
        for (uint32_t ki = d.out_col_offsets[c]; ki < k1; ki++) {
            const uint32_t k = d.out_col_interactions[ki];
            fp3 pre[LOGUP_G];
            fp3 run = fp3{1, 0, 0};
#pragma unroll
            for (int j = 0; j < LOGUP_G; j++) {
                if (j < cnt) run = j == 0 ? fingerprint(d, k, mv, r0) : f3_mul(run, fingerprint(d, k, mv, r0 + j));
                pre[j] = run;
            }
            if (cnt == 0) continue;
            fp3 inv = f3_inv(run);  // (prod of the thread's fingerprints)^-1
#pragma unroll
            for (int j = LOGUP_G - 1; j >= 1; j--) {
                if (j < cnt) {
                    const fp3 f = fingerprint(d, k, mv, r0 + j);
                    const fp3 r = f3_mul(inv, pre[j - 1]);
                    inv = f3_mul(inv, f);
                    term[j] = f3_add(term[j], f3_mul_base(r, multiplicity(d, k, mv, r0 + j)));
                }
            }
            term[0] = f3_add(term[0], f3_mul_base(inv, multiplicity(d, k, mv, r0)));
        }
#pragma unroll
        for (int j = 0; j < LOGUP_G; j++) {
            if (j < cnt) {
                const uint64_t row = r0 + j;
                if (c + 1 < C) {
                    uint64_t *o = aux + 3 * (row * C + c);
                    o[0] = term[j].c0;
                    o[1] = term[j].c1;
                    o[2] = term[j].c2;
                }
                if (terms_out) {
                    uint64_t *o = terms_out + 3 * ((uint64_t)c * n + row);
                    o[0] = term[j].c0;
                    o[1] = term[j].c1;
                    o[2] = term[j].c2;
                }
                rowsum[j] = f3_add(rowsum[j], term[j]);
            }
        }
    }
}

// Block-local inclusive prefix of the row sums into `scan`, the block total
// into `totals` (every thread of the block calls it).
__device__ __forceinline__ void logup_block_scan(fp3 rowsum[LOGUP_G], uint64_t r0, int cnt, uint64_t *__restrict__ scan,
                                                 uint64_t *__restrict__ totals) {
    __shared__ fp3 s_tot[LOGUP_BLOCK];
    const uint32_t t = threadIdx.x;
    // inclusive prefix over the thread's rows
#pragma unroll
    for (int j = 1; j < LOGUP_G; j++) rowsum[j] = f3_add(rowsum[j - 1], rowsum[j]);
    // rows past the end add nothing, so the last prefix is the thread total
    s_tot[t] = rowsum[LOGUP_G - 1];
    __syncthreads();
    // Hillis-Steele inclusive scan of the thread totals
    for (uint32_t off = 1; off < LOGUP_BLOCK; off <<= 1) {
        const fp3 v = t >= off ? s_tot[t - off] : fp3{0, 0, 0};
        __syncthreads();
        if (t >= off) s_tot[t] = f3_add(s_tot[t], v);
        __syncthreads();
    }
    const fp3 before = t ? s_tot[t - 1] : fp3{0, 0, 0};
#pragma unroll
    for (int j = 0; j < LOGUP_G; j++) {
        if (j < cnt) {
            const fp3 v = f3_add(before, rowsum[j]);
            uint64_t *o = scan + 3 * (r0 + j);
            o[0] = v.c0;
            o[1] = v.c1;
            o[2] = v.c2;
        }
    }
    if (t == LOGUP_BLOCK - 1) {
        uint64_t *o = totals + 3 * (uint64_t)blockIdx.x;
        o[0] = s_tot[t].c0;
        o[1] = s_tot[t].c1;
        o[2] = s_tot[t].c2;
    }
}

// Pass 1: every output column of the thread's rows, then the block scan.
__global__ __launch_bounds__(LOGUP_BLOCK) void logup_rows_kernel(LogupDev d, MainView mv, uint64_t n,
                                                                 uint64_t *__restrict__ aux,
                                                                 uint64_t *__restrict__ terms_out,
                                                                 uint64_t *__restrict__ scan,
                                                                 uint64_t *__restrict__ totals) {
    const uint64_t r0 = ((uint64_t)blockIdx.x * LOGUP_BLOCK + threadIdx.x) * LOGUP_G;
    const int cnt = r0 >= n ? 0 : (int)((n - r0) < (uint64_t)LOGUP_G ? (n - r0) : (uint64_t)LOGUP_G);
    fp3 rowsum[LOGUP_G];
#pragma unroll
    for (int j = 0; j < LOGUP_G; j++) rowsum[j] = fp3{0, 0, 0};
    logup_terms(d, mv, n, r0, cnt, 0, d.num_out_cols, aux, terms_out, rowsum);
    logup_block_scan(rowsum, r0, cnt, scan, totals);
}

// Pass 1 of a table with few rows and many output columns, split in two:
// (a) block (x, y) computes columns [y*cps, (y+1)*cps) of its rows and
// stores their row sums in partial[y][row]; (b) the row sums over the
// slices, then the block scan of pass 1.
__global__ __launch_bounds__(LOGUP_BLOCK) void logup_slice_kernel(LogupDev d, MainView mv, uint64_t n, uint32_t cps,
                                                                  uint64_t *__restrict__ aux,
                                                                  uint64_t *__restrict__ terms_out,
                                                                  uint64_t *__restrict__ partial) {
    const uint64_t r0 = ((uint64_t)blockIdx.x * LOGUP_BLOCK + threadIdx.x) * LOGUP_G;
    const int cnt = r0 >= n ? 0 : (int)((n - r0) < (uint64_t)LOGUP_G ? (n - r0) : (uint64_t)LOGUP_G);
    const uint32_t c0 = blockIdx.y * cps;
    const uint32_t c1 = c0 + cps < d.num_out_cols ? c0 + cps : d.num_out_cols;
    fp3 rowsum[LOGUP_G];
#pragma unroll
    for (int j = 0; j < LOGUP_G; j++) rowsum[j] = fp3{0, 0, 0};
    logup_terms(d, mv, n, r0, cnt, c0, c1, aux, terms_out, rowsum);
#pragma unroll
    for (int j = 0; j < LOGUP_G; j++) {
        if (j < cnt) {
            uint64_t *o = partial + 3 * ((uint64_t)blockIdx.y * n + r0 + j);
            o[0] = rowsum[j].c0;
            o[1] = rowsum[j].c1;
            o[2] = rowsum[j].c2;
        }
    }
}

__global__ __launch_bounds__(LOGUP_BLOCK) void logup_slice_scan_kernel(const uint64_t *__restrict__ partial,
                                                                       uint32_t slices, uint64_t n,
                                                                       uint64_t *__restrict__ scan,
                                                                       uint64_t *__restrict__ totals) {
    const uint64_t r0 = ((uint64_t)blockIdx.x * LOGUP_BLOCK + threadIdx.x) * LOGUP_G;
    const int cnt = r0 >= n ? 0 : (int)((n - r0) < (uint64_t)LOGUP_G ? (n - r0) : (uint64_t)LOGUP_G);
    fp3 rowsum[LOGUP_G];
#pragma unroll
    for (int j = 0; j < LOGUP_G; j++) {
        rowsum[j] = fp3{0, 0, 0};
        if (j < cnt) {
            for (uint32_t y = 0; y < slices; y++) {
                const uint64_t *q = partial + 3 * ((uint64_t)y * n + r0 + j);
                rowsum[j] = f3_add(rowsum[j], fp3{q[0], q[1], q[2]});
            }
        }
    }
    logup_block_scan(rowsum, r0, cnt, scan, totals);
}

// Pass 2 (one block): exclusive scan of the block totals in place; out[0..3)
// = L, out[3..6) = L * n^-1.
__global__ __launch_bounds__(1024) void logup_totals_kernel(uint64_t *__restrict__ totals, uint32_t nblocks,
                                                            uint64_t inv_n, uint64_t *__restrict__ out) {
    __shared__ fp3 s[1024];
    const uint32_t t = threadIdx.x;
    fp3 carry = fp3{0, 0, 0};
    for (uint32_t base = 0; base < nblocks; base += 1024) {
        const uint32_t i = base + t;
        const fp3 v = i < nblocks ? fp3{totals[3 * i], totals[3 * i + 1], totals[3 * i + 2]} : fp3{0, 0, 0};
        s[t] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {
            const fp3 u = t >= off ? s[t - off] : fp3{0, 0, 0};
            __syncthreads();
            if (t >= off) s[t] = f3_add(s[t], u);
            __syncthreads();
        }
        if (i < nblocks) {
            const fp3 ex = f3_add(carry, t ? s[t - 1] : fp3{0, 0, 0});
            totals[3 * i] = ex.c0;
            totals[3 * i + 1] = ex.c1;
            totals[3 * i + 2] = ex.c2;
        }
        carry = f3_add(carry, s[1023]);
        __syncthreads();
    }
    if (t == 0) {
        const fp3 off = f3_mul_base(carry, inv_n);
        out[0] = carry.c0;
        out[1] = carry.c1;
        out[2] = carry.c2;
        out[3] = off.c0;
        out[4] = off.c1;
        out[5] = off.c2;
    }
}

// Pass 3: acc(i) = (sum of row sums before i) - i * L/N into aux column C-1.
__global__ __launch_bounds__(256) void logup_acc_kernel(const uint64_t *__restrict__ scan,
                                                        const uint64_t *__restrict__ block_prefix,
                                                        const uint64_t *__restrict__ lo, uint64_t n, uint32_t C,
                                                        uint64_t *__restrict__ aux) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = i / LOGUP_ROWS_PER_BLOCK;
    fp3 ex = fp3{block_prefix[3 * b], block_prefix[3 * b + 1], block_prefix[3 * b + 2]};
    if (i % LOGUP_ROWS_PER_BLOCK) ex = f3_add(ex, fp3{scan[3 * (i - 1)], scan[3 * (i - 1) + 1], scan[3 * (i - 1) + 2]});
    const fp3 acc = f3_sub(ex, f3_mul_base(fp3{lo[3], lo[4], lo[5]}, i));
    uint64_t *o = aux + 3 * (i * C + C - 1);
    o[0] = acc.c0;
    o[1] = acc.c1;
    o[2] = acc.c2;
}

namespace {

template <class T>
size_t put(std::vector<uint8_t> &blob, const T *v, size_t n) {
    size_t at = (blob.size() + 15) & ~(size_t)15;
    blob.resize(at + (n ? n : 1) * sizeof(T));
    if (n) std::memcpy(blob.data() + at, v, n * sizeof(T));
    return at;
}

}  // namespace

int logup_aux_impl(const uint64_t *dev_main, size_t row_stride, size_t col_stride, size_t n, size_t num_cols,
                   const lvm_logup_desc *d, const uint64_t *alpha, size_t num_alpha, const uint64_t z[3],
                   uint64_t *dev_aux, uint64_t *dev_terms, uint64_t L_out[3]) {
    if (!dev_main || !d || !alpha || !z || !dev_aux || !L_out) return fail(LVM_ERR_ARG, "null argument");
    if (n == 0) return fail(LVM_ERR_ARG, "no rows");
    if (n >= (1ull << 40)) return fail(LVM_ERR_ARG, "too many rows");
    const size_t K = d->num_interactions, C = d->num_out_cols;
    if (K == 0 || C == 0) return fail(LVM_ERR_ARG, "no interactions or no output columns");
    if (!d->bus_ids || !d->elem_offsets || !d->term_offsets || !d->out_col_offsets || !d->out_col_interactions ||
        !d->mult_const || !d->mult_term_offsets)
        return fail(LVM_ERR_ARG, "descriptor array missing");
    // CSR consistency: every index the kernels follow stays inside its array
    const size_t E = d->elem_offsets[K], Tm = d->mult_term_offsets[K], O = d->out_col_offsets[C];
    for (size_t k = 0; k < K; k++)
        if (d->elem_offsets[k] > d->elem_offsets[k + 1] || d->mult_term_offsets[k] > d->mult_term_offsets[k + 1])
            return fail(LVM_ERR_ARG, "descriptor offsets not ascending");
    if (d->elem_offsets[0] != 0 || d->mult_term_offsets[0] != 0 || d->term_offsets[0] != 0 ||
        d->out_col_offsets[0] != 0)
        return fail(LVM_ERR_ARG, "descriptor offsets must start at 0");
    for (size_t e = 0; e < E; e++) {
        if (d->term_offsets[e] > d->term_offsets[e + 1]) return fail(LVM_ERR_ARG, "descriptor offsets not ascending");
        if (d->elem_alpha_idx[e] >= num_alpha) return fail(LVM_ERR_ARG, "alpha index beyond the given powers");
    }
    const size_t Tt = d->term_offsets[E];
    for (size_t t = 0; t < Tt; t++)
        if (d->term_col[t] >= num_cols) return fail(LVM_ERR_ARG, "fingerprint term column out of range");
    for (size_t t = 0; t < Tm; t++)
        if (d->mult_term_col[t] >= num_cols) return fail(LVM_ERR_ARG, "multiplicity term column out of range");
    for (size_t c = 0; c < C; c++)
        if (d->out_col_offsets[c] > d->out_col_offsets[c + 1]) return fail(LVM_ERR_ARG, "descriptor offsets not ascending");
    for (size_t i = 0; i < O; i++)
        if (d->out_col_interactions[i] >= K) return fail(LVM_ERR_ARG, "output column interaction out of range");
    if ((E && (!d->elem_alpha_idx || !d->elem_const)) || (Tt && (!d->term_coef || !d->term_col)) ||
        (Tm && (!d->mult_term_coef || !d->mult_term_col)))
        return fail(LVM_ERR_ARG, "descriptor array missing");

    Device *dv = device_state();
    if (!dv) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = lib_stream();

    std::vector<uint8_t> blob;
    const size_t o_bus = put(blob, d->bus_ids, K), o_eoff = put(blob, d->elem_offsets, K + 1),
                 o_ea = put(blob, d->elem_alpha_idx, E), o_ec = put(blob, d->elem_const, E),
                 o_toff = put(blob, d->term_offsets, E + 1), o_tc = put(blob, d->term_coef, Tt),
                 o_tcol = put(blob, d->term_col, Tt), o_ooff = put(blob, d->out_col_offsets, C + 1),
                 o_oint = put(blob, d->out_col_interactions, O), o_mc = put(blob, d->mult_const, K),
                 o_moff = put(blob, d->mult_term_offsets, K + 1), o_mcoef = put(blob, d->mult_term_coef, Tm),
                 o_mcol = put(blob, d->mult_term_col, Tm), o_alpha = put(blob, alpha, 3 * num_alpha);
    const uint64_t nblocks = (n + LOGUP_ROWS_PER_BLOCK - 1) / LOGUP_ROWS_PER_BLOCK;
    // Few rows and many output columns (a wide chip of a short run): the
    // columns are split over grid.y so that some 2^18 threads work, each slice
    // a contiguous run of columns; every column is still computed once.
    const uint64_t row_threads = (n + LOGUP_G - 1) / LOGUP_G;
    uint64_t slices = 1, cps = C;
    if (row_threads < (1u << 17) && C > 1) {
        slices = std::min<uint64_t>(C, ((1u << 18) + row_threads - 1) / row_threads);
        slices = std::min<uint64_t>(slices, 65535);
        cps = (C + slices - 1) / slices;
        slices = (C + cps - 1) / cps;  // no empty slice
    }
    uint8_t *db = nullptr;
    uint64_t *scan = nullptr, *totals = nullptr, *lo = nullptr, *partial = nullptr;
    auto cleanup = [&]() {
        lvm_free(db);
        lvm_free(scan);
        lvm_free(totals);
        lvm_free(lo);
        lvm_free(partial);
    };
    if (lvm_malloc(&db, blob.size()) != hipSuccess || lvm_malloc(&scan, n * 24) != hipSuccess ||
        lvm_malloc(&totals, nblocks * 24) != hipSuccess || lvm_malloc(&lo, 6 * 8) != hipSuccess ||
        (slices > 1 && lvm_malloc(&partial, slices * n * 24) != hipSuccess)) {
        cleanup();
        return fail(LVM_ERR_NOMEM, "device allocation for the LogUp build failed");
    }
    StageScope ss(st);
    hipError_t e = stage_upload(db, blob.data(), blob.size(), st);
    LogupDev ld;
    ld.bus_ids = (const uint64_t *)(db + o_bus);
    ld.elem_offsets = (const uint32_t *)(db + o_eoff);
    ld.elem_alpha_idx = (const uint32_t *)(db + o_ea);
    ld.elem_const = (const uint64_t *)(db + o_ec);
    ld.term_offsets = (const uint32_t *)(db + o_toff);
    ld.term_coef = (const uint64_t *)(db + o_tc);
    ld.term_col = (const uint32_t *)(db + o_tcol);
    ld.out_col_offsets = (const uint32_t *)(db + o_ooff);
    ld.out_col_interactions = (const uint32_t *)(db + o_oint);
    ld.mult_const = (const uint64_t *)(db + o_mc);
    ld.mult_term_offsets = (const uint32_t *)(db + o_moff);
    ld.mult_term_coef = (const uint64_t *)(db + o_mcoef);
    ld.mult_term_col = (const uint32_t *)(db + o_mcol);
    ld.alpha = (const uint64_t *)(db + o_alpha);
    ld.num_out_cols = (uint32_t)C;
    ld.z0 = z[0];
    ld.z1 = z[1];
    ld.z2 = z[2];
    MainView mv{dev_main, row_stride, col_stride};
    // n^-1 (n < p)
    const uint64_t inv_n = inv((uint64_t)n);
    {
        ProfScope ps("logup_terms", st);
        if (slices == 1) {
            hipLaunchKernelGGL(logup_rows_kernel, dim3((unsigned)nblocks), dim3(LOGUP_BLOCK), 0, st, ld, mv,
                               (uint64_t)n, dev_aux, dev_terms, scan, totals);
        } else {
            hipLaunchKernelGGL(logup_slice_kernel, dim3((unsigned)nblocks, (unsigned)slices), dim3(LOGUP_BLOCK), 0, st,
                               ld, mv, (uint64_t)n, (uint32_t)cps, dev_aux, dev_terms, partial);
            hipLaunchKernelGGL(logup_slice_scan_kernel, dim3((unsigned)nblocks), dim3(LOGUP_BLOCK), 0, st, partial,
                               (uint32_t)slices, (uint64_t)n, scan, totals);
        }
    }
    {
        ProfScope ps("logup_accumulate", st);
        hipLaunchKernelGGL(logup_totals_kernel, dim3(1), dim3(1024), 0, st, totals, (uint32_t)nblocks, inv_n, lo);
        hipLaunchKernelGGL(logup_acc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, scan, totals, lo,
                           (uint64_t)n, (uint32_t)C, dev_aux);
    }
    if (e == hipSuccess) e = hipGetLastError();
    uint64_t hl[6];
    if (e == hipSuccess) e = stage_download(hl, lo, 48, st);
    if (e == hipSuccess) e = stage_sync(st);
    cleanup();
    if (e != hipSuccess) return fail(LVM_ERR_DEVICE, std::string("LogUp build failed: ") + hipGetErrorString(e));
    L_out[0] = hl[0];
    L_out[1] = hl[1];
    L_out[2] = hl[2];
    return LVM_OK;
}

}  // namespace lvm

extern "C" {

int lvm_logup_aux_dev(const uint64_t *dev_main, size_t row_stride, size_t col_stride, size_t num_rows,
                      size_t num_cols, const lvm_logup_desc *desc, const uint64_t *alpha_powers, size_t num_alpha,
                      const uint64_t z[3], uint64_t *dev_aux_out, uint64_t *dev_terms_out, uint64_t L_out[3]) {
    return lvm::logup_aux_impl(dev_main, row_stride, col_stride, num_rows, num_cols, desc, alpha_powers, num_alpha,
                               z, dev_aux_out, dev_terms_out, L_out);
}

int lvm_logup_term_columns(const uint64_t *main_cols, size_t num_rows, size_t num_cols, const lvm_logup_desc *desc,
                           const uint64_t *alpha_powers, size_t num_alpha, const uint64_t z[3], uint64_t *terms_out) {
    using namespace lvm;
    if (!main_cols || !terms_out || !desc) return fail(LVM_ERR_ARG, "null argument");
    if (!device_state()) return LVM_ERR_DEVICE;
    hipStream_t st = lib_stream();
    const size_t C = desc->num_out_cols;
    uint64_t *dm = nullptr, *daux = nullptr, *dterms = nullptr;
    auto cleanup = [&]() {
        lvm_free(dm);
        lvm_free(daux);
        lvm_free(dterms);
    };
    if (lvm_malloc(&dm, num_rows * num_cols * 8) != hipSuccess || lvm_malloc(&daux, num_rows * C * 24) != hipSuccess ||
        lvm_malloc(&dterms, num_rows * C * 24) != hipSuccess) {
        cleanup();
        return fail(LVM_ERR_NOMEM, "device allocation for the LogUp terms failed");
    }
    if (hipMemcpyAsync(dm, main_cols, num_rows * num_cols * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
        cleanup();
        return fail(LVM_ERR_DEVICE, "upload of the main columns failed");
    }
    uint64_t L[3];
    // column-major host layout: (row, col) at col * num_rows + row
    int rc = logup_aux_impl(dm, 1, num_rows, num_rows, num_cols, desc, alpha_powers, num_alpha, z, daux, dterms, L);
    if (rc == LVM_OK) {
        hipError_t e = hipMemcpyAsync(terms_out, dterms, num_rows * C * 24, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(LVM_ERR_DEVICE, "download of the term columns failed");
    }
    cleanup();
    return rc;
}

}  // extern "C"
