// C ABI entry points (include/lvm_gpu.h). Every entry point validates its
// arguments, runs on the current device's library stream, and reports
// failure through a status code plus lvm_last_error(); nothing aborts.
#include <cstring>
#include <string>
#include <vector>

#include "lvm_gpu.h"
#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {
void launch_gl_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n, bool ext3,
                  hipStream_t st);
void launch_gather_bitrev_rows(const uint64_t *src, uint64_t *dst, uint64_t rows, uint32_t log_rows, uint64_t w,
                               hipStream_t st);
void launch_coset_weights(uint64_t n, uint64_t offset, uint64_t scale, uint64_t *out, hipStream_t st);
void launch_fill_splitmix(uint64_t *out, uint64_t n, uint64_t m, uint64_t cfg, uint64_t table, hipStream_t st);
void launch_copy16(void *dst, const void *src, uint64_t n16, hipStream_t st);
void launch_open(const uint64_t *nodes, uint64_t L, const uint64_t *lde, uint64_t w, const uint64_t *pos,
                 uint64_t count, uint32_t depth, uint64_t *paths, uint64_t *rows, hipStream_t st);
}  // namespace lvm

using namespace lvm;

namespace {

bool is_pow2(size_t x) { return x && !(x & (x - 1)); }
unsigned log2u(size_t x) { return (unsigned)__builtin_ctzll(x); }

// RAII device buffer freed in stream order
struct DevBuf {
    void *p = nullptr;
    hipStream_t st = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) lvm_free(p);
    }
    int alloc(size_t bytes, hipStream_t s) {
        st = s;
        if (bytes == 0) bytes = 8;
        hipError_t e = lvm_malloc_bytes(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(LVM_ERR_NOMEM, std::string("device allocation of ") + std::to_string(bytes) +
                                           " bytes failed: " + hipGetErrorString(e));
        }
        return LVM_OK;
    }
    void *release() {
        void *q = p;
        p = nullptr;
        return q;
    }
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

#define TRY(x)                       \
    do {                             \
        int _rc = (x);               \
        if (_rc != LVM_OK) return _rc; \
    } while (0)

Device *dev_or_fail() {
    Device *d = device_state();
    return d;
}

int check_fault() {
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    return LVM_OK;
}

// Shape of a coset LDE: checked before any device work (a table of one row
// has no transform; the domain must fit the 2^32 two-adic subgroup).
int check_lde_shape(size_t n, size_t w, size_t blowup) {
    if (!is_pow2(n) || n < 2) return fail(LVM_ERR_ARG, "trace rows must be a power of two >= 2");
    if (!is_pow2(blowup)) return fail(LVM_ERR_ARG, "blowup must be a power of two");
    if (w == 0) return fail(LVM_ERR_ARG, "no columns");
    if (w > (1u << 30)) return fail(LVM_ERR_ARG, "too many columns");
    if (log2u(n) + log2u(blowup) > 32) return fail(LVM_ERR_ARG, "LDE domain exceeds the 2^32 two-adic subgroup");
    return LVM_OK;
}

// LDE + commit from device-resident trace rows (n x w words) and weights.
// Coset LDE of a device trace and its Merkle tree over the row-pair leaves of
// columns [split, w) (all columns when split = 0). With split > 0 a second
// tree over columns [0, split) is built too (a preprocessed table's fixed
// columns): its root goes to pre_root_out and, when pre_nodes_host is set,
// its node buffer to the host.
int lde_core(Device *d, const uint64_t *d_trace, size_t n, size_t w, size_t blowup, const uint64_t *d_weights,
             lvm_lde **out, uint8_t root_out[32], size_t split = 0, uint8_t *pre_nodes_host = nullptr,
             uint8_t *pre_root_out = nullptr) {
    TRY(check_lde_shape(n, w, blowup));
    const size_t N = n * blowup;
    TRY(check_fault());
    hipStream_t st = thread_stream(d);
    const size_t L = N / 2;
    DevBuf coef, lde, nodes, pre;
    TRY(coef.alloc(n * w * 8, st));
    TRY(lde.alloc(N * w * 8, st));
    TRY(nodes.alloc((2 * L - 1) * 32, st));
    if (split) TRY(pre.alloc((2 * L - 1) * 32, st));
    {
        ProfScope ps("ntt_lde", st);
        ntt_inverse_rows(d_trace, (uint32_t)w, coef.as<uint64_t>(), (uint32_t)w, (uint32_t)w, log2u(n), d->tb, st);
        TRY(ntt_coset_expand(coef.as<uint64_t>(), (uint32_t)w, lde.as<uint64_t>(), (uint32_t)w, (uint32_t)w,
                             log2u(n), log2u(blowup), d_weights, d->tb, st));
    }
    {
        ProfScope ps("keccak_leaves", st);
        keccak_row_pair_leaves(lde.as<uint64_t>(), w, split, w, L, nodes.as<uint8_t>() + (L - 1) * 32, st);
        if (split) keccak_row_pair_leaves(lde.as<uint64_t>(), w, 0, split, L, pre.as<uint8_t>() + (L - 1) * 32, st);
    }
    {
        ProfScope ps("merkle_inner", st);
        merkle_build_inner(nodes.as<uint8_t>(), L, st);
        if (split) merkle_build_inner(pre.as<uint8_t>(), L, st);
    }
    LVM_HIP(hipGetLastError());
    uint8_t root[32];
    {
        StageScope ss(st);
        LVM_HIP(stage_download(root, nodes.as<uint8_t>(), 32, st));
        if (split && pre_root_out) LVM_HIP(stage_download(pre_root_out, pre.as<uint8_t>(), 32, st));
        LVM_HIP(stage_sync(st));
        if (split && pre_nodes_host) {
            LVM_HIP(hipMemcpyAsync(pre_nodes_host, pre.p, (2 * L - 1) * 32, hipMemcpyDeviceToHost, st));
            LVM_HIP(hipStreamSynchronize(st));
        }
    }
    lvm_lde *h = new lvm_lde();
    h->device = d->id;
    h->lde = (uint64_t *)lde.release();
    h->nodes = (uint8_t *)nodes.release();
    h->n = n;
    h->N = N;
    h->w = w;
    h->blowup = blowup;
    h->L = L;
    std::memcpy(h->root, root, 32);
    if (root_out) std::memcpy(root_out, root, 32);
    *out = h;
    return LVM_OK;
}

int lde_from_host(const uint64_t *trace, size_t n, size_t w, size_t blowup, const uint64_t *weights,
                  uint64_t *host_lde_out, uint8_t root_out[32], lvm_lde **handle_out) {
    if (!trace || !weights || !handle_out) return fail(LVM_ERR_ARG, "null argument");
    TRY(check_lde_shape(n, w, blowup));
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    hipStream_t st = thread_stream(d);
    DevBuf dt, dw;
    TRY(dt.alloc(n * w * 8, st));
    TRY(dw.alloc(n * 8, st));
    LVM_HIP(hipMemcpyAsync(dt.p, trace, n * w * 8, hipMemcpyHostToDevice, st));
    LVM_HIP(hipMemcpyAsync(dw.p, weights, n * 8, hipMemcpyHostToDevice, st));
    lvm_lde *h = nullptr;
    TRY(lde_core(d, dt.as<uint64_t>(), n, w, blowup, dw.as<uint64_t>(), &h, root_out));
    if (host_lde_out) {
        int rc = lvm_lde_download(h, host_lde_out);
        if (rc != LVM_OK) {
            lvm_lde_release(h);
            return rc;
        }
    }
    *handle_out = h;
    return LVM_OK;
}

}  // namespace

extern "C" {

int lvm_init(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(LVM_ERR_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(LVM_ERR_ARG, "device index out of range");
    LVM_HIP(hipSetDevice(device));
    if (!device_state()) return LVM_ERR_DEVICE;
    return grind_scratch_setup(device);
}

int lvm_device_count(int *count) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return LVM_OK;
}

const char *lvm_last_error(void) { return last_error(); }

int lvm_trim(void) {
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    lvm_release_cached();
    return LVM_OK;
}

int lvm_sync(void) {
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

void lvm_inject_fault_after(long calls) { inject_fault_after(calls); }

int lvm_dev_alloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(LVM_ERR_ARG, "null ptr");
    if (!dev_or_fail()) return LVM_ERR_DEVICE;
    hipError_t e = lvm_malloc(ptr, bytes);
    if (e == hipSuccess) e = hipStreamSynchronize(lib_stream());
    if (e != hipSuccess) return fail(LVM_ERR_NOMEM, hipGetErrorString(e));
    return LVM_OK;
}

int lvm_dev_free(void *ptr) {
    if (!ptr) return LVM_OK;
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    LVM_HIP(lvm_free(ptr));
    return LVM_OK;
}

int lvm_memcpy_htod(void *dst, const void *src, size_t bytes) {
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    LVM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, thread_stream(d)));
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    LVM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, thread_stream(d)));
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

static int field_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n, bool ext3) {
    if (!a || !out || op < 0 || op > 5) return fail(LVM_ERR_ARG, "bad field op arguments");
    if ((op <= 2) && !b) return fail(LVM_ERR_ARG, "binary op needs b");
    if (n == 0) return LVM_OK;
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    hipStream_t st = thread_stream(d);
    size_t words = ext3 ? 3 * n : n;
    DevBuf da, db, dout;
    TRY(da.alloc(words * 8, st));
    TRY(dout.alloc(words * 8, st));
    LVM_HIP(hipMemcpyAsync(da.p, a, words * 8, hipMemcpyHostToDevice, st));
    if (b && op <= 2) {
        TRY(db.alloc(words * 8, st));
        LVM_HIP(hipMemcpyAsync(db.p, b, words * 8, hipMemcpyHostToDevice, st));
    }
    launch_gl_op(op, da.as<uint64_t>(), db.as<uint64_t>(), dout.as<uint64_t>(), n, ext3, st);
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipMemcpyAsync(out, dout.p, words * 8, hipMemcpyDeviceToHost, st));
    LVM_HIP(hipStreamSynchronize(st));
    return LVM_OK;
}

int lvm_gl_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
    return field_op(op, a, b, out, n, false);
}

int lvm_fp3_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
    return field_op(op, a, b, out, n, true);
}

static int ntt_host(const uint64_t *in, uint64_t *out, size_t n, size_t m, bool inverse) {
    if (!in || !out || !is_pow2(n) || m == 0) return fail(LVM_ERR_ARG, "bad NTT arguments");
    if (log2u(n) > 32) return fail(LVM_ERR_ARG, "NTT size exceeds 2^32");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    TRY(check_fault());
    hipStream_t st = thread_stream(d);
    DevBuf a, b;
    TRY(a.alloc(n * m * 8, st));
    TRY(b.alloc(n * m * 8, st));
    LVM_HIP(hipMemcpyAsync(a.p, in, n * m * 8, hipMemcpyHostToDevice, st));
    if (n == 1) {
        LVM_HIP(hipMemcpyAsync(out, a.p, m * 8, hipMemcpyDeviceToHost, st));
        LVM_HIP(hipStreamSynchronize(st));
        return LVM_OK;
    }
    if (inverse)
        ntt_inverse_rows(a.as<uint64_t>(), (uint32_t)m, b.as<uint64_t>(), (uint32_t)m, (uint32_t)m, log2u(n), d->tb, st);
    else
        ntt_forward_rows(a.as<uint64_t>(), (uint32_t)m, b.as<uint64_t>(), (uint32_t)m, (uint32_t)m, log2u(n), d->tb, st);
    launch_gather_bitrev_rows(b.as<uint64_t>(), a.as<uint64_t>(), n, log2u(n), m, st);
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipMemcpyAsync(out, a.p, n * m * 8, hipMemcpyDeviceToHost, st));
    LVM_HIP(hipStreamSynchronize(st));
    return LVM_OK;
}

int lvm_ntt_forward(const uint64_t *in, uint64_t *out, size_t n, size_t m) { return ntt_host(in, out, n, m, false); }
int lvm_ntt_inverse(const uint64_t *in, uint64_t *out, size_t n, size_t m) { return ntt_host(in, out, n, m, true); }

int lvm_coset_lde_row_major_with_merkle_tree_keep(const uint64_t *trace, size_t n, size_t m, size_t blowup,
                                                   const uint64_t *weights, uint64_t *host_lde_out,
                                                   uint8_t root_out[32], lvm_lde **handle_out) {
    return lde_from_host(trace, n, m, blowup, weights, host_lde_out, root_out, handle_out);
}

int lvm_coset_lde_ext3_row_major_with_merkle_tree_keep(const uint64_t *trace, size_t n, size_t m,
                                                        size_t blowup, const uint64_t *weights,
                                                        uint64_t *host_lde_out, uint8_t root_out[32],
                                                        lvm_lde **handle_out) {
    // An Fp3 column is three base columns under a base-field DFT and base
    // weights, and its 24-byte leaf encoding is the three components in order.
    return lde_from_host(trace, n, 3 * m, blowup, weights, host_lde_out, root_out, handle_out);
}

int lvm_coset_lde_dev(const uint64_t *dev_trace, size_t n, size_t m, size_t blowup, const uint64_t *dev_weights,
                      uint8_t root_out[32], lvm_lde **handle_out) {
    if (!dev_trace || !dev_weights || !handle_out) return fail(LVM_ERR_ARG, "null argument");
    TRY(check_lde_shape(n, m, blowup));
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    return lde_core(d, dev_trace, n, m, blowup, dev_weights, handle_out, root_out);
}

int lvm_coset_lde_split_trees_dev(const uint64_t *dev_trace, size_t n, size_t m, size_t blowup,
                                  const uint64_t *dev_weights, size_t split_col, uint8_t *pre_nodes_out,
                                  uint8_t pre_root_out[32], uint8_t root_out[32], lvm_lde **handle_out) {
    if (!dev_trace || !dev_weights || !handle_out || !pre_root_out) return fail(LVM_ERR_ARG, "null argument");
    TRY(check_lde_shape(n, m, blowup));
    if (split_col == 0 || split_col >= m) return fail(LVM_ERR_ARG, "split column must fall inside the row");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    return lde_core(d, dev_trace, n, m, blowup, dev_weights, handle_out, root_out, split_col, pre_nodes_out,
                    pre_root_out);
}

namespace {
// Row-pair leaves and tree over N bit-reversed rows of w words that the new
// handle takes over (rows is released into it on success).
int commit_owned_rows(Device *d, hipStream_t st, DevBuf &rows, size_t N, size_t w, size_t trace_rows,
                      uint8_t root_out[32], lvm_lde **handle_out) {
    const size_t L = N / 2;
    DevBuf nodes;
    TRY(nodes.alloc((2 * L - 1) * 32, st));
    {
        ProfScope ps("keccak_leaves", st);
        keccak_row_pair_leaves(rows.as<uint64_t>(), w, 0, w, L, nodes.as<uint8_t>() + (L - 1) * 32, st);
    }
    {
        ProfScope ps("merkle_inner", st);
        merkle_build_inner(nodes.as<uint8_t>(), L, st);
    }
    LVM_HIP(hipGetLastError());
    uint8_t root[32];
    {
        StageScope ss(st);
        LVM_HIP(stage_download(root, nodes.as<uint8_t>(), 32, st));
        LVM_HIP(stage_sync(st));
    }
    lvm_lde *h = new lvm_lde();
    h->device = d->id;
    h->lde = (uint64_t *)rows.release();
    h->nodes = (uint8_t *)nodes.release();
    h->n = trace_rows;
    h->N = N;
    h->w = w;
    h->blowup = N / trace_rows;
    h->L = L;
    std::memcpy(h->root, root, 32);
    if (root_out) std::memcpy(root_out, root, 32);
    *handle_out = h;
    return LVM_OK;
}
}  // namespace

int lvm_commit_rows_dev(const uint64_t *dev_rows, size_t N, size_t w, size_t trace_rows, uint8_t root_out[32],
                        lvm_lde **handle_out) {
    if (!dev_rows || !handle_out) return fail(LVM_ERR_ARG, "null argument");
    if (!is_pow2(N) || N < 2 || log2u(N) > 32) return fail(LVM_ERR_ARG, "rows must be a power of two in [2, 2^32]");
    if (!is_pow2(trace_rows) || trace_rows > N) return fail(LVM_ERR_ARG, "trace rows must be a power of two <= rows");
    if (w == 0 || w > (1u << 30)) return fail(LVM_ERR_ARG, "bad row width");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    TRY(check_fault());
    hipStream_t st = thread_stream(d);
    DevBuf rows;
    TRY(rows.alloc(N * w * 8, st));
    LVM_HIP(hipMemcpyAsync(rows.p, dev_rows, N * w * 8, hipMemcpyDeviceToDevice, st));
    return commit_owned_rows(d, st, rows, N, w, trace_rows, root_out, handle_out);
}

int lvm_lde_release(lvm_lde *h) {
    if (!h) return LVM_OK;
    lvm_free(h->lde);
    lvm_free(h->nodes);
    delete h;
    return LVM_OK;
}

int lvm_lde_root(const lvm_lde *h, uint8_t root_out[32]) {
    if (!h || !root_out) return fail(LVM_ERR_ARG, "null argument");
    std::memcpy(root_out, h->root, 32);
    return LVM_OK;
}

int lvm_lde_shape(const lvm_lde *h, size_t *lde_rows, size_t *words_per_row, size_t *trace_rows) {
    if (!h) return fail(LVM_ERR_ARG, "null handle");
    if (lde_rows) *lde_rows = h->N;
    if (words_per_row) *words_per_row = h->w;
    if (trace_rows) *trace_rows = h->n;
    return LVM_OK;
}

int lvm_lde_download(const lvm_lde *h, uint64_t *host_out) {
    if (!h || !host_out) return fail(LVM_ERR_ARG, "null argument");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    hipStream_t st = thread_stream(d);
    DevBuf tmp;
    TRY(tmp.alloc(h->N * h->w * 8, st));
    launch_gather_bitrev_rows(h->lde, tmp.as<uint64_t>(), h->N, log2u(h->N), h->w, st);
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipMemcpyAsync(host_out, tmp.p, h->N * h->w * 8, hipMemcpyDeviceToHost, st));
    LVM_HIP(hipStreamSynchronize(st));
    return LVM_OK;
}

int lvm_lde_device_ptr(const lvm_lde *h, const uint64_t **dev_ptr) {
    if (!h || !dev_ptr) return fail(LVM_ERR_ARG, "null argument");
    *dev_ptr = h->lde;
    return LVM_OK;
}

int lvm_lde_tree_nodes(const lvm_lde *h, uint8_t *nodes_out, size_t *num_nodes) {
    if (!h) return fail(LVM_ERR_ARG, "null handle");
    if (num_nodes) *num_nodes = 2 * h->L - 1;
    if (!nodes_out) return LVM_OK;
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    LVM_HIP(hipMemcpyAsync(nodes_out, h->nodes, (2 * h->L - 1) * 32, hipMemcpyDeviceToHost, thread_stream(d)));
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_lde_open(const lvm_lde *h, const uint64_t *positions, size_t count, uint8_t *paths_out, uint64_t *rows_out) {
    if (!h || (!positions && count) || (!paths_out && count)) return fail(LVM_ERR_ARG, "null argument");
    for (size_t i = 0; i < count; i++)
        if (positions[i] >= h->L) return fail(LVM_ERR_ARG, "opening position out of range");
    if (count == 0) return LVM_OK;
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    hipStream_t st = thread_stream(d);
    const uint32_t depth = log2u(h->L);
    DevBuf dp, dpaths, drows;
    TRY(dp.alloc(count * 8, st));
    TRY(dpaths.alloc(count * (depth ? depth : 1) * 32, st));
    if (rows_out) TRY(drows.alloc(count * 2 * h->w * 8, st));
    StageScope ss(st);
    LVM_HIP(stage_upload(dp.p, positions, count * 8, st));
    launch_open((const uint64_t *)h->nodes, h->L, h->lde, h->w, dp.as<uint64_t>(), count, depth,
                dpaths.as<uint64_t>(), rows_out ? drows.as<uint64_t>() : nullptr, st);
    LVM_HIP(hipGetLastError());
    if (depth) LVM_HIP(stage_download(paths_out, dpaths.p, count * depth * 32, st));
    if (rows_out) LVM_HIP(stage_download(rows_out, drows.p, count * 2 * h->w * 8, st));
    LVM_HIP(stage_sync(st));
    return LVM_OK;
}

int lvm_coset_weights(size_t n, uint64_t offset, uint64_t *dev_out) {
    if (!dev_out || !is_pow2(n)) return fail(LVM_ERR_ARG, "bad weights arguments");
    if (canon(offset) == 0) return fail(LVM_ERR_ARG, "coset offset must be non-zero");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    launch_coset_weights(n, offset, inv((uint64_t)n), dev_out, thread_stream(d));
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_composition_commit_d2(const uint64_t *dev_h, size_t N, size_t trace_rows, uint64_t coset_offset,
                              uint8_t root_out[32], lvm_lde **parts_out) {
    if (!dev_h || !parts_out) return fail(LVM_ERR_ARG, "null argument");
    if (!is_pow2(N) || N < 4) return fail(LVM_ERR_ARG, "LDE size must be a power of two >= 4");
    if (log2u(N) > 32) return fail(LVM_ERR_ARG, "LDE domain exceeds the 2^32 subgroup");
    if (!is_pow2(trace_rows) || trace_rows < 2 || trace_rows > N / 2)
        return fail(LVM_ERR_ARG, "trace rows must be a power of two in [2, N/2]");
    if (canon(coset_offset) == 0) return fail(LVM_ERR_ARG, "coset offset must be non-zero");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    TRY(check_fault());
    hipStream_t st = thread_stream(d);
    const size_t half = N / 2;
    DevBuf halves, w;
    TRY(halves.alloc(half * 6 * 8, st));
    TRY(w.alloc(half * 8, st));
    {
        ProfScope ps("d2_split", st);
        launch_d2_split(dev_h, log2u(N), coset_offset, d->tb, halves.as<uint64_t>(), st);
    }
    // weights g^-j / (N/2): interpolate on the g^2-coset, evaluate on the g-coset
    // (crypto/stark/src/prover.rs:408-447 CompositionLdeTwiddles)
    launch_coset_weights(half, inv(coset_offset), inv((uint64_t)half), w.as<uint64_t>(), st);
    LVM_HIP(hipGetLastError());
    TRY(lde_core(d, halves.as<uint64_t>(), half, 6, 2, w.as<uint64_t>(), parts_out, root_out));
    // the parts interpolate the trace degree bound: their out-of-domain values
    // read the n trace-coset rows (crypto/stark/src/prover.rs:2012-2100)
    (*parts_out)->n = trace_rows;
    (*parts_out)->blowup = N / trace_rows;
    return LVM_OK;
}

int lvm_composition_commit_parts(const uint64_t *dev_h, size_t N, size_t trace_rows, size_t num_parts,
                                 uint64_t coset_offset, uint8_t root_out[32], lvm_lde **parts_out) {
    if (!dev_h || !parts_out) return fail(LVM_ERR_ARG, "null argument");
    if (!is_pow2(N) || N < 4 || log2u(N) > 32) return fail(LVM_ERR_ARG, "LDE size must be a power of two in [4, 2^32]");
    if (!is_pow2(trace_rows) || trace_rows < 2 || trace_rows > N / 2)
        return fail(LVM_ERR_ARG, "trace rows must be a power of two in [2, N/2]");
    if (num_parts == 0 || num_parts > 8) return fail(LVM_ERR_ARG, "composition parts must be 1..8");
    if (canon(coset_offset) == 0) return fail(LVM_ERR_ARG, "coset offset must be non-zero");
    if (num_parts == 1) return lvm_commit_rows_dev(dev_h, N, 3, trace_rows, root_out, parts_out);
    if (num_parts == 2) return lvm_composition_commit_d2(dev_h, N, trace_rows, coset_offset, root_out, parts_out);
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    TRY(check_fault());
    hipStream_t st = thread_stream(d);
    const unsigned log_N = log2u(N);
    const uint32_t P = (uint32_t)num_parts;
    DevBuf nat, coef, wt, scat, lde;
    TRY(nat.alloc(N * 3 * 8, st));
    TRY(coef.alloc(N * 3 * 8, st));
    TRY(wt.alloc(N * 8, st));
    TRY(scat.alloc(N * 3 * P * 8, st));
    TRY(lde.alloc(N * 3 * P * 8, st));
    // offset^-i for the part shifts, and the per-row factor N^-1 offset^(t(1-P))
    uint64_t gi[8];
    const uint64_t ginv = inv(coset_offset);
    gi[0] = 1;
    for (uint32_t i = 1; i < P; i++) gi[i] = mul(gi[i - 1], ginv);
    const uint64_t step = pow(ginv, P - 1);
    {
        ProfScope ps("composition_parts", st);
        launch_gather_bitrev_rows(dev_h, nat.as<uint64_t>(), N, log_N, 3, st);
        ntt_inverse_rows(nat.as<uint64_t>(), 3, coef.as<uint64_t>(), 3, 3, log_N, d->tb, st);
        launch_coset_weights(N, step, inv((uint64_t)N), wt.as<uint64_t>(), st);
        launch_parts_scatter(coef.as<uint64_t>(), log_N, P, wt.as<uint64_t>(), gi, scat.as<uint64_t>(), st);
        ntt_forward_rows(scat.as<uint64_t>(), 3 * P, lde.as<uint64_t>(), 3 * P, 3 * P, log_N, d->tb, st);
    }
    LVM_HIP(hipGetLastError());
    return commit_owned_rows(d, st, lde, N, 3 * P, trace_rows, root_out, parts_out);
}

int lvm_fill_splitmix_table(uint64_t *dev_out, size_t n, size_t m, uint64_t cfg, uint64_t table) {
    if (!dev_out && n * m) return fail(LVM_ERR_ARG, "null output");
    if (m && n > ((size_t)1 << 40) / m) return fail(LVM_ERR_ARG, "table too large");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    launch_fill_splitmix(dev_out, n, m, cfg, table, thread_stream(d));
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_dev_copy(void *dst, const void *src, size_t bytes) {
    if ((!dst || !src) && bytes) return fail(LVM_ERR_ARG, "null pointer");
    if ((bytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return fail(LVM_ERR_ARG, "copy needs 16-byte aligned pointers and sizes");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    {
        ProfScope ps("dev_copy", thread_stream(d));
        launch_copy16(dst, src, bytes / 16, thread_stream(d));
    }
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_build_merkle_tree(const uint8_t *leaves, size_t num_leaves, uint8_t *nodes_out) {
    if (!leaves || !nodes_out || num_leaves == 0) return fail(LVM_ERR_ARG, "bad merkle arguments");
    Device *d = dev_or_fail();
    if (!d) return LVM_ERR_DEVICE;
    TRY(check_fault());
    hipStream_t st = thread_stream(d);
    const size_t L = padded_leaves(num_leaves);
    std::vector<uint8_t> padded(L * 32);
    std::memcpy(padded.data(), leaves, num_leaves * 32);
    for (size_t i = num_leaves; i < L; i++) std::memcpy(&padded[i * 32], leaves + (num_leaves - 1) * 32, 32);
    DevBuf nodes;
    TRY(nodes.alloc((2 * L - 1) * 32, st));
    LVM_HIP(hipMemcpyAsync(nodes.as<uint8_t>() + (L - 1) * 32, padded.data(), L * 32, hipMemcpyHostToDevice, st));
    merkle_build_inner(nodes.as<uint8_t>(), L, st);
    LVM_HIP(hipGetLastError());
    LVM_HIP(hipMemcpyAsync(nodes_out, nodes.p, (2 * L - 1) * 32, hipMemcpyDeviceToHost, st));
    LVM_HIP(hipStreamSynchronize(st));
    return LVM_OK;
}

int lvm_prof_enable(int on) {
    prof_set(on != 0);
    return LVM_OK;
}

int lvm_prof_reset(void) {
    prof_reset();
    return LVM_OK;
}

int lvm_prof_query(const char *name, uint64_t *count, double *total_ms) {
    if (!count || !total_ms) return fail(LVM_ERR_ARG, "null argument");
    return prof_query(name, count, total_ms);
}

}  // extern "C"
