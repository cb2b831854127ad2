// Internal declarations shared by the HIP translation units of liblvm_gpu.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "field.hpp"
#include "lvm_gpu.h"

// Resident LDE table + Merkle tree behind the opaque `lvm_lde` handle.
struct lvm_lde {
    int device;
    uint64_t *lde;  // N rows x w words, rows in bit-reversed order
    size_t n, N, w, blowup;
    uint8_t *nodes;  // (2L-1) x 32 bytes, root first
    size_t L;
    uint8_t root[32];
};

namespace lvm {

// Device twiddle tables (built once per device by the library).
//  *_lo[x] = w^x, *_hi[x] = w^(x * 2^16) for w the primitive 2^32-th root of
//  unity of crypto/math/src/field/goldilocks.rs:540 (fwd) or its inverse (inv);
//  *_r[e] = w_256^e (e < 128).
struct Tables {
    uint64_t *fwd_lo, *fwd_hi, *inv_lo, *inv_hi, *fwd_r, *inv_r;
};

// status codes of the C ABI (LVM_OK, LVM_ERR_*) come from lvm_gpu.h

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// per-device library state
struct Device {
    int id = -1;
    hipStream_t stream = nullptr;  // the first host thread's library stream
    std::atomic<bool> stream_claimed{false};
    Tables tb{};
};
Device *device_state();  // current device's state, initialised on first use; null on failure
// The calling host thread's library stream on device d: every thread that
// calls the library gets its own (default priority, non-blocking), so tables
// proved from several threads overlap on the GPU. Streams outlive their
// threads (an ending thread's stream is drained and reused by a later one). Work handed from one
// thread to another is complete when the producing call returns (every
// entry point that yields a handle or host data synchronises its stream).
hipStream_t thread_stream(Device *d);
// The calling thread's library stream if it already has one, else the
// device's first stream; creates nothing (for the allocation-free entry point).
hipStream_t thread_stream_if_any(Device *d);
// Resolves the per-device address of the nonce search's fixed scratch array
// (grinding.hip); lvm_init calls it with the device current.
int grind_scratch_setup(int device);

// fault injection for the fallback tests (never set in production)
bool take_injected_fault();

// Device allocations (context.cpp): a cache of freed blocks in size classes
// over the stream-ordered pool. A freed block is reused at once on the stream
// that freed it, and on another stream once the event recorded at its release
// has completed; the driver pool is only touched on a cache miss (its frees
// cost 0.1-0.3 ms of host time each when many blocks are live). When the
// device runs out, the cache is returned to the pool and the allocation retried.
hipStream_t lib_stream();
hipError_t lvm_malloc_bytes(void **p, size_t bytes);
hipError_t lvm_free(void *p);
void lvm_release_cached();
template <class T>
inline hipError_t lvm_malloc(T **p, size_t bytes) {
    return lvm_malloc_bytes((void **)p, bytes);
}

// Small host<->device transfers through a per-thread pinned staging area: a
// copy of pageable memory is staged by the runtime at tens of microseconds
// each, a copy of pinned memory is a plain DMA. stage_upload copies the host
// data into the area at the call and queues the DMA; stage_download queues the
// DMA into the area and the bytes reach dst_host at stage_sync, which
// synchronises the stream. A StageScope around an entry point drops (after a
// stream synchronisation) the downloads it queued but never delivered, so an
// early error return leaves nothing pending. Transfers above 4 MiB go direct.
hipError_t stage_upload(void *dst_dev, const void *src_host, size_t bytes, hipStream_t st);
hipError_t stage_download(void *dst_host, const void *src_dev, size_t bytes, hipStream_t st);
hipError_t stage_sync(hipStream_t st);
struct StageScope {
    explicit StageScope(hipStream_t st);
    ~StageScope();
    hipStream_t st;
    size_t mark;
};

#define LVM_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) return ::lvm::fail(LVM_ERR_DEVICE, std::string(#expr) + ": " +    \
                                                   hipGetErrorString(_e));                              \
    } while (0)

// ---- NTT (ntt.hip) ----
// Arguments of one NTT pass over one table (layout and remaps described in ntt.hip).
struct PassArgs {
    const uint64_t *in;
    uint64_t *out;
    uint32_t in_pitch, out_pitch;  // row strides in u64
    uint32_t m;                    // number of columns
    uint32_t log_msub;             // log2 of M (sub-problem rows at this pass)
    uint32_t log_s;                // log2 of the row stride S
    uint32_t in_bitrev_bits;       // 0: natural input rows; else bitrev over this many bits
    uint64_t total_groups;         // rows_logical / R
    uint64_t out_coset_rows;       // rows per coset block in `out` (n); 0 = single block
    uint32_t coset_bits;           // log2(blowup) for the per-coset output offset
    uint32_t shift_log;            // log2 N of the coset shift w_N^(coset*j); 0 = none
    const uint64_t *weights;       // nullable pre-multiply weights[logical row]
    const uint64_t *tw_lo;         // w_{2^32}^(+-x), x < 2^16
    const uint64_t *tw_hi;         // w_{2^32}^(+-x*2^16), x < 2^16
    const uint64_t *shift_lo;      // forward tables for the coset shift
    const uint64_t *shift_hi;
    const uint64_t *tw_r;          // w_256^(+-e), e < 256
    uint32_t cosets_per_block;     // cosets produced per block (grid.z = cosets / this)
    uint32_t inverse;              // 1: inverse transform (tw_* hold inverse powers)
    uint32_t in_canonical;         // 1: input values are below p (an earlier pass wrote them)
};

void ntt_inverse_rows(const uint64_t *in, uint32_t in_pitch, uint64_t *out, uint32_t out_pitch, uint32_t m,
                      unsigned log_n, const Tables &tb, hipStream_t st);
void ntt_forward_rows(const uint64_t *in, uint32_t in_pitch, uint64_t *out, uint32_t out_pitch, uint32_t m,
                      unsigned log_n, const Tables &tb, hipStream_t st);
int ntt_coset_expand(const uint64_t *coef, uint32_t coef_pitch, uint64_t *lde, uint32_t lde_pitch, uint32_t m,
                      unsigned log_n, unsigned log_blowup, const uint64_t *weights, const Tables &tb,
                      hipStream_t st);

// ---- Keccak / Merkle (merkle.hip) ----
// Row-pair leaves over words [w0, w1) of each row of a row-major table whose
// rows are already stored in bit-reversed order: leaf i hashes rows 2i, 2i+1.
void keccak_row_pair_leaves(const uint64_t *rows, uint64_t pitch, uint64_t w0, uint64_t w1, uint64_t num_leaves,
                            uint8_t *leaves_out, hipStream_t st);
// inner nodes of a (2L-1)*32-byte tree whose leaves sit at [L-1, 2L-1)
void merkle_build_inner(uint8_t *nodes, uint64_t L, hipStream_t st);
// Keccak(BE24(e[2j]) || BE24(e[2j+1])) over an Fp3 array of `len` elements
void keccak_fp3_pair_leaves(const uint64_t *evals, uint64_t len, uint8_t *leaves_out, hipStream_t st);

size_t padded_leaves(size_t n);

// ---- composition (compose.hip) ----
// degree-2 split of N bit-reversed Fp3 composition values into N/2
// natural-order rows [H0 | H1]
void launch_d2_split(const uint64_t *hb, unsigned log_N, uint64_t coset_offset, const Tables &tb, uint64_t *halves,
                     hipStream_t st);
// three or more composition parts: natural rows t of the P part coefficient
// columns (Fp3, scaled for evaluation on the coset) from the bit-reversed
// unnormalised inverse DFT `coef` of the N composition values; P <= 8
void launch_parts_scatter(const uint64_t *coef, unsigned log_N, uint32_t P, const uint64_t *wt,
                          const uint64_t *offset_inv_pows, uint64_t *out, hipStream_t st);

// ---- work lists over several tables (batched stages) ----
// Work-list slot of block b: the t with prefix[t] <= b < prefix[t+1]
// (prefix ascending, prefix[0] = 0), by binary search over `count` slots.
// Workgroups are dispatched to the 8 XCDs round-robin by block index, and
// each XCD has its own L2. This bijection on [0, nb) gives XCD x the
// contiguous logical range [x*q, (x+1)*q) (q = nb / 8; the nb % 8 tail blocks
// keep their index), so neighbouring logical tiles that share cache lines
// (the column groups of one row tile) are served by one L2.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3;
    return b < (q << 3) ? (b & 7u) * q + (b >> 3) : b;
}

__device__ __forceinline__ uint32_t find_slot(const uint32_t *__restrict__ prefix, uint32_t count, uint32_t b) {
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
int plan_passes(unsigned log_n, int *r);              // radices of a size-2^log_n transform
bool set_ntt_plan(unsigned log_n, const int *r, int np);  // np = 0 restores the default
uint32_t pass_blocks(int r, const PassArgs &a);       // blocks of one table's pass
void launch_pass_batched(int r, const PassArgs *d_args, const uint32_t *d_prefix, uint32_t count, uint32_t blocks,
                         bool multi, hipStream_t st);
// One table's coset-LDE pass sequence (inverse passes into `coef`, then the
// coset expansion into `lde`); returns the number of passes.
int lde_pass_list(const uint64_t *trace, uint32_t m, unsigned log_n, unsigned log_blowup, const uint64_t *weights,
                  uint64_t *coef, uint64_t *lde, const Tables &tb, int *r_out, PassArgs *a_out, uint32_t *cos_out);

struct LeafJob {
    const uint64_t *rows;  // bit-reversed row-major LDE, `pitch` words per row
    uint64_t pitch;
    uint64_t num_leaves;
    uint64_t *leaves;      // num_leaves x 32 bytes
};
struct TreeStep {
    uint64_t *nodes;
    uint64_t cstart, pstart;  // first child / first parent index of the level
    uint32_t span;            // parents per block at the first level (<= 256)
    uint32_t levels;          // levels this step builds
};
void launch_leaves_batched(const LeafJob *d_jobs, const uint32_t *d_prefix, uint32_t count, uint32_t blocks,
                           hipStream_t st);
void launch_tree_steps_batched(const TreeStep *d_steps, const uint32_t *d_prefix, uint32_t count, uint32_t blocks,
                               hipStream_t st);
void launch_gather_roots(uint64_t *const *d_trees, uint32_t count, uint64_t *d_roots, hipStream_t st);
// One FRI fold of each of count codewords in shared launches (fri.hip);
// zetas: count x 3; roots_out (count x 32) when committing.
int fri_fold_batch(lvm_fri *const *fs, size_t count, const uint64_t *zetas, bool commit, uint8_t *roots_out);
// the launches of one tree of L leaves (a power of two) at `nodes`: (step, blocks)
std::vector<std::pair<TreeStep, uint32_t>> tree_steps(uint8_t *nodes, uint64_t L);

// Constraint programs (compose.hip interpreter, compose_jit.cpp compiled form).
struct CompArgs;
// Node computing each constraint (given, or the last writer of a pinned root slot).
bool root_nodes_of(const lvm_program *pg, std::vector<uint32_t> &root_node);
bool jit_enabled();
// Boundary constraints of a composition grouped by step, steps in
// first-seen order (the order of CompArgs::b_col).
struct BoundarySteps {
    uint32_t num_steps = 0;
    std::vector<uint32_t> step_off;  // num_steps + 1
    std::vector<uint32_t> col, is_aux;
    std::vector<uint64_t> step_row;  // trace row of each step
};
// false when there are more than COMP_MAX_STEPS distinct steps (mod trace_rows)
bool boundary_steps_of(const uint32_t *b_col, const uint32_t *b_is_aux, const uint64_t *b_step, size_t num_boundary,
                       uint64_t trace_rows, BoundarySteps &bs, std::vector<uint32_t> *order);
// Compiles the program's kernel (or finds it compiled); false when the
// generator does not take the program or the compile fails.
bool jit_prepare(const lvm_program *pg, const std::vector<uint32_t> &root_node, const uint32_t *root_group, int mode);
// Launches the compiled kernel on the current device; false on failure.
bool jit_launch(const lvm_program *pg, const std::vector<uint32_t> &root_node, const uint32_t *root_group, int mode,
                CompArgs &args, uint64_t blocks, hipStream_t st);
// The composition (mode 0) of several tables' programs as one kernel: compile
// (cached per program set), then launch with table t's arguments at d_args[t]
// and its blocks [d_prefix[t], d_prefix[t+1]) (128 natural rows each).
bool jit_prepare_batch(const std::vector<const lvm_program *> &pgs, const std::vector<std::vector<uint32_t>> &root_nodes,
                       const std::vector<const uint32_t *> &root_groups);
bool jit_launch_batch(const std::vector<const lvm_program *> &pgs, const std::vector<std::vector<uint32_t>> &root_nodes,
                      const std::vector<const uint32_t *> &root_groups, const CompArgs *d_args,
                      const uint32_t *d_prefix, uint32_t count, uint64_t blocks, hipStream_t st);
}  // namespace lvm
