/*
 * lvm_gpu.h — C ABI of the MI355X (gfx950) STARK proving hot path.
 *
 * Drop-in boundary for the GPU backend that crypto/stark calls through its
 * `try_*` dispatch functions (crypto/stark/src/gpu_lde.rs, logup_gpu.rs,
 * constraint_ir/gpu_interp.rs), replacing the CUDA backend crate
 * crypto/math-cuda. Every entry point returns an int status (0 = success);
 * a non-zero status means "decline": the caller runs its CPU arm, exactly as
 * the reference's `Option`/`Result<_, DriverError>` contract does
 * (crypto/stark/src/gpu_lde.rs:313-335). Nothing in the library aborts.
 *
 * Conventions (same limbs and byte orders as the reference):
 *  - base field elements: raw u64 limbs of Goldilocks p = 2^64 - 2^32 + 1,
 *    possibly non-canonical on input; outputs compare equal canonically;
 *  - an Fp3 element (Fp[w]/(w^3 - 2)) is 3 consecutive u64, components 0,1,2;
 *  - host tables are row-major (row r = `m` consecutive elements);
 *  - Merkle nodes are 32-byte Keccak-256 digests; a node buffer holds 2L-1
 *    nodes, root first, leaves at [L-1, 2L-1)
 *    (crypto/crypto/src/merkle_tree/merkle.rs:201-224).
 * Device-resident LDE tables keep rows in bit-reversed order (the order the
 * Merkle leaves, FRI and the barycentric sub-coset read them in); the host
 * copies returned by the download calls are in natural order like the
 * reference's host LDE.
 * All calls are thread-safe; calls on one device are ordered on the
 * library's stream for that device.
 */
#ifndef LVM_GPU_H
#define LVM_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- status / lifecycle ---------------- */
#define LVM_OK 0
#define LVM_ERR_ARG 1
#define LVM_ERR_DEVICE 2
#define LVM_ERR_NOMEM 3
#define LVM_ERR_INJECTED 4
#define LVM_ERR_STATE 5

/* Select and initialise a device (tables, stream). Replaces the backend
 * singleton `backend()` of crypto/math-cuda/src/device.rs:592-619. */
int lvm_init(int device);
/* Number of visible GPUs (0 when none). */
int lvm_device_count(int *count);
/* Thread-local message of the last failing call. */
const char *lvm_last_error(void);
/* Wait for all work the library queued on the current device. */
int lvm_sync(void);
/* Test-only fault injection (crypto/math-cuda/src/faults.rs:33-51): the
 * `calls`-th fallible entry point from now fails with LVM_ERR_INJECTED;
 * negative disarms. */
void lvm_inject_fault_after(long calls);

/* ---------------- device memory (for resident pipelines) ---------------- */
int lvm_dev_alloc(void **ptr, size_t bytes);
int lvm_dev_free(void *ptr);
int lvm_memcpy_htod(void *dst, const void *src, size_t bytes);
int lvm_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain 16-byte-vector kernel (16-byte aligned
 * pointers and size); also the calibration kernel for HBM traffic counters. */
int lvm_dev_copy(void *dst, const void *src, size_t bytes);

/* ---------------- field arithmetic test kernels ----------------
 * crypto/math-cuda/src/lib.rs:34-197 (gl_{add,sub,mul,neg}, ext3_{mul,add,sub}).
 * op: 0 add, 1 sub, 2 mul, 3 neg, 4 inv, 5 square. Host arrays of n
 * elements (n Fp3 elements = 3n u64 for lvm_fp3_op). */
int lvm_gl_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);
int lvm_fp3_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);

/* ---------------- NTT ----------------
 * crypto/math-cuda/src/ntt.rs:50 `forward` and :57 `inverse`: natural order
 * in and out, `n` a power of two. The inverse is unnormalised (no 1/n), like
 * the reference's inverse twiddle pass (crypto/math/src/fft/two_half_fft.rs:
 * 51-81 with `inverse = true`). Batched over `m` row-major columns. */
int lvm_ntt_forward(const uint64_t *in, uint64_t *out, size_t n, size_t m);
int lvm_ntt_inverse(const uint64_t *in, uint64_t *out, size_t n, size_t m);

/* ---------------- coset LDE + Keccak Merkle commitment ---------------- */
typedef struct lvm_lde lvm_lde; /* GpuLdeBase / GpuLdeExt3 + GpuMerkleTree (math-cuda/src/lde.rs:923-988) */

/* crypto/math-cuda/src/lde.rs:674 `coset_lde_row_major_with_merkle_tree_keep`
 * (row_major, predev, n, m, blowup_factor, weights, retain_host_lde):
 * row-major n x m trace -> coset LDE on offset*<w_{n*B}> (the offset enters
 * through weights[i] = n^-1 * offset^i, crypto/stark/src/prover.rs:451-467),
 * row-pair bit-reversed Keccak leaves (crypto/stark/src/prover.rs:895-961),
 * full tree kept on the device. `host_lde_out` (nullable) receives the
 * natural-order row-major (n*B) x m LDE when non-null (retain_host_lde).
 * `trace` is a host pointer; lvm_coset_lde_dev below takes a device one
 * (the reference's `predev`). */
int lvm_coset_lde_row_major_with_merkle_tree_keep(const uint64_t *trace, size_t n, size_t m, size_t blowup,
                                                   const uint64_t *weights, uint64_t *host_lde_out,
                                                   uint8_t root_out[32], lvm_lde **handle_out);
/* Same with the trace and weights already in device memory (row-major n x m). */
int lvm_coset_lde_dev(const uint64_t *dev_trace, size_t n, size_t m, size_t blowup, const uint64_t *dev_weights,
                      uint8_t root_out[32], lvm_lde **handle_out);
/* Fp3 tables: crypto/math-cuda/src/lde.rs:855 `coset_lde_ext3_row_major_with_merkle_tree_keep`.
 * `trace` holds n rows of m Fp3 elements (3m u64 per row); leaves stream the
 * 24-byte BE encoding of each element (extensions_goldilocks.rs:497-503). */
int lvm_coset_lde_ext3_row_major_with_merkle_tree_keep(const uint64_t *trace, size_t n, size_t m,
                                                        size_t blowup, const uint64_t *weights,
                                                        uint64_t *host_lde_out, uint8_t root_out[32],
                                                        lvm_lde **handle_out);
int lvm_lde_release(lvm_lde *h);
int lvm_lde_root(const lvm_lde *h, uint8_t root_out[32]);
/* shape: rows of the LDE, base columns (Fp3 columns count 3), trace rows */
int lvm_lde_shape(const lvm_lde *h, size_t *lde_rows, size_t *words_per_row, size_t *trace_rows);
/* natural-order row-major host copy of the LDE ((n*B) x words_per_row) */
int lvm_lde_download(const lvm_lde *h, uint64_t *host_out);
/* device pointer of the LDE table (rows in bit-reversed order) */
int lvm_lde_device_ptr(const lvm_lde *h, const uint64_t **dev_ptr);
/* the whole node buffer ((2L-1) x 32 bytes) */
int lvm_lde_tree_nodes(const lvm_lde *h, uint8_t *nodes_out, size_t *num_nodes);
/* crypto/math-cuda/src/merkle.rs:358 `gather_merkle_paths_dev`: the
 * authentication path of each leaf position (depth log2 L nodes each) and
 * the two bit-reversed LDE rows br(2*pos), br(2*pos+1) of each opening
 * (crypto/stark/src/prover.rs:2690-2712). rows_out may be null. */
int lvm_lde_open(const lvm_lde *h, const uint64_t *positions, size_t count, uint8_t *paths_out,
                 uint64_t *rows_out);

/* Coset weights w_i = n^-1 * offset^i, i < n, written to device memory
 * (crypto/stark/src/prover.rs:451-467 `LdeTwiddles::new` coset_weights). */
int lvm_coset_weights(size_t n, uint64_t offset, uint64_t *dev_out);

/* ---------------- Keccak Merkle trees on host data ----------------
 * crypto/math-cuda/src/merkle.rs:316 `build_merkle_tree_on_device`: leaves
 * (num_leaves x 32 bytes) padded to a power of two by repeating the last
 * leaf (crypto/crypto/src/merkle_tree/utils.rs:49-54); nodes_out holds
 * (2L-1) x 32 bytes. */
int lvm_build_merkle_tree(const uint8_t *leaves, size_t num_leaves, uint8_t *nodes_out);

/* ---------------- profiling ---------------- */
/* When enabled, the library brackets its kernels with HIP events on the
 * stream they run on; query returns the count and summed duration of the
 * records with the given kernel name (all names when null). */
int lvm_prof_enable(int on);
int lvm_prof_reset(void);
int lvm_prof_query(const char *name, uint64_t *count, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
