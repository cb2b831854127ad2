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
#define LVM_ERR_UNSUPPORTED 6

/* Select and initialise a device (tables, stream). Replaces the backend
 * singleton `backend()` of crypto/math-cuda/src/device.rs:592-619. */
int lvm_init(int device);
/* Number of visible GPUs (0 when none). */
int lvm_device_count(int *count);
/* Thread-local message of the last failing call. */
const char *lvm_last_error(void);
/* Device buffers released by the library are kept in size classes and
 * reused (a freed block is reused on the stream that released it at once,
 * on other streams once that release has completed); this hands every
 * cached block of the current device back to the driver pool. No counterpart
 * in the reference (its buffers are freed on drop, math-cuda device.rs). */
int lvm_trim(void);
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
/* crypto/math-cuda/src/lde.rs:726 `coset_lde_row_major_split_trees` for a
 * preprocessed table (device trace: the reference's `predev`): one coset LDE
 * of all m columns and two subset trees over it, the fixed (precomputed)
 * columns [0, split_col) and the multiplicity columns [split_col, m), leaves =
 * the row-pair bit-reversed Keccak leaves of the column range
 * (crypto/stark/src/prover.rs:895-961). The handle holds the LDE (all
 * columns; lvm_lde_open returns whole rows) and the multiplicity tree, whose
 * root goes to root_out; pre_root_out receives the precomputed tree's root
 * and pre_nodes_out (nullable: the reference's build_precomputed = false) its
 * node buffer, (2L-1) x 32 bytes with L = n*blowup/2, for host-side openings
 * as in the reference. */
int lvm_coset_lde_split_trees_dev(const uint64_t *dev_trace, size_t n, size_t m, size_t blowup,
                                  const uint64_t *dev_weights, size_t split_col, uint8_t *pre_nodes_out,
                                  uint8_t pre_root_out[32], uint8_t root_out[32], lvm_lde **handle_out);
/* Fp3 tables: crypto/math-cuda/src/lde.rs:855 `coset_lde_ext3_row_major_with_merkle_tree_keep`.
 * `trace` holds n rows of m Fp3 elements (3m u64 per row); leaves stream the
 * 24-byte BE encoding of each element (extensions_goldilocks.rs:497-503). */
int lvm_coset_lde_ext3_row_major_with_merkle_tree_keep(const uint64_t *trace, size_t n, size_t m,
                                                        size_t blowup, const uint64_t *weights,
                                                        uint64_t *host_lde_out, uint8_t root_out[32],
                                                        lvm_lde **handle_out);
/* crypto/math-cuda/src/merkle.rs:546 `build_comp_poly_tree_from_evals_ext3_keep`
 * (with constraint_interp.rs:514 `comp_h_to_slabs`): commit device rows that
 * already are an LDE (N rows x w words, bit-reversed row order, e.g. the
 * composition values of lvm_eval_composition when the composition has one
 * part) with row-pair leaves; the handle keeps a copy. trace_rows: the degree
 * bound the rows interpolate (barycentric evaluation reads every
 * (N/trace_rows)-th row). */
int lvm_commit_rows_dev(const uint64_t *dev_rows, size_t N, size_t w, size_t trace_rows, uint8_t root_out[32],
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
/* Synthetic input (no counterpart in the reference; workload generation for
 * benchmarks and the proving service): row-major n x m table whose column c
 * is the SplitMix64 stream seeded with 0x1A3BDA7A ^ (cfg << 48) ^ (table << 32) ^ c
 * (the generator of crypto/stark/src/gpu_lde.rs:3409-3418), each draw reduced mod p. */
int lvm_fill_splitmix_table(uint64_t *dev_out, size_t n, size_t m, uint64_t cfg, uint64_t table);

/* ---------------- batched ragged commitment of many tables ----------------
 * The R1 commitment of all tables of a proof (crypto/stark/src/prover.rs
 * :3299-3492 `multi_prove`, which calls the per-table
 * `coset_lde_row_major_with_merkle_tree_keep`, math-cuda/src/lde.rs:674, once
 * per table) as one launch per stage over every table: each NTT pass shape,
 * the Keccak leaves and each Merkle tree step run as a single work-list
 * launch whose blocks are spread over the tables, and the whole sequence is
 * captured once as a HIP graph and replayed. Results are those of
 * lvm_coset_lde_dev on each table. Shapes are fixed when the plan is made;
 * inputs may be rebound to other buffers of the same shapes. */
typedef struct lvm_table_desc {
    const uint64_t *trace;   /* device, n x m row-major */
    const uint64_t *weights; /* device, the n LDE weights n^-1 * offset^i */
    size_t n, m, blowup;
} lvm_table_desc;
typedef struct lvm_batch lvm_batch;
/* use_graph: 1 = capture the launch sequence as a HIP graph, 0 = plain launches */
int lvm_batch_new(const lvm_table_desc *tables, size_t count, int use_graph, lvm_batch **out);
/* point the plan at other inputs with the same shapes (takes effect at the next run) */
int lvm_batch_bind(lvm_batch *b, const lvm_table_desc *tables, size_t count);
/* run every stage for all tables; roots_out (nullable) receives count x 32 bytes */
int lvm_batch_run(lvm_batch *b, uint8_t *roots_out);
/* hand table `index`'s LDE and tree over as an lvm_lde handle; the plan gets
 * fresh output buffers for that table at its next run */
int lvm_batch_take(lvm_batch *b, size_t index, lvm_lde **out);
int lvm_batch_release(lvm_batch *b);
/* the plan's launches in order: profiler name (the name each launch is
 * recorded under when profiling is on and the plan runs without a graph),
 * compulsory HBM bytes (inputs read once, outputs written once) and blocks */
int lvm_batch_launch_count(const lvm_batch *b, size_t *count);
int lvm_batch_launch_info(const lvm_batch *b, size_t index, char *name_out, size_t name_cap, double *bytes_out,
                          uint32_t *blocks_out);
/* one-shot form: plan, run, hand every table over (handles_out: count entries) */
int lvm_batch_lde_commit(const lvm_table_desc *tables, size_t count, uint8_t *roots_out, lvm_lde **handles_out);
/* NTT pass split for size 2^log_n (radices 1..8 summing to log_n; count 0
 * restores the default); used by every later transform of that size. */
int lvm_set_ntt_plan(unsigned log_n, const int *radices, int count);
int lvm_get_ntt_plan(unsigned log_n, int *radices_out, int *count_out);

/* ---------------- Keccak Merkle trees on host data ----------------
 * crypto/math-cuda/src/merkle.rs:316 `build_merkle_tree_on_device`: leaves
 * (num_leaves x 32 bytes) padded to a power of two by repeating the last
 * leaf (crypto/crypto/src/merkle_tree/utils.rs:49-54); nodes_out holds
 * (2L-1) x 32 bytes. */
int lvm_build_merkle_tree(const uint8_t *leaves, size_t num_leaves, uint8_t *nodes_out);

/* ---------------- batch inverse / barycentric OOD / DEEP ---------------- */
/* crypto/math-cuda/src/inverse.rs:53 `batch_inverse_ext3` (host arrays) and
 * :157 `batch_inverse_ext3_dev` (device arrays, in != out): elementwise Fp3
 * inverses by Montgomery's trick (crypto/math/src/field/element.rs:64-111). */
int lvm_batch_inverse_ext3(const uint64_t *in, uint64_t *out, size_t n);
int lvm_batch_inverse_ext3_dev(const uint64_t *dev_in, uint64_t *dev_out, size_t n);
/* crypto/math-cuda/src/barycentric.rs:379 `barycentric_base_multi_on_device`:
 * every column (word) of the resident table evaluated at K <= 8 Fp3 points by
 * the barycentric formula of crypto/stark/src/trace.rs:831-1110 over the
 * trace sub-coset offset*<w_n>; out: K x words_per_row Fp3 (point-major). For
 * an Fp3 table the three words of a column give its components c0, c1, c2
 * with value c0 + c1*w + c2*w^2. */
int lvm_barycentric_base_multi(const lvm_lde *h, uint64_t coset_offset, const uint64_t *z, size_t K, uint64_t *out);
/* crypto/math-cuda/src/deep.rs:316 `deep_composition_ext3_fully_resident_keep`
 * (formula of crypto/stark/src/prover.rs:2400-2604): the DEEP codeword over
 * the LDE coset of `main` (base columns) plus optional `aux` (Fp3 columns)
 * and `parts` (composition parts, Fp3 columns) tables on the same domain,
 * with K <= 4 trace poles z_points and pole z_parts (= z^P) for the parts.
 * trace_ood / gammas_tr: (main cols + aux cols) x K Fp3 (column-major,
 * (j*K + k)); parts_ood / gammas_h: parts Fp3. dev_out (device, N x 3 u64)
 * receives the codeword in bit-reversed order, the order FRI consumes. */
int lvm_deep_composition(const lvm_lde *main, const lvm_lde *aux, const lvm_lde *parts, uint64_t coset_offset,
                         const uint64_t *z_points, size_t K, const uint64_t *z_parts, const uint64_t *trace_ood,
                         const uint64_t *parts_ood, const uint64_t *gammas_tr, const uint64_t *gammas_h,
                         uint64_t *dev_out);

/* ---------------- LogUp auxiliary trace ----------------
 * Flat CSR descriptor of one table's bus interactions, field for field the
 * reference's crypto/math-cuda/src/logup.rs:26-44 `LogupDescriptor` (built by
 * crypto/stark/src/logup_gpu.rs:256-290 `build_fingerprint_descriptor`;
 * canonical coefficients):
 *   fp_k(i)   = z - (bus_ids[k] + sum_e alpha[elem_alpha_idx[e]] *
 *               (elem_const[e] + sum_t term_coef[t] * main(i, term_col[t])))
 *   m_k(i)    = mult_const[k] + sum_t mult_term_coef[t] * main(i, mult_term_col[t])
 *   term_c(i) = sum over out_col_interactions of column c of m_k(i) / fp_k(i)
 * with e over [elem_offsets[k], elem_offsets[k+1]), t over the term CSR. */
typedef struct lvm_logup_desc {
    size_t num_interactions;
    const uint64_t *bus_ids;
    const uint32_t *elem_offsets;
    const uint32_t *elem_alpha_idx;
    const uint64_t *elem_const;
    const uint32_t *term_offsets;
    const uint64_t *term_coef;
    const uint32_t *term_col;
    size_t num_out_cols; /* committed pairs + 1 (the absorbed column) */
    const uint32_t *out_col_offsets;
    const uint32_t *out_col_interactions;
    const uint64_t *mult_const;
    const uint32_t *mult_term_offsets;
    const uint64_t *mult_term_coef;
    const uint32_t *mult_term_col;
} lvm_logup_desc;
/* crypto/math-cuda/src/logup.rs:311 `logup_aux_resident`: the whole aux
 * build on the device from a resident main trace whose element (row, col)
 * is dev_main[row*row_stride + col*col_stride] (row-major: stride m, 1;
 * the reference's column-major: 1, num_rows). dev_aux_out receives the
 * row-major aux trace, num_out_cols Fp3 per row (committed term columns, then
 * the accumulated column: acc[0] = 0, acc[i+1] = acc[i] + sum_c term_c(i) - L/N,
 * crypto/stark/src/lookup.rs:1790-1834); dev_terms_out (nullable) every term
 * column [(c*num_rows + row)*3 + limb]; L_out the table contribution L.
 * alpha_powers: num_alpha Fp3 values (host). Descriptor arrays are host
 * arrays; every index in them is checked against the sizes given. */
int lvm_logup_aux_dev(const uint64_t *dev_main, size_t row_stride, size_t col_stride, size_t num_rows,
                      size_t num_cols, const lvm_logup_desc *desc, const uint64_t *alpha_powers, size_t num_alpha,
                      const uint64_t z[3], uint64_t *dev_aux_out, uint64_t *dev_terms_out, uint64_t L_out[3]);
/* crypto/math-cuda/src/logup.rs:133 `logup_term_columns`: host column-major
 * main (num_cols x num_rows) -> host term columns [(c*num_rows + row)*3 + limb]. */
int lvm_logup_term_columns(const uint64_t *main_cols, size_t num_rows, size_t num_cols, const lvm_logup_desc *desc,
                           const uint64_t *alpha_powers, size_t num_alpha, const uint64_t z[3], uint64_t *terms_out);

/* ---------------- constraint evaluation / composition polynomial ----------------
 * Program: the flat constraint program of crypto/stark/src/constraint_ir/
 * device.rs (`DeviceProgram`), packed as in crypto/stark/src/constraint_ir/
 * gpu_interp.rs:30-38: nodes 2 u64 each (op | a << 32, b | res << 32);
 * roots one u64 each (slot | ext_bit << 31); ext tables 3 u64 per element.
 * These are the inputs of crypto/math-cuda/src/constraint_interp.rs:53
 * `eval_constraints_on_device`. */
typedef struct lvm_program {
    const uint64_t *nodes;
    size_t num_nodes, num_base_slots, num_ext_slots;
    const uint64_t *base_consts;
    size_t num_base_consts;
    const uint64_t *ext_consts;
    size_t num_ext_consts;
    const uint64_t *roots;
    size_t num_roots;
    const uint64_t *rap_challenges;
    size_t num_rap;
    const uint64_t *alpha_powers;
    size_t num_alpha;
    const uint64_t *table_offset; /* 3 u64, nullable when unused */
    /* nullable: index of the node computing each constraint. When given, root
     * slots need not be pinned (each value is consumed when its node runs);
     * when null the roots' slots must be pinned, as the reference lowering does. */
    const uint32_t *root_nodes;
} lvm_program;
/* Composition inputs (crypto/math-cuda/src/constraint_interp.rs:157-177
 * `CompositionAccum`), with the zerofiers described by their parameters
 * instead of precomputed vectors: constraint c belongs to zerofier group
 * root_group[c], whose inverse zerofier is prod_{j=1..e}(x - w^(n-j)) / (x^n - 1)
 * with e = group_exemptions[g] (crypto/stark/src/constraints/zerofier.rs);
 * boundary constraint b contributes beta_b (t_b(x) - value_b) / (x - w^step_b)
 * (crypto/stark/src/domain.rs:112-131). x runs over coset_offset * <w_N>. */
typedef struct lvm_composition {
    size_t trace_rows;          /* n */
    uint64_t coset_offset;
    const uint64_t *beta;       /* num_roots Fp3 */
    const uint32_t *root_group; /* num_roots */
    const uint32_t *group_exemptions;
    size_t num_groups;          /* 1..4 */
    size_t num_boundary;
    const uint32_t *b_col;
    const uint32_t *b_is_aux;   /* 0 main column, 1 aux (Fp3) column */
    const uint64_t *b_step;     /* trace row of the constraint */
    const uint64_t *b_value;    /* Fp3 each */
    const uint64_t *b_beta;     /* Fp3 each */
} lvm_composition;
/* crypto/math-cuda/src/constraint_interp.rs:353/:409 `eval_composition_on_device(_keep)`:
 * H over the LDE of the resident main (base) and aux (Fp3, nullable) tables;
 * a frame offset k reads the row k * next_step further (next_step = blowup for
 * one trace step). dev_h_out: N Fp3 in device memory, rows in the tables'
 * bit-reversed order, or natural order when natural_order != 0. */
int lvm_eval_composition(const lvm_program *program, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                         const lvm_composition *comp, int natural_order, uint64_t *dev_h_out);
/* lvm_eval_composition for `count` tables (table t: programs[t], mains[t],
 * auxs[t] (nullable), next_steps[t], comps[t], output dev_h_out[t]) in
 * shared launches: the natural-order column copies of every table in one
 * launch per element width, the boundary step inverses of every table in one
 * launch, and every table's constraint program compiled into one kernel that
 * runs once over all tables' LDE rows; the per-table composition of
 * crypto/stark/src/prover.rs:3974-3977 (constraint_interp.rs:353 per table)
 * as one stage. Each output equals lvm_eval_composition's for that table; if
 * the run-time compiler is off or declines a program, the tables are
 * evaluated one by one. */
int lvm_eval_composition_batch(const lvm_program *const *programs, const lvm_lde *const *mains,
                               const lvm_lde *const *auxs, const size_t *next_steps,
                               const lvm_composition *const *comps, size_t count, int natural_order,
                               uint64_t *const *dev_h_out);
/* crypto/math-cuda/src/constraint_interp.rs:53 `eval_constraints_on_device`:
 * dev_evals_out [(c*N + row)*3 + limb] in natural row order; base-rooted
 * constraints carry their value in component 0. */
int lvm_eval_constraints(const lvm_program *program, const lvm_lde *main, const lvm_lde *aux, size_t next_step,
                         uint64_t *dev_evals_out);

/* Constraint programs run as kernels compiled at run time (hipRTC): the
 * program becomes straight-line code, compiled once per distinct program and
 * cached for the life of the process; programs the generator does not take
 * run on the interpreter kernel, with the same results. No counterpart in the
 * reference (its CUDA path interprets the program, constraint_interp.rs).
 * lvm_prepare_constraints compiles ahead of the first evaluation (comp null:
 * the per-constraint form of lvm_eval_constraints); it needs no device and
 * returns LVM_ERR_UNSUPPORTED for a program that is not compiled.
 * lvm_set_constraint_jit(0) makes every evaluation use the interpreter
 * (default on; the environment variable LVM_CONSTRAINT_JIT=0 does the same). */
int lvm_prepare_constraints(const lvm_program *program, const lvm_composition *comp);
int lvm_set_constraint_jit(int on);

/* Degree-2 decomposition and commitment of the composition polynomial
 * (crypto/stark/src/prover.rs:1605-1680 decompose_and_extend_d2 +
 * extend_half_to_lde; crypto/math-cuda/src/constraint_interp.rs:469
 * `decompose_d2_into_slabs` and merkle.rs:494 `build_comp_poly_tree_from_slabs_dev`):
 * dev_h holds the N composition values in bit-reversed row order (as
 * lvm_eval_composition writes them); H0(x^2) = (H(x) + H(-x))/2 and
 * H1(x^2) = (H(x) - H(-x))/(2x) are extended from the offset^2-coset of size
 * N/2 to the offset-coset of size N and committed with row-pair leaves over
 * [H0 | H1] (6 words per row). parts_out: the parts LDE (N rows x 6 words)
 * whose trace rows are `trace_rows` (the table's n: the parts' out-of-domain
 * values read every (N/n)-th row, prover.rs:2012-2100). */
int lvm_composition_commit_d2(const uint64_t *dev_h, size_t N, size_t trace_rows, uint64_t coset_offset,
                              uint8_t root_out[32], lvm_lde **parts_out);
/* The composition polynomial's parts for any part count P = degree bound / n
 * (crypto/stark/src/prover.rs:1683-1900), committed with row-pair leaves over
 * the P Fp3 columns of each row: P = 1 the values themselves
 * (lvm_commit_rows_dev, constraint_interp.rs:514 `comp_h_to_slabs`); P = 2
 * lvm_composition_commit_d2; P = 3..8 the interpolant's coefficients split by
 * stride P (`break_in_parts`, crypto/math/src/polynomial.rs:137-150) and each
 * part evaluated on the offset-coset of size N
 * (crypto/math-cuda/src/lde.rs:1651 `evaluate_poly_coset_batch_ext3_into_keep`
 * via crypto/stark/src/gpu_lde.rs:2262). dev_h: N composition values in
 * bit-reversed row order; parts_out: N rows x 3P words, trace rows n. */
int lvm_composition_commit_parts(const uint64_t *dev_h, size_t N, size_t trace_rows, size_t num_parts,
                                 uint64_t coset_offset, uint8_t root_out[32], lvm_lde **parts_out);

/* ---------------- grinding ----------------
 * crypto/math-cuda/src/grinding.rs:29 `generate_nonce_gpu`: the smallest
 * nonce < max_nonce with BE u64 of Keccak(Keccak(PREFIX || seed || g) ||
 * BE(nonce))[0..8] < 2^(64-g) (crypto/stark/src/grinding.rs:21-93); status
 * LVM_ERR_STATE when there is none below max_nonce. */
int lvm_generate_nonce(const uint8_t seed[32], uint8_t grinding_factor, uint64_t max_nonce, uint64_t *nonce_out);
/* No counterpart in the reference: lvm_generate_nonce for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, result and status
 * codes; it calls no hipMalloc/hipFree and creates no stream. Its scratch (the
 * block minima of one batch of 2^24 candidates, 512 KiB + 8 B) is a fixed
 * device array each GPU holds until the process ends; searches of any length
 * run batch by batch through it, so no input size exceeds it. Needs lvm_init
 * on the current device first (LVM_ERR_STATE otherwise). Calls that overlap
 * on one device share that array and race. */
int lvm_generate_nonce_noalloc(const uint8_t seed[32], uint8_t grinding_factor, uint64_t max_nonce,
                               uint64_t *nonce_out);

/* ---------------- Fiat-Shamir transcript (host) ----------------
 * crypto/crypto/src/fiat_shamir/default_transcript.rs:31-146
 * (`DefaultTranscript`): Keccak-256 duplex, sample() = finalize + reverse +
 * re-absorb, 8-byte big-endian candidates, rejection sampling. Host-side so
 * that every challenge can follow the root the device just produced. */
typedef struct lvm_transcript lvm_transcript;
int lvm_transcript_new(const uint8_t *data, size_t len, lvm_transcript **out);
int lvm_transcript_clone(const lvm_transcript *t, lvm_transcript **out);
int lvm_transcript_release(lvm_transcript *t);
int lvm_transcript_append_bytes(lvm_transcript *t, const uint8_t *data, size_t len);
int lvm_transcript_append_base(lvm_transcript *t, uint64_t v);
int lvm_transcript_append_fp3(lvm_transcript *t, const uint64_t v[3]);
/* count Fp3 elements appended in order (the same bytes as count single appends). */
int lvm_transcript_append_fp3_many(lvm_transcript *t, const uint64_t *v, size_t count);
/* Host helper: out[i] = base^i for i < count (canonical Fp3), the coefficient
 * powers of crypto/stark/src/prover.rs:2159-2182 (gamma) and :4453-4474 (beta). */
int lvm_fp3_powers(const uint64_t base[3], size_t count, uint64_t *out);
int lvm_transcript_sample(lvm_transcript *t, uint8_t out[32]);
int lvm_transcript_state(const lvm_transcript *t, uint8_t out[32]);
int lvm_transcript_sample_base(lvm_transcript *t, uint64_t *out);
int lvm_transcript_sample_fp3(lvm_transcript *t, uint64_t out[3]);
int lvm_transcript_sample_u64(lvm_transcript *t, uint64_t upper_bound, uint64_t *out);

/* ---------------- FRI ----------------
 * crypto/math-cuda/src/fri.rs:64/:87 `FriCommitState::new`/`new_dev`: the
 * Fp3 codeword (len elements, bit-reversed order, 3 u64 each) on the coset
 * offset*<w_len>; evals_on_device: 0 host source, 1 device source (copied),
 * 2 device buffer from lvm_dev_alloc that the FRI state takes over (no copy;
 * the caller must not free it). */
typedef struct lvm_fri lvm_fri;
int lvm_fri_new(const uint64_t *evals, size_t len, uint64_t coset_offset, int evals_on_device, lvm_fri **out);
/* fri.rs:110 `fold_and_commit_layer`: fold with zeta
 * (crypto/stark/src/fri/fri_functions.rs:15-29), hash the pair leaves,
 * build the layer tree, return its root; the layer stays on the device. */
int lvm_fri_fold_and_commit_layer(lvm_fri *f, const uint64_t zeta[3], uint8_t root_out[32]);
/* the final fold (no commitment) */
int lvm_fri_fold(lvm_fri *f, const uint64_t zeta[3]);
/* current codeword (host copy when host_out non-null), its length and coset offset */
int lvm_fri_current(const lvm_fri *f, uint64_t *host_out, size_t *len, uint64_t *offset);
int lvm_fri_num_layers(const lvm_fri *f, size_t *count);
/* committed layer `layer`: evaluations (len x 3 u64) and its (2L-1) x 32 node buffer */
int lvm_fri_layer(const lvm_fri *f, size_t layer, uint64_t *evals_out, uint8_t *nodes_out, size_t *len);
/* crypto/stark/src/fri/mod.rs:153-205 `query_phase` on the resident layers:
 * per query and layer the symmetric evaluation (count x layers x 3 u64,
 * layer-major) and the authentication path (layer-major, depth_l x 32 bytes
 * per query). */
int lvm_fri_query(const lvm_fri *f, const uint64_t *iotas, size_t count, uint64_t *evals_sym_out, uint8_t *paths_out);
int lvm_fri_release(lvm_fri *f);
/* crypto/stark/src/fri/mod.rs:27-151 `commit_phase_from_evaluations` on the
 * device codeword held by `f`, with the transcript on the host
 * (crypto/stark/src/gpu_lde.rs:3164-3289 `fri_commit_gpu_drive`): layout per
 * crypto/stark/src/fri/terminal.rs:45-55, roots (num_committed x 32 bytes)
 * and the 2^effective_k terminal coefficients (Fp3) are written out and
 * absorbed. On a device error the transcript is restored to its state at
 * entry (gpu_lde.rs:3196-3229) and a non-zero status is returned. */
int lvm_fri_commit_phase(lvm_fri *f, lvm_transcript *t, unsigned blowup_log, unsigned k, uint8_t *roots_out,
                         size_t *num_committed_out, uint64_t *coeffs_out, size_t *num_coeffs_out);
/* lvm_fri_commit_phase for `count` tables at once, table t on fs[t] with its
 * own transcript ts[t] and blowup 2^blowup_log[t] (all with terminal parameter
 * k): per layer depth, one launch folds (and hashes the leaves of) every table
 * still committing at that depth, one launch per tree step builds all their
 * layer trees and one copy brings the roots back, where each table's
 * transcript absorbs its root and samples its next zeta. Every table's roots,
 * coefficients and transcript end as lvm_fri_commit_phase leaves them
 * (roots_out[t]: room for its committed layers x 32 bytes; coeffs_out[t]: for
 * 2^k Fp3). The multi-table drive of crypto/stark/src/gpu_lde.rs:3164-3289
 * batched over the tables of crypto/stark/src/prover.rs:3974-3977. On an
 * error every transcript is restored to its state at entry. */
int lvm_fri_commit_phase_batch(lvm_fri *const *fs, lvm_transcript *const *ts, size_t count,
                               const unsigned *blowup_log, unsigned k, uint8_t *const *roots_out,
                               size_t *num_committed_out, uint64_t *const *coeffs_out, size_t *num_coeffs_out);

/* ---------------- rounds of a committed table ---------------- */
/* Caller-provided outputs of lvm_prove_committed_table, for a table of m
 * base columns (the LDE's words per row), N = n * blowup LDE rows, Q queries,
 * K = 2 frame points; layer l of FRI (l < num_fri_roots) has N / 2^(l+1)
 * values and a tree of depth log2(N) - l - 2. */
typedef struct lvm_table_rounds {
    uint64_t z[3];            /* out-of-domain point */
    uint64_t *ood;            /* K x m x 3: the columns at z and z * w_n */
    uint8_t *fri_roots;       /* room for log2(N) x 32 bytes */
    size_t num_fri_roots;     /* set by the call */
    uint64_t *fri_coeffs;     /* room for 2^(log2(blowup) + fri_k) x 3 */
    size_t num_fri_coeffs;    /* set by the call */
    uint64_t *iotas;          /* Q query indexes */
    uint64_t *fri_sym;        /* layers x Q x 3 */
    uint8_t *fri_paths;       /* per layer, Q paths of its depth (layer-major) */
    uint8_t *trace_paths;     /* Q x (log2(N) - 1) x 32 */
    uint64_t *trace_rows;     /* Q x 2 x m */
} lvm_table_rounds;
/* Rounds 3-4 of a table whose LDE `h` is committed and whose root the
 * transcript has absorbed: the out-of-domain point and values
 * (crypto/stark/src/prover.rs:4547-4560), DEEP weights, codeword and FRI
 * commit phase (prover.rs:2159-2333, fri/mod.rs:27-151), query indexes, FRI
 * query phase and trace openings, in one call. The same sequence the host
 * can drive through lvm_barycentric_base_multi, lvm_deep_composition,
 * lvm_fri_commit_phase, lvm_fri_query and lvm_lde_open, without host round
 * trips between them. No counterpart in the reference, which drives these
 * steps from its prover loop (crypto/stark/src/prover.rs:2159-2333). */
int lvm_prove_committed_table(const lvm_lde *h, lvm_transcript *t, size_t n, unsigned blowup, uint64_t offset,
                              unsigned fri_k, size_t num_queries, lvm_table_rounds *rounds);

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
