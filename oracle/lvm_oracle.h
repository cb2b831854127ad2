/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 * C restatement of the reference CPU prover's STARK hot path. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as the
 * checker or as the timed CPU baseline; the product library never links it.
 * Each function cites the reference file:line it restates.
 *
 * Conventions shared with the reference:
 *  - base field elements are raw u64 limbs (possibly non-canonical);
 *  - an Fp3 element is 3 consecutive u64 (components 0, 1, 2);
 *  - row-major buffers hold `rows * cols` elements, row r at [r*cols, (r+1)*cols).
 */
#ifndef LVM_ORACLE_H
#define LVM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- threads ---- */
void ora_set_threads(int t); /* t <= 0: all online cores */
int ora_get_threads(void);

/* ---- field (vector forms for the tests) ---- */
enum { ORA_ADD = 0, ORA_SUB = 1, ORA_MUL = 2, ORA_NEG = 3, ORA_INV = 4, ORA_SQR = 5, ORA_CANON = 6 };
void ora_gl_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);
void ora_fp3_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);
uint64_t ora_gl_root_of_unity(unsigned order);
uint64_t ora_gl_pow(uint64_t a, uint64_t e);

/* ---- Keccak ---- */
typedef struct {
    uint64_t st[25];
    size_t pos;
} ora_keccak_ctx;
void ora_keccak_f1600(uint64_t st[25]);
void ora_keccak_init(ora_keccak_ctx *c);
void ora_keccak_update(ora_keccak_ctx *c, const uint8_t *data, size_t len);
void ora_keccak_final(ora_keccak_ctx *c, uint8_t out[32]);
void ora_keccak_final_pad(ora_keccak_ctx *c, uint8_t out[32], uint8_t pad);
void ora_keccak256(const uint8_t *data, size_t len, uint8_t out[32]);
void ora_keccak_pad(const uint8_t *data, size_t len, uint8_t pad, uint8_t out[32]);

/* ---- Fiat-Shamir transcript (crypto/crypto/src/fiat_shamir/default_transcript.rs) ---- */
typedef struct ora_transcript ora_transcript;
ora_transcript *ora_transcript_new(const uint8_t *data, size_t len);
ora_transcript *ora_transcript_clone(const ora_transcript *t);
void ora_transcript_free(ora_transcript *t);
void ora_transcript_append_bytes(ora_transcript *t, const uint8_t *data, size_t len);
void ora_transcript_append_base(ora_transcript *t, uint64_t v);
void ora_transcript_append_fp3(ora_transcript *t, const uint64_t v[3]);
void ora_transcript_sample(ora_transcript *t, uint8_t out[32]);
void ora_transcript_state(const ora_transcript *t, uint8_t out[32]);
uint64_t ora_transcript_sample_base(ora_transcript *t);
void ora_transcript_sample_fp3(ora_transcript *t, uint64_t out[3]);
uint64_t ora_transcript_sample_u64(ora_transcript *t, uint64_t upper_bound);

/* ---- FFT / LDE ---- */
void ora_bitrev_permute_row_major(uint64_t *buf, size_t n, size_t m, size_t elem_words);
/* natural -> natural batched row-major FFT; inverse=1 uses omega^-1, no 1/n */
int ora_fft_two_half(uint64_t *buf, size_t n, size_t m, int inverse);
/* w_i = n^-1 * offset^i, i < n (crypto/stark/src/prover.rs:451-467) */
void ora_lde_weights(size_t n, uint64_t offset, uint64_t *out);
/* out: (n*blowup) x m row-major natural-order LDE on offset*<omega_N> */
int ora_coset_lde_row_major(const uint64_t *trace, size_t n, size_t m, size_t blowup,
                            const uint64_t *weights, uint64_t *out);

/* ---- Merkle ---- */
/* Leaves of the row-pair bit-reversed commitment over columns [c0, c1) of a
 * row-major buffer of `rows` rows x `cols` elements (elem_words = 1 base, 3 Fp3).
 * nodes_out: (2L-1)*32 bytes, L = rows/2 rounded up to a power of two. */
int ora_commit_rows_bit_reversed(const uint64_t *data, size_t rows, size_t cols, size_t elem_words,
                                 size_t c0, size_t c1, uint8_t *nodes_out);
size_t ora_merkle_padded_leaves(size_t num_leaves);
int ora_merkle_build(const uint8_t *leaves, size_t num_leaves, uint8_t *nodes_out);
/* authentication path of leaf `pos`; returns the path length (log2 L) */
size_t ora_merkle_path(const uint8_t *nodes, size_t leaves_len, size_t pos, uint8_t *path_out);
int ora_merkle_verify(const uint8_t root[32], const uint8_t leaf[32], size_t pos, const uint8_t *path,
                      size_t path_len);

/* ---- FRI ---- */
void ora_fri_inv_twiddles(uint64_t offset, size_t domain_size, uint64_t *out);
void ora_fri_fold(const uint64_t *evals, size_t len, const uint64_t zeta[3], const uint64_t *inv_tw,
                  uint64_t *out);
void ora_fri_update_twiddles(const uint64_t *tw, size_t len, uint64_t *out);
void ora_fri_layout(unsigned lde_log, unsigned blowup_log, unsigned k, unsigned *total_folds,
                    size_t *num_committed, size_t *terminal_len, unsigned *effective_k);
void ora_fri_layer_leaves(const uint64_t *evals, size_t len, uint8_t *leaves_out);
/* Full commit phase (crypto/stark/src/fri/mod.rs:27-151). evals: len Fp3 in
 * bit-reversed order (consumed/overwritten). roots_out: num_committed*32.
 * coeffs_out: 2^effective_k Fp3. layer_evals_out (nullable): concatenated
 * folded layers (len/2 + len/4 + ... for each committed layer). */
int ora_fri_commit(uint64_t *evals, size_t len, ora_transcript *t, uint64_t offset,
                   unsigned blowup_log, unsigned k, uint8_t *roots_out, uint64_t *coeffs_out,
                   uint64_t *layer_evals_out);
/* coset iFFT of the 2^k sub-coset (crypto/stark/src/fri/terminal.rs:72-103) */
void ora_fri_terminal_coeffs(const uint64_t *codeword_bitrev, size_t len, uint64_t terminal_offset,
                             unsigned k, uint64_t *coeffs_out);

/* ---- barycentric OOD (crypto/stark/src/trace.rs:831-1110) ---- */
/* lde: row-major natural (n*blowup) x m base LDE; z: Fp3; out: m Fp3 */
void ora_barycentric_base(const uint64_t *lde, size_t n, size_t m, size_t blowup, uint64_t offset,
                          const uint64_t z[3], uint64_t *out);
/* frame points z * w^(offset*step + s) (trace.rs:1112-1132) */
void ora_frame_points(const uint64_t z[3], const size_t *frame_offsets, size_t nf, uint64_t trace_root,
                      size_t step_size, uint64_t *out);

/* ---- batch inverse (crypto/math/src/field/element.rs:64-111) ---- */
int ora_batch_inverse_fp3(uint64_t *buf, size_t n);
int ora_batch_inverse_base(uint64_t *buf, size_t n);

/* ---- grinding (crypto/stark/src/grinding.rs) ---- */
int ora_is_valid_nonce(const uint8_t seed[32], uint64_t nonce, uint8_t grinding_factor);
uint64_t ora_generate_nonce(const uint8_t seed[32], uint8_t grinding_factor, uint64_t limit_search);

/* ---- DEEP composition (crypto/stark/src/prover.rs:2400-2604) ---- */
/* main: N x m base, aux: N x ma Fp3 (nullable), parts: N x np Fp3 (nullable),
 * all row-major natural order; trace_ood / gammas_tr: (m+ma) x K Fp3;
 * parts_ood / gammas_h: np Fp3; out: N Fp3 natural order. */
int ora_deep_composition(const uint64_t *main, size_t m, const uint64_t *aux, size_t ma, const uint64_t *parts,
                         size_t np, size_t N, uint64_t offset, const uint64_t *z_points, size_t K,
                         const uint64_t *z_parts, const uint64_t *trace_ood, const uint64_t *parts_ood,
                         const uint64_t *gammas_tr, const uint64_t *gammas_h, uint64_t *out);

/* ---- LogUp (crypto/stark/src/lookup.rs, logup_gpu.rs) ---- */
/* Flat CSR descriptor of a table's bus interactions (logup_gpu.rs:100-142). */
typedef struct {
    size_t num_interactions;
    const uint64_t *bus_ids;       /* [K] */
    const uint32_t *elem_offsets;  /* [K+1] */
    const uint32_t *elem_alpha_idx;
    const uint64_t *elem_const;
    const uint32_t *term_offsets;  /* [E+1] */
    const uint64_t *term_coef;
    const uint32_t *term_col;
    size_t num_out_cols;           /* committed pairs + 1 */
    const uint32_t *out_col_offsets;
    const uint32_t *out_col_interactions;
    const uint64_t *mult_const;    /* [K] */
    const uint32_t *mult_term_offsets; /* [K+1] */
    const uint64_t *mult_term_coef;
    const uint32_t *mult_term_col;
} ora_logup_desc;
/* main value (row, col) = main[row*row_stride + col*col_stride]; alpha: Fp3 powers.
 * fp_out: [(k*n + row)*3 + limb] */
int ora_logup_fingerprints(const ora_logup_desc *d, const uint64_t *main, size_t row_stride, size_t col_stride,
                           size_t n, const uint64_t *alpha, const uint64_t z[3], uint64_t *fp_out);
/* terms_out (nullable): [(c*n + row)*3 + limb] for all num_out_cols columns;
 * aux_out: row-major n x num_out_cols Fp3 (committed columns, then the
 * accumulated column); L_out: the table contribution. */
int ora_logup_aux(const ora_logup_desc *d, const uint64_t *main, size_t row_stride, size_t col_stride, size_t n,
                  const uint64_t *alpha, const uint64_t z[3], uint64_t *terms_out, uint64_t *aux_out,
                  uint64_t L_out[3]);

/* ---- constraint programs and the composition polynomial ----
 * Flat program (crypto/stark/src/constraint_ir/device.rs): nodes are 2 u64
 * each (op | a << 32, b | res << 32); ext tables hold 3 u64 per element. */
typedef struct {
    const uint64_t *nodes;
    size_t num_nodes, num_base_slots, num_ext_slots;
    const uint64_t *base_consts, *ext_consts;
    const uint64_t *roots;
    size_t num_roots;
    const uint64_t *rap, *alpha, *table_offset;
    const uint32_t *root_nodes; /* nullable: node computing each constraint (else: pinned root slots) */
} ora_program;
typedef struct {
    size_t trace_rows;                /* n; N / n is the blowup */
    uint64_t coset_offset;
    const uint64_t *beta;             /* num_roots Fp3 */
    const uint32_t *root_group;       /* zerofier group of each constraint */
    const uint32_t *group_exemptions; /* end exemptions of each group */
    size_t num_groups;
    size_t num_boundary;
    const uint32_t *b_col, *b_is_aux;
    const uint64_t *b_step;
    const uint64_t *b_value, *b_beta; /* Fp3 each */
} ora_composition;
/* main: N x m natural row-major; aux: N x ma Fp3 natural row-major.
 * evals_out: [(c*N + i)*3 + limb]; h_out: N Fp3, natural order. */
int ora_eval_constraints(const ora_program *p, const uint64_t *main, size_t m, const uint64_t *aux, size_t ma,
                         size_t N, size_t next_step, uint64_t *evals_out);
int ora_eval_composition(const ora_program *p, const uint64_t *main, size_t m, const uint64_t *aux, size_t ma,
                         size_t N, size_t next_step, const ora_composition *c, uint64_t *h_out);

/* degree-2 split of a natural-order composition codeword (prover.rs:1605-1654):
 * halves_out (N/2) x 6 words [H0 | H1] natural order */
int ora_decompose_d2(const uint64_t *h, size_t N, uint64_t offset, uint64_t *halves_out);

/* ---- pipelines used by the CPU baseline ---- */
/* LDE + row-pair commit of a base table: out_lde (nullable) row-major natural. */
int ora_lde_commit(const uint64_t *trace, size_t n, size_t m, size_t blowup, uint64_t offset,
                   uint64_t *out_lde, uint8_t *nodes_out);

#ifdef __cplusplus
}
#endif
#endif
