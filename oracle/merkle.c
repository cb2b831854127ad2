/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 *
 * Keccak-256 Merkle commitment, restating
 *   crypto/stark/src/prover.rs:895-961  commit_rows_bit_reversed_subset: leaf i =
 *        Keccak( BE(canon(row[br(2i)][c0..c1])) || BE(canon(row[br(2i+1)][c0..c1])) )
 *        (ROWS_PER_LEAF = 2, crypto/stark/src/commitment.rs:42; Fp3 elements stream
 *        components 0,1,2, extensions_goldilocks.rs:497-503)
 *   crypto/crypto/src/merkle_tree/merkle.rs:201-224  build_from_hashed_leaves:
 *        pad to a power of two by repeating the last leaf, nodes[0..L-1) inner
 *        (root at 0), leaves at [L-1, 2L-1)
 *   crypto/crypto/src/merkle_tree/utils.rs:7-21, :49-54, :68-96  sibling/parent,
 *        complete_until_power_of_two, level-by-level build
 *   crypto/crypto/src/merkle_tree/backends/field_element_vector.rs:77-99
 *        parent = Keccak(left32 || right32)
 *   crypto/crypto/src/merkle_tree/merkle.rs:273-283 + proof path walk (auth path:
 *        sibling at each level from the leaf up)
 */
#include <stdlib.h>
#include <string.h>
#include "gl.h"
#include "lvm_oracle.h"
#include "par.h"

static inline void put_be64(uint8_t *o, uint64_t v) {
    for (int i = 0; i < 8; i++) o[i] = (uint8_t)(v >> (56 - 8 * i));
}

size_t ora_merkle_padded_leaves(size_t L) {
    size_t p = 1;
    while (p < L) p <<= 1;
    return p;
}

typedef struct {
    uint8_t *nodes;
    size_t begin;
} lvlctx;

static void level_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    lvlctx *c = (lvlctx *)vctx;
    for (size_t i = b; i < e; i++) {
        size_t parent = c->begin + i;
        size_t left = 2 * parent + 1;
        ora_keccak256(c->nodes + left * 32, 64, c->nodes + parent * 32);
    }
}

/* nodes must hold (2L-1)*32 bytes with leaves already at [L-1, 2L-1), L a power of two */
static void build_inner(uint8_t *nodes, size_t L) {
    /* utils.rs:68-96: level begin = L-1, parents of level [b, 2b] are [b/2, b) */
    size_t level_begin = L - 1;
    while (level_begin > 0) {
        size_t new_begin = level_begin / 2;
        size_t count = level_begin - new_begin;
        lvlctx c = {nodes, new_begin};
        par_for(count, 256, level_body, &c);
        level_begin = new_begin;
    }
}

int ora_merkle_build(const uint8_t *leaves, size_t num_leaves, uint8_t *nodes_out) {
    if (num_leaves == 0) return -1;
    size_t L = ora_merkle_padded_leaves(num_leaves);
    memcpy(nodes_out + (L - 1) * 32, leaves, num_leaves * 32);
    for (size_t i = num_leaves; i < L; i++)
        memcpy(nodes_out + (L - 1 + i) * 32, leaves + (num_leaves - 1) * 32, 32);
    build_inner(nodes_out, L);
    return 0;
}

typedef struct {
    const uint64_t *data;
    size_t rows, cols, ew, c0, c1;
    unsigned log_rows;
    uint8_t *leaves;
} leafctx;

static void leaf_body(void *vctx, size_t b, size_t e, int th) {
    (void)th;
    leafctx *c = (leafctx *)vctx;
    size_t sub = (c->c1 - c->c0) * c->ew;
    size_t bytes = 2 * sub * 8;
    uint8_t *buf = (uint8_t *)malloc(bytes);
    for (size_t i = b; i < e; i++) {
        size_t off = 0;
        for (size_t k = 0; k < 2; k++) {
            size_t r = (size_t)bitrev(2 * i + k, c->log_rows);
            const uint64_t *row = c->data + (r * c->cols + c->c0) * c->ew;
            for (size_t j = 0; j < sub; j++, off += 8) put_be64(buf + off, gl_canon(row[j]));
        }
        ora_keccak256(buf, bytes, c->leaves + i * 32);
    }
    free(buf);
}

int ora_commit_rows_bit_reversed(const uint64_t *data, size_t rows, size_t cols, size_t ew,
                                 size_t c0, size_t c1, uint8_t *nodes_out) {
    if (cols == 0 || rows == 0 || c1 <= c0) return -1;
    if (rows & (rows - 1)) return -2;
    size_t num_leaves = rows / 2;
    if (num_leaves == 0) return -3;
    size_t L = ora_merkle_padded_leaves(num_leaves);
    leafctx c = {data, rows, cols, ew, c0, c1, (unsigned)__builtin_ctzll(rows), nodes_out + (L - 1) * 32};
    par_for(num_leaves, 256, leaf_body, &c);
    build_inner(nodes_out, L);
    return 0;
}

size_t ora_merkle_path(const uint8_t *nodes, size_t leaves_len, size_t pos, uint8_t *path_out) {
    size_t idx = pos + leaves_len - 1, k = 0;
    while (idx != 0) {
        size_t sib = (idx % 2 == 0) ? idx - 1 : idx + 1;
        memcpy(path_out + 32 * k++, nodes + sib * 32, 32);
        idx = (idx % 2 == 0) ? (idx - 1) / 2 : idx / 2;
    }
    return k;
}

/* crypto/crypto/src/merkle_tree/proof.rs verify: hash up with the sibling on
 * the side the index parity dictates */
int ora_merkle_verify(const uint8_t root[32], const uint8_t leaf[32], size_t pos, const uint8_t *path,
                      size_t path_len) {
    uint8_t cur[32], buf[64];
    memcpy(cur, leaf, 32);
    size_t idx = pos;
    for (size_t k = 0; k < path_len; k++) {
        if (idx % 2 == 0) {
            memcpy(buf, cur, 32);
            memcpy(buf + 32, path + 32 * k, 32);
        } else {
            memcpy(buf, path + 32 * k, 32);
            memcpy(buf + 32, cur, 32);
        }
        ora_keccak256(buf, 64, cur);
        idx >>= 1;
    }
    return memcmp(cur, root, 32) == 0;
}

int ora_lde_commit(const uint64_t *trace, size_t n, size_t m, size_t blowup, uint64_t offset,
                   uint64_t *out_lde, uint8_t *nodes_out) {
    size_t N = n * blowup;
    uint64_t *w = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *lde = out_lde ? out_lde : (uint64_t *)malloc(N * m * sizeof(uint64_t));
    ora_lde_weights(n, offset, w);
    int rc = ora_coset_lde_row_major(trace, n, m, blowup, w, lde);
    if (rc == 0) rc = ora_commit_rows_bit_reversed(lde, N, m, 1, 0, m, nodes_out);
    free(w);
    if (!out_lde) free(lde);
    return rc;
}
