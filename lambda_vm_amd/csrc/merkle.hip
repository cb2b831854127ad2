// Keccak-256 Merkle commitment kernels for gfx950.
//
// Leaf layout (crypto/stark/src/prover.rs:895-961, commitment.rs:42, :81-91):
// leaf i = Keccak( BE(canon(row[br(2i)])) || BE(canon(row[br(2i+1)])) ) over
// the committed columns in order; an Fp3 element streams components 0,1,2
// (extensions_goldilocks.rs:497-503). The LDE tables here are stored with
// their rows already in bit-reversed order, so leaf i reads physical rows
// 2i and 2i+1. A big-endian 8-byte element absorbed into a little-endian
// Keccak lane is bswap64(canonical value); 136-byte blocks are 17 lanes.
//
// Tree layout (crypto/crypto/src/merkle_tree/merkle.rs:201-224, utils.rs:
// 68-96): nodes[0..L-1) inner with the root at 0, leaves at [L-1, 2L-1),
// parent(i) = Keccak(node[2i+1] || node[2i+2]).
#include "keccak.hpp"
#include "lvm_internal.hpp"

namespace lvm {

// XOR the big-endian encoding of a canonicalised element into a lane held as
// 32-bit halves: the little-endian lane of BE(v) is bswap64(v).
__device__ __forceinline__ void absorb_be(uint32_t &l, uint32_t &h, uint64_t v) {
    v = canon(v);
    l ^= __builtin_bswap32((uint32_t)(v >> 32));
    h ^= __builtin_bswap32((uint32_t)v);
}

// One thread per leaf. Words [w0, w1) of rows 2i and 2i+1 (row stride
// `pitch` words) are absorbed in order, 17 lanes per permutation. CONTIG:
// the committed words are whole rows, so the leaf's data is one contiguous
// run of 2*pitch words.
template <bool CONTIG>
__device__ __forceinline__ void row_pair_leaf(const uint64_t *__restrict__ rows, uint64_t pitch, uint64_t w0,
                                              uint64_t w1, uint64_t i, uint64_t *__restrict__ leaves) {
    const uint64_t width = w1 - w0;
    const uint64_t total = 2 * width;
    const uint64_t *r0 = rows + (2 * i) * pitch + w0;
    const uint64_t *r1 = rows + (2 * i + 1) * pitch + w0;
    uint32_t l[25], h[25];
#pragma unroll
    for (int k = 0; k < 25; k++) l[k] = h[k] = 0;
    const uint64_t *cur = r0;
    uint64_t left_in_row = width;
    bool second = false;
    uint64_t e = 0;
    while (e + 17 <= total) {
#pragma unroll
        for (int j = 0; j < 17; j++) {
            absorb_be(l[j], h[j], *cur++);
            if (!CONTIG) {
                if (--left_in_row == 0 && !second) {
                    second = true;
                    cur = r1;
                    left_in_row = width;
                }
            }
        }
        keccak_f1600_32(l, h);
        e += 17;
    }
    const uint64_t rem = total - e;
#pragma unroll
    for (int j = 0; j < 17; j++) {
        if ((uint64_t)j < rem) {
            absorb_be(l[j], h[j], *cur++);
            if (!CONTIG) {
                if (--left_in_row == 0 && !second) {
                    second = true;
                    cur = r1;
                    left_in_row = width;
                }
            }
        } else if ((uint64_t)j == rem) {
            l[j] ^= 0x01;
        }
    }
    h[16] ^= 0x80000000u;
    keccak_f1600_32(l, h);
    uint64_t *o = leaves + 4 * i;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = ((uint64_t)h[k] << 32) | l[k];
}

template <bool CONTIG>
__global__ __launch_bounds__(256) void keccak_row_pair_leaves_kernel(const uint64_t *__restrict__ rows,
                                                                    uint64_t pitch, uint64_t w0, uint64_t w1,
                                                                    uint64_t num_leaves,
                                                                    uint64_t *__restrict__ leaves) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_leaves) return;
    row_pair_leaf<CONTIG>(rows, pitch, w0, w1, i, leaves);
}

// Leaves of several tables in one launch: blocks [prefix[t], prefix[t+1])
// hash table t's leaves, 256 per block.
__global__ __launch_bounds__(256) void keccak_leaves_batched_kernel(const LeafJob *__restrict__ jobs,
                                                                   const uint32_t *__restrict__ prefix,
                                                                   uint32_t count) {
    const uint32_t b = blockIdx.x;
    const uint32_t t = find_slot(prefix, count, b);
    const LeafJob j = jobs[t];
    const uint64_t i = (uint64_t)(b - prefix[t]) * 256 + threadIdx.x;
    if (i >= j.num_leaves) return;
    row_pair_leaf<true>(j.rows, j.pitch, 0, j.pitch, i, j.leaves);
}

// FRI layer leaves: Keccak(BE24(e[2j]) || BE24(e[2j+1])), 6 lanes, one permutation
// (crypto/crypto/src/merkle_tree/backends/field_element_vector.rs:129-134).
__global__ __launch_bounds__(256) void keccak_fp3_pair_leaves_kernel(const uint64_t *__restrict__ evals,
                                                                    uint64_t num_leaves,
                                                                    uint64_t *__restrict__ leaves) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= num_leaves) return;
    uint32_t l[25], h[25];
#pragma unroll
    for (int k = 0; k < 25; k++) l[k] = h[k] = 0;
    const uint64_t *p = evals + 6 * j;
#pragma unroll
    for (int k = 0; k < 6; k++) absorb_be(l[k], h[k], p[k]);
    l[6] = 0x01;
    h[16] = 0x80000000u;
    keccak_f1600_32(l, h);
    uint64_t *o = leaves + 4 * j;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = ((uint64_t)h[k] << 32) | l[k];
}

// One level of parents [begin, begin+count).
__global__ __launch_bounds__(256) void merkle_level_kernel(uint64_t *nodes, uint64_t begin, uint64_t count) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint64_t p = begin + k;
    const uint64_t *l = nodes + 4 * (2 * p + 1);
    const uint64_t *r = nodes + 4 * (2 * p + 2);
    uint64_t lv[4] = {l[0], l[1], l[2], l[3]}, rv[4] = {r[0], r[1], r[2], r[3]}, o[4];
    keccak_pair_node(lv, rv, o);
    uint64_t *d = nodes + 4 * p;
    d[0] = o[0];
    d[1] = o[1];
    d[2] = o[2];
    d[3] = o[3];
}

// Up to 9 tree levels per launch: block b takes the 2*span children
// [cstart + 2*span*b, ...) of one level (span <= 256 parents), writes their
// parents, keeps them in LDS and keeps hashing pairs, at most `levels`
// levels in all, writing every level it produces to the node array.
__device__ __forceinline__ void merkle_subtree_body(uint64_t *nodes, uint64_t cstart, uint64_t pstart, uint32_t span,
                                                    uint32_t levels, uint64_t b) {
    __shared__ uint64_t dig[256][4];
    const uint32_t t = threadIdx.x;
    uint64_t pbase = pstart + (uint64_t)span * b;  // first parent of this block at the current level
    if (t < span) {
        const uint64_t *l = nodes + 4 * (cstart + 2 * (uint64_t)span * b + 2 * t);
        uint64_t lv[4] = {l[0], l[1], l[2], l[3]}, rv[4] = {l[4], l[5], l[6], l[7]}, o[4];
        keccak_pair_node<true>(lv, rv, o);
        uint64_t *d = nodes + 4 * (pbase + t);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            d[k] = o[k];
            dig[t][k] = o[k];
        }
    }
    // next levels: the block's nodes at level start L0 = pbase; the parent of
    // node x is (x - 1) / 2, and a block's nodes stay contiguous
    uint32_t cnt = span;
    for (uint32_t lv_done = 1; cnt > 1 && lv_done < levels; lv_done++) {
        __syncthreads();
        const uint32_t half = cnt / 2;
        const uint64_t next = (pbase - 1) / 2;
        uint64_t o[4];
        if (t < half) {
            uint64_t lv[4] = {dig[2 * t][0], dig[2 * t][1], dig[2 * t][2], dig[2 * t][3]};
            uint64_t rv[4] = {dig[2 * t + 1][0], dig[2 * t + 1][1], dig[2 * t + 1][2], dig[2 * t + 1][3]};
            keccak_pair_node<true>(lv, rv, o);
        }
        __syncthreads();
        if (t < half) {
            uint64_t *d = nodes + 4 * (next + t);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                d[k] = o[k];
                dig[t][k] = o[k];
            }
        }
        pbase = next;
        cnt = half;
    }
}

__global__ __launch_bounds__(256) void merkle_subtree_kernel(uint64_t *nodes, uint64_t cstart, uint64_t pstart,
                                                             uint32_t span) {
    merkle_subtree_body(nodes, cstart, pstart, span, 32, blockIdx.x);
}

// One tree step for each of several trees: blocks [prefix[t], prefix[t+1])
// run step t (a single wide level with levels = 1, or a subtree top).
__global__ __launch_bounds__(256) void merkle_steps_batched_kernel(const TreeStep *__restrict__ steps,
                                                                   const uint32_t *__restrict__ prefix,
                                                                   uint32_t count) {
    const uint32_t b = blockIdx.x;
    const uint32_t t = find_slot(prefix, count, b);
    const TreeStep s = steps[t];
    merkle_subtree_body(s.nodes, s.cstart, s.pstart, s.span, s.levels, b - prefix[t]);
}

// Copy each tree's root (node 0) into roots[t].
__global__ void gather_roots_kernel(uint64_t *const *__restrict__ trees, uint32_t count, uint64_t *__restrict__ roots) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint64_t *r = trees[t];
#pragma unroll
    for (int k = 0; k < 4; k++) roots[4 * t + k] = r[k];
}

size_t padded_leaves(size_t n) {
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

void keccak_row_pair_leaves(const uint64_t *rows, uint64_t pitch, uint64_t w0, uint64_t w1, uint64_t num_leaves,
                            uint8_t *leaves_out, hipStream_t st) {
    if (num_leaves == 0) return;
    dim3 grid((unsigned)((num_leaves + 255) / 256));
    if (w0 == 0 && w1 == pitch)
        hipLaunchKernelGGL(keccak_row_pair_leaves_kernel<true>, grid, dim3(256), 0, st, rows, pitch, w0, w1,
                           num_leaves, (uint64_t *)leaves_out);
    else
        hipLaunchKernelGGL(keccak_row_pair_leaves_kernel<false>, grid, dim3(256), 0, st, rows, pitch, w0, w1,
                           num_leaves, (uint64_t *)leaves_out);
}

void keccak_fp3_pair_leaves(const uint64_t *evals, uint64_t len, uint8_t *leaves_out, hipStream_t st) {
    uint64_t nl = len / 2;
    if (nl == 0) return;
    dim3 grid((unsigned)((nl + 255) / 256));
    hipLaunchKernelGGL(keccak_fp3_pair_leaves_kernel, grid, dim3(256), 0, st, evals, nl, (uint64_t *)leaves_out);
}

void merkle_build_inner(uint8_t *nodes_b, uint64_t L, hipStream_t st) {
    uint64_t *nodes = (uint64_t *)nodes_b;
    // children of the current level: C nodes starting at C - 1 (C a power of two)
    uint64_t C = L;
    // wide levels: one fully parallel launch per level
    while (C > (1u << 14)) {
        const uint64_t count = C / 2;
        hipLaunchKernelGGL(merkle_level_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, nodes,
                           count - 1, count);
        C = count;
    }
    // narrow top: up to 9 levels per launch through LDS
    while (C > 1) {
        const uint32_t span = (uint32_t)(C / 2 < 256 ? C / 2 : 256);
        const uint64_t blocks = C / (2 * span);
        hipLaunchKernelGGL(merkle_subtree_kernel, dim3((unsigned)blocks), dim3(256), 0, st, nodes, C - 1, C / 2 - 1,
                           span);
        C = blocks;
    }
}

void launch_leaves_batched(const LeafJob *d_jobs, const uint32_t *d_prefix, uint32_t count, uint32_t blocks,
                           hipStream_t st) {
    if (blocks) hipLaunchKernelGGL(keccak_leaves_batched_kernel, dim3(blocks), dim3(256), 0, st, d_jobs, d_prefix, count);
}

void launch_tree_steps_batched(const TreeStep *d_steps, const uint32_t *d_prefix, uint32_t count, uint32_t blocks,
                               hipStream_t st) {
    if (blocks)
        hipLaunchKernelGGL(merkle_steps_batched_kernel, dim3(blocks), dim3(256), 0, st, d_steps, d_prefix, count);
}

void launch_gather_roots(uint64_t *const *d_trees, uint32_t count, uint64_t *d_roots, hipStream_t st) {
    if (count) hipLaunchKernelGGL(gather_roots_kernel, dim3((count + 63) / 64), dim3(64), 0, st, d_trees, count, d_roots);
}

}  // namespace lvm
