// Small kernels: field-arithmetic test kernels (crypto/math-cuda/kernels/
// arith.cu semantics), bit-reversed row gathers, Merkle path gathers.
#include "lvm_internal.hpp"

namespace lvm {

__global__ void gl_op_kernel(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t x = a[i], y = b ? b[i] : 0, r;
    switch (op) {
    case 0: r = add(x, y); break;
    case 1: r = sub(x, y); break;
    case 2: r = mul(x, y); break;
    case 3: r = neg(x); break;
    case 4: r = inv(x); break;
    default: r = sqr(x); break;
    }
    out[i] = r;
}

__global__ void fp3_op_kernel(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fp3 x{a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    fp3 y{0, 0, 0};
    if (b) y = fp3{b[3 * i], b[3 * i + 1], b[3 * i + 2]};
    fp3 r;
    switch (op) {
    case 0: r = f3_add(x, y); break;
    case 1: r = f3_sub(x, y); break;
    case 2: r = f3_mul(x, y); break;
    case 3: r = f3_neg(x); break;
    case 4: r = f3_inv(x); break;
    default: r = f3_sqr(x); break;
    }
    out[3 * i] = r.c0;
    out[3 * i + 1] = r.c1;
    out[3 * i + 2] = r.c2;
}

void launch_gl_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n, bool ext3,
                  hipStream_t st) {
    dim3 grid((unsigned)((n + 255) / 256));
    if (ext3)
        hipLaunchKernelGGL(fp3_op_kernel, grid, dim3(256), 0, st, op, a, b, out, n);
    else
        hipLaunchKernelGGL(gl_op_kernel, grid, dim3(256), 0, st, op, a, b, out, n);
}

// dst natural row r <- src physical row bitrev(r) (rows of `w` words)
__global__ void gather_bitrev_rows_kernel(const uint64_t *src, uint64_t *dst, uint64_t rows, uint32_t log_rows,
                                          uint64_t w) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * w) return;
    uint64_t r = idx / w, c = idx - r * w;
    uint64_t pr = bitrev((uint32_t)r, log_rows);
    dst[idx] = src[pr * w + c];
}

void launch_gather_bitrev_rows(const uint64_t *src, uint64_t *dst, uint64_t rows, uint32_t log_rows, uint64_t w,
                               hipStream_t st) {
    uint64_t total = rows * w;
    if (!total) return;
    hipLaunchKernelGGL(gather_bitrev_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, src,
                       dst, rows, log_rows, w);
}

// Merkle authentication paths (sibling of each node from the leaf up,
// crypto/crypto/src/merkle_tree/merkle.rs:291-308) and the two opened rows
// per position (physical rows 2*pos and 2*pos+1 of the bit-reversed table).
__global__ void open_kernel(const uint64_t *nodes, uint64_t L, const uint64_t *lde, uint64_t w,
                            const uint64_t *pos, uint64_t count, uint32_t depth, uint64_t *paths,
                            uint64_t *rows) {
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    uint64_t p = pos[q];
    uint64_t idx = p + L - 1;
    for (uint32_t k = 0; k < depth; k++) {
        uint64_t sib = (idx % 2 == 0) ? idx - 1 : idx + 1;
        for (int j = 0; j < 4; j++) paths[(q * depth + k) * 4 + j] = nodes[sib * 4 + j];
        idx = (idx % 2 == 0) ? (idx - 1) / 2 : idx / 2;
    }
    if (rows) {
        for (uint64_t j = 0; j < 2 * w; j++) rows[q * 2 * w + j] = lde[2 * p * w + j];
    }
}

void launch_open(const uint64_t *nodes, uint64_t L, const uint64_t *lde, uint64_t w, const uint64_t *pos,
                 uint64_t count, uint32_t depth, uint64_t *paths, uint64_t *rows, hipStream_t st) {
    if (!count) return;
    hipLaunchKernelGGL(open_kernel, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, st, nodes, L, lde, w, pos,
                       count, depth, paths, rows);
}

// Plain 16-byte-vector device copy (grid-stride), the calibration kernel for
// HBM traffic counters.
// One 16-byte element per lane over a full grid (blocks walk the buffer in
// order): 6.2 TB/s on 1 GiB, where a grid-stride loop over 16384 blocks
// measured 4.7 TB/s.
__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void launch_copy16(void *dst, const void *src, uint64_t n16, hipStream_t st) {
    // launches of at most 2^31 blocks (2^43 bytes) each
    const uint64_t chunk = (uint64_t)256 << 31;
    for (uint64_t off = 0; off < n16; off += chunk) {
        const uint64_t n = n16 - off < chunk ? n16 - off : chunk;
        hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           (const uint4 *)src + off, (uint4 *)dst + off, n);
    }
}

// w_i = scale * offset^i (the coset weights n^-1 * g^i of
// crypto/stark/src/prover.rs:451-467 with scale = n^-1)
__global__ void coset_weights_kernel(uint64_t n, uint64_t offset, uint64_t scale, uint64_t *out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = mul(scale, pow(offset, i));
}

void launch_coset_weights(uint64_t n, uint64_t offset, uint64_t scale, uint64_t *out, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(coset_weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, offset, scale,
                       out);
}

// Synthetic table: element (r, c) = SplitMix64 draw r of the stream seeded
// with 0x1A3BDA7A ^ (cfg << 48) ^ (table << 32) ^ c, reduced mod p (draw k is
// mix(seed + (k+1) * 0x9E3779B97F4A7C15)). Row-major n x m.
__global__ void fill_splitmix_kernel(uint64_t *out, uint64_t n, uint64_t m, uint64_t cfg, uint64_t table) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * m) return;
    const uint64_t r = i / m, c = i % m;
    const uint64_t seed = 0x1A3BDA7Aull ^ (cfg << 48) ^ (table << 32) ^ c;
    uint64_t z = seed + (r + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[i] = canon(z);
}

void launch_fill_splitmix(uint64_t *out, uint64_t n, uint64_t m, uint64_t cfg, uint64_t table, hipStream_t st) {
    const uint64_t total = n * m;
    if (!total) return;
    hipLaunchKernelGGL(fill_splitmix_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, out, n, m, cfg,
                       table);
}

}  // namespace lvm
