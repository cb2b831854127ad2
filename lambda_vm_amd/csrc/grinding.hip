// Proof-of-work nonce search on gfx950 (crypto/stark/src/grinding.rs; the
// reference's GPU search is crypto/math-cuda/src/grinding.rs:29
// `generate_nonce_gpu`).
//
//   inner = Keccak(PREFIX || seed || g)             (grinding.rs:84-93, host)
//   nonce valid iff BE u64 of Keccak(inner || BE(nonce))[0..8] < 2^(64-g)
//                                                   (grinding.rs:66-80)
// The result is the smallest valid nonce, the deterministic answer of the
// reference's sequential search. One thread per candidate; the 40-byte
// message fits one Keccak-f (lanes 0-3 the inner digest, lane 4 the nonce,
// padding in lanes 5 and 16). Batches of candidates are searched in
// increasing order; each block reduces its smallest hit through LDS and
// writes it with a plain store, so the first batch with a hit gives the
// answer without atomics.
#include <cstring>
#include <string>
#include <vector>

#include "keccak.hpp"
#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

void host_keccak256(const uint8_t *p, size_t n, uint8_t out[32]);

constexpr int GRIND_BLOCK = 256;

__global__ __launch_bounds__(GRIND_BLOCK) void grind_kernel(const uint64_t *__restrict__ inner, uint64_t base,
                                                            uint64_t count, uint64_t limit,
                                                            uint64_t *__restrict__ block_min) {
    __shared__ uint64_t s_min[GRIND_BLOCK];
    const uint64_t i = (uint64_t)blockIdx.x * GRIND_BLOCK + threadIdx.x;
    const uint64_t nonce = base + i;
    uint64_t found = ~0ull;
    if (i < count) {
        uint32_t l[25], h[25];
#pragma unroll
        for (int k = 0; k < 25; k++) l[k] = h[k] = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t v = inner[k];
            l[k] = (uint32_t)v;
            h[k] = (uint32_t)(v >> 32);
        }
        // BE(nonce) as a little-endian lane is bswap64(nonce)
        const uint64_t nb = bswap64(nonce);
        l[4] = (uint32_t)nb;
        h[4] = (uint32_t)(nb >> 32);
        l[5] = 0x01;           // byte 40: first padding byte
        h[16] = 0x80000000u;   // byte 135: last padding byte
        keccak_f1600_32(l, h);
        const uint64_t lane0 = ((uint64_t)h[0] << 32) | l[0];
        if (bswap64(lane0) < limit) found = nonce;
    }
    s_min[threadIdx.x] = found;
    __syncthreads();
    for (int off = GRIND_BLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) {
            const uint64_t o = s_min[threadIdx.x + off];
            if (o < s_min[threadIdx.x]) s_min[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) block_min[blockIdx.x] = s_min[0];
}

}  // namespace lvm

extern "C" int lvm_generate_nonce(const uint8_t seed[32], uint8_t grinding_factor, uint64_t max_nonce,
                                  uint64_t *nonce_out) {
    using namespace lvm;
    if (!seed || !nonce_out) return fail(LVM_ERR_ARG, "null argument");
    if (grinding_factor < 1 || grinding_factor > 64) return fail(LVM_ERR_ARG, "grinding factor must be 1..64");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    hipStream_t st = lib_stream();
    static const uint8_t PREFIX[8] = {0x01, 0x23, 0x45, 0x67, 0x89, 0xab, 0xcd, 0xed};
    uint8_t msg[41], inner_b[32];
    std::memcpy(msg, PREFIX, 8);
    std::memcpy(msg + 8, seed, 32);
    msg[40] = grinding_factor;
    host_keccak256(msg, 41, inner_b);
    uint64_t inner[4];
    std::memcpy(inner, inner_b, 32);  // little-endian lanes
    const uint64_t limit = grinding_factor == 64 ? 1 : (1ull << (64 - grinding_factor));
    constexpr uint64_t BATCH = 1ull << 24;
    constexpr uint64_t BLOCKS = BATCH / GRIND_BLOCK;
    uint64_t *d_inner = nullptr, *d_min = nullptr;
    if (lvm_malloc(&d_inner, 32) != hipSuccess || lvm_malloc(&d_min, BLOCKS * 8) != hipSuccess) {
        lvm_free(d_inner);
        return fail(LVM_ERR_NOMEM, "device allocation for the nonce search failed");
    }
    std::vector<uint64_t> mins(BLOCKS);
    StageScope ss(st);
    hipError_t e = stage_upload(d_inner, inner, 32, st);
    uint64_t best = ~0ull;
    for (uint64_t base = 0; e == hipSuccess && base < max_nonce; base += BATCH) {
        const uint64_t count = max_nonce - base < BATCH ? max_nonce - base : BATCH;
        const uint64_t blocks = (count + GRIND_BLOCK - 1) / GRIND_BLOCK;
        {
            ProfScope ps("grinding", st);
            hipLaunchKernelGGL(grind_kernel, dim3((unsigned)blocks), dim3(GRIND_BLOCK), 0, st, d_inner, base, count,
                               limit, d_min);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = stage_download(mins.data(), d_min, blocks * 8, st);
        if (e == hipSuccess) e = stage_sync(st);
        if (e != hipSuccess) break;
        for (uint64_t b = 0; b < blocks; b++) best = mins[b] < best ? mins[b] : best;
        if (best != ~0ull) break;
        if (base + BATCH < base) break;  // wrapped
    }
    lvm_free(d_inner);
    lvm_free(d_min);
    if (e != hipSuccess) return fail(LVM_ERR_DEVICE, std::string("nonce search failed: ") + hipGetErrorString(e));
    if (best == ~0ull) return fail(LVM_ERR_STATE, "no valid nonce below the search limit");
    *nonce_out = best;
    return LVM_OK;
}
