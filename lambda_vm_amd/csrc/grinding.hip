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
//
// lvm_generate_nonce_noalloc runs the same kernels with its block minima in a
// fixed device array (`scratch` below) and reduces them on the device, so a
// search allocates no memory.
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "keccak.hpp"
#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

void host_keccak256(const uint8_t *p, size_t n, uint8_t out[32]);

constexpr int GRIND_BLOCK = 256;
// candidates per launch, and the blocks that take them
constexpr uint64_t GRIND_BATCH = 1ull << 24;
constexpr uint64_t GRIND_BLOCKS = GRIND_BATCH / GRIND_BLOCK;

// the inner digest as four little-endian lanes, passed by value
struct GrindDigest {
    uint64_t w[4];
};

// Scratch of the allocation-free search: words [0, GRIND_BLOCKS) the block
// minima of one batch, word GRIND_BLOCKS the batch minimum (512 KiB + 8 B per
// device, held until the process ends). Host code never names it outside
// HIP_SYMBOL(): kernels get its device address from g_scratch_dev.
static __device__ __align__(256) uint64_t scratch[GRIND_BLOCKS + 1];

__global__ __launch_bounds__(GRIND_BLOCK) void grind_kernel(GrindDigest inner, uint64_t base, uint64_t count,
                                                            uint64_t limit, uint64_t *__restrict__ block_min) {
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
            const uint64_t v = inner.w[k];
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

// out[0] = min(block_min[0..blocks)); one block of 1024 threads
__global__ __launch_bounds__(1024) void grind_min_kernel(const uint64_t *__restrict__ block_min, uint32_t blocks,
                                                         uint64_t *__restrict__ out) {
    __shared__ uint64_t s_min[1024];
    uint64_t m = ~0ull;
    for (uint32_t b = threadIdx.x; b < blocks; b += 1024) m = block_min[b] < m ? block_min[b] : m;
    s_min[threadIdx.x] = m;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) {
            const uint64_t o = s_min[threadIdx.x + off];
            if (o < s_min[threadIdx.x]) s_min[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s_min[0];
}

namespace {
constexpr int GRIND_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint64_t *g_scratch_dev[GRIND_MAX_DEVICES];  // device address of `scratch` per device, set by grind_scratch_setup

// Keccak(PREFIX || seed || g) as little-endian lanes (grinding.rs:84-93)
GrindDigest inner_digest(const uint8_t seed[32], uint8_t grinding_factor) {
    static const uint8_t PREFIX[8] = {0x01, 0x23, 0x45, 0x67, 0x89, 0xab, 0xcd, 0xed};
    uint8_t msg[41], inner_b[32];
    std::memcpy(msg, PREFIX, 8);
    std::memcpy(msg + 8, seed, 32);
    msg[40] = grinding_factor;
    host_keccak256(msg, 41, inner_b);
    GrindDigest dg;
    std::memcpy(dg.w, inner_b, 32);
    return dg;
}
}  // namespace

// Resolves the device address of `scratch` on the current device (called by
// lvm_init with `device` current; once per device).
int grind_scratch_setup(int device) {
    if (device < 0 || device >= GRIND_MAX_DEVICES) return fail(LVM_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return LVM_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(LVM_ERR_DEVICE, std::string("address of the nonce-search scratch: ") + hipGetErrorString(e));
    g_scratch_dev[device] = (uint64_t *)p;
    return LVM_OK;
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
    const GrindDigest inner = inner_digest(seed, grinding_factor);
    const uint64_t limit = grinding_factor == 64 ? 1 : (1ull << (64 - grinding_factor));
    constexpr uint64_t BATCH = GRIND_BATCH;
    constexpr uint64_t BLOCKS = GRIND_BLOCKS;
    uint64_t *d_min = nullptr;
    if (lvm_malloc(&d_min, BLOCKS * 8) != hipSuccess)
        return fail(LVM_ERR_NOMEM, "device allocation for the nonce search failed");
    std::vector<uint64_t> mins(BLOCKS);
    StageScope ss(st);
    hipError_t e = hipSuccess;
    uint64_t best = ~0ull;
    for (uint64_t base = 0; e == hipSuccess && base < max_nonce; base += BATCH) {
        const uint64_t count = max_nonce - base < BATCH ? max_nonce - base : BATCH;
        const uint64_t blocks = (count + GRIND_BLOCK - 1) / GRIND_BLOCK;
        {
            ProfScope ps("grinding", st);
            hipLaunchKernelGGL(grind_kernel, dim3((unsigned)blocks), dim3(GRIND_BLOCK), 0, st, inner, base, count,
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
    lvm_free(d_min);
    if (e != hipSuccess) return fail(LVM_ERR_DEVICE, std::string("nonce search failed: ") + hipGetErrorString(e));
    if (best == ~0ull) return fail(LVM_ERR_STATE, "no valid nonce below the search limit");
    *nonce_out = best;
    return LVM_OK;
}

// The same search with no allocation: block minima in the device's `scratch`
// array, reduced on the device; one 8-byte copy per batch to the host. Needs
// lvm_init on this device first (it resolves the scratch address). Calls that
// overlap on one device share its scratch and race.
extern "C" int lvm_generate_nonce_noalloc(const uint8_t seed[32], uint8_t grinding_factor, uint64_t max_nonce,
                                          uint64_t *nonce_out) {
    using namespace lvm;
    if (!seed || !nonce_out) return fail(LVM_ERR_ARG, "null argument");
    if (grinding_factor < 1 || grinding_factor > 64) return fail(LVM_ERR_ARG, "grinding factor must be 1..64");
    int id = 0;
    if (hipGetDevice(&id) != hipSuccess) return fail(LVM_ERR_DEVICE, "hipGetDevice failed (no GPU visible?)");
    uint64_t *sc = nullptr;
    if (id >= 0 && id < GRIND_MAX_DEVICES) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        sc = g_scratch_dev[id];
    }
    if (!sc) return fail(LVM_ERR_STATE, "nonce-search scratch not set up on this device: call lvm_init first");
    Device *d = device_state();  // initialised by lvm_init: no allocation here
    if (!d) return LVM_ERR_DEVICE;
    hipStream_t st = thread_stream_if_any(d);
    const GrindDigest inner = inner_digest(seed, grinding_factor);
    const uint64_t limit = grinding_factor == 64 ? 1 : (1ull << (64 - grinding_factor));
    hipError_t e = hipSuccess;
    uint64_t best = ~0ull;
    for (uint64_t base = 0; base < max_nonce; base += GRIND_BATCH) {
        const uint64_t count = max_nonce - base < GRIND_BATCH ? max_nonce - base : GRIND_BATCH;
        const uint64_t blocks = (count + GRIND_BLOCK - 1) / GRIND_BLOCK;  // <= GRIND_BLOCKS
        hipLaunchKernelGGL(grind_kernel, dim3((unsigned)blocks), dim3(GRIND_BLOCK), 0, st, inner, base, count, limit,
                           sc);
        hipLaunchKernelGGL(grind_min_kernel, dim3(1), dim3(1024), 0, st, (const uint64_t *)sc, (uint32_t)blocks,
                           sc + GRIND_BLOCKS);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&best, sc + GRIND_BLOCKS, 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess || best != ~0ull) break;
        if (base + GRIND_BATCH < base) break;  // wrapped
    }
    if (e != hipSuccess) return fail(LVM_ERR_DEVICE, std::string("nonce search failed: ") + hipGetErrorString(e));
    if (best == ~0ull) return fail(LVM_ERR_STATE, "no valid nonce below the search limit");
    *nonce_out = best;
    return LVM_OK;
}
