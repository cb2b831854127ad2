// FRI commit phase on gfx950: every layer stays in HBM (codeword + Merkle
// tree) and only the 32-byte root crosses to the host per layer, where the
// transcript turns it into the next challenge.
//
// Fold (crypto/stark/src/fri/fri_functions.rs:15-29), codeword in bit-reversed
// order, pairs (2j, 2j+1) conjugate:
//     out[j] = (lo + hi) + inv_tw[j] * (zeta * (lo - hi))
// with inv_tw[j] = 1 / (g_l * w_{N_l}^bitrev(j)) on layer l (offset g_l =
// g^(2^l), length N_l): the reference builds these by batch inversion and
// squares them per layer (fri_functions.rs:37-61); here each thread forms its
// twiddle as g_l^-1 * w_{N_l}^-bitrev(j) from the two-level root table.
// One thread produces folded elements 2j and 2j+1 and hashes them into leaf
// j = Keccak(BE24(e[2j]) || BE24(e[2j+1])) (field_element_vector.rs:129-134)
// in the same pass, so a committed layer is one read of its parent codeword.
#include <cstring>
#include <vector>

#include "keccak.hpp"
#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

__device__ __forceinline__ fp3 ld3(const uint64_t *p) { return fp3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void st3(uint64_t *p, fp3 v) {
    p[0] = v.c0;
    p[1] = v.c1;
    p[2] = v.c2;
}

// One fold of a codeword of 2*half Fp3 values into `out` (half values);
// `leaves` (nullable) receives the half/2 layer leaves.
struct FoldJob {
    const uint64_t *in;
    uint64_t *out;
    uint64_t *leaves;
    uint64_t half;
    uint64_t z0, z1, z2;
    uint64_t ginv;  // inverse of the codeword's coset offset
    uint32_t log_half;
    uint32_t pad;
};

__device__ __forceinline__ fp3 fold_one(const uint64_t *in, uint64_t j, fp3 zeta, uint64_t ginv, uint32_t log_half,
                                        const uint64_t *ilo, const uint64_t *ihi) {
    fp3 lo = ld3(in + 6 * j), hi = ld3(in + 6 * j + 3);
    fp3 sum = f3_add(lo, hi);
    fp3 diff = f3_sub(lo, hi);
    // w_{N_l}^-bitrev(j) = w_{2^32}^-(bitrev(j) << (32 - log N_l)), log N_l = log_half + 1
    uint32_t e = log_half ? (bitrev((uint32_t)j, log_half) << (31 - log_half)) : 0u;
    uint64_t tw = mul(ginv, mul(ilo[e & 0xFFFFu], ihi[e >> 16]));
    return f3_add(sum, f3_mul_base(f3_mul(zeta, diff), tw));
}

// Thread t of one fold: out has half elements; leaves (nullable) receives
// half/2 digests.
__device__ __forceinline__ void fri_fold_body(const FoldJob &j, uint64_t t, const uint64_t *ilo,
                                              const uint64_t *ihi) {
    const uint64_t *__restrict__ in = j.in;
    uint64_t *__restrict__ out = j.out;
    uint64_t *__restrict__ leaves = j.leaves;
    const uint64_t half = j.half, ginv = j.ginv;
    const uint32_t log_half = j.log_half;
    const fp3 zeta{j.z0, j.z1, j.z2};
    if (leaves) {
        // thread t: folded elements 2t, 2t+1 and leaf t
        if (2 * t >= half) return;
        fp3 a = fold_one(in, 2 * t, zeta, ginv, log_half, ilo, ihi);
        fp3 b = fold_one(in, 2 * t + 1, zeta, ginv, log_half, ilo, ihi);
        st3(out + 6 * t, a);
        st3(out + 6 * t + 3, b);
        uint32_t l[25], h[25];
#pragma unroll
        for (int k = 0; k < 25; k++) l[k] = h[k] = 0;
        const uint64_t v[6] = {a.c0, a.c1, a.c2, b.c0, b.c1, b.c2};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            uint64_t c = canon(v[k]);
            l[k] = __builtin_bswap32((uint32_t)(c >> 32));
            h[k] = __builtin_bswap32((uint32_t)c);
        }
        l[6] = 0x01;
        h[16] = 0x80000000u;
        keccak_f1600_32<true>(l, h);  // small layers are latency-bound: rounds unrolled
        uint64_t *o = leaves + 4 * t;
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = ((uint64_t)h[k] << 32) | l[k];
    } else {
        if (t >= half) return;
        st3(out + 3 * t, fold_one(in, t, zeta, ginv, log_half, ilo, ihi));
    }
}

__global__ __launch_bounds__(256) void fri_fold_kernel(FoldJob j, const uint64_t *ilo, const uint64_t *ihi) {
    fri_fold_body(j, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, ilo, ihi);
}

// The same fold for several codewords in one launch: blocks [prefix[t],
// prefix[t+1]) run jobs[t].
__global__ __launch_bounds__(256) void fri_fold_batched_kernel(const FoldJob *__restrict__ jobs,
                                                               const uint32_t *__restrict__ prefix, uint32_t count,
                                                               const uint64_t *ilo, const uint64_t *ihi) {
    const uint32_t b = blockIdx.x;
    const uint32_t t = find_slot(prefix, count, b);
    const FoldJob j = jobs[t];
    fri_fold_body(j, (uint64_t)(b - prefix[t]) * blockDim.x + threadIdx.x, ilo, ihi);
}

// Symmetric evaluations and authentication paths of one layer for every
// query (crypto/stark/src/fri/mod.rs:153-205): index_l = iota >> l,
// eval_sym = e_l[index_l ^ 1], path of leaf index_l >> 1.
__global__ void fri_query_layer_kernel(const uint64_t *evals, const uint64_t *nodes, uint64_t L, uint32_t depth,
                                       const uint64_t *iotas, uint64_t count, uint32_t layer, uint64_t *sym_out,
                                       uint64_t *paths_out) {
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    uint64_t idx = iotas[q] >> layer;
    const uint64_t *e = evals + 3 * (idx ^ 1);
    sym_out[3 * q] = e[0];
    sym_out[3 * q + 1] = e[1];
    sym_out[3 * q + 2] = e[2];
    uint64_t node = (idx >> 1) + L - 1;
    for (uint32_t k = 0; k < depth; k++) {
        uint64_t sib = (node % 2 == 0) ? node - 1 : node + 1;
        for (int w = 0; w < 4; w++) paths_out[(q * depth + k) * 4 + w] = nodes[sib * 4 + w];
        node = (node % 2 == 0) ? (node - 1) / 2 : node / 2;
    }
}

}  // namespace lvm

using namespace lvm;

struct FriLayer {
    uint64_t *evals = nullptr;  // len Fp3
    uint64_t len = 0;
    uint8_t *nodes = nullptr;   // (2L-1) x 32
    uint64_t L = 0;
    uint8_t root[32];
};

struct lvm_fri {
    int device;
    uint64_t *cur = nullptr;  // current codeword (owned unless it is a committed layer's buffer)
    bool cur_owned = true;
    uint64_t len = 0;
    uint64_t offset = 0;  // coset offset of the current codeword
    std::vector<FriLayer> layers;
};

namespace {
bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
unsigned log2u(uint64_t x) { return (unsigned)__builtin_ctzll(x); }

void free_fri(lvm_fri *f) {
    if (f->cur && f->cur_owned) lvm_free(f->cur);
    for (auto &l : f->layers) {
        lvm_free(l.evals);
        lvm_free(l.nodes);
    }
    delete f;
}
}  // namespace

extern "C" {

int lvm_fri_new(const uint64_t *evals, size_t len, uint64_t coset_offset, int evals_on_device, lvm_fri **out) {
    if (!evals || !out || !is_pow2(len) || len < 2) return fail(LVM_ERR_ARG, "FRI codeword length must be a power of two >= 2");
    if (log2u(len) > 32) return fail(LVM_ERR_ARG, "FRI codeword exceeds 2^32");
    if (canon(coset_offset) == 0) return fail(LVM_ERR_ARG, "coset offset must be non-zero");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    lvm_fri *f = new lvm_fri();
    f->device = d->id;
    f->len = len;
    f->offset = coset_offset;
    if (evals_on_device == 2) {
        // adopt a library allocation (lvm_dev_alloc) without copying it
        f->cur = (uint64_t *)evals;
        *out = f;
        return LVM_OK;
    }
    if (lvm_malloc(&f->cur, len * 24) != hipSuccess) {
        delete f;
        return fail(LVM_ERR_NOMEM, "FRI codeword allocation failed");
    }
    hipError_t e = hipMemcpyAsync(f->cur, evals, len * 24, evals_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  thread_stream(d));
    if (e == hipSuccess) e = hipStreamSynchronize(thread_stream(d));
    if (e != hipSuccess) {
        free_fri(f);
        return fail(LVM_ERR_DEVICE, std::string("FRI upload: ") + hipGetErrorString(e));
    }
    *out = f;
    return LVM_OK;
}

static int fri_fold_impl(lvm_fri *f, const uint64_t zeta[3], bool commit, uint8_t root_out[32]) {
    if (!f || !zeta) return fail(LVM_ERR_ARG, "null argument");
    if (f->len < 2) return fail(LVM_ERR_STATE, "codeword already folded to one element");
    if (commit && f->len < 4) return fail(LVM_ERR_STATE, "a committed layer needs at least two folded elements");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = thread_stream(d);
    const uint64_t half = f->len / 2;
    uint64_t *out = nullptr;
    LVM_HIP(lvm_malloc(&out, half * 24));
    FriLayer layer;
    if (commit) {
        layer.L = padded_leaves(half / 2);
        if (lvm_malloc(&layer.nodes, (2 * layer.L - 1) * 32) != hipSuccess) {
            lvm_free(out);
            return fail(LVM_ERR_NOMEM, "FRI tree allocation failed");
        }
    }
    const uint64_t ginv = inv(f->offset);
    const uint32_t log_half = log2u(half);
    {
        ProfScope ps(commit ? "fri_fold_leaves" : "fri_fold", st);
        uint64_t threads = commit ? half / 2 : half;
        const FoldJob j{f->cur, out, commit ? (uint64_t *)(layer.nodes + (layer.L - 1) * 32) : (uint64_t *)nullptr,
                        half, zeta[0], zeta[1], zeta[2], ginv, log_half, 0};
        hipLaunchKernelGGL(fri_fold_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, j,
                           d->tb.inv_lo, d->tb.inv_hi);
    }
    if (commit) {
        ProfScope ps("fri_merkle", st);
        merkle_build_inner(layer.nodes, layer.L, st);
    }
    StageScope ss(st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && commit) e = stage_download(layer.root, layer.nodes, 32, st);
    if (e == hipSuccess) e = stage_sync(st);
    if (e != hipSuccess) {
        lvm_free(out);
        if (layer.nodes) lvm_free(layer.nodes);
        return fail(LVM_ERR_DEVICE, std::string("FRI fold: ") + hipGetErrorString(e));
    }
    if (f->cur_owned) lvm_free(f->cur);
    f->cur = out;
    f->len = half;
    f->offset = sqr(f->offset);
    if (commit) {
        layer.evals = out;
        layer.len = half;
        f->cur_owned = false;
        f->layers.push_back(layer);
        if (root_out) std::memcpy(root_out, layer.root, 32);
    } else {
        f->cur_owned = true;
    }
    return LVM_OK;
}

}  // extern "C"

namespace lvm {

// One fold of every codeword fs[0..count) (zeta t at zetas[3t]) on the
// calling thread's stream: one work-list launch for the folds (and their
// leaves), one launch per tree step over all layer trees, one copy of all
// roots. Each state then holds what fri_fold_impl would leave. On a device
// error no state changes.
int fri_fold_batch(lvm_fri *const *fs, size_t count, const uint64_t *zetas, bool commit, uint8_t *roots_out) {
    if (!fs || !zetas || (commit && !roots_out)) return fail(LVM_ERR_ARG, "null argument");
    if (count == 0) return LVM_OK;
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    for (size_t t = 0; t < count; t++) {
        if (!fs[t]) return fail(LVM_ERR_ARG, "null FRI state");
        if (fs[t]->device != d->id) return fail(LVM_ERR_ARG, "FRI states of another device");
        if (fs[t]->len < (commit ? 4u : 2u)) return fail(LVM_ERR_STATE, "codeword too short for this fold");
    }
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = thread_stream(d);
    std::vector<uint64_t *> outs(count, nullptr);
    std::vector<FriLayer> layers(count);
    uint8_t *arena = nullptr;
    auto release = [&]() {
        for (size_t t = 0; t < count; t++) {
            lvm_free(outs[t]);
            lvm_free(layers[t].nodes);
        }
        lvm_free(arena);
    };
    // descriptors: fold jobs + block prefix, then per tree step the steps + prefix, then the tree pointers
    std::vector<FoldJob> jobs(count);
    std::vector<uint32_t> prefix(1, 0);
    uint64_t total = 0;
    std::vector<std::vector<std::pair<TreeStep, uint32_t>>> tsteps(count);
    size_t max_steps = 0;
    for (size_t t = 0; t < count; t++) {
        lvm_fri *f = fs[t];
        const uint64_t half = f->len / 2;
        if (lvm_malloc(&outs[t], half * 24) != hipSuccess) {
            release();
            return fail(LVM_ERR_NOMEM, "FRI layer allocation failed");
        }
        if (commit) {
            layers[t].L = padded_leaves(half / 2);
            if (lvm_malloc(&layers[t].nodes, (2 * layers[t].L - 1) * 32) != hipSuccess) {
                release();
                return fail(LVM_ERR_NOMEM, "FRI tree allocation failed");
            }
            tsteps[t] = tree_steps(layers[t].nodes, layers[t].L);
            max_steps = std::max(max_steps, tsteps[t].size());
        }
        jobs[t] = FoldJob{f->cur, outs[t],
                          commit ? (uint64_t *)(layers[t].nodes + (layers[t].L - 1) * 32) : (uint64_t *)nullptr,
                          half, zetas[3 * t], zetas[3 * t + 1], zetas[3 * t + 2], inv(f->offset), log2u(half), 0};
        total += ((commit ? half / 2 : half) + 255) / 256;
        if (total >= (1ull << 31)) {
            release();
            return fail(LVM_ERR_ARG, "batched fold exceeds 2^31 blocks");
        }
        prefix.push_back((uint32_t)total);
    }
    std::vector<uint8_t> blob;
    auto put = [&](const void *p, size_t bytes) {
        const size_t off = (blob.size() + 15) & ~(size_t)15;
        blob.resize(off + bytes);
        std::memcpy(blob.data() + off, p, bytes);
        return off;
    };
    const size_t o_jobs = put(jobs.data(), count * sizeof(FoldJob));
    const size_t o_prefix = put(prefix.data(), prefix.size() * 4);
    struct StepLaunch {
        size_t steps_off, prefix_off;
        uint32_t count, blocks;
    };
    std::vector<StepLaunch> steps_l;
    for (size_t s = 0; s < max_steps; s++) {
        std::vector<TreeStep> steps;
        std::vector<uint32_t> sp(1, 0);
        uint64_t sb = 0;
        for (size_t t = 0; t < count; t++) {
            if (s >= tsteps[t].size()) continue;
            steps.push_back(tsteps[t][s].first);
            sb += tsteps[t][s].second;
            sp.push_back((uint32_t)sb);
        }
        steps_l.push_back({put(steps.data(), steps.size() * sizeof(TreeStep)), put(sp.data(), sp.size() * 4),
                           (uint32_t)steps.size(), (uint32_t)sb});
    }
    std::vector<uint64_t *> trees(count);
    for (size_t t = 0; t < count; t++) trees[t] = (uint64_t *)layers[t].nodes;
    const size_t o_trees = put(trees.data(), count * sizeof(uint64_t *));
    const size_t o_roots = (blob.size() + 15) & ~(size_t)15;  // the gathered roots
    blob.resize(o_roots + count * 32);
    if (lvm_malloc(&arena, blob.size()) != hipSuccess) {
        release();
        return fail(LVM_ERR_NOMEM, "FRI descriptor allocation failed");
    }
    StageScope ss(st);
    hipError_t e = stage_upload(arena, blob.data(), blob.size(), st);
    if (e == hipSuccess) {
        ProfScope ps(commit ? "fri_fold_leaves" : "fri_fold", st);
        hipLaunchKernelGGL(fri_fold_batched_kernel, dim3((unsigned)total), dim3(256), 0, st,
                           (const FoldJob *)(arena + o_jobs), (const uint32_t *)(arena + o_prefix), (uint32_t)count,
                           d->tb.inv_lo, d->tb.inv_hi);
    }
    if (e == hipSuccess && commit) {
        ProfScope ps("fri_merkle", st);
        for (const StepLaunch &l : steps_l)
            launch_tree_steps_batched((const TreeStep *)(arena + l.steps_off), (const uint32_t *)(arena + l.prefix_off),
                                      l.count, l.blocks, st);
        launch_gather_roots((uint64_t *const *)(arena + o_trees), (uint32_t)count, (uint64_t *)(arena + o_roots), st);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && commit) e = stage_download(roots_out, arena + o_roots, count * 32, st);
    if (e == hipSuccess) e = stage_sync(st);
    if (e != hipSuccess) {
        release();
        return fail(LVM_ERR_DEVICE, std::string("batched FRI fold: ") + hipGetErrorString(e));
    }
    lvm_free(arena);
    for (size_t t = 0; t < count; t++) {
        lvm_fri *f = fs[t];
        if (f->cur_owned) lvm_free(f->cur);
        f->cur = outs[t];
        f->len /= 2;
        f->offset = sqr(f->offset);
        if (commit) {
            FriLayer &l = layers[t];
            std::memcpy(l.root, roots_out + 32 * t, 32);
            l.evals = outs[t];
            l.len = f->len;
            f->cur_owned = false;
            f->layers.push_back(l);
        } else {
            f->cur_owned = true;
        }
    }
    return LVM_OK;
}

}  // namespace lvm

extern "C" {

int lvm_fri_fold_and_commit_layer(lvm_fri *f, const uint64_t zeta[3], uint8_t root_out[32]) {
    return fri_fold_impl(f, zeta, true, root_out);
}

int lvm_fri_fold(lvm_fri *f, const uint64_t zeta[3]) { return fri_fold_impl(f, zeta, false, nullptr); }

int lvm_fri_current(const lvm_fri *f, uint64_t *host_out, size_t *len, uint64_t *offset) {
    if (!f) return fail(LVM_ERR_ARG, "null handle");
    if (len) *len = f->len;
    if (offset) *offset = f->offset;
    if (!host_out) return LVM_OK;
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    StageScope ss(thread_stream(d));
    LVM_HIP(stage_download(host_out, f->cur, f->len * 24, thread_stream(d)));
    LVM_HIP(stage_sync(thread_stream(d)));
    return LVM_OK;
}

int lvm_fri_num_layers(const lvm_fri *f, size_t *count) {
    if (!f || !count) return fail(LVM_ERR_ARG, "null argument");
    *count = f->layers.size();
    return LVM_OK;
}

int lvm_fri_layer(const lvm_fri *f, size_t layer, uint64_t *evals_out, uint8_t *nodes_out, size_t *len) {
    if (!f || layer >= f->layers.size()) return fail(LVM_ERR_ARG, "bad layer index");
    const FriLayer &l = f->layers[layer];
    if (len) *len = l.len;
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (evals_out) LVM_HIP(hipMemcpyAsync(evals_out, l.evals, l.len * 24, hipMemcpyDeviceToHost, thread_stream(d)));
    if (nodes_out) LVM_HIP(hipMemcpyAsync(nodes_out, l.nodes, (2 * l.L - 1) * 32, hipMemcpyDeviceToHost, thread_stream(d)));
    LVM_HIP(hipStreamSynchronize(thread_stream(d)));
    return LVM_OK;
}

int lvm_fri_query(const lvm_fri *f, const uint64_t *iotas, size_t count, uint64_t *evals_sym_out, uint8_t *paths_out) {
    if (!f || (count && (!iotas || !evals_sym_out || !paths_out))) return fail(LVM_ERR_ARG, "null argument");
    if (count == 0 || f->layers.empty()) return LVM_OK;
    const uint64_t len0 = f->layers[0].len * 2;
    for (size_t i = 0; i < count; i++)
        if (iotas[i] >= len0 / 2) return fail(LVM_ERR_ARG, "query index out of range");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    hipStream_t st = thread_stream(d);
    // per layer: count sym evals (3 u64) and count paths (depth_l x 32 bytes)
    size_t path_words = 0;
    for (auto &l : f->layers) path_words += count * log2u(l.L) * 4;
    // released on every exit (stream-ordered: a block is reused on this stream
    // only after the work queued before its release)
    struct Owned {
        uint64_t *p = nullptr;
        ~Owned() {
            if (p) lvm_free(p);
        }
    } dio, dsym, dpaths;
    LVM_HIP(lvm_malloc(&dio.p, count * 8));
    LVM_HIP(lvm_malloc(&dsym.p, f->layers.size() * count * 24));
    LVM_HIP(lvm_malloc(&dpaths.p, (path_words ? path_words : 1) * 8));
    StageScope ss(st);
    LVM_HIP(stage_upload(dio.p, iotas, count * 8, st));
    size_t poff = 0;
    for (size_t li = 0; li < f->layers.size(); li++) {
        const FriLayer &l = f->layers[li];
        uint32_t depth = log2u(l.L);
        hipLaunchKernelGGL(fri_query_layer_kernel, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, st, l.evals,
                           (const uint64_t *)l.nodes, l.L, depth, dio.p, count, (uint32_t)li,
                           dsym.p + li * count * 3, dpaths.p + poff);
        poff += count * depth * 4;
    }
    LVM_HIP(hipGetLastError());
    LVM_HIP(stage_download(evals_sym_out, dsym.p, f->layers.size() * count * 24, st));
    if (path_words) LVM_HIP(stage_download(paths_out, dpaths.p, path_words * 8, st));
    LVM_HIP(stage_sync(st));
    return LVM_OK;
}

int lvm_fri_release(lvm_fri *f) {
    if (!f) return LVM_OK;
    Device *d = device_state();
    if (d) hipStreamSynchronize(thread_stream(d));
    free_fri(f);
    return LVM_OK;
}

}  // extern "C"
