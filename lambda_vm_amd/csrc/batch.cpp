// Batched ragged commitment: the coset LDE, Keccak leaves and Merkle trees
// of many tables of different heights and widths, one launch per stage step.
//
// The reference commits a proof's tables one by one
// (crypto/stark/src/prover.rs:3299-3492 `multi_prove` -> per-table
// `coset_lde_row_major_with_merkle_tree_keep`, crypto/math-cuda/src/lde.rs
// :674), so each table pays its own launches and synchronisations and a small
// table leaves most of the GPU idle. Here a plan fixes the shapes once and
// lays out, in one device arena, the work lists of every launch:
//  - NTT: step s of every table's pass sequence (inverse passes, then the
//    coset expansion), grouped by pass shape; one launch per shape and step;
//  - leaves: every table's row-pair leaves in one launch;
//  - Merkle: step s of every table's tree (a wide level, or a subtree top of
//    up to 9 levels through LDS) in one launch per step;
//  - the roots gathered into one array and copied to pinned host memory.
// Blocks of a launch are assigned to tables by a prefix sum over their block
// counts, largest table first so the small ones fill the tail. The kernels
// read their arguments from the arena, so the launch sequence is captured
// once as a HIP graph and replayed; rebinding inputs or handing outputs over
// only rewrites the arena.
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "lvm_gpu.h"
#include "lvm_internal.hpp"
#include "prof.hpp"

using namespace lvm;

struct lvm_batch {
    int device = -1;
    bool use_graph = false;
    bool ran = false;
    bool dirty = true;  // the host arena changed since the last upload
    std::vector<lvm_table_desc> in;
    std::vector<size_t> order;  // tables by decreasing cell count
    std::vector<uint64_t *> coef, lde;
    std::vector<uint8_t *> nodes;

    enum Kind { NTT, LEAVES, TREE, ROOTS };
    struct Launch {
        Kind kind;
        int r;
        bool multi;
        size_t args_off, prefix_off;
        uint32_t count, blocks;
        std::string name;  // profiler name of the launch
        double bytes;      // compulsory HBM bytes (each input read once, each output written once)
    };
    std::vector<Launch> launches;
    size_t roots_off = 0, trees_off = 0;

    uint8_t *h_arena = nullptr;  // pinned staging of the descriptors
    uint8_t *d_arena = nullptr;
    size_t arena_bytes = 0;
    uint8_t *h_roots = nullptr;  // pinned, count x 32 bytes

    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

bool is_pow2(size_t x) { return x && !(x & (x - 1)); }
unsigned log2u(size_t x) { return (unsigned)__builtin_ctzll(x); }

// Appends descriptors at 16-byte aligned offsets; with a null base it only
// measures, so the same walk sizes the arena and then fills it.
struct Writer {
    uint8_t *base;
    size_t off = 0;
    template <class T>
    size_t put(const T *v, size_t n) {
        off = (off + 15) & ~(size_t)15;
        const size_t at = off;
        if (base && n) std::memcpy(base + at, v, n * sizeof(T));
        off += n * sizeof(T);
        return at;
    }
    size_t reserve(size_t bytes) {
        off = (off + 15) & ~(size_t)15;
        const size_t at = off;
        off += bytes;
        return at;
    }
};

}  // namespace

// Tree steps of one table with L leaves (power of two): wide levels one per
// step while more than 2^14 children remain, then subtree tops.
std::vector<std::pair<TreeStep, uint32_t>> lvm::tree_steps(uint8_t *nodes_b, uint64_t L) {
    std::vector<std::pair<TreeStep, uint32_t>> out;  // (step, blocks)
    uint64_t *nodes = (uint64_t *)nodes_b;
    uint64_t C = L;
    while (C > (1u << 14)) {
        TreeStep s{nodes, C - 1, C / 2 - 1, 256, 1};
        out.push_back({s, (uint32_t)(C / 512)});
        C /= 2;
    }
    while (C > 1) {
        const uint32_t span = (uint32_t)(C / 2 < 256 ? C / 2 : 256);
        const uint64_t blocks = C / (2 * span);
        TreeStep s{nodes, C - 1, C / 2 - 1, span, 32};
        out.push_back({s, (uint32_t)blocks});
        C = blocks;
    }
    return out;
}

namespace {

// Walks the whole launch sequence, writing descriptors through `w` and, when
// `record` is set, the launch list of the plan.
int layout(lvm_batch *b, Writer &w, bool record, const Tables &tb) {
    const size_t T = b->in.size();
    // NTT pass lists
    std::vector<std::vector<int>> R(T);
    std::vector<std::vector<PassArgs>> A(T);
    std::vector<std::vector<uint32_t>> COS(T);
    size_t max_steps = 0;
    for (size_t t = 0; t < T; t++) {
        const lvm_table_desc &d = b->in[t];
        int r[16];
        PassArgs a[16];
        uint32_t cos[16];
        const int k = lde_pass_list(d.trace, (uint32_t)d.m, log2u(d.n), log2u(d.blowup), d.weights, b->coef[t],
                                    b->lde[t], tb, r, a, cos);
        if (k <= 0) return fail(LVM_ERR_ARG, "table too small for a coset LDE");
        R[t].assign(r, r + k);
        A[t].assign(a, a + k);
        COS[t].assign(cos, cos + k);
        max_steps = std::max(max_steps, (size_t)k);
    }
    for (size_t s = 0; s < max_steps; s++) {
        // groups by (radix, multi) in order of first appearance
        std::vector<std::pair<int, bool>> keys;
        std::map<std::pair<int, bool>, std::vector<size_t>> groups;
        for (size_t t : b->order) {
            if (s >= R[t].size()) continue;
            auto key = std::make_pair(R[t][s], COS[t][s] > 1);
            if (!groups.count(key)) keys.push_back(key);
            groups[key].push_back(t);
        }
        for (auto &key : keys) {
            const auto &ts = groups[key];
            std::vector<PassArgs> args;
            std::vector<uint32_t> prefix(1, 0);
            uint64_t total = 0;
            for (size_t t : ts) {
                args.push_back(A[t][s]);
                total += pass_blocks(key.first, A[t][s]);
                if (total >= (1ull << 31)) return fail(LVM_ERR_ARG, "batched launch exceeds 2^31 blocks");
                prefix.push_back((uint32_t)total);
            }
            const size_t ao = w.put(args.data(), args.size());
            const size_t po = w.put(prefix.data(), prefix.size());
            if (record) {
                double bytes = 0;
                for (size_t t : ts) {
                    const PassArgs &a = A[t][s];
                    const double rows = (double)(a.total_groups << key.first);  // logical rows of the pass
                    const double out_rows = a.cosets_per_block > 1 ? rows * a.cosets_per_block : rows;
                    bytes += 8.0 * a.m * (rows + out_rows);
                }
                std::string name = "batch_ntt_s" + std::to_string(s) + "_r" + std::to_string(key.first) +
                                   (key.second ? "_coset" : "");
                b->launches.push_back({lvm_batch::NTT, key.first, key.second, ao, po, (uint32_t)ts.size(),
                                       (uint32_t)total, name, bytes});
            }
        }
    }
    // leaves
    {
        std::vector<LeafJob> jobs;
        std::vector<uint32_t> prefix(1, 0);
        uint64_t total = 0;
        for (size_t t : b->order) {
            const lvm_table_desc &d = b->in[t];
            const uint64_t L = d.n * d.blowup / 2;
            jobs.push_back({b->lde[t], d.m, L, (uint64_t *)(b->nodes[t] + (L - 1) * 32)});
            total += (L + 255) / 256;
            if (total >= (1ull << 31)) return fail(LVM_ERR_ARG, "batched launch exceeds 2^31 blocks");
            prefix.push_back((uint32_t)total);
        }
        const size_t jo = w.put(jobs.data(), jobs.size());
        const size_t po = w.put(prefix.data(), prefix.size());
        if (record) {
            double bytes = 0;
            for (const LeafJob &j : jobs) bytes += 16.0 * j.pitch * j.num_leaves + 32.0 * j.num_leaves;
            b->launches.push_back(
                {lvm_batch::LEAVES, 0, false, jo, po, (uint32_t)T, (uint32_t)total, "batch_leaves", bytes});
        }
    }
    // tree steps
    {
        std::vector<std::vector<std::pair<TreeStep, uint32_t>>> st(T);
        size_t max_tree = 0;
        for (size_t t = 0; t < T; t++) {
            const lvm_table_desc &d = b->in[t];
            st[t] = tree_steps(b->nodes[t], d.n * d.blowup / 2);
            max_tree = std::max(max_tree, st[t].size());
        }
        for (size_t s = 0; s < max_tree; s++) {
            std::vector<TreeStep> steps;
            std::vector<uint32_t> prefix(1, 0);
            uint64_t total = 0;
            double bytes = 0;
            for (size_t t : b->order) {
                if (s >= st[t].size()) continue;
                const TreeStep &ts = st[t][s].first;
                steps.push_back(ts);
                total += st[t][s].second;
                prefix.push_back((uint32_t)total);
                // parents built: span per block at the first level, halving per further level
                uint64_t parents = ts.span, cnt = ts.span;
                for (uint32_t lv = 1; lv < ts.levels && cnt > 1; lv++) parents += (cnt /= 2);
                bytes += 96.0 * parents * st[t][s].second;
            }
            const size_t so = w.put(steps.data(), steps.size());
            const size_t po = w.put(prefix.data(), prefix.size());
            if (record)
                b->launches.push_back({lvm_batch::TREE, 0, false, so, po, (uint32_t)steps.size(), (uint32_t)total,
                                       "batch_merkle_s" + std::to_string(s), bytes});
        }
    }
    // roots, in table index order
    {
        std::vector<uint64_t *> trees(T);
        for (size_t t = 0; t < T; t++) trees[t] = (uint64_t *)b->nodes[t];
        const size_t to = w.put(trees.data(), T);
        const size_t ro = w.reserve(T * 32);
        if (record) {
            b->trees_off = to;
            b->roots_off = ro;
            b->launches.push_back({lvm_batch::ROOTS, 0, false, to, ro, (uint32_t)T, 0, "batch_roots", 64.0 * T});
        }
    }
    return LVM_OK;
}

void enqueue(lvm_batch *b, hipStream_t st, bool profile) {
    uint8_t *base = b->d_arena;
    for (const auto &l : b->launches) {
        switch (l.kind) {
        case lvm_batch::NTT: {
            ProfScope ps(profile ? l.name.c_str() : nullptr, st);
            launch_pass_batched(l.r, (const PassArgs *)(base + l.args_off), (const uint32_t *)(base + l.prefix_off),
                                l.count, l.blocks, l.multi, st);
            break;
        }
        case lvm_batch::LEAVES: {
            ProfScope ps(profile ? l.name.c_str() : nullptr, st);
            launch_leaves_batched((const LeafJob *)(base + l.args_off), (const uint32_t *)(base + l.prefix_off),
                                  l.count, l.blocks, st);
            break;
        }
        case lvm_batch::TREE: {
            ProfScope ps(profile ? l.name.c_str() : nullptr, st);
            launch_tree_steps_batched((const TreeStep *)(base + l.args_off), (const uint32_t *)(base + l.prefix_off),
                                      l.count, l.blocks, st);
            break;
        }
        case lvm_batch::ROOTS:
            launch_gather_roots((uint64_t *const *)(base + l.args_off), l.count, (uint64_t *)(base + l.prefix_off),
                                st);
            break;
        }
    }
    hipMemcpyAsync(b->h_roots, base + b->roots_off, b->in.size() * 32, hipMemcpyDeviceToHost, st);
}

int validate(const lvm_table_desc *tables, size_t count) {
    if (!tables || count == 0) return fail(LVM_ERR_ARG, "no tables");
    for (size_t t = 0; t < count; t++) {
        const lvm_table_desc &d = tables[t];
        if (!d.trace || !d.weights) return fail(LVM_ERR_ARG, "null trace or weights in table " + std::to_string(t));
        if (!is_pow2(d.n) || d.n < 2) return fail(LVM_ERR_ARG, "trace rows must be a power of two >= 2");
        if (!is_pow2(d.blowup)) return fail(LVM_ERR_ARG, "blowup must be a power of two");
        if (d.m == 0 || d.m > (1u << 30)) return fail(LVM_ERR_ARG, "bad column count");
        if (log2u(d.n) + log2u(d.blowup) > 32) return fail(LVM_ERR_ARG, "LDE domain exceeds the 2^32 subgroup");
    }
    return LVM_OK;
}

int alloc_outputs(lvm_batch *b, size_t t) {
    const lvm_table_desc &d = b->in[t];
    const size_t N = d.n * d.blowup, L = N / 2;
    if (!b->lde[t] && lvm_malloc(&b->lde[t], N * d.m * 8) != hipSuccess) {
        b->lde[t] = nullptr;
        return fail(LVM_ERR_NOMEM, "device allocation of a batched LDE failed");
    }
    if (!b->nodes[t] && lvm_malloc(&b->nodes[t], (2 * L - 1) * 32) != hipSuccess) {
        b->nodes[t] = nullptr;
        return fail(LVM_ERR_NOMEM, "device allocation of a batched tree failed");
    }
    return LVM_OK;
}

void destroy(lvm_batch *b) {
    if (!b) return;
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->cap_stream) hipStreamDestroy(b->cap_stream);
    for (auto p : b->coef) lvm_free(p);
    for (auto p : b->lde) lvm_free(p);
    for (auto p : b->nodes) lvm_free(p);
    lvm_free(b->d_arena);
    hipStreamSynchronize(lib_stream());
    if (b->h_arena) hipHostFree(b->h_arena);
    if (b->h_roots) hipHostFree(b->h_roots);
    delete b;
}

// Refill the staged arena from the current pointers and mark it for upload.
int refill(lvm_batch *b, const Tables &tb) {
    Writer w{b->h_arena};
    std::vector<lvm_batch::Launch> saved;
    b->launches.swap(saved);
    int rc = layout(b, w, true, tb);
    // the walk is deterministic: the launch list it records matches the first
    bool same = rc == LVM_OK && saved.size() == b->launches.size() && w.off <= b->arena_bytes;
    for (size_t i = 0; same && i < saved.size(); i++)
        same = saved[i].args_off == b->launches[i].args_off && saved[i].blocks == b->launches[i].blocks;
    b->launches.swap(saved);
    if (rc != LVM_OK) return rc;
    if (!same) return fail(LVM_ERR_STATE, "batched plan layout changed");
    b->dirty = true;
    return LVM_OK;
}

}  // namespace

extern "C" {

int lvm_batch_new(const lvm_table_desc *tables, size_t count, int use_graph, lvm_batch **out) {
    if (!out) return fail(LVM_ERR_ARG, "null output");
    *out = nullptr;
    int rc = validate(tables, count);
    if (rc != LVM_OK) return rc;
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    lvm_batch *b = new lvm_batch();
    b->device = d->id;
    b->use_graph = use_graph != 0;
    b->in.assign(tables, tables + count);
    b->coef.assign(count, nullptr);
    b->lde.assign(count, nullptr);
    b->nodes.assign(count, nullptr);
    b->order.resize(count);
    for (size_t t = 0; t < count; t++) b->order[t] = t;
    std::stable_sort(b->order.begin(), b->order.end(), [&](size_t x, size_t y) {
        return tables[x].n * tables[x].blowup * tables[x].m > tables[y].n * tables[y].blowup * tables[y].m;
    });
    for (size_t t = 0; t < count; t++) {
        if (lvm_malloc(&b->coef[t], tables[t].n * tables[t].m * 8) != hipSuccess) {
            b->coef[t] = nullptr;
            destroy(b);
            return fail(LVM_ERR_NOMEM, "device allocation of a batched coefficient buffer failed");
        }
        rc = alloc_outputs(b, t);
        if (rc != LVM_OK) {
            destroy(b);
            return rc;
        }
    }
    Writer sizing{nullptr};
    rc = layout(b, sizing, true, d->tb);
    if (rc != LVM_OK) {
        destroy(b);
        return rc;
    }
    b->arena_bytes = sizing.off;
    if (hipHostMalloc((void **)&b->h_arena, b->arena_bytes, 0) != hipSuccess ||
        hipHostMalloc((void **)&b->h_roots, count * 32, 0) != hipSuccess ||
        lvm_malloc(&b->d_arena, b->arena_bytes) != hipSuccess) {
        destroy(b);
        return fail(LVM_ERR_NOMEM, "allocation of the batched plan arena failed");
    }
    rc = refill(b, d->tb);
    if (rc != LVM_OK) {
        destroy(b);
        return rc;
    }
    if (hipStreamSynchronize(lib_stream()) != hipSuccess) {
        destroy(b);
        return fail(LVM_ERR_DEVICE, "batched plan allocation failed");
    }
    if (b->use_graph) {
        hipError_t e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            enqueue(b, b->cap_stream, false);
            e = hipStreamEndCapture(b->cap_stream, &b->graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            destroy(b);
            return fail(LVM_ERR_DEVICE, std::string("graph capture of the batched plan failed: ") +
                                            hipGetErrorString(e));
        }
    }
    *out = b;
    return LVM_OK;
}

int lvm_batch_bind(lvm_batch *b, const lvm_table_desc *tables, size_t count) {
    if (!b) return fail(LVM_ERR_ARG, "null plan");
    if (count != b->in.size()) return fail(LVM_ERR_ARG, "table count differs from the plan");
    int rc = validate(tables, count);
    if (rc != LVM_OK) return rc;
    for (size_t t = 0; t < count; t++)
        if (tables[t].n != b->in[t].n || tables[t].m != b->in[t].m || tables[t].blowup != b->in[t].blowup)
            return fail(LVM_ERR_ARG, "table " + std::to_string(t) + " has another shape than the plan");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    // the staged arena may still feed an upload in flight
    if (hipStreamSynchronize(lib_stream()) != hipSuccess) return fail(LVM_ERR_DEVICE, "stream sync failed");
    std::vector<lvm_table_desc> old = b->in;
    b->in.assign(tables, tables + count);
    rc = refill(b, d->tb);
    if (rc != LVM_OK) b->in = old;
    return rc;
}

int lvm_batch_run(lvm_batch *b, uint8_t *roots_out) {
    if (!b) return fail(LVM_ERR_ARG, "null plan");
    Device *d = device_state();
    if (!d) return LVM_ERR_DEVICE;
    if (d->id != b->device) return fail(LVM_ERR_STATE, "plan belongs to another device");
    if (take_injected_fault()) return fail(LVM_ERR_INJECTED, "injected fault");
    hipStream_t st = lib_stream();
    bool need_refill = false;
    for (size_t t = 0; t < b->in.size(); t++) {
        if (!b->lde[t] || !b->nodes[t]) {
            int rc = alloc_outputs(b, t);
            if (rc != LVM_OK) return rc;
            need_refill = true;
        }
    }
    if (need_refill) {
        LVM_HIP(hipStreamSynchronize(st));
        int rc = refill(b, d->tb);
        if (rc != LVM_OK) return rc;
    }
    if (b->dirty) {
        LVM_HIP(hipMemcpyAsync(b->d_arena, b->h_arena, b->arena_bytes, hipMemcpyHostToDevice, st));
        b->dirty = false;
    }
    if (b->use_graph) {
        LVM_HIP(hipGraphLaunch(b->exec, st));
    } else {
        enqueue(b, st, true);
        LVM_HIP(hipGetLastError());
    }
    LVM_HIP(hipStreamSynchronize(st));
    b->ran = true;
    if (roots_out) std::memcpy(roots_out, b->h_roots, b->in.size() * 32);
    return LVM_OK;
}

int lvm_batch_take(lvm_batch *b, size_t index, lvm_lde **out) {
    if (!b || !out) return fail(LVM_ERR_ARG, "null argument");
    if (index >= b->in.size()) return fail(LVM_ERR_ARG, "table index out of range");
    if (!b->ran || !b->lde[index]) return fail(LVM_ERR_STATE, "table has no result to hand over");
    const lvm_table_desc &d = b->in[index];
    lvm_lde *h = new lvm_lde();
    h->device = b->device;
    h->lde = b->lde[index];
    h->nodes = b->nodes[index];
    h->n = d.n;
    h->N = d.n * d.blowup;
    h->w = d.m;
    h->blowup = d.blowup;
    h->L = h->N / 2;
    std::memcpy(h->root, b->h_roots + 32 * index, 32);
    b->lde[index] = nullptr;
    b->nodes[index] = nullptr;
    *out = h;
    return LVM_OK;
}

int lvm_batch_release(lvm_batch *b) {
    destroy(b);
    return LVM_OK;
}

int lvm_batch_lde_commit(const lvm_table_desc *tables, size_t count, uint8_t *roots_out, lvm_lde **handles_out) {
    lvm_batch *b = nullptr;
    int rc = lvm_batch_new(tables, count, 0, &b);
    if (rc != LVM_OK) return rc;
    rc = lvm_batch_run(b, roots_out);
    for (size_t t = 0; rc == LVM_OK && handles_out && t < count; t++) {
        rc = lvm_batch_take(b, t, &handles_out[t]);
        if (rc != LVM_OK)
            for (size_t u = 0; u < t; u++) {
                lvm_lde_release(handles_out[u]);
                handles_out[u] = nullptr;
            }
    }
    destroy(b);
    return rc;
}

int lvm_batch_launch_count(const lvm_batch *b, size_t *count) {
    if (!b || !count) return fail(LVM_ERR_ARG, "null argument");
    *count = b->launches.size();
    return LVM_OK;
}

int lvm_batch_launch_info(const lvm_batch *b, size_t index, char *name_out, size_t name_cap, double *bytes_out,
                          uint32_t *blocks_out) {
    if (!b) return fail(LVM_ERR_ARG, "null plan");
    if (index >= b->launches.size()) return fail(LVM_ERR_ARG, "launch index out of range");
    const auto &l = b->launches[index];
    if (name_out && name_cap) {
        const size_t k = std::min(name_cap - 1, l.name.size());
        std::memcpy(name_out, l.name.data(), k);
        name_out[k] = 0;
    }
    if (bytes_out) *bytes_out = l.bytes;
    if (blocks_out) *blocks_out = l.blocks;
    return LVM_OK;
}

int lvm_set_ntt_plan(unsigned log_n, const int *radices, int count) {
    if (count > 0 && !radices) return fail(LVM_ERR_ARG, "null radices");
    if (!set_ntt_plan(log_n, radices, count)) return fail(LVM_ERR_ARG, "radices must lie in 1..8 and sum to log_n");
    return LVM_OK;
}

int lvm_get_ntt_plan(unsigned log_n, int *radices_out, int *count_out) {
    if (!radices_out || !count_out) return fail(LVM_ERR_ARG, "null output");
    if (log_n < 1 || log_n > 32) return fail(LVM_ERR_ARG, "log_n out of range");
    *count_out = plan_passes(log_n, radices_out);
    return LVM_OK;
}

}  // extern "C"
