// Run-time compiled constraint kernels.
//
// The interpreter (compose.hip) decodes every node of the constraint program
// at every LDE row: scalar loads of the node words, a dispatch branch, slot
// traffic through LDS. Here the program is turned into straight-line HIP
// source once (one local value per node, operands resolved through the slot
// assignment at generation time, constants as literals, each constraint's
// value folded into its zerofier group's sum where its node is computed),
// compiled for gfx950 with hipRTC and cached per program text for the life of
// the process. The per-row walk then runs from registers, and the compiler
// schedules the trace reads ahead of their uses. Arithmetic is the same
// field.hpp code as the interpreter and performs the same operations per
// node, so both produce the same values.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "compose_common.hpp"
#include "lvm_internal.hpp"

extern const char *const lvm_jit_field_src;

namespace lvm {

namespace {

std::string u64lit(uint64_t v) {
    char b[32];
    std::snprintf(b, sizeof b, "0x%016llxull", (unsigned long long)v);
    return b;
}

std::string fp3lit(const uint64_t *v) { return "fp3{" + u64lit(v[0]) + ", " + u64lit(v[1]) + ", " + u64lit(v[2]) + "}"; }

struct Operand {
    std::string expr;
    bool ext;
};

// Nodes per load window (trace reads of a window are issued together); the
// environment variable LVM_JIT_LOAD_WINDOW overrides it for measurement.
size_t jit_load_window() {
    static const size_t w = [] {
        const char *e = std::getenv("LVM_JIT_LOAD_WINDOW");
        const long v = e ? std::strtol(e, nullptr, 10) : 0;
        return v > 0 ? (size_t)v : (size_t)48;
    }();
    return w;
}

// Straight-line per-row code of the program for natural LDE index r (the
// statements after `r` is set, with `a` the CompArgs and `N` the LDE size in
// scope); empty on a program this generator does not take (then the
// interpreter runs it).
// Largest program (in nodes) that is compiled: the run-time compiler's time
// grows faster than linearly with the straight-line code (a 1687-node program,
// the VM's 486-column KECCAK table with 84 interactions, takes 267 s at -O1 and
// over 15 min at -O3), so larger programs run on the interpreter kernel.
constexpr size_t JIT_MAX_NODES = 1024;

std::string generate_row(const lvm_program *pg, const std::vector<uint32_t> &root_node, const uint32_t *root_group,
                         int mode) {
    const size_t n = pg->num_nodes;
    if (n > JIT_MAX_NODES) return "";
    std::vector<std::vector<uint32_t>> emits(n);
    for (size_t c = 0; c < pg->num_roots; c++) {
        if (root_node[c] >= n || (mode == 0 && root_group[c] >= (uint32_t)COMP_MAX_GROUPS)) return "";
        emits[root_node[c]].push_back((uint32_t)c);
    }
    uint32_t max_off = 0;
    bool acc_group[COMP_MAX_GROUPS] = {false, false, false, false};
    std::vector<std::string> base_slot(pg->num_base_slots), ext_slot(pg->num_ext_slots);
    std::string body;
    body.reserve(n * 96);
    auto operand = [&](uint32_t enc, Operand &o) -> bool {
        const uint32_t kind = enc >> 29, x = enc & 0x1FFFFFFFu;
        switch (kind) {
        case K_BASE_SLOT:
            if (x >= base_slot.size() || base_slot[x].empty()) return false;
            o = {base_slot[x], false};
            return true;
        case K_EXT_SLOT:
            if (x >= ext_slot.size() || ext_slot[x].empty()) return false;
            o = {ext_slot[x], true};
            return true;
        case K_BASE_CONST:
            if (x >= pg->num_base_consts) return false;
            o = {u64lit(pg->base_consts[x]), false};
            return true;
        case K_EXT_CONST:
            if (x >= pg->num_ext_consts) return false;
            o = {fp3lit(pg->ext_consts + 3 * (size_t)x), true};
            return true;
        case K_RAP:
            if (x >= pg->num_rap) return false;
            o = {"ld3(a.rap + " + std::to_string(3 * (size_t)x) + ")", true};
            return true;
        case K_ALPHA:
            if (x >= pg->num_alpha) return false;
            o = {"ld3(a.alpha + " + std::to_string(3 * (size_t)x) + ")", true};
            return true;
        case 6:
            if (!pg->table_offset) return false;
            o = {"ld3(a.toff)", true};
            return true;
        default: return false;
        }
    };
    auto as_ext = [](const Operand &o) { return o.ext ? o.expr : "embed(" + o.expr + ")"; };
    // per-node source; trace reads are then moved to the front of their
    // window of jit_load_window() nodes so that a window's loads are in flight
    // together (they depend on nothing, so the order of values is unchanged)
    std::vector<std::string> node_text(n);
    std::vector<bool> is_load(n, false);
    for (size_t k = 0; k < n; k++) {
        const uint64_t w0 = pg->nodes[2 * k], w1 = pg->nodes[2 * k + 1];
        const uint32_t op = (uint32_t)w0, x = (uint32_t)(w0 >> 32), y = (uint32_t)w1, res = (uint32_t)(w1 >> 32);
        const bool rext = res >> 31;
        const uint32_t slot = res & 0x7FFFFFFFu;
        if (slot >= (rext ? ext_slot.size() : base_slot.size())) return "";
        const std::string v = "v" + std::to_string(k);
        std::string expr;
        bool ext = false;
        switch (op) {
        case OP_CONST_BASE:
            if (x >= pg->num_base_consts) return "";
            expr = u64lit(pg->base_consts[x]);
            break;
        case OP_CONST_EXT:
            if (x >= pg->num_ext_consts) return "";
            expr = fp3lit(pg->ext_consts + 3 * (size_t)x);
            ext = true;
            break;
        case OP_VAR: {
            const uint32_t col = x & 0xFFFF, is_main = (y >> 16) & 1, off = (y >> 8) & 0xFF;
            max_off = off > max_off ? off : max_off;
            const std::string idx = "(((uint64_t)" + std::to_string(col) + " << a.log_N) + f" + std::to_string(off) + ")";
            if (is_main) {
                expr = "a.main[" + idx + "]";
            } else {
                expr = "ld3(a.aux + 3 * " + idx + ")";
                ext = true;
            }
            break;
        }
        case OP_RAP:
        case OP_ALPHA:
        case OP_OFFSET: {
            Operand o;
            if (x > 0x1FFFFFFFu) return "";
            if (!operand((op == OP_RAP ? K_RAP : op == OP_ALPHA ? K_ALPHA : 6u) << 29 | x, o)) return "";
            expr = o.expr;
            ext = true;
            break;
        }
        case OP_ADD:
        case OP_SUB:
        case OP_MUL: {
            Operand p, q;
            if (!operand(x, p) || !operand(y, q)) return "";
            if (!p.ext && !q.ext) {
                expr = std::string(op == OP_ADD ? "add(" : op == OP_SUB ? "sub(" : "mul(") + p.expr + ", " + q.expr + ")";
            } else {
                ext = true;
                if (op == OP_MUL && !p.ext)
                    expr = "f3_mul_base(" + q.expr + ", " + p.expr + ")";
                else if (op == OP_MUL && !q.ext)
                    expr = "f3_mul_base(" + p.expr + ", " + q.expr + ")";
                else
                    expr = std::string(op == OP_ADD ? "f3_add(" : op == OP_SUB ? "f3_sub(" : "f3_mul(") + as_ext(p) +
                           ", " + as_ext(q) + ")";
            }
            break;
        }
        case OP_NEG: {
            Operand p;
            if (!operand(x, p)) return "";
            ext = p.ext;
            expr = (p.ext ? "f3_neg(" : "neg(") + p.expr + ")";
            break;
        }
        case OP_EMBED: {
            Operand p;
            if (!operand(x, p)) return "";
            ext = true;
            expr = as_ext(p);
            break;
        }
        default: return "";
        }
        if (ext != rext) return "";  // result class differs from its slot class: leave it to the interpreter
        std::string &t = node_text[k];
        is_load[k] = op == OP_VAR;
        t += std::string("        const ") + (ext ? "fp3 " : "uint64_t ") + v + " = " + expr + ";\n";
        (ext ? ext_slot : base_slot)[slot] = v;
        for (uint32_t c : emits[k]) {
            const std::string beta = "ld3(a.beta + " + std::to_string(3 * (size_t)c) + ")";
            if (mode == 0 && ext) {
                const std::string g = "g" + std::to_string(root_group[c]);
                t += "        " + g + " = f3_add(" + g + ", f3_mul(" + v + ", " + beta + "));\n";
            } else if (mode == 0) {
                // a base-valued constraint enters its group's sum through three
                // delayed-reduction dot products (one reduction per row)
                const uint32_t g = root_group[c];
                acc_group[g] = true;
                for (int q = 0; q < 3; q++)
                    t += "        acc_mad(A" + std::to_string(g) + "_" + std::to_string(q) + ", a.beta[" +
                            std::to_string(3 * (size_t)c + q) + "], " + v + ");\n";
            } else {
                t += "        store3(a.out + 3 * ((uint64_t)" + std::to_string(c) + " * N + r), " +
                        (ext ? v : "embed(" + v + ")") + ");\n";
            }
        }
    }
    const size_t window = jit_load_window();
    for (size_t w0 = 0; w0 < n; w0 += window) {
        const size_t w1 = std::min(n, w0 + window);
        for (size_t k = w0; k < w1; k++)
            if (is_load[k]) body += node_text[k];
        for (size_t k = w0; k < w1; k++)
            if (!is_load[k]) body += node_text[k];
    }
    std::string src;
    for (uint32_t off = 0; off <= max_off; off++)
        src += "        const uint64_t f" + std::to_string(off) + " = frame_row(a, r, " + std::to_string(off) + ");\n";
    if (mode == 0) {
        src += "        fp3 g0 = fp3{0, 0, 0}, g1 = fp3{0, 0, 0}, g2 = fp3{0, 0, 0}, g3 = fp3{0, 0, 0};\n";
        auto decl = [&](const std::string &A) { src += "        acc128 " + A + ";\n        acc_zero(" + A + ");\n"; };
        for (int g = 0; g < COMP_MAX_GROUPS; g++)
            if (acc_group[g])
                for (int q = 0; q < 3; q++) decl("A" + std::to_string(g) + "_" + std::to_string(q));
    }
    src += body;
    if (mode == 0) {
        for (int g = 0; g < COMP_MAX_GROUPS; g++)
            if (acc_group[g]) {
                const std::string A = "A" + std::to_string(g) + "_";
                src += "        g" + std::to_string(g) + " = f3_add(g" + std::to_string(g) + ", fp3{acc_reduce(" + A +
                       "0), acc_reduce(" + A + "1), acc_reduce(" + A + "2)});\n";
            }
        src += "        const fp3 gs[COMP_MAX_GROUPS] = {g0, g1, g2, g3};\n";
        src += "        const fp3 h = composition_tail(a, r, gs);\n";
        src += "        store3(a.out + 3 * (a.natural_out ? (uint64_t)r : (uint64_t)bitrev(r, a.log_N)), h);\n";
    }
    return src;
}

std::string source_prelude() {
    std::string src = "typedef unsigned long long uint64_t;\ntypedef unsigned int uint32_t;\n";
    src += lvm_jit_field_src;
    src += "\nnamespace lvm {\n__device__ __forceinline__ void store3(uint64_t *o, const fp3 &v) { o[0] = v.c0; o[1] = v.c1; "
           "o[2] = v.c2; }\n}\n";
    return src;
}

// The single-table kernel: a grid-stride loop over the natural rows.
std::string generate(const lvm_program *pg, const std::vector<uint32_t> &root_node, const uint32_t *root_group,
                     int mode) {
    const std::string row = generate_row(pg, root_node, root_group, mode);
    if (row.empty()) return "";
    std::string src = source_prelude();
    src += "extern \"C\" __global__ __launch_bounds__(128) void lvm_constraint_kernel(lvm::CompArgs a) {\n";
    src += "    using namespace lvm;\n    const uint64_t N = 1ull << a.log_N;\n";
    src += "    for (uint64_t q = (uint64_t)blockIdx.x * 128 + threadIdx.x; q < N; q += a.total_threads) {\n";
    src += "        const uint32_t r = (uint32_t)q;\n";
    src += row;
    src += "    }\n}\n";
    return src;
}

// Several tables' programs in one kernel: blocks [prefix[t], prefix[t+1])
// take table t's natural rows, 128 per block, one per thread; the body of
// each distinct program is a device function, the block's table picks it.
std::string generate_batch(const std::vector<const lvm_program *> &pgs,
                           const std::vector<std::vector<uint32_t>> &root_nodes,
                           const std::vector<const uint32_t *> &root_groups) {
    std::string src = source_prelude();
    std::vector<std::string> bodies;
    std::vector<size_t> fn_of(pgs.size());
    for (size_t t = 0; t < pgs.size(); t++) {
        const std::string row = generate_row(pgs[t], root_nodes[t], root_groups[t], 0);
        if (row.empty()) return "";
        size_t f = 0;
        while (f < bodies.size() && bodies[f] != row) f++;
        if (f == bodies.size()) bodies.push_back(row);
        fn_of[t] = f;
    }
    for (size_t f = 0; f < bodies.size(); f++) {
        src += "__device__ __forceinline__ void lvm_program_" + std::to_string(f) +
               "(const lvm::CompArgs &a, const uint32_t r) {\n    using namespace lvm;\n"
               "    const uint64_t N = 1ull << a.log_N;\n    (void)N;\n    {\n";
        src += bodies[f];
        src += "    }\n}\n";
    }
    src += "extern \"C\" __global__ __launch_bounds__(128) void lvm_constraint_batch_kernel(\n"
           "        const lvm::CompArgs *__restrict__ args, const uint32_t *__restrict__ prefix, uint32_t count) {\n"
           "    const uint32_t b = blockIdx.x;\n"
           "    uint32_t lo = 0, hi = count - 1;\n"
           "    while (lo < hi) {\n"
           "        const uint32_t mid = (lo + hi + 1) / 2;\n"
           "        if (prefix[mid] <= b) lo = mid; else hi = mid - 1;\n"
           "    }\n"
           "    const lvm::CompArgs &a = args[lo];\n"
           "    const uint64_t q = (uint64_t)(b - prefix[lo]) * 128 + threadIdx.x;\n"
           "    if (q >> a.log_N) return;\n"
           "    switch (lo) {\n";
    for (size_t t = 0; t < pgs.size(); t++)
        src += "    case " + std::to_string(t) + ": lvm_program_" + std::to_string(fn_of[t]) + "(a, (uint32_t)q); break;\n";
    src += "    default: break;\n    }\n}\n";
    return src;
}

// hipRTC compile to a gfx950 code object; empty on failure (the log goes to
// stderr when LVM_JIT_LOG is set).
std::vector<char> compile(const std::string &src) {
    std::vector<char> code;
    if (const char *dump = std::getenv("LVM_JIT_DUMP")) {
        // write the generated source for inspection (one file per program)
        static int seq = 0;
        const std::string path = std::string(dump) + "." + std::to_string(seq++) + ".hip";
        if (FILE *f = std::fopen(path.c_str(), "w")) {
            std::fwrite(src.data(), 1, src.size(), f);
            std::fclose(f);
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "lvm_constraints.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return code;
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    if (hiprtcCompileProgram(prog, 7, opts) == HIPRTC_SUCCESS) {
        size_t size = 0;
        if (hiprtcGetCodeSize(prog, &size) == HIPRTC_SUCCESS && size) {
            code.resize(size);
            if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) code.clear();
        }
    } else if (std::getenv("LVM_JIT_LOG")) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        std::fprintf(stderr, "constraint kernel compile failed:\n%s\n", log.c_str());
    }
    hiprtcDestroyProgram(&prog);
    return code;
}

// One compiled program: its code object, and the loaded kernel per device
// (each GPU loads its own module the first time it runs the program).
struct JitEntry {
    std::vector<char> code;
    std::unordered_map<int, hipFunction_t> fn;
};

// The kernel `name` of the entry's module on the current device (its module
// is loaded on first use and stays loaded). Called with g_jit_mu held.
hipFunction_t function_of(JitEntry *e, const char *name) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto it = e->fn.find(dev);
    if (it != e->fn.end()) return it->second;
    hipModule_t mod = nullptr;
    if (hipModuleLoadData(&mod, e->code.data()) != hipSuccess) return nullptr;
    hipFunction_t fn = nullptr;
    if (hipModuleGetFunction(&fn, mod, name) != hipSuccess) {
        hipModuleUnload(mod);
        return nullptr;
    }
    e->fn.emplace(dev, fn);
    return fn;
}

std::mutex g_jit_mu;
std::unordered_map<std::string, std::unique_ptr<JitEntry>> g_jit;  // keyed by generated source
int g_jit_mode = -1;  // -1: unset (LVM_CONSTRAINT_JIT decides, default on), 0: off, 1: on

// Cached entry for the source, compiling it on first sight; null when it
// does not compile. Called with g_jit_mu held.
JitEntry *entry_for(const std::string &src) {
    auto it = g_jit.find(src);
    if (it == g_jit.end()) {
        auto e = std::make_unique<JitEntry>();
        e->code = compile(src);
        it = g_jit.emplace(src, std::move(e)).first;
    }
    return it->second->code.empty() ? nullptr : it->second.get();
}

}  // namespace

bool root_nodes_of(const lvm_program *pg, std::vector<uint32_t> &root_node) {
    root_node.assign(pg->num_roots, 0);
    for (size_t c = 0; c < pg->num_roots; c++) {
        if (pg->root_nodes) {
            root_node[c] = pg->root_nodes[c];
            if (root_node[c] >= pg->num_nodes) return false;
        } else {
            // pinned root slots: the last node writing the root's slot
            size_t k = pg->num_nodes;
            while (k > 0 && (pg->nodes[2 * (k - 1) + 1] >> 32) != (uint32_t)pg->roots[c]) k--;
            if (k == 0) return false;
            root_node[c] = (uint32_t)(k - 1);
        }
    }
    return true;
}

bool boundary_steps_of(const uint32_t *b_col, const uint32_t *b_is_aux, const uint64_t *b_step, size_t num_boundary,
                       uint64_t trace_rows, BoundarySteps &jb, std::vector<uint32_t> *order_out) {
    jb = BoundarySteps{};
    if (num_boundary && !trace_rows) return false;
    std::vector<uint32_t> idx(num_boundary);
    for (size_t b = 0; b < num_boundary; b++) {
        const uint64_t s = b_step[b] % trace_rows;
        size_t k = 0;
        while (k < jb.step_row.size() && jb.step_row[k] != s) k++;
        if (k == jb.step_row.size()) {
            if (jb.step_row.size() == (size_t)COMP_MAX_STEPS) return false;
            jb.step_row.push_back(s);
        }
        idx[b] = (uint32_t)k;
    }
    jb.num_steps = (uint32_t)jb.step_row.size();
    std::vector<uint32_t> order;
    jb.step_off.assign(jb.num_steps + 1, 0);
    for (uint32_t k = 0; k < jb.num_steps; k++) {
        jb.step_off[k] = (uint32_t)order.size();
        for (size_t b = 0; b < num_boundary; b++)
            if (idx[b] == k) order.push_back((uint32_t)b);
    }
    jb.step_off[jb.num_steps] = (uint32_t)order.size();
    for (uint32_t b : order) {
        jb.col.push_back(b_col[b]);
        jb.is_aux.push_back(b_is_aux[b] ? 1u : 0u);
    }
    if (order_out) *order_out = std::move(order);
    return true;
}

bool jit_enabled() {
    std::lock_guard<std::mutex> lk(g_jit_mu);
    if (g_jit_mode < 0) {
        const char *e = std::getenv("LVM_CONSTRAINT_JIT");
        g_jit_mode = (e && e[0] == '0') ? 0 : 1;
    }
    return g_jit_mode == 1;
}

bool jit_prepare(const lvm_program *pg, const std::vector<uint32_t> &root_node, const uint32_t *root_group,
                 int mode) {
    const std::string src = generate(pg, root_node, root_group, mode);
    if (src.empty()) return false;
    std::lock_guard<std::mutex> lk(g_jit_mu);
    return entry_for(src) != nullptr;
}

bool jit_launch(const lvm_program *pg, const std::vector<uint32_t> &root_node, const uint32_t *root_group, int mode,
                CompArgs &args, uint64_t blocks, hipStream_t st) {
    const std::string src = generate(pg, root_node, root_group, mode);
    if (src.empty()) return false;
    hipFunction_t fn = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_jit_mu);
        JitEntry *e = entry_for(src);
        if (!e || !(fn = function_of(e, "lvm_constraint_kernel"))) return false;
    }
    void *params[] = {&args};
    return hipModuleLaunchKernel(fn, (unsigned)blocks, 1, 1, COMP_BLOCK, 1, 1, 0, st, params, nullptr) == hipSuccess;
}

bool jit_prepare_batch(const std::vector<const lvm_program *> &pgs, const std::vector<std::vector<uint32_t>> &root_nodes,
                       const std::vector<const uint32_t *> &root_groups) {
    const std::string src = generate_batch(pgs, root_nodes, root_groups);
    if (src.empty()) return false;
    std::lock_guard<std::mutex> lk(g_jit_mu);
    return entry_for(src) != nullptr;
}

bool jit_launch_batch(const std::vector<const lvm_program *> &pgs, const std::vector<std::vector<uint32_t>> &root_nodes,
                      const std::vector<const uint32_t *> &root_groups, const CompArgs *d_args,
                      const uint32_t *d_prefix, uint32_t count, uint64_t blocks, hipStream_t st) {
    const std::string src = generate_batch(pgs, root_nodes, root_groups);
    if (src.empty()) return false;
    hipFunction_t fn = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_jit_mu);
        JitEntry *e = entry_for(src);
        if (!e || !(fn = function_of(e, "lvm_constraint_batch_kernel"))) return false;
    }
    void *params[] = {&d_args, &d_prefix, &count};
    return hipModuleLaunchKernel(fn, (unsigned)blocks, 1, 1, COMP_BLOCK, 1, 1, 0, st, params, nullptr) == hipSuccess;
}

}  // namespace lvm

extern "C" {

int lvm_set_constraint_jit(int on) {
    std::lock_guard<std::mutex> lk(lvm::g_jit_mu);
    lvm::g_jit_mode = on ? 1 : 0;
    return LVM_OK;
}

int lvm_prepare_constraints(const lvm_program *program, const lvm_composition *comp) {
    if (!program || !program->nodes || !program->roots || program->num_roots == 0)
        return lvm::fail(LVM_ERR_ARG, "empty program");
    if (comp && !comp->root_group) return lvm::fail(LVM_ERR_ARG, "composition arrays missing");
    std::vector<uint32_t> root_node;
    if (!lvm::root_nodes_of(program, root_node)) return lvm::fail(LVM_ERR_ARG, "cannot locate the constraint nodes");
    if (!lvm::jit_prepare(program, root_node, comp ? comp->root_group : nullptr, comp ? 0 : 1))
        return lvm::fail(LVM_ERR_UNSUPPORTED, "constraint program could not be compiled");
    return LVM_OK;
}

}  // extern "C"
