// Optional per-kernel event timing: when enabled, launch sites bracket their
// kernels with HIP events recorded on the launching stream, and callers read
// back per-name totals.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lvm {
bool prof_enabled();
void prof_set(bool on);
void *prof_begin(const char *name, hipStream_t st);
void prof_end(void *h, hipStream_t st);
int prof_query(const char *name, uint64_t *count, double *total_ms);
void prof_reset();
const char *last_error();
void inject_fault_after(long calls);

struct ProfScope {
    void *h;
    hipStream_t st;
    // a null name records nothing
    ProfScope(const char *name, hipStream_t s) : h(name ? prof_begin(name, s) : nullptr), st(s) {}
    ~ProfScope() { prof_end(h, st); }
};
}  // namespace lvm
