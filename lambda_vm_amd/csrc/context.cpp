// Library state: per-device twiddle tables, per-thread streams, thread-local error
// strings, test-only fault injection, and the kernel event profiler.
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include <cstring>

#include "lvm_internal.hpp"
#include "prof.hpp"

namespace lvm {

static thread_local std::string g_last_error;
static std::atomic<long> g_fault_countdown{-1};

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

bool take_injected_fault() {
    long v = g_fault_countdown.load();
    while (v >= 0) {
        if (g_fault_countdown.compare_exchange_weak(v, v - 1)) return v == 0;
    }
    return false;
}

void inject_fault_after(long calls) { g_fault_countdown.store(calls); }

static constexpr uint64_t ROOT_2_32 = 1753635133440165772ULL;  // goldilocks.rs:540

static std::mutex g_dev_mu;
static std::vector<Device *> g_devices;

static int build_tables(Device *d) {
    const size_t T = 1u << 16;
    std::vector<uint64_t> lo(T), hi(T), ilo(T), ihi(T), r(256), ir(256);
    const uint64_t w = ROOT_2_32, wi = inv(ROOT_2_32);
    uint64_t w16 = pow(w, T), wi16 = pow(wi, T);
    uint64_t a = 1, b = 1, c = 1, e = 1;
    for (size_t x = 0; x < T; x++) {
        lo[x] = a; a = mul(a, w);
        hi[x] = b; b = mul(b, w16);
        ilo[x] = c; c = mul(c, wi);
        ihi[x] = e; e = mul(e, wi16);
    }
    uint64_t w256 = pow(w, 1ull << 24), wi256 = pow(wi, 1ull << 24);
    uint64_t p = 1, q = 1;
    for (int k = 0; k < 256; k++) {
        r[k] = canon(p); p = mul(p, w256);
        ir[k] = canon(q); q = mul(q, wi256);
    }
    for (size_t x = 0; x < T; x++) {  // canonical tables (the NTT's inputs stay canonical)
        lo[x] = canon(lo[x]);
        hi[x] = canon(hi[x]);
        ilo[x] = canon(ilo[x]);
        ihi[x] = canon(ihi[x]);
    }
    uint64_t *buf = nullptr;
    const size_t total = 4 * T + 512;
    LVM_HIP(hipMalloc(&buf, total * sizeof(uint64_t)));
    d->tb.fwd_lo = buf;
    d->tb.fwd_hi = buf + T;
    d->tb.inv_lo = buf + 2 * T;
    d->tb.inv_hi = buf + 3 * T;
    d->tb.fwd_r = buf + 4 * T;
    d->tb.inv_r = buf + 4 * T + 256;
    LVM_HIP(hipMemcpy(d->tb.fwd_lo, lo.data(), T * 8, hipMemcpyHostToDevice));
    LVM_HIP(hipMemcpy(d->tb.fwd_hi, hi.data(), T * 8, hipMemcpyHostToDevice));
    LVM_HIP(hipMemcpy(d->tb.inv_lo, ilo.data(), T * 8, hipMemcpyHostToDevice));
    LVM_HIP(hipMemcpy(d->tb.inv_hi, ihi.data(), T * 8, hipMemcpyHostToDevice));
    LVM_HIP(hipMemcpy(d->tb.fwd_r, r.data(), 256 * 8, hipMemcpyHostToDevice));
    LVM_HIP(hipMemcpy(d->tb.inv_r, ir.data(), 256 * 8, hipMemcpyHostToDevice));
    return LVM_OK;
}

Device *device_state() {
    int id = 0;
    if (hipGetDevice(&id) != hipSuccess) {
        set_error("hipGetDevice failed (no GPU visible?)");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if ((int)g_devices.size() <= id) g_devices.resize(id + 1, nullptr);
    if (g_devices[id]) return g_devices[id];
    Device *d = new Device();
    d->id = id;
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreateWithFlags failed");
        delete d;
        return nullptr;
    }
    {
        // keep freed pool memory cached instead of trimming it at every sync
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, id) == hipSuccess) {
            uint64_t threshold = UINT64_MAX;
            hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
        }
    }
    if (build_tables(d) != LVM_OK) {
        hipStreamDestroy(d->stream);
        delete d;
        return nullptr;
    }
    g_devices[id] = d;
    return d;
}

namespace {
// Library streams are never destroyed: events recorded on them (the buffer
// cache keeps one per released block) stay valid for the life of the process.
// A thread that ends hands its streams back for the next thread. The hand-back
// makes no HIP call: a thread's destructors may run at process exit, after the
// runtime (or a profiler attached to it) has shut down. Work still queued on a
// handed-back stream stays ordered before whatever the next owner queues.
// The list and its lock are never destroyed, for the same reason.
std::mutex &g_stream_mu = *new std::mutex;
std::vector<std::vector<hipStream_t>> &g_spare_streams = *new std::vector<std::vector<hipStream_t>>;  // per device

struct ThreadStreams {
    std::vector<hipStream_t> s;
    std::vector<bool> owned;
    ~ThreadStreams() {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        for (size_t i = 0; i < s.size(); i++)
            if (s[i] && owned[i]) {
                if (g_spare_streams.size() <= i) g_spare_streams.resize(i + 1);
                g_spare_streams[i].push_back(s[i]);
            }
    }
};
thread_local ThreadStreams t_streams;

hipStream_t take_spare_stream(int id) {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if ((int)g_spare_streams.size() <= id || g_spare_streams[id].empty()) return nullptr;
    hipStream_t s = g_spare_streams[id].back();
    g_spare_streams[id].pop_back();
    return s;
}
}  // namespace

hipStream_t thread_stream_if_any(Device *d) {
    if (!d) return nullptr;
    auto &ts = t_streams;
    if ((int)ts.s.size() > d->id && ts.s[d->id]) return ts.s[d->id];
    return d->stream;
}

hipStream_t thread_stream(Device *d) {
    if (!d) return nullptr;
    auto &ts = t_streams;
    if ((int)ts.s.size() <= d->id) {
        ts.s.resize(d->id + 1, nullptr);
        ts.owned.resize(d->id + 1, false);
    }
    if (!ts.s[d->id]) {
        bool expected = false;
        if (d->stream_claimed.compare_exchange_strong(expected, true)) {
            ts.s[d->id] = d->stream;  // the first thread keeps the device's stream
        } else {
            hipStream_t s = take_spare_stream(d->id);
            if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return d->stream;
            ts.s[d->id] = s;
            ts.owned[d->id] = true;
        }
    }
    return ts.s[d->id];
}

hipStream_t lib_stream() { return thread_stream(device_state()); }

// ---------------- device allocation cache ----------------
namespace {
struct CachedBlock {
    void *p;
    hipStream_t st;  // stream that released it
    hipEvent_t ev;   // recorded at the release on st
};
struct AllocCache {
    std::mutex mu;
    std::unordered_map<void *, size_t> live;                   // block -> its class size
    std::unordered_map<size_t, std::vector<CachedBlock>> idle;  // class size -> released blocks
    std::vector<hipEvent_t> spare_events;
};
std::mutex g_cache_mu;
std::vector<std::unique_ptr<AllocCache>> g_caches;  // per device id

AllocCache *cache_for(int id) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    if ((int)g_caches.size() <= id) g_caches.resize(id + 1);
    if (!g_caches[id]) g_caches[id] = std::make_unique<AllocCache>();
    return g_caches[id].get();
}

// size classes: powers of two up to 1 MiB, then four classes per octave
size_t class_of(size_t b) {
    if (b < 256) b = 256;
    size_t c = 256;
    while (c < b) c <<= 1;
    if (c <= (size_t(1) << 20)) return c;
    const size_t step = c / 8;  // c/2 < b <= c: round up to a multiple of c/8
    return (b + step - 1) / step * step;
}
}  // namespace

hipError_t lvm_malloc_bytes(void **p, size_t bytes) {
    Device *d = device_state();
    if (!d) return hipErrorNoDevice;
    hipStream_t st = thread_stream(d);
    AllocCache *c = cache_for(d->id);
    const size_t cls = class_of(bytes);
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->idle.find(cls);
        if (it != c->idle.end()) {
            auto &v = it->second;
            for (size_t i = v.size(); i-- > 0;) {
                bool ready = v[i].st == st;
                if (!ready) {
                    const hipError_t q = hipEventQuery(v[i].ev);
                    ready = q == hipSuccess;
                    // a pending event is not an error: clear the thread's last
                    // error so a later hipGetLastError() check does not see it
                    if (q == hipErrorNotReady) (void)hipGetLastError();
                }
                if (ready) {
                    *p = v[i].p;
                    c->spare_events.push_back(v[i].ev);
                    v[i] = v.back();
                    v.pop_back();
                    c->live[*p] = cls;
                    return hipSuccess;
                }
            }
        }
    }
    hipError_t e = hipMallocAsync(p, cls, st);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        lvm_release_cached();
        e = hipMallocAsync(p, cls, st);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    c->live[*p] = cls;
    return hipSuccess;
}

hipError_t lvm_free(void *p) {
    if (!p) return hipSuccess;
    Device *d = device_state();
    if (!d) return hipErrorNoDevice;
    hipStream_t st = thread_stream(d);
    AllocCache *c = cache_for(d->id);
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->live.find(p);
    if (it == c->live.end()) return hipFreeAsync(p, st);
    const size_t cls = it->second;
    c->live.erase(it);
    hipEvent_t ev = nullptr;
    if (!c->spare_events.empty()) {
        ev = c->spare_events.back();
        c->spare_events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        return hipFreeAsync(p, st);
    }
    hipError_t e = hipEventRecord(ev, st);
    if (e != hipSuccess) {
        hipEventDestroy(ev);
        return hipFreeAsync(p, st);
    }
    c->idle[cls].push_back(CachedBlock{p, st, ev});
    return hipSuccess;
}

// Return every cached block of the current device to the pool.
void lvm_release_cached() {
    Device *d = device_state();
    if (!d) return;
    hipStream_t st = thread_stream(d);
    AllocCache *c = cache_for(d->id);
    std::lock_guard<std::mutex> lk(c->mu);
    for (auto &kv : c->idle)
        for (auto &b : kv.second) {
            hipEventSynchronize(b.ev);
            hipFreeAsync(b.p, st);
            c->spare_events.push_back(b.ev);
        }
    c->idle.clear();
    hipStreamSynchronize(st);
}

// ---------------- pinned staging ----------------
namespace {
constexpr size_t STAGE_DIRECT = 4u << 20;  // larger transfers bypass the area
struct PendingDownload {
    void *dst;
    size_t off, bytes;
};
struct Stage {
    uint8_t *buf = nullptr;
    size_t cap = 0, head = 0;
    hipStream_t last = nullptr;  // stream of the copies queued since the area was last drained
    std::vector<PendingDownload> pend;
};
// Pinned areas outlive their threads: a thread that ends hands its area to
// a spare list (no HIP call: a thread's destructors may run after the runtime
// has shut down) and a new thread takes one from there before allocating, so
// threads created per proof do not accumulate pinned memory.
struct SpareArea {
    uint8_t *buf;
    size_t cap;
};
std::mutex &g_area_mu = *new std::mutex;
std::vector<SpareArea> &g_spare_areas = *new std::vector<SpareArea>;

struct ThreadStage : Stage {
    ~ThreadStage() {
        if (!buf) return;
        std::lock_guard<std::mutex> lk(g_area_mu);
        g_spare_areas.push_back(SpareArea{buf, cap});
    }
};
thread_local ThreadStage t_stage;

void deliver(Stage &sg) {
    for (auto &pd : sg.pend) std::memcpy(pd.dst, sg.buf + pd.off, pd.bytes);
    sg.pend.clear();
    sg.head = 0;
    sg.last = nullptr;
}

// A region of `bytes` in the area; drains or grows it first when needed.
hipError_t stage_region(Stage &sg, size_t bytes, hipStream_t st, size_t *off) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (sg.head && (sg.last != st || sg.head + need > sg.cap)) {
        // queued copies still use the area: wait for them and deliver
        hipError_t e = hipStreamSynchronize(sg.last);
        if (e != hipSuccess) return e;
        deliver(sg);
    }
    if (!sg.buf) {
        // a spare area left by an ended thread
        std::lock_guard<std::mutex> lk(g_area_mu);
        for (size_t i = 0; i < g_spare_areas.size(); i++)
            if (g_spare_areas[i].cap >= need) {
                sg.buf = g_spare_areas[i].buf;
                sg.cap = g_spare_areas[i].cap;
                g_spare_areas.erase(g_spare_areas.begin() + (long)i);
                break;
            }
    }
    if (need > sg.cap) {
        if (sg.buf) hipHostFree(sg.buf);
        sg.buf = nullptr;
        size_t cap = sg.cap ? 2 * sg.cap : (size_t)256 << 10;
        while (cap < need) cap *= 2;
        hipError_t e = hipHostMalloc((void **)&sg.buf, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            sg.cap = 0;
            return e;
        }
        sg.cap = cap;
    }
    *off = sg.head;
    sg.head += need;
    sg.last = st;
    return hipSuccess;
}
}  // namespace

hipError_t stage_upload(void *dst_dev, const void *src_host, size_t bytes, hipStream_t st) {
    if (!bytes) return hipSuccess;
    if (bytes > STAGE_DIRECT) return hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, st);
    Stage &sg = t_stage;
    size_t off = 0;
    hipError_t e = stage_region(sg, bytes, st, &off);
    if (e != hipSuccess) return e;
    std::memcpy(sg.buf + off, src_host, bytes);
    return hipMemcpyAsync(dst_dev, sg.buf + off, bytes, hipMemcpyHostToDevice, st);
}

hipError_t stage_download(void *dst_host, const void *src_dev, size_t bytes, hipStream_t st) {
    if (!bytes) return hipSuccess;
    if (bytes > STAGE_DIRECT) return hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, st);
    Stage &sg = t_stage;
    size_t off = 0;
    hipError_t e = stage_region(sg, bytes, st, &off);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(sg.buf + off, src_dev, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) sg.pend.push_back(PendingDownload{dst_host, off, bytes});
    return e;
}

hipError_t stage_sync(hipStream_t st) {
    Stage &sg = t_stage;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (sg.head && sg.last && sg.last != st) {
        e = hipStreamSynchronize(sg.last);
        if (e != hipSuccess) return e;
    }
    deliver(sg);
    return hipSuccess;
}

StageScope::StageScope(hipStream_t s) : st(s), mark(t_stage.pend.size()) {}

StageScope::~StageScope() {
    Stage &sg = t_stage;
    if (sg.pend.size() > mark) {
        // an early return left downloads queued: let the copies land in the
        // area, then forget them (their destinations may no longer exist)
        if (sg.last) hipStreamSynchronize(sg.last);
        sg.pend.resize(mark);
        if (sg.pend.empty()) {
            sg.head = 0;
            sg.last = nullptr;
        }
    }
}

// ---------------- profiler ----------------
struct ProfRec {
    std::string name;  // copied: launch sites may pass names they later free
    hipEvent_t a, b;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;

bool prof_enabled() { return g_prof_on; }

void prof_set(bool on) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on;
}

void *prof_begin(const char *name, hipStream_t st) {
    if (!g_prof_on) return nullptr;
    ProfRec *r = new ProfRec{name, nullptr, nullptr};
    hipEventCreate(&r->a);
    hipEventCreate(&r->b);
    hipEventRecord(r->a, st);
    return r;
}

void prof_end(void *h, hipStream_t st) {
    if (!h) return;
    ProfRec *r = (ProfRec *)h;
    hipEventRecord(r->b, st);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(*r);
    delete r;
}

// Sum of event-measured durations for records whose name matches; clears
// nothing. Synchronises on the recorded events.
int prof_query(const char *name, uint64_t *count, double *total_ms) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    uint64_t c = 0;
    double t = 0;
    for (auto &r : g_prof) {
        if (name && r.name != name) continue;
        float ms = 0;
        if (hipEventSynchronize(r.b) != hipSuccess) return LVM_ERR_DEVICE;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return LVM_ERR_DEVICE;
        c++;
        t += ms;
    }
    *count = c;
    *total_ms = t;
    return LVM_OK;
}

void prof_reset() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &r : g_prof) {
        hipEventSynchronize(r.b);
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    g_prof.clear();
}

}  // namespace lvm
