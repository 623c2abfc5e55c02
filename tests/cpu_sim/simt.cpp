// CPU build of libqo: work-items as fibers, synchronous HIP runtime over host
// memory.  See hip/hip_runtime.h for the execution model.
#include "hip/hip_runtime.h"

#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#if defined(__has_feature)
#if __has_feature(memory_sanitizer)
#define SIMT_MSAN 1
#endif
#if __has_feature(address_sanitizer)
#define SIMT_ASAN 1
extern "C" void __sanitizer_start_switch_fiber(void **fake_stack_save, const void *bottom, size_t size);
extern "C" void __sanitizer_finish_switch_fiber(void *fake_stack_save, const void **bottom_old, size_t *size_old);
#endif
#endif

extern "C" void simt_switch(void **save_sp, void *new_sp);
asm(R"(
.text
.globl simt_switch
.type simt_switch,@function
simt_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size simt_switch, .-simt_switch
)");

namespace simt {

Idx g_block_idx{0, 0, 0}, g_block_dim{1, 1, 1}, g_grid_dim{1, 1, 1};

namespace {

constexpr size_t kStack = 512 * 1024;
constexpr unsigned kMaxThreads = 1024;
constexpr size_t kLdsBytes = 160 * 1024;
constexpr uint64_t kSpinLimit = 200000000ull;

enum State { READY, BLOCKED, DONE };

struct Fiber {
    void *sp = nullptr;
    char *stack = nullptr;
    Idx tid{0, 0, 0};
    State st = READY;
    uint64_t result = 0;
    void *fake = nullptr;
};

struct Wave {
    unsigned live = 0, arrived = 0;
    int op = -1, line = 0, width = 64;
    uint64_t present = 0;
    uint64_t val[64];
    int64_t arg[64];
    std::vector<Fiber *> waiters;
};

struct Block {
    unsigned nthreads = 0, live = 0, arrived = 0, done = 0;
    std::vector<Fiber *> bar_waiters;
    Wave waves[kMaxThreads / 64];
    std::deque<Fiber *> ready;
    uint64_t spins = 0;
};

Fiber g_fibers[kMaxThreads];
Block g_blk;
Fiber *g_cur = nullptr;
void *g_sched_sp = nullptr;
const void *g_sched_bottom = nullptr;
size_t g_sched_size = 0;
const std::function<void()> *g_body = nullptr;
const char *g_kernel = "";
alignas(256) unsigned char g_lds[kLdsBytes];

[[noreturn]] void die(const char *fmt, const char *a = "", long b = 0, long c = 0) {
    fprintf(stderr, "[simt] kernel %s (block %u): ", g_kernel, g_block_idx.x);
    fprintf(stderr, fmt, a, b, c);
    fprintf(stderr, "\n");
    fflush(stderr);
    abort();
}

void to_sched(bool exiting) {
    Fiber *f = g_cur;
#ifdef SIMT_ASAN
    __sanitizer_start_switch_fiber(exiting ? nullptr : &f->fake, g_sched_bottom, g_sched_size);
#else
    (void)exiting;
#endif
    simt_switch(&f->sp, g_sched_sp);
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(f->fake, nullptr, nullptr);
#endif
}

void wave_complete(Wave &w) {
    for (Fiber *f : w.waiters) {
        const unsigned lane = f->tid.x % 64;
        uint64_t r = 0;
        switch (w.op) {
            case W_BALLOT: {
                uint64_t m = 0;
                for (unsigned l = 0; l < 64; ++l)
                    if (((w.present >> l) & 1) && w.val[l]) m |= uint64_t(1) << l;
                r = m;
                break;
            }
            case W_SHFL: {
                const unsigned src = (unsigned)(w.arg[lane] & 63);
                r = ((w.present >> src) & 1) ? w.val[src] : 0;
                break;
            }
            case W_SHFL_UP: {
                const unsigned d = (unsigned)w.arg[lane];
                r = (lane >= d && ((w.present >> (lane - d)) & 1)) ? w.val[lane - d] : w.val[lane];
                break;
            }
            case W_SHFL_XOR: {
                const unsigned src = (lane ^ (unsigned)w.arg[lane]) & 63;
                r = ((w.present >> src) & 1) ? w.val[src] : w.val[lane];
                break;
            }
            case W_FIRST: r = w.val[__builtin_ctzll(w.present)]; break;
            case W_SYNC: break;
        }
        f->result = r;
        f->st = READY;
        g_blk.ready.push_back(f);
    }
    w.waiters.clear();
    w.arrived = 0;
    w.present = 0;
    w.op = -1;
}

void barrier_check() {
    if (g_blk.arrived && g_blk.arrived == g_blk.live) {
        for (Fiber *f : g_blk.bar_waiters) {
            f->st = READY;
            g_blk.ready.push_back(f);
        }
        g_blk.bar_waiters.clear();
        g_blk.arrived = 0;
    }
}

void fiber_entry() {
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
#endif
    (*g_body)();
    Fiber *f = g_cur;
    f->st = DONE;
    ++g_blk.done;
    --g_blk.live;
    Wave &w = g_blk.waves[f->tid.x / 64];
    --w.live;
    if (w.arrived && w.arrived == w.live) wave_complete(w);
    barrier_check();
    to_sched(true);
    die("finished fiber resumed");
}

char *stack_of(unsigned i) {
    static char *base = nullptr;
    if (!base) {
        void *p = mmap(nullptr, kStack * kMaxThreads, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                       -1, 0);
        if (p == MAP_FAILED) die("cannot map fiber stacks");
        base = static_cast<char *>(p);
    }
    return base + (size_t)i * kStack;
}

void run_block(unsigned nthreads) {
    Block &b = g_blk;
    b.nthreads = b.live = nthreads;
    b.arrived = b.done = 0;
    b.spins = 0;
    b.bar_waiters.clear();
    b.ready.clear();
    for (unsigned w = 0; w < (nthreads + 63) / 64; ++w) {
        Wave &wv = b.waves[w];
        wv.live = std::min(64u, nthreads - w * 64);
        wv.arrived = 0;
        wv.present = 0;
        wv.op = -1;
        wv.waiters.clear();
    }
    for (unsigned i = 0; i < nthreads; ++i) {
        Fiber &f = g_fibers[i];
        f.stack = stack_of(i);
        f.tid = Idx{i, 0, 0};
        f.st = READY;
        f.fake = nullptr;
        uintptr_t top = (reinterpret_cast<uintptr_t>(f.stack) + kStack) & ~uintptr_t(15);
        void **s = reinterpret_cast<void **>(top);
        *--s = nullptr;                                     // fiber_entry's return address (never used)
        *--s = reinterpret_cast<void *>(&fiber_entry);      // popped by ret
        for (int r = 0; r < 6; ++r) *--s = nullptr;         // rbp rbx r12 r13 r14 r15
        f.sp = s;
        b.ready.push_back(&f);
    }
    while (b.done < nthreads) {
        if (b.ready.empty()) {
            std::string msg;
            for (unsigned w = 0; w < (nthreads + 63) / 64; ++w)
                if (b.waves[w].arrived)
                    msg += " wave " + std::to_string(w) + ": " + std::to_string(b.waves[w].arrived) + "/" +
                           std::to_string(b.waves[w].live) + " lanes at the collective of line " +
                           std::to_string(b.waves[w].line) + ";";
            die("deadlock: %s barrier %ld/%ld", msg.c_str(), (long)b.arrived, (long)b.live);
        }
        Fiber *f = b.ready.front();
        b.ready.pop_front();
        g_cur = f;
#ifdef SIMT_ASAN
        void *fake = nullptr;
        __sanitizer_start_switch_fiber(&fake, f->stack, kStack);
#endif
        simt_switch(&g_sched_sp, f->sp);
#ifdef SIMT_ASAN
        __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
        g_cur = nullptr;
    }
}

// ---- streams, capture, events ----
struct Graph {
    std::vector<std::function<void()>> ops;
};
std::mutex g_mu;
std::map<void *, Graph *> g_capturing;  // stream -> graph being recorded
hipError_t g_last = hipSuccess;

bool record_if_capturing(void *st, std::function<void()> op) {
    auto it = g_capturing.find(st);
    if (it == g_capturing.end()) return false;
    it->second->ops.push_back(std::move(op));
    return true;
}

}  // namespace

Idx &thread_idx() { return g_cur->tid; }

unsigned lane_of_current() { return g_cur->tid.x % 64; }

unsigned char *dyn_lds() { return g_lds; }

void syncthreads() {
    Fiber *f = g_cur;
    f->st = BLOCKED;
    g_blk.bar_waiters.push_back(f);
    ++g_blk.arrived;
    barrier_check();
    to_sched(false);
}

void sleep_yield() {
    if (++g_blk.spins > kSpinLimit) die("spin wait did not end after %s%ld sleeps", "", (long)kSpinLimit);
    g_blk.ready.push_back(g_cur);
    to_sched(false);
}

uint64_t wave_op(WaveOp op, uint64_t v, int64_t a, int width, int line) {
    Fiber *f = g_cur;
    Wave &w = g_blk.waves[f->tid.x / 64];
    const unsigned lane = f->tid.x % 64;
    if (w.arrived == 0) {
        w.op = op;
        w.line = line;
        w.width = width;
    } else if (w.op != op || w.line != line) {
        die("divergent wave collectives: line %s%ld and line %ld", "", (long)w.line, (long)line);
    }
    w.val[lane] = v;
    w.arg[lane] = a;
    w.present |= uint64_t(1) << lane;
    ++w.arrived;
    f->st = BLOCKED;
    w.waiters.push_back(f);
    if (w.arrived == w.live) wave_complete(w);
    to_sched(false);
    return f->result;
}

void launch(const LaunchCfg &c, const char *name, std::function<void()> body) {
    const unsigned nt = c.block.x * c.block.y * c.block.z;
    const uint64_t nb = (uint64_t)c.grid.x * c.grid.y * c.grid.z;
    if (nt == 0 || nt > kMaxThreads || nb == 0 || c.block.y != 1 || c.block.z != 1 || c.grid.y != 1 ||
        c.grid.z != 1 || c.lds > kLdsBytes) {
        fprintf(stderr, "[simt] invalid launch of %s: grid %u block %u lds %zu\n", name, c.grid.x, c.block.x, c.lds);
        g_last = hipErrorInvalidConfiguration;
        return;
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        LaunchCfg cc = c;
        if (record_if_capturing(c.stream, [cc, name, body]() { launch(cc, name, body); })) return;
    }
    if (getenv("SIMT_TRACE")) fprintf(stderr, "[simt] %s<<<%u, %u, %zu>>>\n", name, c.grid.x, c.block.x, c.lds);
    g_kernel = name;
    g_body = &body;
    g_grid_dim = Idx{c.grid.x, 1, 1};
    g_block_dim = Idx{c.block.x, 1, 1};
    for (unsigned bx = 0; bx < c.grid.x; ++bx) {
        g_block_idx = Idx{bx, 0, 0};
        run_block(nt);
    }
    g_body = nullptr;
    g_kernel = "";
}

// Canaries after carved buffers (see Carver::take in the CPU build): while
// enabled, every buffer a Carver hands out is followed by 256 bytes of
// pattern; a kernel that writes past its buffer into the next one damages it.
namespace {
constexpr unsigned char kCanary = 0xCB;
std::mutex g_canary_mu;
std::vector<std::pair<unsigned char *, size_t>> g_canaries;
std::vector<size_t> g_canary_buf;  // bytes of the buffer each canary follows
std::vector<std::string> g_canary_at;  // file:line of the take() that carved that buffer
std::map<char *, size_t> g_allocs;  // device allocation -> bytes
bool g_canary_on = false;
}  // namespace
bool canary_on() { return g_canary_on; }
void canary_set(void *p, size_t n, size_t buf_bytes, const char *file, int line) {
    std::lock_guard<std::mutex> lk(g_canary_mu);
    // A later carve of the same memory (a workspace reused by the next step)
    // retires the canaries its buffer and canary cover.
    const unsigned char *lo = static_cast<unsigned char *>(p) - buf_bytes, *hi = static_cast<unsigned char *>(p) + n;
    size_t k = 0;
    for (size_t i = 0; i < g_canaries.size(); ++i) {
        const unsigned char *c = g_canaries[i].first;
        if (c < hi && c + g_canaries[i].second > lo) continue;
        g_canaries[k] = g_canaries[i];
        g_canary_buf[k] = g_canary_buf[i];
        g_canary_at[k] = g_canary_at[i];
        ++k;
    }
    g_canaries.resize(k);
    g_canary_buf.resize(k);
    g_canary_at.resize(k);
    memset(p, kCanary, n);
    g_canaries.emplace_back(static_cast<unsigned char *>(p), n);
    g_canary_buf.push_back(buf_bytes);
    const char *slash = strrchr(file, '/');
    g_canary_at.push_back(std::string(slash ? slash + 1 : file) + ":" + std::to_string(line));
}
void canary_retire(const void *p, size_t n) {
    std::lock_guard<std::mutex> lk(g_canary_mu);
    const unsigned char *lo = static_cast<const unsigned char *>(p), *hi = lo + n;
    size_t k = 0;
    for (size_t i = 0; i < g_canaries.size(); ++i) {
        const unsigned char *c = g_canaries[i].first;
        if (c < hi && c + g_canaries[i].second > lo) continue;
        g_canaries[k] = g_canaries[i];
        g_canary_buf[k] = g_canary_buf[i];
        g_canary_at[k] = g_canary_at[i];
        ++k;
    }
    g_canaries.resize(k);
    g_canary_buf.resize(k);
    g_canary_at.resize(k);
}
// Forgets the canaries inside an allocation that is being freed.
static void canary_forget(char *p) {
    std::lock_guard<std::mutex> lk(g_canary_mu);
    auto a = g_allocs.find(p);
    if (a == g_allocs.end()) return;
    const size_t n = a->second;
    g_allocs.erase(a);
    size_t k = 0;
    for (size_t i = 0; i < g_canaries.size(); ++i) {
        const char *c = (const char *)g_canaries[i].first;
        if (c >= p && c < p + n) continue;
        g_canaries[k] = g_canaries[i];
        g_canary_buf[k] = g_canary_buf[i];
        g_canary_at[k] = g_canary_at[i];
        ++k;
    }
    g_canaries.resize(k);
    g_canary_buf.resize(k);
    g_canary_at.resize(k);
}
}  // namespace simt

extern "C" void qo_sim_canary_enable(int on) { simt::g_canary_on = on != 0; }
// Number of damaged canaries since the last check (each reported on stderr with
// its allocation and offset); clears the list.
extern "C" long qo_sim_canary_check(void) {
    std::lock_guard<std::mutex> lk(simt::g_canary_mu);
    long bad = 0;
    for (size_t ci = 0; ci < simt::g_canaries.size(); ++ci) {
        const auto &c = simt::g_canaries[ci];
        for (size_t i = 0; i < c.second; ++i)
            if (c.first[i] != simt::kCanary) {
                char *base = nullptr;
                for (const auto &a : simt::g_allocs)
                    if ((char *)c.first >= a.first && (char *)c.first < a.first + a.second) base = a.first;
                if (bad < 16)
                    fprintf(stderr,
                            "[canary] damaged at allocation %p + %zu (byte %zu of the canary after a buffer of %zu "
                            "bytes carved at %s; canary %zu of %zu)\n",
                            (void *)base, base ? (size_t)((char *)c.first - base) : (size_t)0, i,
                            simt::g_canary_buf[ci], simt::g_canary_at[ci].c_str(), ci, simt::g_canaries.size());
                ++bad;
                break;
            }
    }
    fprintf(stderr, "[canary] %zu checked, %ld damaged\n", simt::g_canaries.size(), bad);
    simt::g_canaries.clear();
    simt::g_canary_buf.clear();
    simt::g_canary_at.clear();
    return bad;
}

using namespace simt;

// ---- runtime API ----
struct ihipStream_t {
    int id;
};
struct ihipEvent_t {
    std::chrono::steady_clock::time_point t;
};
struct ihipGraph {
    Graph g;
};
struct hipGraphExec {
    Graph g;
};

hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
    memset(p, 0, sizeof(*p));
    snprintf(p->name, sizeof(p->name), "CPU build (simt fibers)");
    snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950-cpu");
    p->multiProcessorCount = 256;
    p->sharedMemPerBlock = 160 * 1024;
    p->totalGlobalMem = size_t(16) << 30;
    p->warpSize = 64;
    return hipSuccess;
}
hipError_t hipMalloc(void **p, size_t bytes) {
    const size_t n = (bytes + 255) / 256 * 256;
    *p = aligned_alloc(256, n ? n : 256);
    if (!*p) return hipErrorOutOfMemory;
#if !defined(SIMT_MSAN)
    // Device memory starts with no defined contents (a MemorySanitizer build
    // leaves it uninitialised instead, so that reads of it are reported).
    static const int fill = getenv("SIMT_FILL") ? atoi(getenv("SIMT_FILL")) : 0xA5;
    memset(*p, fill, n ? n : 256);
#endif
    {
        std::lock_guard<std::mutex> lk(simt::g_canary_mu);
        simt::g_allocs[static_cast<char *>(*p)] = n ? n : 256;
    }
    return hipSuccess;
}
hipError_t hipFree(void *p) {
    simt::canary_forget(static_cast<char *>(p));
    free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    if (n) memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (record_if_capturing(st, [d, s, n, k]() { hipMemcpy(d, s, n, k); })) return hipSuccess;
    }
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemset(void *d, int v, size_t n) {
    memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) { return hipMemset(d, v, n); }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    static std::atomic<int> next{16};
    *s = new ihipStream_t{next++};
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s) {
    if (s != hipStreamPerThread && s) delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
hipError_t hipGetLastError() {
    hipError_t e = g_last;
    g_last = hipSuccess;
    return e;
}
hipError_t hipPeekAtLastError() { return g_last; }
const char *hipGetErrorString(hipError_t e) {
    switch (e) {
        case hipSuccess: return "no error";
        case hipErrorInvalidValue: return "invalid value";
        case hipErrorOutOfMemory: return "out of memory";
        case hipErrorNotReady: return "not ready";
        case hipErrorInvalidConfiguration: return "invalid configuration";
        default: return "error";
    }
}
hipError_t hipEventCreate(hipEvent_t *e) {
    *e = new ihipEvent_t{std::chrono::steady_clock::now()};
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    e->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
// Work runs when it is launched, so a wait between streams has nothing to wait for.
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *status) {
    std::lock_guard<std::mutex> lk(g_mu);
    *status = g_capturing.count(s) ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_capturing[s] = new Graph;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_capturing.find(s);
    if (it == g_capturing.end()) return hipErrorInvalidValue;
    *g = new ihipGraph{std::move(*it->second)};
    delete it->second;
    g_capturing.erase(it);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
    *e = new hipGraphExec{g->g};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto &op : e->g.ops) op();
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}
