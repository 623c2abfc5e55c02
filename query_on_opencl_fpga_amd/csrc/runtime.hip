// libqo runtime: error state, device selection, per-stream workspaces, memory
// and stream plumbing, fill kernels and the autotuned parameter table.
#include <dlfcn.h>

#include <atomic>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>

#include "qo_batch.hpp"
#include "qo_device.hpp"
#include "qo_internal.hpp"

namespace qo {

// The device control block: static device memory (one copy per device,
// initialised when the code object loads), so set-up allocates nothing.
__device__ DevCtl qo_dev_ctl = {0u, kDefaultSpinLimit, {}};

namespace {
thread_local std::string g_last_error;
std::atomic<int> g_inject{0};
std::mutex g_mu;
bool g_init = false;
DeviceProps g_props;
constexpr int kMaxDevices = 64;
DevCtl *g_ctl_dev[kMaxDevices];  // each device's copy of qo_dev_ctl
struct Ws {
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::map<std::pair<int, StreamKey>, Ws> g_ws;  // (slot, stream) -> workspace
void load_tune_file(const char *arch);
}  // namespace

StreamKey stream_key_of(hipStream_t s) {
    // hipStreamPerThread is one handle for every host thread, but each thread
    // launches on a device stream of its own: state bound to that handle is
    // kept per thread, so two threads never share a workspace or a counter.
    if (s == hipStreamPerThread) return StreamKey{s, std::this_thread::get_id()};
    return StreamKey{s, std::thread::id()};
}

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

bool consume_injected_fault() {
    int v = g_inject.load();
    while (v > 0) {
        if (g_inject.compare_exchange_weak(v, v - 1)) return true;
    }
    return false;
}

static int init_locked(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return set_error(QO_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count) return set_error(QO_ERR_INVALID_ARG, "device %d out of range", device);
    QO_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    QO_HIP(hipGetDeviceProperties(&p, device));
    g_props.device = device;
    g_props.cu_count = p.multiProcessorCount;
    g_props.lds_per_block = (int)p.sharedMemPerBlock;
    g_props.hbm_bytes = p.totalGlobalMem;
    snprintf(g_props.arch, sizeof(g_props.arch), "%s", p.gcnArchName);
    // Some drivers report an empty marketing name: fall back to the ISA.
    if (p.name[0]) snprintf(g_props.name, sizeof(g_props.name), "%s", p.name);
    else snprintf(g_props.name, sizeof(g_props.name), "AMD %s (%d CUs)", p.gcnArchName, p.multiProcessorCount);
    load_tune_file(g_props.arch);
    QO_REQUIRE(device < kMaxDevices, "device %d beyond the %d the library tracks", device, kMaxDevices);
    if (!g_ctl_dev[device])
        QO_HIP(hipGetSymbolAddress(reinterpret_cast<void **>(&g_ctl_dev[device]), HIP_SYMBOL(qo_dev_ctl)));
    QO_TRY(filter_noalloc_setup(device));
    g_init = true;
    return QO_OK;
}

int ensure_init() {
    if (g_init) {
        // Threads other than the initialising one start on device 0 by default.
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != g_props.device) QO_HIP(hipSetDevice(g_props.device));
        return QO_OK;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init) return QO_OK;
    return init_locked(0);
}

const DeviceProps &props() { return g_props; }

DevCtl *dev_ctl() { return g_ctl_dev[g_props.device]; }

int check_device_error() {
    DevCtl *g_ctl = dev_ctl();
    if (!g_ctl) return QO_OK;
    uint32_t err = 0;
    QO_HIP(hipMemcpy(&err, &g_ctl->error, sizeof(err), hipMemcpyDeviceToHost));
    if (!err) return QO_OK;
    const uint32_t zero = 0;
    QO_HIP(hipMemcpy(&g_ctl->error, &zero, sizeof(zero), hipMemcpyHostToDevice));
    return set_error(QO_ERR_DEVICE_WAIT,
                     "a device-side wait ran out of its bound (error bits 0x%x): results of the calls since the "
                     "last check are undefined",
                     err);
}

int sync_and_check(hipStream_t st) {
    QO_HIP(hipStreamSynchronize(st));
    return check_device_error();
}

hipStream_t stream_of(qo_stream s) {
    if (s) return static_cast<hipStream_t>(s);
    return hipStreamPerThread;
}

void *workspace(hipStream_t s, size_t bytes, int *status) { return workspace_slot(0, s, bytes, status); }

void *workspace_slot(int slot, hipStream_t s, size_t bytes, int *status) {
    *status = QO_OK;
    if (bytes == 0) bytes = 256;
    std::lock_guard<std::mutex> lk(g_mu);
    Ws &w = g_ws[{slot, stream_key_of(s)}];
    if (w.bytes >= bytes) return w.ptr;
    if (w.ptr) {
        // The old workspace may still be read by queued work on this stream.
        (void)hipStreamSynchronize(s);
        (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
    }
    size_t grow = bytes + bytes / 4;
    hipError_t e = hipMalloc(&w.ptr, grow);
    if (e != hipSuccess) {
        w.ptr = nullptr;
        *status = set_error(QO_ERR_NO_MEMORY, "workspace of %zu bytes: %s", grow, hipGetErrorString(e));
        return nullptr;
    }
    w.bytes = grow;
    return w.ptr;
}

__global__ void k_fill_u32(uint32_t *dst, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

__global__ void k_fill_u64(uint64_t *dst, uint64_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

int launch_fill_u32(void *dst, uint32_t value, size_t count, hipStream_t st) {
    if (count == 0) return QO_OK;
    unsigned g = grid_for(count, 256, (unsigned)props().cu_count * 8);
    k_fill_u32<<<g, 256, 0, st>>>(static_cast<uint32_t *>(dst), value, count);
    QO_LAUNCHED();
    return QO_OK;
}

// Plain streaming copy, 16 bytes per lane per iteration (HBM calibration kernel).
// NT selects the non-temporal loads the streaming operators use.
template <bool NT>
__device__ __forceinline__ uint4 copy_ld(const uint4 *p) {
    if constexpr (NT) {
        typedef uint32_t v4 __attribute__((ext_vector_type(4)));
        v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(p));
        return uint4{v.x, v.y, v.z, v.w};
    } else {
        return *p;
    }
}

// Device copy in 32 KiB block-contiguous chunks (each block copies chunks
// b, b + grid, ...; 8 x 16 bytes in flight per thread).  A grid-strided copy
// whose wave loads sit a grid apart moves 4.4-4.9 TB/s of read + write traffic
// on a 4 GiB copy; the chunked one 5.2 TB/s.
constexpr int kCopyU = 8;
constexpr size_t kCopyChunk = 256 * kCopyU;  // uint4 per chunk

template <bool NT>
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
    const size_t nch = n16 / kCopyChunk;
    for (size_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const size_t base = c * kCopyChunk + threadIdx.x;
        uint4 a[kCopyU];
#pragma unroll
        for (int u = 0; u < kCopyU; ++u) a[u] = copy_ld<NT>(src + base + u * 256);
#pragma unroll
        for (int u = 0; u < kCopyU; ++u) dst[base + u * 256] = a[u];
    }
    if (blockIdx.x == 0)
        for (size_t i = nch * kCopyChunk + threadIdx.x; i < n16; i += 256) dst[i] = src[i];
}

int launch_fill_u64(void *dst, uint64_t value, size_t count, hipStream_t st) {
    if (count == 0) return QO_OK;
    unsigned g = grid_for(count, 256, (unsigned)props().cu_count * 8);
    k_fill_u64<<<g, 256, 0, st>>>(static_cast<uint64_t *>(dst), value, count);
    QO_LAUNCHED();
    return QO_OK;
}

// ---------------------------------------------------------------------------
// Launch-parameter table: per operator and size class (floor(log2 n)).  An
// entry comes from the tune file loaded at init (written by the autotuner,
// qo_autotune + qo_tune_save), from qo_tune_set / qo_tune_load, or, for a
// class no entry covers, from the nearest tuned class of the same operator;
// an operator with no tuned class at all uses the defaults below (measured
// sweeps on MI355X at the benchmark sizes).
// ---------------------------------------------------------------------------
namespace {
TuneParams g_tune[TUNE_NUM_OPS][64];
bool g_tune_set[TUNE_NUM_OPS][64];
std::mutex g_tune_mu;

}  // namespace

TuneParams tune_builtin(int op) {
    switch (op) {
        case TUNE_FILTER_AGG: return {256, 4, 1};      // block, loads in flight per column, blocks per CU
        case TUNE_FILTER: return {256, 4, 0};          // block, uint4 loads per thread
        case TUNE_PROBE: return {256, 4, 50};          // block, int4 groups per thread, table load %
        case TUNE_GROUPBY: return {1024, 2048, 1};     // block, LDS slots, blocks per CU
        case TUNE_SORT: return {256, 16, 0};           // block, keys per thread
        case TUNE_JOIN_PART: return {256, 32, 0};      // block, LDS table slice KiB, mode
        case TUNE_GROUPBY_PART: return {512, 1024, 0}; // block, groups per fine bucket, mode
        case TUNE_JOIN_RANGE: return {256, 256, 0};    // block, largest key-range table MiB, mode
        case TUNE_BATCH: return {256, (int)kProbeTicketTiles, (int)kBuildTile};  // block, tiles per ticket, build tile
        case TUNE_BATCH_LOAD: return {256, 4, 50};     // block, -, batch hash table load %
        default: return {256, 4, 0};
    }
}

namespace {
// Loads <directory of libqo.so>/tune/<arch>.tune, or $QO_TUNE_FILE ("none": no file).
void load_tune_file(const char *arch) {
    std::string path;
    const char *env = getenv("QO_TUNE_FILE");
    if (env) {
        if (!*env || !strcmp(env, "none")) return;
        path = env;
    } else {
        Dl_info info;
        if (!dladdr(reinterpret_cast<void *>(&qo_init), &info) || !info.dli_fname) return;
        std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        std::string isa = arch;
        isa = isa.substr(0, isa.find(':'));
        path = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/tune/" + isa + ".tune";
    }
    if (FILE *f = fopen(path.c_str(), "r")) {
        fclose(f);
        (void)qo_tune_load(path.c_str());
    }
}
}  // namespace

TuneParams tune_get(int op, size_t n) {
    const int cls = size_class(n);
    if (op < 0 || op >= TUNE_NUM_OPS) return tune_builtin(op);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune_set[op][cls]) return g_tune[op][cls];
    for (int d = 1; d < 64; ++d) {  // nearest tuned class, the smaller one first
        if (cls - d >= 0 && g_tune_set[op][cls - d]) return g_tune[op][cls - d];
        if (cls + d < 64 && g_tune_set[op][cls + d]) return g_tune[op][cls + d];
    }
    return tune_builtin(op);
}

}  // namespace qo

using namespace qo;

extern "C" {

const char *qo_version(void) { return "qo-mi355x 0.1 (gfx950)"; }

const char *qo_last_error(void) { return g_last_error.c_str(); }

void qo_clear_error(void) { g_last_error.clear(); }

int qo_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return init_locked(device);
}

int qo_get_device_info(qo_device_info *out) {
    QO_TRY(ensure_init());
    QO_REQUIRE(out, "null output");
    memset(out, 0, sizeof(*out));
    snprintf(out->name, sizeof(out->name), "%s", g_props.name);
    snprintf(out->arch, sizeof(out->arch), "%s", g_props.arch);
    out->cu_count = g_props.cu_count;
    out->lds_bytes_per_block = g_props.lds_per_block;
    out->wave_size = 64;
    out->hbm_bytes = g_props.hbm_bytes;
    out->xcd_count = 8;
    return QO_OK;
}

int qo_malloc(void **ptr, size_t bytes) {
    QO_TRY(ensure_init());
    QO_REQUIRE(ptr, "null pointer");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return set_error(QO_ERR_NO_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return QO_OK;
}

int qo_free(void *ptr) {
    QO_TRY(ensure_init());
    if (ptr) QO_HIP(hipFree(ptr));
    return QO_OK;
}

// Host copies wait for all queued device work first, so they observe results of
// operators launched on any stream (including the per-thread default stream).
int qo_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    QO_TRY(ensure_init());
    if (bytes) {
        QO_HIP(hipDeviceSynchronize());
        QO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    }
    return QO_OK;
}

int qo_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    QO_TRY(ensure_init());
    if (bytes) {
        QO_HIP(hipDeviceSynchronize());
        QO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }
    return QO_OK;
}

int qo_memcpy_d2d(void *dst, const void *src, size_t bytes, qo_stream s) {
    QO_TRY(ensure_init());
    if (bytes) QO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_of(s)));
    return QO_OK;
}

static int copy_impl(void *dst, const void *src, size_t bytes, bool nt, qo_stream s) {
    QO_TRY(ensure_init());
    if (bytes == 0) return QO_OK;
    hipStream_t st = stream_of(s);
    if ((uintptr_t)dst % 16 || (uintptr_t)src % 16 || bytes % 16) {
        QO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
        return QO_OK;
    }
    unsigned g = grid_for(bytes / 16, (unsigned)kCopyChunk, (unsigned)g_props.cu_count * 32);
    if (nt) k_copy16<true><<<g, 256, 0, st>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), bytes / 16);
    else k_copy16<false><<<g, 256, 0, st>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), bytes / 16);
    QO_LAUNCHED();
    return QO_OK;
}

int qo_copy(void *dst, const void *src, size_t bytes, qo_stream s) { return copy_impl(dst, src, bytes, false, s); }

int qo_copy_nt(void *dst, const void *src, size_t bytes, qo_stream s) { return copy_impl(dst, src, bytes, true, s); }

int qo_fill_u32(void *dst, uint32_t value, size_t count, qo_stream s) {
    QO_TRY(ensure_init());
    return launch_fill_u32(dst, value, count, stream_of(s));
}

int qo_stream_create(qo_stream *s) {
    QO_TRY(ensure_init());
    hipStream_t st;
    QO_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = st;
    return QO_OK;
}

int qo_stream_destroy(qo_stream s) {
    QO_TRY(ensure_init());
    hipStream_t st = static_cast<hipStream_t>(s);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        (void)hipStreamSynchronize(st);
        for (int slot = 0; slot < kWorkspaceSlots; ++slot) {
            auto it = g_ws.find({slot, stream_key_of(st)});
            if (it == g_ws.end()) continue;
            if (it->second.ptr) (void)hipFree(it->second.ptr);
            g_ws.erase(it);
        }
    }
    QO_HIP(hipStreamDestroy(st));
    return QO_OK;
}

int qo_stream_sync(qo_stream s) {
    QO_TRY(ensure_init());
    return sync_and_check(stream_of(s));
}

int qo_device_sync(void) {
    QO_TRY(ensure_init());
    QO_HIP(hipDeviceSynchronize());
    return check_device_error();
}

int qo_debug_set_spin_limit(uint32_t sleeps) {
    QO_TRY(ensure_init());
    QO_HIP(hipDeviceSynchronize());
    QO_HIP(hipMemcpy(&dev_ctl()->spin_limit, &sleeps, sizeof(sleeps), hipMemcpyHostToDevice));
    return QO_OK;
}

int qo_event_create(qo_event *e) {
    QO_TRY(ensure_init());
    hipEvent_t ev;
    QO_HIP(hipEventCreate(&ev));
    *e = ev;
    return QO_OK;
}

int qo_event_destroy(qo_event e) {
    QO_TRY(ensure_init());
    QO_HIP(hipEventDestroy(static_cast<hipEvent_t>(e)));
    return QO_OK;
}

int qo_event_record(qo_event e, qo_stream s) {
    QO_TRY(ensure_init());
    QO_HIP(hipEventRecord(static_cast<hipEvent_t>(e), stream_of(s)));
    return QO_OK;
}

int qo_event_sync(qo_event e) {
    QO_TRY(ensure_init());
    QO_HIP(hipEventSynchronize(static_cast<hipEvent_t>(e)));
    return QO_OK;
}

int qo_event_query(qo_event e) {
    QO_TRY(ensure_init());
    hipError_t r = hipEventQuery(static_cast<hipEvent_t>(e));
    if (r == hipSuccess) return 1;
    if (r == hipErrorNotReady) return 0;
    set_error(QO_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(r));
    return -QO_ERR_HIP;
}

int qo_event_elapsed_ms(qo_event a, qo_event b, float *ms) {
    QO_TRY(ensure_init());
    QO_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
    return QO_OK;
}

int qo_release_workspaces(void) {
    QO_TRY(ensure_init());
    std::lock_guard<std::mutex> lk(g_mu);
    QO_HIP(hipDeviceSynchronize());
    for (auto &kv : g_ws)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    g_ws.clear();
    return QO_OK;
}

int qo_debug_inject_fault(int count) {
    g_inject.store(count < 0 ? 0 : count);
    return QO_OK;
}

int qo_tune_get(int op, int size_class, int *block, int *items, int *param) {
    QO_REQUIRE(op >= 0 && op < TUNE_NUM_OPS && size_class >= 0 && size_class < 64, "bad tune key");
    TuneParams p = tune_get(op, size_t(1) << size_class);
    if (block) *block = p.block;
    if (items) *items = p.items;
    if (param) *param = p.param;
    return QO_OK;
}

int qo_tune_reset(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int op = 0; op < TUNE_NUM_OPS; ++op)
        for (int c = 0; c < 64; ++c) g_tune_set[op][c] = false;
    return QO_OK;
}

int qo_tune_set(int op, int size_class, int block, int items, int param) {
    QO_REQUIRE(op >= 0 && op < TUNE_NUM_OPS && size_class >= 0 && size_class < 64, "bad tune key");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[op][size_class] = {block, items, param};
    g_tune_set[op][size_class] = true;
    return QO_OK;
}

// Text format, one entry per line: "op size_class block items param".
int qo_tune_load(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return set_error(QO_ERR_NOT_FOUND, "cannot open %s", path);
    int op, cls, b, it, pa;
    int n = 0;
    while (fscanf(f, "%d %d %d %d %d", &op, &cls, &b, &it, &pa) == 5) {
        if (op < 0 || op >= TUNE_NUM_OPS || cls < 0 || cls >= 64) continue;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[op][cls] = {b, it, pa};
        g_tune_set[op][cls] = true;
        ++n;
    }
    fclose(f);
    return QO_OK;
}

int qo_tune_save(const char *path) {
    FILE *f = fopen(path, "w");
    if (!f) return set_error(QO_ERR_NOT_FOUND, "cannot open %s", path);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int op = 0; op < TUNE_NUM_OPS; ++op)
        for (int c = 0; c < 64; ++c)
            if (g_tune_set[op][c])
                fprintf(f, "%d %d %d %d %d\n", op, c, g_tune[op][c].block, g_tune[op][c].items, g_tune[op][c].param);
    fclose(f);
    return QO_OK;
}

}  // extern "C"
