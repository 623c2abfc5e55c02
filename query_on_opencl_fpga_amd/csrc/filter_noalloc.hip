// Stable filter compaction that allocates nothing: qo_filter_records_noalloc.
//
// The compaction's scratch (the decoupled look-back's status word per tile and
// the tile ticket) lives in `scratch`, a static device array of fixed size,
// instead of the per-stream workspace that qo_filter_records may grow with
// hipMalloc.  For callers that must not allocate: code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc.
//
// `scratch` names a host-side stand-in in host code; its device address is
// taken once per device with hipGetSymbolAddress when the library initialises
// that device (qo_init, or the first call of any entry point), and only that
// address is ever passed to kernels or fills.  Every call on a device shares the
// device's one array: calls that overlap on one device race (they must be
// ordered, e.g. on one stream); a call clears the words it uses with a fill
// kernel before its compaction kernel runs.
#include "qo_internal.hpp"

namespace qo {

// 2^17 tiles of 16384 records (16-byte aligned input; 8192 otherwise) plus the
// ticket: inputs up to 2^31 records (2^30 unaligned); 1 MiB + 256 bytes.
constexpr size_t kNoAllocTiles = size_t(1) << 17;
constexpr size_t kNoAllocBytes = kNoAllocTiles * 8 + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each device
}  // namespace

int filter_noalloc_setup(int device) {
    QO_REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond the %d the library tracks", device,
               kMaxDevices);
    if (g_scratch_ptr[device]) return QO_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(QO_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return QO_OK;
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_filter_records_noalloc(const qo_record *rin, size_t n, int32_t lo, int32_t hi, qo_record *rout,
                              uint64_t *d_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(d_count, "null count pointer");
    hipStream_t st = stream_of(s);
    if (n == 0) return launch_fill_u64(d_count, 0, 1, st);
    QO_REQUIRE(rin && rout, "null pointer");
    const size_t need = filter_records_ws_bytes(rin, n);
    QO_REQUIRE(need <= kNoAllocBytes,
               "qo_filter_records_noalloc: %zu records need %zu scratch bytes, more than its fixed %zu "
               "(at most 2^31 16-byte-aligned records or 2^30 otherwise)",
               n, need, kNoAllocBytes);
    void *ptr = g_scratch_ptr[props().device];
    QO_REQUIRE(ptr, "scratch of device %d not set up", props().device);
    return launch_filter_records_ws(rin, n, lo, hi, rout, d_count, ptr, st);
}

}  // extern "C"
