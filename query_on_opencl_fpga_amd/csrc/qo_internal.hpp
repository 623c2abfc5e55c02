// Internal host-side runtime shared by every translation unit of libqo.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <thread>

#include "../../include/qo.h"

namespace qo {

// Error state (thread-local message, never exits).
int set_error(int code, const char *fmt, ...);
bool consume_injected_fault();

struct DeviceProps {
    int device = 0;
    int cu_count = 256;
    int lds_per_block = 65536;
    size_t hbm_bytes = 0;
    char name[64] = {0};
    char arch[32] = {0};
};
const DeviceProps &props();
int ensure_init();

struct DevCtl;
// The device control block of the current device (bounded-wait limit and the
// device error word).
DevCtl *dev_ctl();
// Reads and clears the device error word; QO_ERR_DEVICE_WAIT when a kernel's
// bounded wait ran out since the last check.  Call after synchronising the
// work in question.
int check_device_error();
// hipStreamSynchronize(st) followed by check_device_error().
int sync_and_check(hipStream_t st);
// Set-up of qo_filter_records_noalloc on `device` (filter_noalloc.hip): takes
// the device address of its static scratch array.  Called by the library's
// init with that device current.
int filter_noalloc_setup(int device);

// The stream to launch on: a caller stream, or the library's per-thread stream.
hipStream_t stream_of(qo_stream s);

// Key of per-stream library state (workspaces, counters): the stream, and for
// hipStreamPerThread also the calling thread.
struct StreamKey {
    hipStream_t s;
    std::thread::id t;
    bool operator<(const StreamKey &o) const { return s != o.s ? s < o.s : t < o.t; }
};
StreamKey stream_key_of(hipStream_t s);

// Per-stream grow-only workspace.  Operations on one stream are ordered by the
// stream, so one workspace per stream is never used by two operations at once.
void *workspace(hipStream_t s, size_t bytes, int *status);
// A second grow-only workspace per stream (slot 1) for a compound operation
// whose own temporaries must outlive the operators it calls, which take slot 0
// (qo_plan_execute).
constexpr int kWorkspaceSlots = 2;
void *workspace_slot(int slot, hipStream_t s, size_t bytes, int *status);

// Carves aligned sub-buffers out of a workspace; a null base only measures.
struct Carver {
    char *base;
    size_t off = 0;
    explicit Carver(void *b) : base(static_cast<char *>(b)) {}
    template <class T>
    T *take(size_t count, size_t align = 256) {
        off = (off + align - 1) / align * align;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

// Launch parameters per operator and size class: defaults, or the values the
// autotuner measured best (qo_autotune) and the tune file holds.
struct TuneParams {
    int block;
    int items;
    int param;
};
enum TuneOp {
    TUNE_FILTER_AGG = 0,
    TUNE_FILTER = 1,
    TUNE_PROBE = 2,
    TUNE_GROUPBY = 3,
    TUNE_SORT = 4,
    TUNE_JOIN_PART = 5,     // partitioned build/probe: items = LDS table slice KiB, param = PartMode
    TUNE_GROUPBY_PART = 6,  // partitioned group-by: items = groups per fine bucket, param = PartMode
    TUNE_JOIN_RANGE = 7,    // key-range join table: items = largest such table in MiB, param = PartMode
    TUNE_BATCH = 8,         // batched join (class of the mean |R|): items = S tiles per probe ticket,
                            // param = R rows per build tile
    TUNE_BATCH_LOAD = 9,    // batched join (class of the mean |R|): param = hash table load %
    TUNE_NUM_OPS = 10
};
// Partitioned-path modes: by size (the default), always, never.
// PART_LDS (join only): partition, and probe hash-mixed tables through LDS slices.
// PART_L2 (join only): by size, with every key-range table probed through L2
// (no key-range LDS slices).
enum PartMode { PART_AUTO = 0, PART_ALWAYS = 1, PART_NEVER = 2, PART_LDS = 3, PART_L2 = 4 };
TuneParams tune_get(int op, size_t n);
// The built-in default of an operator (used where no class is tuned).
TuneParams tune_builtin(int op);

inline int size_class(size_t n) {
    int c = 0;
    while (c < 63 && (size_t(1) << (c + 1)) <= n) ++c;
    return c;
}

inline unsigned grid_for(size_t work, unsigned per_block, unsigned cap) {
    size_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

// Internal launchers used by several operators (all asynchronous on `st`).
int launch_fill_u32(void *dst, uint32_t value, size_t count, hipStream_t st);
int launch_fill_u64(void *dst, uint64_t value, size_t count, hipStream_t st);

// Stable filter compaction of Records with caller-provided workspace
// (filter_records_ws_bytes bytes: the look-back status words and the ticket;
// cleared here by a fill kernel).
size_t filter_records_ws_bytes(const qo_record *rin, size_t n);
int launch_filter_records_ws(const qo_record *rin, size_t n, int32_t lo, int32_t hi, qo_record *rout,
                             uint64_t *d_count, void *ws, hipStream_t st);

// Exclusive scan with caller-provided workspace (see scan_workspace_bytes).
size_t scan_workspace_bytes(size_t n);
int launch_scan_excl_u32(const uint32_t *in, uint32_t *out, size_t n, void *ws, hipStream_t st);

// LSD radix sort of 64-bit keys with optional 32-bit values over bits [lo_bit, hi_bit).
size_t radix_workspace_bytes(size_t n, bool with_vals);
// Stable split of n Records (as little-endian uint64) into `parts` partitions
// by pid = hi32 % parts (1 <= parts <= 65536), in -> out (distinct arrays);
// bound (may be null) receives {start, end} per partition.
size_t split_records_ws_bytes(size_t n, uint32_t parts);
int launch_split_records(const uint64_t *in, size_t n, uint32_t parts, uint64_t *out, uint2 *bound, void *ws,
                         hipStream_t st);
int launch_radix_sort_u64(uint64_t *keys, uint32_t *vals, size_t n, int lo_bit, int hi_bit, void *ws,
                          hipStream_t st);
// Same, reading the element count from device memory (grid sized for `capacity`).
int launch_radix_sort_u64_dev(uint64_t *keys, uint32_t *vals, const uint64_t *d_n, size_t capacity,
                              int lo_bit, int hi_bit, void *ws, hipStream_t st);

// Sort of join rows by (key, s_payload) with caller workspace.
size_t sort_join_rows_ws_bytes(size_t m);
int sort_join_rows_ws(int32_t *key, int32_t *rpay, int32_t *spay, size_t m, void *ws, hipStream_t st);

// Hash group-by with caller workspace (groupby.hip); with filt, only tuples
// with lo <= filt[i] <= hi are aggregated (the filter is fused).
size_t groupby_ws_bytes(size_t group_capacity, size_t n, bool filtered = false);
int launch_groupby(const int32_t *key, const int32_t *val, size_t n, size_t group_capacity, int32_t *okey,
                   int64_t *ocount, int64_t *osum, int32_t *omin, int32_t *omax, uint64_t *d_num, void *ws,
                   hipStream_t st, const int32_t *filt = nullptr, int32_t flo = 0, int32_t fhi = 0);

// Join table helpers (join.hip).
size_t join_table_slots(size_t table_bytes);
// Build and probe take the table size so they can pick the partitioned path
// (workspace sized by the matching *_ws_bytes; may be nullptr when that is 0).
size_t join_build_ws_bytes(size_t table_bytes, size_t nr);
int launch_join_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                      void *ws, hipStream_t st);
size_t join_probe_ws_bytes(size_t table_bytes, size_t ns);
// Partitioned build and probes, join_part.hip.  The build: *used = false when
// the table is small enough for the direct path (nothing launched).  The
// probe launches the hash-mixed table's per-slice LDS probe and the key-range
// table's split + L2 probe where planned (each runs only for its table mode);
// *hash_part / *range_part say which were launched.
size_t join_part_ws_bytes(size_t table_bytes, size_t n, bool build);
void join_part_strategy(size_t table_bytes, size_t n, int mode, int *partitioned, int *bits, int *splits);
int launch_join_part_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                           void *ws, hipStream_t st, bool *used);
// Key-range table (decided on the device): the same split of R, by the table's
// own hash, then every slice built in LDS with plain stores and written out
// whole; the kernels return at once for a hash-mixed table.  *used: launched
// (the plan's size rule; the workspace is the partitioned build's).
int launch_range_part_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                            void *ws, hipStream_t st, bool *used);
// Sorts m sort-ready join rows (k64 = join_sort_key(key, s_payload), v32 =
// r_payload) with workspace rws (radix_workspace_bytes(m, true)) and writes
// them out as three int32 columns.  Only bits [0, hi_bit) of k64 are sorted:
// the caller passes 64 unless every key shares the bits above hi_bit.  With
// `flag` (one device word) the key bits are sorted first and runs of equal key
// by s_payload after, which synchronises the stream once to read the flag.
int sort_packed_join_rows(uint64_t *k64, uint32_t *v32, size_t m, void *rws, int32_t *key, int32_t *rpay,
                          int32_t *spay, hipStream_t st, int hi_bit = 64, uint32_t *flag = nullptr);
// Output columns of a probe: three int32 columns, or (k64, v32 set) the
// sort-ready (join_sort_key(key, s_payload), r_payload) rows.
struct JoinOut {
    int32_t *okey = nullptr, *orpay = nullptr, *ospay = nullptr;
    size_t cap = 0;
    uint64_t *k64 = nullptr;
    uint32_t *v32 = nullptr;
};
// hash_only: the table is known to be hash-mixed (the key-range routes are not
// launched).
int launch_join_part_probe(const void *table, size_t table_bytes, const int32_t *skey, const int32_t *spay,
                           const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, const JoinOut &out,
                           uint64_t *d_count, void *ws, hipStream_t st, bool *hash_part, bool *range_part,
                           bool hash_only = false);
int launch_join_probe(const void *table, size_t table_bytes, const int32_t *skey, const int32_t *spay,
                      const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, const JoinOut &out,
                      uint64_t *d_count, void *ws, hipStream_t st);

}  // namespace qo

#define QO_HIP(call)                                                                                \
    do {                                                                                            \
        hipError_t qo_e_ = (call);                                                                  \
        if (qo_e_ != hipSuccess)                                                                    \
            return ::qo::set_error(QO_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,           \
                                   hipGetErrorString(qo_e_));                                       \
    } while (0)

#define QO_LAUNCHED() QO_HIP(hipGetLastError())

#define QO_REQUIRE(cond, ...)                                                                       \
    do {                                                                                            \
        if (!(cond)) return ::qo::set_error(QO_ERR_INVALID_ARG, __VA_ARGS__);                      \
    } while (0)

#define QO_TRY(expr)                                                                                \
    do {                                                                                            \
        int qo_rc_ = (expr);                                                                        \
        if (qo_rc_ != QO_OK) return qo_rc_;                                                         \
    } while (0)

#define QO_ENTER()                                                                                  \
    do {                                                                                            \
        QO_TRY(::qo::ensure_init());                                                                \
        if (::qo::consume_injected_fault())                                                         \
            return ::qo::set_error(QO_ERR_INJECTED, "injected fault (qo_debug_inject_fault)");     \
    } while (0)
