// Reference-named entry points (qo_compat.h) over the native operators.  Each
// call is synchronous on the calling thread's stream, like the reference's
// clFinish after every launch (host/host/common.cpp:28-31).
#include <cstdlib>
#include <cstring>

#include "../../include/qo_compat.h"
#include "qo_internal.hpp"

struct qo_mem_s {
    void *ptr;
    size_t bytes;
    bool owned;
};

namespace qo {
int launch_group_startpos(const qo_record *sorted, size_t n, int32_t *startPos, uint64_t *d_num, void *ws,
                          hipStream_t st);
size_t group_startpos_ws_bytes(size_t n);
int launch_aggregate_groups(const qo_record *s, size_t n, const int32_t *startPos, size_t G, int op, int32_t *out,
                            void *ws, hipStream_t st);
size_t aggregate_ws_bytes(size_t G, size_t n);

// The reference exits on any OpenCL error (AOCL_Utils.cpp:228-251); here that is opt-in.
static void compat_fail(const char *where) {
    const char *e = getenv("QO_COMPAT_EXIT_ON_ERROR");
    if (e && e[0] == '1') {
        fprintf(stderr, "ERROR: %s: %s\n", where, qo_last_error());
        exit(1);
    }
}

__global__ void k_gather_records(const qo_record *rin, size_t n, const qo_record *ragg, size_t nagg, qo_record *out,
                                 uint32_t *bad) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint32_t loc = rin[i].x;
        if (loc < nagg) out[i] = ragg[loc];
        else {
            out[i] = qo_record{0, 0};
            atomicExch(bad, 1u);
        }
    }
}
}  // namespace qo

using namespace qo;

#define QO_COMPAT_CHECK(expr, where, ret)                                                                           \
    do {                                                                                                            \
        if ((expr) != QO_OK) {                                                                                      \
            compat_fail(where);                                                                                     \
            return ret;                                                                                             \
        }                                                                                                           \
    } while (0)

extern "C" {

qo_mem qo_mem_create(size_t bytes) {
    void *p = nullptr;
    if (qo_malloc(&p, bytes) != QO_OK) {
        compat_fail("qo_mem_create");
        return nullptr;
    }
    return new qo_mem_s{p, bytes, true};
}

qo_mem qo_mem_wrap(void *device_ptr, size_t bytes) { return new qo_mem_s{device_ptr, bytes, false}; }

void *qo_mem_ptr(qo_mem m) { return m ? m->ptr : nullptr; }

size_t qo_mem_size(qo_mem m) { return m ? m->bytes : 0; }

void qo_mem_free(qo_mem m) {
    if (!m) return;
    if (m->owned && m->ptr) qo_free(m->ptr);
    delete m;
}

int qo_mem_write(qo_mem m, const void *host, size_t bytes) {
    if (!m || bytes > m->bytes) return set_error(QO_ERR_INVALID_ARG, "write of %zu bytes into a smaller buffer", bytes);
    return qo_memcpy_h2d(m->ptr, host, bytes);
}

int qo_mem_read(void *host, qo_mem m, size_t bytes) {
    if (!m || bytes > m->bytes) return set_error(QO_ERR_INVALID_ARG, "read of %zu bytes from a smaller buffer", bytes);
    return qo_memcpy_d2h(host, m->ptr, bytes);
}

// ---- scan -------------------------------------------------------------------
void initScan(int rLen, ScanPara *SP) {
    if (!SP) return;
    // The reference's plan (testScan.cpp:26-75): pass = floor(log N / log 256),
    // one less when that is integral; its per-level buffers are not needed by
    // the single-pass scan, so outputBuffer / blockSumBuffer / tempBuffer stay
    // NULL and `ws` holds the look-back workspace instead.
    SP->outputBuffer = nullptr;
    SP->blockSumBuffer = nullptr;
    SP->tempBuffer = nullptr;
    SP->gLength = rLen > 0 ? (unsigned)rLen : 0u;
    SP->blockSize = 256;
    unsigned pass = 0;
    for (double x = rLen > 0 ? (double)rLen : 1.0; x >= 256.0; x /= 256.0) ++pass;
    if (rLen > 1 && pass > 0) {
        double v = 1.0;
        for (unsigned i = 0; i < pass; ++i) v *= 256.0;
        if (v == (double)rLen) --pass;
    }
    SP->pass = pass;
    SP->ws = nullptr;
    if (qo_malloc(&SP->ws, scan_workspace_bytes(rLen > 0 ? (size_t)rLen : 1)) != QO_OK) compat_fail("initScan");
}

void closeScan(ScanPara *SP) {
    if (!SP) return;
    if (SP->ws) qo_free(SP->ws);
    SP->ws = nullptr;
}

// Exclusive scan of `rLen` int32 from d_Src (starting at element `offset`) into d_Dst.
static void scan_compat(const char *who, qo_mem d_Src, size_t offset, int rLen, qo_mem d_Dst, ScanPara *SP) {
    if (rLen <= 0) return;
    if (ensure_init() != QO_OK || !d_Src || !d_Dst || (offset + (size_t)rLen) * 4 > d_Src->bytes ||
        (size_t)rLen * 4 > d_Dst->bytes) {
        set_error(QO_ERR_INVALID_ARG, "%s: bad arguments", who);
        compat_fail(who);
        return;
    }
    hipStream_t st = stream_of(nullptr);
    bool own = !(SP && SP->ws && (int)SP->gLength >= rLen);
    void *ws = own ? nullptr : SP->ws;
    if (own && qo_malloc(&ws, scan_workspace_bytes((size_t)rLen)) != QO_OK) {
        compat_fail(who);
        return;
    }
    int rc = launch_scan_excl_u32(static_cast<const uint32_t *>(d_Src->ptr) + offset,
                                  static_cast<uint32_t *>(d_Dst->ptr), (size_t)rLen, ws, st);
    if (rc == QO_OK) rc = sync_and_check(st);
    if (own) qo_free(ws);
    if (rc != QO_OK) compat_fail(who);
}

void scanImpl(qo_mem d_Src, int rLen, qo_mem d_Dst, int *, void *, void *, int *, double *, ScanPara *SP, int) {
    scan_compat("scanImpl", d_Src, 0, rLen, d_Dst, SP);
}

void scanImpl_offset(qo_mem d_Src, int offset, int rLen, qo_mem d_Dst, int *, void *, void *, int *, double *,
                     ScanPara *SP, int) {
    if (offset < 0) {
        set_error(QO_ERR_INVALID_ARG, "scanImpl_offset: negative offset");
        compat_fail("scanImpl_offset");
        return;
    }
    scan_compat("scanImpl_offset", d_Src, (size_t)offset, rLen, d_Dst, SP);
}

// ---- selection ----------------------------------------------------------------
void filterImpl(qo_mem d_Rin, int /*beginPos: unused by filter_op.cl:6-31,218-238*/, int rLen, qo_mem *d_Rout,
                int *outSize, int, int, int smallKey, int largeKey, int *, void *, void *, int *, double *, int) {
    if (!d_Rin || !d_Rout || !outSize || rLen < 0) {
        set_error(QO_ERR_INVALID_ARG, "filterImpl: bad arguments");
        compat_fail("filterImpl");
        return;
    }
    *d_Rout = qo_mem_create(sizeof(Record) * (size_t)(rLen > 0 ? rLen : 1));
    if (!*d_Rout) return;
    uint64_t *d_count = nullptr;
    QO_COMPAT_CHECK(qo_malloc(reinterpret_cast<void **>(&d_count), 8), "filterImpl", );
    int rc = qo_filter_records(static_cast<const qo_record *>(d_Rin->ptr), (size_t)rLen, smallKey, largeKey,
                               static_cast<qo_record *>((*d_Rout)->ptr), d_count, nullptr);
    uint64_t cnt = 0;
    if (rc == QO_OK) rc = qo_memcpy_d2h(&cnt, d_count, 8);
    if (rc == QO_OK) rc = check_device_error();
    qo_free(d_count);
    QO_COMPAT_CHECK(rc, "filterImpl", );
    *outSize = (int)cnt;
}

int range_selection(qo_mem d_Rin, int rLen, int rangeSmallKey, int rangeLargeKey, qo_mem *d_Rout, int numThreadPB,
                    int numBlock, int *index, void *eventList, void *Kernel, int *Flag_CPU_GPU, double *burden,
                    int _CPU_GPU) {
    int outSize = -1;
    filterImpl(d_Rin, 0, rLen, d_Rout, &outSize, numThreadPB, numBlock, rangeSmallKey, rangeLargeKey, index,
               eventList, Kernel, Flag_CPU_GPU, burden, _CPU_GPU);
    return outSize;
}

int CL_RangeSelection(Record *h_Rin, int rLen, int rangeSmallKey, int rangeLargeKey, Record **h_Rout, int numThreadPB,
                      int numBlock, int _CPU_GPU) {
    if (!h_Rout || rLen < 0 || (rLen > 0 && !h_Rin)) {
        set_error(QO_ERR_INVALID_ARG, "CL_RangeSelection: bad arguments");
        compat_fail("CL_RangeSelection");
        return -1;
    }
    qo_mem d_Rin = qo_mem_create(sizeof(Record) * (size_t)(rLen > 0 ? rLen : 1));
    if (!d_Rin) return -1;
    if (qo_mem_write(d_Rin, h_Rin, sizeof(Record) * (size_t)rLen) != QO_OK) {
        qo_mem_free(d_Rin);
        compat_fail("CL_RangeSelection");
        return -1;
    }
    qo_mem d_Rout = nullptr;
    int outSize = range_selection(d_Rin, rLen, rangeSmallKey, rangeLargeKey, &d_Rout, numThreadPB, numBlock, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, _CPU_GPU);
    if (outSize >= 0) {
        *h_Rout = static_cast<Record *>(malloc(sizeof(Record) * (size_t)(outSize > 0 ? outSize : 1)));
        if (qo_mem_read(*h_Rout, d_Rout, sizeof(Record) * (size_t)outSize) != QO_OK) {
            compat_fail("CL_RangeSelection");
            outSize = -1;
        }
    }
    qo_mem_free(d_Rin);
    qo_mem_free(d_Rout);
    return outSize;
}

int point_selection(qo_mem d_Rin, int rLen, int matchingKeyValue, qo_mem *d_Rout, int numThreadPB, int numBlock,
                    int *index, void *eventList, void *Kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU) {
    return range_selection(d_Rin, rLen, matchingKeyValue, matchingKeyValue, d_Rout, numThreadPB, numBlock, index,
                           eventList, Kernel, Flag_CPU_GPU, burden, _CPU_GPU);
}

int CL_PointSelection(Record *h_Rin, int rLen, int matchingKeyValue, Record **h_Rout, int numThreadPB, int numBlock,
                      int _CPU_GPU) {
    return CL_RangeSelection(h_Rin, rLen, matchingKeyValue, matchingKeyValue, h_Rout, numThreadPB, numBlock,
                             _CPU_GPU);
}

int CL_PointSelectionOnly(qo_mem d_Rin, int rLen, int matchingKeyValue, qo_mem *d_Rout, int numThreadPB,
                          int numBlock, int _CPU_GPU) {
    return range_selection(d_Rin, rLen, matchingKeyValue, matchingKeyValue, d_Rout, numThreadPB, numBlock, nullptr,
                           nullptr, nullptr, nullptr, nullptr, _CPU_GPU);
}

int CL_RangeSelectionOnly(qo_mem d_Rin, int rLen, int rangeSmallKey, int rangeLargeKey, qo_mem *d_Rout,
                          int numThreadPB, int numBlock, int _CPU_GPU) {
    return range_selection(d_Rin, rLen, rangeSmallKey, rangeLargeKey, d_Rout, numThreadPB, numBlock, nullptr, nullptr,
                           nullptr, nullptr, nullptr, _CPU_GPU);
}

// ---- split / scatter / gather ---------------------------------------------------------
// True when buffer m holds at least `count` elements of `elem` bytes.
static bool holds(qo_mem m, size_t count, size_t elem) { return m && m->ptr && m->bytes >= count * elem; }

void splitImpl(qo_mem d_R, int rLen, int numPart, qo_mem d_S, qo_mem d_bound, int, int, int /*SPLIT_TYPE*/, int *,
               void *, void *, int *, double *, int) {
    if (rLen < 0 || numPart < 1 || (rLen > 0 && (!holds(d_R, rLen, 8) || !holds(d_S, rLen, 8))) ||
        (d_bound && !holds(d_bound, numPart, 8))) {
        set_error(QO_ERR_INVALID_ARG, "splitImpl: bad arguments");
        compat_fail("splitImpl");
        return;
    }
    int rc = qo_split_records(d_R ? static_cast<const qo_record *>(d_R->ptr) : nullptr, (size_t)rLen,
                              (uint32_t)numPart, d_S ? static_cast<qo_record *>(d_S->ptr) : nullptr,
                              d_bound ? static_cast<qo_record *>(d_bound->ptr) : nullptr, nullptr);
    if (rc == QO_OK) rc = sync_and_check(stream_of(nullptr));
    QO_COMPAT_CHECK(rc, "splitImpl", );
}

void scatterImpl(qo_mem d_R, int rLen, qo_mem d_loc, qo_mem d_S, int, int, int *, void *, void *, int *, double *,
                 int) {
    if (rLen < 0 || (rLen > 0 && (!holds(d_R, rLen, 8) || !holds(d_loc, rLen, 4) || !holds(d_S, rLen, 8)))) {
        set_error(QO_ERR_INVALID_ARG, "scatterImpl: bad arguments");
        compat_fail("scatterImpl");
        return;
    }
    if (rLen == 0) return;
    int rc = qo_scatter_records(static_cast<const qo_record *>(d_R->ptr), (size_t)rLen,
                                static_cast<const int32_t *>(d_loc->ptr), static_cast<qo_record *>(d_S->ptr), nullptr);
    if (rc == QO_OK) rc = sync_and_check(stream_of(nullptr));
    QO_COMPAT_CHECK(rc, "scatterImpl", );
}

void scatterImpl_forPart(qo_mem d_R, int rLen, int /*numPart*/, qo_mem d_loc, qo_mem d_S, int numThreadPB,
                         int numBlock, int *index, void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden,
                         int _CPU_GPU) {
    scatterImpl(d_R, rLen, d_loc, d_S, numThreadPB, numBlock, index, eventList, kernel, Flag_CPU_GPU, burden,
                _CPU_GPU);
}

void gatherImpl(qo_mem d_R, int rLen, qo_mem d_loc, qo_mem d_S, int sLen, int, int, int *, void *, void *, int *,
                double *, int) {
    if (rLen < 0 || sLen < 0 || (rLen > 0 && !holds(d_R, rLen, 8)) ||
        (sLen > 0 && (!holds(d_loc, sLen, 4) || !holds(d_S, sLen, 8)))) {
        set_error(QO_ERR_INVALID_ARG, "gatherImpl: bad arguments");
        compat_fail("gatherImpl");
        return;
    }
    if (sLen == 0) return;
    int rc = qo_gather_records(d_R ? static_cast<const qo_record *>(d_R->ptr) : nullptr, (size_t)rLen,
                               static_cast<const int32_t *>(d_loc->ptr), (size_t)sLen,
                               static_cast<qo_record *>(d_S->ptr), nullptr);
    if (rc == QO_OK) rc = sync_and_check(stream_of(nullptr));
    QO_COMPAT_CHECK(rc, "gatherImpl", );
}

// ---- sort -------------------------------------------------------------------------
void radixSortImpl_sm(qo_mem d_R, int rLen, int, int, int, int *, void *, void *, int *, double *, int) {
    if (!d_R || rLen < 0) {
        set_error(QO_ERR_INVALID_ARG, "radixSortImpl_sm: bad arguments");
        compat_fail("radixSortImpl_sm");
        return;
    }
    QO_COMPAT_CHECK(qo_sort_records(static_cast<qo_record *>(d_R->ptr), (size_t)rLen, nullptr), "radixSortImpl_sm", );
    QO_COMPAT_CHECK(qo_stream_sync(nullptr), "radixSortImpl_sm", );
}

void radixSortImpl(qo_mem d_R, int rLen, int keybits, int numThreadPB, int numBlock, int *index, void *eventList,
                   void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU) {
    radixSortImpl_sm(d_R, rLen, keybits, numThreadPB, numBlock, index, eventList, kernel, Flag_CPU_GPU, burden,
                     _CPU_GPU);
}

void CL_RadixSortOnly(qo_mem d_Rin, int rLen, int numThread, int numBlock, int _CPU_GPU) {
    radixSortImpl_sm(d_Rin, rLen, 32, numThread, numBlock, nullptr, nullptr, nullptr, nullptr, nullptr, _CPU_GPU);
}

void CL_BitonicSortOnly(qo_mem d_Rin, int rLen, qo_mem d_Rout, int numThread, int numBlock, int _CPU_GPU) {
    if (!d_Rin || !d_Rout || rLen < 0 || sizeof(Record) * (size_t)rLen > d_Rout->bytes) {
        set_error(QO_ERR_INVALID_ARG, "CL_BitonicSortOnly: bad arguments");
        compat_fail("CL_BitonicSortOnly");
        return;
    }
    if (rLen > 0) QO_COMPAT_CHECK(qo_memcpy_d2d(d_Rout->ptr, d_Rin->ptr, sizeof(Record) * (size_t)rLen, nullptr),
                                  "CL_BitonicSortOnly", );
    radixSortImpl_sm(d_Rin, rLen, 32, numThread, numBlock, nullptr, nullptr, nullptr, nullptr, nullptr, _CPU_GPU);
}

void CL_RadixSort(Record *h_Rin, int rLen, Record *h_Rout, int numThread, int numBlock, int _CPU_GPU) {
    if (rLen <= 0) return;
    qo_mem d = qo_mem_create(sizeof(Record) * (size_t)rLen);
    if (!d) return;
    if (qo_mem_write(d, h_Rin, sizeof(Record) * (size_t)rLen) == QO_OK) {
        radixSortImpl_sm(d, rLen, 32, numThread, numBlock, nullptr, nullptr, nullptr, nullptr, nullptr, _CPU_GPU);
        if (qo_mem_read(h_Rout, d, sizeof(Record) * (size_t)rLen) != QO_OK) compat_fail("CL_RadixSort");
    } else {
        compat_fail("CL_RadixSort");
    }
    qo_mem_free(d);
}

// ---- group-by ------------------------------------------------------------------------
int groupByImpl(qo_mem d_Rin, int rLen, qo_mem d_Rout, qo_mem *d_startPos, int, int, int *, void *, void *, int *,
                double *, int) {
    if (!d_Rin || !d_Rout || !d_startPos || rLen < 0) {
        set_error(QO_ERR_INVALID_ARG, "groupByImpl: bad arguments");
        compat_fail("groupByImpl");
        return -1;
    }
    if (ensure_init() != QO_OK) return -1;
    hipStream_t st = stream_of(nullptr);
    const size_t n = (size_t)rLen;
    if (d_Rin->ptr != d_Rout->ptr && n)
        QO_COMPAT_CHECK(qo_memcpy_d2d(d_Rout->ptr, d_Rin->ptr, n * sizeof(Record), nullptr), "groupByImpl", -1);
    QO_COMPAT_CHECK(qo_sort_records(static_cast<qo_record *>(d_Rout->ptr), n, nullptr), "groupByImpl", -1);
    int32_t *startPos = nullptr;
    QO_COMPAT_CHECK(qo_malloc(reinterpret_cast<void **>(&startPos), (n ? n : 1) * 4), "groupByImpl", -1);
    int status = QO_OK;
    void *ws = workspace(st, group_startpos_ws_bytes(n) + 256, &status);
    if (!ws) {
        qo_free(startPos);
        compat_fail("groupByImpl");
        return -1;
    }
    Carver cv(ws);
    uint64_t *d_num = cv.take<uint64_t>(1);
    void *sws = cv.take<char>(group_startpos_ws_bytes(n));
    uint64_t num = 0;
    int rc = launch_group_startpos(static_cast<const qo_record *>(d_Rout->ptr), n, startPos, d_num, sws, st);
    if (rc == QO_OK && hipMemcpyAsync(&num, d_num, 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = set_error(QO_ERR_HIP, "groupByImpl readback");
    if (rc == QO_OK) rc = sync_and_check(st);
    if (rc != QO_OK) {
        qo_free(startPos);
        compat_fail("groupByImpl");
        return -1;
    }
    *d_startPos = new qo_mem_s{startPos, (n ? n : 1) * 4, true};
    return (int)num;
}

int CL_GroupBy(Record *h_Rin, int rLen, Record *h_Rout, int **h_startPos, int numThread, int numBlock, int _CPU_GPU) {
    if (rLen < 0 || !h_startPos || (rLen > 0 && (!h_Rin || !h_Rout))) {
        set_error(QO_ERR_INVALID_ARG, "CL_GroupBy: bad arguments");
        compat_fail("CL_GroupBy");
        return -1;
    }
    const size_t bytes = sizeof(Record) * (size_t)(rLen > 0 ? rLen : 1);
    qo_mem d_Rin = qo_mem_create(bytes), d_Rout = qo_mem_create(bytes);
    qo_mem d_startPos = nullptr;
    int numGroup = -1;
    if (d_Rin && d_Rout && qo_mem_write(d_Rin, h_Rin, sizeof(Record) * (size_t)rLen) == QO_OK) {
        numGroup = groupByImpl(d_Rin, rLen, d_Rout, &d_startPos, numThread, numBlock, nullptr, nullptr, nullptr,
                               nullptr, nullptr, _CPU_GPU);
        if (numGroup >= 0) {
            *h_startPos = static_cast<int *>(malloc(sizeof(int) * (size_t)(numGroup > 0 ? numGroup : 1)));
            if (qo_mem_read(*h_startPos, d_startPos, sizeof(int) * (size_t)numGroup) != QO_OK ||
                qo_mem_read(h_Rout, d_Rout, sizeof(Record) * (size_t)rLen) != QO_OK) {
                compat_fail("CL_GroupBy");
                numGroup = -1;
            }
        }
    }
    qo_mem_free(d_Rin);
    qo_mem_free(d_Rout);
    qo_mem_free(d_startPos);
    return numGroup;
}

// ---- aggregation -----------------------------------------------------------------------
static int agg_range(const qo_record *s, size_t rLen, const int32_t *startPos, size_t first, size_t count,
                     size_t numGroups, int op, int32_t *out) {
    if (count == 0) return QO_OK;
    QO_TRY(ensure_init());
    hipStream_t st = stream_of(nullptr);
    // End of the last group in this range: rLen for the final group, else the next start.
    size_t last_end = rLen;
    if (first + count < numGroups) {
        int32_t nxt = 0;
        QO_HIP(hipMemcpy(&nxt, startPos + first + count, 4, hipMemcpyDeviceToHost));
        last_end = (size_t)(nxt > 0 ? nxt : 0);
    }
    int status = QO_OK;
    void *ws = workspace(st, aggregate_ws_bytes(count, (size_t)rLen), &status);
    if (!ws) return status;
    QO_TRY(launch_aggregate_groups(s, last_end, startPos + first, count, op, out + first, ws, st));
    QO_HIP(hipStreamSynchronize(st));
    return QO_OK;
}

void parallelAggregate_init(qo_mem d_S, qo_mem d_startPos, qo_mem d_aggResults, int OPERATOR, int blockOffset,
                            int numGroups, int numBlock, int, int, int rLen, int *, void *, void *, int *, double *,
                            int) {
    if (!d_S || !d_startPos || !d_aggResults || blockOffset < 0 || numBlock < 0 || blockOffset + numBlock > numGroups) {
        set_error(QO_ERR_INVALID_ARG, "parallelAggregate_init: bad arguments");
        compat_fail("parallelAggregate_init");
        return;
    }
    QO_COMPAT_CHECK(agg_range(static_cast<const qo_record *>(d_S->ptr), (size_t)rLen,
                              static_cast<const int32_t *>(d_startPos->ptr), (size_t)blockOffset, (size_t)numBlock,
                              (size_t)numGroups, OPERATOR, static_cast<int32_t *>(d_aggResults->ptr)),
                    "parallelAggregate_init", );
}

void aggAfterGroupByImpl(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg, qo_mem d_aggResults,
                         int OPERATOR, int, int *, void *, void *, int *, double *, int) {
    if (!d_Rin || !d_startPos || !d_aggResults || rLen < 0 || numGroups < 0) {
        set_error(QO_ERR_INVALID_ARG, "aggAfterGroupByImpl: bad arguments");
        compat_fail("aggAfterGroupByImpl");
        return;
    }
    const qo_record *src = static_cast<const qo_record *>(d_Rin->ptr);
    void *gathered = nullptr;
    if (d_Ragg && rLen > 0) {
        QO_COMPAT_CHECK(qo_malloc(&gathered, sizeof(Record) * (size_t)rLen + 256), "aggAfterGroupByImpl", );
        hipStream_t st = stream_of(nullptr);
        uint32_t *bad = reinterpret_cast<uint32_t *>(static_cast<char *>(gathered) + sizeof(Record) * (size_t)rLen);
        int rc = launch_fill_u32(bad, 0, 1, st);
        if (rc == QO_OK) {
            unsigned g = grid_for((size_t)rLen, 256, (unsigned)props().cu_count * 8);
            k_gather_records<<<g, 256, 0, st>>>(src, (size_t)rLen, static_cast<const qo_record *>(d_Ragg->ptr),
                                                d_Ragg->bytes / sizeof(Record), static_cast<qo_record *>(gathered),
                                                bad);
            uint32_t hbad = 0;
            if (hipGetLastError() != hipSuccess || hipMemcpy(&hbad, bad, 4, hipMemcpyDeviceToHost) != hipSuccess)
                rc = set_error(QO_ERR_HIP, "gather failed");
            else if (hbad)
                rc = set_error(QO_ERR_INVALID_ARG, "gather location outside Ragg");
        }
        if (rc != QO_OK) {
            qo_free(gathered);
            compat_fail("aggAfterGroupByImpl");
            return;
        }
        src = static_cast<const qo_record *>(gathered);
    }
    int rc = agg_range(src, (size_t)rLen, static_cast<const int32_t *>(d_startPos->ptr), 0, (size_t)numGroups,
                       (size_t)numGroups, OPERATOR, static_cast<int32_t *>(d_aggResults->ptr));
    if (gathered) qo_free(gathered);
    if (rc != QO_OK) compat_fail("aggAfterGroupByImpl");
}

void agg_max_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU) {
    aggAfterGroupByImpl(d_Rin, rLen, d_startPos, numGroups, d_Ragg, d_aggResults, REDUCE_MAX, numThread, index,
                        eventList, kernel, Flag_CPU_GPU, burden, _CPU_GPU);
}
void agg_min_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU) {
    aggAfterGroupByImpl(d_Rin, rLen, d_startPos, numGroups, d_Ragg, d_aggResults, REDUCE_MIN, numThread, index,
                        eventList, kernel, Flag_CPU_GPU, burden, _CPU_GPU);
}
void agg_sum_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU) {
    aggAfterGroupByImpl(d_Rin, rLen, d_startPos, numGroups, d_Ragg, d_aggResults, REDUCE_SUM, numThread, index,
                        eventList, kernel, Flag_CPU_GPU, burden, _CPU_GPU);
}
void agg_avg_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU) {
    aggAfterGroupByImpl(d_Rin, rLen, d_startPos, numGroups, d_Ragg, d_aggResults, REDUCE_AVERAGE, numThread, index,
                        eventList, kernel, Flag_CPU_GPU, burden, _CPU_GPU);
}

static void cl_agg(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg, int *h_aggResults, int op) {
    if (numGroups <= 0) return;
    const size_t n = (size_t)(rLen > 0 ? rLen : 1);
    qo_mem d_Rin = qo_mem_create(sizeof(Record) * n);
    qo_mem d_sp = qo_mem_create(sizeof(int) * (size_t)numGroups);
    qo_mem d_res = qo_mem_create(sizeof(int) * (size_t)numGroups);
    qo_mem d_Ragg = h_Ragg ? qo_mem_create(sizeof(Record) * n) : nullptr;
    bool ok = d_Rin && d_sp && d_res && (!h_Ragg || d_Ragg);
    ok = ok && qo_mem_write(d_Rin, h_Rin, sizeof(Record) * (size_t)rLen) == QO_OK;
    ok = ok && qo_mem_write(d_sp, h_startPos, sizeof(int) * (size_t)numGroups) == QO_OK;
    if (ok && h_Ragg) ok = qo_mem_write(d_Ragg, h_Ragg, sizeof(Record) * (size_t)rLen) == QO_OK;
    if (ok) {
        aggAfterGroupByImpl(d_Rin, rLen, d_sp, numGroups, d_Ragg, d_res, op, 256, nullptr, nullptr, nullptr, nullptr,
                            nullptr, 0);
        if (qo_mem_read(h_aggResults, d_res, sizeof(int) * (size_t)numGroups) != QO_OK) compat_fail("CL_agg");
    } else {
        compat_fail("CL_agg");
    }
    qo_mem_free(d_Rin);
    qo_mem_free(d_sp);
    qo_mem_free(d_res);
    qo_mem_free(d_Ragg);
}

void CL_agg_max_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg, int *h_aggResults,
                             int, int) {
    cl_agg(h_Rin, rLen, h_startPos, numGroups, h_Ragg, h_aggResults, REDUCE_MAX);
}
void CL_agg_min_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg, int *h_aggResults,
                             int, int) {
    cl_agg(h_Rin, rLen, h_startPos, numGroups, h_Ragg, h_aggResults, REDUCE_MIN);
}
void CL_agg_sum_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg, int *h_aggResults,
                             int, int) {
    cl_agg(h_Rin, rLen, h_startPos, numGroups, h_Ragg, h_aggResults, REDUCE_SUM);
}
void CL_agg_avg_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg, int *h_aggResults,
                             int, int) {
    cl_agg(h_Rin, rLen, h_startPos, numGroups, h_Ragg, h_aggResults, REDUCE_AVERAGE);
}

// ---- hash join --------------------------------------------------------------------------
int buildHashTable(uint32_t *rTableOnDevice, uint32_t *rHashTable, uint32_t size, uint32_t rHashTableBucketNum,
                   uint32_t hashBucketSize, uint32_t *rHashCount) {
    int rc = qo_bucket_build(rTableOnDevice, size, rHashTable, rHashTableBucketNum, hashBucketSize, rHashCount,
                             nullptr);
    if (rc != QO_OK) {
        compat_fail("buildHashTable");
        return -1;
    }
    return 0;
}

int probeHashTable(uint32_t *rHashTable, uint32_t *sTableOnDevice, uint32_t *matchedTable, uint32_t size,
                   uint32_t rHashTableBucketNum, uint32_t hashBucketSize, uint32_t *rHashCount) {
    uint64_t *d_count = nullptr;
    if (qo_malloc(reinterpret_cast<void **>(&d_count), 8) != QO_OK) {
        compat_fail("probeHashTable");
        return -1;
    }
    uint64_t cnt = 0;
    int rc = qo_bucket_probe(rHashTable, rHashCount, rHashTableBucketNum, hashBucketSize, sTableOnDevice, size,
                             matchedTable, (size_t)size, d_count, nullptr);
    if (rc == QO_OK) rc = qo_memcpy_d2h(&cnt, d_count, 8);
    if (rc == QO_OK) rc = check_device_error();
    qo_free(d_count);
    if (rc != QO_OK) {
        compat_fail("probeHashTable");
        return -1;
    }
    if (cnt > size) {
        set_error(QO_ERR_CAPACITY, "probe produced %llu pairs; matchedTable holds %u", (unsigned long long)cnt, size);
        compat_fail("probeHashTable");
        return -1;
    }
    return (int)cnt;
}

}  // extern "C"
