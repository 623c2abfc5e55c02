// Hash join: open-addressing build, probe with the S filter fused, plus the
// reference-shaped bucket-chained build/probe kept for drop-in callers.
//
// Reference path replaced (kernel/hj_filter_sort/1/hj_filter_sort_op.cl):
// buildHashTable (:25-60) is one 128-item work-group inserting under local
// spin-locks into fixed-capacity buckets; probeHashTable (:62-109) scans a
// bucket per S tuple and writes (s_val, r_val) pairs into per-work-item slots
// that are never compacted.
//
// Two strategies, picked from the table size:
//  * direct (tables up to one L2, or small inputs), here: a grid-wide CAS
//    insert, and a probe that streams S once, filters, looks up and appends
//    each tile's matches at a range reserved with one atomic add;
//  * partitioned (larger tables), join_part.hip: the inputs are split by key
//    hash so that every lookup and insert happens in an LDS copy of one table
//    slice.
// Output order is unspecified: rows appear in per-tile reservation order.
#include <algorithm>

#include "qo_internal.hpp"
#include "qo_join.hpp"

namespace qo {

size_t join_table_slots(size_t table_bytes) {
    if (table_bytes < sizeof(JoinHeader) + 8) return 0;
    return (table_bytes - sizeof(JoinHeader)) / 8;
}

__global__ void k_join_init(JoinHeader *hdr, uint64_t cap, uint64_t nr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        hdr->cap = cap;
        hdr->has_dups = 0;
        hdr->sentinel_count = 0;
        hdr->overflow = 0;
        hdr->nr = nr;
    }
}

__global__ __launch_bounds__(256) void k_join_build(const int32_t *__restrict__ rkey, const int32_t *__restrict__ rpay,
                                                    size_t nr, uint64_t *slots, JoinHeader *hdr, uint64_t cap) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nr; i += stride) join_insert(slots, cap, hdr, rkey[i], rpay[i]);
}

size_t join_build_ws_bytes(size_t table_bytes, size_t nr) { return join_part_ws_bytes(table_bytes, nr, true); }

int launch_join_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                      void *ws, hipStream_t st) {
    const size_t cap = join_table_slots(table_bytes);
    QO_REQUIRE(cap >= 2 && cap <= 0xFFFFFFFFull, "join table of %zu bytes has an unsupported size", table_bytes);
    QO_REQUIRE(nr < cap, "join table with %zu slots cannot hold %zu rows", cap, nr);
    JoinHeader *hdr = static_cast<JoinHeader *>(table);
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    k_join_init<<<1, 64, 0, st>>>(hdr, cap, nr);
    QO_LAUNCHED();
    // The partitioned build writes every slot itself.
    bool used = false;
    QO_TRY(launch_join_part_build(table, table_bytes, rkey, rpay, nr, ws, st, &used));
    if (used) return QO_OK;
    QO_TRY(launch_fill_u64(slots, kEmptySlot, cap, st));
    if (!nr) return QO_OK;
    unsigned g = grid_for(nr, 256, (unsigned)props().cu_count * 16);
    k_join_build<<<g, 256, 0, st>>>(rkey, rpay, nr, slots, hdr, cap);
    QO_LAUNCHED();
    return QO_OK;
}

// Probe.  One block per tile of kProbeTile S tuples; *d_count accumulates the
// matches (it is zeroed first, in stream order).
template <bool HAS_FILTER>
__global__ __launch_bounds__(256) void k_join_probe(const JoinHeader *__restrict__ hdr,
                                                    const uint64_t *__restrict__ slots,
                                                    const int32_t *__restrict__ skey, const int32_t *__restrict__ spay,
                                                    const int32_t *__restrict__ sfilt, size_t ns, int32_t flo,
                                                    int32_t fhi, bool vec, SoaEmit emit, uint64_t *d_count) {
    __shared__ uint32_t s_tot[kProbeJ * 4];
    __shared__ uint64_t s_base;
    probe_tile<HAS_FILTER>(hdr, slots, skey, spay, sfilt, ns, flo, fhi, vec, blockIdx.x,
                           reinterpret_cast<unsigned long long *>(d_count), emit, s_tot, &s_base);
}

size_t join_probe_ws_bytes(size_t table_bytes, size_t ns) {
    return std::max<size_t>(256, join_part_ws_bytes(table_bytes, ns, false));
}

int launch_join_probe(const void *table, size_t table_bytes, const int32_t *skey, const int32_t *spay,
                      const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, int32_t *okey, int32_t *orpay,
                      int32_t *ospay, size_t cap, uint64_t *d_count, void *ws, hipStream_t st) {
    QO_TRY(launch_fill_u64(d_count, 0, 1, st));
    if (ns == 0) return QO_OK;
    const size_t tiles = (ns + kProbeTile - 1) / kProbeTile;
    QO_REQUIRE(tiles < 0xFFFFFFFFull, "probe input too large");
    const JoinHeader *hdr = static_cast<const JoinHeader *>(table);
    const uint64_t *slots = reinterpret_cast<const uint64_t *>(hdr + 1);
    const bool vec = ((uintptr_t)skey % 16 == 0) && ((uintptr_t)spay % 16 == 0) && (!sfilt || (uintptr_t)sfilt % 16 == 0);
    SoaEmit emit{okey, orpay, ospay, cap};
    bool used = false;
    QO_TRY(launch_join_part_probe(table, table_bytes, skey, spay, sfilt, ns, flo, fhi, okey, orpay, ospay, cap,
                                  d_count, ws, st, &used));
    if (used) return QO_OK;
    if (sfilt)
        k_join_probe<true><<<(unsigned)tiles, 256, 0, st>>>(hdr, slots, skey, spay, sfilt, ns, flo, fhi, vec, emit,
                                                            d_count);
    else
        k_join_probe<false><<<(unsigned)tiles, 256, 0, st>>>(hdr, slots, skey, spay, sfilt, ns, flo, fhi, vec, emit,
                                                             d_count);
    QO_LAUNCHED();
    return QO_OK;
}

// ---------------------------------------------------------------------------
// Reference-shaped bucket-chained table.
// ---------------------------------------------------------------------------
__global__ void k_bucket_build(const uint32_t *__restrict__ rTable, uint32_t size, uint32_t *rHashTable, uint32_t B,
                               uint32_t bucketSize, uint32_t *rHashCount, uint32_t *overflow) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < size; i += stride) {
        uint2 kv = reinterpret_cast<const uint2 *>(rTable)[i];
        uint32_t h = kv.x & (B - 1);
        uint32_t c = atomicAdd(&rHashCount[h], 1u);
        if (c < bucketSize) {
            size_t at = ((size_t)h * bucketSize + c) * 2;
            rHashTable[at] = kv.x;
            rHashTable[at + 1] = kv.y;
        } else {
            atomicExch(overflow, 1u);
        }
    }
}

template <int J>
__global__ __launch_bounds__(256) void k_bucket_probe(const uint32_t *__restrict__ rHashTable,
                                                      const uint32_t *__restrict__ rHashCount, uint32_t B,
                                                      uint32_t bucketSize, const uint32_t *__restrict__ sTable,
                                                      uint32_t size, uint32_t *__restrict__ matched, size_t cap_pairs,
                                                      uint64_t *status, uint32_t *ticket, uint64_t *d_count,
                                                      uint32_t num_tiles) {
    constexpr int NW = 4;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tot[J * NW];
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base_i = (size_t)tile * (J * 256);
    uint2 kv[J];
    uint32_t c[J], ex[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        size_t i = base_i + (size_t)j * 256 + threadIdx.x;
        c[j] = 0;
        if (i < size) {
            kv[j] = reinterpret_cast<const uint2 *>(sTable)[i];
            uint32_t h = kv[j].x & (B - 1);
            uint32_t cnt = min(rHashCount[h], bucketSize);
            for (uint32_t q = 0; q < cnt; ++q) c[j] += rHashTable[((size_t)h * bucketSize + q) * 2] == kv[j].x;
        }
    }
    const uint32_t total = tile_offsets<J, NW>(c, ex, s_tot);
    if (threadIdx.x < kWave) {
        uint64_t e = lookback(status, tile, total);
        if (threadIdx.x == 0) s_excl = e;
    }
    __syncthreads();
    const uint64_t base = s_excl;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        if (!c[j]) continue;
        uint64_t pos = base + ex[j];
        uint32_t h = kv[j].x & (B - 1);
        uint32_t cnt = min(rHashCount[h], bucketSize);
        for (uint32_t q = 0; q < cnt; ++q) {
            size_t at = ((size_t)h * bucketSize + q) * 2;
            if (rHashTable[at] == kv[j].x) {
                if (pos < cap_pairs) {
                    matched[2 * pos] = kv[j].y;
                    matched[2 * pos + 1] = rHashTable[at + 1];
                }
                ++pos;
            }
        }
    }
    if (tile == num_tiles - 1 && threadIdx.x == 0) *d_count = base + total;
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_join_strategy(size_t table_bytes, size_t n, int *partitioned, int *bits, int *splits) {
    QO_REQUIRE(partitioned && bits && splits, "null argument");
    join_part_strategy(table_bytes, n, partitioned, bits, splits);
    return QO_OK;
}

size_t qo_join_table_bytes(size_t build_rows) {
    // Load factor (percent) is a tuned parameter of the probe; 50% by default.
    int load = tune_get(TUNE_PROBE, build_rows).param;
    if (load < 10 || load > 90) load = 50;
    size_t cap = (size_t)((double)build_rows * 100.0 / load) + 1;
    if (cap < 64) cap = 64;
    return sizeof(JoinHeader) + cap * 8;
}

int qo_hash_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                  qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && (nr == 0 || (rkey && rpay)), "null pointer");
    QO_REQUIRE((uintptr_t)table % 256 == 0, "join table must be 256-byte aligned");
    hipStream_t st = stream_of(s);
    void *ws = nullptr;
    if (const size_t wb = join_build_ws_bytes(table_bytes, nr)) {
        int status = QO_OK;
        ws = workspace(st, wb, &status);
        if (!ws) return status;
    }
    return launch_join_build(table, table_bytes, rkey, rpay, nr, ws, st);
}

int qo_hash_probe(const void *table, size_t table_bytes, const int32_t *skey, const int32_t *spay,
                  const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, int32_t *okey, int32_t *orpay,
                  int32_t *ospay, size_t out_capacity, uint64_t *d_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && d_count && (ns == 0 || (skey && spay)), "null pointer");
    QO_REQUIRE(out_capacity == 0 || (okey && orpay && ospay), "null output column");
    QO_REQUIRE(join_table_slots(table_bytes) >= 2, "join table of %zu bytes is too small", table_bytes);
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, join_probe_ws_bytes(table_bytes, ns), &status);
    if (!ws) return status;
    return launch_join_probe(table, table_bytes, skey, spay, sfilt, ns, flo, fhi, okey, orpay, ospay, out_capacity,
                             d_count, ws, st);
}

int qo_hj_filter_sort(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                      const int32_t *skey, const int32_t *spay, const int32_t *sfilt, size_t ns, int32_t flo,
                      int32_t fhi, int32_t *okey, int32_t *orpay, int32_t *ospay, size_t out_capacity,
                      uint64_t *h_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && h_count, "null pointer");
    hipStream_t st = stream_of(s);
    // One workspace serves the three stages in stream order: build buckets,
    // probe buckets, then the sort (after the count has been read back).
    const size_t stage_bytes = std::max({join_build_ws_bytes(table_bytes, nr), join_probe_ws_bytes(table_bytes, ns),
                                         sort_join_rows_ws_bytes(out_capacity)});
    int status = QO_OK;
    void *ws = workspace(st, stage_bytes + 512, &status);
    if (!ws) return status;
    Carver cv(ws);
    uint64_t *d_count = cv.take<uint64_t>(1);
    void *sws = cv.take<char>(stage_bytes);
    QO_TRY(launch_join_build(table, table_bytes, rkey, rpay, nr, sws, st));
    QO_TRY(launch_join_probe(table, table_bytes, skey, spay, sfilt, ns, flo, fhi, okey, orpay, ospay, out_capacity,
                             d_count, sws, st));
    uint64_t m = 0;
    QO_HIP(hipMemcpyAsync(&m, d_count, 8, hipMemcpyDeviceToHost, st));
    QO_HIP(hipStreamSynchronize(st));
    *h_count = m;
    if (m > out_capacity) return set_error(QO_ERR_CAPACITY, "join produced %llu rows, capacity %zu",
                                           (unsigned long long)m, out_capacity);
    QO_TRY(sort_join_rows_ws(okey, orpay, ospay, m, sws, st));
    QO_HIP(hipStreamSynchronize(st));
    return QO_OK;
}

int qo_bucket_build(const uint32_t *rTable, uint32_t size, uint32_t *rHashTable, uint32_t B, uint32_t bucketSize,
                    uint32_t *rHashCount, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(B >= 1 && (B & (B - 1)) == 0, "bucket count must be a power of two");
    QO_REQUIRE(bucketSize >= 1 && rHashTable && rHashCount && (size == 0 || rTable), "bad bucket-table arguments");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    uint32_t *flag = static_cast<uint32_t *>(workspace(st, 256, &status));
    if (!flag) return status;
    QO_TRY(launch_fill_u32(flag, 0, 1, st));
    QO_TRY(launch_fill_u32(rHashCount, 0, B, st));
    if (size) {
        unsigned g = grid_for(size, 256, (unsigned)props().cu_count * 16);
        k_bucket_build<<<g, 256, 0, st>>>(rTable, size, rHashTable, B, bucketSize, rHashCount, flag);
        QO_LAUNCHED();
    }
    uint32_t over = 0;
    QO_HIP(hipMemcpyAsync(&over, flag, 4, hipMemcpyDeviceToHost, st));
    QO_HIP(hipStreamSynchronize(st));
    if (over) return set_error(QO_ERR_CAPACITY, "a bucket received more than %u tuples", bucketSize);
    return QO_OK;
}

int qo_bucket_probe(const uint32_t *rHashTable, const uint32_t *rHashCount, uint32_t B, uint32_t bucketSize,
                    const uint32_t *sTable, uint32_t size, uint32_t *matchedTable, size_t capacity_pairs,
                    uint64_t *d_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(B >= 1 && (B & (B - 1)) == 0 && rHashTable && rHashCount && d_count, "bad bucket-probe arguments");
    hipStream_t st = stream_of(s);
    if (size == 0) return launch_fill_u64(d_count, 0, 1, st);
    QO_REQUIRE(sTable && (capacity_pairs == 0 || matchedTable), "null pointer");
    constexpr int J = 8;
    const size_t tiles = ((size_t)size + J * 256 - 1) / (J * 256);
    int status = QO_OK;
    void *ws = workspace(st, tiles * 8 + 512, &status);
    if (!ws) return status;
    Carver cv(ws);
    uint64_t *stat = cv.take<uint64_t>(tiles);
    uint32_t *ticket = cv.take<uint32_t>(1);
    QO_TRY(launch_fill_u32(ws, 0, cv.off / 4, st));
    k_bucket_probe<J><<<(unsigned)tiles, 256, 0, st>>>(rHashTable, rHashCount, B, bucketSize, sTable, size,
                                                       matchedTable, capacity_pairs, stat, ticket, d_count,
                                                       (uint32_t)tiles);
    QO_LAUNCHED();
    return QO_OK;
}

}  // extern "C"
