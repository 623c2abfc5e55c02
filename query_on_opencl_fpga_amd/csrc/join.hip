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
        hdr->mode = kJoinHashMix;
        hdr->nr = nr;
        hdr->slots_alloc = cap;
        hdr->kmin = 0;
        hdr->shift = 0;
        hdr->rmin = INT32_MAX;
        hdr->rmax = INT32_MIN;
        hdr->occupied = 0;
    }
}

// Slots [0, hdr->cap) := empty, when the table's mode is run_if.
__global__ __launch_bounds__(256) void k_fill_slots(JoinHeader *hdr, uint32_t run_if) {
    if (hdr->mode != run_if) return;
    const uint64_t cap = hdr->cap;
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    const uint4 e{~0u, ~0u, ~0u, ~0u};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap / 2; i += stride)
        reinterpret_cast<uint4 *>(slots)[i] = e;
    if ((cap & 1u) && blockIdx.x == 0 && threadIdx.x == 0) slots[cap - 1] = kEmptySlot;
}

// ---------------------------------------------------------------------------
// Key-range table (direct-mapped: home slot = key - kmin).
// ---------------------------------------------------------------------------
// Smallest and largest build key, one atomic pair per block.
__global__ __launch_bounds__(256) void k_key_range(const int32_t *__restrict__ rkey, size_t nr, JoinHeader *hdr) {
    __shared__ int32_t s_mn[4], s_mx[4];
    int32_t mn = INT32_MAX, mx = INT32_MIN;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t from = 0;
    if (((uintptr_t)rkey & 15u) == 0) {
        // Four 16-byte loads in flight per thread: one block per CU (its min/max
        // are two same-address atomics, which serialise across blocks).
        const size_t nv = nr / 4;
        const int4 *k4 = reinterpret_cast<const int4 *>(rkey);
        size_t i = t;
        for (; i + 3 * stride < nv; i += 4 * stride) {
            int4 k[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) k[u] = ld_stream(k4 + i + u * stride);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                mn = min(mn, min(min(k[u].x, k[u].y), min(k[u].z, k[u].w)));
                mx = max(mx, max(max(k[u].x, k[u].y), max(k[u].z, k[u].w)));
            }
        }
        for (; i < nv; i += stride) {
            const int4 k = ld_stream(k4 + i);
            mn = min(mn, min(min(k.x, k.y), min(k.z, k.w)));
            mx = max(mx, max(max(k.x, k.y), max(k.z, k.w)));
        }
        from = nv * 4;
    }
    for (size_t i = from + t; i < nr; i += stride) {
        mn = min(mn, rkey[i]);
        mx = max(mx, rkey[i]);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    if (lane_id() == 0) {
        s_mn[wave_id()] = mn;
        s_mx[wave_id()] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            mn = min(mn, s_mn[w]);
            mx = max(mx, s_mx[w]);
        }
        mn = min(mn, s_mn[0]);
        mx = max(mx, s_mx[0]);
        atomicMin(&hdr->rmin, mn);
        atomicMax(&hdr->rmax, mx);
    }
}

// Picks the key-range table when the key range fits 2^p <= the allocated slots
// (and max_bytes): cap = 2^p, h = (key - kmin) << (32 - p).
__global__ void k_range_decide(JoinHeader *hdr, uint64_t max_bytes) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || hdr->nr == 0) return;
    const int64_t range = (int64_t)hdr->rmax - (int64_t)hdr->rmin + 1;
    int p = 1;
    while (p < 31 && (int64_t(1) << p) < range) ++p;
    const uint64_t cap = uint64_t(1) << p;
    if ((int64_t)cap < range || cap > hdr->slots_alloc || cap * 8 > max_bytes) return;
    hdr->kmin = hdr->rmin;
    hdr->shift = 32u - (uint32_t)p;
    hdr->cap = cap;
    hdr->mode = kJoinKeyRange;
}

// Every build tuple to its home slot with a plain store (no probing): two
// tuples with one key leave fewer occupied slots than tuples, which the count
// below detects.
__global__ __launch_bounds__(256) void k_range_scatter(const int32_t *__restrict__ rkey,
                                                       const int32_t *__restrict__ rpay, size_t nr, JoinHeader *hdr) {
    if (hdr->mode != kJoinKeyRange) return;
    const uint32_t kmin = (uint32_t)hdr->kmin;
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t sentinel = 0;
    auto put = [&](int32_t k, int32_t v) {
        const uint64_t w = join_pack(k, v);
        if (w == kEmptySlot) ++sentinel;
        else slots[(uint32_t)k - kmin] = w;
    };
    size_t from = 0;
    if ((((uintptr_t)rkey | (uintptr_t)rpay) & 15u) == 0) {
        const size_t nv = nr / 4;
        for (size_t i = t; i < nv; i += stride) {
            const int4 k = ld_stream(reinterpret_cast<const int4 *>(rkey) + i);
            const int4 v = ld_stream(reinterpret_cast<const int4 *>(rpay) + i);
            put(k.x, v.x);
            put(k.y, v.y);
            put(k.z, v.z);
            put(k.w, v.w);
        }
        from = nv * 4;
    }
    for (size_t i = from + t; i < nr; i += stride) put(rkey[i], rpay[i]);
    sentinel = wave_sum(sentinel);
    if (lane_id() == 0 && sentinel) atomicAdd(&hdr->sentinel_count, sentinel);
}

// Occupied slots of a key-range table, one atomic per block.
__global__ __launch_bounds__(256) void k_range_count(JoinHeader *hdr) {
    __shared__ uint32_t s_c[4];
    if (hdr->mode != kJoinKeyRange) return;
    const uint64_t cap = hdr->cap;  // a power of two >= 2
    const uint4 *slots = reinterpret_cast<const uint4 *>(hdr + 1);
    uint32_t c = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap / 2; i += stride) {
        const uint4 q = ld_stream(slots + i);
        c += ((q.x & q.y) != ~0u) + ((q.z & q.w) != ~0u);
    }
    c = wave_sum(c);
    if (lane_id() == 0) s_c[wave_id()] = c;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&hdr->occupied, (unsigned long long)(s_c[0] + s_c[1] + s_c[2] + s_c[3]));
}

// A key-range build that lost tuples to duplicate keys falls back to the
// hash-mixed table (built next).
__global__ void k_range_commit(JoinHeader *hdr) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || hdr->mode != kJoinKeyRange) return;
    if (hdr->occupied + hdr->sentinel_count == hdr->nr) return;
    hdr->mode = kJoinHashMix;
    hdr->cap = hdr->slots_alloc;
    hdr->sentinel_count = 0;
    hdr->kmin = 0;
    hdr->shift = 0;
}

// ---------------------------------------------------------------------------
// Hash-mixed table, direct build (small tables).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_join_build(const int32_t *__restrict__ rkey, const int32_t *__restrict__ rpay,
                                                    size_t nr, uint64_t *slots, JoinHeader *hdr, uint64_t cap) {
    if (hdr->mode != kJoinHashMix) return;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nr; i += stride) join_insert(slots, cap, hdr, rkey[i], rpay[i]);
}

size_t join_build_ws_bytes(size_t table_bytes, size_t nr) { return join_part_ws_bytes(table_bytes, nr, true); }

// Build: the key-range table when the keys allow it (decided on the device),
// else the hash-mixed table, partitioned or direct.  Every step after the
// decision runs or returns at once according to the table's mode, so the
// sequence is asynchronous and the same for every input.
int launch_join_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                      void *ws, hipStream_t st) {
    const size_t cap = join_table_slots(table_bytes);
    QO_REQUIRE(cap >= 2 && cap <= 0xFFFFFFFFull, "join table of %zu bytes has an unsupported size", table_bytes);
    QO_REQUIRE(nr < cap, "join table with %zu slots cannot hold %zu rows", cap, nr);
    JoinHeader *hdr = static_cast<JoinHeader *>(table);
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    k_join_init<<<1, 64, 0, st>>>(hdr, cap, nr);
    QO_LAUNCHED();
    const unsigned cus = (unsigned)props().cu_count;
    const unsigned fill_grid = grid_for(cap / 2 + 1, 256, cus * 8);
    const TuneParams tr = tune_get(TUNE_JOIN_RANGE, nr);
    if (nr && tr.param != PART_NEVER) {
        k_key_range<<<cus, 256, 0, st>>>(rkey, nr, hdr);
        QO_LAUNCHED();
        k_range_decide<<<1, 64, 0, st>>>(hdr, (uint64_t)std::max(1, tr.items) << 20);
        QO_LAUNCHED();
        // Large tables: slices assembled in LDS and written whole.  Else every
        // tuple stored at its slot, which leaves partial lines to merge.
        bool sliced = false;
        QO_TRY(launch_range_part_build(table, table_bytes, rkey, rpay, nr, ws, st, &sliced));
        if (!sliced) {
            k_fill_slots<<<fill_grid, 256, 0, st>>>(hdr, kJoinKeyRange);
            QO_LAUNCHED();
            k_range_scatter<<<grid_for(nr / 4 + 1, 256, cus * 8), 256, 0, st>>>(rkey, rpay, nr, hdr);
            QO_LAUNCHED();
            k_range_count<<<cus * 2, 256, 0, st>>>(hdr);
            QO_LAUNCHED();
        }
        k_range_commit<<<1, 64, 0, st>>>(hdr);
        QO_LAUNCHED();
    }
    // Hash-mixed table (each kernel returns at once for a key-range table).
    bool used = false;
    QO_TRY(launch_join_part_build(table, table_bytes, rkey, rpay, nr, ws, st, &used));
    if (used) return QO_OK;
    k_fill_slots<<<fill_grid, 256, 0, st>>>(hdr, kJoinHashMix);
    QO_LAUNCHED();
    if (!nr) return QO_OK;
    unsigned g = grid_for(nr, 256, cus * 16);
    k_join_build<<<g, 256, 0, st>>>(rkey, rpay, nr, slots, hdr, cap);
    QO_LAUNCHED();
    return QO_OK;
}

// Probe.  One block per tile of kProbeTile S tuples; *d_count accumulates the
// matches (it is zeroed first, in stream order).  With a gate the kernel runs
// only for the table mode it names.
template <bool HAS_FILTER>
__global__ __launch_bounds__(256) void k_join_probe(const JoinHeader *__restrict__ hdr,
                                                    const uint64_t *__restrict__ slots,
                                                    const int32_t *__restrict__ skey, const int32_t *__restrict__ spay,
                                                    const int32_t *__restrict__ sfilt, size_t ns, int32_t flo,
                                                    int32_t fhi, bool vec, SoaEmit emit, uint64_t *d_count,
                                                    Gate gate) {
    __shared__ uint32_t s_tot[tile_offsets_words(kProbeJ, 4)];
    __shared__ uint64_t s_base;
    __shared__ JoinStage stage;
    if (gate.off()) return;
    probe_tile<HAS_FILTER>(hdr, slots, skey, spay, sfilt, ns, flo, fhi, vec, blockIdx.x,
                           reinterpret_cast<unsigned long long *>(d_count), emit, s_tot, &s_base, &stage);
}

size_t join_probe_ws_bytes(size_t table_bytes, size_t ns) {
    return std::max<size_t>(256, join_part_ws_bytes(table_bytes, ns, false));
}

// Probe: large tables through their partitioned probe (per-slice LDS for a
// hash-mixed table, split + L2-resident slices for a key-range table), small
// ones directly; the table's mode picks among the launched paths on the device.
int launch_join_probe(const void *table, size_t table_bytes, const int32_t *skey, const int32_t *spay,
                      const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, const JoinOut &out,
                      uint64_t *d_count, void *ws, hipStream_t st) {
    QO_TRY(launch_fill_u64(d_count, 0, 1, st));
    if (ns == 0) return QO_OK;
    const size_t tiles = (ns + kProbeTile - 1) / kProbeTile;
    QO_REQUIRE(tiles < 0xFFFFFFFFull, "probe input too large");
    const JoinHeader *hdr = static_cast<const JoinHeader *>(table);
    const uint64_t *slots = reinterpret_cast<const uint64_t *>(hdr + 1);
    bool hash_part = false, range_part = false;
    QO_TRY(launch_join_part_probe(table, table_bytes, skey, spay, sfilt, ns, flo, fhi, out, d_count, ws, st,
                                  &hash_part, &range_part));
    if (hash_part && range_part) return QO_OK;
    // The direct probe serves the mode(s) no partitioned probe was launched for.
    const Gate gate = hash_part ? Gate{&hdr->mode, kJoinKeyRange}
                                : range_part ? Gate{&hdr->mode, kJoinHashMix} : Gate{};
    const bool vec = ((uintptr_t)skey % 16 == 0) && ((uintptr_t)spay % 16 == 0) && (!sfilt || (uintptr_t)sfilt % 16 == 0);
    const SoaEmit emit{out.okey, out.orpay, out.ospay, out.cap, out.k64, out.v32};
    if (sfilt)
        k_join_probe<true><<<(unsigned)tiles, 256, 0, st>>>(hdr, slots, skey, spay, sfilt, ns, flo, fhi, vec, emit,
                                                            d_count, gate);
    else
        k_join_probe<false><<<(unsigned)tiles, 256, 0, st>>>(hdr, slots, skey, spay, sfilt, ns, flo, fhi, vec, emit,
                                                             d_count, gate);
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
                                                      uint32_t num_tiles, DevCtl *ctl) {
    constexpr int NW = 4;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tot[tile_offsets_words(J, NW)];
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= num_tiles) return;  // whole block; one ticket per launched block, so never taken
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
        uint64_t e = lookback(status, tile, total, ctl);
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

int qo_join_strategy(size_t table_bytes, size_t n, int mode, int *partitioned, int *bits, int *splits) {
    QO_REQUIRE(partitioned && bits && splits, "null argument");
    QO_REQUIRE(mode == 0 || mode == 1, "table mode must be 0 or 1");
    join_part_strategy(table_bytes, n, mode, partitioned, bits, splits);
    return QO_OK;
}

int qo_join_table_info(const void *table, int *mode, uint64_t *slots_used, int *has_dups, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && mode && slots_used && has_dups, "null argument");
    hipStream_t st = stream_of(s);
    JoinHeader h;
    QO_HIP(hipStreamSynchronize(st));
    QO_HIP(hipMemcpy(&h, table, sizeof(h), hipMemcpyDeviceToHost));
    *mode = (int)h.mode;
    *slots_used = h.cap;
    *has_dups = (h.has_dups | h.sentinel_count) ? 1 : 0;
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
    JoinOut out;
    out.okey = okey;
    out.orpay = orpay;
    out.ospay = ospay;
    out.cap = out_capacity;
    return launch_join_probe(table, table_bytes, skey, spay, sfilt, ns, flo, fhi, out, d_count, ws, st);
}

int qo_hj_filter_sort(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                      const int32_t *skey, const int32_t *spay, const int32_t *sfilt, size_t ns, int32_t flo,
                      int32_t fhi, int32_t *okey, int32_t *orpay, int32_t *ospay, size_t out_capacity,
                      uint64_t *h_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && h_count, "null pointer");
    hipStream_t st = stream_of(s);
    // The probe writes sort-ready rows (join_sort_key(key, s_payload), r_payload)
    // that the sort orders in place and unpacks into the output columns, so no
    // separate pack pass runs.  Build buckets, probe buckets and the sort's
    // workspace share one region in stream order.
    const size_t stage_bytes = std::max({join_build_ws_bytes(table_bytes, nr), join_probe_ws_bytes(table_bytes, ns),
                                         radix_workspace_bytes(out_capacity, true)});
    const size_t rows = out_capacity ? out_capacity : 1;
    int status = QO_OK;
    void *ws = workspace(st, stage_bytes + rows * 12 + 2048, &status);
    if (!ws) return status;
    Carver cv(ws);
    uint64_t *d_count = cv.take<uint64_t>(1);
    uint32_t *d_flag = cv.take<uint32_t>(2);
    uint64_t *k64 = cv.take<uint64_t>(rows);
    uint32_t *v32 = cv.take<uint32_t>(rows);
    void *sws = cv.take<char>(stage_bytes);
    QO_TRY(launch_join_build(table, table_bytes, rkey, rpay, nr, sws, st));
    JoinOut out;
    out.cap = out_capacity;
    out.k64 = k64;
    out.v32 = v32;
    QO_TRY(launch_join_probe(table, table_bytes, skey, spay, sfilt, ns, flo, fhi, out, d_count, sws, st));
    uint64_t m = 0;
    int32_t range[2] = {INT32_MAX, INT32_MIN};  // build key range, when the build measured it
    JoinHeader *hdr = static_cast<JoinHeader *>(table);
    QO_HIP(hipMemcpyAsync(&m, d_count, 8, hipMemcpyDeviceToHost, st));
    QO_HIP(hipMemcpyAsync(range, &hdr->rmin, sizeof(range), hipMemcpyDeviceToHost, st));
    QO_HIP(hipStreamSynchronize(st));
    *h_count = m;
    if (m > out_capacity) return set_error(QO_ERR_CAPACITY, "join produced %llu rows, capacity %zu",
                                           (unsigned long long)m, out_capacity);
    // Every output key is a build key, so key bits that all build keys share
    // need no radix pass: with keys in [rmin, rmax] only the bits below the
    // highest bit where (rmin, rmax) differ vary (as sort keys, key ^ 2^31).
    int hi_bit = 64;
    if (range[0] <= range[1]) {
        const uint32_t d = ((uint32_t)range[0] ^ 0x80000000u) ^ ((uint32_t)range[1] ^ 0x80000000u);
        hi_bit = 32 + (d ? 32 - __builtin_clz(d) : 0);
    }
    QO_TRY(sort_packed_join_rows(k64, v32, m, sws, okey, orpay, ospay, st, hi_bit, d_flag));
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
                                                       (uint32_t)tiles, dev_ctl());
    QO_LAUNCHED();
    return QO_OK;
}

}  // extern "C"
