// Hash join: open-addressing build, probe with the S filter fused and a stable
// (S-order) compaction of matches, plus the reference-shaped bucket-chained
// build/probe kept for drop-in callers.
//
// Reference path replaced (kernel/hj_filter_sort/1/hj_filter_sort_op.cl):
// buildHashTable (:25-60) is one 128-item work-group inserting under local
// spin-locks into fixed-capacity buckets; probeHashTable (:62-109) scans a
// bucket per S tuple and writes (s_val, r_val) pairs into per-work-item slots
// that are never compacted.  Here the build is a grid-wide CAS insert into a
// bandwidth-sized power-of-two table and the probe streams S once, filters,
// looks up, and appends matches in S order through a decoupled look-back.
#include <algorithm>

#include "qo_internal.hpp"
#include "qo_join.hpp"

namespace qo {

size_t join_table_capacity_log2(size_t table_bytes) {
    if (table_bytes < sizeof(JoinHeader) + 8) return 0;
    size_t slots = (table_bytes - sizeof(JoinHeader)) / 8;
    size_t lg = 0;
    while ((size_t(1) << (lg + 1)) <= slots) ++lg;
    return lg;
}

__global__ void k_join_init(JoinHeader *hdr, uint32_t log2cap, uint64_t nr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        hdr->log2cap = log2cap;
        hdr->has_dups = 0;
        hdr->sentinel_count = 0;
        hdr->overflow = 0;
        hdr->nr = nr;
    }
}

__global__ __launch_bounds__(256) void k_join_build(const int32_t *__restrict__ rkey, const int32_t *__restrict__ rpay,
                                                    size_t nr, uint64_t *slots, JoinHeader *hdr, uint64_t mask) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nr; i += stride) join_insert(slots, mask, hdr, rkey[i], rpay[i]);
}

int launch_join_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                      hipStream_t st) {
    const size_t lg = join_table_capacity_log2(table_bytes);
    QO_REQUIRE(lg >= 1 && lg <= 40, "join table of %zu bytes is too small", table_bytes);
    const size_t cap = size_t(1) << lg;
    QO_REQUIRE(nr < cap, "join table with %zu slots cannot hold %zu rows", cap, nr);
    JoinHeader *hdr = static_cast<JoinHeader *>(table);
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    k_join_init<<<1, 64, 0, st>>>(hdr, (uint32_t)lg, nr);
    QO_LAUNCHED();
    QO_TRY(launch_fill_u64(slots, kEmptySlot, cap, st));
    if (nr) {
        unsigned g = grid_for(nr, 256, (unsigned)props().cu_count * 16);
        k_join_build<<<g, 256, 0, st>>>(rkey, rpay, nr, slots, hdr, cap - 1);
        QO_LAUNCHED();
    }
    return QO_OK;
}

// Probe.  Tile = 256 threads x J int4 groups (4 S tuples each), striped.  The
// common case (unique build keys) resolves each tuple with one chain walk and
// keeps the payload in registers; tables with duplicate keys take the general
// path that counts, then re-walks to emit every match.
template <int J, bool HAS_FILTER, bool VEC>
__global__ __launch_bounds__(256) void k_join_probe(const JoinHeader *__restrict__ hdr,
                                                    const uint64_t *__restrict__ slots,
                                                    const int32_t *__restrict__ skey, const int32_t *__restrict__ spay,
                                                    const int32_t *__restrict__ sfilt, size_t ns, int32_t flo,
                                                    int32_t fhi, int32_t *__restrict__ okey,
                                                    int32_t *__restrict__ orpay, int32_t *__restrict__ ospay,
                                                    size_t cap, uint64_t *status, uint32_t *ticket,
                                                    uint64_t *d_count, uint32_t num_tiles) {
    constexpr int NW = 4;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tot[J * NW];
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    const uint64_t mask = (uint64_t(1) << hdr->log2cap) - 1;
    const bool general = (hdr->has_dups | hdr->sentinel_count) != 0;
    const uint32_t sentinel = hdr->sentinel_count;
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t vbase = (size_t)tile * (J * 256);

    int32_t key[J][4], sp[J][4], rp[J][4];
    uint32_t pass[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const size_t v = vbase + (size_t)j * 256 + threadIdx.x;
        const size_t e0 = v * 4;
        pass[j] = 0;
        if (VEC && e0 + 4 <= ns) {
            int4 kk = ld_stream(reinterpret_cast<const int4 *>(skey) + v);
            int4 pp = ld_stream(reinterpret_cast<const int4 *>(spay) + v);
            key[j][0] = kk.x; key[j][1] = kk.y; key[j][2] = kk.z; key[j][3] = kk.w;
            sp[j][0] = pp.x; sp[j][1] = pp.y; sp[j][2] = pp.z; sp[j][3] = pp.w;
            if (HAS_FILTER) {
                int4 ff = ld_stream(reinterpret_cast<const int4 *>(sfilt) + v);
                pass[j] = ((ff.x >= flo) & (ff.x <= fhi)) | (((ff.y >= flo) & (ff.y <= fhi)) << 1) |
                          (((ff.z >= flo) & (ff.z <= fhi)) << 2) | (((ff.w >= flo) & (ff.w <= fhi)) << 3);
            } else {
                pass[j] = 15u;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const size_t e = e0 + k;
                if (e < ns) {
                    key[j][k] = skey[e];
                    sp[j][k] = spay[e];
                    bool ok = true;
                    if (HAS_FILTER) {
                        int32_t f = sfilt[e];
                        ok = (f >= flo) & (f <= fhi);
                    }
                    pass[j] |= ok ? (1u << k) : 0u;
                }
            }
        }
    }

    uint32_t c[J], ex[J];
    uint32_t hit[J];
    if (!general) {
        // Issue every first-slot load before resolving any chain.
        uint64_t hs[J][4], sv[J][4];
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                hs[j][k] = hash32((uint32_t)key[j][k]) & mask;
                sv[j][k] = (pass[j] >> k) & 1u ? slots[hs[j][k]] : kEmptySlot;
            }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if ((pass[j] >> k) & 1u)
                    if (join_find(slots, mask, key[j][k], hs[j][k], sv[j][k], rp[j][k])) hit[j] |= 1u << k;
            c[j] = __popc(hit[j]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            c[j] = 0;
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if ((pass[j] >> k) & 1u) {
                    uint32_t m = join_count(slots, mask, sentinel, key[j][k]);
                    c[j] += m;
                    if (m) hit[j] |= 1u << k;
                }
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
        uint64_t pos = base + ex[j];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!((hit[j] >> k) & 1u)) continue;
            if (!general) {
                if (pos < cap) {
                    okey[pos] = key[j][k];
                    orpay[pos] = rp[j][k];
                    ospay[pos] = sp[j][k];
                }
                ++pos;
            } else {
                const int32_t kk = key[j][k];
                uint64_t h = hash32((uint32_t)kk) & mask;
                for (uint64_t probe = 0; probe <= mask; ++probe) {
                    uint64_t s = slots[h];
                    if (s == kEmptySlot) break;
                    if ((uint32_t)s == (uint32_t)kk) {
                        if (pos < cap) {
                            okey[pos] = kk;
                            orpay[pos] = (int32_t)(s >> 32);
                            ospay[pos] = sp[j][k];
                        }
                        ++pos;
                    }
                    h = (h + 1) & mask;
                }
                if (kk == -1)
                    for (uint32_t q = 0; q < sentinel; ++q) {
                        if (pos < cap) {
                            okey[pos] = kk;
                            orpay[pos] = -1;
                            ospay[pos] = sp[j][k];
                        }
                        ++pos;
                    }
            }
        }
    }
    if (tile == num_tiles - 1 && threadIdx.x == 0) *d_count = base + total;
}

constexpr int kProbeJ = 4;
constexpr size_t kProbeTile = (size_t)kProbeJ * 256 * 4;

size_t join_probe_ws_bytes(size_t ns) { return ((ns + kProbeTile - 1) / kProbeTile + 1) * 8 + 512; }

int launch_join_probe(const void *table, const int32_t *skey, const int32_t *spay, const int32_t *sfilt, size_t ns,
                      int32_t flo, int32_t fhi, int32_t *okey, int32_t *orpay, int32_t *ospay, size_t cap,
                      uint64_t *d_count, void *ws, hipStream_t st) {
    if (ns == 0) return launch_fill_u64(d_count, 0, 1, st);
    const size_t tiles = (ns + kProbeTile - 1) / kProbeTile;
    QO_REQUIRE(tiles < 0xFFFFFFFFull, "probe input too large");
    const JoinHeader *hdr = static_cast<const JoinHeader *>(table);
    const uint64_t *slots = reinterpret_cast<const uint64_t *>(hdr + 1);
    Carver cv(ws);
    uint64_t *stat = cv.take<uint64_t>(tiles);
    uint32_t *ticket = cv.take<uint32_t>(1);
    QO_TRY(launch_fill_u32(ws, 0, cv.off / 4, st));
    const bool vec = ((uintptr_t)skey % 16 == 0) && ((uintptr_t)spay % 16 == 0) && (!sfilt || (uintptr_t)sfilt % 16 == 0);
#define QO_PROBE(F, V)                                                                                              \
    k_join_probe<kProbeJ, F, V><<<(unsigned)tiles, 256, 0, st>>>(hdr, slots, skey, spay, sfilt, ns, flo, fhi, okey, \
                                                                 orpay, ospay, cap, stat, ticket, d_count,          \
                                                                 (uint32_t)tiles)
    if (sfilt) {
        if (vec) QO_PROBE(true, true);
        else QO_PROBE(true, false);
    } else {
        if (vec) QO_PROBE(false, true);
        else QO_PROBE(false, false);
    }
#undef QO_PROBE
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

size_t qo_join_table_bytes(size_t build_rows) {
    size_t cap = 64;
    while (cap < 2 * build_rows + 1) cap <<= 1;
    return sizeof(JoinHeader) + cap * 8;
}

int qo_hash_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                  qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && (nr == 0 || (rkey && rpay)), "null pointer");
    QO_REQUIRE((uintptr_t)table % 256 == 0, "join table must be 256-byte aligned");
    return launch_join_build(table, table_bytes, rkey, rpay, nr, stream_of(s));
}

int qo_hash_probe(const void *table, const int32_t *skey, const int32_t *spay, const int32_t *sfilt, size_t ns,
                  int32_t flo, int32_t fhi, int32_t *okey, int32_t *orpay, int32_t *ospay, size_t out_capacity,
                  uint64_t *d_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && d_count && (ns == 0 || (skey && spay)), "null pointer");
    QO_REQUIRE(out_capacity == 0 || (okey && orpay && ospay), "null output column");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, join_probe_ws_bytes(ns), &status);
    if (!ws) return status;
    return launch_join_probe(table, skey, spay, sfilt, ns, flo, fhi, okey, orpay, ospay, out_capacity, d_count, ws,
                             st);
}

int qo_hj_filter_sort(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                      const int32_t *skey, const int32_t *spay, const int32_t *sfilt, size_t ns, int32_t flo,
                      int32_t fhi, int32_t *okey, int32_t *orpay, int32_t *ospay, size_t out_capacity,
                      uint64_t *h_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(table && h_count, "null pointer");
    hipStream_t st = stream_of(s);
    QO_TRY(launch_join_build(table, table_bytes, rkey, rpay, nr, st));
    int status = QO_OK;
    size_t ws_bytes = std::max(join_probe_ws_bytes(ns) + 256, sort_join_rows_ws_bytes(out_capacity) + 256);
    void *ws = workspace(st, ws_bytes, &status);
    if (!ws) return status;
    Carver cv(ws);
    uint64_t *d_count = cv.take<uint64_t>(1);
    void *pws = cv.take<char>(join_probe_ws_bytes(ns));
    QO_TRY(launch_join_probe(table, skey, spay, sfilt, ns, flo, fhi, okey, orpay, ospay, out_capacity, d_count, pws,
                             st));
    uint64_t m = 0;
    QO_HIP(hipMemcpyAsync(&m, d_count, 8, hipMemcpyDeviceToHost, st));
    QO_HIP(hipStreamSynchronize(st));
    *h_count = m;
    if (m > out_capacity) return set_error(QO_ERR_CAPACITY, "join produced %llu rows, capacity %zu",
                                           (unsigned long long)m, out_capacity);
    // The sort workspace reuses the whole buffer: the probe is finished.
    QO_TRY(sort_join_rows_ws(okey, orpay, ospay, m, ws, st));
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
