// Group-by aggregation.
//
// 1. Hash aggregation (qo_groupby_agg): every block keeps an LDS-resident
//    partial table (key, COUNT, SUM, MIN, MAX) with linear probing; tuples whose
//    key does not find an LDS slot within a few probes update a global
//    open-addressing table directly.  At the end each block merges its LDS table
//    into the global one (or, when it saw every tuple of its keys, appends its
//    groups as finished rows); the rows are sorted by key and gathered into the
//    SoA outputs.
// 2. The reference's sort-based pipeline for drop-in callers: group start
//    positions over Records sorted by .y (scanGroupLabel_kernel +
//    groupByImpl_write_kernel, kernel/aggregation/1/group_by_cu_op_dp.cl:293-327,
//    505-528) and the per-group reduction with the reference's operator
//    semantics (parallelAggregate_kernel, :547-716).
#include <algorithm>
#include <type_traits>
#include <climits>

#include "qo_device.hpp"
#include "qo_internal.hpp"
#include "qo_split.hpp"

namespace qo {

constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;

struct alignas(32) GEntry {
    uint32_t key;
    int32_t mn;
    int32_t mx;
    uint32_t pad;
    unsigned long long cnt;
    unsigned long long sum;
};
static_assert(sizeof(GEntry) == 32, "entry size");

__global__ void k_gtable_init(GEntry *t, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        GEntry e;
        e.key = kEmptyKey;
        e.mn = INT_MAX;
        e.mx = INT_MIN;
        e.pad = 0;
        e.cnt = 0;
        e.sum = 0;
        t[i] = e;
    }
}

// Home slot of `key` in a table of 2^bits slots: multiplicative (Fibonacci)
// hashing, the top bits of key * 2^32/phi.  Consecutive keys land almost one
// per slot (group keys are often a dense range), and other key sets spread
// like under any mixing hash.
QO_DEV uint32_t fib_slot(uint32_t key, int bits) { return (key * 0x9E3779B9u) >> (32 - bits); }

// Entry for `key` in the global table (index mask+1 is reserved for key -1,
// whose bit pattern is the empty marker).  nullptr when the table is full.
QO_DEV GEntry *g_entry(GEntry *t, uint64_t mask, uint32_t key) {
    if (key == kEmptyKey) return &t[mask + 1];
    uint64_t h = fib_slot(key, __popcll(mask));
    {
        const uint32_t k = __hip_atomic_load(&t[h].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) return &t[h];
    }
    for (uint64_t p = 0; p <= mask; ++p) {
        uint32_t k = __hip_atomic_load(&t[h].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) return &t[h];
        if (k == kEmptyKey) {
            uint32_t prev = atomicCAS(&t[h].key, kEmptyKey, key);
            if (prev == kEmptyKey || prev == key) return &t[h];
        }
        h = (h + 1) & mask;
    }
    return nullptr;
}

QO_DEV void g_update(GEntry *t, uint64_t mask, uint32_t key, uint64_t cnt, uint64_t sum, int32_t mn, int32_t mx,
                     uint32_t *overflow) {
    GEntry *e = g_entry(t, mask, key);
    if (!e) {
        atomicExch(overflow, 1u);
        return;
    }
    atomicAdd(&e->cnt, (unsigned long long)cnt);
    atomicAdd(&e->sum, (unsigned long long)sum);
    atomicMin(&e->mn, mn);
    atomicMax(&e->mx, mx);
}

// LDS slot for `key` within `probes` probes, or -1.  The home slot is the
// next log2(C) bits of the multiplicative hash after the `skip` top bits (a
// split by those bits put the tuple in this table).  The home slot is tested
// first on its own, or with TWO the home slot and the one after it, read
// together: in a table at most half full nearly every key sits in one of them
// (2^20 dense keys over 1024 tables of 2048 slots: 73 % at home, 27 % one slot
// on, 0.02 % further), so the wave enters the probe loop only when one of its
// lanes is further out.  `lkey` holds C + 2 words; the two past the end stay
// kEmptyKey, which no looked-up key equals, so the second read needs no wrap.
// TWO costs LDS bank cycles where keys do sit at home (a few groups).
template <int C, bool TWO = false>
QO_DEV int l_slot(uint32_t *lkey, uint32_t key, int probes, int skip) {
    static_assert((C & (C - 1)) == 0, "power-of-two table");
    constexpr int kBits = __builtin_ctz(C);
    uint32_t h = ((key * 0x9E3779B9u) << skip) >> (32 - kBits);
    if (TWO) {
        typedef uint32_t Pair __attribute__((ext_vector_type(2), aligned(4)));
        const Pair kk = *reinterpret_cast<const Pair *>(lkey + h);  // one two-word LDS read
        if (kk.x == key) return (int)h;
        if (kk.y == key) return (int)(h + 1);
    } else if (lkey[h] == key) {
        return (int)h;
    }
    for (int p = 0; p < probes; ++p) {
        uint32_t k = lkey[h];
        if (k == key) return (int)h;
        if (k == kEmptyKey) {
            uint32_t prev = atomicCAS(&lkey[h], kEmptyKey, key);
            if (prev == kEmptyKey || prev == key) return (int)h;
        }
        h = (h + 1) & (C - 1);
    }
    return -1;
}

// Block-local aggregation table; key[C] and key[C + 1] are padding (always
// kEmptyKey) so that l_slot can read two slots from any home slot.
template <int C>
struct LTable {
    uint32_t key[C + 2];
    uint32_t cnt[C];
    int32_t mn[C];
    int32_t mx[C];
    unsigned long long sum[C];
};

template <int C>
QO_DEV void ltable_init(LTable<C> &L) {
    if (threadIdx.x < 2) L.key[C + threadIdx.x] = kEmptyKey;
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        L.key[i] = kEmptyKey;
        L.cnt[i] = 0;
        L.mn[i] = INT_MAX;
        L.mx[i] = INT_MIN;
        L.sum[i] = 0;
    }
}

// One tuple into the block's LDS table, or into the global table when its key
// finds no LDS slot within `probes` probes (or is the empty-marker pattern).
//
// COMBINE (for inputs where one key can fill whole waves, such as the buckets
// of a heavy hitter): when every lane of the wave is active and at least a
// quarter of the lanes hold the key of the first lane, those lanes combine
// (COUNT, SUM, MIN, MAX) across the wave first and one of them updates the
// table, instead of all of them updating one LDS address one after another.
template <int C, bool COMBINE = false, bool TWO = false>
QO_DEV void ltable_add(LTable<C> &L, int32_t k, int32_t v, int probes, GEntry *gt, uint64_t gmask,
                       uint32_t *overflow, int skip = 0) {
    uint32_t cnt = 1;
    uint64_t sum = (uint64_t)(int64_t)v;
    int32_t mn = v, mx = v;
    if (COMBINE && ballot(true) == ~uint64_t(0)) {
        const int32_t k0 = __builtin_amdgcn_readfirstlane(k);
        const bool in = k == k0;
        const uint64_t m = ballot(in);
        if (popc(m) >= kWave / 4) {  // uniform across the wave
            const uint64_t s2 = wave_sum(in ? sum : uint64_t(0));
            const int32_t mn2 = wave_min(in ? v : INT_MAX);
            const int32_t mx2 = wave_max(in ? v : INT_MIN);
            if (in) {
                if (lane_id() != (unsigned)__builtin_ctzll(m)) return;
                cnt = popc(m);
                sum = s2;
                mn = mn2;
                mx = mx2;
            }
        }
    }
    int s = (uint32_t)k == kEmptyKey ? -1 : l_slot<C, TWO>(L.key, (uint32_t)k, probes, skip);
    if (s >= 0) {
        atomicAdd(&L.cnt[s], cnt);
        atomicAdd(&L.sum[s], (unsigned long long)sum);
        atomicMin(&L.mn[s], mn);
        atomicMax(&L.mx[s], mx);
    } else {
        g_update(gt, gmask, (uint32_t)k, cnt, sum, mn, mx, overflow);
    }
}

// Merges the block's LDS partials into the global table (after a barrier).
// Block b walks its slots starting at (97 b) mod C, so blocks that finish
// together update different global entries instead of queueing on the same
// ones (measured on 2^28 tuples, 1024 groups: 0.46 -> 0.43 ms).
template <int C>
QO_DEV void ltable_flush(const LTable<C> &L, GEntry *gt, uint64_t gmask, uint32_t *overflow) {
    const int rot = (int)((blockIdx.x * 97u) & (C - 1));
    for (int j = threadIdx.x; j < C; j += blockDim.x) {
        const int i = (j + rot) & (C - 1);
        if (L.cnt[i]) g_update(gt, gmask, L.key[i], L.cnt[i], L.sum[i], L.mn[i], L.mx[i], overflow);
    }
}

// Appends the block's LDS groups as finished rows (the block saw every tuple
// of these keys): one reservation on *d_rows per block; rows past `cap` are
// counted but not written.  After a barrier; NT threads.
template <int C, int NT>
QO_DEV void ltable_append(const LTable<C> &L, GEntry *rows, unsigned long long *d_rows, size_t cap, uint32_t *s_w,
                          unsigned long long *s_base) {
    static_assert(C % NT == 0, "table size");
    constexpr int kPer = C / NT;
    const int i0 = threadIdx.x * kPer;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) c += L.cnt[i0 + k] != 0;
    const uint32_t ex = block_excl_scan<NT / kWave>(c, s_w);
    if (threadIdx.x == NT - 1) *s_base = atomicAdd(d_rows, (unsigned long long)(ex + c));
    __syncthreads();
    uint64_t pos = *s_base + ex;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = i0 + k;
        if (!L.cnt[i]) continue;
        if (pos < cap) {
            GEntry e;
            e.key = L.key[i];
            e.mn = L.mn[i];
            e.mx = L.mx[i];
            e.pad = 0;
            e.cnt = L.cnt[i];
            e.sum = L.sum[i];
            rows[pos] = e;
        }
        ++pos;
    }
}

// Dominant keys of a block: the most common key among 64 sampled tuples (one
// wave; kEmptyKey for lanes without a tuple) and how often it occurs there,
// and the most common other key and its count.
QO_DEV uint32_t sample_dominant(uint32_t k, unsigned *count, uint32_t *second = nullptr, unsigned *count2 = nullptr) {
    uint32_t best = k, best2 = k;
    unsigned bc = 0, bc2 = 0;
#pragma unroll
    for (int c = 0; c < kWave; c += 4) {
        const uint32_t cand = (uint32_t)__shfl((int)k, c, kWave);
        const unsigned n = popc(ballot(k == cand));
        if (n > bc) {
            if (cand != best) {
                bc2 = bc;
                best2 = best;
            }
            bc = n;
            best = cand;
        } else if (n > bc2 && cand != best) {
            bc2 = n;
            best2 = cand;
        }
    }
    *count = bc;
    if (second) {
        *second = best2;
        *count2 = best2 == best ? 0u : bc2;
    }
    return best;
}

// Per-thread COUNT / SUM / MIN / MAX of the dominant key's tuples.
struct DomAcc {
    uint32_t c = 0;
    uint64_t s = 0;
    int32_t mn = INT_MAX, mx = INT_MIN;
    QO_DEV void add(int32_t x) {
        ++c;
        s += (uint64_t)(int64_t)x;
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
    }
};

// Block totals of the dominant key, added once to its LDS entry (or to the
// global table when it has none).  All NT threads call it; the caller
// synchronises before reading the table.
template <int C, int NT>
QO_DEV void dom_merge(LTable<C> &L, const DomAcc &a, uint32_t dom, int probes, int skip, GEntry *gt, uint64_t gmask,
                      uint32_t *overflow) {
    __shared__ uint64_t s_ds[NT / kWave];
    __shared__ uint32_t s_dc[NT / kWave];
    __shared__ int32_t s_dmn[NT / kWave], s_dmx[NT / kWave];
    const uint64_t wc = wave_sum((uint64_t)a.c), ws = wave_sum(a.s);
    const int32_t wmn = wave_min(a.mn), wmx = wave_max(a.mx);
    if (lane_id() == 0) {
        s_dc[wave_id()] = (uint32_t)wc;
        s_ds[wave_id()] = ws;
        s_dmn[wave_id()] = wmn;
        s_dmx[wave_id()] = wmx;
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
        const unsigned i = threadIdx.x;
        const bool in = i < NT / kWave;
        const uint32_t tc = (uint32_t)wave_sum(in ? (uint64_t)s_dc[in ? i : 0] : 0ull);
        const uint64_t ts = wave_sum(in ? s_ds[in ? i : 0] : 0ull);
        const int32_t tmn = wave_min(in ? s_dmn[in ? i : 0] : INT_MAX);
        const int32_t tmx = wave_max(in ? s_dmx[in ? i : 0] : INT_MIN);
        if (i == 0 && tc) {
            const int sl = dom == kEmptyKey ? -1 : l_slot<C, true>(L.key, dom, probes, skip);
            if (sl >= 0) {
                atomicAdd(&L.cnt[sl], tc);
                atomicAdd(&L.sum[sl], (unsigned long long)ts);
                atomicMin(&L.mn[sl], tmn);
                atomicMax(&L.mx[sl], tmx);
            } else {
                g_update(gt, gmask, dom, tc, ts, tmn, tmx, overflow);
            }
        }
    }
}

// Direct hash group-by: every block aggregates a strided share of the input in
// its LDS table and merges it into the global table at the end.  A block whose
// 64-tuple sample holds a key at least 4 times (a skewed input) sums that
// key's tuples, and a second such key's, in registers instead of queueing its
// lanes on one LDS address.
// FILT: only tuples with lo <= filt[i] <= hi are aggregated (the filter is
// fused: its column is read beside key and value, no compaction pass).
template <int C, bool VEC, int NT, bool FILT = false>
__global__ __launch_bounds__(NT) void k_groupby_hash(const int32_t *__restrict__ key, const int32_t *__restrict__ val,
                                                      size_t n, GEntry *gt, uint64_t gmask, int probes,
                                                      uint32_t *overflow, const int32_t *__restrict__ filt = nullptr,
                                                      int32_t flo = 0, int32_t fhi = 0) {
    extern __shared__ unsigned char s_raw[];
    LTable<C> &L = *reinterpret_cast<LTable<C> *>(s_raw);
    __shared__ uint32_t s_dom, s_dom2, s_heavy;
    if (threadIdx.x < kWave) {
        unsigned bc, bc2;
        uint32_t d2;
        const uint32_t d =
            sample_dominant((uint32_t)key[((size_t)blockIdx.x * kWave + threadIdx.x) % n], &bc, &d2, &bc2);
        if (threadIdx.x == 0) {
            s_dom = d;
            s_dom2 = bc2 >= 4 ? d2 : d;
            s_heavy = bc >= 4;
        }
    }
    ltable_init<C>(L);
    __syncthreads();
    const uint32_t dom = s_dom, dom2 = s_dom2;
    auto body = [&](auto heavy_tag) {
        constexpr bool kHeavy = decltype(heavy_tag)::value;
        DomAcc acc, acc2;
        auto add_one = [&](int32_t k, int32_t v) {
            if (kHeavy && (uint32_t)k == dom) acc.add(v);
            else if (kHeavy && (uint32_t)k == dom2) acc2.add(v);
            else ltable_add<C>(L, k, v, probes, gt, gmask, overflow);
        };
        auto add = [&](int32_t k, int32_t v, int32_t f) {
            if (!FILT || (f >= flo && f <= fhi)) add_one(k, v);
        };
        const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        const size_t stride = (size_t)gridDim.x * blockDim.x;
        size_t from = 0;
        if (VEC) {
            const size_t nv = n / 4;
            const int4 *k4 = reinterpret_cast<const int4 *>(key), *v4 = reinterpret_cast<const int4 *>(val);
            size_t i = tid;
            // Four rounds of loads in flight per thread (2^28 tuples, 1024 groups:
            // 0.454 ms against 0.480 ms with two).
            constexpr int kRounds = 4;
            const int4 *f4 = reinterpret_cast<const int4 *>(filt);
            for (; i + (kRounds - 1) * stride < nv; i += kRounds * stride) {
                int4 kk[kRounds], vv[kRounds], ff[kRounds];
#pragma unroll
                for (int u = 0; u < kRounds; ++u) {
                    kk[u] = ld_stream(k4 + i + u * stride);
                    vv[u] = ld_stream(v4 + i + u * stride);
                    if (FILT) ff[u] = ld_stream(f4 + i + u * stride);
                }
#pragma unroll
                for (int u = 0; u < kRounds; ++u) {
                    add(kk[u].x, vv[u].x, FILT ? ff[u].x : 0);
                    add(kk[u].y, vv[u].y, FILT ? ff[u].y : 0);
                    add(kk[u].z, vv[u].z, FILT ? ff[u].z : 0);
                    add(kk[u].w, vv[u].w, FILT ? ff[u].w : 0);
                }
            }
            for (; i < nv; i += stride) {
                const int4 kk = ld_stream(k4 + i), vv = ld_stream(v4 + i);
                const int4 ff = FILT ? ld_stream(f4 + i) : int4{0, 0, 0, 0};
                add(kk.x, vv.x, ff.x);
                add(kk.y, vv.y, ff.y);
                add(kk.z, vv.z, ff.z);
                add(kk.w, vv.w, ff.w);
            }
            from = nv * 4;
        }
        for (size_t e = from + tid; e < n; e += stride) add(key[e], val[e], FILT ? filt[e] : 0);
        if (kHeavy) {
            dom_merge<C, NT>(L, acc, dom, probes, 0, gt, gmask, overflow);
            if (dom2 != dom) {  // block-uniform
                __syncthreads();  // dom_merge's staging is reused
                dom_merge<C, NT>(L, acc2, dom2, probes, 0, gt, gmask, overflow);
            }
        }
    };
    if (s_heavy) body(std::true_type{});
    else body(std::false_type{});
    __syncthreads();
    ltable_flush<C>(L, gt, gmask, overflow);
}

// ---------------------------------------------------------------------------
// Partitioned hash group-by (more groups than one LDS table holds).  The
// (key, value) tuples are split exactly by the multiplicative key hash
// (qo_split.hpp: a histogram pass sizes every bucket, so skewed keys cannot
// overflow) into 2^L fine buckets of about `target` groups each; the LDS table
// of a bucket is indexed by the next hash bits.  One work-group per chunk of a fine
// bucket aggregates in its LDS table and merges its partials into the global
// table, so global atomics fall from four per tuple to four per (chunk, group);
// a heavy key's bucket is simply cut into more chunks.
// ---------------------------------------------------------------------------
constexpr uint32_t kGbChunk = 1u << 18;  // tuples per work item of a large bucket
constexpr uint32_t kGbWhole = 1u << 19;  // buckets up to this size are one work item
constexpr uint32_t kChunkU = 8;          // tuples in flight per thread of k_groupby_chunks

// ---------------------------------------------------------------------------
// One-pass exact partition into up to kOnePassParts fine buckets: per-tile
// bucket counts in tile-major rows, a column scan of the rows into the output
// offset of every (tile, bucket), and one placement pass whose tiles run in
// XCD-contiguous order, so the short runs that consecutive tiles append to a
// bucket are merged into whole lines by one L2.  No reservation atomics and no
// second split pass; the bucket of a key is the top L bits of key * 2^32/phi,
// as in the two-pass exact split.
// ---------------------------------------------------------------------------
constexpr uint32_t kOnePassParts = 1024;
constexpr int kPartThreads = 1024;
constexpr int kPartJ = 16;  // tuples per thread and tile
constexpr uint32_t kPartTile = (uint32_t)kPartThreads * kPartJ;
constexpr uint32_t kPartScanBlocks = 128;
constexpr uint32_t kPartMinSpan = 64;

QO_DEV uint32_t fib_part(int32_t key, int L) { return ((uint32_t)key * 0x9E3779B9u) >> (32 - L); }

// rows[t * nb + f] = tuples of tile t in bucket f.  VEC: the key column is
// 16-byte aligned.
template <bool VEC, bool FILT = false>
__global__ __launch_bounds__(kPartThreads) void k_part_hist(const int32_t *__restrict__ key, size_t n, int L,
                                                            uint32_t nb, uint32_t *__restrict__ rows,
                                                            const int32_t *__restrict__ filt = nullptr,
                                                            int32_t flo = 0, int32_t fhi = 0) {
    __shared__ uint32_t s_c[kOnePassParts];
    for (uint32_t i = threadIdx.x; i < nb; i += kPartThreads) s_c[i] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kPartTile;
    const uint32_t rem = (uint32_t)min((size_t)kPartTile, n - base);
    const int32_t *src = key + base;
    if (VEC && !FILT && rem == kPartTile) {
        // Full tile, 16-byte aligned: kPartJ / 4 int4 loads per thread.
        int4 q[kPartJ / 4];
#pragma unroll
        for (int j = 0; j < kPartJ / 4; ++j)
            q[j] = ld_stream(reinterpret_cast<const int4 *>(src) + (uint32_t)j * kPartThreads + threadIdx.x);
#pragma unroll
        for (int j = 0; j < kPartJ / 4; ++j) {
            atomicAdd(&s_c[fib_part(q[j].x, L)], 1u);
            atomicAdd(&s_c[fib_part(q[j].y, L)], 1u);
            atomicAdd(&s_c[fib_part(q[j].z, L)], 1u);
            atomicAdd(&s_c[fib_part(q[j].w, L)], 1u);
        }
    } else {
        int32_t k[kPartJ];
        bool ok[kPartJ];
#pragma unroll
        for (int j = 0; j < kPartJ; ++j) {
            const uint32_t i = (uint32_t)j * kPartThreads + threadIdx.x;
            k[j] = i < rem ? ld_stream(src + i) : 0;
            const int32_t f = FILT && i < rem ? ld_stream(filt + base + i) : 0;
            ok[j] = i < rem && (!FILT || (f >= flo && f <= fhi));
        }
#pragma unroll
        for (int j = 0; j < kPartJ; ++j)
            if (ok[j]) atomicAdd(&s_c[fib_part(k[j], L)], 1u);
    }
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < nb; f += kPartThreads) rows[(size_t)blockIdx.x * nb + f] = s_c[f];
}

// bsum[b * nb + f] = bucket-f tuples of tiles [b * span, (b + 1) * span).
__global__ __launch_bounds__(kPartThreads) void k_part_reduce(const uint32_t *__restrict__ rows, uint32_t T, uint32_t nb,
                                                              uint32_t span, uint32_t *__restrict__ bsum) {
    const uint32_t f = threadIdx.x;
    if (f >= nb) return;
    const uint32_t t0 = blockIdx.x * span, t1 = min(T, t0 + span);
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t t = t0;
    for (; t + 8 <= t1; t += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] += rows[(size_t)(t + u) * nb + f];
    for (; t < t1; ++t) c[0] += rows[(size_t)t * nb + f];
    bsum[(size_t)blockIdx.x * nb + f] = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7];
}

// One block: total[f] = tuples in bucket f, starts[f] = where bucket f begins
// (starts[nb] = n), and bsum[b][f] := starts[f] + bucket-f tuples of the tiles
// of blocks before b (in place).
__global__ __launch_bounds__(kPartThreads) void k_part_starts(uint32_t *bsum, uint32_t nbk, uint32_t nb,
                                                              uint32_t *total, uint32_t *starts) {
    __shared__ uint32_t s_w[kPartThreads / kWave];
    const uint32_t f = threadIdx.x;
    uint32_t tot = 0;
    if (f < nb) {
        uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t b = 0;
        for (; b + 8 <= nbk; b += 8)
#pragma unroll
            for (int u = 0; u < 8; ++u) c[u] += bsum[(size_t)(b + u) * nb + f];
        for (; b < nbk; ++b) c[0] += bsum[(size_t)b * nb + f];
        tot = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7];
    }
    const uint32_t start = block_excl_scan<kPartThreads / kWave>(tot, s_w);
    if (f >= nb) return;
    total[f] = tot;
    starts[f] = start;
    if (f == nb - 1) starts[nb] = start + tot;
    uint32_t run = start;
    for (uint32_t b = 0; b < nbk; ++b) {
        const uint32_t v = bsum[(size_t)b * nb + f];
        bsum[(size_t)b * nb + f] = run;
        run += v;
    }
}

// rows[t][f] := output offset of the first bucket-f tuple of tile t.
__global__ __launch_bounds__(kPartThreads) void k_part_apply(uint32_t *__restrict__ rows, uint32_t T, uint32_t nb,
                                                             uint32_t span, const uint32_t *__restrict__ bsum) {
    const uint32_t f = threadIdx.x;
    if (f >= nb) return;
    const uint32_t t0 = blockIdx.x * span, t1 = min(T, t0 + span);
    uint32_t run = bsum[(size_t)blockIdx.x * nb + f];
    uint32_t t = t0;
    for (; t + 8 <= t1; t += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = rows[(size_t)(t + u) * nb + f];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            rows[(size_t)(t + u) * nb + f] = run;
            run += v[u];
        }
    }
    for (; t < t1; ++t) {
        const uint32_t v = rows[(size_t)t * nb + f];
        rows[(size_t)t * nb + f] = run;
        run += v;
    }
}

// Places tile xcd_tile(blockIdx.x): ranks its tuples by bucket (LDS atomics;
// order inside a bucket is free for aggregation), stages them in bucket order
// and writes each bucket's run at the offset from its row.
template <bool FILT = false>
__global__ __launch_bounds__(kPartThreads) void k_part_place(const int32_t *__restrict__ key,
                                                             const int32_t *__restrict__ val, size_t n, int L,
                                                             uint32_t nb, uint32_t T, const uint32_t *__restrict__ offs,
                                                             uint64_t *__restrict__ out,
                                                             const int32_t *__restrict__ filt = nullptr,
                                                             int32_t flo = 0, int32_t fhi = 0) {
    __shared__ uint64_t s_stage[kPartTile];
    __shared__ uint32_t s_cnt[kOnePassParts], s_loff[kOnePassParts], s_g[kOnePassParts];
    __shared__ uint32_t s_w[kPartThreads / kWave];
    const uint32_t t = xcd_tile(blockIdx.x, T);
    const size_t base = (size_t)t * kPartTile;
    const uint32_t rem = (uint32_t)min((size_t)kPartTile, n - base);
    for (uint32_t i = threadIdx.x; i < nb; i += kPartThreads) s_cnt[i] = 0;
    __syncthreads();
    int32_t k[kPartJ], v[kPartJ];
    uint32_t r[kPartJ];
    uint32_t live = 0;  // bit j: tuple j of this thread is in the tile and passes the filter
#pragma unroll
    for (int j = 0; j < kPartJ; ++j) {
        const uint32_t i = (uint32_t)j * kPartThreads + threadIdx.x;
        k[j] = i < rem ? ld_stream(key + base + i) : 0;
        v[j] = i < rem ? ld_stream(val + base + i) : 0;
        const int32_t f = FILT && i < rem ? ld_stream(filt + base + i) : 0;
        live |= (i < rem && (!FILT || (f >= flo && f <= fhi))) ? (1u << j) : 0u;
    }
#pragma unroll
    for (int j = 0; j < kPartJ; ++j) r[j] = (live >> j) & 1u ? atomicAdd(&s_cnt[fib_part(k[j], L)], 1u) : 0u;
    __syncthreads();
    {
        const uint32_t f = threadIdx.x;
        const uint32_t c = f < nb ? s_cnt[f] : 0u;
        const uint32_t ex = block_excl_scan<kPartThreads / kWave>(c, s_w);
        if (f < nb) {
            s_loff[f] = ex;
            s_g[f] = offs[(size_t)t * nb + f];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPartJ; ++j)
        if ((live >> j) & 1u)
            s_stage[s_loff[fib_part(k[j], L)] + r[j]] = ((uint64_t)(uint32_t)v[j] << 32) | (uint32_t)k[j];
    __syncthreads();
    const uint32_t placed = s_loff[nb - 1] + s_cnt[nb - 1];  // the tile's passing tuples
    for (uint32_t i = threadIdx.x; i < placed; i += kPartThreads) {
        const uint64_t e = s_stage[i];
        const uint32_t f = fib_part((int32_t)(uint32_t)e, L);
        out[s_g[f] + (i - s_loff[f])] = e;
    }
}

struct OnePassBufs {
    uint32_t *rows = nullptr, *bsum = nullptr, *total = nullptr, *starts = nullptr;
    uint64_t *fine = nullptr;
    uint32_t T = 0, span = 0, nbk = 0;
};

static OnePassBufs onepass_carve(size_t n, uint32_t nb, Carver &cv) {
    OnePassBufs b;
    b.T = (uint32_t)((n + kPartTile - 1) / kPartTile);
    b.span = std::max(kPartMinSpan, (b.T + kPartScanBlocks - 1) / kPartScanBlocks);
    b.nbk = (b.T + b.span - 1) / b.span;
    b.rows = cv.take<uint32_t>((size_t)std::max(1u, b.T) * nb);
    b.bsum = cv.take<uint32_t>((size_t)std::max(1u, b.nbk) * nb);
    b.total = cv.take<uint32_t>(nb + 1);
    b.starts = cv.take<uint32_t>(nb + 1);
    b.fine = cv.take<uint64_t>(n ? n : 1);
    return b;
}

static int launch_onepass_partition(const OnePassBufs &b, const int32_t *key, const int32_t *val, size_t n, int L,
                                    uint32_t nb, hipStream_t st, const int32_t *filt = nullptr, int32_t flo = 0,
                                    int32_t fhi = 0) {
    QO_REQUIRE(n < 0xFFFFFFFFull && nb <= kOnePassParts && (1u << L) == nb, "one-pass partition too large");
    if (b.T == 0) return QO_OK;
    if (filt) k_part_hist<false, true><<<b.T, kPartThreads, 0, st>>>(key, n, L, nb, b.rows, filt, flo, fhi);
    else if ((reinterpret_cast<uintptr_t>(key) & 15) == 0) k_part_hist<true><<<b.T, kPartThreads, 0, st>>>(key, n, L, nb, b.rows);
    else k_part_hist<false><<<b.T, kPartThreads, 0, st>>>(key, n, L, nb, b.rows);
    QO_LAUNCHED();
    k_part_reduce<<<b.nbk, kPartThreads, 0, st>>>(b.rows, b.T, nb, b.span, b.bsum);
    QO_LAUNCHED();
    k_part_starts<<<1, kPartThreads, 0, st>>>(b.bsum, b.nbk, nb, b.total, b.starts);
    QO_LAUNCHED();
    k_part_apply<<<b.nbk, kPartThreads, 0, st>>>(b.rows, b.T, nb, b.span, b.bsum);
    QO_LAUNCHED();
    if (filt) k_part_place<true><<<b.T, kPartThreads, 0, st>>>(key, val, n, L, nb, b.T, b.rows, b.fine, filt, flo, fhi);
    else k_part_place<false><<<b.T, kPartThreads, 0, st>>>(key, val, n, L, nb, b.T, b.rows, b.fine);
    QO_LAUNCHED();
    return QO_OK;
}

// Work items per fine bucket: one for a bucket of at most kGbWhole tuples,
// else ceil(hist[f] / kGbChunk); cch[nb] = 0.
__global__ void k_chunk_counts(const uint32_t *hist, uint32_t nb, uint32_t *cch) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > nb) return;
    const uint32_t h = f < nb ? hist[f] : 0u;
    cch[f] = h == 0 ? 0u : h <= kGbWhole ? 1u : (h + kGbChunk - 1) / kGbChunk;
}

// Work item w of fine bucket f (cpre[f] <= w < cpre[f + 1]).  A bucket that
// is one work item holds every tuple of its keys, so the block appends its
// groups as finished rows; a chunk of a larger bucket merges its partials
// into the global table.  NT = 1024: two 48 KB blocks fill a CU (32 waves), so
// a grid of 2 * 256 * k work items runs in k full rounds.
constexpr int kChunkThreads = 1024;

template <int C, int NT>
__global__ __launch_bounds__(NT, 8) void k_groupby_chunks(const uint64_t *__restrict__ fine, const uint32_t *offp,
                                                        const uint32_t *off, const uint32_t *cpre, uint32_t nb, int L_bits, GEntry *gt,
                                                        uint64_t gmask, int probes, uint32_t *overflow, GEntry *rows,
                                                        unsigned long long *d_rows, size_t rows_cap) {
    extern __shared__ unsigned char s_raw[];
    LTable<C> &L = *reinterpret_cast<LTable<C> *>(s_raw);
    __shared__ uint32_t s_w[NT / kWave];
    __shared__ unsigned long long s_base;
    const uint32_t w = blockIdx.x;
    if (w >= cpre[nb]) return;  // whole block: the grid is an upper bound
    int lo = 0, hi = (int)nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cpre[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    const uint64_t size = off[lo + 1] - off[lo];  // bucket lo: [offp[lo], offp[lo] + size) of `fine`
    const bool whole = size <= kGbWhole;
    const uint64_t e0 = offp[lo] + (whole ? 0 : (uint64_t)(w - cpre[lo]) * kGbChunk);
    const uint64_t e1 = offp[lo] + size;
    const uint32_t m = whole ? (uint32_t)size : (uint32_t)(e1 - e0 < kGbChunk ? e1 - e0 : kGbChunk);
    const uint64_t *src = fine + e0;
    // The block takes the most common key of its first 64 tuples as the
    // dominant key K.  When K is at least 4 of them (a skewed bucket, such as
    // a heavy key's), tuples of K are summed in registers and merged once at
    // the end and only the others go to the LDS table.  (2^28 tuples, 2^20
    // groups, Zipf z = 1: the key of rank 1 alone is 7 % of all tuples, and
    // every key of rank up to ~1000 is a sizeable share of its bucket.)
    // A second key that is also at least 4 of the sample gets registers of its
    // own.
    __shared__ uint32_t s_dom, s_dom2, s_heavy;
    if (threadIdx.x < kWave) {
        unsigned bc, bc2;
        uint32_t d2;
        const uint32_t d = sample_dominant(threadIdx.x < m ? (uint32_t)src[threadIdx.x] : kEmptyKey, &bc, &d2, &bc2);
        if (threadIdx.x == 0) {
            s_dom = d;
            s_dom2 = bc2 >= 4 ? d2 : d;
            s_heavy = bc >= 4;
        }
    }
    ltable_init<C>(L);
    __syncthreads();
    const bool heavy = s_heavy;
    // kChunkU tuples per thread and round, all loaded before any is added, so
    // a thread keeps that many loads in flight.
    constexpr uint32_t kRound = NT * kChunkU;
    if (!heavy) {
        for (uint32_t r = 0; r < m; r += kRound) {
            uint64_t v[kChunkU];
#pragma unroll
            for (int u = 0; u < (int)kChunkU; ++u) {
                const uint32_t i = r + u * NT + threadIdx.x;
                v[u] = i < m ? ld_stream(src + i) : 0ull;
            }
#pragma unroll
            for (int u = 0; u < (int)kChunkU; ++u)
                if (r + u * NT + threadIdx.x < m)
                    ltable_add<C, false, true>(L, (int32_t)(uint32_t)v[u], (int32_t)(v[u] >> 32), probes, gt, gmask,
                                               overflow, L_bits);
        }
    } else {
        const uint32_t dom = s_dom, dom2 = s_dom2;
        DomAcc acc, acc2;
        for (uint32_t r = 0; r < m; r += kRound) {
            uint64_t v[kChunkU];
#pragma unroll
            for (int u = 0; u < (int)kChunkU; ++u) {
                const uint32_t i = r + u * NT + threadIdx.x;
                v[u] = i < m ? ld_stream(src + i) : 0ull;
            }
#pragma unroll
            for (int u = 0; u < (int)kChunkU; ++u) {
                if (r + u * NT + threadIdx.x >= m) continue;
                const uint32_t k = (uint32_t)v[u];
                const int32_t x = (int32_t)(v[u] >> 32);
                if (k == dom) acc.add(x);
                else if (k == dom2) acc2.add(x);
                else ltable_add<C, false, true>(L, (int32_t)k, x, probes, gt, gmask, overflow, L_bits);
            }
        }
        dom_merge<C, NT>(L, acc, dom, probes, L_bits, gt, gmask, overflow);
        if (dom2 != dom) {  // block-uniform
            __syncthreads();  // dom_merge's staging is reused
            dom_merge<C, NT>(L, acc2, dom2, probes, L_bits, gt, gmask, overflow);
        }
    }
    __syncthreads();
    if (whole) ltable_append<C, NT>(L, rows, d_rows, rows_cap, s_w, &s_base);
    else ltable_flush<C>(L, gt, gmask, overflow);
}

// Occupied global entries appended to `rows`: kCollectJ entries per thread,
// one reservation per block and round.
constexpr int kCollectJ = 8;
__global__ __launch_bounds__(256) void k_gtable_collect(const GEntry *t, size_t entries, GEntry *rows,
                                                        unsigned long long *d_rows, size_t cap) {
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_base;
    constexpr size_t kRound = 256 * kCollectJ;
    for (size_t r0 = (size_t)blockIdx.x * kRound; r0 < entries; r0 += (size_t)gridDim.x * kRound) {
        const size_t i0 = r0 + (size_t)threadIdx.x * kCollectJ;
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < kCollectJ; ++k) c += (i0 + k < entries) && t[i0 + k].cnt != 0;
        const uint32_t ex = block_excl_scan256(c, s_w);
        if (threadIdx.x == 255) s_base = ex + c ? atomicAdd(d_rows, (unsigned long long)(ex + c)) : 0ull;
        __syncthreads();
        uint64_t pos = s_base + ex;
#pragma unroll
        for (int k = 0; k < kCollectJ; ++k) {
            if (i0 + k >= entries || t[i0 + k].cnt == 0) continue;
            if (pos < cap) rows[pos] = t[i0 + k];
            ++pos;
        }
        __syncthreads();  // s_w and s_base are reused by the next round
    }
}

// Sort keys of the finished rows: (key ^ signbit) << 32 | row.
__global__ void k_rows_list(const GEntry *rows, const unsigned long long *d_rows, size_t cap, uint64_t *list) {
    const uint64_t n = *d_rows < cap ? *d_rows : cap;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        list[i] = ((uint64_t)(rows[i].key ^ 0x80000000u) << 32) | (uint32_t)i;
}

__global__ void k_gtable_gather(const GEntry *rows, const uint64_t *list, const uint64_t *d_count, size_t cap,
                                const uint32_t *overflow, int32_t *okey, int64_t *ocount, int64_t *osum,
                                int32_t *omin, int32_t *omax, uint64_t *d_num) {
    size_t g = d_count[0];
    size_t lim = g < cap ? g : cap;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < lim; i += stride) {
        uint64_t v = list[i];
        const GEntry &e = rows[(uint32_t)v];
        okey[i] = (int32_t)((uint32_t)(v >> 32) ^ 0x80000000u);
        ocount[i] = (int64_t)e.cnt;
        osum[i] = (int64_t)e.sum;
        omin[i] = e.mn;
        omax[i] = e.mx;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_num = *overflow ? ~uint64_t(0) : g;
}

struct GroupbyPlan {
    uint64_t gmask;
    size_t entries;
};

static GroupbyPlan groupby_plan(size_t group_capacity) {
    size_t cap = 64;
    while (cap < 2 * group_capacity + 1) cap <<= 1;
    return {cap - 1, cap + 1};
}

struct GbPartPlan {
    bool on = false;
    SplitPlan s;
};

// Partition when the groups outgrow one LDS table: 2^L fine buckets of about
// `target` groups each (tune items), L <= 14.
// With a filter the partition stays one pass (at most kOnePassParts buckets,
// whose LDS tables spill to the global table beyond 2^20 groups): the filter is
// fused into that pass.
static GbPartPlan groupby_part_plan(size_t group_capacity, size_t n, bool filtered) {
    GbPartPlan p;
    const TuneParams tp = tune_get(TUNE_GROUPBY_PART, n);
    if (tp.param == PART_NEVER || n == 0 || n >= 0xFFFFFFFFull || group_capacity < 2) return p;
    if (tp.param != PART_ALWAYS && (group_capacity <= 1024 || n < (size_t(1) << 20))) return p;
    const size_t target = (size_t)std::max(16, tp.items);
    const int lmax = filtered ? __builtin_ctz(kOnePassParts) : 2 * kMaxSplitBits;
    int L = 1;
    while (L < lmax && (group_capacity >> L) > target) ++L;
    p.on = true;
    p.s = split_plan(L, n);
    return p;
}

struct GroupbyBufs {
    GEntry *gt;
    uint64_t *counters;  // [0] groups found (finished rows), [1] table-full flag (low word)
    GEntry *rows;        // finished groups, unordered
    uint64_t *list;
    void *rws;
    ExactSplitBufs sb;   // partitioned path, two-pass split (more than kOnePassParts buckets)
    OnePassBufs op;      // partitioned path, one-pass partition
    uint32_t *cch = nullptr, *cpre = nullptr;
    void *scan_ws = nullptr;
};

static GroupbyBufs groupby_carve(const GroupbyPlan &p, const GbPartPlan &pp, size_t group_capacity, size_t n,
                                 Carver &cv) {
    GroupbyBufs b;
    b.gt = cv.take<GEntry>(p.entries);
    b.counters = cv.take<uint64_t>(4);
    b.rows = cv.take<GEntry>(group_capacity ? group_capacity : 1);
    b.list = cv.take<uint64_t>(group_capacity ? group_capacity : 1);
    b.rws = cv.take<char>(radix_workspace_bytes(group_capacity, false));
    if (pp.on) {
        const size_t nb = size_t(1) << pp.s.L;
        if (nb <= kOnePassParts) b.op = onepass_carve(n, (uint32_t)nb, cv);
        else b.sb = exact_split_carve(pp.s, n, cv);
        b.cch = cv.take<uint32_t>(nb + 1);
        b.cpre = cv.take<uint32_t>(nb + 1);
        b.scan_ws = cv.take<char>(scan_workspace_bytes(nb + 1));
    }
    return b;
}

size_t groupby_ws_bytes(size_t group_capacity, size_t n, bool filtered) {
    Carver cv(nullptr);
    groupby_carve(groupby_plan(group_capacity), groupby_part_plan(group_capacity, n, filtered), group_capacity, n,
                  cv);
    return cv.off + 256;
}

int launch_groupby(const int32_t *key, const int32_t *val, size_t n, size_t group_capacity, int32_t *okey,
                   int64_t *ocount, int64_t *osum, int32_t *omin, int32_t *omax, uint64_t *d_num, void *ws,
                   hipStream_t st, const int32_t *filt, int32_t flo, int32_t fhi) {
    GroupbyPlan p = groupby_plan(group_capacity);
    const GbPartPlan pp = groupby_part_plan(group_capacity, n, filt != nullptr);
    Carver cv(ws);
    const GroupbyBufs b = groupby_carve(p, pp, group_capacity, n, cv);
    GEntry *gt = b.gt;
    uint64_t *counters = b.counters;
    uint64_t *list = b.list;
    void *rws = b.rws;
    uint32_t *overflow = reinterpret_cast<uint32_t *>(counters + 1);
    QO_TRY(launch_fill_u64(counters, 0, 4, st));
    const unsigned ig = grid_for(p.entries, 256, (unsigned)props().cu_count * 16);
    k_gtable_init<<<ig, 256, 0, st>>>(gt, p.entries);
    QO_LAUNCHED();
    if (n && pp.on) {
        const uint32_t nb = 1u << pp.s.L;
        const bool onepass = nb <= kOnePassParts;
        // Fine buckets back to back: bucket f at [offp[f], offp[f] + off[f + 1] - off[f]) of `fine`.
        const uint32_t *hist, *offp, *off;
        const uint64_t *fine;
        if (onepass) {
            QO_TRY(launch_onepass_partition(b.op, key, val, n, pp.s.L, nb, st, filt, flo, fhi));
            hist = b.op.total;
            offp = off = b.op.starts;
            fine = b.op.fine;
        } else {
            QO_TRY(launch_exact_splits(pp.s, b.sb, key, val, n, st));
            hist = b.sb.hist;
            offp = b.sb.offp;
            off = b.sb.off;
            fine = b.sb.fine;
        }
        k_chunk_counts<<<(nb + 1 + 255) / 256, 256, 0, st>>>(hist, nb, b.cch);
        QO_LAUNCHED();
        QO_TRY(launch_scan_excl_u32(b.cch, b.cpre, nb + 1, b.scan_ws, st));
        // Upper bound on the chunk count: every bucket one partial chunk.
        const uint64_t items = (uint64_t)nb + (n + kGbChunk - 1) / kGbChunk;
        QO_REQUIRE(items < 0x7FFFFFFFull, "group-by input too large");
        k_groupby_chunks<2048, kChunkThreads><<<(unsigned)items, kChunkThreads, sizeof(LTable<2048>), st>>>(
            fine, offp, off, b.cpre, nb, pp.s.L, gt, p.gmask, 64, overflow, b.rows,
            reinterpret_cast<unsigned long long *>(counters), group_capacity);
        QO_LAUNCHED();
    } else if (n) {
        TuneParams tp = tune_get(TUNE_GROUPBY, n);
        const bool vec = ((uintptr_t)key % 16 == 0) && ((uintptr_t)val % 16 == 0) && ((uintptr_t)filt % 16 == 0);
        int C = tp.items >= 4096 ? 4096 : 2048;
        const size_t lds4096 = sizeof(LTable<4096>);
        if (C == 4096 && (size_t)props().lds_per_block < lds4096) C = 2048;
        // One 1024-thread block per CU by default: every block merges its whole
        // LDS table into the global one at the end, and those global atomics
        // (blocks x groups x 4) cost more than the occupancy of more blocks buys
        // (2^28 tuples, 1024 groups: 0.68 ms with 3 x 512 threads per CU, 0.43
        // ms with 1 x 1024).
        constexpr int block = 1024;
        const unsigned grid = grid_for(n, block * 4, (unsigned)props().cu_count * (unsigned)std::max(1, tp.param));
        // Low cardinality: the LDS table holds every group.  High cardinality:
        // it caches the keys seen first and the rest go straight to global.
        const int probes = group_capacity * 2 <= (size_t)C ? C : 8;
#define QO_GBH(CC, V, F)                                                                                    \
    k_groupby_hash<CC, V, block, F><<<grid, block, sizeof(LTable<CC>), st>>>(key, val, n, gt, p.gmask, probes,  \
                                                                             overflow, filt, flo, fhi)
        if (C == 4096) {
            if (filt) {
                if (vec) QO_GBH(4096, true, true);
                else QO_GBH(4096, false, true);
            } else {
                if (vec) QO_GBH(4096, true, false);
                else QO_GBH(4096, false, false);
            }
        } else {
            if (filt) {
                if (vec) QO_GBH(2048, true, true);
                else QO_GBH(2048, false, true);
            } else {
                if (vec) QO_GBH(2048, true, false);
                else QO_GBH(2048, false, false);
            }
        }
#undef QO_GBH
        QO_LAUNCHED();
    }
    unsigned long long *d_rows = reinterpret_cast<unsigned long long *>(counters);
    const unsigned cg = grid_for(p.entries, 256 * kCollectJ, (unsigned)props().cu_count * 4);
    k_gtable_collect<<<cg, 256, 0, st>>>(gt, p.entries, b.rows, d_rows, group_capacity);
    QO_LAUNCHED();
    const unsigned gg = grid_for(group_capacity, 256, (unsigned)props().cu_count * 8);
    k_rows_list<<<gg, 256, 0, st>>>(b.rows, d_rows, group_capacity, list);
    QO_LAUNCHED();
    if (group_capacity > 1) QO_TRY(launch_radix_sort_u64_dev(list, nullptr, counters, group_capacity, 32, 64, rws, st));
    k_gtable_gather<<<gg, 256, 0, st>>>(b.rows, list, counters, group_capacity, overflow, okey, ocount, osum, omin,
                                        omax, d_num);
    QO_LAUNCHED();
    return QO_OK;
}

// ---------------------------------------------------------------------------
// Sort-based pipeline (reference semantics).
// ---------------------------------------------------------------------------
// Group start positions over Records sorted by .y: record i starts a group
// when i == 0 or its .y differs from record i - 1's.  Tiles of J * NT * 2
// records (16-byte loads of two records; the record before each pair is read
// with its own load, an L1 hit), 32-bit in-tile offsets, one ticketed
// decoupled look-back per tile; the starts are written in order.
template <int J, int NT, bool VEC>
__global__ __launch_bounds__(NT) void k_group_startpos(const qo_record *__restrict__ s, size_t n, int32_t *startPos,
                                                       uint64_t *status, uint32_t *ticket, uint64_t *d_num,
                                                       uint32_t num_tiles, DevCtl *ctl) {
    constexpr int NW = NT / kWave;
    constexpr int V = VEC ? 2 : 1;
    constexpr uint32_t kTile = J * NT * V;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tot[tile_offsets_words(J, NW)];
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= num_tiles) return;  // whole block; one ticket per launched block, so never taken
    const size_t ebase = (size_t)tile * kTile;
    const qo_record *src = s + ebase;
    const uint32_t rem = (uint32_t)min((size_t)kTile, n - ebase);
    uint32_t fl[J], c[J], ex[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const uint32_t e0 = ((uint32_t)j * NT + threadIdx.x) * V;
        uint32_t y[V];
        uint32_t prev = 0;
        fl[j] = 0;
        if (e0 < rem) {
            if (VEC && e0 + 2 <= rem) {
                const uint4 q = ld_stream(reinterpret_cast<const uint4 *>(src) + (e0 >> 1));
                y[0] = q.y;
                y[V - 1] = q.w;
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) y[k] = e0 + k < rem ? src[e0 + k].y : 0u;
            }
            const bool first = ebase + e0 == 0;
            if (!first) prev = src[(int64_t)e0 - 1].y;
            fl[j] = (first || y[0] != prev) ? 1u : 0u;
            if (V == 2 && e0 + 1 < rem && y[V - 1] != y[0]) fl[j] |= 2u;
        }
        c[j] = __popc(fl[j]);
    }
    const uint32_t total = tile_offsets<J, NW>(c, ex, s_tot);
    if (threadIdx.x < kWave) {
        const uint64_t e = lookback(status, tile, total, ctl);
        if (threadIdx.x == 0) s_excl = e;
    }
    __syncthreads();
    int32_t *dst = startPos + s_excl;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const uint32_t e0 = ((uint32_t)j * NT + threadIdx.x) * V;
        uint32_t pos = ex[j];
        if (fl[j] & 1u) dst[pos++] = (int32_t)(ebase + e0);
        if (fl[j] & 2u) dst[pos] = (int32_t)(ebase + e0 + 1);
    }
    if (tile == num_tiles - 1 && threadIdx.x == 0) *d_num = s_excl + total;
}

constexpr int kStartJ = 8;
constexpr int kStartThreads = 1024;

// Work units of kAggLanes lanes: group g, chunk q covers startPos[g] + q*kAggChunk.
constexpr int kAggChunk = 128;
constexpr int kAggLanes = 8;

// Range of group g, clamped to the array (start positions come from the caller).
QO_DEV void group_range(const int32_t *startPos, uint32_t g, uint32_t G, size_t n, int64_t &s, int64_t &e) {
    s = startPos[g];
    e = (g == G - 1) ? (int64_t)n : startPos[g + 1];
    if (s < 0) s = 0;
    if (e > (int64_t)n) e = (int64_t)n;
}

__global__ void k_agg_chunks(const int32_t *startPos, uint32_t G, size_t n, uint32_t *chunks) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    int64_t s, e;
    group_range(startPos, (uint32_t)g, G, n, s, e);
    int64_t len = e - s;
    chunks[g] = len > 0 ? (uint32_t)((len + kAggChunk - 1) / kAggChunk) : 0u;
}

QO_DEV int32_t agg_identity(int op) { return op == QO_REDUCE_MIN ? (1 << 30) : 0; }

// Owner group of every work unit below umax: units [cofs[g], cofs[g] +
// chunks[g]) belong to group g (one thread per group writes its units).
__global__ void k_agg_units(const uint32_t *cofs, const uint32_t *chunks, uint32_t G, uint32_t *unit_group,
                            uint32_t umax) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    for (uint32_t q = 0; q < chunks[g] && cofs[g] + q < umax; ++q) unit_group[cofs[g] + q] = (uint32_t)g;
}

// Group of work unit w: from the owner table, or, for units past it (start
// positions that overlap, which only a malformed input produces), the last g
// with cofs[g] <= w.  An empty group shares its cofs with the following group,
// so the last such g is the non-empty group that owns unit w.
QO_DEV uint32_t unit_owner(const uint32_t *unit_group, uint32_t umax, const uint32_t *cofs, uint32_t G, size_t w) {
    if (w < umax) return unit_group[w];
    uint32_t lo = 0, hi = G - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (cofs[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// kAggLanes lanes per work unit, kWave / kAggLanes units per wave at once:
// each lane reduces every kAggLanes-th row of its unit (all its loads issued
// before any is used), the unit's lanes combine by shuffles, and the first
// lane stores the result of a group that is one unit or merges a partial
// with an atomic.  `acc` holds the identity beforehand.  With one wave per
// 512-row unit, groups of ~64 rows left each wave mostly waiting on its
// unit's index loads and atomic (0.32 ms per aggregate over 2^26 Records).
__global__ __launch_bounds__(256) void k_agg_groups(const qo_record *__restrict__ s, size_t n,
                                                    const int32_t *__restrict__ startPos, uint32_t G,
                                                    const uint32_t *__restrict__ cofs, const uint32_t *__restrict__ chunks,
                                                    const uint32_t *__restrict__ unit_group, uint32_t umax, int op,
                                                    int32_t *acc) {
    constexpr int kPerLane = kAggChunk / kAggLanes;
    constexpr int kUnitsPerWave = kWave / kAggLanes;
    const uint32_t total = cofs[G - 1] + chunks[G - 1];
    const unsigned lane = lane_id(), sub = lane / kAggLanes, sl = lane % kAggLanes;
    const size_t waves = (size_t)gridDim.x * (blockDim.x / kWave);
    for (size_t w0 = ((size_t)blockIdx.x * (blockDim.x / kWave) + wave_id()) * kUnitsPerWave; w0 < total;
         w0 += waves * kUnitsPerWave) {
        const size_t w = w0 + sub;
        const bool active = w < total;
        uint32_t g = 0;
        int64_t a = 0, b = 0;
        if (active) {
            g = unit_owner(unit_group, umax, cofs, G, w);
            int64_t gs, ge;
            group_range(startPos, g, G, n, gs, ge);
            a = gs + (int64_t)(w - cofs[g]) * kAggChunk;
            b = a + kAggChunk < ge ? a + kAggChunk : ge;
        }
        int32_t v[kPerLane];
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            const int64_t i = a + sl + (int64_t)k * kAggLanes;
            v[k] = i < b ? (int32_t)s[i].x : agg_identity(op);
        }
        int32_t local = agg_identity(op);
        uint32_t usum = 0;
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            if (op == QO_REDUCE_MAX) local = v[k] > local ? v[k] : local;
            else if (op == QO_REDUCE_MIN) local = v[k] < local ? v[k] : local;
            else usum += (uint32_t)v[k];
        }
        // Combine the unit's lanes (xor offsets below kAggLanes stay inside the unit).
#pragma unroll
        for (int o = kAggLanes / 2; o > 0; o >>= 1) {
            if (op == QO_REDUCE_MAX) local = max(local, __shfl_xor(local, o, kWave));
            else if (op == QO_REDUCE_MIN) local = min(local, __shfl_xor(local, o, kWave));
            else usum += __shfl_xor(usum, o, kWave);
        }
        if (!active || sl != 0) continue;
        if (op != QO_REDUCE_MAX && op != QO_REDUCE_MIN) local = (int32_t)usum;
        if (chunks[g] == 1) acc[g] = local;  // the group's only unit: its result is final
        else if (op == QO_REDUCE_MAX) atomicMax(&acc[g], local);
        else if (op == QO_REDUCE_MIN) atomicMin(&acc[g], local);
        else atomicAdd(reinterpret_cast<uint32_t *>(&acc[g]), usum);
    }
}

__global__ void k_agg_average(const int32_t *startPos, uint32_t G, size_t n, const int32_t *sum, int32_t *out) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    // Count as the reference divides by it: end - start, unclamped.
    int64_t s = startPos[g];
    int64_t e = (g == G - 1) ? (int64_t)n : startPos[g + 1];
    int32_t cnt = (int32_t)(e - s);
    out[g] = cnt > 0 ? sum[g] / cnt : 0;
}

// Work units: at most one partial 512-row chunk per group beyond n / 512 full ones.
static size_t agg_units_max(size_t n, size_t G) { return n / kAggChunk + G + 1; }

size_t aggregate_ws_bytes(size_t G, size_t n) {
    return G * 12 + agg_units_max(n, G) * 4 + scan_workspace_bytes(G) + 2048;
}

int launch_aggregate_groups(const qo_record *s, size_t n, const int32_t *startPos, size_t G, int op, int32_t *out,
                            void *ws, hipStream_t st) {
    if (G == 0) return QO_OK;
    Carver cv(ws);
    uint32_t *chunks = cv.take<uint32_t>(G);
    uint32_t *cofs = cv.take<uint32_t>(G);
    int32_t *sum = cv.take<int32_t>(G);
    uint32_t *unit_group = cv.take<uint32_t>(agg_units_max(n, G));
    void *sws = cv.take<char>(scan_workspace_bytes(G));
    int32_t *acc = op == QO_REDUCE_AVERAGE ? sum : out;
    QO_TRY(launch_fill_u32(acc, op == QO_REDUCE_MIN ? (1u << 30) : 0u, G, st));
    const unsigned gb = grid_for(G, 256, 1u << 30);
    k_agg_chunks<<<gb, 256, 0, st>>>(startPos, (uint32_t)G, n, chunks);
    QO_LAUNCHED();
    QO_TRY(launch_scan_excl_u32(chunks, cofs, G, sws, st));
    const size_t umax = std::min<size_t>(agg_units_max(n, G), 0xFFFFFFFFull);
    k_agg_units<<<gb, 256, 0, st>>>(cofs, chunks, (uint32_t)G, unit_group, (uint32_t)umax);
    QO_LAUNCHED();
    const unsigned grid = grid_for((size_t)n / kAggChunk + G, 4 * (kWave / kAggLanes), (unsigned)props().cu_count * 16);
    k_agg_groups<<<grid, 256, 0, st>>>(s, n, startPos, (uint32_t)G, cofs, chunks, unit_group, (uint32_t)umax, op,
                                       acc);
    QO_LAUNCHED();
    if (op == QO_REDUCE_AVERAGE) {
        k_agg_average<<<gb, 256, 0, st>>>(startPos, (uint32_t)G, n, sum, out);
        QO_LAUNCHED();
    }
    return QO_OK;
}

int launch_group_startpos(const qo_record *sorted, size_t n, int32_t *startPos, uint64_t *d_num, void *ws,
                          hipStream_t st) {
    if (n == 0) return launch_fill_u64(d_num, 0, 1, st);
    QO_REQUIRE(n <= 0x7FFFFFFFull, "start positions are int32: at most 2^31 - 1 records");
    constexpr int J = kStartJ, NT = kStartThreads;
    const bool vec = ((uintptr_t)sorted % 16) == 0;
    const size_t per_tile = (size_t)J * NT * (vec ? 2 : 1);
    const size_t tiles = (n + per_tile - 1) / per_tile;
    Carver cv(ws);
    uint64_t *stat = cv.take<uint64_t>(tiles);
    uint32_t *ticket = cv.take<uint32_t>(1);
    QO_TRY(launch_fill_u32(ws, 0, cv.off / 4, st));
    if (vec)
        k_group_startpos<J, NT, true><<<(unsigned)tiles, NT, 0, st>>>(sorted, n, startPos, stat, ticket, d_num,
                                                                      (uint32_t)tiles, dev_ctl());
    else
        k_group_startpos<J, NT, false><<<(unsigned)tiles, NT, 0, st>>>(sorted, n, startPos, stat, ticket, d_num,
                                                                       (uint32_t)tiles, dev_ctl());
    QO_LAUNCHED();
    return QO_OK;
}

// Sized for the unaligned launch (one record per load), which has the most tiles.
size_t group_startpos_ws_bytes(size_t n) { return (n / ((size_t)kStartJ * kStartThreads) + 2) * 8 + 512; }

}  // namespace qo

using namespace qo;

extern "C" {

int qo_groupby_agg(const int32_t *key, const int32_t *val, size_t n, size_t group_capacity, int32_t *okey,
                   int64_t *ocount, int64_t *osum, int32_t *omin, int32_t *omax, uint64_t *d_numGroups,
                   qo_stream s) {
    return qo_groupby_agg_filtered(key, val, nullptr, n, 0, 0, group_capacity, okey, ocount, osum, omin, omax,
                                   d_numGroups, s);
}

int qo_groupby_agg_filtered(const int32_t *key, const int32_t *val, const int32_t *filt, size_t n, int32_t lo,
                            int32_t hi, size_t group_capacity, int32_t *okey, int64_t *ocount, int64_t *osum,
                            int32_t *omin, int32_t *omax, uint64_t *d_numGroups, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(d_numGroups && (n == 0 || (key && val)), "null pointer");
    QO_REQUIRE(group_capacity == 0 || (okey && ocount && osum && omin && omax), "null output column");
    QO_REQUIRE(group_capacity < (size_t(1) << 31), "group capacity too large");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, groupby_ws_bytes(group_capacity, n, filt != nullptr), &status);
    if (!ws) return status;
    return launch_groupby(key, val, n, group_capacity, okey, ocount, osum, omin, omax, d_numGroups, ws, st, filt, lo,
                          hi);
}

int qo_group_startpos(const qo_record *sorted, size_t n, int32_t *startPos, uint64_t *d_numGroups, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(d_numGroups && (n == 0 || (sorted && startPos)), "null pointer");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, group_startpos_ws_bytes(n), &status);
    if (!ws) return status;
    return launch_group_startpos(sorted, n, startPos, d_numGroups, ws, st);
}

int qo_aggregate_groups(const qo_record *sorted, size_t n, const int32_t *startPos, size_t numGroups, int op,
                        int32_t *out, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(op >= QO_REDUCE_SUM && op <= QO_REDUCE_AVERAGE, "unknown operator %d", op);
    if (numGroups == 0) return QO_OK;
    QO_REQUIRE(sorted && startPos && out && numGroups < 0xFFFFFFFFull, "bad aggregate arguments");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, aggregate_ws_bytes(numGroups, n), &status);
    if (!ws) return status;
    return launch_aggregate_groups(sorted, n, startPos, numGroups, op, out, ws, st);
}

}  // extern "C"
