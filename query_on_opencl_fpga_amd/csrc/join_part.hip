// Partitioned join build and probe for tables larger than one L2.
//
// A lookup into a table that misses L2 moves a 128-byte line from MALL/HBM to
// fetch an 8-byte slot: random lookups into a large table run at a small
// fraction of HBM bandwidth.  The home slot of a key is monotone in the table's
// 32-bit hash (qo_join.hpp), so the keys whose hash has top L bits f have their
// home slots in the slice [f*cap >> L, (f+1)*cap >> L) of the table; splitting
// the input by those bits gives each partition a contiguous table slice.
//
//  * Split: one pass over the input (filter fused) sends every tuple to one of
//    2^L1 buckets by its top L1 hash bits: each 8192-tuple tile is ranked in
//    LDS, reserves one range per bucket with one atomic add, and writes
//    contiguous runs (qo_split.hpp).  A second split by the next L2 bits is
//    used where slices must be smaller still.
//  * L2 probe (both table modes, slices up to a few MiB): one split into 2^L
//    buckets (L <= 7); the probe blocks of a bucket all run on one XCD (work
//    group i runs on XCD i % 8) and walk that XCD's buckets one after another,
//    so the XCD's L2 holds the current slice and a lookup is an L2 hit.
//  * LDS-slice probe (hash-mixed tables too large for the L2 probe): two splits
//    into slices that fit LDS (tens of KiB); one work-group per slice loads it
//    (plus a margin for chains that run past the slice end) and looks every
//    tuple up there; a chain that leaves the loaded range continues in global
//    memory.
//  * Build (hash-mixed table): one work-group per fine bucket inserts its
//    tuples into an empty LDS slice (compare-and-swap in LDS) and writes the
//    whole slice out, which also initialises the table.  Tuples whose chain
//    reaches the slice end are set aside and inserted into the finished table
//    with global compare-and-swap, as are the tuples of a full bucket.
// Matches are appended at a range reserved per 4096 tuples.  The table format
// is the one of the direct path: either path can probe a table built by
// either.  Which launched path runs is decided on the device from the table's
// mode; a path planned for the other mode returns at once.
#include <algorithm>

#include "qo_internal.hpp"
#include "qo_join.hpp"
#include "qo_split.hpp"

namespace qo {

constexpr uint32_t kSliceMargin = 256;              // slots past a slice end also kept in LDS by the probe
constexpr uint32_t kMaxSliceSlots = 8192;           // LDS slots of one work-group (64 KiB)
constexpr uint32_t kRangeSliceSlots = 16384;        // key-range build: LDS slots of one slice (128 KiB)
constexpr int kPairJ = 4;                           // pair quads per thread per probe step
constexpr uint32_t kPairStep = kPairJ * 1024;       // pairs per probe step

struct PartPlan {
    bool on = false;
    int L = 0, L2 = 0;       // fine bucket (table slice) bits; second split bits
    SplitPlan s;
    uint32_t slice_max = 0;  // slots of the largest slice
};

// slice_slots: the LDS slots of one slice for a build that keeps no probe
// margin (the key-range build); 0: the tuned slice of the hash-mixed paths.
static PartPlan part_plan(uint64_t cap, size_t n, uint64_t slice_slots = 0) {
    PartPlan p;
    const TuneParams tp = tune_get(TUNE_JOIN_PART, n);
    if (tp.param == PART_NEVER || n == 0 || n >= 0xFFFFFFFFull || cap < 2) return p;
    if (tp.param != PART_ALWAYS && tp.param != PART_LDS && (cap * 8 <= (uint64_t(4) << 20) || n < (size_t(1) << 20)))
        return p;
    // Slice target in slots (tune items = KiB per slice).
    const uint64_t target = slice_slots ? slice_slots - 2
                                        : std::min<uint64_t>(kMaxSliceSlots - kSliceMargin - 2,
                                                             (uint64_t)std::max(1, tp.items) * 128);
    const uint64_t room = slice_slots ? slice_slots : kMaxSliceSlots - kSliceMargin;
    int L = 1;
    while (L < 2 * kMaxSplitBits && (cap >> L) > target) ++L;
    if ((cap >> L) + 2 > room) return p;  // needs more than two splits
    p.on = true;
    p.L = L;
    p.s = split_plan(L, n);
    p.L2 = p.s.L2;
    p.slice_max = (uint32_t)((cap >> L) + 2);
    return p;
}

constexpr uint64_t kL2SliceBytes = uint64_t(1) << 20;     // key-range table bytes per L2 probe bucket
constexpr uint64_t kL2SliceMaxBytes = uint64_t(3) << 20;  // larger hash-mixed slices take the LDS path

// L2 probe plan: one split into 2^L buckets (L in [3, 7]) whose table slices
// stay in the L2 of the XCD that probes them.  It always serves a key-range
// table (at most the largest power of two <= slots, ~1 MiB slices); it serves
// a hash-mixed table (every slot) when that table's slices stay under
// kL2SliceMaxBytes, else the hash-mixed table takes the LDS-slice probe.
struct L2Plan {
    bool on = false;
    bool hash_fits = false;
    SplitPlan s;
    uint32_t nb = 0;   // buckets (kXcdRegions regions each)
    uint32_t tpb = 0;  // probe tiles per region (upper bound)
};

static L2Plan l2_plan(uint64_t slots, size_t n) {
    L2Plan r;
    const TuneParams tp = tune_get(TUNE_JOIN_PART, n);
    if (tp.param == PART_NEVER || tp.param == PART_LDS || n == 0 || n >= 0xFFFFFFFFull || slots < 2) return r;
    if (tp.param != PART_ALWAYS && (slots * 8 <= (uint64_t(4) << 20) || n < (size_t(1) << 20))) return r;
    uint64_t cap = 2;
    while (cap * 2 <= slots) cap *= 2;
    int L = 3;
    while (L < kMaxSplitBits && ((cap * 8) >> L) > kL2SliceBytes) ++L;
    r.on = true;
    r.hash_fits = ((slots * 8) >> L) <= kL2SliceMaxBytes;
    r.s = split_plan(L, n, kXcdRegions);
    r.nb = 1u << L;
    r.tpb = (uint32_t)((r.s.bcapA + kPairStep - 1) / kPairStep);  // per region
    return r;
}

// Key-range LDS route: the table slice of a bucket is held in LDS as 4-byte
// payloads plus presence bits (a key-range slot names its key), 2^15 slots per
// slice, so each lookup is an LDS read instead of an L1 -> L2 request.
constexpr int kRangeLdsSliceBits = 15;
constexpr uint32_t kRangeLdsSlice = 1u << kRangeLdsSliceBits;
constexpr int kRangeLdsMinCapBits = 21;  // smaller key-range tables keep the L2 probe
constexpr int kRangeLdsMaxBits = 9;      // at most 512 buckets: one split pass

// Which probe serves a table, decided on the device from its header: a
// key-range table of 2^21 .. 2^(15 + lds_bits) slots without (-1, -1) build
// tuples (one slot per key) takes the LDS route, other key-range tables the L2 route, hash-mixed
// tables the hash route.
// A key-range table whose capacity the LDS route serves (build and probe).
QO_DEV bool range_lds_sized(const JoinHeader *h, uint32_t lds_bits) {
    const uint32_t capbits = 32u - h->shift;
    return lds_bits != 0 && capbits >= (uint32_t)kRangeLdsMinCapBits &&
           capbits <= (uint32_t)kRangeLdsSliceBits + lds_bits;
}

enum JoinRoute : uint32_t { kRouteHash = 0, kRouteRangeL2 = 1, kRouteRangeLds = 2 };
struct RouteGate {
    const JoinHeader *hdr = nullptr;  // null: always runs
    uint32_t mask = 0;                // bit r set: runs for route r
    uint32_t lds_bits = 0;            // 0: no table takes the LDS route
    QO_DEV static uint32_t route(const JoinHeader *h, uint32_t lds_bits) {
        if (h->mode != kJoinKeyRange) return kRouteHash;
        const bool lds = (h->sentinel_count | h->has_dups) == 0 && range_lds_sized(h, lds_bits);
        return lds ? kRouteRangeLds : kRouteRangeL2;
    }
    QO_DEV bool off() const { return hdr && !((mask >> route(hdr, lds_bits)) & 1u); }
};

// Key-range builds: `lds` set, the LDS-route build of the tables it serves by
// capacity; clear, the sliced build of the other key-range tables.
struct RangeBuildGate {
    JoinHeader *hdr;
    uint32_t lds_bits;
    bool lds;
    QO_DEV bool off() const { return hdr->mode != kJoinKeyRange || range_lds_sized(hdr, lds_bits) != lds; }
};

// LDS of the key-range slice probe: the slice's payloads and presence bits,
// and one step's matches staged as (slot in the slice, s_payload).
// The probe's work-groups are 16 waves, one pair quad per thread per step of
// kPairStep pairs: the slice takes most of the LDS, so a CU runs one
// work-group, and a wave waiting for its stores or loads leaves 15 others.
constexpr int kLdsProbeThreads = 1024;
constexpr int kLdsProbeJ = kPairStep / 4 / kLdsProbeThreads;
static_assert(kLdsProbeJ * 4 * kLdsProbeThreads == (int)kPairStep, "probe step");
struct RangeLdsProbeLds {
    uint32_t pay[kRangeLdsSlice];
    uint32_t bits[kRangeLdsSlice / 32];
    uint16_t loc[kPairStep];
    int32_t spay[kPairStep];
    uint32_t tot[tile_offsets_words(kLdsProbeJ, kLdsProbeThreads / kWave)];
    uint64_t excl;
};
// LDS of the key-range slice build: the slice's payloads and presence bits.
struct RangeLdsBuildLds {
    uint32_t pay[kRangeLdsSlice];
    uint32_t bits[kRangeLdsSlice / 32];
    uint32_t occ, sentinel;
};

// Host plan of the LDS route.  The table's capacity is known only on the
// device (a power of two >= the key range, <= the slots allocated), so the
// regions are sized for the largest capacity that takes the route, 2^(15 +
// bits) slots: 2^bits buckets of 8 regions (one per XCD) of bcap pairs; a
// table of 2^(15 + L) slots, L < bits, has 2^L buckets whose regions are
// 2^(bits - L) times larger, in the same memory.
struct RangeLdsPlan {
    bool on = false;
    uint32_t bits = 0;
    uint64_t bcap = 0;
    uint32_t gpb = 1;  // probe work-groups per bucket at 2^bits buckets (1, 2, 4 or 8)
};

static RangeLdsPlan range_lds_plan(uint64_t slots, size_t n) {
    RangeLdsPlan r;
    const TuneParams tp = tune_get(TUNE_JOIN_PART, n);
    if (tp.param == PART_NEVER || tp.param == PART_LDS || tp.param == PART_L2 || n == 0 || n >= 0xFFFFFFFFull)
        return r;
    if (tp.param != PART_ALWAYS && (slots * 8 <= (uint64_t(4) << 20) || n < (size_t(1) << 20))) return r;
    if (sizeof(RangeLdsProbeLds) > (size_t)props().lds_per_block) return r;
    int capbits = 0;  // the largest capacity a key-range table can take: 2^capbits <= slots
    while (capbits < 31 && (uint64_t(2) << capbits) <= slots) ++capbits;
    if (capbits < kRangeLdsMinCapBits) return r;
    r.on = true;
    r.bits = (uint32_t)std::min(capbits - kRangeLdsSliceBits, kRangeLdsMaxBits);
    const uint64_t regions = (uint64_t(1) << r.bits) * kXcdRegions;
    const uint64_t per = (n + regions - 1) / regions;
    r.bcap = (per + per / 16 + 1024 + 31) / 32 * 32;
    const uint32_t want = ((uint32_t)props().cu_count * 2) >> r.bits;
    while (r.gpb < kXcdRegions && r.gpb < want) r.gpb *= 2;
    return r;
}

struct RangeLdsBufs {
    uint32_t *ctr = nullptr;  // one cursor per region, kCtrStride apart
    uint64_t *bucket = nullptr;
    size_t ctr_words = 0;
};

static RangeLdsBufs range_lds_carve(const RangeLdsPlan &r, Carver &cv) {
    RangeLdsBufs b;
    const size_t regions = (size_t(1) << r.bits) * kXcdRegions;
    b.ctr_words = regions * kCtrStride;
    b.ctr = cv.take<uint32_t>(b.ctr_words);
    b.bucket = cv.take<uint64_t>(regions * r.bcap);
    return b;
}

// The probe's plans: the L2 plan, and the LDS-slice plan only for a
// hash-mixed table the L2 plan does not serve.
static void probe_plans(uint64_t slots, size_t n, L2Plan &r, PartPlan &p) {
    r = l2_plan(slots, n);
    p = (r.on && r.hash_fits) ? PartPlan{} : part_plan(slots, n);
}

// Workspace carve-up (same order for sizing and binding).
struct PartBufs {
    SplitBufs sb;
    unsigned long long *spill_n = nullptr;
    uint64_t *spill = nullptr;  // build only: nr entries
};

static PartBufs part_carve(const PartPlan &p, size_t n, bool build, Carver &cv) {
    PartBufs b;
    b.sb = split_carve(p.s, cv);
    if (build) {
        b.spill_n = cv.take<unsigned long long>(1);
        b.spill = cv.take<uint64_t>(n);
    }
    return b;
}

void join_part_strategy(size_t table_bytes, size_t n, int mode, int *partitioned, int *bits, int *splits) {
    L2Plan r;
    PartPlan p;
    probe_plans(join_table_slots(table_bytes), n, r, p);
    *partitioned = *bits = *splits = 0;
    const RangeLdsPlan rl = range_lds_plan(join_table_slots(table_bytes), n);
    if (rl.on && mode == (int)kJoinKeyRange) {  // bits: for the largest key-range table of these slots
        *partitioned = 2;
        *bits = (int)rl.bits;
        *splits = 1;
    } else if (r.on && (mode == (int)kJoinKeyRange || r.hash_fits)) {
        *partitioned = 1;
        *bits = r.s.L;
        *splits = 1;
    } else if (mode != (int)kJoinKeyRange && p.on) {
        *partitioned = 1;
        *bits = p.L;
        *splits = 2;  // the LDS-slice probe serves only tables too large for one split
    }
}

size_t join_part_ws_bytes(size_t table_bytes, size_t n, bool build) {
    const uint64_t slots = join_table_slots(table_bytes);
    L2Plan r;
    PartPlan p;
    if (build) p = part_plan(slots, n);
    else probe_plans(slots, n, r, p);
    size_t bytes = 0;
    if (p.on) {
        Carver cv(nullptr);
        part_carve(p, n, build, cv);
        bytes = cv.off + 256;
    }
    if (build) {  // the key-range builds' plans share the workspace
        const PartPlan pr = part_plan(slots, n, kRangeSliceSlots);
        if (pr.on) {
            Carver cv(nullptr);
            part_carve(pr, n, true, cv);
            bytes = std::max(bytes, cv.off + 256);
            const RangeLdsPlan rl = range_lds_plan(slots, n);
            if (rl.on) {
                Carver cl(nullptr);
                range_lds_carve(rl, cl);
                cl.take<unsigned long long>(1);
                cl.take<uint64_t>(n);
                bytes = std::max(bytes, cl.off + 256);
            }
        }
    }
    if (r.on) {
        // The probes share the workspace: only one of them runs.
        Carver cv(nullptr);
        split_carve(r.s, cv);
        bytes = std::max(bytes, cv.off + 256);
    }
    const RangeLdsPlan rl = build ? RangeLdsPlan{} : range_lds_plan(slots, n);
    if (rl.on) {
        Carver cv(nullptr);
        range_lds_carve(rl, cv);
        bytes = std::max(bytes, cv.off + 256);
    }
    return bytes;
}

// Tuples that find their bucket full.  Probe: looked up and emitted at once.
struct ProbeOverflow {
    const JoinHeader *hdr;
    unsigned long long *counter;
    SoaEmit emit;
    QO_DEV void operator()(uint64_t v) const {
        const uint64_t *slots = reinterpret_cast<const uint64_t *>(hdr + 1);
        probe_one(slots, hdr->cap, JoinHash::of(hdr), hdr->sentinel_count, (int32_t)(uint32_t)v, (int32_t)(v >> 32),
                  counter, emit);
    }
};

// ---------------------------------------------------------------------------
// Per-slice probe and build in LDS.
// ---------------------------------------------------------------------------
QO_DEV void slice_bounds(uint32_t f, int L, uint64_t cap, uint64_t &lo, uint64_t &hi) {
    lo = ((uint64_t)f * cap) >> L;
    hi = ((uint64_t)(f + 1) * cap) >> L;
}

// Copies slots [lo, lo + len) of the table to LDS with every load of the
// work-group in flight at once (16-byte loads; `lo` is even, the slot array
// 16-byte aligned).  len <= kMaxSliceSlots.
QO_DEV void slice_load(uint64_t *s_slots, const uint64_t *__restrict__ slots, uint64_t lo, uint32_t len) {
    constexpr int kIters = kMaxSliceSlots / 2 / 256;
    const uint4 *src = reinterpret_cast<const uint4 *>(slots + lo);
    const uint32_t npairs = len / 2;
    uint4 q[kIters];
#pragma unroll
    for (int u = 0; u < kIters; ++u) {
        const uint32_t i = (uint32_t)u * 256 + threadIdx.x;
        if (i < npairs) q[u] = ld_stream(src + i);
    }
#pragma unroll
    for (int u = 0; u < kIters; ++u) {
        const uint32_t i = (uint32_t)u * 256 + threadIdx.x;
        if (i < npairs) reinterpret_cast<uint4 *>(s_slots)[i] = q[u];
    }
    if ((len & 1u) && threadIdx.x == 0) s_slots[len - 1] = slots[lo + len - 1];
}

// Finds `key` (of a hash-mixed table) starting in the LDS copy of slots
// [lo, lo + len), continuing in global memory if the chain leaves it.  On a hit
// sv is the matching slot.
QO_DEV bool slice_find(const uint64_t *s_slots, uint32_t len, uint64_t lo, const uint64_t *__restrict__ slots,
                       uint64_t cap, int32_t key, uint64_t &sv) {
    uint64_t h = join_home(hash32((uint32_t)key), cap);
    uint64_t probes = 0;
    if (h >= lo && h < lo + len) {
        for (uint32_t i = (uint32_t)(h - lo); i < len; ++i, ++probes) {
            const uint64_t s = s_slots[i];
            if (s == kEmptySlot) return false;
            if ((uint32_t)s == (uint32_t)key) {
                sv = s;
                return true;
            }
        }
        h = lo + len == cap ? 0 : lo + len;
    }
    for (; probes < cap; ++probes) {
        const uint64_t s = slots[h];
        if (s == kEmptySlot) return false;
        if ((uint32_t)s == (uint32_t)key) {
            sv = s;
            return true;
        }
        h = join_next(h, cap);
    }
    return false;
}

// Probe of fine bucket f = blockIdx.x of a hash-mixed table (does nothing for
// a key-range table).  Dynamic LDS: (slice_max + margin) slots.
__global__ __launch_bounds__(256) void k_slice_probe(const uint64_t *__restrict__ bucket, uint64_t bcap,
                                                     const uint32_t *ctr, int L, const JoinHeader *__restrict__ hdr,
                                                     unsigned long long *counter, SoaEmit emit) {
    extern __shared__ __attribute__((aligned(16))) uint64_t s_slots[];
    __shared__ uint32_t s_tot[tile_offsets_words(kPairJ, 4)];
    __shared__ uint64_t s_excl;
    if (hdr->mode != kJoinHashMix) return;
    const uint32_t f = blockIdx.x;
    uint64_t n = ctr[f * kCtrStride];
    if (n > bcap) n = bcap;
    if (n == 0) return;  // whole block: no barrier has been reached
    const uint64_t *slots = reinterpret_cast<const uint64_t *>(hdr + 1);
    const uint64_t cap = hdr->cap;
    const uint64_t *src = bucket + (size_t)f * bcap;
    if ((hdr->has_dups | hdr->sentinel_count) != 0) {
        // Duplicate keys or (-1, -1) build tuples: whole chains, in global memory.
        const uint32_t sentinel = hdr->sentinel_count;
        for (uint64_t e0 = 0; e0 < n; e0 += kPairStep) {
            const uint32_t m = (uint32_t)(n - e0 < kPairStep ? n - e0 : kPairStep);
            int32_t key[kPairJ][4], sp[kPairJ][4];
            uint32_t pass[kPairJ];
            load_pairs<kPairJ>(src + e0, m, key, sp, pass);
            lookup_emit<kPairJ>(slots, cap, JoinHash::of(hdr), true, sentinel, key, sp, pass, counter, emit, s_tot,
                                &s_excl);
        }
        return;
    }
    uint64_t lo, hi;
    slice_bounds(f, L, cap, lo, hi);
    lo &= ~uint64_t(1);  // even start: 16-byte loads
    const uint64_t end = hi + kSliceMargin < cap ? hi + kSliceMargin : cap;
    const uint32_t len = (uint32_t)(end - lo);
    slice_load(s_slots, slots, lo, len);
    __syncthreads();
    for (uint64_t e0 = 0; e0 < n; e0 += kPairStep) {
        const uint32_t m = (uint32_t)(n - e0 < kPairStep ? n - e0 : kPairStep);
        int32_t key[kPairJ][4], sp[kPairJ][4];
        uint32_t pass[kPairJ];
        load_pairs<kPairJ>(src + e0, m, key, sp, pass);
        uint64_t sv[kPairJ][4];
        uint32_t hit[kPairJ], c[kPairJ], ex[kPairJ];
#pragma unroll
        for (int j = 0; j < kPairJ; ++j) {
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                sv[j][k] = 0;
                if ((pass[j] >> k) & 1u)
                    if (slice_find(s_slots, len, lo, slots, cap, key[j][k], sv[j][k])) hit[j] |= 1u << k;
            }
            c[j] = __popc(hit[j]);
        }
        const uint32_t total = tile_offsets<kPairJ, 4>(c, ex, s_tot);
        if (threadIdx.x == 0) s_excl = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
        __syncthreads();
        const uint64_t base = s_excl;
#pragma unroll
        for (int j = 0; j < kPairJ; ++j) {
            uint64_t pos = base + ex[j];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if ((hit[j] >> k) & 1u) emit(pos++, key[j][k], (int32_t)(sv[j][k] >> 32), sp[j][k]);
        }
        __syncthreads();  // s_tot and s_excl are reused by the next step
    }
}

// Build of fine bucket f = blockIdx.x of a hash-mixed table (nothing when the
// build chose a key-range table): insert into an empty LDS slice, write the
// whole slice [lo, hi) out.  Dynamic LDS: slice_max slots.
__global__ __launch_bounds__(256) void k_slice_build(const uint64_t *__restrict__ bucket, uint64_t bcap,
                                                     const uint32_t *ctr, int L, JoinHeader *hdr, uint64_t *spill,
                                                     unsigned long long *spill_n) {
    extern __shared__ __attribute__((aligned(16))) uint64_t s_slots[];
    __shared__ uint32_t s_dup, s_sentinel;
    if (hdr->mode != kJoinHashMix) return;
    const uint32_t f = blockIdx.x;
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    const uint64_t cap = hdr->cap;
    uint64_t lo, hi;
    slice_bounds(f, L, cap, lo, hi);
    const uint32_t len = (uint32_t)(hi - lo);
    for (uint32_t i = threadIdx.x; i < len; i += 256) s_slots[i] = kEmptySlot;
    if (threadIdx.x == 0) s_dup = s_sentinel = 0;
    __syncthreads();
    uint64_t n = ctr[f * kCtrStride];
    if (n > bcap) n = bcap;
    const uint64_t *src = bucket + (size_t)f * bcap;
    for (uint64_t i = threadIdx.x; i < n; i += 256) {
        const uint64_t v = ld_stream(src + i);
        if (v == kEmptySlot) {  // the (-1, -1) build tuple is counted, not stored
            atomicAdd(&s_sentinel, 1u);
            continue;
        }
        const int32_t key = (int32_t)(uint32_t)v;
        const uint64_t h = join_home(hash32((uint32_t)key), cap);
        bool placed = false;
        if (h >= lo && h < hi) {
            for (uint32_t j = (uint32_t)(h - lo); j < len; ++j) {
                const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long *>(&s_slots[j]),
                                                          (unsigned long long)kEmptySlot, (unsigned long long)v);
                if (prev == kEmptySlot) {
                    placed = true;
                    break;
                }
                if ((uint32_t)prev == (uint32_t)key) s_dup = 1u;
            }
        }
        if (!placed) spill[atomicAdd(spill_n, 1ull)] = v;  // chain reached the slice end
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < len; i += 256) st_stream(slots + lo + i, s_slots[i]);
    if (threadIdx.x == 0) {
        if (s_dup) atomicExch(&hdr->has_dups, 1u);
        if (s_sentinel) atomicAdd(&hdr->sentinel_count, s_sentinel);
    }
}

// Tuples set aside by the build, inserted into the finished table.
__global__ void k_spill_insert(const uint64_t *spill, const unsigned long long *spill_n, JoinHeader *hdr) {
    if (hdr->mode != kJoinHashMix) return;
    const uint64_t n = *spill_n;
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    const uint64_t cap = hdr->cap;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = spill[i];
        join_insert(slots, cap, hdr, (int32_t)(uint32_t)v, (int32_t)(v >> 32));
    }
}

// Key-range table, partitioned build: slice f of the table is assembled in LDS
// from bucket f of R (a split by the table's own hash, whose top bits are the
// top bits of key - kmin) with a plain store at slot key - kmin, then written
// out whole, empty slots included, so the table needs no fill and every line is
// written once.  A duplicate key overwrites its slot, and the occupied count
// (added to the header) then falls short of the tuples, as the key-range
// commit expects; a tuple whose slot lies outside the slice (never, unless the
// table has fewer slots than buckets) goes to the spill list.
__global__ __launch_bounds__(256) void k_range_slice_build(const uint64_t *__restrict__ bucket, uint64_t bcap,
                                                           const uint32_t *ctr, int L, JoinHeader *hdr,
                                                           uint64_t *spill, unsigned long long *spill_n,
                                                           uint32_t lds_bits) {
    extern __shared__ __attribute__((aligned(16))) uint64_t s_slots[];
    __shared__ uint32_t s_occ, s_sentinel;
    if (RangeBuildGate{hdr, lds_bits, false}.off()) return;  // the LDS-route build takes this table
    const uint32_t f = blockIdx.x;
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    const uint32_t kmin = (uint32_t)hdr->kmin;
    uint64_t lo, hi;
    slice_bounds(f, L, hdr->cap, lo, hi);
    const uint32_t len = (uint32_t)(hi - lo);
    for (uint32_t i = threadIdx.x; i < len; i += 256) s_slots[i] = kEmptySlot;
    if (threadIdx.x == 0) s_occ = s_sentinel = 0;
    __syncthreads();
    uint64_t n = ctr[f * kCtrStride];
    if (n > bcap) n = bcap;
    const uint64_t *src = bucket + (size_t)f * bcap;  // 16-byte aligned: bcap is a multiple of 32
    auto put = [&](uint64_t v) {
        if (v == kEmptySlot) {  // the (-1, -1) build tuple is counted, not stored
            atomicAdd(&s_sentinel, 1u);
            return;
        }
        const uint64_t slot = (uint32_t)v - kmin;
        if (slot >= lo && slot < hi) s_slots[slot - lo] = v;
        else spill[atomicAdd(spill_n, 1ull)] = v;
    };
    // Two tuples per 16-byte load, then the odd one.
    for (uint64_t i = threadIdx.x; 2 * i + 1 < n; i += 256) {
        const uint4 q = ld_stream(reinterpret_cast<const uint4 *>(src) + i);
        put(((uint64_t)q.y << 32) | q.x);
        put(((uint64_t)q.w << 32) | q.z);
    }
    if ((n & 1) && threadIdx.x == 0) put(ld_stream(src + n - 1));
    __syncthreads();
    uint32_t occ = 0;
    if (((lo | len) & 1) == 0) {
        // Two slots per 16-byte store (the slot array is 16-byte aligned).
        uint4 *dst = reinterpret_cast<uint4 *>(slots + lo);
        for (uint32_t i = threadIdx.x; 2 * i < len; i += 256) {
            const uint64_t a = s_slots[2 * i], b = s_slots[2 * i + 1];
            occ += (a != kEmptySlot) + (b != kEmptySlot);
            dst[i] = uint4{(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32)};
        }
    } else {
        for (uint32_t i = threadIdx.x; i < len; i += 256) {
            const uint64_t e = s_slots[i];
            occ += e != kEmptySlot;
            st_stream(slots + lo + i, e);
        }
    }
    occ = wave_sum(occ);
    if (lane_id() == 0 && occ) atomicAdd(&s_occ, occ);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_occ) atomicAdd(&hdr->occupied, (unsigned long long)s_occ);
        if (s_sentinel) atomicAdd(&hdr->sentinel_count, s_sentinel);
    }
}

// Key-range tuples set aside by the split or the slice build, stored at their
// slot in the finished table; a slot already taken (a duplicate key) is not
// counted as occupied, so the commit falls back to the hash-mixed table.
__global__ void k_range_spill(const uint64_t *spill, const unsigned long long *spill_n, JoinHeader *hdr) {
    if (hdr->mode != kJoinKeyRange) return;
    const uint64_t n = *spill_n;
    uint64_t *slots = reinterpret_cast<uint64_t *>(hdr + 1);
    const uint32_t kmin = (uint32_t)hdr->kmin;
    uint32_t occ = 0, sentinel = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = spill[i];
        if (v == kEmptySlot) {
            ++sentinel;
            continue;
        }
        const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long *>(slots + ((uint32_t)v - kmin)),
                                        (unsigned long long)kEmptySlot, (unsigned long long)v);
        occ += prev == kEmptySlot;
    }
    occ = wave_sum(occ);
    sentinel = wave_sum(sentinel);
    if (lane_id() == 0) {
        if (occ) atomicAdd(&hdr->occupied, (unsigned long long)occ);
        if (sentinel) atomicAdd(&hdr->sentinel_count, sentinel);
    }
}

// L2 probe from the S buckets (either table mode), each bucket kXcdRegions regions
// (XcdLayout).  Block i works on XCD x = i % 8: with q = i / 8 it takes bucket
// b = (q / (kXcdRegions * tpb)) * 8 + x, region (q / tpb) % kXcdRegions and its tile
// q % tpb, so each XCD walks its own buckets one after another, every region
// of a bucket in turn, and its L2 holds the current table slice.
__global__ __launch_bounds__(256) void k_l2_probe(const uint64_t *__restrict__ bucket, uint64_t bcap,
                                                  const uint32_t *ctr, uint32_t nb, uint32_t tpb,
                                                  const JoinHeader *__restrict__ hdr, unsigned long long *counter,
                                                  SoaEmit emit, RouteGate gate) {
    __shared__ uint32_t s_tot[tile_offsets_words(kPairJ, 4)];
    __shared__ uint64_t s_excl;
    __shared__ JoinStageT<kPairJ * 1024> stage;
    if (gate.off()) return;
    const uint32_t x = blockIdx.x & 7u, q = blockIdx.x >> 3;
    const uint32_t per_bucket = kXcdRegions * tpb;
    const uint32_t b = (q / per_bucket) * 8u + x, rem = q % per_bucket, t = rem % tpb;
    if (b >= nb) return;
    const uint32_t reg = b * kXcdRegions + rem / tpb;
    uint64_t n = ctr[reg * kCtrStride];
    if (n > bcap) n = bcap;
    const uint64_t e0 = (uint64_t)t * kPairStep;
    if (e0 >= n) return;  // whole block: no barrier has been reached
    const uint32_t m = (uint32_t)(n - e0 < kPairStep ? n - e0 : kPairStep);
    int32_t key[kPairJ][4], sp[kPairJ][4];
    uint32_t pass[kPairJ];
    load_pairs<kPairJ>(bucket + (size_t)reg * bcap + e0, m, key, sp, pass);
    const uint64_t *slots = reinterpret_cast<const uint64_t *>(hdr + 1);
    lookup_emit<kPairJ>(slots, hdr->cap, JoinHash::of(hdr), (hdr->has_dups | hdr->sentinel_count) != 0,
                        hdr->sentinel_count, key, sp, pass, counter, emit, s_tot, &s_excl, &stage);
}

// Split of the S tuples (filter fused) or the R tuples for the LDS route: 2^L
// buckets by the top L bits of key - kmin, L = log2(cap) - 15 read from the table, so bucket f
// holds the keys of slice f; 8 regions per bucket, one per XCD (XcdLayout), of
// bcap << (lds_bits - L) pairs each.  Keys outside the table's range cannot
// match and are dropped here.
template <bool HAS_FILTER, class Overflow, class G>
__global__ __launch_bounds__(kSplitThreads) void k_range_lds_split(const int32_t *__restrict__ kcol,
                                                                   const int32_t *__restrict__ pcol,
                                                                   const int32_t *__restrict__ fcol, size_t n,
                                                                   int32_t flo, int32_t fhi, bool vec, uint64_t bcap,
                                                                   uint64_t *__restrict__ bucket, uint32_t *ctr,
                                                                   Overflow overflow, G gate) {
    __shared__ SplitLdsT<(1u << kRangeLdsMaxBits)> sh;
    if (gate.off()) return;
    const JoinHash hs = JoinHash::of(gate.hdr);
    const int L = (int)(32u - hs.shift) - kRangeLdsSliceBits;  // in [6, lds_bits] on this route
    int32_t key[kSplitJ][4], pay[kSplitJ][4];
    uint32_t pass[kSplitJ];
    load_s_tile<kSplitJ, HAS_FILTER, kSplitThreads>(kcol, pcol, fcol, n, flo, fhi, vec, blockIdx.x, key, pay, pass);
#pragma unroll
    for (int j = 0; j < kSplitJ; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t h;
            if (!hs(key[j][k], h)) pass[j] &= ~(1u << k);
        }
    split_block<kSplitJ>(key, pay, pass, hs, 32 - L, 1u << L, 0, bucket, XcdLayout{bcap << (gate.lds_bits - L)}, ctr,
                         sh, overflow);
}

// Build of slice f = blockIdx.x of the LDS route: every tuple of bucket f (its
// 8 regions) sets its presence bit and payload in an LDS copy of the slice
// (an LDS atomic-or finds a second tuple of one key, which is not counted as
// occupied, so the commit falls back to the hash-mixed table), and the slice
// is written out whole with 16-byte stores, key = kmin + slot.  (-1, -1)
// tuples are counted, not stored.  Dynamic LDS: sizeof(RangeLdsBuildLds).
__global__ __launch_bounds__(kLdsProbeThreads) void k_range_lds_build(const uint64_t *__restrict__ bucket,
                                                                      uint64_t bcap, const uint32_t *ctr,
                                                                      RangeBuildGate gate, uint64_t *spill,
                                                                      unsigned long long *spill_n) {
    extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
    RangeLdsBuildLds &s = *reinterpret_cast<RangeLdsBuildLds *>(s_raw);
    if (gate.off()) return;
    JoinHeader *hdr = gate.hdr;
    const int L = (int)(32u - hdr->shift) - kRangeLdsSliceBits;
    const uint64_t rcap = bcap << (gate.lds_bits - L);
    const uint32_t f = blockIdx.x;
    if (f >= (1u << L)) return;
    for (uint32_t i = threadIdx.x; i < kRangeLdsSlice / 32; i += kLdsProbeThreads) s.bits[i] = 0;
    if (threadIdx.x == 0) s.occ = s.sentinel = 0;
    __syncthreads();
    const uint32_t kbase = (uint32_t)hdr->kmin + (f << kRangeLdsSliceBits);
    uint32_t occ = 0, sentinel = 0;
    auto put = [&](uint32_t key, uint32_t pay) {
        const uint64_t v = ((uint64_t)pay << 32) | key;
        if (v == kEmptySlot) {
            ++sentinel;
            return;
        }
        const uint32_t l = key - kbase;
        if (l >= kRangeLdsSlice) {  // not this slice's key (a split never sends one): stored by k_range_spill
            spill[atomicAdd(spill_n, 1ull)] = v;
            return;
        }
        const uint32_t bit = 1u << (l & 31u);
        if (!(atomicOr(&s.bits[l >> 5], bit) & bit)) {
            s.pay[l] = pay;
            ++occ;
        }
    };
    for (uint32_t r = 0; r < kXcdRegions; ++r) {
        const size_t reg = (size_t)f * kXcdRegions + r;
        uint64_t n = ctr[reg * kCtrStride];
        if (n > rcap) n = rcap;
        const uint64_t *src = bucket + reg * rcap;  // 16-byte aligned: bcap is a multiple of 32
        for (uint64_t i = threadIdx.x; 2 * i + 1 < n; i += kLdsProbeThreads) {
            const uint4 q = ld_stream(reinterpret_cast<const uint4 *>(src) + i);
            put(q.x, q.y);
            put(q.z, q.w);
        }
        if ((n & 1) && threadIdx.x == 0) {
            const uint64_t v = ld_stream(src + n - 1);
            put((uint32_t)v, (uint32_t)(v >> 32));
        }
    }
    __syncthreads();
    uint4 *dst = reinterpret_cast<uint4 *>(reinterpret_cast<uint64_t *>(hdr + 1) + ((uint64_t)f << kRangeLdsSliceBits));
    for (uint32_t i = threadIdx.x; i < kRangeLdsSlice / 2; i += kLdsProbeThreads) {
        const uint32_t w = s.bits[i >> 4], sh = (2u * i) & 31u;
        const bool b0 = (w >> sh) & 1u, b1 = (w >> (sh + 1)) & 1u;
        dst[i] = uint4{b0 ? kbase + 2 * i : ~0u, b0 ? s.pay[2 * i] : ~0u, b1 ? kbase + 2 * i + 1 : ~0u,
                       b1 ? s.pay[2 * i + 1] : ~0u};
    }
    occ = wave_sum(occ);
    sentinel = wave_sum(sentinel);
    if (lane_id() == 0) {
        if (occ) atomicAdd(&s.occ, occ);
        if (sentinel) atomicAdd(&s.sentinel, sentinel);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s.occ) atomicAdd(&hdr->occupied, (unsigned long long)s.occ);
        if (s.sentinel) atomicAdd(&hdr->sentinel_count, s.sentinel);
    }
}

// Bit j of the low 16 bits of x to bit 2j.
QO_DEV uint32_t spread_bits16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// Probe of the LDS route: work-group (f, g) copies slice f of the table (slots
// [f << 15, (f + 1) << 15)) to LDS as payloads plus presence bits and looks up
// the pairs of its share of bucket f's 8 regions there.  A hit is staged as its
// slot in the slice and its s_payload; the rows are written out with the key
// and r_payload taken from the slot, one step later, while the next step's
// output range is being reserved.  The next two steps of 4096 pairs are loaded
// before the current one is looked up.  Dynamic LDS: sizeof(RangeLdsProbeLds).
__global__ __launch_bounds__(kLdsProbeThreads) void k_range_lds_probe(const uint64_t *__restrict__ bucket, uint64_t bcap,
                                                         const uint32_t *ctr, RouteGate gate,
                                                         unsigned long long *counter, SoaEmit emit) {
    extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
    RangeLdsProbeLds &s = *reinterpret_cast<RangeLdsProbeLds *>(s_raw);
    if (gate.off()) return;
    const JoinHeader *hdr = gate.hdr;
    const int L = (int)(32u - hdr->shift) - kRangeLdsSliceBits;
    const uint64_t rcap = bcap << (gate.lds_bits - L);  // pairs per region at this table's 2^L buckets
    // The grid is (1, 2, 4 or 8) << lds_bits work-groups: 2^L buckets of gpb each.
    const uint32_t gpb = std::min<uint32_t>(kXcdRegions, gridDim.x >> L);
    const uint32_t f = blockIdx.x / gpb, g = blockIdx.x % gpb;
    if (f >= (1u << L)) return;
    const uint32_t rpb = kXcdRegions / gpb, r0 = f * kXcdRegions + g * rpb;
    auto region_n = [&](uint32_t i) -> uint32_t {
        const uint64_t c = ctr[(size_t)(r0 + i) * kCtrStride];
        return (uint32_t)(c < rcap ? c : rcap);
    };
    uint32_t ri = 0, e = 0, rn = region_n(0);
    auto seek = [&](uint32_t &i, uint32_t &pos, uint32_t &cnt) {
        while (i < rpb && pos >= cnt) {
            ++i;
            pos = 0;
            cnt = i < rpb ? region_n(i) : 0u;
        }
    };
    seek(ri, e, rn);
    if (ri >= rpb) return;  // no pairs (whole block): no barrier has been reached

    // The slice: 16-byte loads of two slots, 16 in flight per thread.  Each
    // wave's 64 loads cover 128 consecutive slots, whose presence bits are 4
    // words built from two ballots (even and odd slots).
    const uint4 *src4 = reinterpret_cast<const uint4 *>(reinterpret_cast<const uint64_t *>(hdr + 1) +
                                                        ((uint64_t)f << kRangeLdsSliceBits));
    const unsigned lane = lane_id();
    constexpr uint32_t kQuadRounds = kRangeLdsSlice / 2 / kLdsProbeThreads;
    constexpr uint32_t kBatch = kQuadRounds < 16 ? kQuadRounds : 16;
    for (uint32_t u0 = 0; u0 < kQuadRounds; u0 += kBatch) {
        uint4 q[kBatch];
#pragma unroll
        for (uint32_t v = 0; v < kBatch; ++v) q[v] = ld_stream(src4 + (u0 + v) * kLdsProbeThreads + threadIdx.x);
#pragma unroll
        for (uint32_t v = 0; v < kBatch; ++v) {
            const uint32_t i = (u0 + v) * kLdsProbeThreads + threadIdx.x;  // slots 2i, 2i + 1
            reinterpret_cast<uint2 *>(s.pay)[i] = uint2{q[v].y, q[v].w};
            const uint64_t m0 = ballot(!(q[v].x == ~0u && q[v].y == ~0u));
            const uint64_t m1 = ballot(!(q[v].z == ~0u && q[v].w == ~0u));
            if (lane < 4)
                s.bits[(i - lane) / 16 + lane] = spread_bits16((uint32_t)(m0 >> (16 * lane))) |
                                                 (spread_bits16((uint32_t)(m1 >> (16 * lane))) << 1);
        }
    }
    __syncthreads();

    const uint32_t kbase = (uint32_t)hdr->kmin + (f << kRangeLdsSliceBits);
    auto load = [&](uint32_t i, uint32_t pos, uint32_t cnt, int32_t (&k)[kLdsProbeJ][4], int32_t (&p)[kLdsProbeJ][4],
                    uint32_t (&ps)[kLdsProbeJ]) {
        load_pairs<kLdsProbeJ, kLdsProbeThreads>(bucket + (size_t)(r0 + i) * rcap + pos, cnt - pos < kPairStep ? cnt - pos : kPairStep, k,
                           p, ps);
    };
    auto row = [&](uint32_t i, int32_t &k, int32_t &r, int32_t &q) {
        const uint32_t l = s.loc[i];
        k = (int32_t)(kbase + l);
        r = (int32_t)s.pay[l];
        q = s.spay[i];
    };
    uint32_t staged = 0;  // rows of the previous step in the stage, written out at s.excl
    // Steps t (key), t + 1 (key2) and t + 2 (key3) in registers: the loads of
    // two steps are in flight while one is looked up.
    int32_t key[kLdsProbeJ][4], sp[kLdsProbeJ][4], key2[kLdsProbeJ][4], sp2[kLdsProbeJ][4];
    uint32_t pass[kLdsProbeJ], pass2[kLdsProbeJ];
    load(ri, e, rn, key, sp, pass);
    uint32_t ni = ri, ne = e + kPairStep, nn = rn;  // step t + 1
    seek(ni, ne, nn);
    if (ni < rpb) load(ni, ne, nn, key2, sp2, pass2);
    for (;;) {
        const bool more = ni < rpb;
        uint32_t mi = ni, me = ne + kPairStep, mn = nn;  // step t + 2
        if (more) seek(mi, me, mn);
        const bool more2 = more && mi < rpb;
        int32_t key3[kLdsProbeJ][4], sp3[kLdsProbeJ][4];
        uint32_t pass3[kLdsProbeJ];
        if (more2) load(mi, me, mn, key3, sp3, pass3);
        uint32_t hit[kLdsProbeJ], c[kLdsProbeJ], ex[kLdsProbeJ], loc[kLdsProbeJ][4];
#pragma unroll
        for (int j = 0; j < kLdsProbeJ; ++j) {
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t l = (uint32_t)key[j][k] - kbase;
                loc[j][k] = l;
                if (((pass[j] >> k) & 1u) && l < kRangeLdsSlice && ((s.bits[l >> 5] >> (l & 31u)) & 1u))
                    hit[j] |= 1u << k;
            }
            c[j] = __popc(hit[j]);
        }
        const uint32_t total = tile_offsets<kLdsProbeJ, kLdsProbeThreads / kWave>(c, ex, s.tot);
        // The step's output range is reserved now and used after the previous
        // step's rows are written out, which hides the reservation's wait
        // behind those stores (one atomic counter serves every work-group).
        unsigned long long mine = 0;
        if (threadIdx.x == 0 && total) mine = atomicAdd(counter, (unsigned long long)total);
        if (staged) emit.flush_rows<kLdsProbeThreads>(s.excl, staged, row);
        __syncthreads();  // the stage and s.excl are free
        if (threadIdx.x == 0) s.excl = mine;
#pragma unroll
        for (int j = 0; j < kLdsProbeJ; ++j) {
            uint32_t pos = ex[j];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!((hit[j] >> k) & 1u)) continue;
                s.loc[pos] = (uint16_t)loc[j][k];
                s.spay[pos] = sp[j][k];
                ++pos;
            }
        }
        __syncthreads();
        staged = total;
        if (!more) break;
#pragma unroll
        for (int j = 0; j < kLdsProbeJ; ++j) {
            pass[j] = pass2[j];
            pass2[j] = pass3[j];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                key[j][k] = key2[j][k];
                sp[j][k] = sp2[j][k];
                key2[j][k] = key3[j][k];
                sp2[j][k] = sp3[j][k];
            }
        }
        ni = mi;
        ne = me;
        nn = mn;
        if (!more2) ni = rpb;
    }
    emit.flush_rows<kLdsProbeThreads>(s.excl, staged, row);
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
int launch_join_part_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                           void *ws, hipStream_t st, bool *used) {
    *used = false;
    const uint64_t cap = join_table_slots(table_bytes);
    const PartPlan p = part_plan(cap, nr);
    if (!p.on || !ws) return QO_OK;
    Carver cv(ws);
    const PartBufs b = part_carve(p, nr, true, cv);
    JoinHeader *hdr = static_cast<JoinHeader *>(table);
    QO_TRY(launch_fill_u32(b.sb.ctr, 0, b.sb.ctr_words, st));
    QO_TRY(launch_fill_u64(b.spill_n, 0, 1, st));
    QO_TRY(launch_splits(p.s, b.sb, rkey, rpay, nullptr, nr, 0, 0, AppendOverflow{b.spill, b.spill_n},
                         Gate{&hdr->mode, kJoinHashMix}, HashSrc{}, st));
    const uint64_t *fb;
    uint64_t fcap;
    const uint32_t *fctr;
    split_fine(p.s, b.sb, fb, fcap, fctr);
    k_slice_build<<<1u << p.L, 256, (size_t)p.slice_max * 8, st>>>(fb, fcap, fctr, p.L, hdr, b.spill, b.spill_n);
    QO_LAUNCHED();
    k_spill_insert<<<(unsigned)props().cu_count * 4, 256, 0, st>>>(b.spill, b.spill_n, hdr);
    QO_LAUNCHED();
    *used = true;
    return QO_OK;
}

int launch_range_part_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay, size_t nr,
                            void *ws, hipStream_t st, bool *used) {
    *used = false;
    const uint64_t cap = join_table_slots(table_bytes);
    const PartPlan p = part_plan(cap, nr, kRangeSliceSlots);
    if (!p.on || !ws || (size_t)p.slice_max * 8 > (size_t)props().lds_per_block) return QO_OK;
    JoinHeader *hdr = static_cast<JoinHeader *>(table);
    const RangeLdsPlan rl = range_lds_plan(cap, nr);
    const uint32_t lds_bits = rl.on ? rl.bits : 0u;
    if (rl.on) {  // tables of 2^21 slots and more: one split by key, 2^15-slot slices built in LDS
        Carver cv(ws);
        const RangeLdsBufs lb = range_lds_carve(rl, cv);
        unsigned long long *spill_n = cv.take<unsigned long long>(1);
        uint64_t *spill = cv.take<uint64_t>(nr);
        const RangeBuildGate gate{hdr, lds_bits, true};
        QO_TRY(launch_fill_u32(lb.ctr, 0, lb.ctr_words, st));
        QO_TRY(launch_fill_u64(spill_n, 0, 1, st));
        const bool vec = ((uintptr_t)rkey % 16 == 0) && ((uintptr_t)rpay % 16 == 0);
        const size_t tiles = (nr + kSplitTile - 1) / kSplitTile;
        QO_REQUIRE(tiles < 0x7FFFFFFFull, "build input too large");
        k_range_lds_split<false><<<(unsigned)tiles, kSplitThreads, 0, st>>>(
            rkey, rpay, nullptr, nr, 0, 0, vec, rl.bcap, lb.bucket, lb.ctr, AppendOverflow{spill, spill_n}, gate);
        QO_LAUNCHED();
        k_range_lds_build<<<1u << rl.bits, kLdsProbeThreads, sizeof(RangeLdsBuildLds), st>>>(lb.bucket, rl.bcap,
                                                                                           lb.ctr, gate, spill,
                                                                                           spill_n);
        QO_LAUNCHED();
        k_range_spill<<<(unsigned)props().cu_count * 4, 256, 0, st>>>(spill, spill_n, hdr);
        QO_LAUNCHED();
    }
    // With more than 2^20 build rows and at most 2^(15 + bits) slots, every
    // key-range table the LDS build does not take has a repeated key (its
    // key range is under 2^21): it falls back to the hash-mixed table, built or
    // not, so the sliced build is not launched.
    int capbits = 0;
    while (capbits < 31 && (uint64_t(2) << capbits) <= cap) ++capbits;
    if (rl.on && nr > (size_t(1) << 20) && capbits <= kRangeLdsSliceBits + kRangeLdsMaxBits) {
        *used = true;
        return QO_OK;
    }
    // Other key-range tables: split by the table's own hash (for a key-range
    // table its top bits are those of key - kmin, so bucket f holds the keys
    // of table slice f), 16384-slot slices of full slots in LDS.
    Carver cv(ws);
    const PartBufs b = part_carve(p, nr, true, cv);
    QO_TRY(launch_fill_u32(b.sb.ctr, 0, b.sb.ctr_words, st));
    QO_TRY(launch_fill_u64(b.spill_n, 0, 1, st));
    QO_TRY(launch_splits(p.s, b.sb, rkey, rpay, nullptr, nr, 0, 0, AppendOverflow{b.spill, b.spill_n},
                         RangeBuildGate{hdr, lds_bits, false}, HashSrc{hdr}, st));
    const uint64_t *fb;
    uint64_t fcap;
    const uint32_t *fctr;
    split_fine(p.s, b.sb, fb, fcap, fctr);
    k_range_slice_build<<<1u << p.L, 256, (size_t)p.slice_max * 8, st>>>(fb, fcap, fctr, p.L, hdr, b.spill,
                                                                         b.spill_n, lds_bits);
    QO_LAUNCHED();
    k_range_spill<<<(unsigned)props().cu_count * 4, 256, 0, st>>>(b.spill, b.spill_n, hdr);
    QO_LAUNCHED();
    *used = true;
    return QO_OK;
}

int launch_join_part_probe(const void *table, size_t table_bytes, const int32_t *skey, const int32_t *spay,
                           const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, const JoinOut &out,
                           uint64_t *d_count, void *ws, hipStream_t st, bool *hash_part, bool *range_part,
                           bool hash_only) {
    *hash_part = *range_part = false;
    if (!ws) return QO_OK;
    const uint64_t slots = join_table_slots(table_bytes);
    const JoinHeader *hdr = static_cast<const JoinHeader *>(table);
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(d_count);
    const SoaEmit emit{out.okey, out.orpay, out.ospay, out.cap, out.k64, out.v32};
    L2Plan r;
    PartPlan p;
    probe_plans(slots, ns, r, p);
    if (p.on) {  // hash-mixed table too large for the L2 plan: per-slice LDS tables
        Carver cv(ws);
        const PartBufs b = part_carve(p, ns, false, cv);
        QO_TRY(launch_fill_u32(b.sb.ctr, 0, b.sb.ctr_words, st));
        QO_TRY(launch_splits(p.s, b.sb, skey, spay, sfilt, ns, flo, fhi, ProbeOverflow{hdr, counter, emit},
                             Gate{&hdr->mode, kJoinHashMix}, HashSrc{}, st));
        const uint64_t *fb;
        uint64_t fcap;
        const uint32_t *fctr;
        split_fine(p.s, b.sb, fb, fcap, fctr);
        k_slice_probe<<<1u << p.L, 256, (size_t)(p.slice_max + kSliceMargin) * 8, st>>>(fb, fcap, fctr, p.L, hdr,
                                                                                          counter, emit);
        QO_LAUNCHED();
        *hash_part = true;
    }
    const RangeLdsPlan rl = hash_only ? RangeLdsPlan{} : range_lds_plan(slots, ns);
    if (rl.on) {  // key-range tables whose slices fit LDS: split by key, look up in LDS
        Carver cv(ws);
        const RangeLdsBufs b = range_lds_carve(rl, cv);
        const RouteGate gate{hdr, 1u << kRouteRangeLds, rl.bits};
        QO_TRY(launch_fill_u32(b.ctr, 0, b.ctr_words, st));
        const bool vec = ((uintptr_t)skey % 16 == 0) && ((uintptr_t)spay % 16 == 0) &&
                         (!sfilt || (uintptr_t)sfilt % 16 == 0);
        const size_t tiles = (ns + kSplitTile - 1) / kSplitTile;
        QO_REQUIRE(tiles < 0x7FFFFFFFull, "probe input too large");
        const ProbeOverflow ovf{hdr, counter, emit};
        if (sfilt)
            k_range_lds_split<true><<<(unsigned)tiles, kSplitThreads, 0, st>>>(skey, spay, sfilt, ns, flo, fhi, vec,
                                                                              rl.bcap, b.bucket, b.ctr, ovf, gate);
        else
            k_range_lds_split<false><<<(unsigned)tiles, kSplitThreads, 0, st>>>(skey, spay, sfilt, ns, flo, fhi, vec,
                                                                               rl.bcap, b.bucket, b.ctr, ovf, gate);
        QO_LAUNCHED();
        k_range_lds_probe<<<rl.gpb << rl.bits, kLdsProbeThreads, sizeof(RangeLdsProbeLds), st>>>(b.bucket, rl.bcap, b.ctr, gate,
                                                                                   counter, emit);
        QO_LAUNCHED();
        *range_part = true;
    }
    if (r.on && (r.hash_fits || !hash_only)) {  // one split by the table's own hash, L2-resident slices
        // Key-range tables the LDS route does not take, and hash-mixed ones
        // whose slices stay small enough.
        const RouteGate gate{hdr, (1u << kRouteRangeL2) | (r.hash_fits ? 1u << kRouteHash : 0u), rl.on ? rl.bits : 0u};
        Carver cv(ws);
        const SplitBufs b = split_carve(r.s, cv);
        QO_TRY(launch_fill_u32(b.ctr, 0, b.ctr_words, st));
        QO_TRY(launch_splits(r.s, b, skey, spay, sfilt, ns, flo, fhi, ProbeOverflow{hdr, counter, emit}, gate,
                             HashSrc{hdr}, st));
        const uint64_t grid = (uint64_t)((r.nb + 7) / 8) * 8 * kXcdRegions * r.tpb;
        QO_REQUIRE(grid < 0x7FFFFFFFull, "probe input too large");
        k_l2_probe<<<(unsigned)grid, 256, 0, st>>>(b.bucketA, r.s.bcapA, b.ctr, r.nb, r.tpb, hdr, counter, emit, gate);
        QO_LAUNCHED();
        *range_part = true;
        if (r.hash_fits) *hash_part = true;
    }
    return QO_OK;
}

}  // namespace qo
