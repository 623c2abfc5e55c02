// Hash split of (key, payload) tuples into 2^L buckets by the top L bits of a
// 32-bit key hash: the partitioning step shared by the partitioned join and
// the partitioned group-by.  The hash is hash32(key), or the hash of a join
// table read from its header on the device (key-range tables split by key).
//
// Split A streams columns (an optional range filter fused) into 2^L1 buckets;
// split B splits each A bucket by the next L2 bits (L = L1 + L2, skipped when
// L2 = 0).  A split tile of 8192 tuples is ranked in LDS, reserves one range
// per non-empty bucket with one atomic add and writes contiguous runs.  Bucket
// capacities are fixed up front from the input size; tuples that find their
// bucket full go to an Overflow functor (skewed keys).
#pragma once
#include <algorithm>

#include "qo_internal.hpp"
#include "qo_join.hpp"

namespace qo {

constexpr int kSplitThreads = 512;                          // threads per split block
constexpr int kSplitJ = 4;                                  // int4 groups (4 tuples each) per thread
constexpr uint32_t kSplitTile = kSplitJ * 4 * kSplitThreads;  // tuples per split tile
constexpr int kMaxSplitBits = 7;                 // buckets per split pass <= 128
constexpr uint32_t kMaxSplit = 1u << kMaxSplitBits;
constexpr uint32_t kCtrStride = 16;              // bucket counters 64 bytes apart

struct SplitPlan {
    int L = 0, L1 = 0, L2 = 0;      // fine bucket bits = L1 (split A) + L2 (split B)
    uint64_t bcapA = 0, bcapB = 0;  // pairs per region of split A / per bucket of split B
    uint32_t subs = 1;              // split A regions per bucket (kXcdRegions: XcdLayout, one split only)
};
constexpr uint32_t kXcds = 8;        // work-group w runs on XCD w % 8
constexpr uint32_t kXcdRegions = kXcds;  // regions per bucket of an XcdLayout split (16: no faster)

// Plan for n tuples into 2^L fine buckets, 1 <= L <= 2 * kMaxSplitBits.  One
// split when it suffices; else the first split takes the smaller fan-out (it
// is the filtered one) so that both write runs of similar length.
inline SplitPlan split_plan(int L, size_t n, uint32_t subs = 1) {
    SplitPlan p;
    p.L = L;
    p.L1 = L <= kMaxSplitBits ? L : L / 2;
    p.L2 = L - p.L1;
    p.subs = p.L2 == 0 ? subs : 1;
    // Hash buckets of uniform keys stay within a few standard deviations of
    // their mean; the slack leaves only skewed inputs to the overflow path.
    const uint64_t regions = (uint64_t(1) << p.L1) * p.subs;
    const uint64_t perA = (n + regions - 1) / regions;
    p.bcapA = (perA + perA / 16 + 1024 + 31) / 32 * 32;
    const uint64_t perB = (n + (uint64_t(1) << L) - 1) >> L;
    p.bcapB = (perB + perB / 8 + 256 + 31) / 32 * 32;
    return p;
}

struct SplitBufs {
    uint32_t *ctr = nullptr;      // (2^L1 + 2^L) counters, kCtrStride apart: A buckets, then fine buckets
    uint64_t *bucketA = nullptr;
    uint64_t *bucketB = nullptr;  // only when L2 > 0
    size_t ctr_words = 0;
};

// Carves the split buffers out of a workspace (nullptr base: sizing only).
inline SplitBufs split_carve(const SplitPlan &p, Carver &cv) {
    SplitBufs b;
    b.ctr_words = ((size_t(1) << p.L1) * p.subs + (size_t(1) << p.L)) * kCtrStride;
    b.ctr = cv.take<uint32_t>(b.ctr_words);
    b.bucketA = cv.take<uint64_t>((size_t(1) << p.L1) * p.subs * p.bcapA);
    if (p.L2 > 0) b.bucketB = cv.take<uint64_t>((size_t(1) << p.L) * p.bcapB);
    return b;
}

// The 2^L fine buckets: split B's when there is one, else split A's.
inline void split_fine(const SplitPlan &p, const SplitBufs &b, const uint64_t *&bucket, uint64_t &bcap,
                       const uint32_t *&ctr) {
    if (p.L2 > 0) {
        bucket = b.bucketB;
        bcap = p.bcapB;
        ctr = b.ctr + ((size_t)1 << p.L1) * kCtrStride;
    } else {
        bucket = b.bucketA;
        bcap = p.bcapA;
        ctr = b.ctr;
    }
}

// Overflow that appends the tuple to a list (capacity: the input size; a tuple
// overflows at most once).
struct AppendOverflow {
    uint64_t *list;
    unsigned long long *n;
    QO_DEV void operator()(uint64_t v) const { list[atomicAdd(n, 1ull)] = v; }
};


// Where bucket b's tuples go in a split's output array: [start(b), start(b) +
// cap(b)); count(b, ctr) is how many it holds after the split ran.
// Fixed: equal capacity per bucket, fill counts in the counters.
struct FixedLayout {
    uint64_t bcap;
    QO_DEV uint32_t slot(uint32_t b) const { return b; }
    QO_DEV uint64_t start(uint32_t b) const { return (uint64_t)b * bcap; }
    QO_DEV uint64_t cap(uint32_t) const { return bcap; }
    QO_DEV uint64_t count(uint32_t b, const uint32_t *ctr) const {
        const uint64_t c = ctr[b * kCtrStride];
        return c < bcap ? c : bcap;
    }
};
// Per XCD: bucket b has one region and one cursor per XCD, and a block
// appends to region (b, its XCD), slot b * 8 + blockIdx % 8 (work-group w runs
// on XCD w % 8).  Each cursor then takes the reservations of one XCD's
// blocks: same-address device atomics serialise, and with one cursor per
// bucket every split tile of 2^28 tuples queued 128 reservations behind those
// of ~500 other tiles.
struct XcdLayout {
    uint64_t bcap;  // per region
    QO_DEV uint32_t slot(uint32_t b) const { return b * kXcdRegions + (blockIdx.x & (kXcdRegions - 1)); }
    QO_DEV uint64_t start(uint32_t s) const { return (uint64_t)s * bcap; }
    QO_DEV uint64_t cap(uint32_t) const { return bcap; }
};
// Exact: a bucket of this split covers fine buckets [b << sh, (b + 1) << sh).
// Its region starts at offp (the exclusive scan of the fine-bucket counts,
// each rounded up to even, so that every region starts 16-byte aligned) and it
// holds as many tuples as the exclusive scan `off` of the exact counts says.
// Nothing overflows.
struct ExactLayout {
    const uint32_t *offp, *off;
    int sh;
    QO_DEV uint32_t slot(uint32_t b) const { return b; }
    QO_DEV uint64_t start(uint32_t b) const { return offp[(size_t)b << sh]; }
    QO_DEV uint64_t cap(uint32_t b) const { return offp[(size_t)(b + 1) << sh] - offp[(size_t)b << sh]; }
    QO_DEV uint64_t count(uint32_t b, const uint32_t *) const {
        return off[(size_t)(b + 1) << sh] - off[(size_t)b << sh];
    }
};

// Where a split kernel takes its hash from: a join table's header, or (null)
// hash32, or with `fib` the multiplicative hash key * 2^32/phi.
struct HashSrc {
    const JoinHeader *hdr = nullptr;
    uint32_t fib = 0;
    QO_DEV JoinHash get() const {
        return hdr ? JoinHash::of(hdr) : JoinHash{fib ? kHashFib : kJoinHashMix, 0u, 0};
    }
};

// LDS of a split block into at most MAXP (<= kSplitThreads) buckets.
template <uint32_t MAXP>
struct SplitLdsT {
    static_assert(MAXP <= (uint32_t)kSplitThreads, "one bucket per thread at most");
    uint64_t stage[kSplitTile];
    uint64_t gpos[MAXP];  // where the part's run starts in `bucket`
    uint32_t cnt[MAXP], loff[MAXP], room[MAXP];
    uint32_t w[kSplitThreads / kWave];
};
using SplitLds = SplitLdsT<kMaxSplit>;

// Splits the block's live tuples (J groups of 4 per thread, bit k of pass[j]
// marks element k of group j) into `parts` buckets by
// part = (hs(key) >> shift) & (parts - 1); bucket index = bucket0 + part,
// placed by `out` in `bucket`; ctr holds one fill cursor per bucket.
template <int J, class Layout, class Overflow, class Lds>
QO_DEV void split_block(const int32_t (&key)[J][4], const int32_t (&pay)[J][4], const uint32_t (&pass)[J],
                        const JoinHash &hs, int shift, uint32_t parts, uint32_t bucket0,
                        uint64_t *__restrict__ bucket, const Layout &out, uint32_t *ctr, Lds &sh,
                        const Overflow &overflow) {
    const uint32_t mask = parts - 1;
    for (uint32_t i = threadIdx.x; i < parts; i += kSplitThreads) sh.cnt[i] = 0;
    __syncthreads();
    // Rank inside the tile: (part << 16) | arrival order within the part.
    uint32_t rank[J][4];
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rank[j][k] = 0;
            if ((pass[j] >> k) & 1u) {
                uint32_t h;
                hs(key[j][k], h);
                const uint32_t p = (h >> shift) & mask;
                rank[j][k] = (p << 16) | atomicAdd(&sh.cnt[p], 1u);
            }
        }
    __syncthreads();
    // Part offsets inside the tile and one bucket reservation per non-empty part.
    const uint32_t c0 = threadIdx.x < parts ? sh.cnt[threadIdx.x] : 0u;
    const uint32_t ex = block_excl_scan<kSplitThreads / kWave>(c0, sh.w);
    if (threadIdx.x < parts) {
        // The run's start and how many of its tuples fit (the rest overflow).
        const uint32_t b = out.slot(bucket0 + threadIdx.x);
        const uint64_t g = c0 ? atomicAdd(&ctr[b * kCtrStride], c0) : 0u;
        const uint64_t cap = out.cap(b);
        sh.loff[threadIdx.x] = ex;
        sh.gpos[threadIdx.x] = out.start(b) + g;
        sh.room[threadIdx.x] = (uint32_t)(g < cap ? (cap - g < c0 ? cap - g : c0) : 0u);
    }
    __syncthreads();
    const uint32_t total = sh.loff[parts - 1] + sh.cnt[parts - 1];
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if ((pass[j] >> k) & 1u)
                sh.stage[sh.loff[rank[j][k] >> 16] + (rank[j][k] & 0xFFFFu)] = join_pack(key[j][k], pay[j][k]);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < total; i += kSplitThreads) {
        const uint64_t v = sh.stage[i];
        uint32_t h;
        hs((int32_t)(uint32_t)v, h);
        const uint32_t p = (h >> shift) & mask;
        const uint32_t r = i - sh.loff[p];
        // Plain stores: runs from neighbouring tiles complete each other's
        // lines in L2 before they are written back.
        if (r < sh.room[p]) bucket[sh.gpos[p] + r] = v;
        else overflow(v);
    }
}

// Split A: columns (key, payload[, filter]) -> 2^L1 buckets.
template <bool HAS_FILTER, class Layout, class Overflow, class G = Gate>
__global__ __launch_bounds__(kSplitThreads) void k_split_cols(const int32_t *__restrict__ kcol, const int32_t *__restrict__ pcol,
                                                    const int32_t *__restrict__ fcol, size_t n, int32_t flo,
                                                    int32_t fhi, bool vec, int shift, uint32_t parts,
                                                    uint64_t *__restrict__ bucket, Layout out, uint32_t *ctr,
                                                    Overflow overflow, G gate, HashSrc hsrc) {
    __shared__ SplitLds sh;
    if (gate.off()) return;
    int32_t key[kSplitJ][4], pay[kSplitJ][4];
    uint32_t pass[kSplitJ];
    load_s_tile<kSplitJ, HAS_FILTER, kSplitThreads>(kcol, pcol, fcol, n, flo, fhi, vec, blockIdx.x, key, pay, pass);
    split_block<kSplitJ>(key, pay, pass, hsrc.get(), shift, parts, 0, bucket, out, ctr, sh, overflow);
}

// Loads up to J*4*NT (key, payload) pairs from `src` (16-byte aligned) with NT
// threads: pair quad j of thread t holds pairs 2v, 2v+1 for
// v = (2j + h)*NT + t, h = 0, 1.
template <int J, int NT = 256>
QO_DEV void load_pairs(const uint64_t *__restrict__ src, uint32_t m, int32_t (&key)[J][4], int32_t (&pay)[J][4],
                       uint32_t (&pass)[J]) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
        pass[j] = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t v = (uint32_t)(2 * j + h) * NT + threadIdx.x;
            const uint32_t e = 2 * v;
            key[j][2 * h] = pay[j][2 * h] = key[j][2 * h + 1] = pay[j][2 * h + 1] = 0;
            if (e + 2 <= m) {
                const uint4 q = ld_stream(reinterpret_cast<const uint4 *>(src) + v);
                key[j][2 * h] = (int32_t)q.x;
                pay[j][2 * h] = (int32_t)q.y;
                key[j][2 * h + 1] = (int32_t)q.z;
                pay[j][2 * h + 1] = (int32_t)q.w;
                pass[j] |= 3u << (2 * h);
            } else if (e < m) {
                const uint64_t w = src[e];
                key[j][2 * h] = (int32_t)(uint32_t)w;
                pay[j][2 * h] = (int32_t)(w >> 32);
                pass[j] |= 1u << (2 * h);
            }
        }
    }
}

// Split B on m <= kSplitTile pairs of input bucket p1 starting at input[s0]:
// -> buckets (p1 << L2) + part.  A run starting at an odd position (exact
// layout) is read with element loads.
template <class Layout, class Overflow>
QO_DEV void split_pairs_run(const uint64_t *__restrict__ input, uint64_t s0, uint32_t m, uint32_t p1,
                            const JoinHash &hs, int shift, uint32_t parts, uint64_t *__restrict__ bucket,
                            const Layout &out, uint32_t *ctr, SplitLds &sh, const Overflow &overflow) {
    int32_t key[kSplitJ][4], pay[kSplitJ][4];
    uint32_t pass[kSplitJ];
    if ((s0 & 1u) == 0) {
        load_pairs<kSplitJ, kSplitThreads>(input + s0, m, key, pay, pass);
    } else {
#pragma unroll
        for (int j = 0; j < kSplitJ; ++j) {
            pass[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t e = 2 * ((uint32_t)(2 * j + (k >> 1)) * kSplitThreads + threadIdx.x) + (k & 1);
                key[j][k] = pay[j][k] = 0;
                if (e < m) {
                    const uint64_t w = input[s0 + e];
                    key[j][k] = (int32_t)(uint32_t)w;
                    pay[j][k] = (int32_t)(w >> 32);
                    pass[j] |= 1u << k;
                }
            }
        }
    }
    split_block<kSplitJ>(key, pay, pass, hs, shift, parts, p1 * parts, bucket, out, ctr, sh, overflow);
}

// Split B on tile t of input bucket p1 (placed by `in`).
template <class InLayout, class Layout, class Overflow>
QO_DEV void split_pairs_tile(const uint64_t *__restrict__ input, const InLayout &in, const uint32_t *in_ctr,
                             uint32_t p1, uint32_t t, const JoinHash &hs, int shift, uint32_t parts,
                             uint64_t *__restrict__ bucket, const Layout &out, uint32_t *ctr, SplitLds &sh,
                             const Overflow &overflow) {
    const uint64_t n = in.count(p1, in_ctr);
    const uint64_t e0 = (uint64_t)t * kSplitTile;
    if (e0 >= n) return;  // whole block: no barrier has been reached
    const uint32_t m = (uint32_t)(n - e0 < kSplitTile ? n - e0 : kSplitTile);
    split_pairs_run(input, in.start(p1) + e0, m, p1, hs, shift, parts, bucket, out, ctr, sh, overflow);
}

// Split B with fixed-capacity input buckets: block = (bucket, tile).
template <class Layout, class Overflow, class G = Gate>
__global__ __launch_bounds__(kSplitThreads) void k_split_pairs(const uint64_t *__restrict__ input, FixedLayout in,
                                                     const uint32_t *in_ctr, uint32_t tiles_per_bucket, int shift,
                                                     uint32_t parts, uint64_t *__restrict__ bucket, Layout out,
                                                     uint32_t *ctr, Overflow overflow, G gate, HashSrc hsrc) {
    __shared__ SplitLds sh;
    if (gate.off()) return;
    split_pairs_tile(input, in, in_ctr, blockIdx.x / tiles_per_bucket, blockIdx.x % tiles_per_bucket, hsrc.get(),
                     shift, parts, bucket, out, ctr, sh, overflow);
}

// Tile map of split B of the exact split: entry i describes the tile of block
// i, the tiles of all input buckets in order (bucket b has
// ceil(size(b) / kSplitTile) tiles), as tile_map_entry(start, pairs, bucket);
// entries past the last tile are ~0.  One block of kSplitThreads threads
// (nin <= kSplitThreads).  The block of split B thus finds its input with one
// load instead of a chain of dependent loads.
// Bits 0-31 start, 32-47 pairs (<= kSplitTile), 48-63 input bucket.
static_assert(kSplitTile < (1u << 16) && kMaxSplit <= (1u << 16), "tile map fields");
__host__ __device__ inline uint64_t tile_map_entry(uint64_t s0, uint32_t m, uint32_t b) {
    return s0 | ((uint64_t)m << 32) | ((uint64_t)b << 48);
}
static __global__ __launch_bounds__(kSplitThreads) void k_exact_tile_map(ExactLayout in, uint32_t nin, uint64_t *map,
                                                                         uint32_t entries) {
    __shared__ uint32_t s_w[kSplitThreads / kWave];
    __shared__ uint32_t s_first[kSplitThreads + 1];
    const uint32_t nt =
        threadIdx.x < nin ? (uint32_t)((in.count(threadIdx.x, nullptr) + kSplitTile - 1) / kSplitTile) : 0u;
    const uint32_t first = block_excl_scan<kSplitThreads / kWave>(nt, s_w);
    s_first[threadIdx.x] = first;
    if (threadIdx.x == kSplitThreads - 1) s_first[kSplitThreads] = first + nt;
    __syncthreads();
    for (uint32_t b = 0; b < nin; ++b) {
        const uint32_t f = s_first[b], c = s_first[b + 1] - f;
        for (uint32_t t = threadIdx.x; t < c; t += kSplitThreads)
            if (f + t < entries) {
                const uint64_t e0 = (uint64_t)t * kSplitTile, n = in.count(b, nullptr);
                map[f + t] = tile_map_entry(in.start(b) + e0, (uint32_t)(n - e0 < kSplitTile ? n - e0 : kSplitTile), b);
            }
    }
    for (uint32_t i = s_first[kSplitThreads] + threadIdx.x; i < entries; i += kSplitThreads) map[i] = ~uint64_t(0);
}

// Split B of the exact split: block i takes (input bucket, tile) from map[i].
template <class Overflow>
__global__ __launch_bounds__(kSplitThreads) void k_split_pairs_exact(const uint64_t *__restrict__ input,
                                                           const uint64_t *__restrict__ map, int shift,
                                                           uint32_t parts, uint64_t *__restrict__ bucket,
                                                           ExactLayout out, uint32_t *ctr, Overflow overflow) {
    __shared__ SplitLds sh;
    // XCD-contiguous order: one XCD handles whole input buckets, so the runs
    // appended to one output bucket come from one L2.
    const uint64_t e = map[xcd_tile(blockIdx.x, gridDim.x)];
    if (e == ~uint64_t(0)) return;  // past the last tile (the grid is an upper bound)
    split_pairs_run(input, e & 0xFFFFFFFFull, (uint32_t)(e >> 32) & 0xFFFFu, (uint32_t)(e >> 48),
                    HashSrc{nullptr, 1}.get(), shift, parts, bucket, out, ctr, sh, overflow);
}

// Launches split A (and B) with fixed bucket capacities: kcol/pcol (and fcol
// when not null, with the range [flo, fhi]) -> fine buckets.  The counters must
// have been zeroed.  The kernels run only where `gate` lets them and hash as
// `hsrc` says.
template <class Overflow, class G>
inline int launch_splits(const SplitPlan &p, const SplitBufs &b, const int32_t *kcol, const int32_t *pcol,
                         const int32_t *fcol, size_t n, int32_t flo, int32_t fhi, const Overflow &overflow,
                         G gate, HashSrc hsrc, hipStream_t st) {
    const bool vec = ((uintptr_t)kcol % 16 == 0) && ((uintptr_t)pcol % 16 == 0) && (!fcol || (uintptr_t)fcol % 16 == 0);
    const uint32_t partsA = 1u << p.L1;
    const size_t tiles = (n + kSplitTile - 1) / kSplitTile;
    QO_REQUIRE(tiles < 0x7FFFFFFFull, "input too large for a hash split");
    const FixedLayout la{p.bcapA}, lb{p.bcapB};
    if (p.subs > 1) {
        QO_REQUIRE(p.L2 == 0 && p.subs == kXcdRegions, "per-XCD regions need one split");
        const XcdLayout lx{p.bcapA};
        if (fcol)
            k_split_cols<true><<<(unsigned)tiles, kSplitThreads, 0, st>>>(kcol, pcol, fcol, n, flo, fhi, vec, 32 - p.L1,
                                                                         partsA, b.bucketA, lx, b.ctr, overflow, gate,
                                                                         hsrc);
        else
            k_split_cols<false><<<(unsigned)tiles, kSplitThreads, 0, st>>>(kcol, pcol, fcol, n, flo, fhi, vec, 32 - p.L1,
                                                                          partsA, b.bucketA, lx, b.ctr, overflow, gate,
                                                                          hsrc);
        QO_LAUNCHED();
        return QO_OK;
    }
    if (fcol)
        k_split_cols<true><<<(unsigned)tiles, kSplitThreads, 0, st>>>(kcol, pcol, fcol, n, flo, fhi, vec, 32 - p.L1, partsA,
                                                            b.bucketA, la, b.ctr, overflow, gate, hsrc);
    else
        k_split_cols<false><<<(unsigned)tiles, kSplitThreads, 0, st>>>(kcol, pcol, fcol, n, flo, fhi, vec, 32 - p.L1, partsA,
                                                             b.bucketA, la, b.ctr, overflow, gate, hsrc);
    QO_LAUNCHED();
    if (p.L2 > 0) {
        const uint32_t tpb = (uint32_t)((p.bcapA + kSplitTile - 1) / kSplitTile);
        QO_REQUIRE((uint64_t)tpb * partsA < 0x7FFFFFFFull, "input too large for a hash split");
        k_split_pairs<<<tpb * partsA, kSplitThreads, 0, st>>>(b.bucketA, la, b.ctr, tpb, 32 - p.L, 1u << p.L2, b.bucketB, lb,
                                                   b.ctr + (size_t)partsA * kCtrStride, overflow, gate, hsrc);
        QO_LAUNCHED();
    }
    return QO_OK;
}

// ---------------------------------------------------------------------------
// Exact split (by the multiplicative hash, key * 2^32/phi, so that a dense key
// range spreads evenly over the buckets and, within a bucket, over the next
// hash bits): a histogram pass over the keys sizes every bucket exactly, so
// skewed keys cannot overflow.  Output: the fine buckets back to back in
// `fine` (bucket f at [off[f], off[f + 1])), with `mid` holding split A's.
// ---------------------------------------------------------------------------
// Histogram of fine buckets (top L bits of the multiplicative hash) over the key
// column.
// Dynamic LDS: nb counters.
static __global__ __launch_bounds__(512) void k_split_hist(const int32_t *__restrict__ key, size_t n, int shift, uint32_t nb,
                                                           uint32_t *hist) {
    extern __shared__ uint32_t s_hist[];
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) s_hist[i] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const bool vec = ((uintptr_t)key % 16) == 0;
    size_t from = 0;
    if (vec) {
        const size_t nv = n / 4;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
            const int4 k = ld_stream(reinterpret_cast<const int4 *>(key) + i);
            atomicAdd(&s_hist[((uint32_t)k.x * 0x9E3779B9u) >> shift], 1u);
            atomicAdd(&s_hist[((uint32_t)k.y * 0x9E3779B9u) >> shift], 1u);
            atomicAdd(&s_hist[((uint32_t)k.z * 0x9E3779B9u) >> shift], 1u);
            atomicAdd(&s_hist[((uint32_t)k.w * 0x9E3779B9u) >> shift], 1u);
        }
        from = nv * 4;
    }
    for (size_t i = from + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&s_hist[((uint32_t)key[i] * 0x9E3779B9u) >> shift], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        if (s_hist[i]) atomicAdd(&hist[i], s_hist[i]);
}

struct ExactSplitBufs {
    uint32_t *ctr = nullptr;   // fill cursors: 2^L1 (split A), then 2^L (split B), kCtrStride apart
    uint32_t *hist = nullptr;  // 2^L + 1 fine-bucket counts (last entry stays 0)
    uint32_t *off = nullptr;   // 2^L + 1 exclusive offsets of the counts; off[2^L] = n
    uint32_t *hpad = nullptr;  // counts rounded up to even
    uint32_t *offp = nullptr;  // 2^L + 1 exclusive offsets of hpad: where each fine bucket starts
    void *scan_ws = nullptr;
    uint64_t *mid = nullptr;   // split A output (n + 2^L pairs), when L2 > 0
    uint64_t *fine = nullptr;  // fine buckets (n + 2^L pairs)
    uint64_t *tmap = nullptr;  // split B tile map (tiles + 2^L1 entries), when L2 > 0
    size_t ctr_words = 0;
};

inline ExactSplitBufs exact_split_carve(const SplitPlan &p, size_t n, Carver &cv) {
    ExactSplitBufs b;
    const size_t nb = size_t(1) << p.L;
    b.ctr_words = ((size_t(1) << p.L1) + nb) * kCtrStride;
    b.ctr = cv.take<uint32_t>(b.ctr_words);
    b.hist = cv.take<uint32_t>(nb + 1);
    b.off = cv.take<uint32_t>(nb + 1);
    b.hpad = cv.take<uint32_t>(nb + 1);
    b.offp = cv.take<uint32_t>(nb + 1);
    b.scan_ws = cv.take<char>(scan_workspace_bytes(nb + 1));
    if (p.L2 > 0) {
        b.mid = cv.take<uint64_t>(n + nb);
        b.tmap = cv.take<uint64_t>((n + kSplitTile - 1) / kSplitTile + (size_t(1) << p.L1));
    }
    b.fine = cv.take<uint64_t>(n + nb);
    return b;
}

static __global__ void k_round_even(const uint32_t *h, uint32_t m, uint32_t *hp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) hp[i] = (h[i] + 1u) & ~1u;
}

// Histogram, offsets and the exact split(s) of (kcol, pcol) into b.fine.
inline int launch_exact_splits(const SplitPlan &p, const ExactSplitBufs &b, const int32_t *kcol,
                               const int32_t *pcol, size_t n, hipStream_t st) {
    const uint32_t nb = 1u << p.L;
    QO_REQUIRE(n + nb < 0xFFFFFFFFull && nb <= 16384, "exact split too large");
    QO_TRY(launch_fill_u32(b.ctr, 0, b.ctr_words, st));
    QO_TRY(launch_fill_u32(b.hist, 0, nb + 1, st));
    const unsigned hg = grid_for(n, 512 * 16, (unsigned)props().cu_count * 2);
    k_split_hist<<<hg, 512, (size_t)nb * 4, st>>>(kcol, n, 32 - p.L, nb, b.hist);
    QO_LAUNCHED();
    QO_TRY(launch_scan_excl_u32(b.hist, b.off, nb + 1, b.scan_ws, st));
    k_round_even<<<(nb + 1 + 255) / 256, 256, 0, st>>>(b.hist, nb + 1, b.hpad);
    QO_LAUNCHED();
    QO_TRY(launch_scan_excl_u32(b.hpad, b.offp, nb + 1, b.scan_ws, st));
    const bool vec = ((uintptr_t)kcol % 16 == 0) && ((uintptr_t)pcol % 16 == 0);
    const uint32_t partsA = 1u << p.L1;
    const size_t tiles = (n + kSplitTile - 1) / kSplitTile;
    const ExactLayout la{b.offp, b.off, p.L2}, lf{b.offp, b.off, 0};
    AppendOverflow none{nullptr, nullptr};  // exact regions: never called
    uint64_t *outA = p.L2 > 0 ? b.mid : b.fine;
    k_split_cols<false><<<(unsigned)tiles, kSplitThreads, 0, st>>>(kcol, pcol, nullptr, n, 0, 0, vec, 32 - p.L1, partsA, outA,
                                                         la, b.ctr, none, Gate{}, HashSrc{nullptr, 1});
    QO_LAUNCHED();
    if (p.L2 > 0) {
        // At most one partial tile per A bucket beyond n / kSplitTile full ones.
        const size_t tiles_b = tiles + partsA;
        QO_REQUIRE(tiles_b < 0x7FFFFFFFull, "input too large for a hash split");
        QO_REQUIRE(partsA <= kSplitThreads, "split A fan-out too large");
        k_exact_tile_map<<<1, kSplitThreads, 0, st>>>(la, partsA, b.tmap, (uint32_t)tiles_b);
        QO_LAUNCHED();
        k_split_pairs_exact<<<(unsigned)tiles_b, kSplitThreads, 0, st>>>(b.mid, b.tmap, 32 - p.L, 1u << p.L2, b.fine,
                                                               lf, b.ctr + (size_t)partsA * kCtrStride, none);
        QO_LAUNCHED();
    }
    return QO_OK;
}

}  // namespace qo
