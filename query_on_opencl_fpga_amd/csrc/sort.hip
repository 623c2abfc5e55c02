// Sorting: a stable LSD radix sort over 64-bit keys (8-bit digits) with an
// optional 32-bit value.  Every pass ranks a 4096-key tile with wave-level
// match masks (8 ballots per key), stages the tile in LDS in digit order and
// writes contiguous runs per digit.
//
// Every pass is reduce-then-scan: per-tile digit counts in tile-major rows
// (LDS atomics, one read of the keys), a column scan of the rows that turns
// them into the output offset of every (tile, digit), and the placement
// kernel, which reads its tile's 256 offsets as one coalesced row and walks
// the tiles in XCD-contiguous order.  A batch of many arrays (one per query)
// sorts in the same launches: a segment owns a contiguous tile range, the
// column scan restarts at every segment's first tile, and the digit starts of
// every pass per segment come from one histogram up front.  A segment's
// element count may live in device memory (written by an earlier kernel), so a
// whole batch can be captured in one graph without host round trips.
//
// Reference path replaced: radixSortImpl_sm (host/host/testSort.cpp:24-136)
// driving bitonicSortShared / bitonicMergeGlobal (kernel/aggregation/1/
// group_by_cu_op_dp.cl:89-291): O(log^2 N) in-place passes over the array.
// Here a Record sort is 4 passes over .y; ties keep input order (stable).
#include <algorithm>

#include "qo_device.hpp"
#include "qo_internal.hpp"
#include "qo_sort.hpp"

namespace qo {

// Lanes of this wave that are active and hold the same 8-bit digit.
QO_DEV uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint64_t bb = ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    return m;
}

// Digit of key k in the pass at `shift`: key bits [shift, shift + 8), or, with
// MOD (a split into `mod` partitions), bits [shift, shift + 8) of the
// partition id hi32(k) % mod, so that the passes order keys by partition id.
template <bool MOD>
QO_DEV uint32_t key_digit(uint64_t k, int shift, uint32_t mod) {
    if constexpr (MOD) return (((uint32_t)(k >> 32) % mod) >> shift) & 255u;
    else return (uint32_t)(k >> shift) & 255u;
}

// Index of the segment that owns global tile `tile` (0 for a single segment).
QO_DEV int seg_index_of(const RadixSeg *segs, int nseg, uint32_t tile) {
    if (nseg <= 1) return 0;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

QO_DEV size_t seg_count(const RadixSeg &s) {
    size_t v = s.d_n ? (size_t)*s.d_n : s.cap;
    return v < s.cap ? v : s.cap;
}

// ---------------------------------------------------------------------------
// Reduce-then-scan passes for a single array.
// ---------------------------------------------------------------------------
constexpr int kHistTiles = 2;             // tiles per block of k_tile_hist
constexpr uint32_t kRowsBlocks = 512;     // at most this many column-scan blocks
constexpr uint32_t kRowsMinSpan = 64;     // tiles per column-scan block, at least
constexpr uint32_t kRowsOneBlock = 256;   // up to this many tiles, one block scans the rows

// Digit counts of one tile (rem keys at src) into 256 LDS counters; `vec`:
// src is 16-byte aligned, so two keys are read per load.
template <bool MOD = false>
QO_DEV void hist_tile(const uint64_t *__restrict__ src, uint32_t rem, int shift, uint32_t *s_row, bool vec,
                      uint32_t mod = 0) {
    if (vec) {
        constexpr int PER = kSortTile / 512;
        uint4 q[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t v = j * 256 + threadIdx.x;
            q[j] = (2 * v + 2 <= rem) ? ld_stream(reinterpret_cast<const uint4 *>(src) + v) : uint4{0, 0, 0, 0};
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t v = j * 256 + threadIdx.x;
            if (2 * v + 2 <= rem) {
                const uint64_t k0 = ((uint64_t)q[j].y << 32) | q[j].x;
                const uint64_t k1 = ((uint64_t)q[j].w << 32) | q[j].z;
                atomicAdd(&s_row[key_digit<MOD>(k0, shift, mod)], 1u);
                atomicAdd(&s_row[key_digit<MOD>(k1, shift, mod)], 1u);
            } else if (2 * v < rem) {
                atomicAdd(&s_row[key_digit<MOD>(src[2 * v], shift, mod)], 1u);
            }
        }
    } else {
        for (uint32_t i = threadIdx.x; i < rem; i += 256) atomicAdd(&s_row[key_digit<MOD>(src[i], shift, mod)], 1u);
    }
}

// rows[t * 256 + d] = keys of global tile t whose digit is d (zero rows for
// tiles past their segment's element count).  A single array passes its
// segment as seg0 with segs = null; a batch passes the segment list.
template <bool MOD = false>
__global__ __launch_bounds__(256) void k_tile_hist(const uint64_t *__restrict__ keys, RadixSeg seg0,
                                                   const RadixSeg *segs, int nseg, uint32_t ntile, int shift,
                                                   uint32_t *__restrict__ rows, const uint32_t *run_if = nullptr,
                                                   uint32_t mod = 0) {
    if (run_if && *run_if == 0) return;  // gated off: runs only when *run_if != 0
    __shared__ uint32_t s_h[kHistTiles][256];
    for (int i = threadIdx.x; i < kHistTiles * 256; i += 256) (&s_h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t t0 = blockIdx.x * kHistTiles;
#pragma unroll
    for (int tt = 0; tt < kHistTiles; ++tt) {
        const uint32_t t = t0 + tt;
        if (t >= ntile) break;
        const RadixSeg sg = segs ? segs[seg_index_of(segs, nseg, t)] : seg0;
        const size_t n = seg_count(sg);
        const size_t tb = (size_t)(t - sg.tile0) * kSortTile;
        if (tb >= n) continue;
        const uint64_t *src = keys + sg.base + tb;
        hist_tile<MOD>(src, (uint32_t)std::min<size_t>(kSortTile, n - tb), shift, s_h[tt],
                       (reinterpret_cast<uintptr_t>(src) & 15) == 0, mod);
    }
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < kHistTiles; ++tt)
        if (t0 + tt < ntile) rows[(size_t)(t0 + tt) * 256 + threadIdx.x] = s_h[tt][threadIdx.x];
}

// Column-scan layout: block b of the scan owns tiles [b * span, (b + 1) * span).
struct RowsPlan {
    uint32_t span, nb;
};
static RowsPlan rows_plan(uint32_t ntile) {
    RowsPlan r;
    r.span = std::max(kRowsMinSpan, (ntile + kRowsBlocks - 1) / kRowsBlocks);
    r.nb = (ntile + r.span - 1) / r.span;
    return r;
}

// bsum[b * 256 + d] = digit-d keys of the tiles of block b.
__global__ __launch_bounds__(256) void k_rows_reduce(const uint32_t *__restrict__ rows, uint32_t ntile, uint32_t span,
                                                     uint32_t *__restrict__ bsum) {
    const uint32_t t0 = blockIdx.x * span, t1 = std::min(ntile, t0 + span);
    const uint32_t *col = rows + threadIdx.x;
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // independent sums: loads stay in flight together
    uint32_t t = t0;
    for (; t + 8 <= t1; t += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] += col[(size_t)(t + u) * 256];
    for (; t < t1; ++t) c[0] += col[(size_t)t * 256];
    bsum[(size_t)blockIdx.x * 256 + threadIdx.x] = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7];
}

// One block of 1024 threads: bsum[b][d] := start of digit d in the output +
// digit-d keys of blocks before b (exclusive, in place).  Thread (q, d) with
// q = threadIdx / 256 handles quarter q of the blocks for digit d.
__global__ __launch_bounds__(1024) void k_rows_starts(uint32_t *bsum, uint32_t nb) {
    __shared__ uint32_t s_part[4][256];
    __shared__ uint32_t s_w[16];
    const uint32_t d = threadIdx.x & 255u, q = threadIdx.x >> 8;
    const uint32_t per = (nb + 3) / 4, b0 = std::min(nb, q * per), b1 = std::min(nb, b0 + per);
    uint32_t *col = bsum + d;
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t b = b0;
    for (; b + 8 <= b1; b += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] += col[(size_t)(b + u) * 256];
    for (; b < b1; ++b) c[0] += col[(size_t)b * 256];
    s_part[q][d] = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7];
    __syncthreads();
    const uint32_t p0 = s_part[0][d], p1 = s_part[1][d], p2 = s_part[2][d], p3 = s_part[3][d];
    // Threads 0..255 come first in the block, so their exclusive scan is the digit start.
    const uint32_t start = block_excl_scan<16>(q == 0 ? p0 + p1 + p2 + p3 : 0u, s_w);
    __syncthreads();
    if (q == 0) s_part[0][d] = start;
    __syncthreads();
    uint32_t run = s_part[0][d] + (q > 0 ? p0 : 0u) + (q > 1 ? p1 : 0u) + (q > 2 ? p2 : 0u);
    b = b0;
    for (; b + 8 <= b1; b += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = col[(size_t)(b + u) * 256];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            col[(size_t)(b + u) * 256] = run;
            run += v[u];
        }
    }
    for (; b < b1; ++b) {
        const uint32_t v = col[(size_t)b * 256];
        col[(size_t)b * 256] = run;
        run += v;
    }
}

// rows[t][d] := output offset of the first digit-d key of tile t (in place).
__global__ __launch_bounds__(256) void k_rows_apply(uint32_t *__restrict__ rows, uint32_t ntile, uint32_t span,
                                                    const uint32_t *__restrict__ bsum) {
    const uint32_t t0 = blockIdx.x * span, t1 = std::min(ntile, t0 + span);
    uint32_t *col = rows + threadIdx.x;
    uint32_t run = bsum[(size_t)blockIdx.x * 256 + threadIdx.x];
    uint32_t t = t0;
    for (; t + 8 <= t1; t += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = col[(size_t)(t + u) * 256];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            col[(size_t)(t + u) * 256] = run;
            run += v[u];
        }
    }
    for (; t < t1; ++t) {
        const uint32_t v = col[(size_t)t * 256];
        col[(size_t)t * 256] = run;
        run += v;
    }
}

// ---- segmented column scan (a batch: one array per segment) ----
// Walks global tiles t0, t0 + 1, ... and tells which of them start a segment.
struct SegWalk {
    const RadixSeg *segs;
    int nseg, si;
    QO_DEV SegWalk(const RadixSeg *s, int n, uint32_t t0) : segs(s), nseg(n), si(seg_index_of(s, n, t0)) {}
    QO_DEV bool starts(uint32_t t) {
        while (si + 1 < nseg && segs[si + 1].tile0 <= t) ++si;
        return segs[si].tile0 == t;
    }
};

// bsum[b][d] = digit-d keys of block b's tiles after the last segment start
// inside the block (all of them when none starts there); bflag[b] = a segment
// starts inside the block.
__global__ __launch_bounds__(256) void k_rows_reduce_seg(const uint32_t *__restrict__ rows, uint32_t ntile,
                                                         uint32_t span, const RadixSeg *segs, int nseg,
                                                         uint32_t *__restrict__ bsum, uint32_t *__restrict__ bflag,
                                                         const uint32_t *run_if = nullptr) {
    if (run_if && *run_if == 0) return;  // gated off: runs only when *run_if != 0
    const uint32_t t0 = blockIdx.x * span, t1 = std::min(ntile, t0 + span);
    const uint32_t *col = rows + threadIdx.x;
    SegWalk wk(segs, nseg, t0);
    uint32_t tail = 0, flag = 0;
    for (uint32_t t = t0; t < t1; ++t) {
        if (wk.starts(t)) {
            tail = 0;
            flag = 1;
        }
        tail += col[(size_t)t * 256];
    }
    bsum[(size_t)blockIdx.x * 256 + threadIdx.x] = tail;
    if (threadIdx.x == 0) bflag[blockIdx.x] = flag;
}

// One block of 1024 threads: bsum[b][d] := digit-d keys of the blocks before b
// back to the start of b's first segment (exclusive carry, in place).  Thread
// (q, d) walks quarter q of the blocks; the quarters' (flag, sum) pairs combine
// in order.
__global__ __launch_bounds__(1024) void k_rows_starts_seg(uint32_t *bsum, const uint32_t *bflag, uint32_t nb,
                                                          const uint32_t *run_if = nullptr) {
    if (run_if && *run_if == 0) return;  // gated off: runs only when *run_if != 0
    __shared__ uint32_t s_part[4][256];
    __shared__ uint32_t s_flag[4];
    const uint32_t d = threadIdx.x & 255u, q = threadIdx.x >> 8;
    const uint32_t per = (nb + 3) / 4, b0 = std::min(nb, q * per), b1 = std::min(nb, b0 + per);
    uint32_t *col = bsum + d;
    uint32_t agg = 0, f = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t v = col[(size_t)b * 256];
        if (bflag[b]) {
            agg = v;
            f = 1;
        } else {
            agg += v;
        }
    }
    s_part[q][d] = agg;
    if (d == 0) s_flag[q] = f;
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t qq = 0; qq < q; ++qq) c = s_flag[qq] ? s_part[qq][d] : c + s_part[qq][d];
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t v = col[(size_t)b * 256];
        col[(size_t)b * 256] = c;
        c = bflag[b] ? v : c + v;
    }
}

// rows[t][d] := digit-d keys of the tiles of t's segment before t (in place).
__global__ __launch_bounds__(256) void k_rows_apply_seg(uint32_t *__restrict__ rows, uint32_t ntile, uint32_t span,
                                                        const RadixSeg *segs, int nseg,
                                                        const uint32_t *__restrict__ bsum,
                                                        const uint32_t *run_if = nullptr) {
    if (run_if && *run_if == 0) return;  // gated off: runs only when *run_if != 0
    const uint32_t t0 = blockIdx.x * span, t1 = std::min(ntile, t0 + span);
    uint32_t *col = rows + threadIdx.x;
    SegWalk wk(segs, nseg, t0);
    uint32_t run = bsum[(size_t)blockIdx.x * 256 + threadIdx.x];
    for (uint32_t t = t0; t < t1; ++t) {
        if (wk.starts(t)) run = 0;
        const uint32_t v = col[(size_t)t * 256];
        col[(size_t)t * 256] = run;
        run += v;
    }
}

// Places global tile xcd_tile(blockIdx.x): ranks its keys by digit (stable),
// stages them in LDS in digit order and writes each digit's run at the offset
// from its row (plus, for a batch, the digit's start in the segment's output).
template <bool HAS_VAL, bool UNPACK = false, bool MOD = false>
__global__ __launch_bounds__(256) void k_radix_place(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                     uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                     RadixSeg seg0, const RadixSeg *segs, int nseg, uint32_t ntile,
                                                     int shift, const uint32_t *__restrict__ offs,
                                                     const uint32_t *__restrict__ dstart, uint32_t sstride,
                                                     RowOut out0 = RowOut{}, const RowOut *outs = nullptr,
                                                     const uint32_t *run_if = nullptr, uint32_t mod = 0) {
    if (run_if && *run_if == 0) return;  // gated off: runs only when *run_if != 0
    __shared__ uint32_t s_cnt[4][256];
    __shared__ uint32_t s_dpre[256];
    __shared__ uint32_t s_gofs[256];
    __shared__ uint32_t s_w[4];
    __shared__ uint64_t s_k[kSortTile];
    __shared__ uint32_t s_v[HAS_VAL ? kSortTile : 1];
    const uint32_t t = xcd_tile(blockIdx.x, ntile);
    const int si = segs ? seg_index_of(segs, nseg, t) : 0;
    const RadixSeg sg = segs ? segs[si] : seg0;
    const size_t n = seg_count(sg);
    const size_t tb = (size_t)(t - sg.tile0) * kSortTile;
    if (tb >= n) return;
    const uint32_t rem = (uint32_t)std::min<size_t>(kSortTile, n - tb);
    const uint64_t *ks = kin + sg.base + tb;
    const uint32_t *vs = HAS_VAL ? vin + sg.base + tb : nullptr;
    for (int i = threadIdx.x; i < 1024; i += 256) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const unsigned w = wave_id(), lane = lane_id();
    const uint32_t wb = w * (kSortTile / 4);
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = wb + j * kWave + lane;
        const bool valid = idx < rem;
        k[j] = valid ? ks[idx] : ~uint64_t(0);
        if (HAS_VAL) v[j] = valid ? vs[idx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = wb + j * kWave + lane;
        const bool valid = idx < rem;
        const uint32_t d = key_digit<MOD>(k[j], shift, mod);
        const uint64_t m = match_digit(d, valid);
        const uint32_t pre = s_cnt[w][d];
        const uint32_t r = popc(m & lanemask_lt());
        wave_lockstep();  // every lane has read its digit's count before the leaders advance it
        if (valid && r == 0) s_cnt[w][d] = pre + popc(m);
        rank[j] = pre + r;
    }
    __syncthreads();
    {
        const unsigned d = threadIdx.x;
        const uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
        s_gofs[d] = offs[(size_t)t * 256 + d] + (dstart ? dstart[(size_t)si * sstride + d] : 0u);
        s_dpre[d] = block_excl_scan256(c0 + c1 + c2 + c3, s_w);
        s_cnt[0][d] = 0;
        s_cnt[1][d] = c0;
        s_cnt[2][d] = c0 + c1;
        s_cnt[3][d] = c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = wb + j * kWave + lane;
        if (idx < rem) {
            const uint32_t d = key_digit<MOD>(k[j], shift, mod);
            const uint32_t p = s_dpre[d] + s_cnt[w][d] + rank[j];
            s_k[p] = k[j];
            if (HAS_VAL) s_v[p] = v[j];
        }
    }
    __syncthreads();
    if constexpr (UNPACK) {
        // Last pass of a join-row sort: the segment's output columns.
        const RowOut o = outs ? outs[si] : out0;
        for (uint32_t i = threadIdx.x; i < rem; i += 256) {
            const uint64_t kk = s_k[i];
            const uint32_t d = key_digit<MOD>(kk, shift, mod);
            const uint32_t gp = s_gofs[d] + (i - s_dpre[d]);
            o.key[gp] = join_sort_key_key(kk);
            o.spay[gp] = join_sort_key_spay(kk);
            o.rpay[gp] = (int32_t)s_v[i];
        }
        return;
    }
    uint64_t *kd = kout + sg.base;
    uint32_t *vd = HAS_VAL ? vout + sg.base : nullptr;
    for (uint32_t i = threadIdx.x; i < rem; i += 256) {
        const uint64_t kk = s_k[i];
        const uint32_t d = key_digit<MOD>(kk, shift, mod);
        const uint32_t gp = s_gofs[d] + (i - s_dpre[d]);
        kd[gp] = kk;
        if (HAS_VAL) vd[gp] = s_v[i];
    }
}

// Workspace of the reduce-then-scan passes over `tiles` tiles.
struct RowsBufs {
    uint32_t *rows, *bsum;
};
static RowsBufs rows_carve(uint32_t tiles, Carver &cv) {
    RowsBufs b;
    b.rows = cv.take<uint32_t>((size_t)tiles * 256);
    // rows_plan(t).nb <= min(kRowsBlocks, ceil(t / kRowsMinSpan)) for every t <= tiles.
    b.bsum = cv.take<uint32_t>((size_t)std::min(kRowsBlocks, (tiles + kRowsMinSpan - 1) / kRowsMinSpan) * 256);
    return b;
}

static int launch_radix_rows(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt, const RadixSeg &sg,
                             int lo_bit, int hi_bit, const RowsBufs &b, hipStream_t st, bool *in_alt,
                             const RowOut *unpack = nullptr) {
    *in_alt = false;
    QO_REQUIRE(lo_bit >= 0 && hi_bit <= 64 && lo_bit < hi_bit, "bad bit range");
    if (sg.ntile == 0) return QO_OK;
    const RowsPlan rp = rows_plan(sg.ntile);
    const unsigned hblocks = (sg.ntile + kHistTiles - 1) / kHistTiles;
    uint64_t *ksrc = keys, *kdst = kalt;
    uint32_t *vsrc = vals, *vdst = valt;
    for (int sh = lo_bit; sh < hi_bit; sh += 8) {
        k_tile_hist<<<hblocks, 256, 0, st>>>(ksrc, sg, nullptr, 1, sg.ntile, sh, b.rows);
        QO_LAUNCHED();
        if (sg.ntile <= kRowsOneBlock) {
            // Few tiles: one block scans the rows themselves (one launch, not three).
            k_rows_starts<<<1, 1024, 0, st>>>(b.rows, sg.ntile);
            QO_LAUNCHED();
        } else {
            k_rows_reduce<<<rp.nb, 256, 0, st>>>(b.rows, sg.ntile, rp.span, b.bsum);
            QO_LAUNCHED();
            k_rows_starts<<<1, 1024, 0, st>>>(b.bsum, rp.nb);
            QO_LAUNCHED();
            k_rows_apply<<<rp.nb, 256, 0, st>>>(b.rows, sg.ntile, rp.span, b.bsum);
            QO_LAUNCHED();
        }
        if (vals && unpack && sh + 8 >= hi_bit)
            k_radix_place<true, true><<<sg.ntile, 256, 0, st>>>(ksrc, vsrc, kdst, vdst, sg, nullptr, 1, sg.ntile, sh,
                                                                 b.rows, nullptr, 0, *unpack, nullptr);
        else if (vals)
            k_radix_place<true><<<sg.ntile, 256, 0, st>>>(ksrc, vsrc, kdst, vdst, sg, nullptr, 1, sg.ntile, sh,
                                                           b.rows, nullptr, 0);
        else
            k_radix_place<false><<<sg.ntile, 256, 0, st>>>(ksrc, nullptr, kdst, nullptr, sg, nullptr, 1, sg.ntile,
                                                            sh, b.rows, nullptr, 0);
        QO_LAUNCHED();
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
        *in_alt = !*in_alt;
    }
    return QO_OK;
}

// ---------------------------------------------------------------------------
// A batch of arrays (one per query) in the same launches: the digit starts of
// every pass per segment come from one histogram over the keys up front
// (chunks of kSweepChunkTiles tiles); each pass then takes the tile
// histogram, a segmented column scan and the placement above.
// ---------------------------------------------------------------------------
constexpr int kSweepMaxPass = 8;

// Segmented digit counts: block c counts tiles [t0, t1) of segment `seg` of
// chunk c (segment-local tile numbers), one row of npass*256 per chunk.
__global__ __launch_bounds__(256) void k_sweep_counts_seg(const uint64_t *__restrict__ keys, const RadixSeg *segs,
                                                          const SweepChunk *chunks, int lo_bit, int npass,
                                                          uint32_t *partial, const uint32_t *run_if = nullptr) {
    if (run_if && *run_if == 0) return;  // gated off: runs only when *run_if != 0
    __shared__ uint32_t s_h[kSweepMaxPass * 256];
    for (int i = threadIdx.x; i < npass * 256; i += 256) s_h[i] = 0;
    __syncthreads();
    const SweepChunk ch = chunks[blockIdx.x];
    const RadixSeg sg = segs[ch.seg];
    const size_t n = seg_count(sg);
    const size_t e1 = std::min<size_t>((size_t)ch.t1 * kSortTile, n);
    for (size_t i = (size_t)ch.t0 * kSortTile + threadIdx.x; i < e1; i += 256) {
        const uint64_t k = ld_stream(keys + sg.base + i);
        for (int p = 0; p < npass; ++p) atomicAdd(&s_h[p * 256 + ((k >> (lo_bit + 8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npass * 256; i += 256) partial[(size_t)blockIdx.x * npass * 256 + i] = s_h[i];
}

// Per segment and pass: exclusive digit starts from the segment's chunk rows
// [chunk0[s], chunk0[s + 1]).  Block b = s * npass + p.
__global__ __launch_bounds__(256) void k_sweep_starts_seg(const uint32_t *partial, const uint32_t *chunk0, int npass,
                                                          uint32_t *starts, const uint32_t *run_if = nullptr) {
    if (run_if && *run_if == 0) return;  // gated off: runs only when *run_if != 0
    __shared__ uint32_t s_w[4];
    const int s = blockIdx.x / npass, p = blockIdx.x % npass;
    uint32_t c = 0;
    for (uint32_t r = chunk0[s]; r < chunk0[s + 1]; ++r) c += partial[(size_t)r * npass * 256 + p * 256 + threadIdx.x];
    starts[(size_t)blockIdx.x * 256 + threadIdx.x] = block_excl_scan256(c, s_w);
}

static uint32_t radix_tiles(size_t n) { return (uint32_t)((n + kSortTile - 1) / kSortTile); }

// Workspace of launch_radix_rows_segments (same order as its carve-up).
size_t radix_rows_seg_bytes(uint32_t total_tiles, int nseg, uint32_t nchunks) {
    Carver cv(nullptr);
    cv.take<uint32_t>((size_t)nchunks * kSweepMaxPass * 256);
    cv.take<uint32_t>((size_t)nseg * kSweepMaxPass * 256);
    cv.take<uint32_t>((size_t)total_tiles * 256);
    const uint32_t nbmax = std::min(kRowsBlocks, (total_tiles + kRowsMinSpan - 1) / kRowsMinSpan);
    cv.take<uint32_t>((size_t)nbmax * 256);
    cv.take<uint32_t>(nbmax);
    return cv.off + 256;
}

int launch_radix_rows_segments(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt,
                               const RadixSeg *d_segs, int nseg, uint32_t total_tiles, const SweepChunk *d_chunks,
                               uint32_t nchunks, const uint32_t *d_chunk0, int lo_bit, int hi_bit, void *ws,
                               hipStream_t st, bool *in_alt, const RowOut *d_outs, const uint32_t *run_if) {
    *in_alt = false;
    if (total_tiles == 0 || nchunks == 0) return QO_OK;
    QO_REQUIRE(!d_outs || vals, "unpacked output needs values");
    const int npass = (hi_bit - lo_bit + 7) / 8;
    QO_REQUIRE(npass >= 1 && npass <= kSweepMaxPass && lo_bit >= 0 && hi_bit <= 64, "bad bit range");
    QO_REQUIRE(nseg >= 1 && (size_t)nseg * npass < 0x7FFFFFFFull, "bad segment count");
    QO_REQUIRE((size_t)total_tiles * 256 < 0xFFFFFFFFull, "sort input too large");
    Carver cv(ws);
    uint32_t *partial = cv.take<uint32_t>((size_t)nchunks * npass * 256);
    uint32_t *starts = cv.take<uint32_t>((size_t)nseg * npass * 256);
    uint32_t *rows = cv.take<uint32_t>((size_t)total_tiles * 256);
    const RowsPlan rp = rows_plan(total_tiles);
    uint32_t *bsum = cv.take<uint32_t>((size_t)rp.nb * 256);
    uint32_t *bflag = cv.take<uint32_t>(rp.nb);
    // Digit starts of every pass per segment, from one read of the keys.
    k_sweep_counts_seg<<<nchunks, 256, 0, st>>>(keys, d_segs, d_chunks, lo_bit, npass, partial, run_if);
    QO_LAUNCHED();
    k_sweep_starts_seg<<<(unsigned)(nseg * npass), 256, 0, st>>>(partial, d_chunk0, npass, starts, run_if);
    QO_LAUNCHED();
    const RadixSeg none{0, 0, nullptr, 0, 0};
    const unsigned hblocks = (total_tiles + kHistTiles - 1) / kHistTiles;
    uint64_t *ksrc = keys, *kdst = kalt;
    uint32_t *vsrc = vals, *vdst = valt;
    for (int p = 0; p < npass; ++p) {
        const int sh = lo_bit + 8 * p;
        k_tile_hist<<<hblocks, 256, 0, st>>>(ksrc, none, d_segs, nseg, total_tiles, sh, rows, run_if);
        QO_LAUNCHED();
        k_rows_reduce_seg<<<rp.nb, 256, 0, st>>>(rows, total_tiles, rp.span, d_segs, nseg, bsum, bflag, run_if);
        QO_LAUNCHED();
        k_rows_starts_seg<<<1, 1024, 0, st>>>(bsum, bflag, rp.nb, run_if);
        QO_LAUNCHED();
        k_rows_apply_seg<<<rp.nb, 256, 0, st>>>(rows, total_tiles, rp.span, d_segs, nseg, bsum, run_if);
        QO_LAUNCHED();
        if (vals && d_outs && p == npass - 1)
            k_radix_place<true, true><<<total_tiles, 256, 0, st>>>(ksrc, vsrc, kdst, vdst, none, d_segs, nseg,
                                                                   total_tiles, sh, rows, starts + p * 256,
                                                                   (uint32_t)npass * 256, RowOut{}, d_outs, run_if);
        else if (vals)
            k_radix_place<true><<<total_tiles, 256, 0, st>>>(ksrc, vsrc, kdst, vdst, none, d_segs, nseg, total_tiles,
                                                             sh, rows, starts + p * 256, (uint32_t)npass * 256,
                                                             RowOut{}, nullptr, run_if);
        else
            k_radix_place<false><<<total_tiles, 256, 0, st>>>(ksrc, nullptr, kdst, nullptr, none, d_segs, nseg,
                                                              total_tiles, sh, rows, starts + p * 256,
                                                              (uint32_t)npass * 256, RowOut{}, nullptr, run_if);
        QO_LAUNCHED();
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
        *in_alt = !*in_alt;
    }
    return QO_OK;
}

static void radix_carve(size_t n, bool with_vals, Carver &cv, uint64_t **kalt, uint32_t **valt, RowsBufs *rb) {
    const uint32_t tiles = radix_tiles(n) + 1;
    *kalt = cv.take<uint64_t>(n ? n : 1);
    *valt = with_vals ? cv.take<uint32_t>(n ? n : 1) : nullptr;
    *rb = rows_carve(tiles, cv);
}

size_t radix_workspace_bytes(size_t n, bool with_vals) {
    Carver cv(nullptr);
    uint64_t *kalt;
    uint32_t *valt;
    RowsBufs rb;
    radix_carve(n, with_vals, cv, &kalt, &valt, &rb);
    return cv.off + 256;
}

static int radix_impl(uint64_t *keys, uint32_t *vals, size_t n, const uint64_t *d_n, int lo_bit, int hi_bit,
                      void *ws, hipStream_t st, const RowOut *unpack = nullptr) {
    if (n == 0) return QO_OK;
    QO_REQUIRE(n < 0xFFFFFFFFull, "sort input too large");
    QO_REQUIRE(lo_bit >= 0 && hi_bit <= 64 && lo_bit < hi_bit, "bad bit range");
    const uint32_t tiles = radix_tiles(n);
    Carver cv(ws);
    uint64_t *kalt = nullptr;
    uint32_t *valt = nullptr;
    RowsBufs rb{};
    radix_carve(n, vals != nullptr, cv, &kalt, &valt, &rb);
    RadixSeg seg{0, n, d_n, 0, tiles};
    bool in_alt = false;
    QO_TRY(launch_radix_rows(keys, vals, kalt, valt, seg, lo_bit, hi_bit, rb, st, &in_alt, unpack));
    if (in_alt && !unpack) {
        // Odd number of passes: the result is in the alternate buffers.  The
        // count may only be known on the device, so copy the full capacity.
        QO_HIP(hipMemcpyAsync(keys, kalt, n * 8, hipMemcpyDeviceToDevice, st));
        if (vals) QO_HIP(hipMemcpyAsync(vals, valt, n * 4, hipMemcpyDeviceToDevice, st));
    }
    return QO_OK;
}

int launch_radix_sort_u64(uint64_t *keys, uint32_t *vals, size_t n, int lo_bit, int hi_bit, void *ws,
                          hipStream_t st) {
    return radix_impl(keys, vals, n, nullptr, lo_bit, hi_bit, ws, st);
}

int launch_radix_sort_u64_dev(uint64_t *keys, uint32_t *vals, const uint64_t *d_n, size_t capacity, int lo_bit,
                              int hi_bit, void *ws, hipStream_t st) {
    return radix_impl(keys, vals, capacity, d_n, lo_bit, hi_bit, ws, st);
}

// ---------------------------------------------------------------------------
// Split of Records into `parts` partitions by pid = R.y % parts (stable within a
// partition): the radix passes above with the partition id as the key, one pass
// for parts <= 256, two (low, then high byte of the pid) up to 65536; then the
// partition bounds by binary search over the split output.
// ---------------------------------------------------------------------------
constexpr uint32_t kSplitMaxParts = 65536;

// bound[p] = {start, end} of partition p: the first output position whose pid
// is >= p, and the same for p + 1 (n for the last partition).
__global__ __launch_bounds__(256) void k_split_bounds(const uint64_t *__restrict__ out, size_t n, uint32_t parts,
                                                      uint2 *__restrict__ bound) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= parts) return;
    auto first_at_least = [&](uint32_t q) {
        size_t lo = 0, hi = n;
        while (lo < hi) {
            const size_t mid = lo + (hi - lo) / 2;
            if ((uint32_t)(out[mid] >> 32) % parts < q) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    bound[p] = uint2{(uint32_t)first_at_least(p), p + 1 < parts ? (uint32_t)first_at_least(p + 1) : (uint32_t)n};
}

size_t split_records_ws_bytes(size_t n, uint32_t parts) {
    Carver cv(nullptr);
    if (parts > 256) cv.take<uint64_t>(n ? n : 1);
    rows_carve(radix_tiles(n) + 1, cv);
    return cv.off + 256;
}

int launch_split_records(const uint64_t *in, size_t n, uint32_t parts, uint64_t *out, uint2 *bound, void *ws,
                         hipStream_t st) {
    QO_REQUIRE(parts >= 1 && parts <= kSplitMaxParts, "numPart %u outside [1, %u]", parts, kSplitMaxParts);
    QO_REQUIRE(n < 0xFFFFFFFFull, "split input too large");
    if (n > 0) {
        Carver cv(ws);
        uint64_t *alt = parts > 256 ? cv.take<uint64_t>(n) : nullptr;
        const uint32_t tiles = radix_tiles(n);
        const RowsBufs b = rows_carve(tiles + 1, cv);
        const RadixSeg sg{0, n, nullptr, 0, tiles};
        const RowsPlan rp = rows_plan(tiles);
        const unsigned hblocks = (tiles + kHistTiles - 1) / kHistTiles;
        const int npass = parts > 256 ? 2 : 1;
        for (int pass = 0; pass < npass; ++pass) {
            const uint64_t *src = pass == 0 ? in : alt;
            uint64_t *dst = pass == npass - 1 ? out : alt;
            const int sh = 8 * pass;
            k_tile_hist<true><<<hblocks, 256, 0, st>>>(src, sg, nullptr, 1, tiles, sh, b.rows, nullptr, parts);
            QO_LAUNCHED();
            if (tiles <= kRowsOneBlock) {
                k_rows_starts<<<1, 1024, 0, st>>>(b.rows, tiles);
                QO_LAUNCHED();
            } else {
                k_rows_reduce<<<rp.nb, 256, 0, st>>>(b.rows, tiles, rp.span, b.bsum);
                QO_LAUNCHED();
                k_rows_starts<<<1, 1024, 0, st>>>(b.bsum, rp.nb);
                QO_LAUNCHED();
                k_rows_apply<<<rp.nb, 256, 0, st>>>(b.rows, tiles, rp.span, b.bsum);
                QO_LAUNCHED();
            }
            k_radix_place<false, false, true><<<tiles, 256, 0, st>>>(src, nullptr, dst, nullptr, sg, nullptr, 1, tiles,
                                                                      sh, b.rows, nullptr, 0, RowOut{}, nullptr,
                                                                      nullptr, parts);
            QO_LAUNCHED();
        }
    }
    if (bound) {
        k_split_bounds<<<(parts + 255) / 256, 256, 0, st>>>(out, n, parts, bound);
        QO_LAUNCHED();
    }
    return QO_OK;
}

// (key, s_payload) as signed int32 -> order-preserving unsigned 64-bit key.
__global__ void k_pack_join(const int32_t *key, const int32_t *spay, const int32_t *rpay, size_t m, uint64_t *k64,
                            uint32_t *v32) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) {
        k64[i] = join_sort_key(key[i], spay[i]);
        v32[i] = (uint32_t)rpay[i];
    }
}

__global__ void k_unpack_join(const uint64_t *k64, const uint32_t *v32, size_t m, int32_t *key, int32_t *spay,
                              int32_t *rpay) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) {
        uint64_t k = k64[i];
        key[i] = join_sort_key_key(k);
        spay[i] = join_sort_key_spay(k);
        rpay[i] = (int32_t)v32[i];
    }
}

int sort_join_rows_ws(int32_t *key, int32_t *rpay, int32_t *spay, size_t m, void *ws, hipStream_t st) {
    if (m < 2) return QO_OK;
    Carver cv(ws);
    uint64_t *k64 = cv.take<uint64_t>(m);
    uint32_t *v32 = cv.take<uint32_t>(m);
    void *rws = cv.take<char>(radix_workspace_bytes(m, true));
    unsigned g = grid_for(m, 256, (unsigned)props().cu_count * 8);
    uint32_t *flag = cv.take<uint32_t>(1);
    k_pack_join<<<g, 256, 0, st>>>(key, spay, rpay, m, k64, v32);
    QO_LAUNCHED();
    return sort_packed_join_rows(k64, v32, m, rws, key, rpay, spay, st, 64, flag);
}

size_t sort_join_rows_ws_bytes(size_t m) { return m * 12 + 2048 + radix_workspace_bytes(m, true); }

// Runs of equal key after a stable sort of the packed join rows on their key
// bits (the high 32): each block owns kFixTile positions and loads kFixRun more,
// so that every run that starts in its range and has at most kFixRun rows lies
// in LDS whole.  One ballot per 64-position chunk marks where runs start; a
// row finds its run's bounds from its own chunk's marks and its neighbours',
// counts the run's rows that order before it (smaller s_payload, or equal and
// earlier: stable) and writes itself to the output columns at run start + that
// count.  Rows of runs that start outside the block's range are left to the
// block that owns the start; a run longer than kFixRun sets *long_flag and is
// not written.
constexpr uint32_t kFixTile = 4096;
constexpr uint32_t kFixRun = 64;
constexpr uint32_t kFixWin = kFixTile + kFixRun;
constexpr uint32_t kFixThreads = 512;  // four blocks per CU (34 KB LDS each), 8 waves per SIMD
constexpr uint32_t kFixRounds = (kFixWin + kFixThreads - 1) / kFixThreads;
constexpr uint32_t kFixChunks = kFixWin / kWave + 2;  // marks of the window's chunks and one past it

// Block work of the run pass: positions [t0, t0 + kFixTile) of m rows (k, v).
QO_DEV void fix_runs_tile(const uint64_t *__restrict__ k, const uint32_t *__restrict__ v, size_t m, size_t t0,
                          const RowOut &out, uint32_t *long_flag) {
    static_assert(kFixWin % kWave == 0, "window of whole chunks");
    __shared__ uint32_t s_hi[kFixWin];
    __shared__ uint32_t s_lo[kFixWin + 4];  // 4 words of slack for the unrolled rank loop
    __shared__ uint64_t s_mark[kFixChunks];
    const uint32_t own = (uint32_t)(m - t0 < kFixTile ? m - t0 : kFixTile);
    const uint32_t len = (uint32_t)(m - t0 < kFixWin ? m - t0 : kFixWin);
    const bool data_end = t0 + len == m;  // the window reaches the last row
#pragma unroll
    for (uint32_t u = 0; u < kFixRounds; ++u) {
        const uint32_t i = u * kFixThreads + threadIdx.x;
        if (i < len) {
            const uint64_t x = k[t0 + i];
            s_hi[i] = (uint32_t)(x >> 32);
            s_lo[i] = (uint32_t)x;
        }
    }
    const uint32_t prev_key = t0 > 0 ? (uint32_t)(k[t0 - 1] >> 32) : 0u;
    __syncthreads();
    // Run-start marks: position p starts a run if its key differs from the row
    // before it (the first row of all rows starts one); past the last row every
    // position counts as a start, past a cut-off window none does.
    const unsigned lane = lane_id();
    for (uint32_t c = wave_id(); c < kFixChunks; c += kFixThreads / kWave) {
        const uint32_t p = c * kWave + lane;
        bool start;
        if (p < len) start = p == 0 ? (t0 == 0 || prev_key != s_hi[0]) : s_hi[p] != s_hi[p - 1];
        else start = data_end;
        const uint64_t mk = ballot(start);
        if (lane == 0) s_mark[c] = mk;
    }
    __syncthreads();
    bool long_run = false;
    for (uint32_t i = threadIdx.x; i < len; i += kFixThreads) {
        const uint32_t c = i / kWave;  // uniform across the wave
        const uint64_t at_or_below = lane == kWave - 1 ? ~0ull : (2ull << lane) - 1;
        const uint64_t mc = s_mark[c];
        uint32_t s;
        if (mc & at_or_below) {
            s = c * kWave + 63u - (uint32_t)__builtin_clzll(mc & at_or_below);
        } else if (c == 0) {
            continue;  // the run began before the block's range
        } else if (s_mark[c - 1]) {
            s = (c - 1) * kWave + 63u - (uint32_t)__builtin_clzll(s_mark[c - 1]);
        } else {
            long_run = true;  // no start within 64 positions back
            continue;
        }
        if (s >= own) continue;  // the next block's run
        const uint64_t above = mc & ~at_or_below;
        uint32_t e;
        if (above) {
            e = c * kWave + (uint32_t)__builtin_ctzll(above);
        } else if (s_mark[c + 1]) {
            e = (c + 1) * kWave + (uint32_t)__builtin_ctzll(s_mark[c + 1]);
        } else {
            long_run = true;
            continue;
        }
        if (e - s > kFixRun) {
            long_run = true;
            continue;
        }
        const uint32_t xl = s_lo[i];
        uint32_t rank = 0;
        for (uint32_t j = s; j < e; j += 4) {
            const uint32_t y0 = s_lo[j], y1 = s_lo[j + 1], y2 = s_lo[j + 2], y3 = s_lo[j + 3];
            rank += (y0 < xl || (y0 == xl && j < i)) ? 1u : 0u;
            rank += (j + 1 < e && (y1 < xl || (y1 == xl && j + 1 < i))) ? 1u : 0u;
            rank += (j + 2 < e && (y2 < xl || (y2 == xl && j + 2 < i))) ? 1u : 0u;
            rank += (j + 3 < e && (y3 < xl || (y3 == xl && j + 3 < i))) ? 1u : 0u;
        }
        const size_t dst = t0 + s + rank;
        const uint64_t x = ((uint64_t)s_hi[i] << 32) | xl;
        out.key[dst] = join_sort_key_key(x);
        out.spay[dst] = join_sort_key_spay(x);
        out.rpay[dst] = (int32_t)v[t0 + i];
    }
    if (long_run) atomicOr(long_flag, 1u);
}

__global__ __launch_bounds__(kFixThreads) void k_fix_runs(const uint64_t *__restrict__ k, const uint32_t *__restrict__ v,
                                                          size_t m, RowOut out, uint32_t *long_flag) {
    fix_runs_tile(k, v, m, (size_t)blockIdx.x * kFixTile, out, long_flag);
}

// Segmented form: block b is global tile b of the segments (kSortTile rows per
// tile; n = *d_n clamped to cap rows per segment).
__global__ __launch_bounds__(kFixThreads) void k_fix_runs_seg(const uint64_t *__restrict__ k,
                                                              const uint32_t *__restrict__ v, const RadixSeg *segs,
                                                              int nseg, const RowOut *outs, uint32_t *long_flag) {
    const int si = seg_index_of(segs, nseg, blockIdx.x);
    const RadixSeg sg = segs[si];
    uint64_t n = sg.d_n ? *sg.d_n : sg.cap;
    if (n > sg.cap) n = sg.cap;
    const size_t t0 = (size_t)(blockIdx.x - sg.tile0) * kFixTile;
    if (t0 >= n) return;  // whole block
    fix_runs_tile(k + sg.base, v + sg.base, n, t0, outs[si], long_flag);
}

int launch_fix_runs_segments(const uint64_t *keys, const uint32_t *vals, const RadixSeg *d_segs, int nseg,
                             uint32_t total_tiles, const RowOut *d_outs, uint32_t *long_flag, hipStream_t st) {
    static_assert(kFixTile == (uint32_t)kSortTile, "run pass tiles are sort tiles");
    if (total_tiles == 0) return QO_OK;
    k_fix_runs_seg<<<total_tiles, kFixThreads, 0, st>>>(keys, vals, d_segs, nseg, d_outs, long_flag);
    QO_LAUNCHED();
    return QO_OK;
}

int sort_packed_join_rows(uint64_t *k64, uint32_t *v32, size_t m, void *rws, int32_t *key, int32_t *rpay,
                          int32_t *spay, hipStream_t st, int hi_bit, uint32_t *flag) {
    if (m == 0) return QO_OK;
    if (m == 1) {
        k_unpack_join<<<1, 256, 0, st>>>(k64, v32, m, key, spay, rpay);
        QO_LAUNCHED();
        return QO_OK;
    }
    hi_bit = std::max(1, std::min(64, hi_bit));
    // The last radix pass writes the output columns (no separate unpack pass).
    const RowOut out{key, rpay, spay};
    if (!flag || hi_bit <= 32) return radix_impl(k64, v32, m, nullptr, 0, hi_bit, rws, st, &out);
    // Radix passes over the key bits only (stable: a key's rows keep the probe's
    // order), then one pass that sorts every run of equal key by s_payload and
    // writes the output columns: for 2^24 build keys 3 passes + 1 instead of 7.
    QO_REQUIRE(m < 0xFFFFFFFFull, "sort input too large");
    Carver cv(rws);
    uint64_t *kalt = nullptr;
    uint32_t *valt = nullptr;
    RowsBufs rb{};
    radix_carve(m, true, cv, &kalt, &valt, &rb);
    const RadixSeg seg{0, m, nullptr, 0, radix_tiles(m)};
    bool in_alt = false;
    QO_TRY(launch_radix_rows(k64, v32, kalt, valt, seg, 32, hi_bit, rb, st, &in_alt, nullptr));
    QO_TRY(launch_fill_u32(flag, 0, 1, st));
    const size_t blocks = (m + kFixTile - 1) / kFixTile;
    k_fix_runs<<<(unsigned)blocks, kFixThreads, 0, st>>>(in_alt ? kalt : k64, in_alt ? valt : v32, m, out, flag);
    QO_LAUNCHED();
    uint32_t long_run = 0;
    QO_HIP(hipMemcpyAsync(&long_run, flag, 4, hipMemcpyDeviceToHost, st));
    QO_HIP(hipStreamSynchronize(st));
    if (!long_run) return QO_OK;
    // Some key has more than kFixRun rows: sort all rows on the full width
    // (the result does not depend on the rows' current order).
    if (in_alt) {
        QO_HIP(hipMemcpyAsync(k64, kalt, m * 8, hipMemcpyDeviceToDevice, st));
        QO_HIP(hipMemcpyAsync(v32, valt, m * 4, hipMemcpyDeviceToDevice, st));
    }
    return radix_impl(k64, v32, m, nullptr, 0, hi_bit, rws, st, &out);
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_sort_records(qo_record *data, size_t n, qo_stream s) {
    QO_ENTER();
    if (n < 2) return QO_OK;
    QO_REQUIRE(data && ((uintptr_t)data % 8) == 0, "records must be 8-byte aligned device memory");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, radix_workspace_bytes(n, false), &status);
    if (!ws) return status;
    // A Record {x, y} read as little-endian uint64 has y in bits 32..63.
    return launch_radix_sort_u64(reinterpret_cast<uint64_t *>(data), nullptr, n, 32, 64, ws, st);
}

int qo_sort_join_rows(int32_t *key, int32_t *rpay, int32_t *spay, size_t m, qo_stream s) {
    QO_ENTER();
    if (m < 2) return QO_OK;
    QO_REQUIRE(key && rpay && spay, "null pointer");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, sort_join_rows_ws_bytes(m), &status);
    if (!ws) return status;
    return sort_join_rows_ws(key, rpay, spay, m, ws, st);
}

}  // extern "C"
