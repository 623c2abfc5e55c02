// Sorting: a stable LSD radix sort over 64-bit keys (8-bit digits) with an
// optional 32-bit value, three launches per digit: per-tile digit histograms,
// one exclusive scan over the digit-major histogram, and a scatter that ranks
// keys with wave-level match masks (8 ballots per key), stages the tile in LDS
// in digit order and writes contiguous runs.
//
// The sort is segmented: many independent arrays (one per query of a batch)
// are sorted by the same launches.  A segment owns a contiguous tile range; its
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

// Segment that owns global tile `tile` (segments are ordered by tile0).
QO_DEV RadixSeg seg_of(const RadixSeg &seg0, const RadixSeg *segs, int nseg, uint32_t tile) {
    if (nseg <= 1) return seg0;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return segs[lo];
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

__global__ __launch_bounds__(256) void k_radix_hist(const uint64_t *__restrict__ kin, RadixSeg seg0,
                                                    const RadixSeg *segs, int nseg, int shift, uint32_t *hist) {
    __shared__ uint32_t s_h[4][256];
    const uint32_t tile = blockIdx.x;
    const RadixSeg sg = seg_of(seg0, segs, nseg, tile);
    const uint32_t lt = tile - sg.tile0;
    const size_t n = seg_count(sg);
    for (int i = threadIdx.x; i < 1024; i += 256) (&s_h[0][0])[i] = 0;
    __syncthreads();
    const unsigned w = wave_id(), lane = lane_id();
    const size_t wbase = (size_t)lt * kSortTile + (size_t)w * (kSortTile / 4);
    if ((size_t)lt * kSortTile < n) {
#pragma unroll 4
        for (int j = 0; j < kSortItems; ++j) {
            size_t idx = wbase + (size_t)j * kWave + lane;
            bool valid = idx < n;
            uint32_t d = valid ? (uint32_t)(kin[sg.base + idx] >> shift) & 255u : 0u;
            uint64_t m = match_digit(d, valid);
            if (valid && popc(m & lanemask_lt()) == 0) s_h[w][d] += popc(m);
        }
    }
    __syncthreads();
    const unsigned d = threadIdx.x;
    hist[(size_t)sg.tile0 * 256 + (size_t)d * sg.ntile + lt] = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
}

template <bool HAS_VAL>
__global__ __launch_bounds__(256) void k_radix_scatter(const uint64_t *__restrict__ kin,
                                                       const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
                                                       uint32_t *__restrict__ vout, RadixSeg seg0,
                                                       const RadixSeg *segs, int nseg, int shift,
                                                       const uint32_t *__restrict__ offs) {
    __shared__ uint32_t s_cnt[4][256];
    __shared__ uint32_t s_dpre[256];
    __shared__ uint32_t s_gofs[256];
    __shared__ uint32_t s_w[4];
    __shared__ uint64_t s_k[kSortTile];
    __shared__ uint32_t s_v[HAS_VAL ? kSortTile : 1];
    const uint32_t tile = blockIdx.x;
    const RadixSeg sg = seg_of(seg0, segs, nseg, tile);
    const uint32_t lt = tile - sg.tile0;
    const size_t n = seg_count(sg);
    if ((size_t)lt * kSortTile >= n) return;
    for (int i = threadIdx.x; i < 1024; i += 256) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const unsigned w = wave_id(), lane = lane_id();
    const size_t wbase = (size_t)lt * kSortTile + (size_t)w * (kSortTile / 4);
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        size_t idx = wbase + (size_t)j * kWave + lane;
        bool valid = idx < n;
        k[j] = valid ? kin[sg.base + idx] : ~uint64_t(0);
        if (HAS_VAL) v[j] = valid ? vin[sg.base + idx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        size_t idx = wbase + (size_t)j * kWave + lane;
        bool valid = idx < n;
        uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
        uint64_t m = match_digit(d, valid);
        uint32_t pre = s_cnt[w][d];
        uint32_t r = popc(m & lanemask_lt());
        if (valid && r == 0) s_cnt[w][d] = pre + popc(m);
        rank[j] = pre + r;
    }
    __syncthreads();
    {
        const unsigned d = threadIdx.x;
        uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
        uint32_t tot = c0 + c1 + c2 + c3;
        const size_t h0 = (size_t)sg.tile0 * 256;
        s_gofs[d] = offs[h0 + (size_t)d * sg.ntile + lt] - offs[h0];
        uint32_t ex = block_excl_scan256(tot, s_w);
        s_dpre[d] = ex;
        s_cnt[0][d] = 0;
        s_cnt[1][d] = c0;
        s_cnt[2][d] = c0 + c1;
        s_cnt[3][d] = c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        size_t idx = wbase + (size_t)j * kWave + lane;
        if (idx < n) {
            uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
            uint32_t p = s_dpre[d] + s_cnt[w][d] + rank[j];
            s_k[p] = k[j];
            if (HAS_VAL) s_v[p] = v[j];
        }
    }
    __syncthreads();
    const size_t tstart = (size_t)lt * kSortTile;
    const uint32_t tcount = (uint32_t)((n - tstart) < (size_t)kSortTile ? (n - tstart) : (size_t)kSortTile);
    for (uint32_t i = threadIdx.x; i < tcount; i += 256) {
        uint64_t kk = s_k[i];
        uint32_t d = (uint32_t)(kk >> shift) & 255u;
        size_t gp = sg.base + (size_t)s_gofs[d] + (i - s_dpre[d]);
        kout[gp] = kk;
        if (HAS_VAL) vout[gp] = s_v[i];
    }
}

size_t radix_hist_bytes(uint32_t total_tiles) { return (size_t)total_tiles * 256 * 4 + 256; }

// ---------------------------------------------------------------------------
// One-sweep passes for a single array: the digit counts of every pass come
// from one histogram over the keys up front; each pass is then a single
// kernel that ranks its tile, publishes the tile's per-digit counts and finds
// the counts of all earlier tiles with a decoupled look-back per digit (thread
// d of the block walks digit d), so no per-pass histogram read or scan runs.
// Tiles are numbered by a ticket: every tile a block waits for belongs to a
// block that is already running.
// ---------------------------------------------------------------------------
constexpr int kSweepMaxPass = 8;
constexpr int kSweepCountBlocks = 512;

// Per-block digit counts of every pass (LDS), one row of npass*256 per block.
__global__ __launch_bounds__(256) void k_sweep_counts(const uint64_t *__restrict__ keys, RadixSeg sg, int lo_bit,
                                                      int npass, uint32_t *partial) {
    __shared__ uint32_t s_h[kSweepMaxPass * 256];
    for (int i = threadIdx.x; i < npass * 256; i += 256) s_h[i] = 0;
    __syncthreads();
    const size_t n = seg_count(sg);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const uint64_t k = ld_stream(keys + sg.base + i);
        for (int p = 0; p < npass; ++p) atomicAdd(&s_h[p * 256 + ((k >> (lo_bit + 8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npass * 256; i += 256) partial[(size_t)blockIdx.x * npass * 256 + i] = s_h[i];
}

// counts[p*256 + d] = sum over blocks of partial[b][p][d]: exclusive digit starts
// per pass (the digit order is the output order), written in place.
__global__ __launch_bounds__(256) void k_sweep_starts(const uint32_t *partial, int nblocks, int npass,
                                                      uint32_t *starts) {
    __shared__ uint32_t s_w[4];
    const int p = blockIdx.x;
    if (p >= npass) return;
    const uint32_t *col = partial + p * 256 + threadIdx.x;
    const size_t row = (size_t)npass * 256;
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // independent sums: loads stay in flight together
    int b = 0;
    for (; b + 8 <= nblocks; b += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] += col[(size_t)(b + u) * row];
    for (; b < nblocks; ++b) c[0] += col[(size_t)b * row];
    const uint32_t t = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7];
    starts[p * 256 + threadIdx.x] = block_excl_scan256(t, s_w);
}

// Segmented digit counts: block c counts tiles [t0, t1) of segment `seg` of
// chunk c (segment-local tile numbers), one row of npass*256 per chunk.
__global__ __launch_bounds__(256) void k_sweep_counts_seg(const uint64_t *__restrict__ keys, const RadixSeg *segs,
                                                          const SweepChunk *chunks, int lo_bit, int npass,
                                                          uint32_t *partial) {
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
                                                          uint32_t *starts) {
    __shared__ uint32_t s_w[4];
    const int s = blockIdx.x / npass, p = blockIdx.x % npass;
    uint32_t c = 0;
    for (uint32_t r = chunk0[s]; r < chunk0[s + 1]; ++r) c += partial[(size_t)r * npass * 256 + p * 256 + threadIdx.x];
    starts[(size_t)blockIdx.x * 256 + threadIdx.x] = block_excl_scan256(c, s_w);
}

template <bool HAS_VAL>
__global__ __launch_bounds__(256) void k_radix_sweep(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                     uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                     RadixSeg seg0, const RadixSeg *segs, int nseg, int shift,
                                                     const uint32_t *__restrict__ starts, uint32_t sstride,
                                                     uint64_t *status, uint32_t *ticket) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[4][256];
    __shared__ uint32_t s_dpre[256];
    __shared__ uint32_t s_gofs[256];
    __shared__ uint32_t s_w[4];
    __shared__ uint64_t s_k[kSortTile];
    __shared__ uint32_t s_v[HAS_VAL ? kSortTile : 1];
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    for (int i = threadIdx.x; i < 1024; i += 256) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;  // global tile; the segment's tiles are [sg.tile0, sg.tile0 + sg.ntile)
    const int si = seg_index_of(segs, nseg, tile);
    const RadixSeg sg = segs ? segs[si] : seg0;
    const uint32_t lt = tile - sg.tile0;
    const uint32_t *dstart = starts + (size_t)si * sstride;
    const size_t n = seg_count(sg);
    if ((size_t)lt * kSortTile >= n) return;  // whole block; no later tile of the segment waits for it
    const unsigned w = wave_id(), lane = lane_id();
    const size_t wbase = (size_t)lt * kSortTile + (size_t)w * (kSortTile / 4);
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        size_t idx = wbase + (size_t)j * kWave + lane;
        bool valid = idx < n;
        k[j] = valid ? kin[sg.base + idx] : ~uint64_t(0);
        if (HAS_VAL) v[j] = valid ? vin[sg.base + idx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        size_t idx = wbase + (size_t)j * kWave + lane;
        bool valid = idx < n;
        uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
        uint64_t m = match_digit(d, valid);
        uint32_t pre = s_cnt[w][d];
        uint32_t r = popc(m & lanemask_lt());
        if (valid && r == 0) s_cnt[w][d] = pre + popc(m);
        rank[j] = pre + r;
    }
    __syncthreads();
    {
        const unsigned d = threadIdx.x;
        uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
        uint32_t tot = c0 + c1 + c2 + c3;
        // Publish this tile's count of digit d, then add up the earlier tiles'.
        uint64_t *my = status + (size_t)tile * 256 + d;
        uint64_t excl = 0;
        if (lt == 0) {
            st_agent(my, kFlagInc | tot);
        } else {
            st_agent(my, kFlagAgg | tot);
            for (int64_t tp = (int64_t)tile - 1; tp >= (int64_t)sg.tile0; --tp) {
                uint64_t st;
                while (((st = ld_agent(status + (size_t)tp * 256 + d)) >> 62) == 0) __builtin_amdgcn_s_sleep(1);
                excl += st & kValMask;
                if ((st >> 62) == 2) break;
            }
            st_agent(my, kFlagInc | ((excl + tot) & kValMask));
        }
        s_gofs[d] = dstart[d] + (uint32_t)excl;
        uint32_t ex = block_excl_scan256(tot, s_w);
        s_dpre[d] = ex;
        s_cnt[0][d] = 0;
        s_cnt[1][d] = c0;
        s_cnt[2][d] = c0 + c1;
        s_cnt[3][d] = c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        size_t idx = wbase + (size_t)j * kWave + lane;
        if (idx < n) {
            uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
            uint32_t p = s_dpre[d] + s_cnt[w][d] + rank[j];
            s_k[p] = k[j];
            if (HAS_VAL) s_v[p] = v[j];
        }
    }
    __syncthreads();
    const size_t tstart = (size_t)lt * kSortTile;
    const uint32_t tcount = (uint32_t)((n - tstart) < (size_t)kSortTile ? (n - tstart) : (size_t)kSortTile);
    for (uint32_t i = threadIdx.x; i < tcount; i += 256) {
        uint64_t kk = s_k[i];
        uint32_t d = (uint32_t)(kk >> shift) & 255u;
        size_t gp = sg.base + (size_t)s_gofs[d] + (i - s_dpre[d]);
        kout[gp] = kk;
        if (HAS_VAL) vout[gp] = s_v[i];
    }
}

// Workspace of the one-sweep passes over `tiles` tiles and npass passes.
struct SweepBufs {
    uint32_t *partial, *starts, *ticket;
    uint64_t *status;
};
static SweepBufs sweep_carve(uint32_t tiles, int npass, Carver &cv) {
    SweepBufs b;
    b.partial = cv.take<uint32_t>((size_t)kSweepCountBlocks * npass * 256);
    b.starts = cv.take<uint32_t>((size_t)npass * 256);
    b.ticket = cv.take<uint32_t>(64);
    b.status = cv.take<uint64_t>((size_t)tiles * 256);
    return b;
}

static int launch_radix_sweep(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt, const RadixSeg &sg,
                              int lo_bit, int hi_bit, const SweepBufs &b, hipStream_t st, bool *in_alt) {
    *in_alt = false;
    const int npass = (hi_bit - lo_bit + 7) / 8;
    QO_REQUIRE(npass >= 1 && npass <= kSweepMaxPass, "bad bit range");
    // At least 64K keys per counting block (capacity bound when the count is on the device).
    const int cblocks = (int)std::min<uint64_t>(kSweepCountBlocks, std::max<uint64_t>(1, (sg.cap + 65535) / 65536));
    k_sweep_counts<<<cblocks, 256, 0, st>>>(keys, sg, lo_bit, npass, b.partial);
    QO_LAUNCHED();
    k_sweep_starts<<<npass, 256, 0, st>>>(b.partial, cblocks, npass, b.starts);
    QO_LAUNCHED();
    uint64_t *ksrc = keys, *kdst = kalt;
    uint32_t *vsrc = vals, *vdst = valt;
    for (int p = 0; p < npass; ++p) {
        QO_TRY(launch_fill_u64(b.status, 0, (size_t)sg.ntile * 256, st));
        QO_TRY(launch_fill_u32(b.ticket, 0, 1, st));
        const int sh = lo_bit + 8 * p;
        if (vals)
            k_radix_sweep<true><<<sg.ntile, 256, 0, st>>>(ksrc, vsrc, kdst, vdst, sg, nullptr, 1, sh,
                                                          b.starts + p * 256, 0, b.status, b.ticket);
        else
            k_radix_sweep<false><<<sg.ntile, 256, 0, st>>>(ksrc, nullptr, kdst, nullptr, sg, nullptr, 1, sh,
                                                           b.starts + p * 256, 0, b.status, b.ticket);
        QO_LAUNCHED();
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
        *in_alt = !*in_alt;
    }
    return QO_OK;
}

int launch_radix_segments(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt, const RadixSeg &seg0,
                          const RadixSeg *d_segs, int nseg, uint32_t total_tiles, int lo_bit, int hi_bit,
                          uint32_t *hist, void *scan_ws, hipStream_t st, bool *in_alt) {
    *in_alt = false;
    if (total_tiles == 0) return QO_OK;
    QO_REQUIRE(lo_bit >= 0 && hi_bit <= 64 && lo_bit < hi_bit, "bad bit range");
    QO_REQUIRE((size_t)total_tiles * 256 < 0xFFFFFFFFull, "sort input too large");
    uint64_t *ksrc = keys, *kdst = kalt;
    uint32_t *vsrc = vals, *vdst = valt;
    for (int sh = lo_bit; sh < hi_bit; sh += 8) {
        k_radix_hist<<<total_tiles, 256, 0, st>>>(ksrc, seg0, d_segs, nseg, sh, hist);
        QO_LAUNCHED();
        QO_TRY(launch_scan_excl_u32(hist, hist, (size_t)total_tiles * 256, scan_ws, st));
        if (vals)
            k_radix_scatter<true><<<total_tiles, 256, 0, st>>>(ksrc, vsrc, kdst, vdst, seg0, d_segs, nseg, sh, hist);
        else
            k_radix_scatter<false><<<total_tiles, 256, 0, st>>>(ksrc, nullptr, kdst, nullptr, seg0, d_segs, nseg, sh,
                                                               hist);
        QO_LAUNCHED();
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
        *in_alt = !*in_alt;
    }
    return QO_OK;
}

static uint32_t radix_tiles(size_t n) { return (uint32_t)((n + kSortTile - 1) / kSortTile); }

size_t radix_sweep_seg_bytes(uint32_t total_tiles, int nseg, uint32_t nchunks) {
    Carver cv(nullptr);
    cv.take<uint32_t>((size_t)nchunks * kSweepMaxPass * 256);
    cv.take<uint32_t>((size_t)nseg * kSweepMaxPass * 256);
    cv.take<uint32_t>(64);
    cv.take<uint64_t>((size_t)total_tiles * 256 + 1);
    return cv.off + 256;
}

int launch_radix_sweep_segments(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt,
                                const RadixSeg *d_segs, int nseg, uint32_t total_tiles, const SweepChunk *d_chunks,
                                uint32_t nchunks, const uint32_t *d_chunk0, int lo_bit, int hi_bit, void *ws,
                                hipStream_t st, bool *in_alt) {
    *in_alt = false;
    if (total_tiles == 0 || nchunks == 0) return QO_OK;
    const int npass = (hi_bit - lo_bit + 7) / 8;
    QO_REQUIRE(npass >= 1 && npass <= kSweepMaxPass && lo_bit >= 0 && hi_bit <= 64, "bad bit range");
    QO_REQUIRE(nseg >= 1 && (size_t)nseg * npass < 0x7FFFFFFFull, "bad segment count");
    Carver cv(ws);
    uint32_t *partial = cv.take<uint32_t>((size_t)nchunks * npass * 256);
    uint32_t *starts = cv.take<uint32_t>((size_t)nseg * npass * 256);
    uint32_t *ticket = cv.take<uint32_t>(64);
    uint64_t *status = cv.take<uint64_t>((size_t)total_tiles * 256 + 1);
    k_sweep_counts_seg<<<nchunks, 256, 0, st>>>(keys, d_segs, d_chunks, lo_bit, npass, partial);
    QO_LAUNCHED();
    k_sweep_starts_seg<<<(unsigned)(nseg * npass), 256, 0, st>>>(partial, d_chunk0, npass, starts);
    QO_LAUNCHED();
    const RadixSeg none{0, 0, nullptr, 0, 0};
    uint64_t *ksrc = keys, *kdst = kalt;
    uint32_t *vsrc = vals, *vdst = valt;
    for (int p = 0; p < npass; ++p) {
        QO_TRY(launch_fill_u64(status, 0, (size_t)total_tiles * 256, st));
        QO_TRY(launch_fill_u32(ticket, 0, 1, st));
        const int sh = lo_bit + 8 * p;
        if (vals)
            k_radix_sweep<true><<<total_tiles, 256, 0, st>>>(ksrc, vsrc, kdst, vdst, none, d_segs, nseg, sh,
                                                             starts + p * 256, (uint32_t)npass * 256, status, ticket);
        else
            k_radix_sweep<false><<<total_tiles, 256, 0, st>>>(ksrc, nullptr, kdst, nullptr, none, d_segs, nseg, sh,
                                                              starts + p * 256, (uint32_t)npass * 256, status, ticket);
        QO_LAUNCHED();
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
        *in_alt = !*in_alt;
    }
    return QO_OK;
}

// Key + value sorts take the one-sweep passes; key-only sorts, whose tiles are
// light enough that the per-digit look-back costs more than the histogram read
// it replaces (measured: a 2^26-Record sort 2.41 vs 1.98 ms), keep the
// histogram / scan / scatter passes.
static void radix_carve(size_t n, bool with_vals, Carver &cv, uint64_t **kalt, uint32_t **valt, uint32_t **hist,
                        void **scan_ws, SweepBufs *sb) {
    const uint32_t tiles = radix_tiles(n) + 1;
    *kalt = cv.take<uint64_t>(n ? n : 1);
    *valt = with_vals ? cv.take<uint32_t>(n ? n : 1) : nullptr;
    if (with_vals) {
        *sb = sweep_carve(tiles, kSweepMaxPass, cv);
    } else {
        *hist = cv.take<uint32_t>((size_t)tiles * 256);
        *scan_ws = cv.take<char>(scan_workspace_bytes((size_t)tiles * 256));
    }
}

size_t radix_workspace_bytes(size_t n, bool with_vals) {
    Carver cv(nullptr);
    uint64_t *kalt;
    uint32_t *valt, *hist;
    void *scan_ws;
    SweepBufs sb;
    radix_carve(n, with_vals, cv, &kalt, &valt, &hist, &scan_ws, &sb);
    return cv.off + 256;
}

static int radix_impl(uint64_t *keys, uint32_t *vals, size_t n, const uint64_t *d_n, int lo_bit, int hi_bit,
                      void *ws, hipStream_t st) {
    if (n == 0) return QO_OK;
    QO_REQUIRE(n < 0xFFFFFFFFull, "sort input too large");
    QO_REQUIRE(lo_bit >= 0 && hi_bit <= 64 && lo_bit < hi_bit, "bad bit range");
    const uint32_t tiles = radix_tiles(n);
    Carver cv(ws);
    uint64_t *kalt = nullptr;
    uint32_t *valt = nullptr, *hist = nullptr;
    void *scan_ws = nullptr;
    SweepBufs sb{};
    radix_carve(n, vals != nullptr, cv, &kalt, &valt, &hist, &scan_ws, &sb);
    RadixSeg seg{0, n, d_n, 0, tiles};
    bool in_alt = false;
    if (vals)
        QO_TRY(launch_radix_sweep(keys, vals, kalt, valt, seg, lo_bit, hi_bit, sb, st, &in_alt));
    else
        QO_TRY(launch_radix_segments(keys, nullptr, kalt, nullptr, seg, nullptr, 1, tiles, lo_bit, hi_bit, hist,
                                     scan_ws, st, &in_alt));
    if (in_alt) {
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
    k_pack_join<<<g, 256, 0, st>>>(key, spay, rpay, m, k64, v32);
    QO_LAUNCHED();
    QO_TRY(launch_radix_sort_u64(k64, v32, m, 0, 64, rws, st));
    k_unpack_join<<<g, 256, 0, st>>>(k64, v32, m, key, spay, rpay);
    QO_LAUNCHED();
    return QO_OK;
}

size_t sort_join_rows_ws_bytes(size_t m) { return m * 12 + 1024 + radix_workspace_bytes(m, true); }

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
