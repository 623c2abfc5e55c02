// Selection operators: fused filter+aggregate, stable filter compaction and the
// exclusive scan.
//
// Reference path replaced (kernel/filter/filter_op.cl): filterImpl_map_kernel
// (:6-31) marks rows, ScanLargeArrays/prefixSum/blockAddition (:51-207) scan the
// marks in 2*pass+1 launches, filterImpl_outSize_kernel (:211-215) is read back
// by the host before filterImpl_write_kernel (:218-238) scatters.  Here one
// kernel does all of it: predicate, wave ballot ranks, a decoupled look-back
// across tiles, and a stable scatter; the count stays on the device.
#include <map>
#include <mutex>

#include "qo_device.hpp"
#include "qo_internal.hpp"

namespace qo {

// ---------------------------------------------------------------------------
// Per-stream self-resetting counters for single-launch reductions.
// ---------------------------------------------------------------------------
static uint32_t *sync_counters(hipStream_t st, int *status);

// ---------------------------------------------------------------------------
// Fused filter + aggregate.  Each thread streams VW-tuple vectors (UNROLL
// 16-byte loads of each column issued before any is used), keeps COUNT and an
// exact int64 SUM, the block reduces through LDS, and the last block to finish
// (ticket on a self-resetting counter) folds the per-block partials and writes
// the result.  Few waves with several loads in flight stream fastest here: on
// 2^28 tuples, one 256-thread block per CU with 4 loads per column in flight
// reads 7.1 TB/s, eight such blocks per CU 4.8 TB/s (the default launch is the
// first; 8 loads in flight need 110 VGPRs and read 3.8 TB/s).
// ---------------------------------------------------------------------------
template <bool AOS>
struct FaTraits;

template <>
struct FaTraits<false> {  // SoA key / value columns, int4 = 4 tuples
    static constexpr int VW = 4;
    struct V {
        int4 k, v;
    };
    static QO_DEV V load(const int32_t *k, const int32_t *v, size_t i) {
        return V{ld_stream(reinterpret_cast<const int4 *>(k) + i), ld_stream(reinterpret_cast<const int4 *>(v) + i)};
    }
    static QO_DEV void add(const V &q, int32_t lo, int32_t hi, uint64_t &c, int64_t &s) {
        const int4 &kk = q.k, &vv = q.v;
        bool p0 = (kk.x >= lo) & (kk.x <= hi), p1 = (kk.y >= lo) & (kk.y <= hi);
        bool p2 = (kk.z >= lo) & (kk.z <= hi), p3 = (kk.w >= lo) & (kk.w <= hi);
        c += (unsigned)p0 + (unsigned)p1 + (unsigned)p2 + (unsigned)p3;
        s += (int64_t)(p0 ? vv.x : 0) + (int64_t)(p1 ? vv.y : 0) + (int64_t)(p2 ? vv.z : 0) +
             (int64_t)(p3 ? vv.w : 0);
    }
    static QO_DEV void one(const int32_t *k, const int32_t *v, size_t e, int32_t lo, int32_t hi, uint64_t &c,
                           int64_t &s) {
        int32_t kk = k[e];
        bool p = (kk >= lo) & (kk <= hi);
        c += p;
        s += p ? (int64_t)v[e] : 0;
    }
};

template <>
struct FaTraits<true> {  // AoS Records {x = value, y = key}, int4 = 2 records
    static constexpr int VW = 2;
    struct V {
        int4 q;
    };
    static QO_DEV V load(const int32_t *r, const int32_t *, size_t i) {
        return V{ld_stream(reinterpret_cast<const int4 *>(r) + i)};
    }
    static QO_DEV void add(const V &v, int32_t lo, int32_t hi, uint64_t &c, int64_t &s) {
        const int4 &q = v.q;
        bool p0 = (q.y >= lo) & (q.y <= hi), p1 = (q.w >= lo) & (q.w <= hi);
        c += (unsigned)p0 + (unsigned)p1;
        s += (int64_t)(p0 ? q.x : 0) + (int64_t)(p1 ? q.z : 0);
    }
    static QO_DEV void one(const int32_t *r, const int32_t *, size_t e, int32_t lo, int32_t hi, uint64_t &c,
                           int64_t &s) {
        int32_t kk = r[2 * e + 1];
        bool p = (kk >= lo) & (kk <= hi);
        c += p;
        s += p ? (int64_t)r[2 * e] : 0;
    }
};

template <bool AOS, int UNROLL, bool VEC>
__global__ __launch_bounds__(256) void k_filter_agg(const int32_t *__restrict__ a, const int32_t *__restrict__ b,
                                                    size_t n, int32_t lo, int32_t hi, uint64_t *partials,
                                                    uint32_t *done, int64_t *out) {
    using Tr = FaTraits<AOS>;
    uint64_t cnt = 0;
    int64_t sum = 0;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t scalar_from = 0;
    if (VEC) {
        const size_t nv = n / Tr::VW;
        size_t i = tid;
        for (; i + (UNROLL - 1) * stride < nv; i += UNROLL * stride) {
            typename Tr::V q[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) q[u] = Tr::load(a, b, i + u * stride);
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) Tr::add(q[u], lo, hi, cnt, sum);
        }
        for (; i < nv; i += stride) Tr::add(Tr::load(a, b, i), lo, hi, cnt, sum);
        scalar_from = nv * Tr::VW;
    }
    for (size_t e = scalar_from + tid; e < n; e += stride) Tr::one(a, b, e, lo, hi, cnt, sum);

    __shared__ uint64_t s_c[4], s_s[4];
    __shared__ bool s_last;
    cnt = wave_sum(cnt);
    uint64_t usum = wave_sum((uint64_t)sum);
    if (lane_id() == 0) {
        s_c[wave_id()] = cnt;
        s_s[wave_id()] = usum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t c = 0, s = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) {
            c += s_c[w];
            s += s_s[w];
        }
        partials[2 * blockIdx.x] = c;
        partials[2 * blockIdx.x + 1] = s;
        __threadfence();
        s_last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    uint64_t c = 0, s = 0;
    for (unsigned bidx = threadIdx.x; bidx < gridDim.x; bidx += blockDim.x) {
        c += __hip_atomic_load(&partials[2 * bidx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s += __hip_atomic_load(&partials[2 * bidx + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    c = wave_sum(c);
    s = wave_sum(s);
    __syncthreads();
    if (lane_id() == 0) {
        s_c[wave_id()] = c;
        s_s[wave_id()] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t tc = 0, ts = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) {
            tc += s_c[w];
            ts += s_s[w];
        }
        out[0] = (int64_t)tc;
        out[1] = (int64_t)ts;
        out[2] = (int64_t)(int32_t)(uint32_t)ts;
        *done = 0;  // ready for the next launch on this stream
    }
}

static int launch_filter_agg(bool aos, const int32_t *a, const int32_t *b, size_t n, int32_t lo, int32_t hi,
                             int64_t *d_out, hipStream_t st) {
    TuneParams tp = tune_get(TUNE_FILTER_AGG, n);
    const int vw = aos ? 2 : 4;
    bool vec = ((uintptr_t)a % 16 == 0) && (aos || (uintptr_t)b % 16 == 0);
    unsigned cap = (unsigned)props().cu_count * (unsigned)(tp.param > 0 ? tp.param : 8);
    unsigned grid = grid_for(n / vw + 1, 256u * (unsigned)tp.items, cap);
    int status = QO_OK;
    uint32_t *cnt = sync_counters(st, &status);
    if (!cnt) return status;
    void *ws = workspace(st, (size_t)grid * 16, &status);
    if (!ws) return status;
    uint64_t *partials = static_cast<uint64_t *>(ws);
    uint32_t *done = cnt + 0;
#define QO_FA(AOSV, U, V) k_filter_agg<AOSV, U, V><<<grid, 256, 0, st>>>(a, b, n, lo, hi, partials, done, d_out)
    if (aos) {
        if (!vec) QO_FA(true, 1, false);
        else if (tp.items >= 8) QO_FA(true, 8, true);
        else if (tp.items >= 4) QO_FA(true, 4, true);
        else QO_FA(true, 2, true);
    } else {
        if (!vec) QO_FA(false, 1, false);
        else if (tp.items >= 8) QO_FA(false, 8, true);
        else if (tp.items >= 4) QO_FA(false, 4, true);
        else QO_FA(false, 2, true);
    }
#undef QO_FA
    QO_LAUNCHED();
    return QO_OK;
}

// ---------------------------------------------------------------------------
// Stable filter compaction of Records (single pass).  A tile is 256 threads x
// J loads of VEC records, striped so every load instruction is coalesced.
// ---------------------------------------------------------------------------
template <int VEC>
struct RecVec;
template <>
struct RecVec<2> {
    using T = uint4;
};
template <>
struct RecVec<1> {
    using T = uint2;
};

// Tiles of J * NT * VEC records (16384 with the launch below): one ticket per
// tile keeps the same-address ticket atomics few.  The selected records are
// staged in LDS in input order before the look-back and written out row by row
// across the block.  In-tile offsets are 32-bit from the tile's base pointer:
// with 64-bit per-element indexing the kernel needed 140 VGPRs and one block
// per CU, which cost 1.44 ms instead of 0.86 ms on 2^28 records.
constexpr int kFiltJ = 8;
constexpr int kFiltThreads = 1024;

template <int J, int VEC, int NT>
__global__ __launch_bounds__(NT) void k_filter_records(const qo_record *__restrict__ rin, size_t n, int32_t lo,
                                                       int32_t hi, qo_record *__restrict__ rout, uint64_t *status,
                                                       uint32_t *ticket, uint64_t *d_count, uint32_t num_tiles,
                                                       DevCtl *ctl) {
    constexpr int NW = NT / kWave;
    constexpr uint32_t kTile = J * NT * VEC;
    using V = typename RecVec<VEC>::T;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tot[tile_offsets_words(J, NW)];
    __shared__ uint64_t s_excl;
    __shared__ qo_record s_out[kTile];
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= num_tiles) return;  // whole block; one ticket per launched block, so never taken
    const size_t ebase = (size_t)tile * kTile;
    const qo_record *src = rin + ebase;
    const uint32_t rem = (uint32_t)min((size_t)kTile, n - ebase);
    qo_record r[J][VEC];
    uint32_t c[J], ex[J], fl[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const uint32_t v = (uint32_t)j * NT + threadIdx.x;
        const uint32_t e0 = v * VEC;
        if (e0 + VEC <= rem) {
            const V q = ld_stream(reinterpret_cast<const V *>(src) + v);
            if constexpr (VEC == 2) {
                r[j][0] = {q.x, q.y};
                r[j][1] = {q.z, q.w};
            } else {
                r[j][0] = {q.x, q.y};
            }
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) r[j][k] = (e0 + k < rem) ? src[e0 + k] : qo_record{0, 0};
        }
        fl[j] = 0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const int32_t key = (int32_t)r[j][k].y;
            fl[j] |= ((key >= lo) & (key <= hi) & (e0 + k < rem)) ? (1u << k) : 0u;
        }
        c[j] = __popc(fl[j]);
    }
    const uint32_t total = tile_offsets<J, NW>(c, ex, s_tot);
    // Stage the selected records at their tile-local rank (input order).
#pragma unroll
    for (int j = 0; j < J; ++j) {
        uint32_t pos = ex[j];
#pragma unroll
        for (int k = 0; k < VEC; ++k)
            if (fl[j] & (1u << k)) s_out[pos++] = r[j][k];
    }
    if (threadIdx.x < kWave) {
        const uint64_t e = lookback(status, tile, total, ctl);
        if (threadIdx.x == 0) s_excl = e;
    }
    __syncthreads();
    const uint64_t base = s_excl;
    qo_record *dst = rout + base;
    for (uint32_t i = threadIdx.x; i < total; i += NT) dst[i] = s_out[i];
    if (tile == num_tiles - 1 && threadIdx.x == 0) *d_count = base + total;
}

// Tiles of a compaction of n records starting at `rin` (its alignment picks the
// tile width).
static size_t filter_tiles(const qo_record *rin, size_t n) {
    const size_t per_tile = (size_t)kFiltJ * kFiltThreads * ((((uintptr_t)rin % 16) == 0) ? 2 : 1);
    return (n + per_tile - 1) / per_tile;
}

size_t filter_records_ws_bytes(const qo_record *rin, size_t n) { return filter_tiles(rin, n) * 8 + 256; }

int launch_filter_records_ws(const qo_record *rin, size_t n, int32_t lo, int32_t hi, qo_record *rout,
                             uint64_t *d_count, void *ws, hipStream_t st) {
    constexpr int J = kFiltJ, NT = kFiltThreads;
    const bool vec = ((uintptr_t)rin % 16) == 0;
    const size_t tiles = filter_tiles(rin, n);
    QO_REQUIRE(tiles < 0xFFFFFFFFull, "input too large");
    Carver cv(ws);
    uint64_t *stat = cv.take<uint64_t>(tiles);
    uint32_t *ticket = cv.take<uint32_t>(1);
    QO_TRY(launch_fill_u32(ws, 0, cv.off / 4, st));
    if (vec)
        k_filter_records<J, 2, NT><<<(unsigned)tiles, NT, 0, st>>>(rin, n, lo, hi, rout, stat, ticket, d_count,
                                                                   (uint32_t)tiles, dev_ctl());
    else
        k_filter_records<J, 1, NT><<<(unsigned)tiles, NT, 0, st>>>(rin, n, lo, hi, rout, stat, ticket, d_count,
                                                                   (uint32_t)tiles, dev_ctl());
    QO_LAUNCHED();
    return QO_OK;
}

static int launch_filter_records(const qo_record *rin, size_t n, int32_t lo, int32_t hi, qo_record *rout,
                                 uint64_t *d_count, hipStream_t st) {
    int status = QO_OK;
    void *ws = workspace(st, filter_records_ws_bytes(rin, n), &status);
    if (!ws) return status;
    return launch_filter_records_ws(rin, n, lo, hi, rout, d_count, ws, st);
}

// ---------------------------------------------------------------------------
// Exclusive scan of 32-bit integers (modulo 2^32), single pass.
// ---------------------------------------------------------------------------
constexpr int kScanJ = 8;
constexpr int kScanThreads = 1024;

template <int J, bool VEC, int NT>
__global__ __launch_bounds__(NT) void k_scan_excl(const uint32_t *in, uint32_t *out, size_t n, uint64_t *status,
                                                  uint32_t *ticket, uint32_t num_tiles, DevCtl *ctl) {
    constexpr int NW = NT / kWave;
    constexpr int VW = VEC ? 4 : 1;
    constexpr uint32_t kTile = J * NT * VW;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tot[tile_offsets_words(J, NW)];
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= num_tiles) return;  // whole block; one ticket per launched block, so never taken
    const size_t ebase = (size_t)tile * kTile;
    const uint32_t *src = in + ebase;
    uint32_t *dst = out + ebase;
    const uint32_t rem = (uint32_t)min((size_t)kTile, n - ebase);
    uint32_t x[J][VW];
    uint32_t c[J], ex[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const uint32_t v = (uint32_t)j * NT + threadIdx.x;
        const uint32_t e0 = v * VW;
        if (VEC && e0 + VW <= rem) {
            const uint4 q = reinterpret_cast<const uint4 *>(src)[v];
            x[j][0] = q.x;
            if constexpr (VW == 4) {
                x[j][1] = q.y;
                x[j][2] = q.z;
                x[j][3] = q.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < VW; ++k) x[j][k] = (e0 + k < rem) ? src[e0 + k] : 0u;
        }
        uint32_t sacc = 0;
#pragma unroll
        for (int k = 0; k < VW; ++k) sacc += x[j][k];
        c[j] = sacc;
    }
    const uint32_t total = tile_offsets<J, NW>(c, ex, s_tot);
    if (threadIdx.x < kWave) {
        const uint64_t e = lookback(status, tile, total, ctl);
        if (threadIdx.x == 0) s_excl = e;
    }
    __syncthreads();
    const uint32_t base = (uint32_t)s_excl;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const uint32_t v = (uint32_t)j * NT + threadIdx.x;
        const uint32_t e0 = v * VW;
        uint32_t run = base + ex[j];
        uint32_t o[VW];
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            o[k] = run;
            run += x[j][k];
        }
        if (VEC && e0 + VW <= rem) {
            if constexpr (VW == 4) reinterpret_cast<uint4 *>(dst)[v] = uint4{o[0], o[1], o[2], o[3]};
            else dst[e0] = o[0];
        } else {
#pragma unroll
            for (int k = 0; k < VW; ++k)
                if (e0 + k < rem) dst[e0 + k] = o[k];
        }
    }
}

size_t scan_workspace_bytes(size_t n) {
    // Sized for the unaligned path (one element per load), which has the most tiles.
    const size_t per_tile = (size_t)kScanJ * kScanThreads;
    const size_t tiles = (n + per_tile - 1) / per_tile + 1;
    return tiles * 8 + 512;
}

int launch_scan_excl_u32(const uint32_t *in, uint32_t *out, size_t n, void *ws, hipStream_t st) {
    if (n == 0) return QO_OK;
    constexpr int J = kScanJ, NT = kScanThreads;
    const bool vec = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
    const size_t per_tile = (size_t)J * NT * (vec ? 4 : 1);
    const size_t tiles = (n + per_tile - 1) / per_tile;
    QO_REQUIRE(tiles < 0xFFFFFFFFull, "scan input too large");
    Carver cv(ws);
    uint64_t *stat = cv.take<uint64_t>(tiles);
    uint32_t *ticket = cv.take<uint32_t>(1);
    QO_TRY(launch_fill_u32(ws, 0, cv.off / 4, st));
    if (vec) k_scan_excl<J, true, NT><<<(unsigned)tiles, NT, 0, st>>>(in, out, n, stat, ticket, (uint32_t)tiles, dev_ctl());
    else k_scan_excl<J, false, NT><<<(unsigned)tiles, NT, 0, st>>>(in, out, n, stat, ticket, (uint32_t)tiles, dev_ctl());
    QO_LAUNCHED();
    return QO_OK;
}

// ---------------------------------------------------------------------------
namespace {
std::mutex g_cnt_mu;
std::map<StreamKey, uint32_t *> g_cnt;
}  // namespace

static uint32_t *sync_counters(hipStream_t st, int *status) {
    *status = QO_OK;
    std::lock_guard<std::mutex> lk(g_cnt_mu);
    const StreamKey key = stream_key_of(st);
    auto it = g_cnt.find(key);
    if (it != g_cnt.end()) return it->second;
    uint32_t *p = nullptr;
    hipError_t e = hipMalloc(&p, 4096);
    if (e == hipSuccess) e = hipMemset(p, 0, 4096);
    if (e != hipSuccess) {
        *status = set_error(QO_ERR_NO_MEMORY, "counter block: %s", hipGetErrorString(e));
        return nullptr;
    }
    g_cnt[key] = p;
    return p;
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_filter_agg(const int32_t *key, const int32_t *val, size_t n, int32_t lo, int32_t hi, int64_t *d_out,
                  qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(d_out && (n == 0 || (key && val)), "null pointer");
    return launch_filter_agg(false, key, val, n, lo, hi, d_out, stream_of(s));
}

int qo_filter_agg_records(const qo_record *rin, size_t n, int32_t lo, int32_t hi, int64_t *d_out, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(d_out && (n == 0 || rin), "null pointer");
    return launch_filter_agg(true, reinterpret_cast<const int32_t *>(rin), nullptr, n, lo, hi, d_out, stream_of(s));
}

int qo_filter_records(const qo_record *rin, size_t n, int32_t lo, int32_t hi, qo_record *rout, uint64_t *d_count,
                      qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(d_count, "null count pointer");
    hipStream_t st = stream_of(s);
    if (n == 0) return launch_fill_u64(d_count, 0, 1, st);
    QO_REQUIRE(rin && rout, "null pointer");
    return launch_filter_records(rin, n, lo, hi, rout, d_count, st);
}

int qo_scan_exclusive_i32(const int32_t *in, int32_t *out, size_t n, qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(in && out, "null pointer");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, scan_workspace_bytes(n), &status);
    if (!ws) return status;
    return launch_scan_excl_u32(reinterpret_cast<const uint32_t *>(in), reinterpret_cast<uint32_t *>(out), n, ws,
                                st);
}

}  // extern "C"
