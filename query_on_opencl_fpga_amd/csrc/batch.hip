// Batched hj_filter_sort: many independent queries of uneven size executed by
// one hipGraph launch.
//
// Every stage is one kernel over the tiles of ALL queries.  Build and probe are
// persistent kernels pulling tiles from a work queue (an atomic ticket); a tile
// maps to its query through the per-query first-tile table, so a 4K-row query
// and a 4M-row query share the same launch.  The probe writes sort-ready
// (key, s_payload) 64-bit keys with the r_payload beside them; a segmented
// LSD radix sort (one segment per query, element counts read from device
// memory, per-query digit starts up front, then per pass a tile histogram, a
// segmented column scan and the placement) orders every query's matches; its
// last pass writes each query's output columns.
// The stage sequence (fills included, no memset nodes) is captured once into a
// graph at creation and replayed by qo_batch_launch.
//
// Reference path: buildHashTable / probeHashTable + the sort kernels of the
// hj_filter_sort image (kernel/hj_filter_sort/1/hj_filter_sort_op.cl:25-109,
// 362-564), which the reference never composes on the host; here one query of
// the batch is exactly qo_hj_filter_sort.
#include <algorithm>
#include <cstring>
#include <vector>

#include "qo_internal.hpp"
#include "qo_join.hpp"
#include "qo_sort.hpp"
#include "qo_batch.hpp"

namespace qo {

// Last query whose first R tile (WHICH = 0) or S tile (WHICH = 1) is <= t.
template <int WHICH>
QO_DEV int query_of(const BQuery *q, int nq, uint32_t t) {
    int lo = 0, hi = nq - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        uint32_t t0 = WHICH == 0 ? q[mid].rtile0 : q[mid].stile0;
        if (t0 <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_b_init(const BQuery *q, int nq) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    JoinHeader *h = q[i].hdr;
    h->cap = q[i].table_cap;
    h->has_dups = 0;
    h->sentinel_count = 0;
    h->overflow = 0;
    h->mode = kJoinHashMix;
    h->nr = q[i].nr;
    h->slots_alloc = q[i].table_cap;
    h->kmin = 0;
    h->shift = 0;
    h->rmin = INT32_MAX;
    h->rmax = INT32_MIN;
    h->occupied = 0;
    *q[i].count = 0;
}

// Persistent loop: thread 0 draws tickets; the whole block works on each tile.
#define QO_TILE_LOOP(TICKET, TOTAL, T)                                                                              \
    __shared__ uint32_t s_ticket;                                                                                   \
    for (;;) {                                                                                                      \
        if (threadIdx.x == 0) s_ticket = atomicAdd((TICKET), 1u);                                                  \
        __syncthreads();                                                                                            \
        const uint32_t T = s_ticket;                                                                                \
        __syncthreads();                                                                                            \
        if (T >= (TOTAL)) break;

#define QO_TILE_LOOP_END }

__global__ __launch_bounds__(256) void k_b_build(const BQuery *q, int nq, uint32_t total, uint32_t *ticket,
                                                 uint32_t build_tile) {
    QO_TILE_LOOP(ticket, total, t)
    const int qi = query_of<0>(q, nq, t);
    const BQuery &Q = q[qi];
    const uint64_t r0 = (uint64_t)(t - Q.rtile0) * build_tile;
    const uint64_t r1 = r0 + build_tile < Q.nr ? r0 + build_tile : Q.nr;
    uint64_t *slots = reinterpret_cast<uint64_t *>(Q.hdr + 1);
    for (uint64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) join_insert(slots, Q.table_cap, Q.hdr, Q.rkey[i], Q.rpay[i]);
    QO_TILE_LOOP_END
}

// One ticket covers `group` consecutive S tiles (of any queries): fewer
// same-address ticket atomics for a batch of ~10^6 tiles.  The launch bound
// asks for 3 waves per SIMD (the 48 KB output stage allows no more): at 184
// VGPRs the kernel ran 2 and took 48.3 ms per 512-query batch, with the bound
// (168 VGPRs, a few spilled) 39.7 ms.
__global__ __launch_bounds__(256, 3) void k_b_probe(const BQuery *q, int nq, uint32_t total, uint32_t *ticket,
                                                 uint64_t *mkeys, uint32_t *mvals, uint32_t group) {
    __shared__ uint32_t s_tot[tile_offsets_words(kProbeJ, 4)];
    __shared__ uint64_t s_base;
    __shared__ JoinStage stage;
    const uint32_t groups = (total + group - 1) / group;
    QO_TILE_LOOP(ticket, groups, tg)
    for (uint32_t t = tg * group; t < total && t < (tg + 1) * group; ++t) {
    const int qi = query_of<1>(q, nq, t);
    const BQuery &Q = q[qi];
    const uint64_t *slots = reinterpret_cast<const uint64_t *>(Q.hdr + 1);
    // Sort-ready rows (join_sort_key, r_payload); every query's rows start at a
    // multiple of 32, so the staged tiles are written 4 rows per thread.
    const SoaEmit emit{nullptr, nullptr, nullptr, Q.out_cap, mkeys + Q.m_base, mvals + Q.m_base};
    if (Q.sfilt)
        probe_tile<true>(Q.hdr, slots, Q.skey, Q.spay, Q.sfilt, Q.ns, Q.flo, Q.fhi, Q.vec != 0, t - Q.stile0, Q.count,
                         emit, s_tot, &s_base, &stage);
    else
        probe_tile<false>(Q.hdr, slots, Q.skey, Q.spay, Q.sfilt, Q.ns, Q.flo, Q.fhi, Q.vec != 0, t - Q.stile0, Q.count,
                          emit, s_tot, &s_base, &stage);
    __syncthreads();  // s_tot, s_base and the stage are reused by the next tile
    }
    QO_TILE_LOOP_END
}

// ---------------------------------------------------------------------------
// Plan of one join batch: sizes, then (after binding an arena) device pointers.
// ---------------------------------------------------------------------------
int join_plan_init(const qo_join_query *queries, int nq, JoinBatchPlan &p) {
    QO_REQUIRE(nq >= 1 && queries, "bad batch arguments");
    p.nq = nq;
    p.hq.assign(nq, BQuery{});
    p.hs.assign(nq, RadixSeg{});
    p.hchunks.clear();
    p.hchunk0.assign(1, 0u);
    p.houts.assign(nq, RowOut{});
    p.total_slots = p.total_match = p.fill_slots = 0;
    p.part_q.clear();
    p.pbuild_q.clear();
    p.part_ws_bytes = 0;
    uint64_t rt = 0, stt = 0, mt = 0;
    // Launch parameters by the size class of the batch's mean build side.
    uint64_t rsum = 0;
    for (int i = 0; i < nq; ++i) rsum += queries[i].nr;
    const size_t rmean = std::max<size_t>(1, (size_t)(rsum / (uint64_t)nq));
    const int load = std::max(10, std::min(90, tune_get(TUNE_BATCH_LOAD, rmean).param));
    const TuneParams tb = tune_get(TUNE_BATCH, rmean);
    p.build_tile = (uint32_t)std::max(256, std::min(1 << 16, tb.param));
    p.probe_group = (uint32_t)std::max(1, std::min(64, tb.items));
    for (int i = 0; i < nq; ++i) {
        const qo_join_query &u = queries[i];
        QO_REQUIRE((u.nr == 0 || (u.rkey && u.rpay)) && (u.ns == 0 || (u.skey && u.spay)),
                   "query %d: null input column", i);
        QO_REQUIRE(u.out_capacity == 0 || (u.okey && u.orpay && u.ospay), "query %d: null output column", i);
        BQuery &q = p.hq[i];
        memset(&q, 0, sizeof(q));
        q.rkey = u.rkey; q.rpay = u.rpay; q.skey = u.skey; q.spay = u.spay; q.sfilt = u.sfilt;
        q.nr = u.nr; q.ns = u.ns; q.flo = u.flo; q.fhi = u.fhi;
        q.okey = u.okey; q.orpay = u.orpay; q.ospay = u.ospay; q.out_cap = u.out_capacity;
        p.houts[i] = RowOut{u.okey, u.orpay, u.ospay};
        // Rounded to 32 slots so every query's 256-byte header stays 256-byte aligned.
        q.table_cap = (std::max<uint64_t>(64, (uint64_t)((double)u.nr * 100.0 / load) + 1) + 31) / 32 * 32;
        QO_REQUIRE(q.table_cap <= 0xFFFFFFFFull, "query %d: build side too large", i);
        // Tables too large for direct lookups take the partitioned probe (the
        // batch's tables are hash-mixed: header of 32 slots, then the slots).
        const size_t tbytes = (size_t)(32 + q.table_cap) * 8;
        int part = 0, bits = 0, splits = 0;
        join_part_strategy(tbytes, (size_t)u.ns, (int)kJoinHashMix, &part, &bits, &splits);
        const bool partitioned = part != 0 && u.ns > 0;
        if (partitioned) {
            p.part_q.push_back(i);
            p.part_ws_bytes = std::max(p.part_ws_bytes, join_part_ws_bytes(tbytes, (size_t)u.ns, false));
        }
        const size_t bws = u.nr ? join_build_ws_bytes(tbytes, (size_t)u.nr) : 0;
        if (bws) {
            p.pbuild_q.push_back(i);
            p.part_ws_bytes = std::max(p.part_ws_bytes, bws);
        }
        q.vec = ((uintptr_t)u.skey % 16 == 0 && (uintptr_t)u.spay % 16 == 0 &&
                 (!u.sfilt || (uintptr_t)u.sfilt % 16 == 0)) ? 1u : 0u;
        q.rtile0 = (uint32_t)rt;
        q.nrtile = bws ? 0u : (uint32_t)((u.nr + p.build_tile - 1) / p.build_tile);
        q.stile0 = (uint32_t)stt;
        q.nstile = partitioned ? 0u : (uint32_t)((u.ns + kProbeTile - 1) / kProbeTile);
        q.mtile0 = (uint32_t)mt;
        q.nmtile = (uint32_t)((u.out_capacity + kSortTile - 1) / kSortTile);
        q.m_base = p.total_match;
        rt += q.nrtile;
        stt += q.nstile;
        mt += q.nmtile;
        for (uint32_t t0 = 0; t0 < q.nmtile; t0 += kSweepChunkTiles)
            p.hchunks.push_back(SweepChunk{(uint32_t)i, t0, std::min(q.nmtile, t0 + kSweepChunkTiles)});
        p.hchunk0.push_back((uint32_t)p.hchunks.size());
        p.total_slots += 32 + q.table_cap;  // 256-byte header = 32 slots
        if (!bws) p.fill_slots += 32 + q.table_cap;
        p.total_match += (u.out_capacity + 31) / 32 * 32;
    }
    QO_REQUIRE(rt < 0xFFFFFFFFull && stt < 0xFFFFFFFFull && mt * 256 < 0xFFFFFFFFull, "batch too large");
    p.rt = (uint32_t)rt;
    p.stt = (uint32_t)stt;
    p.mt = (uint32_t)mt;
    p.small_words = 64 + (size_t)nq * 2;
    return QO_OK;
}

static void plan_carve(JoinBatchPlan &p, Carver &cv) {
    p.slots_all = cv.take<uint64_t>(p.total_slots);
    p.small = cv.take<uint32_t>(p.small_words);
    p.d_q = cv.take<BQuery>(p.nq);
    p.d_segs = cv.take<RadixSeg>(p.nq);
    p.mkeys = cv.take<uint64_t>(p.total_match);
    p.mvals = cv.take<uint32_t>(p.total_match);
    p.kalt = cv.take<uint64_t>(p.total_match);
    p.valt = cv.take<uint32_t>(p.total_match);
    p.d_chunks = cv.take<SweepChunk>(p.hchunks.size() + 1);
    p.d_chunk0 = cv.take<uint32_t>(p.hchunk0.size());
    p.d_outs = cv.take<RowOut>(p.nq);
    p.sort_ws = cv.take<char>(radix_rows_seg_bytes(p.mt, p.nq, (uint32_t)p.hchunks.size()));
    p.part_ws = p.part_ws_bytes ? cv.take<char>(p.part_ws_bytes) : nullptr;
}

size_t join_plan_bytes(JoinBatchPlan &p) {
    Carver cv(nullptr);
    plan_carve(p, cv);
    return cv.off + 256;
}

void join_plan_bind(JoinBatchPlan &p, void *arena) {
    Carver cv(arena);
    plan_carve(p, cv);
    // Tables the persistent build fills come first, those of partitioned builds after.
    uint64_t slot_off = 0, pslot_off = p.fill_slots;
    size_t pb = 0;
    p.counts = reinterpret_cast<unsigned long long *>(p.small + 64);
    for (int i = 0; i < p.nq; ++i) {
        BQuery &q = p.hq[i];
        const bool pbuild = pb < p.pbuild_q.size() && p.pbuild_q[pb] == i;
        pb += pbuild;
        uint64_t &off = pbuild ? pslot_off : slot_off;
        q.hdr = reinterpret_cast<JoinHeader *>(p.slots_all + off);
        off += 32 + q.table_cap;
        q.count = p.counts + i;
        p.hs[i] = RadixSeg{q.m_base, q.out_cap, reinterpret_cast<const uint64_t *>(q.count), q.mtile0, q.nmtile};
    }
}

int join_plan_upload(const JoinBatchPlan &p, hipStream_t st) {
    QO_HIP(hipMemcpyAsync(p.d_q, p.hq.data(), sizeof(BQuery) * p.nq, hipMemcpyHostToDevice, st));
    QO_HIP(hipMemcpyAsync(p.d_segs, p.hs.data(), sizeof(RadixSeg) * p.nq, hipMemcpyHostToDevice, st));
    if (!p.hchunks.empty())
        QO_HIP(hipMemcpyAsync(p.d_chunks, p.hchunks.data(), sizeof(SweepChunk) * p.hchunks.size(),
                              hipMemcpyHostToDevice, st));
    QO_HIP(hipMemcpyAsync(p.d_chunk0, p.hchunk0.data(), sizeof(uint32_t) * p.hchunk0.size(), hipMemcpyHostToDevice,
                          st));
    QO_HIP(hipMemcpyAsync(p.d_outs, p.houts.data(), sizeof(RowOut) * p.nq, hipMemcpyHostToDevice, st));
    return QO_OK;
}

int join_plan_launch(const JoinBatchPlan &p, hipStream_t st) {
    const unsigned persist = (unsigned)props().cu_count * 8;
    if (p.fill_slots) QO_TRY(launch_fill_u64(p.slots_all, kEmptySlot, p.fill_slots, st));
    QO_TRY(launch_fill_u32(p.small, 0, p.small_words, st));
    k_b_init<<<(p.nq + 255) / 256, 256, 0, st>>>(p.d_q, p.nq);
    QO_LAUNCHED();
    uint32_t *tickets = p.small;
    if (p.rt) {
        k_b_build<<<std::min<unsigned>(p.rt, persist), 256, 0, st>>>(p.d_q, p.nq, p.rt, tickets + 0, p.build_tile);
        QO_LAUNCHED();
    }
    for (int i : p.pbuild_q) {
        const BQuery &q = p.hq[i];
        bool used = false;
        QO_TRY(launch_join_part_build(q.hdr, (size_t)(32 + q.table_cap) * 8, q.rkey, q.rpay, q.nr, p.part_ws, st,
                                      &used));
        QO_REQUIRE(used, "query %d: the partitioned build did not take its table", i);
    }
    if (p.stt) {
        k_b_probe<<<std::min<unsigned>(p.stt, persist), 256, 0, st>>>(p.d_q, p.nq, p.stt, tickets + 1, p.mkeys,
                                                                      p.mvals, p.probe_group);
        QO_LAUNCHED();
    }
    for (int i : p.part_q) {
        const BQuery &q = p.hq[i];
        JoinOut out;
        out.cap = q.out_cap;
        out.k64 = p.mkeys + q.m_base;
        out.v32 = p.mvals + q.m_base;
        bool hp = false, rp = false;
        QO_TRY(launch_join_part_probe(q.hdr, (size_t)(32 + q.table_cap) * 8, q.skey, q.spay, q.sfilt, q.ns, q.flo,
                                      q.fhi, out, reinterpret_cast<uint64_t *>(q.count), p.part_ws, st, &hp, &rp,
                                      true));
        QO_REQUIRE(hp, "query %d: the partitioned probe did not take its hash-mixed table", i);
    }
    if (p.mt) {
        // Radix passes over the key bits only, then one pass that orders every
        // run of equal key by s_payload and writes the query's output columns.
        // Should some key have more than 64 rows, the full-width sort (its
        // kernels gated on that flag, so they return at once otherwise)
        // rewrites the outputs.
        uint32_t *long_flag = p.small + 2;
        bool in_alt = false;
        QO_TRY(launch_radix_rows_segments(p.mkeys, p.mvals, p.kalt, p.valt, p.d_segs, p.nq, p.mt, p.d_chunks,
                                           (uint32_t)p.hchunks.size(), p.d_chunk0, 32, 64, p.sort_ws, st, &in_alt));
        uint64_t *ks = in_alt ? p.kalt : p.mkeys, *ka = in_alt ? p.mkeys : p.kalt;
        uint32_t *vs = in_alt ? p.valt : p.mvals, *va = in_alt ? p.mvals : p.valt;
        QO_TRY(launch_fix_runs_segments(ks, vs, p.d_segs, p.nq, p.mt, p.d_outs, long_flag, st));
        bool alt2 = false;
        QO_TRY(launch_radix_rows_segments(ks, vs, ka, va, p.d_segs, p.nq, p.mt, p.d_chunks,
                                           (uint32_t)p.hchunks.size(), p.d_chunk0, 0, 64, p.sort_ws, st, &alt2,
                                           p.d_outs, long_flag));
    }
    return QO_OK;
}

}  // namespace qo

struct qo_batch {
    void *arena = nullptr;
    qo::JoinBatchPlan plan;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
};

using namespace qo;

extern "C" {

int qo_batch_create(const qo_join_query *queries, int nq, qo_batch **out) {
    QO_ENTER();
    QO_REQUIRE(out, "null output");
    *out = nullptr;
    qo_batch *b = new qo_batch;
    int rc = join_plan_init(queries, nq, b->plan);
    if (rc != QO_OK) {
        delete b;
        return rc;
    }
    const size_t bytes = join_plan_bytes(b->plan);
    hipError_t e = hipMalloc(&b->arena, bytes);
    if (e != hipSuccess) {
        delete b;
        return set_error(QO_ERR_NO_MEMORY, "batch arena of %zu bytes: %s", bytes, hipGetErrorString(e));
    }
    join_plan_bind(b->plan, b->arena);
    if (hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess)
        rc = set_error(QO_ERR_HIP, "batch stream");
    if (rc == QO_OK) rc = join_plan_upload(b->plan, b->cap_stream);
    if (rc == QO_OK && hipStreamSynchronize(b->cap_stream) != hipSuccess)
        rc = set_error(QO_ERR_HIP, "batch descriptor upload failed");
    if (rc == QO_OK) {
        e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) rc = set_error(QO_ERR_HIP, "begin capture: %s", hipGetErrorString(e));
    }
    if (rc == QO_OK) {
        int rec = join_plan_launch(b->plan, b->cap_stream);
        e = hipStreamEndCapture(b->cap_stream, &b->graph);
        if (rec != QO_OK) rc = rec;
        else if (e != hipSuccess) rc = set_error(QO_ERR_HIP, "end capture: %s", hipGetErrorString(e));
    }
    if (rc == QO_OK) {
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = set_error(QO_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    if (rc != QO_OK) {
        qo_batch_destroy(b);
        return rc;
    }
    *out = b;
    return QO_OK;
}

int qo_batch_launch(qo_batch *b, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(b && b->exec, "batch not created");
    QO_HIP(hipGraphLaunch(b->exec, stream_of(s)));
    return QO_OK;
}

int qo_batch_counts(qo_batch *b, uint64_t *h_counts) {
    QO_TRY(ensure_init());
    QO_REQUIRE(b && h_counts, "null argument");
    QO_HIP(hipDeviceSynchronize());
    QO_HIP(hipMemcpy(h_counts, b->plan.counts, 8 * (size_t)b->plan.nq, hipMemcpyDeviceToHost));
    return QO_OK;
}

int qo_batch_destroy(qo_batch *b) {
    if (!b) return QO_OK;
    QO_TRY(ensure_init());
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->arena) (void)hipFree(b->arena);
    delete b;
    return QO_OK;
}

}  // extern "C"
