// Join batch plan shared by the graph-captured batch API and the serving loop.
#pragma once
#include <vector>

#include "qo_internal.hpp"
#include "qo_join.hpp"
#include "qo_sort.hpp"

namespace qo {

constexpr uint32_t kBuildTile = 4096;       // default R rows per build tile (tune op TUNE_BATCH, param)
constexpr uint32_t kProbeTicketTiles = 4;   // default S tiles per probe ticket (tune op TUNE_BATCH, items)

// Device descriptor of one query of a batch.
struct BQuery {
    const int32_t *rkey, *rpay, *skey, *spay, *sfilt;
    uint64_t nr, ns;
    int32_t flo, fhi;
    int32_t *okey, *orpay, *ospay;
    uint64_t out_cap;
    JoinHeader *hdr;        // table header; slots follow it
    uint64_t table_cap;     // slots
    uint64_t m_base;        // first element of this query in the match arrays
    unsigned long long *count;
    uint32_t rtile0, nrtile;
    uint32_t stile0, nstile;
    uint32_t mtile0, nmtile;
    uint32_t vec;           // S columns are 16-byte aligned
    uint32_t pad;
};

struct JoinBatchPlan {
    int nq = 0;
    std::vector<BQuery> hq;
    std::vector<RadixSeg> hs;
    std::vector<SweepChunk> hchunks;  // up-front digit-count blocks of the sort
    std::vector<uint32_t> hchunk0;    // query q owns chunks [hchunk0[q], hchunk0[q + 1])
    std::vector<RowOut> houts;        // query q's output columns (the sort's last pass writes them)
    uint64_t total_slots = 0, total_match = 0;
    uint32_t rt = 0, stt = 0, mt = 0;
    uint32_t build_tile = kBuildTile;        // R rows per build tile
    uint32_t probe_group = kProbeTicketTiles;  // S tiles per probe ticket
    size_t small_words = 0;
    // Queries whose tables outgrow the direct probe: probed one after another by
    // the plain path's partitioned probe (split of S by table slice, L2- or
    // LDS-resident slices), sharing one workspace; the persistent probe skips them.
    std::vector<int> part_q;
    // Queries whose build side is large enough for the plain path's partitioned
    // build (split of R by table slice, each slice built in LDS and written out
    // whole); the persistent build skips them.  Their tables are placed after
    // all others and need no fill, since every slot is written by the build.
    std::vector<int> pbuild_q;
    uint64_t fill_slots = 0;  // slots of the tables before them (filled empty)
    size_t part_ws_bytes = 0;
    // Bound by join_plan_bind:
    uint64_t *slots_all = nullptr;
    uint32_t *small = nullptr;  // tickets, then per-query match counts
    BQuery *d_q = nullptr;
    RadixSeg *d_segs = nullptr;
    uint64_t *mkeys = nullptr, *kalt = nullptr;
    uint32_t *mvals = nullptr, *valt = nullptr;
    SweepChunk *d_chunks = nullptr;
    uint32_t *d_chunk0 = nullptr;
    RowOut *d_outs = nullptr;
    void *sort_ws = nullptr;
    void *part_ws = nullptr;
    unsigned long long *counts = nullptr;
};

// Sizes the plan (host only).
int join_plan_init(const qo_join_query *queries, int nq, JoinBatchPlan &p);
// Device bytes the plan needs.
size_t join_plan_bytes(JoinBatchPlan &p);
// Places the plan's buffers in `arena` and fills the per-query device pointers.
void join_plan_bind(JoinBatchPlan &p, void *arena);
// Copies the descriptors to the device (the plan must outlive the copy).
int join_plan_upload(const JoinBatchPlan &p, hipStream_t st);
// Enqueues every stage (fills, build, probe, segmented sort, unpack) on st.
int join_plan_launch(const JoinBatchPlan &p, hipStream_t st);

}  // namespace qo
