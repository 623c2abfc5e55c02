// Autotuner: searches an operator's launch parameters at one size class on
// the device and keeps the fastest in the parameter table (per operator and
// size class, floor(log2 n)).  qo_tune_save writes the table as the tune file
// that libqo loads at init (tune/<arch>.tune beside the library).
//
// For each operator it builds a synthetic workload of the size class in HBM,
// times every candidate with HIP events (median of `reps` runs after one
// warm-up; the batch is re-planned per candidate since its launch parameters
// are fixed when the graph is captured) and sets the winner.  The candidates
// are the knobs each operator reads from the table:
//   TUNE_FILTER_AGG   loads in flight per column {2, 4, 8} x blocks per CU {1, 2, 4, 8}
//   TUNE_GROUPBY      LDS table slots {2048, 4096} x blocks per CU {1, 2, 4}
//                     (direct path; 1024 groups, Zipf z = 0.5)
//   TUNE_GROUPBY_PART groups per fine bucket {512, 1024, 2048} (2^20 groups, z = 0.5)
//   TUNE_BATCH        S tiles per probe ticket {1, 2, 4, 8} x R rows per build
//                     tile {2048, 4096, 8192} (joins of |R| = 2^class, |S| = 8|R|)
//   TUNE_BATCH_LOAD   batch hash table load {33, 50, 66, 80} % (same batch workload)
//   TUNE_JOIN_PART    LDS table slice of the partitioned hash-mixed build {16, 24, 32, 48, 60} KiB
//                     (build and probe of |R| = 2^class sparse 31-bit keys, |S| = 4|R|, half selected)
//
// The reference's analogue is the per-variant cost table its planner reads
// (host/dynamic_programming.cpp:133-185), measured offline on the FPGA; here
// the measurement runs on the device that will use the parameters.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "qo_internal.hpp"
#include "qo_measure.hpp"

namespace qo {
namespace {

struct Cand {
    int block, items, param;
};

std::vector<Cand> candidates(int op) {
    std::vector<Cand> c;
    switch (op) {
        case TUNE_FILTER_AGG:
            for (int items : {2, 4, 8})
                for (int bpc : {1, 2, 4, 8}) c.push_back({256, items, bpc});
            break;
        case TUNE_GROUPBY:
            for (int slots : {2048, 4096})
                for (int bpc : {1, 2, 4}) c.push_back({1024, slots, bpc});
            break;
        case TUNE_GROUPBY_PART:
            for (int g : {512, 1024, 2048}) c.push_back({512, g, PART_AUTO});
            break;
        case TUNE_BATCH:
            for (int tiles : {1, 2, 4, 8})
                for (int bt : {2048, 4096, 8192}) c.push_back({256, tiles, bt});
            break;
        case TUNE_BATCH_LOAD:
            for (int load : {33, 50, 66, 80}) c.push_back({256, 4, load});
            break;
        case TUNE_JOIN_PART:
            for (int kib : {16, 24, 32, 48, 60}) c.push_back({256, kib, PART_AUTO});
            break;
        default: break;
    }
    return c;
}

int gen_zipf(int32_t *out, size_t n, uint32_t G, double z, uint64_t seed, hipStream_t st) {
    std::vector<uint64_t> t(G);
    QO_TRY(qo_zipf_table(t.data(), G, z));
    DevBuf dt;
    QO_TRY(dt.alloc(G * 8));
    QO_HIP(hipMemcpy(dt.p, t.data(), G * 8, hipMemcpyHostToDevice));
    QO_TRY(qo_gen_zipf_u32(reinterpret_cast<uint32_t *>(out), n, dt.as<uint64_t>(), G, seed, 5, seed + 7, st));
    QO_HIP(hipStreamSynchronize(st));
    return QO_OK;
}

// The workload of one (operator, size class): prepare() after the candidate's
// parameters are set (the batch is planned and captured there), run() once per
// timed step, release() after the candidate.
struct Workload {
    std::vector<DevBuf> bufs;
    std::vector<qo_join_query> jq;
    qo_batch *batch = nullptr;
    std::function<int(hipStream_t)> run;
    std::function<int()> prepare = [] { return QO_OK; };
    std::function<void()> release = [] {};
    int32_t *col(size_t i) const { return bufs[i].as<int32_t>(); }
    ~Workload() {
        if (batch) qo_batch_destroy(batch);
    }
};

int make_workload(int op, size_t n, hipStream_t st, Workload &w) {
    const uint64_t seed = 0x7E57ull + (uint64_t)op;
    if (op == TUNE_FILTER_AGG) {
        w.bufs = std::vector<DevBuf>(3);
        QO_TRY(w.bufs[0].alloc(n * 4));
        QO_TRY(w.bufs[1].alloc(n * 4));
        QO_TRY(w.bufs[2].alloc(64));
        QO_TRY(qo_gen_uniform_u32(w.bufs[0].as<uint32_t>(), n, seed, 0, 0, 1u << 31, st));
        QO_TRY(qo_gen_uniform_u32(w.bufs[1].as<uint32_t>(), n, seed, 1, 0, 1000, st));
        int32_t *k = w.col(0), *v = w.col(1);
        int64_t *o = w.bufs[2].as<int64_t>();
        w.run = [k, v, o, n](hipStream_t s) { return qo_filter_agg(k, v, n, 0, (1 << 30) - 1, o, s); };
        return QO_OK;
    }
    if (op == TUNE_GROUPBY || op == TUNE_GROUPBY_PART) {
        const uint32_t G = op == TUNE_GROUPBY ? 1024u : (1u << 20);
        w.bufs = std::vector<DevBuf>(8);
        QO_TRY(w.bufs[0].alloc(n * 4));
        QO_TRY(w.bufs[1].alloc(n * 4));
        const size_t sz[6] = {4, 8, 8, 4, 4, 1};
        for (int i = 0; i < 6; ++i) QO_TRY(w.bufs[2 + i].alloc(i == 5 ? 8 : G * sz[i]));
        QO_TRY(gen_zipf(w.col(0), n, G, 0.5, seed, st));
        QO_TRY(qo_gen_uniform_u32(w.bufs[1].as<uint32_t>(), n, seed, 1, 0, 1u << 16, st));
        int32_t *k = w.col(0), *v = w.col(1);
        int32_t *ok = w.col(2), *omn = w.col(5), *omx = w.col(6);
        int64_t *oc = w.bufs[3].as<int64_t>(), *os = w.bufs[4].as<int64_t>();
        uint64_t *num = w.bufs[7].as<uint64_t>();
        w.run = [=](hipStream_t s) { return qo_groupby_agg(k, v, n, G, ok, oc, os, omn, omx, num, s); };
        return QO_OK;
    }
    if (op == TUNE_JOIN_PART) {
        // Hash-mixed table (sparse keys), built and probed once per step.
        const size_t ns = std::min<size_t>(4 * n, size_t(1) << 28);
        const size_t tb = qo_join_table_bytes(n);
        w.bufs = std::vector<DevBuf>(10);
        for (int i = 0; i < 2; ++i) QO_TRY(w.bufs[i].alloc(n * 4));
        for (int i = 2; i < 8; ++i) QO_TRY(w.bufs[i].alloc(ns * 4));
        QO_TRY(w.bufs[8].alloc(tb));
        QO_TRY(w.bufs[9].alloc(64));
        QO_TRY(qo_gen_bij_u32(w.bufs[0].as<uint32_t>(), n, 31, seed, st));
        QO_TRY(qo_gen_uniform_u32(w.bufs[1].as<uint32_t>(), n, seed, 1, 0, 0, st));
        QO_TRY(qo_gen_fk_u32(w.bufs[2].as<uint32_t>(), ns, seed, 2, n, 31, seed, st));
        QO_TRY(qo_gen_uniform_u32(w.bufs[3].as<uint32_t>(), ns, seed, 3, 0, 0, st));
        QO_TRY(qo_gen_uniform_u32(w.bufs[4].as<uint32_t>(), ns, seed, 4, 0, 100, st));
        QO_HIP(hipStreamSynchronize(st));
        int32_t *rk = w.col(0), *rp = w.col(1), *sk = w.col(2), *sp = w.col(3), *sf = w.col(4);
        int32_t *ok = w.col(5), *orp = w.col(6), *osp = w.col(7);
        void *table = w.bufs[8].p;
        uint64_t *cnt = w.bufs[9].as<uint64_t>();
        w.run = [=](hipStream_t s) {
            QO_TRY(qo_hash_build(table, tb, rk, rp, n, s));
            return qo_hash_probe(table, tb, sk, sp, sf, ns, INT32_MIN, 49, ok, orp, osp, ns, cnt, s);
        };
        return QO_OK;
    }
    if (op == TUNE_BATCH || op == TUNE_BATCH_LOAD) {
        // Queries of |R| = n, |S| = 8n, filter passing 25 %; at most 2^28 S rows in all.
        const size_t ns = 8 * n;
        const int nq = (int)std::max<size_t>(1, std::min<size_t>(64, (size_t(1) << 28) / ns));
        w.bufs = std::vector<DevBuf>((size_t)nq * 8);
        w.jq.resize(nq);
        for (int q = 0; q < nq; ++q) {
            DevBuf *b = &w.bufs[(size_t)q * 8];
            for (int i = 0; i < 2; ++i) QO_TRY(b[i].alloc(n * 4));
            for (int i = 2; i < 8; ++i) QO_TRY(b[i].alloc(ns * 4));
            const uint64_t s = seed + 100 + (uint64_t)q;
            QO_TRY(qo_gen_bij_u32(b[0].as<uint32_t>(), n, 31, s, st));
            QO_TRY(qo_gen_uniform_u32(b[1].as<uint32_t>(), n, s, 1, 0, 0, st));
            QO_TRY(qo_gen_fk_u32(b[2].as<uint32_t>(), ns, s, 2, n, 31, s, st));
            QO_TRY(qo_gen_uniform_u32(b[3].as<uint32_t>(), ns, s, 3, 0, 0, st));
            QO_TRY(qo_gen_uniform_u32(b[4].as<uint32_t>(), ns, s, 4, 0, 100, st));
            qo_join_query &j = w.jq[q];
            j = qo_join_query{};
            j.rkey = b[0].as<int32_t>(); j.rpay = b[1].as<int32_t>();
            j.skey = b[2].as<int32_t>(); j.spay = b[3].as<int32_t>(); j.sfilt = b[4].as<int32_t>();
            j.nr = n; j.ns = ns; j.flo = INT32_MIN; j.fhi = 24;
            j.okey = b[5].as<int32_t>(); j.orpay = b[6].as<int32_t>(); j.ospay = b[7].as<int32_t>();
            j.out_capacity = ns;
        }
        QO_HIP(hipStreamSynchronize(st));
        // The candidate's parameters are read when the batch is planned.
        w.prepare = [&w] {
            if (w.batch) qo_batch_destroy(w.batch);
            w.batch = nullptr;
            return qo_batch_create(w.jq.data(), (int)w.jq.size(), &w.batch);
        };
        w.release = [&w] {
            if (w.batch) qo_batch_destroy(w.batch);
            w.batch = nullptr;
        };
        w.run = [&w](hipStream_t s) { return qo_batch_launch(w.batch, s); };
        return QO_OK;
    }
    return set_error(QO_ERR_INVALID_ARG, "operator %d has no tuning workload", op);
}

}  // namespace
}  // namespace qo

using namespace qo;

extern "C" {

int qo_autotune(int op, int size_class, int reps, int *block, int *items, int *param, double *best_ms,
                double *default_ms) {
    QO_ENTER();
    QO_REQUIRE(size_class >= 4 && size_class <= 32, "size class %d outside [4, 32]", size_class);
    const std::vector<Cand> cands = candidates(op);
    QO_REQUIRE(!cands.empty(), "operator %d has no tunable candidates", op);
    reps = std::max(1, std::min(reps, 50));
    const size_t n = size_t(1) << size_class;
    Timer tm;
    QO_TRY(tm.init());
    Workload w;
    QO_TRY(make_workload(op, n, tm.st, w));
    // The built-in defaults first (the reference point of the reported gain).
    const TuneParams d = tune_builtin(op);
    QO_TRY(qo_tune_set(op, size_class, d.block, d.items, d.param));
    double t_def = 0;
    QO_TRY(w.prepare());
    QO_TRY(tm.median(reps, w.run, &t_def));
    w.release();
    double best = 1e30;
    Cand win = cands[0];
    for (const Cand &c : cands) {
        if (op == TUNE_GROUPBY && c.items == 4096 && (size_t)props().lds_per_block < 4096 * 24 + 8) continue;
        QO_TRY(qo_tune_set(op, size_class, c.block, c.items, c.param));
        double t = 0;
        QO_TRY(w.prepare());
        QO_TRY(tm.median(reps, w.run, &t));
        w.release();
        if (t < best) {
            best = t;
            win = c;
        }
    }
    QO_TRY(qo_tune_set(op, size_class, win.block, win.items, win.param));
    if (block) *block = win.block;
    if (items) *items = win.items;
    if (param) *param = win.param;
    if (best_ms) *best_ms = best;
    if (default_ms) *default_ms = t_def;
    return QO_OK;
}

}  // extern "C"
