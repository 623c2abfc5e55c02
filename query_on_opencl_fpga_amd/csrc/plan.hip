// Cost-model planner for filter -> group-by -> aggregate queries over Records.
//
// Reference: the FPGA planner of host/dynamic_programming.cpp.  There, an
// operator chain (selection = map, scan, write; group + aggregation = sort,
// group label, scan, write, aggregate; their concatenation) is cut into
// consecutive segments, each segment gets the kernel variants that minimise
// its time under the device's resource budget (recursive_time_compute,
// :331-386, from the per-variant unit costs of kernel_primitive, :133-185),
// and a dynamic programme over the chain's prefixes adds the cost of ending a
// segment there (the reconfiguration and the intermediate's round trip,
// dp_for_selection, :874-996).
//
// Here the chain's steps are FILTER, GROUP+AGGREGATE (one step: the group-by
// only exists to aggregate) or REDUCE, and a segment is one of libqo's
// implementations of a contiguous run of steps: the one-pass filter
// compaction, the sort-based group-by pipeline (the reference's), the hash
// group-by with the reference's aggregate semantics, the hash group-by with the
// filter fused into its kernels, and the no-group reduce.  An implementation's
// cost is a sum of primitive costs, each (fixed microseconds + bytes moved /
// measured bandwidth); the primitive table holds MI355X measurements and
// qo_plan_calibrate re-measures it on the current device.  A segment boundary
// costs what the producing implementation writes and the next one reads, which
// those byte counts already hold.  The dynamic programme over prefixes picks the
// cheapest cover; each segment also reports the launch parameters the tune
// table gives its operator at its size class.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "qo_device.hpp"
#include "qo_internal.hpp"
#include "qo_measure.hpp"

namespace qo {

// Primitive costs: fixed launch cost (us) and bandwidth over the bytes the
// primitive moves (GB/s).  Defaults: qo_plan_calibrate at 2^26 rows on one
// MI355X (the fixed cost is the time at 2^12 rows).
struct PrimCost {
    double fixed_us, gbps;
};
static PrimCost g_cost[QO_COST_NUM] = {
    {13.0, 4300.0},  // QO_COST_FILTER: filter compaction, 8 B read + 8 B per selected row written
    {15.0, 4000.0},  // QO_COST_SORT_PASS: one radix pass, 24 B per row (histogram read, read, write)
    {12.0, 3300.0},  // QO_COST_STARTPOS: group start positions, 8 B per row
    {28.0, 850.0},   // QO_COST_AGG: one per-group aggregate, 8 B per row (1024 groups)
    {6.0, 5000.0},   // QO_COST_UNPACK: Records -> columns (or a Record copy), 16 B per row
    {81.0, 3400.0},  // QO_COST_HASH_SMALL: hash group-by, at most 1024 groups, 8 B per row
    {245.0, 1260.0}, // QO_COST_HASH_LARGE: hash group-by, more groups (partitioned), 8 B per row
    {10.0, 6300.0},  // QO_COST_REDUCE: no-group reduce, 8 B per row
    {75.0, 3960.0},  // QO_COST_FHASH_SMALL: filtered hash group-by (half selected), <= 1024 groups
    {222.0, 1300.0}, // QO_COST_FHASH_LARGE: filtered hash group-by (half selected), more groups
};
static std::mutex g_cost_mu;

static double prim_ms(int p, double bytes) {
    const PrimCost c = g_cost[p];
    return c.fixed_us * 1e-3 + bytes / (c.gbps * 1e9) * 1e3;
}

// Steps of a chain.
enum Step { STEP_FILTER = 0, STEP_GROUP_AGG = 1, STEP_REDUCE = 2 };

struct Chain {
    int steps[3];
    int n = 0;
};

static Chain chain_of(const qo_query &q) {
    Chain c;
    if (q.filter) c.steps[c.n++] = STEP_FILTER;
    if (q.group) c.steps[c.n++] = STEP_GROUP_AGG;
    else if (q.op >= 0) c.steps[c.n++] = STEP_REDUCE;
    return c;
}

// Rows reaching step i of the chain.
static double rows_at(const qo_query &q, const Chain &c, int i, double n) {
    double m = n;
    for (int k = 0; k < i; ++k)
        if (c.steps[k] == STEP_FILTER) m *= q.selectivity;
    return m;
}

// Steps an implementation covers, as a contiguous run starting at `first`;
// returns the run length (0: it does not fit there).
static int covers(int impl, const Chain &c, int first) {
    auto at = [&](int k) { return first + k < c.n ? c.steps[first + k] : -1; };
    switch (impl) {
        case QO_IMPL_FILTER: return at(0) == STEP_FILTER ? 1 : 0;
        case QO_IMPL_SORT_GROUP_AGG:
        case QO_IMPL_HASH_GROUP_AGG: return at(0) == STEP_GROUP_AGG ? 1 : 0;
        case QO_IMPL_FILTER_HASH_GROUP_AGG: return at(0) == STEP_FILTER && at(1) == STEP_GROUP_AGG ? 2 : 0;
        case QO_IMPL_REDUCE: return at(0) == STEP_REDUCE ? 1 : 0;
        default: return 0;
    }
}

// Estimated ms of `impl` on m input rows (n = rows of the original input).
// `owns_input`: the rows are a temporary of the plan (sorted in place), not
// the caller's Records (copied before an in-place sort).
static double impl_ms(int impl, const qo_query &q, double m, bool owns_input) {
    const double G = std::max(1.0, std::min(q.groups, std::max(m, 1.0)));
    const int hash = G <= 1024.0 ? QO_COST_HASH_SMALL : QO_COST_HASH_LARGE;
    switch (impl) {
        case QO_IMPL_FILTER: return prim_ms(QO_COST_FILTER, 8 * m + 8 * m * q.selectivity);
        case QO_IMPL_SORT_GROUP_AGG: {
            double t = owns_input ? 0.0 : prim_ms(QO_COST_UNPACK, 16 * m);
            for (int p = 0; p < 4; ++p) t += prim_ms(QO_COST_SORT_PASS, 24 * m);
            return t + prim_ms(QO_COST_STARTPOS, 8 * m + 4 * G) + prim_ms(QO_COST_AGG, 8 * m + 8 * G);
        }
        case QO_IMPL_HASH_GROUP_AGG:
            return prim_ms(QO_COST_UNPACK, 16 * m) + prim_ms(hash, 8 * m + 64 * G);
        case QO_IMPL_FILTER_HASH_GROUP_AGG: {
            // The fused kernels read every row but update the tables for the
            // selected ones only: time taken as linear in the selectivity s
            // between the measured filtered group-by (s = 1/2) and the plain
            // one (s = 1), and never below the filtered one's fixed cost.
            const int fhash = G <= 1024.0 ? QO_COST_FHASH_SMALL : QO_COST_FHASH_LARGE;
            const double t1 = prim_ms(hash, 8 * m + 64 * G), th = prim_ms(fhash, 8 * m + 64 * G);
            const double t = std::max(g_cost[fhash].fixed_us * 1e-3, t1 + 2.0 * (1.0 - q.selectivity) * (th - t1));
            return prim_ms(QO_COST_UNPACK, 16 * m) + t;
        }
        case QO_IMPL_REDUCE: return prim_ms(QO_COST_REDUCE, 8 * m);
        default: return 1e30;
    }
}

// The tune-table operator that sets an implementation's launch parameters.
static int tune_op_of(int impl, const qo_query &q, double m) {
    switch (impl) {
        case QO_IMPL_FILTER: return TUNE_FILTER;
        case QO_IMPL_SORT_GROUP_AGG: return TUNE_SORT;
        case QO_IMPL_HASH_GROUP_AGG:
        case QO_IMPL_FILTER_HASH_GROUP_AGG:
            return std::min(q.groups, std::max(m, 1.0)) <= 1024.0 ? TUNE_GROUPBY : TUNE_GROUPBY_PART;
        default: return -1;
    }
}

static int plan_impl(const qo_query &q, size_t n, unsigned allowed, qo_plan *out) {
    QO_REQUIRE(out, "null plan");
    QO_REQUIRE(q.selectivity >= 0.0 && q.selectivity <= 1.0, "selectivity %g outside [0, 1]", q.selectivity);
    QO_REQUIRE(q.op >= -1 && q.op <= 3, "op %d is not -1 (none) or REDUCE_SUM/MAX/MIN/AVERAGE", q.op);
    QO_REQUIRE(!(q.group && q.op < 0), "a grouped query needs an aggregate operator");
    const Chain c = chain_of(q);
    QO_REQUIRE(c.n > 0, "empty query: no filter, group or reduce");
    if (allowed == 0) allowed = (1u << QO_IMPL_NUM) - 1;
    std::lock_guard<std::mutex> lk(g_cost_mu);
    // best[i]: cheapest cover of steps [0, i); from[i] / via[i]: its last segment.
    double best[4];
    int from[4], via[4];
    best[0] = 0.0;
    for (int i = 1; i <= c.n; ++i) {
        best[i] = INFINITY;
        from[i] = via[i] = -1;
        for (int j = 0; j < i; ++j) {
            if (!std::isfinite(best[j])) continue;
            const double m = rows_at(q, c, j, (double)n);
            for (int impl = 0; impl < QO_IMPL_NUM; ++impl) {
                if (!(allowed >> impl & 1u) || covers(impl, c, j) != i - j) continue;
                const double t = best[j] + impl_ms(impl, q, m, j > 0);
                if (t < best[i]) {
                    best[i] = t;
                    from[i] = j;
                    via[i] = impl;
                }
            }
        }
    }
    if (!std::isfinite(best[c.n]))
        return set_error(QO_ERR_NOT_FOUND, "no allowed implementation set covers this query (mask 0x%x)", allowed);
    qo_plan p{};
    int segs[QO_PLAN_MAX_SEGS], ns = 0;
    for (int i = c.n; i > 0; i = from[i]) segs[ns++] = i;
    p.nseg = ns;
    for (int s = 0; s < ns; ++s) {
        const int end = segs[ns - 1 - s], beg = from[end];
        const double m = rows_at(q, c, beg, (double)n);
        p.impl[s] = via[end];
        p.first_step[s] = beg;
        p.last_step[s] = end - 1;
        p.seg_ms[s] = best[end] - best[beg];
        p.tune_op[s] = tune_op_of(via[end], q, m);
        if (p.tune_op[s] >= 0) {
            const TuneParams tp = tune_get(p.tune_op[s], (size_t)std::max(m, 1.0));
            p.block[s] = tp.block;
            p.items[s] = tp.items;
            p.param[s] = tp.param;
        }
    }
    p.est_ms = best[c.n];
    *out = p;
    return QO_OK;
}

// ---- execution -------------------------------------------------------------------------
// keys[g] = sorted[startPos[g]].y.
__global__ void k_plan_keys_at(const qo_record *sorted, const int32_t *startPos, size_t G, uint32_t *keys) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < G) keys[g] = sorted[startPos[g]].y;
}

// Hash group-by rows (ascending by signed key) -> the sort pipeline's output:
// groups ascending by unsigned key (the signed order rotated: keys with the
// sign bit set go last; nneg of them lead the input) and the reference's
// aggregate of .x per group: SUM wraps to int32, MAX starts at 0, MIN at
// 2^30, AVERAGE is the wrapped sum divided by the count, truncated.
__global__ void k_plan_epilogue(const int32_t *okey, const int64_t *ocount, const int64_t *osum,
                                const int32_t *omin, const int32_t *omax, const uint64_t *d_num, int op,
                                uint32_t *keys, int32_t *aggs) {
    const size_t G = (size_t)*d_num;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    // nneg = groups with a negative signed key (a prefix of the signed order).
    size_t lo = 0, hi = G;
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (okey[mid] < 0) lo = mid + 1;
        else hi = mid;
    }
    const size_t nneg = lo;
    const size_t src = g < G - nneg ? nneg + g : g - (G - nneg);
    keys[g] = (uint32_t)okey[src];
    int32_t v;
    if (op == 1) v = std::max<int32_t>(0, omax[src]);
    else if (op == 2) v = std::min<int32_t>(1 << 30, omin[src]);
    else {
        const int32_t s = (int32_t)(uint32_t)(uint64_t)osum[src];
        v = op == 3 ? (int32_t)((int64_t)s / ocount[src]) : s;
    }
    aggs[g] = v;
}

// Device -> host read of a few bytes ordered after the work on st.
static int read_back(void *dst, const void *src, size_t bytes, hipStream_t st) {
    QO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    return sync_and_check(st);
}

// Temporaries of one execution, carved from the stream's second workspace
// (the operators a plan calls use the first).
struct Temps {
    Carver cv{nullptr};
    size_t cap_bytes = 0;
    template <class T>
    T *get(size_t count, int *status) {
        T *p = cv.take<T>(std::max<size_t>(count, 1));
        if (cv.off > cap_bytes) {
            *status = set_error(QO_ERR_INVALID_ARG, "plan temporaries exceed their workspace");
            return nullptr;
        }
        return p;
    }
};

// Upper bound of the temporaries qo_plan_execute takes for n rows: every
// segment's buffers at n rows (a filter only shrinks what follows), each
// rounded up to the carver's 256-byte alignment.
static size_t plan_temp_bytes(const qo_plan &p, size_t n, size_t cap) {
    size_t b = 256 + 8;
    const size_t rows = std::max<size_t>(n, 1), c = std::max<size_t>(cap, 1);
    for (int s = 0; s < p.nseg; ++s) {
        switch (p.impl[s]) {
            case QO_IMPL_FILTER: b += 256 + 8 * rows; break;
            case QO_IMPL_SORT_GROUP_AGG: b += 2 * 256 + 12 * rows; break;
            case QO_IMPL_HASH_GROUP_AGG:
            case QO_IMPL_FILTER_HASH_GROUP_AGG: b += 7 * 256 + 8 * rows + 28 * c; break;
            default: break;
        }
    }
    return b;
}

// One primitive's measurement: time fn at n_small and n_large rows; fixed
// cost = the small-size time (its bytes are negligible), bandwidth = the
// large-size bytes over the time above that.
static int fit_cost(Timer &tm, int reps, const std::function<int(hipStream_t, size_t)> &fn, double bytes_small,
                    double bytes_large, size_t n_small, size_t n_large, PrimCost *out) {
    double t1 = 0, t2 = 0;
    QO_TRY(tm.median(reps, [&](hipStream_t st) { return fn(st, n_small); }, &t1));
    QO_TRY(tm.median(reps, [&](hipStream_t st) { return fn(st, n_large); }, &t2));
    if (t2 > t1 && bytes_large > bytes_small) {
        out->fixed_us = t1 * 1e3;
        out->gbps = (bytes_large - bytes_small) / ((t2 - t1) * 1e-3) / 1e9;
    }
    return QO_OK;
}

static int calibrate_impl(int cls, int reps) {
    const size_t n = size_t(1) << cls, n0 = size_t(1) << 12;
    const uint32_t g_small = 1024, g_large = (uint32_t)std::min<size_t>(size_t(1) << 20, n / 8);
    Timer tm;
    QO_TRY(tm.init());
    DevBuf rec, rec2, work, out, cx, cy, sp, okey, ocnt, osum, omin, omax, cnt;
    QO_TRY(rec.alloc(n * 8));
    QO_TRY(rec2.alloc(n * 8));
    QO_TRY(work.alloc(n * 8));
    QO_TRY(out.alloc(n * 8));
    QO_TRY(cx.alloc(n * 4));
    QO_TRY(cy.alloc(n * 4));
    QO_TRY(sp.alloc(n * 4));
    QO_TRY(okey.alloc((size_t)g_large * 4));
    QO_TRY(ocnt.alloc((size_t)g_large * 8));
    QO_TRY(osum.alloc((size_t)g_large * 8));
    QO_TRY(omin.alloc((size_t)g_large * 4));
    QO_TRY(omax.alloc((size_t)g_large * 4));
    QO_TRY(cnt.alloc(64));
    qo_stream s = tm.st;
    // Records: .x uniform 32-bit, .y uniform in [0, 1024) (rec) or [0, g_large) (rec2).
    QO_TRY(qo_gen_uniform_u32(cx.as<uint32_t>(), n, 0x9A11, 0, 0, 0, s));
    QO_TRY(qo_gen_uniform_u32(cy.as<uint32_t>(), n, 0x9A11, 1, 0, g_small, s));
    QO_TRY(qo_pack_records(cx.as<uint32_t>(), cy.as<uint32_t>(), n, rec.as<qo_record>(), s));
    QO_TRY(qo_gen_uniform_u32(cy.as<uint32_t>(), n, 0x9A11, 2, 0, g_large, s));
    QO_TRY(qo_pack_records(cx.as<uint32_t>(), cy.as<uint32_t>(), n, rec2.as<qo_record>(), s));
    QO_TRY(sync_and_check(tm.st));
    PrimCost c[QO_COST_NUM];
    {
        std::lock_guard<std::mutex> lk(g_cost_mu);
        std::copy(g_cost, g_cost + QO_COST_NUM, c);
    }
    const qo_record *R = rec.as<qo_record>();
    const double dn = (double)n, dn0 = (double)n0;
    // Filter: about half the rows selected (.y < 512 of [0, 1024)).
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) {
        return qo_filter_records(R, m, 0, 511, out.as<qo_record>(), cnt.as<uint64_t>(), st); }, 12 * dn0, 12 * dn,
        n0, n, &c[QO_COST_FILTER]));
    // The sort's four passes, the copy that restores the input taken off.
    double tc1 = 0, tc2 = 0, ts1 = 0, ts2 = 0;
    QO_TRY(tm.median(reps, [&](hipStream_t st) { return qo_copy(work.p, R, n0 * 8, st); }, &tc1));
    QO_TRY(tm.median(reps, [&](hipStream_t st) { return qo_copy(work.p, R, n * 8, st); }, &tc2));
    auto copy_sort = [&](hipStream_t st, size_t m) -> int {
        QO_TRY(qo_copy(work.p, R, m * 8, st));
        return qo_sort_records(work.as<qo_record>(), m, st);
    };
    QO_TRY(tm.median(reps, [&](hipStream_t st) { return copy_sort(st, n0); }, &ts1));
    QO_TRY(tm.median(reps, [&](hipStream_t st) { return copy_sort(st, n); }, &ts2));
    const double p1 = std::max(1e-6, (ts1 - tc1) / 4), p2 = (ts2 - tc2) / 4;
    if (p2 > p1) c[QO_COST_SORT_PASS] = PrimCost{p1 * 1e3, 24 * (dn - dn0) / ((p2 - p1) * 1e-3) / 1e9};
    // Start positions and one aggregate on the sorted rows (work holds n sorted rows).
    QO_TRY(copy_sort(tm.st, n));
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) {
        return qo_group_startpos(work.as<qo_record>(), m, sp.as<int32_t>(), cnt.as<uint64_t>(), st); },
        8 * dn0, 8 * dn, n0, n, &c[QO_COST_STARTPOS]));
    QO_TRY(qo_group_startpos(work.as<qo_record>(), n, sp.as<int32_t>(), cnt.as<uint64_t>(), tm.st));
    uint64_t groups = 0;
    QO_HIP(hipMemcpyAsync(&groups, cnt.p, 8, hipMemcpyDeviceToHost, tm.st));
    QO_TRY(sync_and_check(tm.st));
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) {
        // Small size: the first m rows as one group (startPos[0] = 0); full size: every group.
        return qo_aggregate_groups(work.as<qo_record>(), m, sp.as<int32_t>(), m == n ? groups : 1, 0,
                                   out.as<int32_t>(), st); },
        8 * dn0, 8 * dn + 8 * (double)groups, n0, n, &c[QO_COST_AGG]));
    // Records -> columns.
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) {
        return qo_unpack_records(R, m, cx.as<uint32_t>(), cy.as<uint32_t>(), st); }, 16 * dn0, 16 * dn, n0, n,
        &c[QO_COST_UNPACK]));
    // Hash group-by: 1024 groups, then g_large groups.
    QO_TRY(qo_unpack_records(R, n, cx.as<uint32_t>(), cy.as<uint32_t>(), tm.st));
    auto hash = [&](hipStream_t st, size_t m, size_t cap) {
        return qo_groupby_agg(cy.as<int32_t>(), cx.as<int32_t>(), m, cap, okey.as<int32_t>(), ocnt.as<int64_t>(),
                              osum.as<int64_t>(), omin.as<int32_t>(), omax.as<int32_t>(), cnt.as<uint64_t>(), st);
    };
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) { return hash(st, m, g_small); },
                    8 * dn0 + 64.0 * g_small, 8 * dn + 64.0 * g_small, n0, n, &c[QO_COST_HASH_SMALL]));
    // The same with the filter fused (.y below half the key range: half selected).
    auto fhash = [&](hipStream_t st, size_t m, size_t cap, int32_t hi) {
        return qo_groupby_agg_filtered(cy.as<int32_t>(), cx.as<int32_t>(), cy.as<int32_t>(), m, 0, hi, cap,
                                       okey.as<int32_t>(), ocnt.as<int64_t>(), osum.as<int64_t>(),
                                       omin.as<int32_t>(), omax.as<int32_t>(), cnt.as<uint64_t>(), st);
    };
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) { return fhash(st, m, g_small, g_small / 2 - 1); },
                    8 * dn0 + 64.0 * g_small, 8 * dn + 64.0 * g_small, n0, n, &c[QO_COST_FHASH_SMALL]));
    QO_TRY(qo_unpack_records(rec2.as<qo_record>(), n, cx.as<uint32_t>(), cy.as<uint32_t>(), tm.st));
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) { return hash(st, m, g_large); },
                    8 * dn0 + 64.0 * std::min<double>(dn0, g_large), 8 * dn + 64.0 * g_large, n0, n,
                    &c[QO_COST_HASH_LARGE]));
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) {
        return fhash(st, m, g_large, (int32_t)(g_large / 2 - 1)); },
        8 * dn0 + 64.0 * std::min<double>(dn0, g_large), 8 * dn + 64.0 * g_large, n0, n, &c[QO_COST_FHASH_LARGE]));
    // Reduce.
    QO_TRY(fit_cost(tm, reps, [&](hipStream_t st, size_t m) {
        return qo_reduce_records(R, m, 0, cnt.as<uint32_t>() + 8, st); }, 8 * dn0, 8 * dn, n0, n,
        &c[QO_COST_REDUCE]));
    std::lock_guard<std::mutex> lk(g_cost_mu);
    std::copy(c, c + QO_COST_NUM, g_cost);
    return QO_OK;
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_plan_calibrate(int size_class, int reps) {
    QO_ENTER();
    QO_REQUIRE(size_class >= 16 && size_class <= 28, "size class %d outside [16, 28]", size_class);
    QO_REQUIRE(reps >= 1 && reps <= 50, "reps %d outside [1, 50]", reps);
    return calibrate_impl(size_class, reps);
}

int qo_plan_query(const qo_query *q, size_t n, unsigned allowed_impls, qo_plan *plan) {
    QO_REQUIRE(q, "null query");
    return plan_impl(*q, n, allowed_impls, plan);
}

int qo_plan_cost_set(int prim, double fixed_us, double gbps) {
    QO_REQUIRE(prim >= 0 && prim < QO_COST_NUM, "cost primitive %d outside [0, %d)", prim, QO_COST_NUM);
    QO_REQUIRE(fixed_us >= 0.0 && gbps > 0.0, "cost needs fixed_us >= 0 and gbps > 0");
    std::lock_guard<std::mutex> lk(g_cost_mu);
    g_cost[prim] = PrimCost{fixed_us, gbps};
    return QO_OK;
}

int qo_plan_cost_get(int prim, double *fixed_us, double *gbps) {
    QO_REQUIRE(prim >= 0 && prim < QO_COST_NUM, "cost primitive %d outside [0, %d)", prim, QO_COST_NUM);
    QO_REQUIRE(fixed_us && gbps, "null output");
    std::lock_guard<std::mutex> lk(g_cost_mu);
    *fixed_us = g_cost[prim].fixed_us;
    *gbps = g_cost[prim].gbps;
    return QO_OK;
}

int qo_plan_execute(const qo_plan *plan, const qo_query *q, const qo_record *rin, size_t n, size_t group_capacity,
                    qo_record *d_sel, uint32_t *d_keys, int32_t *d_aggs, uint64_t *d_count, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(plan && q && d_count, "null argument");
    QO_REQUIRE(plan->nseg >= 1 && plan->nseg <= QO_PLAN_MAX_SEGS, "bad plan");
    QO_REQUIRE(n == 0 || (rin && (uintptr_t)rin % 8 == 0), "Records must be 8-byte aligned device memory");
    const Chain c = chain_of(*q);
    QO_REQUIRE(plan->last_step[plan->nseg - 1] == c.n - 1 && plan->first_step[0] == 0,
               "the plan does not cover this query's steps");
    const bool select_only = !q->group && q->op < 0;
    QO_REQUIRE(!select_only || d_sel, "a selection needs d_sel (n Records)");
    QO_REQUIRE(select_only || d_aggs, "an aggregate needs d_aggs");
    QO_REQUIRE(!q->group || (d_keys && group_capacity > 0), "a grouped query needs d_keys and a capacity");
    hipStream_t st = stream_of(s);
    Temps tmp;
    int status = QO_OK;
    tmp.cap_bytes = plan_temp_bytes(*plan, n, group_capacity);
    tmp.cv.base = static_cast<char *>(workspace_slot(1, st, tmp.cap_bytes, &status));
    if (!tmp.cv.base) return status;
    const qo_record *cur = rin;  // the rows entering the next segment
    size_t m = n;
    bool owned = false;          // cur is a temporary of this execution
    uint64_t *d_m = tmp.get<uint64_t>(1, &status);
    if (!d_m) return status;
    for (int sidx = 0; sidx < plan->nseg; ++sidx) {
        const int impl = plan->impl[sidx];
        QO_REQUIRE(covers(impl, c, plan->first_step[sidx]) == plan->last_step[sidx] - plan->first_step[sidx] + 1,
                   "segment %d: implementation %d does not cover its steps", sidx, impl);
        switch (impl) {
            case QO_IMPL_FILTER: {
                const bool last = sidx == plan->nseg - 1;
                qo_record *out = last ? d_sel : tmp.get<qo_record>(m, &status);
                if (!out) return status;
                QO_TRY(qo_filter_records(cur, m, q->lo, q->hi, out, last ? d_count : d_m, s));
                uint64_t h = 0;
                QO_TRY(read_back(&h, last ? d_count : d_m, 8, st));  // m sizes what follows
                cur = out;
                m = (size_t)h;
                owned = true;
                break;
            }
            case QO_IMPL_SORT_GROUP_AGG: {
                qo_record *work = const_cast<qo_record *>(cur);
                if (!owned) {
                    work = tmp.get<qo_record>(m, &status);
                    if (!work) return status;
                    if (m) QO_TRY(qo_copy(work, cur, m * 8, s));
                }
                int32_t *sp = tmp.get<int32_t>(m, &status);
                if (!sp) return status;
                uint64_t G = 0;
                if (m) {
                    QO_TRY(qo_sort_records(work, m, s));
                    QO_TRY(qo_group_startpos(work, m, sp, d_m, s));
                    QO_TRY(read_back(&G, d_m, 8, st));
                }
                QO_TRY(launch_fill_u64(d_count, G, 1, st));
                QO_REQUIRE(G <= group_capacity, "%llu groups exceed the capacity %zu", (unsigned long long)G,
                           group_capacity);
                if (G) {
                    QO_TRY(qo_aggregate_groups(work, m, sp, G, q->op, d_aggs, s));
                    k_plan_keys_at<<<grid_for(G, 256, 1u << 20), 256, 0, st>>>(work, sp, G, d_keys);
                    QO_LAUNCHED();
                }
                break;
            }
            case QO_IMPL_HASH_GROUP_AGG:
            case QO_IMPL_FILTER_HASH_GROUP_AGG: {
                uint32_t *kx = tmp.get<uint32_t>(m, &status);
                uint32_t *ky = kx ? tmp.get<uint32_t>(m, &status) : nullptr;
                if (!ky) return status;
                const size_t cap = group_capacity;
                int32_t *okey = tmp.get<int32_t>(cap, &status);
                int64_t *ocnt = okey ? tmp.get<int64_t>(cap, &status) : nullptr;
                int64_t *osum = ocnt ? tmp.get<int64_t>(cap, &status) : nullptr;
                int32_t *omin = osum ? tmp.get<int32_t>(cap, &status) : nullptr;
                int32_t *omax = omin ? tmp.get<int32_t>(cap, &status) : nullptr;
                if (!omax) return status;
                uint64_t G = 0;
                if (m) {
                    QO_TRY(qo_unpack_records(cur, m, kx, ky, s));
                    const int32_t *key = reinterpret_cast<const int32_t *>(ky);
                    const int32_t *val = reinterpret_cast<const int32_t *>(kx);
                    if (impl == QO_IMPL_FILTER_HASH_GROUP_AGG)
                        QO_TRY(qo_groupby_agg_filtered(key, val, key, m, q->lo, q->hi, cap, okey, ocnt, osum, omin,
                                                       omax, d_m, s));
                    else
                        QO_TRY(qo_groupby_agg(key, val, m, cap, okey, ocnt, osum, omin, omax, d_m, s));
                    QO_TRY(read_back(&G, d_m, 8, st));
                }
                QO_TRY(launch_fill_u64(d_count, G, 1, st));
                QO_REQUIRE(G <= cap, "%llu groups exceed the capacity %zu", (unsigned long long)G, cap);
                if (G) {
                    k_plan_epilogue<<<grid_for(G, 256, 1u << 20), 256, 0, st>>>(okey, ocnt, osum, omin, omax, d_m,
                                                                                q->op, d_keys, d_aggs);
                    QO_LAUNCHED();
                }
                break;
            }
            case QO_IMPL_REDUCE: {
                QO_TRY(qo_reduce_records(cur, m, q->op, reinterpret_cast<uint32_t *>(d_aggs), s));
                QO_TRY(launch_fill_u64(d_count, m, 1, st));
                break;
            }
            default: return set_error(QO_ERR_INVALID_ARG, "segment %d: unknown implementation %d", sidx, impl);
        }
    }
    return sync_and_check(st);
}

}  // extern "C"
