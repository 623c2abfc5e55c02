/*
 * qo_oracle.c -- CPU restatement of the reference's relational operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library (libqo.so) never links,
 * loads or calls anything here.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the Xtra-Computing/Query_on_OpenCL_FPGA tree).  It is plain single-threaded
 * C99 so that it is easy to audit.
 *
 * Pinning (what checks each function independently of this file):
 *   or_rand64 / or_gen_*          : published splitmix64 test vector; GPU generator
 *                                   equality; uniformity and bijectivity properties.
 *   or_scan_excl_i32              : closed forms (all-ones -> iota, wrap at 2^32),
 *                                   hand-made golden vectors.
 *   or_filter_records             : closed form (count, order-preserved subsequence),
 *                                   hand-made golden vectors, brute force.
 *   or_filter_agg*                : closed form COUNT/SUM, hand-made golden vectors.
 *   or_sort_records               : invariants (non-decreasing .y, multiset equal,
 *                                   stable) and golden vectors.
 *   or_group_startpos             : closed form from sorted keys, golden vectors.
 *   or_parallel_aggregate         : hand-computed quirk vectors (MAX init 0,
 *                                   MIN init 2^30, int32 wrap, AVG truncation).
 *   or_groupby_agg                : brute force dictionary in Python, golden vectors.
 *   or_hash_join                  : closed form for permutation keys (one match per
 *                                   selected S row), brute-force nested loop.
 *   or_sort_join_rows             : invariants + golden vectors.
 *   or_split_records              : validateSplit's condition, brute force (pure-Python
 *                                   stable bucketing), bounds = closed form from counts.
 *   or_gather_records /
 *   or_scatter_records            : brute force; gather o scatter = identity on a
 *                                   permutation; numpy fancy indexing.
 *   or_project_records            : validateProjection's rule restated in numpy.
 *   or_reduce_records             : closed forms (sum of 0..n-1 mod 2^32, max/min of
 *                                   a permutation), hand-computed unsigned quirks.
 *   or_unpack_records             : numpy column view of the interleaved array.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct { uint32_t x, y; } or_record; /* reference Record = uint2 {x, y}: host/host/common.h:27-28 */

/* ------------------------------------------------------------------------- */
/* Deterministic counter-based generator (this project's own; shared spec with  */
/* the GPU generator in libqo).  splitmix64 is Steele/Lea/Flood (2014).        */
/* ------------------------------------------------------------------------- */
#define OR_GAMMA 0x9E3779B97F4A7C15ull

uint64_t or_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* splitmix64 stream keyed by (seed, stream) evaluated at position i. */
uint64_t or_rand64(uint64_t seed, uint64_t stream, uint64_t i)
{
    uint64_t key = or_mix64(seed + OR_GAMMA * (stream + 1));
    return or_mix64(key + OR_GAMMA * (i + 1));
}

static inline uint32_t or_bounded(uint64_t r, uint32_t range)
{
    return (uint32_t)(((r >> 32) * (uint64_t)range) >> 32);
}

void or_gen_uniform_u32(uint32_t *out, size_t n, uint64_t seed, uint64_t stream,
                        uint32_t lo, uint32_t range)
{
    for (size_t i = 0; i < n; ++i)
        out[i] = lo + (range ? or_bounded(or_rand64(seed, stream, i), range)
                             : (uint32_t)(or_rand64(seed, stream, i) >> 32));
}

/* Keyed bijection on [0, 2^bits): rounds of odd multiply, xorshift, add. */
uint32_t or_bij(uint32_t x, uint32_t bits, uint64_t seed)
{
    uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
    uint64_t v = x;
    uint32_t sh = bits / 2 + 1;
    for (int r = 0; r < 3; ++r) {
        uint64_t m = or_mix64(seed + 2 * r + 1) | 1ull;
        uint64_t c = or_mix64(seed + 2 * r + 2);
        v = (v * m) & mask;
        v ^= v >> sh;
        v = (v + c) & mask;
    }
    return (uint32_t)v;
}

static uint32_t or_bits_for(uint64_t n)
{
    uint32_t b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

/* Element i of a seeded permutation of [0, n) (cycle walking on the bijection). */
uint32_t or_perm(uint64_t i, uint64_t n, uint64_t seed)
{
    uint32_t bits = or_bits_for(n);
    uint32_t y = or_bij((uint32_t)i, bits, seed);
    while (y >= n) y = or_bij(y, bits, seed);
    return y;
}

void or_gen_perm(uint32_t *out, size_t n, uint64_t seed)
{
    for (size_t i = 0; i < n; ++i) out[i] = or_perm(i, n, seed);
}

/* Fixed-point Zipf CDF: t[r] = floor(2^53 * sum_{j<=r} (j+1)^-z / H); t[G-1] = 2^53. */
void or_zipf_table(uint64_t *t, uint32_t G, double z)
{
    double h = 0.0;
    for (uint32_t r = 0; r < G; ++r) h += pow((double)r + 1.0, -z);
    double c = 0.0;
    for (uint32_t r = 0; r < G; ++r) {
        c += pow((double)r + 1.0, -z);
        double f = c / h;
        if (f > 1.0) f = 1.0;
        t[r] = (uint64_t)(f * 9007199254740992.0);
    }
    t[G - 1] = 9007199254740992ull;
}

/* rank = first r with u < t[r], u = 53 random bits; key = perm_G(rank). */
void or_gen_zipf_keys(uint32_t *out, size_t n, const uint64_t *t, uint32_t G,
                      uint64_t seed, uint64_t stream, uint64_t perm_seed)
{
    for (size_t i = 0; i < n; ++i) {
        uint64_t u = or_rand64(seed, stream, i) >> 11;
        uint32_t lo = 0, hi = G - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (u < t[mid]) hi = mid; else lo = mid + 1;
        }
        out[i] = or_perm(lo, G, perm_seed);
    }
}

/* ------------------------------------------------------------------------- */
/* Scan: kernel/filter/filter_op.cl:51-207 (ScanLargeArrays_kernel,            */
/* prefixSum_kernel, blockAddition_kernel) compose to an exclusive prefix sum   */
/* in int32 arithmetic; host/host/Validate.cpp:4-37 states the same closed form.*/
/* Overflow wraps modulo 2^32 (computed in uint32 to keep C well defined).      */
/* ------------------------------------------------------------------------- */
void or_scan_excl_i32(const int32_t *in, int32_t *out, size_t n)
{
    uint32_t acc = 0;
    for (size_t i = 0; i < n; ++i) {
        uint32_t v = (uint32_t)in[i];
        out[i] = (int32_t)acc;
        acc += v;
    }
}

/* ------------------------------------------------------------------------- */
/* Filter: map (filter_op.cl:6-31: mark = smallKey <= (int)R.y <= largeKey),   */
/* exclusive scan of marks, outSize = mark[n-1] + scan[n-1] (filter_op.cl:     */
/* 211-215), stable scatter out[scan[i]] = R[i] (filter_op.cl:218-238).         */
/* beginPos is unused by both kernels, so it is not a parameter here.           */
/* ------------------------------------------------------------------------- */
uint64_t or_filter_records(const or_record *in, size_t n, int32_t lo, int32_t hi,
                           or_record *out)
{
    if (n == 0) return 0;
    int32_t *mark = (int32_t *)malloc(n * sizeof(int32_t));
    int32_t *pos = (int32_t *)malloc(n * sizeof(int32_t));
    for (size_t i = 0; i < n; ++i) {
        int32_t key = (int32_t)in[i].y;
        mark[i] = (key >= lo && key <= hi) ? 1 : 0;
    }
    or_scan_excl_i32(mark, pos, n);
    uint64_t outSize = (uint64_t)(uint32_t)(mark[n - 1] + pos[n - 1]);
    if (out)
        for (size_t i = 0; i < n; ++i)
            if (mark[i]) out[(uint32_t)pos[i]] = in[i];
    free(mark);
    free(pos);
    return outSize;
}

/* Fused filter -> aggregate over one group (the chain testAggAfterGB.cpp:313-427
 * intends): COUNT = outSize, SUM over .x of the selected rows as exact int64,
 * and the reference's int32 REDUCE_SUM (group_by_cu_op_dp.cl:547-716 with one
 * group, wraps mod 2^32).  out = {count, sum64, (int64)sum32}. */
void or_filter_agg_records(const or_record *in, size_t n, int32_t lo, int32_t hi,
                           int64_t out[3])
{
    uint64_t cnt = 0;
    int64_t sum = 0;
    uint32_t sum32 = 0;
    for (size_t i = 0; i < n; ++i) {
        int32_t key = (int32_t)in[i].y;
        if (key >= lo && key <= hi) {
            ++cnt;
            sum += (int64_t)(int32_t)in[i].x;
            sum32 += in[i].x;
        }
    }
    out[0] = (int64_t)cnt;
    out[1] = sum;
    out[2] = (int64_t)(int32_t)sum32;
}

/* Same over two int32 columns (key column = Record.y, value column = Record.x). */
void or_filter_agg(const int32_t *key, const int32_t *val, size_t n, int32_t lo, int32_t hi,
                   int64_t out[3])
{
    uint64_t cnt = 0;
    int64_t sum = 0;
    uint32_t sum32 = 0;
    for (size_t i = 0; i < n; ++i) {
        if (key[i] >= lo && key[i] <= hi) {
            ++cnt;
            sum += val[i];
            sum32 += (uint32_t)val[i];
        }
    }
    out[0] = (int64_t)cnt;
    out[1] = sum;
    out[2] = (int64_t)(int32_t)sum32;
}

/* ------------------------------------------------------------------------- */
/* Stable LSD radix sort helper: sorts 64-bit keys, carrying a 32-bit index.    */
/* ------------------------------------------------------------------------- */
static void or_radix_u64(uint64_t *key, uint32_t *idx, size_t n, int lo_bit, int hi_bit)
{
    uint64_t *k2 = (uint64_t *)malloc(n * sizeof(uint64_t) + 8);
    uint32_t *i2 = (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    size_t cnt[256];
    for (int sh = lo_bit; sh < hi_bit; sh += 8) {
        memset(cnt, 0, sizeof(cnt));
        for (size_t i = 0; i < n; ++i) cnt[(key[i] >> sh) & 255]++;
        size_t s = 0;
        for (int d = 0; d < 256; ++d) { size_t c = cnt[d]; cnt[d] = s; s += c; }
        for (size_t i = 0; i < n; ++i) {
            size_t p = cnt[(key[i] >> sh) & 255]++;
            k2[p] = key[i];
            i2[p] = idx[i];
        }
        memcpy(key, k2, n * sizeof(uint64_t));
        memcpy(idx, i2, n * sizeof(uint32_t));
    }
    free(k2);
    free(i2);
}

/* ------------------------------------------------------------------------- */
/* Sort: radixSortImpl_sm (host/host/testSort.cpp:24-136) drives the bitonic    */
/* network of group_by_cu_op_dp.cl:89-291, ascending by .y compared as uint     */
/* (uint2 compare at :132).  The reference validator (Validate.cpp:470-487)     */
/* checks only non-decreasing .y; the network's tie order is not a contract.    */
/* This restatement fixes the tie rule as STABLE (equal .y keep input order).   */
/* ------------------------------------------------------------------------- */
void or_sort_records(or_record *a, size_t n)
{
    if (n < 2) return;
    uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *ix = (uint32_t *)malloc(n * sizeof(uint32_t));
    or_record *tmp = (or_record *)malloc(n * sizeof(or_record));
    for (size_t i = 0; i < n; ++i) { k[i] = a[i].y; ix[i] = (uint32_t)i; }
    or_radix_u64(k, ix, n, 0, 32);
    for (size_t i = 0; i < n; ++i) tmp[i] = a[ix[i]];
    memcpy(a, tmp, n * sizeof(or_record));
    free(k); free(ix); free(tmp);
}

/* ------------------------------------------------------------------------- */
/* Group boundaries: scanGroupLabel_kernel (group_by_cu_op_dp.cl:293-327):      */
/* label[0]=1, label[i+1] = (R[i].y != R[i+1].y); exclusive scan; numGroups =   */
/* label[n-1]+scan[n-1] (:505-509); startPos[scan[i]] = i where label[i]==1     */
/* (:511-528).  Restated by validateGroupBy (Validate.cpp:257-316).             */
/* ------------------------------------------------------------------------- */
uint32_t or_group_startpos(const or_record *s, size_t n, int32_t *startPos)
{
    if (n == 0) return 0;
    int32_t *label = (int32_t *)malloc(n * sizeof(int32_t));
    int32_t *pos = (int32_t *)malloc(n * sizeof(int32_t));
    label[0] = 1;
    for (size_t i = 0; i + 1 < n; ++i) label[i + 1] = (s[i].y != s[i + 1].y) ? 1 : 0;
    or_scan_excl_i32(label, pos, n);
    uint32_t numGroups = (uint32_t)(label[n - 1] + pos[n - 1]);
    if (startPos)
        for (size_t i = 0; i < n; ++i)
            if (label[i] == 1) startPos[pos[i]] = (int32_t)i;
    free(label);
    free(pos);
    return numGroups;
}

/* ------------------------------------------------------------------------- */
/* parallelAggregate_kernel (group_by_cu_op_dp.cl:547-716): group idx covers    */
/* [startPos[idx], idx==G-1 ? rLen : startPos[idx+1]) and reduces (int).x.     */
/* REDUCE_SUM=0 int32 (wraps), REDUCE_MAX=1 starting from 0, REDUCE_MIN=2       */
/* starting from TEST_MAX = 1<<30, REDUCE_AVERAGE=3 = int32 sum / (end-start)   */
/* with C truncation.  The per-lane partials and tree reduction combine with    */
/* the same operator and the same identity, so the result equals one           */
/* sequential fold from the identity.                                          */
/* ------------------------------------------------------------------------- */
void or_parallel_aggregate(const or_record *s, size_t n, const int32_t *startPos, uint32_t G,
                           int op, int32_t *out)
{
    for (uint32_t g = 0; g < G; ++g) {
        int64_t start = startPos[g];
        int64_t end = (g == G - 1) ? (int64_t)n : startPos[g + 1];
        int32_t acc = (op == 2) ? (1 << 30) : 0;
        uint32_t usum = 0;
        for (int64_t i = start; i < end; ++i) {
            int32_t v = (int32_t)s[i].x;
            if (op == 1) { if (v > acc) acc = v; }
            else if (op == 2) { if (v < acc) acc = v; }
            else usum += (uint32_t)v;
        }
        if (op == 0) acc = (int32_t)usum;
        if (op == 3) {
            int32_t sum32 = (int32_t)usum;
            int64_t cnt = end - start;
            acc = cnt > 0 ? (int32_t)(sum32 / (int32_t)cnt) : 0;
        }
        out[g] = acc;
    }
}

/* ------------------------------------------------------------------------- */
/* Group-by aggregation with exact outputs: the sort -> label -> startPos ->    */
/* parallelAggregate pipeline above, keyed by a signed int32 column, producing  */
/* per group (key, COUNT int64, SUM int64, MIN int32, MAX int32) ascending by   */
/* key.  COUNT = end - start as in the reference; SUM widened to int64; MIN and */
/* MAX start from the int32 extremes (no 0 / 2^30 clamp).  Returns G.          */
/* ------------------------------------------------------------------------- */
size_t or_groupby_agg(const int32_t *key, const int32_t *val, size_t n,
                      int32_t *okey, int64_t *ocount, int64_t *osum, int32_t *omin, int32_t *omax)
{
    if (n == 0) return 0;
    uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *ix = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) { k[i] = (uint32_t)key[i] ^ 0x80000000u; ix[i] = (uint32_t)i; }
    or_radix_u64(k, ix, n, 0, 32);
    size_t g = 0;
    size_t i = 0;
    while (i < n) {
        size_t j = i;
        int64_t sum = 0;
        int32_t mn = INT32_MAX, mx = INT32_MIN;
        while (j < n && k[j] == k[i]) {
            int32_t v = val[ix[j]];
            sum += v;
            if (v < mn) mn = v;
            if (v > mx) mx = v;
            ++j;
        }
        okey[g] = (int32_t)((uint32_t)k[i] ^ 0x80000000u);
        ocount[g] = (int64_t)(j - i);
        osum[g] = sum;
        omin[g] = mn;
        omax[g] = mx;
        ++g;
        i = j;
    }
    free(k); free(ix);
    return g;
}

/* ------------------------------------------------------------------------- */
/* Hash join: buildHashTable (kernel/hj_filter_sort/1/hj_filter_sort_op.cl:     */
/* 25-60) puts each R tuple into bucket key & (B-1) of a bucket-chained table;  */
/* probeHashTable (:62-109) scans the bucket of each S key and emits every     */
/* equal-key R entry.  The S filter is the filter kernels of the same image     */
/* (:117-349) applied to the filter column (lo <= f <= hi).  This restatement   */
/* chains buckets in R index order (the reference's order depends on lock      */
/* order) and emits matches in S order, then chain order.  Output rows are      */
/* (key, r_payload, s_payload).  Returns the match count; writes at most cap.  */
/* ------------------------------------------------------------------------- */
size_t or_hash_join(const int32_t *rkey, const int32_t *rpay, size_t nr,
                    const int32_t *skey, const int32_t *spay, const int32_t *sfilt, size_t ns,
                    int32_t flo, int32_t fhi,
                    int32_t *okey, int32_t *orpay, int32_t *ospay, size_t cap)
{
    uint32_t bits = or_bits_for(nr ? nr : 1);
    size_t B = (size_t)1 << bits;
    int64_t *head = (int64_t *)malloc(B * sizeof(int64_t));
    int64_t *tail = (int64_t *)malloc(B * sizeof(int64_t));
    int64_t *next = (int64_t *)malloc((nr ? nr : 1) * sizeof(int64_t));
    for (size_t b = 0; b < B; ++b) { head[b] = -1; tail[b] = -1; }
    for (size_t i = 0; i < nr; ++i) {
        size_t h = (uint32_t)rkey[i] & (B - 1);
        next[i] = -1;
        if (tail[h] < 0) head[h] = (int64_t)i; else next[tail[h]] = (int64_t)i;
        tail[h] = (int64_t)i;
    }
    size_t m = 0;
    for (size_t j = 0; j < ns; ++j) {
        if (sfilt && !(sfilt[j] >= flo && sfilt[j] <= fhi)) continue;
        size_t h = (uint32_t)skey[j] & (B - 1);
        for (int64_t e = head[h]; e >= 0; e = next[e]) {
            if (rkey[e] == skey[j]) {
                if (m < cap) { okey[m] = skey[j]; orpay[m] = rpay[e]; ospay[m] = spay[j]; }
                ++m;
            }
        }
    }
    free(head); free(tail); free(next);
    return m;
}

/* ------------------------------------------------------------------------- */
/* Sort of join output (the hj_filter_sort image's sort kernels,               */
/* hj_filter_sort_op.cl:362-564, applied to the probe output): ascending by     */
/* (key, s_payload) as signed int32; rows equal in both keep their input order. */
/* ------------------------------------------------------------------------- */
void or_sort_join_rows(int32_t *key, int32_t *rpay, int32_t *spay, size_t m)
{
    if (m < 2) return;
    uint64_t *k = (uint64_t *)malloc(m * sizeof(uint64_t));
    uint32_t *ix = (uint32_t *)malloc(m * sizeof(uint32_t));
    int32_t *t = (int32_t *)malloc(m * sizeof(int32_t));
    for (size_t i = 0; i < m; ++i) {
        k[i] = ((uint64_t)((uint32_t)key[i] ^ 0x80000000u) << 32) | ((uint32_t)spay[i] ^ 0x80000000u);
        ix[i] = (uint32_t)i;
    }
    or_radix_u64(k, ix, m, 0, 64);
    for (size_t i = 0; i < m; ++i) t[i] = key[ix[i]];
    memcpy(key, t, m * sizeof(int32_t));
    for (size_t i = 0; i < m; ++i) t[i] = rpay[ix[i]];
    memcpy(rpay, t, m * sizeof(int32_t));
    for (size_t i = 0; i < m; ++i) t[i] = spay[ix[i]];
    memcpy(spay, t, m * sizeof(int32_t));
    free(k); free(ix); free(t);
}

/* Bucket-chained probe restated with interleaved inputs, matching the
 * reference kernel signatures (rTable / sTable as (key, val) uint pairs,
 * hj_filter_sort_op.cl:40-41,86-87) and emitting compacted (s_val, r_val)
 * pairs (:102-103).  Returns the number of pairs. */
size_t or_probe_pairs(const uint32_t *rTable, size_t nr, const uint32_t *sTable, size_t ns,
                      uint32_t *matched, size_t cap_pairs)
{
    int32_t *rk = (int32_t *)malloc((nr + 1) * sizeof(int32_t));
    int32_t *rv = (int32_t *)malloc((nr + 1) * sizeof(int32_t));
    int32_t *sk = (int32_t *)malloc((ns + 1) * sizeof(int32_t));
    int32_t *sv = (int32_t *)malloc((ns + 1) * sizeof(int32_t));
    for (size_t i = 0; i < nr; ++i) { rk[i] = (int32_t)rTable[2 * i]; rv[i] = (int32_t)rTable[2 * i + 1]; }
    for (size_t i = 0; i < ns; ++i) { sk[i] = (int32_t)sTable[2 * i]; sv[i] = (int32_t)sTable[2 * i + 1]; }
    size_t m = or_hash_join(rk, rv, nr, sk, sv, NULL, ns, 0, 0, NULL, NULL, NULL, 0);
    int32_t *ok = (int32_t *)malloc((m + 1) * sizeof(int32_t));
    int32_t *orr = (int32_t *)malloc((m + 1) * sizeof(int32_t));
    int32_t *os = (int32_t *)malloc((m + 1) * sizeof(int32_t));
    or_hash_join(rk, rv, nr, sk, sv, NULL, ns, 0, 0, ok, orr, os, m);
    for (size_t i = 0; i < m && i < cap_pairs; ++i) {
        matched[2 * i] = (uint32_t)os[i];
        matched[2 * i + 1] = (uint32_t)orr[i];
    }
    free(rk); free(rv); free(sk); free(sv); free(ok); free(orr); free(os);
    return m;
}

/* ------------------------------------------------------------------------- */
/* Split, scatter, gather, projection, reduce (the operators around the hot   */
/* path; their reference kernels are absent, so each follows the host driver */
/* and the validator that fixes its result).                                  */
/* ------------------------------------------------------------------------- */

/* Split into numPart partitions by pid = R.y % numPart (Record is uint2,
 * host/host/common.h:26, so the remainder is unsigned): splitImpl
 * (host/host/testSplit.cpp:156-209: map pid, per-partition counts, scan,
 * write, bounds, scatter), checked by validateSplit (host/host/Validate.cpp:
 * 510-527: pid non-decreasing along the output).  Records keep input order
 * within a partition (a counting sort).  bound[p] = {start, end} of partition
 * p in the output (getBound_int, testSplit.cpp:121-140: psSum at the
 * partition's first slot, the next partition's, or rLen for the last). */
void or_split_records(const or_record *in, size_t n, uint32_t numPart, or_record *out, or_record *bound)
{
    size_t *start = (size_t *)calloc((size_t)numPart + 1, sizeof(size_t));
    for (size_t i = 0; i < n; ++i) start[in[i].y % numPart + 1]++;
    for (uint32_t p = 0; p < numPart; ++p) start[p + 1] += start[p];
    if (bound)
        for (uint32_t p = 0; p < numPart; ++p) {
            bound[p].x = (uint32_t)start[p];
            bound[p].y = (uint32_t)start[p + 1];
        }
    for (size_t i = 0; i < n; ++i) out[start[in[i].y % numPart]++] = in[i];
    free(start);
}

/* S[pos] = R[loc[pos]] for pos < sLen (optGather: the windows [from, to) of
 * gatherImpl, host/host/testScatter.cpp:105-130, cover [0, rLen) between
 * them).  A loc outside [0, rLen) leaves S[pos] as it was. */
void or_gather_records(const or_record *R, size_t rLen, const int32_t *loc, size_t sLen, or_record *S)
{
    for (size_t pos = 0; pos < sLen; ++pos)
        if (loc[pos] >= 0 && (size_t)loc[pos] < rLen) S[pos] = R[loc[pos]];
}

/* S[loc[i]] = R[i] for i < rLen (optScatter over the windows of scatterImpl,
 * host/host/testScatter.cpp:61-81; S holds rLen records).  A loc outside
 * [0, rLen) writes nothing.  With repeated targets the reference's order is
 * unspecified; this restatement lets the last i win. */
void or_scatter_records(const or_record *R, size_t rLen, const int32_t *loc, or_record *S)
{
    for (size_t i = 0; i < rLen; ++i)
        if (loc[i] >= 0 && (size_t)loc[i] < rLen) S[loc[i]] = R[i];
}

/* Projection checked by validateProjection (host/host/Validate.cpp:39-71):
 * out[i] = {proj[i].x, R[proj[i].x].y}.  A row id beyond rLen keeps the
 * projection record's own .y. */
void or_project_records(const or_record *R, size_t rLen, const or_record *proj, size_t pLen, or_record *out)
{
    for (size_t i = 0; i < pLen; ++i) {
        out[i].x = proj[i].x;
        out[i].y = proj[i].x < rLen ? R[proj[i].x].y : proj[i].y;
    }
}

/* Reduce without grouping over R.y, as validateReduce computes it
 * (host/host/Validate.cpp:372-468; operator codes common.h:89-92):
 * SUM (0): unsigned sum mod 2^32; MAX (1) / MIN (2): from element 0, compared
 * as unsigned (the int accumulator meets the uint .y in an unsigned
 * comparison); AVERAGE (3): the unsigned sum mod 2^32 divided by rLen.
 * n == 0 gives 0. */
uint32_t or_reduce_records(const or_record *R, size_t n, int op)
{
    if (n == 0) return 0;
    uint32_t acc = op == 1 || op == 2 ? R[0].y : 0u;
    for (size_t i = op == 1 || op == 2 ? 1 : 0; i < n; ++i) {
        const uint32_t v = R[i].y;
        if (op == 1) acc = v > acc ? v : acc;
        else if (op == 2) acc = v < acc ? v : acc;
        else acc += v;
    }
    return op == 3 ? (uint32_t)(acc / (uint64_t)n) : acc;
}

/* Columns of Records: x[i] = R[i].x, y[i] = R[i].y (the AoS -> SoA step that
 * feeds the column operators; Record layout host/host/common.h:26). */
void or_unpack_records(const or_record *R, size_t n, uint32_t *x, uint32_t *y)
{
    for (size_t i = 0; i < n; ++i) {
        x[i] = R[i].x;
        y[i] = R[i].y;
    }
}
