/*
 * qo_cpu_par.c -- multi-threaded CPU restatement of the reference's operator
 * algorithms, used as the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY.  Only bench.py's cpu_baseline leg and tests/ (which
 * check it against the single-threaded oracle in qo_oracle.c) load this
 * library; the product library never links, loads or calls it.
 *
 * The algorithms are the reference's, split over host threads the way an
 * OpenCL CPU runtime splits work-groups over cores (paths relative to the
 * Xtra-Computing/Query_on_OpenCL_FPGA tree):
 *   pr_filter_agg    : filter map + aggregate, filter_aggregation_cu_op_dp.cl:82-107
 *                      and parallelAggregate (:609-778) for one group: per-thread
 *                      partial COUNT/SUM, then a reduction.
 *   pr_hash_join     : buildHashTable / probeHashTable, hj_filter_sort_op.cl:25-109:
 *                      bucket = key & (B-1) with B = 2^ceil(log2 |R|), chained
 *                      buckets built concurrently (lock-free push), filtered S
 *                      probed in per-thread ranges; output (key, r_payload,
 *                      s_payload), per thread in S order.
 *   pr_groupby_agg   : sort -> group label -> start positions -> per-group
 *                      aggregate (group_by_cu_op_dp.cl:89-716): a parallel
 *                      stable LSD radix sort on the key replaces the bitonic
 *                      network (the same sorted order up to ties), then groups
 *                      are reduced in per-thread ranges cut at group starts.
 *   pr_sort_join_rows: the hj_filter_sort image's sort (hj_filter_sort_op.cl:362-564)
 *                      as a parallel LSD radix sort by (key, s_payload).
 * Results equal the single-threaded oracle's (qo_oracle.c); join output order
 * differs (per-thread S ranges concatenated in order, chain order may differ).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define PR_MAX_THREADS 256

int pr_threads(int t)
{
    if (t > 0) return t > PR_MAX_THREADS ? PR_MAX_THREADS : t;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1) n = 1;
    return n > PR_MAX_THREADS ? PR_MAX_THREADS : (int)n;
}

typedef void (*pr_fn)(void *arg, int tid, int nt);
typedef struct { pr_fn fn; void *arg; int tid, nt; } pr_job;

static void *pr_tramp(void *p)
{
    pr_job *j = (pr_job *)p;
    j->fn(j->arg, j->tid, j->nt);
    return NULL;
}

/* Runs fn(arg, tid, nt) on nt threads (tid 0 on the caller) and joins them. */
static void pr_run(pr_fn fn, void *arg, int nt)
{
    pthread_t th[PR_MAX_THREADS];
    pr_job jobs[PR_MAX_THREADS];
    for (int t = 0; t < nt; ++t) { jobs[t].fn = fn; jobs[t].arg = arg; jobs[t].tid = t; jobs[t].nt = nt; }
    for (int t = 1; t < nt; ++t) pthread_create(&th[t], NULL, pr_tramp, &jobs[t]);
    fn(arg, 0, nt);
    for (int t = 1; t < nt; ++t) pthread_join(th[t], NULL);
}

static size_t pr_lo(size_t n, int t, int nt) { return n * (size_t)t / (size_t)nt; }

/* ---------------------------------------------------------------- filter_agg */
typedef struct {
    const int32_t *key, *val;
    size_t n;
    int32_t lo, hi;
    uint64_t cnt[PR_MAX_THREADS];
    int64_t sum[PR_MAX_THREADS];
    uint32_t sum32[PR_MAX_THREADS];
} fa_arg;

static void fa_body(void *p, int t, int nt)
{
    fa_arg *a = (fa_arg *)p;
    const size_t i0 = pr_lo(a->n, t, nt), i1 = pr_lo(a->n, t + 1, nt);
    uint64_t c = 0;
    int64_t s = 0;
    uint32_t s32 = 0;
    for (size_t i = i0; i < i1; ++i) {
        const int32_t k = a->key[i];
        if (k >= a->lo && k <= a->hi) { ++c; s += a->val[i]; s32 += (uint32_t)a->val[i]; }
    }
    a->cnt[t] = c; a->sum[t] = s; a->sum32[t] = s32;
}

void pr_filter_agg(const int32_t *key, const int32_t *val, size_t n, int32_t lo, int32_t hi, int64_t out[3],
                   int threads)
{
    fa_arg *a = (fa_arg *)calloc(1, sizeof(fa_arg));
    const int nt = pr_threads(threads);
    a->key = key; a->val = val; a->n = n; a->lo = lo; a->hi = hi;
    pr_run(fa_body, a, nt);
    uint64_t c = 0; int64_t s = 0; uint32_t s32 = 0;
    for (int t = 0; t < nt; ++t) { c += a->cnt[t]; s += a->sum[t]; s32 += a->sum32[t]; }
    out[0] = (int64_t)c; out[1] = s; out[2] = (int64_t)(int32_t)s32;
    free(a);
}

/* ----------------------------------------------------- parallel radix sort */
typedef struct {
    uint64_t *key, *k2;
    uint32_t *val, *v2;
    size_t n;
    int shift;
    size_t (*hist)[256];  /* per thread */
} rs_arg;

static void rs_count(void *p, int t, int nt)
{
    rs_arg *a = (rs_arg *)p;
    size_t *h = a->hist[t];
    memset(h, 0, 256 * sizeof(size_t));
    const size_t i0 = pr_lo(a->n, t, nt), i1 = pr_lo(a->n, t + 1, nt);
    for (size_t i = i0; i < i1; ++i) h[(a->key[i] >> a->shift) & 255]++;
}

static void rs_scatter(void *p, int t, int nt)
{
    rs_arg *a = (rs_arg *)p;
    size_t *h = a->hist[t];  /* holds this thread's start per digit */
    const size_t i0 = pr_lo(a->n, t, nt), i1 = pr_lo(a->n, t + 1, nt);
    for (size_t i = i0; i < i1; ++i) {
        const size_t q = h[(a->key[i] >> a->shift) & 255]++;
        a->k2[q] = a->key[i];
        if (a->val) a->v2[q] = a->val[i];
    }
}

/* Stable LSD sort of key (with val) over bits [lo_bit, hi_bit); the result is
 * left in key / val. */
static void pr_radix(uint64_t *key, uint32_t *val, size_t n, int lo_bit, int hi_bit, int nt)
{
    if (n < 2 || lo_bit >= hi_bit) return;
    rs_arg a;
    a.key = key; a.val = val; a.n = n;
    a.k2 = (uint64_t *)malloc(n * sizeof(uint64_t));
    a.v2 = val ? (uint32_t *)malloc(n * sizeof(uint32_t)) : NULL;
    a.hist = (size_t (*)[256])malloc((size_t)nt * sizeof(*a.hist));
    uint64_t *korig = key;
    uint32_t *vorig = val;
    for (int sh = lo_bit; sh < hi_bit; sh += 8) {
        a.shift = sh;
        pr_run(rs_count, &a, nt);
        size_t s = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < nt; ++t) { const size_t c = a.hist[t][d]; a.hist[t][d] = s; s += c; }
        pr_run(rs_scatter, &a, nt);
        uint64_t *tk = a.key; a.key = a.k2; a.k2 = tk;
        uint32_t *tv = a.val; a.val = a.v2; a.v2 = tv;
    }
    if (a.key != korig) {
        memcpy(korig, a.key, n * sizeof(uint64_t));
        if (vorig) memcpy(vorig, a.val, n * sizeof(uint32_t));
        free(a.key);
        free(a.val);
    } else {
        free(a.k2);
        free(a.v2);
    }
    free(a.hist);
}

/* Highest bit (exclusive) in which the 64-bit keys differ from key[0], so that
 * passes over bits that every key shares are skipped. */
static int pr_hi_bit(const uint64_t *key, size_t n)
{
    uint64_t diff = 0;
    for (size_t i = 1; i < n; ++i) diff |= key[i] ^ key[0];
    int b = 0;
    while (b < 64 && (diff >> b)) ++b;
    return b;
}

/* ---------------------------------------------------------------- group-by */
typedef struct {
    const int32_t *val;
    const uint64_t *k;
    const uint32_t *ix;
    size_t n;
    size_t start[PR_MAX_THREADS + 1];  /* per-thread ranges cut at group starts */
    size_t groups[PR_MAX_THREADS];
    size_t goff[PR_MAX_THREADS];
    int32_t *okey, *omin, *omax;
    int64_t *ocount, *osum;
    int pass;  /* 0: count groups, 1: write them */
} gb_arg;

static void gb_body(void *p, int t, int nt)
{
    gb_arg *a = (gb_arg *)p;
    (void)nt;
    size_t i = a->start[t];
    const size_t end = a->start[t + 1];
    size_t g = a->pass ? a->goff[t] : 0;
    while (i < end) {
        size_t j = i;
        int64_t sum = 0;
        int32_t mn = INT32_MAX, mx = INT32_MIN;
        while (j < a->n && a->k[j] == a->k[i]) {
            if (a->pass) {
                const int32_t v = a->val[a->ix[j]];
                sum += v;
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
            ++j;
        }
        if (a->pass) {
            a->okey[g] = (int32_t)((uint32_t)a->k[i] ^ 0x80000000u);
            a->ocount[g] = (int64_t)(j - i);
            a->osum[g] = sum; a->omin[g] = mn; a->omax[g] = mx;
        }
        ++g;
        i = j;
    }
    if (!a->pass) a->groups[t] = g;
}

typedef struct { uint64_t *k; uint32_t *ix; const int32_t *key; size_t n; } gbk_arg;
static void gbk_body(void *p, int t, int nt)
{
    gbk_arg *a = (gbk_arg *)p;
    for (size_t i = pr_lo(a->n, t, nt); i < pr_lo(a->n, t + 1, nt); ++i) {
        a->k[i] = (uint32_t)a->key[i] ^ 0x80000000u;
        a->ix[i] = (uint32_t)i;
    }
}

/* Per group (key, COUNT, SUM int64, MIN, MAX), ascending by signed key; the
 * outputs of or_groupby_agg.  Returns the number of groups. */
size_t pr_groupby_agg(const int32_t *key, const int32_t *val, size_t n, int32_t *okey, int64_t *ocount,
                      int64_t *osum, int32_t *omin, int32_t *omax, int threads)
{
    if (n == 0) return 0;
    const int nt = pr_threads(threads);
    uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *ix = (uint32_t *)malloc(n * sizeof(uint32_t));
    gbk_arg ka = {k, ix, key, n};
    pr_run(gbk_body, &ka, nt);
    pr_radix(k, ix, n, 0, pr_hi_bit(k, n), nt);
    gb_arg *a = (gb_arg *)calloc(1, sizeof(gb_arg));
    a->val = val; a->k = k; a->ix = ix; a->n = n;
    a->okey = okey; a->ocount = ocount; a->osum = osum; a->omin = omin; a->omax = omax;
    for (int t = 0; t <= nt; ++t) {
        size_t s = pr_lo(n, t, nt);
        while (s > 0 && s < n && k[s] == k[s - 1]) ++s;  /* move the cut to a group start */
        a->start[t] = s;
    }
    for (int t = 1; t <= nt; ++t)
        if (a->start[t] < a->start[t - 1]) a->start[t] = a->start[t - 1];
    a->pass = 0;
    pr_run(gb_body, a, nt);
    size_t g = 0;
    for (int t = 0; t < nt; ++t) { a->goff[t] = g; g += a->groups[t]; }
    a->pass = 1;
    pr_run(gb_body, a, nt);
    free(a); free(k); free(ix);
    return g;
}

/* ----------------------------------------------------------------- join */
typedef struct {
    const int32_t *rkey, *rpay, *skey, *spay, *sfilt;
    size_t nr, ns;
    int32_t flo, fhi;
    uint64_t B;
    int64_t *head, *next;   /* chained buckets: head[h] -> next[i] -> ... */
    size_t cnt[PR_MAX_THREADS], off[PR_MAX_THREADS];
    int32_t *okey, *orpay, *ospay;
    size_t cap;
    int pass;
} hj_arg;

static void hj_init(void *p, int t, int nt)
{
    hj_arg *a = (hj_arg *)p;
    for (size_t b = pr_lo(a->B, t, nt); b < pr_lo(a->B, t + 1, nt); ++b) a->head[b] = -1;
}

static void hj_build(void *p, int t, int nt)
{
    hj_arg *a = (hj_arg *)p;
    for (size_t i = pr_lo(a->nr, t, nt); i < pr_lo(a->nr, t + 1, nt); ++i) {
        const size_t h = (uint32_t)a->rkey[i] & (a->B - 1);
        int64_t old = __atomic_load_n(&a->head[h], __ATOMIC_RELAXED);
        do a->next[i] = old;
        while (!__atomic_compare_exchange_n(&a->head[h], &old, (int64_t)i, 1, __ATOMIC_RELEASE, __ATOMIC_RELAXED));
    }
}

static void hj_probe(void *p, int t, int nt)
{
    hj_arg *a = (hj_arg *)p;
    size_t m = a->pass ? a->off[t] : 0;
    for (size_t j = pr_lo(a->ns, t, nt); j < pr_lo(a->ns, t + 1, nt); ++j) {
        if (a->sfilt && !(a->sfilt[j] >= a->flo && a->sfilt[j] <= a->fhi)) continue;
        const int32_t sk = a->skey[j];
        for (int64_t e = a->head[(uint32_t)sk & (a->B - 1)]; e >= 0; e = a->next[e]) {
            if (a->rkey[e] != sk) continue;
            if (a->pass && m < a->cap) { a->okey[m] = sk; a->orpay[m] = a->rpay[e]; a->ospay[m] = a->spay[j]; }
            ++m;
        }
    }
    if (!a->pass) a->cnt[t] = m;
}

/* Filtered hash join; returns the match count, writes at most cap rows. */
size_t pr_hash_join(const int32_t *rkey, const int32_t *rpay, size_t nr, const int32_t *skey, const int32_t *spay,
                    const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, int32_t *okey, int32_t *orpay,
                    int32_t *ospay, size_t cap, int threads)
{
    const int nt = pr_threads(threads);
    hj_arg *a = (hj_arg *)calloc(1, sizeof(hj_arg));
    a->rkey = rkey; a->rpay = rpay; a->skey = skey; a->spay = spay; a->sfilt = sfilt;
    a->nr = nr; a->ns = ns; a->flo = flo; a->fhi = fhi;
    a->B = 1;
    while (a->B < nr) a->B <<= 1;
    a->head = (int64_t *)malloc(a->B * sizeof(int64_t));
    a->next = (int64_t *)malloc((nr ? nr : 1) * sizeof(int64_t));
    a->okey = okey; a->orpay = orpay; a->ospay = ospay; a->cap = okey ? cap : 0;
    pr_run(hj_init, a, nt);
    pr_run(hj_build, a, nt);
    a->pass = 0;
    pr_run(hj_probe, a, nt);
    size_t m = 0;
    for (int t = 0; t < nt; ++t) { a->off[t] = m; m += a->cnt[t]; }
    if (a->cap) {
        a->pass = 1;
        pr_run(hj_probe, a, nt);
    }
    free(a->head); free(a->next); free(a);
    return m;
}

/* ------------------------------------------------------------ join-row sort */
typedef struct { int32_t *key, *rpay, *spay, *t; uint64_t *k; uint32_t *ix; size_t n; int step; } js_arg;
static void js_body(void *p, int t, int nt)
{
    js_arg *a = (js_arg *)p;
    const size_t i0 = pr_lo(a->n, t, nt), i1 = pr_lo(a->n, t + 1, nt);
    if (a->step == 0) {
        for (size_t i = i0; i < i1; ++i) {
            a->k[i] = ((uint64_t)((uint32_t)a->key[i] ^ 0x80000000u) << 32) | ((uint32_t)a->spay[i] ^ 0x80000000u);
            a->ix[i] = (uint32_t)i;
        }
    } else {
        for (size_t i = i0; i < i1; ++i) a->t[i] = a->rpay[a->ix[i]];
    }
}

/* Sorts join rows ascending by (key, s_payload), stable (as or_sort_join_rows). */
void pr_sort_join_rows(int32_t *key, int32_t *rpay, int32_t *spay, size_t m, int threads)
{
    if (m < 2) return;
    const int nt = pr_threads(threads);
    js_arg a = {key, rpay, spay, NULL, (uint64_t *)malloc(m * 8), (uint32_t *)malloc(m * 4), m, 0};
    pr_run(js_body, &a, nt);
    pr_radix(a.k, a.ix, m, 0, pr_hi_bit(a.k, m), nt);
    a.t = (int32_t *)malloc(m * 4);
    a.step = 1;
    pr_run(js_body, &a, nt);
    memcpy(rpay, a.t, m * 4);
    for (size_t i = 0; i < m; ++i) {
        key[i] = (int32_t)((uint32_t)(a.k[i] >> 32) ^ 0x80000000u);
        spay[i] = (int32_t)((uint32_t)a.k[i] ^ 0x80000000u);
    }
    free(a.k); free(a.ix); free(a.t);
}
