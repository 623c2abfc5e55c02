/*
 * qo.h -- native C ABI of libqo, the MI355X (gfx950) relational operator library.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are HIP
 * device memory; `qo_stream` is a hipStream_t passed as an opaque pointer
 * (NULL selects the library's own per-thread stream).  Operators are
 * asynchronous on the given stream unless stated otherwise; counts that the
 * reference read back mid-pipeline are written to device memory instead.
 * All functions return a qo_status; on failure qo_last_error() describes it.
 * No function ever exits the process.
 *
 * Record layout is the reference's 8-byte AoS tuple (host/host/common.h:27-28):
 * .x = value / row id, .y = key.  Column APIs take SoA int32 columns.
 */
#ifndef QO_H
#define QO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qo_record { uint32_t x, y; } qo_record;
typedef void *qo_stream;
typedef void *qo_event;

typedef enum qo_status {
    QO_OK = 0,
    QO_ERR_INVALID_ARG = 1,
    QO_ERR_HIP = 2,
    QO_ERR_CAPACITY = 3,   /* output buffer or hash table too small; count is still written */
    QO_ERR_NO_MEMORY = 4,
    QO_ERR_BUDGET = 5,     /* server admission: request can never fit the memory budget */
    QO_ERR_INJECTED = 6,   /* fault injection flag set by qo_debug_inject_fault */
    QO_ERR_NOT_FOUND = 7,
    QO_ERR_CANCELLED = 8,
    QO_ERR_DEVICE_WAIT = 9, /* a kernel's bounded device-side wait ran out (results undefined) */
    QO_ERR_TIMEOUT = 10     /* a host-side deadline passed before the device finished */
} qo_status;

/* Aggregate operators, same codes as the reference (host/host/common.h:89-92). */
enum { QO_REDUCE_SUM = 0, QO_REDUCE_MAX = 1, QO_REDUCE_MIN = 2, QO_REDUCE_AVERAGE = 3 };

typedef struct qo_device_info {
    char name[64];
    char arch[32];
    int cu_count;
    int lds_bytes_per_block;
    int wave_size;
    size_t hbm_bytes;
    int xcd_count;
} qo_device_info;

/* ---- runtime ---------------------------------------------------------- */
const char *qo_version(void);
const char *qo_last_error(void);
void qo_clear_error(void);
int qo_init(int device);
int qo_get_device_info(qo_device_info *out);
int qo_malloc(void **ptr, size_t bytes);
int qo_free(void *ptr);
int qo_memcpy_h2d(void *dst, const void *src, size_t bytes);
int qo_memcpy_d2h(void *dst, const void *src, size_t bytes);
int qo_memcpy_d2d(void *dst, const void *src, size_t bytes, qo_stream s);
int qo_fill_u32(void *dst, uint32_t value, size_t count, qo_stream s);
/* Device-to-device streaming copy (a plain read+write kernel; also the HBM
 * bandwidth calibration kernel of the benchmark). */
int qo_copy(void *dst, const void *src, size_t bytes, qo_stream s);
/* Same with non-temporal loads (the load type the streaming operators use). */
int qo_copy_nt(void *dst, const void *src, size_t bytes, qo_stream s);
int qo_stream_create(qo_stream *s);
int qo_stream_destroy(qo_stream s);
/* Both synchronise, then report QO_ERR_DEVICE_WAIT when a kernel's bounded
 * device-side wait (the look-back of the single-pass scans) ran out since the
 * last check. */
int qo_stream_sync(qo_stream s);
int qo_device_sync(void);
int qo_event_create(qo_event *e);
int qo_event_destroy(qo_event e);
int qo_event_record(qo_event e, qo_stream s);
int qo_event_sync(qo_event e);
int qo_event_query(qo_event e);      /* 1 = complete, 0 = pending, <0 error */
int qo_event_elapsed_ms(qo_event start, qo_event stop, float *ms);
/* Release cached per-stream workspaces (they grow on demand). */
int qo_release_workspaces(void);
/* Fault injection for failure-handling tests: the next `count` operator calls
 * return QO_ERR_INJECTED before launching anything. */
int qo_debug_inject_fault(int count);
/* Sleeps a look-back wave may wait for a predecessor tile before it gives up
 * (default 2^22, about a second); for tests of the bounded wait. */
int qo_debug_set_spin_limit(uint32_t sleeps);

/* ---- deterministic synthetic data (identical to the CPU generator spec) ---- */
/* out[i] = lo + floor(hi32(r_i) * range / 2^32), r_i = splitmix64 stream (seed, stream_id)
 * at position i; range == 0 means the full 32-bit draw. */
int qo_gen_uniform_u32(uint32_t *out, size_t n, uint64_t seed, uint64_t stream_id,
                       uint32_t lo, uint32_t range, qo_stream s);
/* out[i] = i-th element of a seeded permutation of [0, n). */
int qo_gen_perm_u32(uint32_t *out, size_t n, uint64_t perm_seed, qo_stream s);
/* out[i] = bij_bits(i): keyed bijection on [0, 2^bits) (unique keys). */
int qo_gen_bij_u32(uint32_t *out, size_t n, uint32_t bits, uint64_t seed, qo_stream s);
/* Foreign keys drawn uniformly over the build keys: out[i] = keymap(u_i), u_i uniform in
 * [0, nr); keymap = permutation of [0, nr) (bits == 0) or bij_bits (bits > 0). */
int qo_gen_fk_u32(uint32_t *out, size_t n, uint64_t seed, uint64_t stream_id, uint64_t nr,
                  uint32_t bits, uint64_t key_seed, qo_stream s);
/* Host: fixed-point Zipf CDF table (G entries of uint64). */
int qo_zipf_table(uint64_t *host_table, uint32_t G, double z);
/* Zipf(z) ranks over G groups mapped through a seeded permutation of [0, G);
 * d_table is the qo_zipf_table output copied to device memory. */
int qo_gen_zipf_u32(uint32_t *out, size_t n, const uint64_t *d_table, uint32_t G,
                    uint64_t seed, uint64_t stream_id, uint64_t perm_seed, qo_stream s);
/* Interleave two columns into Records: out[i] = {x[i], y[i]}. */
int qo_pack_records(const uint32_t *x, const uint32_t *y, size_t n, qo_record *out, qo_stream s);

/* ---- filter ------------------------------------------------------------- */
/* Fused filter + aggregate (COUNT, SUM of values where lo <= key <= hi).
 * d_out[0] = COUNT, d_out[1] = SUM as int64, d_out[2] = SUM wrapped to int32
 * (the reference's REDUCE_SUM), sign-extended. */
int qo_filter_agg(const int32_t *key, const int32_t *val, size_t n, int32_t lo, int32_t hi,
                  int64_t *d_out, qo_stream s);
int qo_filter_agg_records(const qo_record *rin, size_t n, int32_t lo, int32_t hi,
                          int64_t *d_out, qo_stream s);
/* Stable selection lo <= (int)R.y <= hi with compaction into rout (capacity n). */
int qo_filter_records(const qo_record *rin, size_t n, int32_t lo, int32_t hi,
                      qo_record *rout, uint64_t *d_count, qo_stream s);
/* No counterpart in the reference: the same compaction for callers that
 * cannot allocate (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Allocates nothing: its scratch (look-back status
 * words and ticket) is a fixed static device array of 1 MiB + 256 bytes, which
 * bounds n to 2^31 records (16-byte-aligned rin) or 2^30 (otherwise); larger n
 * fails with QO_ERR_INVALID_ARG.  Every call on a device shares that device's
 * array, so calls that overlap on one device race: order them (one stream). */
int qo_filter_records_noalloc(const qo_record *rin, size_t n, int32_t lo, int32_t hi, qo_record *rout,
                              uint64_t *d_count, qo_stream s);
/* Exclusive prefix sum of int32 (wraps modulo 2^32). in == out is allowed. */
int qo_scan_exclusive_i32(const int32_t *in, int32_t *out, size_t n, qo_stream s);

/* ---- sort ---------------------------------------------------------------- */
/* Sort Records ascending by .y (uint32), stable, in place. */
int qo_sort_records(qo_record *data, size_t n, qo_stream s);
/* Sort join rows ascending by (key, s_payload) as signed int32, stable, in place.
 * Synchronizes stream s once: the rows are sorted on the key first and each
 * key's run by s_payload after, and a key with more than 64 rows (read back
 * from the device) sends all rows through the full-width sort. */
int qo_sort_join_rows(int32_t *key, int32_t *rpay, int32_t *spay, size_t m, qo_stream s);

/* ---- sort-based group-by (reference pipeline) ------------------------------ */
/* On Records sorted by .y: startPos[g] = first index of group g; *d_numGroups = G. */
int qo_group_startpos(const qo_record *sorted, size_t n, int32_t *startPos,
                      uint64_t *d_numGroups, qo_stream s);
/* Per-group reduction of (int).x with the reference's operator semantics. */
int qo_aggregate_groups(const qo_record *sorted, size_t n, const int32_t *startPos,
                        size_t numGroups, int op, int32_t *out, qo_stream s);

/* ---- split, scatter, gather, projection, reduce ------------------------------ */
/* Split into num_part partitions (1..65536) by pid = R.y % num_part, stable
 * within a partition (splitImpl, host/host/testSplit.cpp:156-209; validateSplit,
 * Validate.cpp:510-527).  rin and rout are distinct n-Record arrays; d_bound
 * (num_part Records, may be NULL) receives {start, end} of every partition
 * (getBound_int, testSplit.cpp:121-140). */
int qo_split_records(const qo_record *rin, size_t n, uint32_t num_part, qo_record *rout,
                     qo_record *d_bound, qo_stream s);
/* S[pos] = R[loc[pos]] for pos < s_len (gatherImpl, testScatter.cpp:105-130);
 * a loc outside [0, r_len) leaves S[pos] unchanged.  Out of place: S != R. */
int qo_gather_records(const qo_record *R, size_t r_len, const int32_t *loc, size_t s_len,
                      qo_record *S, qo_stream s);
/* S[loc[i]] = R[i] for i < r_len (scatterImpl, testScatter.cpp:61-81; S holds
 * r_len Records); a loc outside [0, r_len) writes nothing.  When several i
 * target one position, one of them lands there, in no specified order (as in
 * the reference). */
int qo_scatter_records(const qo_record *R, size_t r_len, const int32_t *loc, qo_record *S, qo_stream s);
/* out[i] = {proj[i].x, R[proj[i].x].y} (validateProjection, Validate.cpp:39-71);
 * a row id proj[i].x >= r_len keeps proj[i].y.  out may be proj, not R. */
int qo_project_records(const qo_record *R, size_t r_len, const qo_record *proj, size_t p_len,
                       qo_record *out, qo_stream s);
/* Reduce R.y without grouping into *d_result (device memory), as validateReduce
 * computes it (Validate.cpp:372-468): op 0 SUM (unsigned, wraps mod 2^32),
 * 1 MAX and 2 MIN (unsigned), 3 AVERAGE (the wrapped sum / n); n == 0 gives 0. */
int qo_reduce_records(const qo_record *rin, size_t n, int op, uint32_t *d_result, qo_stream s);
/* Columns of Records: x[i] = R[i].x, y[i] = R[i].y (inverse of qo_pack_records). */
int qo_unpack_records(const qo_record *R, size_t n, uint32_t *x, uint32_t *y, qo_stream s);

/* ---- hash group-by aggregation -------------------------------------------- */
/* Per distinct key: (key, COUNT, SUM int64, MIN, MAX) ascending by key (signed).
 * group_capacity bounds the number of distinct keys (output arrays hold it).
 * The call is asynchronous: *d_numGroups (device memory) receives the number of
 * groups, a value above group_capacity when there were more distinct keys
 * (only the first group_capacity rows are then written; ~0 if the internal
 * table overflowed).  The host checks it after synchronizing. */
int qo_groupby_agg(const int32_t *key, const int32_t *val, size_t n, size_t group_capacity,
                   int32_t *okey, int64_t *ocount, int64_t *osum, int32_t *omin, int32_t *omax,
                   uint64_t *d_numGroups, qo_stream s);
/* Filtered group-by aggregation (the filter -> group-by -> aggregate chain of
 * host/host/testAggAfterGB.cpp:313-427 and kernel/filter_aggregation): only
 * rows with lo <= filt[i] <= hi are grouped and aggregated, the predicate
 * evaluated inside the group-by kernels (no compaction pass).  filt may be the
 * key column itself (the reference filters Record.y, the group key).  Same
 * outputs and capacity rules as qo_groupby_agg. */
int qo_groupby_agg_filtered(const int32_t *key, const int32_t *val, const int32_t *filt, size_t n, int32_t lo,
                            int32_t hi, size_t group_capacity, int32_t *okey, int64_t *ocount, int64_t *osum,
                            int32_t *omin, int32_t *omax, uint64_t *d_numGroups, qo_stream s);

/* ---- hash join ------------------------------------------------------------ */
/* Bytes of device memory a build table for `build_rows` rows needs. */
size_t qo_join_table_bytes(size_t build_rows);
/* Strategy the probe of n rows against a table of table_bytes takes for a
 * table of the given mode (0 hash-mixed, 1 key-range; see qo_join_table_info):
 * *partitioned = 1 for a partitioned probe, *bits = log2 of the partition
 * count, *splits = 1 for one split and L2-resident table slices (either mode),
 * 2 for two splits and LDS-resident slices (hash-mixed tables too large for
 * the first); *partitioned = 2 for a key-range table split by key into slices
 * of 2^15 slots whose payloads are looked up in LDS, *bits = log2 of the
 * partition count for a table of the largest capacity these slots allow
 * (capacity / 2^15 partitions for a smaller one; tables under 2^21 slots or
 * with (-1, -1) build tuples take the L2-resident slices); *partitioned = 0
 * for the direct probe.  Host-only. */
int qo_join_strategy(size_t table_bytes, size_t n, int mode, int *partitioned, int *bits, int *splits);
/* Builds the table in `table` (256-byte aligned device memory of table_bytes,
 * e.g. qo_join_table_bytes(nr)).  When the build keys span a range that fits a
 * power of two of the table's slots (and the TUNE op 7 limit), the table is
 * direct-mapped (home slot = key - min key), the layout of the reference's
 * key & (B - 1) buckets; duplicate keys fall back to a hash-mixed table, which
 * is built partition by partition through a per-stream workspace when it is
 * larger than one L2.  The choice is made on the device: the call stays
 * asynchronous. */
int qo_hash_build(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay,
                  size_t nr, qo_stream s);
/* Probe with the S filter flo <= sfilt <= fhi fused (sfilt may be NULL: no filter).
 * table_bytes is the size the table was built with (it selects the direct or the
 * partitioned probe).  Matches are written as (key, r_payload, s_payload) in an
 * unspecified row order; *d_count = matches (when it exceeds out_capacity only
 * out_capacity rows are written). */
int qo_hash_probe(const void *table, size_t table_bytes, const int32_t *skey, const int32_t *spay,
                  const int32_t *sfilt, size_t ns, int32_t flo, int32_t fhi, int32_t *okey,
                  int32_t *orpay, int32_t *ospay, size_t out_capacity, uint64_t *d_count, qo_stream s);
/* Which table a build made (synchronizes stream s): *mode = 0 hash-mixed,
 * 1 key-range; *slots_used = slots in use; *has_dups = 1 when some key has
 * several build tuples (the probe then walks whole chains). */
int qo_join_table_info(const void *table, int *mode, uint64_t *slots_used, int *has_dups, qo_stream s);
/* Filter S, build on R, probe, sort output by (key, s_payload).  Synchronous:
 * *h_count receives the match count.  Rows equal in (key, s_payload), which
 * only duplicate R keys produce, come in no specified order. */
int qo_hj_filter_sort(void *table, size_t table_bytes, const int32_t *rkey, const int32_t *rpay,
                      size_t nr, const int32_t *skey, const int32_t *spay, const int32_t *sfilt,
                      size_t ns, int32_t flo, int32_t fhi, int32_t *okey, int32_t *orpay,
                      int32_t *ospay, size_t out_capacity, uint64_t *h_count, qo_stream s);
/* Order-sensitive digest of m join rows (three int32 columns), on the device:
 * d_digest[0] = sum of mix(key, r_payload, s_payload, row index), d_digest[1] =
 * sum of keys (int64, wrapping), d_digest[2] = rows that order before their
 * predecessor by (key, s_payload) (0 when sorted).  With d_count the row count
 * is min(*d_count, m).  Two outputs are equal, order included, when their
 * digests are (up to 64-bit mixing collisions). */
int qo_join_rows_digest(const int32_t *key, const int32_t *rpay, const int32_t *spay, size_t m,
                        const uint64_t *d_count, uint64_t *d_digest, qo_stream s);

/* Reference-shaped bucket-chained table (hj_filter_sort_op.cl:25-109): B buckets
 * (power of two) of bucketSize (key, val) slots plus per-bucket counts.  Input
 * and output tables are interleaved uint pairs.  Overflowing a bucket gives
 * QO_ERR_CAPACITY.  The probe writes compacted (s_val, r_val) pairs. */
int qo_bucket_build(const uint32_t *rTable, uint32_t size, uint32_t *rHashTable, uint32_t B,
                    uint32_t bucketSize, uint32_t *rHashCount, qo_stream s);
int qo_bucket_probe(const uint32_t *rHashTable, const uint32_t *rHashCount, uint32_t B,
                    uint32_t bucketSize, const uint32_t *sTable, uint32_t size,
                    uint32_t *matchedTable, size_t capacity_pairs, uint64_t *d_count, qo_stream s);

/* ---- batched hj_filter_sort (one hipGraph launch for many queries) ---------- */
typedef struct qo_join_query {
    const int32_t *rkey, *rpay;
    const int32_t *skey, *spay, *sfilt;
    uint64_t nr, ns;
    int32_t flo, fhi;
    int32_t *okey, *orpay, *ospay;
    uint64_t out_capacity;
} qo_join_query;
typedef struct qo_batch qo_batch;
/* Plan a batch: allocates the batch's tables, workspaces and the captured graph. */
int qo_batch_create(const qo_join_query *queries, int nq, qo_batch **out);
/* Launch the whole batch (one graph launch) on stream s. */
int qo_batch_launch(qo_batch *b, qo_stream s);
/* After completion: per-query match counts (host array of nq). */
int qo_batch_counts(qo_batch *b, uint64_t *h_counts);
int qo_batch_destroy(qo_batch *b);

/* ---- cost-model planner ------------------------------------------------------- */
/* A filter -> group-by -> aggregate query over Records, the chains the
 * reference's planner schedules (host/dynamic_programming.cpp:874-1023:
 * selection, group + aggregation, filter + group + aggregation): an optional
 * range filter lo <= (int)R.y <= hi, then a per-group aggregate of (int)R.x
 * grouped by R.y (group = 1; op 0 SUM, 1 MAX, 2 MIN, 3 AVERAGE with the
 * reference's group semantics: SUM wraps, MAX starts at 0, MIN at 2^30,
 * AVERAGE truncates), or the no-group reduce of R.y (group = 0, op >= 0), or
 * nothing more (group = 0, op = -1: the selection itself). */
typedef struct {
    int filter;          /* 1: apply lo <= (int)R.y <= hi first */
    int32_t lo, hi;
    int group;           /* 1: group by R.y */
    int op;              /* 0..3, or -1 for a selection */
    double selectivity;  /* expected fraction of rows passing the filter (cost model only) */
    double groups;       /* expected number of groups (cost model only) */
} qo_query;
/* Implementations a plan segment can use. */
enum {
    QO_IMPL_FILTER = 0,                /* one-pass stable compaction (selected Records materialized) */
    QO_IMPL_SORT_GROUP_AGG = 1,        /* radix sort, group start positions, per-group aggregate */
    QO_IMPL_HASH_GROUP_AGG = 2,        /* Records -> columns, hash group-by, reference-semantics epilogue */
    QO_IMPL_FILTER_HASH_GROUP_AGG = 3, /* the same with the filter inside the group-by kernels */
    QO_IMPL_REDUCE = 4,                /* no-group reduce */
    QO_IMPL_NUM = 5
};
/* Cost-model primitives: each costs fixed_us + bytes moved / gbps. */
enum {
    QO_COST_FILTER = 0,     /* filter compaction: 8 B read + 8 B per selected row */
    QO_COST_SORT_PASS = 1,  /* one radix pass: 24 B per row */
    QO_COST_STARTPOS = 2,   /* group start positions: 8 B per row */
    QO_COST_AGG = 3,        /* one per-group aggregate: 8 B per row */
    QO_COST_UNPACK = 4,     /* Records -> columns or a Record copy: 16 B per row */
    QO_COST_HASH_SMALL = 5, /* hash group-by, <= 1024 groups: 8 B per row */
    QO_COST_HASH_LARGE = 6, /* hash group-by, more groups: 8 B per row */
    QO_COST_REDUCE = 7,     /* no-group reduce: 8 B per row */
    QO_COST_FHASH_SMALL = 8, /* hash group-by with the filter fused, half the rows selected, <= 1024 groups */
    QO_COST_FHASH_LARGE = 9, /* the same, more groups */
    QO_COST_NUM = 10
};
#define QO_PLAN_MAX_SEGS 4
typedef struct {
    int nseg;                          /* segments, in execution order */
    int impl[QO_PLAN_MAX_SEGS];        /* QO_IMPL_* */
    int first_step[QO_PLAN_MAX_SEGS];  /* steps covered (0-based positions in the chain filter?, group/reduce) */
    int last_step[QO_PLAN_MAX_SEGS];
    double seg_ms[QO_PLAN_MAX_SEGS];   /* estimated ms per segment */
    double est_ms;                     /* estimated ms of the whole query */
    int tune_op[QO_PLAN_MAX_SEGS];     /* tune-table operator setting the segment's launch (-1: none) */
    int block[QO_PLAN_MAX_SEGS], items[QO_PLAN_MAX_SEGS], param[QO_PLAN_MAX_SEGS];
} qo_plan;
/* Cheapest cover of the query's steps by implementations in allowed_impls (a
 * bit mask of QO_IMPL_*; 0 = all) for n input rows: a dynamic programme over
 * the chain's prefixes (dp_for_selection, host/dynamic_programming.cpp:874-996)
 * with each segment costed from the primitive table.  Host only: needs no device. */
int qo_plan_query(const qo_query *q, size_t n, unsigned allowed_impls, qo_plan *plan);
/* The primitive cost table (defaults measured on MI355X; qo_plan_calibrate
 * re-measures it on the current device).  Host only. */
int qo_plan_cost_set(int prim, double fixed_us, double gbps);
int qo_plan_cost_get(int prim, double *fixed_us, double *gbps);
/* Times every primitive on synthetic data of 2^size_class rows on the device
 * (median of reps) and stores the measured costs. */
int qo_plan_calibrate(int size_class, int reps);
/* Runs a plan on n device Records; synchronizes stream s (intermediate sizes
 * size the later segments).  Outputs: a selection writes d_sel (n Records) and
 * *d_count = rows; a grouped query writes d_keys[g] (R.y, ascending unsigned),
 * d_aggs[g] and *d_count = groups (more than group_capacity fails with
 * QO_ERR_INVALID_ARG); a reduce writes its uint32 result to d_aggs[0] and
 * *d_count = rows reduced. */
int qo_plan_execute(const qo_plan *plan, const qo_query *q, const qo_record *rin, size_t n, size_t group_capacity,
                    qo_record *d_sel, uint32_t *d_keys, int32_t *d_aggs, uint64_t *d_count, qo_stream s);

/* ---- autotuner -------------------------------------------------------------- */
/* Operators: 0 filter_agg, 1 filter, 2 probe (param = table load %), 3 groupby, 4 sort,
 * 5 partitioned join (items = LDS table slice KiB per partition, param = 0 when
 * the table outgrows one L2 / 1 always / 2 never / 3 always, hash-mixed probes
 * through LDS slices / 4 when the table outgrows one L2, key-range probes through
 * L2 only), 6 partitioned group-by (items =
 * groups per fine bucket, param = 0 auto / 1 always / 2 never), 7 key-range join
 * table (items = largest such table in MiB, param = 0 when the keys allow /
 * 2 never), 8 batched join (class of the batch's mean |R|; items = S tiles per
 * probe ticket, param = R rows per build tile), 9 batched join table load %
 * (class of the mean |R|).  Size class = floor(log2 n).
 * An operator's class with no entry takes the nearest class that has one, and
 * an operator with none its built-in default.  qo_init loads the tune file
 * tune/<arch>.tune beside libqo.so (or $QO_TUNE_FILE; "none" loads nothing). */
int qo_tune_get(int op, int size_class, int *block, int *items, int *param);
int qo_tune_set(int op, int size_class, int block, int items, int param);
int qo_tune_load(const char *path);
int qo_tune_save(const char *path);
/* Clears every entry (all operators back to their built-in defaults). */
int qo_tune_reset(void);
/* Searches operator `op`'s launch parameters at size class `size_class` on the
 * device: builds a synthetic workload of that size in HBM, times the built-in
 * default and every candidate (median of `reps` runs after a warm-up, HIP
 * events) and sets the fastest into the table.  Tunable operators: 0, 3, 6, 8
 * and 9.  Outputs the chosen parameters and both times (ms). */
int qo_autotune(int op, int size_class, int reps, int *block, int *items, int *param, double *best_ms,
                double *default_ms);

/* ---- serving loop ------------------------------------------------------------ */
/* A request loop over tables resident in HBM, run on the caller's thread by
 * qo_server_run / drain / poll.  Each request has an arrival time; the loop
 * takes every request that has arrived (arrival order, up to
 * max_batch_queries, within the device-memory budget) as ONE batch on the
 * server's stream, waits for it, and starts over with what arrived meanwhile
 * (continuous batching).  Requests of a batch share launches (one multi-query
 * filter kernel, the batched join pipeline), so their work overlaps on the
 * device.  A request larger than the whole budget fails with QO_ERR_BUDGET.
 * With `plain` set the loop runs each request alone, in arrival order, through
 * the per-query operator API (the baseline).  Results are summaries computed on
 * the device.  A server is used from one thread at a time. */
typedef struct qo_server qo_server;
typedef struct qo_server_config {
    uint64_t memory_budget_bytes;  /* device-memory admission budget */
    int max_inflight_batches;      /* must be 1: one batch in flight per server */
    int max_batch_queries;         /* requests launched together in one batch */
    int plain;                     /* 1: no batching or overlap (per-query path) */
    double batch_timeout_ms;       /* host deadline per batch (<= 0: 60000); a batch that
                                      misses it fails with QO_ERR_TIMEOUT */
} qo_server_config;
enum { QO_Q_FILTER = 0, QO_Q_GROUPBY = 1, QO_Q_JOIN_SORT = 2 };
/* Columns of a resident table (int32 device arrays owned by the caller; a
 * column an operator does not use may be NULL). */
typedef struct qo_table_desc {
    const int32_t *key;      /* unique keys: join build side */
    const int32_t *fk;       /* foreign keys: join probe side */
    const int32_t *grp;      /* group-by key */
    const int32_t *val;      /* value / payload */
    const int32_t *filt;     /* filter column */
    uint64_t n;              /* rows */
    uint64_t group_capacity; /* bound on the distinct grp values */
} qo_table_desc;
typedef struct qo_request {
    int kind;           /* QO_Q_* */
    int table_a;        /* filter / group-by input; join build side (key, val) */
    int table_b;        /* join probe side (fk, val, filt) */
    int32_t lo, hi;     /* predicate lo <= filt <= hi (filter and join) */
} qo_request;
typedef struct qo_result {
    uint64_t id;
    int status;         /* QO_OK or the error of this request alone */
    int kind;
    uint64_t count;     /* filter: selected rows; group-by: groups; join: matches */
    int64_t sum;        /* filter: SUM(val); group-by: sum of the group SUMs; join: SUM(key) */
    int64_t checksum;   /* order-sensitive checksum of the materialised, sorted output */
    double submit_ms, start_ms, end_ms;  /* server clock (ms since creation) */
} qo_result;
int qo_server_create(const qo_server_config *cfg, qo_server **out);
int qo_server_add_table(qo_server *sv, const qo_table_desc *table, int *table_id);
/* Queues a request arriving now (server clock). */
int qo_server_submit(qo_server *sv, const qo_request *req, uint64_t *id);
/* Queues a request arriving at `arrival_ms` on the server clock (may be in the
 * future: the loop does not start it earlier; submit_ms reports it). */
int qo_server_submit_at(qo_server *sv, const qo_request *req, double arrival_ms, uint64_t *id);
/* Serves until every queued request has finished (waiting for arrivals that
 * are not due yet). */
int qo_server_run(qo_server *sv);
/* Copies up to `max` finished results.  When none is ready and timeout_ms != 0,
 * serves first until one is, or until timeout_ms has passed (timeout_ms < 0:
 * no limit; returns 0 when nothing is queued).  Returns the number copied
 * (>= 0) or a negative qo_status. */
int qo_server_poll(qo_server *sv, qo_result *out, int max, int timeout_ms);
/* Same as qo_server_run. */
int qo_server_drain(qo_server *sv);
int qo_server_destroy(qo_server *sv);
/* Runs one request synchronously on the plain per-query path (same result). */
int qo_server_run_plain(qo_server *sv, const qo_request *req, qo_result *out);
/* Server clock now (ms since creation). */
double qo_server_now_ms(qo_server *sv);

#ifdef __cplusplus
}
#endif
#endif /* QO_H */
