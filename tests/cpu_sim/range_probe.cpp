// CPU probe driver: joins whose key-range table is probed through LDS slices
// (the default plan for tables of 2^21 slots and more), through L2 (tune
// operator 5, param 4) and without partitioning (param 2), each output equal
// as a row multiset to a plain hash join on the host.  Cases: a permutation
// table of 2^21 slots; sparse keys in a table allocated for 2^22 slots (the
// device capacity below the planned one, so buckets take doubled regions);
// the same permutation plus a (-1, -1) build tuple (not on the LDS route);
// a table of 2^19 slots (below the LDS route's minimum).  The first case also
// runs qo_hj_filter_sort (the sort-ready output form of the probe) against
// the host rows in (key, s_payload) order.
//
//     range_probe [log2 |S|]
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "qo.h"

#define CK(x)                                                                         \
    do {                                                                              \
        int rc_ = (x);                                                                \
        if (rc_ != QO_OK) {                                                           \
            fprintf(stderr, "%s:%d %s -> %d (%s)\n", __FILE__, __LINE__, #x, rc_,     \
                    qo_last_error());                                                 \
            exit(1);                                                                  \
        }                                                                             \
    } while (0)

typedef std::tuple<int32_t, int32_t, int32_t> Row;  // key, r_payload, s_payload

static void set_all(int op, int param) {
    for (int cls = 0; cls <= 40; ++cls) {
        int b = 0, it = 0, pa = 0;
        CK(qo_tune_get(op, cls, &b, &it, &pa));
        CK(qo_tune_set(op, cls, b, it, param));
    }
}

static uint64_t g_x = 88172645463325252ull;
static uint64_t rnd() {
    g_x ^= g_x << 13;
    g_x ^= g_x >> 7;
    g_x ^= g_x << 17;
    return g_x;
}

struct Case {
    std::vector<int32_t> rk, rp, sk, sp, sf;
    size_t table_bytes;
};

static void *dev(const std::vector<int32_t> &v) {
    void *p = nullptr;
    CK(qo_malloc(&p, std::max<size_t>(4, v.size() * 4)));
    if (!v.empty()) CK(qo_memcpy_h2d(p, v.data(), v.size() * 4));
    return p;
}

// Host join of the filtered S (filter value in [0, 49]).
static std::vector<Row> host_join(const Case &c) {
    std::unordered_multimap<int32_t, int32_t> m;
    m.reserve(c.rk.size() * 2);
    for (size_t i = 0; i < c.rk.size(); ++i) m.emplace(c.rk[i], c.rp[i]);
    std::vector<Row> out;
    for (size_t i = 0; i < c.sk.size(); ++i) {
        if (c.sf[i] < 0 || c.sf[i] > 49) continue;
        auto r = m.equal_range(c.sk[i]);
        for (auto it = r.first; it != r.second; ++it) out.emplace_back(c.sk[i], it->second, c.sp[i]);
    }
    std::sort(out.begin(), out.end(), [](const Row &a, const Row &b) {
        return std::make_tuple(std::get<0>(a), std::get<2>(a), std::get<1>(a)) <
               std::make_tuple(std::get<0>(b), std::get<2>(b), std::get<1>(b));
    });
    return out;
}

// Build + probe with tune operator 5 at `param`; rows sorted like host_join's.
static std::vector<Row> run_probe(const Case &c, int param, int *mode) {
    set_all(5, param);
    void *rk = dev(c.rk), *rp = dev(c.rp), *sk = dev(c.sk), *sp = dev(c.sp), *sf = dev(c.sf);
    void *table = nullptr, *ok = nullptr, *orp = nullptr, *osp = nullptr, *cnt = nullptr;
    const size_t cap = c.sk.size() + 16;
    CK(qo_malloc(&table, c.table_bytes));
    CK(qo_malloc(&ok, cap * 4));
    CK(qo_malloc(&orp, cap * 4));
    CK(qo_malloc(&osp, cap * 4));
    CK(qo_malloc(&cnt, 8));
    CK(qo_hash_build(table, c.table_bytes, (int32_t *)rk, (int32_t *)rp, c.rk.size(), nullptr));
    uint64_t used = 0;
    int dups = 0;
    CK(qo_join_table_info(table, mode, &used, &dups, nullptr));
    CK(qo_fill_u32(cnt, 0, 2, nullptr));
    CK(qo_hash_probe(table, c.table_bytes, (int32_t *)sk, (int32_t *)sp, (int32_t *)sf, c.sk.size(), 0, 49,
                     (int32_t *)ok, (int32_t *)orp, (int32_t *)osp, cap, (uint64_t *)cnt, nullptr));
    CK(qo_device_sync());
    uint64_t m = 0;
    CK(qo_memcpy_d2h(&m, cnt, 8));
    std::vector<Row> out;
    if (m <= cap) {
        std::vector<int32_t> a(m), b(m), d(m);
        if (m) {
            CK(qo_memcpy_d2h(a.data(), ok, m * 4));
            CK(qo_memcpy_d2h(b.data(), orp, m * 4));
            CK(qo_memcpy_d2h(d.data(), osp, m * 4));
        }
        for (uint64_t i = 0; i < m; ++i) out.emplace_back(a[i], b[i], d[i]);
        std::sort(out.begin(), out.end(), [](const Row &x, const Row &y) {
            return std::make_tuple(std::get<0>(x), std::get<2>(x), std::get<1>(x)) <
                   std::make_tuple(std::get<0>(y), std::get<2>(y), std::get<1>(y));
        });
    } else {
        out.assign(m, Row{0, 0, 0});  // more rows than S tuples: reported as a size mismatch
    }
    for (void *p : {rk, rp, sk, sp, sf, table, ok, orp, osp, cnt}) CK(qo_free(p));
    return out;
}

// qo_hj_filter_sort on the default plan: rows in (key, s_payload) order.
static std::vector<Row> run_sorted(const Case &c) {
    set_all(5, 0);
    void *rk = dev(c.rk), *rp = dev(c.rp), *sk = dev(c.sk), *sp = dev(c.sp), *sf = dev(c.sf);
    void *table = nullptr, *ok = nullptr, *orp = nullptr, *osp = nullptr;
    const size_t cap = c.sk.size() + 16;
    CK(qo_malloc(&table, c.table_bytes));
    CK(qo_malloc(&ok, cap * 4));
    CK(qo_malloc(&orp, cap * 4));
    CK(qo_malloc(&osp, cap * 4));
    uint64_t m = 0;
    CK(qo_hj_filter_sort(table, c.table_bytes, (int32_t *)rk, (int32_t *)rp, c.rk.size(), (int32_t *)sk,
                         (int32_t *)sp, (int32_t *)sf, c.sk.size(), 0, 49, (int32_t *)ok, (int32_t *)orp,
                         (int32_t *)osp, cap, &m, nullptr));
    std::vector<Row> out;
    if (m <= cap) {
        std::vector<int32_t> a(m), b(m), d(m);
        if (m) {
            CK(qo_memcpy_d2h(a.data(), ok, m * 4));
            CK(qo_memcpy_d2h(b.data(), orp, m * 4));
            CK(qo_memcpy_d2h(d.data(), osp, m * 4));
        }
        for (uint64_t i = 0; i < m; ++i) out.emplace_back(a[i], b[i], d[i]);
    }
    for (void *p : {rk, rp, sk, sp, sf, table, ok, orp, osp}) CK(qo_free(p));
    return out;
}

// R: keys `keys` (shuffled), random payloads; S: ns keys uniform over the key
// range widened by 1000 on each side, plus some -1 keys.
static Case make_case(std::vector<int32_t> keys, size_t ns, size_t table_bytes) {
    Case c;
    for (size_t i = keys.size(); i > 1; --i) std::swap(keys[i - 1], keys[rnd() % i]);
    c.rk = keys;
    c.rp.resize(keys.size());
    for (auto &v : c.rp) v = (int32_t)(uint32_t)rnd();
    int32_t lo = *std::min_element(keys.begin(), keys.end()), hi = *std::max_element(keys.begin(), keys.end());
    const uint64_t span = (uint64_t)((int64_t)hi - lo) + 2001;
    c.sk.resize(ns);
    c.sp.resize(ns);
    c.sf.resize(ns);
    for (size_t i = 0; i < ns; ++i) {
        c.sk[i] = (i % 97 == 5) ? -1 : (int32_t)((int64_t)lo - 1000 + (int64_t)(rnd() % span));
        c.sp[i] = (int32_t)(uint32_t)rnd();
        c.sf[i] = (int32_t)(rnd() % 100);
    }
    c.table_bytes = table_bytes ? table_bytes : qo_join_table_bytes(keys.size());
    return c;
}

int main(int argc, char **argv) {
    const int lgs = argc > 1 ? atoi(argv[1]) : 20;
    const size_t ns = size_t(1) << lgs;
    CK(qo_init(0));
    int fails = 0;
    std::vector<int32_t> perm(3u << 19);
    for (size_t i = 0; i < perm.size(); ++i) perm[i] = (int32_t)i;
    std::vector<int32_t> sparse(1u << 16);
    for (size_t i = 0; i < sparse.size(); ++i) sparse[i] = (int32_t)(i * 24) - 5000;
    std::vector<int32_t> with_sentinel = perm;
    with_sentinel.push_back(-1);
    std::vector<int32_t> small(1u << 19);
    for (size_t i = 0; i < small.size(); ++i) small[i] = (int32_t)i + 70000;
    struct Spec {
        const char *name;
        std::vector<int32_t> *keys;
        size_t table_bytes;
    } specs[] = {{"permutation, 2^21 slots", &perm, 0},
                 {"sparse keys, table for 2^22 slots", &sparse, 256 + (size_t(8) << 22)},
                 {"permutation with a (-1, -1) tuple", &with_sentinel, 0},
                 {"2^19 slots", &small, 0}};
    for (int ci = 0; ci < 4; ++ci) {
        Case c = make_case(*specs[ci].keys, ns, specs[ci].table_bytes);
        if (ci == 2) c.rp[std::find(c.rk.begin(), c.rk.end(), -1) - c.rk.begin()] = -1;
        const std::vector<Row> want = host_join(c);
        int modes[3];
        const std::vector<Row> lds = run_probe(c, 0, &modes[0]), l2 = run_probe(c, 4, &modes[1]),
                               direct = run_probe(c, 2, &modes[2]);
        const bool ok = modes[0] == 1 && modes[1] == 1 && modes[2] == 1 && lds == want && l2 == want && direct == want;
        printf("%s: %zu rows, key-range table %s; default plan %s, L2 %s, direct %s\n", specs[ci].name, want.size(),
               modes[0] == 1 ? "yes" : "NO", lds == want ? "same rows" : "DIFFERENT",
               l2 == want ? "same rows" : "DIFFERENT", direct == want ? "same rows" : "DIFFERENT");
        fails += !ok;
        if (ci == 0) {
            const std::vector<Row> sorted = run_sorted(c);
            const bool sok = sorted == want;
            printf("%s: hj_filter_sort %s\n", specs[ci].name, sok ? "same rows in order" : "DIFFERENT");
            fails += !sok;
        }
    }
    set_all(5, 0);
    CK(qo_release_workspaces());
    printf(fails ? "FAIL\n" : "PASS\n");
    return fails ? 1 : 0;
}
