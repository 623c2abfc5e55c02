// CPU build driver: batched hj_filter_sort with the plain path's partitioned
// build and probe of large tables left to their rule, forced for every query
// and switched off (tune operator 5, param 0 / 1 / 2), each query's three
// output columns compared with the same query run alone through
// qo_hj_filter_sort.  Under the ASan build every device access of the batch
// kernels is bounds-checked.
//
//     batch_parts [scale]
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

static int32_t *dev(uint64_t n) {
    void *p = nullptr;
    CK(qo_malloc(&p, (n ? n : 1) * 4));
    return static_cast<int32_t *>(p);
}

static void set_all(int op, int param) {
    for (int cls = 0; cls <= 40; ++cls) {
        int b = 0, it = 0, pa = 0;
        CK(qo_tune_get(op, cls, &b, &it, &pa));
        CK(qo_tune_set(op, cls, b, it, param));
    }
}

int main(int argc, char **argv) {
    const double scale = argc > 1 ? atof(argv[1]) : 1.0;
    // |R| per query (|S| = 3|R| + 1000; the last query has an empty build side).
    const uint64_t nr_base[] = {5000, 20000, 70000, 3000, 150000, 9000, 40000, 1, 60000, 0};
    const int nq = (int)(sizeof(nr_base) / sizeof(nr_base[0]));
    CK(qo_init(0));
    std::vector<qo_join_query> q(nq);
    std::vector<int32_t *> sfilt(nq);
    for (int i = 0; i < nq; ++i) {
        const uint64_t nr = nr_base[i] > 1 ? (uint64_t)(nr_base[i] * scale) : nr_base[i];
        const uint64_t ns = 3 * nr + 1000;
        int32_t *rk = dev(nr), *rp = dev(nr), *sk = dev(ns), *sp = dev(ns), *sf = dev(ns);
        if (nr) {
            CK(qo_gen_bij_u32((uint32_t *)rk, nr, 31, 100 + i, nullptr));
            CK(qo_gen_uniform_u32((uint32_t *)rp, nr, 100 + i, 1, 0, 0, nullptr));
            CK(qo_gen_fk_u32((uint32_t *)sk, ns, 100 + i, 2, nr, 31, 100 + i, nullptr));
        } else {
            CK(qo_gen_uniform_u32((uint32_t *)sk, ns, 100 + i, 2, 0, 0, nullptr));
        }
        CK(qo_gen_uniform_u32((uint32_t *)sp, ns, 100 + i, 3, 0, 0, nullptr));
        CK(qo_gen_uniform_u32((uint32_t *)sf, ns, 100 + i, 4, 0, 100, nullptr));
        sfilt[i] = sf;
        qo_join_query &j = q[i];
        memset(&j, 0, sizeof(j));
        j.rkey = rk; j.rpay = rp; j.skey = sk; j.spay = sp; j.sfilt = i % 3 == 2 ? nullptr : sf;
        j.nr = nr; j.ns = ns; j.flo = INT32_MIN; j.fhi = 10 + 9 * i;
        j.okey = dev(ns); j.orpay = dev(ns); j.ospay = dev(ns);
        j.out_capacity = ns;
    }
    CK(qo_device_sync());
    // Expected rows: every query alone on the plain path.
    std::vector<std::vector<int32_t>> want(nq);
    std::vector<uint64_t> want_n(nq);
    for (int i = 0; i < nq; ++i) {
        const qo_join_query &j = q[i];
        const size_t tb = qo_join_table_bytes(j.nr);
        void *table = nullptr;
        CK(qo_malloc(&table, tb));
        int32_t *ok = dev(j.ns), *orp = dev(j.ns), *osp = dev(j.ns);
        uint64_t m = 0;
        CK(qo_hj_filter_sort(table, tb, j.rkey, j.rpay, j.nr, j.skey, j.spay, j.sfilt, j.ns, j.flo, j.fhi, ok, orp,
                             osp, j.ns, &m, nullptr));
        CK(qo_device_sync());
        want_n[i] = m;
        want[i].resize(3 * m);
        CK(qo_memcpy_d2h(want[i].data(), ok, m * 4));
        CK(qo_memcpy_d2h(want[i].data() + m, orp, m * 4));
        CK(qo_memcpy_d2h(want[i].data() + 2 * m, osp, m * 4));
        for (void *p : {table, (void *)ok, (void *)orp, (void *)osp}) CK(qo_free(p));
    }
    int fails = 0;
    const char *mode_name[] = {"by rule", "forced", "off"};
    for (int part : {0, 1, 2}) {
        set_all(5, part);
        qo_batch *b = nullptr;
        CK(qo_batch_create(q.data(), nq, &b));
        CK(qo_batch_launch(b, nullptr));
        CK(qo_device_sync());
        std::vector<uint64_t> n(nq);
        CK(qo_batch_counts(b, n.data()));
        int bad = 0;
        for (int i = 0; i < nq; ++i) {
            const uint64_t m = n[i];
            if (m != want_n[i]) {
                ++bad;
                continue;
            }
            std::vector<int32_t> got(3 * m);
            CK(qo_memcpy_d2h(got.data(), q[i].okey, m * 4));
            CK(qo_memcpy_d2h(got.data() + m, q[i].orpay, m * 4));
            CK(qo_memcpy_d2h(got.data() + 2 * m, q[i].ospay, m * 4));
            bad += got != want[i];
        }
        CK(qo_batch_destroy(b));
        printf("partitioned build and probe %s: %d of %d queries differ\n", mode_name[part], bad, nq);
        fflush(stdout);
        fails += bad;
    }
    for (const qo_join_query &j : q)
        for (const void *p : {(const void *)j.rkey, (const void *)j.rpay, (const void *)j.skey, (const void *)j.spay,
                              (const void *)j.okey, (const void *)j.orpay, (const void *)j.ospay})
            CK(qo_free(const_cast<void *>(p)));
    for (int i = 0; i < nq; ++i) CK(qo_free(sfilt[i]));
    CK(qo_release_workspaces());
    printf(fails ? "FAIL\n" : "PASS\n");
    return fails ? 1 : 0;
}
