// CPU build driver: the served mixed batch of the serving tests (four tables
// of 50K-3M rows, 24 mixed requests in batches of 8), each result compared
// with the same request run alone on the plain path.
//
//     serve_mix [scale] [requests] [canary | canary-part]
//
// With "canary" every buffer carved from a batch's arena (and from the
// workspaces carved inside it) is followed by 256 bytes of pattern, checked
// after each batch: a kernel that writes past its buffer damages it.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qo.h"

extern "C" void qo_sim_canary_enable(int on);
extern "C" long qo_sim_canary_check(void);

#define CK(x)                                                                         \
    do {                                                                              \
        int rc_ = (x);                                                                \
        if (rc_ != QO_OK) {                                                           \
            fprintf(stderr, "%s:%d %s -> %d (%s)\n", __FILE__, __LINE__, #x, rc_,     \
                    qo_last_error());                                                 \
            exit(1);                                                                  \
        }                                                                             \
    } while (0)

struct Tab {
    uint64_t n, G;
    int32_t *key, *fk, *grp, *val, *filt;
};

static int32_t *dev(uint64_t n) {
    void *p = nullptr;
    CK(qo_malloc(&p, n * 4));
    return static_cast<int32_t *>(p);
}

int main(int argc, char **argv) {
    const uint64_t sizes[4][2] = {{50000, 16}, {200000, 1000}, {1000000, 50000}, {3000000, 1u << 20}};
    const double scale = argc > 1 ? atof(argv[1]) : 1.0;
    const uint64_t seed = 0x5E21;
    CK(qo_init(0));
    std::vector<Tab> t(4);
    for (int i = 0; i < 4; ++i) {
        Tab &T = t[i];
        T.n = (uint64_t)(sizes[i][0] * scale);
        T.G = sizes[i][1];
        T.key = dev(T.n); T.fk = dev(T.n); T.grp = dev(T.n); T.val = dev(T.n); T.filt = dev(T.n);
        CK(qo_gen_bij_u32((uint32_t *)T.key, T.n, 31, seed, nullptr));
        CK(qo_gen_fk_u32((uint32_t *)T.fk, T.n, seed + i, 1, T.n, 31, seed, nullptr));
        std::vector<uint64_t> zt(T.G);
        CK(qo_zipf_table(zt.data(), (uint32_t)T.G, 0.8));
        void *dz = nullptr;
        CK(qo_malloc(&dz, T.G * 8));
        CK(qo_memcpy_h2d(dz, zt.data(), T.G * 8));
        CK(qo_gen_zipf_u32((uint32_t *)T.grp, T.n, (const uint64_t *)dz, (uint32_t)T.G, seed + i, 2, seed + 7, nullptr));
        CK(qo_device_sync());
        CK(qo_free(dz));
        CK(qo_gen_uniform_u32((uint32_t *)T.val, T.n, seed + i, 3, 0, 1u << 16, nullptr));
        CK(qo_gen_uniform_u32((uint32_t *)T.filt, T.n, seed + i, 4, 0, 1000, nullptr));
    }
    CK(qo_device_sync());
    fprintf(stderr, "[serve_mix] tables ready\n");
    qo_server_config cfg{(uint64_t)200e9, 1, 8, 0, 60000.0};
    qo_server *sv = nullptr;
    CK(qo_server_create(&cfg, &sv));
    int ids[4];
    for (int i = 0; i < 4; ++i) {
        qo_table_desc d{t[i].key, t[i].fk, t[i].grp, t[i].val, t[i].filt, t[i].n, t[i].G};
        CK(qo_server_add_table(sv, &d, &ids[i]));
    }
    // The 24 requests of request_mix(ids, 24, seed=1): kind, a, b, hi (lo = INT_MIN).
    const int mix[24][4] = {{1, 3, -1, 477}, {0, 3, -1, 476}, {0, 3, -1, 240}, {2, 1, 2, 234}, {0, 2, -1, 439},
                            {1, 3, -1, 417}, {2, 1, 1, 105}, {1, 3, -1, 110}, {1, 3, -1, 141}, {0, 0, -1, 387},
                            {0, 1, -1, 268}, {2, 0, 2, 482}, {1, 2, -1, 466}, {0, 1, -1, 195}, {1, 3, -1, 181},
                            {1, 0, -1, 255}, {1, 3, -1, 397}, {0, 3, -1, 373}, {1, 2, -1, 215}, {1, 0, -1, 396},
                            {1, 3, -1, 411}, {2, 1, 2, 311}, {1, 3, -1, 388}, {0, 2, -1, 418}};
    const int nreq = argc > 2 ? atoi(argv[2]) : 24;
    // "canary": one batch of 8 per run, with a canary after every buffer carved
    // from the batch's arena, checked after each batch (the kernels of this
    // loop run as they are launched, so a workspace carved again by a later
    // step has retired the earlier canaries it covers).
    // "canary-part": the same with the partitioned join build and probe forced
    // for every table (tune operator 5, param 1).
    const bool canary = argc > 3 && !strncmp(argv[3], "canary", 6);
    if (argc > 3 && !strcmp(argv[3], "canary-part"))
        for (int cls = 0; cls <= 40; ++cls) {
            int b = 0, it = 0, pa = 0;
            CK(qo_tune_get(5, cls, &b, &it, &pa));
            CK(qo_tune_set(5, cls, b, it, 1));
        }
    std::vector<qo_request> reqs;
    long damaged = 0;
    for (int i = 0; i < nreq; ++i) {
        qo_request r{mix[i][0], ids[mix[i][1]], mix[i][2] < 0 ? -1 : ids[mix[i][2]], INT32_MIN, mix[i][3]};
        uint64_t id;
        CK(qo_server_submit(sv, &r, &id));
        reqs.push_back(r);
        if (canary && (i % 8 == 7 || i + 1 == nreq)) {
            qo_sim_canary_enable(1);
            CK(qo_server_run(sv));
            qo_sim_canary_enable(0);
            damaged += qo_sim_canary_check();
        }
    }
    if (canary) printf("canaries: %ld damaged\n", damaged);
    else CK(qo_server_run(sv));
    std::vector<qo_result> res(nreq);
    const int got = qo_server_poll(sv, res.data(), nreq, 0);
    fprintf(stderr, "[serve_mix] %d results\n", got);
    int bad = got == nreq ? 0 : 1;
    for (int i = 0; i < got; ++i) {
        const qo_result &r = res[i];
        qo_result alone{};
        CK(qo_server_run_plain(sv, &reqs[r.id - 1], &alone));
        const bool ok = r.status == 0 && r.count == alone.count && r.sum == alone.sum && r.checksum == alone.checksum;
        printf("req %2llu kind %d status %d count %llu sum %lld cs %lld | alone count %llu sum %lld cs %lld %s\n",
               (unsigned long long)r.id, r.kind, r.status, (unsigned long long)r.count, (long long)r.sum,
               (long long)r.checksum, (unsigned long long)alone.count, (long long)alone.sum,
               (long long)alone.checksum, ok ? "ok" : "MISMATCH");
        bad += !ok;
    }
    CK(qo_server_destroy(sv));
    bad += damaged != 0;
    printf("%s\n", bad ? "FAIL" : "PASS");
    return bad ? 1 : 0;
}
