// CPU build driver: the key-range join table built by the partitioned build
// (R split by slot range, every slice assembled in LDS and written out whole)
// against the same table built by storing every tuple at its slot (the
// partitioned build switched off with tune operator 5, param 2): identical
// slots, mode, capacity, occupied and sentinel counts.  Cases: a permutation
// of [0, n); keys from a negative base with (-1, -1) tuples among them (counted,
// not stored); keys spread over twice their count (empty slots between); and
// one duplicate key, which both builds must turn into a hash-mixed table.
//
//     range_build [log2 n]
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

static void set_all(int op, int param) {
    for (int cls = 0; cls <= 40; ++cls) {
        int b = 0, it = 0, pa = 0;
        CK(qo_tune_get(op, cls, &b, &it, &pa));
        CK(qo_tune_set(op, cls, b, it, param));
    }
}

// Header words compared (the 256-byte header, then the slots).
struct Head {
    uint64_t cap;
    uint32_t has_dups, sentinel_count, overflow, mode;
};

static std::vector<uint64_t> build(const std::vector<int32_t> &k, const std::vector<int32_t> &v, int part_mode,
                                   Head *h) {
    const size_t n = k.size(), tb = qo_join_table_bytes(n);
    void *dk = nullptr, *dv = nullptr, *table = nullptr;
    CK(qo_malloc(&dk, n * 4));
    CK(qo_malloc(&dv, n * 4));
    CK(qo_malloc(&table, tb));
    CK(qo_memcpy_h2d(dk, k.data(), n * 4));
    CK(qo_memcpy_h2d(dv, v.data(), n * 4));
    set_all(5, part_mode);
    CK(qo_hash_build(table, tb, static_cast<int32_t *>(dk), static_cast<int32_t *>(dv), n, nullptr));
    CK(qo_device_sync());
    std::vector<uint64_t> all(tb / 8);
    CK(qo_memcpy_d2h(all.data(), table, tb));
    memcpy(h, all.data(), sizeof(Head));
    for (void *p : {dk, dv, table}) CK(qo_free(p));
    // Key-range tables use slots [0, cap); a hash-mixed one every slot.
    return std::vector<uint64_t>(all.begin() + 32, all.begin() + 32 + h->cap);
}

int main(int argc, char **argv) {
    const int lg = argc > 1 ? atoi(argv[1]) : 20;
    const size_t n = size_t(1) << lg;
    CK(qo_init(0));
    uint64_t x = 88172645463325252ull;
    auto rnd = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    int fails = 0;
    const char *names[] = {"permutation", "negative base with (-1, -1) tuples", "keys over twice their count",
                           "one duplicate key"};
    for (int c = 0; c < 4; ++c) {
        std::vector<int32_t> k(n), v(n);
        for (size_t i = 0; i < n; ++i) k[i] = (int32_t)i;
        for (size_t i = n - 1; i > 0; --i) std::swap(k[i], k[rnd() % (i + 1)]);
        for (size_t i = 0; i < n; ++i) v[i] = (int32_t)(uint32_t)rnd();
        if (c == 1)
            for (size_t i = 0; i < n; ++i) {
                k[i] -= 1000;
                if (k[i] == -1) v[i] = -1;  // the one tuple whose packed form is the empty slot
            }
        if (c == 2)
            for (size_t i = 0; i < n; ++i) k[i] *= 2;
        if (c == 3) k[n / 3] = k[n / 2];
        Head hs, hd;
        const std::vector<uint64_t> sliced = build(k, v, 0, &hs), direct = build(k, v, 2, &hd);
        bool ok = hs.mode == hd.mode && hs.cap == hd.cap && hs.sentinel_count == hd.sentinel_count &&
                  hs.has_dups == hd.has_dups && hs.overflow == hd.overflow;
        if (c < 3) ok = ok && hs.mode == 1 && sliced == direct;  // key-range: identical slots
        else ok = ok && hs.mode == 0;                            // duplicate: hash-mixed fallback
        printf("%s: mode %u cap %llu sentinels %u, %s\n", names[c], hs.mode, (unsigned long long)hs.cap,
               hs.sentinel_count, ok ? "same table" : "DIFFERENT");
        fails += !ok;
    }
    set_all(5, 0);
    CK(qo_release_workspaces());
    printf(fails ? "FAIL\n" : "PASS\n");
    return fails ? 1 : 0;
}
