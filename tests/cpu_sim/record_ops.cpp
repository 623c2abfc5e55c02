// CPU build driver: split, gather, scatter, projection, reduce and
// Records -> columns on one input, in the order and shape of bench.py's
// operator rows (.x a permutation of [0, n), .y uniform 32-bit), each checked
// against a plain loop.  Under the ASan build every device access of those
// kernels is bounds-checked.
//
//     record_ops <log2 n>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qo.h"

static int fails = 0;
#define EXPECT(c)                                                                       \
    do {                                                                                \
        if (!(c)) {                                                                     \
            fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
            ++fails;                                                                    \
        }                                                                               \
    } while (0)
#define OK(call) EXPECT((call) == QO_OK)

int main(int argc, char **argv) {
    const int lg = argc > 1 ? atoi(argv[1]) : 16;
    const size_t n = size_t(1) << lg;
    OK(qo_init(0));
    void *x = nullptr, *y = nullptr, *src = nullptr, *dst = nullptr, *bound = nullptr, *res = nullptr;
    void *cx = nullptr, *cy = nullptr;
    OK(qo_malloc(&x, n * 4));
    OK(qo_malloc(&y, n * 4));
    OK(qo_malloc(&src, n * 8));
    OK(qo_malloc(&dst, n * 8));
    OK(qo_malloc(&bound, 65536 * 8));
    OK(qo_malloc(&res, 8));
    OK(qo_malloc(&cx, n * 4));
    OK(qo_malloc(&cy, n * 4));
    OK(qo_gen_perm_u32(static_cast<uint32_t *>(x), n, 7, nullptr));
    OK(qo_gen_uniform_u32(static_cast<uint32_t *>(y), n, 11, 4, 0, 0, nullptr));
    OK(qo_pack_records(static_cast<uint32_t *>(x), static_cast<uint32_t *>(y), n, static_cast<qo_record *>(src),
                       nullptr));
    OK(qo_stream_sync(nullptr));
    const qo_record *R = static_cast<const qo_record *>(src);
    qo_record *D = static_cast<qo_record *>(dst);
    const int32_t *loc = static_cast<const int32_t *>(x);
    // Split into 32 and 1000 partitions: stable by .y % parts, bounds by partition.
    for (uint32_t parts : {32u, 1000u}) {
        OK(qo_split_records(R, n, parts, D, static_cast<qo_record *>(bound), nullptr));
        OK(qo_stream_sync(nullptr));
        std::vector<size_t> start(parts + 1, 0);
        for (size_t i = 0; i < n; ++i) start[R[i].y % parts + 1]++;
        for (uint32_t p = 0; p < parts; ++p) start[p + 1] += start[p];
        const qo_record *B = static_cast<const qo_record *>(bound);
        for (uint32_t p = 0; p < parts; ++p) EXPECT(B[p].x == start[p] && B[p].y == start[p + 1]);
        std::vector<size_t> cur(start.begin(), start.end() - 1);
        size_t bad = 0;
        for (size_t i = 0; i < n; ++i) {
            const size_t at = cur[R[i].y % parts]++;
            bad += D[at].x != R[i].x || D[at].y != R[i].y;
        }
        EXPECT(bad == 0);
        printf("split %u partitions: %zu mismatches\n", parts, bad);
    }
    // Gather, scatter and projection with the permutation as locations / row ids.
    OK(qo_gather_records(R, n, loc, n, D, nullptr));
    OK(qo_stream_sync(nullptr));
    size_t bad = 0;
    for (size_t i = 0; i < n; ++i) bad += memcmp(&D[i], &R[loc[i]], 8) != 0;
    EXPECT(bad == 0);
    OK(qo_scatter_records(R, n, loc, D, nullptr));
    OK(qo_stream_sync(nullptr));
    bad = 0;
    for (size_t i = 0; i < n; ++i) bad += memcmp(&D[loc[i]], &R[i], 8) != 0;
    EXPECT(bad == 0);
    OK(qo_project_records(R, n, R, n, D, nullptr));
    OK(qo_stream_sync(nullptr));
    bad = 0;
    for (size_t i = 0; i < n; ++i) bad += D[i].x != R[i].x || D[i].y != R[R[i].x].y;
    EXPECT(bad == 0);
    // Reduce (SUM) and Records -> columns.
    OK(qo_reduce_records(R, n, 0, static_cast<uint32_t *>(res), nullptr));
    OK(qo_stream_sync(nullptr));
    uint32_t sum = 0;
    for (size_t i = 0; i < n; ++i) sum += R[i].y;
    EXPECT(*static_cast<uint32_t *>(res) == sum);
    OK(qo_unpack_records(R, n, static_cast<uint32_t *>(cx), static_cast<uint32_t *>(cy), nullptr));
    OK(qo_stream_sync(nullptr));
    bad = 0;
    for (size_t i = 0; i < n; ++i)
        bad += static_cast<uint32_t *>(cx)[i] != R[i].x || static_cast<uint32_t *>(cy)[i] != R[i].y;
    EXPECT(bad == 0);
    for (void *p : {x, y, src, dst, bound, res, cx, cy}) qo_free(p);
    printf(fails ? "FAIL\n" : "PASS\n");
    return fails ? 1 : 0;
}
