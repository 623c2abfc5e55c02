// CPU build driver: host-side runtime checks that need no GPU.
//  1. Per-thread state of the NULL stream: two host threads that use the
//     library's per-thread stream get distinct workspaces (and a thread gets
//     its own back), while an explicit stream keeps one workspace.
//  2. The bounded look-back wait: a tile whose predecessor never publishes
//     gives up after the spin limit, sets the device error word, and the next
//     synchronising call reports QO_ERR_DEVICE_WAIT (then the word is clear).
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>

#include "qo.h"
#include "qo_device.hpp"
#include "qo_internal.hpp"

static int fails = 0;
#define EXPECT(c)                                                              \
    do {                                                                       \
        if (!(c)) {                                                            \
            fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
            ++fails;                                                           \
        }                                                                      \
    } while (0)

int main() {
    EXPECT(qo_init(0) == QO_OK);
    int st = 0;
    void *a1 = nullptr, *a2 = nullptr, *b1 = nullptr;
    // Both threads are alive while they take their workspaces (a thread id is
    // only unique among running threads).
    std::atomic<int> arrived{0};
    auto meet = [&] {
        ++arrived;
        while (arrived.load() < 2) std::this_thread::yield();
    };
    std::thread ta([&] {
        meet();
        a1 = qo::workspace(hipStreamPerThread, 1 << 16, &st);
        a2 = qo::workspace(hipStreamPerThread, 1 << 12, &st);
        ++arrived;
        while (arrived.load() < 4) std::this_thread::yield();
    });
    std::thread tb([&] {
        meet();
        b1 = qo::workspace(hipStreamPerThread, 1 << 16, &st);
        ++arrived;
        while (arrived.load() < 4) std::this_thread::yield();
    });
    ta.join();
    tb.join();
    EXPECT(a1 && b1 && a1 == a2);
    EXPECT(a1 != b1);
    qo_stream s = nullptr;
    EXPECT(qo_stream_create(&s) == QO_OK);
    void *e1 = nullptr, *e2 = nullptr;
    std::thread tc([&] { e1 = qo::workspace(static_cast<hipStream_t>(s), 1 << 12, &st); });
    tc.join();
    std::thread td([&] { e2 = qo::workspace(static_cast<hipStream_t>(s), 1 << 12, &st); });
    td.join();
    EXPECT(e1 && e1 == e2);
    EXPECT(qo_stream_destroy(s) == QO_OK);

    // Bounded look-back: tile 1 waits for tile 0, which never publishes.
    uint64_t *status = nullptr;
    uint64_t *out = nullptr;
    EXPECT(qo_malloc(reinterpret_cast<void **>(&status), 2 * sizeof(uint64_t)) == QO_OK);
    EXPECT(qo_malloc(reinterpret_cast<void **>(&out), sizeof(uint64_t)) == QO_OK);
    EXPECT(qo_fill_u32(status, 0, 4, nullptr) == QO_OK);
    EXPECT(qo_debug_set_spin_limit(1000) == QO_OK);
    qo::DevCtl *ctl = qo::dev_ctl();
    ::simt::launch(::simt::cfg(1, 64), "lookback_probe", [=]() {
        const uint64_t e = qo::lookback(status, 1, 5, ctl);
        if (threadIdx.x == 0) *out = e;
    });
    EXPECT(qo_device_sync() == QO_ERR_DEVICE_WAIT);
    EXPECT(qo_device_sync() == QO_OK);  // the error word was cleared
    uint64_t h = 1;
    EXPECT(qo_memcpy_d2h(&h, out, 8) == QO_OK);
    EXPECT(h == 0);
    // With tile 0 published, the same wait returns its prefix.
    EXPECT(qo_debug_set_spin_limit(1000) == QO_OK);
    ::simt::launch(::simt::cfg(1, 64), "lookback_tile0", [=]() { (void)qo::lookback(status, 0, 7, ctl); });
    ::simt::launch(::simt::cfg(1, 64), "lookback_probe", [=]() {
        const uint64_t e = qo::lookback(status, 1, 5, ctl);
        if (threadIdx.x == 0) *out = e;
    });
    EXPECT(qo_device_sync() == QO_OK);
    EXPECT(qo_memcpy_d2h(&h, out, 8) == QO_OK);
    EXPECT(h == 7);
    qo_free(status);
    qo_free(out);
    printf("%s\n", fails ? "FAIL" : "PASS");
    return fails ? 1 : 0;
}
