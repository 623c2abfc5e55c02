// Device buffers and stream timing for the on-device measurements of the
// autotuner and the planner's calibration.
#pragma once
#include <algorithm>
#include <functional>
#include <vector>

#include "qo_internal.hpp"

namespace qo {

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        QO_HIP(hipMalloc(&p, bytes ? bytes : 256));
        return QO_OK;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

struct Timer {
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (st) (void)hipStreamDestroy(st);
    }
    int init() {
        QO_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        QO_HIP(hipEventCreate(&a));
        QO_HIP(hipEventCreate(&b));
        return QO_OK;
    }
    // Median ms of `reps` runs of fn(st) after one warm-up run.
    int median(int reps, const std::function<int(hipStream_t)> &fn, double *ms) {
        QO_TRY(fn(st));
        QO_TRY(sync_and_check(st));
        std::vector<double> t;
        for (int r = 0; r < reps; ++r) {
            QO_HIP(hipEventRecord(a, st));
            QO_TRY(fn(st));
            QO_HIP(hipEventRecord(b, st));
            QO_TRY(sync_and_check(st));
            float x = 0;
            QO_HIP(hipEventElapsedTime(&x, a, b));
            t.push_back(x);
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return QO_OK;
    }
};

}  // namespace qo
