// Split, scatter, gather, projection and the Record -> column step: the
// operators that move Records between the hot-path kernels.
//
// Reference paths replaced (their OpenCL kernels are absent from the
// reference; the host drivers and validators fix the results):
//   splitImpl        host/host/testSplit.cpp:156-209 (mapPart, countHist, scan,
//                    writeHist, getBound, scatterImpl_forPart), checked by
//                    validateSplit (host/host/Validate.cpp:510-527);
//   scatterImpl      host/host/testScatter.cpp:61-81 (optScatter over windows);
//   gatherImpl       host/host/testScatter.cpp:105-130 (optGather over windows);
//   projection       validateProjection (host/host/Validate.cpp:39-71).
// The reference splits one window [from, to) of target positions per launch
// to keep its scatter's writes cache-local on a GPU of small caches; the
// windows only partition the work, so one pass over all positions gives the
// same result.  Gather, scatter and projection are latency-bound random
// accesses: each thread keeps kItems independent accesses in flight.
#include "qo_device.hpp"
#include "qo_internal.hpp"

namespace qo {

constexpr int kItems = 8;  // elements per thread per grid-stride step

static unsigned move_grid(size_t n) {
    return grid_for((n + kItems - 1) / kItems, 256, (unsigned)props().cu_count * 32);
}

// S[pos] = R[loc[pos]] for loc[pos] in [0, rLen); other positions keep S.
__global__ __launch_bounds__(256) void k_gather_by_loc(const uint2 *__restrict__ R, size_t rLen,
                                                        const int32_t *__restrict__ loc, size_t sLen,
                                                        uint2 *__restrict__ S) {
    const size_t step = (size_t)gridDim.x * 256 * kItems;
    for (size_t base = (size_t)blockIdx.x * 256 * kItems + threadIdx.x; base < sLen; base += step) {
        int32_t l[kItems];
        uint2 v[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const size_t pos = base + (size_t)j * 256;
            l[j] = pos < sLen ? loc[pos] : -1;
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (l[j] >= 0 && (size_t)l[j] < rLen) v[j] = R[l[j]];
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (l[j] >= 0 && (size_t)l[j] < rLen) S[base + (size_t)j * 256] = v[j];
    }
}

// S[loc[i]] = R[i] for loc[i] in [0, rLen).  Repeated targets: one of the
// records aimed at a position lands there (as in the reference, no order).
__global__ __launch_bounds__(256) void k_scatter_by_loc(const uint2 *__restrict__ R, size_t rLen,
                                                         const int32_t *__restrict__ loc, uint2 *__restrict__ S) {
    const size_t step = (size_t)gridDim.x * 256 * kItems;
    for (size_t base = (size_t)blockIdx.x * 256 * kItems + threadIdx.x; base < rLen; base += step) {
        int32_t l[kItems];
        uint2 v[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const size_t i = base + (size_t)j * 256;
            const bool in = i < rLen;
            l[j] = in ? loc[i] : -1;
            v[j] = in ? ld_stream(R + i) : uint2{0, 0};
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (l[j] >= 0 && (size_t)l[j] < rLen) S[l[j]] = v[j];
    }
}

// out[i] = {proj[i].x, R[proj[i].x].y}; a row id beyond rLen keeps proj[i].y.
__global__ __launch_bounds__(256) void k_project_records(const uint2 *__restrict__ R, size_t rLen,
                                                         const uint2 *__restrict__ proj, size_t pLen,
                                                         uint2 *__restrict__ out) {
    const size_t step = (size_t)gridDim.x * 256 * kItems;
    for (size_t base = (size_t)blockIdx.x * 256 * kItems + threadIdx.x; base < pLen; base += step) {
        uint2 p[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const size_t i = base + (size_t)j * 256;
            p[j] = i < pLen ? ld_stream(proj + i) : uint2{0xFFFFFFFFu, 0};
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (p[j].x < rLen) p[j].y = R[p[j].x].y;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const size_t i = base + (size_t)j * 256;
            if (i < pLen) out[i] = p[j];
        }
    }
}

// x[i] = R[i].x, y[i] = R[i].y: 4 Records (two 16-byte loads) per thread and
// 16-byte column stores when all three arrays are 16-byte aligned.
__global__ __launch_bounds__(256) void k_unpack_records(const uint2 *__restrict__ R, size_t n,
                                                        uint32_t *__restrict__ x, uint32_t *__restrict__ y,
                                                        bool vec) {
    const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    if (vec) {
        const size_t n4 = n / 4;
        const uint4 *r4 = reinterpret_cast<const uint4 *>(R);
        for (size_t q = tid; q < n4; q += stride) {
            const uint4 a = ld_stream(r4 + 2 * q), b = ld_stream(r4 + 2 * q + 1);
            reinterpret_cast<uint4 *>(x)[q] = uint4{a.x, a.z, b.x, b.z};
            reinterpret_cast<uint4 *>(y)[q] = uint4{a.y, a.w, b.y, b.w};
        }
        for (size_t i = n4 * 4 + tid; i < n; i += stride) {
            x[i] = R[i].x;
            y[i] = R[i].y;
        }
    } else {
        for (size_t i = tid; i < n; i += stride) {
            const uint2 r = R[i];
            x[i] = r.x;
            y[i] = r.y;
        }
    }
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_split_records(const qo_record *rin, size_t n, uint32_t num_part, qo_record *rout, qo_record *d_bound,
                     qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(num_part >= 1, "numPart must be at least 1");
    QO_REQUIRE((n == 0 || (rin && rout)), "null pointer");
    QO_REQUIRE(n == 0 || ((uintptr_t)rin % 8 == 0 && (uintptr_t)rout % 8 == 0), "Records must be 8-byte aligned");
    QO_REQUIRE(rin != rout || n == 0, "split is out of place: rin and rout must differ");
    hipStream_t st = stream_of(s);
    int status = QO_OK;
    void *ws = workspace(st, split_records_ws_bytes(n, num_part), &status);
    if (!ws) return status;
    return launch_split_records(reinterpret_cast<const uint64_t *>(rin), n, num_part,
                                reinterpret_cast<uint64_t *>(rout), reinterpret_cast<uint2 *>(d_bound), ws, st);
}

int qo_gather_records(const qo_record *R, size_t r_len, const int32_t *loc, size_t s_len, qo_record *S,
                      qo_stream s) {
    QO_ENTER();
    if (s_len == 0) return QO_OK;
    QO_REQUIRE(loc && S && (R || r_len == 0), "null pointer");
    QO_REQUIRE(R != S || r_len == 0, "gather is out of place: R and S must differ");
    k_gather_by_loc<<<move_grid(s_len), 256, 0, stream_of(s)>>>(reinterpret_cast<const uint2 *>(R), r_len, loc,
                                                                  s_len, reinterpret_cast<uint2 *>(S));
    QO_LAUNCHED();
    return QO_OK;
}

int qo_scatter_records(const qo_record *R, size_t r_len, const int32_t *loc, qo_record *S, qo_stream s) {
    QO_ENTER();
    if (r_len == 0) return QO_OK;
    QO_REQUIRE(R && loc && S, "null pointer");
    QO_REQUIRE(R != S, "scatter is out of place: R and S must differ");
    k_scatter_by_loc<<<move_grid(r_len), 256, 0, stream_of(s)>>>(reinterpret_cast<const uint2 *>(R), r_len, loc,
                                                                   reinterpret_cast<uint2 *>(S));
    QO_LAUNCHED();
    return QO_OK;
}

int qo_project_records(const qo_record *R, size_t r_len, const qo_record *proj, size_t p_len, qo_record *out,
                       qo_stream s) {
    QO_ENTER();
    if (p_len == 0) return QO_OK;
    QO_REQUIRE(proj && out && (R || r_len == 0), "null pointer");
    QO_REQUIRE(out != R || r_len == 0, "projection writes out while reading R: out and R must differ");
    k_project_records<<<move_grid(p_len), 256, 0, stream_of(s)>>>(
        reinterpret_cast<const uint2 *>(R), r_len, reinterpret_cast<const uint2 *>(proj), p_len,
        reinterpret_cast<uint2 *>(out));
    QO_LAUNCHED();
    return QO_OK;
}

int qo_unpack_records(const qo_record *R, size_t n, uint32_t *x, uint32_t *y, qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(R && x && y, "null pointer");
    const bool vec = ((uintptr_t)R % 16 == 0) && ((uintptr_t)x % 16 == 0) && ((uintptr_t)y % 16 == 0);
    k_unpack_records<<<grid_for((n + 3) / 4, 256, (unsigned)props().cu_count * 16), 256, 0, stream_of(s)>>>(
        reinterpret_cast<const uint2 *>(R), n, x, y, vec);
    QO_LAUNCHED();
    return QO_OK;
}

}  // extern "C"
