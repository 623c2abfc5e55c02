// Reduce without grouping over R.y: SUM, MAX, MIN, AVERAGE.
//
// Reference: the results validateReduce checks (host/host/Validate.cpp:
// 372-468; operator codes host/host/common.h:89-92).  Record is uint2, so .y
// is unsigned: SUM wraps mod 2^32, MAX and MIN compare unsigned (the int
// accumulator there meets the uint .y in an unsigned comparison, and starting
// from element 0 equals starting from 0 / 0xFFFFFFFF for n >= 1), AVERAGE is
// the wrapped sum divided by n.  Integer +, max and min are associative and
// commutative, so any combining order gives those bits exactly.
//
// One streaming pass: each block folds a grid-stride share of the Records
// (16-byte loads of two Records, 8 in flight per thread, as many blocks per
// CU as the filter-aggregate stream takes), writes one partial, and the last
// block to finish (a ticket counter) folds the partials and writes the result.
#include <algorithm>

#include "qo_device.hpp"
#include "qo_internal.hpp"

namespace qo {

constexpr unsigned kReduceBlocks = 2048;  // partial slots in the workspace: at most 8 blocks per CU

template <int OP>
QO_DEV uint32_t fold(uint32_t a, uint32_t b) {
    if constexpr (OP == 1) return a > b ? a : b;
    else if constexpr (OP == 2) return a < b ? a : b;
    else return a + b;  // SUM and AVERAGE
}
template <int OP>
constexpr uint32_t fold_identity() {
    return OP == 2 ? 0xFFFFFFFFu : 0u;
}

template <int OP>
QO_DEV uint32_t block_fold(uint32_t v, uint32_t *s_w) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fold<OP>(v, (uint32_t)__shfl_xor((int)v, o));
    if (lane_id() == 0) s_w[wave_id()] = v;
    __syncthreads();
    v = fold_identity<OP>();
    for (unsigned w = 0; w < blockDim.x / kWave; ++w) v = fold<OP>(v, s_w[w]);
    return v;
}

// UNROLL 16-byte loads (two Records each) in flight per thread before any is
// folded; grid-stride over the 16-byte groups.
template <int OP, int UNROLL>
__global__ __launch_bounds__(256) void k_reduce_records(const uint2 *__restrict__ R, size_t n, bool vec,
                                                        uint32_t *__restrict__ partial, uint32_t *ticket,
                                                        uint32_t *__restrict__ result) {
    __shared__ uint32_t s_w[4];
    __shared__ bool s_last;
    uint32_t acc = fold_identity<OP>();
    const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    if (vec) {
        const uint4 *r4 = reinterpret_cast<const uint4 *>(R);
        const size_t n2 = n / 2;
        size_t q = tid;
        for (; q + (UNROLL - 1) * stride < n2; q += UNROLL * stride) {
            uint4 v[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) v[u] = ld_stream(r4 + q + u * stride);
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) acc = fold<OP>(acc, fold<OP>(v[u].y, v[u].w));
        }
        for (; q < n2; q += stride) {
            const uint4 a = ld_stream(r4 + q);
            acc = fold<OP>(acc, fold<OP>(a.y, a.w));
        }
        if (tid == 0 && (n & 1)) acc = fold<OP>(acc, R[n - 1].y);
    } else {
        for (size_t i = tid; i < n; i += stride) acc = fold<OP>(acc, R[i].y);
    }
    acc = block_fold<OP>(acc, s_w);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = acc;
        __threadfence();
        s_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    // Last block: every other block's partial is visible (each fenced before
    // taking its ticket).
    __threadfence();
    uint32_t v = fold_identity<OP>();
    for (unsigned b = threadIdx.x; b < gridDim.x; b += 256)
        v = fold<OP>(v, __hip_atomic_load(partial + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    __syncthreads();  // s_w reuse
    v = block_fold<OP>(v, s_w);
    if (threadIdx.x == 0) *result = OP == 3 ? (uint32_t)((uint64_t)v / n) : v;
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_reduce_records(const qo_record *rin, size_t n, int op, uint32_t *d_result, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(op >= 0 && op <= 3, "reduce operator %d is not SUM (0), MAX (1), MIN (2) or AVERAGE (3)", op);
    QO_REQUIRE(d_result, "null result pointer");
    hipStream_t st = stream_of(s);
    if (n == 0) return launch_fill_u32(d_result, 0, 1, st);
    QO_REQUIRE(rin && (uintptr_t)rin % 8 == 0, "Records must be 8-byte aligned device memory");
    // The streaming shape of the filter-aggregate kernel (tune operator 0: blocks
    // per CU), 8 loads in flight per thread.
    const TuneParams tp = tune_get(TUNE_FILTER_AGG, n);
    const unsigned per_cu = (unsigned)(tp.param > 0 ? tp.param : 1);
    const unsigned cap = std::min<unsigned>(kReduceBlocks, (unsigned)props().cu_count * per_cu);
    const unsigned grid = grid_for((n + 15) / 16, 256, cap);
    int status = QO_OK;
    void *ws = workspace(st, (size_t)kReduceBlocks * 4 + 256, &status);
    if (!ws) return status;
    uint32_t *ticket = reinterpret_cast<uint32_t *>(ws);
    uint32_t *partial = ticket + 64;
    QO_TRY(launch_fill_u32(ticket, 0, 1, st));
    const uint2 *R = reinterpret_cast<const uint2 *>(rin);
    const bool vec = (uintptr_t)rin % 16 == 0;
    switch (op) {
        case 0: k_reduce_records<0, 8><<<grid, 256, 0, st>>>(R, n, vec, partial, ticket, d_result); break;
        case 1: k_reduce_records<1, 8><<<grid, 256, 0, st>>>(R, n, vec, partial, ticket, d_result); break;
        case 2: k_reduce_records<2, 8><<<grid, 256, 0, st>>>(R, n, vec, partial, ticket, d_result); break;
        default: k_reduce_records<3, 8><<<grid, 256, 0, st>>>(R, n, vec, partial, ticket, d_result); break;
    }
    QO_LAUNCHED();
    return QO_OK;
}

}  // extern "C"
