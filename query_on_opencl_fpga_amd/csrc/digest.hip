// Order-sensitive digest of materialised join rows, computed on the device:
// lets a caller compare two join outputs of any size (a batched query with
// its plain run, a served request with its lone run) without copying them
// to the host.
//
// digest[0] = sum over rows i of mix(key, r_payload, s_payload, i)   (order-sensitive)
// digest[1] = sum over rows of key as int64 (two's complement, mod 2^64)
// digest[2] = rows i > 0 that order before row i - 1 by (key, s_payload)
//             (0 for an output sorted as hj_filter_sort sorts it)
#include "qo_internal.hpp"
#include "qo_device.hpp"
#include "qo_rng.hpp"

namespace qo {

QO_DEV uint64_t row_digest_term(int32_t key, int32_t rpay, int32_t spay, uint64_t i) {
    const uint64_t a = ((uint64_t)(uint32_t)key << 32) | (uint32_t)rpay;
    return mix64(mix64(a + kGamma * (i + 1)) ^ (uint32_t)spay);
}

__global__ __launch_bounds__(256) void k_rows_digest(const int32_t *__restrict__ key, const int32_t *__restrict__ rpay,
                                                     const int32_t *__restrict__ spay, const uint64_t *d_n,
                                                     uint64_t n_host, unsigned long long *digest) {
    uint64_t n = d_n ? *d_n : n_host;
    if (n > n_host) n = n_host;
    uint64_t cs = 0, sum = 0, bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t k = key[i], s = spay[i];
        cs += row_digest_term(k, rpay[i], s, i);
        sum += (uint64_t)(int64_t)k;
        if (i > 0) {
            const int32_t pk = key[i - 1], ps = spay[i - 1];
            bad += (k < pk || (k == pk && s < ps)) ? 1u : 0u;
        }
    }
    cs = wave_sum(cs);
    sum = wave_sum(sum);
    bad = wave_sum(bad);
    if (lane_id() == 0) {
        atomicAdd(&digest[0], (unsigned long long)cs);
        atomicAdd(&digest[1], (unsigned long long)sum);
        atomicAdd(&digest[2], (unsigned long long)bad);
    }
}

}  // namespace qo

using namespace qo;

extern "C" {

int qo_join_rows_digest(const int32_t *key, const int32_t *rpay, const int32_t *spay, size_t m,
                        const uint64_t *d_count, uint64_t *d_digest, qo_stream s) {
    QO_ENTER();
    QO_REQUIRE(d_digest && (m == 0 || (key && rpay && spay)), "null pointer");
    hipStream_t st = stream_of(s);
    QO_TRY(launch_fill_u64(d_digest, 0, 3, st));
    if (m == 0) return QO_OK;
    k_rows_digest<<<grid_for(m, 256 * 8, (unsigned)props().cu_count * 8), 256, 0, st>>>(
        key, rpay, spay, d_count, m, reinterpret_cast<unsigned long long *>(d_digest));
    QO_LAUNCHED();
    return QO_OK;
}

}  // extern "C"
