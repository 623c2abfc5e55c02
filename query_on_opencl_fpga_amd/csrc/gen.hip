// Deterministic synthetic-table generators (device side of the generator spec
// in qo_rng.hpp).  Used to build resident inputs without host round trips.
#include <cmath>
#include <vector>

#include "qo_internal.hpp"
#include "qo_rng.hpp"

namespace qo {

__global__ void k_gen_uniform(uint32_t *out, size_t n, uint64_t key, uint32_t lo, uint32_t range) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint64_t r = rand_at(key, i);
        out[i] = lo + (range ? bounded(r, range) : static_cast<uint32_t>(r >> 32));
    }
}

__global__ void k_gen_perm(uint32_t *out, size_t n, Bij b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = perm_apply(b, static_cast<uint32_t>(i), n);
}

__global__ void k_gen_bij(uint32_t *out, size_t n, Bij b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = bij_apply(b, static_cast<uint32_t>(i));
}

__global__ void k_gen_fk(uint32_t *out, size_t n, uint64_t key, uint64_t nr, bool use_perm, Bij b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint32_t u = bounded(rand_at(key, i), static_cast<uint32_t>(nr));
        out[i] = use_perm ? perm_apply(b, u, nr) : bij_apply(b, u);
    }
}

__global__ void k_gen_zipf(uint32_t *out, size_t n, const uint64_t *__restrict__ t, uint32_t G, uint64_t key,
                           Bij b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint64_t u = rand_at(key, i) >> 11;
        uint32_t lo = 0, hi = G - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (u < t[mid]) hi = mid;
            else lo = mid + 1;
        }
        out[i] = perm_apply(b, lo, G);
    }
}

// out[i] = {x[i], y[i]}; with all three arrays 16-byte aligned, 4 Records per
// thread: one 16-byte load per column, two 16-byte stores.
__global__ void k_pack_records(const uint32_t *x, const uint32_t *y, size_t n, qo_record *out, bool vec) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t from = 0;
    if (vec) {
        const size_t n4 = n / 4;
        for (size_t q = tid; q < n4; q += stride) {
            const uint4 a = reinterpret_cast<const uint4 *>(x)[q], b = reinterpret_cast<const uint4 *>(y)[q];
            uint4 *o = reinterpret_cast<uint4 *>(out) + 2 * q;
            o[0] = uint4{a.x, b.x, a.y, b.y};
            o[1] = uint4{a.z, b.z, a.w, b.w};
        }
        from = n4 * 4;
    }
    for (size_t i = from + tid; i < n; i += stride) out[i] = qo_record{x[i], y[i]};
}

static unsigned gen_grid(size_t n) { return grid_for(n, 256, (unsigned)props().cu_count * 16); }

}  // namespace qo

using namespace qo;

extern "C" {

int qo_gen_uniform_u32(uint32_t *out, size_t n, uint64_t seed, uint64_t stream_id, uint32_t lo, uint32_t range,
                       qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(out, "null output");
    k_gen_uniform<<<gen_grid(n), 256, 0, stream_of(s)>>>(out, n, stream_key(seed, stream_id), lo, range);
    QO_LAUNCHED();
    return QO_OK;
}

int qo_gen_perm_u32(uint32_t *out, size_t n, uint64_t perm_seed, qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(out && n <= (size_t(1) << 32), "bad permutation arguments");
    k_gen_perm<<<gen_grid(n), 256, 0, stream_of(s)>>>(out, n, make_bij(bits_for(n), perm_seed));
    QO_LAUNCHED();
    return QO_OK;
}

int qo_gen_bij_u32(uint32_t *out, size_t n, uint32_t bits, uint64_t seed, qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(out && bits >= 1 && bits <= 32 && (bits == 32 || n <= (size_t(1) << bits)),
               "bad bijection arguments");
    k_gen_bij<<<gen_grid(n), 256, 0, stream_of(s)>>>(out, n, make_bij(bits, seed));
    QO_LAUNCHED();
    return QO_OK;
}

int qo_gen_fk_u32(uint32_t *out, size_t n, uint64_t seed, uint64_t stream_id, uint64_t nr, uint32_t bits,
                  uint64_t key_seed, qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(out && nr >= 1 && nr <= 0xFFFFFFFFull && bits <= 32, "bad foreign-key arguments");
    bool use_perm = bits == 0;
    Bij b = make_bij(use_perm ? bits_for(nr) : bits, key_seed);
    k_gen_fk<<<gen_grid(n), 256, 0, stream_of(s)>>>(out, n, stream_key(seed, stream_id), nr, use_perm, b);
    QO_LAUNCHED();
    return QO_OK;
}

int qo_zipf_table(uint64_t *t, uint32_t G, double z) {
    QO_REQUIRE(t && G >= 1, "bad zipf arguments");
    double h = 0.0;
    for (uint32_t r = 0; r < G; ++r) h += pow((double)r + 1.0, -z);
    double c = 0.0;
    for (uint32_t r = 0; r < G; ++r) {
        c += pow((double)r + 1.0, -z);
        double f = c / h;
        if (f > 1.0) f = 1.0;
        t[r] = (uint64_t)(f * 9007199254740992.0);
    }
    t[G - 1] = 9007199254740992ull;
    return QO_OK;
}

int qo_gen_zipf_u32(uint32_t *out, size_t n, const uint64_t *d_table, uint32_t G, uint64_t seed,
                    uint64_t stream_id, uint64_t perm_seed, qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(out && d_table && G >= 1, "bad zipf arguments");
    k_gen_zipf<<<gen_grid(n), 256, 0, stream_of(s)>>>(out, n, d_table, G, stream_key(seed, stream_id),
                                                      make_bij(bits_for(G), perm_seed));
    QO_LAUNCHED();
    return QO_OK;
}

int qo_pack_records(const uint32_t *x, const uint32_t *y, size_t n, qo_record *out, qo_stream s) {
    QO_ENTER();
    if (n == 0) return QO_OK;
    QO_REQUIRE(x && y && out, "null pointer");
    const bool vec = ((uintptr_t)x % 16 == 0) && ((uintptr_t)y % 16 == 0) && ((uintptr_t)out % 16 == 0);
    k_pack_records<<<gen_grid(n), 256, 0, stream_of(s)>>>(x, y, n, out, vec);
    QO_LAUNCHED();
    return QO_OK;
}

}  // extern "C"
