// Counter-based deterministic generator used for synthetic tables, callable
// on host and device.  Value i of stream (seed, stream_id) is the i-th output
// of a splitmix64 sequence (Steele, Lea, Flood 2014) whose state starts at
// mix64(seed + gamma*(stream_id+1)); every element is computed independently.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define QO_HD __host__ __device__ __forceinline__

namespace qo {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

QO_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

QO_HD uint64_t stream_key(uint64_t seed, uint64_t stream_id) { return mix64(seed + kGamma * (stream_id + 1)); }

QO_HD uint64_t rand_at(uint64_t key, uint64_t i) { return mix64(key + kGamma * (i + 1)); }

// floor(hi32(r) * range / 2^32)
QO_HD uint32_t bounded(uint64_t r, uint32_t range) {
    return static_cast<uint32_t>(((r >> 32) * static_cast<uint64_t>(range)) >> 32);
}

// Keyed bijection on [0, 2^bits): three rounds of (odd multiply, xorshift, add).
struct Bij {
    uint64_t mul[3];
    uint64_t add[3];
    uint64_t mask;
    uint32_t shift;
};

QO_HD Bij make_bij(uint32_t bits, uint64_t seed) {
    Bij b;
    b.mask = bits >= 32 ? 0xFFFFFFFFull : ((1ull << bits) - 1);
    b.shift = bits / 2 + 1;
    for (int r = 0; r < 3; ++r) {
        b.mul[r] = mix64(seed + 2 * r + 1) | 1ull;
        b.add[r] = mix64(seed + 2 * r + 2);
    }
    return b;
}

QO_HD uint32_t bij_apply(const Bij &b, uint32_t x) {
    uint64_t v = x;
    for (int r = 0; r < 3; ++r) {
        v = (v * b.mul[r]) & b.mask;
        v ^= v >> b.shift;
        v = (v + b.add[r]) & b.mask;
    }
    return static_cast<uint32_t>(v);
}

QO_HD uint32_t bits_for(uint64_t n) {
    uint32_t b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

// Element i of a seeded permutation of [0, n): cycle-walk the bijection on
// [0, 2^bits_for(n)) until the value lands below n.
QO_HD uint32_t perm_apply(const Bij &b, uint32_t i, uint64_t n) {
    uint32_t y = bij_apply(b, i);
    while (y >= n) y = bij_apply(b, y);
    return y;
}

}  // namespace qo
