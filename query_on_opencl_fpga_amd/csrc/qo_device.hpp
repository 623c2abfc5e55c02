// Device-side primitives for gfx950 (wave64): lane masks, wave scans and
// reductions, block-level striped offsets and a decoupled look-back across
// tiles handed out by a ticket counter.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define QO_DEV __device__ __forceinline__

namespace qo {

constexpr int kWave = 64;

QO_DEV unsigned lane_id() { return __lane_id(); }
QO_DEV unsigned wave_id() { return threadIdx.x / kWave; }
QO_DEV uint64_t lanemask_lt() { return (uint64_t(1) << lane_id()) - 1; }
QO_DEV uint64_t ballot(bool p) { return __ballot(p); }
QO_DEV unsigned popc(uint64_t m) { return __popcll(m); }

// Every lane of the wave finishes the statements before this point before any
// lane runs the ones after it.  The lanes of a wave run in lockstep, so this
// emits no instruction; it keeps the compiler from moving memory operations
// across it and marks code that relies on the lockstep (a lane reading a word
// that another lane of its wave writes next).
QO_DEV void wave_lockstep() { __builtin_amdgcn_wave_barrier(); }

template <class T>
QO_DEV T wave_incl_scan(T v) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T t = __shfl_up(v, o, kWave);
        if (lane_id() >= (unsigned)o) v += t;
    }
    return v;
}

template <class T>
QO_DEV T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <class T>
QO_DEV T wave_min(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const T t = __shfl_xor(v, o, kWave);
        v = t < v ? t : v;
    }
    return v;
}

template <class T>
QO_DEV T wave_max(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const T t = __shfl_xor(v, o, kWave);
        v = t > v ? t : v;
    }
    return v;
}

// Device-side switch for kernels whose use is decided on the device: a kernel
// given a gate with a mode word runs only when *mode == run_if.
struct Gate {
    const uint32_t *mode = nullptr;
    uint32_t run_if = 0;
    QO_DEV bool off() const { return mode && *mode != run_if; }
};

// Streaming loads that should not displace cache-resident data.  The HIP vector
// structs are loaded through the equivalent native clang vector types.
typedef int32_t nv4i __attribute__((ext_vector_type(4)));
typedef uint32_t nv4u __attribute__((ext_vector_type(4)));
typedef uint32_t nv2u __attribute__((ext_vector_type(2)));

template <class T>
QO_DEV T ld_stream(const T *p) { return __builtin_nontemporal_load(p); }
QO_DEV int4 ld_stream(const int4 *p) {
    nv4i v = __builtin_nontemporal_load(reinterpret_cast<const nv4i *>(p));
    return int4{v.x, v.y, v.z, v.w};
}
QO_DEV uint4 ld_stream(const uint4 *p) {
    nv4u v = __builtin_nontemporal_load(reinterpret_cast<const nv4u *>(p));
    return uint4{v.x, v.y, v.z, v.w};
}
QO_DEV uint2 ld_stream(const uint2 *p) {
    nv2u v = __builtin_nontemporal_load(reinterpret_cast<const nv2u *>(p));
    return uint2{v.x, v.y};
}
template <class T>
QO_DEV void st_stream(T *p, T v) { __builtin_nontemporal_store(v, p); }

QO_DEV uint64_t ld_agent(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
QO_DEV void st_agent(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Striped tile offsets.  A block of NW waves holds J groups of per-lane counts
// c[j]; group j of thread t precedes group j of thread t+1 and all of group j+1
// (order (j, wave, lane)).  On return excl[j] is the exclusive offset of the
// lane's items inside the tile and the tile total is returned.  s_tot must hold
// tile_offsets_words(J, NW) entries; the caller must __syncthreads() before
// reusing it.  Wave 0 scans the J*NW wave totals (P per lane) so that no thread
// keeps them all in registers.
constexpr int tile_offsets_words(int J, int NW) { return J * NW + 1; }

template <int J, int NW, class T>
QO_DEV T tile_offsets(const T (&c)[J], T (&excl)[J], T *s_tot) {
    constexpr int K = J * NW;
    constexpr int P = (K + kWave - 1) / kWave;
    const unsigned w = wave_id();
    const unsigned lane = lane_id();
#pragma unroll
    for (int j = 0; j < J; ++j) {
        T inc = wave_incl_scan(c[j]);
        excl[j] = inc - c[j];
        if (lane == kWave - 1) s_tot[j * NW + w] = inc;
    }
    __syncthreads();
    if (w == 0) {
        T v[P];
        T sum = 0;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const unsigned idx = lane * P + p;
            v[p] = idx < (unsigned)K ? s_tot[idx] : T(0);
            sum += v[p];
        }
        const T inc = wave_incl_scan(sum);
        T run = inc - sum;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const unsigned idx = lane * P + p;
            if (idx < (unsigned)K) s_tot[idx] = run;
            run += v[p];
        }
        if (lane == kWave - 1) s_tot[K] = inc;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < J; ++j) excl[j] += s_tot[j * NW + w];
    return s_tot[K];
}

// Decoupled look-back status word: [63:62] flag (0 empty, 1 aggregate,
// 2 inclusive prefix), [61:0] value.
constexpr uint64_t kFlagAgg = uint64_t(1) << 62;
constexpr uint64_t kFlagInc = uint64_t(2) << 62;
constexpr uint64_t kValMask = (uint64_t(1) << 62) - 1;

// Library-wide device control block (one per device, allocated at init):
// kernels with a bounded wait read the bound from it and report a wait that
// ran out into `error` (the host turns that into QO_ERR_DEVICE_WAIT).
struct DevCtl {
    uint32_t error;       // kDevErr* bits, sticky until the host reads them
    uint32_t spin_limit;  // sleeps a look-back may wait before giving up
    uint32_t pad[62];
};
constexpr uint32_t kDevErrLookback = 1u;
constexpr uint32_t kDefaultSpinLimit = 1u << 22;  // s_sleep(1) rounds: about a second

// Called by ONE whole wave of the block that owns `tile`.  Publishes the tile
// aggregate, accumulates predecessors 64 at a time and publishes the inclusive
// prefix.  Returns the exclusive prefix (sum of all earlier tiles, modulo 2^62).
// Tiles must be numbered by a ticket counter, so every predecessor is already
// resident and the wait ends; it is bounded all the same: after
// ctl->spin_limit sleeps the wave sets kDevErrLookback in ctl->error and
// returns 0 (the call's results are then undefined, and the kernel ends).
QO_DEV uint64_t lookback(uint64_t *status, uint64_t tile, uint64_t aggregate, DevCtl *ctl) {
    const unsigned lane = lane_id();
    if (tile == 0) {
        if (lane == 0) st_agent(&status[0], kFlagInc | (aggregate & kValMask));
        return 0;
    }
    if (lane == 0) st_agent(&status[tile], kFlagAgg | (aggregate & kValMask));
    uint64_t excl = 0;
    int64_t base = (int64_t)tile - 1;
    const uint32_t limit = ctl->spin_limit;
    uint32_t sleeps = 0;  // wave-uniform: the loop's branches depend on ballots only
    while (true) {
        int64_t idx = base - (int64_t)lane;
        uint64_t st = (idx >= 0) ? ld_agent(&status[idx]) : kFlagInc;
        uint64_t flag = st >> 62;
        uint64_t inc_mask = ballot(flag == 2);
        uint64_t empty_mask = ballot(flag == 0);
        unsigned first_inc = inc_mask ? (unsigned)__builtin_ctzll(inc_mask) : 64u;
        uint64_t relevant = first_inc >= 63 ? ~uint64_t(0) : ((uint64_t(2) << first_inc) - 1);
        if (empty_mask & relevant) {
            if (++sleeps > limit) {
                if (lane == 0) atomicOr(&ctl->error, kDevErrLookback);
                return 0;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t v = (lane <= first_inc) ? (st & kValMask) : 0;
        excl += wave_sum(v);
        if (first_inc < 64) break;
        base -= kWave;
    }
    excl &= kValMask;
    if (lane == 0) st_agent(&status[tile], kFlagInc | ((excl + aggregate) & kValMask));
    return excl;
}

// Exclusive scan of one value per thread over a block of NW waves.  s_w holds
// NW words; the caller must __syncthreads() before reusing it.
template <int NW>
QO_DEV uint32_t block_excl_scan(uint32_t v, uint32_t *s_w) {
    uint32_t inc = wave_incl_scan(v);
    if (lane_id() == kWave - 1) s_w[wave_id()] = inc;
    __syncthreads();
    uint32_t add = 0;
    for (unsigned w = 0; w < wave_id(); ++w) add += s_w[w];
    return inc - v + add;
}
QO_DEV uint32_t block_excl_scan256(uint32_t v, uint32_t *s_w) { return block_excl_scan<4>(v, s_w); }

// Tile of work-group b when each XCD walks a contiguous tile range in order:
// work-group b runs on XCD b mod 8, which owns tiles [x*q + min(x, r), ...)
// with q = ntile / 8, r = ntile mod 8.  Writes that continue one another
// (adjacent runs of one output region) then come from one XCD, and its L2
// merges them into whole lines before they reach HBM.  A bijection of
// [0, ntile) for b < ntile.
QO_DEV uint32_t xcd_tile(uint32_t b, uint32_t ntile) {
    const uint32_t x = b & 7u, i = b >> 3, q = ntile >> 3, r = ntile & 7u;
    return x * q + (x < r ? x : r) + i;
}

// 32-bit finaliser of MurmurHash3 (Appleby): spreads sequential keys.
QO_DEV uint32_t hash32(uint32_t k) {
    k ^= k >> 16;
    k *= 0x85EBCA6Bu;
    k ^= k >> 13;
    k *= 0xC2B2AE35u;
    k ^= k >> 16;
    return k;
}

}  // namespace qo
