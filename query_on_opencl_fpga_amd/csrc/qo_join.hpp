// Open-addressing join table shared by the plain and batched join paths.
//
// Layout in device memory: a 256-byte JoinHeader followed by `cap` 8-byte
// slots (any count; the table size is picked for bandwidth, not rounded to a
// power of two).  A key's home slot is floor(hash32(key) * cap / 2^32) and
// collisions probe linearly, wrapping at cap.  A slot holds
// (r_payload << 32) | key; kEmptySlot marks a free slot.  The one tuple whose
// packed form equals kEmptySlot (key = -1, payload = -1) is counted in the
// header instead of being stored.
#pragma once
#include "qo_device.hpp"

namespace qo {

constexpr uint64_t kEmptySlot = ~uint64_t(0);

struct alignas(256) JoinHeader {
    uint64_t cap;
    uint32_t has_dups;        // some key was inserted twice: probe emits every match
    uint32_t sentinel_count;  // number of (-1, -1) build tuples
    uint32_t overflow;        // an insert found no free slot
    uint32_t pad0;
    uint64_t nr;
    uint64_t pad[28];
};
static_assert(sizeof(JoinHeader) == 256, "header size");

QO_DEV uint64_t join_pack(int32_t key, int32_t pay) { return ((uint64_t)(uint32_t)pay << 32) | (uint32_t)key; }

QO_DEV uint64_t join_home(int32_t key, uint64_t cap) { return ((uint64_t)hash32((uint32_t)key) * cap) >> 32; }

QO_DEV uint64_t join_next(uint64_t h, uint64_t cap) { return h + 1 == cap ? 0 : h + 1; }

QO_DEV void join_insert(uint64_t *slots, uint64_t cap, JoinHeader *hdr, int32_t key, int32_t pay) {
    const uint64_t packed = join_pack(key, pay);
    if (packed == kEmptySlot) {
        atomicAdd(&hdr->sentinel_count, 1u);
        return;
    }
    uint64_t h = join_home(key, cap);
    for (uint64_t probe = 0; probe < cap; ++probe) {
        uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long *>(&slots[h]),
                                  (unsigned long long)kEmptySlot, (unsigned long long)packed);
        if (prev == kEmptySlot) return;
        if ((uint32_t)prev == (uint32_t)key && hdr->has_dups == 0) atomicExch(&hdr->has_dups, 1u);
        h = join_next(h, cap);
    }
    atomicExch(&hdr->overflow, 1u);
}

// Number of matches of `key` (walks the whole chain; general path).
QO_DEV uint32_t join_count(const uint64_t *slots, uint64_t cap, uint32_t sentinel, int32_t key) {
    uint32_t c = 0;
    uint64_t h = join_home(key, cap);
    for (uint64_t probe = 0; probe < cap; ++probe) {
        uint64_t s = slots[h];
        if (s == kEmptySlot) break;
        c += ((uint32_t)s == (uint32_t)key);
        h = join_next(h, cap);
    }
    if (key == -1) c += sentinel;
    return c;
}

// First match of `key` in a table built from unique keys, given its home slot h
// and that slot's value s; returns whether found.
QO_DEV bool join_find(const uint64_t *slots, uint64_t cap, int32_t key, uint64_t h, uint64_t s, int32_t &rpay) {
    for (uint64_t probe = 0; probe < cap; ++probe) {
        if (s == kEmptySlot) return false;
        if ((uint32_t)s == (uint32_t)key) {
            rpay = (int32_t)(s >> 32);
            return true;
        }
        h = join_next(h, cap);
        s = slots[h];
    }
    return false;
}

constexpr int kProbeJ = 4;                                   // int4 groups per thread
constexpr size_t kProbeTile = (size_t)kProbeJ * 256 * 4;     // S tuples per tile

// Probe one tile of S: tuples [tile*kProbeTile, ...) of (skey, spay, sfilt).
// Filters, looks up, reserves the tile's output range with one atomic add on
// `counter` (the running match count) and emits the tile's matches there in S
// order (for one S tuple, in table-chain order).  Tiles land in completion
// order: a look-back across tiles would make every tile wait for the slowest
// lookup of its predecessors.  Emit(pos, key, r_payload, s_payload).
// Must be called by all 256 threads of the block; s_tot holds kProbeJ*4 words.
template <bool HAS_FILTER, class Emit>
QO_DEV void probe_tile(const JoinHeader *__restrict__ hdr, const uint64_t *__restrict__ slots,
                       const int32_t *__restrict__ skey, const int32_t *__restrict__ spay,
                       const int32_t *__restrict__ sfilt, size_t ns, int32_t flo, int32_t fhi, bool vec,
                       uint32_t tile, unsigned long long *counter, const Emit &emit, uint32_t *s_tot,
                       uint64_t *s_excl) {
    constexpr int J = kProbeJ;
    constexpr int NW = 4;
    const uint64_t cap_slots = hdr->cap;
    const bool general = (hdr->has_dups | hdr->sentinel_count) != 0;
    const uint32_t sentinel = hdr->sentinel_count;
    const size_t vbase = (size_t)tile * (J * 256);

    int32_t key[J][4], sp[J][4], rp[J][4];
    uint32_t pass[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const size_t v = vbase + (size_t)j * 256 + threadIdx.x;
        const size_t e0 = v * 4;
        pass[j] = 0;
        if (vec && e0 + 4 <= ns) {
            int4 kk = ld_stream(reinterpret_cast<const int4 *>(skey) + v);
            int4 pp = ld_stream(reinterpret_cast<const int4 *>(spay) + v);
            key[j][0] = kk.x; key[j][1] = kk.y; key[j][2] = kk.z; key[j][3] = kk.w;
            sp[j][0] = pp.x; sp[j][1] = pp.y; sp[j][2] = pp.z; sp[j][3] = pp.w;
            if (HAS_FILTER) {
                int4 ff = ld_stream(reinterpret_cast<const int4 *>(sfilt) + v);
                pass[j] = ((ff.x >= flo) & (ff.x <= fhi)) | (((ff.y >= flo) & (ff.y <= fhi)) << 1) |
                          (((ff.z >= flo) & (ff.z <= fhi)) << 2) | (((ff.w >= flo) & (ff.w <= fhi)) << 3);
            } else {
                pass[j] = 15u;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const size_t e = e0 + k;
                key[j][k] = 0;
                sp[j][k] = 0;
                if (e < ns) {
                    key[j][k] = skey[e];
                    sp[j][k] = spay[e];
                    bool ok = true;
                    if (HAS_FILTER) {
                        int32_t f = sfilt[e];
                        ok = (f >= flo) & (f <= fhi);
                    }
                    pass[j] |= ok ? (1u << k) : 0u;
                }
            }
        }
    }

    uint32_t c[J], ex[J];
    uint32_t hit[J];
    if (!general) {
        // Issue every first-slot load before resolving any chain.
        uint64_t hs[J][4], sv[J][4];
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                hs[j][k] = join_home(key[j][k], cap_slots);
                sv[j][k] = (pass[j] >> k) & 1u ? slots[hs[j][k]] : kEmptySlot;
            }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if ((pass[j] >> k) & 1u)
                    if (join_find(slots, cap_slots, key[j][k], hs[j][k], sv[j][k], rp[j][k])) hit[j] |= 1u << k;
            c[j] = __popc(hit[j]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            c[j] = 0;
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if ((pass[j] >> k) & 1u) {
                    uint32_t m = join_count(slots, cap_slots, sentinel, key[j][k]);
                    c[j] += m;
                    if (m) hit[j] |= 1u << k;
                }
        }
    }

    const uint32_t total = tile_offsets<J, NW>(c, ex, s_tot);
    if (threadIdx.x == 0) *s_excl = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
    __syncthreads();
    const uint64_t base = *s_excl;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        uint64_t pos = base + ex[j];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!((hit[j] >> k) & 1u)) continue;
            if (!general) {
                emit(pos++, key[j][k], rp[j][k], sp[j][k]);
            } else {
                const int32_t kk = key[j][k];
                uint64_t h = join_home(kk, cap_slots);
                for (uint64_t probe = 0; probe < cap_slots; ++probe) {
                    uint64_t s = slots[h];
                    if (s == kEmptySlot) break;
                    if ((uint32_t)s == (uint32_t)kk) emit(pos++, kk, (int32_t)(s >> 32), sp[j][k]);
                    h = join_next(h, cap_slots);
                }
                if (kk == -1)
                    for (uint32_t q = 0; q < sentinel; ++q) emit(pos++, kk, -1, sp[j][k]);
            }
        }
    }
}

// Emits match rows into three SoA int32 columns (capacity-checked).
struct SoaEmit {
    int32_t *okey, *orpay, *ospay;
    size_t cap;
    QO_DEV void operator()(uint64_t pos, int32_t k, int32_t rp, int32_t sp) const {
        if (pos < cap) {
            okey[pos] = k;
            orpay[pos] = rp;
            ospay[pos] = sp;
        }
    }
};

}  // namespace qo
