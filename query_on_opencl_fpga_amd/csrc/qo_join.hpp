// Open-addressing join table shared by the plain and batched join paths.
//
// Layout in device memory: a 256-byte JoinHeader followed by 8-byte slots.  A
// slot holds (r_payload << 32) | key; kEmptySlot marks a free slot.  The one
// tuple whose packed form equals kEmptySlot (key = -1, payload = -1) is counted
// in the header instead of being stored.  Collisions probe linearly, wrapping
// at `cap` (the slots in use).
//
// A key's home slot is floor(h * cap / 2^32) for a 32-bit hash h, chosen by the
// build and recorded in the header (JoinHash):
//  * hash-mixed table: h = hash32(key), cap = every allocated slot (any count);
//  * key-range table: when the build keys span a range of at most 2^p <= the
//    allocated slots, h = (key - kmin) << (32 - p) and cap = 2^p, so the home
//    slot IS key - kmin: with unique keys every lookup is one load and no
//    chain, the direct-mapped layout of the reference's table (bucket =
//    key & (B - 1), hj_filter_sort_op.cl:36,89).  Keys outside the range
//    cannot be in such a table.  A build that finds a duplicate key falls back
//    to the hash-mixed table (chains of a dense key range would run long).
//
// Because the home slot is monotone in h, the keys whose hash has top bits p
// (of L) all have home slots in [p*cap/2^L, (p+1)*cap/2^L]: a contiguous 1/2^L
// of the table.  The partitioned build and probe of the hash-mixed table use
// this to work on one table slice at a time in LDS.
#pragma once
#include "qo_device.hpp"
#include "qo_sort.hpp"

namespace qo {

constexpr uint64_t kEmptySlot = ~uint64_t(0);
constexpr uint32_t kJoinHashMix = 0;   // JoinHeader::mode
constexpr uint32_t kJoinKeyRange = 1;
constexpr uint32_t kHashFib = 2;       // multiplicative hash: splits of the group-by only, never a table mode

struct alignas(256) JoinHeader {
    uint64_t cap;             // slots in use
    uint32_t has_dups;        // some key was inserted twice: probe emits every match
    uint32_t sentinel_count;  // number of (-1, -1) build tuples
    uint32_t overflow;        // an insert found no free slot
    uint32_t mode;            // kJoinHashMix or kJoinKeyRange
    uint64_t nr;
    uint64_t slots_alloc;     // slots allocated after the header
    int32_t kmin;             // key-range table: home slot = key - kmin
    uint32_t shift;           // key-range table: h = (key - kmin) << shift, cap = 2^(32 - shift)
    int32_t rmin, rmax;       // build key range (reduced by the build)
    unsigned long long occupied;  // key-range build: occupied slots
    uint64_t pad[24];
};
static_assert(sizeof(JoinHeader) == 256, "header size");

// The table's hash, read from its header once per kernel.
struct JoinHash {
    uint32_t mode, shift;
    int32_t kmin;
    QO_DEV static JoinHash of(const JoinHeader *h) { return JoinHash{h->mode, h->shift, h->kmin}; }
    // Sets h; false when `key` lies outside a key-range table (it cannot be stored there).
    QO_DEV bool operator()(int32_t key, uint32_t &h) const {
        if (mode == kHashFib) {
            h = (uint32_t)key * 0x9E3779B9u;
            return true;
        }
        if (mode != kJoinKeyRange) {
            h = hash32((uint32_t)key);
            return true;
        }
        const uint32_t d = (uint32_t)key - (uint32_t)kmin;
        h = d << shift;
        return (d >> (32u - shift)) == 0u;  // shift is in [1, 31]
    }
    // Slots a lookup may have to read: a key-range table holds each key at its
    // home slot only (a build with duplicate keys falls back to hash mixing),
    // while hash-mixed chains run until an empty slot.
    QO_DEV uint64_t walk(uint64_t cap) const { return mode == kJoinKeyRange ? 1 : cap; }
};

QO_DEV uint64_t join_pack(int32_t key, int32_t pay) { return ((uint64_t)(uint32_t)pay << 32) | (uint32_t)key; }

QO_DEV uint64_t join_home(uint32_t h, uint64_t cap) { return ((uint64_t)h * cap) >> 32; }

QO_DEV uint64_t join_next(uint64_t h, uint64_t cap) { return h + 1 == cap ? 0 : h + 1; }

// Insert into a hash-mixed table.
QO_DEV void join_insert(uint64_t *slots, uint64_t cap, JoinHeader *hdr, int32_t key, int32_t pay) {
    const uint64_t packed = join_pack(key, pay);
    if (packed == kEmptySlot) {
        atomicAdd(&hdr->sentinel_count, 1u);
        return;
    }
    uint64_t h = join_home(hash32((uint32_t)key), cap);
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
QO_DEV uint32_t join_count(const uint64_t *slots, uint64_t cap, const JoinHash &hs, uint32_t sentinel, int32_t key) {
    uint32_t c = 0;
    uint32_t hv;
    if (hs(key, hv)) {
        uint64_t h = join_home(hv, cap);
        const uint64_t walk = hs.walk(cap);
        for (uint64_t probe = 0; probe < walk; ++probe) {
            uint64_t s = slots[h];
            if (s == kEmptySlot) break;
            c += ((uint32_t)s == (uint32_t)key);
            h = join_next(h, cap);
        }
    }
    if (key == -1) c += sentinel;
    return c;
}

constexpr int kProbeJ = 4;                                   // int4 groups per thread
constexpr size_t kProbeTile = (size_t)kProbeJ * 256 * 4;     // S tuples per tile

// Emits every match of `key` from position `pos` on, walking the whole chain
// (table-chain order, then the (-1, -1) build tuples counted in the header).
template <class Emit>
QO_DEV uint64_t emit_chain(const uint64_t *slots, uint64_t cap, const JoinHash &hs, uint32_t sentinel, int32_t key,
                           int32_t sp, uint64_t pos, const Emit &emit) {
    uint32_t hv;
    if (hs(key, hv)) {
        uint64_t h = join_home(hv, cap);
        const uint64_t walk = hs.walk(cap);
        for (uint64_t probe = 0; probe < walk; ++probe) {
            uint64_t s = slots[h];
            if (s == kEmptySlot) break;
            if ((uint32_t)s == (uint32_t)key) emit(pos++, key, (int32_t)(s >> 32), sp);
            h = join_next(h, cap);
        }
    }
    if (key == -1)
        for (uint32_t q = 0; q < sentinel; ++q) emit(pos++, key, -1, sp);
    return pos;
}

// One S tuple on its own (the rare overflow path of the partitioned probe):
// counts its matches, reserves them with one atomic add and emits them.
template <class Emit>
QO_DEV void probe_one(const uint64_t *slots, uint64_t cap, const JoinHash &hs, uint32_t sentinel, int32_t key,
                      int32_t sp, unsigned long long *counter, const Emit &emit) {
    const uint32_t m = join_count(slots, cap, hs, sentinel, key);
    if (!m) return;
    const uint64_t pos = atomicAdd(counter, (unsigned long long)m);
    emit_chain(slots, cap, hs, sentinel, key, sp, pos, emit);
}

// LDS staging of one tile's matches (unique-key path), so that the tile is
// written out with consecutive threads on consecutive rows (coalesced stores).
template <uint32_t ROWS>
struct JoinStageT {
    static constexpr uint32_t kRows = ROWS;
    int32_t key[ROWS], rpay[ROWS], spay[ROWS];
};
constexpr uint32_t kStageRows = 1024 * 4;
using JoinStage = JoinStageT<kStageRows>;

// Looks up the block's tuples (J groups of 4 per thread; bit k of pass[j]
// marks element k of group j live), reserves the block's output range with one
// atomic add on `counter` and emits the matches there in (group, wave, lane,
// element) order; for one tuple, in table-chain order.  Emit(pos, key,
// r_payload, s_payload).  With `stage` (J * 1024 <= kStageRows) the unique-key
// path writes the matches to LDS first and emits them row by row across the
// block; without it, and on the general path, each thread emits its own
// matches.  Must be called by all 256 threads of the block;
// s_tot holds tile_offsets_words(J, 4) words.
template <int J, class Emit, class Stage = JoinStage>
QO_DEV void lookup_emit(const uint64_t *__restrict__ slots, uint64_t cap_slots, const JoinHash &hsh, bool general,
                        uint32_t sentinel,
                        const int32_t (&key)[J][4], const int32_t (&sp)[J][4], const uint32_t (&pass)[J],
                        unsigned long long *counter, const Emit &emit, uint32_t *s_tot, uint64_t *s_excl,
                        Stage *stage = nullptr) {
    uint32_t c[J], ex[J];
    uint32_t hit[J];
    // On a hit of the unique-key path sv holds the matching slot (its payload
    // is emitted from there).  Slot indices fit 32 bits (cap < 2^32).
    uint64_t sv[J][4];
    if (!general) {
        // Chains are resolved in rounds: every unresolved tuple of the thread
        // issues its next slot load in the same round, so a wave waits one
        // memory round trip per chain step, not one per tuple and step.
        const uint32_t cap32 = (uint32_t)cap_slots;
        const uint64_t walk = hsh.walk(cap_slots);
        uint32_t hs[J][4];
        uint32_t todo[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            todo[j] = pass[j];
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t hv;
                const bool in = hsh(key[j][k], hv);
                hs[j][k] = (uint32_t)join_home(hv, cap_slots);
                sv[j][k] = ((pass[j] >> k) & 1u) && in ? slots[hs[j][k]] : kEmptySlot;
            }
        }
        for (uint32_t round = 1;; ++round) {
            uint32_t left = 0;
#pragma unroll
            for (int j = 0; j < J; ++j) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (!((todo[j] >> k) & 1u)) continue;
                    if (sv[j][k] == kEmptySlot) {
                        todo[j] &= ~(1u << k);
                    } else if ((uint32_t)sv[j][k] == (uint32_t)key[j][k]) {
                        hit[j] |= 1u << k;
                        todo[j] &= ~(1u << k);
                    } else {
                        hs[j][k] = hs[j][k] + 1 == cap32 ? 0u : hs[j][k] + 1;
                    }
                }
                left |= todo[j];
            }
            if (!left || round >= walk) break;  // a chain never exceeds the table
#pragma unroll
            for (int j = 0; j < J; ++j)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if ((todo[j] >> k) & 1u) sv[j][k] = slots[hs[j][k]];
        }
#pragma unroll
        for (int j = 0; j < J; ++j) c[j] = __popc(hit[j]);
    } else {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            c[j] = 0;
            hit[j] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                sv[j][k] = 0;
                if ((pass[j] >> k) & 1u) {
                    uint32_t m = join_count(slots, cap_slots, hsh, sentinel, key[j][k]);
                    c[j] += m;
                    if (m) hit[j] |= 1u << k;
                }
            }
        }
    }

    const uint32_t total = tile_offsets<J, 4>(c, ex, s_tot);
    if (threadIdx.x == 0) *s_excl = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
    if (stage && !general) {
        static_assert(J * 1024 <= Stage::kRows, "tile larger than the stage");
#pragma unroll
        for (int j = 0; j < J; ++j) {
            uint32_t pos = ex[j];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!((hit[j] >> k) & 1u)) continue;
                stage->key[pos] = key[j][k];
                stage->rpay[pos] = (int32_t)(sv[j][k] >> 32);
                stage->spay[pos] = sp[j][k];
                ++pos;
            }
        }
        __syncthreads();
        const uint64_t base = *s_excl;
        if constexpr (Emit::kFlush) {
            emit.flush(base, total, stage->key, stage->rpay, stage->spay);
        } else {
            for (uint32_t i = threadIdx.x; i < total; i += 256)
                emit(base + i, stage->key[i], stage->rpay[i], stage->spay[i]);
        }
        return;
    }
    __syncthreads();
    const uint64_t base = *s_excl;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        uint64_t pos = base + ex[j];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!((hit[j] >> k) & 1u)) continue;
            if (!general) emit(pos++, key[j][k], (int32_t)(sv[j][k] >> 32), sp[j][k]);
            else pos = emit_chain(slots, cap_slots, hsh, sentinel, key[j][k], sp[j][k], pos, emit);
        }
    }
}

// Loads J int4 groups per thread (NT threads) of S tile `tile` (tuples
// [tile*J*4*NT, ...)) and evaluates the filter: bit k of pass[j] marks element
// k of group j.
template <int J, bool HAS_FILTER, int NT = 256>
QO_DEV void load_s_tile(const int32_t *__restrict__ skey, const int32_t *__restrict__ spay,
                        const int32_t *__restrict__ sfilt, size_t ns, int32_t flo, int32_t fhi, bool vec,
                        uint32_t tile, int32_t (&key)[J][4], int32_t (&sp)[J][4], uint32_t (&pass)[J]) {
    const size_t vbase = (size_t)tile * (J * NT);
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const size_t v = vbase + (size_t)j * NT + threadIdx.x;
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
}

// Probe one tile of S: tuples [tile*kProbeTile, ...) of (skey, spay, sfilt).
// Filters, looks up, reserves the tile's output range with one atomic add on
// `counter` (the running match count) and emits the tile's matches there in S
// order (for one S tuple, in table-chain order).  Tiles land in completion
// order: a look-back across tiles would make every tile wait for the slowest
// lookup of its predecessors.  Must be called by all 256 threads of the block;
// s_tot holds tile_offsets_words(kProbeJ, 4) words.
template <bool HAS_FILTER, class Emit>
QO_DEV void probe_tile(const JoinHeader *__restrict__ hdr, const uint64_t *__restrict__ slots,
                       const int32_t *__restrict__ skey, const int32_t *__restrict__ spay,
                       const int32_t *__restrict__ sfilt, size_t ns, int32_t flo, int32_t fhi, bool vec,
                       uint32_t tile, unsigned long long *counter, const Emit &emit, uint32_t *s_tot,
                       uint64_t *s_excl, JoinStage *stage = nullptr) {
    const uint64_t cap_slots = hdr->cap;
    const bool general = (hdr->has_dups | hdr->sentinel_count) != 0;
    const uint32_t sentinel = hdr->sentinel_count;
    int32_t key[kProbeJ][4], sp[kProbeJ][4];
    uint32_t pass[kProbeJ];
    load_s_tile<kProbeJ, HAS_FILTER>(skey, spay, sfilt, ns, flo, fhi, vec, tile, key, sp, pass);
    lookup_emit<kProbeJ>(slots, cap_slots, JoinHash::of(hdr), general, sentinel, key, sp, pass, counter, emit, s_tot,
                         s_excl, stage);
}

// Emits match rows into three SoA int32 columns (capacity-checked).
// Join output row `pos` (rows past `cap` are counted, not written): three
// int32 columns, or, when k64 is set, the sort-ready form of a join that is
// sorted next: k64 = join_sort_key(key, s_payload), v32 = r_payload.
struct SoaEmit {
    static constexpr bool kFlush = true;  // writes a staged tile with flush()
    int32_t *okey, *orpay, *ospay;
    size_t cap;
    uint64_t *k64 = nullptr;
    uint32_t *v32 = nullptr;
    QO_DEV void operator()(uint64_t pos, int32_t k, int32_t rp, int32_t sp) const {
        if (pos >= cap) return;
        if (k64) {
            k64[pos] = join_sort_key(k, sp);
            v32[pos] = (uint32_t)rp;
        } else {
            okey[pos] = k;
            orpay[pos] = rp;
            ospay[pos] = sp;
        }
    }
    // Rows [base, base + total) from LDS columns, written by all 256 threads of
    // the block: the rows before the first 4-aligned output row and those past
    // the last full group of 4 one per thread, the rest 4 rows per thread with
    // 16-byte stores to each column (two to k64 in the sort-ready form), so a
    // wave writes 1 KiB runs per column with a quarter of the store
    // instructions.  Needs 16-byte-aligned columns; otherwise, and for rows
    // past `cap`, rows go one per thread.
    QO_DEV void flush(uint64_t base, uint32_t total, const int32_t *key, const int32_t *rp,
                      const int32_t *sp) const {
        flush_rows(base, total, [&](uint32_t i, int32_t &k, int32_t &r, int32_t &s) {
            k = key[i];
            r = rp[i];
            s = sp[i];
        });
    }
    // The same for rows that `row(i, key, r_payload, s_payload)` reads, by
    // all NT threads of the block.
    template <int NT = 256, class Row>
    QO_DEV void flush_rows(uint64_t base, uint32_t total, const Row &row) const {
        const bool vec = base + total <= cap &&
                         ((k64 ? (reinterpret_cast<uintptr_t>(k64) | reinterpret_cast<uintptr_t>(v32))
                               : (reinterpret_cast<uintptr_t>(okey) | reinterpret_cast<uintptr_t>(orpay) |
                                  reinterpret_cast<uintptr_t>(ospay))) &
                          15u) == 0;
        auto one = [&](uint32_t i) {
            int32_t k, r, s;
            row(i, k, r, s);
            (*this)(base + i, k, r, s);
        };
        if (!vec) {
            for (uint32_t i = threadIdx.x; i < total; i += NT) one(i);
            return;
        }
        const uint32_t head = std::min<uint32_t>(total, (uint32_t)((4u - (base & 3u)) & 3u));
        const uint32_t groups = (total - head) / 4u, tail0 = head + groups * 4u;
        for (uint32_t i = threadIdx.x; i < head; i += NT) one(i);
        for (uint32_t i = tail0 + threadIdx.x; i < total; i += NT) one(i);
        for (uint32_t g = threadIdx.x; g < groups; g += NT) {
            const uint32_t i = head + 4u * g;
            const uint64_t pos = base + i;  // a multiple of 4
            int32_t key[4], rp[4], sp[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) row(i + e, key[e], rp[e], sp[e]);
            if (k64) {
                uint64_t kk[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) kk[e] = join_sort_key(key[e], sp[e]);
                uint4 *k4 = reinterpret_cast<uint4 *>(k64 + pos);
                k4[0] = uint4{(uint32_t)kk[0], (uint32_t)(kk[0] >> 32), (uint32_t)kk[1], (uint32_t)(kk[1] >> 32)};
                k4[1] = uint4{(uint32_t)kk[2], (uint32_t)(kk[2] >> 32), (uint32_t)kk[3], (uint32_t)(kk[3] >> 32)};
                *reinterpret_cast<uint4 *>(v32 + pos) =
                    uint4{(uint32_t)rp[0], (uint32_t)rp[1], (uint32_t)rp[2], (uint32_t)rp[3]};
            } else {
                *reinterpret_cast<int4 *>(okey + pos) = int4{key[0], key[1], key[2], key[3]};
                *reinterpret_cast<int4 *>(orpay + pos) = int4{rp[0], rp[1], rp[2], rp[3]};
                *reinterpret_cast<int4 *>(ospay + pos) = int4{sp[0], sp[1], sp[2], sp[3]};
            }
        }
    }
};

}  // namespace qo
