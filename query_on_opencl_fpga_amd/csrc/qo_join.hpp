// Open-addressing join table shared by the plain and batched join paths.
//
// Layout in device memory: a 256-byte JoinHeader followed by 2^log2cap 8-byte
// slots.  A slot holds (r_payload << 32) | key; kEmptySlot marks a free slot.
// The one tuple whose packed form equals kEmptySlot (key = -1, payload = -1)
// is counted in the header instead of being stored.
#pragma once
#include "qo_device.hpp"

namespace qo {

constexpr uint64_t kEmptySlot = ~uint64_t(0);

struct alignas(256) JoinHeader {
    uint32_t log2cap;
    uint32_t has_dups;        // some key was inserted twice: probe emits every match
    uint32_t sentinel_count;  // number of (-1, -1) build tuples
    uint32_t overflow;        // an insert found no free slot
    uint64_t nr;
    uint64_t pad[29];
};
static_assert(sizeof(JoinHeader) == 256, "header size");

QO_DEV uint64_t join_pack(int32_t key, int32_t pay) { return ((uint64_t)(uint32_t)pay << 32) | (uint32_t)key; }

QO_DEV void join_insert(uint64_t *slots, uint64_t mask, JoinHeader *hdr, int32_t key, int32_t pay) {
    const uint64_t packed = join_pack(key, pay);
    if (packed == kEmptySlot) {
        atomicAdd(&hdr->sentinel_count, 1u);
        return;
    }
    uint64_t h = hash32((uint32_t)key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long *>(&slots[h]),
                                  (unsigned long long)kEmptySlot, (unsigned long long)packed);
        if (prev == kEmptySlot) return;
        if ((uint32_t)prev == (uint32_t)key && hdr->has_dups == 0) atomicExch(&hdr->has_dups, 1u);
        h = (h + 1) & mask;
    }
    atomicExch(&hdr->overflow, 1u);
}

// Number of matches of `key` (walks the whole chain; general path).
QO_DEV uint32_t join_count(const uint64_t *slots, uint64_t mask, uint32_t sentinel, int32_t key) {
    uint32_t c = 0;
    uint64_t h = hash32((uint32_t)key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        uint64_t s = slots[h];
        if (s == kEmptySlot) break;
        c += ((uint32_t)s == (uint32_t)key);
        h = (h + 1) & mask;
    }
    if (key == -1) c += sentinel;
    return c;
}

// First match of `key` in a table built from unique keys; returns whether found.
QO_DEV bool join_find(const uint64_t *slots, uint64_t mask, int32_t key, uint64_t h, uint64_t s, int32_t &rpay) {
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        if (s == kEmptySlot) return false;
        if ((uint32_t)s == (uint32_t)key) {
            rpay = (int32_t)(s >> 32);
            return true;
        }
        h = (h + 1) & mask;
        s = slots[h];
    }
    return false;
}

}  // namespace qo
