// Segmented radix sort interface and the join-row sort key shared by the plain
// and batched paths.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace qo {

constexpr int kSortItems = 16;
constexpr int kSortTile = 256 * kSortItems;

// One independently sorted array: elements [base, base + n) of the shared key /
// value arrays, n = *d_n (clamped to cap) or cap when d_n is null.  The
// segment owns global tiles [tile0, tile0 + ntile), ntile = ceil(cap / tile).
struct RadixSeg {
    uint64_t base;
    uint64_t cap;
    const uint64_t *d_n;
    uint32_t tile0;
    uint32_t ntile;
};

// Output columns of a sorted join-row segment: the last pass of a join-row
// sort writes (key, r_payload, s_payload) here instead of the packed rows.
struct RowOut {
    int32_t *key, *rpay, *spay;
};

// Tiles [t0, t1) (segment-local) of segment `seg`: one block of the up-front
// digit count of a segmented sort.
struct SweepChunk {
    uint32_t seg, t0, t1;
};
constexpr uint32_t kSweepChunkTiles = 16;
// Workspace of launch_radix_rows_segments.
size_t radix_rows_seg_bytes(uint32_t total_tiles, int nseg, uint32_t nchunks);
// Segmented sort by bits [lo_bit, hi_bit): digit starts per segment from the
// chunks (segment s owns chunks [chunk0[s], chunk0[s + 1])), then per pass a
// tile histogram, a segmented column scan and the placement.  On return
// *in_alt tells whether the data ended in kalt / valt.
// With d_outs (one RowOut per segment, device memory) the keys are join
// sort keys with r_payload values and the last pass writes each segment's
// rows unpacked into its RowOut; the packed result is then not kept.  With
// run_if every kernel returns at once unless *run_if != 0 (decided on the
// device, for stream captures).
int launch_radix_rows_segments(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt,
                               const RadixSeg *d_segs, int nseg, uint32_t total_tiles, const SweepChunk *d_chunks,
                               uint32_t nchunks, const uint32_t *d_chunk0, int lo_bit, int hi_bit, void *ws,
                               hipStream_t st, bool *in_alt, const RowOut *d_outs = nullptr,
                               const uint32_t *run_if = nullptr);
// After a segmented sort of join rows on their key bits only [32, 64): sorts
// every run of equal key by s_payload and writes each segment's rows to its
// RowOut (segment tiles of kSortTile rows); a run longer than 64 rows sets
// *long_flag and is not written.
int launch_fix_runs_segments(const uint64_t *keys, const uint32_t *vals, const RadixSeg *d_segs, int nseg,
                             uint32_t total_tiles, const RowOut *d_outs, uint32_t *long_flag, hipStream_t st);
// Join rows sort ascending by (key, s_payload) as signed int32: flip the sign
// bits so unsigned 64-bit order is the signed lexicographic order.
__host__ __device__ __forceinline__ uint64_t join_sort_key(int32_t key, int32_t spay) {
    return ((uint64_t)((uint32_t)key ^ 0x80000000u) << 32) | ((uint32_t)spay ^ 0x80000000u);
}
__host__ __device__ __forceinline__ int32_t join_sort_key_key(uint64_t k) {
    return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
}
__host__ __device__ __forceinline__ int32_t join_sort_key_spay(uint64_t k) {
    return (int32_t)((uint32_t)k ^ 0x80000000u);
}

}  // namespace qo
