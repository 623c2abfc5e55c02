/*
 * qo_compat.h -- drop-in replacements for the reference's host operator entry
 * points (Xtra-Computing/Query_on_OpenCL_FPGA, host/host/ *.cpp), backed by
 * libqo's gfx950 kernels.
 *
 * Names, parameter order and meanings follow the reference.  `cl_mem` becomes
 * the opaque `qo_mem` (device pointer + byte size); the reference's vestigial
 * scheduling parameters (index, eventList, Kernel, Flag_CPU_GPU, burden,
 * _CPU_GPU) are accepted and ignored; numThread/numBlock are hints only.
 * Every call is synchronous, like the reference's clFinish after each launch.
 * On error a call returns -1 (int results) or leaves its outputs untouched and
 * sets qo_last_error(); with QO_COMPAT_EXIT_ON_ERROR=1 in the environment it
 * prints the error and exits like the reference's checkError.
 */
#ifndef QO_COMPAT_H
#define QO_COMPAT_H

#include "qo.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef qo_record Record;                 /* host/host/common.h:27-28 (uint2 {x, y}) */
typedef struct qo_mem_s *qo_mem;          /* stands where the reference passes cl_mem */

#ifndef REDUCE_SUM
#define REDUCE_SUM (0)
#define REDUCE_MAX (1)
#define REDUCE_MIN (2)
#define REDUCE_AVERAGE (3)
#endif
#ifndef TEST_MAX
#define TEST_MAX (1 << 30)                /* host/host/common.h:29 */
#endif

/* Device buffer handles (the role of clCreateBuffer / clReleaseMemObject). */
qo_mem qo_mem_create(size_t bytes);
qo_mem qo_mem_wrap(void *device_ptr, size_t bytes);   /* caller keeps ownership of the memory */
void *qo_mem_ptr(qo_mem m);
size_t qo_mem_size(qo_mem m);
void qo_mem_free(qo_mem m);
int qo_mem_write(qo_mem m, const void *host, size_t bytes);   /* cl_writebuffer, common.cpp:93 */
int qo_mem_read(void *host, qo_mem m, size_t bytes);          /* cl_readbuffer, common.cpp:64 */

/* Scan plan (host/host/testScan.h:4-12; testScan.cpp:26-75, 229-270).  The
 * reference's fields come first, in its order; initScan fills pass, gLength
 * and blockSize as the reference does, while outputBuffer, blockSumBuffer and
 * tempBuffer stay NULL (the single-pass scan needs no per-level buffers) and
 * `ws` holds its workspace. */
typedef struct ScanPara {
    qo_mem *outputBuffer;    /* always NULL here */
    qo_mem *blockSumBuffer;  /* always NULL here */
    qo_mem tempBuffer;       /* always NULL here */
    unsigned pass;
    unsigned gLength;
    unsigned blockSize;
    void *ws;
} ScanPara;
void initScan(int rLen, ScanPara *SP);
void closeScan(ScanPara *SP);
/* Exclusive int32 prefix sum of d_Src into d_Dst (testScan.cpp:278). */
void scanImpl(qo_mem d_Src, int rLen, qo_mem d_Dst, int *index, void *eventList, void *kernel,
              int *Flag_CPU_GPU, double *burden, ScanPara *SP, int _CPU_GPU);
/* The offset overload scanImpl(d_Src, offset, rLen, d_Dst, ...) of testScan.h:16
 * (declared there, defined nowhere in the reference) under a distinct C name:
 * scans elements [offset, offset + rLen) of d_Src into d_Dst[0, rLen). */
void scanImpl_offset(qo_mem d_Src, int offset, int rLen, qo_mem d_Dst, int *index, void *eventList, void *kernel,
                     int *Flag_CPU_GPU, double *burden, ScanPara *SP, int _CPU_GPU);

/* Selection (host/host/testFilter.cpp:92, host/host/Selection.cpp:6,19).
 * *d_Rout is allocated by the callee (rLen Records) and freed by the caller. */
void filterImpl(qo_mem d_Rin, int beginPos, int rLen, qo_mem *d_Rout, int *outSize, int numThread,
                int numBlock, int smallKey, int largeKey, int *index, void *eventList, void *Kernel,
                int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
int range_selection(qo_mem d_Rin, int rLen, int rangeSmallKey, int rangeLargeKey, qo_mem *d_Rout,
                    int numThreadPB, int numBlock, int *index, void *eventList, void *Kernel,
                    int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
/* *h_Rout is malloc'ed by the callee with outSize Records; the caller frees it. */
int CL_RangeSelection(Record *h_Rin, int rLen, int rangeSmallKey, int rangeLargeKey, Record **h_Rout,
                      int numThreadPB, int numBlock, int _CPU_GPU);

/* Point and device-buffer selection (host/host/Selection.cpp:53-129, a block the
 * reference compiles out with #if 0): point_selection is filterImpl with
 * smallKey = largeKey = matchingKeyValue; the *Only forms take device buffers
 * and allocate *d_Rout (rLen Records) for the caller to free. */
int point_selection(qo_mem d_Rin, int rLen, int matchingKeyValue, qo_mem *d_Rout, int numThreadPB, int numBlock,
                    int *index, void *eventList, void *Kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
int CL_PointSelection(Record *h_Rin, int rLen, int matchingKeyValue, Record **h_Rout, int numThreadPB,
                      int numBlock, int _CPU_GPU);
int CL_PointSelectionOnly(qo_mem d_Rin, int rLen, int matchingKeyValue, qo_mem *d_Rout, int numThreadPB,
                          int numBlock, int _CPU_GPU);
int CL_RangeSelectionOnly(qo_mem d_Rin, int rLen, int rangeSmallKey, int rangeLargeKey, qo_mem *d_Rout,
                          int numThreadPB, int numBlock, int _CPU_GPU);

/* Split (host/host/testSplit.h:6, testSplit.cpp:156-209): d_S receives d_R
 * split into numPart partitions by .y % numPart (stable within a partition);
 * d_bound, when not NULL, numPart Records {start, end}.  SPLIT_TYPE is accepted
 * and ignored, as the reference's splitImpl maps with SPLIT whatever it is
 * given (testSplit.cpp:159). */
void splitImpl(qo_mem d_R, int rLen, int numPart, qo_mem d_S, qo_mem d_bound, int numThreadPB, int numBlock,
               int SPLIT_TYPE, int *index, void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden,
               int _CPU_GPU);
/* Scatter d_S[d_loc[i]] = d_R[i] and gather d_S[i] = d_R[d_loc[i]]
 * (host/host/testScatter.h:2-7, testScatter.cpp:61-130); locations outside
 * [0, rLen) are skipped.  scatterImpl_forPart is the same scatter (numPart only
 * set the reference's number of windows). */
void scatterImpl(qo_mem d_R, int rLen, qo_mem d_loc, qo_mem d_S, int numThreadPB, int numBlock, int *index,
                 void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
void scatterImpl_forPart(qo_mem d_R, int rLen, int numPart, qo_mem d_loc, qo_mem d_S, int numThreadPB,
                         int numBlock, int *index, void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden,
                         int _CPU_GPU);
void gatherImpl(qo_mem d_R, int rLen, qo_mem d_loc, qo_mem d_S, int sLen, int numThreadPB, int numBlock,
                int *index, void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);

/* Sort ascending by .y (host/host/testSort.cpp:24, host/host/Sort.cpp:34).  Stable. */
void radixSortImpl_sm(qo_mem d_R, int rLen, int meaningless, int numThreadPB, int numBlock, int *index,
                      void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
void CL_RadixSort(Record *h_Rin, int rLen, Record *h_Rout, int numThread, int numBlock, int _CPU_GPU);
/* radixSortImpl (host/host/testSort.cpp:274): the same sort as radixSortImpl_sm. */
void radixSortImpl(qo_mem d_R, int rLen, int keybits, int numThreadPB, int numBlock, int *index, void *eventList,
                   void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
/* CL_RadixSortOnly (host/host/Sort.cpp:11): sorts the device buffer in place. */
void CL_RadixSortOnly(qo_mem d_Rin, int rLen, int numThread, int numBlock, int _CPU_GPU);
/* CL_BitonicSortOnly (host/host/Sort.cpp:20): as in the reference, copies d_Rin
 * into d_Rout first and then sorts d_Rin in place (d_Rout keeps the input). */
void CL_BitonicSortOnly(qo_mem d_Rin, int rLen, qo_mem d_Rout, int numThread, int numBlock, int _CPU_GPU);

/* Group-by (host/host/testGroupBy.cpp:81, host/host/GroupBy.cpp:7).  groupByImpl
 * sorts d_Rout in place (after copying d_Rin into it when they differ) and
 * allocates *d_startPos with numGroups ints; returns numGroups. */
int groupByImpl(qo_mem d_Rin, int rLen, qo_mem d_Rout, qo_mem *d_startPos, int numThread, int numBlock,
                int *index, void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
int CL_GroupBy(Record *h_Rin, int rLen, Record *h_Rout, int **h_startPos, int numThread, int numBlock,
               int _CPU_GPU);

/* Aggregation after group-by (host/host/testAggAfterGB.cpp:69,92; AggAfterGroupBy.cpp:30-170).
 * With d_Ragg / h_Ragg non-NULL the rows are first gathered S[i] = Ragg[Rin[i].x]
 * (gatherBeforeAgg, testAggAfterGB.cpp:58-67); with NULL the groups reduce
 * Rin[i].x directly, as testAggAfterGroupByImpl does.  Results are int32 with
 * the reference's semantics (SUM wraps, MAX from 0, MIN from TEST_MAX, AVG
 * truncates). */
void parallelAggregate_init(qo_mem d_S, qo_mem d_startPos, qo_mem d_aggResults, int OPERATOR, int blockOffset,
                            int numGroups, int numBlock, int numThread, int sharedMemSize, int rLen, int *index,
                            void *eventList, void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
void aggAfterGroupByImpl(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                         qo_mem d_aggResults, int OPERATOR, int numThread, int *index, void *eventList,
                         void *kernel, int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
/* Device-buffer forms (host/host/AggAfterGroupBy.cpp:9-24): aggAfterGroupByImpl
 * with REDUCE_MAX / MIN / SUM / AVERAGE. */
void agg_max_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
void agg_min_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
void agg_sum_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
void agg_avg_afterGroupBy(qo_mem d_Rin, int rLen, qo_mem d_startPos, int numGroups, qo_mem d_Ragg,
                          qo_mem d_aggResults, int numThread, int *index, void *eventList, void *kernel,
                          int *Flag_CPU_GPU, double *burden, int _CPU_GPU);
void CL_agg_max_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg,
                             int *h_aggResults, int numThread, int _CPU_GPU);
void CL_agg_min_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg,
                             int *h_aggResults, int numThread, int _CPU_GPU);
void CL_agg_sum_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg,
                             int *h_aggResults, int numThread, int _CPU_GPU);
void CL_agg_avg_afterGroupBy(Record *h_Rin, int rLen, int *h_startPos, int numGroups, Record *h_Ragg,
                             int *h_aggResults, int numThread, int _CPU_GPU);

/* Hash join with the kernel-shaped parameters of kernel/hj_filter_sort/1/
 * hj_filter_sort_op.cl:27 and :64 (device pointers; interleaved (key, val) uint
 * pairs; B a power of two).  buildHashTable zeroes rHashCount itself.
 * probeHashTable writes COMPACTED (s_val, r_val) pairs and returns their count
 * (matchedTable must hold 2*size uints, as the reference sizes it). */
int buildHashTable(uint32_t *rTableOnDevice, uint32_t *rHashTable, uint32_t size, uint32_t rHashTableBucketNum,
                   uint32_t hashBucketSize, uint32_t *rHashCount);
int probeHashTable(uint32_t *rHashTable, uint32_t *sTableOnDevice, uint32_t *matchedTable, uint32_t size,
                   uint32_t rHashTableBucketNum, uint32_t hashBucketSize, uint32_t *rHashCount);

#ifdef __cplusplus
}
#endif
#endif /* QO_COMPAT_H */
