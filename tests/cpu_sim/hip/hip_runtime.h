// CPU build of libqo's HIP sources (test infrastructure).
//
// This header stands in for <hip/hip_runtime.h> when csrc/*.hip is compiled
// as host C++ by tests/cpu_sim/build.py.  Kernels run on the CPU: every
// work-item is a fiber, a work-group runs its fibers to completion before the
// next work-group starts, __syncthreads is a fiber barrier and the wave
// intrinsics (ballot, shuffles, readfirstlane) are collectives over the 64
// fibers of a wave.  The HIP runtime calls act synchronously on host memory,
// so a build with -fsanitize=address checks every device access a kernel
// makes.  It models no timing and no concurrency between work-groups: a spin
// wait that only another running work-group could end is reported as a hang.
#pragma once

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <type_traits>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 0

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

struct alignas(16) int4 { int x, y, z, w; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
struct alignas(8) int2 { int x, y; };
struct alignas(8) uint2 { unsigned x, y; };

namespace simt {

struct Idx {
    unsigned x, y, z;
};
Idx &thread_idx();
extern Idx g_block_idx, g_block_dim, g_grid_dim;

void syncthreads();
void sleep_yield();
unsigned lane_of_current();
// Wave collective: every live lane of the current wave contributes `v`
// (and `a`); returns this lane's result of op over the wave.
enum WaveOp { W_BALLOT, W_SHFL, W_SHFL_UP, W_SHFL_XOR, W_FIRST, W_SYNC };
uint64_t wave_op(WaveOp op, uint64_t v, int64_t a, int width, int line);
unsigned char *dyn_lds();

template <class T>
inline uint64_t bits_of(T v) {
    static_assert(sizeof(T) <= 8, "wave values are at most 64 bits");
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}
template <class T>
inline T from_bits(uint64_t b) {
    T v;
    std::memcpy(&v, &b, sizeof(T));
    return v;
}

struct LaunchCfg {
    dim3 grid, block;
    size_t lds;
    void *stream;
};
inline LaunchCfg cfg(dim3 g, dim3 b, size_t lds = 0, void *st = nullptr) { return LaunchCfg{g, b, lds, st}; }
// Runs (or, under stream capture, records) a kernel: body() is called once per
// work-item.
void launch(const LaunchCfg &c, const char *name, std::function<void()> body);
// Canaries after carved buffers (qo_sim_canary_enable / qo_sim_canary_check).
bool canary_on();
void canary_set(void *p, size_t n, size_t buf_bytes, const char *file = "", int line = 0);
// Forgets the canaries inside [p, p + n) (memory about to be written whole).
void canary_retire(const void *p, size_t n);

}  // namespace simt

#define threadIdx (::simt::thread_idx())
#define blockIdx (::simt::g_block_idx)
#define blockDim (::simt::g_block_dim)
#define gridDim (::simt::g_grid_dim)

inline void __syncthreads() { ::simt::syncthreads(); }
inline void __threadfence() {}
inline unsigned __lane_id() { return ::simt::lane_of_current(); }
inline uint64_t __ballot_impl(bool p, int line) { return ::simt::wave_op(::simt::W_BALLOT, p ? 1 : 0, 0, 64, line); }
#define __ballot(p) __ballot_impl((p), __LINE__)

template <class T>
inline T __shfl_impl(T v, int src, int width, int line) {
    return ::simt::from_bits<T>(::simt::wave_op(::simt::W_SHFL, ::simt::bits_of(v), src, width, line));
}
template <class T>
inline T __shfl_up_impl(T v, unsigned d, int width, int line) {
    return ::simt::from_bits<T>(::simt::wave_op(::simt::W_SHFL_UP, ::simt::bits_of(v), d, width, line));
}
template <class T>
inline T __shfl_xor_impl(T v, int m, int width, int line) {
    return ::simt::from_bits<T>(::simt::wave_op(::simt::W_SHFL_XOR, ::simt::bits_of(v), m, width, line));
}
template <class T>
inline T __readfirstlane_impl(T v, int line) {
    return ::simt::from_bits<T>(::simt::wave_op(::simt::W_FIRST, ::simt::bits_of(v), 0, 64, line));
}
#define __shfl(v, s, ...) __shfl_impl((v), (s), 64, __LINE__)
#define __shfl_up(v, d, ...) __shfl_up_impl((v), (d), 64, __LINE__)
#define __shfl_xor(v, m, ...) __shfl_xor_impl((v), (m), 64, __LINE__)
#define __builtin_amdgcn_readfirstlane(v) __readfirstlane_impl((v), __LINE__)
#define __builtin_amdgcn_s_sleep(n) ::simt::sleep_yield()
#define __builtin_amdgcn_wave_barrier() ((void)::simt::wave_op(::simt::W_SYNC, 0, 0, 64, __LINE__))

inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __clz(int v) { return v ? __builtin_clz((unsigned)v) : 32; }
inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }

template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
    typedef typename std::common_type<A, B>::type C;
    return (C)a < (C)b ? (C)a : (C)b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
    typedef typename std::common_type<A, B>::type C;
    return (C)a > (C)b ? (C)a : (C)b;
}

// Atomics: work-items switch only at barriers and wave collectives, so a
// read-modify-write is indivisible here.
template <class T, class U>
inline T atomicAdd(T *p, U v) {
    T o = *p;
    *p = (T)(o + (T)v);
    return o;
}
template <class T, class U>
inline T atomicSub(T *p, U v) {
    T o = *p;
    *p = (T)(o - (T)v);
    return o;
}
template <class T, class U, class V>
inline T atomicCAS(T *p, U cmp, V val) {
    T o = *p;
    if (o == (T)cmp) *p = (T)val;
    return o;
}
template <class T, class U>
inline T atomicExch(T *p, U v) {
    T o = *p;
    *p = (T)v;
    return o;
}
template <class T, class U>
inline T atomicMin(T *p, U v) {
    T o = *p;
    if ((T)v < o) *p = (T)v;
    return o;
}
template <class T, class U>
inline T atomicMax(T *p, U v) {
    T o = *p;
    if ((T)v > o) *p = (T)v;
    return o;
}
template <class T, class U>
inline T atomicOr(T *p, U v) {
    T o = *p;
    *p = (T)(o | (T)v);
    return o;
}
template <class T>
inline T simt_atomic_load(const T *p) {
    return *p;
}
template <class T, class U>
inline void simt_atomic_store(T *p, U v) {
    *p = (T)v;
}
#define __hip_atomic_load(p, order, scope) simt_atomic_load(p)
#define __hip_atomic_store(p, v, order, scope) simt_atomic_store((p), (v))

// ---- runtime API (synchronous, host memory) ----
typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorInvalidConfiguration = 9,
    hipErrorLaunchFailure = 719,
} hipError_t;
typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
} hipMemcpyKind;
typedef enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2
} hipStreamCaptureMode;
typedef enum hipStreamCaptureStatus {
    hipStreamCaptureStatusNone = 0,
    hipStreamCaptureStatusActive = 1,
    hipStreamCaptureStatusInvalidated = 2
} hipStreamCaptureStatus;
#define hipEventDisableTiming 2
struct ihipStream_t;
typedef ihipStream_t *hipStream_t;
struct ihipEvent_t;
typedef ihipEvent_t *hipEvent_t;
struct ihipGraph;
typedef ihipGraph *hipGraph_t;
struct hipGraphExec;
typedef hipGraphExec *hipGraphExec_t;
struct hipGraphNode;
typedef hipGraphNode *hipGraphNode_t;
#define hipStreamPerThread ((hipStream_t)2)
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
#define HIP_SYMBOL(x) (&(x))

struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t sharedMemPerBlock;
    size_t totalGlobalMem;
    int warpSize;
};

hipError_t hipGetDeviceCount(int *n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipMalloc(void **p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T **p, size_t bytes) {
    return hipMalloc(reinterpret_cast<void **>(p), bytes);
}
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
template <class T>
inline hipError_t hipHostMalloc(T **p, size_t bytes, unsigned flags = 0) {
    return hipHostMalloc(reinterpret_cast<void **>(p), bytes, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemset(void *d, int v, size_t n);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamQuery(hipStream_t s);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags = 0);
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *status);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err, char *log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
