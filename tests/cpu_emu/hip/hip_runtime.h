// CPU emulation of the parts of HIP the library uses, so that the library's
// kernel and host sources build with g++ and run on the CPU under
// AddressSanitizer / UndefinedBehaviorSanitizer.
//
// Execution model.  A launch runs its workgroups on up to EMU_RESIDENT
// concurrent OS threads, which take workgroup indices from a counter as a
// GPU's dispatcher does (so a kernel whose workgroups wait on each other,
// within the resident limit, makes progress).  Inside a workgroup every
// thread is a fiber (ucontext) of that OS thread; a fiber runs until it
// reaches __syncthreads, a wave operation (__shfl_xor, __shfl_up, __ballot)
// or its end.  A barrier is released when every thread of the workgroup that
// has not ended waits at one; a wave operation when every live lane of its
// 64-lane wave waits at one.  Anything else is a deadlock of the kernel --
// for example a wave operation reached by only some lanes of a wave, or a
// barrier reached by only some waves -- and the emulation reports it and
// aborts.  __shared__ variables are thread_local statics: one copy per OS
// thread, i.e. per running workgroup.  Device memory is heap memory (the
// sanitizer checks every access against the allocation), streams execute
// synchronously in issue order, and graph capture records the captured
// operations for replay.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

// ------------------------------------------------------------ qualifiers
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __align__(n) __attribute__((aligned(n)))
#define __restrict__ __restrict

using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::isfinite;
using std::isinf;
using std::isnan;
using std::ldexp;
using std::sqrt;

template <class A, class B> inline typename std::common_type<A, B>::type min(A a, B b)
{
    typedef typename std::common_type<A, B>::type C;
    return (C)b < (C)a ? (C)b : (C)a;
}
template <class A, class B> inline typename std::common_type<A, B>::type max(A a, B b)
{
    typedef typename std::common_type<A, B>::type C;
    return (C)a < (C)b ? (C)b : (C)a;
}

// ------------------------------------------------------------ types
struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

struct double2 {
    double x, y;
};
struct float2 {
    float x, y;
};
struct int2 {
    int x, y;
};

typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorInvalidDevice = 101,
    hipErrorLaunchFailure = 719,
    hipErrorUnknown = 999,
} hipError_t;

enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
};

enum hipDeviceAttribute_t { hipDeviceAttributeWallClockRate = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };

#define hipStreamDefault 0u
#define hipStreamNonBlocking 1u
#define hipEventDefault 0u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u

struct emu_stream;
struct emu_event;
struct emu_graph;
typedef emu_stream *hipStream_t;
typedef emu_event *hipEvent_t;
typedef emu_graph *hipGraph_t;
typedef emu_graph *hipGraphExec_t;
typedef void *hipGraphNode_t;

struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t totalGlobalMem;
    int warpSize;
    int maxThreadsPerBlock;
};

// ------------------------------------------------------------ runtime API
hipError_t hipMalloc(void **p, size_t n);
template <class T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags = 0);
template <class T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned flags = 0)
{
    return hipHostMalloc((void **)p, n, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind k, hipStream_t s = nullptr);
hipError_t hipMemset(void *p, int v, size_t n);
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t s = nullptr);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int d);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *err, char *log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t x);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
namespace emu {
constexpr int EMU_OCCUPANCY_REPORTED = 2;
}
template <class F> inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F, int, size_t)
{
    *n = emu::EMU_OCCUPANCY_REPORTED;  // workgroups per CU
    return hipSuccess;
}
#define HIP_SYMBOL(x) ((const void *)&(x))

// ------------------------------------------------------------ emulation core
namespace emu {

constexpr int WAVE = 64;
constexpr int EMU_CUS = 4;        // multiProcessorCount reported
constexpr int EMU_RESIDENT = 16;  // workgroups of one launch that run at once (>= CUs x occupancy)

struct Fiber;
struct Block;

// the running fiber and workgroup of this OS thread
extern thread_local Fiber *t_fiber;
extern thread_local Block *t_block;

dim3 &thread_idx();
dim3 &block_idx();
dim3 &block_dim();
dim3 &grid_dim();

// synchronisation points (a fiber yields to its workgroup's scheduler)
int sync_block(int pred, int op);  // op 0: none, 1: and, 2: or, 3: count
unsigned long long wave_exchange(unsigned long long v, int op, int arg);  // op 0 xor-shfl, 1 up-shfl, 2 ballot

// run body() for every thread of every workgroup of the grid
void run_grid(dim3 grid, dim3 block, const std::function<void()> &body);
// enqueue an operation on a stream (runs now, or is recorded while capturing)
hipError_t submit(hipStream_t s, std::function<void()> op);
long long now_ticks();  // wall clock at 100 MHz

template <class T> inline unsigned long long bits_of(T v)
{
    static_assert(sizeof(T) <= 8, "wave operations move at most 8 bytes");
    unsigned long long u = 0;
    std::memcpy(&u, &v, sizeof(T));
    return u;
}
template <class T> inline T from_bits(unsigned long long u)
{
    T v;
    std::memcpy(&v, &u, sizeof(T));
    return v;
}

template <class... P, class... A>
hipError_t launch(void (*k)(P...), dim3 grid, dim3 block, size_t, hipStream_t s, A &&...a)
{
    std::tuple<typename std::decay<P>::type...> args((P)a...);
    return submit(s, [k, grid, block, args]() {
        run_grid(grid, block, [&]() { std::apply(k, args); });
    });
}

template <class... P, class... A>
hipError_t launch_ext(void (*k)(P...), dim3 grid, dim3 block, size_t sh, hipStream_t s, hipEvent_t start,
                      hipEvent_t stop, unsigned, A &&...a)
{
    if (start) (void)hipEventRecord(start, s);
    hipError_t e = launch(k, grid, block, sh, s, std::forward<A>(a)...);
    if (stop) (void)hipEventRecord(stop, s);
    return e;
}

}  // namespace emu

#define threadIdx (emu::thread_idx())
#define blockIdx (emu::block_idx())
#define blockDim (emu::block_dim())
#define gridDim (emu::grid_dim())

#define hipLaunchKernelGGL(k, grid, block, sh, st, ...) (void)emu::launch(k, grid, block, sh, st, ##__VA_ARGS__)
#define hipExtLaunchKernelGGL(k, grid, block, sh, st, ev0, ev1, fl, ...) \
    (void)emu::launch_ext(k, grid, block, sh, st, ev0, ev1, fl, ##__VA_ARGS__)

inline void __syncthreads() { (void)emu::sync_block(0, 0); }
inline int __syncthreads_and(int p) { return emu::sync_block(p, 1); }
inline int __syncthreads_or(int p) { return emu::sync_block(p, 2); }
inline int __syncthreads_count(int p) { return emu::sync_block(p, 3); }

template <class T> inline T __shfl_xor(T v, int mask, int width = 64)
{
    (void)width;
    return emu::from_bits<T>(emu::wave_exchange(emu::bits_of(v), 0, mask));
}
template <class T> inline T __shfl_up(T v, unsigned d, int width = 64)
{
    (void)width;
    return emu::from_bits<T>(emu::wave_exchange(emu::bits_of(v), 1, (int)d));
}
inline unsigned long long __ballot(int p) { return emu::wave_exchange(p ? 1ull : 0ull, 2, 0); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }

inline long long __double_as_longlong(double d) { return emu::from_bits<long long>(emu::bits_of(d)); }
inline double __longlong_as_double(long long v) { return emu::from_bits<double>(emu::bits_of(v)); }
inline long long wall_clock64() { return emu::now_ticks(); }
inline long long clock64() { return emu::now_ticks(); }

inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(order)
inline void __threadfence_block() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }

template <class T> inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicMax(T *p, T v)
{
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
template <class T> inline T atomicExch(T *p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicCAS(T *p, T cmp, T v)
{
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;
}

// scoped atomics of HIP-Clang: the scope has no meaning on the CPU
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#define __hip_atomic_load(p, order, scope) __atomic_load_n(p, order)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n(p, v, order)
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add(p, v, order)
#define __builtin_amdgcn_s_sleep(n) emu_yield()
void emu_yield();

// the kernels apply it only to values that are uniform over the wave
#define __builtin_amdgcn_readfirstlane(x) (x)

// AMDGPU math builtins the kernels call: exact CPU operations
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt(x))
#define __builtin_amdgcn_rsqf(x) (1.0f / std::sqrt(x))
#define __builtin_amdgcn_sqrtf(x) (std::sqrt(x))
