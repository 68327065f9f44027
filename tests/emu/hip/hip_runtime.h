// Host emulation of the HIP subset used by mpsft_amd/csrc, for running the
// library's kernels on the CPU under AddressSanitizer (test infrastructure:
// it finds out-of-bounds indexing, short LDS allocations and grid/shape
// mismatches without a GPU).  It is never part of the shipped library.
//
// Model: a launch runs its workgroups one after another; the threads of a
// workgroup are fibers on one device worker thread, switched at barriers.
// __syncthreads is a barrier over the threads still running; wave
// operations (__shfl, __ballot, __any) are a barrier over the 64 lanes of
// the wave plus an exchange buffer.  Dynamic LDS is a
// heap block of exactly the launch's byte count, so ASan reports any read
// or write past it.  Launches, graph launches, hipMemsetAsync and event
// records queue on the worker, which runs them in order while the host
// goes on; synchronous calls and hipMemcpyAsync wait for the queue first.
// Stream capture records launches into a graph that hipGraphLaunch queues.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define HIP_DYNAMIC_SHARED(type, var) type *var = static_cast<type *>(emu::DynShared());

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
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
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline double2 make_double2(double x, double y) { return double2{x, y}; }
inline float2 make_float2(float x, float y) { return float2{x, y}; }

using std::max;
using std::min;

namespace emu {
struct Index {
  unsigned x, y, z;
};
extern thread_local Index tid, bid;
extern thread_local dim3 bdim, gdim;
void *DynShared();
void SyncThreads();
double ShflD(double v, int src);
int ShflI(int v, int src);
uint64_t Ballot(int pred);
void Launch(void *stream, dim3 grid, dim3 block, size_t lds, std::function<void()> body);
}  // namespace emu

#define threadIdx emu::tid
#define blockIdx emu::bid
#define blockDim emu::bdim
#define gridDim emu::gdim

inline void __syncthreads() { emu::SyncThreads(); }
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline double __shfl(double v, int src, int = 64) { return emu::ShflD(v, src & 63); }
inline int __shfl(int v, int src, int = 64) { return emu::ShflI(v, src & 63); }
inline double __shfl_xor(double v, int mask, int = 64) {
  const int lane = int(emu::tid.x & 63);
  return emu::ShflD(v, (lane ^ mask) & 63);
}
inline uint64_t __ballot(int pred) { return emu::Ballot(pred); }
inline int __any(int pred) { return emu::Ballot(pred) != 0; }
inline int __clz(int v) { return v ? __builtin_clz((unsigned)v) : 32; }
inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }
inline unsigned __brev(unsigned v) {
  unsigned r = 0;
  for (int i = 0; i < 32; ++i) r |= ((v >> i) & 1u) << (31 - i);
  return r;
}

inline double sinpi(double x) { return std::sin(M_PI * std::fmod(x, 2.0)); }
inline void sincospi(double x, double *s, double *c) {
  const double r = std::fmod(x, 2.0);
  *s = std::sin(M_PI * r);
  *c = std::cos(M_PI * r);
}

inline double atomicAdd(double *p, double v) {
  uint64_t old = __atomic_load_n(reinterpret_cast<uint64_t *>(p), __ATOMIC_SEQ_CST);
  for (;;) {
    double cur;
    std::memcpy(&cur, &old, 8);
    const double nv = cur + v;
    uint64_t want;
    std::memcpy(&want, &nv, 8);
    if (__atomic_compare_exchange_n(reinterpret_cast<uint64_t *>(p), &old, want, false,
                                    __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST))
      return cur;
  }
}
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned atomicAdd(unsigned *p, unsigned v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicMax(int *p, int v) {
  int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (old < v &&
         !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}
inline int atomicOr(int *p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicOr(unsigned *p, unsigned v) {
  return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
#define __HIP_MEMORY_SCOPE_SYSTEM 0
#define __HIP_MEMORY_SCOPE_AGENT 1
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), (order))
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), (order))
#define __builtin_nontemporal_load(p) (*(p))

// ---------------------------------------------------------------- runtime
typedef int hipError_t;
enum : int {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotSupported = 801,
};
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4,
};
enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2,
};
constexpr unsigned hipStreamNonBlocking = 1;
constexpr unsigned hipHostMallocDefault = 0;
constexpr unsigned hipHostMallocMapped = 2;
constexpr unsigned hipHostMallocCoherent = 0x40000000;

typedef struct EmuStream *hipStream_t;
typedef struct EmuEvent *hipEvent_t;
typedef struct EmuGraph *hipGraph_t;
typedef struct EmuGraph *hipGraphExec_t;
typedef struct EmuGraphNode *hipGraphNode_t;

#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
  emu::Launch((stream), (grid), (block), (lds), [=]() { kernel(__VA_ARGS__); })

hipError_t hipGetLastError();
hipError_t hipGetDeviceCount(int *count);
hipError_t hipGetDevice(int *device);
// A file-scope __device__ array is host memory here: its address is itself.
#define HIP_SYMBOL(sym) (reinterpret_cast<const void *>(&(sym)))
#define __align__(n) __attribute__((aligned(n)))
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
hipError_t hipMalloc(void **p, size_t bytes);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags = 0);
hipError_t hipHostFree(void *p);
hipError_t hipHostGetDevicePointer(void **dev, void *host, unsigned flags);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                          hipStream_t s);
hipError_t hipMemcpy2D(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                       size_t height, hipMemcpyKind kind);
hipError_t hipMemset(void *dst, int value, size_t bytes);
hipError_t hipMemsetAsync(void *dst, int value, size_t bytes, hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err_node,
                               char *log, size_t log_size);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
