// gather_probe: random 8-byte gathers from a 256 MiB array under the load
// forms a gather-fold could use -- plain global loads, non-temporal loads and
// buffer loads with each combination of the sc0 / sc1 / nt cache-policy bits --
// and, for reference, the same loads in sequence (a = 1).  Prints each form's
// time and its rate counted in logical bytes and in 128-byte lines.  Run
// under rocprofv3 --pmc TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum
// TCC_EA0_RDREQ_128B_sum it shows the size of the HBM read requests each
// form produces.  Built by build() into tests/gather_probe.
//
// Then two measurements behind the gather-fold's design: plain random
// gathers from arrays of 8 MiB to 1 GiB (does a smaller footprint, one
// signal's x in the Infinity Cache, raise the line rate?), and the
// transpose of a gather -- a contiguous read whose samples are written to
// random places, alone and as a filter over a 4x larger contiguous read
// (the first phase of a form that would stream x instead of gathering it).
//
// Every index is (a u + c) & (n - 1) < n, and buffer loads are also range-
// checked by the hardware against the array's size.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kPer = 16;  // gathers per thread
constexpr int kThreads = 256;

template <int V>
__device__ __forceinline__ unsigned long long Load(const unsigned long long *p,
                                                   __amdgpu_buffer_rsrc_t r, uint32_t i) {
  if constexpr (V == 0) {
    return p[i];
  } else if constexpr (V == 1) {
    return __builtin_nontemporal_load(p + i);
  } else {
    // cache-policy operand: bit 0 sc0, bit 1 nt, bit 4 sc1
    constexpr int kAux[] = {0, 1, 2, 3, 16, 17, 18, 19};
    return __builtin_bit_cast(unsigned long long,
                              __builtin_amdgcn_raw_buffer_load_b64(r, i * 8u, 0, kAux[V - 2]));
  }
}

template <int V>
__global__ void __launch_bounds__(kThreads)
    k_gather(const unsigned long long *p, uint32_t n, uint32_t a, uint32_t c,
             unsigned long long *out) {
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<unsigned long long *>(p), 0, int(n * 8u),
                                        0x00020000);
  const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
  unsigned long long s = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) s ^= Load<V>(p, r, (a * (t * kPer + j) + c) & (n - 1));
  out[t] = s;
}

// contiguous read of the first kPer * grid samples, each written to a
// random place of out (n elements)
__global__ void __launch_bounds__(kThreads)
    k_scatter(const unsigned long long *in, uint32_t n, uint32_t a, uint32_t c,
              unsigned long long *out) {
  const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = t + j * gridDim.x * kThreads;
    out[(a * i + c) & (n - 1)] = in[i];
  }
}

// contiguous read of every sample of in (n elements); those whose permuted
// index falls in the first quarter are written there
__global__ void __launch_bounds__(kThreads)
    k_filter_scatter(const unsigned long long *in, uint32_t n, uint32_t a, uint32_t c,
                     unsigned long long *out) {
  const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = t + j * gridDim.x * kThreads;
    const uint32_t v = (a * i + c) & (n - 1);
    const unsigned long long x = in[i];
    if (v < n / 4) out[v] = x;
  }
}

#define CHECK(x)                                                                   \
  do {                                                                             \
    const hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                        \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);   \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

template <int V>
int Time(const char *name, const unsigned long long *p, uint32_t n, uint32_t a, uint32_t gathers,
         unsigned long long *out) {
  const uint32_t blocks = gathers / kPer / kThreads;
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_gather<V>, dim3(blocks), dim3(kThreads), 0, 0, p, n, a, 12345u, out);
  CHECK(hipGetLastError());
  constexpr int kReps = 5;
  CHECK(hipEventRecord(e0));
  for (int i = 0; i < kReps; ++i)
    hipLaunchKernelGGL(k_gather<V>, dim3(blocks), dim3(kThreads), 0, 0, p, n, a, 12345u + i, out);
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double sec = ms / kReps / 1e3;
  std::printf("%-24s a=%-10u %8.1f us  logical %7.1f GB/s  as 128-byte lines %7.1f GB/s\n", name, a,
              sec * 1e6, gathers * 8.0 / sec / 1e9, gathers * 128.0 / sec / 1e9);
  std::fflush(stdout);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

// Best of 5 timed launches (after one warm-up) of kernel K over `blocks`.
template <typename K, typename... Args>
int Best(float *us, K kernel, uint32_t blocks, Args... args) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  float best = 1e30f;
  for (int rep = 0; rep < 6; ++rep) {
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kThreads), 0, 0, args...);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0 && ms < best) best = ms;
  }
  *us = best * 1e3f;
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

}  // namespace

int main() {
  const uint32_t n = 1u << 25;        // 256 MiB of 8-byte samples
  const uint32_t gathers = 1u << 23;  // per launch
  unsigned long long *p = nullptr, *out = nullptr;
  CHECK(hipMalloc(&p, size_t(n) * 8));
  CHECK(hipMalloc(&out, size_t(gathers / kPer) * 8));
  CHECK(hipMemset(p, 1, size_t(n) * 8));
  CHECK(hipDeviceSynchronize());
  for (const uint32_t a : {1u, 2654435761u}) {  // odd: a permutation of [0, n)
    if (Time<0>("global", p, n, a, gathers, out) || Time<1>("global nt", p, n, a, gathers, out) ||
        Time<2>("buffer", p, n, a, gathers, out) || Time<3>("buffer sc0", p, n, a, gathers, out) ||
        Time<4>("buffer nt", p, n, a, gathers, out) ||
        Time<5>("buffer sc0 nt", p, n, a, gathers, out) ||
        Time<6>("buffer sc1", p, n, a, gathers, out) ||
        Time<7>("buffer sc0 sc1", p, n, a, gathers, out) ||
        Time<8>("buffer sc1 nt", p, n, a, gathers, out) ||
        Time<9>("buffer sc0 sc1 nt", p, n, a, gathers, out))
      return 1;
  }
  CHECK(hipFree(p));
  CHECK(hipFree(out));

  // plain random gathers by array size, 2^23 per launch
  const uint32_t nmax = 1u << 27;  // 1 GiB
  unsigned long long *big = nullptr, *dst = nullptr;
  CHECK(hipMalloc(&big, size_t(nmax) * 8));
  CHECK(hipMalloc(&dst, size_t(n) * 8));
  CHECK(hipMemset(big, 1, size_t(nmax) * 8));
  CHECK(hipMemset(dst, 0, size_t(n) * 8));
  CHECK(hipMalloc(&out, size_t(gathers / kPer) * 8));
  CHECK(hipDeviceSynchronize());
  const uint32_t ar = 2654435761u, blocks = gathers / kPer / kThreads;
  for (uint32_t logn = 20; logn <= 27; ++logn) {
    float us = 0;
    if (Best(&us, k_gather<0>, blocks, (const unsigned long long *)big, 1u << logn, ar, 12345u,
             out))
      return 1;
    std::printf("size %5u MiB  random gathers %8.1f us  as 128-byte lines %7.1f GB/s\n",
                (1u << logn) >> 17, us, gathers * 128.0 / us / 1e3);
  }
  // scatters: 2^23 samples read in order, written at random into n; then
  // all n read in order, a quarter written at random
  float us = 0;
  if (Best(&us, k_scatter, blocks, (const unsigned long long *)big, n, ar, 12345u, dst)) return 1;
  std::printf("scatter %u contiguous reads to random 8-byte writes %8.1f us  %6.1f G samples/s\n",
              gathers, us, gathers / us / 1e3);
  if (Best(&us, k_filter_scatter, n / kPer / kThreads, (const unsigned long long *)big, n, ar,
           12345u, dst))
    return 1;
  std::printf("filter %u contiguous reads, %u random 8-byte writes %8.1f us  %6.1f G samples/s\n",
              n, n / 4, us, (n / 4) / us / 1e3);
  CHECK(hipFree(big));
  CHECK(hipFree(dst));
  CHECK(hipFree(out));
  std::printf("done\n");
  return 0;
}
