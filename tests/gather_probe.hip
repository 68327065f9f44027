// gather_probe: random 8-byte gathers from a 256 MiB array under the load
// forms a gather-fold could use -- plain global loads, non-temporal loads and
// buffer loads with each combination of the sc0 / sc1 / nt cache-policy bits --
// and, for reference, the same loads in sequence (a = 1).  Prints each form's
// time and its rate counted in logical bytes and in 128-byte lines.  Run
// under rocprofv3 --pmc TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum
// TCC_EA0_RDREQ_128B_sum it shows the size of the HBM read requests each
// form produces.  Built by build() into tests/gather_probe.
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
  std::printf("done\n");
  return 0;
}
