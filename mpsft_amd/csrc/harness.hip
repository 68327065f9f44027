// harness.hip -- the reference's solve harness on the device: PostAnalyze
// (src/demo1.cc:61-112), the accuracy report of a solve against the signal's
// x-hat.  Nothing here is on the recovery loop's path.
//
// PostAnalyze compares the found modes with the truth and with x-hat:
//   per found key k: err = |c_k - xhat_k|, a hit when k is a true mode, a
//   miss otherwise; hit/miss counts, mean err and rms err of each kind;
//   over every location i: err_i = |xh2_i - xhat_i| with xh2 the found
//   coefficients (0 elsewhere): mean, rms and max.
// The solver's found-mode map is dense (coef over all n locations, zero
// where nothing was found), so xh2 is read directly.  Each workgroup sums
// a fixed, strided share of the work and reduces it in a fixed tree; one
// workgroup then adds the workgroups' partial sums in index order, so the
// result is the same on every run (bitwise).  Bound: HBM, one pass over
// coef and x-hat (2 x 16 B per location).
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace mpsft {

// The no-allocation form (mpsft_solver_post_analyze_noalloc): the partial
// sums of kPaBlocks workgroups in one file-scope device array, 40 KiB on
// every device the library runs on, held until the process ends.  Every
// call on a device shares it (each call rewrites all of it before reading
// it); calls that overlap on one device race.
static __device__ __align__(256) double scratch[kPaBlocks * kPaFields];

namespace {

// The array's device address on each device, looked up once per device on
// its first call (the array's name is a host-side stand-in, not a device
// address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

hipError_t PostAnalyzeScratch(double **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return hipSuccess;
}

namespace {

// Partial sums per workgroup (kPaFields of them): hit, hit l1, hit l2,
// miss, miss l1, miss l2, l1, l2, max, truth keys out of order.

__device__ __forceinline__ double Err(double2 a, double2 b) { return hypot(a.x - b.x, a.y - b.y); }

// true when k is one of truth[0..m) (ascending)
__device__ __forceinline__ bool InSorted(const int64_t *truth, int64_t m, int64_t k) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (truth[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < m && truth[lo] == k;
}

// Tree reduction of v[kPaFields] over the workgroup (fields 8 by max, the
// rest by sum) in LDS; thread 0 ends with the totals.
__device__ void BlockReduce(double (&v)[kPaFields], double (*sh)[kThreads]) {
#pragma unroll
  for (int f = 0; f < kPaFields; ++f) sh[f][threadIdx.x] = v[f];
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
#pragma unroll
      for (int f = 0; f < kPaFields; ++f) {
        const double o = sh[f][threadIdx.x + w];
        sh[f][threadIdx.x] = f == 8 ? fmax(sh[f][threadIdx.x], o) : sh[f][threadIdx.x] + o;
      }
    __syncthreads();
  }
#pragma unroll
  for (int f = 0; f < kPaFields; ++f) v[f] = sh[f][0];
}

__global__ void __launch_bounds__(kThreads)
    k_post_analyze(Ctx c, const int64_t *truth, int64_t m, const double2 *xhat,
                   double *partial) {
  __shared__ double sh[kPaFields][kThreads];
  double v[kPaFields];
#pragma unroll
  for (int f = 0; f < kPaFields; ++f) v[f] = 0.0;
  const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * blockDim.x;
  const int64_t cnt = min<int64_t>(c.st->count, c.cap);
  for (int64_t i = gid; i < cnt; i += gsize) {
    const int64_t k = c.keys[i];
    const double e = Err(c.coef[k], xhat[k]);
    const int o = InSorted(truth, m, k) ? 0 : 3;
    v[o] += 1.0;
    v[o + 1] += e;
    v[o + 2] += e * e;
  }
  for (int64_t i = gid; i < c.n; i += gsize) {
    const double e = Err(c.coef[i], xhat[i]);
    v[6] += e;
    v[7] += e * e;
    v[8] = fmax(v[8], e);
  }
  for (int64_t i = gid; i + 1 < m; i += gsize)
    if (!(truth[i] < truth[i + 1])) v[9] = 1.0;
  BlockReduce(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int f = 0; f < kPaFields; ++f) partial[int64_t(blockIdx.x) * kPaFields + f] = v[f];
}

// stats[0..8] = {hit, hit l1, hit l2, miss, miss l1, miss l2, l1, l2, max}
// as PostAnalyze prints them (means and rms: a kind with no entries gives
// 0/0, as there); stats[9] = 1 when the truth keys were not ascending.
__global__ void __launch_bounds__(kThreads)
    k_post_analyze_final(const double *partial, int32_t blocks, int64_t n, double *stats) {
  __shared__ double sh[kPaFields][kThreads];
  double v[kPaFields];
#pragma unroll
  for (int f = 0; f < kPaFields; ++f) v[f] = 0.0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x)
#pragma unroll
    for (int f = 0; f < kPaFields; ++f) {
      const double o = partial[int64_t(b) * kPaFields + f];
      v[f] = f == 8 ? fmax(v[f], o) : v[f] + o;
    }
  BlockReduce(v, sh);
  if (threadIdx.x != 0) return;
  stats[0] = v[0];
  stats[1] = v[1] / v[0];
  stats[2] = sqrt(v[2] / v[0]);
  stats[3] = v[3];
  stats[4] = v[4] / v[3];
  stats[5] = sqrt(v[5] / v[3]);
  stats[6] = v[6] / double(n);
  stats[7] = sqrt(v[7] / double(n));
  stats[8] = v[8];
  stats[9] = v[9] > 0 ? 1.0 : 0.0;
}

}  // namespace

hipError_t LaunchPostAnalyze(const Ctx &c, const int64_t *truth, int64_t m, const double2 *xhat,
                             double *partial, int32_t blocks, double *stats, hipStream_t s) {
  hipLaunchKernelGGL(k_post_analyze, dim3(blocks), dim3(kThreads), 0, s, c, truth, m, xhat,
                     partial);
  hipLaunchKernelGGL(k_post_analyze_final, dim3(1), dim3(kThreads), 0, s, partial, blocks, c.n,
                     stats);
  return hipGetLastError();
}

}  // namespace mpsft
