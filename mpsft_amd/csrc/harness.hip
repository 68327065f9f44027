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
// where nothing was found, and a presence bit each), so xh2 is read
// directly.  Each workgroup sums
// a fixed, strided share of the work and reduces it in a fixed tree; one
// workgroup then adds the workgroups' partial sums in index order, so the
// result is the same on every run (bitwise).  Bound: HBM, one pass over
// coef and x-hat (2 x 16 B per location).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <random>
#include <sstream>
#include <unordered_set>
#include <vector>

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
  // One pass over the locations in a fixed order (not the key list, whose
  // order depends on how identify's workgroups raced), so the sums do not
  // depend on the run.
  for (int64_t i = gid; i < c.n; i += gsize) {
    const double e = Err(c.coef[i], xhat[i]);
    v[6] += e;
    v[7] += e * e;
    v[8] = fmax(v[8], e);
    if ((c.present[i >> 5] >> (i & 31)) & 1u) {  // a found mode
      const int o = InSorted(truth, m, i) ? 0 : 3;
      v[o] += 1.0;
      v[o + 1] += e;
      v[o + 2] += e * e;
    }
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

// ---------------------------------------------------------------------------
// The signal of Demo1 (src/demo1.cc:27-34): GenerateModeMap then
// GenerateXhatAlt (src/gen.cc:27-38, 78-89) from the reference's global
// std::mt19937 (src/base.cc:32-36).  x-hat = the modes plus
// Cplex(RandomNormal(), RandomNormal()) * sigma/sqrt(2n) at every location,
// the imaginary part drawn first (g++ evaluates the arguments right to
// left).  std::normal_distribution<double> is the Marsaglia polar method of
// libstdc++: an attempt draws two uniforms (generate_canonical<double, 53>:
// two 32-bit words each, (w0 + w1 2^32) / 2^64 in double arithmetic, 1
// replaced by the double below it), x = 2u1 - 1, y = 2u2 - 1, accepted when
// 0 < x^2 + y^2 <= 1; an accepted attempt gives y*m, then (cached) x*m with
// m = sqrt(-2 ln r2 / r2).  With an even number of normals drawn before
// (two per mode), location i's sample is exactly the i-th accepted attempt
// of the engine's stream after the mode map: re = x m, im = y m.
//
// The whole file is compiled without FP contraction (build.py): the host's
// normal_distribution computes r2 = x x + y y, which a fused multiply-add
// would round differently from the reference's g++ build.
//
// On the device: one wavefront runs the engine from its state (624 words
// and a position) and writes the raw 32-bit stream; then every attempt (4
// words) is tested in parallel, the accepted ones are ranked by a prefix
// count, and the i-th accepted attempt writes sample i.  Bound: the
// engine's twist, 624 words per step of four dependent phases.

constexpr uint32_t kMtN = 624, kMtM = 397;
constexpr uint32_t kMtUpper = 0x80000000u, kMtLower = 0x7fffffffu, kMtA = 0x9908b0dfu;

__device__ __forceinline__ uint32_t MtTwist(uint32_t cur, uint32_t next, uint32_t far) {
  const uint32_t y = (cur & kMtUpper) | (next & kMtLower);
  return far ^ (y >> 1) ^ ((y & 1u) ? kMtA : 0u);
}

__device__ __forceinline__ uint32_t MtTemper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// words out[0..count) of the stream of an engine in state (mt, pos); one
// workgroup of 64 threads.
__global__ void __launch_bounds__(64)
    k_mt19937(const uint32_t *state, uint32_t pos, uint32_t *out, int64_t count) {
  __shared__ uint32_t mt[kMtN];
  for (uint32_t i = threadIdx.x; i < kMtN; i += blockDim.x) mt[i] = state[i];
  __syncthreads();
  int64_t done = 0;
  // the rest of the current block
  for (uint32_t i = pos + threadIdx.x; i < kMtN && done + (i - pos) < count; i += blockDim.x)
    out[i - pos] = MtTemper(mt[i]);
  done = min<int64_t>(count, int64_t(kMtN - pos));
  while (done < count) {
    // The twist in four phases; each reads only words that are final
    // before it (new for index < i, old for index > i), then writes.
    //   [0, 227): mt[i+397] old;  [227, 454): mt[i-227] new from the first;
    //   [454, 623): mt[i-227] new from the second;  623: mt[0], mt[396] new.
    const uint32_t lo[4] = {0, kMtN - kMtM, 2 * (kMtN - kMtM), kMtN - 1};
    const uint32_t hi[4] = {kMtN - kMtM, 2 * (kMtN - kMtM), kMtN - 1, kMtN};
    for (int ph = 0; ph < 4; ++ph) {
      uint32_t v[4];
      int nv = 0;
      for (uint32_t i = lo[ph] + threadIdx.x; i < hi[ph]; i += blockDim.x, ++nv) {
        const uint32_t far = mt[i < kMtN - kMtM ? i + kMtM : i - (kMtN - kMtM)];
        v[nv] = MtTwist(mt[i], mt[(i + 1) % kMtN], far);
      }
      __syncthreads();
      nv = 0;
      for (uint32_t i = lo[ph] + threadIdx.x; i < hi[ph]; i += blockDim.x, ++nv) mt[i] = v[nv];
      __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < kMtN && done + i < count; i += blockDim.x)
      out[done + i] = MtTemper(mt[i]);
    done += kMtN;
    __syncthreads();
  }
}

// generate_canonical<double, 53> over two 32-bit words.
__device__ __forceinline__ double Canonical(uint32_t w0, uint32_t w1) {
#pragma clang fp contract(off)
  const double sum = double(w0) + double(w1) * 4294967296.0;
  const double r = sum / 18446744073709551616.0;  // 2^64
  return r >= 1.0 ? 0.99999999999999988898 : r;   // the double below 1
}

__device__ __forceinline__ bool PolarAttempt(const uint32_t *w, double *x, double *y, double *r2) {
#pragma clang fp contract(off)
  *x = 2.0 * Canonical(w[0], w[1]) - 1.0;
  *y = 2.0 * Canonical(w[2], w[3]) - 1.0;
  *r2 = *x * *x + *y * *y;
  return !(*r2 > 1.0 || *r2 == 0.0);
}

constexpr int kScanTile = 4 * kThreads;  // attempts per workgroup of the ranking scan

// acc[t] = accepted attempts in tile t (kScanTile attempts per workgroup)
__global__ void __launch_bounds__(kThreads)
    k_polar_count(const uint32_t *words, int64_t attempts, int32_t *acc) {
  __shared__ int32_t sh[kThreads];
  int32_t c = 0;
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  for (int j = threadIdx.x; j < kScanTile; j += blockDim.x) {
    double x, y, r2;
    if (base + j < attempts && PolarAttempt(words + 4 * (base + j), &x, &y, &r2)) ++c;
  }
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[blockIdx.x] = sh[0];
}

// acc[t] <- accepted attempts before tile t (exclusive prefix over tiles),
// acc[tiles] <- the total; one thread (a few thousand tiles per 2^22
// samples).
__global__ void k_polar_offsets(int32_t *acc, int32_t tiles) {
  if (threadIdx.x != 0) return;
  int32_t s = 0;
  for (int32_t t = 0; t < tiles; ++t) {
    const int32_t v = acc[t];
    acc[t] = s;
    s += v;
  }
  acc[tiles] = s;
}

// xhat[r] = (x m, y m) * scale for the r-th accepted attempt, r < n.  Each
// workgroup ranks its tile's attempts in order (a prefix count in LDS).
__global__ void __launch_bounds__(kThreads)
    k_polar_samples(const uint32_t *words, int64_t attempts, const int32_t *acc, int64_t n,
                    double scale, double2 *xhat) {
#pragma clang fp contract(off)
  __shared__ int32_t sh[kThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  constexpr int kPer = kScanTile / kThreads;  // consecutive attempts per thread
  double xs[kPer], ys[kPer], r2s[kPer];
  bool ok[kPer];
  int32_t c = 0;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int64_t j = base + int64_t(threadIdx.x) * kPer + e;
    ok[e] = j < attempts && PolarAttempt(words + 4 * j, &xs[e], &ys[e], &r2s[e]);
    c += ok[e] ? 1 : 0;
  }
  // exclusive prefix of the per-thread counts (Hillis-Steele in LDS)
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    const int32_t v = int(threadIdx.x) >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t r = int64_t(acc[blockIdx.x]) + sh[threadIdx.x] - c;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    if (!ok[e]) continue;
    if (r < n) {
      const double m = sqrt(-2.0 * log(r2s[e]) / r2s[e]);
      xhat[r] = make_double2(xs[e] * m * scale, ys[e] * m * scale);
    }
    ++r;
  }
}

// xhat[k] += coef for every mode (location k: the mode plus its noise).
__global__ void __launch_bounds__(kThreads)
    k_add_modes(const int64_t *keys, const double2 *coefs, int64_t m, double2 *xhat) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) {
    const double2 c = coefs[i];
    const double2 v = xhat[keys[i]];
    xhat[keys[i]] = make_double2(c.x + v.x, c.y + v.y);
  }
}

}  // namespace

namespace {

// RandomInt32 (src/base.cc:53) as libstdc++ computed it before GCC 11:
// the engine's outputs below 2^31.
int32_t RandomInt32(std::mt19937 &eng) {
  uint32_t r;
  do {
    r = uint32_t(eng());
  } while (r >= 0x80000000u);
  return int32_t(r);
}

template <typename T>
int Alloc(T **p, size_t count) {
  return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T)) == hipSuccess
             ? MPSFT_OK
             : MPSFT_ERR_NOMEM;
}

}  // namespace

int GenerateSignal(int64_t n, int32_t k, double sigma, uint32_t seed, double2 *xhat,
                   int64_t *keys, double *coefs, hipStream_t s) {
  if (n < 1 || n > 0x7fffffffLL || k < 0 || k > n || !(sigma >= 0) || !xhat ||
      (k > 0 && (!keys || !coefs)))
    return MPSFT_ERR_INVALID_ARG;
  // GenerateModeMap (src/gen.cc:27-38) on the host: k distinct locations
  // UnsafePosMod(RandomInt32(), n), coefficients Cplex(RandomNormal(),
  // RandomNormal()) (imaginary part first) over their modulus.
  std::mt19937 eng(seed);
  std::normal_distribution<double> nd;
  std::unordered_set<int64_t> seen;
  for (int32_t i = 0; i < k;) {
    const int64_t idx = RandomInt32(eng) % n;
    if (seen.count(idx)) continue;
    const double im = nd(eng);
    const double re = nd(eng);
    const double a = std::hypot(re, im);  // std::abs of a complex<double> (glibc cabs)
    seen.insert(idx);
    keys[i] = idx;
    coefs[2 * i] = re / a;
    coefs[2 * i + 1] = im / a;
    ++i;
  }
  // The engine's state after the mode map: 624 words, then the position.
  std::vector<uint32_t> state(kMtN);
  uint32_t pos = 0;
  {
    std::ostringstream os;
    os << eng;
    std::istringstream is(os.str());
    for (auto &w : state) is >> w;
    is >> pos;
    if (!is || pos > kMtN) return MPSFT_ERR_UNSUPPORTED;
  }
  const double scale = sigma / std::sqrt(2.0 * double(n));
  int rc = MPSFT_OK;
  uint32_t *d_state = nullptr, *d_words = nullptr;
  int32_t *d_acc = nullptr;
  int64_t *d_keys = nullptr;
  double2 *d_coefs = nullptr;
  // Attempts for n accepted ones (acceptance pi/4): 1.35 n plus a margin,
  // more when the draw falls short.
  int64_t attempts = int64_t(1.35 * double(n)) + 4096;
  for (int tries = 0; tries < 4 && rc == MPSFT_OK; ++tries, attempts *= 2) {
    const int32_t tiles = int32_t((attempts + kScanTile - 1) / kScanTile);
    if ((rc = Alloc(&d_state, kMtN)) != MPSFT_OK || (rc = Alloc(&d_words, 4 * attempts)) != MPSFT_OK ||
        (rc = Alloc(&d_acc, size_t(tiles) + 1)) != MPSFT_OK)
      break;
    int32_t total = 0;
    if (hipMemcpyAsync(d_state, state.data(), kMtN * sizeof(uint32_t), hipMemcpyHostToDevice, s) !=
        hipSuccess) {
      rc = MPSFT_ERR_HIP;
      break;
    }
    hipLaunchKernelGGL(k_mt19937, dim3(1), dim3(64), 0, s, d_state, pos, d_words, 4 * attempts);
    hipLaunchKernelGGL(k_polar_count, dim3(tiles), dim3(kThreads), 0, s, d_words, attempts, d_acc);
    hipLaunchKernelGGL(k_polar_offsets, dim3(1), dim3(64), 0, s, d_acc, tiles);
    if (hipMemcpyAsync(&total, d_acc + tiles, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = MPSFT_ERR_HIP;
      break;
    }
    if (total >= n) {
      hipLaunchKernelGGL(k_polar_samples, dim3(tiles), dim3(kThreads), 0, s, d_words, attempts,
                         d_acc, n, scale, xhat);
      if (k > 0) {
        if ((rc = Alloc(&d_keys, k)) != MPSFT_OK || (rc = Alloc(&d_coefs, k)) != MPSFT_OK) break;
        if (hipMemcpyAsync(d_keys, keys, size_t(k) * sizeof(int64_t), hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipMemcpyAsync(d_coefs, coefs, size_t(k) * sizeof(double2), hipMemcpyHostToDevice,
                           s) != hipSuccess) {
          rc = MPSFT_ERR_HIP;
          break;
        }
        hipLaunchKernelGGL(k_add_modes, dim3((k + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                           d_keys, d_coefs, int64_t(k), xhat);
      }
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        rc = MPSFT_ERR_HIP;
      break;
    }
    (void)hipFree(d_state);
    (void)hipFree(d_words);
    (void)hipFree(d_acc);
    d_state = nullptr;
    d_words = nullptr;
    d_acc = nullptr;
    if (tries == 3) rc = MPSFT_ERR_HIP;
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_state);
  (void)hipFree(d_words);
  (void)hipFree(d_acc);
  (void)hipFree(d_keys);
  (void)hipFree(d_coefs);
  return rc;
}

hipError_t LaunchPostAnalyze(const Ctx &c, const int64_t *truth, int64_t m, const double2 *xhat,
                             double *partial, int32_t blocks, double *stats, hipStream_t s) {
  hipLaunchKernelGGL(k_post_analyze, dim3(blocks), dim3(kThreads), 0, s, c, truth, m, xhat,
                     partial);
  hipLaunchKernelGGL(k_post_analyze_final, dim3(1), dim3(kThreads), 0, s, partial, blocks, c.n,
                     stats);
  return hipGetLastError();
}

}  // namespace mpsft
