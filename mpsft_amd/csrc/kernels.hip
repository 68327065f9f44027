// kernels.hip -- HIP kernels of the MPSFT recovery loop for gfx950 (CDNA4).
//
// Layout in HBM (all row-major, interleaved complex):
//   x        : n samples, double2 (FP64) or float2 (MIXED), read by gathers
//   fold     : [trial][row][m]  folded, windowed samples, m < B (double2)
//   A        : [trial][row][b]  bin coefficients after BinInFreq (double2)
//   term_bin : [trial][mode]    bin a found mode falls in this iteration
//   term_val : [trial][row][mode] what BinInFreq subtracts there
//   coef     : n double2, dense coefficients of the found modes
//   present  : n bits, membership of the found-mode map
//   keys     : found-mode locations in insertion order
// Nothing here is a dense contraction, so nothing uses MFMA: the gather is
// bound by random-access memory traffic, the bin FFT by LDS latency and the
// identify/estimate by transcendental latency.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mpsft {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kTwM = 1 << kTwLogM;     // twiddle table period

__device__ __forceinline__ int64_t ModN(int64_t v, int64_t n, int64_t mask) {
  if (mask >= 0) return v & mask;
  const int64_t r = v % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ double2 Cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * b with each product rounded on its own (no fused multiply-add), as
// the reference's complex<double> product rounds (src/freqid.cc:27-48 via
// Eigen, restated by the oracle without contraction).  The pencil's vote
// Im(conj(u0) u2 + conj(u1) u0) and its sigma_2 = |u0^2 - u1 u2| / sigma_1
// cancel heavily on a bin holding one strong mode, and both are compared
// against thresholds; Cmul (contracted) serves everything else.
__device__ __forceinline__ double2 CmulRn(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 Conj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 Add(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 Sub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 Scale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
// a * (-i) and a * (+i)
__device__ __forceinline__ double2 MulNegI(double2 a) { return make_double2(a.y, -a.x); }
__device__ __forceinline__ double2 MulPosI(double2 a) { return make_double2(-a.y, a.x); }

// e^{2 pi i turns}; sincospi reduces its argument exactly.
__device__ __forceinline__ double2 CisTurns(double turns) {
  double s, c;
  sincospi(2.0 * turns, &s, &c);
  return make_double2(c, s);
}
// e^{2 pi i r/n} for an integer r in [0, n), as Sinusoid(double(r)/double(n))
// (src/sincos.h:42-45).
__device__ __forceinline__ double2 CisFrac(int64_t r, int64_t n) {
  return CisTurns(double(r) / double(n));
}

// A sample of x (double2 or float2) in double precision.
__device__ __forceinline__ double2 Widen(double2 v) { return v; }

// A sample of x read by the gather-fold.  Every random 8- or 16-byte
// gather is served as a 128-byte line.  The single-signal kernels load with
// the nontemporal hint (no L2 allocation: measured faster, N=2^22 S=1000
// solve 546 -> 525 us, N=2^26 S=4000 1377 -> 1340 us); the batched
// gather-fold, whose launch gathers from 64 signals' lines, loads them
// plainly so that a line gathered again by the same XCD hits its L2
// (64 x N=2^22 S=1000: 170 -> 154 us per solve).
// x comes to the kernels as a pointer stored in device memory (the loop
// state), which the compiler cannot place in an address space: its loads
// would be flat loads, which also count against the LDS wait counter, so
// every LDS wait after the gathers are issued would wait for them too.
// They are made global loads explicitly, through pointers cast to the
// global address space.
typedef const unsigned long long __attribute__((address_space(1))) *GlobalU64;
typedef const double __attribute__((address_space(1))) *GlobalF64;

typedef const unsigned __int128 __attribute__((address_space(1))) *GlobalU128;
template <bool kNt>
__device__ __forceinline__ float2 LoadX(const float2 *p) {
  const unsigned long long v = kNt ? __builtin_nontemporal_load((GlobalU64)p) : *(GlobalU64)p;
  float2 r;
  __builtin_memcpy(&r, &v, sizeof(r));
  return r;
}
template <bool kNt>
__device__ __forceinline__ double2 LoadX(const double2 *p) {
  if (kNt) {
    const GlobalF64 q = (GlobalF64)p;
    return make_double2(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
  }
  const unsigned __int128 v = *(GlobalU128)p;
  double2 r;
  __builtin_memcpy(&r, &v, sizeof(r));
  return r;
}
__device__ __forceinline__ double2 Widen(float2 v) { return make_double2(double(v.x), double(v.y)); }

// Window::SampleInTime (src/window.cc:71-75): w sinc(pi t w) e^{-2 (pi t
// sigma_f)^2}, with SincPi (src/sincos.cc:69-88) = sin(x)/x and its Taylor
// terms near 0.  sin(pi |t| w) is evaluated as sinpi(|t| w): the same value
// up to rounding, with exact argument reduction (the generic sin carries a
// large-argument path that needs private memory).
__device__ __forceinline__ double SampleInTime(int64_t i, double width, double sigma_f) {
  const double taylor_0_bound = 2.220446049250313080847e-16;
  const double taylor_2_bound = 1.490116119384765625e-08;  // sqrt(eps)
  const double taylor_n_bound = 1.220703125e-04;           // eps^(1/4)
  const double t = double(i);
  const double at = fabs(t);
  const double x = at * kPi * width;
  double sinc;
  if (x >= taylor_n_bound) {
    sinc = sinpi(at * width) / x;
  } else {
    sinc = 1.0;
    if (x >= taylor_0_bound) {
      const double x2 = x * x;
      sinc -= x2 / 6.0;
      if (x >= taylor_2_bound) sinc += (x2 * x2) / 120.0;
    }
  }
  const double u = t * kPi * sigma_f;
  return width * sinc * exp(-2.0 * u * u);
}

// Window::SampleInFreq (src/window.cc:77-81).
__device__ __forceinline__ double SampleInFreq(double xi, double width, double d) {
  const double c = 0.5 * width;
  return 0.5 * (erf((xi + c) * d) - erf((xi - c) * d));
}

__device__ __forceinline__ uint32_t BitRev(uint32_t v, int logM) {
  return logM ? (__brev(v) >> (32 - logM)) : 0u;
}

// Row u of the bin matrix has tau = q + off(u)*B with off(0) = 0,
// off(2k+1) = 2^k, off(2k+2) = -2^k (GetTau, src/binner.cc:42-50).
__device__ __forceinline__ int64_t RowOffset(int row) {
  if (row == 0) return 0;
  if (row & 1) return int64_t(1) << ((row - 1) / 2);
  return -(int64_t(1) << (row / 2 - 1));
}

// ---------------------------------------------------------------------------
// Power-of-two FFTs in LDS, radix 4 (two radix-2 stages per pass, held in
// registers) with a final radix-2 stage when logM is odd.  `count`
// independent transforms of M points sit back to back in a.  Twiddles come
// from a table w[j] = e^{-2 pi i j / M}, j < M/2, read as tw[j * tstride].
struct Twiddles {
  const double2 *tw;
  int stride;
  bool inverse;
  __device__ __forceinline__ double2 operator()(int j) const {
    double2 w = tw[j * stride];
    if (inverse) w.y = -w.y;
    return w;
  }
};

// Decimation in time: bit-reversed input, natural-order output.
__device__ void FftDit(double2 *a, int logM, const Twiddles &W, int count = 1) {
  const int M = 1 << logM;
  int s = 1;
  for (; s + 1 <= logM; s += 2) {
    const int half = 1 << (s - 1);
    const int sh1 = logM - s, sh2 = logM - s - 1;  // W_len^k = W_M^(k << sh)
    for (int t = threadIdx.x; t < count * (M / 4); t += blockDim.x) {
      const int idx = t & (M / 4 - 1);
      const int k = idx & (half - 1);
      const int base = ((t >> (logM - 2)) << logM) + ((idx >> (s - 1)) << (s + 1)) + k;
      double2 e0 = a[base], e1 = a[base + half], e2 = a[base + 2 * half], e3 = a[base + 3 * half];
      const double2 w1 = W(k << sh1);
      const double2 w2 = W(k << sh2);
      const double2 w3 = W.inverse ? MulPosI(w2) : MulNegI(w2);  // W_len2^(k + half)
      double2 v = Cmul(e1, w1);
      e1 = Sub(e0, v);
      e0 = Add(e0, v);
      v = Cmul(e3, w1);
      e3 = Sub(e2, v);
      e2 = Add(e2, v);
      v = Cmul(e2, w2);
      a[base + 2 * half] = Sub(e0, v);
      a[base] = Add(e0, v);
      v = Cmul(e3, w3);
      a[base + 3 * half] = Sub(e1, v);
      a[base + half] = Add(e1, v);
    }
    __syncthreads();
  }
  if (s == logM) {  // last radix-2 stage
    const int half = 1 << (s - 1);
    for (int t = threadIdx.x; t < count * (M / 2); t += blockDim.x) {
      const int idx = t & (M / 2 - 1);
      const int k = idx & (half - 1);
      const int i = ((t >> (logM - 1)) << logM) + ((idx >> (s - 1)) << s) + k;
      const double2 u = a[i], v = Cmul(a[i + half], W(k));
      a[i] = Add(u, v);
      a[i + half] = Sub(u, v);
    }
    __syncthreads();
  }
}

// Decimation in frequency: natural input, bit-reversed output.
__device__ void FftDif(double2 *a, int logM, const Twiddles &W, int count = 1) {
  const int M = 1 << logM;
  int s = logM;
  for (; s - 1 >= 1; s -= 2) {
    // stages s (len2 = 2^s) then s-1 (len = 2^(s-1)), half = 2^(s-2)
    const int half = 1 << (s - 2);
    const int sh2 = logM - s, sh1 = logM - s + 1;
    for (int t = threadIdx.x; t < count * (M / 4); t += blockDim.x) {
      const int idx = t & (M / 4 - 1);
      const int k = idx & (half - 1);
      const int base = ((t >> (logM - 2)) << logM) + ((idx >> (s - 2)) << s) + k;
      const double2 e0 = a[base], e1 = a[base + half], e2 = a[base + 2 * half],
                    e3 = a[base + 3 * half];
      const double2 w2 = W(k << sh2);
      const double2 w3 = W.inverse ? MulPosI(w2) : MulNegI(w2);
      const double2 w1 = W(k << sh1);
      const double2 f0 = Add(e0, e2), f2 = Cmul(Sub(e0, e2), w2);
      const double2 f1 = Add(e1, e3), f3 = Cmul(Sub(e1, e3), w3);
      a[base] = Add(f0, f1);
      a[base + half] = Cmul(Sub(f0, f1), w1);
      a[base + 2 * half] = Add(f2, f3);
      a[base + 3 * half] = Cmul(Sub(f2, f3), w1);
    }
    __syncthreads();
  }
  if (s == 1) {
    for (int idx = threadIdx.x; idx < count * (M / 2); idx += blockDim.x) {
      const int i = 2 * idx;
      const double2 u = a[i], v = a[i + 1];
      a[i] = Add(u, v);
      a[i + 1] = Sub(u, v);
    }
    __syncthreads();
  }
}

// Bluestein chirp of the forward B-point DFT: c_m = e^{-pi i m^2 / B} with
// m^2 reduced modulo 2B exactly.
__device__ __forceinline__ double2 Chirp(int64_t m, int32_t B) {
  const int64_t r = (m * m) % (2 * int64_t(B));
  double s, c;
  sincospi(double(r) / double(B), &s, &c);
  return make_double2(c, -s);
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_build_twiddles(double2 *tw) {
  const int half = kTwM / 2;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < half; k += gridDim.x * blockDim.x) {
    double s, c;
    sincospi(-2.0 * double(k) / double(kTwM), &s, &c);
    tw[k] = make_double2(c, s);
  }
}

// One workgroup per B: chirp c_m (m < B) and the spectrum of the Bluestein
// filter h (h_j = conj(c_j), h_{M-j} = conj(c_j)), pre-scaled by 1/M.
__global__ void __launch_bounds__(kThreads)
    k_build_tables(const int32_t *blist, const BParams *btab, double2 *table,
                   const double2 *tw) {
  HIP_DYNAMIC_SHARED(double2, lds)
  const int32_t B = blist[blockIdx.x];
  const BParams bp = btab[B];
  double2 *chirp = table + bp.off;
  double2 *filt = chirp + B;
  const int M = bp.M;
  for (int m = threadIdx.x; m < B; m += blockDim.x) chirp[m] = Chirp(m, B);
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (j < B)
      v = Conj(Chirp(j, B));
    else if (j > M - B)
      v = Conj(Chirp(M - j, B));
    lds[BitRev(j, bp.logM)] = v;
  }
  __syncthreads();
  FftDit(lds, bp.logM, Twiddles{tw, kTwM / M, false});
  const double inv = 1.0 / double(M);
  for (int k = threadIdx.x; k < M; k += blockDim.x) filt[k] = Scale(lds[k], inv);
}

// Clears the found-mode map: only the listed keys (their coefficient and
// presence word), or everything after an overflow.
// tid / gsize: this thread's index and the thread count over the map.
__device__ __forceinline__ void ClearModesBody(const Ctx &c, int32_t full, int64_t tid,
                                               int64_t gsize) {
  const DevState *st = c.st;
  if (full || (st->flags & kFlagOverflow)) {
    for (int64_t i = tid; i < c.n; i += gsize) c.coef[i] = make_double2(0.0, 0.0);
    for (int64_t i = tid; i < (c.n + 31) / 32; i += gsize) c.present[i] = 0u;
    return;
  }
  const int64_t cnt = min<int64_t>(st->count, c.cap);
  for (int64_t i = tid; i < cnt; i += gsize) {
    const int64_t k = c.keys[i];
    c.coef[k] = make_double2(0.0, 0.0);
    c.present[k >> 5] = 0u;
  }
}

__global__ void __launch_bounds__(kThreads) k_clear_modes(Ctx c, int32_t full) {
  ClearModesBody(c, full, int64_t(blockIdx.x) * blockDim.x + threadIdx.x,
                 int64_t(gridDim.x) * blockDim.x);
}

__device__ __forceinline__ void FillParams(IterParams *ip, const BParams &bp, int32_t bits) {
  ip->B = bp.B;
  ip->bits = bits;
  ip->rows = 1 + 2 * bits;
  ip->f = bp.f;
  ip->P = bp.P;
  ip->p2 = (bp.P - 1) / 2;
  ip->M = bp.M;
  ip->logM = bp.logM;
  ip->R = bp.R;
  ip->slices = bp.slices;
  ip->slice_shift = bp.slice_shift;
  ip->tab_off = bp.off;
  ip->width = bp.width;
  ip->sigma_f = bp.sigma_f;
  ip->dfreq = bp.dfreq;
}

// Fixed per-call settings; iteration -1 sits in slot 1 as "found", so the
// first iteration (slot 0) starts with stale = 0.
__device__ __forceinline__ void InitBody(const Ctx &c, const InitArgs &a) {
  DevState *st = c.st;
  st->x = a.x;
  st->mode = a.mode;
  st->trials = a.trials;
  st->k = a.k;
  st->min_bins = a.min_bins;
  st->max_stale = a.max_stale;
  st->max_iter = a.max_iter;
  st->bits_override = a.bits_override;
  st->B_fixed = a.B;
  st->win_thr = a.win_thr;
  st->bin_thr_fixed = a.bin_thr;
  st->sv_thr_fixed = a.sv_thr;
  if (!a.keep_modes) {
    st->count = 0;
    st->flags = 0;
  }
  st->found_iter[0] = 0;
  st->found_iter[1] = 1;
  st->iters = 0;
  st->stop_stale = 0;
  st->L[1].it = -1;
  st->L[1].stale = 0;
  st->L[1].done = 0;
  st->L[0] = st->L[1];  // no stale "stopped" from an earlier call
  st->cnt_base[0] = st->cnt_base[1] = 0;
  st->P[1].B = a.B;
  st->P[1].done = 0;
  if (c.progress)
    for (int i = 0; i < 6; ++i)
      __hip_atomic_store(&c.progress[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_init(Ctx c, InitArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  InitBody(c, a);
}

// ModeMap clear and loop-state initialisation in one launch of one
// workgroup: the clear reads the count that the initialisation resets.
__global__ void __launch_bounds__(kThreads) k_reset(Ctx c, InitArgs a) {
  ClearModesBody(c, 0, threadIdx.x, blockDim.x);
  __syncthreads();
  if (threadIdx.x == 0) InitBody(c, a);
}

// The next iteration's parameters, as Demo1::Run derives them
// (src/demo1.cc:45-58): stale += (previous iteration found nothing); stop
// at max_stale or max_iter; otherwise B = max((k - |found|)*2 + 1, min_bins)
// with thresholds 3 sigma/sqrt(B), sigma/sqrt(B), and the iteration's
// transform and q's.  Reads only state that no kernel of this iteration
// writes before k_bin_time ends, so every workgroup computes the same.
__device__ void NextIteration(const Ctx &c, int p, IterParams *ip, LoopCounters *lc) {
  const DevState *st = c.st;
  const LoopCounters prev = st->L[1 - p];
  if (prev.done) {
    *lc = prev;
    ip->done = 1;
    ip->it = prev.it;
    return;
  }
  lc->it = prev.it + 1;
  lc->stale = prev.stale + (st->found_iter[1 - p] ? 0 : 1);
  lc->pad = 0;
  const bool replay = st->mode == kModeReplay;  // a recorded solve's iterations, no stale stop
  const bool done = (st->flags & kFlagOverflow) || (!replay && lc->stale >= st->max_stale) ||
                    lc->it >= st->max_iter;
  lc->done = done ? 1 : 0;
  ip->done = lc->done;
  ip->it = lc->it;
  if (done) return;
  int32_t B = st->B_fixed;
  if (st->mode == kModeSolve) {
    const int64_t want = (int64_t(st->k) - int64_t(st->count)) * 2 + 1;
    B = int32_t(want > st->min_bins ? want : st->min_bins);
  } else if (replay) {
    B = c.log[lc->it];  // the recorded solve's bins
  }
  const BParams bp = c.btab[B];
  FillParams(ip, bp, st->bits_override > 0 ? st->bits_override : bp.bits);
  if (st->mode == kModeSolve) {
    ip->bin_thr = bp.bin_thr;
    ip->sv_thr = bp.sv_thr;
  } else {
    ip->bin_thr = st->bin_thr_fixed;
    ip->sv_thr = st->sv_thr_fixed;
  }
  ip->tf = c.tfs[lc->it];
  // ip->q is filled by PublishIteration only: workgroups load the q of
  // their own trial, keeping the struct out of private memory.
}

// The solve's result in the host-mapped progress words: iterations, stale
// count and flags, then the iteration and the stop flag, which the host
// polls.  The release store orders the result words before the flag, so a
// host that sees the flag reads this solve's result.
__device__ __forceinline__ void PublishStop(const Ctx &c, int32_t it, int32_t stale,
                                            int32_t flags) {
  __hip_atomic_store(&c.progress[3], it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&c.progress[4], stale, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&c.progress[5], flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&c.progress[0], it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&c.progress[1], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Workgroup 0 of k_bin_time: publish the iteration, log the previous one,
// report progress to the host.  PublishLoop: all of it but the copy of ip
// into slot p (for a caller that built ip there).
__device__ void PublishLoop(const Ctx &c, int p, const IterParams &ip, const LoopCounters &lc);
__device__ void PublishIteration(const Ctx &c, int p, const IterParams &ip,
                                 const LoopCounters &lc) {
  c.st->P[p] = ip;
  PublishLoop(c, p, ip, lc);
}

__device__ void PublishLoop(const Ctx &c, int p, const IterParams &ip, const LoopCounters &lc) {
  DevState *st = c.st;
  const LoopCounters prev = st->L[1 - p];
  if (!ip.done)
    for (int t = 0; t < st->trials; ++t)
      st->P[p].q[t] = c.qs[int64_t(lc.it) * c.trials_cap + t];
  st->L[p] = lc;
  st->found_iter[p] = 0;
  st->cnt_base[p] = st->count;
  if (prev.done) return;
  if (st->mode == kModeSolve && prev.it >= 0 && prev.it < c.log_cap) {
    c.log[prev.it] = st->P[1 - p].B;
    c.log[c.log_cap + prev.it] = st->count;
  }
  if (lc.done) {
    st->iters = lc.it;
    st->stop_stale = lc.stale;
    // A batch counts its stopped signals; the next batched launch's first
    // workgroup alone copies the count to the host.
    if (c.batch_done) atomicAdd(c.batch_done, 1);
  }
  if (c.progress) {
    if (!lc.done) {
      __hip_atomic_store(&c.progress[2], ip.B, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&c.progress[0], lc.it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
      PublishStop(c, lc.it, lc.stale, st->flags);
    }
  }
}

// Inserts keys/coefs into an empty map (ModeMap assignment).
__global__ void __launch_bounds__(kThreads)
    k_set_modes(Ctx c, const int64_t *keys, const double2 *coefs, int64_t m) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) {
    const int64_t k = keys[i];
    atomicOr(&c.present[k >> 5], 1u << (k & 31));
    c.coef[k] = coefs[i];
    c.keys[i] = k;
  }
  if (i == 0) c.st->count = int32_t(m);
}

__global__ void __launch_bounds__(kThreads)
    k_get_coefs(Ctx c, double2 *out, int64_t m) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) out[i] = c.coef[c.keys[i]];
}

// ---------------------------------------------------------------------------
// Lag layout shared by every row of an iteration.  Row `row` reads samples
// u = q - p2 + lambda*B + rho for lambda = off(row) + k, k < f.  Rows with
// 2^k < f overlap row 0: their lambdas form one run [-2^(K-1), 2^(K-1) + f)
// of len0 lags ("segment 0"); every other row owns a run of f lags.  The
// union, nlam lags, is staged once per residue.
struct Lags {
  int K, len0, nlam, f;
  float inv_f;
  __device__ __forceinline__ Lags(int bits, int f_) : f(f_), inv_f(1.0f / float(f_)) {
    K = 0;
    while (K < bits && (1 << K) < f) ++K;
    len0 = (K ? (1 << K) : 0) + f;
    nlam = len0 + 2 * (bits - K) * f;
  }
  // lambda of staged index li
  __device__ __forceinline__ int64_t Lambda(int li) const {
    if (li < len0) return K ? li - (1 << (K - 1)) : li;
    const int h = li - len0;
    // h / f through the float reciprocal, corrected to the exact quotient
    int seg = int(float(h) * inv_f);
    if ((seg + 1) * f <= h) ++seg;
    else if (seg * f > h) --seg;
    const int kk = K + seg / 2;
    const int64_t lo = (seg & 1) ? -(int64_t(1) << kk) : (int64_t(1) << kk);
    return lo + (h - seg * f);
  }
  // staged index of the first lag of row `row`
  __device__ __forceinline__ int RowStart(int row) const {
    const int kk = row == 0 ? -1 : (row - 1) / 2;
    if (kk < K) return (K ? (1 << (K - 1)) : 0) + int(RowOffset(row));
    const int seg = 2 * (kk - K) + ((row & 1) ? 0 : 1);
    return len0 + seg * f;
  }
};

// k_bin_time: the permuted, windowed gather-fold of BinInTime
// (src/binner.cc:81-111 semantics), without its DFT.  Factoring the row's
// own phase e^{2 pi i b tau/n} out (applied after the DFT), row `row` folds
//   fold[row][m(rho)] = sum_{k<f} x[(a u + c) mod n] * tap(k*B + rho - p2),
//   u = q - p2 + (off(row) + k)*B + rho,   m(rho) = (rho - (B-1)/2) mod B,
// tap(t) = wt(t) e^{2 pi i [(b t mod n)/n + frac(-t/(2B))]}.
// A tile = (trial, block of R residues rho) covers every row:
//   phase A gathers the nlam x R distinct samples into LDS (each once),
//   phase T computes the f x R taps into LDS (each once per iteration),
//   phase B forms every (row, rho) sum, split over `parts` threads.
// With ip.slices = kSlices the tile is split over the 8 XCDs: its unit s
// gathers only the samples in the s-th eighth of x (index >> slice_shift ==
// s; the others read as 0) and writes partial fold s.  Units go to XCDs
// round-robin (unit s of every tile on XCD s), so each XCD's L2 serves one
// eighth of x, and a line of x is fetched by one XCD only, however many of
// the tile's gathers it serves.  The bin FFT adds the partials in the fixed
// order s = 0..7.
// Workgroups below kPrepBlocks instead compute the BinInFreq terms of every
// found mode for this iteration.
// blk: this workgroup's block index within the signal's launch.
template <typename Tx, bool kNt = true>
__device__ __forceinline__ void TileFold(const Ctx &c, const IterParams &ip, int32_t trials,
                                         int32_t tile, double2 *lds);

// The BinInFreq factor of mode k0 (k1 = its image under the transform) in
// row 0 for offset q: wf(xi) e^{2 pi i (q k1 + c k0)/n}, and its bin.  Every
// form of an iteration builds a term as coef (base) or coef (base g) in
// this order, so all of them subtract bit-identical terms.
__device__ __forceinline__ double2 FreqBase(const Ctx &c, const IterParams &ip, int64_t q,
                                            int64_t k0, int64_t k1, int32_t *bin) {
  const int64_t n = c.n, mask = c.mask;
  const int32_t B = ip.B;
  const int32_t b = int32_t(k1 * B / n);
  const double xi = (double(b) + 0.5) / double(B) - double(k1) / double(n);
  const double wf = SampleInFreq(xi, ip.width, ip.dfreq);
  *bin = b;
  return Scale(CisFrac(ModN(q * k1 + ip.tf.c * k0, n, mask), n), wf);
}

// g = e^{2 pi i (B 2^bit k1)/n}: rows 2bit+1 and 2bit+2 take base g and
// base conj(g).
__device__ __forceinline__ double2 BitTwiddle(const Ctx &c, int32_t B, int bit, int64_t k1) {
  return CisFrac(ModN((int64_t(B) << bit) * k1, c.n, c.mask), c.n);
}

// BinInFreqV1 (src/binner.cc:565-591) per found mode and trial: bin,
// base = wf e^{2 pi i (q k1 + c k0)/n}, and per bit
// g = e^{2 pi i (B 2^bit k1)/n}; the terms are coef base (row 0),
// coef (base g) (row 2bit+1) and coef (base conj(g)) (row 2bit+2).
// Workgroup blk of nblk.
__device__ __forceinline__ void FreqTerms(const Ctx &c, const IterParams &ip, int32_t trials,
                                          int32_t blk, int32_t nblk) {
  const int32_t B = ip.B, bits = ip.bits;
  const DevTf tf = ip.tf;
  const int64_t n = c.n, mask = c.mask;
  const int32_t cnt = min(c.st->count, c.cap);
  for (int64_t i = int64_t(blk) * blockDim.x + threadIdx.x; i < int64_t(cnt) * trials;
       i += int64_t(nblk) * blockDim.x) {
    const int32_t trial = int32_t(i / cnt);
    const int32_t mi = int32_t(i - int64_t(trial) * cnt);
    const int64_t k0 = c.keys[mi];
    const double2 cf = c.coef[k0];
    const int64_t q = c.qs[int64_t(ip.it) * c.trials_cap + trial];
    const int64_t k1 = ModN(tf.a * k0 + tf.b, n, mask);
    int32_t bin;
    const double2 base = FreqBase(c, ip, q, k0, k1, &bin);
    c.term_bin[int64_t(trial) * c.cap + mi] = bin;
    double2 *val = c.term_val + int64_t(trial) * c.rows_cap * c.cap + mi;
    val[0] = Cmul(cf, base);
    for (int32_t bit = 0; bit < bits; ++bit) {
      const double2 g = BitTwiddle(c, B, bit, k1);
      val[int64_t(2 * bit + 1) * c.cap] = Cmul(cf, Cmul(base, g));
      val[int64_t(2 * bit + 2) * c.cap] = Cmul(cf, Cmul(base, Conj(g)));
    }
  }
}

template <typename Tx>
__device__ __forceinline__ void BinTimeBody(const Ctx &c, int p, int blk, double2 *lds) {
  const DevState *st = c.st;
  IterParams ip;
  LoopCounters lc;
  NextIteration(c, p, &ip, &lc);
  if (blk == 0 && threadIdx.x == 0) PublishIteration(c, p, ip, lc);
  if (ip.done) return;
  const int32_t trials = st->trials;
  if (blk < kPrepBlocks) {
    FreqTerms(c, ip, trials, blk, kPrepBlocks);
    return;
  }

  TileFold<Tx>(c, ip, trials, blk - kPrepBlocks, lds);
}

// The gather-fold of one tile (trial, block of R residues) of an iteration
// with parameters ip (see BinTimeBody).
template <typename Tx, bool kNt>
__device__ __forceinline__ void TileFold(const Ctx &c, const IterParams &ip, int32_t trials,
                                         int32_t tile, double2 *lds) {
  const DevState *st = c.st;
  const int32_t B = ip.B, f = ip.f, bits = ip.bits, p2 = ip.p2, rows = ip.rows;
  const DevTf tf = ip.tf;
  const int64_t n = c.n, mask = c.mask;
  const int32_t R = ip.R;
  const int32_t nblk = (B + R - 1) / R;
  const int32_t S = ip.slices;
  const int32_t slice = tile % S;  // this unit's eighth of x (0 unsliced)
  tile /= S;
  if (tile >= trials * nblk) return;
  const int32_t trial = tile / nblk;
  const int32_t rho0 = (tile - trial * nblk) * R;
  const int32_t nrho = min(R, B - rho0);
  const Lags L(bits, f);
  const int64_t q = c.qs[int64_t(ip.it) * c.trials_cap + trial];
  const Tx *x = static_cast<const Tx *>(st->x);

  // [nlam][R] samples as stored in x (complex64 samples take half the LDS
  // and widen exactly at use), then [f][R] taps and [kThreads] partial sums
  // in double2 (16-byte aligned).
  Tx *X = reinterpret_cast<Tx *>(lds);
  const int xs = sizeof(Tx) == sizeof(double2) ? L.nlam * R : (L.nlam * R + 1) & ~1;
  double2 *T = reinterpret_cast<double2 *>(X + xs);
  double2 *red = T + f * R;

  // Phase A: gather every distinct sample of the tile once, and phase T:
  // the taps tap(t), t = k*B + rho - p2 (Window::SampleInTime and the phase
  // of src/binner.cc:100-104 with tau factored out).  Each thread issues
  // the loads of up to kGather samples into registers, computes its taps
  // while they are in flight, and only then writes the samples to LDS;
  // storing each sample as it arrives would serialise the loads.
  constexpr int kGather = 8;
  const int logR = R == 1 ? 0 : R == 2 ? 1 : R == 4 ? 2 : 3;
  const int total = L.nlam << logR;
  const int64_t ubase = q - p2 + rho0;
  for (int base = 0; base < total; base += kGather * kThreads) {
    Tx g[kGather];
#pragma unroll
    for (int k = 0; k < kGather; ++k) {
      const int idx = base + int(threadIdx.x) + k * kThreads;
      const int li = idx >> logR, ri = idx & (R - 1);
      g[k] = Tx{};
      if (idx < total && ri < nrho) {
        const int64_t u = ubase + L.Lambda(li) * B + ri;
        const int64_t j = ModN(tf.a * u + tf.c, n, mask);
        if (S == 1 || int32_t(j >> ip.slice_shift) == slice) g[k] = LoadX<kNt>(x + j);
      }
    }
    if (base == 0) {
      const double width = ip.width, sigma_f = ip.sigma_f;
      const double delta = -0.5 / double(B);
      for (int idx = threadIdx.x; idx < (f << logR); idx += blockDim.x) {
        const int k = idx >> logR;
        const int ri = idx & (R - 1);
        if (ri < nrho) {
          const int64_t t = int64_t(k) * B + rho0 + ri - p2;
          // the caller's tap table is indexed as Window::wt() (t < 0 at t + P)
          const double w = c.wtab ? c.wtab[t >= 0 ? t : t + ip.P]
                                  : (t == 0) ? width : SampleInTime(t, width, sigma_f);
          const int64_t r = ModN(tf.b * t, n, mask);
          const double d = delta * double(t);
          const double turns = double(r) / double(n) + (d - floor(d));
          T[idx] = Scale(CisTurns(turns), w);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kGather; ++k) {
      const int idx = base + int(threadIdx.x) + k * kThreads;
      if (idx < total && (idx & (R - 1)) < nrho) X[idx] = g[k];
    }
  }
  __syncthreads();

  // Phase B: per (row, rho) sums over k, split in `parts` when pairs are few.
  const int32_t pairs = rows * nrho;
  int32_t parts = 1;
  while (parts * 2 * pairs <= kThreads && parts * 2 <= f) parts *= 2;
  const int32_t h = (B - 1) / 2;
  double2 *fold_t = c.fold + slice * c.fold_stride + int64_t(trial) * c.rows_cap * c.B_cap;
  for (int base = 0; base < pairs * parts; base += blockDim.x) {
    const int p = base + threadIdx.x;
    const int pair = p / parts;
    const int part = p - pair * parts;
    double2 acc = make_double2(0.0, 0.0);
    int row = 0, ri = 0;
    if (pair < pairs) {
      row = pair / nrho;
      ri = pair - row * nrho;
      const int s0 = L.RowStart(row);
      for (int k = part; k < f; k += parts)
        acc = Add(acc, Cmul(Widen(X[(s0 + k) * R + ri]), T[k * R + ri]));
    }
    if (parts > 1) {
      red[threadIdx.x] = acc;
      __syncthreads();
      if (pair < pairs && part == 0)
        for (int j = 1; j < parts; ++j) acc = Add(acc, red[threadIdx.x + j]);
      __syncthreads();
    }
    if (pair < pairs && part == 0) {
      const int rho = rho0 + ri;
      const int m = rho >= h ? rho - h : rho - h + B;
      fold_t[int64_t(row) * c.B_cap + m] = acc;
    }
  }
}

template <typename Tx>
__global__ void __launch_bounds__(kThreads) k_bin_time(Ctx c, int p) {
  HIP_DYNAMIC_SHARED(double2, lds)
  BinTimeBody<Tx>(c, p, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// Register-resident Stockham FFT, the large-M path of k_bin_fft (M = 2048,
// 4096): one first pass of radix R0 = 2^(logM mod 4) (16 when 4 divides
// logM), then radix-16 passes.  In a pass of radix R with Ns = product of
// the earlier radices, butterfly j < M/R takes x[j + r M/R] (r < R) into
// registers, multiplies input r by W_{Ns R}^{k r} (k = j mod Ns), transforms
// them, and writes output r to (j/Ns) Ns R + k + r Ns (natural order in and
// out).  A thread holds at most 16 values (M <= 16 kFftThreads).  Between
// passes the array sits in LDS with one element of padding per 16, which
// spreads the early passes' strided writes over the banks.  The
// 4096-point transform is 3 LDS round trips instead of 6 radix-4 passes.
__device__ __forceinline__ int Pad(int i) { return i + (i >> 4); }


constexpr double kCos8 = 0.92387953251128675613;   // cos(pi/8)
constexpr double kSin8 = 0.38268343236508977173;   // sin(pi/8)
constexpr double kHalfR2 = 0.70710678118654752440;  // cos(pi/4)

// cos and sin of 2 pi m/16, for a compile-time m.
__device__ __forceinline__ constexpr double Cos16(int m) {
  switch (m & 15) {
    case 0: return 1.0;
    case 1: case 15: return kCos8;
    case 2: case 14: return kHalfR2;
    case 3: case 13: return kSin8;
    case 4: case 12: return 0.0;
    case 5: case 11: return -kSin8;
    case 6: case 10: return -kHalfR2;
    case 7: case 9: return -kCos8;
    default: return -1.0;
  }
}
__device__ __forceinline__ constexpr double Sin16(int m) { return Cos16(m - 4); }

// a * e^{-/+ 2 pi i m/16} (forward / inverse); m is a compile-time constant
// after unrolling.
template <bool INV>
__device__ __forceinline__ double2 Rot16(double2 a, int m) {
  const double cs = Cos16(m), sn = INV ? Sin16(m) : -Sin16(m);
  return make_double2(a.x * cs - a.y * sn, a.x * sn + a.y * cs);
}

template <bool INV>
__device__ __forceinline__ void Dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3) {
  const double2 t0 = Add(a0, a2), t1 = Sub(a0, a2), t2 = Add(a1, a3);
  const double2 d = Sub(a1, a3);
  const double2 t3 = INV ? MulPosI(d) : MulNegI(d);
  a0 = Add(t0, t2);
  a2 = Sub(t0, t2);
  a1 = Add(t1, t3);
  a3 = Sub(t1, t3);
}

// In-place R-point DFT of v[0..R-1], as R = R1 R2 with n = n1 + R1 n2,
// k = k1 + R2 k2:
//   X[k1 + R2 k2] = sum_n1 W_R1^(n1 k2) W_R^(n1 k1) sum_n2 W_R2^(n2 k1) x[n1 + R1 n2].
// Inputs in natural order; output k sits in register DftSlot<R>(k).
template <int R>
__device__ __forceinline__ constexpr int DftSlot(int k) {
  if constexpr (R == 8) return 2 * (k & 3) + (k >> 2);   // X[k1 + 4 k2] in v[2 k1 + k2]
  if constexpr (R == 16) return 4 * (k & 3) + (k >> 2);  // X[k1 + 4 k2] in v[4 k1 + k2]
  return k;
}

template <int R, bool INV>
__device__ __forceinline__ void Dft(double2 *v) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = Add(a, b);
    v[1] = Sub(a, b);
  } else if constexpr (R == 4) {
    Dft4<INV>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {  // R1 = 2, R2 = 4
    Dft4<INV>(v[0], v[2], v[4], v[6]);
    Dft4<INV>(v[1], v[3], v[5], v[7]);
    v[3] = Rot16<INV>(v[3], 2);  // n1 = 1: W_8^k1 = W_16^(2 k1)
    v[5] = Rot16<INV>(v[5], 4);
    v[7] = Rot16<INV>(v[7], 6);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      const double2 x0 = v[2 * k1], x1 = v[2 * k1 + 1];
      v[2 * k1] = Add(x0, x1);
      v[2 * k1 + 1] = Sub(x0, x1);
    }
  } else {  // R == 16: R1 = R2 = 4
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) Dft4<INV>(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]);
#pragma unroll
    for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) v[n1 + 4 * k1] = Rot16<INV>(v[n1 + 4 * k1], n1 * k1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) Dft4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
  }
}

// Loads the inputs of this thread's butterflies of a radix-R pass:
// v[b R + r] = a[j + r M/R], j = tid + b T.
template <int R>
__device__ __forceinline__ void PassLoad(const double2 *a, double2 (&v)[16], int logM) {
  constexpr int NB = 16 / R;
  const int nb = (1 << logM) / R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = int(threadIdx.x) + b * kFftThreads;
    if (j < nb)
#pragma unroll
      for (int r = 0; r < R; ++r) v[b * R + r] = a[Pad(j + r * nb)];
  }
}

// Twist (input r by w^r, w = W_{Ns R}^k, given for b = 0 in w1), transform
// and store this thread's butterflies of a radix-R pass.  Passes with
// Ns > 1 have one butterfly per thread.
template <int R, bool INV>
__device__ __forceinline__ void PassComputeStore(double2 *a, double2 (&v)[16], int logM, int Ns,
                                                 double2 w1) {
  constexpr int NB = 16 / R;
  const int nb = (1 << logM) / R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = int(threadIdx.x) + b * kFftThreads;
    if (j < nb) {
      double2 *u = v + b * R;
      if (Ns > 1) {
        // w^r by a squaring tree (depth <= 4 products) rather than a chain
        double2 wp[R];
        wp[1] = INV ? Conj(w1) : w1;
#pragma unroll
        for (int r = 2; r < R; ++r) wp[r] = (r & 1) ? Cmul(wp[r - 1], wp[1]) : Cmul(wp[r / 2], wp[r / 2]);
#pragma unroll
        for (int r = 1; r < R; ++r) u[r] = Cmul(u[r], wp[r]);
      }
      Dft<R, INV>(u);
      const int k = j & (Ns - 1);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) a[Pad(base + r * Ns)] = u[DftSlot<R>(r)];
    }
  }
}

// Runtime radix to template dispatch: load, or twist + transform + store.
__device__ __forceinline__ void LoadR(const double2 *a, double2 (&v)[16], int logM, int R) {
  switch (R) {
    case 2: PassLoad<2>(a, v, logM); break;
    case 4: PassLoad<4>(a, v, logM); break;
    case 8: PassLoad<8>(a, v, logM); break;
    default: PassLoad<16>(a, v, logM); break;
  }
}

template <bool INV>
__device__ __forceinline__ void ComputeStoreR(double2 *a, double2 (&v)[16], int logM, int R, int Ns,
                                              double2 w1) {
  switch (R) {
    case 2: PassComputeStore<2, INV>(a, v, logM, Ns, w1); break;
    case 4: PassComputeStore<4, INV>(a, v, logM, Ns, w1); break;
    case 8: PassComputeStore<8, INV>(a, v, logM, Ns, w1); break;
    default: PassComputeStore<16, INV>(a, v, logM, Ns, w1); break;
  }
}

// ---------------------------------------------------------------------------
// BinInFreq's subtraction (src/binner.cc:565-591) in an order-independent
// form.  The reference subtracts a bin's terms one by one in ModeMap order;
// here every term v of a workgroup's bins is added, as two signed integer
// limbs per component on the grid 2^(E-94) (E: the smallest power of two
// above every term component of the workgroup), into per-bin accumulators:
// x = v 2^(94-E) = h 2^47 + l with h = trunc(x / 2^47) exact and l = the
// integer nearest x - h 2^47 (the only rounding: below 2^-94 of the largest
// term).  Integer addition is associative, so each bin's sum, the one
// rounding that turns it into a double and the bin value A = DFT - sum are
// the same whatever order the terms (and the found-mode keys behind them)
// arrive in, and the same in every form of an iteration that sees the same
// terms.  Exact for up to 2^15 terms in one bin.
struct TermGrid {
  double scale, inv;  // 2^(94-E), 2^(E-94)
  bool any;           // some term is nonzero
};
constexpr int kLimbs = 4;  // per bin: re h, re l, im h, im l

// The workgroup's largest term component (v) -> the grid.  Every thread
// calls it.
__device__ __forceinline__ TermGrid MakeTermGrid(double v) {
  __shared__ double wmax[kFftThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = 0.0;
  for (int w = 0; w < int(blockDim.x >> 6); ++w) m = fmax(m, wmax[w]);
  TermGrid g;
  g.any = m > 0.0;
  const int e = g.any ? ilogb(m) + 1 : 0;
  g.scale = ldexp(1.0, 94 - e);
  g.inv = ldexp(1.0, e - 94);
  return g;
}

__device__ __forceinline__ double MaxComp(double2 v) { return fmax(fabs(v.x), fabs(v.y)); }

__device__ __forceinline__ void LimbAdd(unsigned long long *acc, double2 v, const TermGrid &g) {
  const double comp[2] = {v.x, v.y};
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const double x = comp[j] * g.scale;
    const double h = trunc(x * 0x1p-47);
    const double l = rint(x - h * 0x1p47);
    atomicAdd(&acc[2 * j], static_cast<unsigned long long>(static_cast<long long>(h)));
    atomicAdd(&acc[2 * j + 1], static_cast<unsigned long long>(static_cast<long long>(l)));
  }
}

// The bin's accumulated terms as a double2 (one rounding per component).
__device__ __forceinline__ double2 LimbValue(const unsigned long long *acc, const TermGrid &g) {
  double out[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    long long h = static_cast<long long>(acc[2 * j]);
    long long l = static_cast<long long>(acc[2 * j + 1]);
    const long long carry = l >> 47;  // floor(l / 2^47): 0 <= l < 2^47 after it
    l -= carry * (1ll << 47);
    h += carry;
    out[j] = (double(h) * 0x1p47 + double(l)) * g.inv;
  }
  return make_double2(out[0], out[1]);
}

// A bin FFT's outputs X_k, k = tid + r kFftThreads < B (at k in LDS, or at
// Pad(k) on the large path), moved into registers so that the BinInFreq
// limbs can take the FFT array's place; the caller has synchronised after
// the FFT.
// Which bin-FFT paths a kernel carries: both (the path picked per launch
// from M), or one of them alone, which then runs with the registers and LDS
// of that path only (the small path's kernels reach several wavefronts per
// SIMD; with both paths the large one's 256 VGPRs set one).
enum : int { kFftBoth = 0, kFftSmallOnly = 1, kFftLargeOnly = 2 };
template <int kPaths>
constexpr int OutPer() {  // B <= 512 on the small path, <= 2048 on the large one
  return kPaths == kFftSmallOnly ? 512 / kFftThreads : 2048 / kFftThreads;
}
template <int kPaths>
__device__ __forceinline__ bool FftPathMatches(const IterParams &ip) {
  const bool small = ip.logM <= kFftSmallLogM;
  return kPaths == kFftBoth || (kPaths == kFftSmallOnly) == small;
}

template <int N>
__device__ __forceinline__ void TakeOutputs(const double2 *a, bool small, int32_t B,
                                            double2 (&xo)[N]) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const int k = threadIdx.x + r * kFftThreads;
    xo[r] = k < B ? a[small ? k : Pad(k)] : make_double2(0.0, 0.0);
  }
  __syncthreads();
}

// A[k] = X_k - (the bin's term sum), for the outputs TakeOutputs took.
template <int N>
__device__ __forceinline__ void StoreBins(double2 *A, int32_t B, const double2 (&xo)[N],
                                          const unsigned long long *limbs, const TermGrid &g) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const int k = threadIdx.x + r * kFftThreads;
    if (k < B) A[k] = g.any ? Sub(xo[r], LimbValue(limbs + kLimbs * k, g)) : xo[r];
  }
}

// ---------------------------------------------------------------------------
// k_bin_fft: per (trial, row) the forward B-point DFT of the fold by
// Bluestein (X_k = c_k sum_j (v_j c_j) conj(c_{k-j})) over two M-point
// FFTs (M = 2^logM >= 2B-1), then the row phase e^{2 pi i b tau/n} and the
// BinInFreqV1 subtraction (src/binner.cc:565-591) of every found mode
// (terms from k_bin_time).  Every global load is issued up front.  Two
// paths, chosen by M for the whole launch:
//   M <= 1024 (B <= 512): in-place radix-4 passes in LDS (bit-reversed
//     DIT forward, DIF inverse), twiddles staged in LDS; up to 256
//     butterflies per pass keep every thread busy at small M;
//   M = 2048, 4096: the register Stockham passes above, the forward
//     transform starting from registers and the filter product applied as
//     the inverse transform's first pass loads its inputs.
// Either path leaves X_k (chirp and row phase applied) in LDS, at k on the
// small path and at Pad(k) on the large one.

// Fold entry k: with a gather-fold split over XCD slices (parts =
// kSlices), the sum of the partial folds in the fixed order 0..7.
__device__ __forceinline__ double2 FoldAt(const double2 *f, int k, int32_t parts,
                                          int64_t stride) {
  double2 v = f[k];
  if (parts > 1) {
#pragma unroll
    for (int s = 1; s < kSlices; ++s) v = Add(v, f[s * stride + k]);
  }
  return v;
}

// CONJ_IN: transform conj(fold) (the adjoint pass feeds conjugated bins,
// never split); otherwise fold is partial fold 0 of ip.slices.
template <bool CONJ_IN>
__device__ __forceinline__ void BinFftSmall(const Ctx &c, const IterParams &ip, const double2 *fold,
                                            double2 rp, double2 *a) {
  constexpr int T = kFftThreads;
  constexpr int kPerB = 512 / T;   // fold / chirp entries per thread (B <= 512)
  constexpr int kPerM = 1024 / T;  // filter entries per thread
  constexpr int kPerTw = 512 / T;  // twiddles staged per thread
  const int32_t B = ip.B, logM = ip.logM;
  const int M = 1 << logM;
  const int tid = threadIdx.x;
  const double2 *chirp = c.table + ip.tab_off;
  const double2 *filt = chirp + B;
  double2 fv[kPerB], cv[kPerB], hv[kPerM], twv[kPerTw];
#pragma unroll
  for (int r = 0; r < kPerB; ++r) {
    const int k = tid + r * T;
    fv[r] = cv[r] = make_double2(0.0, 0.0);
    if (k < B) {
      fv[r] = CONJ_IN ? Conj(fold[k]) : FoldAt(fold, k, ip.slices, c.fold_stride);
      cv[r] = chirp[k];
    }
  }
#pragma unroll
  for (int r = 0; r < kPerM; ++r) {
    const int k = tid + r * T;
    hv[r] = k < M ? filt[k] : make_double2(0.0, 0.0);
  }
#pragma unroll
  for (int r = 0; r < kPerTw; ++r) {
    const int j = tid + r * T;
    twv[r] = j < M / 2 ? c.tw[j << (kTwLogM - logM)] : make_double2(0.0, 0.0);
  }
  double2 *twl = a + M;
#pragma unroll
  for (int r = 0; r < kPerTw; ++r) {
    const int j = tid + r * T;
    if (j < M / 2) twl[j] = twv[r];
  }
  const Twiddles Wf{twl, 1, false};
#pragma unroll
  for (int r = 0; r < kPerM; ++r) {
    const int m = tid + r * T;
    if (m < M) {
      double2 v = make_double2(0.0, 0.0);
      if (r < kPerB && m < B) v = Cmul(fv[r], cv[r]);
      a[BitRev(m, logM)] = v;
    }
  }
  __syncthreads();
  FftDit(a, logM, Wf);
#pragma unroll
  for (int r = 0; r < kPerM; ++r) {
    const int k = tid + r * T;
    if (k < M) a[k] = Cmul(a[k], hv[r]);
  }
  __syncthreads();
  Twiddles Wi = Wf;
  Wi.inverse = true;
  FftDif(a, logM, Wi);
  double2 vals[kPerB];
#pragma unroll
  for (int r = 0; r < kPerB; ++r) {
    const int k = tid + r * T;
    if (k < B) vals[r] = Cmul(Cmul(a[BitRev(k, logM)], cv[r]), rp);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kPerB; ++r) {
    const int k = tid + r * T;
    if (k < B) a[k] = vals[r];
  }
}

template <bool CONJ_IN>
__device__ __forceinline__ void BinFftLarge(const Ctx &c, const IterParams &ip, const double2 *fold,
                                            double2 rp, double2 *a) {
  constexpr int T = kFftThreads;
  const int32_t B = ip.B, logM = ip.logM;
  const int M = 1 << logM;
  const int tid = threadIdx.x;
  const double2 *chirp = c.table + ip.tab_off;
  const double2 *filt = chirp + B;
  const int r0 = logM & 3;
  const int R0 = r0 ? (1 << r0) : 16;
  const int lr0 = r0 ? r0 : 4;
  const int nb0 = M / R0, NB0 = 16 / R0;
  const int npass = (logM >> 2) + (r0 ? 1 : 0);  // passes after the first are radix 16

  // Prefetch, in the first pass's layout (element b R0 + r is index
  // j + r nb0, j = tid + b T): fold x chirp and the filter spectrum.  The
  // chirp of the last step is loaded again (into the filter's registers)
  // once the filter product is formed.
  double2 v[16], hv[16], wv[2];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int b = e >> lr0, r = e & (R0 - 1);
    const int j = tid + b * T;
    const int m = j + r * nb0;
    v[e] = hv[e] = make_double2(0.0, 0.0);
    if (b < NB0 && j < nb0) {
      if (m < B)
        v[e] = Cmul(CONJ_IN ? Conj(fold[m]) : FoldAt(fold, m, ip.slices, c.fold_stride), chirp[m]);
      hv[e] = filt[m];
    }
  }
  // w1 of the radix-16 passes 1, 2 for butterfly j = tid:
  // W_{Ns 16}^(j mod Ns), Ns = R0 16^(q-1).
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    wv[q] = make_double2(1.0, 0.0);
    const int logNs = lr0 + 4 * q;  // pass q + 1
    if (q + 1 < npass && tid < M / 16) {
      const int k = tid & ((1 << logNs) - 1);
      wv[q] = c.tw[k << (kTwLogM - logNs - 4)];
    }
  }
  // Forward transform.
  ComputeStoreR<false>(a, v, logM, R0, 1, make_double2(1.0, 0.0));
  __syncthreads();
#pragma unroll
  for (int q = 1; q < 3; ++q) {
    if (q < npass) {
      PassLoad<16>(a, v, logM);
      __syncthreads();
      PassComputeStore<16, false>(a, v, logM, R0 << (4 * (q - 1)), wv[q - 1]);
      __syncthreads();
    }
  }
  // Inverse transform of (spectrum x filter); the filter is pre-scaled by 1/M.
  LoadR(a, v, logM, R0);
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    v[e] = Cmul(v[e], hv[e]);
    const int b = e >> lr0, r = e & (R0 - 1);
    const int j = tid + b * T;
    const int m = j + r * nb0;
    hv[e] = (b < NB0 && j < nb0 && m < B) ? chirp[m] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  ComputeStoreR<true>(a, v, logM, R0, 1, make_double2(1.0, 0.0));
  __syncthreads();
#pragma unroll
  for (int q = 1; q < 3; ++q) {
    if (q < npass) {
      PassLoad<16>(a, v, logM);
      __syncthreads();
      PassComputeStore<16, true>(a, v, logM, R0 << (4 * (q - 1)), wv[q - 1]);
      __syncthreads();
    }
  }
  // X_k = c_k y_k, times the row phase, in place (each thread its own k's).
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int b = e >> lr0, r = e & (R0 - 1);
    const int j = tid + b * T;
    const int m = j + r * nb0;
    if (b < NB0 && j < nb0 && m < B) a[Pad(m)] = Cmul(Cmul(a[Pad(m)], hv[e]), rp);
  }
}

template <int kPaths>
__device__ __forceinline__ void BinFftBody(const Ctx &c, int p, int blk, double2 *lds) {
  constexpr int T = kFftThreads;
  constexpr int kPerTerm = 2;  // BinInFreq terms prefetched per thread
  const DevState *st = c.st;
  const IterParams &ip = st->P[p];
  if (ip.done || ip.logM > kMaxLogM || !FftPathMatches<kPaths>(ip)) return;
  const int32_t trials = st->trials, rows = ip.rows;
  const int32_t trial = blk / c.rows_cap;
  const int32_t row = blk - trial * c.rows_cap;
  if (trial >= trials || row >= rows) return;
  const int32_t B = ip.B;
  const int tid = threadIdx.x;
  const int64_t n = c.n, mask = c.mask;
  const double2 *fold = c.fold + (int64_t(trial) * c.rows_cap + row) * c.B_cap;
  const int32_t cnt = min(st->count, c.cap);
  const int32_t *tb = c.term_bin + int64_t(trial) * c.cap;
  const double2 *tv = c.term_val + (int64_t(trial) * c.rows_cap + row) * c.cap;
  double2 tval[kPerTerm];
  int tbin[kPerTerm];
#pragma unroll
  for (int r = 0; r < kPerTerm; ++r) {
    const int i = tid + r * T;
    tbin[r] = 0;
    tval[r] = make_double2(0.0, 0.0);
    if (i < cnt) {
      tbin[r] = tb[i];
      tval[r] = tv[i];
    }
  }
  const int64_t tau = ModN(ip.q[trial] + RowOffset(row) * B, n, mask);
  const double2 rp = CisFrac(ModN(ip.tf.b * tau, n, mask), n);
  double2 *a = lds;
  const bool small = ip.logM <= kFftSmallLogM;
  if constexpr (kPaths == kFftSmallOnly) {
    BinFftSmall<false>(c, ip, fold, rp, a);
  } else if constexpr (kPaths == kFftLargeOnly) {
    BinFftLarge<false>(c, ip, fold, rp, a);
  } else {
    if (small)
      BinFftSmall<false>(c, ip, fold, rp, a);
    else
      BinFftLarge<false>(c, ip, fold, rp, a);
  }
  __syncthreads();
  // BinInFreq: every bin's terms summed exactly (TermGrid), then A = X - sum.
  // The limbs (2B double2 slots <= M) take the FFT array's place once the
  // outputs are in registers.
  double2 xo[OutPer<kPaths>()];
  TakeOutputs(a, small, B, xo);
  unsigned long long *limbs = reinterpret_cast<unsigned long long *>(a);
  double mx = 0.0;
#pragma unroll
  for (int r = 0; r < kPerTerm; ++r)
    if (tid + r * T < cnt) mx = fmax(mx, MaxComp(tval[r]));
  for (int i = tid + kPerTerm * T; i < cnt; i += T) mx = fmax(mx, MaxComp(tv[i]));
  for (int k = tid; k < kLimbs * B; k += T) limbs[k] = 0ull;
  const TermGrid g = MakeTermGrid(mx);
  if (g.any) {
#pragma unroll
    for (int r = 0; r < kPerTerm; ++r)
      if (tid + r * T < cnt) LimbAdd(limbs + kLimbs * tbin[r], tval[r], g);
    for (int i = tid + kPerTerm * T; i < cnt; i += T) LimbAdd(limbs + kLimbs * tb[i], tv[i], g);
  }
  __syncthreads();
  StoreBins(c.A + (int64_t(trial) * c.rows_cap + row) * c.B_cap, B, xo, limbs, g);
}

__global__ void __launch_bounds__(kFftThreads) k_bin_fft(Ctx c, int p) {
  HIP_DYNAMIC_SHARED(double2, lds)
  BinFftBody<kFftBoth>(c, p, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// The same Bluestein DFT for B > kMaxLdsBins, where the M-point FFT does not
// fit one workgroup's LDS.  M = N1 * N2 is split four-step style:
//   forward  X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) W_M^(n2 k1)
//                            sum_n1 W_N1^(n1 k1) u[N2 n1 + n2]
//   inverse  y[N2 n1 + n2] = sum_k1 W_N1^(-n1 k1) W_M^(-n2 k1)
//                            sum_k2 W_N2^(-n2 k2) U[k1 + N1 k2]
// k_big_cols : per column group, the N1-point FFTs over n1 and the W_M^(n2 k1)
//              twist, stored k1-major ([k1][n2]) in the scratch slot;
// k_big_rows : per k1 group, the N2-point FFT over n2, the product with the
//              filter spectrum, the inverse N2-point FFT and the conjugate
//              twist, in place;
// k_big_out  : per column group, the inverse N1-point FFTs over k1, then as
//              k_bin_fft: chirp, row phase, BinInFreq subtraction, store A.
// The filter spectrum H = FFT(h)/M is not tabulated per B: k_big_cols
// transforms h as one more row (the last scratch slot) and every k_big_rows
// workgroup finishes the transform of the k1 group it needs.  Each workgroup
// holds kBigTile elements; the scratch slot of a row holds M of them.
struct BigGeom {
  int logN1, logN2, N1, N2, M;
  __device__ explicit BigGeom(int logM)
      : logN1(logM / 2), logN2(logM - logM / 2), N1(1 << (logM / 2)),
        N2(1 << (logM - logM / 2)), M(1 << logM) {}
};

// W_M^j = e^{-2 pi i j / M} for 0 <= j < M from the half-period table.
__device__ __forceinline__ double2 TwFull(const double2 *tw, int logM, int j) {
  const int half = 1 << (logM - 1);
  const int stride = kTwM >> logM;
  if (j < half) return tw[j * stride];
  const double2 w = tw[(j - half) * stride];
  return make_double2(-w.x, -w.y);
}

// Resolves blockIdx.x into (trial, row, group) of a four-step kernel;
// false when the workgroup has nothing to do this iteration.
__device__ __forceinline__ bool BigSlot(const Ctx &c, const IterParams &ip, int32_t trials,
                                        bool with_filter, int32_t *slot, int32_t *trial,
                                        int32_t *row, int32_t *g) {
  const int32_t gpr = 1 << (ip.logM - 10);  // kBigTile = 2^10
  const int32_t s = blockIdx.x / c.big_gpr;
  *g = blockIdx.x - s * c.big_gpr;
  *slot = s;
  if (*g >= gpr) return false;
  if (s == c.trials_cap * c.rows_cap) {  // the filter row
    *trial = -1;
    *row = -1;
    return with_filter;
  }
  *trial = s / c.rows_cap;
  *row = s - *trial * c.rows_cap;
  return *trial < trials && *row < ip.rows;
}

__global__ void __launch_bounds__(kThreads) k_big_cols(Ctx c, int p) {
  __shared__ double2 a[kBigTile];
  const DevState *st = c.st;
  const IterParams &ip = st->P[p];
  if (ip.done || ip.logM <= kMaxLogM) return;
  int32_t slot, trial, row, g;
  if (!BigSlot(c, ip, st->trials, true, &slot, &trial, &row, &g)) return;
  const BigGeom G(ip.logM);
  const int32_t B = ip.B;
  const int logC = 10 - G.logN1;  // columns per workgroup = kBigTile / N1
  const int C = 1 << logC;
  const int n2_0 = g * C;
  const double2 *fold =
      trial >= 0 ? c.fold + (int64_t(trial) * c.rows_cap + row) * c.B_cap : nullptr;
  const double inv_m = 1.0 / double(G.M);
  for (int e = threadIdx.x; e < kBigTile; e += kThreads) {
    const int col = e & (C - 1), n1 = e >> logC;
    const int m = (n1 << G.logN2) + n2_0 + col;
    double2 v = make_double2(0.0, 0.0);
    if (fold) {
      if (m < B) v = Cmul(fold[m], Chirp(m, B));
    } else if (m < B) {
      v = Scale(Conj(Chirp(m, B)), inv_m);
    } else if (m > G.M - B) {
      v = Scale(Conj(Chirp(G.M - m, B)), inv_m);
    }
    a[(col << G.logN1) + BitRev(n1, G.logN1)] = v;
  }
  __syncthreads();
  FftDit(a, G.logN1, Twiddles{c.tw, kTwM >> G.logN1, false}, C);
  double2 *out = c.big + int64_t(slot) * c.big_gpr * kBigTile;
  for (int e = threadIdx.x; e < kBigTile; e += kThreads) {
    const int col = e & (C - 1), k1 = e >> logC;
    const int n2 = n2_0 + col;
    out[(k1 << G.logN2) + n2] = Cmul(a[(col << G.logN1) + k1], TwFull(c.tw, ip.logM, n2 * k1));
  }
}

__global__ void __launch_bounds__(kThreads) k_big_rows(Ctx c, int p) {
  __shared__ double2 a[2 * kBigTile];  // [data rows | filter rows]
  const DevState *st = c.st;
  const IterParams &ip = st->P[p];
  if (ip.done || ip.logM <= kMaxLogM) return;
  int32_t slot, trial, row, g;
  if (!BigSlot(c, ip, st->trials, false, &slot, &trial, &row, &g)) return;
  const BigGeom G(ip.logM);
  const int logC = 10 - G.logN2;  // k1 rows per workgroup = kBigTile / N2
  const int C = 1 << logC;
  const int64_t stride = int64_t(c.big_gpr) * kBigTile;
  double2 *y = c.big + slot * stride + g * kBigTile;
  const double2 *h = c.big + int64_t(c.trials_cap) * c.rows_cap * stride + g * kBigTile;
  for (int e = threadIdx.x; e < kBigTile; e += kThreads) {
    a[e] = y[e];
    a[kBigTile + e] = h[e];
  }
  __syncthreads();
  const Twiddles Wf{c.tw, kTwM >> G.logN2, false};
  FftDif(a, G.logN2, Wf, 2 * C);
  for (int e = threadIdx.x; e < kBigTile; e += kThreads) a[e] = Cmul(a[e], a[kBigTile + e]);
  __syncthreads();
  Twiddles Wi = Wf;
  Wi.inverse = true;
  FftDit(a, G.logN2, Wi, C);
  for (int e = threadIdx.x; e < kBigTile; e += kThreads) {
    const int k1 = g * C + (e >> G.logN2), n2 = e & (G.N2 - 1);
    y[e] = Cmul(a[e], Conj(TwFull(c.tw, ip.logM, n2 * k1)));
  }
}

__global__ void __launch_bounds__(kThreads) k_big_out(Ctx c, int p) {
  __shared__ double2 a[kBigTile];
  const DevState *st = c.st;
  const IterParams &ip = st->P[p];
  if (ip.done || ip.logM <= kMaxLogM) return;
  int32_t slot, trial, row, g;
  if (!BigSlot(c, ip, st->trials, false, &slot, &trial, &row, &g)) return;
  const BigGeom G(ip.logM);
  const int32_t B = ip.B;
  const int logC = 10 - G.logN1;
  const int C = 1 << logC;
  const int n2_0 = g * C;
  const int64_t n = c.n, mask = c.mask;
  const double2 *z = c.big + int64_t(slot) * c.big_gpr * kBigTile;
  for (int e = threadIdx.x; e < kBigTile; e += kThreads) {
    const int col = e & (C - 1), k1 = e >> logC;
    a[(col << G.logN1) + BitRev(k1, G.logN1)] = z[(k1 << G.logN2) + n2_0 + col];
  }
  __syncthreads();
  FftDit(a, G.logN1, Twiddles{c.tw, kTwM >> G.logN1, true}, C);
  const int64_t tau = ModN(ip.q[trial] + RowOffset(row) * B, n, mask);
  const double2 rp = CisFrac(ModN(ip.tf.b * tau, n, mask), n);
  for (int e = threadIdx.x; e < kBigTile; e += kThreads) {
    const int col = e & (C - 1), n1 = e >> logC;
    const int m = (n1 << G.logN2) + n2_0 + col;
    if (m < B) {
      const int i = (col << G.logN1) + n1;
      a[i] = Cmul(Cmul(a[i], Chirp(m, B)), rp);
    }
  }
  // BinInFreqV1 (src/binner.cc:565-591): the found modes whose bin lies in
  // this workgroup's columns, summed exactly per bin (TermGrid) and
  // subtracted once.  The outputs move to registers and the limbs take the
  // tile's place in LDS, half of the tile's bins at a time (4 limbs per bin:
  // half a tile of limbs fills the array).
  constexpr int kPer = kBigTile / kThreads;
  constexpr int kHalf = kBigTile / 2;
  const int32_t cnt = min(st->count, c.cap);
  const int32_t *tb = c.term_bin + int64_t(trial) * c.cap;
  const double2 *tv = c.term_val + (int64_t(trial) * c.rows_cap + row) * c.cap;
  __syncthreads();
  double2 xo[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int e = threadIdx.x + r * kThreads;
    const int col = e & (C - 1), n1 = e >> logC;
    xo[r] = a[(col << G.logN1) + n1];
  }
  double mx = 0.0;
  for (int i = threadIdx.x; i < cnt; i += kThreads) {
    const int col = (tb[i] & (G.N2 - 1)) - n2_0;
    if (col >= 0 && col < C) mx = fmax(mx, MaxComp(tv[i]));
  }
  const TermGrid tg = MakeTermGrid(mx);  // its barrier: every output is out of LDS
  unsigned long long *limbs = reinterpret_cast<unsigned long long *>(a);
  double2 *A = c.A + (int64_t(trial) * c.rows_cap + row) * c.B_cap;
  for (int h = 0; h < 2; ++h) {
    if (tg.any) {
      for (int k = threadIdx.x; k < kLimbs * kHalf; k += kThreads) limbs[k] = 0ull;
      __syncthreads();
      for (int i = threadIdx.x; i < cnt; i += kThreads) {
        const int bin = tb[i];
        const int col = (bin & (G.N2 - 1)) - n2_0;
        const int li = (col << G.logN1) + (bin >> G.logN2) - h * kHalf;
        if (col >= 0 && col < C && li >= 0 && li < kHalf) LimbAdd(limbs + kLimbs * li, tv[i], tg);
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int e = threadIdx.x + r * kThreads;
      const int col = e & (C - 1), n1 = e >> logC;
      const int m = (n1 << G.logN2) + n2_0 + col;
      const int li = (col << G.logN1) + n1 - h * kHalf;
      if (m < B && li >= 0 && li < kHalf)
        A[m] = tg.any ? Sub(xo[r], LimbValue(limbs + kLimbs * li, tg)) : xo[r];
    }
    if (tg.any) __syncthreads();  // the limbs are read before the next half clears them
  }
}

// ---------------------------------------------------------------------------
// k_identify: Iterate's per-bin body (src/iterate.cc:113-175), one
// wavefront per bin.  Each lane loads one or more (trial, row) entries of
// the bin's column of A (one memory round trip); lanes read each other's
// entries through cross-lane shuffles.  Then: the energy test; IdentifyFreq
// (src/iterate.cc:44-76) with lane j voting for bit j through the
// closed-form 2x2 pencil; k1 rounding and the window test; the coefficient
// estimate reduced across the lanes; un-permutation and mm[k0] += coef.  FP
// contraction is off so the decision arithmetic rounds as the reference's
// does.
// mm[k0] += v.  A key not yet present is appended to the workgroup's list
// (LDS); AppendKeys later reserves the list's slots in c.keys with one
// global atomic per workgroup instead of one per key (a single counter
// updated by every inserting wave serialises them: ~1500 inserts took
// ~30 us at B = 8001).
constexpr int kIdentWaves = kThreads / 64;
__device__ __forceinline__ void InsertMode(const Ctx &c, int64_t k0, double2 v, int32_t *ncnt,
                                           int64_t *nkeys) {
  const uint32_t bit = 1u << (k0 & 31);
  const uint32_t old = atomicOr(&c.present[k0 >> 5], bit);
  if (!(old & bit)) {
    const int32_t li = atomicAdd(ncnt, 1);  // LDS
    if (li < kIdentKeys) {
      nkeys[li] = k0;
    } else {  // more new keys than the list holds
      const int32_t idx = atomicAdd(&c.st->count, 1);
      if (idx < c.cap)
        c.keys[idx] = k0;
      else
        atomicOr(&c.st->flags, kFlagOverflow);
    }
  }
  atomicAdd(&c.coef[k0].x, v.x);
  atomicAdd(&c.coef[k0].y, v.y);
}

// After the workgroup's bins: reserve slots for its new keys and store them.
__device__ __forceinline__ void AppendKeys(const Ctx &c, int32_t *ncnt, int32_t *nbase,
                                           const int64_t *nkeys) {
  __syncthreads();
  const int32_t m = min(*ncnt, kIdentKeys);
  if (threadIdx.x == 0 && m > 0) *nbase = atomicAdd(&c.st->count, m);
  __syncthreads();
  if (int(threadIdx.x) < m) {
    const int32_t idx = *nbase + int(threadIdx.x);
    if (idx < c.cap)
      c.keys[idx] = nkeys[threadIdx.x];
    else
      atomicOr(&c.st->flags, kFlagOverflow);
  }
}

__device__ __forceinline__ double WaveSum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ double WaveMin(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  return v;
}

// Entry `src` of a column held as col[s] = entry (lane + 64 s).  Every lane
// of the wave must call it (the shuffles read other lanes' registers).
template <int NS>
__device__ __forceinline__ double2 ColFetch(const double2 (&col)[NS], int src) {
  double2 r = make_double2(0.0, 0.0);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double x = __shfl(col[s].x, src & 63);
    const double y = __shfl(col[s].y, src & 63);
    if ((src >> 6) == s) r = make_double2(x, y);
  }
  return r;
}

// blk, nblk: this workgroup's block index and the block count of the
// signal's launch (bins are distributed over nblk x waves wavefronts).
// NS: column slices of 64 lanes (1 when trials x rows <= 64, else
// kColSlices); kIo: the per-bin outputs of the identify-and-estimate entry
// point are compiled in (c.idout is used only then).  A solve's kernels run
// with one slice and no outputs, which leaves them registers for more
// wavefronts per SIMD.
template <int NS, bool kIo>
__device__ __forceinline__ void IdentifyBody(const Ctx &c, int p, int blk, int nblk) {
#pragma clang fp contract(off)
  DevState *st = c.st;
  const IterParams &ip = st->P[p];
  if (ip.done) return;
  const int32_t B = ip.B, bits = ip.bits, trials = st->trials;
  const int32_t rows = 1 + 2 * bits;
  const int32_t npairs = trials * rows;  // <= 64 * kColSlices (host-checked)
  const int64_t n = c.n, mask = c.mask;
  const DevTf tf = ip.tf;
  const double bin_thr = ip.bin_thr, sv_thr = ip.sv_thr, win_thr = st->win_thr;
  const double width = ip.width, dfreq = ip.dfreq;
  const int64_t rstride = c.B_cap, tstride = int64_t(c.rows_cap) * c.B_cap;
  const int lane = threadIdx.x & 63;
  const int waves = blockDim.x >> 6;
  __shared__ int32_t ncnt, nbase, nany;
  __shared__ int64_t nkeys[kIdentKeys];
  if (threadIdx.x == 0) {
    ncnt = 0;
    nany = 0;
  }
  __syncthreads();

  const bool replay = st->mode == kModeReplay;
  int64_t *rec = (c.rec && ip.it < c.rec_cap) ? c.rec + int64_t(ip.it) * c.B_cap : nullptr;
  // With per-bin outputs (the identify-and-estimate entry point) every bin
  // reports its decision, k1, k0, coefficient and decision margin, and the
  // found-mode map is left alone.
  const IdOut *const io = kIo ? c.idout : nullptr;
  for (int32_t b = blk * waves + (threadIdx.x >> 6); b < B; b += nblk * waves) {
    const double2 *Ab = c.A + b;
    double2 col[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int p = lane + 64 * s;
      col[s] = make_double2(0.0, 0.0);
      if (p < npairs) {
        const int t = p / rows;
        col[s] = Ab[t * tstride + (p - t * rows) * rstride];
      }
    }
    int64_t k1;
    double wf;
    if (replay) {  // the recorded decision for this bin
      k1 = rec ? rec[b] : -1;
      if (k1 < 0) continue;
      const double xi0 = double(k1) / double(n) - (double(b) + 0.5) / double(B);
      wf = SampleInFreq(xi0, width, dfreq);
    } else {
    if (rec && lane == 0) rec[b] = -1;
    // the smallest relative distance of a decision from its threshold
    // (Iterate's comparisons, src/iterate.cc:121,134,142 and
    // src/iterate.cc:58,62,68); only with per-bin outputs
    double margin = 1e300;
    if (bin_thr > 0) {
      double e = 0;
      for (int t = 0; t < trials; ++t) {
        const double2 v = ColFetch(col, t * rows);
        e += v.x * v.x + v.y * v.y;
      }
      e /= trials;
      const double thr2 = bin_thr * bin_thr;
      if (io) margin = fmin(margin, fabs(e - thr2) / thr2);
      if (e < thr2) {
        if (io && lane == 0) io->Record(b, kDecideEnergy, -1, margin);
        continue;
      }
    }
    // IdentifyFreq: lane j < bits takes bit j (offset B 2^j).  MatPencil
    // (src/freqid.cc:27-48) in closed form: the vote is
    // Im(conj(u0) u2 + conj(u1) u0) > 0 and sigma2 = |u0^2 - u1 u2|/sigma1.
    const int jb = lane < bits ? lane : 0;
    int32_t count = 0;
    bool rej = false;
    for (int t = 0; t < trials; ++t) {
      const double2 u0 = ColFetch(col, t * rows);
      const double2 u1 = ColFetch(col, t * rows + 2 * jb + 1);
      const double2 u2 = ColFetch(col, t * rows + 2 * jb + 2);
      const double alpha = (u0.x * u0.x + u0.y * u0.y) + (u1.x * u1.x + u1.y * u1.y);
      const double gamma = (u2.x * u2.x + u2.y * u2.y) + (u0.x * u0.x + u0.y * u0.y);
      const double2 beta = Add(CmulRn(Conj(u0), u2), CmulRn(Conj(u1), u0));
      if (beta.y > 0) ++count;
      double s2 = 0.0;
      if (sv_thr > 0) {
        const double hh = 0.5 * (alpha - gamma);
        const double lam1 =
            0.5 * (alpha + gamma) + sqrt(hh * hh + (beta.x * beta.x + beta.y * beta.y));
        const double s1 = sqrt(lam1);
        const double2 det = Sub(CmulRn(u0, u0), CmulRn(u1, u2));
        s2 = s1 > 0 ? hypot(det.x, det.y) / s1 : 0.0;
        if (s2 * 0.5 > sv_thr) rej = true;
      }
      if (io && lane < bits) {
        const double ab = hypot(beta.x, beta.y);
        if (ab > 0) margin = fmin(margin, fabs(beta.y) / ab);
        if (sv_thr > 0) margin = fmin(margin, fabs(s2 * 0.5 - sv_thr) / sv_thr);
      }
    }
    if (io) margin = WaveMin(margin);
    const bool vote = lane < bits && count > trials / 2;
    if (__any(lane < bits && rej)) {
      if (io && lane == 0) io->Record(b, kDecideSv, -1, margin);
      continue;
    }
    const uint64_t vm = __ballot(vote);
    int32_t xi_sum = 0;
    for (int j = 0; j < bits; ++j) xi_sum = (xi_sum << 1) | int32_t((vm >> j) & 1);
    const double xi1 = double(2 * xi_sum + 1) / double(1 << (bits + 1));
    // k1 = std::round(n (b + xi1) / B)  (round half away from zero).
    const double kr = double(n) * (double(b) + xi1) / double(B);
    k1 = int64_t(round(kr));
    const double xi0 = double(k1) / double(n) - (double(b) + 0.5) / double(B);
    wf = SampleInFreq(xi0, width, dfreq);
    if (io) {
      margin = fmin(margin, fabs(fabs(kr - floor(kr)) - 0.5));
      margin = fmin(margin, fabs(fabs(wf) - win_thr) / win_thr);
    }
    if (fabs(wf) < win_thr) {
      if (io && lane == 0) io->Record(b, kDecideWindow, k1, margin);
      continue;
    }
    if (rec && lane == 0) rec[b] = k1;
    if (io && lane == 0) io->Record(b, kDecideFound, k1, margin);
    }
    // Coefficient estimate (src/iterate.cc:147-166): sum over trials and rows
    // of A * e^{-2 pi i (q + off*B) k1 / n}.
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int p = lane + 64 * s;
      if (p < npairs) {
        const int t = p / rows;
        const int row = p - t * rows;
        const int64_t q = ip.q[t];
        const double2 fac = Conj(CisFrac(ModN(q * k1, n, mask), n));
        double2 f = fac;
        if (row > 0) {
          const int bit = (row - 1) / 2;
          const double2 fac2 = Conj(CisFrac(ModN((int64_t(B) << bit) * k1, n, mask), n));
          f = Cmul(fac, (row & 1) ? fac2 : Conj(fac2));
        }
        acc = Add(acc, Cmul(col[s], f));
      }
    }
    acc.x = WaveSum(acc.x);
    acc.y = WaveSum(acc.y);
    if (lane == 0) {
      const double denom = double(trials * (1 + 2 * bits));
      double2 coef = make_double2(acc.x / denom, acc.y / denom);
      // Undo the permutation: k0 = a_inv (k1 - b) mod n; coef *= e^{-2 pi i c k0/n}.
      const int64_t k0 = ModN(tf.a_inv * ModN(k1 - tf.b, n, mask), n, mask);
      coef = Cmul(coef, Conj(CisFrac(ModN(tf.c * k0, n, mask), n)));
      coef = make_double2(coef.x / wf, coef.y / wf);
      if (io) {
        io->k0[b] = k0;
        io->coef[b] = coef;
      } else {
        InsertMode(c, k0, coef, &ncnt, nkeys);
      }
      nany = 1;
    }
  }
  AppendKeys(c, &ncnt, &nbase, nkeys);
  if (threadIdx.x == 0 && nany) atomicOr(&st->found_iter[p], 1);
}

template <int NS, bool kIo>
__global__ void __launch_bounds__(kThreads) k_identify(Ctx c, int p) {
  IdentifyBody<NS, kIo>(c, p, blockIdx.x, gridDim.x);
}

// Column slices a launch needs for trials x rows entries per bin.
inline int ColSlicesFor(int64_t trials_rows) { return trials_rows <= 64 ? 1 : kColSlices; }

// ---------------------------------------------------------------------------
// The late iterations of a solve.  Once an iteration starts with
// 2 (k - |found|) + 1 <= min_bins, every later one has B = min_bins (|found|
// only grows), so iteration i's gather-fold no longer depends on iteration
// i-1's identify.  An iteration then takes two launches instead of three:
//   k_late_fold(i): identify of i-1 (blocks below nb_ident) beside the
//                   gather-fold of i and, in kPrepBlocks blocks, the
//                   BinInFreq factors of i for the modes found before i-1
//                   (wf e^{2 pi i (q k1 + c k0)/n} times the row's phase:
//                   the term without the coefficient, which i-1's identify
//                   may still change);
//   k_late_fft(i):  i's loop control (stale count, stop), then its bin DFTs;
//                   the terms are the factors times the final coefficients,
//                   and for the modes i-1 found the whole term.
// Slots as in the three-launch form: iteration i (parity p) reads loop
// state of slot 1-p and writes slot p.  cnt_base[p] is the number of modes
// at the start of the iteration of parity p.

// Iteration parameters of a late iteration (false: it does not exist).
__device__ __forceinline__ bool LateGeometry(const Ctx &c, int p, IterParams *ip) {
  const DevState *st = c.st;
  const LoopCounters prev = st->L[1 - p];
  if (prev.done) return false;
  const int32_t it = prev.it + 1;
  if (it >= st->max_iter) return false;
  const BParams bp = c.btab[st->min_bins];
  FillParams(ip, bp, bp.bits);
  ip->bin_thr = bp.bin_thr;
  ip->sv_thr = bp.sv_thr;
  ip->tf = c.tfs[it];
  ip->it = it;
  ip->done = 0;
  return true;
}

// The BinInFreq factor of mode k0 for (q, row) and its bin.
__device__ __forceinline__ double2 RowFactor(const Ctx &c, const IterParams &ip, int64_t q, int row,
                                             int64_t k0, int64_t k1, int32_t *bin) {
  const double2 base = FreqBase(c, ip, q, k0, k1, bin);
  if (row == 0) return base;
  const double2 g = BitTwiddle(c, ip.B, (row - 1) / 2, k1);
  return Cmul(base, (row & 1) ? g : Conj(g));
}

template <typename Tx, int NS>
__global__ void __launch_bounds__(kThreads)
    k_late_fold(Ctx c, int p, int32_t nb_ident, int32_t with_identify) {
  HIP_DYNAMIC_SHARED(double2, lds)
  const int32_t blk = blockIdx.x;
  if (blk < nb_ident) {
    if (with_identify) IdentifyBody<NS, false>(c, 1 - p, blk, nb_ident);
    return;
  }
  IterParams ip;
  if (!LateGeometry(c, p, &ip)) return;
  DevState *st = c.st;
  const int32_t trials = st->trials;
  const int32_t pb = blk - nb_ident;
  if (pb == 0 && threadIdx.x == 0) {  // publish the geometry for k_late_fft
    st->P[p] = ip;
    for (int t = 0; t < trials; ++t) st->P[p].q[t] = c.qs[int64_t(ip.it) * c.trials_cap + t];
    if (c.progress) {
      __hip_atomic_store(&c.progress[2], ip.B, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&c.progress[0], ip.it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  if (pb < kPrepBlocks) {
    // factors of the modes found before iteration i-1 (final keys)
    const int32_t cnt = min(st->cnt_base[1 - p], c.cap);
    const int32_t rows = ip.rows;
    const int64_t n = c.n, mask = c.mask;
    for (int64_t i = int64_t(pb) * blockDim.x + threadIdx.x; i < int64_t(cnt) * trials;
         i += int64_t(kPrepBlocks) * blockDim.x) {
      const int32_t trial = int32_t(i / cnt);
      const int32_t mi = int32_t(i - int64_t(trial) * cnt);
      const int64_t k0 = c.keys[mi];
      const int64_t q = c.qs[int64_t(ip.it) * c.trials_cap + trial];
      const int64_t k1 = ModN(ip.tf.a * k0 + ip.tf.b, n, mask);
      int32_t bin;
      const double2 base = FreqBase(c, ip, q, k0, k1, &bin);
      double2 *val = c.term_val + int64_t(trial) * c.rows_cap * c.cap + mi;
      val[0] = base;
      for (int32_t bit = 0; 2 * bit + 1 < rows; ++bit) {
        const double2 g = BitTwiddle(c, ip.B, bit, k1);
        val[int64_t(2 * bit + 1) * c.cap] = Cmul(base, g);
        val[int64_t(2 * bit + 2) * c.cap] = Cmul(base, Conj(g));
      }
      c.term_bin[int64_t(trial) * c.cap + mi] = bin;
    }
    return;
  }
  TileFold<Tx>(c, ip, trials, pb - kPrepBlocks, lds);
}

template <int kPaths>
__global__ void __launch_bounds__(kFftThreads) k_late_fft(Ctx c, int p) {
  HIP_DYNAMIC_SHARED(double2, lds)
  constexpr int T = kFftThreads;
  constexpr int kPerTerm = 4;  // factors prefetched per thread (S up to 1024 at once)
  DevState *st = c.st;
  // Loop control of this iteration (NextIteration's; identify of the
  // previous iteration has finished).
  const LoopCounters prev = st->L[1 - p];
  if (prev.done) {
    // Past the stop: carry it into this parity's slots, so the launches
    // still queued behind this one (the next k_late_fold's identify and
    // geometry read them) see a stopped loop too.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->L[p] = prev;
      st->P[p].done = 1;
      st->found_iter[p] = 0;
    }
    return;
  }
  LoopCounters lc;
  lc.it = prev.it + 1;
  lc.stale = prev.stale + (st->found_iter[1 - p] ? 0 : 1);
  lc.pad = 0;
  const bool done =
      (st->flags & kFlagOverflow) || lc.stale >= st->max_stale || lc.it >= st->max_iter;
  lc.done = done ? 1 : 0;
  const int32_t cnt_pre = min(st->cnt_base[1 - p], c.cap);  // modes with factors
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // PublishIteration's bookkeeping, for the slot k_late_fold filled.
    st->L[p] = lc;
    st->P[p].done = lc.done;  // read by the next launch's identify
    st->found_iter[p] = 0;
    st->cnt_base[p] = st->count;
    if (st->mode == kModeSolve && prev.it >= 0 && prev.it < c.log_cap) {
      c.log[prev.it] = st->P[1 - p].B;
      c.log[c.log_cap + prev.it] = st->count;
    }
    if (lc.done) {
      st->iters = lc.it;
      st->stop_stale = lc.stale;
    }
    if (c.progress && lc.done) PublishStop(c, lc.it, lc.stale, st->flags);
  }
  if (done) return;
  const IterParams &ip = st->P[p];
  const int32_t trials = st->trials, rows = ip.rows;
  const int32_t trial = blockIdx.x / c.rows_cap;
  const int32_t row = blockIdx.x - trial * c.rows_cap;
  if (trial >= trials || row >= rows) return;
  const int32_t B = ip.B;
  const int tid = threadIdx.x;
  const int64_t n = c.n, mask = c.mask;
  const double2 *fold = c.fold + (int64_t(trial) * c.rows_cap + row) * c.B_cap;
  const int32_t cnt = min(st->count, c.cap);
  const int32_t *tb = c.term_bin + int64_t(trial) * c.cap;
  const double2 *tv = c.term_val + (int64_t(trial) * c.rows_cap + row) * c.cap;
  // prefetch: factor, bin and key of the first modes
  double2 fac[kPerTerm];
  int tbin[kPerTerm];
  int64_t kk[kPerTerm];
#pragma unroll
  for (int r = 0; r < kPerTerm; ++r) {
    const int i = tid + r * T;
    fac[r] = make_double2(0.0, 0.0);
    tbin[r] = 0;
    kk[r] = -1;
    if (i < cnt) kk[r] = c.keys[i];
    if (i < cnt_pre) {
      fac[r] = tv[i];
      tbin[r] = tb[i];
    }
  }
  const int64_t q = ip.q[trial];
  const int64_t tau = ModN(q + RowOffset(row) * B, n, mask);
  const double2 rp = CisFrac(ModN(ip.tf.b * tau, n, mask), n);
  // their coefficients, final since the previous launch (identify of the
  // previous iteration), load while the DFT runs
  double2 cfv[kPerTerm];
#pragma unroll
  for (int r = 0; r < kPerTerm; ++r) cfv[r] = kk[r] >= 0 ? c.coef[kk[r]] : make_double2(0.0, 0.0);
  double2 *a = lds;
  const bool small = ip.logM <= kFftSmallLogM;
  if constexpr (kPaths == kFftSmallOnly) {
    BinFftSmall<false>(c, ip, fold, rp, a);
  } else {
    if (small)
      BinFftSmall<false>(c, ip, fold, rp, a);
    else
      BinFftLarge<false>(c, ip, fold, rp, a);
  }
  __syncthreads();
  // BinInFreq, as k_bin_fft's epilogue (exact per-bin sums of the terms):
  // the terms of the first modes are formed once, in registers, and the
  // rest (more than kPerTerm x T modes) are formed again for the second pass.
  double2 xo[OutPer<kPaths>()];
  TakeOutputs(a, small, B, xo);
  unsigned long long *limbs = reinterpret_cast<unsigned long long *>(a);
  auto term = [&](int i, int64_t k0, const double2 &cf, int32_t *bin) {
    if (i < cnt_pre) {
      *bin = tb[i];
      return Cmul(cf, tv[i]);
    }
    const int64_t k1 = ModN(ip.tf.a * k0 + ip.tf.b, n, mask);
    return Cmul(cf, RowFactor(c, ip, q, row, k0, k1, bin));
  };
  double2 vr[kPerTerm];
  double mx = 0.0;
#pragma unroll
  for (int r = 0; r < kPerTerm; ++r) {
    const int i = tid + r * T;
    vr[r] = make_double2(0.0, 0.0);
    if (i < cnt) {
      if (i < cnt_pre) {
        vr[r] = Cmul(cfv[r], fac[r]);
      } else {
        const int64_t k1 = ModN(ip.tf.a * kk[r] + ip.tf.b, n, mask);
        vr[r] = Cmul(cfv[r], RowFactor(c, ip, q, row, kk[r], k1, &tbin[r]));
      }
      mx = fmax(mx, MaxComp(vr[r]));
    }
  }
  for (int i = tid + kPerTerm * T; i < cnt; i += T) {
    int32_t bin;
    const int64_t k0 = c.keys[i];
    mx = fmax(mx, MaxComp(term(i, k0, c.coef[k0], &bin)));
  }
  for (int k = tid; k < kLimbs * B; k += T) limbs[k] = 0ull;
  const TermGrid g = MakeTermGrid(mx);
  if (g.any) {
#pragma unroll
    for (int r = 0; r < kPerTerm; ++r)
      if (tid + r * T < cnt) LimbAdd(limbs + kLimbs * tbin[r], vr[r], g);
    for (int i = tid + kPerTerm * T; i < cnt; i += T) {
      int32_t bin;
      const int64_t k0 = c.keys[i];
      const double2 v = term(i, k0, c.coef[k0], &bin);
      LimbAdd(limbs + kLimbs * bin, v, g);
    }
  }
  __syncthreads();
  StoreBins(c.A + (int64_t(trial) * c.rows_cap + row) * c.B_cap, B, xo, limbs, g);
}

// BinInFreqV1 (src/binner.cc:565-591) on a caller-owned bin matrix (rows x
// B, row stride B), one thread per mode, with each bin's terms summed
// exactly as in the solve (TermGrid) so the result does not depend on the
// order of the modes or of the threads: a pass for the grid exponent (the
// largest term component over all modes, by atomicMax), a pass adding the
// limbs into limbs[(row B + bin) kLimbs ...], and one subtracting each
// bin's sum.
template <typename Fn>
__device__ __forceinline__ void DirectTerms(int64_t n, int64_t mask, const BParams &bp,
                                            const DevTf &tf, int64_t q, int64_t k0s,
                                            double2 coef, Fn fn) {
  const int32_t B = bp.B;
  const int64_t k0 = ModN(k0s, n, mask);
  const int64_t k1 = ModN(tf.a * k0 + tf.b, n, mask);
  const int32_t bin = int32_t(k1 * B / n);
  const double xi = (double(bin) + 0.5) / double(B) - double(k1) / double(n);
  const double wf = SampleInFreq(xi, bp.width, bp.dfreq);
  const double2 base = Cmul(Scale(coef, wf), CisFrac(ModN(q * k1 + tf.c * k0, n, mask), n));
  fn(0, bin, base);
  for (int32_t bit = 0; bit < bp.bits; ++bit) {
    const double2 g = CisFrac(ModN((int64_t(B) << bit) * k1, n, mask), n);
    fn(2 * bit + 1, bin, Cmul(base, g));
    fn(2 * bit + 2, bin, Cmul(base, Conj(g)));
  }
}

__global__ void __launch_bounds__(kThreads)
    k_freq_direct_max(int64_t n, int64_t mask, BParams bp, DevTf tf, int64_t q,
                      const int64_t *k0s, const double2 *coefs, int64_t m, int32_t *emax) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double mx = 0.0;
  if (i < m)
    DirectTerms(n, mask, bp, tf, q, k0s[i], coefs[i],
                [&](int, int32_t, double2 v) { mx = fmax(mx, MaxComp(v)); });
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0 && mx > 0.0) atomicMax(emax, ilogb(mx) + 1);
}

// The grid of exponent e (every term component below 2^e); e = kNoTerms
// when there is no nonzero term.
constexpr int32_t kNoTerms = -(1 << 30);
__device__ __forceinline__ TermGrid TermGridOf(int32_t e) {
  TermGrid g;
  g.any = e != kNoTerms;
  g.scale = g.any ? ldexp(1.0, 94 - e) : 1.0;
  g.inv = g.any ? ldexp(1.0, e - 94) : 1.0;
  return g;
}

__global__ void __launch_bounds__(kThreads)
    k_freq_direct_add(int64_t n, int64_t mask, BParams bp, DevTf tf, int64_t q,
                      const int64_t *k0s, const double2 *coefs, int64_t m, const int32_t *emax,
                      unsigned long long *limbs) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const TermGrid g = TermGridOf(*emax);
  if (i >= m || !g.any) return;
  DirectTerms(n, mask, bp, tf, q, k0s[i], coefs[i], [&](int row, int32_t bin, double2 v) {
    LimbAdd(limbs + kLimbs * (int64_t(row) * bp.B + bin), v, g);
  });
}

__global__ void __launch_bounds__(kThreads)
    k_freq_direct_sub(const unsigned long long *limbs, const int32_t *emax, int64_t count,
                      double2 *out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const TermGrid g = TermGridOf(*emax);
  if (i < count && g.any) out[i] = Sub(out[i], LimbValue(limbs + kLimbs * i, g));
}

// limbs[0..count) = 0 and *emax = kNoTerms (a kernel rather than a memset).
__global__ void __launch_bounds__(kThreads)
    k_freq_direct_clear(unsigned long long *limbs, int64_t count, int32_t *emax) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int64_t j = i; j < count; j += int64_t(gridDim.x) * blockDim.x) limbs[j] = 0ull;
  if (i == 0) *emax = kNoTerms;
}

__global__ void __launch_bounds__(kThreads)
    k_mat_pencil(const double2 *u, int64_t m, double *sigma, uint8_t *vote) {
#pragma clang fp contract(off)
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double2 u0 = u[3 * i], u1 = u[3 * i + 1], u2 = u[3 * i + 2];
  const double alpha = (u0.x * u0.x + u0.y * u0.y) + (u1.x * u1.x + u1.y * u1.y);
  const double gamma = (u2.x * u2.x + u2.y * u2.y) + (u0.x * u0.x + u0.y * u0.y);
  const double2 beta = Add(CmulRn(Conj(u0), u2), CmulRn(Conj(u1), u0));
  const double hh = 0.5 * (alpha - gamma);
  const double lam1 = 0.5 * (alpha + gamma) + sqrt(hh * hh + (beta.x * beta.x + beta.y * beta.y));
  const double s1 = sqrt(lam1);
  const double2 det = Sub(CmulRn(u0, u0), CmulRn(u1, u2));
  sigma[2 * i] = s1;
  sigma[2 * i + 1] = s1 > 0 ? hypot(det.x, det.y) / s1 : 0.0;
  vote[i] = beta.y > 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// The batched form: one launch per stage covers every signal of a batch.
// map[blockIdx.x] = {signal, block within that signal's share}; each
// workgroup runs the single-signal body on its signal's Ctx.  Signals that
// have stopped exit at once.
__global__ void __launch_bounds__(kThreads)
    k_batch_reset(const Ctx *ctxs, const InitArgs *args, int32_t *counters) {
  const Ctx &c = ctxs[blockIdx.x];
  ClearModesBody(c, 0, threadIdx.x, blockDim.x);
  __syncthreads();
  if (threadIdx.x == 0) InitBody(c, args[blockIdx.x]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counters[0] = 0;  // batched iterations started
    counters[1] = 0;  // signals stopped
  }
}

// ---- the batched iteration: plan, then three launches over packed units
//
// k_batch_signal starts every signal's next iteration (NextIteration and
// PublishIteration, as k_bin_time's first workgroup does for one signal),
// one wavefront per signal, and k_batch_plan (one workgroup) packs the
// work of the three launches that follow: the units of
// signal i occupy [off[i], off[i] + cnt[i]) of each launch's map, so a
// launch's workgroups visit only units that exist this iteration (signals
// that stopped, and the tiles of a B smaller than the largest, cost
// nothing).  Units (kinds 0..3): k_bin_time: PrepUnits term blocks then
// trials x ceil(B/R) tiles; k_bin_fft: trials x rows (trial, row) DFTs, as
// kind 1 on the small FFT path (M <= 1024) or kind 3 on the large one, each
// path a kernel of its own; k_identify (kind 2): ceil(B / ident_bins) blocks
// of ident_bins bins (ident_bins / kIdentWaves per wavefront).
constexpr int kPlanThreads = 1024;

// Workgroups computing BinInFreq terms for `terms` (modes x trials), a
// multiple of 8 so that tile units keep their XCD (unit index mod 8).
__device__ __forceinline__ int32_t PrepUnits(int64_t terms) {
  const int32_t u = int32_t(min<int64_t>(kPrepBlocks, (terms + kThreads - 1) / kThreads));
  return (u + 7) & ~7;
}

// Signal blockIdx.x's loop control, on the first lane of a wavefront of
// its own (every signal's chain of dependent loads runs at once, spread
// over the CUs): NextIteration and PublishLoop into slot p, then the
// signal's unit counts and rows for k_batch_plan.
__global__ void __launch_bounds__(64)
    k_batch_signal(const Ctx *ctxs, int32_t count, int p, BatchWork w, int32_t ident_bins) {
  const int32_t i = blockIdx.x;
  if (threadIdx.x != 0 || i >= count) return;
  const Ctx &c = ctxs[i];
  // the iteration's parameters are built in place, in slot p
  IterParams &ip = c.st->P[p];
  LoopCounters lc;
  NextIteration(c, p, &ip, &lc);
  PublishLoop(c, p, ip, lc);
  int32_t u[kUnitKinds] = {0, 0, 0, 0};
  if (!ip.done) {
    const int32_t trials = c.st->trials;
    // a multiple of 8 units per signal: every signal's units start on XCD 0
    u[0] = (PrepUnits(int64_t(min(c.st->count, c.cap)) * trials) +
            trials * ((ip.B + ip.R - 1) / ip.R) * ip.slices + 7) & ~7;
    u[ip.logM <= kFftSmallLogM ? 1 : 3] = trials * ip.rows;
    u[2] = (ip.B + ident_bins - 1) / ident_bins;
  }
  for (int k = 0; k < kUnitKinds; ++k) w.cnt[k * count + i] = u[k];
  w.rows[i] = ip.rows;
}

__global__ void __launch_bounds__(kPlanThreads)
    k_batch_plan(const Ctx *ctxs, int32_t count, BatchWork w, int32_t *counters, int32_t *host) {
  __shared__ int32_t part[kUnitKinds][kPlanThreads];
  // with at most kPlanThreads signals (one per thread): each signal's unit
  // counts, rows and rows_cap, for the map fill below
  __shared__ int32_t s_cnt[kUnitKinds][kPlanThreads];
  __shared__ int32_t s_rows[2][kPlanThreads];
  const int tid = threadIdx.x;
  // signals [i0, i1) belong to this thread; k_batch_signal wrote their
  // counts and rows
  const int32_t per = (count + kPlanThreads - 1) / kPlanThreads;
  const int32_t i0 = min(count, tid * per), i1 = min(count, i0 + per);
  int32_t sum[kUnitKinds] = {0, 0, 0, 0};
  for (int32_t i = i0; i < i1; ++i) {
    for (int k = 0; k < kUnitKinds; ++k) {
      const int32_t u = w.cnt[k * count + i];
      sum[k] += u;
      if (per == 1) s_cnt[k][tid] = u;
    }
    if (per == 1) {
      s_rows[0][tid] = w.rows[i];
      s_rows[1][tid] = ctxs[i].rows_cap;
    }
  }
  // inclusive scan of the sums: within each wavefront by shuffles, then of
  // the wavefronts' totals by the first wavefront (two barriers in all)
  {
    __shared__ int32_t wtot[kUnitKinds][kPlanThreads / 64];
    const int lane = tid & 63, wv = tid >> 6;
    int32_t inc[kUnitKinds];
    for (int k = 0; k < kUnitKinds; ++k) {
      int32_t v = sum[k];
      for (int d = 1; d < 64; d <<= 1) {
        const int32_t t = __shfl(v, lane - d);  // every lane takes part
        if (lane >= d) v += t;
      }
      inc[k] = v;
      if (lane == 63) wtot[k][wv] = v;
    }
    __syncthreads();
    if (wv == 0) {
      for (int k = 0; k < kUnitKinds; ++k) {
        int32_t v = lane < kPlanThreads / 64 ? wtot[k][lane] : 0;
        for (int d = 1; d < kPlanThreads / 64; d <<= 1) {
          const int32_t t = __shfl(v, lane - d);
          if (lane >= d) v += t;
        }
        if (lane < kPlanThreads / 64) wtot[k][lane] = v;
      }
    }
    __syncthreads();
    for (int k = 0; k < kUnitKinds; ++k) part[k][tid] = inc[k] + (wv > 0 ? wtot[k][wv - 1] : 0);
  }
  int32_t base[kUnitKinds];
  for (int k = 0; k < kUnitKinds; ++k) base[k] = part[k][tid] - sum[k];
  for (int32_t i = i0; i < i1; ++i)
    for (int k = 0; k < kUnitKinds; ++k) {
      w.off[k * count + i] = base[k];
      base[k] += w.cnt[k * count + i];
    }
  if (tid == kPlanThreads - 1)
    for (int k = 0; k < kUnitKinds; ++k) w.total[k] = part[k][tid];
  // with more signals than threads, the counts and offsets above are read
  // below by other threads of this workgroup (device-scope loads)
  if (per > 1) __threadfence();
  __syncthreads();
  if (tid == 0) {
    // The single writer of the host's copy: batched iterations started and
    // signals stopped so far (counters[1] only grows, and each launch
    // publishes after the previous one has finished).
    const int32_t v = atomicAdd(&counters[0], 1) + 1;
    const int32_t d = __hip_atomic_load(&counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the large-path bin-FFT units of this iteration: a signal's B never
    // grows during a solve, so once none are left none come back
    __hip_atomic_store(&host[2], part[3][kPlanThreads - 1], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&host[1], d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&host[0], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  // the maps: one wavefront per signal
  const int lane = tid & 63;
  for (int32_t i = tid >> 6; i < count; i += kPlanThreads / 64) {
    int32_t off[kUnitKinds], cnt[kUnitKinds], rows, rows_cap;
    if (per == 1) {  // signal i belonged to thread i: its counts and inclusive sums in LDS
      for (int k = 0; k < kUnitKinds; ++k) {
        cnt[k] = s_cnt[k][i];
        off[k] = part[k][i] - cnt[k];
      }
      rows = s_rows[0][i];
      rows_cap = s_rows[1][i];
    } else {
      for (int k = 0; k < kUnitKinds; ++k) {
        off[k] = __hip_atomic_load(&w.off[k * count + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        cnt[k] = __hip_atomic_load(&w.cnt[k * count + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      rows = __hip_atomic_load(&w.rows[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      rows_cap = ctxs[i].rows_cap;
    }
    for (int32_t u = lane; u < cnt[0]; u += 64) w.map[0][off[0] + u] = make_int2(i, u);
    for (int k = 1; k < kUnitKinds; k += 2)  // the bin-FFT units, small or large path
      for (int32_t u = lane; u < cnt[k]; u += 64) {
        const int32_t trial = u / rows;
        w.map[k][off[k] + u] = make_int2(i, trial * rows_cap + (u - trial * rows));
      }
    for (int32_t u = lane; u < cnt[2]; u += 64) w.map[2][off[2] + u] = make_int2(i, u);
  }
}

// One k_bin_time unit of a published iteration: BinInFreq terms (the first
// PrepUnits units) or a tile of the gather-fold.
template <typename Tx>
__device__ __forceinline__ void BinTimeUnit(const Ctx &c, int p, int32_t unit, double2 *lds) {
  const DevState *st = c.st;
  const IterParams &ip = st->P[p];
  const int32_t trials = st->trials;
  const int32_t nprep = PrepUnits(int64_t(min(st->count, c.cap)) * trials);
  if (unit < nprep)
    FreqTerms(c, ip, trials, unit, nprep);
  else
    TileFold<Tx, false>(c, ip, trials, unit - nprep, lds);
}

template <typename Tx>
__global__ void __launch_bounds__(kThreads)
    k_bin_time_batch(const Ctx *ctxs, const int2 *map, const int32_t *total, int p) {
  HIP_DYNAMIC_SHARED(double2, lds)
  const int32_t tot = total[0];
  for (int32_t u = blockIdx.x; u < tot; u += gridDim.x) {
    if (u != int32_t(blockIdx.x)) __syncthreads();  // the previous unit's LDS reads are done
    const int2 m = map[u];
    BinTimeUnit<Tx>(ctxs[m.x], p, m.y, lds);
  }
}

template <int kPaths>
__global__ void __launch_bounds__(kFftThreads)
    k_bin_fft_batch(const Ctx *ctxs, const int2 *map, const int32_t *total, int p) {
  HIP_DYNAMIC_SHARED(double2, lds)
  const int32_t tot = total[kPaths == kFftSmallOnly ? 1 : 3];
  for (int32_t u = blockIdx.x; u < tot; u += gridDim.x) {
    if (u != int32_t(blockIdx.x)) __syncthreads();
    const int2 m = map[u];
    BinFftBody<kPaths>(ctxs[m.x], p, m.y, lds);
  }
}

template <int NS>
__global__ void __launch_bounds__(kThreads, NS == 1 ? 3 : 1)
    k_identify_batch(const Ctx *ctxs, const int2 *map, const int32_t *total, const int32_t *cnt,
                     int p) {
  const int32_t tot = total[2];
  for (int32_t u = blockIdx.x; u < tot; u += gridDim.x) {
    if (u != int32_t(blockIdx.x)) __syncthreads();
    const int2 m = map[u];
    IdentifyBody<NS, false>(ctxs[m.x], p, m.y, cnt[m.x]);
  }
}

// Per signal: what the host needs to report the batched solve.
__global__ void __launch_bounds__(kThreads)
    k_batch_results(const Ctx *ctxs, int32_t count, BatchResultDev *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const DevState *st = ctxs[i].st;
  BatchResultDev r;
  r.iters = st->iters;
  r.stop_stale = st->stop_stale;
  r.flags = st->flags;
  r.stopped = (st->L[0].done | st->L[1].done) ? 1 : 0;
  r.last_it = max(st->L[0].it, st->L[1].it);
  r.count = min(st->count, ctxs[i].cap);
  out[i] = r;
}

// ---------------------------------------------------------------------------
// Adjoint of the recovery map with the decisions of a recorded solve held
// fixed (support, transforms, bins, identified k1's).  The coefficients are
// then linear in x: per iteration i, A_i = D_i x - T_i c^(i) (BinInTime and
// BinInFreq) and c^(i+1) = c^(i) + E_i A_i (the estimates of the bins that
// identified a mode).  Going back over the iterations with lam = dL/dc
// (dense over k0, the conjugate-Wirtinger convention of torch):
//   Abar    = E_i^H lam                          AdjBinsBody
//   lam    -= T_i^H Abar                         AdjTermsBody
//   fbar_i  = conj(rp) DFT^H Abar (per row)      AdjFftBody, kept per iteration
// (with iteration i's conjugated taps, AdjTapsBody) and at the end, for every
// sample j at once,
//   xbar[j] = sum_i fold_i^H fbar_i              AdjGatherBody.
// The last step is the fold read backwards as a gather: sample j was
// staged by iteration i at u = a_inv (j - c) mod n, i.e. at every lag
// lambda and residue rho with lambda B + rho = u - q + p2 (mod n) whose
// lambda lies in some row's window, so each output is written once, with
// no atomics and no accumulation buffer.  Only lam entries at found keys
// are ever read, so only those are cleared.

// e^{-2 pi i (q + off(row) B) k1 / n}: the estimate's factor for one entry
// (src/iterate.cc:147-166).
__device__ __forceinline__ double2 EstFactor(int64_t q, int32_t B, int row, int64_t k1, int64_t n,
                                             int64_t mask) {
  const double2 fac = Conj(CisFrac(ModN(q * k1, n, mask), n));
  if (row == 0) return fac;
  const int bit = (row - 1) / 2;
  const double2 fac2 = Conj(CisFrac(ModN((int64_t(B) << bit) * k1, n, mask), n));
  return Cmul(fac, (row & 1) ? fac2 : Conj(fac2));
}

// Abar[t][row][b] = conj(d coef / d A[t][row][b]) lam[k0] for every bin that
// identified a mode in this iteration, 0 elsewhere.  One wavefront per bin.
__device__ __forceinline__ void AdjBinsBody(const Ctx &c, const IterParams &ip, int32_t trials,
                                            const double2 *lam, const int64_t *rec_row, int blk,
                                            int nblk) {
  const int32_t B = ip.B, rows = ip.rows;
  const int32_t npairs = trials * rows;
  const int64_t n = c.n, mask = c.mask;
  const int64_t rstride = c.B_cap, tstride = int64_t(c.rows_cap) * c.B_cap;
  const int lane = threadIdx.x & 63;
  const int waves = blockDim.x >> 6;
  const double denom = double(trials * rows);
  const DevTf tf = ip.tf;
  const double width = ip.width, dfreq = ip.dfreq;
  for (int32_t b = blk * waves + (threadIdx.x >> 6); b < B; b += nblk * waves) {
    const int64_t k1 = rec_row[b];
    double2 g = make_double2(0.0, 0.0), cz = make_double2(0.0, 0.0);
    double wf = 1.0;
    if (k1 >= 0) {
      const int64_t k0 = ModN(tf.a_inv * ModN(k1 - tf.b, n, mask), n, mask);
      g = lam[k0];
      cz = Conj(CisFrac(ModN(tf.c * k0, n, mask), n));
      const double xi0 = double(k1) / double(n) - (double(b) + 0.5) / double(B);
      wf = SampleInFreq(xi0, width, dfreq);
    }
    for (int pi = lane; pi < npairs; pi += 64) {
      const int t = pi / rows, row = pi - t * rows;
      double2 v = make_double2(0.0, 0.0);
      if (k1 >= 0) {
        const double2 h = Scale(Cmul(EstFactor(ip.q[t], B, row, k1, n, mask), cz), 1.0 / (wf * denom));
        v = Cmul(Conj(h), g);
      }
      c.A[t * tstride + row * rstride + b] = v;
    }
  }
}

// lam[k0] -= sum_{t,row} conj(term factor) Abar at k0's bin, for every mode
// in the map before this iteration (keys[0..count_before)).  Threads tid of
// a grid of gsize.
__device__ __forceinline__ void AdjTermsBody(const Ctx &c, const IterParams &ip, int32_t trials,
                                             double2 *lam, int32_t count_before,
                                             const int64_t *rec_row, int64_t tid, int64_t gsize) {
  const int32_t B = ip.B, bits = ip.bits;
  const int64_t n = c.n, mask = c.mask;
  const DevTf tf = ip.tf;
  const int64_t rstride = c.B_cap, tstride = int64_t(c.rows_cap) * c.B_cap;
  for (int64_t i = tid; i < count_before; i += gsize) {
    const int64_t k0 = c.keys[i];
    const int64_t k1 = ModN(tf.a * k0 + tf.b, n, mask);
    const int32_t bin = int32_t(k1 * B / n);
    if (rec_row[bin] < 0) continue;  // Abar is 0 there
    const double xi = (double(bin) + 0.5) / double(B) - double(k1) / double(n);
    const double wf = SampleInFreq(xi, ip.width, ip.dfreq);
    double2 acc = make_double2(0.0, 0.0);
    for (int t = 0; t < trials; ++t) {
      const double2 base = Scale(CisFrac(ModN(ip.q[t] * k1 + tf.c * k0, n, mask), n), wf);
      const double2 *Ab = c.A + t * tstride + bin;
      acc = Add(acc, Cmul(Conj(base), Ab[0]));
      for (int32_t bit = 0; bit < bits; ++bit) {
        const double2 gv = CisFrac(ModN((int64_t(B) << bit) * k1, n, mask), n);
        acc = Add(acc, Cmul(Conj(Cmul(base, gv)), Ab[(2 * bit + 1) * rstride]));
        acc = Add(acc, Cmul(Conj(Cmul(base, Conj(gv))), Ab[(2 * bit + 2) * rstride]));
      }
    }
    lam[k0] = Sub(lam[k0], acc);
  }
}

// fbar[t][row] = conj(rp) DFT^H Abar[t][row] = conj(DFT(conj(Abar)) rp): the
// forward bin FFT on conj(Abar), output conjugated (the Bluestein matrix is
// symmetric).  Block blk = trial * rows_cap + row; fbar holds
// trials_cap x rows_cap x B_cap entries.
template <int kPaths>
__device__ __forceinline__ void AdjFftBody(const Ctx &c, const IterParams &ip, int32_t trials,
                                           double2 *fbar, int blk, double2 *lds) {
  const int32_t trial = blk / c.rows_cap;
  const int32_t row = blk - trial * c.rows_cap;
  if (trial >= trials || row >= ip.rows) return;
  const int32_t B = ip.B;
  const int64_t n = c.n, mask = c.mask;
  const int64_t tau = ModN(ip.q[trial] + RowOffset(row) * B, n, mask);
  const double2 rp = CisFrac(ModN(ip.tf.b * tau, n, mask), n);
  const int64_t off = (int64_t(trial) * c.rows_cap + row) * c.B_cap;
  double2 *a = lds;
  // kFftSmallOnly: the launch's signals all have M <= 2^kFftSmallLogM, and
  // the kernel carries the small path's registers only
  const bool small = kPaths == kFftSmallOnly || ip.logM <= kFftSmallLogM;
  if constexpr (kPaths == kFftSmallOnly) {
    BinFftSmall<true>(c, ip, c.A + off, rp, a);
  } else {
    if (small)
      BinFftSmall<true>(c, ip, c.A + off, rp, a);
    else
      BinFftLarge<true>(c, ip, c.A + off, rp, a);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < B; k += kFftThreads) fbar[off + k] = Conj(a[small ? k : Pad(k)]);
}

// taps[k B + rho] = conj(tap(k B + rho - p2)), k < f, rho < B: the taps of
// k_bin_time, conjugated.  Threads tid of a grid of gsize.
__device__ __forceinline__ void AdjTapsBody(const IterParams &ip, int64_t n, int64_t mask,
                                            double2 *taps, int64_t tid, int64_t gsize) {
  const int32_t B = ip.B;
  const double delta = -0.5 / double(B);
  for (int64_t i = tid; i < int64_t(ip.f) * B; i += gsize) {
    const int64_t t = i - ip.p2;  // i = k B + rho
    const double w = (t == 0) ? ip.width : SampleInTime(t, ip.width, ip.sigma_f);
    const int64_t r = ModN(ip.tf.b * t, n, mask);
    const double d = delta * double(t);
    const double turns = double(r) / double(n) + (d - floor(d));
    taps[i] = Conj(Scale(CisTurns(turns), w));
  }
}

// Sum over the rows whose window [off(row), off(row) + f) holds lag lam of
// taps[k B + rho] fbar[row][m(rho)], k = lam - off(row).  Row 0 has offset
// 0, row 2b+1 offset 2^b, row 2b+2 offset -2^b (b < bits).  tp = taps + rho,
// fm = fbar + m(rho).
__device__ __forceinline__ double2 AdjLagTerms(int32_t f, int32_t bits, int64_t B, int64_t lam,
                                               const double2 *tp, const double2 *fm,
                                               int64_t rstride) {
  double2 acc = make_double2(0.0, 0.0);
  if (lam >= 0) {
    if (lam < f) acc = Add(acc, Cmul(tp[lam * B], fm[0]));
    if (lam >= 1) {  // offsets 2^b in (lam - f, lam]
      for (int b = 63 - __clzll(lam); b >= 0 && (int64_t(1) << b) > lam - f; --b)
        if (b < bits)
          acc = Add(acc, Cmul(tp[(lam - (int64_t(1) << b)) * B], fm[(2 * b + 1) * rstride]));
    }
    for (int b = 0; b < bits && (int64_t(1) << b) < f - lam; ++b)  // offsets -2^b
      acc = Add(acc, Cmul(tp[(lam + (int64_t(1) << b)) * B], fm[(2 * b + 2) * rstride]));
  } else {  // offsets -2^b with 2^b in [-lam, f - lam)
    int b = 63 - __clzll(-lam);
    if ((int64_t(1) << b) < -lam) ++b;
    for (; b < bits && (int64_t(1) << b) < f - lam; ++b)
      acc = Add(acc, Cmul(tp[(lam + (int64_t(1) << b)) * B], fm[(2 * b + 2) * rstride]));
  }
  return acc;
}

// out[j] = sum over the recorded iterations of fold^H fbar at sample j, for
// kGatherJ samples per thread (iteration parameters loaded once for them).
constexpr int kGatherJ = 4;

// AdjLagTerms in 32-bit integers (every lag and offset of an n <= 2^30
// signal fits), with no loop whose length depends on the lag.  With K the
// smallest power with 2^K >= f (at most bits), segment 0 =
// [-2^(K-1), 2^(K-1) + f) holds the windows of row 0 and of the rows with
// offsets +-2^b, b < K, and meets those of offsets +-2^K: a lag there is
// checked against row 0 and offsets +-2^b for b <= K (a loop the whole
// wave runs alike).  A lag outside segment 0 lies in at most one window:
// offset 2^b with b = floor(log2 lam) for lam > 0, offset -2^b with
// b = ceil(log2 -lam) for lam < 0.
__device__ __forceinline__ double2 AdjLagTerms32(int32_t f, int32_t bits, int32_t K, int32_t B,
                                                 int32_t lam, const double2 *tp, const double2 *fm,
                                                 int32_t rstride) {
  double2 acc = make_double2(0.0, 0.0);
  const int32_t s0 = K ? (1 << (K - 1)) : 0;
  if (lam >= -s0 && lam < s0 + f) {
    if (lam >= 0 && lam < f) acc = Cmul(tp[lam * B], fm[0]);
    for (int b = 0; b <= K && b < bits; ++b) {
      const int32_t kp = lam - (1 << b), kn = lam + (1 << b);
      if (kp >= 0 && kp < f) acc = Add(acc, Cmul(tp[kp * B], fm[(2 * b + 1) * rstride]));
      if (kn >= 0 && kn < f) acc = Add(acc, Cmul(tp[kn * B], fm[(2 * b + 2) * rstride]));
    }
    return acc;
  }
  int32_t b, k, row;
  if (lam > 0) {
    b = 31 - __clz(lam);
    k = lam - (1 << b);
    row = 2 * b + 1;
  } else {
    const int32_t nl = -lam;
    b = 31 - __clz(nl);
    if ((1 << b) < nl) ++b;
    k = lam + (b < 31 ? (1 << b) : 0);
    row = 2 * b + 2;
  }
  if (b < bits && k >= 0 && k < f) acc = Cmul(tp[k * B], fm[row * rstride]);
  return acc;
}

// The lags of one recorded iteration, in sample offsets w = lam B + rho
// from the trial's base (lam the lag, rho the residue), for n = 2^k <=
// 2^30.  AdjLagTerms32 is non-zero only in segment 0 (lags [-s0, s0 + f),
// s0 = 2^(K-1)) and in the windows [2^b, 2^b + f) and [-2^b, -2^b + f) of
// the rows b in [K, bits) outside it.  k_adj_lags tabulates AdjLagTerms32
// over exactly those offsets, per (iteration, trial):
//   segment 0, w in [-s0 B, (s0 + f) B):  index w + s0 B
//   row +2^b, w = 2^b B + off, off < f B: index pos0 + b f B + off
//   row -2^b, w = -2^b B + off:          index neg0 + b f B + off
// (E entries in all, E <= (rows_cap + 2) f B since 2 s0 < 2 f and
// 2 bits + 1 <= rows_cap), and the gather looks a sample's offsets up
// there instead of recomputing the lag sums.
struct LagGeom {
  int32_t B, f, bits, K, lb;        // lb = floor(log2 B)
  uint32_t sB, seg_end, fB;         // s0 B, (s0 + f) B, f B
  int32_t pos0, neg0, E;            // table layout
  int32_t top;                      // hi - n: offsets v < top also stand for w = v + n
};

__device__ __forceinline__ LagGeom MakeLagGeom(const IterParams &ip, int64_t n) {
  LagGeom L;
  L.B = ip.B;
  L.f = ip.f;
  L.bits = ip.bits;
  L.K = 0;  // the smallest power with 2^K >= f, at most bits
  while (L.K < L.bits && (1 << L.K) < L.f) ++L.K;
  L.lb = 31 - __clz(L.B);
  const int32_t s0 = L.K ? (1 << (L.K - 1)) : 0;
  L.sB = uint32_t(s0) * uint32_t(L.B);
  L.fB = uint32_t(L.f) * uint32_t(L.B);
  L.seg_end = L.sB + L.fB;
  const int32_t L0 = int32_t(2 * L.sB + L.fB);
  L.pos0 = L0 - L.K * int32_t(L.fB);
  L.neg0 = L0 + (L.bits - 2 * L.K) * int32_t(L.fB);
  L.E = L0 + 2 * (L.bits - L.K) * int32_t(L.fB);
  L.top = int32_t(int64_t((1 << (L.bits - 1)) + L.f) * L.B - n);
  return L;
}

// Table index of offset w = v >= 0 (v < n + f B < 2^31), or -1 outside
// every window.  Outside segment 0 the window is row +2^b with
// 2^b B <= v < 2^(b+1) B (b = floor(log2 lam)), found from the leading bit
// of v.  Selects only, no branches: the lanes of a wavefront take
// different cases.
__device__ __forceinline__ int32_t PosRowIndex(const LagGeom &L, uint32_t v) {
  int32_t b = max(31 - __clz(int32_t(v | 1u)) - L.lb, 0);
  const uint32_t t0 = uint32_t(L.B) << b;
  const int32_t below = v < t0 ? 1 : 0;
  b -= below;  // >= K outside segment 0 (v >= (s0 + f) B > 2^K B)
  const uint32_t off = v - (t0 >> below);
  return ((b < L.bits) & (off < L.fB)) ? L.pos0 + b * int32_t(L.fB) + int32_t(off) : -1;
}
__device__ __forceinline__ int32_t PosLagIndex(const LagGeom &L, uint32_t v) {
  return v < L.seg_end ? int32_t(v + L.sB) : PosRowIndex(L, v);
}

// Table index of offset w = -z (0 < z <= n), or -1.  Outside segment 0
// the window is row -2^b with 2^(b-1) B < z <= 2^b B (b = ceil(log2 -lam)).
__device__ __forceinline__ int32_t NegLagIndex(const LagGeom &L, uint32_t z) {
  const uint32_t zp = z - 1;
  int32_t d = max(31 - __clz(int32_t(zp | 1u)) - L.lb, 0);
  d -= zp < (uint32_t(L.B) << d) ? 1 : 0;  // zp < B: d = -1
  const int32_t b = d + 1;                  // B << b <= 2 zp < 2^31 when b >= 1
  const uint32_t off = (uint32_t(L.B) << min(b, 30)) - z;
  const int32_t row =
      ((b < L.bits) & (off < L.fB)) ? L.neg0 + b * int32_t(L.fB) + int32_t(off) : -1;
  return z <= L.sB ? int32_t(L.sB - z) : row;
}

// The table is E / B slots of B residues: slot sigma holds lag lam (offsets
// w = lam B + rho, rho < B): lags -s0 .. s0 + f - 1 (segment 0), then
// 2^b + k and -2^b + k (k < f) for the rows b = K .. bits - 1.
__device__ __forceinline__ int32_t SlotLag(const LagGeom &L, int32_t sigma) {
  const int32_t s0 = L.K ? (1 << (L.K - 1)) : 0;
  const int32_t n0 = 2 * s0 + L.f;
  if (sigma < n0) return sigma - s0;
  const int32_t r = (sigma - n0) / L.f, k = (sigma - n0) - r * L.f;
  const int32_t rows = L.bits - L.K;
  return r < rows ? (1 << (L.K + r)) + k : k - (1 << (L.K + r - rows));
}

// AdjGatherBody for n = 2^k <= 2^30 with the lag tables of k_adj_lags: per
// sample and recorded (iteration, trial), the offset v = u - base of its
// permuted index u (the low bits of 32-bit products) is looked up as
// w = v - n, w = v and, for v < top, w = v + n: every representative of v
// in the lag range [lo, hi) = [-2^(bits-1) B, (2^(bits-1) + f) B), which
// lies in [-n, n + f B) since 2^(bits-1) B <= n (bits = 1 + floor(log2(n /
// B))), in increasing order.  top <= f B <= (s0 + f) B, so for v < top the
// offset w = v is in segment 0 and only w = v + n needs the general lookup.
template <typename Tout>
__device__ __forceinline__ void AdjGatherBody32(const Ctx &c, const AdjSig &g, const AdjStep *steps,
                                                int32_t count, int32_t si, Tout *out, int64_t tid,
                                                int64_t gsize) {
  const uint32_t n = uint32_t(c.n);
  const uint32_t m32 = n - 1;
  for (int64_t j0 = tid; j0 < int64_t(n); j0 += gsize * kGatherJ) {
    double2 acc[kGatherJ];
#pragma unroll
    for (int q = 0; q < kGatherJ; ++q) acc[q] = make_double2(0.0, 0.0);
    for (int32_t s = 0; s < g.iters; ++s) {
      const AdjStep &st = steps[int64_t(s) * count + si];
      if (!st.active) continue;
      const IterParams &ip = st.ip;
      const LagGeom L = MakeLagGeom(ip, n);
      const uint32_t a_inv = uint32_t(ip.tf.a_inv), cc = uint32_t(ip.tf.c);
      for (int32_t t = 0; t < st.trials; ++t) {
        const uint32_t base = uint32_t(ip.q[t] - ip.p2) & m32;
        const double2 *tab = g.lags + (int64_t(st.it) * c.trials_cap + t) * g.lags_stride;
        // every lookup of the kGatherJ samples is issued before any is
        // added; only lanes with a window load (a miss adds +0, which
        // leaves a sum that started at +0 unchanged)
        int32_t idx[kGatherJ][3];
        double2 val[kGatherJ][3];
#pragma unroll
        for (int q = 0; q < kGatherJ; ++q) {
          const int64_t j = j0 + q * gsize;
          const uint32_t u = (a_inv * ((uint32_t(j) - cc) & m32)) & m32;
          const uint32_t v = (u - base) & m32;
          const bool in = j < int64_t(n), wrap = int32_t(v) < L.top;
          const uint32_t w = wrap ? v + n : v;
          const int32_t p = w < L.seg_end ? int32_t(w + L.sB) : PosRowIndex(L, w);
          idx[q][0] = in ? NegLagIndex(L, n - v) : -1;
          idx[q][1] = !in ? -1 : (wrap ? int32_t(v + L.sB) : p);
          idx[q][2] = (in & wrap) ? p : -1;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            val[q][r] = make_double2(0.0, 0.0);
            if (idx[q][r] >= 0) val[q][r] = LoadX<false>(tab + idx[q][r]);
          }
        }
#pragma unroll
        for (int q = 0; q < kGatherJ; ++q)
#pragma unroll
          for (int r = 0; r < 3; ++r) acc[q] = Add(acc[q], val[q][r]);
      }
    }
#pragma unroll
    for (int q = 0; q < kGatherJ; ++q) {
      const int64_t j = j0 + q * gsize;
      if (j < int64_t(n)) {
        out[j].x = acc[q].x;
        out[j].y = acc[q].y;
      }
    }
  }
}

template <typename Tout>
__device__ __forceinline__ void AdjGatherBody(const Ctx &c, const AdjSig &g, const AdjStep *steps,
                                              int32_t count, int32_t si, Tout *out, int64_t tid,
                                              int64_t gsize) {
  const int64_t n = c.n, mask = c.mask;
  const int64_t rstride = c.B_cap;
  const int64_t tstride = int64_t(c.rows_cap) * c.B_cap;
  const int64_t istride = int64_t(c.trials_cap) * tstride;
  for (int64_t j0 = tid; j0 < n; j0 += gsize * kGatherJ) {
    double2 acc[kGatherJ];
#pragma unroll
    for (int q = 0; q < kGatherJ; ++q) acc[q] = make_double2(0.0, 0.0);
    for (int32_t s = 0; s < g.iters; ++s) {
      const AdjStep &st = steps[int64_t(s) * count + si];
      if (!st.active) continue;
      const IterParams &ip = st.ip;
      const int64_t B = ip.B;
      const int32_t f = ip.f, bits = ip.bits;
      const int32_t h = (ip.B - 1) / 2;
      const double invB = 1.0 / double(B);
      const int64_t half = int64_t(1) << (bits - 1);
      const int64_t lo = -half * B, hi = (half + f) * B;  // lag range times B
      const double2 *taps = g.taps + int64_t(st.it) * g.taps_stride;
      const int64_t a_inv = ip.tf.a_inv, cc = ip.tf.c;
      for (int32_t t = 0; t < st.trials; ++t) {
        const int64_t base = ModN(ip.q[t] - ip.p2, n, mask);
        const double2 *fb = g.fbar + int64_t(st.it) * istride + t * tstride;
#pragma unroll
        for (int q = 0; q < kGatherJ; ++q) {
          const int64_t j = j0 + q * gsize;
          if (j >= n) continue;
          const int64_t u = ModN(a_inv * ModN(j - cc, n, mask), n, mask);
          const int64_t v = ModN(u - base, n, mask);
          for (int64_t w = v >= hi ? v - n : v - n + (v - n < lo ? n : 0); w < hi; w += n) {
            // lam B + rho = w, w = v mod n, lo <= w < hi
            int64_t lam = int64_t(floor(double(w) * invB));
            int64_t rho = w - lam * B;
            if (rho < 0) {
              --lam;
              rho += B;
            } else if (rho >= B) {
              ++lam;
              rho -= B;
            }
            const int64_t m = rho >= h ? rho - h : rho - h + B;
            acc[q] = Add(acc[q], AdjLagTerms(f, bits, B, lam, taps + rho, fb + m, rstride));
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kGatherJ; ++q) {
      const int64_t j = j0 + q * gsize;
      if (j < n) {
        out[j].x = acc[q].x;
        out[j].y = acc[q].y;
      }
    }
  }
}

// ---- Kernels.  Every signal of a batch (or the one signal of a solver)
// at once: map[blockIdx.x] = {signal, local block}, or with map null signal
// 0 and the block index itself.  sig[i] holds signal i's buffers, steps
// (count entries per reverse step) its iteration parameters.

__device__ __forceinline__ int2 AdjBlock(const int2 *map) {
  return map ? map[blockIdx.x] : make_int2(0, int(blockIdx.x));
}

// Per signal: the IterParams of every reverse step, as NextIteration derives
// them in a replay (bins from the log, the transform and q's drawn).
__global__ void __launch_bounds__(kThreads)
    k_adj_steps(const Ctx *ctxs, const AdjSig *sig, int32_t count, int32_t max_steps,
                AdjStep *steps) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(count) * max_steps) return;
  const int32_t j = int32_t(i / count), s = int32_t(i - int64_t(j) * count);
  const Ctx &c = ctxs[s];
  AdjStep &st = steps[i];
  const int32_t it = sig[s].iters - 1 - j;
  st.active = it >= 0 ? 1 : 0;
  st.trials = sig[s].trials;
  st.count_before = 0;
  st.it = it;
  if (it < 0) return;
  const int32_t B = c.log[it];
  FillParams(&st.ip, c.btab[B], c.btab[B].bits);
  st.ip.bin_thr = 0;
  st.ip.sv_thr = 0;
  st.ip.tf = c.tfs[it];
  for (int t = 0; t < kMaxTrials; ++t)
    st.ip.q[t] = t < st.trials ? c.qs[int64_t(it) * c.trials_cap + t] : 0;
  st.ip.it = it;
  st.ip.done = 0;
  st.count_before = it > 0 ? c.log[c.log_cap + it - 1] : 0;
  // An iteration whose bins identified nothing has Abar = 0: it changes
  // neither lam nor xbar, and every stage skips it.
  const int64_t *rec_row = sig[s].rec + int64_t(it) * c.B_cap;
  int32_t live = 0;
  for (int32_t b = 0; b < B && !live; ++b) live = rec_row[b] >= 0;
  st.active = live;
}

// lam = 0 at every found key (the only entries read), `per` blocks a signal.
__global__ void __launch_bounds__(kThreads)
    k_adj_clear(const Ctx *ctxs, const AdjSig *sig, int32_t per) {
  const int32_t si = blockIdx.x / per;
  const Ctx &c = ctxs[si];
  const int32_t cnt = min(c.st->count, c.cap);
  double2 *lam = sig[si].lam;
  for (int32_t i = (blockIdx.x % per) * blockDim.x + threadIdx.x; i < cnt; i += per * blockDim.x)
    lam[c.keys[i]] = make_double2(0.0, 0.0);
}

// lam[keys[i]] += g[i]; keys outside [0, n) are not frequencies of the
// signal and carry nothing.
__global__ void __launch_bounds__(kThreads) k_adj_seed(const AdjSig *sig, int32_t per) {
  const AdjSig &g = sig[blockIdx.x / per];
  for (int64_t i = int64_t(blockIdx.x % per) * blockDim.x + threadIdx.x; i < g.m;
       i += int64_t(per) * blockDim.x) {
    const int64_t k = g.keys[i];
    if (k < 0 || k >= g.n) continue;
    atomicAdd(&g.lam[k].x, g.grad[i].x);
    atomicAdd(&g.lam[k].y, g.grad[i].y);
  }
}

__global__ void __launch_bounds__(kThreads)
    k_adj_bins(const Ctx *ctxs, const AdjSig *sig, const AdjStep *steps, const int2 *map,
               const int32_t *nblk) {
  const int2 m = AdjBlock(map);
  const AdjStep &st = steps[m.x];
  if (!st.active) return;
  const Ctx &c = ctxs[m.x];
  const AdjSig &g = sig[m.x];
  const int nb = map ? nblk[m.x] : int(gridDim.x);
  AdjBinsBody(c, st.ip, st.trials, g.lam, g.rec + int64_t(st.it) * c.B_cap, m.y, nb);
  AdjTapsBody(st.ip, c.n, c.mask, g.taps + int64_t(st.it) * g.taps_stride,
              int64_t(m.y) * blockDim.x + threadIdx.x, int64_t(nb) * blockDim.x);
}

// Blocks below kPrepBlocks of each signal update lam (the terms), the
// others transform one (trial, row) each.
template <int kPaths>
__global__ void __launch_bounds__(kFftThreads)
    k_adj_fft_terms(const Ctx *ctxs, const AdjSig *sig, const AdjStep *steps, const int2 *map) {
  HIP_DYNAMIC_SHARED(double2, lds)
  const int2 m = AdjBlock(map);
  const AdjStep &st = steps[m.x];
  if (!st.active) return;
  const Ctx &c = ctxs[m.x];
  const AdjSig &g = sig[m.x];
  if (m.y < kPrepBlocks) {
    AdjTermsBody(c, st.ip, st.trials, g.lam, st.count_before, g.rec + int64_t(st.it) * c.B_cap,
                 int64_t(m.y) * blockDim.x + threadIdx.x, int64_t(kPrepBlocks) * blockDim.x);
    return;
  }
  const int64_t istride = int64_t(c.trials_cap) * c.rows_cap * c.B_cap;
  AdjFftBody<kPaths>(c, st.ip, st.trials, g.fbar + int64_t(st.it) * istride, m.y - kPrepBlocks,
                     lds);
}

// The lag tables of every (signal, step, trial) with a table (AdjSig.lags):
// `per` blocks per (step, signal), in the steps array's order.
__global__ void __launch_bounds__(kThreads)
    k_adj_lags(const Ctx *ctxs, const AdjSig *sig, const AdjStep *steps, int32_t count,
               int32_t max_steps, int32_t per) {
  const int64_t unit = blockIdx.x / per;
  if (unit >= int64_t(count) * max_steps) return;
  const AdjStep &st = steps[unit];
  const int32_t si = int32_t(unit % count);
  const AdjSig &g = sig[si];
  if (!st.active || !g.lags) return;
  const Ctx &c = ctxs[si];
  const LagGeom L = MakeLagGeom(st.ip, c.n);
  if (L.E > g.lags_stride) return;  // cannot happen: E <= (rows_cap + 2) f B
  const int64_t tstride = int64_t(c.rows_cap) * c.B_cap;
  const int64_t istride = int64_t(c.trials_cap) * tstride;
  const double2 *taps = g.taps + int64_t(st.it) * g.taps_stride;
  const int32_t B = L.B, h = (B - 1) / 2, slots = L.E / B;
  for (int32_t t = 0; t < st.trials; ++t) {
    const double2 *fb = g.fbar + int64_t(st.it) * istride + t * tstride;
    double2 *tab = g.lags + (int64_t(st.it) * c.trials_cap + t) * g.lags_stride;
    // one slot per wavefront at a time, its residues over the lanes
    const int32_t waves = int32_t(blockDim.x >> 6), lane = int32_t(threadIdx.x & 63);
    for (int32_t sigma = int32_t(blockIdx.x % per) * waves + int32_t(threadIdx.x >> 6);
         sigma < slots; sigma += per * waves) {
      const int32_t lam = SlotLag(L, sigma);  // the same for the whole wavefront
      for (int32_t rho = lane; rho < B; rho += 64) {
        const int32_t m = rho >= h ? rho - h : rho - h + B;
        tab[int64_t(sigma) * B + rho] =
            AdjLagTerms32(L.f, L.bits, L.K, B, lam, taps + rho, fb + m, c.B_cap);
      }
    }
  }
}

// The gather of every signal with lag tables (k_adj_gather_tab) and of
// every other signal (k_adj_gather): two kernels, so that the table path's
// register use is its own.
template <typename Tout>
__global__ void __launch_bounds__(kThreads)
    k_adj_gather_tab(const Ctx *ctxs, const AdjSig *sig, const AdjStep *steps, int32_t count,
                     int32_t per) {
  const int32_t si = blockIdx.x / per;
  const AdjSig &g = sig[si];
  if (!g.lags) return;
  const int64_t tid = int64_t(blockIdx.x % per) * blockDim.x + threadIdx.x;
  AdjGatherBody32(ctxs[si], g, steps, count, si, static_cast<Tout *>(g.out), tid,
                  int64_t(per) * blockDim.x);
}

template <typename Tout>
__global__ void __launch_bounds__(kThreads)
    k_adj_gather(const Ctx *ctxs, const AdjSig *sig, const AdjStep *steps, int32_t count,
                 int32_t per) {
  const int32_t si = blockIdx.x / per;
  const AdjSig &g = sig[si];
  if (g.lags) return;
  const int64_t tid = int64_t(blockIdx.x % per) * blockDim.x + threadIdx.x;
  AdjGatherBody(ctxs[si], g, steps, count, si, static_cast<Tout *>(g.out), tid,
                int64_t(per) * blockDim.x);
}

// One workgroup per descriptor of a staging area (TfScatterDesc).
__global__ void __launch_bounds__(kThreads) k_tf_scatter(const unsigned char *base) {
  const TfScatterDesc d = reinterpret_cast<const TfScatterDesc *>(base)[blockIdx.x];
  for (int r = 0; r < 2; ++r) {
    unsigned long long *dst = static_cast<unsigned long long *>(d.dst[r]);
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(base + d.off[r]);
    for (int64_t i = threadIdx.x; i < d.words[r]; i += blockDim.x) dst[i] = src[i];
  }
}

int32_t Blocks(int64_t work) {
  const int64_t b = (work + kThreads - 1) / kThreads;
  return int32_t(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

}  // namespace

hipError_t LaunchTfScatter(const unsigned char *staging, int32_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tf_scatter, dim3(count), dim3(kThreads), 0, s, staging);
  return hipGetLastError();
}

hipError_t LaunchBuildTwiddles(double2 *tw, hipStream_t s) {
  hipLaunchKernelGGL(k_build_twiddles, dim3(Blocks(kTwM / 2)), dim3(kThreads), 0, s, tw);
  return hipGetLastError();
}

hipError_t LaunchBuildTables(const int32_t *blist, int32_t nb, const BParams *btab,
                             double2 *table, const double2 *tw, int32_t M_cap,
                             hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_build_tables, dim3(nb), dim3(kThreads), size_t(M_cap) * sizeof(double2), s, blist, btab, table, tw);
  return hipGetLastError();
}

hipError_t LaunchClearModes(const Ctx &c, int32_t full, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_modes, dim3(1024), dim3(kThreads), 0, s, c, full);
  return hipGetLastError();
}

hipError_t LaunchReset(const Ctx &c, const InitArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(kThreads), 0, s, c, a);
  return hipGetLastError();
}

hipError_t LaunchInit(const Ctx &c, const InitArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, s, c, a);
  return hipGetLastError();
}

hipError_t LaunchBinTime(const Ctx &c, int parity, int32_t tiles, size_t lds, hipStream_t s) {
  if (c.mixed)
    hipLaunchKernelGGL(k_bin_time<float2>, dim3(kPrepBlocks + tiles), dim3(kThreads), lds, s, c, parity);
  else
    hipLaunchKernelGGL(k_bin_time<double2>, dim3(kPrepBlocks + tiles), dim3(kThreads), lds, s, c, parity);
  return hipGetLastError();
}

hipError_t LaunchBinFft(const Ctx &c, int parity, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_fft, dim3(c.trials_cap * c.rows_cap), dim3(kFftThreads), lds, s, c, parity);
  return hipGetLastError();
}

hipError_t LaunchBigFft(const Ctx &c, int parity, hipStream_t s) {
  if (c.big_gpr <= 0) return hipSuccess;
  const int32_t rows = c.trials_cap * c.rows_cap;
  hipLaunchKernelGGL(k_big_cols, dim3((rows + 1) * c.big_gpr), dim3(kThreads), 0, s, c, parity);
  hipLaunchKernelGGL(k_big_rows, dim3(rows * c.big_gpr), dim3(kThreads), 0, s, c, parity);
  hipLaunchKernelGGL(k_big_out, dim3(rows * c.big_gpr), dim3(kThreads), 0, s, c, parity);
  return hipGetLastError();
}

hipError_t LaunchLateIteration(const Ctx &c, int parity, int32_t nb_ident, int32_t tiles,
                               size_t lds_time, size_t lds_fft, int with_identify, int small_fft,
                               hipStream_t s) {
  const dim3 grid(nb_ident + kPrepBlocks + tiles);
  const bool one = ColSlicesFor(int64_t(c.trials_cap) * c.rows_cap) == 1;
  if (c.mixed && one)
    hipLaunchKernelGGL((k_late_fold<float2, 1>), grid, dim3(kThreads), lds_time, s, c, parity,
                       nb_ident, with_identify);
  else if (c.mixed)
    hipLaunchKernelGGL((k_late_fold<float2, kColSlices>), grid, dim3(kThreads), lds_time, s, c,
                       parity, nb_ident, with_identify);
  else if (one)
    hipLaunchKernelGGL((k_late_fold<double2, 1>), grid, dim3(kThreads), lds_time, s, c, parity,
                       nb_ident, with_identify);
  else
    hipLaunchKernelGGL((k_late_fold<double2, kColSlices>), grid, dim3(kThreads), lds_time, s, c,
                       parity, nb_ident, with_identify);
  // the late B's own bin-FFT path: the small one alone where it applies
  if (small_fft)
    hipLaunchKernelGGL(k_late_fft<kFftSmallOnly>, dim3(c.trials_cap * c.rows_cap),
                       dim3(kFftThreads), lds_fft, s, c, parity);
  else
    hipLaunchKernelGGL(k_late_fft<kFftBoth>, dim3(c.trials_cap * c.rows_cap), dim3(kFftThreads),
                       lds_fft, s, c, parity);
  return hipGetLastError();
}

hipError_t LaunchPublish(const Ctx &c, int parity, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_time<double2>, dim3(1), dim3(kThreads), 0, s, c, parity);
  return hipGetLastError();
}

hipError_t LaunchIdentify(const Ctx &c, int parity, int32_t blocks, hipStream_t s) {
  const bool one = ColSlicesFor(int64_t(c.trials_cap) * c.rows_cap) == 1;
  if (c.idout) {
    if (one)
      hipLaunchKernelGGL((k_identify<1, true>), dim3(blocks), dim3(kThreads), 0, s, c, parity);
    else
      hipLaunchKernelGGL((k_identify<kColSlices, true>), dim3(blocks), dim3(kThreads), 0, s, c,
                         parity);
  } else {
    if (one)
      hipLaunchKernelGGL((k_identify<1, false>), dim3(blocks), dim3(kThreads), 0, s, c, parity);
    else
      hipLaunchKernelGGL((k_identify<kColSlices, false>), dim3(blocks), dim3(kThreads), 0, s, c,
                         parity);
  }
  return hipGetLastError();
}

hipError_t LaunchSetModes(const Ctx &c, const int64_t *keys, const double2 *coefs,
                          int64_t m, hipStream_t s) {
  hipLaunchKernelGGL(k_set_modes, dim3(Blocks(m)), dim3(kThreads), 0, s, c, keys, coefs, m);
  return hipGetLastError();
}

hipError_t LaunchGetCoefs(const Ctx &c, double2 *out, int64_t m, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_get_coefs, dim3(Blocks(m)), dim3(kThreads), 0, s, c, out, m);
  return hipGetLastError();
}

hipError_t LaunchBinInFreqDirect(int64_t n, int64_t mask, const BParams &bp,
                                 const DevTf &tf, int64_t q, const int64_t *k0,
                                 const double2 *coef, int64_t m, double2 *out,
                                 unsigned long long *limbs, int32_t *emax, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  const int64_t bins = int64_t(1 + 2 * bp.bits) * bp.B;
  hipLaunchKernelGGL(k_freq_direct_clear, dim3(Blocks(kLimbs * bins)), dim3(kThreads), 0, s, limbs,
                     kLimbs * bins, emax);
  hipLaunchKernelGGL(k_freq_direct_max, dim3(Blocks(m)), dim3(kThreads), 0, s, n, mask, bp, tf, q,
                     k0, coef, m, emax);
  hipLaunchKernelGGL(k_freq_direct_add, dim3(Blocks(m)), dim3(kThreads), 0, s, n, mask, bp, tf, q,
                     k0, coef, m, emax, limbs);
  hipLaunchKernelGGL(k_freq_direct_sub, dim3(Blocks(bins)), dim3(kThreads), 0, s, limbs, emax, bins,
                     out);
  return hipGetLastError();
}

hipError_t LaunchMatPencil(const double2 *u, int64_t m, double *sigma, uint8_t *vote,
                           hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mat_pencil, dim3(Blocks(m)), dim3(kThreads), 0, s, u, m, sigma, vote);
  return hipGetLastError();
}

hipError_t LaunchBatchReset(const Ctx *ctxs, const InitArgs *args, int32_t count,
                            int32_t *counters, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_reset, dim3(count), dim3(kThreads), 0, s, ctxs, args, counters);
  return hipGetLastError();
}

hipError_t LaunchBatchStage(const BatchLaunch &b, int parity, int stage, hipStream_t s) {
  const BatchWork &w = b.work;
  switch (stage) {
    case 0:
      hipLaunchKernelGGL(k_batch_signal, dim3(b.count), dim3(64), 0, s, b.ctxs, b.count, parity, w,
                         b.ident_bins);
      hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(kPlanThreads), 0, s, b.ctxs, b.count, w,
                         b.counters, b.host);
      break;
    case 1:
      if (b.mixed)
        hipLaunchKernelGGL(k_bin_time_batch<float2>, dim3(b.grid_time), dim3(kThreads),
                           b.lds_time, s, b.ctxs, w.map[0], w.total, parity);
      else
        hipLaunchKernelGGL(k_bin_time_batch<double2>, dim3(b.grid_time), dim3(kThreads),
                           b.lds_time, s, b.ctxs, w.map[0], w.total, parity);
      break;
    case 2:  // one kernel per bin-FFT path, each over its own units
      hipLaunchKernelGGL(k_bin_fft_batch<kFftSmallOnly>, dim3(b.grid_fft), dim3(kFftThreads),
                         b.lds_fft_small, s, b.ctxs, w.map[1], w.total, parity);
      if (b.large_fft)
        hipLaunchKernelGGL(k_bin_fft_batch<kFftLargeOnly>, dim3(b.grid_fft), dim3(kFftThreads),
                           b.lds_fft, s, b.ctxs, w.map[3], w.total, parity);
      break;
    default:
      if (b.col_slices == 1)
        hipLaunchKernelGGL(k_identify_batch<1>, dim3(b.grid_ident), dim3(kThreads), 0, s,
                           b.ctxs, w.map[2], w.total, w.cnt + 2 * b.count, parity);
      else
        hipLaunchKernelGGL(k_identify_batch<kColSlices>, dim3(b.grid_ident), dim3(kThreads), 0,
                           s, b.ctxs, w.map[2], w.total, w.cnt + 2 * b.count, parity);
      break;
  }
  return hipGetLastError();
}

hipError_t LaunchBatchIteration(const BatchLaunch &b, int parity, hipStream_t s) {
  for (int stage = 0; stage < kBatchStages; ++stage) {
    const hipError_t e = LaunchBatchStage(b, parity, stage, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Blocks per signal of k_batch_modes_out.
constexpr int32_t kModesOutBlocks = 8;

__global__ void __launch_bounds__(kThreads)
    k_batch_modes_out(const Ctx *ctxs, const ModesOut *out) {
  const int32_t si = blockIdx.x / kModesOutBlocks;
  const Ctx &c = ctxs[si];
  const ModesOut o = out[si];
  const int32_t cnt = int32_t(min<int64_t>(min(c.st->count, c.cap), o.cap));
  for (int32_t i = (blockIdx.x % kModesOutBlocks) * blockDim.x + threadIdx.x; i < cnt;
       i += kModesOutBlocks * blockDim.x) {
    const int64_t k = c.keys[i];
    o.keys[i] = k;
    o.coefs[i] = c.coef[k];
  }
}

hipError_t LaunchBatchModesOut(const Ctx *ctxs, const ModesOut *out, int32_t count,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_batch_modes_out, dim3(count * kModesOutBlocks), dim3(kThreads), 0, s, ctxs,
                     out);
  return hipGetLastError();
}

hipError_t LaunchBatchResults(const Ctx *ctxs, int32_t count, BatchResultDev *out,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_batch_results, dim3(Blocks(count)), dim3(kThreads), 0, s, ctxs, count, out);
  return hipGetLastError();
}

hipError_t LaunchAdj(const AdjLaunch &b, hipStream_t s) {
  const int64_t nsteps = int64_t(b.count) * b.max_steps;
  hipLaunchKernelGGL(k_adj_steps, dim3(Blocks(nsteps)), dim3(kThreads), 0, s, b.ctxs, b.sig,
                     b.count, b.max_steps, b.steps);
  hipLaunchKernelGGL(k_adj_clear, dim3(b.count * b.per_small), dim3(kThreads), 0, s, b.ctxs, b.sig,
                     b.per_small);
  hipLaunchKernelGGL(k_adj_seed, dim3(b.count * b.per_small), dim3(kThreads), 0, s, b.sig,
                     b.per_small);
  for (int32_t j = 0; j < b.max_steps; ++j) {
    const AdjStep *st = b.steps + int64_t(j) * b.count;
    hipLaunchKernelGGL(k_adj_bins, dim3(b.nb_bins), dim3(kThreads), 0, s, b.ctxs, b.sig, st,
                       b.map_bins, b.nblk_bins);
    if (b.fft_small_only)
      hipLaunchKernelGGL(k_adj_fft_terms<kFftSmallOnly>, dim3(b.nb_fft), dim3(kFftThreads),
                         b.lds_fft, s, b.ctxs, b.sig, st, b.map_fft);
    else
      hipLaunchKernelGGL(k_adj_fft_terms<kFftBoth>, dim3(b.nb_fft), dim3(kFftThreads), b.lds_fft,
                         s, b.ctxs, b.sig, st, b.map_fft);
  }
  if (b.any_tab)
    hipLaunchKernelGGL(k_adj_lags, dim3(b.count * b.max_steps * b.per_lags), dim3(kThreads), 0,
                       s, b.ctxs, b.sig, b.steps, b.count, b.max_steps, b.per_lags);
  const dim3 grid(b.count * b.per_gather);
  if (b.any_tab && b.mixed)
    hipLaunchKernelGGL(k_adj_gather_tab<float2>, grid, dim3(kThreads), 0, s, b.ctxs, b.sig,
                       b.steps, b.count, b.per_gather);
  if (b.any_tab && !b.mixed)
    hipLaunchKernelGGL(k_adj_gather_tab<double2>, grid, dim3(kThreads), 0, s, b.ctxs, b.sig,
                       b.steps, b.count, b.per_gather);
  if (b.any_gen && b.mixed)
    hipLaunchKernelGGL(k_adj_gather<float2>, grid, dim3(kThreads), 0, s, b.ctxs, b.sig, b.steps,
                       b.count, b.per_gather);
  if (b.any_gen && !b.mixed)
    hipLaunchKernelGGL(k_adj_gather<double2>, grid, dim3(kThreads), 0, s, b.ctxs, b.sig, b.steps,
                       b.count, b.per_gather);
  return hipGetLastError();
}

}  // namespace mpsft
