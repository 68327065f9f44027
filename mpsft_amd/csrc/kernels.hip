// kernels.hip -- HIP kernels of the MPSFT recovery loop for gfx950 (CDNA4).
//
// Layout in HBM (all row-major, interleaved complex):
//   x      : n samples, double2 (FP64) or float2 (MIXED), read by gathers
//   fold   : [trial][row][m]  folded, windowed samples, m < B (double2)
//   A      : [trial][row][b]  bin coefficients after BinInFreq (double2)
//   coef   : n double2, dense coefficients of the found modes
//   present: n bits, membership of the found-mode map
//   keys   : found-mode locations in insertion order
// Nothing here is a dense contraction, so nothing uses MFMA: the gather is
// bound by random-access memory traffic, the bin FFT by LDS latency and the
// identify/estimate by transcendental latency.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mpsft {
namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ int64_t ModN(int64_t v, int64_t n, int64_t mask) {
  if (mask >= 0) return v & mask;
  const int64_t r = v % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ double2 Cmul(double2 a, double2 b) {
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

__device__ __forceinline__ double LoadRe(const double2 *x, int64_t j, double *im) {
  const double2 v = x[j];
  *im = v.y;
  return v.x;
}

template <typename Tx>
__device__ __forceinline__ double2 LoadX(const Tx *x, int64_t j);
template <>
__device__ __forceinline__ double2 LoadX<double2>(const double2 *x, int64_t j) {
  return x[j];
}
template <>
__device__ __forceinline__ double2 LoadX<float2>(const float2 *x, int64_t j) {
  const float2 v = x[j];
  return make_double2(double(v.x), double(v.y));
}

// SincPi (src/sincos.cc:69-88): sin(x)/x with Taylor terms near 0.
__device__ double SincPiRef(double x) {
  const double taylor_0_bound = 2.220446049250313080847e-16;
  const double taylor_2_bound = 1.490116119384765625e-08;  // sqrt(eps)
  const double taylor_n_bound = 1.220703125e-04;           // eps^(1/4)
  if (x < 0) x = -x;
  if (x >= taylor_n_bound) return sin(x) / x;
  double r = 1.0;
  if (x >= taylor_0_bound) {
    const double x2 = x * x;
    r -= x2 / 6.0;
    if (x >= taylor_2_bound) r += (x2 * x2) / 120.0;
  }
  return r;
}

// Window::SampleInTime (src/window.cc:71-75).
__device__ __forceinline__ double SampleInTime(int64_t i, double width, double sigma_f) {
  const double t = double(i);
  const double u = t * kPi * sigma_f;
  return width * SincPiRef(t * kPi * width) * exp(-2.0 * u * u);
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

// In-place radix-2 FFTs over M = 2^logM points in LDS.  tw[k] =
// e^{-2 pi i k / 2^kMaxLogM}.  Dit: bit-reversed in, natural out.
// Dif: natural in, bit-reversed out.  inverse flips the sign.
__device__ void FftDit(double2 *a, int logM, bool inverse, const double2 *tw) {
  const int M = 1 << logM;
  for (int s = 1; s <= logM; ++s) {
    const int half = 1 << (s - 1);
    const int tstride = 1 << (kMaxLogM - s);
    for (int idx = threadIdx.x; idx < M / 2; idx += blockDim.x) {
      const int k = idx & (half - 1);
      const int i = ((idx >> (s - 1)) << s) + k;
      const int j = i + half;
      double2 w = tw[k * tstride];
      if (inverse) w.y = -w.y;
      const double2 u = a[i];
      const double2 v = Cmul(a[j], w);
      a[i] = Add(u, v);
      a[j] = Sub(u, v);
    }
    __syncthreads();
  }
}

__device__ void FftDif(double2 *a, int logM, bool inverse, const double2 *tw) {
  const int M = 1 << logM;
  for (int s = logM; s >= 1; --s) {
    const int half = 1 << (s - 1);
    const int tstride = 1 << (kMaxLogM - s);
    for (int idx = threadIdx.x; idx < M / 2; idx += blockDim.x) {
      const int k = idx & (half - 1);
      const int i = ((idx >> (s - 1)) << s) + k;
      const int j = i + half;
      double2 w = tw[k * tstride];
      if (inverse) w.y = -w.y;
      const double2 u = a[i];
      const double2 v = a[j];
      a[i] = Add(u, v);
      a[j] = Cmul(Sub(u, v), w);
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
  const int half = 1 << (kMaxLogM - 1);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < half; k += gridDim.x * blockDim.x) {
    double s, c;
    sincospi(-2.0 * double(k) / double(1 << kMaxLogM), &s, &c);
    tw[k] = make_double2(c, s);
  }
}

// One workgroup per B: chirp c_m (m < B) and the spectrum of the Bluestein
// filter h (h_j = conj(c_j), h_{M-j} = conj(c_j)), pre-scaled by 1/M.
__global__ void __launch_bounds__(kThreads)
    k_build_tables(const int32_t *blist, const BParams *btab, double2 *table,
                   const double2 *tw) {
  extern __shared__ double2 lds[];
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
  FftDit(lds, bp.logM, false, tw);
  const double inv = 1.0 / double(M);
  for (int k = threadIdx.x; k < M; k += blockDim.x) filt[k] = Scale(lds[k], inv);
}

// Clears the found-mode map: only the listed keys (their coefficient and
// presence word), or everything after an overflow.
__global__ void __launch_bounds__(kThreads) k_clear_modes(Ctx c, int32_t full) {
  const DevState *st = c.st;
  const int64_t gsize = int64_t(gridDim.x) * blockDim.x;
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
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

__device__ void LoadBParams(DevState *st, const BParams &bp, int32_t bits_override) {
  st->B = bp.B;
  st->bits = bits_override > 0 ? bits_override : bp.bits;
  st->rows = 1 + 2 * st->bits;
  st->f = bp.f;
  st->P = bp.P;
  st->p2 = (bp.P - 1) / 2;
  st->M = bp.M;
  st->logM = bp.logM;
  st->tab_off = bp.off;
  st->width = bp.width;
  st->sigma_f = bp.sigma_f;
  st->dfreq = bp.dfreq;
}

__global__ void k_init(Ctx c, InitArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DevState *st = c.st;
  st->x = a.x;
  st->mode = a.mode;
  st->trials = a.trials;
  st->k = a.k;
  st->min_bins = a.min_bins;
  st->max_stale = a.max_stale;
  st->max_iter = a.max_iter;
  st->it = 0;
  st->stale = 0;
  st->done = 0;
  if (!a.keep_modes) {
    st->count = 0;
    st->flags = 0;
  }
  st->found_iter = 0;
  st->found_last = 0;
  st->blocks_done = 0u;
  const BParams bp = c.btab[a.B];
  LoadBParams(st, bp, a.bits_override);
  if (a.use_table_thr) {
    st->bin_thr = bp.bin_thr;
    st->sv_thr = bp.sv_thr;
  } else {
    st->bin_thr = a.bin_thr;
    st->sv_thr = a.sv_thr;
  }
  st->win_thr = a.win_thr;
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
// k_bin_time: the permuted, windowed gather-fold of BinInTime
// (src/binner.cc:81-111 semantics).  With s = t + tau - (q - p2) the row
// offset, every row's samples sit at u = q - p2 + lambda*B + rho, and the
// row with offset off (in units of B) folds
//   fold[row][m(rho)] = sum_{k<f} x[(a u + c) mod n] * tap(k*B + rho - p2),
//   u = q - p2 + (off + k)*B + rho,  m(rho) = (rho - (B-1)/2) mod B,
// where tap(t) = wt(t) e^{2 pi i [(b t mod n)/n + frac(-t/(2B))]}; the row's
// own phase e^{2 pi i b tau/n} is applied after the DFT.  Rows with
// 2^k < f overlap row 0; they form segment 0 and share one staged copy of
// each sample (lambda in [-2^(K-1), 2^(K-1)+f)).  Every other row is its own
// segment.  A tile = (trial, segment, block of kResBlock residues rho):
//   phase A gathers the segment's samples for the block into LDS,
//   phase T computes the block's f taps into LDS,
//   phase B sums per (row, rho) and writes the fold.
// Workgroups below kPrepBlocks instead compute the BinInFreq mode terms.
template <typename Tx>
__global__ void __launch_bounds__(kThreads) k_bin_time(Ctx c) {
  extern __shared__ double2 lds[];
  const DevState *st = c.st;
  if (st->done) return;
  const int32_t it = st->it;
  const int32_t trials = st->trials;
  const int32_t B = st->B, f = st->f, bits = st->bits, p2 = st->p2;
  const DevTf tf = c.tfs[it];
  const int64_t n = c.n, mask = c.mask;

  if (blockIdx.x < kPrepBlocks) {
    // BinInFreqV1 terms (src/binner.cc:565-591) for every found mode.
    const int32_t cnt = min(st->count, c.cap);
    const double width = st->width, dfreq = st->dfreq;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < int64_t(cnt) * trials;
         i += int64_t(kPrepBlocks) * blockDim.x) {
      const int32_t trial = int32_t(i / cnt);
      const int32_t mi = int32_t(i - int64_t(trial) * cnt);
      const int64_t k0 = c.keys[mi];
      const double2 cf = c.coef[k0];
      const int64_t q = c.qs[int64_t(it) * c.trials_cap + trial];
      const int64_t k1 = ModN(tf.a * k0 + tf.b, n, mask);
      const int32_t bin = int32_t(k1 * B / n);
      const double xi = (double(bin) + 0.5) / double(B) - double(k1) / double(n);
      const double wf = SampleInFreq(xi, width, dfreq);
      const double2 ph = CisFrac(ModN(q * k1 + tf.c * k0, n, mask), n);
      ModeTerm t;
      t.k1 = k1;
      t.bin = bin;
      t.pad = 0;
      t.base = Cmul(Scale(cf, wf), ph);
      c.terms[int64_t(trial) * c.cap + mi] = t;
    }
    return;
  }

  int32_t tile = blockIdx.x - kPrepBlocks;
  int K = 0;
  while (K < bits && (1 << K) < f) ++K;
  const int32_t nseg = 1 + 2 * (bits - K);
  const int32_t nblk = (B + kResBlock - 1) / kResBlock;
  const int32_t per_trial = nseg * nblk;
  if (tile >= trials * per_trial) return;
  const int32_t trial = tile / per_trial;
  tile -= trial * per_trial;
  const int32_t seg = tile / nblk;
  const int32_t blk = tile - seg * nblk;

  int32_t lam_lo, len, nrows, row1;
  if (seg == 0) {
    lam_lo = K ? -(1 << (K - 1)) : 0;
    len = (K ? (1 << K) : 0) + f;
    nrows = 1 + 2 * K;
    row1 = 0;
  } else {
    const int32_t kk = K + (seg - 1) / 2;
    const bool neg = ((seg - 1) & 1) != 0;
    lam_lo = neg ? -(1 << kk) : (1 << kk);
    len = f;
    nrows = 1;
    row1 = neg ? 2 * kk + 2 : 2 * kk + 1;
  }
  const int32_t rho0 = blk * kResBlock;
  const int32_t nrho = min(kResBlock, B - rho0);
  const int64_t q = c.qs[int64_t(it) * c.trials_cap + trial];
  const Tx *x = static_cast<const Tx *>(st->x);

  double2 *X = lds;                                   // [len_cap][kResBlock]
  double2 *T = lds + c.len_cap * kResBlock;           // [f_cap][kResBlock]
  double2 *red = T + c.f_cap * kResBlock;             // [kThreads]

  // Phase A: gather.
  const int64_t ubase = q - p2 + int64_t(lam_lo) * B + rho0;
  for (int idx = threadIdx.x; idx < len * nrho; idx += blockDim.x) {
    const int li = idx / nrho;
    const int ri = idx - li * nrho;
    const int64_t u = ubase + int64_t(li) * B + ri;
    const int64_t j = ModN(tf.a * u + tf.c, n, mask);
    X[li * kResBlock + ri] = LoadX<Tx>(x, j);
  }
  // Phase T: taps tap(t), t = k*B + rho - p2.
  {
    const double width = st->width, sigma_f = st->sigma_f;
    const double delta = -0.5 / double(B);
    for (int idx = threadIdx.x; idx < f * nrho; idx += blockDim.x) {
      const int k = idx / nrho;
      const int ri = idx - k * nrho;
      const int64_t t = int64_t(k) * B + rho0 + ri - p2;
      const double w = (t == 0) ? width : SampleInTime(t, width, sigma_f);
      const int64_t r = ModN(tf.b * t, n, mask);
      const double d = delta * double(t);
      const double turns = double(r) / double(n) + (d - floor(d));
      T[k * kResBlock + ri] = Scale(CisTurns(turns), w);
    }
  }
  __syncthreads();

  // Phase B: per (row, rho) sums over k, split in `parts` when few pairs.
  const int32_t pairs = nrows * nrho;
  int32_t parts = 1;
  while (parts * 2 * pairs <= kThreads && parts * 2 <= f) parts *= 2;
  const int32_t h = (B - 1) / 2;
  double2 *fold_t = c.fold + int64_t(trial) * c.rows_cap * c.B_cap;
  if (parts == 1) {
    for (int p = threadIdx.x; p < pairs; p += blockDim.x) {
      const int rr = p / nrho;
      const int ri = p - rr * nrho;
      const int row = row1 + rr;
      const int off = int(RowOffset(row)) - lam_lo;
      double2 acc = make_double2(0.0, 0.0);
      for (int k = 0; k < f; ++k) acc = Add(acc, Cmul(X[(off + k) * kResBlock + ri], T[k * kResBlock + ri]));
      const int rho = rho0 + ri;
      const int m = rho >= h ? rho - h : rho - h + B;
      fold_t[int64_t(row) * c.B_cap + m] = acc;
    }
  } else {
    const int p = threadIdx.x;
    const int pair = p / parts;
    const int part = p - pair * parts;
    double2 acc = make_double2(0.0, 0.0);
    int row = 0, ri = 0;
    if (pair < pairs) {
      const int rr = pair / nrho;
      ri = pair - rr * nrho;
      row = row1 + rr;
      const int off = int(RowOffset(row)) - lam_lo;
      for (int k = part; k < f; k += parts)
        acc = Add(acc, Cmul(X[(off + k) * kResBlock + ri], T[k * kResBlock + ri]));
    }
    red[p] = acc;
    __syncthreads();
    if (pair < pairs && part == 0) {
      double2 s = red[p];
      for (int j = 1; j < parts; ++j) s = Add(s, red[p + j]);
      const int rho = rho0 + ri;
      const int m = rho >= h ? rho - h : rho - h + B;
      fold_t[int64_t(row) * c.B_cap + m] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// k_bin_fft: per (trial, row) the forward B-point DFT of the fold by
// Bluestein (X_k = c_k sum_j (v_j c_j) conj(c_{k-j})), an LDS FFT of
// M = 2^logM >= 2B-1 points each way; then the row phase e^{2 pi i b tau/n}
// and the BinInFreqV1 subtraction (src/binner.cc:565-591) of every found
// mode that falls in this row.
__global__ void __launch_bounds__(kThreads) k_bin_fft(Ctx c) {
  extern __shared__ double2 lds[];
  const DevState *st = c.st;
  if (st->done) return;
  const int32_t trials = st->trials, rows = st->rows;
  const int32_t trial = blockIdx.x / c.rows_cap;
  const int32_t row = blockIdx.x - trial * c.rows_cap;
  if (trial >= trials || row >= rows) return;
  const int32_t B = st->B, logM = st->logM;
  const int M = 1 << logM;
  const int32_t it = st->it;
  const DevTf tf = c.tfs[it];
  const int64_t n = c.n, mask = c.mask;
  const double2 *chirp = c.table + st->tab_off;
  const double2 *filt = chirp + B;
  const double2 *fold = c.fold + (int64_t(trial) * c.rows_cap + row) * c.B_cap;

  for (int m = threadIdx.x; m < M; m += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (m < B) v = Cmul(fold[m], chirp[m]);
    lds[BitRev(m, logM)] = v;
  }
  __syncthreads();
  FftDit(lds, logM, false, c.tw);
  for (int k = threadIdx.x; k < M; k += blockDim.x) lds[k] = Cmul(lds[k], filt[k]);
  __syncthreads();
  FftDif(lds, logM, true, c.tw);

  const int64_t q = c.qs[int64_t(it) * c.trials_cap + trial];
  const int64_t tau = ModN(q + RowOffset(row) * B, n, mask);
  const double2 rp = CisFrac(ModN(tf.b * tau, n, mask), n);
  constexpr int kPer = (1 << kMaxLogM) / 2 / kThreads + 1;  // >= ceil(B/256)
  double2 vals[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int k = threadIdx.x + r * kThreads;
    if (k < B) vals[r] = Cmul(Cmul(lds[BitRev(k, logM)], chirp[k]), rp);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int k = threadIdx.x + r * kThreads;
    if (k < B) lds[k] = vals[r];
  }
  __syncthreads();

  const int32_t cnt = min(st->count, c.cap);
  const ModeTerm *terms = c.terms + int64_t(trial) * c.cap;
  const int64_t off = RowOffset(row) < 0 ? -RowOffset(row) : RowOffset(row);
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
    const ModeTerm t = terms[i];
    double2 contrib = t.base;
    if (row != 0) {
      double2 g = CisFrac(ModN(off * B * t.k1, n, mask), n);
      if ((row & 1) == 0) g = Conj(g);
      contrib = Cmul(t.base, g);
    }
    atomicAdd(&lds[t.bin].x, -contrib.x);
    atomicAdd(&lds[t.bin].y, -contrib.y);
  }
  __syncthreads();
  double2 *A = c.A + (int64_t(trial) * c.rows_cap + row) * c.B_cap;
  for (int k = threadIdx.x; k < B; k += blockDim.x) A[k] = lds[k];
}

// ---------------------------------------------------------------------------
// k_identify: per bin, Iterate's loop body (src/iterate.cc:113-175):
// energy test, IdentifyFreq (src/iterate.cc:44-76) with the closed-form 2x2
// pencil, k1 rounding, window test, coefficient estimate, un-permutation,
// and mm[k0] += coef.  FP contraction is off so the decision arithmetic
// rounds as the reference's does.  The last workgroup to finish advances
// the loop state as Demo1::Run does (src/demo1.cc:45-58).
__device__ __forceinline__ void InsertMode(const Ctx &c, int64_t k0, double2 v) {
  const uint32_t bit = 1u << (k0 & 31);
  const uint32_t old = atomicOr(&c.present[k0 >> 5], bit);
  if (!(old & bit)) {
    const int32_t idx = atomicAdd(&c.st->count, 1);
    if (idx < c.cap)
      c.keys[idx] = k0;
    else
      atomicOr(&c.st->flags, kFlagOverflow);
  }
  atomicAdd(&c.coef[k0].x, v.x);
  atomicAdd(&c.coef[k0].y, v.y);
}

__global__ void __launch_bounds__(kThreads) k_identify(Ctx c) {
#pragma clang fp contract(off)
  DevState *st = c.st;
  if (st->done) return;
  const int32_t B = st->B, bits = st->bits, trials = st->trials, it = st->it;
  const int64_t n = c.n, mask = c.mask;
  const DevTf tf = c.tfs[it];
  const double bin_thr = st->bin_thr, sv_thr = st->sv_thr, win_thr = st->win_thr;
  const double width = st->width, dfreq = st->dfreq;
  const int64_t rstride = c.B_cap, tstride = int64_t(c.rows_cap) * c.B_cap;
  bool any = false;

  for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
    const double2 *Ab = c.A + b;
    if (bin_thr > 0) {
      double e = 0;
      for (int t = 0; t < trials; ++t) {
        const double2 v = Ab[t * tstride];
        e += v.x * v.x + v.y * v.y;
      }
      e /= trials;
      if (e < bin_thr * bin_thr) continue;
    }
    // IdentifyFreq: majority vote per bit, most significant bit first.
    int32_t xi_sum = 0;
    bool reject = false;
    for (int32_t bit = 0; bit < bits && !reject; ++bit) {
      int32_t count = 0;
      for (int t = 0; t < trials; ++t) {
        const double2 u0 = Ab[t * tstride];
        const double2 u1 = Ab[t * tstride + (2 * bit + 1) * rstride];
        const double2 u2 = Ab[t * tstride + (2 * bit + 2) * rstride];
        // MatPencil (src/freqid.cc:27-48) in closed form: vote is
        // Im(conj(u0) u2 + conj(u1) u0) > 0, sigma2 = |u0^2 - u1 u2| / sigma1.
        const double alpha = (u0.x * u0.x + u0.y * u0.y) + (u1.x * u1.x + u1.y * u1.y);
        const double gamma = (u2.x * u2.x + u2.y * u2.y) + (u0.x * u0.x + u0.y * u0.y);
        const double2 beta = Add(Cmul(Conj(u0), u2), Cmul(Conj(u1), u0));
        if (beta.y > 0) ++count;
        if (sv_thr > 0) {
          const double hh = 0.5 * (alpha - gamma);
          const double lam1 =
              0.5 * (alpha + gamma) + sqrt(hh * hh + (beta.x * beta.x + beta.y * beta.y));
          const double s1 = sqrt(lam1);
          const double2 det = Sub(Cmul(u0, u0), Cmul(u1, u2));
          const double s2 = s1 > 0 ? hypot(det.x, det.y) / s1 : 0.0;
          if (s2 * 0.5 > sv_thr) {
            reject = true;
            break;
          }
        }
      }
      xi_sum <<= 1;
      if (count > trials / 2) ++xi_sum;
    }
    if (reject) continue;
    const double xi1 = double(2 * xi_sum + 1) / double(1 << (bits + 1));
    // k1 = std::round(n (b + xi1) / B)  (round half away from zero).
    const int64_t k1 = int64_t(round(double(n) * (double(b) + xi1) / double(B)));
    const double xi0 = double(k1) / double(n) - (double(b) + 0.5) / double(B);
    const double wf = SampleInFreq(xi0, width, dfreq);
    if (fabs(wf) < win_thr) continue;
    // Coefficient estimate (src/iterate.cc:147-166).
    double2 sum = make_double2(0.0, 0.0);
    for (int t = 0; t < trials; ++t) {
      const int64_t q = c.qs[int64_t(it) * c.trials_cap + t];
      const double2 fac = Conj(CisFrac(ModN(q * k1, n, mask), n));
      sum = Add(sum, Cmul(Ab[t * tstride], fac));
      for (int32_t bit = 0; bit < bits; ++bit) {
        const int64_t off = int64_t(B) << bit;
        const double2 fac2 = Conj(CisFrac(ModN(off * k1, n, mask), n));
        const double2 f1 = Cmul(fac, fac2);
        const double2 f2 = Cmul(fac, Conj(fac2));
        sum = Add(sum, Cmul(Ab[t * tstride + (2 * bit + 1) * rstride], f1));
        sum = Add(sum, Cmul(Ab[t * tstride + (2 * bit + 2) * rstride], f2));
      }
    }
    const double denom = double(trials * (1 + 2 * bits));
    double2 coef = make_double2(sum.x / denom, sum.y / denom);
    // Undo the permutation: k0 = a_inv (k1 - b) mod n; coef *= e^{-2 pi i c k0/n}.
    const int64_t k0 = ModN(tf.a_inv * ModN(k1 - tf.b, n, mask), n, mask);
    coef = Cmul(coef, Conj(CisFrac(ModN(tf.c * k0, n, mask), n)));
    coef = make_double2(coef.x / wf, coef.y / wf);
    InsertMode(c, k0, coef);
    any = true;
  }

  if (any) atomicOr(&st->found_iter, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t ticket = atomicAdd(&st->blocks_done, 1u);
    if (ticket == gridDim.x - 1) {
      __threadfence();
      // Last workgroup: every insert of this iteration is visible.
      const int32_t cnt = atomicAdd(&st->count, 0);
      const int32_t fi = atomicOr(&st->found_iter, 0);
      if (it < c.log_cap) {
        c.log[it] = B;
        c.log[c.log_cap + it] = cnt;
      }
      st->found_last = fi;
      if (!fi) st->stale += 1;
      st->it = it + 1;
      st->found_iter = 0;
      st->blocks_done = 0u;
      const bool overflow = (atomicOr(&st->flags, 0) & kFlagOverflow) != 0;
      if (st->mode != kModeSolve || overflow || st->stale >= st->max_stale ||
          st->it >= st->max_iter) {
        st->done = 1;
      } else {
        // B = max((k - |found|)*2 + 1, min_bins) (src/demo1.cc:48-51).
        const int64_t want = (int64_t(st->k) - int64_t(cnt)) * 2 + 1;
        const int32_t Bn = int32_t(want > st->min_bins ? want : st->min_bins);
        const BParams bp = c.btab[Bn];
        LoadBParams(st, bp, 0);
        st->bin_thr = bp.bin_thr;
        st->sv_thr = bp.sv_thr;
      }
    }
  }
}

// BinInFreqV1 (src/binner.cc:565-591) on a caller-owned bin matrix: one
// thread per mode, subtracting from every row with device atomics.
__global__ void __launch_bounds__(kThreads)
    k_bin_in_freq_direct(int64_t n, int64_t mask, BParams bp, DevTf tf, int64_t q,
                         const int64_t *k0s, const double2 *coefs, int64_t m,
                         double2 *out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t B = bp.B;
  const int64_t k0 = ModN(k0s[i], n, mask);
  const int64_t k1 = ModN(tf.a * k0 + tf.b, n, mask);
  const int32_t bin = int32_t(k1 * B / n);
  const double xi = (double(bin) + 0.5) / double(B) - double(k1) / double(n);
  const double wf = SampleInFreq(xi, bp.width, bp.dfreq);
  const double2 base = Cmul(Scale(coefs[i], wf), CisFrac(ModN(q * k1 + tf.c * k0, n, mask), n));
  atomicAdd(&out[bin].x, -base.x);
  atomicAdd(&out[bin].y, -base.y);
  for (int32_t bit = 0; bit < bp.bits; ++bit) {
    const double2 g = CisFrac(ModN((int64_t(B) << bit) * k1, n, mask), n);
    const double2 c1 = Cmul(base, g);
    const double2 c2 = Cmul(base, Conj(g));
    double2 *r1 = out + int64_t(2 * bit + 1) * B + bin;
    double2 *r2 = out + int64_t(2 * bit + 2) * B + bin;
    atomicAdd(&r1->x, -c1.x);
    atomicAdd(&r1->y, -c1.y);
    atomicAdd(&r2->x, -c2.x);
    atomicAdd(&r2->y, -c2.y);
  }
}

__global__ void __launch_bounds__(kThreads)
    k_mat_pencil(const double2 *u, int64_t m, double *sigma, uint8_t *vote) {
#pragma clang fp contract(off)
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double2 u0 = u[3 * i], u1 = u[3 * i + 1], u2 = u[3 * i + 2];
  const double alpha = (u0.x * u0.x + u0.y * u0.y) + (u1.x * u1.x + u1.y * u1.y);
  const double gamma = (u2.x * u2.x + u2.y * u2.y) + (u0.x * u0.x + u0.y * u0.y);
  const double2 beta = Add(Cmul(Conj(u0), u2), Cmul(Conj(u1), u0));
  const double hh = 0.5 * (alpha - gamma);
  const double lam1 = 0.5 * (alpha + gamma) + sqrt(hh * hh + (beta.x * beta.x + beta.y * beta.y));
  const double s1 = sqrt(lam1);
  const double2 det = Sub(Cmul(u0, u0), Cmul(u1, u2));
  sigma[2 * i] = s1;
  sigma[2 * i + 1] = s1 > 0 ? hypot(det.x, det.y) / s1 : 0.0;
  vote[i] = beta.y > 0 ? 1 : 0;
}

int32_t Blocks(int64_t work) {
  const int64_t b = (work + kThreads - 1) / kThreads;
  return int32_t(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

}  // namespace

hipError_t LaunchBuildTwiddles(double2 *tw, hipStream_t s) {
  k_build_twiddles<<<Blocks(1 << (kMaxLogM - 1)), kThreads, 0, s>>>(tw);
  return hipGetLastError();
}

hipError_t LaunchBuildTables(const int32_t *blist, int32_t nb, const BParams *btab,
                             double2 *table, const double2 *tw, int32_t M_cap,
                             hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  k_build_tables<<<nb, kThreads, size_t(M_cap) * sizeof(double2), s>>>(blist, btab, table, tw);
  return hipGetLastError();
}

hipError_t LaunchClearModes(const Ctx &c, int32_t full, hipStream_t s) {
  k_clear_modes<<<1024, kThreads, 0, s>>>(c, full);
  return hipGetLastError();
}

hipError_t LaunchInit(const Ctx &c, const InitArgs &a, hipStream_t s) {
  k_init<<<1, 64, 0, s>>>(c, a);
  return hipGetLastError();
}

hipError_t LaunchBinTime(const Ctx &c, int32_t tiles, size_t lds, hipStream_t s) {
  if (c.mixed)
    k_bin_time<float2><<<kPrepBlocks + tiles, kThreads, lds, s>>>(c);
  else
    k_bin_time<double2><<<kPrepBlocks + tiles, kThreads, lds, s>>>(c);
  return hipGetLastError();
}

hipError_t LaunchBinFft(const Ctx &c, size_t lds, hipStream_t s) {
  k_bin_fft<<<c.trials_cap * c.rows_cap, kThreads, lds, s>>>(c);
  return hipGetLastError();
}

hipError_t LaunchIdentify(const Ctx &c, int32_t blocks, hipStream_t s) {
  k_identify<<<blocks, kThreads, 0, s>>>(c);
  return hipGetLastError();
}

hipError_t LaunchSetModes(const Ctx &c, const int64_t *keys, const double2 *coefs,
                          int64_t m, hipStream_t s) {
  k_set_modes<<<Blocks(m), kThreads, 0, s>>>(c, keys, coefs, m);
  return hipGetLastError();
}

hipError_t LaunchGetCoefs(const Ctx &c, double2 *out, int64_t m, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  k_get_coefs<<<Blocks(m), kThreads, 0, s>>>(c, out, m);
  return hipGetLastError();
}

hipError_t LaunchBinInFreqDirect(int64_t n, int64_t mask, const BParams &bp,
                                 const DevTf &tf, int64_t q, const int64_t *k0,
                                 const double2 *coef, int64_t m, double2 *out,
                                 hipStream_t s) {
  if (m <= 0) return hipSuccess;
  k_bin_in_freq_direct<<<Blocks(m), kThreads, 0, s>>>(n, mask, bp, tf, q, k0, coef, m, out);
  return hipGetLastError();
}

hipError_t LaunchMatPencil(const double2 *u, int64_t m, double *sigma, uint8_t *vote,
                           hipStream_t s) {
  if (m <= 0) return hipSuccess;
  k_mat_pencil<<<Blocks(m), kThreads, 0, s>>>(u, m, sigma, vote);
  return hipGetLastError();
}

}  // namespace mpsft
