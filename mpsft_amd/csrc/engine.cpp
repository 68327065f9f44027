// engine.cpp -- host side of the MPSFT engine: tables, transform draws,
// device buffers and the iteration graph.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "kernels.h"

namespace mpsft {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt2 = 1.41421356237309504880;

bool IsPow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

int64_t Gcd(int64_t a, int64_t b) {
  a = a < 0 ? -a : a;
  b = b < 0 ? -b : b;
  while (b) {
    const int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

int64_t PosMod(int64_t v, int64_t n) {
  const int64_t r = v % n;
  return r < 0 ? r + n : r;
}

template <typename T>
int Alloc(T **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) return MPSFT_OK;
  if (hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return MPSFT_ERR_NOMEM;
  }
  return MPSFT_OK;
}

#define HIP_OK(expr)                                 \
  do {                                               \
    if ((expr) != hipSuccess) return MPSFT_ERR_HIP;  \
  } while (0)

// Distinct lags of one iteration (the union of every row's f lags).
int32_t NumLags(int32_t bits, int32_t f) {
  int K = 0;
  while (K < bits && (1 << K) < f) ++K;
  return (K ? (1 << K) : 0) + f + 2 * (bits - K) * f;
}

// Residues per k_bin_time tile: the largest power of two <= 8 that still
// gives `target` tiles, and keeps the tile's LDS under 64 KiB.  A forced
// width (forced > 0, from the autotuner) is kept within the LDS of one
// workgroup.
int32_t ResiduesPerTile(int32_t B, int32_t nlam, int32_t f, int32_t forced, int32_t target) {
  if (forced > 0) {
    int32_t R = forced;
    while (R > 1 && (size_t(nlam + f) * R + kThreads) * sizeof(double2) > 160 * 1024) R /= 2;
    return R;
  }
  int32_t R = 8;
  while (R > 1 && ((B + R - 1) / R < target ||
                   size_t(nlam + f) * R * sizeof(double2) > 65536))
    R /= 2;
  return R;
}

// Split of an iteration's gather-fold over the XCDs (1 or kSlices), opt-in
// with MPSFT_SLICES=8 where n is a power of two and the bin FFT is the LDS
// one.  Each XCD then gathers only its eighth of x, so its L2 (4 MiB) serves
// every gather into that eighth; at N = 2^22 complex64 (32 MiB) a gather-
// only test runs 1.56x faster that way.  But every slice repeats the
// tile's taps and its dense fold sums, and measured on MI355X the whole
// solve is 1.25x (one signal) to 2x (16 signals batched) slower, so the
// unsplit form is the default.  The split is kept for measurement and is
// checked against the oracle like the default.
bool SlicingOn(int64_t n, int64_t mask) {
  static const int mode = [] {
    const char *v = std::getenv("MPSFT_SLICES");
    return v ? std::atoi(v) : 0;
  }();
  return mode == kSlices && mask >= 0 && n >= 64;
}

int32_t SlicesFor(int64_t n, int64_t mask, const BParams &bp) {
  return SlicingOn(n, mask) && bp.logM <= kMaxLogM ? kSlices : 1;
}

int32_t Log2(int64_t n) {
  int32_t l = 0;
  while ((int64_t(1) << (l + 1)) <= n) ++l;
  return l;
}

// Transforms drawn before a solve's first launch.
constexpr int32_t kFirstDraws = 32;

int32_t ChunkIters() {
  const char *e = std::getenv("MPSFT_LOOKAHEAD");
  if (e) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 4096) return v;
  }
  return 2;
}

// Lookahead once every launch is a late (two-kernel) iteration: B no longer
// changes, so launching further ahead commits no graph choice early, and
// each iteration is short enough that two queued ones do not cover the
// host's launch latency.
// Iterations per late graph (even): once every launch is a late iteration the
// graph kind no longer changes, so longer graphs cost no early commitment
// and cross fewer inter-graph gaps.
int32_t LateGraphIters() {
  const char *e = std::getenv("MPSFT_LATE_GRAPH");
  if (e) {
    const int v = std::atoi(e);
    if (v >= 2 && v <= 64 && v % 2 == 0) return v;
  }
  return 6;
}

int32_t LateLookahead() {
  const char *e = std::getenv("MPSFT_LATE_LOOKAHEAD");
  if (e) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 4096) return v;
  }
  return 3;
}

}  // namespace

// How long a host launch loop spins without a progress report before it
// waits on the stream instead (MPSFT_STALL_MS, default 200 ms).
std::chrono::milliseconds StallWait() {
  const char *e = std::getenv("MPSFT_STALL_MS");
  if (e) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 3600000) return std::chrono::milliseconds(v);
  }
  return std::chrono::milliseconds(200);
}

// NumBits (src/iterate.cc:30-42): 1 + floor(log2(n / bins)), at least 1.
int32_t NumBits(int64_t n, int32_t bins) {
  if (n < bins) return 1;
  int32_t out = 0;
  while (n >= bins) {
    n >>= 1;
    ++out;
  }
  return out;
}

// Window parameters with the reference's expressions (src/window.cc:28-50).
bool WindowParams(int64_t n, int32_t bins, double delta, BParams *bp) {
  if (n <= 0 || bins <= 0 || !(delta > 0) || bins % 2 != 1) return false;
  const double width = 1.0 / (2.0 * bins);
  const double sqrt_c_delta = std::sqrt(-std::log(delta));
  const double sigma_f = 0.5 / (bins * 2.0 * kSqrt2 * sqrt_c_delta);
  const double sigma_t = 1.0 / ((2.0 * kPi) * sigma_f);
  const double tmp = 2.0 * kSqrt2 * sigma_t * sqrt_c_delta + 1;
  int32_t factor = int32_t(std::ceil(tmp / double(bins)));
  if (factor % 2 == 0) ++factor;
  bp->B = bins;
  bp->f = factor;
  bp->P = factor * bins;
  bp->bits = NumBits(n, bins);
  int32_t logM = 1;
  while ((int64_t(1) << logM) < 2 * int64_t(bins) - 1) ++logM;
  bp->logM = logM;
  bp->M = 1 << logM;
  bp->off = 0;
  bp->R = 1;
  bp->slices = 1;
  bp->slice_shift = 0;
  bp->width = width;
  bp->sigma_f = sigma_f;
  bp->dfreq = kSqrt2 * kPi * sigma_t;
  bp->bin_thr = 0;
  bp->sv_thr = 0;
  return true;
}

// Window(n, bins, delta) exposes its taps but not delta (src/window.h:25-55).
// Its taps are wt_t = w sinc(pi t w) e^{-2 (pi t sigma_f)^2} with w = 1/(2
// bins) (src/window.cc:71-75), so each tap away from the sinc's zeros gives
// sigma_f^2 = -ln(wt_t / (w sinc(pi t w))) / (2 pi^2 t^2), and delta follows
// from sigma_f = 1/(4 sqrt(2) bins sqrt(-ln delta)) (src/window.cc:35-37).
// The taps whose Gaussian factor lies in [1e-8, 0.5] are averaged (each
// carries a relative error of about 1e-16 / |ln factor|).  The window of the
// recovered delta must have p taps and reproduce every tap to 1e-9 of w.
bool WindowDeltaFromTaps(int64_t n, int32_t bins, int32_t p, const double *wt, double *delta) {
  if (!wt || bins <= 0 || bins % 2 != 1 || p < bins || p % bins != 0 || (p / bins) % 2 != 1)
    return false;
  const double w = 1.0 / (2.0 * bins);
  if (!(std::fabs(wt[0] - w) <= 1e-12 * w)) return false;
  const int32_t p2 = (p - 1) / 2;
  // Tens of thousands of nearly equal estimates: summed with Neumaier's
  // compensation (a plain running sum drifts by ~1e-12 relative here).
  double sum = 0, comp = 0;
  int32_t used = 0;
  for (int32_t t = 1; t <= p2; ++t) {
    const double s = std::sin(kPi * t * w);
    if (std::fabs(s) < 0.05) continue;  // near a zero of the sinc
    const double g = wt[t] / (w * s / (kPi * t * w));
    if (!(g >= 1e-8 && g <= 0.5)) continue;
    const double e = -std::log(g) / (2.0 * kPi * kPi * double(t) * double(t));
    const double nsum = sum + e;
    comp += std::fabs(sum) >= std::fabs(e) ? (sum - nsum) + e : (e - nsum) + sum;
    sum = nsum;
    ++used;
  }
  if (used == 0 || !(sum > 0)) return false;
  const double sigma_f2 = (sum + comp) / used;
  const double c = 1.0 / (32.0 * double(bins) * double(bins) * sigma_f2);
  const double d = std::exp(-c);
  BParams bp;
  if (!(d > 0 && d < 1) || !WindowParams(n, bins, d, &bp) || bp.P != p) return false;
  // every tap of the recovered window against the given ones
  for (int32_t i = 1; i < p; ++i) {
    const int32_t t = i <= p2 ? i : i - p;
    const double x = kPi * std::fabs(double(t)) * w;
    const double u = kPi * double(t) * bp.sigma_f;
    const double tap = w * (std::sin(x) / x) * std::exp(-2.0 * u * u);
    if (!(std::fabs(tap - wt[i]) <= 1e-9 * w)) return false;
  }
  *delta = d;
  return true;
}

// Inverse modulo n; 0 when a is not a unit.  For n = 2^k and odd a, Newton's
// iteration x <- x (2 - a x) in 64-bit words: x = a is right to 3 bits
// (a^2 = 1 mod 8) and every step doubles them, so five steps give 96 >= 64.
// Otherwise extended Euclid.
int64_t InvMod(int64_t a, int64_t n) {
  if (n > 1 && (n & (n - 1)) == 0 && (a & 1)) {
    const uint64_t u = uint64_t(a);
    uint64_t x = u;
    for (int i = 0; i < 5; ++i) x *= 2 - u * x;
    return int64_t(x & uint64_t(n - 1));
  }
  int64_t r0 = n, r1 = PosMod(a, n);
  int64_t s0 = 0, s1 = 1;
  while (r1 != 0) {
    const int64_t qq = r0 / r1;
    int64_t t = r0 - qq * r1;
    r0 = r1;
    r1 = t;
    t = s0 - qq * s1;
    s0 = s1;
    s1 = t;
  }
  if (r0 != 1) return 0;
  return PosMod(s0, n);
}

DevTf ReduceTf(int64_t n, int64_t a, int64_t b, int64_t c, int64_t a_inv) {
  DevTf t;
  t.a = PosMod(a, n);
  t.b = PosMod(b, n);
  t.c = PosMod(c, n);
  t.a_inv = PosMod(a_inv, n);
  return t;
}

Engine::Engine(int64_t n, int precision, int64_t capacity)
    : n_(n), mask_(IsPow2(n) ? n - 1 : -1), precision_(precision) {
  if (n < 2 || n > 0x7fffffffLL || (precision != MPSFT_FP64 && precision != MPSFT_MIXED) ||
      capacity <= 0 || capacity > n || capacity > 0x7fffffffLL) {
    status_ = MPSFT_ERR_INVALID_ARG;
    return;
  }
  cap_ = int32_t(capacity);
  log_cap_ = 1 << 16;
  int rc = MPSFT_OK;
  rc |= Alloc(&d_state_, 1);
  rc |= Alloc(&d_tw_, size_t(1) << (kTwLogM - 1));
  rc |= Alloc(&d_keys_, cap_);
  rc |= Alloc(&d_coef_, size_t(n_));
  rc |= Alloc(&d_present_, size_t((n_ + 31) / 32));
  rc |= Alloc(&d_log_, size_t(2) * log_cap_);
  rc |= Alloc(&d_tmp_, cap_);
  rc |= Alloc(&d_tmpk_, cap_);
  if (rc != MPSFT_OK) {
    status_ = MPSFT_ERR_NOMEM;
    return;
  }
  if (hipHostMalloc(reinterpret_cast<void **>(&h_state_), sizeof(DevState)) != hipSuccess) {
    status_ = MPSFT_ERR_NOMEM;
    return;
  }
  int32_t *hp = nullptr;
  if (hipHostMalloc(reinterpret_cast<void **>(&hp), 8 * sizeof(int32_t),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void **>(&d_progress_), hp, 0) != hipSuccess) {
    status_ = MPSFT_ERR_NOMEM;
    return;
  }
  h_progress_ = hp;
  for (int i = 0; i < 8; ++i) hp[i] = 0;
  std::memset(h_state_, 0, sizeof(DevState));
  if (hipMemset(d_state_, 0, sizeof(DevState)) != hipSuccess ||
      hipMemset(d_coef_, 0, size_t(n_) * sizeof(double2)) != hipSuccess ||
      hipMemset(d_present_, 0, size_t((n_ + 31) / 32) * sizeof(uint32_t)) != hipSuccess ||
      LaunchBuildTwiddles(d_tw_, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    status_ = MPSFT_ERR_HIP;
    return;
  }
}

Engine::~Engine() {
  // Work a caller left queued may still use these buffers.
  (void)hipDeviceSynchronize();
  for (auto e : graph_exec_)
    if (e) (void)hipGraphExecDestroy(e);
  (void)hipFree(d_state_);
  (void)hipFree(d_btab_);
  (void)hipFree(d_table_);
  (void)hipFree(d_tw_);
  (void)hipFree(d_tfs_);
  (void)hipFree(d_qs_);
  (void)hipFree(d_fold_);
  (void)hipFree(d_A_);
  (void)hipFree(d_big_);
  (void)hipFree(d_rec_);
  (void)hipFree(d_lam_);
  (void)hipFree(d_fbar_);
  (void)hipFree(d_taps_);
  (void)hipFree(d_lags_);
  (void)hipFree(d_ctx1_);
  (void)hipFree(d_sig1_);
  (void)hipFree(d_steps1_);
  (void)hipFree(d_term_bin_);
  (void)hipFree(d_term_val_);
  (void)hipFree(d_keys_);
  (void)hipFree(d_coef_);
  (void)hipFree(d_present_);
  (void)hipFree(d_log_);
  (void)hipFree(d_tmp_);
  (void)hipFree(d_tmpk_);
  (void)hipFree(d_wtab_);
  (void)hipFree(d_idout_);
  (void)hipFree(d_pa_);
  if (h_state_) (void)hipHostFree(h_state_);
  if (p_tfs_) (void)hipHostFree(p_tfs_);
  if (p_qs_) (void)hipHostFree(p_qs_);
  if (h_progress_) (void)hipHostFree(const_cast<int32_t *>(h_progress_));
}

// Builds the window parameters of every B in bs (for window delta) that is
// not built yet, and for B <= kMaxLdsBins its chirp and Bluestein filter.
int Engine::EnsureTables(const std::vector<int32_t> &bs, double delta, int32_t bits_override) {
  if (delta != tab_delta_) {
    hbtab_.clear();
    built_.clear();
    table_used_ = 0;
    tab_delta_ = delta;
    tab_sigma_ = -1;
    ++tab_epoch_;
  }
  int32_t Bmax = 0;
  for (int32_t B : bs) {
    if (B <= 0 || B % 2 != 1) return MPSFT_ERR_INVALID_ARG;
    if (B > kMaxBins) return MPSFT_ERR_UNSUPPORTED;
    Bmax = std::max(Bmax, B);
  }
  if (int32_t(hbtab_.size()) <= Bmax) {
    hbtab_.resize(Bmax + 1);
    built_.resize(Bmax + 1, 0);
  }
  std::vector<int32_t> todo;
  int64_t need = 0;
  int32_t Mmax = 0;
  bool changed = false;
  for (int32_t B : bs) {
    if (built_[B]) continue;
    changed = true;
    BParams bp;
    if (!WindowParams(n_, B, delta, &bp)) return MPSFT_ERR_INVALID_ARG;
    if (bp.logM > kMaxLogMBig) return MPSFT_ERR_UNSUPPORTED;
    const int32_t nlam = NumLags(bits_override > 0 ? bits_override : bp.bits, bp.f);
    bp.R = ResiduesPerTile(B, nlam, bp.f, forced_R_[B >= kSmallTileB ? 0 : 1], tile_target_);
    bp.slices = SlicesFor(n_, mask_, bp);
    bp.slice_shift = bp.slices > 1 ? Log2(n_) - 3 : 0;
    bp.off = 0;
    if (bp.logM <= kMaxLogM) {
      bp.off = table_used_ + need;
      need += B + bp.M;
      todo.push_back(B);
      Mmax = std::max(Mmax, bp.M);
    }
    hbtab_[B] = bp;
    built_[B] = 1;
  }
  if (bits_override != tab_bits_override_) {
    tab_bits_override_ = bits_override;
    ++tab_epoch_;
  }
  if (changed) ++tab_epoch_;
  if (table_used_ + need > table_len_) {
    const int64_t len = std::max<int64_t>(table_used_ + need, 2 * table_len_);
    double2 *nt = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&nt), size_t(len) * sizeof(double2)) != hipSuccess)
      return MPSFT_ERR_NOMEM;
    if (d_table_) {
      HIP_OK(hipMemcpy(nt, d_table_, size_t(table_used_) * sizeof(double2),
                       hipMemcpyDeviceToDevice));
      (void)hipFree(d_table_);
    }
    d_table_ = nt;
    table_len_ = len;
    ++shape_gen_;
  }
  table_used_ += need;
  if (d_btab_len_ < int32_t(hbtab_.size())) {
    if (Alloc(&d_btab_, hbtab_.size()) != MPSFT_OK) return MPSFT_ERR_NOMEM;
    d_btab_len_ = int32_t(hbtab_.size());
    ++shape_gen_;
    tab_sigma_ = -1;
    changed = true;
  }
  // The device copy changes only with new entries (a solve with the same
  // range of bins uploads nothing).
  if (changed)
    HIP_OK(hipMemcpy(d_btab_, hbtab_.data(), hbtab_.size() * sizeof(BParams),
                     hipMemcpyHostToDevice));
  if (!todo.empty()) {
    int32_t *d_list = nullptr;
    if (Alloc(&d_list, todo.size()) != MPSFT_OK) return MPSFT_ERR_NOMEM;
    HIP_OK(hipMemcpy(d_list, todo.data(), todo.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_OK(LaunchBuildTables(d_list, int32_t(todo.size()), d_btab_, d_table_, d_tw_, Mmax,
                             nullptr));
    HIP_OK(hipDeviceSynchronize());
    (void)hipFree(d_list);
  }
  return MPSFT_OK;
}

// Demo1 thresholds per B: 3 sigma / sqrt(B) and sigma / sqrt(B)
// (src/demo1.cc:50-51).
int Engine::SetThresholds(double sigma) {
  if (sigma == tab_sigma_) return MPSFT_OK;
  ++tab_epoch_;
  for (size_t B = 0; B < hbtab_.size(); ++B) {
    if (!built_[B]) continue;
    hbtab_[B].bin_thr = 3.0 * sigma / std::sqrt(double(B));
    hbtab_[B].sv_thr = 1.0 * sigma / std::sqrt(double(B));
  }
  HIP_OK(hipMemcpy(d_btab_, hbtab_.data(), hbtab_.size() * sizeof(BParams),
                   hipMemcpyHostToDevice));
  tab_sigma_ = sigma;
  return MPSFT_OK;
}

// Fixes the k_bin_time tile width (residues per tile) of every B >=
// kSmallTileB and of every smaller B, each back to the default heuristic
// for R <= 0.  Tiling changes only how the work is split, not what is
// summed; the buffers are re-sized by Reserve.
int Engine::SetTileResidues(int32_t R_wide, int32_t R_narrow) {
  if (status_ != MPSFT_OK) return status_;
  for (const int32_t R : {R_wide, R_narrow})
    if (R > 8 || (R > 0 && (R & (R - 1)))) return MPSFT_ERR_INVALID_ARG;
  Drain();
  ++tab_epoch_;
  forced_R_[0] = R_wide > 0 ? R_wide : 0;
  forced_R_[1] = R_narrow > 0 ? R_narrow : 0;
  for (size_t B = 0; B < hbtab_.size(); ++B) {
    if (!built_[B]) continue;
    BParams &bp = hbtab_[B];
    const int32_t bits = tab_bits_override_ > 0 ? tab_bits_override_ : bp.bits;
    bp.R = ResiduesPerTile(int32_t(B), NumLags(bits, bp.f), bp.f,
                           forced_R_[B >= size_t(kSmallTileB) ? 0 : 1], tile_target_);
  }
  if (d_btab_)
    HIP_OK(hipMemcpy(d_btab_, hbtab_.data(), hbtab_.size() * sizeof(BParams),
                     hipMemcpyHostToDevice));
  ++shape_gen_;
  return MPSFT_OK;
}

// Sizes the per-iteration buffers for trials and every B in bs.
int Engine::Reserve(int32_t trials, const std::vector<int32_t> &bs) {
  int32_t Bc = 0, rc = 0, tc = 0, Mc = 0, Mbig = 0;
  size_t lbt = 0, lbt32 = 0, lfft = 0, lfft_small = 0;
  bool small = false, big = false;
  for (int32_t B : bs) {
    const BParams &bp = hbtab_[B];
    const int32_t bits = tab_bits_override_ > 0 ? tab_bits_override_ : bp.bits;
    Bc = std::max(Bc, B);
    rc = std::max(rc, 1 + 2 * bits);
    tc = std::max(tc, (B + bp.R - 1) / bp.R * bp.slices);
    Mc = std::max(Mc, bp.M);
    const size_t staged = size_t(NumLags(bits, bp.f)) * bp.R;
    const size_t rest = (size_t(bp.f) * bp.R + kThreads) * sizeof(double2);
    lbt = std::max(lbt, staged * sizeof(double2) + rest);
    lbt32 = std::max(lbt32, ((staged + 1) & ~size_t(1)) * sizeof(float2) + rest);
    if (bp.logM <= kMaxLogM) {
      small = true;
      // k_bin_fft: M <= 1024 array + M/2 twiddles; larger M the padded
      // array (M + M/16).  The BinInFreq limbs of the B bins (four 8-byte
      // integers per bin: 2B <= M double2 slots) later take the array's place.
      const size_t el = bp.logM <= kFftSmallLogM ? size_t(bp.M + bp.M / 2)
                                                 : size_t(bp.M + bp.M / 16);
      lfft = std::max(lfft, el * sizeof(double2));
      if (bp.logM <= kFftSmallLogM) lfft_small = std::max(lfft_small, el * sizeof(double2));
    } else {
      big = true;
      Mbig = std::max(Mbig, bp.M);
    }
  }
  // k_identify keeps a bin's trials x rows column in 64 x kColSlices lanes.
  if (int64_t(trials) * rc > 64 * kColSlices) return MPSFT_ERR_UNSUPPORTED;
  const bool grow_rows = trials > trials_cap_ || rc > rows_cap_ || Bc > B_cap_;
  if (grow_rows) {
    trials_cap_ = std::max(trials_cap_, trials);
    rows_cap_ = std::max(rows_cap_, rc);
    B_cap_ = std::max(B_cap_, Bc);
    const size_t sz = size_t(trials_cap_) * rows_cap_ * B_cap_;
    // room for kSlices partial folds where iterations may split
    const size_t fold_parts = SlicingOn(n_, mask_) ? kSlices : 1;
    if (Alloc(&d_fold_, sz * fold_parts) != MPSFT_OK || Alloc(&d_A_, sz) != MPSFT_OK ||
        Alloc(&d_term_bin_, size_t(trials_cap_) * cap_) != MPSFT_OK ||
        Alloc(&d_term_val_, size_t(trials_cap_) * rows_cap_ * cap_) != MPSFT_OK)
      return MPSFT_ERR_NOMEM;
    HIP_OK(hipMemset(d_fold_, 0, sz * fold_parts * sizeof(double2)));
    HIP_OK(hipMemset(d_A_, 0, sz * sizeof(double2)));
    tf_valid_ = false;  // q stride follows trials_cap_
    ++shape_gen_;
  }
  if (big) {
    // scratch slots: every (trial, row) plus the filter row, M_cap each
    const int32_t mcap = std::max(big_M_cap_, Mbig);
    const size_t want = (size_t(trials_cap_) * rows_cap_ + 1) * size_t(mcap);
    if (mcap != big_M_cap_ || want != big_len_) {
      if (Alloc(&d_big_, want) != MPSFT_OK) return MPSFT_ERR_NOMEM;
      big_M_cap_ = mcap;
      big_len_ = want;
      ++shape_gen_;
    }
  }
  if (small != fft_small_ || big != fft_big_) {
    fft_small_ = small;
    fft_big_ = big;
    ++shape_gen_;
  }
  if (tc > tiles_cap_ || Mc > M_cap_ || lbt > lds_bin_time_ || lbt32 > lds_bin_time32_ ||
      lfft > lds_fft_ || lfft_small > lds_fft_small_) {
    lds_fft_small_ = std::max(lds_fft_small_, lfft_small);
    tiles_cap_ = std::max(tiles_cap_, tc);
    M_cap_ = std::max(M_cap_, Mc);
    lds_bin_time_ = std::max(lds_bin_time_, lbt);
    lds_bin_time32_ = std::max(lds_bin_time32_, lbt32);
    lds_fft_ = std::max(lds_fft_, lfft);
    ++shape_gen_;
  }
  return MPSFT_OK;
}

// Sizes the transform / q arrays for count iterations; *grown is set when
// the arrays were reallocated (their device content is then gone).  A
// captured graph holds their addresses, so a caller with graph launches in
// flight drains its stream first; the new addresses bump shape_gen_, which
// makes EnsureGraph capture again.
int Engine::GrowTransforms(int32_t count, bool *grown) {
  *grown = false;
  if (count <= tf_cap_ && tf_qstride_ == trials_cap_) return MPSFT_OK;
  const int32_t cap = std::max(count, std::max(1024, 2 * tf_cap_));
  if (Alloc(&d_tfs_, cap) != MPSFT_OK || Alloc(&d_qs_, size_t(cap) * trials_cap_) != MPSFT_OK)
    return MPSFT_ERR_NOMEM;
  // The pinned staging copies: no upload from the old ones is in flight
  // (callers drain the stream before the arrays grow).
  if (p_tfs_) (void)hipHostFree(p_tfs_);
  if (p_qs_) (void)hipHostFree(p_qs_);
  p_tfs_ = nullptr;
  p_qs_ = nullptr;
  if (hipHostMalloc(reinterpret_cast<void **>(&p_tfs_), size_t(cap) * sizeof(DevTf)) !=
          hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&p_qs_),
                    size_t(cap) * trials_cap_ * sizeof(int64_t)) != hipSuccess)
    return MPSFT_ERR_NOMEM;
  tf_cap_ = cap;
  if (tf_qstride_ != trials_cap_) tf_valid_ = false;  // the host copies' q layout changed
  tf_qstride_ = trials_cap_;
  *grown = true;
  ++shape_gen_;
  return MPSFT_OK;
}

// Draws the transforms and q's of iterations [0, count) for seed, in the
// reference's order: Transform(n) draws a, b, c (src/integer.cc:44-49),
// then Iterate draws RandomInt32() % n per trial (src/iterate.cc:103).
// For n = 2^k the multiplier is drawn odd, a = 2 (r mod n/2) + 1; for other
// composite n the draw of a repeats until gcd(a, n) = 1.  Draws already made
// for the same seed are kept (host copies) and only the new ones are
// uploaded, unless the device arrays were reallocated.  The uploads read
// pinned copies and are not waited for: the range they read is rewritten
// only by a later solve's draws (this solve's uploads precede its kernels,
// and its stop is reported by a kernel) or after the arrays grow (callers
// drain the stream first).
int Engine::EnsureTransforms(uint32_t seed, int32_t count, int32_t trials, hipStream_t s) {
  if (defer_upload_) *defer_upload_ = TfUpload{};
  bool grown = false;
  const int rc = GrowTransforms(count, &grown);
  if (rc != MPSFT_OK) return rc;
  const bool fresh = !tf_valid_ || seed != tf_seed_ || trials != tf_trials_;
  if (fresh) {
    rng_.seed(seed);
    h_tfs_.clear();
    h_qs_.clear();
    tf_count_ = 0;
    tf_seed_ = seed;
    tf_trials_ = trials;
  }
  if (!fresh && !grown && count <= tf_count_) return MPSFT_OK;
  const int32_t target = std::max(count, tf_count_);
  const int64_t n = n_;
  while (int32_t(h_tfs_.size()) < target) {
    int64_t a;
    if (IsPow2(n)) {
      a = 2 * (RandomInt32() % (n / 2)) + 1;
    } else {
      a = (RandomInt32() % (n - 1)) + 1;
      while (Gcd(a, n) != 1) a = (RandomInt32() % (n - 1)) + 1;
    }
    const int64_t b = RandomInt32() % n;
    const int64_t c = RandomInt32() % n;
    h_tfs_.push_back(ReduceTf(n, a, b, c, InvMod(a, n)));
    for (int32_t t = 0; t < trials_cap_; ++t)
      h_qs_.push_back(t < trials ? int64_t(RandomInt32() % n) : 0);
  }
  const int32_t from = (fresh || grown) ? 0 : tf_count_;
  const size_t q0 = size_t(from) * trials_cap_, q1 = size_t(target) * trials_cap_;
  if (defer_upload_) {
    *defer_upload_ = TfUpload{h_tfs_.data() + from, d_tfs_ + from, int64_t(target - from),
                              h_qs_.data() + q0, d_qs_ + q0, int64_t(q1 - q0)};
  } else {
    std::copy(h_tfs_.begin() + from, h_tfs_.begin() + target, p_tfs_ + from);
    std::copy(h_qs_.begin() + q0, h_qs_.begin() + q1, p_qs_ + q0);
    HIP_OK(hipMemcpyAsync(d_tfs_ + from, p_tfs_ + from, size_t(target - from) * sizeof(DevTf),
                          hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(d_qs_ + q0, p_qs_ + q0, (q1 - q0) * sizeof(int64_t),
                          hipMemcpyHostToDevice, s));
  }
  tf_count_ = target;
  tf_valid_ = true;
  return MPSFT_OK;
}

Ctx Engine::MakeCtx() const {
  Ctx c;
  c.st = d_state_;
  c.btab = d_btab_;
  c.table = d_table_;
  c.tw = d_tw_;
  c.tfs = d_tfs_;
  c.qs = d_qs_;
  c.n = n_;
  c.mask = mask_;
  c.fold = d_fold_;
  c.fold_stride = int64_t(trials_cap_) * rows_cap_ * B_cap_;
  c.A = A_override_ ? A_override_ : d_A_;
  c.term_bin = d_term_bin_;
  c.term_val = d_term_val_;
  c.keys = d_keys_;
  c.coef = d_coef_;
  c.present = d_present_;
  c.log = d_log_;
  c.log_cap = log_cap_;
  c.progress = d_progress_;
  c.cap = cap_;
  c.trials_cap = trials_cap_;
  c.rows_cap = rows_cap_;
  c.B_cap = B_cap_;
  c.big = d_big_;
  c.big_gpr = fft_big_ ? big_M_cap_ / kBigTile : 0;
  c.mixed = precision_ == MPSFT_MIXED ? 1 : 0;
  c.batch_done = nullptr;
  c.idout = nullptr;
  c.wtab = nullptr;
  // The recording's rows are rec_B_cap_ apart; a context of another B_cap
  // (a later Reserve) must not index them.
  c.rec = ((record_ || rec_valid_) && rec_B_cap_ == B_cap_) ? d_rec_ : nullptr;
  c.rec_cap = rec_cap_;
  return c;
}

size_t Engine::LdsBinTime(bool mixed) const { return mixed ? lds_bin_time32_ : lds_bin_time_; }
size_t Engine::LdsFft() const { return lds_fft_; }
int32_t Engine::TilesCap() const { return trials_cap_ * tiles_cap_; }
int32_t Engine::IdentifyBlocks() const {
  const int32_t waves = kThreads / 64;  // one wavefront per bin
  return std::max<int32_t>(1, (B_cap_ + waves - 1) / waves);
}

// The bin FFT of every B the current call may use: k_bin_fft for
// B <= kMaxLdsBins, the four-step kernels above (unless big is false); each
// exits at once for an iteration whose B is the other kind.
int Engine::LaunchFft(const Ctx &c, int parity, hipStream_t s, bool big) {
  if (fft_small_) HIP_OK(LaunchBinFft(c, parity, LdsFft(), s));
  if (fft_big_ && big) HIP_OK(LaunchBigFft(c, parity, s));
  return MPSFT_OK;
}

int Engine::LaunchIteration(const Ctx &c, int parity, hipStream_t s, bool big) {
  HIP_OK(LaunchBinTime(c, parity, TilesCap(), LdsBinTime(c.mixed != 0), s));
  const int rc = LaunchFft(c, parity, s, big);
  if (rc != MPSFT_OK) return rc;
  HIP_OK(LaunchIdentify(c, parity, IdentifyBlocks(), s));
  return MPSFT_OK;
}

// Captures `iters` iterations (parities 0, 1, ...) into graph_exec_[0] and,
// when the call reserves both bin-FFT kinds, the same without the four-step
// kernels into graph_exec_[1]: a solve's B never grows, so once the device
// reports B <= kMaxLdsBins the host launches the shorter graph.
int Engine::EnsureGraph(int32_t iters, hipStream_t s) {
  const int want = (fft_big_ && fft_small_) ? 2 : 1;
  const int32_t late_iters = LateGraphIters();
  if (graph_exec_[0] && graph_key_ == shape_gen_ && graph_iters_ == iters &&
      (want == 1 || graph_exec_[1]) && graph_late_B_ == late_B_ &&
      (late_B_ == 0 || graph_late_iters_ == late_iters))
    return MPSFT_OK;
  Drain();  // a graph may still run the launches a solve queued past its stop
  for (auto &e : graph_exec_) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
  const Ctx c = MakeCtx();
  for (int k = 0; k < want; ++k) {
    hipStream_t cs;
    HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return MPSFT_ERR_HIP;
    }
    int rc = MPSFT_OK;
    // iters is even: the graph starts at an even iteration, so node i runs
    // with parity i & 1.
    for (int32_t i = 0; i < iters && rc == MPSFT_OK; ++i) rc = LaunchIteration(c, i & 1, cs, k == 0);
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc != MPSFT_OK || ec != hipSuccess) return MPSFT_ERR_HIP;
    const hipError_t ei = hipGraphInstantiate(&graph_exec_[k], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      graph_exec_[k] = nullptr;
      return MPSFT_ERR_HIP;
    }
  }
  if (late_B_ > 0) {
    // The late graphs: launches sized for B = late_B_ only.
    const BParams &bp = hbtab_[late_B_];
    const int32_t waves = kThreads / 64;
    const int32_t nb_ident = std::max<int32_t>(1, (late_B_ + waves - 1) / waves);
    const int32_t tiles = trials_cap_ * ((late_B_ + bp.R - 1) / bp.R) * bp.slices;
    // the late B's bin FFT on its own path: the small one's kernel and LDS
    // (array M and M/2 twiddles) where it applies
    const bool small_fft = bp.logM <= kFftSmallLogM;
    const size_t lds_fft = small_fft ? size_t(bp.M + bp.M / 2) * sizeof(double2) : LdsFft();
    for (int k = 2; k < 4; ++k) {
      hipStream_t cs;
      HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      hipGraph_t g;
      if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return MPSFT_ERR_HIP;
      }
      hipError_t el = hipSuccess;
      for (int32_t i = 0; i < late_iters && el == hipSuccess; ++i)
        el = LaunchLateIteration(c, i & 1, nb_ident, tiles, LdsBinTime(c.mixed != 0), lds_fft,
                                 (k == 3 || i > 0) ? 1 : 0, small_fft ? 1 : 0, cs);
      const hipError_t ec = hipStreamEndCapture(cs, &g);
      (void)hipStreamDestroy(cs);
      if (el != hipSuccess || ec != hipSuccess) return MPSFT_ERR_HIP;
      const hipError_t ei = hipGraphInstantiate(&graph_exec_[k], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        graph_exec_[k] = nullptr;
        return MPSFT_ERR_HIP;
      }
    }
  }
  graph_late_B_ = late_B_;
  graph_late_iters_ = late_iters;
  graph_iters_ = iters;
  graph_key_ = shape_gen_;
  (void)s;
  return MPSFT_OK;
}

// Waits for the launches a solve left queued past its stop (they only read
// and rewrite the loop state); every call that reads device state back,
// changes buffers or launches on another stream does so first.
void Engine::Drain() {
  if (!tail_pending_) return;
  (void)hipStreamSynchronize(tail_stream_);
  tail_pending_ = false;
}

int Engine::ReadState(hipStream_t s) {
  Drain();
  HIP_OK(hipMemcpyAsync(h_state_, d_state_, sizeof(DevState), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return MPSFT_OK;
}

// Everything a solve needs before its first launch: argument checks, the
// window tables of every reachable B, thresholds, buffers and the first
// transform draws; *a gets the device-side settings for k_init.
int Engine::PrepareSolve(const void *x, const mpsft_solve_options &opt, uint32_t seed,
                         hipStream_t s, InitArgs *a, int32_t first_draws) {
  if (status_ != MPSFT_OK) return status_;
  if (!x || opt.trials <= 0 || opt.trials % 2 != 1 || opt.trials > kMaxTrials ||
      !(opt.window_threshold > 0) || opt.k < 0 || opt.min_bins <= 0 || opt.min_bins % 2 != 1 ||
      opt.max_stale_iter <= 0 || !(opt.window_delta > 0) || !(opt.window_delta < 1) ||
      opt.sigma < 0)
    return MPSFT_ERR_INVALID_ARG;
  const int64_t want0 = 2 * int64_t(opt.k) + 1;
  const int32_t B0 = int32_t(std::max<int64_t>(want0, opt.min_bins));
  if (B0 > kMaxBins) return MPSFT_ERR_UNSUPPORTED;
  A_override_ = nullptr;
  int rc = MPSFT_OK;
  // Tables, thresholds and buffers for these options are in place when
  // nothing changed them since the last solve with the same options.
  const bool same = prep_valid_ && prep_gen_ == shape_gen_ && prep_epoch_ == tab_epoch_ &&
                    std::memcmp(&prep_opt_, &opt, sizeof(opt)) == 0;
  if (!same) {
    std::vector<int32_t> bs;
    for (int32_t B = opt.min_bins; B <= B0; B += 2) bs.push_back(B);
    prep_valid_ = false;
    rc = EnsureTables(bs, opt.window_delta, 0);
    if (rc != MPSFT_OK) return rc;
    rc = SetThresholds(opt.sigma);
    if (rc != MPSFT_OK) return rc;
    rc = Reserve(opt.trials, bs);
    if (rc != MPSFT_OK) return rc;
    prep_valid_ = true;
    prep_opt_ = opt;
    prep_gen_ = shape_gen_;
    prep_epoch_ = tab_epoch_;
  }
  const int32_t max_iter = opt.max_iter > 0 ? opt.max_iter : 100000;
  // A few graphs' worth up front; the launch loop draws ahead in doubling
  // batches while the device runs.
  const int32_t first = first_draws > 0 ? first_draws : std::max(kFirstDraws, ChunkIters());
  rc = EnsureTransforms(seed, std::min(max_iter, first), opt.trials, s);
  if (rc != MPSFT_OK) return rc;
  *a = InitArgs{};
  a->x = x;
  a->mode = kModeSolve;
  a->B = B0;
  a->trials = opt.trials;
  a->k = opt.k;
  a->min_bins = opt.min_bins;
  a->max_stale = opt.max_stale_iter;
  a->max_iter = max_iter;
  a->bits_override = 0;
  a->win_thr = opt.window_threshold;
  return MPSFT_OK;
}

// The result of a finished solve from the device state.
int Engine::SolveResult(const mpsft_solve_options &opt, int32_t *iters) {
  const bool stopped = (h_state_->L[0].done | h_state_->L[1].done) != 0;
  if (!stopped) {
    // Only when max_iter stops the launches before the device saw it.
    *iters = std::max(h_state_->L[0].it, h_state_->L[1].it) + 1;
    return MPSFT_ERR_NOT_CONVERGED;
  }
  *iters = h_state_->iters;
  if (h_state_->flags & kFlagOverflow) return MPSFT_ERR_CAPACITY;
  if (h_state_->stop_stale < opt.max_stale_iter) return MPSFT_ERR_NOT_CONVERGED;
  return MPSFT_OK;
}

int Engine::Solve(const void *x, const mpsft_solve_options &opt, uint32_t seed,
                  int32_t *iters, hipStream_t s, double *kernel_ms) {
  if (tail_stream_ != s || kernel_ms) Drain();
  rec_valid_ = false;
  InitArgs a;
  int rc = PrepareSolve(x, opt, seed, s, &a);
  if (rc != MPSFT_OK) return rc;
  if (record_ && (rc = EnsureRecord()) != MPSFT_OK) return rc;
  const int32_t max_iter = a.max_iter;
  const int32_t chunk = ChunkIters();
  Ctx c = MakeCtx();
  // The previous solve has finished; the device writes these again only
  // after k_reset below (which clears them too).
  for (int i = 0; i < 6; ++i) h_progress_[i] = 0;
  HIP_OK(LaunchReset(c, a, s));

  int32_t launched = 0;
  if (kernel_ms) {
    // Instrumented: one iteration at a time, each kernel bracketed by
    // events on the stream it runs on.  The gather-fold and the bin FFT only
    // read what the previous kernels left (both rewrite their whole output),
    // so each runs kTimedRepeats times back to back and the mean per launch
    // is taken: the events' own overhead (several us) then stays out of
    // the per-launch figure.  k_identify inserts modes and runs once.
    constexpr int kTimedRepeats = 8;
    hipEvent_t ev[4];
    for (auto &e : ev) HIP_OK(hipEventCreate(&e));
    for (int i = 0; i < 3; ++i) kernel_ms[i] = 0.0;
    int out = MPSFT_OK;
    for (; launched < max_iter; ++launched) {
      if ((out = EnsureTransforms(seed, std::min(max_iter, launched + 2), opt.trials, s)) !=
          MPSFT_OK)
        break;
      c = MakeCtx();
      const int par = launched & 1;
      bool ok = hipEventRecord(ev[0], s) == hipSuccess;
      for (int r = 0; r < kTimedRepeats && ok; ++r)
        ok = LaunchBinTime(c, par, TilesCap(), LdsBinTime(c.mixed != 0), s) == hipSuccess;
      ok = ok && hipEventRecord(ev[1], s) == hipSuccess;
      for (int r = 0; r < kTimedRepeats && ok; ++r) ok = LaunchFft(c, par, s) == MPSFT_OK;
      ok = ok && hipEventRecord(ev[2], s) == hipSuccess &&
           LaunchIdentify(c, par, IdentifyBlocks(), s) == hipSuccess &&
           hipEventRecord(ev[3], s) == hipSuccess;
      if (!ok) {
        out = MPSFT_ERR_HIP;
        break;
      }
      if ((out = ReadState(s)) != MPSFT_OK) break;
      if (h_state_->L[par].done) break;  // the loop had stopped before this iteration
      for (int i = 0; i < 3; ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) out = MPSFT_ERR_HIP;
        kernel_ms[i] += i < 2 ? ms / kTimedRepeats : ms;
      }
      if (out != MPSFT_OK) break;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (out != MPSFT_OK) {
      (void)hipStreamSynchronize(s);
      return out;
    }
  } else {
    // One graph per two iterations, launched ahead of the device by at most
    // `lookahead` iterations.  The first kernel of every iteration reports
    // {iterations done, stopped} in pinned host memory, so the host paces
    // launches without waiting on the stream and stops once the loop stops.
    // Every return below first waits for what was launched: the kernels
    // read and write this engine's buffers and pinned counters.
    auto fail = [&](int code) {
      (void)hipStreamSynchronize(s);
      return code;
    };
    // Late graphs when the last B of the solve takes the LDS bin FFT.
    late_B_ = (opt.min_bins <= kMaxLdsBins && hbtab_[opt.min_bins].logM <= kMaxLogM)
                  ? opt.min_bins : 0;
    rc = EnsureGraph(2, s);
    if (rc != MPSFT_OK) return fail(rc);
    const int32_t lookahead = chunk < 2 ? 2 : chunk;
    const int32_t late_lookahead = std::max(lookahead, LateLookahead());
    // 0: three-launch iterations; 1: the next graph opens the late form
    // (the iterations launched so far have their identify); 2: late form.
    int late = 0;
    const int32_t first_B = int32_t(std::max<int64_t>(2 * int64_t(opt.k) + 1, opt.min_bins));
    const auto stall = StallWait();
    auto last_change = std::chrono::steady_clock::now();
    int32_t seen = 0, synced = 0;
    bool done = false;
    for (uint64_t spin = 0; !done; ++spin) {
      const int32_t hp0 = h_progress_[0];
      const int32_t progressed = hp0 > synced ? hp0 : synced;
      // Launch up to iteration max_iter: its first kernel sees the cap and
      // stops the loop (logging the last iteration).
      if (h_progress_[1] || launched > max_iter) break;
      if (launched - progressed < (late == 2 ? late_lookahead : lookahead)) {
        const int32_t need = std::min(max_iter, launched + graph_late_iters_ + 1);
        if (need > tf_count_) {
          // Draw ahead in doubling batches.  Growing the arrays moves them,
          // and the graph holds their addresses: drain the stream first, so
          // no launched graph can read the freed arrays, then capture again.
          const int32_t want = std::min(max_iter, std::max(need, 2 * tf_count_));
          if (want > tf_cap_ && hipStreamSynchronize(s) != hipSuccess) return fail(MPSFT_ERR_HIP);
          rc = EnsureTransforms(seed, want, opt.trials, s);
          if (rc == MPSFT_OK) rc = EnsureGraph(2, s);
          if (rc != MPSFT_OK) return fail(rc);
        }
        // progress[2] is the B of the latest iteration the device started;
        // B never grows, so once it is min_bins every later iteration has it.
        // Before the first report it is the first iteration's B, 2k+1 but at
        // least min_bins (NextIteration with nothing found).
        const int32_t bnow = h_progress_[2] > 0 ? h_progress_[2] : first_B;
        if (late == 0 && late_B_ > 0 && bnow == late_B_) late = 1;
        hipGraphExec_t ge = (graph_exec_[1] && bnow > 0 && bnow <= kMaxLdsBins) ? graph_exec_[1]
                                                                                 : graph_exec_[0];
        if (late == 1) {
          ge = graph_exec_[2];
          late = 2;
        } else if (late == 2) {
          ge = graph_exec_[3];
        }
        if (hipGraphLaunch(ge, s) != hipSuccess) return fail(MPSFT_ERR_HIP);
        launched += late == 2 ? graph_late_iters_ : 2;
        continue;
      }
      if (progressed != seen) {
        seen = progressed;
        last_change = std::chrono::steady_clock::now();
      } else if ((spin & 4095) == 0 &&
                 std::chrono::steady_clock::now() - last_change > stall) {
        // No progress report for a while (a long iteration): wait on the
        // stream instead of spinning, then continue from the device state.
        rc = ReadState(s);
        if (rc != MPSFT_OK) return fail(rc);
        synced = std::max(h_state_->L[0].it, h_state_->L[1].it);
        done = (h_state_->L[0].done | h_state_->L[1].done) != 0;
        last_change = std::chrono::steady_clock::now();
      }
    }
    if (h_progress_[1]) {
      // The device reported the stop and its result in pinned memory.  The
      // launches queued past the stop exit at once without touching the
      // results; the next call that needs them idle waits (Drain).
      tail_pending_ = true;
      tail_stream_ = s;
      h_state_->L[0].done = 1;
      h_state_->iters = h_progress_[3];
      h_state_->stop_stale = h_progress_[4];
      h_state_->flags = h_progress_[5];
    } else {
      rc = ReadState(s);
      if (rc != MPSFT_OK) return fail(rc);
    }
  }
  rc = SolveResult(opt, iters);
  FinishRecord(opt, seed, rc, *iters, (h_state_->L[0].done | h_state_->L[1].done) != 0);
  return rc;
}

// stopped: the device stopped the loop, so every iteration's bins are logged.
void Engine::FinishRecord(const mpsft_solve_options &opt, uint32_t seed, int rc, int32_t iters,
                          bool stopped) {
  rec_valid_ = record_ && d_rec_ && stopped &&
               (rc == MPSFT_OK || rc == MPSFT_ERR_NOT_CONVERGED) && iters <= rec_cap_;
  if (!rec_valid_) return;
  ++rec_id_;
  rec_iters_ = iters;
  rec_seed_ = seed;
  rec_opt_ = opt;
}

// The recording buffer: rec_cap_ iterations of B_cap_ bins.
int Engine::EnsureRecord() {
  if (d_rec_ && rec_B_cap_ == B_cap_) return MPSFT_OK;
  rec_cap_ = 4096;
  if (Alloc(&d_rec_, size_t(rec_cap_) * B_cap_) != MPSFT_OK) return MPSFT_ERR_NOMEM;
  rec_B_cap_ = B_cap_;
  ++shape_gen_;  // captured graphs hold the pointer
  return MPSFT_OK;
}

int Engine::SetRecord(bool on) {
  if (status_ != MPSFT_OK) return status_;
  Drain();
  if (record_ != on) ++shape_gen_;
  record_ = on;
  return MPSFT_OK;
}

// The last recorded solve again, on x, with its decisions fixed: the same
// transforms, bins and identified k1's, so the coefficients are a linear
// function of x.  One launch sequence per iteration, no graph.
int Engine::Replay(const void *x, int precision, hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  Drain();
  if (!x) return MPSFT_ERR_INVALID_ARG;
  if (precision != MPSFT_FP64 && precision != MPSFT_MIXED) return MPSFT_ERR_UNSUPPORTED;
  if (!rec_valid_ || !d_rec_ || rec_B_cap_ != B_cap_) return MPSFT_ERR_SHAPE;
  int rc = EnsureTransforms(rec_seed_, rec_iters_ + 1, rec_opt_.trials, s);
  if (rc != MPSFT_OK) return rc;
  Ctx c = MakeCtx();
  c.mixed = precision == MPSFT_MIXED ? 1 : 0;
  HIP_OK(LaunchClearModes(c, 0, s));
  InitArgs a{};
  a.x = x;
  a.mode = kModeReplay;
  a.B = rec_opt_.min_bins;
  a.trials = rec_opt_.trials;
  a.k = rec_opt_.k;
  a.min_bins = rec_opt_.min_bins;
  a.max_stale = rec_opt_.max_stale_iter;
  a.max_iter = rec_iters_;
  a.win_thr = rec_opt_.window_threshold;
  HIP_OK(LaunchInit(c, a, s));
  for (int32_t it = 0; it <= rec_iters_; ++it) {
    rc = LaunchIteration(c, it & 1, s, fft_big_);
    if (rc != MPSFT_OK) return rc;
  }
  return ReadState(s);
}

int Engine::EnsureAdjBuffers(int32_t iters) {
  if (!d_lam_ && Alloc(&d_lam_, size_t(n_)) != MPSFT_OK) return MPSFT_ERR_NOMEM;
  const size_t want = size_t(std::max(iters, 1)) * trials_cap_ * rows_cap_ * B_cap_;
  if (want > fbar_len_) {
    if (Alloc(&d_fbar_, want) != MPSFT_OK) {
      fbar_len_ = 0;
      return MPSFT_ERR_NOMEM;
    }
    fbar_len_ = want;
  }
  int64_t stride = 1;
  for (size_t B = 0; B < hbtab_.size(); ++B)
    if (built_[B]) stride = std::max<int64_t>(stride, int64_t(hbtab_[B].f) * int64_t(B));
  taps_stride_ = stride;
  const size_t twant = size_t(std::max(iters, 1)) * size_t(stride);
  if (twant > taps_len_) {
    if (Alloc(&d_taps_, twant) != MPSFT_OK) {
      taps_len_ = 0;
      return MPSFT_ERR_NOMEM;
    }
    taps_len_ = twant;
  }
  // The per-lag contributions k_adj_lags tabulates for the gather (n a
  // power of two <= 2^30 only; the general gather computes them per sample).
  if (mask_ >= 0 && n_ <= (int64_t(1) << 30)) {
    lags_stride_ = int64_t(rows_cap_ + 2) * stride;
    const size_t lwant = size_t(std::max(iters, 1)) * trials_cap_ * size_t(lags_stride_);
    if (lwant > lags_len_) {
      if (Alloc(&d_lags_, lwant) != MPSFT_OK) {
        lags_len_ = 0;
        return MPSFT_ERR_NOMEM;
      }
      lags_len_ = lwant;
    }
  }
  return MPSFT_OK;
}

int Engine::Adjoint(const int64_t *keys, const void *grad, int64_t m, void *out, int precision,
                    hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  Drain();
  if (!out || m < 0 || (m > 0 && (!grad || !keys))) return MPSFT_ERR_INVALID_ARG;
  if (precision != MPSFT_FP64 && precision != MPSFT_MIXED) return MPSFT_ERR_UNSUPPORTED;
  if (!rec_valid_ || !d_rec_ || rec_B_cap_ != B_cap_) return MPSFT_ERR_SHAPE;
  if (fft_big_) return MPSFT_ERR_UNSUPPORTED;  // the adjoint DFT is k_bin_fft's (B <= kMaxLdsBins)
  int rc = EnsureTransforms(rec_seed_, rec_iters_ + 1, rec_opt_.trials, s);
  if (rc != MPSFT_OK) return rc;
  if ((rc = EnsureAdjBuffers(rec_iters_)) != MPSFT_OK) return rc;
  if (!d_ctx1_ && (Alloc(&d_ctx1_, 1) != MPSFT_OK || Alloc(&d_sig1_, 1) != MPSFT_OK))
    return MPSFT_ERR_NOMEM;
  if (rec_iters_ > steps1_cap_) {
    if (Alloc(&d_steps1_, size_t(rec_iters_)) != MPSFT_OK) return MPSFT_ERR_NOMEM;
    steps1_cap_ = rec_iters_;
  }
  const Ctx c = MakeCtx();
  AdjSig g{};
  g.lam = d_lam_;
  g.fbar = d_fbar_;
  g.taps = d_taps_;
  g.taps_stride = taps_stride_;
  g.lags = lags_len_ ? d_lags_ : nullptr;
  g.lags_stride = lags_stride_;
  g.rec = d_rec_;
  g.keys = keys;
  g.grad = static_cast<const double2 *>(grad);
  g.out = out;
  g.n = n_;
  g.m = m;
  g.iters = rec_iters_;
  g.trials = rec_opt_.trials;
  HIP_OK(hipMemcpyAsync(d_ctx1_, &c, sizeof(Ctx), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(d_sig1_, &g, sizeof(AdjSig), hipMemcpyHostToDevice, s));
  AdjLaunch a{};
  a.ctxs = d_ctx1_;
  a.sig = d_sig1_;
  a.steps = d_steps1_;
  a.count = 1;
  a.max_steps = rec_iters_;
  a.nb_bins = IdentifyBlocks();
  a.nb_fft = kPrepBlocks + trials_cap_ * rows_cap_;
  a.lds_fft = LdsFft();
  a.per_small = 16;
  a.per_gather = int32_t(std::min<int64_t>(4096, (n_ + 1023) / 1024));
  a.per_lags = LagBlocks(lags_stride_);
  a.fft_small_only = M_cap_ <= (1 << kFftSmallLogM) ? 1 : 0;
  a.any_tab = g.lags ? 1 : 0;
  a.any_gen = g.lags ? 0 : 1;
  a.mixed = precision == MPSFT_MIXED ? 1 : 0;
  HIP_OK(LaunchAdj(a, s));
  // c and g are read by the copies above; the output is the caller's next read
  HIP_OK(hipStreamSynchronize(s));
  return MPSFT_OK;
}

int Engine::IterateOnce(const void *x, const mpsft_iterate_options &opt, const DevTf &tf,
                        const int64_t *qs, int32_t *found, hipStream_t s) {
  Drain();
  rec_valid_ = false;  // the recorded solve's map or tables change
  if (status_ != MPSFT_OK) return status_;
  if (!x || !qs || opt.trials <= 0 || opt.trials % 2 != 1 || opt.trials > kMaxTrials ||
      !(opt.window_threshold > 0) ||
      opt.bins <= 0 || opt.bins % 2 != 1 || !(opt.window_delta > 0))
    return MPSFT_ERR_INVALID_ARG;
  const int want_type = precision_ == MPSFT_FP64 ? 5 : 6;
  if (opt.bin_in_time_type != want_type || opt.bin_in_freq_type != want_type)
    return MPSFT_ERR_UNSUPPORTED;
  if (opt.bins > kMaxBins) return MPSFT_ERR_UNSUPPORTED;
  A_override_ = nullptr;
  std::vector<int32_t> bs{opt.bins};
  int rc = EnsureTables(bs, opt.window_delta, 0);
  if (rc != MPSFT_OK) return rc;
  rc = Reserve(opt.trials, bs);
  if (rc != MPSFT_OK) return rc;
  bool grown = false;
  rc = GrowTransforms(1, &grown);
  if (rc != MPSFT_OK) return rc;
  tf_valid_ = false;
  std::vector<int64_t> hq(trials_cap_, 0);
  for (int32_t t = 0; t < opt.trials; ++t) hq[t] = PosMod(qs[t], n_);
  HIP_OK(hipMemcpyAsync(d_tfs_, &tf, sizeof(DevTf), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(d_qs_, hq.data(), hq.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
  Ctx c = MakeCtx();
  InitArgs a{};
  a.x = x;
  a.mode = kModeIterate;
  a.B = opt.bins;
  a.trials = opt.trials;
  a.max_stale = 1;
  a.max_iter = 1;
  a.keep_modes = 1;
  a.bin_thr = opt.bin_threshold;
  a.sv_thr = opt.sv_threshold;
  a.win_thr = opt.window_threshold;
  HIP_OK(LaunchInit(c, a, s));
  rc = LaunchIteration(c, 0, s);
  if (rc != MPSFT_OK) return rc;
  rc = ReadState(s);
  if (rc != MPSFT_OK) return rc;
  *found = h_state_->found_iter[0];
  if (h_state_->flags & kFlagOverflow) return MPSFT_ERR_CAPACITY;
  return MPSFT_OK;
}

int Engine::BinOnly(const void *x, int32_t B, double delta, int32_t bits, const DevTf &tf,
                    int64_t q, double2 *out, hipStream_t s) {
  Drain();
  rec_valid_ = false;  // the recorded solve's map or tables change
  if (status_ != MPSFT_OK) return status_;
  if (!x || !out || bits < 1 || bits > 30) return MPSFT_ERR_INVALID_ARG;
  std::vector<int32_t> bs{B};
  int rc = EnsureTables(bs, delta, bits);
  if (rc != MPSFT_OK) return rc;
  rc = Reserve(1, bs);
  if (rc != MPSFT_OK) return rc;
  // out's layout is rows x B: the engine's row stride must be B (a larger
  // trials_cap_ from an earlier IdentifyOnly leaves trial 0's rows alone)
  if (B_cap_ != B || rows_cap_ != 1 + 2 * bits) return MPSFT_ERR_SHAPE;
  bool grown = false;
  rc = GrowTransforms(1, &grown);
  if (rc != MPSFT_OK) return rc;
  tf_valid_ = false;
  const int64_t hq = PosMod(q, n_);
  HIP_OK(hipMemcpyAsync(d_tfs_, &tf, sizeof(DevTf), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(d_qs_, &hq, sizeof(int64_t), hipMemcpyHostToDevice, s));
  A_override_ = out;
  Ctx c = MakeCtx();
  if (d_wtab_ && wtab_B_ == B) c.wtab = d_wtab_;
  InitArgs a{};
  a.x = x;
  a.mode = kModeBinOnly;
  a.B = B;
  a.trials = 1;
  a.max_stale = 1;
  a.max_iter = 1;
  a.bits_override = bits;
  a.win_thr = 1;
  HIP_OK(LaunchInit(c, a, s));
  HIP_OK(LaunchBinTime(c, 0, TilesCap(), LdsBinTime(c.mixed != 0), s));
  rc = LaunchFft(c, 0, s);
  A_override_ = nullptr;
  if (rc != MPSFT_OK) return rc;
  return MPSFT_OK;
}

int Engine::SetWindowTaps(int32_t B, int32_t P, const double *wt_host) {
  if (status_ != MPSFT_OK) return status_;
  if (!wt_host || P <= 0) return MPSFT_ERR_INVALID_ARG;
  Drain();
  if (Alloc(&d_wtab_, size_t(P)) != MPSFT_OK) return MPSFT_ERR_NOMEM;
  HIP_OK(hipMemcpy(d_wtab_, wt_host, size_t(P) * sizeof(double), hipMemcpyHostToDevice));
  wtab_B_ = B;
  return MPSFT_OK;
}

int Engine::IdentifyOnly(const double2 *A, int32_t trials, int32_t B, double delta, int32_t bits,
                         const DevTf &tf, const int64_t *qs, double bin_thr, double sv_thr,
                         double win_thr, const IdOut &out, hipStream_t s) {
  Drain();
  rec_valid_ = false;
  if (status_ != MPSFT_OK) return status_;
  if (!A || !qs || trials <= 0 || trials % 2 != 1 || trials > kMaxTrials || bits < 1 ||
      bits > 30 || !(win_thr > 0) || !out.decide || !out.k1 || !out.k0 || !out.coef)
    return MPSFT_ERR_INVALID_ARG;
  std::vector<int32_t> bs{B};
  int rc = EnsureTables(bs, delta, bits);
  if (rc != MPSFT_OK) return rc;
  rc = Reserve(trials, bs);
  if (rc != MPSFT_OK) return rc;
  // A's layout is [trial][row][bin] with rows x B per trial: the engine's
  // strides must be exactly those.
  if (B_cap_ != B || rows_cap_ != 1 + 2 * bits) return MPSFT_ERR_SHAPE;
  bool grown = false;
  rc = GrowTransforms(1, &grown);
  if (rc != MPSFT_OK) return rc;
  tf_valid_ = false;
  std::vector<int64_t> hq(trials_cap_, 0);
  for (int32_t t = 0; t < trials; ++t) hq[t] = PosMod(qs[t], n_);
  if (!d_idout_ && Alloc(&d_idout_, 1) != MPSFT_OK) return MPSFT_ERR_NOMEM;
  HIP_OK(hipMemcpyAsync(d_tfs_, &tf, sizeof(DevTf), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(d_qs_, hq.data(), hq.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(d_idout_, &out, sizeof(IdOut), hipMemcpyHostToDevice, s));
  A_override_ = const_cast<double2 *>(A);
  Ctx c = MakeCtx();
  A_override_ = nullptr;
  c.idout = d_idout_;
  InitArgs a{};
  a.x = A;  // not read: no gather runs
  a.mode = kModeIterate;
  a.B = B;
  a.trials = trials;
  a.max_stale = 1;
  a.max_iter = 1;
  a.bits_override = bits;
  a.bin_thr = bin_thr;
  a.sv_thr = sv_thr;
  a.win_thr = win_thr;
  HIP_OK(LaunchReset(c, a, s));
  HIP_OK(LaunchPublish(c, 0, s));
  HIP_OK(LaunchIdentify(c, 0, IdentifyBlocks(), s));
  // the host copies of out and the q's are read by the copies above
  HIP_OK(hipStreamSynchronize(s));
  return MPSFT_OK;
}

int Engine::PostAnalyzeDev(const int64_t *truth, int64_t m, const double2 *xhat, double *partial,
                           int32_t blocks, double *stats_dev, hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  if (m < 0 || (m > 0 && !truth) || !xhat || !partial || !stats_dev || blocks < 1)
    return MPSFT_ERR_INVALID_ARG;
  if (tail_stream_ != s) Drain();  // a solve's queued launches run before, in stream order
  HIP_OK(LaunchPostAnalyze(MakeCtx(), truth, m, xhat, partial, blocks, stats_dev, s));
  return MPSFT_OK;
}

int Engine::PostAnalyze(const int64_t *truth, int64_t m, const double2 *xhat, double *stats,
                        hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  if (!stats) return MPSFT_ERR_INVALID_ARG;
  if (!d_pa_ && Alloc(&d_pa_, size_t(kPaBlocks + 1) * kPaFields) != MPSFT_OK)
    return MPSFT_ERR_NOMEM;
  double *dstats = d_pa_ + size_t(kPaBlocks) * kPaFields;
  int rc = PostAnalyzeDev(truth, m, xhat, d_pa_, kPaBlocks, dstats, s);
  if (rc != MPSFT_OK) return rc;
  double h[kPaFields];
  HIP_OK(hipMemcpyAsync(h, dstats, sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (h[9] != 0) return MPSFT_ERR_INVALID_ARG;  // truth keys not ascending
  for (int i = 0; i < 9; ++i) stats[i] = h[i];
  return MPSFT_OK;
}

int Engine::ModesCount(int64_t *count) {
  const int rc = ReadState(nullptr);
  if (rc != MPSFT_OK) return rc;
  *count = std::min<int64_t>(h_state_->count, cap_);
  return MPSFT_OK;
}

int Engine::ModesGet(int64_t *keys, double *coefs, int64_t cap, int64_t *count) {
  int64_t m = 0;
  int rc = ModesCount(&m);
  if (rc != MPSFT_OK) return rc;
  *count = m;
  if (cap < m) return MPSFT_ERR_SHAPE;
  if (m == 0) return MPSFT_OK;
  Ctx c = MakeCtx();
  HIP_OK(LaunchGetCoefs(c, d_tmp_, m, nullptr));
  HIP_OK(hipMemcpy(keys, d_keys_, size_t(m) * sizeof(int64_t), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(coefs, d_tmp_, size_t(m) * sizeof(double2), hipMemcpyDeviceToHost));
  return MPSFT_OK;
}

int Engine::ModesSet(const int64_t *keys, const double *coefs, int64_t m) {
  Drain();
  rec_valid_ = false;  // the recorded solve's map or tables change
  if (m < 0 || m > cap_ || (m > 0 && (!keys || !coefs))) return MPSFT_ERR_INVALID_ARG;
  std::vector<int64_t> hk(keys, keys + m);
  for (auto &k : hk) k = PosMod(k, n_);
  // A ModeMap (src/base.h:35) holds each location once: two keys equal
  // modulo n are refused rather than entered twice.
  std::vector<int64_t> sorted(hk);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    return MPSFT_ERR_INVALID_ARG;
  Ctx c = MakeCtx();
  HIP_OK(LaunchClearModes(c, 0, nullptr));
  if (m > 0) {
    HIP_OK(hipMemcpy(d_tmpk_, hk.data(), size_t(m) * sizeof(int64_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_tmp_, coefs, size_t(m) * sizeof(double2), hipMemcpyHostToDevice));
  }
  HIP_OK(LaunchSetModes(c, d_tmpk_, d_tmp_, m, nullptr));
  HIP_OK(hipDeviceSynchronize());
  return MPSFT_OK;
}

int Engine::BinsGet(double *out, int64_t cap, int32_t *trials, int32_t *rows, int32_t *bins) {
  int rc = ReadState(nullptr);
  if (rc != MPSFT_OK) return rc;
  const int32_t T = h_state_->trials, R = h_state_->P[0].rows, B = h_state_->P[0].B;
  *trials = T;
  *rows = R;
  *bins = B;
  if (!d_A_ || T > trials_cap_ || R > rows_cap_ || B > B_cap_) return MPSFT_ERR_SHAPE;
  if (cap < int64_t(2) * T * R * B) return MPSFT_ERR_SHAPE;
  for (int32_t t = 0; t < T; ++t)
    HIP_OK(hipMemcpy2D(out + size_t(2) * t * R * B, size_t(B) * sizeof(double2),
                       d_A_ + size_t(t) * rows_cap_ * B_cap_, size_t(B_cap_) * sizeof(double2),
                       size_t(B) * sizeof(double2), R, hipMemcpyDeviceToHost));
  return MPSFT_OK;
}

int Engine::LogGet(int32_t *bins, int32_t *found, int32_t cap, int32_t *n_iter) {
  int rc = ReadState(nullptr);
  if (rc != MPSFT_OK) return rc;
  const int32_t it = std::min(h_state_->iters, log_cap_);
  *n_iter = it;
  const int32_t m = std::min(it, cap);
  if (m > 0) {
    if (bins) HIP_OK(hipMemcpy(bins, d_log_, size_t(m) * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (found)
      HIP_OK(hipMemcpy(found, d_log_ + log_cap_, size_t(m) * sizeof(int32_t),
                       hipMemcpyDeviceToHost));
  }
  return MPSFT_OK;
}

}  // namespace mpsft

namespace mpsft {

int Engine::TransformsGet(DevTf *tfs, int64_t *qs, int32_t cap, int32_t trials, int32_t *count) {
  int rc = ReadState(nullptr);
  if (rc != MPSFT_OK) return rc;
  if (!tf_valid_ || trials != tf_trials_) {
    *count = 0;
    return MPSFT_OK;
  }
  const int32_t m = std::min(std::min(h_state_->iters, tf_count_), cap);
  for (int32_t i = 0; i < m; ++i) {
    tfs[i] = h_tfs_[i];
    for (int32_t t = 0; t < trials; ++t) qs[int64_t(i) * trials + t] = h_qs_[size_t(i) * trials_cap_ + t];
  }
  *count = m;
  return MPSFT_OK;
}

}  // namespace mpsft
