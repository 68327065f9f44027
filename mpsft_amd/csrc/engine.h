// engine.h -- device data layout and host engine of the MPSFT recovery loop.
//
// One Engine holds the device state of one signal length n.  An iteration
// of Iterate (src/iterate.cc:80-177) is three kernels on one stream:
//   k_bin_time  : derives the iteration's parameters from the loop state
//                 (Demo1::Run, src/demo1.cc:36-59; workgroup 0 publishes
//                 them), then the permuted, windowed gather + fold
//                 (BinInTime without its B-point DFT), plus the per-mode
//                 terms BinInFreq needs;
//   k_bin_fft   : per (trial, row), the B-point forward DFT (Bluestein over
//                 an LDS-resident power-of-two FFT), the row phase, and the
//                 BinInFreq subtraction; for B > kMaxLdsBins the same
//                 Bluestein DFT runs as a four-step FFT in three kernels
//                 (k_big_cols, k_big_rows, k_big_out) through a scratch
//                 buffer that stays in cache;
//   k_identify  : per bin, the energy test, the matrix-pencil votes
//                 (IdentifyFreq), the coefficient estimate and the insert
//                 into the found-mode map.
// Every kernel reads the iteration's parameters (B, bits, transform, q)
// from DevState in device memory, so one fixed launch sequence is captured
// in a hipGraph and replayed for any iteration.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/mpsft.h"
#include "mt.h"

namespace mpsft {

struct InitArgs;

constexpr int kThreads = 256;      // workgroup size of most kernels
constexpr int kFftThreads = 256;   // workgroup size of k_bin_fft
constexpr int kMaxLogM = 12;       // k_bin_fft (one LDS FFT) up to M = 4096, B <= 2048
constexpr int kMaxLdsBins = 2048;  // largest B of k_bin_fft
constexpr int kFftSmallLogM = 10;  // k_bin_fft's in-place radix-4 path up to M = 1024
constexpr int kMaxLogMBig = 16;    // four-step Bluestein FFT up to M = 65536
constexpr int kTwLogM = kMaxLogMBig;  // twiddle table e^{-2 pi i j / 2^kTwLogM}
constexpr int kBigTile = 1024;     // elements of one four-step workgroup
constexpr int kSlices = 8;         // XCDs of MI355X: partial gather-folds of an eighth of x each
constexpr int kIdentKeys = 64;     // new keys a k_identify workgroup lists before appending
constexpr int kPrepBlocks = 16;    // k_bin_time workgroups computing mode terms
constexpr int kMaxBins = 32767;    // largest B (2B-1 <= 2^kMaxLogMBig)
constexpr int kSmallTileB = 256;   // below it a B takes the narrow class's tile width
constexpr int kTargetTiles = 512;  // k_bin_time tiles aimed for per trial
constexpr int kColSlices = 4;      // k_identify holds trials*rows <= 256 entries per bin
constexpr int kMaxTrials = 15;     // trials per iteration (odd)

// Per-B window and bin-FFT parameters, computed on the host with the
// reference's formulas (src/window.cc:28-50, src/iterate.cc:30-42).
struct BParams {
  int32_t B, P, f, bits, M, logM;
  int32_t R;        // residues per k_bin_time tile
  int32_t slices;   // 1, or kSlices: the gather-fold split by eighths of x (one per XCD)
  int32_t slice_shift;  // log2(n) - 3 when slices > 1
  int64_t off;      // offset (in double2) of [chirp(B) | filter(M)] in table
                    // (B <= kMaxLdsBins only; larger B compute both on the fly)
  double width;     // 1/(2B)
  double sigma_f;   // window std-dev in frequency
  double dfreq;     // sqrt(2)*pi*sigma_t  (SampleInFreq)
  double bin_thr;   // 3*sigma/sqrt(B) in a solve
  double sv_thr;    // sigma/sqrt(B) in a solve
};

// Transform reduced modulo n.
struct DevTf {
  int64_t a, b, c, a_inv;
};

enum : int32_t { kModeSolve = 0, kModeIterate = 1, kModeBinOnly = 2, kModeReplay = 3 };
enum : int32_t { kFlagOverflow = 1 };

// Per-bin decisions of Iterate (src/iterate.cc:113-175): below the bin
// energy threshold, rejected by the pencil's sv threshold, rejected by the
// window threshold, or a mode found.
enum : int32_t { kDecideEnergy = 0, kDecideSv = 1, kDecideWindow = 2, kDecideFound = 3 };

// Where k_identify writes each bin's outcome instead of updating the
// found-mode map (device memory, B entries each; margin may be null).
struct IdOut {
  int32_t *decide;
  int64_t *k1;      // the rounded location, -1 when not reached
  int64_t *k0;      // the un-permuted location of a found mode, else -1
  double2 *coef;    // its coefficient estimate, else 0
  double *margin;   // smallest relative distance of a decision from its threshold
  __device__ void Record(int32_t b, int32_t d, int64_t k1v, double m) const {
    decide[b] = d;
    k1[b] = k1v;
    if (margin) margin[b] = m;
    if (d != kDecideFound) {
      k0[b] = -1;
      coef[b] = make_double2(0.0, 0.0);
    }
  }
};

// The parameters of one iteration.  Every k_bin_time workgroup derives them
// from the previous iteration's slot and the counters k_identify left;
// workgroup 0 publishes them in slot (iteration & 1) for the iteration's
// k_bin_fft and k_identify.
struct IterParams {
  DevTf tf;
  int64_t q[kMaxTrials];
  int64_t tab_off;
  double width, sigma_f, dfreq;  // window (src/window.cc:28-50)
  double bin_thr, sv_thr;
  int32_t B, bits, rows, f, P, p2, M, logM, R;
  int32_t slices, slice_shift;  // as BParams
  int32_t it;    // iteration index
  int32_t done;  // the loop stopped before this iteration
};

// New transforms and q's an engine drew but did not upload (a batched solve
// gathers every signal's into one copy): host sources, device destinations.
struct TfUpload {
  const DevTf *tf_src;
  DevTf *tf_dst;
  int64_t ntf;
  const int64_t *q_src;
  int64_t *q_dst;
  int64_t nq;
};

struct LoopCounters {
  int32_t it;     // iteration index (-1 before the first)
  int32_t stale;  // iterations that found nothing (src/demo1.cc:55-57)
  int32_t done;
  int32_t pad;
};

// Device-resident loop state.
struct DevState {
  const void *x;  // input signal (double2 or float2)
  // fixed for a call
  int32_t mode, trials, k, min_bins, max_stale, max_iter, bits_override, B_fixed;
  double win_thr, bin_thr_fixed, sv_thr_fixed;
  // updated during a call
  int32_t count;          // number of distinct found keys
  int32_t flags;
  int32_t found_iter[2];  // per parity: this iteration inserted something
  int32_t iters;          // iterations run when the loop stopped
  int32_t stop_stale;     // stale count when the loop stopped
  LoopCounters L[2];
  IterParams P[2];
  int32_t cnt_base[2];  // per parity: |found| when that iteration started
  int32_t pad_[2];
};

// Everything a kernel needs; passed by value.
struct Ctx {
  DevState *st;
  const BParams *btab;   // indexed by B
  const double2 *table;  // chirps and Bluestein filters
  const double2 *tw;     // e^{-2 pi i k / 2^kTwLogM}, k < 2^(kTwLogM-1)
  const DevTf *tfs;      // per iteration
  const int64_t *qs;     // per iteration x trials_cap
  int64_t n;
  int64_t mask;          // n-1 if n is a power of two, else -1
  double2 *fold;         // kSlices partial folds, each trials_cap x rows_cap x B_cap
  int64_t fold_stride;   // entries between partial folds
  double2 *A;            // trials_cap x rows_cap x B_cap (bin coefficients)
  double2 *big;          // (trials_cap x rows_cap + 1) x M_cap: four-step FFT scratch
  int32_t *term_bin;     // trials_cap x cap: bin of each found mode
  double2 *term_val;     // trials_cap x rows_cap x cap: BinInFreq terms
  int64_t *keys;         // cap
  double2 *coef;         // n (dense found coefficients)
  uint32_t *present;     // ceil(n/32) bits
  int32_t *log;          // 2 x log_cap
  int32_t *progress;     // host-mapped {iteration started, stopped, its B, and at the
                         // stop: iterations, stale count, flags}, or null
  int32_t log_cap;
  int32_t cap;
  int32_t trials_cap, rows_cap, B_cap;
  int32_t big_gpr;       // four-step workgroups per row (M_cap / kBigTile), 0 if unused
  int32_t mixed;
  int64_t *rec;          // recorded solve: k1 of each (iteration, bin), -1 if none; or null
  int32_t rec_cap;       // iterations rec holds (B_cap entries each)
  int32_t *batch_done;   // batched form: device count of stopped signals, or null
  const IdOut *idout;    // per-bin outputs of k_identify instead of map updates, or null
  const double *wtab;    // the caller's window taps wt[i] (src/window.h:36), or null:
                         // computed on the device
};

struct AdjSig;
struct AdjStep;

class Engine {
 public:
  Engine(int64_t n, int precision, int64_t capacity);
  ~Engine();
  int status() const { return status_; }

  // kernel_ms (3 entries, or null): when given, the solve runs without the
  // graph and accumulates the time of k_bin_time, the bin FFT and
  // k_identify (the first two as the mean of repeated launches).
  int Solve(const void *x, const mpsft_solve_options &opt, uint32_t seed,
            int32_t *iters, hipStream_t s, double *kernel_ms = nullptr);
  int IterateOnce(const void *x, const mpsft_iterate_options &opt,
                  const DevTf &tf, const int64_t *qs, int32_t *found,
                  hipStream_t s);
  // BinInTime (bin_only) for a plan: out gets rows x B, row stride B.
  int BinOnly(const void *x, int32_t B, double delta, int32_t bits,
              const DevTf &tf, int64_t q, double2 *out, hipStream_t s);
  // The per-bin body of Iterate (src/iterate.cc:113-175) on a caller's bin
  // matrices A (trials x rows x B, row-major, rows = 1 + 2 bits): each bin's
  // decision, k1, k0, coefficient and margin into out (device arrays).
  int IdentifyOnly(const double2 *A, int32_t trials, int32_t B, double delta, int32_t bits,
                   const DevTf &tf, const int64_t *qs, double bin_thr, double sv_thr,
                   double win_thr, const IdOut &out, hipStream_t s);
  // Window taps given by the caller (Window::wt(), src/window.h:36) for the
  // window of B bins and P taps: BinInTime then reads them instead of
  // computing them on the device.
  int SetWindowTaps(int32_t B, int32_t P, const double *wt_host);

  // PostAnalyze (src/demo1.cc:61-112) of the found modes against the true
  // locations truth[0..m) (device, ascending) and x-hat (device, n double2):
  // stats_dev gets 10 doubles (the nine figures, then 1 if truth was not
  // ascending); partial: `blocks` x kPaFields doubles of scratch.  Writes
  // device memory only (no allocation, no synchronisation).
  int PostAnalyzeDev(const int64_t *truth, int64_t m, const double2 *xhat, double *partial,
                     int32_t blocks, double *stats_dev, hipStream_t s);
  // The same into host memory (9 figures), with scratch the engine holds.
  int PostAnalyze(const int64_t *truth, int64_t m, const double2 *xhat, double *stats,
                  hipStream_t s);

  int ModesCount(int64_t *count);
  int ModesGet(int64_t *keys, double *coefs, int64_t cap, int64_t *count);
  int ModesSet(const int64_t *keys, const double *coefs, int64_t m);
  int BinsGet(double *out, int64_t cap, int32_t *trials, int32_t *rows,
              int32_t *bins);
  int LogGet(int32_t *bins, int32_t *found, int32_t cap, int32_t *n_iter);
  int TransformsGet(DevTf *tfs, int64_t *qs, int32_t cap, int32_t trials, int32_t *count);

  int64_t n() const { return n_; }
  int precision() const { return precision_; }
  // k_bin_time tile width: residues per tile (1, 2, 4, 8), or 0 for the
  // default choice.
  int SetTileResidues(int32_t R) { return SetTileResidues(R, R); }
  int SetTileResidues(int32_t R_wide, int32_t R_narrow);

  // Recording: the next solves keep every iteration's identified bins, so
  // that Replay (the recovery map with those decisions fixed, linear in x)
  // and Adjoint (its transpose) can run on the last recorded solve.
  int SetRecord(bool on);
  // x: device, n samples of `precision` (MPSFT_FP64 or MPSFT_MIXED).
  int Replay(const void *x, int precision, hipStream_t s);
  // keys, grad: device, m keys and their complex128 gradients dL/dconj(c);
  // out: device, n samples of `precision`, gets the recovery map's adjoint
  // applied to them.
  int Adjoint(const int64_t *keys, const void *grad, int64_t m, void *out, int precision,
              hipStream_t s);
  // Identifies the held recording: a new value per recorded solve, 0 when
  // none is held.
  int64_t RecordId() const { return rec_valid_ ? rec_id_ : 0; }

 private:
  friend class Batch;
  // first_draws > 0: the iterations whose transforms are drawn before the
  // first launch (default: a few graphs' worth).
  int PrepareSolve(const void *x, const mpsft_solve_options &opt, uint32_t seed, hipStream_t s,
                   InitArgs *a, int32_t first_draws = 0);
  int SolveResult(const mpsft_solve_options &opt, int32_t *iters);
  int EnsureTables(const std::vector<int32_t> &bs, double delta, int32_t bits_override);
  int SetThresholds(double sigma);
  int Reserve(int32_t trials, const std::vector<int32_t> &bs);
  int EnsureTransforms(uint32_t seed, int32_t count, int32_t trials, hipStream_t s);
  // While set, EnsureTransforms records its upload here instead of copying
  // (the host sources stay valid until the next draw).
  TfUpload *defer_upload_ = nullptr;
  int GrowTransforms(int32_t count, bool *grown);
  Ctx MakeCtx() const;
  int LaunchIteration(const Ctx &c, int parity, hipStream_t s, bool big = true);
  int LaunchFft(const Ctx &c, int parity, hipStream_t s, bool big = true);
  int EnsureGraph(int32_t iters, hipStream_t s);
  int ReadState(hipStream_t s);
  // k_bin_time's dynamic LDS: samples staged as stored in x (complex64 for
  // a mixed launch), taps and the reduction buffer in double2.
  size_t LdsBinTime(bool mixed) const;
  size_t LdsFft() const;
  int32_t TilesCap() const;
  int32_t IdentifyBlocks() const;

  int status_ = MPSFT_OK;
  int64_t n_;
  int64_t mask_;
  int precision_;
  int32_t cap_;

  // host copy of the B table (index = B); built_[B] set when its chirp and
  // filter exist on the device.  bits_[B] may be overridden for a plan.
  std::vector<BParams> hbtab_;
  std::vector<char> built_;
  double tab_delta_ = -1;
  double tab_sigma_ = -1;
  int32_t tab_bits_override_ = 0;

  int32_t trials_cap_ = 0, rows_cap_ = 0, B_cap_ = 0;
  int32_t tiles_cap_ = 0;  // k_bin_time tiles per trial (max over B)
  int32_t M_cap_ = 0;
  bool fft_small_ = false;  // some reserved B runs k_bin_fft
  bool fft_big_ = false;    // some reserved B runs the four-step kernels
  int32_t big_M_cap_ = 0;   // largest four-step M reserved
  size_t big_len_ = 0;      // elements of d_big_
  size_t lds_bin_time_ = 0, lds_bin_time32_ = 0, lds_fft_ = 0;

  // device memory
  DevState *d_state_ = nullptr;
  BParams *d_btab_ = nullptr;
  int32_t d_btab_len_ = 0;
  double2 *d_table_ = nullptr;
  int64_t table_len_ = 0, table_used_ = 0;
  double2 *d_tw_ = nullptr;
  DevTf *d_tfs_ = nullptr;
  int64_t *d_qs_ = nullptr;
  int32_t tf_cap_ = 0;      // iterations the device arrays can hold
  int32_t tf_count_ = 0;    // iterations currently drawn for tf_seed_
  int32_t tf_trials_ = 0;
  int32_t tf_qstride_ = 0;
  bool tf_valid_ = false;
  uint32_t tf_seed_ = 0;
  double2 *d_fold_ = nullptr;
  double2 *d_A_ = nullptr;
  double2 *d_big_ = nullptr;
  int32_t *d_term_bin_ = nullptr;
  double2 *d_term_val_ = nullptr;
  int64_t *d_keys_ = nullptr;
  double2 *d_coef_ = nullptr;
  uint32_t *d_present_ = nullptr;
  int32_t *d_log_ = nullptr;
  int32_t log_cap_ = 0;
  DevState *h_state_ = nullptr;  // pinned mirror of d_state_
  volatile int32_t *h_progress_ = nullptr;  // pinned, written by k_bin_time
  int32_t *d_progress_ = nullptr;           // device address of h_progress_
  double2 *A_override_ = nullptr;
  double *d_wtab_ = nullptr;     // the caller's window taps (SetWindowTaps), or null
  int32_t wtab_B_ = 0;           // the window they belong to
  IdOut *d_idout_ = nullptr;     // IdentifyOnly's output pointers (device copy)
  double *d_pa_ = nullptr;       // PostAnalyze scratch: partial sums, then the stats
  double2 *d_tmp_ = nullptr;     // cap double2 staging for mode I/O
  int64_t *d_tmpk_ = nullptr;    // cap int64 staging for mode I/O

  // host RNG of the transform draws (src/base.cc:32-36): std::mt19937's
  // stream, seeded lazily (a solve draws a few dozen words of the 624)
  LazyMt19937 rng_;
  // RandomInt32 (src/base.cc:53): uniform_int_distribution<int32_t>() over
  // mt19937 as libstdc++ computed it before GCC 11: 32-bit outputs below
  // 2^31, the rest rejected.
  int32_t RandomInt32() {
    uint32_t r;
    do {
      r = uint32_t(rng_());
    } while (r >= 0x80000000u);
    return int32_t(r);
  }
  std::vector<DevTf> h_tfs_;
  std::vector<int64_t> h_qs_;
  // pinned copies of the draws (tf_cap_ and tf_cap_ x trials_cap_ entries):
  // the uploads read them asynchronously, so a solve never waits for them
  DevTf *p_tfs_ = nullptr;
  int64_t *p_qs_ = nullptr;

  // graph of graph_iters_ iterations for the current shapes
  // [1]: without the four-step FFT; [2], [3]: the late iterations at
  // B = late_B_ (two launches each), [2] opening with no identify
  hipGraphExec_t graph_exec_[4] = {nullptr, nullptr, nullptr, nullptr};
  int32_t graph_iters_ = 0;
  int64_t graph_key_ = -1;
  int32_t late_B_ = 0, graph_late_B_ = 0;  // min_bins of the solve (0: no late graphs)
  int32_t graph_late_iters_ = 2;            // iterations per late graph
  int64_t shape_gen_ = 0;
  size_t lds_fft_small_ = 0;  // k_bin_fft's LDS on its small path (largest reserved B)
  // k_bin_time tile width fixed by the autotuner (0: the default per B), for
  // B >= kSmallTileB and for B < kSmallTileB (the late iterations' B)
  int32_t forced_R_[2] = {0, 0};
  // k_bin_time tiles aimed for per trial by the default tile width: fewer
  // (wider tiles) when many signals share the launches of a batch
  int32_t tile_target_ = kTargetTiles;

  // recorded solve
  bool record_ = false;
  int64_t *d_rec_ = nullptr;  // rec_cap_ x B_cap_ k1's
  int32_t rec_cap_ = 0, rec_B_cap_ = 0;
  bool rec_valid_ = false;
  int32_t rec_iters_ = 0;
  int64_t rec_id_ = 0;

  // launches a solve left queued past its stop, on tail_stream_
  bool tail_pending_ = false;
  hipStream_t tail_stream_ = nullptr;
  void Drain();
  // PrepareSolve's key: the options, buffer shapes and table contents of
  // the last full preparation
  bool prep_valid_ = false;
  mpsft_solve_options prep_opt_{};
  int64_t prep_gen_ = -1, tab_epoch_ = 0, prep_epoch_ = -1;
  uint32_t rec_seed_ = 0;
  mpsft_solve_options rec_opt_{};
  double2 *d_lam_ = nullptr, *d_fbar_ = nullptr, *d_taps_ = nullptr, *d_lags_ = nullptr;
  size_t fbar_len_ = 0, taps_len_ = 0, lags_len_ = 0;
  int64_t taps_stride_ = 0;  // max f B over the built tables
  int64_t lags_stride_ = 0;  // (rows_cap + 2) taps_stride: one (iteration, trial)'s lag table
  // one-signal launch state of Adjoint (device): Ctx, AdjSig, steps
  Ctx *d_ctx1_ = nullptr;
  AdjSig *d_sig1_ = nullptr;
  AdjStep *d_steps1_ = nullptr;
  int32_t steps1_cap_ = 0;
  // the adjoint buffers of this signal: lam, and fbar for `iters` iterations
  int EnsureAdjBuffers(int32_t iters);
  int EnsureRecord();
  // After a solve: keeps its decisions for Replay / Adjoint when the solve
  // was recorded and its loop stopped on the device.
  void FinishRecord(const mpsft_solve_options &opt, uint32_t seed, int rc, int32_t iters,
                    bool stopped);
};

// Host-side helpers shared with the C-ABI.
int32_t NumBits(int64_t n, int32_t bins);
bool WindowParams(int64_t n, int32_t bins, double delta, BParams *bp);
// The delta of a window given by its taps (Window::wt(), p() taps of
// Window(n, bins, delta)); false when the taps are not such a window's.
bool WindowDeltaFromTaps(int64_t n, int32_t bins, int32_t p, const double *wt, double *delta);
int64_t InvMod(int64_t a, int64_t n);
DevTf ReduceTf(int64_t n, int64_t a, int64_t b, int64_t c, int64_t a_inv);
std::chrono::milliseconds StallWait();

}  // namespace mpsft
