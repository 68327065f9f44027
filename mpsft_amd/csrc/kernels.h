// kernels.h -- launch wrappers of the HIP kernels (kernels.hip).
#pragma once

#include "engine.h"

namespace mpsft {

// Initial loop state for k_init.
struct InitArgs {
  const void *x;
  int32_t mode;
  int32_t B;          // bins for Iterate / BinInTime (the solve derives its own)
  int32_t trials;
  int32_t k, min_bins, max_stale, max_iter;
  int32_t bits_override;  // > 0: use this bits instead of the table's
  int32_t keep_modes;     // 1: keep the found-mode map (Iterate)
  double bin_thr, sv_thr, win_thr;
};

hipError_t LaunchBuildTwiddles(double2 *tw, hipStream_t s);
hipError_t LaunchBuildTables(const int32_t *blist, int32_t nb, const BParams *btab,
                             double2 *table, const double2 *tw, int32_t M_cap,
                             hipStream_t s);
hipError_t LaunchClearModes(const Ctx &c, int32_t full, hipStream_t s);
hipError_t LaunchInit(const Ctx &c, const InitArgs &a, hipStream_t s);
// LaunchClearModes(c, 0) then LaunchInit in one single-workgroup launch.
hipError_t LaunchReset(const Ctx &c, const InitArgs &a, hipStream_t s);
// One late iteration of a solve (every iteration from here on has
// B = min_bins <= kMaxLdsBins): k_late_fold (the previous iteration's
// identify unless with_identify is 0, beside this one's BinInFreq factors
// and gather-fold over `tiles` tiles) then k_late_fft (the kernel of the
// small bin-FFT path alone when small_fft, with lds_fft sized for it).
hipError_t LaunchLateIteration(const Ctx &c, int parity, int32_t nb_ident, int32_t tiles,
                               size_t lds_time, size_t lds_fft, int with_identify, int small_fft,
                               hipStream_t s);
// parity = iteration & 1: the IterParams slot the kernels use.
hipError_t LaunchBinTime(const Ctx &c, int parity, int32_t tiles, size_t lds, hipStream_t s);
hipError_t LaunchBinFft(const Ctx &c, int parity, size_t lds, hipStream_t s);
// The four-step bin FFT (B > kMaxLdsBins): three kernels.
hipError_t LaunchBigFft(const Ctx &c, int parity, hipStream_t s);
hipError_t LaunchIdentify(const Ctx &c, int parity, int32_t blocks, hipStream_t s);
// Publishes the next iteration's parameters in slot `parity` (k_bin_time's
// first workgroup alone: no gather, no terms of found modes).
hipError_t LaunchPublish(const Ctx &c, int parity, hipStream_t s);

// The batched form: per-stage launches over every signal of a batch.
// BatchWork (device) is rewritten by each batched iteration's plan: per
// kind k (0 k_bin_time, 1 k_bin_fft on the small path, 2 k_identify, 3
// k_bin_fft on the large path) the units of signal i
// (cnt[k count + i] of them from off[k count + i]) and the packed map of
// {signal, unit} entries, total[k] in all.
constexpr int kUnitKinds = 4;  // gather-fold, small-path FFT, identify, large-path FFT
struct BatchWork {
  int32_t *cnt, *off;  // kUnitKinds x count each
  int32_t *rows;       // count: the iteration's rows per signal
  int32_t *total;      // kUnitKinds
  int2 *map[kUnitKinds];  // capacities: the units of every signal at its largest B
};
struct BatchLaunch {
  const Ctx *ctxs;                      // device, one per signal
  const int2 *map_time, *map_fft, *map_ident;  // device, {signal, local block} per block
  const int32_t *nblk_ident;            // device, k_identify blocks per signal
  int32_t nb_time, nb_fft, nb_ident;    // blocks of the largest iteration of every signal
  int32_t grid_time, grid_fft, grid_ident;  // workgroups of the packed launches
  int32_t ident_bins;                   // bins per k_identify unit (a multiple of 4)
  int32_t col_slices;                   // k_identify column slices (1: trials x rows <= 64)
  int32_t count;                        // signals
  BatchWork work;
  size_t lds_time, lds_fft;             // dynamic LDS (max over signals)
  size_t lds_fft_small;                 // the small bin-FFT path's (max over signals)
  int32_t *counters;                    // device {iterations started, signals stopped}
  int32_t *host;                        // host-mapped: the counters, then the iteration's
                                        // large-path bin-FFT units
  int mixed;
  int large_fft = 1;                    // 0: the large-path bin-FFT launch is left out
};
// A staging area in device memory holding `count` descriptors, then their
// payloads: k_tf_scatter copies each descriptor's two payloads (an engine's
// transforms and q's, or a batch's Ctx and InitArgs arrays) to their
// destinations, in 8-byte words (one launch for every signal of a batch).
struct TfScatterDesc {
  void *dst[2];
  int64_t off[2];    // payload offsets from the staging area, in bytes
  int64_t words[2];  // payload sizes in 8-byte words
};
hipError_t LaunchTfScatter(const unsigned char *staging, int32_t count, hipStream_t s);
hipError_t LaunchBatchReset(const Ctx *ctxs, const InitArgs *args, int32_t count,
                            int32_t *counters, hipStream_t s);
hipError_t LaunchBatchIteration(const BatchLaunch &b, int parity, hipStream_t s);
// One stage of a batched iteration: 0 the plan, 1 the gather-fold, 2 the
// bin FFT, 3 identify.
constexpr int kBatchStages = 4;
hipError_t LaunchBatchStage(const BatchLaunch &b, int parity, int stage, hipStream_t s);
struct BatchResultDev {
  int32_t iters, stop_stale, flags, stopped, last_it, count;
};
hipError_t LaunchBatchResults(const Ctx *ctxs, int32_t count, BatchResultDev *out, hipStream_t s);
struct ModesOut {  // where one signal's found keys and coefficients go (device)
  int64_t *keys;
  double2 *coefs;
  int64_t cap;     // entries the buffers hold
};
// Every signal's found modes, in map order, into caller buffers.
hipError_t LaunchBatchModesOut(const Ctx *ctxs, const ModesOut *out, int32_t count, hipStream_t s);

// The adjoint of recorded solves, one signal or every signal of a batch at
// once: per reverse step one launch for the bins and one for the bin DFTs
// and terms, then one gather launch writing every output sample.
struct AdjSig {            // one signal (device memory, per call)
  double2 *lam;            // n: dL/dc over k0 (read only at found keys)
  double2 *fbar;           // iters x trials_cap x rows_cap x B_cap: fold gradients
  double2 *taps;           // iters x taps_stride: each iteration's conjugated taps (f B)
  int64_t taps_stride;
  double2 *lags;           // iters x trials_cap x lags_stride: per-lag contributions
  int64_t lags_stride;     // (k_adj_lags); null: the gather computes them per sample
  const int64_t *rec;      // recorded k1 rows, Ctx.B_cap apart
  const int64_t *keys;     // m seed keys and their gradients
  const double2 *grad;
  void *out;               // n samples of the output precision
  int64_t n, m;
  int32_t iters, trials;
};
struct AdjStep {           // one signal at one reverse step (built on the device)
  IterParams ip;
  int32_t active;          // the signal has this iteration and it identified something
  int32_t trials, count_before, it;
};
struct AdjLaunch {
  const Ctx *ctxs;
  const AdjSig *sig;
  AdjStep *steps;          // max_steps x count
  int32_t count, max_steps;
  // block maps {signal, local block}; null for one signal (block = index)
  const int2 *map_bins, *map_fft;
  const int32_t *nblk_bins;  // bins blocks per signal (with map_bins)
  int32_t nb_bins, nb_fft;   // blocks per launch; map_fft: kPrepBlocks terms
                             // blocks, then trials_cap x rows_cap DFT blocks
  size_t lds_fft;
  int32_t per_small, per_gather;  // blocks per signal: clear / seed, gather
  int32_t per_lags;               // blocks per (signal, step) of k_adj_lags
  int any_tab, any_gen;           // some signal has lag tables / some has none
  int fft_small_only;             // every signal's bin FFTs are on the small path
  int mixed;                      // output precision
};
hipError_t LaunchAdj(const AdjLaunch &b, hipStream_t s);
// Blocks per (signal, step) of k_adj_lags (one lag slot of B residues per
// block at a time) for lag tables of up to `stride` entries.
inline int32_t LagBlocks(int64_t stride) {
  const int64_t b = (stride + 4 * 256 - 1) / (4 * 256);
  return int32_t(b < 1 ? 1 : (b > 64 ? 64 : b));
}
hipError_t LaunchSetModes(const Ctx &c, const int64_t *keys, const double2 *coefs,
                          int64_t m, hipStream_t s);
hipError_t LaunchGetCoefs(const Ctx &c, double2 *out, int64_t m, hipStream_t s);
// BinInFreq on a caller's rows x B matrix; limbs: rows x B x 4 integers and
// emax one int of scratch (the exact per-bin term sums).
hipError_t LaunchBinInFreqDirect(int64_t n, int64_t mask, const BParams &bp,
                                 const DevTf &tf, int64_t q, const int64_t *k0,
                                 const double2 *coef, int64_t m, double2 *out,
                                 unsigned long long *limbs, int32_t *emax, hipStream_t s);
hipError_t LaunchMatPencil(const double2 *u, int64_t m, double *sigma,
                           uint8_t *vote, hipStream_t s);

// PostAnalyze (harness.hip): `blocks` workgroups write their partial sums
// (kPaFields doubles each) to partial, one more reduces them into
// stats[0..9] (device memory).
constexpr int kPaFields = 10;
constexpr int kPaBlocks = 512;  // workgroups of a PostAnalyze pass
hipError_t LaunchPostAnalyze(const Ctx &c, const int64_t *truth, int64_t m, const double2 *xhat,
                             double *partial, int32_t blocks, double *stats, hipStream_t s);
// The signal of Demo1 (src/demo1.cc:27-34) from the reference's RNG seeded
// with seed: GenerateModeMap on the host (k keys and interleaved complex
// coefficients into host arrays, in draw order), GenerateXhatAlt into xhat
// (n double2, device).  Returns an mpsft_status.
int GenerateSignal(int64_t n, int32_t k, double sigma, uint32_t seed, double2 *xhat,
                   int64_t *keys, double *coefs, hipStream_t s);
// The device address of the static PostAnalyze scratch (kPaBlocks x
// kPaFields doubles) on the current device, looked up on its first use.
hipError_t PostAnalyzeScratch(double **ptr);

}  // namespace mpsft
