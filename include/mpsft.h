/*
 * mpsft.h -- C-ABI of the MI355X MPSFT recovery loop (libmpsft_hip.so).
 *
 * Plain pointers and sizes only.  Every device pointer is a HIP device
 * address owned by the caller; every call is ordered on the HIP stream it
 * is given (NULL = the default stream).  Complex arrays are interleaved
 * (re, im): double2 for MPSFT_FP64, float2 for the MPSFT_MIXED input x.
 * No call aborts: errors come back as an mpsft_status.
 *
 * Each entry point replaces a reference (jiaweichiu/mpsft) interface; the
 * file:line of that interface is given beside it (paths under src/).
 */
#ifndef MPSFT_H
#define MPSFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPSFT_VERSION_MAJOR 0
#define MPSFT_VERSION_MINOR 1

/* Status codes.  The reference CHECK-aborts (glog) where these are
 * returned: even bins (src/window.cc:33), wrong row count
 * (src/binner.cc:83,410), unknown binner type (src/binner.cc:74,533),
 * window_threshold <= 0 (src/iterate.cc:84), even trials
 * (src/iterate.cc:86). */
typedef enum mpsft_status {
  MPSFT_OK = 0,
  MPSFT_ERR_INVALID_ARG = 1, /* even bins / trials, bad n, bad threshold */
  MPSFT_ERR_SHAPE = 2,       /* rows / sizes do not match the plan */
  MPSFT_ERR_UNSUPPORTED = 3, /* unknown binner type or precision, B too big */
  MPSFT_ERR_HIP = 4,         /* a HIP runtime call failed */
  MPSFT_ERR_NOMEM = 5,       /* device allocation failed */
  MPSFT_ERR_CAPACITY = 6,    /* found-mode list overflowed its capacity */
  MPSFT_ERR_NOT_CONVERGED = 7 /* solve hit its iteration cap */
} mpsft_status;

/* Arithmetic forms.  FP64: input complex128, every stage in double, equal
 * to the reference.  MIXED: input complex64 (half the bytes in device
 * memory); each gathered sample is widened exactly to double and every
 * stage after it (window products, fold, bin FFT, pencil, estimate, rounding)
 * runs in double, so MIXED equals FP64 run on the complex64-rounded samples.
 * (FP32 window products or bin FFTs leave errors at the noise level per bin
 * and change the pencil's decisions; a double-float form that does not costs
 * more than FP64 on this part.)  Thresholds: pass sigma_eff = sqrt(sigma^2 +
 * sigma_q^2), sigma_q the rms rounding error of the complex64 samples. */
typedef enum mpsft_precision {
  MPSFT_FP64 = 0,
  MPSFT_MIXED = 1
} mpsft_precision;

/* Transform{a, b, c, a_inv} (src/integer.h:75-82): y_t = x_{a t + c}
 * e^{2 pi i b t / n}.  Values may be unreduced or negative, as the
 * reference allows (src/binner_test.cc:124); they are reduced mod n. */
typedef struct mpsft_transform {
  int64_t a, b, c, a_inv;
} mpsft_transform;

const char *mpsft_status_string(int status);
int mpsft_version(void);

/* ---- device memory / stream helpers (for callers without a runtime) --- */
int mpsft_device_count(int *count);
int mpsft_malloc(void **ptr, size_t bytes);
int mpsft_free(void *ptr);
int mpsft_memcpy_htod(void *dst, const void *src, size_t bytes);
int mpsft_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int mpsft_memset(void *dst, int value, size_t bytes);
int mpsft_stream_create(void **stream);
int mpsft_stream_destroy(void *stream);
int mpsft_stream_synchronize(void *stream);
int mpsft_device_synchronize(void);

/* ---- integer helpers ----
 * mpsft_transform_make: Transform(n, a, b, c) (src/integer.cc:51-56) with
 * the inverse found by extended Euclid (equal to the Fermat inverse of
 * src/integer.cc:39-40 for prime n, and valid for odd a when n = 2^k). */
int mpsft_transform_make(int64_t n, int64_t a, int64_t b, int64_t c,
                         mpsft_transform *tf);
/* NumBits (src/iterate.cc:30-42). */
int32_t mpsft_num_bits(int64_t n, int32_t bins);
/* Window::p() (src/window.cc:40-50); 0 on invalid arguments. */
int32_t mpsft_window_p(int64_t n, int32_t bins, double delta);

/* The delta of Window(n, bins, delta) recovered from its p taps wt
 * (Window::wt(), src/window.h:36); MPSFT_ERR_INVALID_ARG when the taps are
 * not such a window's.  Host only. */
int mpsft_window_delta(int64_t n, int32_t bins, int32_t p, const double *wt, double *delta);

/* The signal of Demo1(opt, n, k, sigma) (src/demo1.cc:27-34) after
 * RandomSeed(seed) (src/base.cc:52): GenerateModeMap (src/gen.cc:27-38) on
 * the host, k distinct locations with unit-modulus coefficients, into keys
 * (k int64) and coefs (k interleaved complex128), in draw order; then
 * GenerateXhatAlt (src/gen.cc:78-89) on the device into xhat_dev (n
 * double2): the modes plus complex Gaussian noise of sigma/sqrt(2n) per
 * component, from the same std::mt19937 stream (normals by libstdc++'s
 * polar method).  The signal is x = the unnormalised backward DFT of x-hat
 * (src/demo1.cc:32-33).  n < 2^31.  Synchronises the stream. */
int mpsft_generate_signal(int64_t n, int32_t k, double sigma, uint32_t seed, void *xhat_dev,
                          int64_t *keys, double *coefs, void *stream);

/* ---- plan: one window + binner pair ----
 * Replaces Window(n, bins, delta) (src/window.cc:28) together with
 * BinInTime::Create / BinInFreq::Create (src/binner.cc:58-76, 526-535).
 * bits gives 1 + 2*bits rows (src/binner.h:27-28). */
typedef struct mpsft_plan mpsft_plan;
int mpsft_plan_create(int64_t n, int32_t bins, double delta, int32_t bits,
                      int precision, mpsft_plan **plan);
int mpsft_plan_destroy(mpsft_plan *plan);

/* A plan from a Window the caller already holds (src/window.h:25-55), which
 * exposes n(), bins(), p() and its taps wt() but keeps delta private:
 * wt[i] for i < p, t = i for i <= (p-1)/2 and i - p above (Window::t()).
 * BinInTime then uses these taps exactly; delta is recovered from them (the
 * taps' Gaussian factor gives sigma_f, src/window.cc:35-37) for the window's
 * frequency response in BinInFreq and the identify step.  Taps that are not
 * a Window(n, bins, delta)'s return MPSFT_ERR_INVALID_ARG. */
int mpsft_plan_create_window(int64_t n, int32_t bins, int32_t p, const double *wt,
                             int32_t bits, int precision, mpsft_plan **plan);
/* The plan's window delta (as given, or as recovered from the taps). */
int mpsft_plan_delta(const mpsft_plan *plan, double *delta);

/* BinInTime::Run (src/binner.h:33-34; V0 definition src/binner.cc:81-111,
 * default V4 src/binner.cc:394-521).  x_dev: n complex (double2 for FP64,
 * float2 for MIXED).  out_dev: (1+2*bits) x bins complex double, row-major,
 * overwritten. */
int mpsft_bin_in_time(mpsft_plan *plan, const void *x_dev, int64_t n,
                      const mpsft_transform *tf, int64_t q, void *out_dev,
                      void *stream);

/* BinInFreq::Run (src/binner.h:126-127; default V1 src/binner.cc:565-591).
 * Subtracts the window response of the m modes (k0_dev: int64 locations,
 * coef_dev: double2 coefficients) from out_dev in place. */
int mpsft_bin_in_freq(mpsft_plan *plan, const int64_t *k0_dev,
                      const void *coef_dev, int64_t m,
                      const mpsft_transform *tf, int64_t q, void *out_dev,
                      void *stream);

/* Iterate's per-bin body (src/iterate.cc:113-175: bin energy test,
 * IdentifyFreq with MatPencil votes, src/iterate.cc:44-76, k1 rounding,
 * window test, coefficient estimate and un-permutation) over a caller's
 * bin matrices, one bin per wavefront.  a_dev: trials x (1+2*bits) x bins
 * double2, row-major, after BinInFreq (src/iterate.cc:109-110); qs: the
 * trials' offsets q (host); tf: the iteration's Transform.  Per bin b
 * (device arrays of `bins` entries): decide_dev[b] = 0 below the bin
 * threshold, 1 rejected by the sv threshold, 2 rejected by the window
 * threshold, 3 a mode found; k1_dev[b] the rounded permuted location (-1
 * when not reached); k0_dev[b], coef_dev[b] (double2) the found mode's
 * location and coefficient (-1 and 0 otherwise); margin_dev[b] (may be
 * NULL) the smallest relative distance of the bin's decisions from their
 * thresholds.  A threshold <= 0 turns its test off, as in the reference;
 * window_threshold must be > 0 (src/iterate.cc:84). */
int mpsft_identify_estimate(mpsft_plan *plan, const void *a_dev, int32_t trials,
                            const mpsft_transform *tf, const int64_t *qs,
                            double bin_threshold, double sv_threshold,
                            double window_threshold, int32_t *decide_dev,
                            int64_t *k1_dev, int64_t *k0_dev, void *coef_dev,
                            double *margin_dev, void *stream);

/* MatPencil (src/freqid.cc:27-48), batched on the device: for i < m,
 * u_dev[3i..3i+2] = (u0, u1, u2) double2; writes sigma_dev[2i..2i+1] and
 * vote_dev[i] (1 = second half). */
int mpsft_mat_pencil(const void *u_dev, int64_t m, double *sigma_dev,
                     uint8_t *vote_dev, void *stream);

/* ---- solver: Iterate and Demo1::Run on the device ---- */

/* IterateOptions (src/iterate.h:26-59).  Types: bin_in_time_type 5/6 and
 * bin_in_freq_type 5/6 select the HIP FP64 / MIXED forms; the reference's
 * CPU types 0-4 / 0-1 are not provided by this library. */
typedef struct mpsft_iterate_options {
  int32_t bins;
  double window_delta;
  int32_t trials;
  double bin_threshold;
  double window_threshold;
  double sv_threshold;
  int32_t bin_in_time_type;
  int32_t bin_in_freq_type;
} mpsft_iterate_options;

/* Demo1Options (src/demo1.h:25-33) plus the mode count k and noise sigma
 * of Demo1(opt, n, k, sigma) (src/demo1.cc:27). */
typedef struct mpsft_solve_options {
  int32_t trials;
  int32_t min_bins;
  double window_delta;
  double window_threshold;
  int32_t max_stale_iter;
  int32_t k;
  double sigma;
  int32_t max_iter; /* safety cap on iterations; 0 = 100000 */
} mpsft_solve_options;

typedef struct mpsft_solver mpsft_solver;

/* A solver owns the device state of one signal length n: found-mode
 * storage (a dense n-entry coefficient array plus a key list), per-B
 * window tables and the iteration graph.  capacity bounds the number of
 * distinct found modes (0 = 8*k + 4096). */
int mpsft_solver_create(int64_t n, const mpsft_solve_options *opt,
                        int precision, int64_t capacity, mpsft_solver **s);
int mpsft_solver_destroy(mpsft_solver *s);

/* Demo1::Run (src/demo1.cc:36-59) on x_dev.  The transforms and q's are
 * drawn on the host from a std::mt19937 seeded with seed, in the order of
 * Transform(n) (src/integer.cc:44-49) then RandomInt32() % n per trial
 * (src/iterate.cc:103); for n = 2^k the multiplier is drawn odd.
 * Returns the iteration count in *iters. */
int mpsft_solve(mpsft_solver *s, const void *x_dev, uint32_t seed,
                int32_t *iters, void *stream);

/* Instrumentation, no reference counterpart: the same solve with each
 * kernel bracketed by HIP events on `stream` (no graph), returning per
 * solve the summed milliseconds of the gather-fold, bin-FFT and identify
 * kernels in kernel_ms[0..2].  The gather-fold and the bin FFT rewrite
 * their whole output from unchanged inputs, so each is launched 8 times
 * back to back per iteration and its mean launch time is counted (keeping
 * event overhead out).  Results are identical to mpsft_solve. */
int mpsft_solve_profiled(mpsft_solver *s, const void *x_dev, uint32_t seed,
                         int32_t *iters, double *kernel_ms, void *stream);

/* Iterate (src/iterate.cc:80-177) once with explicit options, transform
 * and q's (qs: trials entries, host memory).  The solver's current found
 * modes are the ModeMap; *found is the return value of Iterate. */
int mpsft_iterate(mpsft_solver *s, const void *x_dev,
                  const mpsft_iterate_options *opt, const mpsft_transform *tf,
                  const int64_t *qs, int32_t *found, void *stream);

/* Found-mode map access (host memory).  set replaces the map; keys are
 * reduced modulo n, and two keys equal modulo n are refused with
 * MPSFT_ERR_INVALID_ARG (a ModeMap, src/base.h:35, holds a key once). */
int mpsft_solver_modes_count(mpsft_solver *s, int64_t *count);
int mpsft_solver_modes_get(mpsft_solver *s, int64_t *keys, double *coefs,
                           int64_t cap, int64_t *count);
int mpsft_solver_modes_set(mpsft_solver *s, const int64_t *keys,
                           const double *coefs, int64_t m);
/* PostAnalyze (src/demo1.cc:61-112) of the last solve on the device: the
 * found modes against the true locations truth_keys_dev[0..m) (device
 * int64, ascending; MPSFT_ERR_INVALID_ARG otherwise) and the signal's
 * x-hat (device, n double2, the xh_ of Demo1, src/demo1.cc:30-33).  stats
 * (host, 9 doubles): hit count, hit mean and rms |c - xhat|, miss count,
 * miss mean and rms, then over all n locations the mean, rms and max of
 * |found - xhat| (found = 0 where nothing was found).  A kind with no
 * entries gives 0/0 (NaN) for its mean and rms, as the reference prints. */
int mpsft_solver_post_analyze(mpsft_solver *s, const int64_t *truth_keys_dev, int64_t m,
                              const void *xhat_dev, double *stats, void *stream);
/* No counterpart in the reference: mpsft_solver_post_analyze for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  It calls no allocator and does not synchronise: its
 * partial sums live in a fixed 40 KiB device array of the library (one per
 * device, kept until the process ends), and stats_dev (device, 10 doubles)
 * gets the nine figures and then 1.0 if the truth keys were not ascending
 * (0.0 otherwise).  Any n and m; calls that overlap on one device race. */
int mpsft_solver_post_analyze_noalloc(mpsft_solver *s, const int64_t *truth_keys_dev,
                                      int64_t m, const void *xhat_dev, double *stats_dev,
                                      void *stream);

/* After an mpsft_iterate: copies the last iteration's bin matrices
 * (trials x rows x bins double2) into host memory. */
int mpsft_solver_bins_get(mpsft_solver *s, double *out, int64_t cap_doubles,
                          int32_t *trials, int32_t *rows, int32_t *bins);
/* Per-iteration log of the last solve: bins and found count. */
int mpsft_solver_log_get(mpsft_solver *s, int32_t *bins, int32_t *found,
                         int32_t cap, int32_t *n_iter);
/* The transforms (reduced mod n) and q's (trials per iteration) the last
 * mpsft_solve drew, for its first min(cap, iterations) iterations. */
int mpsft_solver_transforms_get(mpsft_solver *s, mpsft_transform *tfs,
                                int64_t *qs, int32_t cap, int32_t *count);
/* No reference counterpart: fixes the gather-fold's tile width (residues
 * per workgroup: 1, 2, 4 or 8; 0 = the default choice).  It changes how
 * the work is split, not what is computed. */
int mpsft_solver_set_tile_residues(mpsft_solver *s, int32_t residues);

/* ---- The differentiable form: no reference counterpart.  For callers
 * that train through the recovery (loss on the recovered coefficients,
 * gradient with respect to the signal).  With the decisions of a solve
 * held fixed (its transforms, q's, bins per iteration and the k1 each bin
 * identified), Iterate (src/iterate.cc:80-177) is linear in x: BinInTime,
 * the bin DFT, the BinInFreq subtraction of earlier estimates and the
 * coefficient estimate.  The recovered coefficients are then c = F x.
 *
 * set_record(on): while on, every mpsft_solve keeps those decisions in
 * device memory (4096 iterations of at most the first iteration's bins).
 * replay: the last recorded solve's map F applied to x_dev (n samples,
 * complex128 for MPSFT_FP64 or complex64 for MPSFT_MIXED, whatever the
 * solver's own precision); the found modes are replaced by F x.  Returns
 * MPSFT_ERR_SHAPE when no recorded solve is held (none ran, it did not
 * stop, or the map or tables changed since).
 * adjoint: out_dev (n samples of `precision`) = F^H g, where g holds, for
 * each of the m keys (device int64), the complex128 gradient dL/dconj(c)
 * of a real loss L -- the convention of torch, whose x.grad is then
 * F^H g.  Keys repeated add; keys outside [0, n) or not found carry
 * nothing.  Bins above 2047 in the recorded solve: MPSFT_ERR_UNSUPPORTED. */
int mpsft_solver_set_record(mpsft_solver *s, int32_t on);
/* The solver's n and precision, and an identifier of the recording it
 * holds (a new value per recorded solve, 0 when none is held); any output
 * may be null. */
int mpsft_solver_info(mpsft_solver *s, int64_t *n, int32_t *precision, int64_t *record_id);
int mpsft_solver_replay(mpsft_solver *s, const void *x_dev, int precision, void *stream);
int mpsft_solver_adjoint(mpsft_solver *s, const int64_t *keys_dev, const void *grad_dev,
                         int64_t m, void *out_dev, int precision, void *stream);

/* ---- The batched form: no reference counterpart.  Demo1::Run
 * (src/demo1.cc:36-59) over many independent signals at once, for callers
 * with many small signals: every stage of an iteration is one launch over
 * all signals, each signal keeping its own loop state and stopping on its
 * own.  Signal i has its own n[i], options opts[i] (first bins
 * 2 k + 1 <= 2047) and seed.  All signals share one precision. */
typedef struct mpsft_batch mpsft_batch;
int mpsft_batch_create(int32_t count, const int64_t *n, const mpsft_solve_options *opts,
                       int precision, mpsft_batch **batch);
int mpsft_batch_destroy(mpsft_batch *batch);
/* x_dev[i]: signal i in device memory (complex128 or complex64 as the
 * precision).  iters[i], status[i] (host, may be null): signal i's
 * iteration count and mpsft_status; the return value is the first non-OK
 * status, else MPSFT_OK. */
int mpsft_batch_solve(mpsft_batch *batch, const void *const *x_dev, const uint32_t *seeds,
                      int32_t *iters, int32_t *status, void *stream);
/* Instrumentation, no reference counterpart: the same batched solve run
 * one iteration at a time without a graph, each stage's kernel bracketed
 * by HIP events on `stream`.  kernel_ms[0..3]: summed milliseconds of the
 * plan, gather-fold, bin-FFT and identify kernels over the iterations;
 * *launches: the batched iterations launched.  Results are identical to
 * mpsft_batch_solve. */
int mpsft_batch_solve_profiled(mpsft_batch *batch, const void *const *x_dev,
                               const uint32_t *seeds, int32_t *iters, int32_t *status,
                               double *kernel_ms, int32_t *launches, void *stream);
/* Times whole batched solves of these signals in a coordinate search --
 * k_identify bins per unit (8, 16, 32), then the gather-fold tile width of
 * B >= 256 and then that of smaller B (0: the default width per B, or 1,
 * 2, 4, 8 residues) -- and keeps the fastest setting for later solves;
 * *residues gets the tile width of B >= 256.  Results depend on none of
 * them.  mpsft_solver_set_tile_residues sets one width for every B. */
int mpsft_batch_autotune(mpsft_batch *batch, const void *const *x_dev, const uint32_t *seeds,
                         int32_t *residues, void *stream);
/* Results of signal i of the last batched solve, as the solver calls. */
int mpsft_batch_modes_count(mpsft_batch *batch, int32_t i, int64_t *count);
int mpsft_batch_modes_get(mpsft_batch *batch, int32_t i, int64_t *keys, double *coefs,
                          int64_t cap, int64_t *count);
int mpsft_batch_log_get(mpsft_batch *batch, int32_t i, int32_t *bins, int32_t *found,
                        int32_t cap, int32_t *n_iter);
/* The last batched solve's found-mode counts (host memory, no device
 * access), and every signal's keys (int64) and coefficients (complex128) in
 * map order copied into device buffers keys_dev[i], coefs_dev[i] of at least
 * counts[i] entries (host arrays of device pointers). */
int mpsft_batch_modes_counts(mpsft_batch *batch, int64_t *counts);
int mpsft_batch_modes_device(mpsft_batch *batch, int64_t *const *keys_dev,
                             void *const *coefs_dev, void *stream);
/* The differentiable form over a batch (see mpsft_solver_set_record):
 * with recording on, a batched solve keeps every signal's decisions;
 * replay runs signal i's recorded map on x_dev (as mpsft_solver_replay);
 * adjoint applies every signal's adjoint at once, one launch per stage for
 * all signals: out_dev[i] (n[i] samples of `precision`) = F_i^H g_i for the
 * m[i] keys_dev[i] (int64) and grad_dev[i] (complex128) entries (device
 * pointers; the arrays of pointers and m are host memory).  info gives
 * signal i's n, precision and recording identifier. */
int mpsft_batch_set_record(mpsft_batch *batch, int32_t on);
int mpsft_batch_info(mpsft_batch *batch, int32_t i, int64_t *n, int32_t *precision,
                     int64_t *record_id);
int mpsft_batch_replay(mpsft_batch *batch, int32_t i, const void *x_dev, int precision,
                       void *stream);
int mpsft_batch_adjoint(mpsft_batch *batch, const int64_t *const *keys_dev,
                        const void *const *grad_dev, const int64_t *m, void *const *out_dev,
                        int precision, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MPSFT_H */
