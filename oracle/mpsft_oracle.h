/*
 * mpsft_oracle.h -- CPU FP64 restatement of the MPSFT recovery loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (mpsft_amd/) never links or calls it.
 *
 * Every function restates the semantics of the reference (jiaweichiu/mpsft,
 * C++11) from its source text; citations are file:line under src/.  The
 * reference itself cannot be built here (it needs FFTW3, glog, Eigen3,
 * Boost.SIMD and Bazel, none of which exist in this image), so the three
 * third-party pieces it relies on are re-derived from their mathematics:
 *   - FFTW forward/backward DFT  -> own mixed-radix + Bluestein FFT
 *     (unnormalised, sign convention of FFTW_FORWARD = -1);
 *   - Eigen JacobiSVD of the 2x2 pencil -> closed form;
 *   - Boost.SIMD sincos -> libm sin/cos.
 *
 * Pinning (what checks each function; the tests live in tests/):
 *   rng          : libstdc++ mt19937 + distributions, identical engine; the
 *                  stream is only bounded by the reference tests.
 *   transform    : a*a_inv == 1 (src/integer_test.cc:36-51) + brute force.
 *   window       : src/window_test.cc:30-79 KATs via a long-double DFT.
 *   fft          : src/base_test.cc:66-84 KAT + long-double naive DFT.
 *   bin_in_time  : src/binner_test.cc:55-103 golden (1.12652-0.108838i),
 *                  V0 == V1 (src/binner_test.cc:29-52), brute force in
 *                  long double on tiny n, time == -freq identity.
 *   bin_in_freq  : the time == -freq identity (src/binner_test.cc:98-102,
 *                  140-142) and a closed-form single-mode evaluation.
 *   mat_pencil   : three golden vectors (src/freqid_test.cc:26-60) and
 *                  numpy.linalg.svd on random matrices.
 *   iterate/demo1: src/iterate_test.cc acceptance bounds, exact recovery on
 *                  noiseless signals, and a dense-FFT global check.
 *
 * All complex arrays are interleaved (re, im) doubles.  Mode maps are passed
 * as an opaque handle wrapping std::unordered_map<int32_t, complex<double>>,
 * the reference's ModeMap (src/base.h:35).
 */
#ifndef MPSFT_ORACLE_H
#define MPSFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference CHECK-aborts instead) ---- */
#define ORC_OK 0
#define ORC_ERR_INVALID_ARG 1
#define ORC_ERR_SHAPE 2
#define ORC_ERR_UNSUPPORTED 3

/* ---- threads used by the row-parallel loops (OpenMP); 0 = default ---- */
void orc_set_threads(int32_t threads);
int32_t orc_get_threads(void);

/* ---- global RNG, src/base.cc:32-56 ---- */
void orc_rng_seed(uint32_t seed);
int32_t orc_rng_int32(void);
int64_t orc_rng_int64(void);
double orc_rng_normal(void);
double orc_rng_double(void);

/* ---- integer helpers, src/integer.h:34-66, src/integer.cc:27-56 ---- */
int32_t orc_mul_mod(int64_t a, int64_t b, int64_t d);
int32_t orc_mul_pos_mod(int64_t a, int64_t b, int64_t d);
int32_t orc_mul_add_pos_mod(int64_t a, int64_t b, int64_t c, int64_t d);
/* Transform(n) drawn from the global RNG.  tf = {a, b, c, a_inv}. */
int orc_transform_draw(int32_t n, int32_t tf[4]);
/* Transform(n, a, b, c). */
int orc_transform_make(int32_t n, int32_t a, int32_t b, int32_t c,
                       int32_t tf[4]);
int32_t orc_num_bits(int32_t n, int32_t bins);

/* ---- window, src/window.cc:28-81 ---- */
int orc_window_params(int32_t n, int32_t bins, double delta, int32_t *p,
                      double *width, double *sigma_f, double *sigma_t);
int orc_window_taps(int32_t n, int32_t bins, double delta, double *wt,
                    int32_t *t);
double orc_window_sample_in_time(int32_t n, int32_t bins, double delta,
                                 int32_t i);
double orc_window_sample_in_freq(int32_t n, int32_t bins, double delta,
                                 double xi);

/* ---- FFT, stands in for FFTPlan (src/base.cc:86-98) ---- */
/* sign = -1 forward, +1 backward; unnormalised. */
int orc_fft(const double *in, double *out, int32_t n, int sign);
/* O(n^2) DFT accumulated in long double. */
int orc_dft_naive(const double *in, double *out, int32_t n, int sign);

/* ---- binners, src/binner.cc ---- */
/* version 0 = BinInTimeV0 (src/binner.cc:81-111), the definition;
 * version 1 = BinInTimeV1 (src/binner.cc:140-191), the symmetric form that
 * V2-V4 vectorise.  out: (1+2*bits) x bins complex, row-major. */
int orc_bin_in_time(int version, const double *x, int32_t n, int32_t bins,
                    double delta, int32_t bits, const int32_t tf[4], int32_t q,
                    double *out);
/* version 0 = BinInFreqV0 (src/binner.cc:540-560), 1 = BinInFreqV1
 * (src/binner.cc:565-591).  Subtracts from out in place. */
int orc_bin_in_freq(int version, const int32_t *keys, const double *coefs,
                    int32_t m, int32_t n, int32_t bins, double delta,
                    int32_t bits, const int32_t tf[4], int32_t q, double *out);

/* ---- freqid, src/freqid.cc:27-48 ---- */
/* u = {u0.re, u0.im, u1.re, u1.im, u2.re, u2.im}; returns the vote. */
int orc_mat_pencil(const double *u, double *sigma);

/* ---- mode maps (ModeMap, src/base.h:35) ---- */
typedef struct orc_modemap orc_modemap;
orc_modemap *orc_modemap_create(void);
void orc_modemap_destroy(orc_modemap *mm);
int32_t orc_modemap_size(const orc_modemap *mm);
/* Keys/coefs in the map's own iteration order. */
void orc_modemap_get(const orc_modemap *mm, int32_t *keys, double *coefs);
void orc_modemap_set(orc_modemap *mm, const int32_t *keys, const double *coefs,
                     int32_t m);

/* ---- Iterate, src/iterate.h:26-65, src/iterate.cc:80-177 ---- */
typedef struct orc_iterate_options {
  int32_t bins;
  double window_delta;
  int32_t trials;
  double bin_threshold;
  double window_threshold;
  double sv_threshold;
  int32_t bin_in_time_type; /* 0 or 1 here (V2-V4 reorder V1) */
  int32_t bin_in_freq_type; /* 0 or 1 */
} orc_iterate_options;

/* Record of one Iterate call, for parity checks.  Caller provides storage;
 * any pointer may be NULL.  a_out: trials x rows x bins complex, the bin
 * matrices after BinInFreq; decide_out: per bin, 0 = below bin threshold,
 * 1 = rejected by sv threshold, 2 = rejected by window threshold, 3 = found.
 * k1_out: per bin the permuted location (-1 if not reached). */
typedef struct orc_iterate_trace {
  int32_t tf[4];
  int32_t *q_out;      /* trials */
  double *a_out;       /* trials*rows*bins*2 */
  int32_t *decide_out; /* bins */
  int32_t *k1_out;     /* bins */
  double *margin_out;  /* bins: min relative decision margin */
  int32_t bits;
  int32_t found;
} orc_iterate_trace;

/* Iterate with the transform and q's drawn from the global RNG. */
int orc_iterate(const double *x, int32_t n, const orc_iterate_options *opt,
                orc_modemap *mm, orc_iterate_trace *trace);
/* Iterate with an explicit transform and q's (no RNG draws). */
int orc_iterate_fixed(const double *x, int32_t n,
                      const orc_iterate_options *opt, const int32_t tf[4],
                      const int32_t *qs, orc_modemap *mm,
                      orc_iterate_trace *trace);

/* Iterate's per-bin body alone (src/iterate.cc:113-175) on given bin
 * matrices a (trials x rows x bins complex, after BinInFreq) with the
 * iteration's transform and q's: the decisions, k1 and margins into trace
 * (its a_out and q_out are not used) and the found modes added to mm.
 * Returns 1 if a mode was found, 0 if not, or a negative status. */
int orc_identify(const double *a, int32_t trials, int32_t n,
                 const orc_iterate_options *opt, const int32_t tf[4],
                 const int32_t *qs, orc_modemap *mm, orc_iterate_trace *trace);

/* ---- Demo1::Run, src/demo1.h:25-33, src/demo1.cc:36-59 ---- */
typedef struct orc_demo1_options {
  int32_t trials;
  int32_t min_bins;
  double window_delta;
  double window_threshold;
  int32_t max_stale_iter;
} orc_demo1_options;

/* Runs the solve loop on x with k modes and noise sigma; the thresholds are
 * 3*sigma/sqrt(B) and sigma/sqrt(B) as in src/demo1.cc:50-51.  Transforms
 * and q's come from the global RNG.  bins_log (may be NULL, capacity
 * log_cap) receives B per iteration, found_log the map size after it.
 * Returns the number of iterations run, or a negative status. */
int32_t orc_demo1_run(const double *x, int32_t n, int32_t k, double sigma,
                      const orc_demo1_options *opt, orc_modemap *found,
                      int32_t *bins_log, int32_t *found_log, int32_t log_cap);

/* Per-iteration record of the Demo1::Run loop (any array may be NULL; each
 * holds cap entries): bins, map size after the iteration, Iterate's return
 * value, the stale count after it, and the two thresholds it ran with. */
typedef struct orc_demo1_log {
  int32_t *bins;
  int32_t *found;
  uint8_t *res;
  int32_t *stale;
  double *bin_thr;
  double *sv_thr;
  int32_t cap;
} orc_demo1_log;

/* orc_demo1_run with the full per-iteration record. */
int32_t orc_demo1_run_log(const double *x, int32_t n, int32_t k, double sigma,
                          const orc_demo1_options *opt, orc_modemap *found,
                          orc_demo1_log *log);

/* The same loop with Iterate replaced by a script: iteration i returns
 * script_res[i] and leaves the map at script_size[i] modes.  Runs until the
 * stale stop or the script ends; returns the iterations run.  Pins the
 * schedule (bins, thresholds, stop) independently of any signal. */
int32_t orc_demo1_schedule(int32_t k, double sigma, const orc_demo1_options *opt,
                           const uint8_t *script_res, const int32_t *script_size,
                           int32_t script_len, orc_demo1_log *log);

/* ---- gen, src/gen.cc:27-89 ---- */
int orc_generate_mode_map(int32_t n, int32_t k, orc_modemap *mm);
int orc_evaluate_modes(int32_t n, const orc_modemap *mm, double *out);
int orc_generate_xhat(int32_t n, const orc_modemap *mm, double sigma,
                      double *out);
int orc_generate_xhat_alt(int32_t n, const orc_modemap *mm, double sigma,
                          double *out);

/* ---- PostAnalyze, src/demo1.cc:61-112 ----
 * stats = {hit, hit_l1, hit_l2, miss, miss_l1, miss_l2, l1, l2, max}. */
int orc_post_analyze(int32_t n, const orc_modemap *truth, const double *xh,
                     const orc_modemap *found, double *stats);

#ifdef __cplusplus
}
#endif

#endif
