// capi.cpp -- the extern "C" boundary declared in include/mpsft.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/mpsft.h"
#include "batch.h"
#include "engine.h"
#include "kernels.h"

using mpsft::DevTf;
using mpsft::Engine;

struct mpsft_plan {
  Engine *eng = nullptr;
  int64_t n = 0;
  int32_t bins = 0, bits = 0;
  double delta = 0;
  mpsft::BParams bp{};
  // BinInFreq scratch: 4 integers per (row, bin) and the grid exponent
  unsigned long long *d_limbs = nullptr;
  int32_t *d_emax = nullptr;
};

struct mpsft_batch {
  mpsft::Batch *b;
};

struct mpsft_solver {
  Engine *eng = nullptr;
  mpsft_solve_options opt{};
};

namespace {

hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

int Hip(hipError_t e) { return e == hipSuccess ? MPSFT_OK : MPSFT_ERR_HIP; }

bool TfOk(int64_t n, const mpsft_transform *tf) {
  return tf && n >= 2 && mpsft::InvMod(tf->a, n) != 0;
}

DevTf Reduce(int64_t n, const mpsft_transform *tf) {
  return mpsft::ReduceTf(n, tf->a, tf->b, tf->c, tf->a_inv);
}

}  // namespace

extern "C" {

const char *mpsft_status_string(int status) {
  switch (status) {
    case MPSFT_OK: return "ok";
    case MPSFT_ERR_INVALID_ARG: return "invalid argument";
    case MPSFT_ERR_SHAPE: return "shape mismatch";
    case MPSFT_ERR_UNSUPPORTED: return "unsupported";
    case MPSFT_ERR_HIP: return "HIP runtime error";
    case MPSFT_ERR_NOMEM: return "device allocation failed";
    case MPSFT_ERR_CAPACITY: return "found-mode capacity exceeded";
    case MPSFT_ERR_NOT_CONVERGED: return "iteration cap reached";
    default: return "unknown status";
  }
}

int mpsft_version(void) { return MPSFT_VERSION_MAJOR * 100 + MPSFT_VERSION_MINOR; }

int mpsft_device_count(int *count) {
  if (!count) return MPSFT_ERR_INVALID_ARG;
  *count = 0;
  return Hip(hipGetDeviceCount(count));
}
int mpsft_malloc(void **ptr, size_t bytes) {
  if (!ptr) return MPSFT_ERR_INVALID_ARG;
  return hipMalloc(ptr, bytes) == hipSuccess ? MPSFT_OK : MPSFT_ERR_NOMEM;
}
int mpsft_free(void *ptr) { return Hip(hipFree(ptr)); }
int mpsft_memcpy_htod(void *dst, const void *src, size_t bytes) {
  return Hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
int mpsft_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  return Hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
int mpsft_memset(void *dst, int value, size_t bytes) {
  return Hip(hipMemset(dst, value, bytes));
}
int mpsft_stream_create(void **stream) {
  if (!stream) return MPSFT_ERR_INVALID_ARG;
  hipStream_t s;
  const int rc = Hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = rc == MPSFT_OK ? reinterpret_cast<void *>(s) : nullptr;
  return rc;
}
int mpsft_stream_destroy(void *stream) { return Hip(hipStreamDestroy(S(stream))); }
int mpsft_stream_synchronize(void *stream) { return Hip(hipStreamSynchronize(S(stream))); }
int mpsft_device_synchronize(void) { return Hip(hipDeviceSynchronize()); }

int mpsft_transform_make(int64_t n, int64_t a, int64_t b, int64_t c, mpsft_transform *tf) {
  if (!tf || n < 2) return MPSFT_ERR_INVALID_ARG;
  const int64_t inv = mpsft::InvMod(a, n);
  if (inv == 0) return MPSFT_ERR_INVALID_ARG;
  tf->a = a;
  tf->b = b;
  tf->c = c;
  tf->a_inv = inv;
  return MPSFT_OK;
}

int32_t mpsft_num_bits(int64_t n, int32_t bins) { return mpsft::NumBits(n, bins); }

int32_t mpsft_window_p(int64_t n, int32_t bins, double delta) {
  mpsft::BParams bp;
  return mpsft::WindowParams(n, bins, delta, &bp) ? bp.P : 0;
}

int mpsft_plan_create(int64_t n, int32_t bins, double delta, int32_t bits, int precision,
                      mpsft_plan **plan) {
  if (!plan) return MPSFT_ERR_INVALID_ARG;
  *plan = nullptr;
  mpsft::BParams bp;
  if (!mpsft::WindowParams(n, bins, delta, &bp) || bits < 1 || bits > 30)
    return MPSFT_ERR_INVALID_ARG;
  if (bp.logM > mpsft::kMaxLogMBig) return MPSFT_ERR_UNSUPPORTED;
  if (precision != MPSFT_FP64 && precision != MPSFT_MIXED) return MPSFT_ERR_UNSUPPORTED;
  mpsft_plan *p = new (std::nothrow) mpsft_plan;
  if (!p) return MPSFT_ERR_NOMEM;
  p->eng = new (std::nothrow) Engine(n, precision, 1);
  if (!p->eng || p->eng->status() != MPSFT_OK) {
    const int rc = p->eng ? p->eng->status() : MPSFT_ERR_NOMEM;
    delete p->eng;
    delete p;
    return rc;
  }
  p->n = n;
  p->bins = bins;
  p->bits = bits;
  p->delta = delta;
  p->bp = bp;
  p->bp.bits = bits;
  const size_t limbs = size_t(1 + 2 * bits) * bins * 4;
  if (hipMalloc(reinterpret_cast<void **>(&p->d_limbs), limbs * sizeof(unsigned long long)) !=
          hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&p->d_emax), sizeof(int32_t)) != hipSuccess) {
    mpsft_plan_destroy(p);
    return MPSFT_ERR_NOMEM;
  }
  *plan = p;
  return MPSFT_OK;
}

int mpsft_window_delta(int64_t n, int32_t bins, int32_t p, const double *wt, double *delta) {
  if (!delta) return MPSFT_ERR_INVALID_ARG;
  return mpsft::WindowDeltaFromTaps(n, bins, p, wt, delta) ? MPSFT_OK : MPSFT_ERR_INVALID_ARG;
}

int mpsft_plan_create_window(int64_t n, int32_t bins, int32_t p, const double *wt, int32_t bits,
                             int precision, mpsft_plan **plan) {
  if (!plan || !wt) return MPSFT_ERR_INVALID_ARG;
  *plan = nullptr;
  double delta = 0;
  if (!mpsft::WindowDeltaFromTaps(n, bins, p, wt, &delta)) return MPSFT_ERR_INVALID_ARG;
  int rc = mpsft_plan_create(n, bins, delta, bits, precision, plan);
  if (rc != MPSFT_OK) return rc;
  rc = (*plan)->eng->SetWindowTaps(bins, p, wt);
  if (rc != MPSFT_OK) {
    mpsft_plan_destroy(*plan);
    *plan = nullptr;
  }
  return rc;
}

int mpsft_plan_delta(const mpsft_plan *plan, double *delta) {
  if (!plan || !delta) return MPSFT_ERR_INVALID_ARG;
  *delta = plan->delta;
  return MPSFT_OK;
}

int mpsft_plan_destroy(mpsft_plan *plan) {
  if (!plan) return MPSFT_ERR_INVALID_ARG;
  delete plan->eng;  // waits for the device
  (void)hipFree(plan->d_limbs);
  (void)hipFree(plan->d_emax);
  delete plan;
  return MPSFT_OK;
}

int mpsft_bin_in_time(mpsft_plan *plan, const void *x_dev, int64_t n, const mpsft_transform *tf,
                      int64_t q, void *out_dev, void *stream) {
  if (!plan || !x_dev || !out_dev) return MPSFT_ERR_INVALID_ARG;
  if (n != plan->n) return MPSFT_ERR_SHAPE;
  if (!TfOk(n, tf)) return MPSFT_ERR_INVALID_ARG;
  return plan->eng->BinOnly(x_dev, plan->bins, plan->delta, plan->bits, Reduce(n, tf), q,
                            static_cast<double2 *>(out_dev), S(stream));
}

int mpsft_bin_in_freq(mpsft_plan *plan, const int64_t *k0_dev, const void *coef_dev, int64_t m,
                      const mpsft_transform *tf, int64_t q, void *out_dev, void *stream) {
  if (!plan || !out_dev || m < 0 || (m > 0 && (!k0_dev || !coef_dev)))
    return MPSFT_ERR_INVALID_ARG;
  if (!TfOk(plan->n, tf)) return MPSFT_ERR_INVALID_ARG;
  const int64_t n = plan->n;
  const int64_t mask = (n & (n - 1)) == 0 ? n - 1 : -1;
  int64_t qq = q % n;
  if (qq < 0) qq += n;
  return Hip(mpsft::LaunchBinInFreqDirect(n, mask, plan->bp, Reduce(n, tf), qq, k0_dev,
                                          static_cast<const double2 *>(coef_dev), m,
                                          static_cast<double2 *>(out_dev), plan->d_limbs,
                                          plan->d_emax, S(stream)));
}

int mpsft_identify_estimate(mpsft_plan *plan, const void *a_dev, int32_t trials,
                            const mpsft_transform *tf, const int64_t *qs, double bin_threshold,
                            double sv_threshold, double window_threshold, int32_t *decide_dev,
                            int64_t *k1_dev, int64_t *k0_dev, void *coef_dev, double *margin_dev,
                            void *stream) {
  if (!plan || !a_dev || !qs || !decide_dev || !k1_dev || !k0_dev || !coef_dev)
    return MPSFT_ERR_INVALID_ARG;
  if (!TfOk(plan->n, tf)) return MPSFT_ERR_INVALID_ARG;
  mpsft::IdOut out{decide_dev, k1_dev, k0_dev, static_cast<double2 *>(coef_dev), margin_dev};
  return plan->eng->IdentifyOnly(static_cast<const double2 *>(a_dev), trials, plan->bins,
                                 plan->delta, plan->bits, Reduce(plan->n, tf), qs, bin_threshold,
                                 sv_threshold, window_threshold, out, S(stream));
}

int mpsft_mat_pencil(const void *u_dev, int64_t m, double *sigma_dev, uint8_t *vote_dev,
                     void *stream) {
  if (m < 0 || (m > 0 && (!u_dev || !sigma_dev || !vote_dev))) return MPSFT_ERR_INVALID_ARG;
  return Hip(mpsft::LaunchMatPencil(static_cast<const double2 *>(u_dev), m, sigma_dev, vote_dev,
                                    S(stream)));
}

int mpsft_solver_create(int64_t n, const mpsft_solve_options *opt, int precision,
                        int64_t capacity, mpsft_solver **s) {
  if (!s || !opt) return MPSFT_ERR_INVALID_ARG;
  *s = nullptr;
  if (precision != MPSFT_FP64 && precision != MPSFT_MIXED) return MPSFT_ERR_UNSUPPORTED;
  if (opt->k < 0 || n < 2) return MPSFT_ERR_INVALID_ARG;
  int64_t cap = capacity > 0 ? capacity : 8 * int64_t(opt->k) + 4096;
  if (cap > n) cap = n;
  mpsft_solver *p = new (std::nothrow) mpsft_solver;
  if (!p) return MPSFT_ERR_NOMEM;
  p->eng = new (std::nothrow) Engine(n, precision, cap);
  if (!p->eng || p->eng->status() != MPSFT_OK) {
    const int rc = p->eng ? p->eng->status() : MPSFT_ERR_NOMEM;
    delete p->eng;
    delete p;
    return rc;
  }
  p->opt = *opt;
  *s = p;
  return MPSFT_OK;
}

int mpsft_solver_destroy(mpsft_solver *s) {
  if (!s) return MPSFT_ERR_INVALID_ARG;
  delete s->eng;
  delete s;
  return MPSFT_OK;
}

int mpsft_solve(mpsft_solver *s, const void *x_dev, uint32_t seed, int32_t *iters,
                void *stream) {
  if (!s || !x_dev || !iters) return MPSFT_ERR_INVALID_ARG;
  return s->eng->Solve(x_dev, s->opt, seed, iters, S(stream));
}

int mpsft_solve_profiled(mpsft_solver *s, const void *x_dev, uint32_t seed, int32_t *iters,
                         double *kernel_ms, void *stream) {
  if (!s || !x_dev || !iters || !kernel_ms) return MPSFT_ERR_INVALID_ARG;
  return s->eng->Solve(x_dev, s->opt, seed, iters, S(stream), kernel_ms);
}

int mpsft_iterate(mpsft_solver *s, const void *x_dev, const mpsft_iterate_options *opt,
                  const mpsft_transform *tf, const int64_t *qs, int32_t *found, void *stream) {
  if (!s || !x_dev || !opt || !qs || !found) return MPSFT_ERR_INVALID_ARG;
  if (!TfOk(s->eng->n(), tf)) return MPSFT_ERR_INVALID_ARG;
  return s->eng->IterateOnce(x_dev, *opt, Reduce(s->eng->n(), tf), qs, found, S(stream));
}

int mpsft_solver_post_analyze(mpsft_solver *s, const int64_t *truth_keys_dev, int64_t m,
                              const void *xhat_dev, double *stats, void *stream) {
  if (!s) return MPSFT_ERR_INVALID_ARG;
  return s->eng->PostAnalyze(truth_keys_dev, m, static_cast<const double2 *>(xhat_dev), stats,
                             S(stream));
}

int mpsft_generate_signal(int64_t n, int32_t k, double sigma, uint32_t seed, void *xhat_dev,
                          int64_t *keys, double *coefs, void *stream) {
  return mpsft::GenerateSignal(n, k, sigma, seed, static_cast<double2 *>(xhat_dev), keys, coefs,
                               S(stream));
}

int mpsft_solver_post_analyze_noalloc(mpsft_solver *s, const int64_t *truth_keys_dev, int64_t m,
                                      const void *xhat_dev, double *stats_dev, void *stream) {
  if (!s || !stats_dev) return MPSFT_ERR_INVALID_ARG;
  double *partial = nullptr;
  if (mpsft::PostAnalyzeScratch(&partial) != hipSuccess) return MPSFT_ERR_HIP;
  return s->eng->PostAnalyzeDev(truth_keys_dev, m, static_cast<const double2 *>(xhat_dev),
                                partial, mpsft::kPaBlocks, stats_dev, S(stream));
}

int mpsft_solver_modes_count(mpsft_solver *s, int64_t *count) {
  if (!s || !count) return MPSFT_ERR_INVALID_ARG;
  return s->eng->ModesCount(count);
}

int mpsft_solver_modes_get(mpsft_solver *s, int64_t *keys, double *coefs, int64_t cap,
                           int64_t *count) {
  if (!s || !count || (cap > 0 && (!keys || !coefs))) return MPSFT_ERR_INVALID_ARG;
  return s->eng->ModesGet(keys, coefs, cap, count);
}

int mpsft_solver_modes_set(mpsft_solver *s, const int64_t *keys, const double *coefs,
                           int64_t m) {
  if (!s) return MPSFT_ERR_INVALID_ARG;
  return s->eng->ModesSet(keys, coefs, m);
}

int mpsft_solver_bins_get(mpsft_solver *s, double *out, int64_t cap_doubles, int32_t *trials,
                          int32_t *rows, int32_t *bins) {
  if (!s || !out || !trials || !rows || !bins) return MPSFT_ERR_INVALID_ARG;
  return s->eng->BinsGet(out, cap_doubles, trials, rows, bins);
}

int mpsft_solver_log_get(mpsft_solver *s, int32_t *bins, int32_t *found, int32_t cap,
                         int32_t *n_iter) {
  if (!s || !n_iter) return MPSFT_ERR_INVALID_ARG;
  return s->eng->LogGet(bins, found, cap, n_iter);
}

int mpsft_solver_transforms_get(mpsft_solver *s, mpsft_transform *tfs, int64_t *qs, int32_t cap,
                                int32_t *count) {
  if (!s || !count || (cap > 0 && (!tfs || !qs))) return MPSFT_ERR_INVALID_ARG;
  return s->eng->TransformsGet(reinterpret_cast<DevTf *>(tfs), qs, cap, s->opt.trials, count);
}

int mpsft_solver_set_tile_residues(mpsft_solver *s, int32_t residues) {
  if (!s) return MPSFT_ERR_INVALID_ARG;
  return s->eng->SetTileResidues(residues);
}

int mpsft_solver_set_record(mpsft_solver *s, int32_t on) {
  if (!s) return MPSFT_ERR_INVALID_ARG;
  return s->eng->SetRecord(on != 0);
}

int mpsft_solver_info(mpsft_solver *s, int64_t *n, int32_t *precision, int64_t *record_id) {
  if (!s) return MPSFT_ERR_INVALID_ARG;
  if (n) *n = s->eng->n();
  if (precision) *precision = s->eng->precision();
  if (record_id) *record_id = s->eng->RecordId();
  return MPSFT_OK;
}

int mpsft_solver_replay(mpsft_solver *s, const void *x_dev, int precision, void *stream) {
  if (!s || !x_dev) return MPSFT_ERR_INVALID_ARG;
  return s->eng->Replay(x_dev, precision, S(stream));
}

int mpsft_solver_adjoint(mpsft_solver *s, const int64_t *keys_dev, const void *grad_dev,
                         int64_t m, void *out_dev, int precision, void *stream) {
  if (!s || !out_dev || m < 0 || (m > 0 && (!keys_dev || !grad_dev))) return MPSFT_ERR_INVALID_ARG;
  return s->eng->Adjoint(keys_dev, grad_dev, m, out_dev, precision, S(stream));
}

int mpsft_batch_create(int32_t count, const int64_t *n, const mpsft_solve_options *opts,
                       int precision, mpsft_batch **batch) {
  if (!batch) return MPSFT_ERR_INVALID_ARG;
  *batch = nullptr;
  if (precision != MPSFT_FP64 && precision != MPSFT_MIXED) return MPSFT_ERR_UNSUPPORTED;
  if (count <= 0 || !n || !opts) return MPSFT_ERR_INVALID_ARG;
  for (int32_t i = 0; i < count; ++i)
    if (2 * int64_t(opts[i].k) + 1 > mpsft::kMaxLdsBins) return MPSFT_ERR_UNSUPPORTED;
  mpsft_batch *b = new (std::nothrow) mpsft_batch;
  if (!b) return MPSFT_ERR_NOMEM;
  b->b = new (std::nothrow) mpsft::Batch(count, n, opts, precision);
  if (!b->b || b->b->status() != MPSFT_OK) {
    const int rc = b->b ? b->b->status() : MPSFT_ERR_NOMEM;
    delete b->b;
    delete b;
    return rc;
  }
  *batch = b;
  return MPSFT_OK;
}

int mpsft_batch_destroy(mpsft_batch *batch) {
  if (!batch) return MPSFT_ERR_INVALID_ARG;
  delete batch->b;
  delete batch;
  return MPSFT_OK;
}

int mpsft_batch_solve(mpsft_batch *batch, const void *const *x_dev, const uint32_t *seeds,
                      int32_t *iters, int32_t *status, void *stream) {
  if (!batch || !x_dev || !seeds) return MPSFT_ERR_INVALID_ARG;
  return batch->b->Solve(x_dev, seeds, iters, status, S(stream));
}

int mpsft_batch_solve_profiled(mpsft_batch *batch, const void *const *x_dev,
                               const uint32_t *seeds, int32_t *iters, int32_t *status,
                               double *kernel_ms, int32_t *launches, void *stream) {
  if (!batch || !x_dev || !seeds || !kernel_ms || !launches) return MPSFT_ERR_INVALID_ARG;
  return batch->b->SolveProfiled(x_dev, seeds, iters, status, kernel_ms, launches, S(stream));
}

int mpsft_batch_autotune(mpsft_batch *batch, const void *const *x_dev, const uint32_t *seeds,
                         int32_t *residues, void *stream) {
  if (!batch || !x_dev || !seeds) return MPSFT_ERR_INVALID_ARG;
  return batch->b->Autotune(x_dev, seeds, residues, S(stream));
}

int mpsft_batch_modes_counts(mpsft_batch *batch, int64_t *counts) {
  if (!batch || !counts) return MPSFT_ERR_INVALID_ARG;
  return batch->b->ModesCounts(counts);
}

int mpsft_batch_modes_device(mpsft_batch *batch, int64_t *const *keys_dev,
                             void *const *coefs_dev, void *stream) {
  if (!batch || !keys_dev || !coefs_dev) return MPSFT_ERR_INVALID_ARG;
  return batch->b->ModesToDevice(keys_dev, coefs_dev, S(stream));
}

int mpsft_batch_set_record(mpsft_batch *batch, int32_t on) {
  if (!batch) return MPSFT_ERR_INVALID_ARG;
  return batch->b->SetRecord(on != 0);
}

int mpsft_batch_info(mpsft_batch *batch, int32_t i, int64_t *n, int32_t *precision,
                     int64_t *record_id) {
  if (!batch || i < 0 || i >= batch->b->count()) return MPSFT_ERR_INVALID_ARG;
  Engine &e = batch->b->engine(i);
  if (n) *n = e.n();
  if (precision) *precision = e.precision();
  if (record_id) *record_id = e.RecordId();
  return MPSFT_OK;
}

int mpsft_batch_replay(mpsft_batch *batch, int32_t i, const void *x_dev, int precision,
                       void *stream) {
  if (!batch || !x_dev || i < 0 || i >= batch->b->count()) return MPSFT_ERR_INVALID_ARG;
  return batch->b->engine(i).Replay(x_dev, precision, S(stream));
}

int mpsft_batch_adjoint(mpsft_batch *batch, const int64_t *const *keys_dev,
                        const void *const *grad_dev, const int64_t *m, void *const *out_dev,
                        int precision, void *stream) {
  if (!batch || !keys_dev || !grad_dev || !m || !out_dev) return MPSFT_ERR_INVALID_ARG;
  return batch->b->Adjoint(keys_dev, grad_dev, m, out_dev, precision, S(stream));
}

int mpsft_batch_modes_count(mpsft_batch *batch, int32_t i, int64_t *count) {
  if (!batch || !count || i < 0 || i >= batch->b->count()) return MPSFT_ERR_INVALID_ARG;
  return batch->b->engine(i).ModesCount(count);
}

int mpsft_batch_modes_get(mpsft_batch *batch, int32_t i, int64_t *keys, double *coefs,
                          int64_t cap, int64_t *count) {
  if (!batch || !count || i < 0 || i >= batch->b->count() || (cap > 0 && (!keys || !coefs)))
    return MPSFT_ERR_INVALID_ARG;
  return batch->b->engine(i).ModesGet(keys, coefs, cap, count);
}

int mpsft_batch_log_get(mpsft_batch *batch, int32_t i, int32_t *bins, int32_t *found,
                        int32_t cap, int32_t *n_iter) {
  if (!batch || !n_iter || i < 0 || i >= batch->b->count()) return MPSFT_ERR_INVALID_ARG;
  return batch->b->engine(i).LogGet(bins, found, cap, n_iter);
}

}  // extern "C"
