// batch.h -- the batched form of the MPSFT solve: many independent signals,
// each with its own n, options and seed, recovered together.  Every stage
// of an iteration (gather-fold, bin FFT, identify) is one launch over all
// signals; each signal keeps its own loop state (Demo1::Run,
// src/demo1.cc:36-59) and stops on its own, and a graph of two batched
// iterations is replayed until every signal has stopped.
#pragma once

#include <memory>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace mpsft {

// A host array bound for device memory (a staged upload).
struct Blob {
  const void *src = nullptr;
  void *dst = nullptr;
  size_t bytes = 0;
};

class Batch {
 public:
  Batch(int32_t count, const int64_t *n, const mpsft_solve_options *opts, int precision);
  ~Batch();
  int status() const { return status_; }
  int32_t count() const { return int32_t(eng_.size()); }
  Engine &engine(int32_t i) { return *eng_[i]; }

  // Demo1::Run for every signal: x[i] is signal i (device, n[i] samples),
  // seeds[i] its transform seed.  iters[i] / status[i] (host, may be null)
  // get each signal's iteration count and mpsft_status.
  int Solve(const void *const *x, const uint32_t *seeds, int32_t *iters, int32_t *status,
            hipStream_t s);
  // Tries k_bin_time tile widths 1, 2, 4, 8 on a short batched solve of
  // these signals and keeps the fastest; *chosen (may be null) gets it.
  int Autotune(const void *const *x, const uint32_t *seeds, int32_t *chosen, hipStream_t s);
  // Solve with each stage's launches bracketed by events (no graph):
  // kernel_ms[stage] is the summed milliseconds of that stage's kernel over
  // the iterations (0 plan, 1 gather-fold, 2 bin FFT, 3 identify) and
  // *launches the iterations launched.
  int SolveProfiled(const void *const *x, const uint32_t *seeds, int32_t *iters,
                    int32_t *status, double *kernel_ms, int32_t *launches, hipStream_t s);

  // The differentiable form over the batch: with recording on, a batched
  // solve keeps every signal's decisions (Engine::SetRecord); Adjoint then
  // applies every signal's adjoint at once: out[i] (device, n[i] samples of
  // `precision`) = F_i^H g_i for the m[i] keys / complex128 gradients
  // keys[i], grad[i] (device).
  int SetRecord(bool on);
  // Found-mode counts of the last batched solve (host, no device access),
  // and every signal's keys / coefficients (complex128) copied into device
  // buffers keys[i], coefs[i] of at least counts[i] entries, on stream s.
  int ModesCounts(int64_t *counts) const;
  int ModesToDevice(int64_t *const *keys, void *const *coefs, hipStream_t s);
  int Adjoint(const int64_t *const *keys, const void *const *grad, const int64_t *m,
              void *const *out, int precision, hipStream_t s);

 private:
  int Prepare(const void *const *x, const uint32_t *seeds, hipStream_t s);
  int Build(hipStream_t s);
  int Run(hipStream_t s, int32_t max_iter);
  int RunProfiled(hipStream_t s, int32_t max_iter, double *kernel_ms, int32_t *launches);
  int Finish(int32_t *iters, int32_t *status, hipStream_t s);
  int EnsureAhead(int32_t need, int32_t max_iter, hipStream_t s);
  // Copies every engine's deferred transform upload (uploads_), and up to
  // two host arrays (a, b) to device memory, in one staging copy and one
  // k_tf_scatter launch.
  int FlushUploads(hipStream_t s, const Blob &a = Blob{}, const Blob &b = Blob{});
  Ctx SignalCtx(const Engine &e) const;

  int status_ = MPSFT_OK;
  int precision_;
  std::vector<std::unique_ptr<Engine>> eng_;
  std::vector<mpsft_solve_options> opts_;
  std::vector<InitArgs> args_;
  std::vector<uint32_t> seeds_;
  std::vector<int64_t> built_gen_;  // engines' shape generations the maps were built for

  Ctx *d_ctxs_ = nullptr;
  InitArgs *d_args_ = nullptr;
  int2 *d_map_ = nullptr;           // map_time | map_fft | map_ident
  int32_t *d_nblk_ident_ = nullptr;
  int2 *d_work_map_ = nullptr;      // the packed maps of BatchWork
  int32_t *d_work_i32_ = nullptr;   // BatchWork cnt | off | rows | total
  int32_t *d_counters_ = nullptr;   // {iterations started, signals stopped}
  BatchResultDev *d_results_ = nullptr;
  BatchResultDev *h_results_ = nullptr;  // pinned
  volatile int32_t *h_host_ = nullptr;  // pinned, mapped: device copy of the counters
  int32_t *d_host_ = nullptr;
  BatchLaunch bl_{};
  hipGraphExec_t graph_ = nullptr;
  // the same two iterations without the large-path bin-FFT launch, for the
  // rest of a solve once an iteration has no large-path units
  hipGraphExec_t graph_small_ = nullptr;
  bool graph_stale_ = false;   // launch parameters changed since the graph's capture
  int32_t ident_bins_ = 16;    // k_identify bins per unit (autotuned)
  AdjSig *d_adj_sig_ = nullptr;
  int2 *d_adj_map_fft_ = nullptr;
  int32_t adj_nb_fft_ = 0;
  ModesOut *d_modes_out_ = nullptr;
  bool solved_ = false;  // h_results_ hold a finished batched solve
  AdjStep *d_adj_steps_ = nullptr;
  size_t adj_steps_cap_ = 0;
  // The transform uploads of a solve: each engine's deferred upload, then
  // staged (pinned host, device) and appended within a solve, so a pending
  // copy's bytes are never rewritten; stage_ev_ follows the last copy.
  std::vector<TfUpload> uploads_;
  unsigned char *p_stage_ = nullptr, *d_stage_ = nullptr;
  size_t stage_cap_ = 0, stage_used_ = 0;
  hipEvent_t stage_ev_ = nullptr;
};

}  // namespace mpsft
