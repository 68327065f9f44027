// batch.cpp -- host side of the batched form: one Engine per signal (its
// tables, buffers, transforms and loop state), one launch per stage over
// all of them.
#include "batch.h"

#include <cstring>

#include <algorithm>
#include <chrono>
#include <cstdlib>

namespace mpsft {

namespace {

#define HIP_OK(expr)                                 \
  do {                                               \
    if ((expr) != hipSuccess) return MPSFT_ERR_HIP;  \
  } while (0)

// Iterations whose transforms a batched solve draws before its first launch.
constexpr int32_t kBatchFirstDraws = 4;

// Workgroups of each packed batched launch (default 24576: 96 per CU of
// MI355X's 256, so that a launch of up to that many units gives each
// workgroup one unit and leaves the balancing to the dispatcher; measured
// on 64 signals of n = 2^22, S = 1000: 4096: 118.0, 8192: 116.7-116.9,
// 16384: 116.3-116.6, 24576: 115.9, 32768: 116.0, 65536: 116.4 us per
// solve; on 256 uneven signals 8192 and 24576 within 1%), and bins per
// k_identify unit (default 16: four per wavefront), each overridable for
// measurement by MPSFT_PACKED_GRID / MPSFT_IDENT_BINS.
int32_t EnvInt(const char *name, int32_t def, int32_t lo, int32_t hi) {
  const char *e = std::getenv(name);
  if (!e) return def;
  const int v = std::atoi(e);
  return v >= lo && v <= hi ? v : def;
}

template <typename T>
int AllocDev(T **p, size_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T)) !=
      hipSuccess) {
    *p = nullptr;
    return MPSFT_ERR_NOMEM;
  }
  return MPSFT_OK;
}

}  // namespace

Batch::Batch(int32_t count, const int64_t *n, const mpsft_solve_options *opts, int precision)
    : precision_(precision) {
  if (count <= 0 || !n || !opts || (precision != MPSFT_FP64 && precision != MPSFT_MIXED)) {
    status_ = MPSFT_ERR_INVALID_ARG;
    return;
  }
  for (int32_t i = 0; i < count; ++i) {
    if (opts[i].k < 0 || n[i] < 2) {
      status_ = MPSFT_ERR_INVALID_ARG;
      return;
    }
    int64_t cap = 8 * int64_t(opts[i].k) + 4096;
    if (cap > n[i]) cap = n[i];
    eng_.emplace_back(new Engine(n[i], precision, cap));
    if (eng_.back()->status() != MPSFT_OK) {
      status_ = eng_.back()->status();
      return;
    }
    // wider gather-fold tiles when many signals fill the launches: about
    // 2048 tiles per trial over the batch (measured on 64 config-2 signals:
    // 512 tiles per signal 152.2, 64 per signal 144.7 us per solve)
    eng_.back()->tile_target_ = std::max(16, std::min(kTargetTiles, 2048 / count));
    opts_.push_back(opts[i]);
  }
  args_.resize(count);
  built_gen_.assign(count, -1);
  ident_bins_ = EnvInt("MPSFT_IDENT_BINS", 16, 4, kIdentKeys) / 4 * 4;
  int32_t *hp = nullptr;
  if (AllocDev(&d_ctxs_, count) != MPSFT_OK || AllocDev(&d_args_, count) != MPSFT_OK ||
      AllocDev(&d_counters_, 2) != MPSFT_OK || AllocDev(&d_results_, count) != MPSFT_OK ||
      hipHostMalloc(reinterpret_cast<void **>(&h_results_), count * sizeof(BatchResultDev)) !=
          hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&hp), 4 * sizeof(int32_t),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void **>(&d_host_), hp, 0) != hipSuccess) {
    status_ = MPSFT_ERR_NOMEM;
    return;
  }
  h_host_ = hp;
  hp[0] = hp[1] = hp[3] = 0;
  hp[2] = 1;
}

Batch::~Batch() {
  (void)hipDeviceSynchronize();
  if (graph_) (void)hipGraphExecDestroy(graph_);
  if (graph_small_) (void)hipGraphExecDestroy(graph_small_);
  (void)hipFree(d_ctxs_);
  (void)hipFree(d_args_);
  (void)hipFree(d_map_);
  (void)hipFree(d_nblk_ident_);
  (void)hipFree(d_work_map_);
  (void)hipFree(d_work_i32_);
  (void)hipFree(d_counters_);
  (void)hipFree(d_results_);
  (void)hipFree(d_adj_sig_);
  (void)hipFree(d_adj_map_fft_);
  (void)hipFree(d_modes_out_);
  (void)hipFree(d_adj_steps_);
  (void)hipFree(d_stage_);
  if (p_stage_) (void)hipHostFree(p_stage_);
  if (stage_ev_) (void)hipEventDestroy(stage_ev_);
  if (h_results_) (void)hipHostFree(h_results_);
  if (h_host_) (void)hipHostFree(const_cast<int32_t *>(h_host_));
}

// Per-signal set-up (Engine::PrepareSolve) and the Ctx array the batched
// kernels read.  The batched kernels cover bin FFTs of one workgroup per
// row only (B <= kMaxLdsBins).
int Batch::Prepare(const void *const *x, const uint32_t *seeds, hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  if (!x || !seeds) return MPSFT_ERR_INVALID_ARG;
  seeds_.assign(seeds, seeds + eng_.size());
  // the previous solve's staged uploads have been copied (its end drained
  // the stream; the event covers a solve that stopped early)
  if (stage_ev_ && hipEventSynchronize(stage_ev_) != hipSuccess) return MPSFT_ERR_HIP;
  stage_used_ = 0;
  uploads_.assign(eng_.size(), TfUpload{});
  std::vector<Ctx> ctxs(eng_.size());
  for (size_t i = 0; i < eng_.size(); ++i) {
    Engine &e = *eng_[i];
    e.rec_valid_ = false;
    e.defer_upload_ = &uploads_[i];
    // the first two graphs' iterations only: Run draws the rest while the
    // device works (every signal's first 32 drawn here kept the device idle
    // between solves)
    int rc = e.PrepareSolve(x[i], opts_[i], seeds[i], s, &args_[i], kBatchFirstDraws);
    e.defer_upload_ = nullptr;
    if (rc != MPSFT_OK) return rc;
    if (e.record_ && (rc = e.EnsureRecord()) != MPSFT_OK) return rc;
    if (e.fft_big_) return MPSFT_ERR_UNSUPPORTED;
    if ((e.precision_ == MPSFT_MIXED) != (precision_ == MPSFT_MIXED)) return MPSFT_ERR_INVALID_ARG;
    ctxs[i] = SignalCtx(e);
  }
  // the transforms, the Ctx array and the loop's initial arguments in one
  // staged copy and one scatter
  const int urc = FlushUploads(s, Blob{ctxs.data(), d_ctxs_, ctxs.size() * sizeof(Ctx)},
                               Blob{args_.data(), d_args_, args_.size() * sizeof(InitArgs)});
  if (urc != MPSFT_OK) return urc;
  return Build(s);
}

// Block maps of the three batched launches and the graph of two batched
// iterations; redone only when an engine's shapes changed.
int Batch::Build(hipStream_t s) {
  bool same = graph_ != nullptr && !graph_stale_;
  for (size_t i = 0; i < eng_.size() && same; ++i) same = built_gen_[i] == eng_[i]->shape_gen_;
  if (same) return MPSFT_OK;
  // the maps below are rewritten by blocking copies: the queued work that
  // may read them (an autotune pass's launches, the uploads) ends first
  HIP_OK(hipStreamSynchronize(s));
  graph_stale_ = false;
  std::vector<int2> map_time, map_fft, map_ident;
  std::vector<int32_t> nblk_ident(eng_.size());
  size_t lds_time = 0, lds_fft = 0, lds_fft_small = 0;
  for (size_t i = 0; i < eng_.size(); ++i) {
    Engine &e = *eng_[i];
    const int32_t si = int32_t(i);
    // units of k_bin_time: rounded up to 8 per signal (XCD alignment)
    const int32_t nt = (kPrepBlocks + e.TilesCap() + 7) & ~7;
    for (int32_t b = 0; b < nt; ++b) map_time.push_back(make_int2(si, b));
    const int32_t nf = e.trials_cap_ * e.rows_cap_;
    for (int32_t b = 0; b < nf; ++b) map_fft.push_back(make_int2(si, b));
    nblk_ident[i] = e.IdentifyBlocks();
    for (int32_t b = 0; b < nblk_ident[i]; ++b) map_ident.push_back(make_int2(si, b));
    lds_time = std::max(lds_time, e.LdsBinTime(precision_ == MPSFT_MIXED));
    lds_fft = std::max(lds_fft, e.LdsFft());
    lds_fft_small = std::max(lds_fft_small, e.lds_fft_small_);
    built_gen_[i] = e.shape_gen_;
  }
  // the adjoint's DFT-and-terms launch: kPrepBlocks terms blocks, then the
  // DFT blocks, per signal
  std::vector<int2> map_adj;
  for (size_t i = 0; i < eng_.size(); ++i)
    for (int32_t b = 0; b < kPrepBlocks + eng_[i]->trials_cap_ * eng_[i]->rows_cap_; ++b)
      map_adj.push_back(make_int2(int32_t(i), b));
  if (AllocDev(&d_adj_map_fft_, map_adj.size()) != MPSFT_OK) return MPSFT_ERR_NOMEM;
  HIP_OK(hipMemcpy(d_adj_map_fft_, map_adj.data(), map_adj.size() * sizeof(int2),
                   hipMemcpyHostToDevice));
  adj_nb_fft_ = int32_t(map_adj.size());
  const size_t total = map_time.size() + map_fft.size() + map_ident.size();
  const size_t K = eng_.size();
  if (AllocDev(&d_map_, total) != MPSFT_OK || AllocDev(&d_nblk_ident_, K) != MPSFT_OK ||
      AllocDev(&d_work_map_, total + map_fft.size()) != MPSFT_OK ||
      AllocDev(&d_work_i32_, (2 * kUnitKinds + 1) * K + kUnitKinds) != MPSFT_OK)
    return MPSFT_ERR_NOMEM;
  HIP_OK(hipMemcpy(d_map_, map_time.data(), map_time.size() * sizeof(int2),
                   hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_map_ + map_time.size(), map_fft.data(), map_fft.size() * sizeof(int2),
                   hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_map_ + map_time.size() + map_fft.size(), map_ident.data(),
                   map_ident.size() * sizeof(int2), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_nblk_ident_, nblk_ident.data(), nblk_ident.size() * sizeof(int32_t),
                   hipMemcpyHostToDevice));
  bl_.ctxs = d_ctxs_;
  bl_.map_time = d_map_;
  bl_.map_fft = d_map_ + map_time.size();
  bl_.map_ident = d_map_ + map_time.size() + map_fft.size();
  bl_.nblk_ident = d_nblk_ident_;
  bl_.nb_time = int32_t(map_time.size());
  bl_.nb_fft = int32_t(map_fft.size());
  bl_.nb_ident = int32_t(map_ident.size());
  // The packed launches: at most `grid` workgroups each (24576 by default),
  // each visiting units blockIdx.x, blockIdx.x + gridDim.x, ...
  const int32_t grid = EnvInt("MPSFT_PACKED_GRID", 24576, 64, 1 << 20);
  bl_.ident_bins = ident_bins_;
  int64_t trials_rows = 1;
  for (auto &e : eng_) trials_rows = std::max<int64_t>(trials_rows, int64_t(e->trials_cap_) * e->rows_cap_);
  bl_.col_slices = trials_rows <= 64 ? 1 : kColSlices;
  bl_.grid_time = std::max(8, std::min(bl_.nb_time, grid) & ~7);  // a multiple of 8
  bl_.grid_fft = std::max(1, std::min(bl_.nb_fft, grid));
  bl_.grid_ident = std::max(1, std::min(bl_.nb_ident, grid));
  bl_.count = int32_t(K);
  bl_.work.cnt = d_work_i32_;
  bl_.work.off = d_work_i32_ + kUnitKinds * K;
  bl_.work.rows = d_work_i32_ + 2 * kUnitKinds * K;
  bl_.work.total = d_work_i32_ + (2 * kUnitKinds + 1) * K;
  bl_.work.map[0] = d_work_map_;
  bl_.work.map[1] = d_work_map_ + map_time.size();
  bl_.work.map[2] = d_work_map_ + map_time.size() + map_fft.size();
  bl_.work.map[3] = d_work_map_ + total;  // the large-path FFT units, as many as map_fft
  bl_.lds_time = lds_time;
  bl_.lds_fft = lds_fft;
  bl_.lds_fft_small = lds_fft_small;
  bl_.counters = d_counters_;
  bl_.host = d_host_;
  bl_.mixed = precision_ == MPSFT_MIXED ? 1 : 0;

  for (hipGraphExec_t *ge : {&graph_, &graph_small_})
    if (*ge) {
      (void)hipGraphExecDestroy(*ge);
      *ge = nullptr;
    }
  // graph_ with every launch, graph_small_ without the large-path bin FFT
  bl_.large_fft = 1;
  for (int large = 1; large >= 0; --large) {
    BatchLaunch b = bl_;
    b.large_fft = large;
    hipStream_t cs;
    HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return MPSFT_ERR_HIP;
    }
    const hipError_t e0 = LaunchBatchIteration(b, 0, cs);
    const hipError_t e1 = LaunchBatchIteration(b, 1, cs);
    hipGraph_t g;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (e0 != hipSuccess || e1 != hipSuccess || ec != hipSuccess) return MPSFT_ERR_HIP;
    hipGraphExec_t *ge = large ? &graph_ : &graph_small_;
    const hipError_t ei = hipGraphInstantiate(ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      *ge = nullptr;
      return MPSFT_ERR_HIP;
    }
  }
  (void)s;
  return MPSFT_OK;
}

// Signal i's Ctx in a batch: its engine's, counting its stop in the batch's
// counters, and without the engine's host-mapped progress words, which only
// a single-signal launch loop reads (the plan kernel would otherwise store
// to host memory for every signal on every iteration).
Ctx Batch::SignalCtx(const Engine &e) const {
  Ctx c = e.MakeCtx();
  c.batch_done = d_counters_ + 1;
  c.progress = nullptr;
  return c;
}

int Batch::FlushUploads(hipStream_t s, const Blob &a, const Blob &b) {
  static_assert(sizeof(DevTf) % 8 == 0 && sizeof(Ctx) % 8 == 0 && sizeof(InitArgs) % 8 == 0,
                "staged payloads are copied in 8-byte words");
  auto up16 = [](size_t n) { return (n + 15) & ~size_t(15); };
  const bool blobs = a.bytes > 0 || b.bytes > 0;
  int32_t m = blobs ? 1 : 0;
  size_t bytes = up16(a.bytes) + up16(b.bytes);
  for (const TfUpload &u : uploads_)
    if (u.ntf > 0 || u.nq > 0) {
      ++m;
      bytes += up16(size_t(u.ntf) * sizeof(DevTf)) + up16(size_t(u.nq) * sizeof(int64_t));
    }
  if (m == 0) return MPSFT_OK;
  const size_t head = up16(size_t(m) * sizeof(TfScatterDesc));
  bytes += head;
  if (stage_used_ + bytes > stage_cap_) {
    // a pending copy may still read the staged bytes: wait before reusing them
    if (stage_ev_ && hipEventSynchronize(stage_ev_) != hipSuccess) return MPSFT_ERR_HIP;
    stage_used_ = 0;
    if (bytes > stage_cap_) {
      const size_t cap = std::max(bytes, std::max<size_t>(2 * stage_cap_, 1 << 16));
      (void)hipFree(d_stage_);
      if (p_stage_) (void)hipHostFree(p_stage_);
      d_stage_ = p_stage_ = nullptr;
      stage_cap_ = 0;
      if (hipMalloc(reinterpret_cast<void **>(&d_stage_), cap) != hipSuccess ||
          hipHostMalloc(reinterpret_cast<void **>(&p_stage_), cap) != hipSuccess)
        return MPSFT_ERR_NOMEM;
      stage_cap_ = cap;
    }
  }
  if (!stage_ev_ && hipEventCreate(&stage_ev_) != hipSuccess) return MPSFT_ERR_HIP;
  unsigned char *h = p_stage_ + stage_used_;
  TfScatterDesc *desc = reinterpret_cast<TfScatterDesc *>(h);
  size_t off = head;
  int32_t k = 0;
  // one payload: its bytes staged at off, its descriptor slot filled
  auto put = [&](TfScatterDesc &d, int r, void *dst, const void *src, size_t n) {
    d.dst[r] = dst;
    d.off[r] = int64_t(off);
    d.words[r] = int64_t(n / 8);
    if (n) std::memcpy(h + off, src, n);
    off += up16(n);
  };
  if (blobs) {
    TfScatterDesc &d = desc[k++];
    put(d, 0, a.dst, a.src, a.bytes);
    put(d, 1, b.dst, b.src, b.bytes);
  }
  for (const TfUpload &u : uploads_) {
    if (u.ntf <= 0 && u.nq <= 0) continue;
    TfScatterDesc &d = desc[k++];
    put(d, 0, u.tf_dst, u.tf_src, size_t(u.ntf) * sizeof(DevTf));
    put(d, 1, u.q_dst, u.q_src, size_t(u.nq) * sizeof(int64_t));
  }
  HIP_OK(hipMemcpyAsync(d_stage_ + stage_used_, h, bytes, hipMemcpyHostToDevice, s));
  HIP_OK(LaunchTfScatter(d_stage_ + stage_used_, m, s));
  HIP_OK(hipEventRecord(stage_ev_, s));
  stage_used_ += bytes;
  return MPSFT_OK;
}

// Every signal's transforms and q's drawn for iterations [0, need); when an
// engine's arrays must grow, the stream is drained first and the Ctx array
// re-uploaded (the kernels read it).
int Batch::EnsureAhead(int32_t need, int32_t max_iter, hipStream_t s) {
  int32_t have = max_iter;
  for (auto &e : eng_) have = std::min(have, e->tf_count_);
  if (need <= have) return MPSFT_OK;
  bool grow = false;
  const int32_t want = std::min(max_iter, std::max(need, 2 * have));
  for (auto &e : eng_) grow = grow || want > e->tf_cap_;
  if (grow && hipStreamSynchronize(s) != hipSuccess) return MPSFT_ERR_HIP;
  std::vector<Ctx> ctxs(eng_.size());
  uploads_.assign(eng_.size(), TfUpload{});
  for (size_t i = 0; i < eng_.size(); ++i) {
    eng_[i]->defer_upload_ = &uploads_[i];
    const int rc = eng_[i]->EnsureTransforms(seeds_[i], want, opts_[i].trials, s);
    eng_[i]->defer_upload_ = nullptr;
    if (rc != MPSFT_OK) return rc;
    ctxs[i] = SignalCtx(*eng_[i]);
  }
  const int urc = FlushUploads(s, grow ? Blob{ctxs.data(), d_ctxs_, ctxs.size() * sizeof(Ctx)}
                                      : Blob{});
  if (urc != MPSFT_OK) return urc;
  return MPSFT_OK;
}

// Launches graphs of two batched iterations, at most `lookahead` ahead of
// the device, until every signal has stopped (or max_iter).  Transforms are
// drawn ahead for all signals; when an engine's arrays must grow, the stream
// is drained first and the Ctx array re-uploaded (the graph reads it).
int Batch::Run(hipStream_t s, int32_t max_iter) {
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(s);
    return code;
  };
  const int32_t K = count();
  const int32_t lookahead = 2;
  int32_t launched = 0;
  const auto stall = StallWait();
  auto last_change = std::chrono::steady_clock::now();
  int32_t seen = 0;
  for (uint64_t spin = 0;; ++spin) {
    const int32_t started = h_host_[0];
    if (h_host_[1] >= K || launched > max_iter) break;
    if (launched - started < lookahead) {
      const int rc = EnsureAhead(std::min(max_iter, launched + 3), max_iter, s);
      if (rc != MPSFT_OK) return fail(rc);
      // without the large-path bin FFT once a started iteration had none
      // (a signal's B never grows, so none come back this solve)
      const hipGraphExec_t g = h_host_[2] == 0 && graph_small_ ? graph_small_ : graph_;
      if (hipGraphLaunch(g, s) != hipSuccess) return fail(MPSFT_ERR_HIP);
      launched += 2;
      continue;
    }
    if (started != seen) {
      seen = started;
      last_change = std::chrono::steady_clock::now();
    } else if ((spin & 4095) == 0 &&
               std::chrono::steady_clock::now() - last_change > stall) {
      // No report for a while: wait for the queued work, then go on.
      if (hipStreamSynchronize(s) != hipSuccess) return fail(MPSFT_ERR_HIP);
      last_change = std::chrono::steady_clock::now();
    }
  }
  return MPSFT_OK;
}

int Batch::Solve(const void *const *x, const uint32_t *seeds, int32_t *iters, int32_t *status,
                 hipStream_t s) {
  solved_ = false;
  int rc = Prepare(x, seeds, s);
  if (rc != MPSFT_OK) return rc;
  int32_t max_iter = 0;
  for (auto &a : args_) max_iter = std::max(max_iter, a.max_iter);
  h_host_[0] = 0;
  h_host_[1] = 0;
  h_host_[2] = 1;  // large-path bin-FFT units: assumed until a plan reports none
  HIP_OK(LaunchBatchReset(d_ctxs_, d_args_, count(), d_counters_, s));
  rc = Run(s, max_iter);
  if (rc != MPSFT_OK) return rc;
  return Finish(iters, status, s);
}

int Batch::SolveProfiled(const void *const *x, const uint32_t *seeds, int32_t *iters,
                         int32_t *status, double *kernel_ms, int32_t *launches, hipStream_t s) {
  if (!kernel_ms || !launches) return MPSFT_ERR_INVALID_ARG;
  solved_ = false;
  int rc = Prepare(x, seeds, s);
  if (rc != MPSFT_OK) return rc;
  int32_t max_iter = 0;
  for (auto &a : args_) max_iter = std::max(max_iter, a.max_iter);
  h_host_[0] = 0;
  h_host_[1] = 0;
  h_host_[2] = 1;  // large-path bin-FFT units: assumed until a plan reports none
  HIP_OK(LaunchBatchReset(d_ctxs_, d_args_, count(), d_counters_, s));
  rc = RunProfiled(s, max_iter, kernel_ms, launches);
  if (rc != MPSFT_OK) return rc;
  return Finish(iters, status, s);
}

// The batched iterations one at a time, each stage once between events on
// the stream, until every signal has stopped.
int Batch::RunProfiled(hipStream_t s, int32_t max_iter, double *kernel_ms, int32_t *launches) {
  hipEvent_t ev[kBatchStages + 1];
  for (auto &e : ev) HIP_OK(hipEventCreate(&e));
  for (int i = 0; i < kBatchStages; ++i) kernel_ms[i] = 0.0;
  *launches = 0;
  int out = MPSFT_OK;
  for (int32_t it = 0; it <= max_iter && out == MPSFT_OK; ++it) {
    if ((out = EnsureAhead(std::min(max_iter, it + 2), max_iter, s)) != MPSFT_OK) break;
    bool ok = hipEventRecord(ev[0], s) == hipSuccess;
    for (int st = 0; st < kBatchStages && ok; ++st)
      ok = LaunchBatchStage(bl_, it & 1, st, s) == hipSuccess &&
           hipEventRecord(ev[st + 1], s) == hipSuccess;
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
      out = MPSFT_ERR_HIP;
      break;
    }
    for (int st = 0; st < kBatchStages; ++st) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev[st], ev[st + 1]) != hipSuccess) out = MPSFT_ERR_HIP;
      kernel_ms[st] += ms;
    }
    ++*launches;
    if (h_host_[1] >= count()) break;  // this iteration's plan found every signal stopped
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (out != MPSFT_OK) (void)hipStreamSynchronize(s);
  return out;
}

// Per-signal results of the batched solve just run.
int Batch::Finish(int32_t *iters, int32_t *status, hipStream_t s) {
  HIP_OK(LaunchBatchResults(d_ctxs_, count(), d_results_, s));
  HIP_OK(hipMemcpyAsync(h_results_, d_results_, count() * sizeof(BatchResultDev),
                        hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  solved_ = true;
  int worst = MPSFT_OK;
  for (int32_t i = 0; i < count(); ++i) {
    const BatchResultDev &r = h_results_[i];
    int st = MPSFT_OK;
    int32_t it = r.iters;
    if (!r.stopped) {
      st = MPSFT_ERR_NOT_CONVERGED;
      it = r.last_it + 1;
    } else if (r.flags & kFlagOverflow) {
      st = MPSFT_ERR_CAPACITY;
    } else if (r.stop_stale < opts_[i].max_stale_iter) {
      st = MPSFT_ERR_NOT_CONVERGED;
    }
    eng_[i]->FinishRecord(opts_[i], seeds_[i], st, it, r.stopped != 0);
    if (iters) iters[i] = it;
    if (status) status[i] = st;
    if (st != MPSFT_OK && worst == MPSFT_OK) worst = st;
  }
  return worst;
}

int Batch::ModesCounts(int64_t *counts) const {
  if (status_ != MPSFT_OK) return status_;
  if (!counts) return MPSFT_ERR_INVALID_ARG;
  if (!solved_) return MPSFT_ERR_SHAPE;
  for (int32_t i = 0; i < count(); ++i) counts[i] = h_results_[i].count;
  return MPSFT_OK;
}

int Batch::ModesToDevice(int64_t *const *keys, void *const *coefs, hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  if (!keys || !coefs) return MPSFT_ERR_INVALID_ARG;
  if (!solved_) return MPSFT_ERR_SHAPE;
  std::vector<ModesOut> out(count());
  for (int32_t i = 0; i < count(); ++i) {
    if (h_results_[i].count > 0 && (!keys[i] || !coefs[i])) return MPSFT_ERR_INVALID_ARG;
    out[i].keys = keys[i];
    out[i].coefs = static_cast<double2 *>(coefs[i]);
    out[i].cap = h_results_[i].count;
  }
  if (!d_modes_out_ && AllocDev(&d_modes_out_, count()) != MPSFT_OK) return MPSFT_ERR_NOMEM;
  HIP_OK(hipMemcpyAsync(d_modes_out_, out.data(), out.size() * sizeof(ModesOut),
                        hipMemcpyHostToDevice, s));
  HIP_OK(LaunchBatchModesOut(d_ctxs_, d_modes_out_, count(), s));
  // out is read by the copy above
  HIP_OK(hipStreamSynchronize(s));
  return MPSFT_OK;
}

int Batch::SetRecord(bool on) {
  if (status_ != MPSFT_OK) return status_;
  for (auto &e : eng_) {
    const int rc = e->SetRecord(on);
    if (rc != MPSFT_OK) return rc;
  }
  return MPSFT_OK;
}

int Batch::Adjoint(const int64_t *const *keys, const void *const *grad, const int64_t *m,
                   void *const *out, int precision, hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  if (!keys || !grad || !m || !out) return MPSFT_ERR_INVALID_ARG;
  if (precision != MPSFT_FP64 && precision != MPSFT_MIXED) return MPSFT_ERR_UNSUPPORTED;
  const int32_t K = count();
  if (!graph_) return MPSFT_ERR_SHAPE;  // no batched solve yet
  // A replay since the solve may have grown an engine's arrays: the block
  // maps and the Ctx array must describe the engines as they are now.
  int rc0 = Build(s);
  if (rc0 != MPSFT_OK) return rc0;
  std::vector<Ctx> ctxs(K);
  for (int32_t i = 0; i < K; ++i) {
    ctxs[i] = SignalCtx(*eng_[i]);
  }
  HIP_OK(hipMemcpyAsync(d_ctxs_, ctxs.data(), ctxs.size() * sizeof(Ctx), hipMemcpyHostToDevice, s));
  std::vector<AdjSig> sig(K);
  int32_t max_steps = 0;
  int64_t max_m = 1, max_n = 1, max_lags = 0;
  int any_tab = 0, any_gen = 0, small_only = 1;
  for (int32_t i = 0; i < K; ++i) {
    Engine &e = *eng_[i];
    if (m[i] < 0 || (m[i] > 0 && (!keys[i] || !grad[i])) || !out[i]) return MPSFT_ERR_INVALID_ARG;
    if (!e.rec_valid_ || !e.d_rec_ || e.rec_B_cap_ != e.B_cap_) return MPSFT_ERR_SHAPE;
    const int rc = e.EnsureAdjBuffers(e.rec_iters_);
    if (rc != MPSFT_OK) return rc;
    AdjSig &g = sig[i];
    g.lam = e.d_lam_;
    g.fbar = e.d_fbar_;
    g.taps = e.d_taps_;
    g.taps_stride = e.taps_stride_;
    g.lags = e.lags_len_ ? e.d_lags_ : nullptr;
    g.lags_stride = e.lags_stride_;
    max_lags = std::max(max_lags, e.lags_stride_);
    (g.lags ? any_tab : any_gen) = 1;
    if (e.M_cap_ > (1 << kFftSmallLogM)) small_only = 0;
    g.rec = e.d_rec_;
    g.keys = keys[i];
    g.grad = static_cast<const double2 *>(grad[i]);
    g.out = out[i];
    g.n = e.n_;
    g.m = m[i];
    g.iters = e.rec_iters_;
    g.trials = e.rec_opt_.trials;
    max_steps = std::max(max_steps, e.rec_iters_);
    max_m = std::max(max_m, m[i]);
    max_n = std::max(max_n, e.n_);
  }
  if (!d_adj_sig_ && AllocDev(&d_adj_sig_, K) != MPSFT_OK) return MPSFT_ERR_NOMEM;
  const size_t nsteps = size_t(K) * std::max(max_steps, 1);
  if (nsteps > adj_steps_cap_) {
    if (AllocDev(&d_adj_steps_, nsteps) != MPSFT_OK) return MPSFT_ERR_NOMEM;
    adj_steps_cap_ = nsteps;
  }
  HIP_OK(hipMemcpyAsync(d_adj_sig_, sig.data(), sig.size() * sizeof(AdjSig), hipMemcpyHostToDevice,
                        s));
  AdjLaunch a{};
  a.ctxs = d_ctxs_;
  a.sig = d_adj_sig_;
  a.steps = d_adj_steps_;
  a.count = K;
  a.max_steps = max_steps;
  a.map_bins = bl_.map_ident;
  a.nblk_bins = bl_.nblk_ident;
  a.nb_bins = bl_.nb_ident;
  a.map_fft = d_adj_map_fft_;
  a.nb_fft = adj_nb_fft_;
  a.lds_fft = bl_.lds_fft;
  a.per_small = int32_t(std::min<int64_t>((max_m + kThreads - 1) / kThreads, 16));
  a.per_gather = int32_t(std::min<int64_t>(4096, (max_n + 1023) / 1024));
  a.per_lags = LagBlocks(max_lags);
  a.fft_small_only = small_only;
  a.any_tab = any_tab;
  a.any_gen = any_gen;
  a.mixed = precision == MPSFT_MIXED ? 1 : 0;
  HIP_OK(LaunchAdj(a, s));
  // ctxs and sig are read by the copies above, the outputs by the caller:
  // wait for all.
  HIP_OK(hipStreamSynchronize(s));
  return MPSFT_OK;
}

int Batch::Autotune(const void *const *x, const uint32_t *seeds, int32_t *chosen,
                    hipStream_t s) {
  if (status_ != MPSFT_OK) return status_;
  // Whole batched solves of these signals, timed with events on the launch
  // stream (the first solve of each setting also builds tables and the
  // graph; the best of the next kTimed is its time), in a coordinate search:
  // k_identify bins per unit (8, 16, 32), then the gather-fold tile width of
  // B >= kSmallTileB (0: the default per B, or 1, 2, 4, 8 residues), then
  // that of smaller B (the late iterations').  A setting replaces the best
  // so far only when it is kMargin faster: the settings differ by about a
  // percent on full GPUs, which one solve's timing does not resolve.  None
  // of them changes a result, only how the work is split.
  constexpr int kTimed = 3;
  constexpr float kMargin = 0.98f;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  int rc = MPSFT_OK;
  auto timed = [&](float *ms) {
    *ms = 0;
    for (int rep = 0; rep <= kTimed && rc == MPSFT_OK; ++rep) {
      float t = 0;
      if (hipEventRecord(e0, s) != hipSuccess) rc = MPSFT_ERR_HIP;
      const int r = Solve(x, seeds, nullptr, nullptr, s);
      if (r != MPSFT_OK && r != MPSFT_ERR_NOT_CONVERGED) rc = r;
      if (rc == MPSFT_OK && (hipEventRecord(e1, s) != hipSuccess ||
                             hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        rc = MPSFT_ERR_HIP;
      if (rep > 0 && (rep == 1 || t < *ms)) *ms = t;  // rep 0 warms up
    }
  };
  auto set_widths = [&](int32_t wide, int32_t narrow) {
    for (auto &e : eng_)
      if (rc == MPSFT_OK) rc = e->SetTileResidues(wide, narrow);
  };
  const int32_t kWidths[] = {0, 1, 2, 4, 8};
  const int32_t saved_bins = ident_bins_;
  int32_t bins = saved_bins, wide = 0, narrow = 0;
  float best = 0, ms = 0;
  bool have = false;
  set_widths(0, 0);
  for (const int32_t nb : {saved_bins, 8, 16, 32}) {  // the current setting first
    if (rc != MPSFT_OK) break;
    if (have && nb == saved_bins) continue;
    ident_bins_ = nb;
    graph_stale_ = true;  // the launch parameters changed: capture again
    timed(&ms);
    if (rc == MPSFT_OK && (!have || ms < kMargin * best)) {
      best = ms;
      bins = nb;
      have = true;
    }
  }
  ident_bins_ = bins;
  graph_stale_ = true;
  for (int cls = 0; cls < 2 && rc == MPSFT_OK; ++cls) {
    int32_t keep = 0;
    for (const int32_t R : kWidths) {
      if (rc != MPSFT_OK) break;
      if (R == 0) continue;  // the current setting, timed above
      set_widths(cls == 0 ? R : wide, cls == 0 ? narrow : R);
      timed(&ms);
      if (rc == MPSFT_OK && ms < kMargin * best) {
        best = ms;
        keep = R;
      }
    }
    (cls == 0 ? wide : narrow) = keep;
    set_widths(wide, narrow);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != MPSFT_OK) {
    ident_bins_ = saved_bins;
    graph_stale_ = true;
    return rc;
  }
  if (chosen) *chosen = wide;
  return MPSFT_OK;
}

}  // namespace mpsft
