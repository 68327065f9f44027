// Host emulation runtime for tests/emu/hip/hip_runtime.h (test
// infrastructure, see that header).
#include <hip/hip_runtime.h>

#include <sanitizer/common_interface_defs.h>
#include <ucontext.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace emu {

thread_local Index tid, bid;
thread_local dim3 bdim, gdim;

namespace {

// A workgroup runs as one fiber per thread on the device worker thread;
// fibers switch only at barriers and wave operations, so every access is
// sequential and ASan sees one coherent stack per fiber.
void Yield();

// Barrier over a set of fibers that may shrink: a fiber that returns from
// the kernel leaves every barrier it belonged to, as an ended wave stops
// counting at s_barrier.
struct Barrier {
  int count = 0, arrived = 0;
  uint64_t gen = 0;
  void Reset(int n) {
    count = n;
    arrived = 0;
  }
  void Wait() {
    const uint64_t g = gen;
    if (++arrived >= count) {
      arrived = 0;
      ++gen;
      return;
    }
    while (gen == g) Yield();
  }
  void Drop() {
    --count;
    if (count > 0 && arrived >= count) {
      arrived = 0;
      ++gen;
    }
  }
};

struct Wave {
  Barrier bar;
  double d[64];
  int i[64];
};

struct Block {
  void *lds = nullptr;
  Barrier bar;
  std::vector<Wave> waves;
};

struct Node {
  dim3 grid, block;
  size_t lds;
  std::function<void()> body;
};

constexpr size_t kStack = 256 * 1024;

struct Fiber {
  ucontext_t ctx;
  char *stack = nullptr;
  bool done = false;
};

struct Sched {
  ucontext_t main;
  const void *main_bottom = nullptr;
  size_t main_size = 0;
  std::vector<Fiber> fibers;
  unsigned cur = 0;
  const Node *nd = nullptr;
  unsigned b = 0;
  Block *blk = nullptr;
};

Sched *sched = nullptr;  // the running launch (launches do not nest)

void Enter(unsigned i) {
  tid = Index{i, 0, 0};
  bid = Index{sched->b, 0, 0};
  bdim = sched->nd->block;
  gdim = sched->nd->grid;
}

void Yield() {
  Sched *S = sched;
  Fiber &f = S->fibers[S->cur];
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, S->main_bottom, S->main_size);
  swapcontext(&f.ctx, &S->main);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
}

void Trampoline() {
  Sched *S = sched;
  __sanitizer_finish_switch_fiber(nullptr, &S->main_bottom, &S->main_size);
  const unsigned i = S->cur;
  S->nd->body();
  S->blk->bar.Drop();
  S->blk->waves[i / 64].bar.Drop();
  S->fibers[i].done = true;
  __sanitizer_start_switch_fiber(nullptr, S->main_bottom, S->main_size);
  setcontext(&S->main);
}

void Run(const Node &nd) {
  if (nd.block.x == 0 || nd.block.x > 1024 || nd.block.y != 1 || nd.block.z != 1 ||
      nd.grid.y != 1 || nd.grid.z != 1) {
    std::fprintf(stderr, "emu: unsupported launch shape\n");
    std::abort();
  }
  if (nd.lds > 160 * 1024) {
    std::fprintf(stderr, "emu: dynamic LDS %zu bytes exceeds 160 KiB\n", nd.lds);
    std::abort();
  }
  static std::vector<char *> stacks;  // reused by every launch
  const unsigned nt = nd.block.x;
  while (stacks.size() < nt) stacks.push_back(static_cast<char *>(std::malloc(kStack)));
  Sched S;
  S.nd = &nd;
  S.fibers.resize(nt);
  sched = &S;
  for (unsigned b = 0; b < nd.grid.x; ++b) {
    Block blk;
    if (nd.lds) {
      blk.lds = std::malloc(nd.lds);
      std::memset(blk.lds, 0xff, nd.lds);  // NaN pattern: unwritten LDS shows up
    }
    blk.bar.Reset(int(nt));
    blk.waves = std::vector<Wave>((nt + 63) / 64);
    for (size_t w = 0; w < blk.waves.size(); ++w)
      blk.waves[w].bar.Reset(int(std::min<unsigned>(64, nt - 64 * unsigned(w))));
    S.b = b;
    S.blk = &blk;
    for (unsigned t = 0; t < nt; ++t) {
      Fiber &f = S.fibers[t];
      f.done = false;
      f.stack = stacks[t];
      getcontext(&f.ctx);
      f.ctx.uc_stack.ss_sp = f.stack;
      f.ctx.uc_stack.ss_size = kStack;
      f.ctx.uc_link = nullptr;
      makecontext(&f.ctx, Trampoline, 0);
    }
    for (unsigned live = nt; live > 0;) {
      live = 0;
      for (unsigned t = 0; t < nt; ++t) {
        Fiber &f = S.fibers[t];
        if (f.done) continue;
        S.cur = t;
        Enter(t);
        void *fake = nullptr;
        __sanitizer_start_switch_fiber(&fake, f.stack, kStack);
        swapcontext(&S.main, &f.ctx);
        __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
        if (!f.done) ++live;
      }
    }
    std::free(blk.lds);
  }
  sched = nullptr;
}

}  // namespace

void *DynShared() { return sched->blk->lds; }
void SyncThreads() { sched->blk->bar.Wait(); }

double ShflD(double v, int src) {
  Wave &w = sched->blk->waves[tid.x / 64];
  w.d[tid.x & 63] = v;
  w.bar.Wait();
  const double r = w.d[src];
  w.bar.Wait();
  return r;
}

int ShflI(int v, int src) {
  Wave &w = sched->blk->waves[tid.x / 64];
  w.i[tid.x & 63] = v;
  w.bar.Wait();
  const int r = w.i[src];
  w.bar.Wait();
  return r;
}

uint64_t Ballot(int pred) {
  Wave &w = sched->blk->waves[tid.x / 64];
  const unsigned lanes = std::min<unsigned>(64, bdim.x - (tid.x & ~63u));
  w.i[tid.x & 63] = pred ? 1 : 0;
  w.bar.Wait();
  uint64_t m = 0;
  for (unsigned l = 0; l < lanes; ++l)
    if (w.i[l]) m |= uint64_t(1) << l;
  w.bar.Wait();
  return m;
}

}  // namespace emu

struct EmuStream {
  int unused;
};
struct EmuEvent {
  std::chrono::steady_clock::time_point t;
};
struct EmuGraph {
  std::vector<emu::Node> nodes;
};

namespace {
thread_local EmuGraph *g_capture = nullptr;
thread_local hipStream_t g_capture_stream = nullptr;

// The device: one worker thread runs every stream's work in submission
// order while the host goes on, as a GPU runs queued launches after the
// launching call has returned.  Host code that reads what kernels wrote
// without waiting for them (pinned progress words) sees it late, and
// launches a host queues past a loop's stop run after the host has moved
// on.  Synchronous calls (hipMemcpy, hipMemset, every *Synchronize,
// hipFree) first wait until the queue is empty.
class Device {
 public:
  Device() : worker_([this] { Loop(); }) { worker_.detach(); }
  void Submit(std::function<void()> task) {
    std::lock_guard<std::mutex> lk(mu_);
    q_.push_back(std::move(task));
    cv_.notify_all();
  }
  void Wait() {
    if (std::this_thread::get_id() == id_) return;  // called from a kernel body
    std::unique_lock<std::mutex> lk(mu_);
    idle_cv_.wait(lk, [this] { return q_.empty() && !busy_; });
  }

 private:
  void Loop() {
    id_ = std::this_thread::get_id();
    for (;;) {
      std::function<void()> task;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return !q_.empty(); });
        task = std::move(q_.front());
        q_.pop_front();
        busy_ = true;
      }
      task();
      {
        std::lock_guard<std::mutex> lk(mu_);
        busy_ = false;
        if (q_.empty()) idle_cv_.notify_all();
      }
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, idle_cv_;
  std::deque<std::function<void()>> q_;
  bool busy_ = false;
  std::thread::id id_;
  std::thread worker_;
};

// Device (hipMalloc) and pinned (hipHostMalloc) allocations: an async copy
// between two of them is queued; one with a pageable host side waits for
// the queue and copies before returning, as HIP stages pageable memory.
std::mutex g_alloc_mu;
std::map<uintptr_t, size_t> g_allocs;

void Track(void *p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocs[reinterpret_cast<uintptr_t>(p)] = bytes;
}
void Untrack(void *p) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocs.erase(reinterpret_cast<uintptr_t>(p));
}
bool Tracked(const void *p) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  auto it = g_allocs.upper_bound(a);
  if (it == g_allocs.begin()) return false;
  --it;
  return a < it->first + it->second;
}

Device &Dev() {
  // Never destroyed: the worker thread may outlive static destructors.
  static Device *d = [] {
    Device *p = new Device;
    std::atexit([] { Dev().Wait(); });
    return p;
  }();
  return *d;
}
}  // namespace

void emu::Launch(void *stream, dim3 grid, dim3 block, size_t lds, std::function<void()> body) {
  Node nd{grid, block, lds, std::move(body)};
  if (g_capture && static_cast<void *>(g_capture_stream) == stream) {
    g_capture->nodes.push_back(std::move(nd));
    return;
  }
  Dev().Submit([nd = std::move(nd)] { Run(nd); });
}

hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipGetDeviceCount(int *count) {
  *count = 1;
  return hipSuccess;
}
hipError_t hipGetDevice(int *device) {
  *device = 0;
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
  *p = const_cast<void *>(symbol);
  return hipSuccess;
}
hipError_t hipMalloc(void **p, size_t bytes) {
  *p = nullptr;
  if (posix_memalign(p, 256, bytes ? bytes : 1) != 0) return hipErrorOutOfMemory;
  std::memset(*p, 0xa5, bytes);  // device memory starts as garbage
  Track(*p, bytes ? bytes : 1);
  return hipSuccess;
}
hipError_t hipFree(void *p) {
  Dev().Wait();
  if (p) Untrack(p);
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipHostGetDevicePointer(void **dev, void *host, unsigned) {
  *dev = host;
  return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
  Dev().Wait();
  if (bytes) std::memcpy(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k,
                          hipStream_t) {
  if (g_capture) return hipErrorNotSupported;
  if (Tracked(dst) && Tracked(src)) {
    if (bytes) Dev().Submit([=] { std::memcpy(dst, src, bytes); });
    return hipSuccess;
  }
  return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemcpy2D(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                       size_t height, hipMemcpyKind) {
  Dev().Wait();
  for (size_t r = 0; r < height; ++r)
    std::memcpy(static_cast<char *>(dst) + r * dpitch,
                static_cast<const char *>(src) + r * spitch, width);
  return hipSuccess;
}
hipError_t hipMemset(void *dst, int value, size_t bytes) {
  Dev().Wait();
  if (bytes) std::memset(dst, value, bytes);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void *dst, int value, size_t bytes, hipStream_t) {
  if (g_capture) return hipErrorNotSupported;
  if (bytes) Dev().Submit([=] { std::memset(dst, value, bytes); });
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() {
  Dev().Wait();
  return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) {
  *s = new EmuStream{0};
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) {
  Dev().Wait();
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) {
  Dev().Wait();
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (g_capture || !s) return hipErrorInvalidValue;
  g_capture = new EmuGraph;
  g_capture_stream = s;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!g_capture || s != g_capture_stream) return hipErrorInvalidValue;
  *g = g_capture;
  g_capture = nullptr;
  g_capture_stream = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *,
                               size_t) {
  *e = new EmuGraph(*g);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (const auto &nd : e->nodes) Dev().Submit([nd] { emu::Run(nd); });
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new EmuEvent;
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  Dev().Wait();
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  Dev().Submit([e] { e->t = std::chrono::steady_clock::now(); });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) {
  Dev().Wait();
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
