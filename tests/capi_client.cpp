// capi_client.cpp -- a C++11 caller of libmpsft_hip.so through include/mpsft.h
// alone, compiled with g++ (no HIP headers), as the reference's own C++ code
// would call it.  It runs one window + binner plan on a signal and then a
// full solve:
//   mpsft_plan_create          <- Window(n, bins, delta) + BinInTime::Create
//   mpsft_bin_in_time          <- BinInTime::Run       (src/binner.h:33-34)
//   mpsft_bin_in_freq          <- BinInFreq::Run       (src/binner.h:126-127)
//   mpsft_identify_estimate    <- Iterate's per-bin body (src/iterate.cc:113-175)
//   mpsft_solver_create/solve  <- Demo1::Run           (src/demo1.cc:36-59)
// The test (tests/test_gpu_capi_client.py) writes the input file, runs this
// program and checks what it writes against the oracle.
//
// Input (little-endian): int64 n, bins, bits, q, a, b, c, m, k, seed;
// double delta, bin_threshold, sv_threshold, window_threshold, sigma;
// n complex128 samples; m int64 keys; m complex128 coefficients.
// Output: rows x bins complex128 after BinInTime, the same after BinInFreq,
// per bin int32 decide, int64 k1, int64 k0, complex128 coef; then int32
// iterations, int64 found count, that many int64 keys and complex128
// coefficients, and iterations int32 bins of the solve's log.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mpsft.h"

namespace {

struct Reader {
  FILE *f;
  template <typename T>
  void Get(T *v, size_t count = 1) {
    if (fread(v, sizeof(T), count, f) != count) {
      fprintf(stderr, "capi_client: short input\n");
      exit(2);
    }
  }
};

template <typename T>
void Put(FILE *f, const T *v, size_t count = 1) {
  if (fwrite(v, sizeof(T), count, f) != count) {
    fprintf(stderr, "capi_client: short write\n");
    exit(2);
  }
}

void Check(int status, const char *what) {
  if (status != MPSFT_OK) {
    fprintf(stderr, "capi_client: %s failed: %s\n", what, mpsft_status_string(status));
    exit(1);
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: capi_client <input> <output>\n");
    return 2;
  }
  FILE *in = fopen(argv[1], "rb");
  if (!in) return 2;
  Reader r{in};
  int64_t n, bins, bits, q, a, b, c, m, k, seed;
  double delta, bin_thr, sv_thr, win_thr, sigma;
  r.Get(&n), r.Get(&bins), r.Get(&bits), r.Get(&q), r.Get(&a), r.Get(&b), r.Get(&c);
  r.Get(&m), r.Get(&k), r.Get(&seed);
  r.Get(&delta), r.Get(&bin_thr), r.Get(&sv_thr), r.Get(&win_thr), r.Get(&sigma);
  std::vector<double> x(2 * n), coefs(2 * m);
  std::vector<int64_t> keys(m);
  r.Get(x.data(), x.size());
  r.Get(keys.data(), keys.size());
  r.Get(coefs.data(), coefs.size());
  fclose(in);

  const int64_t rows = 1 + 2 * bits;
  mpsft_transform tf;
  Check(mpsft_transform_make(n, a, b, c, &tf), "mpsft_transform_make");
  mpsft_plan *plan = nullptr;
  Check(mpsft_plan_create(n, int32_t(bins), delta, int32_t(bits), MPSFT_FP64, &plan),
        "mpsft_plan_create");
  void *x_dev = nullptr, *a_dev = nullptr, *k_dev = nullptr, *c_dev = nullptr;
  Check(mpsft_malloc(&x_dev, sizeof(double) * x.size()), "mpsft_malloc");
  Check(mpsft_malloc(&a_dev, sizeof(double) * 2 * rows * bins), "mpsft_malloc");
  Check(mpsft_malloc(&k_dev, sizeof(int64_t) * (m + 1)), "mpsft_malloc");
  Check(mpsft_malloc(&c_dev, sizeof(double) * 2 * (m + 1)), "mpsft_malloc");
  Check(mpsft_memcpy_htod(x_dev, x.data(), sizeof(double) * x.size()), "upload x");
  Check(mpsft_memcpy_htod(k_dev, keys.data(), sizeof(int64_t) * m), "upload keys");
  Check(mpsft_memcpy_htod(c_dev, coefs.data(), sizeof(double) * 2 * m), "upload coefs");

  FILE *out = fopen(argv[2], "wb");
  if (!out) return 2;
  std::vector<double> A(2 * rows * bins);
  Check(mpsft_bin_in_time(plan, x_dev, n, &tf, q, a_dev, nullptr), "mpsft_bin_in_time");
  Check(mpsft_memcpy_dtoh(A.data(), a_dev, sizeof(double) * A.size()), "read bins");
  Put(out, A.data(), A.size());
  Check(mpsft_bin_in_freq(plan, static_cast<const int64_t *>(k_dev), c_dev, m, &tf, q, a_dev,
                          nullptr),
        "mpsft_bin_in_freq");
  Check(mpsft_memcpy_dtoh(A.data(), a_dev, sizeof(double) * A.size()), "read bins");
  Put(out, A.data(), A.size());

  void *dec_dev = nullptr, *k1_dev = nullptr, *k0_dev = nullptr, *coef_dev = nullptr;
  Check(mpsft_malloc(&dec_dev, sizeof(int32_t) * bins), "mpsft_malloc");
  Check(mpsft_malloc(&k1_dev, sizeof(int64_t) * bins), "mpsft_malloc");
  Check(mpsft_malloc(&k0_dev, sizeof(int64_t) * bins), "mpsft_malloc");
  Check(mpsft_malloc(&coef_dev, sizeof(double) * 2 * bins), "mpsft_malloc");
  Check(mpsft_identify_estimate(plan, a_dev, 1, &tf, &q, bin_thr, sv_thr, win_thr,
                                static_cast<int32_t *>(dec_dev), static_cast<int64_t *>(k1_dev),
                                static_cast<int64_t *>(k0_dev), coef_dev, nullptr, nullptr),
        "mpsft_identify_estimate");
  std::vector<int32_t> decide(bins);
  std::vector<int64_t> k1(bins), k0(bins);
  std::vector<double> coef(2 * bins);
  Check(mpsft_memcpy_dtoh(decide.data(), dec_dev, sizeof(int32_t) * bins), "read decide");
  Check(mpsft_memcpy_dtoh(k1.data(), k1_dev, sizeof(int64_t) * bins), "read k1");
  Check(mpsft_memcpy_dtoh(k0.data(), k0_dev, sizeof(int64_t) * bins), "read k0");
  Check(mpsft_memcpy_dtoh(coef.data(), coef_dev, sizeof(double) * 2 * bins), "read coef");
  Put(out, decide.data(), decide.size());
  Put(out, k1.data(), k1.size());
  Put(out, k0.data(), k0.size());
  Put(out, coef.data(), coef.size());

  // Demo1::Run with Demo1Options defaults and the given mode count and noise.
  mpsft_solve_options opt;
  opt.trials = 1;
  opt.min_bins = 101;
  opt.window_delta = delta;
  opt.window_threshold = win_thr;
  opt.max_stale_iter = 3;
  opt.k = int32_t(k);
  opt.sigma = sigma;
  opt.max_iter = 0;
  mpsft_solver *solver = nullptr;
  Check(mpsft_solver_create(n, &opt, MPSFT_FP64, 0, &solver), "mpsft_solver_create");
  int32_t iters = 0;
  Check(mpsft_solve(solver, x_dev, uint32_t(seed), &iters, nullptr), "mpsft_solve");
  int64_t count = 0;
  Check(mpsft_solver_modes_count(solver, &count), "mpsft_solver_modes_count");
  std::vector<int64_t> fkeys(count + 1);
  std::vector<double> fcoefs(2 * (count + 1));
  Check(mpsft_solver_modes_get(solver, fkeys.data(), fcoefs.data(), count, &count),
        "mpsft_solver_modes_get");
  std::vector<int32_t> log_bins(iters + 1);
  int32_t logged = 0;
  Check(mpsft_solver_log_get(solver, log_bins.data(), nullptr, iters, &logged),
        "mpsft_solver_log_get");
  Put(out, &iters);
  Put(out, &count);
  Put(out, fkeys.data(), size_t(count));
  Put(out, fcoefs.data(), size_t(2 * count));
  Put(out, log_bins.data(), size_t(iters));
  fclose(out);

  mpsft_solver_destroy(solver);
  mpsft_plan_destroy(plan);
  for (void *p : {x_dev, a_dev, k_dev, c_dev, dec_dev, k1_dev, k0_dev, coef_dev}) mpsft_free(p);
  return 0;
}
