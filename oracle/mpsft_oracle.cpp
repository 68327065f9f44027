// mpsft_oracle.cpp -- CPU FP64 restatement of the MPSFT recovery loop.
//
// TEST INFRASTRUCTURE ONLY (the checker).  See mpsft_oracle.h for the list of
// reference functions restated here and what pins each of them.  Built with
// g++ -O2 -ffp-contract=off (no -ffast-math) so that every floating-point
// expression rounds as written.
#include "mpsft_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <unordered_map>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Cplex = std::complex<double>;
using ModeMap = std::unordered_map<int32_t, Cplex>;

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt2 = 1.41421356237309504880;

// ---------------------------------------------------------------------------
// Global RNG.  Same engine and distributions as src/base.cc:32-36 (libstdc++
// std::mt19937, uniform_int_distribution<int32_t/int64_t> over their default
// ranges, normal_distribution<double>, uniform_real_distribution<double>).
// Seeding also resets the distributions' cached state, so a seed alone fixes
// the stream (src/base.cc:52 reseeds only the engine).
struct Rng {
  std::mt19937 eng;
  std::uniform_int_distribution<int32_t> uid32;
  std::uniform_int_distribution<int64_t> uid64;
  std::normal_distribution<double> nd;
  std::uniform_real_distribution<double> ud;
  Rng() { eng.seed(123534); }  // src/base.cc:43
};
Rng g_rng;

// uniform_int_distribution<int32_t> over its default range [0, 2^31-1] as
// libstdc++ computed it before GCC 11 (the toolchains of 2017): the engine's
// 32-bit outputs are scaled by uerange = 2^31, scaling = (2^32-1)/2^31 = 1,
// so outputs >= 2^31 are rejected and the rest returned as they are.  GCC 11
// switched this call to a multiply-shift (r >> 1), which would give a
// different stream from the one the reference's tests were written against.
int32_t RandomInt32() {
  uint32_t r;
  do {
    r = uint32_t(g_rng.eng());
  } while (r >= 0x80000000u);
  return int32_t(r);
}
int64_t RandomInt64() { return g_rng.uid64(g_rng.eng); }
double RandomNormal() { return g_rng.nd(g_rng.eng); }
double RandomDouble() { return g_rng.ud(g_rng.eng); }

// The reference writes Cplex(RandomNormal(), RandomNormal())
// (src/gen.cc:35,60,87); g++ evaluates those arguments right to left, so the
// imaginary part takes the first draw.  The order is made explicit here.
Cplex RandomCplexNormal() {
  const double im = RandomNormal();
  const double re = RandomNormal();
  return Cplex(re, im);
}

// ---------------------------------------------------------------------------
// Integer helpers, src/integer.h:34-66.  C++ '%' truncates toward zero.
inline int32_t MulMod(int64_t a, int64_t b, int64_t d) {
  return int32_t((a * b) % d);
}
inline int32_t MulPosMod(int64_t a, int64_t b, int64_t d) {
  const int32_t x = MulMod(a, b, d);
  return x >= 0 ? x : int32_t(x + d);
}
inline int32_t MulAddMod(int64_t a, int64_t b, int64_t c, int64_t d) {
  return int32_t((a * b + c) % d);
}
inline int32_t MulAddPosMod(int64_t a, int64_t b, int64_t c, int64_t d) {
  const int32_t x = MulAddMod(a, b, c, d);
  return x >= 0 ? x : int32_t(x + d);
}

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

// Inverse of a modulo n by the extended Euclidean algorithm.  For prime n it
// equals the Fermat inverse a^(n-2) of src/integer.cc:27-40 (the inverse is
// unique in [0, n)); it is also valid for odd a and n = 2^k.
int32_t InvMod(int64_t a, int64_t n) {
  int64_t r0 = n, r1 = ((a % n) + n) % n;
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
  if (r0 != 1) return 0;  // not invertible
  return int32_t(((s0 % n) + n) % n);
}

struct Transform {
  int32_t a, b, c, a_inv;
};

// Transform(n), src/integer.cc:44-49: draws a, b, c in that order.  The
// reference draws a = r % (n-1) + 1, which is a unit only for prime n.  For
// n = 2^k the multiplier must be odd: a = 2*(r % (n/2)) + 1 (same number of
// draws).  For other n the draw of a is repeated until gcd(a, n) = 1; for
// prime n that never repeats, so the reference stream is unchanged.
bool DrawTransform(int32_t n, Transform *tf) {
  if (n < 2) return false;
  if (IsPow2(n)) {
    tf->a = 2 * (RandomInt32() % (n / 2)) + 1;
  } else {
    tf->a = (RandomInt32() % (n - 1)) + 1;
    while (Gcd(tf->a, n) != 1) tf->a = (RandomInt32() % (n - 1)) + 1;
  }
  tf->b = RandomInt32() % n;
  tf->c = RandomInt32() % n;
  tf->a_inv = InvMod(tf->a, n);
  return true;
}

// src/iterate.cc:30-42.
int32_t NumBits(int32_t n, int32_t bins) {
  if (n < bins) return 1;
  int32_t out = 0;
  while (n >= bins) {
    n >>= 1;
    ++out;
  }
  return out;
}

// src/binner.cc:42-50.
int32_t GetTau(int32_t q, int32_t bins, int32_t idx) {
  if (idx == 0) return q;
  if (idx & 1) return q + bins * (1 << ((idx - 1) / 2));
  return q - bins * (1 << (idx / 2 - 1));
}

// ---------------------------------------------------------------------------
// src/base.h:72, src/sincos.h:42-45, src/sincos.cc:69-88.
inline double PosModOne(double x) { return x - std::floor(x); }

inline Cplex Sinusoid(double freq) {
  freq *= (2.0 * kPi);
  return Cplex(std::cos(freq), std::sin(freq));
}

double SincPi(double x) {
  const double taylor_0_bound = std::numeric_limits<double>::epsilon();
  const double taylor_2_bound = std::sqrt(taylor_0_bound);
  const double taylor_n_bound = std::sqrt(taylor_2_bound);
  if (x < 0) x = -x;
  if (x >= taylor_n_bound) return std::sin(x) / x;
  double r = 1.0;
  if (x >= taylor_0_bound) {
    const double x2 = x * x;
    r -= x2 / 6.0;
    if (x >= taylor_2_bound) r += (x2 * x2) / 120.0;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Window, src/window.cc:28-81.  Gaussian-smoothed boxcar of width 1/(2B) in
// frequency; P = f*B taps in time with f the smallest odd integer covering
// the Gaussian's support at level delta.
struct Window {
  int32_t n = 0, bins = 0, p = 0;
  double delta = 0, width = 0, sigma_f = 0, sigma_t = 0;
  std::vector<double> wt;  // size p, index i <-> t = i (i<=p2) or i-p
  std::vector<int32_t> t;

  bool Init(int32_t n_, int32_t bins_, double delta_) {
    if (n_ <= 0 || bins_ <= 0 || !(delta_ > 0) || bins_ % 2 != 1) return false;
    n = n_;
    bins = bins_;
    delta = delta_;
    width = 1.0 / (2.0 * bins);
    const double sqrt_c_delta = std::sqrt(-std::log(delta));
    sigma_f = 0.5 / (bins * 2.0 * kSqrt2 * sqrt_c_delta);
    sigma_t = 1.0 / ((2.0 * kPi) * sigma_f);
    const double tmp = 2.0 * kSqrt2 * sigma_t * sqrt_c_delta + 1;
    int32_t factor = int32_t(std::ceil(tmp / double(bins)));
    if (factor % 2 == 0) ++factor;
    p = factor * bins;
    const int32_t p2 = (p - 1) / 2;
    t.assign(p, 0);
    for (int32_t i = 1; i <= p2; ++i) t[i] = i;
    for (int32_t i = p2 + 1; i < p; ++i) t[i] = i - p;
    wt.assign(p, 0.0);
    wt[0] = width;
    for (int32_t i = 1; i < p; ++i) wt[i] = SampleInTime(t[i]);
    return true;
  }
  int32_t p2() const { return (p - 1) / 2; }
  double SampleInTime(int32_t i) const {
    const double tt = double(i);
    const double u = tt * kPi * sigma_f;
    return width * SincPi(tt * kPi * width) * std::exp(-2.0 * u * u);
  }
  double SampleInFreq(double xi) const {
    const double c = 0.5 * width;
    const double d = kSqrt2 * kPi * sigma_t;
    return 0.5 * (std::erf((xi + c) * d) - std::erf((xi - c) * d));
  }
};

// ---------------------------------------------------------------------------
// FFT standing in for FFTW (src/base.cc:86-98): unnormalised, out[k] =
// sum_j in[j] exp(sign*2*pi*i*j*k/n).  Mixed radix (Cooley-Tukey, decimation
// in time) when every prime factor of n is at most 64, Bluestein over a
// power-of-two radix-2 FFT otherwise.  Twiddles use exact integer reduction
// of the angle.

Cplex Twiddle(int64_t num, int64_t den, int sign) {
  // exp(sign*2*pi*i*num/den) with num reduced modulo den first.
  num %= den;
  if (num < 0) num += den;
  const double ang = 2.0 * kPi * (double(num) / double(den));
  return Cplex(std::cos(ang), sign * std::sin(ang));
}

void Factorize(int32_t n, std::vector<int32_t> *f) {
  f->clear();
  int32_t m = n;
  // radix 4 first for speed on powers of two
  while (m % 4 == 0) {
    f->push_back(4);
    m /= 4;
  }
  for (int32_t p = 2; int64_t(p) * p <= m; ++p) {
    while (m % p == 0) {
      f->push_back(p);
      m /= p;
    }
  }
  if (m > 1) f->push_back(m);
}

void FftRec(const Cplex *in, int64_t stride, Cplex *out, int32_t n, int sign,
            const int32_t *fac, Cplex *scratch) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int32_t p = fac[0];
  const int32_t m = n / p;
  for (int32_t r = 0; r < p; ++r)
    FftRec(in + r * stride, stride * p, out + int64_t(r) * m, m, sign, fac + 1,
           scratch);
  // Combine: X[k + s*m] = sum_r W_n^{r(k+s*m)} Y_r[k].
  std::vector<Cplex> wp(p);
  for (int32_t s = 0; s < p; ++s) wp[s] = Twiddle(s, p, sign);
  std::vector<Cplex> y(p);
  for (int32_t k = 0; k < m; ++k) {
    for (int32_t r = 0; r < p; ++r)
      y[r] = out[int64_t(r) * m + k] * Twiddle(int64_t(r) * k, n, sign);
    for (int32_t s = 0; s < p; ++s) {
      Cplex acc = 0;
      for (int32_t r = 0; r < p; ++r) acc += y[r] * wp[(int64_t(r) * s) % p];
      scratch[k + int64_t(s) * m] = acc;
    }
  }
  std::copy(scratch, scratch + n, out);
}

void FftPow2(Cplex *a, int32_t n, int sign) {
  // iterative radix-2, in place, bit-reversed input order
  for (int32_t i = 1, j = 0; i < n; ++i) {
    int32_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int32_t len = 2; len <= n; len <<= 1) {
    const int32_t h = len / 2;
    std::vector<Cplex> w(h);
    for (int32_t k = 0; k < h; ++k) w[k] = Twiddle(k, len, sign);
    for (int32_t i = 0; i < n; i += len)
      for (int32_t k = 0; k < h; ++k) {
        const Cplex u = a[i + k];
        const Cplex v = a[i + k + h] * w[k];
        a[i + k] = u + v;
        a[i + k + h] = u - v;
      }
  }
}

void FftAny(const Cplex *in, Cplex *out, int32_t n, int sign) {
  if (n <= 0) return;
  std::vector<int32_t> fac;
  Factorize(n, &fac);
  const int32_t maxp = fac.empty() ? 1 : *std::max_element(fac.begin(), fac.end());
  if (maxp <= 64) {
    std::vector<Cplex> scratch(n);
    std::vector<Cplex> tmp(in, in + n);
    fac.push_back(1);
    FftRec(tmp.data(), 1, out, n, sign, fac.data(), scratch.data());
    return;
  }
  // Bluestein: X_k = c_k * sum_j (x_j c_j) conj(c_{k-j}), c_j =
  // exp(sign*pi*i*j^2/n), with j^2 reduced modulo 2n exactly.
  int32_t m = 1;
  while (m < 2 * n - 1) m <<= 1;
  std::vector<Cplex> chirp(n);
  for (int32_t j = 0; j < n; ++j) {
    const int64_t jj = (int64_t(j) * j) % (2 * int64_t(n));
    chirp[j] = Twiddle(jj, 2 * int64_t(n), sign);
  }
  std::vector<Cplex> a(m, 0.0), b(m, 0.0);
  for (int32_t j = 0; j < n; ++j) a[j] = in[j] * chirp[j];
  b[0] = std::conj(chirp[0]);
  for (int32_t j = 1; j < n; ++j) b[j] = b[m - j] = std::conj(chirp[j]);
  FftPow2(a.data(), m, -1);
  FftPow2(b.data(), m, -1);
  for (int32_t i = 0; i < m; ++i) a[i] *= b[i];
  FftPow2(a.data(), m, +1);
  const double inv = 1.0 / m;
  for (int32_t k = 0; k < n; ++k) out[k] = a[k] * inv * chirp[k];
}

// ---------------------------------------------------------------------------
// BinInTimeV0, src/binner.cc:81-111: the definitional binning.  For each row
// tau, fold y_{t+tau} exp(-2 pi i t/(2B)) wt(t) into t mod B and take the
// forward B-point DFT.
void BinInTimeV0(const Cplex *x, const Window &win, int32_t bits,
                 const Transform &tf, int32_t q, Cplex *out) {
  const int32_t bins = win.bins, n = win.n, p = win.p, p2 = win.p2();
  const double delta = -0.5 / double(bins);
  const int32_t rows = 1 + 2 * bits;
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t u = 0; u < rows; ++u) {
    std::vector<Cplex> scratch(bins, 0.0);
    const int32_t tau = GetTau(q, bins, u);
    for (int32_t i = 0; i < p; ++i) {
      const int32_t t = i <= p2 ? i : i - p;
      const double wt = win.wt[i];
      const int32_t j = MulAddPosMod(tf.a, int64_t(t) + int64_t(tau), tf.c, n);
      const int32_t k = MulMod(tf.b, int64_t(t) + int64_t(tau), n);
      const double freq = double(k) / double(n) + PosModOne(delta * double(t));
      scratch[i % bins] += (x[j] * Sinusoid(freq)) * wt;
    }
    FftAny(scratch.data(), out + int64_t(u) * bins, bins, -1);
  }
}

// BinInTimeV1, src/binner.cc:113-191: rows tau = q +- B*2^bit share one
// sinusoid per tap; z1 = y_{q+o+t}, z2 = conj(y_{q-o-t}) are split into
// s1 = (z1+z2)/2 and s2 = -i(z1-z2)/2, folded and transformed, and recombined
// as R + iI and conj(R) + i conj(I).  V2-V4 are vectorised reorderings of
// exactly this arithmetic.
void BinInTimeV1(const Cplex *x, const Window &win, int32_t bits,
                 const Transform &tf, int32_t q, Cplex *out) {
  const int32_t bins = win.bins, n = win.n, p = win.p;
  const double delta = -0.5 / double(bins);
  {
    std::vector<Cplex> s(bins, 0.0);
    for (int32_t i = 0; i < p; ++i) {
      const int32_t t = win.t[i];
      const double wt = win.wt[i];
      const int32_t j = MulAddPosMod(tf.a, int64_t(t) + int64_t(q), tf.c, n);
      const int32_t k = MulMod(tf.b, int64_t(t) + int64_t(q), n);
      const double freq = double(k) / double(n) + std::fmod(delta * double(t), 1.0);
      s[i % bins] += (x[j] * Sinusoid(freq)) * wt;
    }
    FftAny(s.data(), out, bins, -1);
  }
  const Cplex bq = Sinusoid(double(MulMod(tf.b, q, n)) / double(n));
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t bit = 0; bit < bits; ++bit) {
    std::vector<Cplex> s1(bins, 0.0), s2(bins, 0.0), r1(bins), r2(bins);
    const int32_t offset = bins * (1 << bit);
    for (int32_t i = 0; i < p; ++i) {
      const int32_t t = win.t[i];
      const double wt = win.wt[i];
      const int32_t ts = t + offset;
      const int32_t j1 = MulAddPosMod(tf.a, int64_t(q) + int64_t(ts), tf.c, n);
      const int32_t j2 = MulAddPosMod(tf.a, int64_t(q) - int64_t(ts), tf.c, n);
      const int32_t k = MulMod(tf.b, ts, n);
      const double freq = double(k) / double(n) + std::fmod(delta * double(t), 1.0);
      const Cplex factor = Sinusoid(freq);
      const Cplex z1 = x[j1] * bq * factor * wt;
      const Cplex z2 = std::conj(x[j2]) * std::conj(bq) * factor * wt;
      s1[i % bins] += 0.5 * (z1 + z2);
      const Cplex d = 0.5 * (z1 - z2);
      s2[i % bins] += Cplex(d.imag(), -d.real());  // times -i
    }
    FftAny(s1.data(), r1.data(), bins, -1);
    FftAny(s2.data(), r2.data(), bins, -1);
    Cplex *o1 = out + int64_t(2 * bit + 1) * bins;
    Cplex *o2 = out + int64_t(2 * bit + 2) * bins;
    for (int32_t b = 0; b < bins; ++b) {
      const Cplex cr = r1[b], ci = r2[b];
      o1[b] = cr + Cplex(-ci.imag(), ci.real());  // R + iI
      const Cplex cci = std::conj(ci);
      o2[b] = std::conj(cr) + Cplex(-cci.imag(), cci.real());
    }
  }
}

// BinInFreqV0, src/binner.cc:540-560 and V1, src/binner.cc:565-591: subtract
// the window response of every known mode from its own bin only.
void BinInFreq(int version, const ModeMap &mm, const Window &win,
               int32_t bits, const Transform &tf, int32_t q, Cplex *out) {
  const int32_t bins = win.bins, n = win.n;
  const int32_t rows = 1 + 2 * bits;
  for (const auto &kv : mm) {
    const int32_t k0 = kv.first;
    const int32_t k1 = MulAddPosMod(tf.a, k0, tf.b, n);
    const int32_t bin = int32_t(int64_t(k1) * int64_t(bins) / int64_t(n));
    const double xi = (double(bin) + 0.5) / double(bins) - double(k1) / double(n);
    const double wf = win.SampleInFreq(xi);
    if (version == 0) {
      const Cplex base = kv.second * wf;
      for (int32_t u = 0; u < rows; ++u) {
        const int32_t tau = GetTau(q, bins, u);
        const int32_t s = MulAddMod(tf.c, k0, int64_t(k1) * int64_t(tau), n);
        out[int64_t(u) * bins + bin] -= base * Sinusoid(double(s) / double(n));
      }
    } else {
      const double freq =
          double(MulAddMod(q, k1, int64_t(tf.c) * int64_t(k0), n)) / double(n);
      const Cplex base = wf * kv.second * Sinusoid(freq);
      out[bin] -= base;
      for (int32_t bit = 0; bit < bits; ++bit) {
        const int32_t offset = bins * (1 << bit);
        const Cplex factor = Sinusoid(double(MulMod(offset, k1, n)) / double(n));
        out[int64_t(bit * 2 + 1) * bins + bin] -= base * factor;
        out[int64_t(bit * 2 + 2) * bins + bin] -= base * std::conj(factor);
      }
    }
  }
}

// MatPencil, src/freqid.cc:27-48, in closed form.  For M = [[u0,u2],[u1,u0]]
// the Gram matrix M^H M = [[alpha, beta],[conj(beta), gamma]] has alpha =
// |u0|^2+|u1|^2, gamma = |u2|^2+|u0|^2, beta = conj(u0) u2 + conj(u1) u0.
// sigma1^2 is its larger eigenvalue; sigma2 = |det M| / sigma1.  The dominant
// right singular vector is proportional to (beta, lambda1 - alpha), so
// Im(conj(v0) v1) < 0 exactly when Im(beta) > 0.
bool MatPencil(Cplex u0, Cplex u1, Cplex u2, double *sigma) {
  const double alpha = std::norm(u0) + std::norm(u1);
  const double gamma = std::norm(u2) + std::norm(u0);
  const Cplex beta = std::conj(u0) * u2 + std::conj(u1) * u0;
  const double h = 0.5 * (alpha - gamma);
  const double lam1 = 0.5 * (alpha + gamma) + std::sqrt(h * h + std::norm(beta));
  const double s1 = std::sqrt(lam1);
  const double det = std::abs(u0 * u0 - u1 * u2);
  sigma[0] = s1;
  sigma[1] = s1 > 0 ? det / s1 : 0.0;
  return beta.imag() > 0;
}

struct IterOpts {
  int32_t bins;
  double window_delta;
  int32_t trials;
  double bin_threshold, window_threshold, sv_threshold;
  int32_t bin_in_time_type, bin_in_freq_type;
};

// IdentifyFreq, src/iterate.cc:44-76.  Majority vote per bit over trials,
// most significant bit first (bit 0, offset B, decides the coarsest half).
std::pair<double, bool> IdentifyFreq(const IterOpts &opt, int32_t b,
                                     int32_t bits, int32_t bins,
                                     const std::vector<std::vector<Cplex>> &A,
                                     double *margin) {
  const int32_t trials = opt.trials;
  double sigma[2];
  int32_t xi_sum = 0;
  for (int32_t bit = 0; bit < bits; ++bit) {
    int32_t count = 0;
    for (int32_t trial = 0; trial < trials; ++trial) {
      const std::vector<Cplex> &a = A[trial];
      const Cplex u1 = a[int64_t(2 * bit + 1) * bins + b];
      const Cplex u2 = a[int64_t(2 * bit + 2) * bins + b];
      const Cplex u0 = a[b];
      if (MatPencil(u0, u1, u2, sigma)) ++count;
      if (margin) {
        const Cplex beta = std::conj(u0) * u2 + std::conj(u1) * u0;
        const double ab = std::abs(beta);
        if (ab > 0) *margin = std::min(*margin, std::fabs(beta.imag()) / ab);
        if (opt.sv_threshold > 0)
          *margin = std::min(*margin, std::fabs(sigma[1] * 0.5 - opt.sv_threshold) /
                                          opt.sv_threshold);
      }
      if (opt.sv_threshold > 0 && sigma[1] * 0.5 > opt.sv_threshold)
        return std::make_pair(0.0, false);
    }
    xi_sum <<= 1;
    if (count > trials / 2) ++xi_sum;
  }
  return std::make_pair(double(2 * xi_sum + 1) / double(1 << (bits + 1)), true);
}

int IdentifyBins(int32_t n, const IterOpts &opt, const Window &win, int32_t bits,
                 const Transform &tf, const std::vector<int64_t> &list_q,
                 const std::vector<std::vector<Cplex>> &A, ModeMap *mm,
                 orc_iterate_trace *trace, bool *found_out);

// Iterate, src/iterate.cc:80-177.
int IterateImpl(const Cplex *x, int32_t n, const IterOpts &opt,
                const Transform *tf_in, const int32_t *qs_in, ModeMap *mm,
                orc_iterate_trace *trace, bool *found_out) {
  const int32_t trials = opt.trials;
  const int32_t bins = opt.bins;
  if (!(opt.window_threshold > 0)) return ORC_ERR_INVALID_ARG;
  if (trials <= 0 || trials % 2 != 1) return ORC_ERR_INVALID_ARG;
  Window win;
  if (!win.Init(n, bins, opt.window_delta)) return ORC_ERR_INVALID_ARG;
  if (opt.bin_in_time_type < 0 || opt.bin_in_time_type > 4) return ORC_ERR_UNSUPPORTED;
  if (opt.bin_in_freq_type < 0 || opt.bin_in_freq_type > 1) return ORC_ERR_UNSUPPORTED;
  const int32_t bits = NumBits(n, bins);
  const int32_t rows = 1 + 2 * bits;

  Transform tf;
  if (tf_in) {
    tf = *tf_in;
  } else if (!DrawTransform(n, &tf)) {
    return ORC_ERR_INVALID_ARG;
  }

  std::vector<int64_t> list_q;
  std::vector<std::vector<Cplex>> A(trials);
  for (int32_t trial = 0; trial < trials; ++trial) {
    const int32_t q = qs_in ? qs_in[trial] : RandomInt32() % n;
    list_q.push_back(q);
    A[trial].assign(int64_t(rows) * bins, 0.0);
    if (opt.bin_in_time_type == 0)
      BinInTimeV0(x, win, bits, tf, q, A[trial].data());
    else
      BinInTimeV1(x, win, bits, tf, q, A[trial].data());
    BinInFreq(opt.bin_in_freq_type, *mm, win, bits, tf, q, A[trial].data());
  }
  if (trace) {
    trace->tf[0] = tf.a;
    trace->tf[1] = tf.b;
    trace->tf[2] = tf.c;
    trace->tf[3] = tf.a_inv;
    trace->bits = bits;
    if (trace->q_out)
      for (int32_t t = 0; t < trials; ++t) trace->q_out[t] = int32_t(list_q[t]);
    if (trace->a_out)
      for (int32_t t = 0; t < trials; ++t)
        std::memcpy(trace->a_out + 2 * int64_t(t) * rows * bins, A[t].data(),
                    sizeof(Cplex) * size_t(rows) * bins);
  }

  const int rc = IdentifyBins(n, opt, win, bits, tf, list_q, A, mm, trace, found_out);
  return rc;
}

// Iterate's per-bin body (src/iterate.cc:113-175) on bin matrices A[trial]
// (rows x bins each): energy test, IdentifyFreq, k1 rounding, window test,
// coefficient estimate, un-permutation, mm[k0] += coef.
int IdentifyBins(int32_t n, const IterOpts &opt, const Window &win, int32_t bits,
                 const Transform &tf, const std::vector<int64_t> &list_q,
                 const std::vector<std::vector<Cplex>> &A, ModeMap *mm,
                 orc_iterate_trace *trace, bool *found_out) {
  bool found = false;
  const int32_t trials = opt.trials;
  const int32_t bins = opt.bins;
  int32_t nfound = 0;
  for (int32_t b = 0; b < bins; ++b) {
    double margin = 1e300;
    int32_t decide = 0;
    int32_t k1_rec = -1;
    if (opt.bin_threshold > 0) {
      double bin_energy = 0;
      for (int32_t trial = 0; trial < trials; ++trial) {
        const Cplex coef = A[trial][b];
        bin_energy += coef.real() * coef.real() + coef.imag() * coef.imag();
      }
      bin_energy /= trials;
      const double thr2 = opt.bin_threshold * opt.bin_threshold;
      margin = std::min(margin, std::fabs(bin_energy - thr2) / thr2);
      if (bin_energy < thr2) goto record;
    }
    {
      auto res = IdentifyFreq(opt, b, bits, bins, A, &margin);
      if (!res.second) {
        decide = 1;
        goto record;
      }
      const double xi1 = res.first;
      const double kr = double(n) * (double(b) + xi1) / double(bins);
      margin = std::min(margin, std::fabs(std::fabs(kr - std::floor(kr)) - 0.5));
      const int32_t k1 = int32_t(std::round(kr));
      k1_rec = k1;
      const double xi0 = double(k1) / double(n) - (double(b) + 0.5) / double(bins);
      const double wf = win.SampleInFreq(xi0);
      margin = std::min(margin, std::fabs(std::fabs(wf) - opt.window_threshold) /
                                    opt.window_threshold);
      if (std::abs(wf) < opt.window_threshold) {
        decide = 2;
        goto record;
      }
      Cplex coef_sum = 0;
      for (int32_t trial = 0; trial < trials; ++trial) {
        const std::vector<Cplex> &a = A[trial];
        const double freq = -double(MulMod(list_q[trial], k1, n)) / double(n);
        const Cplex factor = Sinusoid(freq);
        coef_sum += a[b] * factor;
        for (int32_t bit = 0; bit < bits; ++bit) {
          const double freq2 = -double(MulMod(bins * (1 << bit), k1, n)) / double(n);
          const Cplex factor2 = Sinusoid(freq2);
          const Cplex f1 = factor * factor2;
          const Cplex f2 = factor * std::conj(factor2);
          coef_sum += a[int64_t(2 * bit + 1) * bins + b] * f1;
          coef_sum += a[int64_t(2 * bit + 2) * bins + b] * f2;
        }
      }
      Cplex coef = coef_sum / double(trials * (1 + 2 * bits));
      const int32_t k0 = MulPosMod(tf.a_inv, int64_t(k1) - int64_t(tf.b), n);
      coef *= Sinusoid(-double(MulMod(tf.c, k0, n)) / double(n));
      coef /= wf;
      (*mm)[k0] += coef;
      found = true;
      ++nfound;
      decide = 3;
    }
  record:
    if (trace) {
      if (trace->decide_out) trace->decide_out[b] = decide;
      if (trace->k1_out) trace->k1_out[b] = k1_rec;
      if (trace->margin_out) trace->margin_out[b] = margin;
    }
  }
  if (trace) trace->found = nfound;
  *found_out = found;
  return ORC_OK;
}


IterOpts FromC(const orc_iterate_options *o) {
  IterOpts r;
  r.bins = o->bins;
  r.window_delta = o->window_delta;
  r.trials = o->trials;
  r.bin_threshold = o->bin_threshold;
  r.window_threshold = o->window_threshold;
  r.sv_threshold = o->sv_threshold;
  r.bin_in_time_type = o->bin_in_time_type;
  r.bin_in_freq_type = o->bin_in_freq_type;
  return r;
}

}  // namespace

struct orc_modemap {
  ModeMap m;
};

extern "C" {

void orc_set_threads(int32_t threads) {
#ifdef _OPENMP
  omp_set_num_threads(threads > 0 ? threads : omp_get_num_procs());
#else
  (void)threads;
#endif
}

int32_t orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_rng_seed(uint32_t seed) {
  g_rng.eng.seed(seed);
  g_rng.uid32.reset();
  g_rng.uid64.reset();
  g_rng.nd.reset();
  g_rng.ud.reset();
}
int32_t orc_rng_int32(void) { return RandomInt32(); }
int64_t orc_rng_int64(void) { return RandomInt64(); }
double orc_rng_normal(void) { return RandomNormal(); }
double orc_rng_double(void) { return RandomDouble(); }

int32_t orc_mul_mod(int64_t a, int64_t b, int64_t d) { return MulMod(a, b, d); }
int32_t orc_mul_pos_mod(int64_t a, int64_t b, int64_t d) { return MulPosMod(a, b, d); }
int32_t orc_mul_add_pos_mod(int64_t a, int64_t b, int64_t c, int64_t d) {
  return MulAddPosMod(a, b, c, d);
}

int orc_transform_draw(int32_t n, int32_t tf[4]) {
  Transform t;
  if (!DrawTransform(n, &t)) return ORC_ERR_INVALID_ARG;
  tf[0] = t.a;
  tf[1] = t.b;
  tf[2] = t.c;
  tf[3] = t.a_inv;
  return ORC_OK;
}

int orc_transform_make(int32_t n, int32_t a, int32_t b, int32_t c, int32_t tf[4]) {
  if (n < 2) return ORC_ERR_INVALID_ARG;
  tf[0] = a;
  tf[1] = b;
  tf[2] = c;
  tf[3] = InvMod(a, n);
  return ORC_OK;
}

int32_t orc_num_bits(int32_t n, int32_t bins) { return NumBits(n, bins); }

int orc_window_params(int32_t n, int32_t bins, double delta, int32_t *p,
                      double *width, double *sigma_f, double *sigma_t) {
  Window w;
  if (!w.Init(n, bins, delta)) return ORC_ERR_INVALID_ARG;
  if (p) *p = w.p;
  if (width) *width = w.width;
  if (sigma_f) *sigma_f = w.sigma_f;
  if (sigma_t) *sigma_t = w.sigma_t;
  return ORC_OK;
}

int orc_window_taps(int32_t n, int32_t bins, double delta, double *wt, int32_t *t) {
  Window w;
  if (!w.Init(n, bins, delta)) return ORC_ERR_INVALID_ARG;
  std::copy(w.wt.begin(), w.wt.end(), wt);
  std::copy(w.t.begin(), w.t.end(), t);
  return ORC_OK;
}

double orc_window_sample_in_time(int32_t n, int32_t bins, double delta, int32_t i) {
  Window w;
  if (!w.Init(n, bins, delta)) return NAN;
  return w.SampleInTime(i);
}

double orc_window_sample_in_freq(int32_t n, int32_t bins, double delta, double xi) {
  Window w;
  if (!w.Init(n, bins, delta)) return NAN;
  return w.SampleInFreq(xi);
}

int orc_fft(const double *in, double *out, int32_t n, int sign) {
  if (n <= 0 || (sign != 1 && sign != -1)) return ORC_ERR_INVALID_ARG;
  FftAny(reinterpret_cast<const Cplex *>(in), reinterpret_cast<Cplex *>(out), n, sign);
  return ORC_OK;
}

int orc_dft_naive(const double *in, double *out, int32_t n, int sign) {
  if (n <= 0 || (sign != 1 && sign != -1)) return ORC_ERR_INVALID_ARG;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int32_t k = 0; k < n; ++k) {
    long double re = 0, im = 0;
    for (int32_t j = 0; j < n; ++j) {
      const int64_t r = (int64_t(j) * k) % n;
      const long double ang = sign * 2.0L * pi * (long double)r / (long double)n;
      const long double c = cosl(ang), s = sinl(ang);
      re += in[2 * j] * c - in[2 * j + 1] * s;
      im += in[2 * j] * s + in[2 * j + 1] * c;
    }
    out[2 * k] = double(re);
    out[2 * k + 1] = double(im);
  }
  return ORC_OK;
}

int orc_bin_in_time(int version, const double *x, int32_t n, int32_t bins,
                    double delta, int32_t bits, const int32_t tf[4], int32_t q,
                    double *out) {
  Window w;
  if (!w.Init(n, bins, delta)) return ORC_ERR_INVALID_ARG;
  if (bits < 1) return ORC_ERR_SHAPE;
  Transform t{tf[0], tf[1], tf[2], tf[3]};
  const Cplex *xc = reinterpret_cast<const Cplex *>(x);
  Cplex *oc = reinterpret_cast<Cplex *>(out);
  if (version == 0)
    BinInTimeV0(xc, w, bits, t, q, oc);
  else if (version >= 1 && version <= 4)
    BinInTimeV1(xc, w, bits, t, q, oc);
  else
    return ORC_ERR_UNSUPPORTED;
  return ORC_OK;
}

int orc_bin_in_freq(int version, const int32_t *keys, const double *coefs,
                    int32_t m, int32_t n, int32_t bins, double delta,
                    int32_t bits, const int32_t tf[4], int32_t q, double *out) {
  Window w;
  if (!w.Init(n, bins, delta)) return ORC_ERR_INVALID_ARG;
  if (version != 0 && version != 1) return ORC_ERR_UNSUPPORTED;
  ModeMap mm;
  for (int32_t i = 0; i < m; ++i) mm[keys[i]] = Cplex(coefs[2 * i], coefs[2 * i + 1]);
  Transform t{tf[0], tf[1], tf[2], tf[3]};
  BinInFreq(version, mm, w, bits, t, q, reinterpret_cast<Cplex *>(out));
  return ORC_OK;
}

int orc_mat_pencil(const double *u, double *sigma) {
  return MatPencil(Cplex(u[0], u[1]), Cplex(u[2], u[3]), Cplex(u[4], u[5]), sigma) ? 1 : 0;
}

orc_modemap *orc_modemap_create(void) { return new orc_modemap; }
void orc_modemap_destroy(orc_modemap *mm) { delete mm; }
int32_t orc_modemap_size(const orc_modemap *mm) { return int32_t(mm->m.size()); }
void orc_modemap_get(const orc_modemap *mm, int32_t *keys, double *coefs) {
  int32_t i = 0;
  for (const auto &kv : mm->m) {
    keys[i] = kv.first;
    coefs[2 * i] = kv.second.real();
    coefs[2 * i + 1] = kv.second.imag();
    ++i;
  }
}
void orc_modemap_set(orc_modemap *mm, const int32_t *keys, const double *coefs,
                     int32_t m) {
  mm->m.clear();
  for (int32_t i = 0; i < m; ++i) mm->m[keys[i]] = Cplex(coefs[2 * i], coefs[2 * i + 1]);
}

int orc_iterate(const double *x, int32_t n, const orc_iterate_options *opt,
                orc_modemap *mm, orc_iterate_trace *trace) {
  bool found = false;
  const int rc = IterateImpl(reinterpret_cast<const Cplex *>(x), n, FromC(opt),
                             nullptr, nullptr, &mm->m, trace, &found);
  if (rc != ORC_OK) return -rc;
  return found ? 1 : 0;
}

int orc_iterate_fixed(const double *x, int32_t n, const orc_iterate_options *opt,
                      const int32_t tf[4], const int32_t *qs, orc_modemap *mm,
                      orc_iterate_trace *trace) {
  bool found = false;
  Transform t{tf[0], tf[1], tf[2], tf[3]};
  const int rc = IterateImpl(reinterpret_cast<const Cplex *>(x), n, FromC(opt),
                             &t, qs, &mm->m, trace, &found);
  if (rc != ORC_OK) return -rc;
  return found ? 1 : 0;
}

}  // extern "C"

namespace {

// The loop of Demo1::Run (src/demo1.cc:36-59) with its Iterate call given
// as step(io, &res, &size): step runs one Iterate with the options io and
// reports its return value and the map size after it.  Before iteration i:
// bins = max((k - size)*2 + 1, min_bins) (src/demo1.cc:48-49), computed in
// size_t there and converted to int32, which equals the signed value for
// any size < 2^30; thresholds 3 sigma/sqrt(bins) and sigma/sqrt(bins)
// (src/demo1.cc:50-51).  stale counts the iterations whose Iterate returned
// false and is never reset (src/demo1.cc:55-57); the loop runs while
// stale < max_stale_iter.  Both the real solve (orc_demo1_run*) and the
// scripted schedule (orc_demo1_schedule) run this one loop.
template <typename Step>
int32_t Demo1Loop(int32_t k, double sigma, const orc_demo1_options *opt, int32_t start_size,
                  orc_demo1_log *log, int32_t max_iter, Step step) {
  IterOpts io;
  io.window_delta = opt->window_delta;
  io.window_threshold = opt->window_threshold;
  io.trials = opt->trials;
  io.bin_in_time_type = 1;
  io.bin_in_freq_type = 1;
  int32_t stale_iter = 0;
  int32_t size = start_size;
  int32_t i = 0;
  for (; stale_iter < opt->max_stale_iter; ++i) {
    if (max_iter >= 0 && i >= max_iter) break;
    io.bins = std::max<int32_t>(int32_t((int64_t(k) - int64_t(size)) * 2 + 1), opt->min_bins);
    io.bin_threshold = 3.0 * sigma / std::sqrt(double(io.bins));
    io.sv_threshold = 1.0 * sigma / std::sqrt(double(io.bins));
    bool res = false;
    const int rc = step(io, &res, &size);
    if (rc != ORC_OK) return -rc;
    if (log && i < log->cap) {
      if (log->bins) log->bins[i] = io.bins;
      if (log->found) log->found[i] = size;
      if (log->res) log->res[i] = res ? 1 : 0;
      if (log->stale) log->stale[i] = stale_iter + (res ? 0 : 1);
      if (log->bin_thr) log->bin_thr[i] = io.bin_threshold;
      if (log->sv_thr) log->sv_thr[i] = io.sv_threshold;
    }
    if (!res) ++stale_iter;
  }
  return i;
}

}  // namespace

extern "C" {

int orc_identify(const double *a, int32_t trials, int32_t n, const orc_iterate_options *opt,
                 const int32_t tf[4], const int32_t *qs, orc_modemap *mm,
                 orc_iterate_trace *trace) {
  IterOpts io = FromC(opt);
  if (trials != io.trials || trials <= 0 || trials % 2 != 1 || !(io.window_threshold > 0))
    return -ORC_ERR_INVALID_ARG;
  Window win;
  if (!win.Init(n, io.bins, io.window_delta)) return -ORC_ERR_INVALID_ARG;
  const int32_t bits = NumBits(n, io.bins);
  const int64_t per = int64_t(1 + 2 * bits) * io.bins;
  std::vector<std::vector<Cplex>> A(trials);
  std::vector<int64_t> list_q(trials);
  for (int32_t t = 0; t < trials; ++t) {
    const Cplex *src = reinterpret_cast<const Cplex *>(a) + t * per;
    A[t].assign(src, src + per);
    list_q[t] = qs[t];
  }
  Transform t4{tf[0], tf[1], tf[2], tf[3]};
  if (trace) {
    for (int i = 0; i < 4; ++i) trace->tf[i] = tf[i];
    trace->bits = bits;
  }
  bool found = false;
  const int rc = IdentifyBins(n, io, win, bits, t4, list_q, A, &mm->m, trace, &found);
  if (rc != ORC_OK) return -rc;
  return found ? 1 : 0;
}

int32_t orc_demo1_run_log(const double *x, int32_t n, int32_t k, double sigma,
                          const orc_demo1_options *opt, orc_modemap *found,
                          orc_demo1_log *log) {
  found->m.clear();
  return Demo1Loop(k, sigma, opt, 0, log, -1,
                   [&](const IterOpts &io, bool *res, int32_t *size) {
                     const int rc = IterateImpl(reinterpret_cast<const Cplex *>(x), n, io,
                                                nullptr, nullptr, &found->m, nullptr, res);
                     *size = int32_t(found->m.size());
                     return rc;
                   });
}

int32_t orc_demo1_run(const double *x, int32_t n, int32_t k, double sigma,
                      const orc_demo1_options *opt, orc_modemap *found,
                      int32_t *bins_log, int32_t *found_log, int32_t log_cap) {
  orc_demo1_log log{};
  log.bins = bins_log;
  log.found = found_log;
  log.cap = log_cap;
  return orc_demo1_run_log(x, n, k, sigma, opt, found, &log);
}

int32_t orc_demo1_schedule(int32_t k, double sigma, const orc_demo1_options *opt,
                           const uint8_t *script_res, const int32_t *script_size,
                           int32_t script_len, orc_demo1_log *log) {
  int32_t i = 0;
  const int32_t it = Demo1Loop(k, sigma, opt, 0, log, script_len,
                               [&](const IterOpts &, bool *res, int32_t *size) {
                                 *res = script_res[i] != 0;
                                 *size = script_size[i];
                                 ++i;
                                 return ORC_OK;
                               });
  return it;
}

int orc_generate_mode_map(int32_t n, int32_t k, orc_modemap *mm) {
  if (n <= 0 || k < 0 || k > n) return ORC_ERR_INVALID_ARG;
  mm->m.clear();
  while (mm->m.size() < size_t(k)) {
    int32_t idx = RandomInt32() % n;
    if (idx < 0) idx += n;
    if (mm->m.find(idx) != mm->m.end()) continue;
    Cplex coef = RandomCplexNormal();
    coef /= std::abs(coef);
    mm->m[idx] = coef;
  }
  return ORC_OK;
}

int orc_evaluate_modes(int32_t n, const orc_modemap *mm, double *out) {
  Cplex *data = reinterpret_cast<Cplex *>(out);
  std::fill(data, data + n, Cplex(0.0));
  for (const auto &kv : mm->m) {
    for (int32_t t = 0; t < n; ++t) {
      const double angle = 2.0 * kPi * (double(MulMod(kv.first, t, n)) / double(n));
      data[t] += kv.second * Cplex(std::cos(angle), std::sin(angle));
    }
  }
  return ORC_OK;
}

int orc_generate_xhat(int32_t n, const orc_modemap *mm, double sigma, double *out) {
  Cplex *o = reinterpret_cast<Cplex *>(out);
  double noise_energy = 0;
  for (int32_t i = 0; i < n; ++i) {
    o[i] = RandomCplexNormal();
    noise_energy += o[i].real() * o[i].real() + o[i].imag() * o[i].imag();
  }
  for (const auto &kv : mm->m) {
    const Cplex v = o[kv.first];
    noise_energy -= v.real() * v.real() + v.imag() * v.imag();
  }
  const double factor = sigma / std::sqrt(noise_energy);
  for (int32_t i = 0; i < n; ++i) o[i] *= factor;
  for (const auto &kv : mm->m) o[kv.first] = kv.second;
  return ORC_OK;
}

int orc_generate_xhat_alt(int32_t n, const orc_modemap *mm, double sigma, double *out) {
  Cplex *o = reinterpret_cast<Cplex *>(out);
  std::fill(o, o + n, Cplex(0.0));
  for (const auto &kv : mm->m) o[kv.first] = kv.second;
  const double scale = sigma / std::sqrt(2 * n);
  for (int32_t i = 0; i < n; ++i) o[i] += RandomCplexNormal() * scale;
  return ORC_OK;
}

int orc_post_analyze(int32_t n, const orc_modemap *truth, const double *xh,
                     const orc_modemap *found, double *stats) {
  const Cplex *x = reinterpret_cast<const Cplex *>(xh);
  std::vector<Cplex> xh2(n, 0.0);
  int32_t hit = 0, miss = 0;
  double hit_l1 = 0, hit_l2 = 0, miss_l1 = 0, miss_l2 = 0;
  for (const auto &kv : found->m) {
    xh2[kv.first] = kv.second;
    const double err = std::abs(kv.second - x[kv.first]);
    if (truth->m.find(kv.first) == truth->m.end()) {
      ++miss;
      miss_l1 += err;
      miss_l2 += err * err;
    } else {
      ++hit;
      hit_l1 += err;
      hit_l2 += err * err;
    }
  }
  double s1 = 0, s2 = 0, mx = 0;
  for (int32_t i = 0; i < n; ++i) {
    const double err = std::abs(xh2[i] - x[i]);
    s1 += err;
    s2 += err * err;
    mx = std::max(mx, err);
  }
  stats[0] = hit;
  stats[1] = hit_l1 / hit;
  stats[2] = std::sqrt(hit_l2 / hit);
  stats[3] = miss;
  stats[4] = miss_l1 / miss;
  stats[5] = std::sqrt(miss_l2 / miss);
  stats[6] = s1 / n;
  stats[7] = std::sqrt(s2 / n);
  stats[8] = mx;
  return ORC_OK;
}

}  // extern "C"
