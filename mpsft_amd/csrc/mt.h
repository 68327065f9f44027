// mt.h -- std::mt19937's output stream with a lazy first block (host only).
//
// The transforms of a solve are drawn from mt19937 seeded with the solve's
// seed (src/base.cc:32-36, 52), so every draw must equal std::mt19937's.
// std::mt19937 fills its 624-word state and twists all of it before the
// first output; a solve draws a few dozen numbers.  This generator runs the
// same two loops -- the seeding recurrence x[i] = 1812433253 (x[i-1] ^
// (x[i-1] >> 30)) + i, then the twist of word k from x[k], x[k+1] and
// x[k+397] (k < 227) or the already twisted x[k-227] -- in the same order,
// but only as far as the outputs asked for need: output k of the first
// block needs the seeding up to word k + 397 (k < 227) or k + 1, and the
// twist up to word k.  Later blocks twist whole, as std::mt19937 does.
#pragma once

#include <cstdint>

namespace mpsft {

class LazyMt19937 {
 public:
  using result_type = uint32_t;
  explicit LazyMt19937(uint32_t s = 5489u) { seed(s); }

  void seed(uint32_t s) {
    x_[0] = s;
    seeded_ = 1;  // words [0, seeded_) hold the seeding recurrence
    twisted_ = 0; // words [0, twisted_) of the current block are twisted
    pos_ = 0;     // the next output's word
    first_ = true;
  }

  uint32_t operator()() {
    if (pos_ == kN) {  // a later block: twist it whole
      for (int k = 0; k < kN; ++k) Twist(k);
      pos_ = 0;
      first_ = false;
    }
    if (first_) {
      const int need = pos_ < kN - kM ? pos_ + kM : (pos_ + 1 < kN ? pos_ + 1 : kN - 1);
      while (seeded_ <= need) {
        const uint32_t p = x_[seeded_ - 1];
        x_[seeded_] = 1812433253u * (p ^ (p >> 30)) + uint32_t(seeded_);
        ++seeded_;
      }
      while (twisted_ <= pos_) Twist(twisted_++);
    }
    uint32_t z = x_[pos_++];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
  }

 private:
  static constexpr int kN = 624, kM = 397;

  // Word k of a block: the mixed upper bit of x[k] and lower bits of
  // x[k+1] (x[0], already twisted, for k = 623), shifted and conditionally
  // xored with the matrix, xored with x[k+397] (or the twisted x[k-227]).
  void Twist(int k) {
    const uint32_t next = x_[k + 1 < kN ? k + 1 : 0];
    const uint32_t y = (x_[k] & 0x80000000u) | (next & 0x7fffffffu);
    const uint32_t far = x_[k + kM < kN ? k + kM : k + kM - kN];
    x_[k] = far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }

  uint32_t x_[kN];
  int seeded_ = 0, twisted_ = 0, pos_ = 0;
  bool first_ = true;
};

}  // namespace mpsft
