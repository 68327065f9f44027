// LazyMt19937 (mpsft_amd/csrc/mt.h) against std::mt19937: the same words
// for seeds 0..4, 2^32-1 and random ones, over streams that stop inside the
// first block, at its end and several blocks on, and after a reseed
// mid-stream.  Prints "mismatches <count> streams <count>".
#include <cstdio>
#include <random>

#include "../mpsft_amd/csrc/mt.h"

int main() {
  std::mt19937_64 pick(7);
  long bad = 0, streams = 0;
  for (int t = 0; t < 4000; ++t) {
    const uint32_t s = t < 5 ? uint32_t(t) : (t == 5 ? 0xffffffffu : uint32_t(pick()));
    const int len = t < 40 ? 4000 : (t < 80 ? 624 + (t - 40) % 3 - 1 : int(pick() % 1400));
    std::mt19937 a(s);
    mpsft::LazyMt19937 b(s);
    ++streams;
    for (int i = 0; i < len; ++i)
      if (a() != b()) { ++bad; break; }
    const uint32_t s2 = uint32_t(pick());
    a.seed(s2);
    b.seed(s2);
    ++streams;
    for (int i = 0; i < 1300; ++i)
      if (a() != b()) { ++bad; break; }
  }
  std::printf("mismatches %ld streams %ld\n", bad, streams);
  return bad != 0;
}
