// rng.h -- the random numbers the model draws.
//
// The reference seeds std::mt19937 with 12345678 under -DDEBUG
// (src/model.cu:32-38) and draws through libstdc++'s
// uniform_real_distribution<double> and uniform_int_distribution<int>
// (src/model.cu:79, src/anderson.cu:25, src/charge.cu:52,68,
// src/model_lattice.cu:32).  The distributions are implementation-defined,
// so they are written out here, following GCC 11's libstdc++ algorithms,
// which makes the draws independent of the C++ library the host is built
// against:
//   uniform_real(a, b): u = (w0 + w1 * 2^32) / 2^64 with two 32-bit words
//     (generate_canonical<double, 53>), u clamped below 1, then u*(b-a) + a;
//   uniform_int(0, INT_MAX): Lemire's multiply-shift with range 2^31, whose
//     rejection threshold is 0, i.e. w >> 1.
// The engine is the standard MT19937 (Matsumoto & Nishimura 1998).
#pragma once

#include <cmath>
#include <cstdint>

namespace lsqt_host {

class MT19937 {
public:
  explicit MT19937(uint32_t seed = 5489u) { seed_with(seed); }

  void seed_with(uint32_t seed)
  {
    state_[0] = seed;
    for (int i = 1; i < kN; ++i)
      state_[i] = 1812433253u * (state_[i - 1] ^ (state_[i - 1] >> 30)) + (uint32_t)i;
    next_ = kN;
  }

  uint32_t operator()()
  {
    if (next_ >= kN)
      regenerate();
    uint32_t y = state_[next_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }

private:
  static constexpr int kN = 624, kM = 397;

  void regenerate()
  {
    for (int i = 0; i < kN; ++i) {
      uint32_t y = (state_[i] & 0x80000000u) | (state_[(i + 1) % kN] & 0x7fffffffu);
      uint32_t v = state_[(i + kM) % kN] ^ (y >> 1);
      state_[i] = (y & 1u) ? (v ^ 0x9908b0dfu) : v;
    }
    next_ = 0;
  }

  uint32_t state_[kN];
  int next_;
};

inline double uniform_real(MT19937& g, double a, double b)
{
  const double lo = (double)g();
  const double hi = (double)g();
  double u = (lo + hi * 4294967296.0) / 18446744073709551616.0;
  if (u >= 1.0)
    u = std::nextafter(1.0, 0.0);
  return u * (b - a) + a;
}

inline int uniform_int_0_intmax(MT19937& g)
{
  return (int)(g() >> 1);
}

}  // namespace lsqt_host
