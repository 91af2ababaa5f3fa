/*
 * orc_rng.c -- std::mt19937 and the two libstdc++ distributions GPUQT draws
 * from.  Test infrastructure only (see lsqt_oracle.h).
 *
 * Reference call sites: src/model.cu:34 (seed 12345678 under -DDEBUG),
 * src/model.cu:79 and :100-104 (random phases), src/anderson.cu:24-28,
 * src/charge.cu:52,68, src/model_lattice.cu:32.
 *
 * The engine is MT19937 as specified by Matsumoto & Nishimura (1998) and by
 * [rand.eng.mers] of the C++ standard (w=32, n=624, m=397, r=31,
 * a=0x9908b0df, u=11, s=7/b=0x9d2c5680, t=15/c=0xefc60000, l=18,
 * f=1812433253).  The distributions are implementation-defined; these follow
 * the algorithms of GCC 11's libstdc++:
 *   - uniform_real_distribution<double>(a,b): generate_canonical<double,53>
 *     draws k = ceil(53/32) = 2 words, forms (w0 + w1*2^32) in double, divides
 *     by 2^64, clamps a result of 1 to the largest double below 1, then
 *     returns u*(b-a) + a.
 *   - uniform_int_distribution<int>(0, INT_MAX) over a 32-bit engine uses
 *     Lemire's multiply-shift (ACM TOMACS 29(1), 2019) with range 2^31; the
 *     rejection threshold (2^32 mod 2^31) is 0, so the result is w >> 1.
 * tests/test_oracle_rng.py pins the first draws of each against the C++
 * standard library compiled here, and the 10000th draw of the engine
 * against the value the standard fixes (4123659995 for the default seed).
 */
#include "lsqt_oracle.h"

#include <math.h>

void orc_mt_seed(orc_mt19937* g, uint32_t seed)
{
  g->mt[0] = seed;
  for (int i = 1; i < 624; ++i) {
    uint32_t prev = g->mt[i - 1];
    g->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
  }
  g->mti = 624;
}

static void mt_twist(orc_mt19937* g)
{
  for (int i = 0; i < 624; ++i) {
    uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
    uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
    if (y & 1u)
      v ^= 0x9908b0dfu;
    g->mt[i] = v;
  }
  g->mti = 0;
}

uint32_t orc_mt_next(orc_mt19937* g)
{
  if (g->mti >= 624)
    mt_twist(g);
  uint32_t y = g->mt[g->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double orc_uniform_real(orc_mt19937* g, double a, double b)
{
  double w0 = (double)orc_mt_next(g);
  double w1 = (double)orc_mt_next(g);
  double sum = w0 + w1 * 4294967296.0;
  double u = sum / 18446744073709551616.0;
  if (u >= 1.0)
    u = nextafter(1.0, 0.0);
  return u * (b - a) + a;
}

int orc_uniform_int_0_intmax(orc_mt19937* g)
{
  return (int)(orc_mt_next(g) >> 1);
}
