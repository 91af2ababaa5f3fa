// rng_draws.cpp -- prints draws of the C++ standard library's mt19937 and of
// the two distributions the reference uses (src/model.cu:79,
// src/anderson.cu:25, src/model_lattice.cu:32), as JSON.  Built and run by
// tests/golden/make_rng_draws.py with the host g++; the output is the
// fixture tests/golden/rng_draws.json.
#include <climits>
#include <cstdio>
#include <random>

int main()
{
  const double PI = 3.141592653589793;
  printf("{\n");
  {
    std::mt19937 g;
    unsigned long v = 0;
    for (int i = 0; i < 10000; ++i)
      v = g();
    printf("\"default_10000th\": %lu,\n", v);
  }
  {
    std::mt19937 g(12345678);
    printf("\"raw_12345678\": [");
    for (int i = 0; i < 2000; ++i)
      printf("%s%lu", i ? ", " : "", (unsigned long)g());
    printf("],\n");
  }
  {
    std::mt19937 g(12345678);
    std::uniform_real_distribution<double> phase(0, 2 * PI);
    printf("\"phase_12345678\": [");
    for (int i = 0; i < 1500; ++i)
      printf("%s\"%a\"", i ? ", " : "", phase(g));
    printf("],\n");
  }
  {
    std::mt19937 g(777);
    std::uniform_real_distribution<double> w(-0.5 * 8.25, 0.5 * 8.25);
    printf("\"anderson_777\": [");
    for (int i = 0; i < 1500; ++i)
      printf("%s\"%a\"", i ? ", " : "", w(g));
    printf("],\n");
  }
  {
    std::mt19937 g(12345678);
    std::uniform_int_distribution<int> r(0, INT_MAX);
    printf("\"int_12345678\": [");
    for (int i = 0; i < 1500; ++i)
      printf("%s%d", i ? ", " : "", r(g));
    printf("]\n");
  }
  printf("}\n");
  return 0;
}
