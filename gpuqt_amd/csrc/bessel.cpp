// bessel.cpp -- Chebyshev coefficients of the time-evolution operator,
// computed on the host with the C library's Bessel functions exactly as the
// reference does (src/sigma.cu:125-146 for U(t), :171-190 for [X, U(t)]).
// Compiled by the host C++ compiler so that j0/j1/jn are the C library's.
#include "bessel.h"

#include <math.h>

namespace lsqt {

// term m >= 2 is kept while |J_m(t)| >= 1e-15; its coefficient is 2 J_m(t)
static void bessel_tail(double t, std::vector<double>& b, std::vector<int>& m_index)
{
  for (int m = 2; m < 1000000; ++m) {
    double bm = jn(m, t);
    if (bm < 1.0e-15 && bm > -1.0e-15)
      break;
    b.push_back(bm * 2.0);
    m_index.push_back(m);
  }
}

BesselTerms evolve_terms(int direction, double t)
{
  BesselTerms r;
  r.b0 = j0(t);
  r.b1 = 2.0 * j1(t);
  std::vector<int> ms;
  bessel_tail(t, r.bm, ms);
  for (int m : ms) {
    int k = m % 4, label;
    if (k == 0)
      label = 1;
    else if (k == 2)
      label = 2;
    else if ((k == 1 && direction == 1) || (k == 3 && direction == -1))
      label = 3;
    else
      label = 4;
    r.label.push_back(label);
  }
  return r;
}

BesselTerms evolvex_terms(int direction, double t)
{
  BesselTerms r;
  r.b0 = 0.0;
  r.b1 = 2.0 * j1(t);
  std::vector<int> ms;
  bessel_tail(t, r.bm, ms);
  for (int m : ms) {
    int k = m % 4, label;
    if (k == 1)
      label = 3;
    else if (k == 3)
      label = 4;
    else if ((k == 0 && direction == 1) || (k == 2 && direction == -1))
      label = 1;
    else
      label = 2;
    r.label.push_back(label);
  }
  return r;
}

}  // namespace lsqt
