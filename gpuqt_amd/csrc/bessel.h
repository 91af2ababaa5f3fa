// bessel.h -- host-side Bessel coefficients of the evolution (internal).
#pragma once

#include <vector>

namespace lsqt {

struct BesselTerms {
  double b0 = 0.0;          // J_0(t)
  double b1 = 0.0;          // 2 J_1(t)
  std::vector<double> bm;   // 2 J_m(t), m = 2, 3, ... (last term kept)
  std::vector<int> label;   // phase case of each term (1..4)
};

// U(direction * t)   (src/sigma.cu:116-151)
BesselTerms evolve_terms(int direction, double t);
// [X, U(t)] / [U(-t), X]   (src/sigma.cu:157-200)
BesselTerms evolvex_terms(int direction, double t);

}  // namespace lsqt
