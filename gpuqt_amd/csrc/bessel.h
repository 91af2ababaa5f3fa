// bessel.h -- host-side Bessel coefficients of the evolution (internal).
#pragma once

#include <cstddef>
#include <vector>

namespace lsqt {

struct BesselTerms {
  double b0 = 0.0;          // J_0(t)
  double b1 = 0.0;          // 2 J_1(t)
  std::vector<double> bm;   // 2 J_m(t), m = 2, 3, ... (last term kept)
  std::vector<int> label;   // phase case of each term (1..4)
};

// What the launch of term k (k = 0, 1, ... of bm) adds to the accumulated
// state when terms are added in pairs: launch 2i leaves the state alone (its
// term is the next launch's source vector), launch 2i + 1 adds term 2i and
// then term 2i + 1, and a last term of even index is added by its own
// launch.  The same additions in the same order as one term per launch, with
// the state read and written every second launch.
struct DeferredTerm {
  double bm = 0.0;
  int label = 0;  // 0: nothing
  double bm_prev = 0.0;
  int label_prev = 0;
};

inline DeferredTerm deferred_term(const BesselTerms& bt, std::size_t k)
{
  DeferredTerm d;
  if (k % 2 == 1) {
    d.bm_prev = bt.bm[k - 1];
    d.label_prev = bt.label[k - 1];
  } else if (k + 1 < bt.bm.size()) {
    return d;  // added by launch k + 1
  }
  d.bm = bt.bm[k];
  d.label = bt.label[k];
  return d;
}

// U(direction * t)   (src/sigma.cu:116-151)
BesselTerms evolve_terms(int direction, double t);
// [X, U(t)] / [U(-t), X]   (src/sigma.cu:157-200)
BesselTerms evolvex_terms(int direction, double t);

}  // namespace lsqt
