// hamiltonian.h -- class Hamiltonian with the reference's interface
// (src/hamiltonian.h:25-54).  The operator lives on the GPU behind the C ABI;
// each method is one entry point of include/lsqt_hip.h.
#pragma once

#include "common.h"

class Vector;
class Model;
struct LatticeSpec;
struct lsqt_ham;

class Hamiltonian
{
public:
  // Uploads the model's arrays and, like the reference
  // (src/hamiltonian.cu:27-89), releases them from the model afterwards.
  Hamiltonian(Model&);
  ~Hamiltonian();

  void apply(Vector&, Vector&);             // out = H~ in
  void apply_commutator(Vector&, Vector&);  // out = [X, H~] in
  void apply_current(Vector&, Vector&);     // out = V in
  void kernel_polynomial(Vector&, Vector&, Vector&);
  void chebyshev_01(Vector&, Vector&, Vector&, real, real, int);
  void chebyshev_2(Vector&, Vector&, Vector&, Vector&, real, int);
  void chebyshev_1x(Vector&, Vector&, real);
  void chebyshev_2x(Vector&, Vector&, Vector&, Vector&, Vector&, Vector&, Vector&, real, int);

  lsqt_ham* handle() const { return h; }

private:
  Hamiltonian(const Hamiltonian&) = delete;
  Hamiltonian& operator=(const Hamiltonian&) = delete;
  lsqt_ham* h = nullptr;
};
