// sigma.h -- the quantity drivers with the reference's interface
// (src/sigma.h:26-30) and the algorithm functions of src/sigma.cu that they
// are built from.
#pragma once

#include <vector>

#include "common.h"

class Model;
class Hamiltonian;
class Vector;

void find_dos(Model&, Hamiltonian&, Vector&, int);
void find_vac0(Model&, Hamiltonian&, Vector&);
void find_vac(Model&, Hamiltonian&, Vector&);
void find_msd(Model&, Hamiltonian&, Vector&);
void find_spin_polarization(Model&, Hamiltonian&, Vector&);

// mu_m = <left|T_m(H~)|right>, m < number_of_moments, into host arrays
// (src/sigma.cu:33-61; the whole recursion runs on the GPU)
void find_moments_chebyshev(Model&, Hamiltonian&, Vector& left, Vector& right, real* mu_re,
                            real* mu_im);
void apply_damping(Model&, real* mu_re, real* mu_im);                    // src/sigma.cu:76-86
void perform_chebyshev_summation(Model&, real*, real*, real* out);     // src/sigma.cu:89-110
void evolve(Model&, int direction, real time_step_scaled, Hamiltonian&, Vector&);   // :116-151
void evolvex(Model&, int direction, real time_step_scaled, Hamiltonian&, Vector&);  // :157-200

namespace lsqt_host {
// apply_damping + perform_chebyshev_summation on plain arrays: mu_re is
// damped in place (mu_im too, when non-null), out gets n_energy values
void damped_chebyshev_sum(int M, real* mu_re, real* mu_im, int n_energy, const real* energy,
                          real energy_max, real volume, real* out);
// Significant digits of the .out rows: 6 (the stream default the reference
// writes with) unless set otherwise.
void set_output_precision(int digits);
int output_precision();
}  // namespace lsqt_host
