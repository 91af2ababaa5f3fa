// sigma.h -- the quantity drivers with the reference's interface
// (src/sigma.h:26-30) and the algorithm functions of src/sigma.cu that they
// are built from.
#pragma once

#include <vector>

#include "common.h"

class Model;
class Hamiltonian;
class Vector;
struct lsqt_ham;

void find_dos(Model&, Hamiltonian&, Vector&, int);
void find_vac0(Model&, Hamiltonian&, Vector&);
void find_vac(Model&, Hamiltonian&, Vector&);
void find_msd(Model&, Hamiltonian&, Vector&);
void find_spin_polarization(Model&, Hamiltonian&, Vector&);

// mu_m = <left|T_m(H~)|right>, m < number_of_moments, into host arrays
// (src/sigma.cu:33-61; the whole recursion runs on the GPU)
void find_moments_chebyshev(Model&, Hamiltonian&, Vector& left, Vector& right, real* mu_re,
                            real* mu_im);
// the reference's signature: output[m] = mu_m, left on the device
void find_moments_chebyshev(Model&, Hamiltonian&, Vector& left, Vector& right, Vector& output);
void apply_damping(Model&, real* mu_re, real* mu_im);                    // src/sigma.cu:76-86
void perform_chebyshev_summation(Model&, real*, real*, real* out);     // src/sigma.cu:89-110
void evolve(Model&, int direction, real time_step_scaled, Hamiltonian&, Vector&);   // :116-151
void evolvex(Model&, int direction, real time_step_scaled, Hamiltonian&, Vector&);  // :157-200

namespace lsqt_host {
// The same algorithms sequenced method by method on Hamiltonian and Vector,
// as the reference's own functions drive them: the recursion one
// kernel_polynomial and one Vector::inner_product_1 per moment and
// inner_product_2 at the end (src/sigma.cu:33-61); the evolutions one
// chebyshev_2 / chebyshev_2x per Bessel term with the states rotated by
// Vector::swap (src/sigma.cu:116-200).  Their states are bit-identical to the
// fused entry points' above (the moments differ by reduction order only);
// they serve callers that drive the methods themselves.
void find_moments_by_terms(Model&, Hamiltonian&, Vector& left, Vector& right, Vector& output);
void evolve_by_terms(Model&, int direction, real time_step_scaled, Hamiltonian&, Vector& state);
void evolvex_by_terms(Model&, int direction, real time_step_scaled, Hamiltonian&, Vector& state);

// apply_damping + perform_chebyshev_summation on plain arrays: mu_re is
// damped in place (mu_im too, when non-null), out gets n_energy values
void damped_chebyshev_sum(int M, real* mu_re, real* mu_im, int n_energy, const real* energy,
                          real energy_max, real volume, real* out);
// Significant digits of the .out rows: 6 (the stream default the reference
// writes with) unless set otherwise.
void set_output_precision(int digits);
int output_precision();

// Which device path the drivers use (environment LSQT_PATH):
//   batch all directories of the driver file with the same moments, time
//         steps, energy_max, flags and vector count at once, as independent
//         systems of one segmented Hamiltonian on the block path (others,
//         and those asking for spin or LDOS, run one by one on the block
//         path);
//   f64   (default) the reference's arithmetic, bit-faithful states;
//   f32   FP32-stored recursion states, FP64 arithmetic, moments and
//         re-normalised evolved states (lsqt_moments_f32, lsqt_evolve_f32);
//   block all random vectors of a directory at once (lsqt_moments_block_self,
//         lsqt_block_vac / msd): FP32 recursion, FP64 moments and evolved
//         states; the per-vector leftovers (spin, LDOS) use f32.
enum class Path { f64, f32, block, batch };
Path path();
void set_path(Path p);
Path path_from_environment();

// One system of a block run: its Model, whose sites are [start, start + n)
// of the (possibly segmented) Hamiltonian, and its number of random vectors.
struct BlockMember {
  Model* model;
  int start;
  int used;
};
// The block path's drivers: DOS, VAC0, VAC and MSD of R vectors for every
// member at once (host arrays [R][n] over all members' sites; member k's
// first `used` vectors are its random vectors, the rest zero padding), each
// member's rows appended to its own .out files in the reference's order
// (vector by vector; Nt rows per vector for VAC/MSD).  All members share the
// moments, time steps, energy_max and flags of members[0].  h is segmented
// with one segment per member when there is more than one.
void find_block_quantities(std::vector<BlockMember>& members, lsqt_ham* h, int R,
                           const std::vector<real>& re, const std::vector<real>& im);
// Spin polarization (S.out) is included for a single member.  The LDOS of
// all local orbitals (up to 256 per block) on the block path:
void find_block_ldos(Model& model, lsqt_ham* h);
}  // namespace lsqt_host
