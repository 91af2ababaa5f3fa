// model.h -- class Model with the reference's public interface
// (src/model.h:27-104): input parsing, lattice / general model construction,
// disorder, and the random vectors.  The arrays are the row-major ELL arrays
// the reference builds (pair (n, m) at n*max_neighbor + m); construction
// follows the reference's order of parsing and of random draws exactly, so
// the same inputs give the same bits.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "rng.h"

class Vector;

// The contents of a lattice.in (src/model_lattice.cu:165-341) and, with
// vacancies, the index of every pristine site after their removal (-1 for a
// vacancy): what the Hamiltonian needs to expand the neighbour lists itself.
struct LatticeSpec {
  int N_cell[3] = {0, 0, 0};
  int pbc[3] = {0, 0, 0};
  int transport_direction = 0;
  real lattice_constant[3] = {0, 0, 0};
  int N_orbital = 0;
  int max_neighbor = 0;
  std::vector<real> x_cell;            // [N_orbital]
  std::vector<int> number_of_hoppings; // [N_orbital]
  std::vector<int> hop_cell;           // [N_orbital][max_neighbor][3] cell offsets
  std::vector<int> hop_orbital;        // [N_orbital][max_neighbor]
  std::vector<real> hop_re, hop_im;    // [N_orbital][max_neighbor]
  std::vector<int> site_index;         // [pristine sites], empty without vacancies
};

class Model
{
public:
  // Throws lsqt_host::InputError with the reference's message where the
  // reference prints it and exits.
  Model(std::string input_dir);
  // lattice_on_device: a lattice model keeps its description (lattice())
  // instead of expanding the neighbour arrays on the host; the Hamiltonian
  // then expands it on the GPU.  The random draws, the potential and every
  // other member are those of Model(input_dir).
  Model(std::string input_dir, bool lattice_on_device);
  ~Model();
  // the description of a lattice model built with lattice_on_device (its
  // neighbour arrays not yet expanded), else null
  const LatticeSpec* lattice() const;
  // expands a lattice_on_device model's neighbour arrays on the host after
  // all (the arrays of Model(input_dir), bit for bit)
  void expand_lattice_arrays();
  void initialize_state(Vector& random_state, int orbital);
  // the same, into host arrays of number_of_atoms values
  void initialize_state_host(real* re, real* im, int orbital);

  bool calculate_vac0 = false;
  bool calculate_vac = false;
  bool calculate_msd = false;
  bool calculate_spin = false;
  bool calculate_ldos = false;

  int number_of_random_vectors = 1;
  int number_of_atoms = 0;
  int max_neighbor = 0;
  int number_of_pairs = 0;
  int number_of_energy_points = 0;
  int number_of_moments = 1000;
  int number_of_steps_correlation = 0;
  int number_of_local_orbitals = 0;
  std::string input_dir;
  real energy_max = 10;

  real* energy = nullptr;
  real* time_step = nullptr;
  std::vector<int> local_orbitals;

  int* neighbor_number = nullptr;
  int* neighbor_list = nullptr;
  real* xx = nullptr;
  real* potential = nullptr;
  real* hopping_real = nullptr;
  real* hopping_imag = nullptr;

  real volume = 0;

  // releases the Hamiltonian arrays (the Hamiltonian calls this after upload)
  void release_hamiltonian_arrays();

private:
  Model(const Model&) = delete;
  Model& operator=(const Model&) = delete;

  void initialize_parameters();
  void verify_parameters();
  void initialize_energy();
  void initialize_time();
  void initialize_local_orbitals();

  void initialize_model_general();
  void initialize_neighbor();
  void initialize_positions();
  void initialize_potential();
  void initialize_hopping();

  void initialize_lattice_model();
  void draw_vacancies(std::vector<int>& new_index);
  void expand_lattice(const LatticeSpec& L);
  void remove_vacancies(const std::vector<int>& new_index);
  void add_anderson_disorder();
  void add_charged_impurities(const int pbc[3], const real box_length[3],
                              const std::vector<real>& x, const std::vector<real>& y,
                              const std::vector<real>& z);
  void random_permutation_prefix(int max_value, int total_number, int* out);

  bool requires_time = false;
  bool use_lattice_model = false;

  bool has_anderson = false;
  real anderson_strength = 0;
  bool has_charge = false;
  int charge_number = 0;
  real charge_strength = 0;
  real charge_range = 0;
  bool has_vacancy_disorder = false;
  int number_of_vacancies = 0;

  lsqt_host::MT19937 generator;
  bool defer_lattice = false;
  std::unique_ptr<LatticeSpec> spec;  // set while the arrays are not expanded
};
