// model.h -- class Model with the reference's public interface
// (src/model.h:27-104): input parsing, lattice / general model construction,
// disorder, and the random vectors.  The arrays are the row-major ELL arrays
// the reference builds (pair (n, m) at n*max_neighbor + m); construction
// follows the reference's order of parsing and of random draws exactly, so
// the same inputs give the same bits.
#pragma once

#include <string>
#include <vector>

#include "common.h"
#include "rng.h"

class Vector;

class Model
{
public:
  // Throws lsqt_host::InputError with the reference's message where the
  // reference prints it and exits.
  Model(std::string input_dir);
  ~Model();
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
  void add_vacancies();
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
};
