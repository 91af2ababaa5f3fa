// model.cpp -- Model: para.in / energy.in / time_step.in / local_orbitals.in,
// the lattice and general models, vacancies, Anderson disorder and charged
// impurities (reference: src/model.cu, src/model_lattice.cu,
// src/model_general.cu, src/anderson.cu, src/charge.cu).
//
// What must stay as the reference has it, because it fixes the bits of
// every result: the keyword grammar and stream parsing of each file, the site
// index ((nx*Ny + ny)*Nz + nz)*N_orbital + m with x slowest
// (src/model_lattice.cu:148-163), the neighbour order (hoppings in file order
// minus open-boundary skips), the transport offset taken from the orbitals' x
// coordinates whatever the transport direction (src/model_lattice.cu:373-375),
// vacancies as a full shuffle with one draw per pristine site, the RNG order
// vacancies -> Anderson -> impurities -> random phases, impurity positions
// indexed by the pristine lattice, and the cell-list double counting for
// fewer than three cells (src/charge.cu:152-171).
//
// Deliberate differences: inputs on which the reference loops forever or
// writes out of bounds (an impurity cell count of 0, more hoppings than
// max_neighbor, neighbour indices out of range, spin with an odd site count,
// fewer than 2 moments) are reported as errors.
#include "model.h"

#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "text_reader.h"
#include "vector.h"

#define PI 3.141592653589793

using lsqt_host::InputError;

namespace {

void print_started_reading(const std::string& filename)
{
  std::cout << std::endl;
  std::cout << "===========================================================";
  std::cout << std::endl;
  std::cout << "Started reading " + filename << std::endl;
  std::cout << std::endl;
}

void print_finished_reading(const std::string& filename)
{
  std::cout << std::endl;
  std::cout << "Finished reading " + filename << std::endl;
  std::cout << "===========================================================";
  std::cout << std::endl << std::endl;
}

template <typename T>
T* new_zeroed(size_t n)
{
  T* p = new T[n ? n : 1];
  std::memset(p, 0, sizeof(T) * (n ? n : 1));
  return p;
}

}  // namespace

Model::Model(std::string input_dir) : Model(input_dir, false) {}

Model::Model(std::string input_dir, bool lattice_on_device)
  : generator(12345678u), defer_lattice(lattice_on_device)
{
  this->input_dir = input_dir;
  try {
    initialize_parameters();
    if (use_lattice_model)
      initialize_lattice_model();
    else
      initialize_model_general();
    if (calculate_spin && number_of_atoms % 2 != 0)
      throw InputError("Error: spin polarization needs an even number of orbitals");
    initialize_energy();
    if (requires_time)
      initialize_time();
    if (calculate_ldos)
      initialize_local_orbitals();
  } catch (...) {
    release_hamiltonian_arrays();
    delete[] energy;
    delete[] time_step;
    energy = time_step = nullptr;
    throw;
  }
}

Model::~Model()
{
  release_hamiltonian_arrays();
  delete[] energy;
  delete[] time_step;
}

const LatticeSpec* Model::lattice() const
{
  return spec.get();
}

void Model::release_hamiltonian_arrays()
{
  spec.reset();
  delete[] neighbor_number;
  delete[] neighbor_list;
  delete[] potential;
  delete[] hopping_real;
  delete[] hopping_imag;
  delete[] xx;
  neighbor_number = neighbor_list = nullptr;
  potential = hopping_real = hopping_imag = xx = nullptr;
}

// ---- random vectors (src/model.cu:77-109) ----------------------------------
void Model::initialize_state_host(real* re, real* im, int orbital)
{
  if (orbital >= 0) {
    for (int n = 0; n < number_of_atoms; ++n) {
      re[n] = 0.0;
      im[n] = 0.0;
    }
    re[orbital] = sqrt(1.0 * number_of_atoms);
  } else if (calculate_spin) {
    // even sites only: normalised to N/2 to remove the spin degeneracy
    for (int n = 0; n < number_of_atoms; n += 2) {
      real phase = lsqt_host::uniform_real(generator, 0, 2 * PI);
      re[n] = cos(phase);
      im[n] = sin(phase);
      re[n + 1] = 0.0;
      im[n + 1] = 0.0;
    }
  } else {
    for (int n = 0; n < number_of_atoms; ++n) {
      real phase = lsqt_host::uniform_real(generator, 0, 2 * PI);
      re[n] = cos(phase);
      im[n] = sin(phase);
    }
  }
}

void Model::initialize_state(Vector& random_state, int orbital)
{
  std::vector<real> re((size_t)number_of_atoms), im((size_t)number_of_atoms);
  initialize_state_host(re.data(), im.data(), orbital);
  random_state.copy_from_host(re.data(), im.data());
}

// ---- para.in (src/model.cu:243-310) ----------------------------------------
void Model::initialize_parameters()
{
  std::string filename = input_dir + "/para.in";
  std::ifstream input(filename);
  if (!input.is_open())
    throw InputError("Error: cannot open " + filename);
  print_started_reading(filename);

  std::string line;
  while (std::getline(input, line)) {
    std::stringstream ss(line);
    std::string token;
    ss >> token;
    if (token == "")
      continue;
    if (token == "model") {
      ss >> use_lattice_model;
    } else if (token == "anderson_disorder") {
      has_anderson = true;
      ss >> anderson_strength;
    } else if (token == "charged_impurity") {
      has_charge = true;
      ss >> charge_number >> charge_strength >> charge_range;
    } else if (token == "vacancy_disorder") {
      has_vacancy_disorder = true;
      ss >> number_of_vacancies;
    } else if (token == "calculate_vac0") {
      calculate_vac0 = true;
    } else if (token == "calculate_vac") {
      calculate_vac = true;
    } else if (token == "calculate_msd") {
      calculate_msd = true;
    } else if (token == "calculate_spin") {
      calculate_spin = true;
    } else if (token == "calculate_ldos") {
      calculate_ldos = true;
    } else if (token == "number_of_random_vectors") {
      ss >> number_of_random_vectors;
    } else if (token == "number_of_moments") {
      ss >> number_of_moments;
    } else if (token == "energy_max") {
      ss >> energy_max;
    } else {
      std::ostringstream msg;
      msg << "Error: Unknown identifier in " << input_dir + "/para.in: " + line << std::endl
          << "Valid keywords include: " << std::endl
          << "--model" << std::endl
          << "--anderson_disorder" << std::endl
          << "--charged_impurity" << std::endl
          << "--vacancy_disorder" << std::endl
          << "--calculate_vac0" << std::endl
          << "--calculate_vac" << std::endl
          << "--calculate_msd" << std::endl
          << "--calculate_spin" << std::endl
          << "--calculate_ldos" << std::endl
          << "--number_of_random_vectors" << std::endl
          << "--number_of_moments" << std::endl
          << "--energy_max";
      throw InputError(msg.str());
    }
  }
  verify_parameters();
  print_finished_reading(filename);
}

// src/model.cu:128-241
void Model::verify_parameters()
{
  if (calculate_vac || calculate_msd || calculate_spin)
    requires_time = true;

  if (use_lattice_model) {
    std::cout << "- Use lattice model" << std::endl;
    if (calculate_spin)
      throw InputError("Error: lattice model does not support spin calculations yet");
    if (has_anderson)
      std::cout << "- Add Anderson disorder with strength W = " << anderson_strength << std::endl;
    else
      std::cout << "- No Anderson disorder" << std::endl;
    if (has_charge)
      std::cout << "- Add charged impurities with " << std::endl
                << "  N = " << charge_number << std::endl
                << "  W = " << charge_strength << std::endl
                << "  xi = " << charge_range << std::endl;
    else
      std::cout << "- No charged impurity" << std::endl;
    if (has_vacancy_disorder) {
      std::cout << "- Add " << number_of_vacancies << " vacancies" << std::endl;
      if (number_of_vacancies <= 0)
        throw InputError("Error: number of vacancies should > 0");
    } else {
      std::cout << "- No vacancy disorder" << std::endl;
    }
  } else {
    std::cout << "- Use general model" << std::endl;
    if (has_anderson)
      throw InputError("Error: General model does not allowed to add Anderson disorder");
    if (has_vacancy_disorder)
      throw InputError("Error: General model does not allowed to add vacancy disorder");
    if (has_charge)
      throw InputError("Error: General model does not allowed to add charged impurities");
  }

  std::cout << "- DOS will be calculated" << std::endl;
  std::cout << (calculate_vac0 ? "- VAC0 will be calculated" : "- VAC0 will not be calculated")
            << std::endl;
  std::cout << (calculate_vac ? "- VAC will be calculated" : "- VAC will not be calculated")
            << std::endl;
  std::cout << (calculate_msd ? "- MSD will be calculated" : "- MSD will not be calculated")
            << std::endl;
  std::cout << (calculate_spin ? "- spin polarization will be calculated"
                               : "- spin polarization will not be calculated")
            << std::endl;
  if (calculate_spin && calculate_vac0)
    throw InputError("Error: spin and VAC0 cannot be calculated together");
  if (calculate_spin && calculate_vac)
    throw InputError("Error: spin and VAC cannot be calculated together");
  if (calculate_spin && calculate_msd)
    throw InputError("Error: spin and MSD cannot be calculated together");

  std::cout << "- Number of random vectors is " << number_of_random_vectors << std::endl;
  if (number_of_random_vectors <= 0)
    throw InputError("Error: Number of random vectors should > 0");
  std::cout << "- Number of moments is " << number_of_moments << std::endl;
  if (number_of_moments <= 0)
    throw InputError("Error: Number of moments should > 0");
  // the moment recursion and the summation read moments 0 and 1
  // (src/sigma.cu:43-49,97); the reference has no check for M = 1
  if (number_of_moments < 2)
    throw InputError("Error: Number of moments should be >= 2");
  std::cout << "- Energy maximum is " << energy_max << std::endl;
  if (energy_max <= 0)
    throw InputError("Error: Energy maximum should > 0");
}

// ---- count-then-values files (src/model.cu:312-384) ------------------------
void Model::initialize_energy()
{
  std::string filename = input_dir + "/energy.in";
  std::ifstream input(filename);
  if (!input.is_open())
    throw InputError("Error: cannot open " + filename);
  print_started_reading(filename);
  input >> number_of_energy_points;
  if (number_of_energy_points < 0)
    number_of_energy_points = 0;
  std::cout << "- number of energy points = " << number_of_energy_points << std::endl;
  energy = new_zeroed<real>((size_t)number_of_energy_points);
  for (int n = 0; n < number_of_energy_points; ++n)
    input >> energy[n];
  print_finished_reading(filename);
}

void Model::initialize_local_orbitals()
{
  std::string filename = input_dir + "/local_orbitals.in";
  std::ifstream input(filename);
  if (!input.is_open())
    throw InputError("Error: cannot open " + filename);
  print_started_reading(filename);
  input >> number_of_local_orbitals;
  if (number_of_local_orbitals < 0)
    number_of_local_orbitals = 0;
  std::cout << "- number of local orbitals = " << number_of_local_orbitals << std::endl;
  local_orbitals.assign((size_t)number_of_local_orbitals, 0);
  for (int n = 0; n < number_of_local_orbitals; ++n) {
    input >> local_orbitals[n];
    if (local_orbitals[n] < 0 || local_orbitals[n] >= number_of_atoms)
      throw InputError("Error: wrong local orbital index: " + std::to_string(local_orbitals[n]));
  }
  print_finished_reading(filename);
}

void Model::initialize_time()
{
  std::string filename = input_dir + "/time_step.in";
  std::ifstream input(filename);
  if (!input.is_open())
    throw InputError("Error: cannot open " + filename);
  print_started_reading(filename);
  input >> number_of_steps_correlation;
  if (number_of_steps_correlation < 0)
    number_of_steps_correlation = 0;
  std::cout << "- number of time steps = " << number_of_steps_correlation << std::endl;
  time_step = new_zeroed<real>((size_t)number_of_steps_correlation);
  for (int n = 0; n < number_of_steps_correlation; ++n)
    input >> time_step[n];
  print_finished_reading(filename);
}

// ---- general model (src/model_general.cu) ----------------------------------
void Model::initialize_model_general()
{
  initialize_neighbor();
  initialize_positions();
  initialize_potential();
  initialize_hopping();
}

// Each file of the general model is parsed by one template over the input
// type: first from a TextReader (mapped file, exact fast conversion); if that
// meets a token outside the plain number grammar, again from std::ifstream,
// so that unusual inputs behave exactly as with the stream (see
// text_reader.h).  Messages are printed once, after the parse.
namespace {

template <class In>
void parse_neighbors(In& input, const std::string& filename, int& n_atoms, int& max_nb,
                     std::vector<int>& nn, std::vector<int>& nl)
{
  input >> n_atoms >> max_nb;
  if (n_atoms <= 0 || max_nb <= 0 || (long long)n_atoms * max_nb > INT_MAX)
    throw InputError("Error: invalid header in " + filename);
  nn.assign((size_t)n_atoms, 0);
  nl.assign((size_t)n_atoms * max_nb, 0);
  for (int n = 0; n < n_atoms; ++n) {
    input >> nn[n];
    if (nn[n] < 0 || nn[n] > max_nb)
      throw InputError("Error: site " + std::to_string(n) + " has more neighbors than " +
                       std::to_string(max_nb));
    for (int m = 0; m < nn[n]; ++m) {
      int& j = nl[(size_t)n * max_nb + m];
      input >> j;
      if (j < 0 || j >= n_atoms)
        throw InputError("Error: neighbor index out of range at site " + std::to_string(n));
    }
  }
}

template <class In>
void parse_values(In& input, std::vector<real>& v)
{
  for (real& x : v)
    input >> x;
}

template <class In>
int parse_hopping_kind(In& input)
{
  std::string first;
  input >> first;
  if (first == "complex")
    return 1;
  if (first == "real")
    return 2;
  return 0;
}

template <class In>
void parse_hoppings(In& input, int type, int n_atoms, int max_nb, const int* nn,
                    std::vector<real>& hr, std::vector<real>& hi)
{
  hr.assign((size_t)n_atoms * max_nb, 0.0);
  hi.assign((size_t)n_atoms * max_nb, 0.0);
  for (int n = 0; n < n_atoms; ++n)
    for (int m = 0; m < nn[n]; ++m) {
      size_t q = (size_t)n * max_nb + m;
      input >> hr[q];
      if (type == 1)
        input >> hi[q];
    }
}

// runs parse(reader) on the mapped file and, if it met an unusual token,
// parse(stream) instead; false if the file does not exist
template <class F>
bool parse_file(const std::string& filename, F parse)
{
  // LSQT_STREAM_READER=1 parses with the stream only (to check the fast path)
  const char* stream_only = std::getenv("LSQT_STREAM_READER");
  if (!(stream_only && stream_only[0] == '1')) {
    lsqt_host::TextReader fast(filename);
    if (!fast.is_open())
      return false;
    bool done = false;
    try {
      parse(fast);
      done = !fast.unusual();
    } catch (const InputError&) {
      if (!fast.unusual())
        throw;
    }
    if (done)
      return true;
  }
  std::ifstream input(filename);
  if (!input.is_open())
    return false;
  parse(input);
  return true;
}

}  // namespace

void Model::initialize_neighbor()
{
  std::string filename = input_dir + "/neighbor.in";
  std::vector<int> nn, nl;
  int n_atoms = 0, max_nb = 0;
  bool found = false;
  {
    // the file must open before anything is printed (the reference's order)
    std::ifstream probe(filename);
    found = probe.is_open();
  }
  if (!found)
    throw InputError("Error: cannot open " + filename);
  print_started_reading(filename);
  if (!parse_file(filename, [&](auto& in) { parse_neighbors(in, filename, n_atoms, max_nb, nn, nl); }))
    throw InputError("Error: cannot open " + filename);
  number_of_atoms = n_atoms;
  max_neighbor = max_nb;
  number_of_pairs = number_of_atoms * max_neighbor;
  neighbor_number = new_zeroed<int>((size_t)number_of_atoms);
  neighbor_list = new_zeroed<int>((size_t)number_of_pairs);
  std::memcpy(neighbor_number, nn.data(), sizeof(int) * nn.size());
  std::memcpy(neighbor_list, nl.data(), sizeof(int) * nl.size());
  std::cout << "- Number of orbitals is " << number_of_atoms << std::endl;
  std::cout << "- Maximum neighbor number is " << max_neighbor << std::endl;
  print_finished_reading(filename);
}

static real reduce_distance(real d, real box)
{
  if (d > box / 2.0)
    return d - box;
  if (d < -box / 2.0)
    return d + box;
  return d;
}

void Model::initialize_positions()
{
  std::string filename = input_dir + "/position.in";
  {
    std::ifstream probe(filename);
    if (!probe.is_open())
      throw InputError("Error: cannot open " + filename);
  }
  print_started_reading(filename);
  real box = 0.0, vol = 0.0;
  std::vector<real> x;
  parse_file(filename, [&](auto& in) {
    box = vol = 0.0;
    in >> box >> vol;
    x.assign((size_t)number_of_atoms, 0.0);
    parse_values(in, x);
  });
  volume = vol;
  std::cout << "- Box length along the transport direction is " << box << std::endl;
  std::cout << "- System volume is " << volume << std::endl;
  xx = new_zeroed<real>((size_t)number_of_pairs);
  for (int n = 0; n < number_of_atoms; ++n)
    for (int m = 0; m < neighbor_number[n]; ++m) {
      size_t q = (size_t)n * max_neighbor + m;
      xx[q] = reduce_distance(x[neighbor_list[q]] - x[n], box);
    }
  print_finished_reading(filename);
}

void Model::initialize_potential()
{
  std::string filename = input_dir + "/potential.in";
  print_started_reading(filename);
  std::vector<real> v((size_t)number_of_atoms, 0.0);
  bool nonzero = parse_file(filename, [&](auto& in) {
    v.assign((size_t)number_of_atoms, 0.0);
    parse_values(in, v);
  });
  if (!nonzero) {
    std::cout << "- Could not open " + filename << std::endl;
    std::cout << "- Assuming zero on-site potential" << std::endl;
  } else {
    std::cout << "- On-site potential will be read in" << std::endl;
  }
  potential = new_zeroed<real>((size_t)number_of_atoms);
  std::memcpy(potential, v.data(), sizeof(real) * v.size());
  print_finished_reading(filename);
}

void Model::initialize_hopping()
{
  std::string filename = input_dir + "/hopping.in";
  print_started_reading(filename);
  // 1: complex, 2: real, 3: uniform -1 (no file)
  int type = 3;
  std::vector<real> hr, hi;
  bool found = parse_file(filename, [&](auto& in) {
    type = parse_hopping_kind(in);
    if (type == 0)
      throw InputError("- Hoppings can only be real or complex");
    parse_hoppings(in, type, number_of_atoms, max_neighbor, neighbor_number, hr, hi);
  });
  hopping_real = new_zeroed<real>((size_t)number_of_pairs);
  hopping_imag = new_zeroed<real>((size_t)number_of_pairs);
  if (!found) {
    type = 3;
    std::cout << "- Could not open " + filename << std::endl;
    std::cout << "- Assuming uniform hoppings with strength -1" << std::endl;
    for (int n = 0; n < number_of_atoms; ++n)
      for (int m = 0; m < neighbor_number[n]; ++m)
        hopping_real[(size_t)n * max_neighbor + m] = -1.0;
  } else {
    std::cout << (type == 1 ? "- Hoppings have imaginary part" : "- Hoppings are real")
              << std::endl;
    std::memcpy(hopping_real, hr.data(), sizeof(real) * hr.size());
    std::memcpy(hopping_imag, hi.data(), sizeof(real) * hi.size());
  }
  print_finished_reading(filename);
}

// ---- lattice model (src/model_lattice.cu) ----------------------------------
static int find_index(int nx, int ny, int nz, int Nx, int Ny, int Nz, int m, int N_orbital)
{
  if (nx < 0)
    nx += Nx;
  if (nx >= Nx)
    nx -= Nx;
  if (ny < 0)
    ny += Ny;
  if (ny >= Ny)
    ny -= Ny;
  if (nz < 0)
    nz += Nz;
  if (nz >= Nz)
    nz -= Nz;
  return ((nx * Ny + ny) * Nz + nz) * N_orbital + m;
}

// a full shuffle of [0, max_value), one draw per position; keeps the prefix
// (src/model_lattice.cu:26-43, src/charge.cu:46-63)
void Model::random_permutation_prefix(int max_value, int total_number, int* out)
{
  std::vector<int> p((size_t)max_value);
  for (int i = 0; i < max_value; ++i)
    p[i] = i;
  for (int i = 0; i < max_value; ++i) {
    int j = lsqt_host::uniform_int_0_intmax(generator) % (max_value - i) + i;
    std::swap(p[i], p[j]);
  }
  for (int i = 0; i < total_number; ++i)
    out[i] = p[i];
}

void Model::initialize_lattice_model()
{
  std::string filename = input_dir + "/lattice.in";
  print_started_reading(filename);
  std::ifstream input(filename);
  if (!input.is_open())
    throw InputError("Could not open " + filename);

  int N_orbital, transport_direction, N_cell[3];
  real lattice_constant[3], box_length[3];
  int pbc[3];
  const char* axis[3] = {"x", "y", "z"};
  input >> N_cell[0] >> N_cell[1] >> N_cell[2];
  for (int d = 0; d < 3; ++d) {
    std::cout << "- Number of cells in the " << axis[d] << " direction = " << N_cell[d]
              << std::endl;
    if (N_cell[d] <= 0)
      throw InputError(std::string("Error: Number of cells in the ") + axis[d] +
                       " direction should > 0");
  }
  input >> pbc[0] >> pbc[1] >> pbc[2] >> transport_direction;
  for (int d = 0; d < 3; ++d) {
    if (pbc[d] == 1)
      std::cout << "- " << axis[d] << " direction has periodic boundary" << std::endl;
    else if (pbc[d] == 0)
      std::cout << "- " << axis[d] << " direction has open boundary" << std::endl;
    else
      throw InputError(std::string("Error: ") + axis[d] + " direction has wrong boundary");
  }
  if (transport_direction < 0 || transport_direction > 2)
    throw InputError("Error: wrong transport direction");
  std::cout << "- transport in " << axis[transport_direction] << " direction" << std::endl;
  if (pbc[transport_direction] != 1)
    throw InputError("Error: transport direction must be periodic");
  input >> lattice_constant[0] >> lattice_constant[1] >> lattice_constant[2];
  for (int d = 0; d < 3; ++d) {
    std::cout << "- lattice constant in " << axis[d] << " direction = " << lattice_constant[d]
              << std::endl;
    if (lattice_constant[d] <= 0)
      throw InputError(std::string("Error: lattice constant in ") + axis[d] +
                       " direction < 0");
  }
  for (int d = 0; d < 3; ++d)
    box_length[d] = lattice_constant[d] * N_cell[d];
  volume = box_length[0] * box_length[1] * box_length[2];

  input >> N_orbital >> max_neighbor;
  std::cout << "- number of orbitals per cell = " << N_orbital << std::endl;
  if (N_orbital <= 0)
    throw InputError("Error: number of orbitals per cell should > 0");
  std::cout << "- maximum number of hoppings per orbital = " << max_neighbor << std::endl;
  if (max_neighbor <= 0)
    throw InputError("Error: maximum number of hoppings per orbital should > 0");
  long long total = (long long)N_orbital * N_cell[0] * N_cell[1] * N_cell[2];
  if (total * max_neighbor > INT_MAX)
    throw InputError("Error: system too large for 32-bit pair indexing");
  number_of_atoms = (int)total;
  std::cout << "- total number of orbitals = " << number_of_atoms << std::endl;
  number_of_pairs = number_of_atoms * max_neighbor;
  potential = new_zeroed<real>((size_t)number_of_atoms);

  std::unique_ptr<LatticeSpec> L(new LatticeSpec());
  for (int d = 0; d < 3; ++d) {
    L->N_cell[d] = N_cell[d];
    L->pbc[d] = pbc[d];
    L->lattice_constant[d] = lattice_constant[d];
  }
  L->transport_direction = transport_direction;
  L->N_orbital = N_orbital;
  L->max_neighbor = max_neighbor;
  L->x_cell.assign((size_t)N_orbital, 0.0);
  std::vector<real> y_cell(N_orbital), z_cell(N_orbital);
  const int per_cell = N_orbital * max_neighbor;
  L->hop_cell.assign((size_t)per_cell * 3, 0);
  L->hop_orbital.assign((size_t)per_cell, 0);
  L->hop_re.assign((size_t)per_cell, 0.0);
  L->hop_im.assign((size_t)per_cell, 0.0);
  std::cout << std::endl << "\torbital\tx\ty\tz" << std::endl;
  for (int n = 0; n < N_orbital; ++n) {
    input >> L->x_cell[n] >> y_cell[n] >> z_cell[n];
    std::cout << "\t" << n << "\t" << L->x_cell[n] << "\t" << y_cell[n] << "\t" << z_cell[n]
              << std::endl;
  }
  L->number_of_hoppings.assign((size_t)N_orbital, 0);
  for (int m = 0; m < N_orbital; m++) {
    int& count = L->number_of_hoppings[m];
    input >> count;
    std::cout << std::endl
              << "- number of hoppings for orbital " << m << " = " << count << std::endl;
    if (count < 0 || count > max_neighbor)
      throw InputError("Error: orbital " + std::to_string(m) +
                       " has more hoppings than the maximum");
    for (int n = 0; n < count; ++n) {
      int q = m * max_neighbor + n;
      int* c = &L->hop_cell[3 * (size_t)q];
      input >> c[0] >> c[1] >> c[2] >> L->hop_orbital[q] >> L->hop_re[q] >> L->hop_im[q];
      std::cout << "\tH(0,0,0," << m << "; " << c[0] << "," << c[1] << "," << c[2] << ","
                << L->hop_orbital[q] << ") = " << L->hop_re[q] << " + i " << L->hop_im[q]
                << std::endl;
      if (L->hop_orbital[q] < 0 || L->hop_orbital[q] >= N_orbital)
        throw InputError("Error: hopping to a non-existent orbital");
    }
  }

  // impurity positions, indexed by the pristine lattice
  std::vector<real> x, y, z;
  if (has_charge) {
    x.resize((size_t)number_of_atoms);
    y.resize((size_t)number_of_atoms);
    z.resize((size_t)number_of_atoms);
    for (int nx1 = 0; nx1 < N_cell[0]; ++nx1)
      for (int ny1 = 0; ny1 < N_cell[1]; ++ny1)
        for (int nz1 = 0; nz1 < N_cell[2]; ++nz1)
          for (int m = 0; m < N_orbital; ++m) {
            int n1 = find_index(nx1, ny1, nz1, N_cell[0], N_cell[1], N_cell[2], m, N_orbital);
            x[n1] = L->x_cell[m] + lattice_constant[0] * nx1;
            y[n1] = y_cell[m] + lattice_constant[1] * ny1;
            z[n1] = z_cell[m] + lattice_constant[2] * nz1;
          }
  }

  if (!defer_lattice) {
    expand_lattice(*L);
  } else {
    // the condition under which the expansion finds a neighbour index out
    // of range: a hopping that some site keeps (every open direction's
    // offset below the cell count) reaching further than one period of a
    // periodic direction (one wrap, src/model_lattice.cu:148-163)
    for (int m = 0; m < N_orbital; ++m)
      for (int i = 0; i < L->number_of_hoppings[m]; ++i) {
        const int* c = &L->hop_cell[3 * (size_t)(m * max_neighbor + i)];
        bool kept = true, beyond = false;
        for (int d = 0; d < 3; ++d) {
          if (!pbc[d] && std::abs(c[d]) >= N_cell[d])
            kept = false;
          if (pbc[d] && std::abs(c[d]) > N_cell[d])
            beyond = true;
        }
        if (kept && beyond)
          throw InputError("Error: a hopping leaves the lattice by more than one cell");
      }
  }

  if (has_vacancy_disorder) {
    std::vector<int> new_index;
    draw_vacancies(new_index);
    if (!defer_lattice) {
      remove_vacancies(new_index);
    } else {
      L->site_index = new_index;
      number_of_atoms -= number_of_vacancies;
      number_of_pairs = number_of_atoms * max_neighbor;
    }
  }
  if (has_anderson)
    add_anderson_disorder();
  if (has_charge)
    add_charged_impurities(pbc, box_length, x, y, z);
  if (defer_lattice)
    spec = std::move(L);
  print_finished_reading(filename);
}

// the neighbour arrays of the pristine lattice (src/model_lattice.cu:342-386):
// hoppings in file order, those leaving an open boundary skipped
void Model::expand_lattice(const LatticeSpec& L)
{
  const int* N_cell = L.N_cell;
  const int N_orbital = L.N_orbital, K = L.max_neighbor;
  const int N0 = N_orbital * N_cell[0] * N_cell[1] * N_cell[2];
  number_of_atoms = N0;
  number_of_pairs = N0 * K;
  delete[] neighbor_number;
  delete[] neighbor_list;
  delete[] hopping_real;
  delete[] hopping_imag;
  delete[] xx;
  neighbor_number = new_zeroed<int>((size_t)N0);
  neighbor_list = new_zeroed<int>((size_t)number_of_pairs);
  hopping_real = new_zeroed<real>((size_t)number_of_pairs);
  hopping_imag = new_zeroed<real>((size_t)number_of_pairs);
  xx = new_zeroed<real>((size_t)number_of_pairs);
  const int t = L.transport_direction;
  for (int nx1 = 0; nx1 < N_cell[0]; ++nx1)
    for (int ny1 = 0; ny1 < N_cell[1]; ++ny1)
      for (int nz1 = 0; nz1 < N_cell[2]; ++nz1)
        for (int m = 0; m < N_orbital; ++m) {
          int n1 = find_index(nx1, ny1, nz1, N_cell[0], N_cell[1], N_cell[2], m, N_orbital);
          int count = 0;
          for (int i = 0; i < L.number_of_hoppings[m]; ++i) {
            size_t slot = (size_t)n1 * K + count;
            int k = m * K + i;
            const int* c = &L.hop_cell[3 * (size_t)k];
            int nx2 = c[0] + nx1, ny2 = c[1] + ny1, nz2 = c[2] + nz1;
            bool skip_x = !L.pbc[0] && (nx2 < 0 || nx2 >= N_cell[0]);
            bool skip_y = !L.pbc[1] && (ny2 < 0 || ny2 >= N_cell[1]);
            bool skip_z = !L.pbc[2] && (nz2 < 0 || nz2 >= N_cell[2]);
            if (skip_x || skip_y || skip_z)
              continue;
            neighbor_list[slot] = find_index(nx2, ny2, nz2, N_cell[0], N_cell[1], N_cell[2],
                                             L.hop_orbital[k], N_orbital);
            real x12 = L.lattice_constant[t] * c[t];
            x12 += L.x_cell[L.hop_orbital[k]] - L.x_cell[m];
            xx[slot] = x12;
            hopping_real[slot] = L.hop_re[k];
            hopping_imag[slot] = L.hop_im[k];
            ++count;
          }
          neighbor_number[n1] = count;
        }
  for (size_t q = 0; q < (size_t)number_of_pairs; ++q)
    if (neighbor_list[q] < 0 || neighbor_list[q] >= number_of_atoms)
      throw InputError("Error: a hopping leaves the lattice by more than one cell");
}

void Model::expand_lattice_arrays()
{
  if (!spec)
    return;
  std::unique_ptr<LatticeSpec> L = std::move(spec);
  const int N1 = number_of_atoms;
  expand_lattice(*L);
  if (!L->site_index.empty())
    remove_vacancies(L->site_index);
  if (number_of_atoms != N1)
    throw InputError("Error: the lattice and its vacancies do not match the model");
}

// src/model_lattice.cu:26-68: the vacancies are the first
// number_of_vacancies entries of a full shuffle of the pristine sites;
// new_index[n] = index of pristine site n after their removal, -1 for a
// vacancy
void Model::draw_vacancies(std::vector<int>& new_index)
{
  const int N0 = number_of_atoms;
  if (number_of_vacancies > N0)
    throw InputError("Error: more vacancies than sites");
  std::vector<int> vacancy((size_t)number_of_vacancies);
  random_permutation_prefix(N0, number_of_vacancies, vacancy.data());
  new_index.assign((size_t)N0, 0);
  for (int v : vacancy)
    new_index[v] = -1;
  for (int n = 0, c = 0; n < N0; ++n)
    if (new_index[n] == 0)
      new_index[n] = c++;
}

// src/model_lattice.cu:70-146: the arrays without the vacancies; neighbours
// that are vacancies are dropped and the later slots move up
void Model::remove_vacancies(const std::vector<int>& new_index)
{
  const int N0 = number_of_atoms, K = max_neighbor;
  int N1 = 0;
  for (int v : new_index)
    N1 += v >= 0;
  const size_t P1 = (size_t)N1 * K;
  int* nn = new_zeroed<int>((size_t)N1);
  int* nl = new_zeroed<int>(P1);
  real* hr = new_zeroed<real>(P1);
  real* hi = new_zeroed<real>(P1);
  real* x = new_zeroed<real>(P1);
  int count_atom = 0;
  for (int n = 0; n < N0; ++n) {
    if (new_index[n] < 0)
      continue;
    int count_neighbor = 0;
    for (int m = 0; m < neighbor_number[n]; ++m) {
      size_t old = (size_t)n * K + m;
      int k = neighbor_list[old];
      if (new_index[k] >= 0) {
        size_t nw = (size_t)count_atom * K + count_neighbor;
        nl[nw] = new_index[k];
        hr[nw] = hopping_real[old];
        hi[nw] = hopping_imag[old];
        x[nw] = xx[old];
        ++count_neighbor;
      }
    }
    nn[count_atom++] = count_neighbor;
  }
  delete[] neighbor_number;
  delete[] neighbor_list;
  delete[] hopping_real;
  delete[] hopping_imag;
  delete[] xx;
  neighbor_number = nn;
  neighbor_list = nl;
  hopping_real = hr;
  hopping_imag = hi;
  xx = x;
  number_of_atoms = N1;
  number_of_pairs = (int)P1;
}

// src/anderson.cu:22-29
void Model::add_anderson_disorder()
{
  real W2 = anderson_strength * 0.5;
  for (int n = 0; n < number_of_atoms; ++n)
    potential[n] = lsqt_host::uniform_real(generator, -W2, W2);
}

// src/charge.cu:25-231
void Model::add_charged_impurities(const int pbc[3], const real box_length[3],
                                   const std::vector<real>& x, const std::vector<real>& y,
                                   const std::vector<real>& z)
{
  const int N = number_of_atoms, Ni = charge_number;
  const real rc = 5.0 * charge_range;  // exp(-25/2) ~ 1e-6
  const real rc2 = rc * rc;
  if (Ni < 0 || Ni > N)
    throw InputError("Error: invalid number of charged impurities");
  std::vector<int> impurity((size_t)Ni);
  random_permutation_prefix(N, Ni, impurity.data());
  std::vector<real> strength((size_t)Ni);
  real W2 = charge_strength * 0.5;
  for (int i = 0; i < Ni; ++i)
    strength[i] = lsqt_host::uniform_real(generator, -W2, W2);

  int Nc[3];
  for (int d = 0; d < 3; ++d)
    Nc[d] = pbc[d] ? (int)floor(box_length[d] / rc) : 1;
  if (Nc[0] <= 0 || Nc[1] <= 0 || Nc[2] <= 0)
    throw InputError("Error: charged impurity range 5*xi exceeds the box length");
  const int Nxyz = Nc[0] * Nc[1] * Nc[2];
  auto cell_of = [&](real px, real py, real pz, int& cx, int& cy, int& cz) {
    cx = (int)floor(px / rc);
    cy = (int)floor(py / rc);
    cz = (int)floor(pz / rc);
    while (cx < 0)
      cx += Nc[0];
    while (cx >= Nc[0])
      cx -= Nc[0];
    while (cy < 0)
      cy += Nc[1];
    while (cy >= Nc[1])
      cy -= Nc[1];
    while (cz < 0)
      cz += Nc[2];
    while (cz >= Nc[2])
      cz -= Nc[2];
    return cx + Nc[0] * cy + Nc[0] * Nc[1] * cz;
  };
  std::vector<int> cell_count((size_t)Nxyz, 0), cell_start((size_t)Nxyz, 0), contents((size_t)N);
  int cx, cy, cz;
  for (int i = 0; i < N; ++i)
    ++cell_count[cell_of(x[i], y[i], z[i], cx, cy, cz)];
  for (int i = 1; i < Nxyz; ++i)
    cell_start[i] = cell_count[i - 1] + cell_start[i - 1];
  std::fill(cell_count.begin(), cell_count.end(), 0);
  for (int i = 0; i < N; ++i) {
    int c = cell_of(x[i], y[i], z[i], cx, cy, cz);
    contents[cell_start[c] + cell_count[c]++] = i;
  }

  const real xi_factor = -0.5 / (charge_range * charge_range);
  real half[3];
  for (int d = 0; d < 3; ++d)
    half[d] = box_length[d] * 0.5;
  for (int n = 0; n < N; ++n)
    potential[n] = 0.0;
  for (int ni = 0; ni < Ni; ++ni) {
    int n1 = impurity[ni];
    real x1 = x[n1], y1 = y[n1], z1 = z[n1];
    int c1 = cell_of(x1, y1, z1, cx, cy, cz);
    int klim = pbc[2] ? 1 : 0, jlim = pbc[1] ? 1 : 0, ilim = pbc[0] ? 1 : 0;
    for (int k = -klim; k < klim + 1; ++k)
      for (int j = -jlim; j < jlim + 1; ++j)
        for (int i = -ilim; i < ilim + 1; ++i) {
          // src/charge.cu:215-231
          int nb = c1 + k * Nc[0] * Nc[1] + j * Nc[0] + i;
          if (cx + i < 0)
            nb += Nc[0];
          if (cx + i >= Nc[0])
            nb -= Nc[0];
          if (cy + j < 0)
            nb += Nc[1] * Nc[0];
          if (cy + j >= Nc[1])
            nb -= Nc[1] * Nc[0];
          if (cz + k < 0)
            nb += Nxyz;
          if (cz + k >= Nc[2])
            nb -= Nxyz;
          for (int q = 0; q < cell_count[nb]; ++q) {
            int n2 = contents[cell_start[nb] + q];
            real r12[3] = {x[n2] - x1, y[n2] - y1, z[n2] - z1};
            real d12_square = 0.0;
            for (int d = 0; d < 3; ++d) {
              r12[d] = fabs(r12[d]);
              if (pbc[d] == 1 && r12[d] > half[d])
                r12[d] = box_length[d] - r12[d];
              d12_square += r12[d] * r12[d];
            }
            if (d12_square > rc2)
              continue;
            potential[n2] += strength[ni] * exp(d12_square * xi_factor);
          }
        }
  }
}
