// host_capi.cpp -- the C entry points of include/lsqt_host.h.
#include <chrono>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "general_io.h"
#include "hamiltonian.h"
#include "lsqt_hip.h"
#include "lsqt.h"
#include "lsqt_host.h"
#include "model.h"
#include "postprocess.h"
#include "sigma.h"
#include "vector.h"

namespace lsqt_host {
lsqt_lattice abi_lattice(const LatticeSpec& L);
}

struct lsqt_host_model {
  Model* model;
  Hamiltonian* ham = nullptr;  // built on first use (releases the model's arrays)
};

namespace {

int report(const std::exception& e, char* err, int err_len)
{
  if (err && err_len > 0) {
    std::strncpy(err, e.what(), (size_t)err_len - 1);
    err[err_len - 1] = 0;
  }
  return 1;
}

void print_welcome()
{
  std::cout << std::endl;
  std::cout << "***************************************************************\n";
  std::cout << "*                  Welcome to use LSQT                        *\n";
  std::cout << "*          (Linear Scaling Quantum Transport)                 *\n";
  std::cout << "*        (MI355X build: HIP kernels for gfx950)               *\n";
  std::cout << "***************************************************************\n";
  std::cout << std::endl;
}

}  // namespace

namespace {

// `lsqt_gpu --sigma <driver-file>`: the conductivity of every listed
// directory from its existing output rows
int postprocess_all(const char* driver)
{
  std::ifstream input(driver);
  if (!input.is_open()) {
    std::cout << "Failed to open " << driver << std::endl;
    return 1;
  }
  std::string directory;
  while (std::getline(input, directory)) {
    if (directory == "")
      continue;
    try {
      lsqt_host::Conductivity c = lsqt_host::postprocess_directory(directory);
      std::cout << "Conductivity for " << directory << ": " << c.Nt << " time steps x " << c.Ne
                << " energies" << (c.sigma_vac.empty() ? "" : ", sigma_vac.out")
                << (c.sigma_msd.empty() ? "" : ", sigma_msd.out") << std::endl;
    } catch (const lsqt_host::InputError& e) {
      std::cout << e.what() << std::endl;
      return 1;
    }
  }
  return 0;
}

}  // namespace

extern "C" int lsqt_host_main(int argc, char** argv)
{
  if (argc == 3 && std::strcmp(argv[1], "--sigma") == 0)
    return postprocess_all(argv[2]);
  print_welcome();
  if (argc != 2) {
    std::cout << "Usage: src/gpuqt input.txt" << std::endl;
    return 1;
  }
  if (const char* digits = getenv("LSQT_OUTPUT_DIGITS"))
    lsqt_host::set_output_precision(atoi(digits));
  std::ifstream input(argv[1]);
  if (!input.is_open()) {
    std::cout << "Failed to open " << argv[1] << std::endl;
    return 1;
  }
  std::string directory;
  lsqt_host::Path path;
  try {
    path = lsqt_host::path_from_environment();
  } catch (const lsqt_host::InputError& e) {
    std::cout << e.what() << std::endl;
    return 1;
  }
  if (path == lsqt_host::Path::batch) {
    std::vector<std::string> directories;
    while (std::getline(input, directory))
      if (directory != "")
        directories.push_back(directory);
    std::cout << std::endl;
    std::cout << "===============================================================\n";
    std::cout << "Run LSQT simulation for " << directories.size()
              << " directories together (batch path)" << std::endl;
    std::cout << "===============================================================\n";
    auto t0 = std::chrono::steady_clock::now();
    try {
      lsqt_batch(directories);
    } catch (const lsqt_host::InputError& e) {
      std::cout << e.what() << std::endl;
      return 1;
    }
    double used =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << std::endl;
    std::cout << "===============================================================\n";
    std::cout << "Total time used for " << directories.size() << " directories = " << used
              << " s" << std::endl;
    std::cout << "===============================================================\n";
    lsqt_host::release_default_context();
    return 0;
  }
  while (std::getline(input, directory)) {
    if (directory == "")
      continue;
    std::cout << std::endl;
    std::cout << "===============================================================\n";
    std::cout << "Run LSQT simulation for " << directory << std::endl;
    std::cout << "===============================================================\n";
    auto t0 = std::chrono::steady_clock::now();
    try {
      lsqt(directory);
    } catch (const lsqt_host::InputError& e) {
      std::cout << e.what() << std::endl;
      return 1;
    }
    double used =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << std::endl;
    std::cout << "===============================================================\n";
    std::cout << "Total time used for " << directory << " = " << used << " s" << std::endl;
    std::cout << "===============================================================\n";
  }
  lsqt_host::release_default_context();
  return 0;
}

extern "C" int lsqt_host_run(const char* input_dir, char* err, int err_len)
{
  try {
    lsqt(input_dir);
  } catch (const lsqt_host::InputError& e) {
    return report(e, err, err_len);
  }
  return 0;
}

extern "C" int lsqt_host_damped_sum(int M, double* mu_re, double* mu_im, int n_energy,
                                    const double* energy, double energy_max, double volume,
                                    double* out)
{
  if (M < 2 || !mu_re || (n_energy > 0 && (!energy || !out)))
    return 1;
  lsqt_host::damped_chebyshev_sum(M, mu_re, mu_im, n_energy, energy, energy_max, volume, out);
  return 0;
}

extern "C" int lsqt_host_conductivity(int n_energy, int n_time, const double* time_step,
                                      const double* dos, int dos_rows, const double* vac,
                                      int vac_rows, const double* msd, int msd_rows,
                                      double* sigma_vac, double* sigma_msd, double* vac_true,
                                      double* msd_true, char* err, int err_len)
{
  try {
    lsqt_host::Conductivity c = lsqt_host::conductivity(n_energy, n_time, time_step, dos,
                                                        dos_rows, vac, vac_rows, msd, msd_rows);
    auto put = [](const std::vector<double>& v, double* out) {
      if (out && !v.empty())
        std::memcpy(out, v.data(), sizeof(double) * v.size());
    };
    put(c.sigma_vac, sigma_vac);
    put(c.sigma_msd, sigma_msd);
    put(c.vac, vac_true);
    put(c.msd, msd_true);
  } catch (const lsqt_host::InputError& e) {
    return report(e, err, err_len);
  }
  return 0;
}

extern "C" int lsqt_host_write_general(const char* dir, int n, int max_neighbor,
                                       const int* neighbor_number, const int* neighbor_list,
                                       const double* x, double box, double volume,
                                       const double* potential, const double* hopping_real,
                                       const double* hopping_imag, char* err, int err_len)
{
  try {
    if (!dir || n <= 0 || max_neighbor <= 0 || !neighbor_number || !neighbor_list || !x ||
        !hopping_real)
      throw lsqt_host::InputError("Error: invalid arguments for the general-model files");
    for (int i = 0; i < n; ++i)
      if (neighbor_number[i] < 0 || neighbor_number[i] > max_neighbor)
        throw lsqt_host::InputError("Error: site " + std::to_string(i) +
                                    " has more neighbors than " +
                                    std::to_string(max_neighbor));
    lsqt_host::write_general(dir, n, max_neighbor, neighbor_number, neighbor_list, x, box,
                             volume, potential, hopping_real, hopping_imag);
  } catch (const lsqt_host::InputError& e) {
    return report(e, err, err_len);
  }
  return 0;
}

extern "C" int lsqt_host_postprocess(const char* input_dir, char* err, int err_len)
{
  try {
    lsqt_host::postprocess_directory(input_dir);
  } catch (const lsqt_host::InputError& e) {
    return report(e, err, err_len);
  }
  return 0;
}

extern "C" void lsqt_host_set_output_precision(int digits)
{
  lsqt_host::set_output_precision(digits);
}

extern "C" int lsqt_host_model_load(const char* input_dir, lsqt_host_model** out, char* err,
                                    int err_len)
{
  return lsqt_host_model_load_ex(input_dir, 0, out, err, err_len);
}

extern "C" int lsqt_host_model_load_ex(const char* input_dir, int flags, lsqt_host_model** out,
                                       char* err, int err_len)
{
  if (!out)
    return 1;
  *out = nullptr;
  try {
    Model* m = new Model(input_dir, (flags & LSQT_HOST_LATTICE_ON_DEVICE) != 0);
    *out = new lsqt_host_model{m};
  } catch (const lsqt_host::InputError& e) {
    return report(e, err, err_len);
  }
  return 0;
}

extern "C" void lsqt_host_model_free(lsqt_host_model* m)
{
  if (m) {
    delete m->ham;
    delete m->model;
    delete m;
  }
}

extern "C" int lsqt_host_model_info(const lsqt_host_model* h, int* ints, double* reals)
{
  if (!h || !ints || !reals)
    return 1;
  const Model& m = *h->model;
  ints[0] = m.number_of_atoms;
  ints[1] = m.max_neighbor;
  ints[2] = m.number_of_moments;
  ints[3] = m.number_of_random_vectors;
  ints[4] = m.number_of_energy_points;
  ints[5] = m.number_of_steps_correlation;
  ints[6] = m.number_of_local_orbitals;
  ints[7] = (m.calculate_vac0 ? 1 : 0) | (m.calculate_vac ? 2 : 0) | (m.calculate_msd ? 4 : 0) |
            (m.calculate_spin ? 8 : 0) | (m.calculate_ldos ? 16 : 0);
  reals[0] = m.energy_max;
  reals[1] = m.volume;
  return 0;
}

extern "C" int lsqt_host_model_arrays(const lsqt_host_model* h, int* nn, int* nl, double* pot,
                                      double* hr, double* hi, double* xx)
{
  if (!h || !h->model->neighbor_number)
    return 1;
  const Model& m = *h->model;
  size_t N = (size_t)m.number_of_atoms, P = (size_t)m.number_of_pairs;
  if (nn)
    std::memcpy(nn, m.neighbor_number, N * sizeof(int));
  if (nl)
    std::memcpy(nl, m.neighbor_list, P * sizeof(int));
  if (pot)
    std::memcpy(pot, m.potential, N * sizeof(double));
  if (hr)
    std::memcpy(hr, m.hopping_real, P * sizeof(double));
  if (hi)
    std::memcpy(hi, m.hopping_imag, P * sizeof(double));
  if (xx)
    std::memcpy(xx, m.xx, P * sizeof(double));
  return 0;
}

extern "C" int lsqt_host_model_inputs(const lsqt_host_model* h, double* energy,
                                      double* time_step, int* local_orbitals)
{
  if (!h)
    return 1;
  const Model& m = *h->model;
  if (energy)
    std::memcpy(energy, m.energy, sizeof(double) * (size_t)m.number_of_energy_points);
  if (time_step && m.time_step)
    std::memcpy(time_step, m.time_step, sizeof(double) * (size_t)m.number_of_steps_correlation);
  if (local_orbitals && !m.local_orbitals.empty())
    std::memcpy(local_orbitals, m.local_orbitals.data(), sizeof(int) * m.local_orbitals.size());
  return 0;
}

extern "C" int lsqt_host_model_random_state(lsqt_host_model* h, int orbital, double* re,
                                            double* im)
{
  if (!h || !re || !im || orbital >= h->model->number_of_atoms)
    return 1;
  h->model->initialize_state_host(re, im, orbital);
  return 0;
}

namespace {
Hamiltonian& model_hamiltonian(lsqt_host_model* h)
{
  if (!h->ham)
    h->ham = new Hamiltonian(*h->model);
  return *h->ham;
}
}  // namespace

extern "C" int lsqt_host_model_evolve(lsqt_host_model* h, int kind, int direction,
                                      double time_step_scaled, int by_terms, double* re,
                                      double* im)
{
  if (!h || !re || !im || (kind != 0 && kind != 1) || (direction != 1 && direction != -1))
    return 1;
  Model& m = *h->model;
  Hamiltonian& H = model_hamiltonian(h);
  Vector state(m.number_of_atoms);
  state.copy_from_host(re, im);
  if (kind == 0) {
    if (by_terms)
      lsqt_host::evolve_by_terms(m, direction, time_step_scaled, H, state);
    else
      evolve(m, direction, time_step_scaled, H, state);
  } else {
    if (by_terms)
      lsqt_host::evolvex_by_terms(m, direction, time_step_scaled, H, state);
    else
      evolvex(m, direction, time_step_scaled, H, state);
  }
  state.copy_to_host(re, im);
  return 0;
}

extern "C" int lsqt_host_model_moments(lsqt_host_model* h, const double* l_re,
                                       const double* l_im, const double* r_re,
                                       const double* r_im, int by_terms, double* mu_re,
                                       double* mu_im)
{
  if (!h || !l_re || !l_im || !r_re || !r_im || !mu_re || !mu_im)
    return 1;
  Model& m = *h->model;
  if (m.number_of_moments < 2)
    return 1;
  Hamiltonian& H = model_hamiltonian(h);
  const int n = m.number_of_atoms, M = m.number_of_moments;
  Vector left(n), right(n), output(M);
  // copy_from_host takes non-const arrays, as the reference's does
  std::vector<real> buf(l_re, l_re + n), buf_im(l_im, l_im + n);
  left.copy_from_host(buf.data(), buf_im.data());
  buf.assign(r_re, r_re + n);
  buf_im.assign(r_im, r_im + n);
  right.copy_from_host(buf.data(), buf_im.data());
  if (by_terms)
    lsqt_host::find_moments_by_terms(m, H, left, right, output);
  else
    find_moments_chebyshev(m, H, left, right, output);
  output.copy_to_host(mu_re, mu_im);
  return 0;
}

extern "C" int lsqt_host_model_lattice(const lsqt_host_model* h, lsqt_lattice* lattice,
                                       const int** site_index, int* n_pristine)
{
  if (!h || !lattice || !site_index || !n_pristine || !h->model->lattice())
    return 1;
  const LatticeSpec& L = *h->model->lattice();
  *lattice = lsqt_host::abi_lattice(L);
  *site_index = L.site_index.empty() ? nullptr : L.site_index.data();
  *n_pristine = L.N_orbital * L.N_cell[0] * L.N_cell[1] * L.N_cell[2];
  return 0;
}

extern "C" int lsqt_host_model_potential(const lsqt_host_model* h, double* potential)
{
  if (!h || !potential || !h->model->potential)
    return 1;
  std::memcpy(potential, h->model->potential, sizeof(double) * (size_t)h->model->number_of_atoms);
  return 0;
}

extern "C" int lsqt_host_model_expand_lattice(lsqt_host_model* h)
{
  if (!h)
    return 1;
  try {
    h->model->expand_lattice_arrays();
  } catch (const lsqt_host::InputError&) {
    return 1;
  }
  return 0;
}
