// host_capi.cpp -- the C entry points of include/lsqt_host.h.
#include <chrono>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>

#include "lsqt.h"
#include "lsqt_host.h"
#include "model.h"
#include "sigma.h"
#include "vector.h"

struct lsqt_host_model {
  Model* model;
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

extern "C" int lsqt_host_main(int argc, char** argv)
{
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

extern "C" void lsqt_host_set_output_precision(int digits)
{
  lsqt_host::set_output_precision(digits);
}

extern "C" int lsqt_host_model_load(const char* input_dir, lsqt_host_model** out, char* err,
                                    int err_len)
{
  if (!out)
    return 1;
  *out = nullptr;
  try {
    Model* m = new Model(input_dir);
    *out = new lsqt_host_model{m};
  } catch (const lsqt_host::InputError& e) {
    return report(e, err, err_len);
  }
  return 0;
}

extern "C" void lsqt_host_model_free(lsqt_host_model* m)
{
  if (m) {
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
