// lsqt.cpp -- the driver of one input directory (src/lsqt.cu:135-152):
// per random vector DOS, VAC0, VAC, MSD, spin; then the LDOS orbitals.
// Phase timers are wall-clock (the reference times CPU time with clock()).
#include "lsqt.h"

#include <chrono>
#include <iostream>
#include <vector>

#include "hamiltonian.h"
#include "model.h"
#include "sigma.h"
#include "vector.h"

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0)
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void banner(const std::string& text)
{
  std::cout << std::endl;
  std::cout << "===========================================================";
  std::cout << std::endl;
  std::cout << text << std::endl;
  std::cout << std::endl;
}

void banner_end(const std::string& text)
{
  std::cout << std::endl;
  std::cout << text << std::endl;
  std::cout << "===========================================================";
  std::cout << std::endl << std::endl;
}

template <typename F>
void timed(const char* what, F&& f)
{
  auto t0 = std::chrono::steady_clock::now();
  f();
  std::cout << "- Time used for finding " << what << " = " << seconds_since(t0) << " s"
            << std::endl;
}

}  // namespace

// one random vector at a time (src/lsqt.cu:140-150)
void lsqt_vectors(Model& model, Hamiltonian& H, Vector& random_state);

void lsqt(std::string input_directory)
{
  lsqt_host::set_path(lsqt_host::path_from_environment());
  Model model(input_directory);
  Hamiltonian H(model);
  Vector random_state(model.number_of_atoms);
  if (lsqt_host::path() == lsqt_host::Path::block && model.number_of_random_vectors > 0) {
    // every random vector of the directory at once; the RNG draws are the
    // same as vector by vector (nothing else draws in between)
    const int used = model.number_of_random_vectors;
    int R = 1;
    while (R < used)
      R *= 2;
    if (R > 256)
      throw lsqt_host::InputError("Error: the block path takes at most 256 random vectors");
    const size_t n = (size_t)model.number_of_atoms;
    std::vector<real> re(n * R, 0.0), im(n * R, 0.0);
    for (int i = 0; i < used; ++i)
      model.initialize_state_host(re.data() + i * n, im.data() + i * n, -1);
    banner("Started  simulation with " + std::to_string(used) + " random vectors (block path)");
    lsqt_host::find_block_quantities(model, H, R, used, re, im);
    if (model.calculate_spin) {
      timed("spin polarization", [&] {
        for (int i = 0; i < used; ++i) {
          random_state.copy_from_host(re.data() + i * n, im.data() + i * n);
          find_spin_polarization(model, H, random_state);
        }
      });
    }
    banner_end("Finished simulation with " + std::to_string(used) + " random vectors");
  } else {
    lsqt_vectors(model, H, random_state);
  }
  if (model.calculate_ldos) {
    banner("Started LDOS calculation");
    timed("LDOS", [&] {
      for (int i = 0; i < model.number_of_local_orbitals; ++i) {
        model.initialize_state(random_state, model.local_orbitals[i]);
        find_dos(model, H, random_state, 1);
      }
    });
    banner_end("Finished LDOS calculation");
  }
}

void lsqt_vectors(Model& model, Hamiltonian& H, Vector& random_state)
{
  for (int i = 0; i < model.number_of_random_vectors; ++i) {
    banner("Started  simulation with random vector number " + std::to_string(i));
    model.initialize_state(random_state, -1);
    timed("DOS", [&] { find_dos(model, H, random_state, 0); });
    if (model.calculate_vac0)
      timed("VAC0", [&] { find_vac0(model, H, random_state); });
    if (model.calculate_vac)
      timed("VAC", [&] { find_vac(model, H, random_state); });
    if (model.calculate_msd)
      timed("MSD", [&] { find_msd(model, H, random_state); });
    if (model.calculate_spin)
      timed("spin polarization", [&] { find_spin_polarization(model, H, random_state); });
    banner_end("Finished simulation with random vector number " + std::to_string(i));
  }
}
