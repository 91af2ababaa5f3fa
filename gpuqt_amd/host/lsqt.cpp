// lsqt.cpp -- the driver of one input directory (src/lsqt.cu:135-152):
// per random vector DOS, VAC0, VAC, MSD, spin; then the LDOS orbitals.
// Phase timers are wall-clock (the reference times CPU time with clock()).
#include "lsqt.h"

#include <chrono>
#include <iostream>

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

void lsqt(std::string input_directory)
{
  Model model(input_directory);
  Hamiltonian H(model);
  Vector random_state(model.number_of_atoms);
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
