// lsqt.cpp -- the driver of one input directory (src/lsqt.cu:135-152):
// per random vector DOS, VAC0, VAC, MSD, spin; then the LDOS orbitals.
// Phase timers are wall-clock (the reference times CPU time with clock()).
#include "lsqt.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <vector>

#include "hamiltonian.h"
#include "lsqt_hip.h"
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

namespace lsqt_host {
lsqt_lattice abi_lattice(const LatticeSpec& L);

// Lattice models are expanded on the GPU unless LSQT_DEVICE_LATTICE=0 (the
// same Hamiltonian either way)
bool lattice_on_device()
{
  const char* e = getenv("LSQT_DEVICE_LATTICE");
  return !(e && e[0] == '0');
}
}  // namespace lsqt_host

void lsqt(std::string input_directory)
{
  lsqt_host::set_path(lsqt_host::path_from_environment());
  Model model(input_directory, lsqt_host::lattice_on_device());
  Hamiltonian H(model);
  Vector random_state(model.number_of_atoms);
  const bool block = lsqt_host::path() == lsqt_host::Path::block ||
                     lsqt_host::path() == lsqt_host::Path::batch;
  if (block && model.number_of_random_vectors > 0) {
    // the random vectors of the directory in blocks of up to 256 at once;
    // the RNG draws are the same as vector by vector (nothing else draws in
    // between), and each .out file receives its rows in vector order
    const int total = model.number_of_random_vectors;
    const size_t n = (size_t)model.number_of_atoms;
    for (int first = 0; first < total; first += 256) {
      const int used = std::min(256, total - first);
      int R = 1;
      while (R < used)
        R *= 2;
      std::vector<real> re(n * R, 0.0), im(n * R, 0.0);
      for (int i = 0; i < used; ++i)
        model.initialize_state_host(re.data() + i * n, im.data() + i * n, -1);
      banner("Started  simulation with random vectors " + std::to_string(first) + " to " +
             std::to_string(first + used - 1) + " (block path)");
      std::vector<lsqt_host::BlockMember> members{{&model, 0, used}};
      lsqt_host::find_block_quantities(members, H.handle(), R, re, im);
      banner_end("Finished simulation with random vectors " + std::to_string(first) + " to " +
                 std::to_string(first + used - 1));
    }
  } else {
    lsqt_vectors(model, H, random_state);
  }
  if (model.calculate_ldos) {
    banner("Started LDOS calculation");
    timed("LDOS", [&] {
      if (block) {
        lsqt_host::find_block_ldos(model, H.handle());
        return;
      }
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

namespace {

int pow2_at_least(int v)
{
  int R = 1;
  while (R < v)
    R *= 2;
  return R;
}

// settings that systems of one batch must share
std::vector<double> batch_key(const Model& m)
{
  std::vector<double> k = {(double)m.number_of_moments, m.energy_max,
                           (double)m.number_of_steps_correlation,
                           (double)m.calculate_vac0, (double)m.calculate_vac,
                           (double)m.calculate_msd,
                           (double)pow2_at_least(m.number_of_random_vectors)};
  for (int t = 0; t < m.number_of_steps_correlation; ++t)
    k.push_back(m.time_step[t]);
  return k;
}

// one segmented Hamiltonian for the members' models (arrays released after):
// expanded on the GPU when every member is a lattice model that kept its
// description, else from the host arrays (expanded for those that kept it)
lsqt_ham* segmented_hamiltonian(std::vector<lsqt_host::BlockMember>& members, int n_total)
{
  bool all_lattices = true;
  for (auto& mb : members)
    all_lattices = all_lattices && mb.model->lattice() != nullptr;
  if (all_lattices) {
    std::vector<lsqt_lattice> lats;
    std::vector<const int*> index;
    std::vector<real> pot((size_t)n_total);
    bool any_index = false;
    for (auto& mb : members) {
      const LatticeSpec& L = *mb.model->lattice();
      lats.push_back(lsqt_host::abi_lattice(L));
      index.push_back(L.site_index.empty() ? nullptr : L.site_index.data());
      any_index = any_index || !L.site_index.empty();
      std::copy(mb.model->potential, mb.model->potential + mb.model->number_of_atoms,
                pot.begin() + mb.start);
    }
    lsqt_ham* h = nullptr;
    LSQT_CHECK(lsqt_ham_create_lattices(lsqt_host::default_context(), (int)members.size(),
                                        lats.data(), any_index ? index.data() : nullptr,
                                        pot.data(), members[0].model->energy_max, &h));
    for (auto& mb : members)
      mb.model->release_hamiltonian_arrays();
    return h;
  }
  for (auto& mb : members)
    mb.model->expand_lattice_arrays();
  int K = 0;
  for (auto& mb : members)
    K = std::max(K, mb.model->max_neighbor);
  std::vector<int> nn((size_t)n_total), nl((size_t)n_total * K, 0), starts;
  std::vector<real> pot((size_t)n_total), hr((size_t)n_total * K, 0.0),
    hi((size_t)n_total * K, 0.0), xx((size_t)n_total * K, 0.0);
  for (auto& mb : members) {
    Model& m = *mb.model;
    starts.push_back(mb.start);
    for (int i = 0; i < m.number_of_atoms; ++i) {
      const size_t g = (size_t)(mb.start + i);
      nn[g] = m.neighbor_number[i];
      pot[g] = m.potential[i];
      for (int q = 0; q < m.neighbor_number[i]; ++q) {
        const size_t src = (size_t)i * m.max_neighbor + q, dst = g * K + q;
        nl[dst] = m.neighbor_list[src] + mb.start;
        hr[dst] = m.hopping_real[src];
        hi[dst] = m.hopping_imag[src];
        xx[dst] = m.xx[src];
      }
    }
  }
  starts.push_back(n_total);
  lsqt_ham* h = nullptr;
  LSQT_CHECK(lsqt_ham_create_segmented(lsqt_host::default_context(), (int)members.size(),
                                       starts.data(), n_total, K, nn.data(), nl.data(),
                                       pot.data(), hr.data(), hi.data(), xx.data(),
                                       members[0].model->energy_max, &h));
  for (auto& mb : members)
    mb.model->release_hamiltonian_arrays();
  return h;
}

void run_batch(std::vector<lsqt_host::BlockMember>& members, int n_total, int R)
{
  banner("Started  simulation of " + std::to_string(members.size()) +
         " systems together (batch path, " + std::to_string(n_total) + " sites)");
  lsqt_ham* h = segmented_hamiltonian(members, n_total);
  // every system's random vectors, drawn in its own RNG order
  std::vector<real> re((size_t)n_total * R, 0.0), im((size_t)n_total * R, 0.0);
  for (auto& mb : members)
    for (int r = 0; r < mb.used; ++r)
      mb.model->initialize_state_host(re.data() + (size_t)r * n_total + mb.start,
                                      im.data() + (size_t)r * n_total + mb.start, -1);
  lsqt_host::find_block_quantities(members, h, R, re, im);
  lsqt_ham_destroy(h);
  banner_end("Finished simulation of " + std::to_string(members.size()) + " systems");
}

}  // namespace

void lsqt_batch(const std::vector<std::string>& directories)
{
  lsqt_host::set_path(lsqt_host::Path::batch);
  std::vector<std::unique_ptr<Model>> models;
  std::vector<std::string> alone;
  for (const std::string& d : directories) {
    std::unique_ptr<Model> m(new Model(d, lsqt_host::lattice_on_device()));
    if (m->calculate_spin || m->calculate_ldos || m->number_of_random_vectors <= 0 ||
        m->number_of_random_vectors > 256)
      alone.push_back(d);  // run by itself below
    else
      models.push_back(std::move(m));
  }
  // systems with the same settings, in driver-file order within each group
  std::vector<std::vector<double>> keys;
  std::vector<std::vector<Model*>> groups;
  for (auto& m : models) {
    auto k = batch_key(*m);
    size_t g = 0;
    while (g < keys.size() && keys[g] != k)
      ++g;
    if (g == keys.size()) {
      keys.push_back(k);
      groups.emplace_back();
    }
    groups[g].push_back(m.get());
  }
  for (auto& group : groups) {
    const int R = pow2_at_least(group[0]->number_of_random_vectors);
    std::vector<lsqt_host::BlockMember> members;
    long long n_total = 0;
    for (Model* m : group) {
      // one launch covers at most 2^31 values of a block
      if (!members.empty() && (n_total + m->number_of_atoms) * R > INT_MAX) {
        run_batch(members, (int)n_total, R);
        members.clear();
        n_total = 0;
      }
      members.push_back({m, (int)n_total, m->number_of_random_vectors});
      n_total += m->number_of_atoms;
    }
    if (!members.empty())
      run_batch(members, (int)n_total, R);
  }
  models.clear();
  for (const std::string& d : alone)
    lsqt(d);
}
