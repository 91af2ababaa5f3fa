// ref_harness.cpp -- test-only driver linked against the reference's own CPU
// objects (compiled from /root/reference/src by oracle/Makefile into
// oracle/_ref/).  It is never shipped and never linked by product code.
//
// The reference writes its results with 6 significant digits
// (src/sigma.cu:230-233).  This harness calls the reference's functions
// directly -- Model (src/model.h:27-104), Hamiltonian (src/hamiltonian.h),
// Vector (src/vector.h), and the external-linkage algorithm functions of
// src/sigma.cu (find_moments_chebyshev :33, apply_damping :76,
// perform_chebyshev_summation :89, evolve :116, evolvex :157) -- in the order
// the drivers find_dos/find_vac0/find_vac/find_msd/find_spin_polarization and
// lsqt() use them (src/sigma.cu:204-442, src/lsqt.cu:135-152), and prints
// every intermediate result as a C99 hex float, so the oracle and the HIP
// path can be checked bit for bit.
//
// Output: a sequence of records, each a header line "<name> <count>"
// followed by <count> values, one per line.
#include "hamiltonian.h"
#include "model.h"
#include "sigma.h"
#include "vector.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

void find_moments_chebyshev(Model&, Hamiltonian&, Vector&, Vector&, Vector&);
void apply_damping(Model&, real*, real*);
void perform_chebyshev_summation(Model&, real*, real*, real*);
void evolve(Model&, int, real, Hamiltonian&, Vector&);
void evolvex(Model&, int, real, Hamiltonian&, Vector&);

static FILE* out;

static void dump_d(const char* name, const double* v, int n)
{
  fprintf(out, "%s %d\n", name, n);
  for (int i = 0; i < n; ++i)
    fprintf(out, "%a\n", v[i]);
}

static void dump_i(const char* name, const int* v, int n)
{
  fprintf(out, "%s %d\n", name, n);
  for (int i = 0; i < n; ++i)
    fprintf(out, "%d\n", v[i]);
}

static void dump_vec(const std::string& name, Vector& v, int n)
{
  std::vector<double> re(n), im(n);
  v.copy_to_host(re.data(), im.data());
  dump_d((name + "_re").c_str(), re.data(), n);
  dump_d((name + "_im").c_str(), im.data(), n);
}

// one moment set: raw moments, then the damped and summed row
static void moment_set(Model& model, Hamiltonian& H, Vector& L, Vector& R, const std::string& tag)
{
  int M = model.number_of_moments;
  Vector ip(M);
  find_moments_chebyshev(model, H, L, R, ip);
  std::vector<double> re(M), im(M), row(model.number_of_energy_points);
  ip.copy_to_host(re.data(), im.data());
  dump_d((tag + "_mu_re").c_str(), re.data(), M);
  dump_d((tag + "_mu_im").c_str(), im.data(), M);
  apply_damping(model, re.data(), im.data());
  perform_chebyshev_summation(model, re.data(), im.data(), row.data());
  dump_d((tag + "_row").c_str(), row.data(), model.number_of_energy_points);
}

int main(int argc, char** argv)
{
  if (argc != 3) {
    fprintf(stderr, "usage: ref_harness <input_dir> <dump_file>\n");
    return 2;
  }
  std::string dir = argv[1];
  out = fopen(argv[2], "w");
  if (!out)
    return 2;
  Model model(dir);
  int N = model.number_of_atoms, K = model.max_neighbor;

  // model arrays, valid slots only (Hamiltonian's constructor frees them)
  {
    int meta[4] = {N, K, model.number_of_moments, model.number_of_random_vectors};
    dump_i("meta", meta, 4);
    double fmeta[2] = {model.energy_max, model.volume};
    dump_d("fmeta", fmeta, 2);
    dump_i("neighbor_number", model.neighbor_number, N);
    std::vector<int> nl;
    std::vector<double> hr, hi, xx;
    for (int n = 0; n < N; ++n)
      for (int m = 0; m < model.neighbor_number[n]; ++m) {
        int q = n * K + m;
        nl.push_back(model.neighbor_list[q]);
        hr.push_back(model.hopping_real[q]);
        hi.push_back(model.hopping_imag[q]);
        xx.push_back(model.xx[q]);
      }
    dump_i("neighbor_list", nl.data(), (int)nl.size());
    dump_d("hopping_real", hr.data(), (int)hr.size());
    dump_d("hopping_imag", hi.data(), (int)hi.size());
    dump_d("xx", xx.data(), (int)xx.size());
    dump_d("potential", model.potential, N);
  }

  Hamiltonian H(model);
  Vector random_state(N);
  int Nt = model.number_of_steps_correlation;
  for (int i = 0; i < model.number_of_random_vectors; ++i) {
    std::string v = "v" + std::to_string(i);
    model.initialize_state(random_state, -1);
    dump_vec(v + "_random", random_state, N);

    // single operator applications and single evolutions of the random vector
    {
      Vector o(N);
      H.apply(random_state, o);
      dump_vec(v + "_apply", o, N);
      H.apply_commutator(random_state, o);
      dump_vec(v + "_commutator", o, N);
      H.apply_current(random_state, o);
      dump_vec(v + "_current", o, N);
      Vector s1(N), s2(N);
      H.apply(random_state, s1);
      Vector s0(random_state);
      H.kernel_polynomial(s0, s1, s2);
      dump_vec(v + "_kpoly", s2, N);
      if (Nt > 0) {
        real t = model.time_step[0] * model.energy_max;
        Vector e(random_state);
        evolve(model, 1, t, H, e);
        dump_vec(v + "_evolve_p", e, N);
        Vector e2(random_state);
        evolve(model, -1, t, H, e2);
        dump_vec(v + "_evolve_m", e2, N);
        Vector ex(random_state);
        evolvex(model, 1, t, H, ex);
        dump_vec(v + "_evolvex_p", ex, N);
        Vector ex2(random_state);
        evolvex(model, -1, t, H, ex2);
        dump_vec(v + "_evolvex_m", ex2, N);
      }
    }

    // find_dos (src/sigma.cu:204-239)
    moment_set(model, H, random_state, random_state, v + "_dos");

    // find_vac0 (src/sigma.cu:243-274)
    if (model.calculate_vac0) {
      Vector state(N);
      H.apply_current(random_state, state);
      moment_set(model, H, state, state, v + "_vac0");
    }

    // find_vac (src/sigma.cu:278-327)
    if (model.calculate_vac) {
      Vector left(random_state), left_copy(N), right(random_state);
      H.apply_current(left, right);
      for (int m = 0; m < Nt; ++m) {
        H.apply_current(left, left_copy);
        moment_set(model, H, right, left_copy, v + "_vac" + std::to_string(m));
        if (m < Nt - 1) {
          real t = model.time_step[m] * model.energy_max;
          evolve(model, -1, t, H, left);
          evolve(model, -1, t, H, right);
        }
      }
    }

    // find_msd (src/sigma.cu:331-391)
    if (model.calculate_msd) {
      Vector state(random_state), state_x(random_state), state_copy(N);
      real t = model.time_step[0] * model.energy_max;
      evolve(model, 1, t, H, state);
      evolvex(model, 1, t, H, state_x);
      for (int m = 0; m < Nt; ++m) {
        moment_set(model, H, state_x, state_x, v + "_msd" + std::to_string(m));
        if (m < Nt - 1) {
          t = model.time_step[m + 1] * model.energy_max;
          state_copy.copy(state);
          evolvex(model, 1, t, H, state_copy);
          evolve(model, 1, t, H, state_x);
          state_x.add(state_copy);
          evolve(model, 1, t, H, state);
        }
      }
    }

    // find_spin_polarization (src/sigma.cu:395-442)
    if (model.calculate_spin) {
      Vector state(random_state), state_sz(N);
      for (int m = 0; m < Nt; ++m) {
        state_sz.apply_sz(state);
        moment_set(model, H, state, state_sz, v + "_spin" + std::to_string(m));
        if (m < Nt - 1) {
          real t = model.time_step[m] * model.energy_max;
          evolve(model, 1, t, H, state);
        }
      }
    }
  }

  // LDOS (src/lsqt.cu:118-133)
  if (model.calculate_ldos) {
    for (int i = 0; i < model.number_of_local_orbitals; ++i) {
      model.initialize_state(random_state, model.local_orbitals[i]);
      moment_set(model, H, random_state, random_state, "ldos" + std::to_string(i));
    }
  }
  fclose(out);
  return 0;
}
